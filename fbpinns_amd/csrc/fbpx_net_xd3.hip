// Pair kernels for xd = 3.
#define FBPX_INST_XD 3
#define FBPX_INST_JETS FBPX_JETS_XD3
#include "fbpx_inst.inc"
