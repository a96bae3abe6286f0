// Pair kernels for xd = 2.
#define FBPX_INST_XD 2
#define FBPX_INST_JETS FBPX_JETS_XD2
#include "fbpx_inst.inc"
