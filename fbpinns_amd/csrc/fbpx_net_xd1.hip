// Pair kernels for xd = 1.
#define FBPX_INST_XD 1
#define FBPX_INST_JETS FBPX_JETS_XD1
#include "fbpx_inst.inc"
