// clear_fwdk_pw2.hip -- tile set 2 of the fused clear-sky forward / K-matrix
// kernel (16 FP32 / 8 FP64 profiles per workgroup), compiled apart so the three
// tile sets build in parallel.
#define CRTM_FWDK_TILE 2
#include "clear_fwdk.hip"
