// clear_fwdk_pw0.hip -- tile set 0 of the fused clear-sky forward / K-matrix
// kernel (4 FP32 / 2 FP64 profiles per workgroup), compiled apart so the three
// tile sets build in parallel.
#define CRTM_FWDK_TILE 0
#include "clear_fwdk.hip"
