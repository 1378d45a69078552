// rcr_bf_w8.hip -- the FP32 translation kernels with split-BF16 rotations
// and 8 waves per workgroup (rcr_kernels.h), in a translation unit of their
// own so that they compile in parallel.
#include "rcr_kernels.h"

int fmbh_rcr_detail::run_bf_w8(const fmbh_rcr_job *j, int nrhs, hipStream_t st) {
  return dispatch_shape<float, 8, true>(j, nrhs, st);
}
