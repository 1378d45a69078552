// rcr_f64_w8.hip -- the FP64 translation kernels with 8 waves per workgroup
// (rcr_kernels.h), one translation unit per precision and wave count so
// that they compile in parallel.
#include "rcr_kernels.h"

int fmbh_rcr_detail::run_f64_w8(const fmbh_rcr_job *j, int nrhs, hipStream_t st) {
  return dispatch_shape<double, 8>(j, nrhs, st);
}
