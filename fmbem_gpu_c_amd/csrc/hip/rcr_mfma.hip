// rcr_mfma.hip -- host entry points of the rotate-coaxial-rotate translations
// on the gfx950 matrix cores (kernels in rcr_kernels.h): tile-shape rules,
// dispatch by precision and wave count, and the zero-padded operator copies.
#include "rcr_kernels.h"

namespace {

// Even/odd rotation blocks of degree n from the FP64 blocks H_n (column-
// major s x s, s = 2n+1, H_n^{m'm} at [(m+n) s + m'+n]), rounded once to T.
// With H_n^{m'm} = H_n^{-m',-m} and the coordinates e_m = (z_m + z_{-m})/2
// (m > 0), e_0 = z_0, o_m = (z_m - z_{-m})/2 (m > 0), y = H_n z becomes
//   e(y)_{m'} = sum_m E_{m'm} e(z)_m,  E_{m'0} = H^{m'0},
//                                     E_{m'm} = H^{m'm} + H^{m',-m} (m > 0);
//   o(y)_{m'} = sum_m O_{m'm} o(z)_m,  O_{m'm} = H^{m'm} - H^{m',-m}.
// E's index i stands for m = i + 1 (i < n) and m = 0 (i = n), O's for
// |m| = i + 1, so that E and O row i of a tile meet in the same lane.  Each
// block zero-padded to 16 rows / 4 columns, column-major with leading
// dimension the padded rows; E then O.
template <typename T>
__global__ void k_pad_rot(const double *__restrict__ H, size_t hs, const int *__restrict__ hoff,
                          size_t ps, T *__restrict__ Hp) {
  const int n = blockIdx.x, b = blockIdx.y;
  const int s = 2 * n + 1;
  const double *src = H + hs * b + rot_off(n);
  auto h = [&](int mp, int m) { return src[(size_t)(m + n) * s + (mp + n)]; };
  T *dst = Hp + ps * b + hoff[n];
  {
    const int se = n + 1, ld = (se + 15) & ~15, nc = (se + 3) & ~3;
    for (int e = threadIdx.x; e < ld * nc; e += blockDim.x) {
      const int i = e % ld, k = e / ld;
      double v = 0.0;
      if (i < se && k < se) {
        const int mp = i < n ? i + 1 : 0, m = k < n ? k + 1 : 0;
        v = m ? h(mp, m) + h(mp, -m) : h(mp, 0);
      }
      dst[e] = (T)v;
    }
  }
  if (n > 0) {
    const int so = n, ld = (so + 15) & ~15, nc = (so + 3) & ~3;
    T *od = dst + eo_block(n + 1);
    for (int e = threadIdx.x; e < ld * nc; e += blockDim.x) {
      const int i = e % ld, k = e / ld;
      double v = 0.0;
      if (i < so && k < so) v = h(i + 1, k + 1) - h(i + 1, -(k + 1));
      od[e] = (T)v;
    }
  }
}

// split-BF16 copies of the standard rotation blocks H_n (from FP64, via the
// FP32 value h): row-major, rows padded to 16, K to 32, hi = bf16(h),
// lo = bf16(h - hi)
__global__ void k_split_rot(const double *__restrict__ H, size_t hs,
                            const int *__restrict__ hboff, size_t half, __bf16 *__restrict__ Hb) {
  const int n = blockIdx.x, b = blockIdx.y;
  const int s = 2 * n + 1, ld = (s + 15) & ~15, kp = (s + 31) & ~31;
  const double *src = H + hs * b + rot_off(n);
  __bf16 *hi = Hb + 2 * half * b + hboff[n];
  for (int e = threadIdx.x; e < ld * kp; e += blockDim.x) {
    const int r = e / kp, k = e % kp;
    const float v = (r < s && k < s) ? (float)src[(size_t)k * s + r] : 0.0f;
    const __bf16 h = (__bf16)v;
    hi[e] = h;
    hi[e + half] = (__bf16)(v - (float)h);
  }
}

template <typename T>
__global__ void k_pad_coax(const cplx<T> *__restrict__ C, size_t cs, int P, int pin,
                           int pout, const int *__restrict__ coff, size_t ps,
                           cplx<T> *__restrict__ Cp) {
  const int am = blockIdx.x, d = blockIdx.y;
  const int so = pout - am, si = pin - am, ldc = P - am;
  const int ld = (so + 15) & ~15, nc = (si + 3) & ~3;
  const cplx<T> *src = C + cs * d + coax_blk(P, am);
  cplx<T> *dst = Cp + ps * d + coff[am];
  for (int e = threadIdx.x; e < ld * nc; e += blockDim.x) {
    const int i = e % ld, k = e / ld;
    dst[e] = (i < so && k < si) ? src[k * ldc + i] : cmk<T>(0, 0);
  }
}

}  // namespace

extern "C" int fmbh_rcr_shape(int pm, int prec, int *mode, int *nt, int *mt) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  const int slen = 8 * pm + 16; /* whole-block schedule bound */
  *mt = mt_lds(pm);
  if (*mt && x_bytes(pm, 2, tsz) + small_lds(pm, tsz, slen) <= 80 * 1024) {
    *mode = FMBH_RCR_LDS;
    *nt = 2;
  } else if (*mt && x_bytes(pm, 1, tsz) + small_lds(pm, tsz, slen) <= 160 * 1024) {
    *mode = FMBH_RCR_LDS;
    *nt = 1;
  } else {
    *mode = FMBH_RCR_GWS;
    *nt = 2;
    *mt = 4;
  }
  return FMB_OK;
}

/* LDS-mode bytes of a workgroup (0: the block rows need more than 6 tiles,
 * the largest LDS-mode unit) */
extern "C" size_t fmbh_rcr_lds_bytes(int pm, int prec, int nt, int sched_len) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  if (!mt_lds(pm)) return 0;
  return x_bytes(pm, nt, tsz) + small_lds(pm, tsz, sched_len);
}

extern "C" int fmbh_rcr_mt_lds(int pm, int bf) { return mt_lds(pm, bf); }

extern "C" size_t fmbh_rcr_ldso_bytes(int pm, int prec, int nt, int sched_len, int hlen,
                                      size_t clen) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  if (!mt_lds(pm)) return 0;
  return ldso_bytes(pm, nt, tsz, sched_len, hlen, clen);
}

extern "C" size_t fmbh_rcr_gws_bytes(int pm, int prec, int nt) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  return 2 * x_bytes(pm, nt, tsz);
}

extern "C" int fmbh_rcr(const fmbh_rcr_job *j, int prec, int nrhs, void *stream) {
  using namespace fmbh_rcr_detail;
  if (j->nchunks == 0) return FMB_OK;
  if (j->nw != 4 && j->nw != 8)
    return fmb_set_error(FMB_ERR_RANGE, "no RCR tile shape with %d waves", j->nw);
  hipStream_t st = (hipStream_t)stream;
  if (j->bf) {
    if (prec != FMB_MIXED)
      return fmb_set_error(FMB_ERR_ARG, "split-BF16 translations are for the FP32 pass");
    return j->nw == 8 ? run_bf_w8(j, nrhs, st) : run_bf_w4(j, nrhs, st);
  }
  if (prec == FMB_FP64) return j->nw == 8 ? run_f64_w8(j, nrhs, st) : run_f64_w4(j, nrhs, st);
  return j->nw == 8 ? run_f32_w8(j, nrhs, st) : run_f32_w4(j, nrhs, st);
}

extern "C" int fmbh_pad_rot(int prec, const double *d_H64, int nbeta, int Pmax,
                            const int *d_hoff, size_t pstride, void *d_Hp,
                            void *stream) {
  hipStream_t st = (hipStream_t)stream;
  size_t hs = (size_t)rot_off(Pmax);
  if (prec == FMB_FP64)
    hipLaunchKernelGGL(k_pad_rot<double>, dim3(Pmax, nbeta), dim3(256), 0, st,
                       d_H64, hs, d_hoff, pstride, (double *)d_Hp);
  else
    hipLaunchKernelGGL(k_pad_rot<float>, dim3(Pmax, nbeta), dim3(256), 0, st,
                       d_H64, hs, d_hoff, pstride, (float *)d_Hp);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_pad_coax(int prec, const void *d_C, int ncoax, int P, int pin,
                             int pout, const int *d_coff, size_t pstride,
                             void *d_Cp, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  const int mn = pin < pout ? pin : pout;
  size_t cs = (size_t)coax_size(P);
  if (prec == FMB_FP64)
    hipLaunchKernelGGL(k_pad_coax<double>, dim3(mn, ncoax), dim3(256), 0, st,
                       (const cplx<double> *)d_C, cs, P, pin, pout, d_coff, pstride,
                       (cplx<double> *)d_Cp);
  else
    hipLaunchKernelGGL(k_pad_coax<float>, dim3(mn, ncoax), dim3(256), 0, st,
                       (const cplx<float> *)d_C, cs, P, pin, pout, d_coff, pstride,
                       (cplx<float> *)d_Cp);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_split_rot(const double *d_H64, int nbeta, int Pmax, const int *d_hboff,
                              size_t half, void *d_Hb, void *stream) {
  hipLaunchKernelGGL(k_split_rot, dim3(Pmax, nbeta), dim3(256), 0, (hipStream_t)stream, d_H64,
                     (size_t)rot_off(Pmax), d_hboff, half, (__bf16 *)d_Hb);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}
