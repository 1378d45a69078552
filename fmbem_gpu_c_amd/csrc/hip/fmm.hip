// fmm.hip -- the FMM pass on gfx950: gather/scatter between caller and
// Morton element order, P2M, M2M ("SS"), M2L ("SR"), L2L ("RR"), L2P fused
// with the near-field block product.
//
// Translations are rotate-coaxial-rotate (translation.cu:1711-1808,
// 3461-3759): with t = (|t|, theta, phi),
//   a_n^m  = e^{i m phi} x_n^m                     (rotation alpha = phi)
//   b_n^m' = sum_m H_n^{m' m}(theta) a_n^m          (beta = theta, gamma = 0)
//   c_n'^m = sum_n C^{|m|}_{n' n}(|t|) b_n^m        (coaxial block |m|)
//   y_n^m' = e^{-i m' phi} sum_m H_n^{m' m}(theta) c_n^m   (0, theta, phi)
// on n-dependent scaled coefficients (M~ = M / w_n, L~ = w_n L) so that FP32
// operators stay inside the float range.  One translation job (box pair x
// right-hand side) runs on one workgroup of GS lanes: the vector is staged in
// LDS, each lane owns R outputs in registers between the three stages, and
// the rotation / coaxial matrices are read with lanes along rows of the
// column-major reference layouts (coalesced).  M2L jobs are grouped by offset
// (one H, one coaxial matrix per workgroup chunk) and accumulate into the
// target with hardware float atomics.
#include "fmb_device.h"

namespace {

// One rotate-coaxial-rotate translation of vector x (order pin) to order
// pout.  H: rotation blocks for theta (prefix layout), C: coaxial matrix at
// order P >= pin, pout (|m|-block layout), mtab: m-major enumeration of the
// pout^2 outputs (natural indices).  buf: LDS, max(pin,pout)^2 entries;
// ph: LDS, max(pin,pout) entries e^{i m phi}.  Leaves out[r] = output with
// natural index lane + GS*r.
template <typename T, int GS, int R>
__device__ __forceinline__ void rcr_job(const T *__restrict__ H,
                                        const cplx<T> *__restrict__ C, int P,
                                        int pin, int pout,
                                        const cplx<T> *__restrict__ x,
                                        const int *__restrict__ mtab,
                                        cplx<T> *buf, const cplx<T> *ph,
                                        cplx<T> out[R]) {
  const int lane = threadIdx.x;
  const int nin = pin * pin, nout = pout * pout;
  // 1. stage with forward phase e^{i m phi}
  for (int f = lane; f < nin; f += GS) {
    int n = isqrt_fix(f), m = f - n * n - n;
    cplx<T> e = ph[m < 0 ? -m : m];
    if (m < 0) e.im = -e.im;
    buf[f] = cmul(x[f], e);
  }
  __syncthreads();
  // 2. forward rotation, block n: b^{m'} = sum_m H^{m' m} a^m
#pragma unroll
  for (int r = 0; r < R; r++) {
    int f = lane + GS * r;
    cplx<T> s = cmk<T>(0, 0);
    if (f < nin) {
      int n = isqrt_fix(f), mp = f - n * n - n, ld = 2 * n + 1;
      const T *h = H + rot_off(n) + (mp + n);
      const cplx<T> *a = buf + n * n;
      for (int m = 0; m < ld; m++) rfma(s, h[m * ld], a[m]);
    }
    out[r] = s;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; r++) {
    int f = lane + GS * r;
    if (f < nin) buf[f] = out[r];
  }
  __syncthreads();
  // 3. coaxial, m-major outputs: c_{n'}^m = sum_{n>=|m|} C^{|m|}(n', n) b_n^m
#pragma unroll
  for (int r = 0; r < R; r++) {
    int g = lane + GS * r;
    cplx<T> s = cmk<T>(0, 0);
    if (g < nout) {
      int nat = mtab[g];
      int np = isqrt_fix(nat), m = nat - np * np - np, am = m < 0 ? -m : m;
      int ld = P - am;
      const cplx<T> *c = C + coax_blk(P, am) + (np - am);
      for (int n = am; n < pin; n++) cfma(s, c[(n - am) * ld], buf[n * n + n + m]);
    }
    out[r] = s;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; r++) {
    int g = lane + GS * r;
    if (g < nout) buf[mtab[g]] = out[r];
  }
  __syncthreads();
  // 4. back rotation with output phase e^{-i m' phi}
#pragma unroll
  for (int r = 0; r < R; r++) {
    int f = lane + GS * r;
    cplx<T> s = cmk<T>(0, 0);
    if (f < nout) {
      int n = isqrt_fix(f), mp = f - n * n - n, ld = 2 * n + 1;
      const T *h = H + rot_off(n) + (mp + n);
      const cplx<T> *a = buf + n * n;
      for (int m = 0; m < ld; m++) rfma(s, h[m * ld], a[m]);
      cplx<T> e = ph[mp < 0 ? -mp : mp];
      if (mp > 0) e.im = -e.im;
      s = cmul(s, e);
    }
    out[r] = s;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void fill_phase(cplx<T> *ph, int np_, double phi) {
  for (int m = threadIdx.x; m < np_; m += blockDim.x) {
    double s, c;
    sincos(m * phi, &s, &c);
    ph[m] = cmk<T>((T)c, (T)s);
  }
}

// M2L: blocks take chunks (offset, first pair, last pair) of one offset bucket
template <typename T, int GS, int R>
__global__ void __launch_bounds__(GS)
    k_m2l(const int4 *__restrict__ chunks, const OffInfo *__restrict__ info,
          const int *__restrict__ ptgt, const int *__restrict__ psrc,
          const T *__restrict__ Hall, size_t hstride,
          const cplx<T> *__restrict__ Call, size_t cstride, int p,
          const int *__restrict__ mtab, const cplx<T> *__restrict__ Min,
          cplx<T> *__restrict__ Lout, int nrhs) {
  extern __shared__ unsigned char smem[];
  cplx<T> *buf = (cplx<T> *)smem;
  cplx<T> *ph = buf + p * p;
  int4 ch = chunks[blockIdx.x];
  OffInfo oi = info[ch.x];
  const T *H = Hall + hstride * oi.beta;
  const cplx<T> *C = Call + cstride * oi.dist;
  fill_phase<T>(ph, p, oi.phi);
  __syncthreads();
  const int ps = p * p;
  cplx<T> out[R];
  for (int q = ch.y; q < ch.z; q++) {
    int tg = ptgt[q], sr = psrc[q];
    for (int r = 0; r < nrhs; r++) {
      rcr_job<T, GS, R>(H, C, p, p, p, Min + ((size_t)sr * nrhs + r) * ps, mtab,
                        buf, ph, out);
      cplx<T> *dst = Lout + ((size_t)tg * nrhs + r) * ps;
#pragma unroll
      for (int k = 0; k < R; k++) {
        int f = threadIdx.x + GS * k;
        if (f < ps) atomic_add_c(dst + f, out[k]);
      }
    }
  }
}

// M2M: one block per parent, children of level l -> parent at level l-1
template <typename T, int GS, int R>
__global__ void __launch_bounds__(GS)
    k_m2m(const int *__restrict__ cstart, const int *__restrict__ ckeys,
          const int *__restrict__ obeta, const double *__restrict__ ophi,
          const T *__restrict__ Hall, size_t hstride,
          const cplx<T> *__restrict__ C, int P, int pc, int pp,
          const int *__restrict__ mtab, const cplx<T> *__restrict__ Mc,
          cplx<T> *__restrict__ Mp, int nrhs) {
  extern __shared__ unsigned char smem[];
  int Pm = pc > pp ? pc : pp;
  cplx<T> *buf = (cplx<T> *)smem;
  cplx<T> *ph = buf + Pm * Pm;
  int pb = blockIdx.x;
  int c0 = cstart[pb], c1 = cstart[pb + 1];
  for (int r = 0; r < nrhs; r++) {
    cplx<T> acc[R];
#pragma unroll
    for (int k = 0; k < R; k++) acc[k] = cmk<T>(0, 0);
    for (int c = c0; c < c1; c++) {
      int oct = ckeys[c] & 7;
      __syncthreads();
      fill_phase<T>(ph, Pm, ophi[oct]);
      __syncthreads();
      cplx<T> out[R];
      rcr_job<T, GS, R>(Hall + hstride * obeta[oct], C, P, pc, pp,
                        Mc + ((size_t)c * nrhs + r) * pc * pc, mtab, buf, ph,
                        out);
#pragma unroll
      for (int k = 0; k < R; k++) acc[k] = cadd(acc[k], out[k]);
    }
    cplx<T> *dst = Mp + ((size_t)pb * nrhs + r) * pp * pp;
#pragma unroll
    for (int k = 0; k < R; k++) {
      int f = threadIdx.x + GS * k;
      if (f < pp * pp) dst[f] = acc[k];
    }
  }
}

// L2L: one block per child, parent at level l-1 -> child at level l (adds)
template <typename T, int GS, int R>
__global__ void __launch_bounds__(GS)
    k_l2l(const int *__restrict__ parent, const int *__restrict__ ckeys,
          const int *__restrict__ obeta, const double *__restrict__ ophi,
          const T *__restrict__ Hall, size_t hstride,
          const cplx<T> *__restrict__ C, int P, int pp, int pc,
          const int *__restrict__ mtab, const cplx<T> *__restrict__ Lp,
          cplx<T> *__restrict__ Lc, int nrhs) {
  extern __shared__ unsigned char smem[];
  int Pm = pc > pp ? pc : pp;
  cplx<T> *buf = (cplx<T> *)smem;
  cplx<T> *ph = buf + Pm * Pm;
  int cb = blockIdx.x;
  int oct = ckeys[cb] & 7, pb = parent[cb];
  fill_phase<T>(ph, Pm, ophi[oct]);
  __syncthreads();
  for (int r = 0; r < nrhs; r++) {
    cplx<T> out[R];
    rcr_job<T, GS, R>(Hall + hstride * obeta[oct], C, P, pp, pc,
                      Lp + ((size_t)pb * nrhs + r) * pp * pp, mtab, buf, ph,
                      out);
    cplx<T> *dst = Lc + ((size_t)cb * nrhs + r) * pc * pc;
#pragma unroll
    for (int k = 0; k < R; k++) {
      int f = threadIdx.x + GS * k;
      if (f < pc * pc) dst[f] = cadd(dst[f], out[k]);
    }
  }
}

// P2M: one block per leaf box, M[b][r][f] = sum_{e in b} B[e][f] xs[r][e].
// FLIP multiplies coefficient f = (n, m) by (-1)^n: with B = the L2P basis
// it is the transposed L2P of the transposed pass.
template <typename T, bool FLIP>
__global__ void k_p2m(const int *__restrict__ bstart,
                      const cplx<T> *__restrict__ B, const cplx<T> *__restrict__ xs,
                      int N, int ps, cplx<T> *__restrict__ M, int nrhs) {
  int b = blockIdx.x;
  int e0 = bstart[b], e1 = bstart[b + 1];
  for (int r = 0; r < nrhs; r++)
    for (int f = threadIdx.x; f < ps; f += blockDim.x) {
      cplx<T> s = cmk<T>(0, 0);
      for (int e = e0; e < e1; e++)
        cfma(s, B[(size_t)e * ps + f], xs[(size_t)r * N + e]);
      if (FLIP && (isqrt_fix(f) & 1)) s = cmk<T>(-s.re, -s.im);
      M[((size_t)b * nrhs + r) * ps + f] = s;
    }
}

template <typename T>
__global__ void k_gather(const int *__restrict__ perm, const double2 *__restrict__ x,
                         int N, int nrhs, cplx<T> *__restrict__ xs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * nrhs) return;
  int r = (int)(i / N), e = (int)(i % N);
  double2 v = x[(size_t)r * N + perm[e]];
  xs[i] = cmk<T>((T)v.x, (T)v.y);
}

template <typename T>
__global__ void k_scatter(const int *__restrict__ perm,
                          const cplx<T> *__restrict__ ys, int N, int nrhs,
                          double2 *__restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * nrhs) return;
  int r = (int)(i / N), e = (int)(i % N);
  cplx<T> v = ys[i];
  y[(size_t)r * N + perm[e]] = make_double2((double)v.re, (double)v.im);
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

struct Shape {
  int gs, r;
};
inline Shape shape_for(int pm) {
  int q = pm * pm;
  if (q > 512 * 32) return {0, 0};
  if (q <= 256) return {64, 4};
  if (q <= 512) return {64, 8};
  if (q <= 1024) return {256, 4};
  if (q <= 2048) return {256, 8};
  if (q <= 4096) return {256, 16};
  if (q <= 8192) return {512, 16};
  return {512, 32};
}

template <typename K>
int set_lds(K kern, size_t bytes) {
  if (bytes > 160 * 1024)
    return fmb_set_error(FMB_ERR_RANGE, "translation needs %zu B of LDS", bytes);
  if (bytes > 64 * 1024) /* the CU's 160 KB: valid for every captured launch */
    FMB_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024));
  return FMB_OK;
}

}  // namespace

// ---------------------------------------------------------------- launchers

struct fmbh_m2l_plan;  // opaque per-plan launch tables are kept on fmb_plan

extern "C" int fmbh_mmajor_table(int p, int **d_tab, void *stream) {
  int *h = (int *)malloc(sizeof(int) * p * p);
  int g = 0;
  for (int m = -(p - 1); m <= p - 1; m++) {
    int am = m < 0 ? -m : m;
    for (int np = am; np < p; np++) h[g++] = np * np + np + m;
  }
  int rc = fmbh_malloc((void **)d_tab, sizeof(int) * p * p);
  if (!rc) rc = fmbh_h2d(*d_tab, h, sizeof(int) * p * p, stream);
  free(h);
  return rc;
}

extern "C" int fmbh_gather(const fmb_plan *pl, int op, const fmb_c64 *x,
                           int nrhs, void *stream) {
  const fmb_tree *t = pl->tree;
  size_t n = (size_t)t->nelem * nrhs;
  hipStream_t st = (hipStream_t)stream;
  if (pl->ops[op].prec == FMB_FP64)
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(n)), dim3(256), 0, st,
                       t->d_perm, (const double2 *)x, t->nelem, nrhs,
                       (cplx<double> *)pl->ops[op].d_xs);
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(nblk(n)), dim3(256), 0, st,
                       t->d_perm, (const double2 *)x, t->nelem, nrhs,
                       (cplx<float> *)pl->ops[op].d_xs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_scatter(const fmb_plan *pl, int op, fmb_c64 *y, int nrhs,
                            void *stream) {
  const fmb_tree *t = pl->tree;
  size_t n = (size_t)t->nelem * nrhs;
  hipStream_t st = (hipStream_t)stream;
  if (pl->ops[op].prec == FMB_FP64)
    hipLaunchKernelGGL(k_scatter<double>, dim3(nblk(n)), dim3(256), 0, st,
                       t->d_perm, (const cplx<double> *)pl->ops[op].d_ys,
                       t->nelem, nrhs, (double2 *)y);
  else
    hipLaunchKernelGGL(k_scatter<float>, dim3(nblk(n)), dim3(256), 0, st,
                       t->d_perm, (const cplx<float> *)pl->ops[op].d_ys,
                       t->nelem, nrhs, (double2 *)y);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_p2m(const fmb_plan *pl, int op, int nrhs, void *stream) {
  const fmb_tree *t = pl->tree;
  int L = t->lmax, pL = pl->p[L];
  hipStream_t st = (hipStream_t)stream;
  const fmb_opset *o = &pl->ops[op];
  int bs = pL * pL >= 256 ? 256 : (pL * pL >= 128 ? 128 : 64);
  if (o->prec == FMB_FP64)
    hipLaunchKernelGGL((k_p2m<double, false>), dim3(t->nbox[L]), dim3(bs), 0, st,
                       t->d_box_start, (const cplx<double> *)o->d_P2M,
                       (const cplx<double> *)o->d_xs, t->nelem, pL * pL,
                       (cplx<double> *)o->d_M[L], nrhs);
  else
    hipLaunchKernelGGL((k_p2m<float, false>), dim3(t->nbox[L]), dim3(bs), 0, st,
                       t->d_box_start, (const cplx<float> *)o->d_P2M,
                       (const cplx<float> *)o->d_xs, t->nelem, pL * pL,
                       (cplx<float> *)o->d_M[L], nrhs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

// transposed L2P: L'[L] = (-1)^n L2P^T xs, the start of the transposed pass
extern "C" int fmbh_l2p_t(const fmb_plan *pl, int op, int nrhs, void *stream) {
  const fmb_tree *t = pl->tree;
  int L = t->lmax, pL = pl->p[L];
  hipStream_t st = (hipStream_t)stream;
  const fmb_opset *o = &pl->ops[op];
  int bs = pL * pL >= 256 ? 256 : (pL * pL >= 128 ? 128 : 64);
  if (o->prec == FMB_FP64)
    hipLaunchKernelGGL((k_p2m<double, true>), dim3(t->nbox[L]), dim3(bs), 0, st,
                       t->d_box_start, (const cplx<double> *)o->d_L2P,
                       (const cplx<double> *)o->d_xs, t->nelem, pL * pL,
                       (cplx<double> *)o->d_L[L], nrhs);
  else
    hipLaunchKernelGGL((k_p2m<float, true>), dim3(t->nbox[L]), dim3(bs), 0, st,
                       t->d_box_start, (const cplx<float> *)o->d_L2P,
                       (const cplx<float> *)o->d_xs, t->nelem, pL * pL,
                       (cplx<float> *)o->d_L[L], nrhs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

template <typename T, int GS, int R>
static int launch_m2l(const fmb_plan *pl, const fmb_opset *o, int l, int nrhs,
                      hipStream_t st) {
  const fmb_tree *t = pl->tree;
  int p = pl->p[l];
  size_t lds = sizeof(cplx<T>) * (size_t)(p * p + p);
  auto kern = k_m2l<T, GS, R>;
  int rc = set_lds(kern, lds);
  if (rc) return rc;
  const int4 *chunks = (const int4 *)pl->d_m2l_chunks[l];
  hipLaunchKernelGGL(kern, dim3(pl->n_m2l_chunks[l]), dim3(GS), lds, st, chunks,
                     (const OffInfo *)pl->d_off_info, t->d_pair_tgt[l],
                     t->d_pair_src[l], (const T *)o->d_H,
                     (size_t)rot_off(pl->Pmax), (const cplx<T> *)o->d_Csr[l],
                     (size_t)coax_size(p), p, pl->d_tab_mmajor[l],
                     (const cplx<T> *)o->d_M[l], (cplx<T> *)o->d_L[l], nrhs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

template <typename T>
static int dispatch_m2l(const fmb_plan *pl, const fmb_opset *o, int l,
                        int nrhs, hipStream_t st) {
  Shape s = shape_for(pl->p[l]);
#define CASE(G, RR)                                                            \
  if (s.gs == G && s.r == RR) return launch_m2l<T, G, RR>(pl, o, l, nrhs, st);
  CASE(64, 4) CASE(64, 8) CASE(256, 4) CASE(256, 8) CASE(256, 16) CASE(512, 16)
  CASE(512, 32)
#undef CASE
  return fmb_set_error(FMB_ERR_RANGE, "no M2L shape for p=%d", pl->p[l]);
}

extern "C" int fmbh_m2l(const fmb_plan *pl, int op, int l, int nrhs,
                        void *stream) {
  if (pl->n_m2l_chunks[l] == 0) return FMB_OK;
  const fmb_opset *o = &pl->ops[op];
  if (o->prec == FMB_FP64)
    return dispatch_m2l<double>(pl, o, l, nrhs, (hipStream_t)stream);
  return dispatch_m2l<float>(pl, o, l, nrhs, (hipStream_t)stream);
}

template <typename T, int GS, int R>
static int launch_m2m(const fmb_plan *pl, const fmb_opset *o, int l, int nrhs,
                      hipStream_t st) {
  const fmb_tree *t = pl->tree;
  int pc = pl->p[l], pp = pl->p[l - 1], Pm = pc > pp ? pc : pp;
  size_t lds = sizeof(cplx<T>) * (size_t)(Pm * Pm + Pm);
  auto kern = k_m2m<T, GS, R>;
  int rc = set_lds(kern, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(t->nbox[l - 1]), dim3(GS), lds, st,
                     t->d_cstart[l - 1], t->d_keys[l], pl->d_m2m_beta,
                     pl->d_m2m_phi, (const T *)o->d_H,
                     (size_t)rot_off(pl->Pmax), (const cplx<T> *)o->d_Cm2m[l],
                     Pm, pc, pp, pl->d_tab_mmajor[l - 1],
                     (const cplx<T> *)o->d_M[l], (cplx<T> *)o->d_M[l - 1], nrhs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

template <typename T>
static int dispatch_m2m(const fmb_plan *pl, const fmb_opset *o, int l,
                        int nrhs, hipStream_t st) {
  int pm = pl->p[l] > pl->p[l - 1] ? pl->p[l] : pl->p[l - 1];
  Shape s = shape_for(pm);
#define CASE(G, RR)                                                            \
  if (s.gs == G && s.r == RR) return launch_m2m<T, G, RR>(pl, o, l, nrhs, st);
  CASE(64, 4) CASE(64, 8) CASE(256, 4) CASE(256, 8) CASE(256, 16) CASE(512, 16)
  CASE(512, 32)
#undef CASE
  return fmb_set_error(FMB_ERR_RANGE, "no M2M shape for p=%d", pm);
}

extern "C" int fmbh_m2m(const fmb_plan *pl, int op, int l, int nrhs,
                        void *stream) {
  const fmb_opset *o = &pl->ops[op];
  if (o->prec == FMB_FP64)
    return dispatch_m2m<double>(pl, o, l, nrhs, (hipStream_t)stream);
  return dispatch_m2m<float>(pl, o, l, nrhs, (hipStream_t)stream);
}

template <typename T, int GS, int R>
static int launch_l2l(const fmb_plan *pl, const fmb_opset *o, int l, int nrhs,
                      hipStream_t st) {
  const fmb_tree *t = pl->tree;
  int pc = pl->p[l], pp = pl->p[l - 1], Pm = pc > pp ? pc : pp;
  size_t lds = sizeof(cplx<T>) * (size_t)(Pm * Pm + Pm);
  auto kern = k_l2l<T, GS, R>;
  int rc = set_lds(kern, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(t->nbox[l]), dim3(GS), lds, st, t->d_parent[l],
                     t->d_keys[l], pl->d_l2l_beta, pl->d_l2l_phi,
                     (const T *)o->d_H, (size_t)rot_off(pl->Pmax),
                     (const cplx<T> *)o->d_Cl2l[l], Pm, pp, pc,
                     pl->d_tab_mmajor[l], (const cplx<T> *)o->d_L[l - 1],
                     (cplx<T> *)o->d_L[l], nrhs);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

template <typename T>
static int dispatch_l2l(const fmb_plan *pl, const fmb_opset *o, int l,
                        int nrhs, hipStream_t st) {
  int pm = pl->p[l] > pl->p[l - 1] ? pl->p[l] : pl->p[l - 1];
  Shape s = shape_for(pm);
#define CASE(G, RR)                                                            \
  if (s.gs == G && s.r == RR) return launch_l2l<T, G, RR>(pl, o, l, nrhs, st);
  CASE(64, 4) CASE(64, 8) CASE(256, 4) CASE(256, 8) CASE(256, 16) CASE(512, 16)
  CASE(512, 32)
#undef CASE
  return fmb_set_error(FMB_ERR_RANGE, "no L2L shape for p=%d", pm);
}

extern "C" int fmbh_l2l(const fmb_plan *pl, int op, int l, int nrhs,
                        void *stream) {
  const fmb_opset *o = &pl->ops[op];
  if (o->prec == FMB_FP64)
    return dispatch_l2l<double>(pl, o, l, nrhs, (hipStream_t)stream);
  return dispatch_l2l<float>(pl, o, l, nrhs, (hipStream_t)stream);
}
