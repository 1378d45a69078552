// blas.hip -- FP64 complex vector kernels for GMRES (batched over right-hand
// sides) and the plane-wave right-hand side.  Reductions are two-stage with
// a fixed chunk count so results do not depend on scheduling.
#include "fmb_device.h"

namespace {

constexpr int CHUNKS = 128;
constexpr int BS = 256;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// partial[(j*nrhs + r)*CHUNKS + c] = sum over chunk c of conj(V_j^r) w^r
__global__ void __launch_bounds__(BS)
    k_zdots_part(const double2 *__restrict__ V, int nrhs, int N,
                 const double2 *__restrict__ w, double2 *__restrict__ part) {
  __shared__ double sh[BS];
  int c = blockIdx.x, jr = blockIdx.y;
  int j = jr / nrhs, r = jr % nrhs;
  const double2 *v = V + ((size_t)j * nrhs + r) * N;
  const double2 *x = w + (size_t)r * N;
  size_t per = ((size_t)N + CHUNKS - 1) / CHUNKS;
  size_t i0 = per * c, i1 = i0 + per < (size_t)N ? i0 + per : (size_t)N;
  double sr = 0, si = 0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += BS) {
    double2 a = v[i], b = x[i];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  sr = block_sum(sr, sh);
  si = block_sum(si, sh);
  if (threadIdx.x == 0) part[(size_t)jr * CHUNKS + c] = make_double2(sr, si);
}

__global__ void k_sum_chunks(const double2 *__restrict__ part, int n,
                             double2 *__restrict__ out) {
  int jr = blockIdx.x * blockDim.x + threadIdx.x;
  if (jr >= n) return;
  double sr = 0, si = 0;
  for (int c = 0; c < CHUNKS; c++) {
    double2 v = part[(size_t)jr * CHUNKS + c];
    sr += v.x;
    si += v.y;
  }
  out[jr] = make_double2(sr, si);
}

__global__ void __launch_bounds__(BS)
    k_znorm_part(const double2 *__restrict__ w, int N, double *__restrict__ part) {
  __shared__ double sh[BS];
  int c = blockIdx.x, r = blockIdx.y;
  const double2 *x = w + (size_t)r * N;
  size_t per = ((size_t)N + CHUNKS - 1) / CHUNKS;
  size_t i0 = per * c, i1 = i0 + per < (size_t)N ? i0 + per : (size_t)N;
  double s = 0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += BS) {
    double2 a = x[i];
    s += a.x * a.x + a.y * a.y;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[(size_t)r * CHUNKS + c] = s;
}

__global__ void k_norm_final(const double *__restrict__ part, int n,
                             double *__restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0;
  for (int c = 0; c < CHUNKS; c++) s += part[(size_t)r * CHUNKS + c];
  out[r] = sqrt(s);
}

__global__ void k_gemv(const double2 *__restrict__ V, int nv,
                       const double2 *__restrict__ h, double2 *__restrict__ w,
                       int nrhs, int N, double sign) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * nrhs) return;
  int r = (int)(t / N);
  size_t i = t % N;
  double sr = 0, si = 0;
  for (int j = 0; j < nv; j++) {
    double2 a = h[(size_t)j * nrhs + r], v = V[((size_t)j * nrhs + r) * N + i];
    sr += a.x * v.x - a.y * v.y;
    si += a.x * v.y + a.y * v.x;
  }
  w[t].x += sign * sr;
  w[t].y += sign * si;
}

__global__ void k_axpby(const double2 *__restrict__ a, const double2 *__restrict__ x,
                        const double2 *__restrict__ b, double2 *__restrict__ y,
                        int nrhs, int N) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * nrhs) return;
  int r = (int)(t / N);
  double2 av = a[r], xv = x[t];
  double2 res = make_double2(av.x * xv.x - av.y * xv.y, av.x * xv.y + av.y * xv.x);
  if (b) {
    double2 bv = b[r], yv = y[t];
    res.x += bv.x * yv.x - bv.y * yv.y;
    res.y += bv.x * yv.y + bv.y * yv.x;
  }
  y[t] = res;
}

__global__ void k_rhs_plane(const double *__restrict__ geo,
                            const int *__restrict__ iperm, const double *__restrict__ dirs,
                            int nrhs, int N, double k, double alre, double alim,
                            double2 *__restrict__ b) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * nrhs) return;
  int r = (int)(t / N), e = (int)(t % N);
  const double *g = geo + 8 * (size_t)iperm[e];
  const double *d = dirs + 3 * r;
  double dx = d[0] * g[0] + d[1] * g[1] + d[2] * g[2];
  double dn = d[0] * g[3] + d[1] * g[4] + d[2] * g[5];
  double s, c;
  sincos(k * dx, &s, &c);
  // e^{ik d.x} (1 + alpha i k d.n)
  double fr = 1.0 - alim * k * dn, fi = alre * k * dn;
  b[t] = make_double2(c * fr - s * fi, c * fi + s * fr);
}

// x0 = x e^{i dk d_r . c_e}, caller order; in place when x0 == x
__global__ void k_sweep_guess(const double *__restrict__ geo,
                              const int *__restrict__ iperm,
                              const double *__restrict__ dirs, int nrhs, int N,
                              double dk, const double2 *x, double2 *x0) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * nrhs) return;
  int r = (int)(t / N), e = (int)(t % N);
  double2 v = x[t];
  if (dirs) {
    const double *g = geo + 8 * (size_t)iperm[e];
    const double *d = dirs + 3 * r;
    double s, c;
    sincos(dk * (d[0] * g[0] + d[1] * g[1] + d[2] * g[2]), &s, &c);
    v = make_double2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
  x0[t] = v;
}

// dJ/dsigma_j = sum_r 2 conj(w_rj), w_rj = -x_rj (c_rj + s_rj)
__global__ void k_adj_combine(const double2 *__restrict__ x, const double2 *__restrict__ c,
                              const double2 *__restrict__ s, int nrhs, int N,
                              double2 *__restrict__ g) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  double gr = 0.0, gi = 0.0;
  for (int r = 0; r < nrhs; r++) {
    size_t i = (size_t)r * N + j;
    double2 xv = x[i], a = make_double2(c[i].x + s[i].x, c[i].y + s[i].y);
    double wr = -(xv.x * a.x - xv.y * a.y), wi = -(xv.x * a.y + xv.y * a.x);
    gr += 2.0 * wr;
    gi -= 2.0 * wi;
  }
  g[j] = make_double2(gr, gi);
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

extern "C" int fmbh_zdots(const fmb_c64 *d_V, int nv, const fmb_c64 *d_w,
                          int nrhs, int N, fmb_c64 *d_h, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  double2 *part;
  int rc = fmbh_malloc((void **)&part, sizeof(double2) * CHUNKS * nv * nrhs);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zdots_part, dim3(CHUNKS, nv * nrhs), dim3(BS), 0, st,
                     (const double2 *)d_V, nrhs, N, (const double2 *)d_w, part);
  FMB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_chunks, dim3(nblk(nv * nrhs)), dim3(256), 0, st,
                     part, nv * nrhs, (double2 *)d_h);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(part);
  return FMB_OK;
}

extern "C" int fmbh_znorms(const fmb_c64 *d_w, int nrhs, int N, double *d_nrm,
                           void *stream) {
  hipStream_t st = (hipStream_t)stream;
  double *part;
  int rc = fmbh_malloc((void **)&part, sizeof(double) * CHUNKS * nrhs);
  if (rc) return rc;
  hipLaunchKernelGGL(k_znorm_part, dim3(CHUNKS, nrhs), dim3(BS), 0, st,
                     (const double2 *)d_w, N, part);
  FMB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_norm_final, dim3(nblk(nrhs)), dim3(256), 0, st, part,
                     nrhs, d_nrm);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(part);
  return FMB_OK;
}

extern "C" int fmbh_zgemv_sub(const fmb_c64 *d_V, int nv, const fmb_c64 *d_h,
                              fmb_c64 *d_w, int nrhs, int N, void *stream) {
  hipLaunchKernelGGL(k_gemv, dim3(nblk((size_t)N * nrhs)), dim3(256), 0,
                     (hipStream_t)stream, (const double2 *)d_V, nv,
                     (const double2 *)d_h, (double2 *)d_w, nrhs, N, -1.0);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_zgemv_add(const fmb_c64 *d_V, int nv, const fmb_c64 *d_y,
                              fmb_c64 *d_x, int nrhs, int N, void *stream) {
  hipLaunchKernelGGL(k_gemv, dim3(nblk((size_t)N * nrhs)), dim3(256), 0,
                     (hipStream_t)stream, (const double2 *)d_V, nv,
                     (const double2 *)d_y, (double2 *)d_x, nrhs, N, 1.0);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_zaxpby(const fmb_c64 *d_a, const fmb_c64 *d_x,
                           const fmb_c64 *d_b, fmb_c64 *d_y, int nrhs, int N,
                           void *stream) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk((size_t)N * nrhs)), dim3(256), 0,
                     (hipStream_t)stream, (const double2 *)d_a,
                     (const double2 *)d_x, (const double2 *)d_b, (double2 *)d_y,
                     nrhs, N);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_rhs_plane(const fmb_plan *pl, const double *h_dirs,
                              int nrhs, fmb_c64 *d_b, void *stream) {
  const fmb_tree *t = pl->tree;
  double *d_dirs;
  int rc = fmbh_malloc((void **)&d_dirs, sizeof(double) * 3 * nrhs);
  if (rc) return rc;
  if ((rc = fmbh_h2d(d_dirs, h_dirs, sizeof(double) * 3 * nrhs, stream))) return rc;
  size_t n = (size_t)t->nelem * nrhs;
  hipLaunchKernelGGL(k_rhs_plane, dim3(nblk(n)), dim3(256), 0,
                     (hipStream_t)stream, t->d_geo, t->d_iperm, d_dirs, nrhs,
                     t->nelem, pl->k, pl->alpha_re, pl->alpha_im,
                     (double2 *)d_b);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  fmbh_free(d_dirs);
  return FMB_OK;
}

extern "C" int fmbh_sweep_guess(const fmb_plan *pl, double dk, const fmb_c64 *d_x,
                                const double *h_dirs, int nrhs, fmb_c64 *d_x0,
                                void *stream) {
  const fmb_tree *t = pl->tree;
  double *d_dirs = NULL;
  int rc;
  if (h_dirs) {
    if ((rc = fmbh_malloc((void **)&d_dirs, sizeof(double) * 3 * nrhs))) return rc;
    if ((rc = fmbh_h2d(d_dirs, h_dirs, sizeof(double) * 3 * nrhs, stream))) return rc;
  }
  size_t n = (size_t)t->nelem * nrhs;
  hipLaunchKernelGGL(k_sweep_guess, dim3(nblk(n)), dim3(256), 0,
                     (hipStream_t)stream, t->d_geo, t->d_iperm, d_dirs, nrhs,
                     t->nelem, dk, (const double2 *)d_x, (double2 *)d_x0);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return fmbh_free(d_dirs);
}

extern "C" int fmbh_adjoint_combine(const fmb_c64 *x, const fmb_c64 *c, const fmb_c64 *s,
                                    int nrhs, int N, fmb_c64 *g, void *stream) {
  hipLaunchKernelGGL(k_adj_combine, dim3(nblk((size_t)N)), dim3(256), 0,
                     (hipStream_t)stream, (const double2 *)x, (const double2 *)c,
                     (const double2 *)s, nrhs, N, (double2 *)g);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}
