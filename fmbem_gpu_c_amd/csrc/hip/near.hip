// near.hip -- near-field blocks of the Burton-Miller operator and the fused
// L2P + near-field product.
//
// For a leaf box b the block holds rows = elements of b, columns = elements
// of the boxes in b's 27-box neighbourhood (the reference's I2 set,
// octree.c:497-505), row-major, in Morton element order.  Entries
//   A_ij = -K_ij - alpha T_ij + sigma_j (S_ij + alpha K'_ij)        (i != j)
//   A_ii = 1/2 + alpha sigma_i/2 - alpha T_ii + sigma_i S_ii
// with the kernels of integral.cu:72-78, 391-415, 602-622 integrated by the
// far rule (integral.cu:80-139 style panel rules), a refined rule when
// |x_i - c_j| < ratio * circumradius_j, and the centroid self terms as
// contour integrals over the three edges (integral.cu:202-249, 666-751):
//   S_ii = (1/4pi) oint (e^{ik rho}-1)/(ik) dtheta
//   T_ii = -(1/4pi) [ oint (e^{ik rho}-1)/rho dtheta + oint dtheta/rho - 2 pi i k ]
// (K_ii = K'_ii = 0 on a flat panel).  Evaluated in FP64, stored in the
// compute precision.  With dsig set the blocks hold dA_ij/dsigma_j instead:
// S_ij + alpha K'_ij (i != j) and alpha/2 + S_ii.
#include <hipcub/hipcub.hpp>

#include "fmb_device.h"

namespace {

__global__ void k_near_sizes(const int *bstart, const int *nstart,
                             const int *nbox, int nb, int *ncol,
                             unsigned long long *bsz) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int c = 0;
  for (int q = nstart[b]; q < nstart[b + 1]; q++) {
    int o = nbox[q];
    c += bstart[o + 1] - bstart[o];
  }
  ncol[b] = c;
  bsz[b] = (unsigned long long)(bstart[b + 1] - bstart[b]) * c;
}

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_near_assemble(const int *__restrict__ bstart, const int *__restrict__ nstart,
                    const int *__restrict__ nbox, const int *__restrict__ ncols,
                    const unsigned long long *__restrict__ boff,
                    const double *__restrict__ tri, const double *__restrict__ geo,
                    double k, double alre, double alim,
                    const double2 *__restrict__ robin,
                    const double *__restrict__ quad, int nq,
                    const double *__restrict__ quadn, int nqn, double ratio,
                    const double *__restrict__ edge, int nedge, T *__restrict__ A,
                    int dsig) {
  __shared__ int cpre[28], cbox[27];
  __shared__ int ncb;
  int b = blockIdx.x;
  if (threadIdx.x == 0) {
    int acc = 0, q0 = nstart[b], q1 = nstart[b + 1];
    ncb = q1 - q0;
    for (int q = q0; q < q1; q++) {
      cpre[q - q0] = acc;
      cbox[q - q0] = nbox[q];
      int o = nbox[q];
      acc += bstart[o + 1] - bstart[o];
    }
    cpre[q1 - q0] = acc;
  }
  __syncthreads();
  int r0 = bstart[b], nr = bstart[b + 1] - r0, nc = ncols[b];
  double2 alpha = make_double2(alre, alim);
  unsigned long long base = boff[b];
  for (int e = threadIdx.x; e < nr * nc; e += blockDim.x) {
    int r = e / nc, cc = e % nc;
    int q = 0;
    while (q + 1 < ncb && cpre[q + 1] <= cc) q++;
    int i = r0 + r, j = bstart[cbox[q]] + (cc - cpre[q]);
    const double *gi = geo + 8 * (size_t)i, *gj = geo + 8 * (size_t)j;
    double2 sj = robin ? robin[j] : make_double2(0, 0);
    double2 v;
    if (i == j) {
      double2 S, Tt;
      self_terms(k, tri + 9 * (size_t)j, gj, edge, nedge, S, Tt);
      double2 aT = cm(alpha, Tt), sS = cm(sj, S), as = cm(alpha, sj);
      if (dsig)
        v = make_double2(0.5 * alpha.x + S.x, 0.5 * alpha.y + S.y);
      else
        v = make_double2(0.5 + 0.5 * as.x - aT.x + sS.x, 0.5 * as.y - aT.y + sS.y);
    } else {
      double dx = gi[0] - gj[0], dy = gi[1] - gj[1], dz = gi[2] - gj[2];
      bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
      double2 o[4];
      panel_sum(k, gi, gi + 3, tri + 9 * (size_t)j, gj + 3, gj[6],
                close ? quadn : quad, close ? nqn : nq, o);
      double2 aT = cm(alpha, o[3]), aKp = cm(alpha, o[2]);
      double2 t = make_double2(o[0].x + aKp.x, o[0].y + aKp.y);
      double2 st = cm(sj, t);
      if (dsig)
        v = t;
      else
        v = make_double2(-o[1].x - aT.x + st.x, -o[1].y - aT.y + st.y);
    }
    A[2 * (base + e)] = (T)v.x;
    A[2 * (base + e) + 1] = (T)v.y;
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// L2P + near: one block per leaf box, one wavefront per (row, rhs) job
template <typename T>
__global__ void __launch_bounds__(256)
    k_l2p_near(const int *__restrict__ bstart, const int *__restrict__ nstart,
               const int *__restrict__ nbox, const int *__restrict__ ncols,
               const unsigned long long *__restrict__ boff,
               const cplx<T> *__restrict__ A, const cplx<T> *__restrict__ E,
               const cplx<T> *__restrict__ Lc, const cplx<T> *__restrict__ xs,
               cplx<T> *__restrict__ ys, int N, int ps, int nrhs, int parts) {
  extern __shared__ int colj[];
  int b = blockIdx.x;
  int r0 = bstart[b], nr = bstart[b + 1] - r0, nc = ncols[b];
  if (threadIdx.x == 0) {
    int c = 0;
    for (int q = nstart[b]; q < nstart[b + 1]; q++) {
      int o = nbox[q];
      for (int e = bstart[o]; e < bstart[o + 1]; e++) colj[c++] = e;
    }
  }
  __syncthreads();
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const cplx<T> *Ab = A + boff[b];
  for (int job = wave; job < nr * nrhs; job += (blockDim.x >> 6)) {
    int r = job % nr, rh = job / nr;
    int i = r0 + r;
    T sr = 0, si = 0;
    const cplx<T> *arow = Ab + (size_t)r * nc;
    const cplx<T> *x = xs + (size_t)rh * N;
    for (int c = (parts & 1) ? lane : nc; c < nc; c += 64) {
      cplx<T> a = arow[c], v = x[colj[c]];
      sr += a.re * v.re - a.im * v.im;
      si += a.re * v.im + a.im * v.re;
    }
    const cplx<T> *e = E + (size_t)i * ps;
    const cplx<T> *l = Lc + ((size_t)b * nrhs + rh) * ps;
    for (int f = (parts & 2) ? lane : ps; f < ps; f += 64) {
      cplx<T> a = e[f], v = l[f];
      sr += a.re * v.re - a.im * v.im;
      si += a.re * v.im + a.im * v.re;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) ys[(size_t)rh * N + i] = cmk<T>(sr, si);
  }
}

// Transposed near field + transposed P2M: one block per leaf box c, one
// wavefront per (element j of c, right-hand side) job.  Column j of A lies in
// the blocks of the boxes b of c's neighbourhood (neighbourhoods are
// symmetric), at the columns where b's block lists box c.  Far part:
// sum_f B[j][f] (-1)^n M'[c][r][f], the transposed pass's coefficients being
// kept multiplied by (-1)^n.
template <typename T>
__global__ void __launch_bounds__(256)
    k_near_t(const int *__restrict__ bstart, const int *__restrict__ nstart,
             const int *__restrict__ nbox, const int *__restrict__ ncols,
             const unsigned long long *__restrict__ boff,
             const cplx<T> *__restrict__ A, const cplx<T> *__restrict__ B,
             const cplx<T> *__restrict__ Mc, const cplx<T> *__restrict__ xs,
             cplx<T> *__restrict__ ys, int N, int ps, int nrhs, int parts) {
  __shared__ int nb_first[27], nb_row0[28], nb_col[27], nb_nc[27];
  __shared__ unsigned long long nb_off[27];
  __shared__ int nnb;
  const int c = blockIdx.x;
  const int e0 = bstart[c], nj = bstart[c + 1] - e0;
  if (threadIdx.x == 0) {
    int q0 = nstart[c], q1 = nstart[c + 1], acc = 0, k = 0;
    for (int q = q0; q < q1 && k < 27; q++) {
      int b = nbox[q], co = 0, found = 0;
      for (int qq = nstart[b]; qq < nstart[b + 1]; qq++) {
        int o = nbox[qq];
        if (o == c) {
          found = 1;
          break;
        }
        co += bstart[o + 1] - bstart[o];
      }
      if (!found) continue; /* not reached: neighbourhoods are symmetric */
      nb_first[k] = bstart[b];
      nb_row0[k] = acc;
      nb_col[k] = co;
      nb_nc[k] = ncols[b];
      nb_off[k] = boff[b];
      acc += bstart[b + 1] - bstart[b];
      k++;
    }
    nb_row0[k] = acc;
    nnb = k;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int R = nb_row0[nnb];
  for (int job = wave; job < nj * nrhs; job += (blockDim.x >> 6)) {
    const int jj = job % nj, rh = job / nj, j = e0 + jj;
    T sr = 0, si = 0;
    const cplx<T> *x = xs + (size_t)rh * N;
    for (int u = (parts & 1) ? lane : R; u < R; u += 64) {
      int q = 0;
      while (q + 1 < nnb && nb_row0[q + 1] <= u) q++;
      const int r = u - nb_row0[q];
      cplx<T> a = A[nb_off[q] + (size_t)r * nb_nc[q] + nb_col[q] + jj];
      cplx<T> v = x[nb_first[q] + r];
      sr += a.re * v.re - a.im * v.im;
      si += a.re * v.im + a.im * v.re;
    }
    const cplx<T> *e = B + (size_t)j * ps;
    const cplx<T> *l = Mc + ((size_t)c * nrhs + rh) * ps;
    for (int f = (parts & 2) ? lane : ps; f < ps; f += 64) {
      cplx<T> a = e[f], v = l[f];
      T g = (isqrt_fix(f) & 1) ? T(-1) : T(1);
      sr += g * (a.re * v.re - a.im * v.im);
      si += g * (a.re * v.im + a.im * v.re);
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) ys[(size_t)rh * N + j] = cmk<T>(sr, si);
  }
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

extern "C" int fmbh_near_layout(fmb_plan *pl, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  int L = t->lmax, nb = t->nbox[L];
  unsigned long long *bsz;
  int rc = fmbh_malloc((void **)&pl->d_near_ncol, sizeof(int) * nb);
  if (!rc) rc = fmbh_malloc((void **)&bsz, sizeof(unsigned long long) * (nb + 1));
  if (!rc)
    rc = fmbh_malloc((void **)&pl->d_near_off,
                     sizeof(unsigned long long) * (nb + 1));
  if (rc) return rc;
  FMB_HIP_CHECK(hipMemsetAsync(bsz, 0, sizeof(unsigned long long) * (nb + 1), st));
  hipLaunchKernelGGL(k_near_sizes, dim3(nblk(nb)), dim3(256), 0, st,
                     t->d_box_start, t->d_near_start, t->d_near_box, nb,
                     pl->d_near_ncol, bsz);
  FMB_LAUNCH_CHECK();
  size_t tb = 0;
  unsigned long long *off = (unsigned long long *)pl->d_near_off;
  FMB_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, bsz, off, nb + 1, st));
  void *tmp;
  if ((rc = fmbh_malloc(&tmp, tb))) return rc;
  FMB_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, bsz, off, nb + 1, st));
  unsigned long long tot;
  if ((rc = fmbh_d2h(&tot, off + nb, sizeof(tot), stream))) return rc;
  pl->near_total = (size_t)tot;
  int *h = (int *)malloc(sizeof(int) * nb);
  rc = fmbh_d2h(h, pl->d_near_ncol, sizeof(int) * nb, stream);
  int mx = 0;
  for (int b = 0; b < nb; b++)
    if (h[b] > mx) mx = h[b];
  free(h);
  pl->near_max_ncol = mx;
  fmbh_free(tmp);
  fmbh_free(bsz);
  if (mx > 16384)
    return fmb_set_error(FMB_ERR_RANGE,
                         "near-field neighbourhood of %d elements exceeds 16384; "
                         "use a deeper tree",
                         mx);
  return rc;
}

static int launch_near(fmb_plan *pl, int op, const fmb_c64 *robin, void *out,
                       int dsig, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  int L = t->lmax, nb = t->nbox[L];
  fmb_opset *o = &pl->ops[op];
  const unsigned long long *off = (const unsigned long long *)pl->d_near_off;
  if (o->prec == FMB_FP64)
    hipLaunchKernelGGL(k_near_assemble<double>, dim3(nb), dim3(256), 0, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, t->d_tri, t->d_geo, pl->k,
                       pl->alpha_re, pl->alpha_im, (const double2 *)robin,
                       pl->d_quad, pl->nq, pl->d_quadn, pl->nqn,
                       pl->prm.near_ratio, pl->d_edge, pl->prm.nedge,
                       (double *)out, dsig);
  else
    hipLaunchKernelGGL(k_near_assemble<float>, dim3(nb), dim3(256), 0, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, t->d_tri, t->d_geo, pl->k,
                       pl->alpha_re, pl->alpha_im, (const double2 *)robin,
                       pl->d_quad, pl->nq, pl->d_quadn, pl->nqn,
                       pl->prm.near_ratio, pl->d_edge, pl->prm.nedge,
                       (float *)out, dsig);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  return FMB_OK;
}

extern "C" int fmbh_near_assemble(fmb_plan *pl, int op, const fmb_c64 *robin,
                                  void *stream) {
  return launch_near(pl, op, robin, pl->ops[op].d_near, 0, stream);
}

extern "C" int fmbh_near_dsigma(fmb_plan *pl, int op, void *out, void *stream) {
  return launch_near(pl, op, NULL, out, 1, stream);
}

extern "C" int fmbh_l2p_near(const fmb_plan *pl, int op, int nrhs, int parts,
                             void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  int L = t->lmax, nb = t->nbox[L], pL = pl->p[L];
  const fmb_opset *o = &pl->ops[op];
  size_t lds = sizeof(int) * (size_t)(pl->near_max_ncol > 0 ? pl->near_max_ncol : 1);
  const unsigned long long *off = (const unsigned long long *)pl->d_near_off;
  if (o->prec == FMB_FP64)
    hipLaunchKernelGGL(k_l2p_near<double>, dim3(nb), dim3(256), lds, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, (const cplx<double> *)o->d_near,
                       (const cplx<double> *)o->d_L2P,
                       (const cplx<double> *)o->d_L[L],
                       (const cplx<double> *)o->d_xs, (cplx<double> *)o->d_ys,
                       t->nelem, pL * pL, nrhs, parts);
  else
    hipLaunchKernelGGL(k_l2p_near<float>, dim3(nb), dim3(256), lds, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, (const cplx<float> *)o->d_near,
                       (const cplx<float> *)o->d_L2P,
                       (const cplx<float> *)o->d_L[L],
                       (const cplx<float> *)o->d_xs, (cplx<float> *)o->d_ys,
                       t->nelem, pL * pL, nrhs, parts);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_near_t(const fmb_plan *pl, int op, const void *near,
                           const void *basis, int nrhs, int parts, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  int L = t->lmax, nb = t->nbox[L], pL = pl->p[L];
  const fmb_opset *o = &pl->ops[op];
  const void *A = near ? near : o->d_near;
  const void *B = basis ? basis : o->d_P2M;
  const unsigned long long *off = (const unsigned long long *)pl->d_near_off;
  if (o->prec == FMB_FP64)
    hipLaunchKernelGGL(k_near_t<double>, dim3(nb), dim3(256), 0, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, (const cplx<double> *)A,
                       (const cplx<double> *)B, (const cplx<double> *)o->d_M[L],
                       (const cplx<double> *)o->d_xs, (cplx<double> *)o->d_ys,
                       t->nelem, pL * pL, nrhs, parts);
  else
    hipLaunchKernelGGL(k_near_t<float>, dim3(nb), dim3(256), 0, st,
                       t->d_box_start, t->d_near_start, t->d_near_box,
                       pl->d_near_ncol, off, (const cplx<float> *)A,
                       (const cplx<float> *)B, (const cplx<float> *)o->d_M[L],
                       (const cplx<float> *)o->d_xs, (cplx<float> *)o->d_ys,
                       t->nelem, pL * pL, nrhs, parts);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}
