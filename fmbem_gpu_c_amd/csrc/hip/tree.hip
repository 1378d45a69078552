// tree.hip -- device Morton octree: centroids, bounding cube, level-10 keys,
// radix sort, leaf depth, level sets, parent/child ranges, M2L pair lists
// grouped by offset, leaf neighbourhoods.
//
// Restates octree.c on the device:
//   triCentroid_d            structs.cu:126-131
//   pnt2boxnum / Morton      octree.c:62-183 (x most significant per digit)
//   deterLmax                octree.c:635-679 (windows of s+1 sorted keys)
//   FMMLevelSet              octree.c:890-906 (parents up to lmin = 2)
//   neighbors / I4           octree.c:237-391, 517-543
// with sorted arrays and binary search in place of the O(n^2) set algebra.
#include <hipcub/hipcub.hpp>

#include "fmb_device.h"

namespace {

__device__ __forceinline__ int bsearch_key(const int *keys, int n, int key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = keys[mid];
    if (v == key) return mid;
    if (v < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

__global__ void k_centroid(const double *V, const int *F, int n, double *ctr,
                           double *part) {
  __shared__ double red[6][256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    const double *a = V + 3 * F[3 * i], *b = V + 3 * F[3 * i + 1],
                 *c = V + 3 * F[3 * i + 2];
    for (int q = 0; q < 3; q++) {
      double v = a[q] + 2.0 / 3.0 * (0.5 * (b[q] + c[q]) - a[q]);
      ctr[3 * i + q] = v;
      lo[q] = hi[q] = v;
    }
  }
  for (int q = 0; q < 3; q++) {
    red[q][threadIdx.x] = lo[q];
    red[3 + q][threadIdx.x] = hi[q];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 3; q++) {
        red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        red[3 + q][threadIdx.x] =
            fmax(red[3 + q][threadIdx.x], red[3 + q][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_morton10(const double *ctr, int n, double p0, double p1,
                           double p2, double D, int *key, int *idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int l = 10, N = 1 << l;
  double pm[3] = {p0, p1, p2};
  int t[3];
  for (int q = 0; q < 3; q++) {
    double s = (ctr[3 * i + q] - pm[q]) / D;
    int v = (int)floor(s * (double)N);
    t[q] = v < 0 ? 0 : (v > N - 1 ? N - 1 : v);
  }
  key[i] = interleave3(t[0], t[1], t[2], l);
  idx[i] = i;
}

__global__ void k_deter_lmax(const int *key, int n, int s, int *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + s >= n) return;
  int a = key[i], b = key[i + s];
  if (a == b) {
    atomicMax(out + 1, 1);
    return;
  }
  int h = 31 - __clz(a ^ b);
  atomicMax(out, 10 - h / 3);
}

__global__ void k_flag_change(const int *key, int n, int shift, int *flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0) ? 0 : ((key[i] >> shift) != (key[i - 1] >> shift));
}

// scan[i] = box index of element i; write box keys and starts
__global__ void k_emit_boxes(const int *key, int n, int shift, const int *scan,
                             int *bkey, int *bstart, int *ebox) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int b = scan[i];
  ebox[i] = b;
  if (i == 0 || (key[i] >> shift) != (key[i - 1] >> shift)) {
    bkey[b] = key[i] >> shift;
    bstart[b] = i;
  }
  if (i == n - 1) bstart[b + 1] = n;
}

__global__ void k_count_pairs(const int *keys, int nb, int l, int *cnt) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int tx, ty, tz;
  deinterleave3(keys[b], l, tx, ty, tz);
  int N = 1 << l, c = 0;
  int t[3] = {tx, ty, tz};
  int lo[3], hi[3];
  for (int q = 0; q < 3; q++) {
    // offsets o = target - source with source parent adjacent to target parent
    lo[q] = (t[q] & 1) ? -2 : -3;
    hi[q] = (t[q] & 1) ? 3 : 2;
  }
  for (int dx = lo[0]; dx <= hi[0]; dx++)
    for (int dy = lo[1]; dy <= hi[1]; dy++)
      for (int dz = lo[2]; dz <= hi[2]; dz++) {
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        int sx = tx - dx, sy = ty - dy, sz = tz - dz;
        if (sx < 0 || sy < 0 || sz < 0 || sx >= N || sy >= N || sz >= N)
          continue;
        if (bsearch_key(keys, nb, interleave3(sx, sy, sz, l)) >= 0) c++;
      }
  cnt[b] = c;
}

__global__ void k_fill_pairs(const int *keys, int nb, int l, const int *start,
                             int *tgt, int *src, int *off, int *hist) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int tx, ty, tz;
  deinterleave3(keys[b], l, tx, ty, tz);
  int N = 1 << l, c = start[b];
  int t[3] = {tx, ty, tz};
  int lo[3], hi[3];
  for (int q = 0; q < 3; q++) {
    lo[q] = (t[q] & 1) ? -2 : -3;
    hi[q] = (t[q] & 1) ? 3 : 2;
  }
  for (int dx = lo[0]; dx <= hi[0]; dx++)
    for (int dy = lo[1]; dy <= hi[1]; dy++)
      for (int dz = lo[2]; dz <= hi[2]; dz++) {
        if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
        int sx = tx - dx, sy = ty - dy, sz = tz - dz;
        if (sx < 0 || sy < 0 || sz < 0 || sx >= N || sy >= N || sz >= N)
          continue;
        int s = bsearch_key(keys, nb, interleave3(sx, sy, sz, l));
        if (s < 0) continue;
        int o = (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
        tgt[c] = b;
        src[c] = s;
        off[c] = o;
        atomicAdd(hist + o, 1);
        c++;
      }
}

__global__ void k_permute2(const int *perm, int n, const int *a, const int *b,
                           int *a2, int *b2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int j = perm[i];
  a2[i] = a[j];
  b2[i] = b[j];
}

__global__ void k_iota(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void k_near_count(const int *keys, int nb, int l, int *cnt) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int tx, ty, tz;
  deinterleave3(keys[b], l, tx, ty, tz);
  int N = 1 << l, c = 0;
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        int x = tx + dx, y = ty + dy, z = tz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) continue;
        if (bsearch_key(keys, nb, interleave3(x, y, z, l)) >= 0) c++;
      }
  cnt[b] = c;
}

__global__ void k_near_fill(const int *keys, int nb, int l, const int *start,
                            int *lst) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int tx, ty, tz;
  deinterleave3(keys[b], l, tx, ty, tz);
  int N = 1 << l, c = start[b];
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        int x = tx + dx, y = ty + dy, z = tz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) continue;
        int s = bsearch_key(keys, nb, interleave3(x, y, z, l));
        if (s >= 0) lst[c++] = s;
      }
}

__global__ void k_sorted_geo(const double *V, const int *F, const int *perm,
                             int n, double *tri, double *geo, int *iperm, int *fv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int e = perm[i];
  iperm[e] = i;
  double P[3][3];
  for (int v = 0; v < 3; v++) fv[3 * i + v] = F[3 * e + v];
  for (int v = 0; v < 3; v++)
    for (int q = 0; q < 3; q++) {
      P[v][q] = V[3 * F[3 * e + v] + q];
      tri[9 * i + 3 * v + q] = P[v][q];
    }
  double c[3], a[3], b[3];
  for (int q = 0; q < 3; q++) {
    c[q] = P[0][q] + 2.0 / 3.0 * (0.5 * (P[1][q] + P[2][q]) - P[0][q]);
    a[q] = P[0][q] - P[2][q];
    b[q] = P[1][q] - P[2][q];
  }
  double nr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                  a[0] * b[1] - a[1] * b[0]};
  double J = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
  double rad = 0.0;
  for (int v = 0; v < 3; v++) {
    double d2 = 0.0;
    for (int q = 0; q < 3; q++) d2 += (P[v][q] - c[q]) * (P[v][q] - c[q]);
    rad = fmax(rad, d2);
  }
  double *g = geo + 8 * i;
  for (int q = 0; q < 3; q++) {
    g[q] = c[q];
    g[3 + q] = nr[q] / J;
  }
  g[6] = 0.5 * J;
  g[7] = sqrt(rad);
}

inline int nblk(long n, int b = 256) { return (int)((n + b - 1) / b); }

int scan_inclusive(const int *in, int *out, int n, hipStream_t st) {
  size_t tb = 0;
  FMB_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, n, st));
  void *tmp;
  int rc = fmbh_malloc(&tmp, tb);
  if (rc) return rc;
  FMB_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out, n, st));
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(tmp);
  return FMB_OK;
}

int scan_exclusive(const int *in, int *out, int n, hipStream_t st) {
  size_t tb = 0;
  FMB_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, st));
  void *tmp;
  int rc = fmbh_malloc(&tmp, tb);
  if (rc) return rc;
  FMB_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, st));
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(tmp);
  return FMB_OK;
}

int sort_pairs(const int *kin, int *kout, const int *vin, int *vout, int n,
               int bits, hipStream_t st) {
  size_t tb = 0;
  FMB_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin,
                                                   vout, n, 0, bits, st));
  void *tmp;
  int rc = fmbh_malloc(&tmp, tb);
  if (rc) return rc;
  FMB_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin,
                                                   vout, n, 0, bits, st));
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(tmp);
  return FMB_OK;
}

}  // namespace

extern "C" int fmbh_centroid_bbox(const double *d_V, const int *d_F, int n,
                                  double *d_ctr, double *lo, double *hi,
                                  void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int nb = nblk(n);
  double *d_part;
  int rc = fmbh_malloc((void **)&d_part, sizeof(double) * 6 * nb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_centroid, dim3(nb), dim3(256), 0, st, d_V, d_F, n,
                     d_ctr, d_part);
  FMB_LAUNCH_CHECK();
  double *h = (double *)malloc(sizeof(double) * 6 * nb);
  rc = fmbh_d2h(h, d_part, sizeof(double) * 6 * nb, stream);
  if (!rc) {
    for (int q = 0; q < 3; q++) {
      lo[q] = INFINITY;
      hi[q] = -INFINITY;
    }
    for (int b = 0; b < nb; b++)
      for (int q = 0; q < 3; q++) {
        lo[q] = fmin(lo[q], h[6 * b + q]);
        hi[q] = fmax(hi[q], h[6 * b + 3 + q]);
      }
  }
  free(h);
  fmbh_free(d_part);
  return rc;
}

extern "C" int fmbh_morton_sort(const double *d_ctr, int n, const double *pmin,
                                double D, int *d_key10, int *d_perm,
                                void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int *k0, *i0;
  int rc = fmbh_malloc((void **)&k0, sizeof(int) * n);
  if (rc) return rc;
  rc = fmbh_malloc((void **)&i0, sizeof(int) * n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_morton10, dim3(nblk(n)), dim3(256), 0, st, d_ctr, n,
                     pmin[0], pmin[1], pmin[2], D, k0, i0);
  FMB_LAUNCH_CHECK();
  rc = sort_pairs(k0, d_key10, i0, d_perm, n, 30, st);
  fmbh_free(k0);
  fmbh_free(i0);
  return rc;
}

extern "C" int fmbh_deter_lmax(const int *d_key10, int n, int s, int *lmax,
                               void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int *d_out;
  int rc = fmbh_malloc((void **)&d_out, 2 * sizeof(int));
  if (rc) return rc;
  FMB_HIP_CHECK(hipMemsetAsync(d_out, 0, 2 * sizeof(int), st));
  if (n > s) {
    hipLaunchKernelGGL(k_deter_lmax, dim3(nblk(n)), dim3(256), 0, st, d_key10,
                       n, s, d_out);
    FMB_LAUNCH_CHECK();
  }
  int h[2];
  rc = fmbh_d2h(h, d_out, sizeof(h), stream);
  fmbh_free(d_out);
  if (rc) return rc;
  if (h[1])
    return fmb_set_error(FMB_ERR_DEPTH,
                         "%d+1 centroids share one level-10 box: leaf depth "
                         "exceeds the level-10 key resolution",
                         s);
  *lmax = h[0];
  return FMB_OK;
}

// leaf level: box keys, element ranges, element -> box
extern "C" int fmbh_leaf_level(fmb_tree *t, const int *d_key10, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int n = t->nelem, L = t->lmax, shift = 3 * (10 - L);
  int *flag, *scan;
  int rc = fmbh_malloc((void **)&flag, sizeof(int) * n);
  if (!rc) rc = fmbh_malloc((void **)&scan, sizeof(int) * n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_flag_change, dim3(nblk(n)), dim3(256), 0, st, d_key10,
                     n, shift, flag);
  FMB_LAUNCH_CHECK();
  if ((rc = scan_inclusive(flag, scan, n, st))) return rc;
  int last;
  if ((rc = fmbh_d2h(&last, scan + n - 1, sizeof(int), stream))) return rc;
  int nb = last + 1;
  t->nbox[L] = nb;
  if ((rc = fmbh_malloc((void **)&t->d_keys[L], sizeof(int) * nb))) return rc;
  if ((rc = fmbh_malloc((void **)&t->d_box_start, sizeof(int) * (nb + 1))))
    return rc;
  if ((rc = fmbh_malloc((void **)&t->d_elem_box, sizeof(int) * n))) return rc;
  hipLaunchKernelGGL(k_emit_boxes, dim3(nblk(n)), dim3(256), 0, st, d_key10, n,
                     shift, scan, t->d_keys[L], t->d_box_start, t->d_elem_box);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(flag);
  fmbh_free(scan);
  return FMB_OK;
}

// level l from level l+1: keys, parent index of each child, child ranges
extern "C" int fmbh_parent_level(fmb_tree *t, int l, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int nc = t->nbox[l + 1];
  int *flag, *scan;
  int rc = fmbh_malloc((void **)&flag, sizeof(int) * nc);
  if (!rc) rc = fmbh_malloc((void **)&scan, sizeof(int) * nc);
  if (rc) return rc;
  hipLaunchKernelGGL(k_flag_change, dim3(nblk(nc)), dim3(256), 0, st,
                     t->d_keys[l + 1], nc, 3, flag);
  FMB_LAUNCH_CHECK();
  if ((rc = scan_inclusive(flag, scan, nc, st))) return rc;
  int last;
  if ((rc = fmbh_d2h(&last, scan + nc - 1, sizeof(int), stream))) return rc;
  int np = last + 1;
  t->nbox[l] = np;
  if ((rc = fmbh_malloc((void **)&t->d_keys[l], sizeof(int) * np))) return rc;
  if ((rc = fmbh_malloc((void **)&t->d_cstart[l], sizeof(int) * (np + 1))))
    return rc;
  if ((rc = fmbh_malloc((void **)&t->d_parent[l + 1], sizeof(int) * nc)))
    return rc;
  hipLaunchKernelGGL(k_emit_boxes, dim3(nblk(nc)), dim3(256), 0, st,
                     t->d_keys[l + 1], nc, 3, scan, t->d_keys[l],
                     t->d_cstart[l], t->d_parent[l + 1]);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(flag);
  fmbh_free(scan);
  return FMB_OK;
}

extern "C" int fmbh_m2l_pairs(fmb_tree *t, int l, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int nb = t->nbox[l];
  int *cnt, *start, *hist;
  int rc = fmbh_malloc((void **)&cnt, sizeof(int) * (nb + 1));
  if (!rc) rc = fmbh_malloc((void **)&start, sizeof(int) * (nb + 1));
  if (!rc) rc = fmbh_malloc((void **)&hist, sizeof(int) * FMB_NOFF);
  if (rc) return rc;
  FMB_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (nb + 1), st));
  FMB_HIP_CHECK(hipMemsetAsync(hist, 0, sizeof(int) * FMB_NOFF, st));
  hipLaunchKernelGGL(k_count_pairs, dim3(nblk(nb)), dim3(256), 0, st,
                     t->d_keys[l], nb, l, cnt);
  FMB_LAUNCH_CHECK();
  if ((rc = scan_exclusive(cnt, start, nb + 1, st))) return rc;
  int np;
  if ((rc = fmbh_d2h(&np, start + nb, sizeof(int), stream))) return rc;
  t->npairs[l] = np;
  int *tg, *sr, *of, *of2, *idx, *idx2;
  size_t b = sizeof(int) * (np > 0 ? np : 1);
  rc = fmbh_malloc((void **)&tg, b);
  if (!rc) rc = fmbh_malloc((void **)&sr, b);
  if (!rc) rc = fmbh_malloc((void **)&of, b);
  if (!rc) rc = fmbh_malloc((void **)&of2, b);
  if (!rc) rc = fmbh_malloc((void **)&idx, b);
  if (!rc) rc = fmbh_malloc((void **)&idx2, b);
  if (!rc) rc = fmbh_malloc((void **)&t->d_pair_tgt[l], b);
  if (!rc) rc = fmbh_malloc((void **)&t->d_pair_src[l], b);
  if (rc) return rc;
  t->h_off_start[l] = (int *)calloc(FMB_NOFF + 1, sizeof(int));
  if (np > 0) {
    hipLaunchKernelGGL(k_fill_pairs, dim3(nblk(nb)), dim3(256), 0, st,
                       t->d_keys[l], nb, l, start, tg, sr, of, hist);
    FMB_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_iota, dim3(nblk(np)), dim3(256), 0, st, idx, np);
    FMB_LAUNCH_CHECK();
    // stable radix sort by offset code keeps target order inside a bucket
    if ((rc = sort_pairs(of, of2, idx, idx2, np, 9, st))) return rc;
    hipLaunchKernelGGL(k_permute2, dim3(nblk(np)), dim3(256), 0, st, idx2, np,
                       tg, sr, t->d_pair_tgt[l], t->d_pair_src[l]);
    FMB_LAUNCH_CHECK();
    int h[FMB_NOFF];
    if ((rc = fmbh_d2h(h, hist, sizeof(h), stream))) return rc;
    int acc = 0;
    for (int o = 0; o < FMB_NOFF; o++) {
      t->h_off_start[l][o] = acc;
      acc += h[o];
    }
    t->h_off_start[l][FMB_NOFF] = acc;
    if (acc != np)
      return fmb_set_error(FMB_ERR_HIP, "pair histogram %d != %d", acc, np);
  }
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(cnt);
  fmbh_free(start);
  fmbh_free(hist);
  fmbh_free(tg);
  fmbh_free(sr);
  fmbh_free(of);
  fmbh_free(of2);
  fmbh_free(idx);
  fmbh_free(idx2);
  return FMB_OK;
}

extern "C" int fmbh_near_lists(fmb_tree *t, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int L = t->lmax, nb = t->nbox[L];
  int *cnt;
  int rc = fmbh_malloc((void **)&cnt, sizeof(int) * (nb + 1));
  if (!rc) rc = fmbh_malloc((void **)&t->d_near_start, sizeof(int) * (nb + 1));
  if (rc) return rc;
  FMB_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (nb + 1), st));
  hipLaunchKernelGGL(k_near_count, dim3(nblk(nb)), dim3(256), 0, st,
                     t->d_keys[L], nb, L, cnt);
  FMB_LAUNCH_CHECK();
  if ((rc = scan_exclusive(cnt, t->d_near_start, nb + 1, st))) return rc;
  int nnz;
  if ((rc = fmbh_d2h(&nnz, t->d_near_start + nb, sizeof(int), stream)))
    return rc;
  t->near_nnzbox = nnz;
  if ((rc = fmbh_malloc((void **)&t->d_near_box, sizeof(int) * nnz))) return rc;
  hipLaunchKernelGGL(k_near_fill, dim3(nblk(nb)), dim3(256), 0, st,
                     t->d_keys[L], nb, L, t->d_near_start, t->d_near_box);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(cnt);
  return FMB_OK;
}

extern "C" int fmbh_sorted_geometry(fmb_tree *t, const double *d_V,
                                    const int *d_F, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int n = t->nelem;
  int rc = fmbh_malloc((void **)&t->d_tri, sizeof(double) * 9 * n);
  if (!rc) rc = fmbh_malloc((void **)&t->d_geo, sizeof(double) * 8 * n);
  if (!rc) rc = fmbh_malloc((void **)&t->d_iperm, sizeof(int) * n);
  if (!rc) rc = fmbh_malloc((void **)&t->d_fv, sizeof(int) * 3 * n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sorted_geo, dim3(nblk(n)), dim3(256), 0, st, d_V, d_F,
                     t->d_perm, n, t->d_tri, t->d_geo, t->d_iperm, t->d_fv);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  return FMB_OK;
}
