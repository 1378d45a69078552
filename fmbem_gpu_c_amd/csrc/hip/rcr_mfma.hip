// rcr_mfma.hip -- rotate-coaxial-rotate translations (M2L, M2M, L2L) on the
// gfx950 matrix cores.
//
// Reference semantics: translation.cu:1711-1808 (one RCR translation) as used
// by OprL's SS / SR / RR passes (translation.cu:3461-3759).  For a box pair
// whose translation vector is t = (|t|, theta, phi) and source coefficients
// x of order pin, the target of order pout receives
//   a_n^m   = e^{i m phi} x_n^m                       n < pin
//   b_n^m'  = sum_m H_n^{m'm}(theta) a_n^m            real (2n+1)^2 blocks
//   c_n'^m  = sum_n C^{|m|}_{n'n}(|t|) b_n^m           n' < pout, n < pin
//   y_n^m'  = e^{-i m' phi} sum_m H_n^{m'm}(theta) c_n^m    n < pout
// with C at order P >= pin, pout (|m|-block layout, translation.cu:745-769).
//
// Every pair of one bucket (M2L: one of the 316 offsets; M2M / L2L: one of
// the 8 child octants) shares H and C, so a batch of V = 8*NT vectors
// (pair x right-hand side) is one block-diagonal GEMM per stage:
//   * X[f][col], f = n^2+n+m, col = 2*vector + {re, im} (16*NT real columns,
//     rows padded by 2 for the LDS banks);
//   * rotation block n: X_n <- H_n X_n, a real (s x s)(s x 16NT) product on
//     v_mfma_{f32,f64}_16x16x4 (A = H from global memory, B = X rows);
//   * coaxial block m: X_m <- C_m X_m, complex, as Cr*B + Ci*B' with
//     B' = (-im, re) formed from the same rows.
// Two workspace modes:
//   LDS: X in LDS, updated in place; each block is owned by one wave (LPT
//        schedule built on the host), which keeps the block's outputs in its
//        accumulators until its inputs are read; one barrier per stage.
//   GWS: X in a per-workgroup global workspace, two buffers ping-ponged
//        between stages, so blocks split into row groups of 16*MT rows that
//        any wave may take; a persistent grid walks the batches.
// Targets receive hardware float atomic adds, one contiguous vector per
// wave instruction.
#include "fmb_device.h"

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;

template <typename T> struct Mfma;
// Operand layouts of both 16x16x4 forms: A[i][k] in lane i + 16k,
// B[k][j] in lane j + 16k.  The result D[i][j] sits in lane j + 16q,
// element e, with row i = row(q, e): 4q + e for f32, q + 4e for f64.
template <> struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int e) { return 4 * q + e; }
};
template <> struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int e) { return q + 4 * e; }
};

template <int NT> struct Row {
  static constexpr int stride = 16 * NT + 2;  // reals per X row
};

struct RcrArgs {
  const int4 *chunks;  // {bucket, first pair, end pair, 0}
  int nchunks, nb;     // chunks, batches per chunk (grid y / item stride)
  const OffInfo *info; // per bucket: beta index, coaxial matrix index, phi
  const int *ptgt, *psrc;
  const void *H;       // [nbeta][rot_off(Pmax)]
  size_t hstride;
  const void *C;       // [ncoax][coax_size(P)]
  size_t cstride;
  int P, pin, pout;
  const int *sched;    // {w1, wc, w3}, [4][w1], [4][wc], [4][w3] unit codes
  const void *Min;     // [box][nrhs][pin^2]
  void *Lout;          // [box][nrhs][pout^2], added to
  int nrhs;
  void *gws;           // GWS mode: per-workgroup 2 x Pm^2 x RS reals
};

// Rotation stage over this wave's units n*16 + g (-1 terminated): row tiles
// [g*MT, g*MT + MT) of block n, read from Xi, written to Xo.
template <typename T, int NT, int MT>
__device__ __forceinline__ void rot_stage(const T *Xi, T *Xo,
                                          const T *__restrict__ H,
                                          const int *__restrict__ units,
                                          int width) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int RS = Row<NT>::stride;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  for (int j = 0; j < width; j++) {
    const int code = units[j];
    if (code < 0) break;
    const int n = code >> 4, t0 = (code & 15) * MT;
    const int s = 2 * n + 1, base = n * n;
    const int nt = min(MT, ((s + 15) >> 4) - t0), nks = (s + 3) >> 2;
    const T *Hn = H + rot_off(n);
    acc_t acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[mt][t] = acc_t{0, 0, 0, 0};
    for (int ks = 0; ks < nks; ks++) {
      const int k = 4 * ks + lk;
      const bool kin = k < s;
      T b[NT];
#pragma unroll
      for (int t = 0; t < NT; t++)
        b[t] = kin ? Xi[(base + k) * RS + 16 * t + li] : T(0);
      const T *hk = Hn + (size_t)k * s;
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        if (mt < nt) {
          const int i = 16 * (t0 + mt) + li;
          const T a = (kin && i < s) ? hk[i] : T(0);
#pragma unroll
          for (int t = 0; t < NT; t++) acc[mt][t] = Mfma<T>::mma(a, b[t], acc[mt][t]);
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      if (mt < nt) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int i = 16 * (t0 + mt) + Mfma<T>::row(lk, e);
          if (i < s) {
#pragma unroll
            for (int t = 0; t < NT; t++) Xo[(base + i) * RS + 16 * t + li] = acc[mt][t][e];
          }
        }
      }
    }
  }
}

// Coaxial stage over this wave's units (m + P)*16 + g: block m maps rows
// (n, m), n = |m| .. pin-1, to rows (n', m), n' = |m| .. pout-1.
template <typename T, int NT, int MT>
__device__ __forceinline__ void coax_stage(const T *Xi, T *Xo,
                                           const cplx<T> *__restrict__ C,
                                           const int *__restrict__ units,
                                           int width, int P, int pin, int pout) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int RS = Row<NT>::stride;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const bool odd = li & 1;
  for (int j = 0; j < width; j++) {
    const int code = units[j];
    if (code < 0) break;
    const int m = (code >> 4) - P, t0 = (code & 15) * MT;
    const int am = m < 0 ? -m : m, so = pout - am, si = pin - am, ld = P - am;
    const int nt = min(MT, ((so + 15) >> 4) - t0), nks = (si + 3) >> 2;
    const cplx<T> *Cm = C + coax_blk(P, am);
    acc_t acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[mt][t] = acc_t{0, 0, 0, 0};
    for (int ks = 0; ks < nks; ks++) {
      const int k = 4 * ks + lk;
      const bool kin = k < si;
      const int nn = am + k, row = nn * nn + nn + m;
      T b[NT], bq[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        cplx<T> z = cmk<T>(0, 0);
        if (kin) z = *(const cplx<T> *)(Xi + row * RS + 16 * t + (li & ~1));
        b[t] = odd ? z.im : z.re;
        bq[t] = odd ? z.re : -z.im;
      }
      const cplx<T> *ck = Cm + (size_t)k * ld;
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        if (mt < nt) {
          const int i = 16 * (t0 + mt) + li;
          cplx<T> a = cmk<T>(0, 0);
          if (kin && i < so) a = ck[i];
#pragma unroll
          for (int t = 0; t < NT; t++) {
            acc[mt][t] = Mfma<T>::mma(a.re, b[t], acc[mt][t]);
            acc[mt][t] = Mfma<T>::mma(a.im, bq[t], acc[mt][t]);
          }
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      if (mt < nt) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int i = 16 * (t0 + mt) + Mfma<T>::row(lk, e);
          if (i < so) {
            const int nn = am + i, row = nn * nn + nn + m;
#pragma unroll
            for (int t = 0; t < NT; t++) Xo[row * RS + 16 * t + li] = acc[mt][t][e];
          }
        }
      }
    }
  }
}

// One batch: vectors u0 .. u0+V-1 of chunk ch.
template <typename T, int NT, int MT, bool GWS>
__device__ __forceinline__ void rcr_batch(const RcrArgs &a, int4 ch, int u0,
                                          T *X0, T *X1, cplx<T> *ph,
                                          long long *vsrc, long long *vdst) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int pin = a.pin, pout = a.pout, P = a.P;
  const int Pm = pin > pout ? pin : pout;
  const int Pi2 = pin * pin, Po2 = pout * pout, Pm2 = Pm * Pm;
  const OffInfo oi = a.info[ch.x];
  const T *H = (const T *)a.H + a.hstride * oi.beta;
  const cplx<T> *C = (const cplx<T> *)a.C + a.cstride * oi.dist;
  const int U = (ch.z - ch.y) * a.nrhs;
  __syncthreads();
  for (int m = tid; m < Pm; m += kThreads) {
    double sn, cs;
    sincos(m * oi.phi, &sn, &cs);
    ph[m] = cmk<T>((T)cs, (T)sn);
  }
  if (tid < V) {
    const int u = u0 + tid;
    long long so = -1, d = -1;
    if (u < U) {
      const int q = ch.y + u / a.nrhs, r = u % a.nrhs;
      so = ((long long)a.psrc[q] * a.nrhs + r) * Pi2;
      d = ((long long)a.ptgt[q] * a.nrhs + r) * Po2;
    }
    vsrc[tid] = so;
    vdst[tid] = d;
  }
  __syncthreads();
  // source coefficients times e^{i m phi}; rows f >= pin^2 and absent
  // vectors are zero
  const cplx<T> *Min = (const cplx<T> *)a.Min;
  for (int v = 0; v < V; v++) {
    const long long so = vsrc[v];
    for (int f = tid; f < Pm2; f += kThreads) {
      cplx<T> z = cmk<T>(0, 0);
      if (so >= 0 && f < Pi2) {
        const int n = isqrt_fix(f), m = f - n * n - n;
        cplx<T> w = ph[m < 0 ? -m : m];
        if (m < 0) w.im = -w.im;
        z = cmul(Min[so + f], w);
      }
      *(cplx<T> *)(X0 + f * RS + 2 * v) = z;
    }
  }
  __syncthreads();
  const int *sc = a.sched;
  const int w1 = sc[0], wc = sc[1], w3 = sc[2];
  const int *u1 = sc + 3 + wave * w1;
  const int *uc = sc + 3 + kWaves * w1 + wave * wc;
  const int *u3 = sc + 3 + kWaves * (w1 + wc) + wave * w3;
  rot_stage<T, NT, MT>(X0, X1, H, u1, w1);
  __syncthreads();
  coax_stage<T, NT, MT>(X1, X0, C, uc, wc, P, pin, pout);
  __syncthreads();
  rot_stage<T, NT, MT>(X0, X1, H, u3, w3);
  __syncthreads();
  cplx<T> *Lout = (cplx<T> *)a.Lout;
  for (int v = 0; v < V; v++) {
    const long long d = vdst[v];
    if (d < 0) continue;
    for (int f = tid; f < Po2; f += kThreads) {
      const int n = isqrt_fix(f), m = f - n * n - n;
      cplx<T> w = ph[m < 0 ? -m : m];
      if (m > 0) w.im = -w.im;
      const cplx<T> z = *(const cplx<T> *)(X1 + f * RS + 2 * v);
      atomic_add_c(Lout + d + f, cmul(z, w));
    }
  }
}

// LDS mode: grid (nchunks, nb), one batch per workgroup.
template <typename T, int NT, int MT>
__global__ void __launch_bounds__(kThreads) k_rcr_lds(RcrArgs a) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  const int4 ch = a.chunks[blockIdx.x];
  const int u0 = blockIdx.y * V;
  if (u0 >= (ch.z - ch.y) * a.nrhs) return;
  T *X = (T *)smem;
  cplx<T> *ph = (cplx<T> *)(X + (size_t)Pm * Pm * RS);
  long long *vsrc = (long long *)(ph + Pm + (Pm & 1));
  rcr_batch<T, NT, MT, false>(a, ch, u0, X, X, ph, vsrc, vsrc + V);
}

// GWS mode: persistent grid over items chunk * nb + batch.
template <typename T, int NT, int MT>
__global__ void __launch_bounds__(kThreads) k_rcr_gws(RcrArgs a) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  const size_t slice = (size_t)Pm * Pm * RS;
  T *X0 = (T *)a.gws + 2 * slice * blockIdx.x;
  T *X1 = X0 + slice;
  cplx<T> *ph = (cplx<T> *)smem;
  long long *vsrc = (long long *)(ph + Pm + (Pm & 1));
  const long long items = (long long)a.nchunks * a.nb;
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    rcr_batch<T, NT, MT, true>(a, ch, b * V, X0, X1, ph, vsrc, vsrc + V);
  }
}

int mt_lds(int pm) {
  int tiles = (2 * pm - 1 + 15) / 16;
  if (tiles <= 2) return 2;
  if (tiles <= 3) return 3;
  if (tiles <= 4) return 4;
  if (tiles <= 6) return 6;
  return 0;
}

size_t small_lds(int pm, int nt, size_t tsz) {
  return (size_t)(pm + (pm & 1)) * 2 * tsz + 2 * 8 * nt * sizeof(long long);
}

size_t x_bytes(int pm, int nt, size_t tsz) {
  return (size_t)pm * pm * (16 * nt + 2) * tsz;
}

template <typename K>
int set_lds(K kern, size_t lds) {
  if (lds > 64 * 1024)
    FMB_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
  return FMB_OK;
}

template <typename T, int NT, int MT>
int launch(const fmbh_rcr_job *j, int nrhs, hipStream_t st) {
  constexpr int V = 8 * NT;
  const int Pm = j->pin > j->pout ? j->pin : j->pout;
  RcrArgs a;
  a.chunks = (const int4 *)j->chunks;
  a.nchunks = j->nchunks;
  a.nb = (j->chunk_pairs * nrhs + V - 1) / V;
  a.info = (const OffInfo *)j->info;
  a.ptgt = j->tgt;
  a.psrc = j->src;
  a.H = j->H;
  a.hstride = j->hstride;
  a.C = j->C;
  a.cstride = j->cstride;
  a.P = j->P;
  a.pin = j->pin;
  a.pout = j->pout;
  a.sched = j->sched;
  a.Min = j->in;
  a.Lout = j->out;
  a.nrhs = nrhs;
  a.gws = j->gws;
  const size_t tsz = sizeof(T);
  if (j->mode == FMBH_RCR_LDS) {
    const size_t lds = x_bytes(Pm, NT, tsz) + small_lds(Pm, NT, tsz);
    auto kern = k_rcr_lds<T, NT, MT>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(kern, dim3(a.nchunks, a.nb), dim3(kThreads), lds, st, a);
  } else {
    const size_t lds = small_lds(Pm, NT, tsz);
    auto kern = k_rcr_gws<T, NT, MT>;
    long long items = (long long)a.nchunks * a.nb;
    int grid = items < j->gws_grid ? (int)items : j->gws_grid;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds, st, a);
  }
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

template <typename T>
int dispatch(const fmbh_rcr_job *j, int nrhs, hipStream_t st) {
#define CASE(A, B) \
  if (j->nt == A && j->mt == B) return launch<T, A, B>(j, nrhs, st);
  CASE(1, 2) CASE(1, 3) CASE(1, 4) CASE(1, 6)
  CASE(2, 2) CASE(2, 3) CASE(2, 4) CASE(2, 6)
#undef CASE
  return fmb_set_error(FMB_ERR_RANGE, "no RCR tile shape nt=%d mt=%d", j->nt, j->mt);
}

}  // namespace

extern "C" int fmbh_rcr_shape(int pm, int prec, int *mode, int *nt, int *mt) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  *mt = mt_lds(pm);
  if (*mt && x_bytes(pm, 2, tsz) + small_lds(pm, 2, tsz) <= 80 * 1024) {
    *mode = FMBH_RCR_LDS;
    *nt = 2;
  } else if (*mt && x_bytes(pm, 1, tsz) + small_lds(pm, 1, tsz) <= 160 * 1024) {
    *mode = FMBH_RCR_LDS;
    *nt = 1;
  } else {
    *mode = FMBH_RCR_GWS;
    *nt = 2;
    *mt = 4;
  }
  return FMB_OK;
}

extern "C" size_t fmbh_rcr_gws_bytes(int pm, int prec, int nt) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  return 2 * x_bytes(pm, nt, tsz);
}

extern "C" int fmbh_rcr(const fmbh_rcr_job *j, int prec, int nrhs, void *stream) {
  if (j->nchunks == 0) return FMB_OK;
  if (prec == FMB_FP64) return dispatch<double>(j, nrhs, (hipStream_t)stream);
  return dispatch<float>(j, nrhs, (hipStream_t)stream);
}
