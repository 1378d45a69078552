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
// The operator blocks are stored zero-padded to 16-row / 4-column multiples
// (column-major, leading dimension = padded rows), so operand loads need no
// guards: rows of X beyond a block are read at a clamped index and meet
// zero operator entries.  Operands of the next KU K-steps are loaded into one
// register set while the MFMAs consume the other.
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

// Values the wave shares (schedule entries, offsets read from LDS) are
// moved to scalar registers so that the branches and addresses derived from
// them stay scalar.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int NT> struct Row {
  static constexpr int stride = 16 * NT + 2;  // reals per X row
};

// K-steps per operand register set
template <typename T, int NTL> struct Ahead {
  static constexpr int ku = (sizeof(T) == 4 || NTL <= 2) ? 4 : 2;
};

__device__ __forceinline__ int pad16(int s) { return (s + 15) & ~15; }

struct RcrArgs {
  const int4 *chunks;  // {bucket, first pair, end pair, 0}
  int nchunks, nb;     // chunks, batches per chunk (grid y / item stride)
  const OffInfo *info; // per bucket: beta index, coaxial matrix index, phi
  const void *phase;   // per bucket: e^{i m phi}, m < pstride
  int pstride;
  const int *ptgt, *psrc;
  const void *H;       // padded rotation blocks, per beta
  size_t hstride;
  const void *C;       // padded coaxial blocks, per coaxial matrix
  size_t cstride;
  int P, pin, pout;
  const int *sched;    // offsets [3][kWaves + 1], then {code, operator offset}
  int sched_len;
  const void *Min;     // [box][nrhs][pin^2]
  void *Lout;          // [box][nrhs][pout^2], added to
  int nrhs;
  void *gws;           // GWS mode: per-workgroup 2 x Pm^2 x RS reals
};

// Row tiles [t0, t0 + NTL) of rotation block n (size s, first row base):
// Xo[rows] = Hn Xi[rows].  Ab: the block (padded, leading dimension pad16(s)).
template <typename T, int NT, int NTL>
__device__ __forceinline__ void rot_unit(const T *Xi, T *Xo,
                                         const T *__restrict__ Ab, int s,
                                         int base, int t0) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int RS = Row<NT>::stride, KU = Ahead<T, NTL>::ku;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int ld = pad16(s), nks = (s + 3) >> 2, ng = (nks + KU - 1) / KU;
  const T *ap = Ab + (unsigned)(lk * ld + 16 * t0 + li);
  acc_t acc[NTL][NT];
#pragma unroll
  for (int mt = 0; mt < NTL; mt++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[mt][t] = acc_t{0, 0, 0, 0};
  T a0[KU][NTL], b0[KU][NT], a1[KU][NTL], b1[KU][NT];
  auto fetch = [&](int g, T(&fa)[KU][NTL], T(&fb)[KU][NT]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int ks = min(g * KU + u, nks - 1);
      const int k = min(4 * ks + lk, s - 1);
      const T *xr = Xi + (base + k) * RS + li;
#pragma unroll
      for (int t = 0; t < NT; t++) fb[u][t] = xr[16 * t];
      const T *ak = ap + (unsigned)(4 * ks * ld);
#pragma unroll
      for (int mt = 0; mt < NTL; mt++) fa[u][mt] = ak[16 * mt];
    }
  };
  auto mma = [&](int g, const T(&fa)[KU][NTL], const T(&fb)[KU][NT]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      if (g * KU + u < nks) {
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
#pragma unroll
          for (int t = 0; t < NT; t++) acc[mt][t] = Mfma<T>::mma(fa[u][mt], fb[u][t], acc[mt][t]);
      }
    }
  };
  fetch(0, a0, b0);
  for (int g = 0; g < ng; g += 2) {
    if (g + 1 < ng) fetch(g + 1, a1, b1);
    mma(g, a0, b0);
    if (g + 1 < ng) {
      if (g + 2 < ng) fetch(g + 2, a0, b0);
      mma(g + 1, a1, b1);
    }
  }
#pragma unroll
  for (int mt = 0; mt < NTL; mt++) {
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

// Row tiles [t0, t0 + NTL) of the coaxial blocks +am and (SG = 2) -am, which
// share the matrix C^{am}: rows (n', +-am), n' = am .. pout-1, from rows
// (n, +-am), n = am .. pin-1.  Ab: the block (padded, leading dimension
// pad16(so)).
template <typename T, int NT, int NTL, int SG>
__device__ __forceinline__ void coax_unit(const T *Xi, T *Xo,
                                          const cplx<T> *__restrict__ Ab,
                                          int am, int pin, int pout, int t0) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int RS = Row<NT>::stride, NB = SG * NT;
  constexpr int KU = (sizeof(T) == 4 && NTL * SG <= 3) ? 4 : 2;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const bool odd = li & 1;
  const int so = pout - am, si = pin - am;
  const int ld = pad16(so), nks = (si + 3) >> 2, ng = (nks + KU - 1) / KU;
  const cplx<T> *ap = Ab + (unsigned)(lk * ld + 16 * t0 + li);
  acc_t acc[NTL][NB];
#pragma unroll
  for (int mt = 0; mt < NTL; mt++)
#pragma unroll
    for (int t = 0; t < NB; t++) acc[mt][t] = acc_t{0, 0, 0, 0};
  cplx<T> a0[KU][NTL], z0[KU][NB], a1[KU][NTL], z1[KU][NB];
  auto fetch = [&](int g, cplx<T>(&fa)[KU][NTL], cplx<T>(&fz)[KU][NB]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int ks = min(g * KU + u, nks - 1);
      const int nn = am + min(4 * ks + lk, si - 1), rb = nn * nn + nn;
#pragma unroll
      for (int sg = 0; sg < SG; sg++) {
        const T *xr = Xi + (rb + (sg ? -am : am)) * RS + (li & ~1);
#pragma unroll
        for (int t = 0; t < NT; t++) fz[u][sg * NT + t] = *(const cplx<T> *)(xr + 16 * t);
      }
      const cplx<T> *ak = ap + (unsigned)(4 * ks * ld);
#pragma unroll
      for (int mt = 0; mt < NTL; mt++) fa[u][mt] = ak[16 * mt];
    }
  };
  auto mma = [&](int g, const cplx<T>(&fa)[KU][NTL], const cplx<T>(&fz)[KU][NB]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      if (g * KU + u < nks) {
        T b[NB], bq[NB];
#pragma unroll
        for (int t = 0; t < NB; t++) {
          b[t] = odd ? fz[u][t].im : fz[u][t].re;
          bq[t] = odd ? fz[u][t].re : -fz[u][t].im;
        }
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
#pragma unroll
          for (int t = 0; t < NB; t++) {
            acc[mt][t] = Mfma<T>::mma(fa[u][mt].re, b[t], acc[mt][t]);
            acc[mt][t] = Mfma<T>::mma(fa[u][mt].im, bq[t], acc[mt][t]);
          }
      }
    }
  };
  fetch(0, a0, z0);
  for (int g = 0; g < ng; g += 2) {
    if (g + 1 < ng) fetch(g + 1, a1, z1);
    mma(g, a0, z0);
    if (g + 1 < ng) {
      if (g + 2 < ng) fetch(g + 2, a0, z0);
      mma(g + 1, a1, z1);
    }
  }
#pragma unroll
  for (int mt = 0; mt < NTL; mt++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int i = 16 * (t0 + mt) + Mfma<T>::row(lk, e);
      if (i < so) {
        const int nn = am + i, rb = nn * nn + nn;
#pragma unroll
        for (int sg = 0; sg < SG; sg++)
#pragma unroll
          for (int t = 0; t < NT; t++)
            Xo[(rb + (sg ? -am : am)) * RS + 16 * t + li] = acc[mt][sg * NT + t][e];
      }
    }
  }
}

// runs unit<NTL = ntl> for a wave-uniform ntl in [1, MT]
template <typename T, int NT, int MT>
__device__ __forceinline__ void rot_any(int ntl, const T *Xi, T *Xo,
                                        const T *Ab, int s, int base, int t0) {
  if constexpr (MT > 1) {
    if (ntl < MT) {
      rot_any<T, NT, MT - 1>(ntl, Xi, Xo, Ab, s, base, t0);
      return;
    }
  }
  rot_unit<T, NT, MT>(Xi, Xo, Ab, s, base, t0);
}

template <typename T, int NT, int MT, int SG>
__device__ __forceinline__ void coax_any(int ntl, const T *Xi, T *Xo,
                                         const cplx<T> *Ab, int am, int pin,
                                         int pout, int t0) {
  if constexpr (MT > 1) {
    if (ntl < MT) {
      coax_any<T, NT, MT - 1, SG>(ntl, Xi, Xo, Ab, am, pin, pout, t0);
      return;
    }
  }
  coax_unit<T, NT, MT, SG>(Xi, Xo, Ab, am, pin, pout, t0);
}

// Rotation stage over this wave's units {n*16 + g, operator offset}: row
// tiles [g*MT, g*MT + MT) of block n.
template <typename T, int NT, int MT>
__device__ __forceinline__ void rot_stage(const T *Xi, T *Xo, const T *H,
                                          const int *units, int count) {
  for (int j = 0; j < count; j++) {
    const int code = uni(units[2 * j]), aoff = uni(units[2 * j + 1]);
    const int n = code >> 4, t0 = (code & 15) * MT;
    const int s = 2 * n + 1;
    const int ntl = min(MT, ((s + 15) >> 4) - t0);
    rot_any<T, NT, MT>(ntl, Xi, Xo, H + aoff, s, n * n, t0);
  }
}

// Coaxial stage over this wave's units {am*16 + g, operator offset}: the
// blocks +am and -am together.
template <typename T, int NT, int MT>
__device__ __forceinline__ void coax_stage(const T *Xi, T *Xo, const cplx<T> *C,
                                           const int *units, int count, int pin,
                                           int pout) {
  for (int j = 0; j < count; j++) {
    const int code = uni(units[2 * j]), aoff = uni(units[2 * j + 1]);
    const int am = code >> 4, t0 = (code & 15) * MT;
    const int ntl = min(MT, ((pout - am + 15) >> 4) - t0);
    if (am == 0)
      coax_any<T, NT, MT, 1>(ntl, Xi, Xo, C + aoff, am, pin, pout, t0);
    else
      coax_any<T, NT, MT, 2>(ntl, Xi, Xo, C + aoff, am, pin, pout, t0);
  }
}

// One batch: vectors u0 .. u0+V-1 of chunk ch.
template <typename T, int NT, int MT, bool GWS>
__device__ __forceinline__ void rcr_batch(const RcrArgs &a, int4 ch, int u0,
                                          T *X0, T *X1, cplx<T> *ph,
                                          const int *sched) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  const int tid = threadIdx.x, wave = uni(tid >> 6);
  const int pin = a.pin, pout = a.pout;
  const int Pm = pin > pout ? pin : pout;
  const int Pi2 = pin * pin, Po2 = pout * pout, Pm2 = Pm * Pm;
  const OffInfo oi = a.info[ch.x];
  const T *H = (const T *)a.H + a.hstride * oi.beta;
  const cplx<T> *C = (const cplx<T> *)a.C + a.cstride * oi.dist;
  const int U = (ch.z - ch.y) * a.nrhs;
  __syncthreads();
  const cplx<T> *phb = (const cplx<T> *)a.phase + (size_t)a.pstride * ch.x;
  for (int m = tid; m < Pm; m += kThreads) ph[m] = phb[m];
  // vector v of the batch: (pair, right-hand side) = divmod(u0 + v, nrhs);
  // its source and target base pointers are wave-uniform
  const cplx<T> *vin[V];
  cplx<T> *vout[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    const int u = u0 + v;
    vin[v] = nullptr;
    vout[v] = nullptr;
    if (u < U) {
      const int q = ch.y + u / a.nrhs, r = u % a.nrhs;
      vin[v] = (const cplx<T> *)a.Min + ((size_t)a.psrc[q] * a.nrhs + r) * Pi2;
      vout[v] = (cplx<T> *)a.Lout + ((size_t)a.ptgt[q] * a.nrhs + r) * Po2;
    }
  }
  __syncthreads();
  // source coefficients times e^{i m phi}; rows f >= pin^2 and absent
  // vectors are zero.  One thread per row f: its V loads are in flight
  // together, the degree and phase of f are computed once.
  for (int f = tid; f < Pm2; f += kThreads) {
    cplx<T> z[V];
#pragma unroll
    for (int v = 0; v < V; v++)
      z[v] = (vin[v] && f < Pi2) ? vin[v][(unsigned)f] : cmk<T>(0, 0);
    const int n = isqrt_fix(f), m = f - n * n - n;
    cplx<T> w = ph[m < 0 ? -m : m];
    if (m < 0) w.im = -w.im;
#pragma unroll
    for (int v = 0; v < V; v++) *(cplx<T> *)(X0 + f * RS + 2 * v) = cmul(z[v], w);
  }
  __syncthreads();
  const int *o1 = sched, *oc = sched + kWaves + 1, *o3 = sched + 2 * (kWaves + 1);
  const int *u1 = sched + 2 * uni(o1[wave]), w1 = uni(o1[wave + 1] - o1[wave]);
  const int *uc = sched + 2 * uni(oc[wave]), wc = uni(oc[wave + 1] - oc[wave]);
  const int *u3 = sched + 2 * uni(o3[wave]), w3 = uni(o3[wave + 1] - o3[wave]);
  rot_stage<T, NT, MT>(X0, X1, H, u1, w1);
  __syncthreads();
  coax_stage<T, NT, MT>(X1, X0, C, uc, wc, pin, pout);
  __syncthreads();
  rot_stage<T, NT, MT>(X0, X1, H, u3, w3);
  __syncthreads();
  // add e^{-i m' phi} y into the targets.  Thread e adds real e of each
  // target vector (part e & 1 of coefficient e >> 1), so that one wave's
  // atomic adds fill whole cache lines: the adds are performed beyond the
  // L2, and half-filled lines cost nearly as much as full ones.
  for (int e = tid; e < 2 * Po2; e += kThreads) {
    const int f = e >> 1, n = isqrt_fix(f), m = f - n * n - n;
    const bool im = e & 1;
    cplx<T> w = ph[m < 0 ? -m : m];
    if (m > 0) w.im = -w.im;
#pragma unroll
    for (int v = 0; v < V; v++)
      if (vout[v]) {
        const cplx<T> x = *(const cplx<T> *)(X1 + f * RS + 2 * v);
        const T y = im ? x.re * w.im + x.im * w.re : x.re * w.re - x.im * w.im;
        unsafeAtomicAdd((T *)vout[v] + (unsigned)e, y);
      }
  }
}

// LDS mode: grid (nchunks, nb), one batch per workgroup.
// FP32 batches ask for 3 waves per SIMD (the LDS of three workgroups fits a
// CU); FP64 keeps its accumulators unspilled instead.
template <typename T, int NT, int MT>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) k_rcr_lds(RcrArgs a) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  const int4 ch = a.chunks[blockIdx.x];
  const int u0 = blockIdx.y * V;
  if (u0 >= (ch.z - ch.y) * a.nrhs) return;
  T *X = (T *)smem;
  cplx<T> *ph = (cplx<T> *)(X + (size_t)Pm * Pm * RS);
  int *sched = (int *)(ph + Pm + (Pm & 1));
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads) sched[i] = a.sched[i];
  rcr_batch<T, NT, MT, false>(a, ch, u0, X, X, ph, sched);
}

// LDS mode on a persistent grid (one workgroup per free slot of the GPU):
// each workgroup loads the schedule once and loops over the items chunk * nb
// + batch, so workgroups are not dispatched (and set up) per batch.
template <typename T, int NT, int MT>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) k_rcr_ldsp(RcrArgs a) {
  constexpr int V = 8 * NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  T *X = (T *)smem;
  cplx<T> *ph = (cplx<T> *)(X + (size_t)Pm * Pm * Row<NT>::stride);
  int *sched = (int *)(ph + Pm + (Pm & 1));
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads) sched[i] = a.sched[i];
  const long long items = (long long)a.nchunks * a.nb;
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    rcr_batch<T, NT, MT, false>(a, ch, b * V, X, X, ph, sched);
  }
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
  int *sched = (int *)(ph + Pm + (Pm & 1));
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads) sched[i] = a.sched[i];
  const long long items = (long long)a.nchunks * a.nb;
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    rcr_batch<T, NT, MT, true>(a, ch, b * V, X0, X1, ph, sched);
  }
}

// zero-padded copies of the operator blocks (see the file comment)
template <typename T>
__global__ void k_pad_rot(const T *__restrict__ H, size_t hs, const int *__restrict__ hoff,
                          size_t ps, T *__restrict__ Hp) {
  const int n = blockIdx.x, b = blockIdx.y;
  const int s = 2 * n + 1, ld = (s + 15) & ~15, nc = (s + 3) & ~3;
  const T *src = H + hs * b + rot_off(n);
  T *dst = Hp + ps * b + hoff[n];
  for (int e = threadIdx.x; e < ld * nc; e += blockDim.x) {
    const int i = e % ld, k = e / ld;
    dst[e] = (i < s && k < s) ? src[k * s + i] : T(0);
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

int mt_lds(int pm) {
  int tiles = (2 * pm - 1 + 15) / 16;
  if (tiles <= 2) return 2;
  if (tiles <= 3) return 3;
  if (tiles <= 4) return 4;
  if (tiles <= 6) return 6;
  return 0;
}

// phase table and a schedule of sched_len ints
size_t small_lds(int pm, size_t tsz, int sched_len) {
  return (size_t)(pm + (pm & 1)) * 2 * tsz + sizeof(int) * (size_t)sched_len;
}

size_t x_bytes(int pm, int nt, size_t tsz) {
  return (size_t)pm * pm * (16 * nt + 2) * tsz;
}

// Raises the kernel's dynamic-LDS cap to the CU's 160 KB (not to this
// launch's size): launches captured into a graph replay after every capture
// has run, so a cap lowered by a later, smaller launch would invalidate them.
template <typename K>
int set_lds(K kern, size_t lds) {
  if (lds > 160 * 1024)
    return fmb_set_error(FMB_ERR_RANGE, "translation batch needs %zu B of LDS", lds);
  if (lds > 64 * 1024)
    FMB_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024));
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
  a.phase = j->phase;
  a.pstride = j->pstride;
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
  a.sched_len = j->sched_len;
  a.Min = j->in;
  a.Lout = j->out;
  a.nrhs = nrhs;
  a.gws = j->gws;
  const size_t tsz = sizeof(T);
  if (j->mode == FMBH_RCR_LDS) {
    const size_t lds = x_bytes(Pm, NT, tsz) + small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_lds<T, NT, MT>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(kern, dim3(a.nchunks, a.nb), dim3(kThreads), lds, st, a);
  } else if (j->mode == FMBH_RCR_LDSP) {
    const size_t lds = x_bytes(Pm, NT, tsz) + small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_ldsp<T, NT, MT>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    int per_cu = 0, ncu = 0, dev = 0;
    FMB_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, lds));
    FMB_HIP_CHECK(hipGetDevice(&dev));
    FMB_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu < 1) per_cu = 1;
    long long items = (long long)a.nchunks * a.nb, cap = (long long)per_cu * ncu;
    int grid = (int)(items < cap ? items : cap);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds, st, a);
  } else {
    const size_t lds = small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_gws<T, NT, MT>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
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

extern "C" int fmbh_rcr_mt_lds(int pm) { return mt_lds(pm); }

extern "C" size_t fmbh_rcr_gws_bytes(int pm, int prec, int nt) {
  const size_t tsz = prec == FMB_FP64 ? sizeof(double) : sizeof(float);
  return 2 * x_bytes(pm, nt, tsz);
}

extern "C" int fmbh_rcr(const fmbh_rcr_job *j, int prec, int nrhs, void *stream) {
  if (j->nchunks == 0) return FMB_OK;
  if (prec == FMB_FP64) return dispatch<double>(j, nrhs, (hipStream_t)stream);
  return dispatch<float>(j, nrhs, (hipStream_t)stream);
}

extern "C" int fmbh_pad_rot(int prec, const void *d_H, int nbeta, int Pmax,
                            const int *d_hoff, size_t pstride, void *d_Hp,
                            void *stream) {
  hipStream_t st = (hipStream_t)stream;
  size_t hs = (size_t)rot_off(Pmax);
  if (prec == FMB_FP64)
    hipLaunchKernelGGL(k_pad_rot<double>, dim3(Pmax, nbeta), dim3(256), 0, st,
                       (const double *)d_H, hs, d_hoff, pstride, (double *)d_Hp);
  else
    hipLaunchKernelGGL(k_pad_rot<float>, dim3(Pmax, nbeta), dim3(256), 0, st,
                       (const float *)d_H, hs, d_hoff, pstride, (float *)d_Hp);
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
