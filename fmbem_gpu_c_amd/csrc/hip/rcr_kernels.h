// rcr_kernels.h -- rotate-coaxial-rotate translations (M2L, M2M, L2L) on the
// gfx950 matrix cores: the kernels and their launch, instantiated per
// precision and wave count in rcr_{f32,f64}_w{4,8}.hip (built in parallel)
// and dispatched by fmbh_rcr (rcr_mfma.hip).
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
//   * rotation of degree n: H_n^{m'm} = H_n^{-m',-m}, so in the even/odd
//     coordinates e_m = (z_m + z_{-m}) / 2 (e_0 = z_0) and o_m = (z_m -
//     z_{-m}) / 2 it splits into an even block E_n ((n+1) x (n+1)) and an
//     odd block O_n (n x n), about half the products of H_n: real (s x s)(s
//     x 16NT) products on v_mfma_{f32,f64}_16x16x4 (A = E_n, O_n from global
//     memory, B = X rows).  The first rotation forms the coordinates from
//     the rows +-m as it reads them and leaves e_m in row n^2+n+m, o_m in
//     row n^2+n-m; the back rotation recombines z_{+-m} = e_m +- o_m in its
//     accumulators;
//   * coaxial block m: X_m <- C_m X_m, complex, as Cr*B + Ci*B' with
//     B' = (-im, re) formed from the same rows.  C^{|m|} acts alike on the
//     rows +m and -m, so it maps even/odd coordinates to even/odd ones.
// The operator blocks are stored zero-padded to 16-row / 4-column multiples
// (column-major, leading dimension = padded rows), so operand loads need no
// guards: rows of X beyond a block are read at a clamped index and meet
// zero operator entries.  Operands of the next KU K-steps are loaded into one
// register set while the MFMAs consume the other; operator blocks in global
// memory through buffer loads (BUF), LDS copies (LDSO) through pointers;
// the batch rows likewise by buffer loads where they sit in the global
// workspace (XG, GWS mode), by LDS reads otherwise.
// Two workspace modes:
//   LDS: X in LDS, updated in place; each block is owned by one wave (LPT
//        schedule built on the host), which keeps the block's outputs in its
//        accumulators until its inputs are read; one barrier per stage.
//   GWS: X in a per-workgroup global workspace, two buffers ping-ponged
//        between stages, so blocks split into row groups of 16*MT rows that
//        any wave may take; a persistent grid walks the batches.
//   LDSO: as LDS, on a persistent grid whose workgroups each take a
//        contiguous range of batches (so of one or two buckets) and hold the
//        bucket's padded H and C blocks in LDS, copied in when the bucket's
//        polar angle or distance changes: operand A is then read from LDS
//        instead of L2 (low orders only, where operators and batch fit the
//        160 KB).
// A workgroup has W = 4 or 8 waves (a tile-shape choice).  Targets receive
// hardware float atomic adds; the lanes of a wave add consecutive reals of
// one target vector, so each instruction fills whole cache lines.

#pragma once
#include "fmb_device.h"

namespace fmbh_rcr_detail {
int run_f32_w4(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
int run_f32_w8(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
int run_f64_w4(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
int run_f64_w8(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
int run_bf_w4(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
int run_bf_w8(const fmbh_rcr_job *j, int nrhs, hipStream_t st);
}  // namespace fmbh_rcr_detail

namespace {

// W waves per workgroup (4 or 8, a tile-shape choice); kThreads<W> lanes.
template <int W> constexpr int kThreads = 64 * W;

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

// reals per X row: 16 NT columns and kRowPad more for the LDS banks; the
// kernels' row stride and every host-side size (x_bytes) use this one value
constexpr int kRowPad = 2;
template <int NT> struct Row {
  static constexpr int stride = 16 * NT + kRowPad;
};

// K-steps per operand register set
template <typename T, int NTL> struct Ahead {
  static constexpr int ku = (sizeof(T) == 4 || NTL <= 2) ? 4 : 2;
};

__device__ __forceinline__ int pad16(int s) { return (s + 15) & ~15; }

// Operator blocks in global memory are read by buffer loads: the block's
// base in a scalar resource (raw, stride 0, a 2 GB range, the gfx9 data
// format word), the lane's byte offset in one VGPR for the whole K loop, the
// K-step's byte offset as a scalar.  This spares the 64-bit address
// arithmetic of a pointer load per operand (fewer VALU instructions; the
// level-7 FP32 M2L 3% faster).  The offsets stay those of the pointer
// loads, inside the padded blocks.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t op_rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7fffffff,
                                           0x00020000);
}
template <typename V> struct OpLoad;
template <> struct OpLoad<float> {
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
  }
};
template <> struct OpLoad<double> {
  static __device__ __forceinline__ double ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
  }
};
template <> struct OpLoad<cplx<float>> {
  static __device__ __forceinline__ cplx<float> ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(cplx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
  }
};
template <> struct OpLoad<cplx<double>> {
  static __device__ __forceinline__ cplx<double> ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(cplx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
  }
};

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
  const void *Hb;      // split-BF16 rotation blocks, per beta: hi, then lo
  size_t hbstride, hbhalf;
  int P, pin, pout;
  const int *sched;    // offsets [3][W + 1], then {code, operator offset}
  int sched_len;
  int hlen;            // LDSO: T values of the padded H blocks n < P
  int xcd;             // LDS: 1-D grid, XCD x takes items [x I / 8, (x+1) I / 8)
  const void *Min;     // [box][nrhs][pin^2]
  void *Lout;          // [box][nrhs][pout^2], added to
  int nrhs;
  void *gws;           // GWS mode: per-workgroup 2 x Pm^2 x RS reals
};

// values of a padded even/odd rotation block of s rows (16-row / 4-column
// multiples)
__device__ __forceinline__ int eo_block(int s) { return ((s + 15) & ~15) * ((s + 3) & ~3); }

// Row tiles [t0, t0 + NTL) of degree n's rotation in even/odd coordinates.
// Index i of the even block (n + 1 rows and columns) stands for m_i = i + 1
// (i < n) and m_n = 0; index i of the odd block (n x n) for |m| = i + 1, so
// that even and odd row i of a tile sit in the same lane and element.
//   DIR 0 (first rotation): the batch rows hold standard coordinates z_m;
//     the operands are e = z_m + z_{-m} and o = z_m - z_{-m} (so e_0 = 2 z_0),
//     and the outputs go to the rows of e_m (n^2 + n + m) and o_m
//     (n^2 + n - m): even/odd coordinates for the coaxial stage.
//   DIR 1 (back rotation): the rows hold even/odd coordinates; the outputs
//     are recombined to y_{+-m} = e_m +- o_m, y_0 = e_0, in standard rows.
// Both blocks run in one K loop (their operand streams in flight together).
// Ab: the even block (padded, leading dimension pad16(n + 1)); the odd block
// follows it.  The odd block has NTL or NTL - 1 of the tiles (none for n = 0)
// and no more K-steps than the even one; its missing tiles and K-steps are
// skipped (wave-uniform) and its operands read at clamped, in-bounds
// positions.
template <typename T, int NT, int NTL, int DIR, bool BUF = false, bool XG = false>
__device__ __forceinline__ void rot_unit(const T *Xi, T *Xo, const T *__restrict__ Ab,
                                         int n, int t0) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int RS = Row<NT>::stride, KU = Ahead<T, NTL>::ku;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int se = n + 1, so = n, c0 = n * n + n;  // c0: the row of m = 0
  const int lde = pad16(se), ldo = pad16(so > 0 ? so : 1);
  const int nke = (se + 3) >> 2, nko = (so + 3) >> 2, ng = (nke + KU - 1) / KU;
  const int nto = min(NTL, ((so + 15) >> 4) - t0);  // odd tiles of this unit
  const T *ape = Ab + (unsigned)(lk * lde + 16 * t0 + li);
  // (a unit past the odd block's tiles reads the even block's operands)
  const T *apo = nto > 0 ? Ab + (unsigned)(eo_block(se) + lk * ldo + 16 * t0 + li) : ape;
  constexpr int TS = (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t ar = op_rsrc(Ab);
  const int vae = TS * (lk * lde + 16 * t0 + li);
  const int vao = nto > 0 ? TS * (eo_block(se) + lk * ldo + 16 * t0 + li) : vae;
  const __amdgpu_buffer_rsrc_t xrs = op_rsrc(Xi);
  // set by the first K-step (every block has one), which multiplies into a
  // zero accumulator operand instead of cleared registers
  acc_t acce[NTL][NT], acco[NTL][NT];
  T ae0[KU][NTL], ao0[KU][NTL], be0[KU][NT], bo0[KU][NT];
  T ae1[KU][NTL], ao1[KU][NTL], be1[KU][NT], bo1[KU][NT];
  auto fetch = [&](int g, T(&fae)[KU][NTL], T(&fao)[KU][NTL], T(&fbe)[KU][NT],
                   T(&fbo)[KU][NT]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int ks = min(g * KU + u, nke - 1);
      const int k = min(4 * ks + lk, n), m = k < n ? k + 1 : 0;
      const T *xp = Xi + (c0 + m) * RS + li, *xm = Xi + (c0 - m) * RS + li;
      const int xop = TS * ((c0 + m) * RS + li), xom = TS * ((c0 - m) * RS + li);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        T p, q;
        if constexpr (XG) {
          p = OpLoad<T>::ld(xrs, xop + TS * 16 * t, 0);
          q = OpLoad<T>::ld(xrs, xom + TS * 16 * t, 0);
        } else {
          p = xp[16 * t];
          q = xm[16 * t];
        }
        if constexpr (DIR == 0) {
          fbe[u][t] = p + q;
          fbo[u][t] = p - q;
        } else {
          fbe[u][t] = p;
          fbo[u][t] = q;
        }
      }
      const int kso = min(ks, max(nko - 1, 0));
      if constexpr (BUF) {
        const int ske = uni(TS * 4 * ks * lde), sko = uni(TS * 4 * kso * ldo);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
          fae[u][mt] = OpLoad<T>::ld(ar, vae + TS * 16 * mt, ske);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
          fao[u][mt] = OpLoad<T>::ld(ar, vao + TS * 16 * (mt < nto ? mt : 0), sko);
      } else {
        const T *ak = ape + (unsigned)(4 * ks * lde);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++) fae[u][mt] = ak[16 * mt];
        const T *ao = apo + (unsigned)(4 * kso * ldo);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++) fao[u][mt] = ao[16 * (mt < nto ? mt : 0)];
      }
    }
  };
  auto mma = [&](int g, const T(&fae)[KU][NTL], const T(&fao)[KU][NTL],
                 const T(&fbe)[KU][NT], const T(&fbo)[KU][NT]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int ks = g * KU + u;
      if (ks == 0) {
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
#pragma unroll
          for (int t = 0; t < NT; t++)
            acce[mt][t] = Mfma<T>::mma(fae[u][mt], fbe[u][t], acc_t{0, 0, 0, 0});
      } else if (ks < nke) {
#pragma unroll
        for (int mt = 0; mt < NTL; mt++)
#pragma unroll
          for (int t = 0; t < NT; t++) acce[mt][t] = Mfma<T>::mma(fae[u][mt], fbe[u][t], acce[mt][t]);
      }
#pragma unroll
      for (int mt = 0; mt < NTL; mt++)
        if (mt < nto) {
          if (ks == 0) {
#pragma unroll
            for (int t = 0; t < NT; t++)
              acco[mt][t] = Mfma<T>::mma(fao[u][mt], fbo[u][t], acc_t{0, 0, 0, 0});
          } else if (ks < nko) {
#pragma unroll
            for (int t = 0; t < NT; t++)
              acco[mt][t] = Mfma<T>::mma(fao[u][mt], fbo[u][t], acco[mt][t]);
          }
        }
    }
  };
  fetch(0, ae0, ao0, be0, bo0);
  for (int g = 0; g < ng; g += 2) {
    if (g + 1 < ng) fetch(g + 1, ae1, ao1, be1, bo1);
    mma(g, ae0, ao0, be0, bo0);
    if (g + 1 < ng) {
      if (g + 2 < ng) fetch(g + 2, ae0, ao0, be0, bo0);
      mma(g + 1, ae1, ao1, be1, bo1);
    }
  }
  // tile row i of both blocks: |m| = i + 1 (i < n, where the odd block has
  // the tile too) or m = 0 (i = n, even block only); padding rows are skipped
#pragma unroll
  for (int mt = 0; mt < NTL; mt++) {
    const bool odd = mt < nto;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int i = 16 * (t0 + mt) + Mfma<T>::row(lk, e);
      if (i > n) continue;
      const int m = i < n ? i + 1 : 0;
      T *yp = Xo + (c0 + m) * RS + li, *ym = Xo + (c0 - m) * RS + li;
      if (i == n || !odd) {
#pragma unroll
        for (int t = 0; t < NT; t++) yp[16 * t] = acce[mt][t][e];
      } else if constexpr (DIR == 0) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          yp[16 * t] = acce[mt][t][e];
          ym[16 * t] = acco[mt][t][e];
        }
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const T ev = acce[mt][t][e], ov = acco[mt][t][e];
          yp[16 * t] = ev + ov;
          ym[16 * t] = ev - ov;
        }
      }
    }
  }
}

// Row tiles [t0, t0 + NTL) of the coaxial blocks +am and (SG = 2) -am, which
// share the matrix C^{am}: rows (n', +-am), n' = am .. pout-1, from rows
// (n, +-am), n = am .. pin-1.  Ab: the block (padded, leading dimension
// pad16(so)).
template <typename T, int NT, int NTL, int SG, bool BUF = false, bool XG = false>
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
  constexpr int CS = (int)sizeof(cplx<T>);
  const __amdgpu_buffer_rsrc_t ar = op_rsrc(Ab);
  const int va = CS * (lk * ld + 16 * t0 + li);
  const __amdgpu_buffer_rsrc_t xrs = op_rsrc(Xi);
  // set by the first K-step (every block has one), as in rot_unit
  acc_t acc[NTL][NB];
  cplx<T> a0[KU][NTL], z0[KU][NB], a1[KU][NTL], z1[KU][NB];
  auto fetch = [&](int g, cplx<T>(&fa)[KU][NTL], cplx<T>(&fz)[KU][NB]) {
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int ks = min(g * KU + u, nks - 1);
      const int nn = am + min(4 * ks + lk, si - 1), rb = nn * nn + nn;
#pragma unroll
      for (int sg = 0; sg < SG; sg++) {
        const T *xr = Xi + (rb + (sg ? -am : am)) * RS + (li & ~1);
        const int xo = (int)sizeof(T) * ((rb + (sg ? -am : am)) * RS + (li & ~1));
#pragma unroll
        for (int t = 0; t < NT; t++) {
          if constexpr (XG)
            fz[u][sg * NT + t] = OpLoad<cplx<T>>::ld(xrs, xo + (int)sizeof(T) * 16 * t, 0);
          else
            fz[u][sg * NT + t] = *(const cplx<T> *)(xr + 16 * t);
        }
      }
      if constexpr (BUF) {
        const int sk = uni(CS * 4 * ks * ld);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++) fa[u][mt] = OpLoad<cplx<T>>::ld(ar, va + CS * 16 * mt, sk);
      } else {
        const cplx<T> *ak = ap + (unsigned)(4 * ks * ld);
#pragma unroll
        for (int mt = 0; mt < NTL; mt++) fa[u][mt] = ak[16 * mt];
      }
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
        if (u == 0 && g == 0) {
#pragma unroll
          for (int mt = 0; mt < NTL; mt++)
#pragma unroll
            for (int t = 0; t < NB; t++) {
              acc[mt][t] = Mfma<T>::mma(fa[u][mt].re, b[t], acc_t{0, 0, 0, 0});
              acc[mt][t] = Mfma<T>::mma(fa[u][mt].im, bq[t], acc[mt][t]);
            }
        } else {
#pragma unroll
          for (int mt = 0; mt < NTL; mt++)
#pragma unroll
            for (int t = 0; t < NB; t++) {
              acc[mt][t] = Mfma<T>::mma(fa[u][mt].re, b[t], acc[mt][t]);
              acc[mt][t] = Mfma<T>::mma(fa[u][mt].im, bq[t], acc[mt][t]);
            }
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
  // rows of a full tile are stored unguarded (a wave-uniform test)
  auto put = [&](int mt, int e, int i) {
    const int nn = am + i, rb = nn * nn + nn;
#pragma unroll
    for (int sg = 0; sg < SG; sg++)
#pragma unroll
      for (int t = 0; t < NT; t++)
        Xo[(rb + (sg ? -am : am)) * RS + 16 * t + li] = acc[mt][sg * NT + t][e];
  };
#pragma unroll
  for (int mt = 0; mt < NTL; mt++) {
    const int r0 = 16 * (t0 + mt);
    if (r0 + 16 <= so) {
#pragma unroll
      for (int e = 0; e < 4; e++) put(mt, e, r0 + Mfma<T>::row(lk, e));
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int i = r0 + Mfma<T>::row(lk, e);
        if (i < so) put(mt, e, i);
      }
    }
  }
}

// runs unit<NTL = ntl> for a wave-uniform ntl in [1, MT]
template <typename T, int NT, int MT, int DIR, bool BUF = false, bool XG = false>
__device__ __forceinline__ void rot_any(int ntl, const T *Xi, T *Xo, const T *Ab, int n,
                                        int t0) {
  if constexpr (MT > 1) {
    if (ntl < MT) {
      rot_any<T, NT, MT - 1, DIR, BUF, XG>(ntl, Xi, Xo, Ab, n, t0);
      return;
    }
  }
  rot_unit<T, NT, MT, DIR, BUF, XG>(Xi, Xo, Ab, n, t0);
}

template <typename T, int NT, int MT, int SG, bool BUF = false, bool XG = false>
__device__ __forceinline__ void coax_any(int ntl, const T *Xi, T *Xo,
                                         const cplx<T> *Ab, int am, int pin,
                                         int pout, int t0) {
  if constexpr (MT > 1) {
    if (ntl < MT) {
      coax_any<T, NT, MT - 1, SG, BUF, XG>(ntl, Xi, Xo, Ab, am, pin, pout, t0);
      return;
    }
  }
  coax_unit<T, NT, MT, SG, BUF, XG>(Xi, Xo, Ab, am, pin, pout, t0);
}

// Rotation stage over this wave's units {n*16 + g, operator offset}: row
// tiles [g*MT, g*MT + MT) of degree n's even and odd blocks.
template <typename T, int NT, int MT, int DIR, bool BUF = false, bool XG = false>
__device__ __forceinline__ void rot_stage(const T *Xi, T *Xo, const T *H,
                                          const int *units, int count) {
  for (int j = 0; j < count; j++) {
    const int code = uni(units[2 * j]), aoff = uni(units[2 * j + 1]);
    const int n = code >> 4, t0 = (code & 15) * MT;
    const int ntl = min(MT, ((n + 1 + 15) >> 4) - t0);
    rot_any<T, NT, MT, DIR, BUF, XG>(ntl, Xi, Xo, H + aoff, n, t0);
  }
}

// ---- split-BF16 rotations (FP32 batches, FP32 accumulation) ----
// v_mfma_f32_16x16x32_bf16: A[i][k] in lane i + 16 (k / 8), element k % 8,
// and B[k][j] in lane j + 16 (k / 8), element k % 8; D as in the FP32 form.
// Each operand is a sum hi + lo of two bf16 values, and a product takes the
// three terms hi.hi + hi.lo + lo.hi (relative error about 2^-16).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Row tiles [t0, t0 + NTL) of the standard rotation block n (size s = 2n+1,
// first row base = n^2) in split-BF16, on the same batch rows as the
// even/odd units (standard coordinates in, standard out): Ab is the block's
// hi part (rows padded to 16, K to 32, row-major); its lo part lies `half`
// values further on.
template <int NT, int NTL>
__device__ __forceinline__ void rot_unit_bf(const float *Xi, float *Xo,
                                            const __bf16 *__restrict__ Ab, size_t half,
                                            int s, int base, int t0) {
  using acc_t = Mfma<float>::acc_t;
  constexpr int RS = Row<NT>::stride;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int kp = (s + 31) & ~31, nks = kp >> 5;
  const __bf16 *ap = Ab + (size_t)(16 * t0 + li) * kp + 8 * lk;
  acc_t acc[NTL][NT];
  for (int ks = 0; ks < nks; ks++) {
    bf16x8 ahi[NTL], alo[NTL], bhi[NT], blo[NT];
#pragma unroll
    for (int mt = 0; mt < NTL; mt++) {
      const __bf16 *a = ap + (size_t)16 * mt * kp + 32 * ks;
      ahi[mt] = *(const bf16x8 *)a;
      alo[mt] = *(const bf16x8 *)(a + half);
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      // rows past the block meet zero operator columns; they are read at a
      // clamped index so that they stay finite
      const int k = min(32 * ks + 8 * lk + e, s - 1);
      const float *xr = Xi + (base + k) * RS + li;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const float x = xr[16 * t];
        const __bf16 h = (__bf16)x;
        bhi[t][e] = h;
        blo[t][e] = (__bf16)(x - (float)h);
      }
    }
    if (ks == 0) {
#pragma unroll
      for (int mt = 0; mt < NTL; mt++)
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[mt], bhi[t],
                                                               acc_t{0, 0, 0, 0}, 0, 0, 0);
    } else {
#pragma unroll
      for (int mt = 0; mt < NTL; mt++)
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[mt], bhi[t], acc[mt][t],
                                                               0, 0, 0);
    }
#pragma unroll
    for (int mt = 0; mt < NTL; mt++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[mt], blo[t], acc[mt][t], 0, 0, 0);
        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[mt], bhi[t], acc[mt][t], 0, 0, 0);
      }
  }
  // rows of a full tile are stored unguarded (a wave-uniform test)
#pragma unroll
  for (int mt = 0; mt < NTL; mt++) {
    const int r0 = 16 * (t0 + mt);
    float *xo = Xo + (base + r0) * RS + li;
    if (r0 + 16 <= s) {
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int t = 0; t < NT; t++) xo[Mfma<float>::row(lk, e) * RS + 16 * t] = acc[mt][t][e];
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int i = Mfma<float>::row(lk, e);
        if (r0 + i < s) {
#pragma unroll
          for (int t = 0; t < NT; t++) xo[i * RS + 16 * t] = acc[mt][t][e];
        }
      }
    }
  }
}

template <int NT, int MT>
__device__ __forceinline__ void rot_any_bf(int ntl, const float *Xi, float *Xo,
                                           const __bf16 *Ab, size_t half, int s, int base,
                                           int t0) {
  if constexpr (MT > 1) {
    if (ntl < MT) {
      rot_any_bf<NT, MT - 1>(ntl, Xi, Xo, Ab, half, s, base, t0);
      return;
    }
  }
  rot_unit_bf<NT, MT>(Xi, Xo, Ab, half, s, base, t0);
}

// the rotation stage in split-BF16; unit operator offsets index the split
// blocks
template <int NT, int MT>
__device__ __forceinline__ void rot_stage_bf(const float *Xi, float *Xo, const __bf16 *Hb,
                                             size_t half, const int *units, int count) {
  for (int j = 0; j < count; j++) {
    const int code = uni(units[2 * j]), aoff = uni(units[2 * j + 1]);
    const int n = code >> 4, t0 = (code & 15) * MT;
    const int s = 2 * n + 1;
    const int ntl = min(MT, ((s + 15) >> 4) - t0);
    rot_any_bf<NT, MT>(ntl, Xi, Xo, Hb + aoff, half, s, n * n, t0);
  }
}

// Coaxial stage over this wave's units {am*16 + g, operator offset}: the
// blocks +am and -am together.
template <typename T, int NT, int MT, bool BUF = false, bool XG = false>
__device__ __forceinline__ void coax_stage(const T *Xi, T *Xo, const cplx<T> *C,
                                           const int *units, int count, int pin,
                                           int pout) {
  for (int j = 0; j < count; j++) {
    const int code = uni(units[2 * j]), aoff = uni(units[2 * j + 1]);
    const int am = code >> 4, t0 = (code & 15) * MT;
    const int ntl = min(MT, ((pout - am + 15) >> 4) - t0);
    if (am == 0)
      coax_any<T, NT, MT, 1, BUF, XG>(ntl, Xi, Xo, C + aoff, am, pin, pout, t0);
    else
      coax_any<T, NT, MT, 2, BUF, XG>(ntl, Xi, Xo, C + aoff, am, pin, pout, t0);
  }
}

// Value x of lane l (wave-uniform) as a scalar.
__device__ __forceinline__ long long lane64(long long x, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), l);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// One batch: vectors u0 .. u0+V-1 of chunk ch (u0 below the chunk's count).
// pre() runs just before the first barrier (the LDS-mode kernel stores its
// schedule there, so that the schedule's loads and the source gather are in
// flight together).
// OPL: the bucket's operators are Hs / Cs (LDS copies) instead of global.
// B32 (FP32 only): the source and target arrays span under 4 GB, so the
// gather reads them by buffer loads and the scatter adds by buffer atomics,
// one scalar byte offset per vector and one VGPR offset per lane, instead
// of 64-bit pointer arithmetic per access.
template <typename T, int NT, int MT, int W, bool GWS, bool BF = false, bool OPL = false,
          bool B32 = false, typename Pre>
__device__ __forceinline__ void rcr_batch(const RcrArgs &a, int4 ch, int u0,
                                          T *X0, T *X1, const int *sched, Pre pre,
                                          const T *Hs = nullptr,
                                          const cplx<T> *Cs = nullptr) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride, NTH = kThreads<W>;
  const int tid = threadIdx.x, wave = uni(tid >> 6), lane = tid & 63;
  const int pin = a.pin, pout = a.pout;
  const int Pm = pin > pout ? pin : pout;
  const int Pi2 = pin * pin, Po2 = pout * pout, Pm2 = Pm * Pm;
  const OffInfo oi = a.info[ch.x];
  const T *H;
  const cplx<T> *C;
  if constexpr (OPL) {
    H = Hs;
    C = Cs;
  } else {
    H = (const T *)a.H + a.hstride * oi.beta;
    C = (const cplx<T> *)a.C + a.cstride * oi.dist;
  }
  // e^{i m phi}, m < Pm, read through the caches where needed
  const cplx<T> *phb = (const cplx<T> *)a.phase + (size_t)a.pstride * ch.x;
  // vector v of the batch: (pair, right-hand side) = divmod(u0 + v, nrhs),
  // present for v < nv.  Lane v of every wave computes vector v's source and
  // target offsets (complex values) with one load of each index array; the
  // loops read them back as wave-uniform values.  Lanes past nv repeat
  // vector 0, so every source address below is in bounds; the columns of
  // absent vectors are computed but never added to a target (a product's
  // column depends only on the same column of its input).
  const int nv = min(V, (ch.z - ch.y) * a.nrhs - u0);
  long long soff, toff;
  {
    const int u = u0 + (lane < nv ? lane : 0);
    const int qq = u / a.nrhs, r = u - qq * a.nrhs, q = ch.y + qq;
    soff = ((long long)a.psrc[q] * a.nrhs + r) * Pi2;
    toff = ((long long)a.ptgt[q] * a.nrhs + r) * Po2;
  }
  constexpr int CS = (int)sizeof(cplx<T>);
  const cplx<T> *vin[V];
  int sin[V];
  __amdgpu_buffer_rsrc_t rin;
  if constexpr (B32) {
    rin = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.Min), (short)0, -1,
                                            0x00020000);
#pragma unroll
    for (int v = 0; v < V; v++) sin[v] = (int)(unsigned)(lane64(soff, v) * CS);
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) vin[v] = (const cplx<T> *)a.Min + lane64(soff, v);
  }
  // source coefficients times e^{i m phi} (times 1/2 for the even/odd
  // rotations, whose first stage forms z_m + z_{-m} and z_m - z_{-m}); rows
  // f >= pin^2 are zero.  Rows f = tid and f = tid + NTH are loaded (at
  // clamped, in-bounds rows) before the barrier that frees X0 from the
  // previous batch; larger orders loop.
  constexpr T kScale = BF ? T(1) : T(0.5);
  auto row = [&](int f, cplx<T> (&z)[V], cplx<T> &w, int &m) {
    const int g = min(f, Pi2 - 1), h = min(f, Pm2 - 1);
    if constexpr (B32) {
#pragma unroll
      for (int v = 0; v < V; v++) z[v] = OpLoad<cplx<T>>::ld(rin, g * CS, sin[v]);
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) z[v] = vin[v][(unsigned)g];
    }
    const int n = isqrt_fix(h);
    m = h - n * n - n;
    w = phb[m < 0 ? -m : m];
  };
  auto put = [&](int f, const cplx<T> (&z)[V], cplx<T> w, int m) {
    w = cmk<T>(kScale * w.re, m < 0 ? -kScale * w.im : kScale * w.im);
    if (f >= Pi2) w = cmk<T>(0, 0);
#pragma unroll
    for (int v = 0; v < V; v++) *(cplx<T> *)(X0 + f * RS + 2 * v) = cmul(z[v], w);
  };
  const bool two = Pm2 > NTH;  // wave-uniform
  cplx<T> z0[V], z1[V], w0, w1;
  int m0, m1;
  row(tid, z0, w0, m0);
  if (two) row(tid + NTH, z1, w1, m1);
  pre();
  __syncthreads();
  if (tid < Pm2) put(tid, z0, w0, m0);
  if (two && tid + NTH < Pm2) put(tid + NTH, z1, w1, m1);
  for (int f = tid + 2 * NTH; f < Pm2; f += NTH) {
    row(f, z0, w0, m0);
    put(f, z0, w0, m0);
  }
  __syncthreads();
  const int *o1 = sched, *oc = sched + W + 1, *o3 = sched + 2 * (W + 1);
  const int *u1 = sched + 2 * uni(o1[wave]), w1c = uni(o1[wave + 1] - o1[wave]);
  const int *uc = sched + 2 * uni(oc[wave]), wcc = uni(oc[wave + 1] - oc[wave]);
  const int *u3 = sched + 2 * uni(o3[wave]), w3c = uni(o3[wave + 1] - o3[wave]);
  if constexpr (BF) {
    const __bf16 *Hb = (const __bf16 *)a.Hb + a.hbstride * oi.beta;
    rot_stage_bf<NT, MT>(X0, X1, Hb, a.hbhalf, u1, w1c);
    __syncthreads();
    coax_stage<T, NT, MT, !OPL, GWS>(X1, X0, C, uc, wcc, pin, pout);
    __syncthreads();
    rot_stage_bf<NT, MT>(X0, X1, Hb, a.hbhalf, u3, w3c);
  } else {
    rot_stage<T, NT, MT, 0, !OPL, GWS>(X0, X1, H, u1, w1c);
    __syncthreads();
    coax_stage<T, NT, MT, !OPL, GWS>(X1, X0, C, uc, wcc, pin, pout);
    __syncthreads();
    rot_stage<T, NT, MT, 1, !OPL, GWS>(X0, X1, H, u3, w3c);
  }
  __syncthreads();
  // add e^{-i m' phi} y into the targets.  Thread e adds real e of each
  // target vector (part e & 1 of coefficient e >> 1), so that one wave's
  // atomic adds fill whole cache lines: the adds are performed beyond the
  // L2, and half-filled lines cost nearly as much as full ones.
  T *vout[V];
  int sout[V];
  __amdgpu_buffer_rsrc_t rout;
  if constexpr (B32) {
    rout = __builtin_amdgcn_make_buffer_rsrc(a.Lout, (short)0, -1, 0x00020000);
#pragma unroll
    for (int v = 0; v < V; v++) sout[v] = (int)(unsigned)(lane64(toff, v) * CS);
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) vout[v] = (T *)((cplx<T> *)a.Lout + lane64(toff, v));
  }
  for (int e = tid; e < 2 * Po2; e += NTH) {
    const int f = e >> 1, n = isqrt_fix(f), m = f - n * n - n;
    const bool im = e & 1;
    cplx<T> w = phb[m < 0 ? -m : m];
    if (m > 0) w.im = -w.im;
#pragma unroll
    for (int v = 0; v < V; v++)
      if (v < nv) {
        const cplx<T> x = *(const cplx<T> *)(X1 + f * RS + 2 * v);
        const T y = im ? x.re * w.im + x.im * w.re : x.re * w.re - x.im * w.im;
        if constexpr (B32)
          __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(y, rout, e * (int)sizeof(T), sout[v], 0);
        else
          unsafeAtomicAdd(vout[v] + (unsigned)e, y);
      }
  }
}

// Waves per SIMD asked of the compiler: FP32 batches fit three workgroups'
// LDS on a CU (3 waves per SIMD at 4 waves per workgroup; two 8-wave
// workgroups give 4); FP64 keeps its accumulators unspilled instead.
template <typename T, int W> constexpr int kWavesPerEu = sizeof(T) == 4 ? (W == 8 ? 4 : 3) : 1;

// LDS mode: grid (nchunks, nb), one batch per workgroup.
template <typename T, int NT, int MT, int W, bool BF = false, bool B32 = false>
__global__ void __launch_bounds__(kThreads<W>)
    __attribute__((amdgpu_waves_per_eu(kWavesPerEu<T, W>))) k_rcr_lds(RcrArgs a) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  int cx = blockIdx.x, by = blockIdx.y;
  if (a.xcd) {
    // workgroups are dealt to the 8 XCDs round-robin: workgroup g runs on
    // XCD g % 8, which takes the contiguous item range x * per .. x * per +
    // per - 1, so that neighbouring chunks share the XCD's L2
    // (the launch keeps the grid, so the item count, below 2^31)
    const unsigned items = (unsigned)a.nchunks * (unsigned)a.nb;
    const unsigned per = (items + 7) / 8;
    const unsigned it = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (it >= items || (blockIdx.x >> 3) >= per) return;
    cx = (int)(it / (unsigned)a.nb);
    by = (int)(it % (unsigned)a.nb);
  }
  const int4 ch = a.chunks[cx];
  const int u0 = by * V;
  if (u0 >= (ch.z - ch.y) * a.nrhs) return;
  T *X = (T *)smem;
  int *sched = (int *)(X + (size_t)Pm * Pm * RS);
  // the schedule's first 2 NTH entries are loaded here and stored by
  // rcr_batch just before its first barrier, with the gather in flight
  constexpr int NTH = kThreads<W>;
  const int t = threadIdx.x, n = a.sched_len;
  const int s0 = t < n ? a.sched[t] : 0, s1 = t + NTH < n ? a.sched[t + NTH] : 0;
  rcr_batch<T, NT, MT, W, false, BF, false, B32>(a, ch, u0, X, X, sched, [&] {
    if (t < n) sched[t] = s0;
    if (t + NTH < n) sched[t + NTH] = s1;
    for (int i = t + 2 * NTH; i < n; i += NTH) sched[i] = a.sched[i];
  });
}

// LDS mode on a persistent grid (one workgroup per free slot of the GPU):
// each workgroup loads the schedule once and loops over the items chunk * nb
// + batch, so workgroups are not dispatched (and set up) per batch.
template <typename T, int NT, int MT, int W, bool BF = false>
__global__ void __launch_bounds__(kThreads<W>)
    __attribute__((amdgpu_waves_per_eu(kWavesPerEu<T, W>))) k_rcr_ldsp(RcrArgs a) {
  constexpr int V = 8 * NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  T *X = (T *)smem;
  int *sched = (int *)(X + (size_t)Pm * Pm * Row<NT>::stride);
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads<W>) sched[i] = a.sched[i];
  const long long items = (long long)a.nchunks * a.nb;
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    rcr_batch<T, NT, MT, W, false, BF>(a, ch, b * V, X, X, sched, [] {});
  }
}

// LDSO mode: persistent grid; workgroup g takes items [g I / G, (g+1) I / G)
// of the I = nchunks * nb items (chunk-major, so the range covers one or two
// buckets) and keeps the current bucket's padded rotation blocks (n < P) and
// coaxial blocks in LDS.  Two waves per SIMD at most: the LDS holds one
// workgroup per CU.
template <typename T, int W> constexpr int kWavesPerEuO = W == 8 ? 2 : 1;

__device__ __forceinline__ size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

template <typename T, int NT, int MT, int W>
__global__ void __launch_bounds__(kThreads<W>)
    __attribute__((amdgpu_waves_per_eu(kWavesPerEuO<T, W>))) k_rcr_ldso(RcrArgs a) {
  constexpr int V = 8 * NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  T *X = (T *)smem;
  int *sched = (int *)(X + (size_t)Pm * Pm * Row<NT>::stride);
  T *Hs = (T *)(smem + align16((size_t)((unsigned char *)(sched + a.sched_len) - smem)));
  cplx<T> *Cs = (cplx<T> *)(Hs + a.hlen + (a.hlen & 1));
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads<W>) sched[i] = a.sched[i];
  const long long items = (long long)a.nchunks * a.nb;
  const long long i0 = blockIdx.x * items / gridDim.x;
  const long long i1 = (blockIdx.x + 1) * items / gridDim.x;
  int beta = -1, dist = -1;
  const int clen = (int)a.cstride;
  for (long long it = i0; it < i1; it++) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    const OffInfo oi = a.info[ch.x];
    if (oi.beta != beta || oi.dist != dist) {
      // every wave is past the previous batch's last operator read
      __syncthreads();
      if (oi.beta != beta) {
        const T *src = (const T *)a.H + a.hstride * oi.beta;
        for (int i = threadIdx.x; i < a.hlen; i += kThreads<W>) Hs[i] = src[i];
      }
      if (oi.dist != dist) {
        const cplx<T> *src = (const cplx<T> *)a.C + a.cstride * oi.dist;
        for (int i = threadIdx.x; i < clen; i += kThreads<W>) Cs[i] = src[i];
      }
      beta = oi.beta;
      dist = oi.dist;
      // rcr_batch opens with a barrier, which publishes the copies
    }
    rcr_batch<T, NT, MT, W, false, false, true>(a, ch, b * V, X, X, sched, [] {}, Hs, Cs);
  }
}

// GWS mode: persistent grid over items chunk * nb + batch.
template <typename T, int NT, int MT, int W>
__global__ void __launch_bounds__(kThreads<W>) k_rcr_gws(RcrArgs a) {
  constexpr int V = 8 * NT, RS = Row<NT>::stride;
  extern __shared__ __align__(16) unsigned char smem[];
  const int Pm = a.pin > a.pout ? a.pin : a.pout;
  const size_t slice = (size_t)Pm * Pm * RS;
  T *X0 = (T *)a.gws + 2 * slice * blockIdx.x;
  T *X1 = X0 + slice;
  int *sched = (int *)smem;
  for (int i = threadIdx.x; i < a.sched_len; i += kThreads<W>) sched[i] = a.sched[i];
  const long long items = (long long)a.nchunks * a.nb;
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = (int)(it / a.nb), b = (int)(it % a.nb);
    const int4 ch = a.chunks[c];
    if (b * V >= (ch.z - ch.y) * a.nrhs) continue;
    rcr_batch<T, NT, MT, W, true>(a, ch, b * V, X0, X1, sched, [] {});
  }
}

// row tiles of the largest whole block (an even rotation block or a coaxial
// block: at most pm rows; a split-BF16 standard rotation block: 2 pm - 1),
// or 0 beyond the LDS-mode units
inline int mt_lds(int pm, int bf = 0) {
  int tiles = ((bf ? 2 * pm - 1 : pm) + 15) / 16;
  if (tiles <= 2) return 2;
  if (tiles <= 3) return 3;
  if (tiles <= 4) return 4;
  if (tiles <= 6) return 6;
  return 0;
}

// the schedule: sched_len ints
inline size_t small_lds(int pm, size_t tsz, int sched_len) {
  (void)pm;
  (void)tsz;
  return sizeof(int) * (size_t)sched_len;
}


inline size_t x_bytes(int pm, int nt, size_t tsz) {
  return (size_t)pm * pm * (16 * nt + kRowPad) * tsz;
}

// LDSO mode: batch, schedule, then (16-byte aligned) hlen rotation
// values and clen complex coaxial values
inline size_t ldso_bytes(int pm, int nt, size_t tsz, int sched_len, int hlen, size_t clen) {
  size_t b = (x_bytes(pm, nt, tsz) + small_lds(pm, tsz, sched_len) + 15) & ~(size_t)15;
  return b + (size_t)(hlen + (hlen & 1)) * tsz + clen * 2 * tsz;
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

// BF: split-BF16 rotations (FP32 batches in LDS mode only)
template <typename T, int NT, int MT, int W, bool BF = false>
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
  a.Hb = j->Hb;
  a.hbstride = j->hbstride;
  a.hbhalf = j->hbhalf;
  a.P = j->P;
  a.pin = j->pin;
  a.pout = j->pout;
  a.sched = j->sched;
  a.sched_len = j->sched_len;
  a.hlen = j->hlen;
  a.xcd = j->xcd;
  a.Min = j->in;
  a.Lout = j->out;
  a.nrhs = nrhs;
  a.gws = j->gws;
  const size_t tsz = sizeof(T);
  if (j->mode == FMBH_RCR_LDS) {
    const size_t lds = x_bytes(Pm, NT, tsz) + small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_lds<T, NT, MT, W, BF>;
    // FP32 arrays under 4 GB: the buffer-addressed gather and scatter
    if constexpr (sizeof(T) == 4 && !BF)
      if (j->in_bytes < (1ull << 32) && j->out_bytes < (1ull << 32))
        kern = k_rcr_lds<T, NT, MT, W, BF, true>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    if (a.xcd) {
      const long long items = (long long)a.nchunks * a.nb, grid = 8 * ((items + 7) / 8);
      if (grid > 0x7fffffffLL) return fmb_set_error(FMB_ERR_RANGE, "translation grid too large");
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kThreads<W>), lds, st, a);
    } else {
      hipLaunchKernelGGL(kern, dim3(a.nchunks, a.nb), dim3(kThreads<W>), lds, st, a);
    }
  } else if (j->mode == FMBH_RCR_LDSP) {
    const size_t lds = x_bytes(Pm, NT, tsz) + small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_ldsp<T, NT, MT, W, BF>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    int per_cu = 0, ncu = 0, dev = 0;
    FMB_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads<W>, lds));
    FMB_HIP_CHECK(hipGetDevice(&dev));
    FMB_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu < 1) per_cu = 1;
    long long items = (long long)a.nchunks * a.nb, cap = (long long)per_cu * ncu;
    int grid = (int)(items < cap ? items : cap);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads<W>), lds, st, a);
  } else if (j->mode == FMBH_RCR_LDSO) {
    if constexpr (BF) {
      return fmb_set_error(FMB_ERR_RANGE, "split-BF16 rotations have no LDSO tile shape");
    } else {
      const size_t lds = ldso_bytes(Pm, NT, tsz, j->sched_len, j->hlen, j->cstride);
      auto kern = k_rcr_ldso<T, NT, MT, W>;
      int rc = set_lds(kern, lds);
      if (rc) return rc;
      int per_cu = 0, ncu = 0, dev = 0;
      FMB_HIP_CHECK(
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads<W>, lds));
      FMB_HIP_CHECK(hipGetDevice(&dev));
      FMB_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
      if (per_cu < 1) per_cu = 1;
      long long items = (long long)a.nchunks * a.nb, cap = (long long)per_cu * ncu;
      int grid = (int)(items < cap ? items : cap);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads<W>), lds, st, a);
    }
  } else if constexpr (BF) {
    return fmb_set_error(FMB_ERR_RANGE, "split-BF16 translations need an LDS tile shape");
  } else {
    const size_t lds = small_lds(Pm, tsz, j->sched_len);
    auto kern = k_rcr_gws<T, NT, MT, W>;
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    long long items = (long long)a.nchunks * a.nb;
    int grid = items < j->gws_grid ? (int)items : j->gws_grid;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads<W>), lds, st, a);
  }
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

// the tile shapes (nt, mt) of one precision and wave count
template <typename T, int W, bool BF = false>
int dispatch_shape(const fmbh_rcr_job *j, int nrhs, hipStream_t st) {
#define CASE(A, B) \
  if (j->nt == A && j->mt == B) return launch<T, A, B, W, BF>(j, nrhs, st);
  CASE(1, 2) CASE(1, 3) CASE(1, 4) CASE(1, 6)
  CASE(2, 2) CASE(2, 3) CASE(2, 4) CASE(2, 6)
#undef CASE
  return fmb_set_error(FMB_ERR_RANGE, "no RCR tile shape nt=%d mt=%d", j->nt, j->mt);
}

}  // namespace
