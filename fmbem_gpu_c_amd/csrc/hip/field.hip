// field.hip -- pressure at field points off the surface from a surface
// solution, by the exterior representation formula
//   p(y) = p_inc(y) + sum_j x_j int_j (dG/dn_y - sigma_j G) dS_y
// with the same kernels, signs and panel rules as the collocation entries
// (near.hip): the refined rule when |y - c_j| < ratio * radius_j.  The
// reference has no field evaluation.  One workgroup per field point reduces
// over all panels in FP64 (for few points, over one slice of them, the
// slices summed in order by k_field_sum); R right-hand sides per pass.
//
// Its transpose, for adjoint objectives: given weights u[r][q] at the field
// points, per element j (caller order)
//   kt[r][j] = sum_q u[r][q] int_j (dG/dn_y - sigma_j G)(y_q, y) dS_y
//   st[r][j] = sum_q u[r][q] int_j G(y_q, y) dS_y
// one thread per element, R right-hand sides per pass.
//
// Through the tree (k_field_box, k_field_tree): the multipole expansions M
// of the upward pass stand in for the panels of every box well separated
// from the point (a treecode with the separation of the M2L lists), the
// panels next to the point are summed directly with the rule above.  The
// coarse levels, whose orders are high, run box-major over chunks of points
// in Morton order; the finer levels per point.
#include "fmb_device.h"

namespace {

constexpr int kT = 256, kR = 8;
constexpr int kFieldSlicePts = 1024; // below this many points k_field slices the panels

__global__ void __launch_bounds__(kT)
    k_field(const double *__restrict__ pts, int npts, const double *__restrict__ tri,
            const double *__restrict__ geo, const int *__restrict__ perm, int N,
            double k, const double2 *__restrict__ robin,
            const double *__restrict__ quad, int nq, const double *__restrict__ quadn,
            int nqn, double ratio, const double2 *__restrict__ x, int nrhs, int r0,
            const double *__restrict__ dirs, double2 *__restrict__ out,
            double2 *__restrict__ part) {
  __shared__ double2 red[kT / 64][kR];
  const int q = blockIdx.x, nsl = gridDim.y;
  const double y[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  const double nz[3] = {0.0, 0.0, 0.0};
  const int nr = min(kR, nrhs - r0);
  double2 acc[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) acc[r] = make_double2(0.0, 0.0);
  for (int j = blockIdx.y * kT + threadIdx.x; j < N; j += nsl * kT) {
    const double *gj = geo + 8 * (size_t)j;
    const double dx = y[0] - gj[0], dy = y[1] - gj[1], dz = y[2] - gj[2];
    const bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
    double2 o[4];
    panel_sum(k, y, nz, tri + 9 * (size_t)j, gj + 3, gj[6], close ? quadn : quad,
              close ? nqn : nq, o);
    double2 g = o[1];
    if (robin) {
      const double2 s = robin[j];
      g.x -= s.x * o[0].x - s.y * o[0].y;
      g.y -= s.x * o[0].y + s.y * o[0].x;
    }
    const int c = perm[j];
#pragma unroll
    for (int r = 0; r < kR; r++) {
      if (r < nr) {
        const double2 v = x[(size_t)(r0 + r) * N + c];
        acc[r].x += g.x * v.x - g.y * v.y;
        acc[r].y += g.x * v.y + g.y * v.x;
      }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kR; r++) {
    double re = acc[r].x, im = acc[r].y;
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) red[w][r] = make_double2(re, im);
  }
  __syncthreads();
  if (threadIdx.x < nr) {
    const int r = threadIdx.x;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kT / 64; i++) {
      s.x += red[i][r].x;
      s.y += red[i][r].y;
    }
    if (part) {
      part[((size_t)blockIdx.y * kR + r) * npts + q] = s;
      return;
    }
    if (dirs) {
      const double *d = dirs + 3 * (r0 + r);
      double sn, cs;
      sincos(k * (d[0] * y[0] + d[1] * y[1] + d[2] * y[2]), &sn, &cs);
      s.x += cs;
      s.y += sn;
    }
    out[(size_t)(r0 + r) * npts + q] = s;
  }
}

// The panel slices of k_field summed in slice order, plus the incident
// waves: one thread per (point, right-hand side).
__global__ void __launch_bounds__(kT)
    k_field_sum(const double2 *__restrict__ part, int nsl, const double *__restrict__ pts,
                int npts, double k, int nrhs, int r0, const double *__restrict__ dirs,
                double2 *__restrict__ out) {
  const int i = blockIdx.x * kT + threadIdx.x;
  const int nr = min(kR, nrhs - r0);
  if (i >= npts * nr) return;
  const int r = i / npts, q = i % npts;
  double2 s = make_double2(0.0, 0.0);
  for (int sl = 0; sl < nsl; sl++) {
    const double2 v = part[((size_t)sl * kR + r) * npts + q];
    s.x += v.x;
    s.y += v.y;
  }
  if (dirs) {
    const double *d = dirs + 3 * (r0 + r);
    double sn, cs;
    sincos(k * (d[0] * pts[3 * q] + d[1] * pts[3 * q + 1] + d[2] * pts[3 * q + 2]), &sn, &cs);
    s.x += cs;
    s.y += sn;
  }
  out[(size_t)(r0 + r) * npts + q] = s;
}

__global__ void __launch_bounds__(kT)
    k_field_t(const double *__restrict__ pts, int npts, const double *__restrict__ tri,
              const double *__restrict__ geo, const int *__restrict__ perm, int N,
              double k, const double2 *__restrict__ robin,
              const double *__restrict__ quad, int nq, const double *__restrict__ quadn,
              int nqn, double ratio, const double2 *__restrict__ u, int nrhs, int r0,
              double2 *__restrict__ kt, double2 *__restrict__ st) {
  const int j = blockIdx.x * kT + threadIdx.x;
  if (j >= N) return;
  const int nr = min(kR, nrhs - r0);
  const double *gj = geo + 8 * (size_t)j;
  const double nz[3] = {0.0, 0.0, 0.0};
  const double2 sg = robin ? robin[j] : make_double2(0.0, 0.0);
  double2 ak[kR], as[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) ak[r] = as[r] = make_double2(0.0, 0.0);
  for (int q = 0; q < npts; q++) {
    const double y[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
    const double dx = y[0] - gj[0], dy = y[1] - gj[1], dz = y[2] - gj[2];
    const bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
    double2 o[4];
    panel_sum(k, y, nz, tri + 9 * (size_t)j, gj + 3, gj[6], close ? quadn : quad,
              close ? nqn : nq, o);
    const double2 g = make_double2(o[1].x - (sg.x * o[0].x - sg.y * o[0].y),
                                   o[1].y - (sg.x * o[0].y + sg.y * o[0].x));
#pragma unroll
    for (int r = 0; r < kR; r++) {
      if (r < nr) {
        const double2 w = u[(size_t)(r0 + r) * npts + q];
        ak[r].x += g.x * w.x - g.y * w.y;
        ak[r].y += g.x * w.y + g.y * w.x;
        as[r].x += o[0].x * w.x - o[0].y * w.y;
        as[r].y += o[0].x * w.y + o[0].y * w.x;
      }
    }
  }
  const int c = perm[j];
#pragma unroll
  for (int r = 0; r < kR; r++) {
    if (r < nr) {
      kt[(size_t)(r0 + r) * N + c] = ak[r];
      st[(size_t)(r0 + r) * N + c] = as[r];
    }
  }
}

// ---------------------------------------------------------------- treecode

constexpr int kCand = 256; // candidates of one level: <= 64 at level 2, <= 27*8 below
constexpr int kFieldChunk = 32768; // points per pass of the box-major sums
constexpr int kBoxMinP = 24;        // box-major from level 2 down while orders reach this
constexpr int kBoxMaxBoxes = 32768; // and the level has at most this many boxes
constexpr int kBoxWG = 64;          // points per box-major workgroup

struct TreeLevels {
  int L;
  int p[FMB_MAXLEV], nbox[FMB_MAXLEV];
  const int *keys[FMB_MAXLEV];
  const int *cstart[FMB_MAXLEV];
  const double *scale[FMB_MAXLEV];
  const double2 *M[FMB_MAXLEV];
  double D, pmin[3];
};

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Pressure at field point blockIdx.x through the tree, right-hand sides
// r0 .. r0+kR-1 of the nrhs the upward pass ran on.  Per level l, wave 0
// sorts the level's candidate boxes: a box whose level-l index differs by
// 2 or more in some coordinate from the index of the level-l box holding
// the point adds its expansion, -sum_{n<p} w_n Mt_n^m S_n^m(y - c) (Mt the
// stored coefficients M / w_n; the P2M basis expands -dG/dn + sigma G, the
// negative of the field kernel); a box next to it passes its children on;
// at the leaf level its panels are summed directly.  Levels up to lb were
// summed box-major by k_field_box into boxsum, read back here; below lb
// groups of 16, 32 or 64 lanes (by the level's order) take one box each,
// and lane m of a group runs column m of S_n^m up in n: the normalised
// Legendre recurrence P_n^m = a (cos(theta) P_{n-1}^m - b P_{n-2}^m) from
// the table leg (a, b per (n, m)) and dg[m] sin(theta)^m, and h_n by the
// upward recurrence h_{n+1} = (2n+1)/x h_n - h_{n-1}, dominated by y_n
// wherever the two recurrences part.  The workgroup's sums are reduced in a
// fixed order; the result goes to the caller's index qperm[q].
__global__ void __launch_bounds__(kT)
    k_field_tree(TreeLevels tl, const double2 *__restrict__ leg,
                 const double *__restrict__ dg, const double *__restrict__ pts,
                 const double *__restrict__ tri, const double *__restrict__ geo,
                 const int *__restrict__ perm, const int *__restrict__ bstart, int N,
                 double k, const double2 *__restrict__ robin,
                 const double *__restrict__ quad, int nq, const double *__restrict__ quadn,
                 int nqn, double ratio, const double2 *__restrict__ x, int nrhs, int r0,
                 const double *__restrict__ dirs, int npts, double2 *__restrict__ out,
                 int ostride, const int *__restrict__ qperm, int lb,
                 const double2 *__restrict__ boxsum) {
  __shared__ int cand[2][kCand];
  __shared__ int farl[kCand];
  __shared__ int nearl[32];
  __shared__ int cnt[3];
  __shared__ double2 red[kT / 64][kR];
  const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double y[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  double u[3];
  for (int a = 0; a < 3; a++) u[a] = (y[a] - tl.pmin[a]) / tl.D;
  const int nr = min(kR, nrhs - r0);
  double2 acc[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) acc[r] = make_double2(0.0, 0.0);
  const int L = tl.L;
  const int nb2 = min(tl.nbox[2], 64);
  if (tid == 0) cnt[0] = nb2;
  if (tid < nb2) cand[0][tid] = tid;
  __syncthreads();
  int cur = 0;
  for (int l = 2; l <= L; l++) {
    const int nc = cnt[0];
    if (nc == 0) break;
    const double sc = ldexp(1.0, l);
    const double ty[3] = {floor(u[0] * sc), floor(u[1] * sc), floor(u[2] * sc)};
    __syncthreads();  // cnt[0] read by every thread before wave 0 rewrites it
    if (w == 0) {
      int nfar = 0, nnext = 0, nnear = 0;
      const unsigned long long lt = (1ull << lane) - 1ull;
      for (int i0 = 0; i0 < nc; i0 += 64) {
        const int i = i0 + lane;
        const bool valid = i < nc;
        const int b = valid ? cand[cur][i] : 0;
        bool far = false;
        if (valid) {
          int tx, tyy, tz;
          deinterleave3(tl.keys[l][b], l, tx, tyy, tz);
          const double sep = fmax(fabs(tx - ty[0]), fmax(fabs(tyy - ty[1]), fabs(tz - ty[2])));
          far = sep >= 2.0;
        }
        const bool nb = valid && !far;
        const unsigned long long fm = __ballot(far);
        if (far) {
          const int o = nfar + __popcll(fm & lt);
          if (o < kCand) farl[o] = b;
        }
        nfar += __popcll(fm);
        if (l < L) {
          const int c0 = nb ? tl.cstart[l][b] : 0, c1 = nb ? tl.cstart[l][b + 1] : 0;
          const int c = c1 - c0;
          int s = c;
          for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(s, o, 64);
            if (lane >= o) s += v;
          }
          const int base = nnext + s - c;
          for (int j = 0; j < c; j++)
            if (base + j < kCand) cand[cur ^ 1][base + j] = c0 + j;
          nnext += __shfl(s, 63, 64);
        } else {
          const unsigned long long nm = __ballot(nb);
          if (nb) {
            const int o = nnear + __popcll(nm & lt);
            if (o < 32) nearl[o] = b;
          }
          nnear += __popcll(nm);
        }
      }
      if (lane == 0) {
        cnt[0] = min(nnext, kCand);
        cnt[1] = min(nfar, kCand);
        cnt[2] = min(nnear, 32);
      }
    }
    __syncthreads();
    // expansions of the well-separated boxes (levels up to lb: summed
    // box-major by k_field_box into boxsum)
    const int p = tl.p[l], ps = p * p;
    const int G = p <= 16 ? 16 : (p <= 32 ? 32 : 64);
    const int gpw = 64 / G, ng = (kT / 64) * gpw;
    const int g = w * gpw + lane / G, lm = lane % G;
    const double *wsc = tl.scale[l];
    const double h = tl.D / sc;
    for (int i = g; l > lb && i < cnt[1]; i += ng) {
      const int b = farl[i];
      int tx, tyy, tz;
      deinterleave3(tl.keys[l][b], l, tx, tyy, tz);
      const double v0 = y[0] - (tl.pmin[0] + (tx + 0.5) * h);
      const double v1 = y[1] - (tl.pmin[1] + (tyy + 0.5) * h);
      const double v2 = y[2] - (tl.pmin[2] + (tz + 0.5) * h);
      const double rho = sqrt(v0 * v0 + v1 * v1), r = sqrt(rho * rho + v2 * v2);
      const double ct = v2 / r, st = rho / r, phi = atan2(v1, v0);
      const double kr = k * r, ikr = 1.0 / kr;
      double sx, cx;
      sincos(kr, &sx, &cx);
      // h_0 = -i e^{ix}/x, h_1 = -e^{ix}(x + i)/x^2
      const double2 h0 = make_double2(sx * ikr, -cx * ikr);
      const double2 h1 = make_double2((sx - cx * kr) * ikr * ikr, -(sx * kr + cx) * ikr * ikr);
      const double2 *Mb = tl.M[l] + ((size_t)b * nrhs + r0) * ps;
      for (int m = lm; m < p; m += G) {
        double stm = 1.0, base = st;
        for (int e = m; e > 0; e >>= 1) {
          if (e & 1) stm *= base;
          base *= base;
        }
        double P = dg[m] * stm, Pp = 0.0;
        double sm, cm;
        sincos(m * phi, &sm, &cm);
        const double2 ep = make_double2(cm, sm), em = make_double2(cm, -sm);
        double2 hc = h0, hp = make_double2(0.0, 0.0);
        for (int n = 0; n < p; n++) {
          if (n >= m) {
            if (n > m) {
              const double2 ab = leg[n * (n + 1) / 2 + m];
              const double Pn = ab.x * (ct * P - ab.y * Pp);
              Pp = P;
              P = Pn;
            }
            const double f = wsc[n] * P;
            const double2 t = make_double2(f * hc.x, f * hc.y);
            const int fp = n * n + n + m, fm = n * n + n - m;
#pragma unroll
            for (int rr = 0; rr < kR; rr++) {
              if (rr < nr) {
                double2 z = cmul2(Mb[(size_t)rr * ps + fp], ep);
                if (m > 0) {
                  const double2 zm = cmul2(Mb[(size_t)rr * ps + fm], em);
                  z.x += zm.x;
                  z.y += zm.y;
                }
                const double2 tz = cmul2(t, z);
                acc[rr].x -= tz.x;
                acc[rr].y -= tz.y;
              }
            }
          }
          if (n == 0) {
            hp = hc;
            hc = h1;
          } else {
            const double c = (2.0 * n + 1.0) * ikr;
            const double2 hn = make_double2(c * hc.x - hp.x, c * hc.y - hp.y);
            hp = hc;
            hc = hn;
          }
        }
      }
    }
    if (l == L) {
      // panels of the leaf boxes next to the point
      const double nz[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < cnt[2]; i++) {
        const int b = nearl[i];
        for (int j = bstart[b] + tid; j < bstart[b + 1]; j += kT) {
          const double *gj = geo + 8 * (size_t)j;
          const double dx = y[0] - gj[0], dy = y[1] - gj[1], dz = y[2] - gj[2];
          const bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
          double2 o[4];
          panel_sum(k, y, nz, tri + 9 * (size_t)j, gj + 3, gj[6], close ? quadn : quad,
                    close ? nqn : nq, o);
          double2 gg = o[1];
          if (robin) {
            const double2 sg = robin[j];
            gg.x -= sg.x * o[0].x - sg.y * o[0].y;
            gg.y -= sg.x * o[0].y + sg.y * o[0].x;
          }
          const int c = perm[j];
#pragma unroll
          for (int rr = 0; rr < kR; rr++) {
            if (rr < nr) {
              const double2 v = x[(size_t)(r0 + rr) * N + c];
              acc[rr].x += gg.x * v.x - gg.y * v.y;
              acc[rr].y += gg.x * v.y + gg.y * v.x;
            }
          }
        }
      }
    }
    cur ^= 1;
    __syncthreads();  // farl, nearl and cnt are rewritten by the next level
  }
  if (tid == 0) {
#pragma unroll
    for (int rr = 0; rr < kR; rr++) {
      if (rr < nr) {
        const double2 v = boxsum[(size_t)rr * npts + q];
        acc[rr].x += v.x;
        acc[rr].y += v.y;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; r++) {
    double re = acc[r].x, im = acc[r].y;
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) red[w][r] = make_double2(re, im);
  }
  __syncthreads();
  if (tid < nr) {
    const int r = tid;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kT / 64; i++) {
      s.x += red[i][r].x;
      s.y += red[i][r].y;
    }
    if (dirs) {
      const double *d = dirs + 3 * (r0 + r);
      double sn, cs;
      sincos(k * (d[0] * y[0] + d[1] * y[1] + d[2] * y[2]), &sn, &cs);
      s.x += cs;
      s.y += sn;
    }
    out[(size_t)(r0 + r) * ostride + qperm[q]] = s;
  }
}

// The expansions of level l box-major: workgroup w evaluates box
// w / nchunk of level l at the WG points of chunk w % nchunk (the points
// come sorted in Morton order, so a chunk is compact), one point per lane,
// for the points the traversal of k_field_tree meets the box at: the box is
// well separated from the point at level l (indices 2 or more apart) and
// its parent is next to the point at level l - 1 (no condition at level 2;
// the parent's ancestors then lie next to the point too).  Column m of the
// box's coefficients (n = m .. p-1, orders +m and -m, every right-hand side)
// is staged in LDS by the whole workgroup and read by every lane at one
// address; per lane the recurrences of k_field_tree run up that column,
// h_n restarted from n = 0.  boxsum[r * npts + q] += -sum w_n Mt_n^m
// S_n^m(y_q - c) for right-hand side r0 + r, by FP64 atomic adds (the order
// of the boxes' contributions is not fixed).  Dynamic LDS: 2 p complex
// values per right-hand side of the launch.
template <int WG>
__global__ void __launch_bounds__(WG)
    k_field_box(TreeLevels tl, int l, const double2 *__restrict__ leg,
                const double *__restrict__ dg, const double *__restrict__ pts, int npts,
                int nchunk, double k, int nrhs, int r0, double2 *__restrict__ boxsum) {
  extern __shared__ double2 col[];  // [rr][sign][n - m]
  const int b = blockIdx.x / nchunk, tid = threadIdx.x;
  const int q = (blockIdx.x % nchunk) * WG + tid;
  int tx, tyy, tz;
  deinterleave3(tl.keys[l][b], l, tx, tyy, tz);
  const double sc = ldexp(1.0, l);
  double y[3] = {0.0, 0.0, 0.0};
  bool active = false;
  if (q < npts) {
    for (int a = 0; a < 3; a++) y[a] = pts[3 * q + a];
    const int t3[3] = {tx, tyy, tz};
    double sep = 0.0, psep = 0.0;
    for (int a = 0; a < 3; a++) {
      const double u = (y[a] - tl.pmin[a]) / tl.D;
      sep = fmax(sep, fabs(t3[a] - floor(u * sc)));
      psep = fmax(psep, fabs((t3[a] >> 1) - floor(u * (0.5 * sc))));
    }
    active = sep >= 2.0 && (l == 2 || psep <= 1.0);
  }
  if (!__syncthreads_or(active)) return;
  const double h = tl.D / sc;
  const double v0 = y[0] - (tl.pmin[0] + (tx + 0.5) * h);
  const double v1 = y[1] - (tl.pmin[1] + (tyy + 0.5) * h);
  const double v2 = y[2] - (tl.pmin[2] + (tz + 0.5) * h);
  const double rho = sqrt(v0 * v0 + v1 * v1), r = sqrt(rho * rho + v2 * v2);
  // inactive lanes run the same steps on a harmless point and add nothing
  const double ct = active ? v2 / r : 1.0, st = active ? rho / r : 0.0;
  const double phi = active ? atan2(v1, v0) : 0.0;
  const double kr = active ? k * r : 1.0, ikr = 1.0 / kr;
  double sx, cx;
  sincos(kr, &sx, &cx);
  const double2 h0 = make_double2(sx * ikr, -cx * ikr);
  const double2 h1 = make_double2((sx - cx * kr) * ikr * ikr, -(sx * kr + cx) * ikr * ikr);
  const int p = tl.p[l], ps = p * p;
  const int nr = min(kR, nrhs - r0);
  const double *wsc = tl.scale[l];
  const double2 *Mb = tl.M[l] + ((size_t)b * nrhs + r0) * ps;
  double2 acc[kR];
#pragma unroll
  for (int rr = 0; rr < kR; rr++) acc[rr] = make_double2(0.0, 0.0);
  double stm = 1.0;
  for (int m = 0; m < p; m++) {
    const int len = p - m;
    __syncthreads();  // the previous column is read
    for (int i = tid; i < 2 * nr * len; i += WG) {
      const int j = i % len, sg = (i / len) & 1, rr = i / (2 * len);
      const int n = m + j;
      col[(rr * 2 + sg) * p + j] = Mb[(size_t)rr * ps + n * n + n + (sg ? -m : m)];
    }
    __syncthreads();
    double P = dg[m] * stm, Pp = 0.0;
    stm *= st;
    double sm, cm;
    sincos(m * phi, &sm, &cm);
    const double2 ep = make_double2(cm, sm), em = make_double2(cm, -sm);
    double2 hc = h0, hp = make_double2(0.0, 0.0);
    for (int n = 0; n < m; n++) {
      if (n == 0) {
        hp = hc;
        hc = h1;
      } else {
        const double c = (2.0 * n + 1.0) * ikr;
        const double2 hn = make_double2(c * hc.x - hp.x, c * hc.y - hp.y);
        hp = hc;
        hc = hn;
      }
    }
    for (int n = m; n < p; n++) {
      if (n > m) {
        const double2 ab = leg[n * (n + 1) / 2 + m];
        const double Pn = ab.x * (ct * P - ab.y * Pp);
        Pp = P;
        P = Pn;
      }
      const double f = wsc[n] * P;
      const double2 t = make_double2(f * hc.x, f * hc.y);
#pragma unroll
      for (int rr = 0; rr < kR; rr++) {
        if (rr < nr) {
          double2 z = cmul2(col[(rr * 2) * p + n - m], ep);
          if (m > 0) {
            const double2 zm = cmul2(col[(rr * 2 + 1) * p + n - m], em);
            z.x += zm.x;
            z.y += zm.y;
          }
          const double2 tz = cmul2(t, z);
          acc[rr].x -= tz.x;
          acc[rr].y -= tz.y;
        }
      }
      if (n == 0) {
        hp = hc;
        hc = h1;
      } else {
        const double c = (2.0 * n + 1.0) * ikr;
        const double2 hn = make_double2(c * hc.x - hp.x, c * hc.y - hp.y);
        hp = hc;
        hc = hn;
      }
    }
  }
  if (active) {
#pragma unroll
    for (int rr = 0; rr < kR; rr++) {
      if (rr < nr) {
        atomicAdd(&boxsum[(size_t)rr * npts + q].x, acc[rr].x);
        atomicAdd(&boxsum[(size_t)rr * npts + q].y, acc[rr].y);
      }
    }
  }
}

}  // namespace

extern "C" int fmbh_field_t(const fmb_plan *pl, const fmb_c64 *u, int nrhs,
                            const double *d_pts, int npts, fmb_c64 *kt, fmb_c64 *st,
                            void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t s = (hipStream_t)stream;
  const int N = t->nelem;
  for (int r0 = 0; r0 < nrhs; r0 += kR) {
    hipLaunchKernelGGL(k_field_t, dim3((N + kT - 1) / kT), dim3(kT), 0, s, d_pts, npts,
                       t->d_tri, t->d_geo, t->d_perm, N, pl->k,
                       (const double2 *)pl->d_robin, pl->d_quad, pl->nq, pl->d_quadn,
                       pl->nqn, pl->prm.near_ratio, (const double2 *)u, nrhs, r0,
                       (double2 *)kt, (double2 *)st);
    FMB_LAUNCH_CHECK();
  }
  return FMB_OK;
}

extern "C" int fmbh_field(const fmb_plan *pl, const fmb_c64 *x, int nrhs,
                          const double *d_pts, int npts, const double *d_dirs,
                          fmb_c64 *out, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  // few points: each point's panels in slices over several workgroups, so
  // that the launch fills the device; the slices are summed in order
  const int nsl = npts >= kFieldSlicePts ? 1 : min(32, (2 * kFieldSlicePts + npts - 1) / npts);
  double2 *d_part = NULL;
  int rc = FMB_OK;
  if (nsl > 1 && (rc = fmbh_malloc((void **)&d_part, sizeof(double2) * (size_t)nsl * kR * npts)))
    return rc;
  for (int r0 = 0; r0 < nrhs && !rc; r0 += kR) {
    hipLaunchKernelGGL(k_field, dim3(npts, nsl), dim3(kT), 0, st, d_pts, npts, t->d_tri,
                       t->d_geo, t->d_perm, t->nelem, pl->k,
                       (const double2 *)pl->d_robin, pl->d_quad, pl->nq, pl->d_quadn,
                       pl->nqn, pl->prm.near_ratio, (const double2 *)x, nrhs, r0,
                       d_dirs, (double2 *)out, d_part);
    if (nsl > 1) {
      const int nr = nrhs - r0 < kR ? nrhs - r0 : kR;
      hipLaunchKernelGGL(k_field_sum, dim3((npts * nr + kT - 1) / kT), dim3(kT), 0, st, d_part,
                         nsl, d_pts, npts, pl->k, nrhs, r0, d_dirs, (double2 *)out);
    }
    if (hipGetLastError() != hipSuccess) rc = fmb_set_error(FMB_ERR_HIP, "k_field launch");
  }
  if (d_part) {
    // the slices are freed only after the launches that read them
    if (hipStreamSynchronize(st) != hipSuccess && !rc)
      rc = fmb_set_error(FMB_ERR_HIP, "k_field");
    fmbh_free(d_part);
  }
  return rc;
}

// Treecode field of one upward pass: x, dirs and out start at the pass's
// first right-hand side; the opset op holds its multipole expansions.
// d_pts holds the points in Morton order, d_qperm[i] the caller's index of
// sorted point i (out is in caller order).
extern "C" int fmbh_field_tree(const fmb_plan *pl, int op, const fmb_c64 *x, int nrhs,
                               const double *d_pts, const int *d_qperm, int npts,
                               const double *d_dirs, fmb_c64 *out, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  const fmb_opset *o = &pl->ops[op];
  if (o->prec != FMB_FP64)
    return fmb_set_error(FMB_ERR_ARG, "field through the tree: FP64 expansions only");
  if (t->lmax < 2 || t->nbox[2] > 64)
    return fmb_set_error(FMB_ERR_ARG, "field through the tree: no level-2 boxes");
  for (int l = 2; l <= t->lmax; l++)
    if (sizeof(double2) * 2 * kR * pl->p[l] > 64 * 1024)
      return fmb_set_error(FMB_ERR_RANGE, "field through the tree: order %d", pl->p[l]);
  TreeLevels tl = {};
  tl.L = t->lmax;
  tl.D = t->D;
  for (int a = 0; a < 3; a++) tl.pmin[a] = t->pmin[a];
  int pmax = 0;
  for (int l = 2; l <= t->lmax; l++) {
    tl.p[l] = pl->p[l];
    tl.nbox[l] = t->nbox[l];
    tl.keys[l] = t->d_keys[l];
    tl.cstart[l] = l < t->lmax ? t->d_cstart[l] : NULL;
    tl.scale[l] = pl->d_scale[l];
    tl.M[l] = (const double2 *)o->d_M[l];
    pmax = pl->p[l] > pmax ? pl->p[l] : pmax;
  }
  // normalised Legendre recurrence: (a, b) per (n, m), P_m^m / sin^m theta
  const int ntri = pmax * (pmax + 1) / 2;
  double *h = (double *)malloc(sizeof(double) * (2 * (size_t)ntri + pmax));
  if (!h) return fmb_set_error(FMB_ERR_NOMEM, "field through the tree: tables");
  for (int n = 0; n < pmax; n++)
    for (int m = 0; m <= n; m++) {
      double a = 0.0, b = 0.0;
      if (n == m + 1) {
        a = sqrt(2.0 * m + 3.0);
      } else if (n > m + 1) {
        a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
        b = sqrt(((double)(n - 1) * (n - 1) - (double)m * m) /
                 (4.0 * (n - 1) * (n - 1) - 1.0));
      }
      h[2 * (n * (n + 1) / 2 + m)] = a;
      h[2 * (n * (n + 1) / 2 + m) + 1] = b;
    }
  double d = 1.0 / sqrt(4.0 * M_PI);
  for (int m = 0; m < pmax; m++) {
    if (m > 0) d *= sqrt((2.0 * m + 1.0) / (2.0 * m));
    h[2 * ntri + m] = d;
  }
  double *d_tab = NULL;
  int rc = fmbh_malloc((void **)&d_tab, sizeof(double) * (2 * (size_t)ntri + pmax));
  if (!rc && hipMemcpy(d_tab, h, sizeof(double) * (2 * (size_t)ntri + pmax),
                       hipMemcpyHostToDevice) != hipSuccess)
    rc = fmb_set_error(FMB_ERR_HIP, "field through the tree: table copy");
  free(h);
  if (rc) return rc;
  // levels 2 .. lb box-major (orders of kBoxMinP and above), the rest per
  // point; the box-major sums of one chunk of points: [kR][chunk]
  int lb = 2;
  while (lb < t->lmax && pl->p[lb + 1] >= kBoxMinP && t->nbox[lb + 1] <= kBoxMaxBoxes) lb++;
  int pb = 0;
  for (int l = 2; l <= lb; l++) pb = pl->p[l] > pb ? pl->p[l] : pb;
  const int chunk = npts < kFieldChunk ? npts : kFieldChunk;
  double2 *d_box = NULL;
  if (!rc) rc = fmbh_malloc((void **)&d_box, sizeof(double2) * (size_t)kR * chunk);
  for (int q0 = 0; q0 < npts && !rc; q0 += chunk) {
    const int nq0 = npts - q0 < chunk ? npts - q0 : chunk;
    const int nchunk = (nq0 + kBoxWG - 1) / kBoxWG;
    for (int r0 = 0; r0 < nrhs && !rc; r0 += kR) {
      const int nr = nrhs - r0 < kR ? nrhs - r0 : kR;
      if (hipMemsetAsync(d_box, 0, sizeof(double2) * (size_t)kR * nq0, st) != hipSuccess) {
        rc = fmb_set_error(FMB_ERR_HIP, "field through the tree: clear");
        break;
      }
      for (int l = 2; l <= lb; l++)
        hipLaunchKernelGGL(k_field_box<kBoxWG>, dim3(nchunk * t->nbox[l]), dim3(kBoxWG),
                           sizeof(double2) * 2 * nr * pb, st, tl, l, (const double2 *)d_tab,
                           d_tab + 2 * ntri, d_pts + 3 * (size_t)q0, nq0, nchunk, pl->k,
                           nrhs, r0, d_box);
      hipLaunchKernelGGL(k_field_tree, dim3(nq0), dim3(kT), 0, st, tl,
                         (const double2 *)d_tab, d_tab + 2 * ntri, d_pts + 3 * (size_t)q0,
                         t->d_tri, t->d_geo, t->d_perm, t->d_box_start, t->nelem, pl->k,
                         (const double2 *)pl->d_robin, pl->d_quad, pl->nq, pl->d_quadn,
                         pl->nqn, pl->prm.near_ratio, (const double2 *)x, nrhs, r0,
                         d_dirs, nq0, (double2 *)out, npts, d_qperm + q0, lb, d_box);
      if (hipGetLastError() != hipSuccess)
        rc = fmb_set_error(FMB_ERR_HIP, "k_field_tree launch");
    }
  }
  // the tables are freed only after the launches that read them
  if (hipStreamSynchronize(st) != hipSuccess && !rc)
    rc = fmb_set_error(FMB_ERR_HIP, "k_field_tree");
  fmbh_free(d_tab);
  fmbh_free(d_box);
  return rc;
}
