/*
 * fmb_device.h -- device-side helpers shared by the gfx950 kernels:
 * complex arithmetic on (re, im) pairs of the compute type, launch checks,
 * FP64 special functions (spherical Bessel, normalised Legendre, R_n^m and
 * its normal derivative) used by operator generation and the bases.
 */
#ifndef FMB_DEVICE_H
#define FMB_DEVICE_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "fmb_hip.h"

#define FMB_HIP_CHECK(call)                                                    \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fmb_set_error(FMB_ERR_HIP, "%s:%d: %s", __FILE__, __LINE__,       \
                           hipGetErrorString(e_));                             \
  } while (0)

#define FMB_LAUNCH_CHECK() FMB_HIP_CHECK(hipGetLastError())

template <typename T> struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmk(T r, T i) {
  cplx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return cmk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* a += b*c */
template <typename T>
__device__ __forceinline__ void cfma(cplx<T> &a, cplx<T> b, cplx<T> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
/* a += h*c, h real */
template <typename T>
__device__ __forceinline__ void rfma(cplx<T> &a, T h, cplx<T> c) {
  a.re = fma(h, c.re, a.re);
  a.im = fma(h, c.im, a.im);
}

/* a += v with hardware float atomics (-munsafe-fp-atomics) */
template <typename T>
__device__ __forceinline__ void atomic_add_c(cplx<T> *dst, cplx<T> v) {
  unsafeAtomicAdd(&dst->re, v.re);
  unsafeAtomicAdd(&dst->im, v.im);
}

/* n with n^2 <= f < (n+1)^2: the degree of coefficient index f = n^2+n+m.
 * For 0 <= f < 2^22 the float square root is within one of n, so one
 * correction each way makes it exact (no data-dependent loop). */
__device__ __forceinline__ int isqrt_fix(int f) {
  int n = (int)sqrtf((float)f);
  n -= n * n > f;
  n += (n + 1) * (n + 1) <= f;
  return n;
}

/* one M2L offset: polar-angle table index, distance class, azimuth phi */
struct OffInfo {
  int beta, dist;
  double phi;
};

/* Morton keys: x is the most significant bit of each octal digit, coarse
 * levels most significant (octree.c:132-142). */
__device__ __forceinline__ int interleave3(int tx, int ty, int tz, int l) {
  int key = 0;
  for (int i = l - 1; i >= 0; i--)
    key = key * 8 + 4 * ((tx >> i) & 1) + 2 * ((ty >> i) & 1) + ((tz >> i) & 1);
  return key;
}

__device__ __forceinline__ void deinterleave3(int key, int l, int &tx, int &ty,
                                              int &tz) {
  int x = 0, y = 0, z = 0;
  for (int i = 0; i < l; i++) {
    int d = (key >> (3 * i)) & 7;
    x |= ((d >> 2) & 1) << i;
    y |= ((d >> 1) & 1) << i;
    z |= (d & 1) << i;
  }
  tx = x;
  ty = y;
  tz = z;
}

static inline __host__ __device__ int rot_off(int n) {
  return n * (4 * n * n - 1) / 3;
}
static inline __host__ __device__ int coax_size(int p) {
  return p * (2 * p * p + 3 * p + 1) / 6;
}
static inline __host__ __device__ int coax_blk(int p, int m) {
  int a = m < 0 ? -m : m;
  return a * (2 * a * a - (3 + 6 * p) * a + 1 + 6 * (p + p * p)) / 6;
}

/* translation.cu:189-212 in FP64 */
__device__ __forceinline__ double acoef(int n, int m) {
  int am = m < 0 ? -m : m;
  if (n < am) return 0.0;
  double nf = n, mf = m;
  return sqrt((nf + 1 + mf) * (nf + 1 - mf) / ((2 * nf + 1) * (2 * nf + 3)));
}
__device__ __forceinline__ double bcoef(int n, int m) {
  double nf = n, mf = m;
  if (m >= 0 && n >= m)
    return sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)));
  if (-n <= m && m < 0)
    return -sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)));
  return 0.0;
}

/* j_n(x), n = 0..nmax: Miller downward recurrence normalised by
 * sum (2n+1) j_n^2 = 1; serial (one lane). */
__device__ inline void dev_sph_jn(int nmax, double x, double *j) {
  if (x == 0.0) {
    j[0] = 1.0;
    for (int n = 1; n <= nmax; n++) j[n] = 0.0;
    return;
  }
  int nk = nmax < 1 ? 1 : nmax;
  double ax = fabs(x);
  double top = nk > ax ? (double)nk : ax;
  int nst = (int)ceil(top) + 40 + (int)ceil(6.0 * cbrt(top));
  double fp1 = 0.0, fc = 1.0, sum = 0.0, f0 = 0.0, f1 = 0.0;
  for (int n = nst; n >= 0; n--) {
    if (n <= nmax) j[n] = fc;
    if (n == 0) f0 = fc;
    if (n == 1) f1 = fc;
    sum += (2.0 * n + 1.0) * fc * fc;
    if (n == 0) break;
    double fm1 = (2.0 * n + 1.0) / x * fc - fp1;
    fp1 = fc;
    fc = fm1;
    if (fabs(fc) > 1e100) {
      const double s = 1e-100;
      fc *= s;
      fp1 *= s;
      sum *= s * s;
      if (n <= 1) f1 *= s;
      for (int q = (n - 1 > nmax ? nmax + 1 : n); q <= nmax; q++) j[q] *= s;
    }
  }
  double c = 1.0 / sqrt(sum);
  double sx, cx;
  sincos(x, &sx, &cx);
  double j0 = sx / x, j1 = sx / (x * x) - cx / x;
  bool neg = (fabs(j0) >= fabs(j1)) ? ((j0 < 0) != (f0 < 0))
                                    : ((j1 < 0) != (f1 < 0));
  if (neg) c = -c;
  for (int n = 0; n <= nmax; n++) j[n] *= c;
}

/* y_n(x) by upward recurrence */
__device__ inline void dev_sph_yn(int nmax, double x, double *y) {
  double sx, cx;
  sincos(x, &sx, &cx);
  y[0] = -cx / x;
  if (nmax == 0) return;
  y[1] = -cx / (x * x) - sx / x;
  for (int n = 1; n < nmax; n++) y[n + 1] = (2.0 * n + 1.0) / x * y[n] - y[n - 1];
}

/* Regular panel integrals at point x (normal nx) over triangle tri (9
 * doubles, normal ny, area) with the rule quad (nq x {b0, b1, b2, w}):
 * out = [S = int G, K = int dG/dn_y, K' = int dG/dn_x, T = int d2G/dn_x dn_y],
 * G = e^{ikr} / (4 pi r), r = |x - y| (integral.cu:72-78, 391-415, 602-622). */
__device__ inline void panel_sum(double k, const double *x, const double *nx,
                          const double *tri, const double *ny, double area,
                          const double *quad, int nq, double2 out[4]) {
  for (int i = 0; i < 4; i++) out[i] = make_double2(0, 0);
  for (int q = 0; q < nq; q++) {
    const double *qq = quad + 4 * q;
    double y[3], d[3];
    for (int a = 0; a < 3; a++) {
      y[a] = qq[0] * tri[a] + qq[1] * tri[3 + a] + qq[2] * tri[6 + a];
      d[a] = x[a] - y[a];
    }
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double s, c;
    sincos(k * r, &s, &c);
    double w = qq[3] * area;
    double g0 = 1.0 / (4.0 * M_PI * r);
    double gr = c * g0, gi = s * g0;
    double rnx = (d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]) / r;
    double rny = -(d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2]) / r;
    // gp = g (-1/r + i k)
    double gpr = -gr / r - k * gi, gpi = -gi / r + k * gr;
    double nn = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
    // d2G = e/(4 pi r^3) [ (3 - k^2 r^2 - 3ikr) rnx rny + (1 - ikr) nn ]
    double ar = (3.0 - k * k * r * r) * rnx * rny + nn;
    double ai = -3.0 * k * r * rnx * rny - k * r * nn;
    double h0 = g0 / (r * r);
    double hr = h0 * (c * ar - s * ai), hi = h0 * (c * ai + s * ar);
    out[0].x += w * gr;
    out[0].y += w * gi;
    out[1].x += w * gpr * rny;
    out[1].y += w * gpi * rny;
    out[2].x += w * gpr * rnx;
    out[2].y += w * gpi * rnx;
    out[3].x += w * hr;
    out[3].y += w * hi;
  }
}

/* Self-panel integrals at the centroid c of a flat triangle (9 doubles) as
 * contour integrals over its three edges (integral.cu:202-249, 666-751):
 *   S = (1/4pi) oint (e^{ik rho}-1)/(ik) dtheta
 *   T = -(1/4pi) [ oint (e^{ik rho}-1)/rho dtheta + oint dtheta/rho - 2 pi i k ]
 * with nedge Gauss points per edge (edge: {x, w} pairs on [-1,1]). */
__device__ inline void self_terms(double k, const double *tri, const double *c,
                           const double *edge, int nedge, double2 &S,
                           double2 &T) {
  double sSr = 0, sSi = 0, sTr = 0, sTi = 0, sinv = 0;
  for (int e = 0; e < 3; e++) {
    const double *P = tri + 3 * e, *Q = tri + 3 * ((e + 1) % 3);
    double ev[3], L = 0;
    for (int i = 0; i < 3; i++) {
      ev[i] = Q[i] - P[i];
      L += ev[i] * ev[i];
    }
    L = sqrt(L);
    double tp = 0;
    for (int i = 0; i < 3; i++) {
      ev[i] /= L;
      tp += (c[i] - P[i]) * ev[i];
    }
    double h = 0;
    for (int i = 0; i < 3; i++) {
      double f = P[i] + tp * ev[i] - c[i];
      h += f * f;
    }
    h = sqrt(h);
    double thP = atan2(-tp, h), thQ = atan2(L - tp, h);
    sinv += (sin(thQ) - sin(thP)) / h;
    double half = 0.5 * (thQ - thP), mid = 0.5 * (thQ + thP);
    for (int q = 0; q < nedge; q++) {
      double th = mid + half * edge[2 * q];
      double rho = h / cos(th);
      double w = half * edge[2 * q + 1];
      double s2, c2;
      sincos(0.5 * k * rho, &s2, &c2);
      // gs = 2 sin(k rho/2) e^{i k rho/2} / k
      double gr, gi;
      if (k == 0.0) {
        gr = rho;
        gi = 0;
      } else {
        gr = 2.0 * s2 * c2 / k;
        gi = 2.0 * s2 * s2 / k;
      }
      sSr += w * gr;
      sSi += w * gi;
      // i k gs / rho
      sTr += w * (-k * gi / rho);
      sTi += w * (k * gr / rho);
    }
  }
  const double f = 1.0 / (4.0 * M_PI);
  S = make_double2(f * sSr, f * sSi);
  T = make_double2(-f * (sTr + sinv), -f * (sTi - 2.0 * M_PI * k));
}


/* ---- spherical basis tables (P2M / L2P bases, shape sensitivities) ---- */
__device__ __forceinline__ int tri_idx(int n, int m) { return n * (n + 1) / 2 + m; }

// Fill Rt[(n,m)] = R_n^m(v), n <= nm, in LDS, in a workgroup of one
// 64-lane wavefront.  Scratch in LDS as given.
__device__ inline void r_table(int nm, double k, const double *v, double *jn,
                        double *diag, double *Pt, double2 *ph, double2 *Rt,
                        double *geo) {
  int lane = threadIdx.x;
  if (lane == 0) {
    double rho2 = v[0] * v[0] + v[1] * v[1];
    double r = sqrt(rho2 + v[2] * v[2]);
    double ct = 1.0, st = 0.0, phi = 0.0;
    if (r > 0) {
      ct = v[2] / r;
      st = sqrt(rho2) / r;
      phi = atan2(v[1], v[0]);
    }
    geo[0] = ct;
    geo[1] = st;
    geo[2] = phi;
    dev_sph_jn(nm, k * r, jn);
    double d = 1.0 / sqrt(4.0 * M_PI);
    diag[0] = d;
    for (int m = 1; m <= nm; m++) {
      d *= sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
      diag[m] = d;
    }
  }
  __syncthreads();
  double ct = geo[0], phi = geo[2];
  for (int m = lane; m <= nm; m += 64) {
    double p2 = 0.0, p1 = diag[m];
    Pt[tri_idx(m, m)] = p1;
    for (int n = m + 1; n <= nm; n++) {
      double pv;
      if (n == m + 1)
        pv = sqrt(2.0 * m + 3.0) * ct * p1;
      else {
        double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
        double b = sqrt((((double)(n - 1) * (n - 1)) - (double)m * m) /
                        (4.0 * (n - 1) * (n - 1) - 1.0));
        pv = a * (ct * p1 - b * p2);
      }
      Pt[tri_idx(n, m)] = pv;
      p2 = p1;
      p1 = pv;
    }
    double s, c;
    sincos(m * phi, &s, &c);
    ph[m] = make_double2(c, s);
  }
  __syncthreads();
  for (int f = lane; f < (nm + 1) * (nm + 1); f += 64) {
    int n = (int)sqrt((double)f);
    while (n * n > f) n--;
    while ((n + 1) * (n + 1) <= f) n++;
    int m = f - n * n - n, am = m < 0 ? -m : m;
    double a = jn[n] * Pt[tri_idx(n, am)];
    double2 e = ph[am];
    Rt[f] = make_double2(a * e.x, m < 0 ? -a * e.y : a * e.y);
  }
  __syncthreads();
}

__device__ __forceinline__ double2 rget(const double2 *Rt, int n, int m) {
  if (n < 0 || (m < 0 ? -m : m) > n) return make_double2(0.0, 0.0);
  return Rt[n * n + n + m];
}

// dR_n^m/dn (integral.cu:284-331)
__device__ __forceinline__ double2 dRdn(const double2 *Rt, int n, int m,
                                        double k, const double *nr) {
  double2 a1 = rget(Rt, n + 1, m + 1), a2 = rget(Rt, n - 1, m + 1);
  double2 b1 = rget(Rt, n + 1, m - 1), b2 = rget(Rt, n - 1, m - 1);
  double2 c1 = rget(Rt, n - 1, m), c2 = rget(Rt, n + 1, m);
  double f0 = bcoef(n + 1, -m - 1), f1 = bcoef(n, m);
  double f2 = bcoef(n + 1, m - 1), f3 = bcoef(n, -m);
  double f4 = acoef(n - 1, m), f5 = acoef(n, m);
  double t0r = f0 * a1.x - f1 * a2.x, t0i = f0 * a1.y - f1 * a2.y;
  double t1r = f2 * b1.x - f3 * b2.x, t1i = f2 * b1.y - f3 * b2.y;
  double t2r = f4 * c1.x - f5 * c2.x, t2i = f4 * c1.y - f5 * c2.y;
  double hx = 0.5 * k * nr[0], hy = 0.5 * k * nr[1], kz = k * nr[2];
  return make_double2(hx * t0r + hy * t0i + hx * t1r - hy * t1i + kz * t2r,
                      hx * t0i - hy * t0r + hx * t1i + hy * t1r + kz * t2i);
}


#endif
