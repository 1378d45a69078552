// ops.hip -- FP64 generation of the rotate-coaxial-rotate operators on the
// device with O(p^3) work and O(p^2) scratch per matrix, and their scaled
// conversion to the compute precision.
//
//   rotation H_n^{m',m}(beta) in the normalisation of rotMatsInit /
//     sparseRotMatGen (translation.cu:1064-1237), computed from Wigner d by the
//     recurrence in the degree (stable; the reference's recurrence in m is
//     not); phases e^{i m alpha} e^{-i m' gamma} (translation.cu:1229) are
//     applied on the fly by the translation kernels.
//   coaxial (R|R), (S|R): initial column (-1)^n' sqrt(2n'+1) f_n'(kt)
//     (translation.cu:555-584, 919-949), sectorial recurrence
//     (translation.cu:689-711), advancing in n (translation.cu:713-745),
//     |m|-block storage (translation.cu:760-769).
// Rotation: one lane per (m', m) column.  Coaxial: one workgroup per matrix
// (sectorial pass) or per |m| block (advancing pass), sequential in m
// (resp. n) with the orthogonal index spread over the workgroup.
#include <mutex>

#include "fmb_device.h"

namespace {

constexpr int NT = 256;

// Wigner d^l_{m'm}(beta) from its closed form at l0 = max(|m'|,|m|)
__device__ double d_top(int j, int m, double beta) {
  double s, c;
  sincos(0.5 * beta, &s, &c);
  double lv = 0.5 * (lgamma(2.0 * j + 1) - lgamma(j + m + 1.0) - lgamma(j - m + 1.0));
  double sg = ((j - m) & 1) ? -1.0 : 1.0;
  if (j + m > 0) {
    if (c == 0.0) return 0.0;
    lv += (j + m) * log(fabs(c));
    if (c < 0 && ((j + m) & 1)) sg = -sg;
  }
  if (j - m > 0) {
    if (s == 0.0) return 0.0;
    lv += (j - m) * log(fabs(s));
    if (s < 0 && ((j - m) & 1)) sg = -sg;
  }
  return sg * exp(lv);
}

__device__ double d_start(int mp, int m, double beta) {
  int j = abs(mp) > abs(m) ? abs(mp) : abs(m);
  if (mp == j) return d_top(j, m, beta);
  if (mp == -j) return ((j + m) & 1 ? -1.0 : 1.0) * d_top(j, -m, beta);
  if (m == j) return ((j - mp) & 1 ? -1.0 : 1.0) * d_top(j, mp, beta);
  return d_top(j, -mp, beta);
}

// H_n^{m',m}(beta) = eps_{m'} (-1)^m eps_m d^n_{m'm}(beta), one lane per
// (m', m) running the three-term recurrence in the degree n.  The reference's
// recurrence in m (translation.cu:1179-1204) loses about a digit per degree
// past n ~ 15 and is not used.
__global__ void __launch_bounds__(NT)
    k_gen_H(const double *betas, int P, double *H) {
  int b = blockIdx.y;
  int W = 2 * P - 1;
  int idx = blockIdx.x * NT + threadIdx.x;
  if (idx >= W * W) return;
  int mp = idx % W - (P - 1), m = idx / W - (P - 1);
  double beta = betas[b];
  double cb = cos(beta);
  double *Hb = H + (size_t)b * rot_off(P);
  int l0 = abs(mp) > abs(m) ? abs(mp) : abs(m);
  double em = (m < 0 && ((-m) & 1)) ? -1.0 : 1.0;
  double ep = (mp > 0 && (mp & 1)) ? -1.0 : 1.0;
  double f = em * ep;
  double dm2 = 0.0, dm1 = d_start(mp, m, beta);
  for (int n = l0; n < P; n++) {
    double v;
    if (n == l0)
      v = dm1;
    else if (n == 1)
      v = cb;
    else {
      double a = (2.0 * n - 1.0) * (n * (n - 1.0) * cb - (double)mp * m);
      double bq = n * sqrt(fmax(0.0, ((n - 1.0) * (n - 1.0) - (double)mp * mp) *
                                         ((n - 1.0) * (n - 1.0) - (double)m * m)));
      double den = (n - 1.0) * sqrt(((double)n * n - (double)mp * mp) *
                                    ((double)n * n - (double)m * m));
      v = (a * dm1 - bq * dm2) / den;
    }
    if (n > l0) {
      dm2 = dm1;
      dm1 = v;
    }
    Hb[rot_off(n) + (m + n) * (2 * n + 1) + (mp + n)] = f * v;
  }
}

// sectorial values S[m][n'] = E(m, n', m), n' in [m, 2P-2-m]
__global__ void __launch_bounds__(NT)
    k_coax_sect(int kind, double k, const double *ts, int P, double2 *S) {
  int d = blockIdx.x;
  int nmax = 2 * P - 2, ld = 2 * P - 1;
  double2 *Sd = S + (size_t)d * P * ld;
  __shared__ double jn[512], yn[512];
  if (threadIdx.x == 0) {
    double x = k * ts[d];
    dev_sph_jn(nmax, x, jn);
    if (kind == FMB_COAX_SR) dev_sph_yn(nmax, x, yn);
  }
  __syncthreads();
  for (int np = threadIdx.x; np <= nmax; np += NT) {
    double s = ((np & 1) ? -1.0 : 1.0) * sqrt(2.0 * np + 1.0);
    Sd[np] = make_double2(s * jn[np], kind == FMB_COAX_SR ? s * yn[np] : 0.0);
  }
  __syncthreads();
  for (int m = 0; m + 1 <= P - 1; m++) {
    double den = bcoef(m + 1, -m - 1);
    const double2 *cur = Sd + (size_t)m * ld;
    double2 *nx = Sd + (size_t)(m + 1) * ld;
    for (int np = m + 1 + threadIdx.x; np <= nmax - (m + 1); np += NT) {
      double a0 = bcoef(np, -m - 1) / den, a1 = bcoef(np + 1, m) / den;
      double2 u = cur[np - 1], v = cur[np + 1];
      nx[np] = make_double2(a0 * u.x - a1 * v.x, a0 * u.y - a1 * v.y);
    }
    __syncthreads();
  }
}

// advancing in n inside block m; writes the |m| block of the sparse matrix
__global__ void __launch_bounds__(NT)
    k_coax_adv(int P, const double2 *S, double2 *out) {
  int d = blockIdx.y, m = blockIdx.x;
  int nmax = 2 * P - 2, ld = 2 * P - 1;
  const double2 *Sm = S + ((size_t)d * P + m) * ld;
  double2 *blk = out + (size_t)d * coax_size(P) + coax_blk(P, m);
  int bl = P - m;
  extern __shared__ double2 sh[];
  double2 *prev = sh, *cur = sh + ld, *nxt = sh + 2 * ld;
  for (int np = threadIdx.x; np < ld; np += NT) {
    prev[np] = make_double2(0.0, 0.0);
    cur[np] = (np >= m && np <= nmax - m) ? Sm[np] : make_double2(0.0, 0.0);
    nxt[np] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int np = m + threadIdx.x; np < P; np += NT)
    blk[(size_t)(m - m) * bl + (np - m)] = cur[np];
  for (int n = m; n + 1 <= P - 1; n++) {
    double an = acoef(n, m);
    for (int np = m + threadIdx.x; np <= nmax - (n + 1); np += NT) {
      double2 v0 = (np - 1 >= m) ? cur[np - 1] : make_double2(0.0, 0.0);
      double2 v1 = cur[np + 1];
      double2 v2 = (n - 1 >= m) ? prev[np] : make_double2(0.0, 0.0);
      double c0 = acoef(np - 1, m), c1 = acoef(np, m), c2 = acoef(n - 1, m);
      nxt[np] = make_double2((c0 * v0.x - c1 * v1.x + c2 * v2.x) / an,
                             (c0 * v0.y - c1 * v1.y + c2 * v2.y) / an);
    }
    __syncthreads();
    double2 *t = prev;
    prev = cur;
    cur = nxt;
    nxt = t;
    for (int np = m + threadIdx.x; np < P; np += NT)
      blk[(size_t)(n + 1 - m) * bl + (np - m)] = cur[np];
    __syncthreads();
  }
  // entries with n' < n from the exact symmetry (E|F)^m_{n'n} =
  // (-1)^{n+n'} (E|F)^m_{nn'}: the advancing step loses accuracy there when
  // kt is small (it subtracts terms (kt)^-2 larger than its result)
  for (int e = threadIdx.x; e < bl * bl; e += NT) {
    int r = e % bl, c = e / bl;
    if (r < c) {
      double2 v = blk[(size_t)r * bl + c];
      double sg = ((r + c) & 1) ? -1.0 : 1.0;
      blk[e] = make_double2(sg * v.x, sg * v.y);
    }
  }
}

template <typename T>
__global__ void k_coax_scale(const double2 *in, int P, const double *rs,
                             const double *cs, int inv_r, int inv_c, T *out) {
  int m = blockIdx.x, bl = P - m;
  size_t off = coax_blk(P, m);
  for (int e = threadIdx.x; e < bl * bl; e += blockDim.x) {
    int np = m + e % bl, n = m + e / bl;
    double r = rs ? (inv_r ? 1.0 / rs[np] : rs[np]) : 1.0;
    double c = cs ? (inv_c ? 1.0 / cs[n] : cs[n]) : 1.0;
    double2 v = in[off + e];
    out[2 * (off + e)] = (T)(r * c * v.x);
    out[2 * (off + e) + 1] = (T)(r * c * v.y);
  }
}

template <typename T>
__global__ void k_convert(const double *in, size_t n, T *out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) out[i] = (T)in[i];
}

__global__ void k_rot_phase(const double *H, int P, const double *ang,
                            double2 *out) {
  int a = blockIdx.y;
  double al = ang[3 * a], ga = ang[3 * a + 2];
  const double *Hb = H + (size_t)a * rot_off(P);
  double2 *o = out + (size_t)a * rot_off(P);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < rot_off(P);
       f += gridDim.x * blockDim.x) {
    // flat sparse index -> (n, m', m)
    int n = 0;
    while (rot_off(n + 1) <= f) n++;
    int e = f - rot_off(n), ld = 2 * n + 1;
    int mp = e % ld - n, m = e / ld - n;
    double s, c;
    sincos(m * al - mp * ga, &s, &c);
    o[f] = make_double2(Hb[f] * c, Hb[f] * s);
  }
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

extern "C" int fmbh_gen_H(const double *h_beta, int nbeta, int P, double *d_H,
                          void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (P < 1 || 2 * P - 2 >= 512)
    return fmb_set_error(FMB_ERR_RANGE, "rotation order %d out of range", P);
  double *d_b;
  int W = 2 * P - 1;
  int rc = fmbh_malloc((void **)&d_b, sizeof(double) * nbeta);
  if (rc) return rc;
  if ((rc = fmbh_h2d(d_b, h_beta, sizeof(double) * nbeta, stream))) return rc;
  hipLaunchKernelGGL(k_gen_H, dim3((W * W + NT - 1) / NT, nbeta), dim3(NT), 0, st,
                     d_b, P, d_H);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(d_b);
  return FMB_OK;
}

extern "C" int fmbh_gen_coax(int kind, double k, const double *h_t, int n,
                             int P, fmb_c64 *d_out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (P < 1 || 2 * P - 2 >= 512)
    return fmb_set_error(FMB_ERR_RANGE, "coaxial order %d out of range", P);
  double *d_t;
  double2 *S;
  int ld = 2 * P - 1;
  int rc = fmbh_malloc((void **)&d_t, sizeof(double) * n);
  if (!rc) rc = fmbh_malloc((void **)&S, sizeof(double2) * (size_t)n * P * ld);
  if (rc) return rc;
  if ((rc = fmbh_h2d(d_t, h_t, sizeof(double) * n, stream))) return rc;
  hipLaunchKernelGGL(k_coax_sect, dim3(n), dim3(NT), 0, st, kind, k, d_t, P, S);
  FMB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_coax_adv, dim3(P, n), dim3(NT),
                     3 * ld * sizeof(double2), st, P, S, (double2 *)d_out);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(d_t);
  fmbh_free(S);
  return FMB_OK;
}

extern "C" int fmbh_coax_scale(int prec, const fmb_c64 *d_in, int P,
                               const double *d_rs, const double *d_cs,
                               int invert_rs, int invert_cs, void *d_out,
                               void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (prec == FMB_FP64)
    hipLaunchKernelGGL(k_coax_scale<double>, dim3(P), dim3(256), 0, st,
                       (const double2 *)d_in, P, d_rs, d_cs, invert_rs,
                       invert_cs, (double *)d_out);
  else
    hipLaunchKernelGGL(k_coax_scale<float>, dim3(P), dim3(256), 0, st,
                       (const double2 *)d_in, P, d_rs, d_cs, invert_rs,
                       invert_cs, (float *)d_out);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_convert_real(int prec, const double *d_in, size_t n,
                                 void *d_out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  int nb = nblk(n);
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  if (prec == FMB_FP64)
    hipLaunchKernelGGL(k_convert<double>, dim3(nb), dim3(256), 0, st, d_in, n,
                       (double *)d_out);
  else
    hipLaunchKernelGGL(k_convert<float>, dim3(nb), dim3(256), 0, st, d_in, n,
                       (float *)d_out);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}

extern "C" int fmbh_rot_phase(const double *d_H, int n, int P,
                              const double *h_angles, fmb_c64 *d_out,
                              void *stream) {
  hipStream_t st = (hipStream_t)stream;
  double *d_a;
  int rc = fmbh_malloc((void **)&d_a, sizeof(double) * 3 * n);
  if (rc) return rc;
  if ((rc = fmbh_h2d(d_a, h_angles, sizeof(double) * 3 * n, stream))) return rc;
  int nb = nblk(rot_off(P));
  hipLaunchKernelGGL(k_rot_phase, dim3(nb, n), dim3(256), 0, st, d_H, P, d_a,
                     (double2 *)d_out);
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  fmbh_free(d_a);
  return FMB_OK;
}

// ---- the coaxial generator without allocations ----
// The sectorial table of fmbh_gen_coax (n * P * (2P - 1) complex values)
// lives in a fixed device array instead of a fresh allocation, for callers
// that cannot allocate.  Every call on a device shares that device's copy;
// each call writes every entry it reads (k_coax_sect fills row m before
// k_coax_adv reads it), so nothing needs clearing between calls, and calls
// that overlap on one device race.
#define FMB_COAX_SCRATCH (8u << 20) /* bytes */
static __device__ __align__(256) double2 scratch[FMB_COAX_SCRATCH / sizeof(double2)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
void *scratch_ptr[kMaxDevices];  // device address of `scratch`, per device
}  // namespace

// the device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address)
static int scratch_address(void **out) {
  int dev = 0;
  FMB_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fmb_set_error(FMB_ERR_RANGE, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> g(scratch_lock);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fmb_set_error(FMB_ERR_HIP, "scratch address: %s", hipGetErrorString(e));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return FMB_OK;
}

extern "C" size_t fmbh_coax_scratch_bytes(void) { return FMB_COAX_SCRATCH; }

extern "C" int fmbh_gen_coax_static(int kind, double k, const double *d_t, int n, int P,
                                    fmb_c64 *d_out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (P < 1 || 2 * P - 2 >= 512)
    return fmb_set_error(FMB_ERR_RANGE, "coaxial order %d out of range", P);
  const size_t need = sizeof(double2) * (size_t)n * P * (2 * P - 1);
  if (need > FMB_COAX_SCRATCH)
    return fmb_set_error(FMB_ERR_RANGE,
                         "%d coaxial matrices of order %d need %zu bytes of scratch, "
                         "above the %u available (n * p * (2p - 1) * 16 bytes)",
                         n, P, need, FMB_COAX_SCRATCH);
  void *S = nullptr;
  int rc = scratch_address(&S);
  if (rc) return rc;
  int ld = 2 * P - 1;
  hipLaunchKernelGGL(k_coax_sect, dim3(n), dim3(NT), 0, st, kind, k, d_t, P, (double2 *)S);
  FMB_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_coax_adv, dim3(P, n), dim3(NT), 3 * ld * sizeof(double2), st, P,
                     (const double2 *)S, (double2 *)d_out);
  FMB_LAUNCH_CHECK();
  return FMB_OK;
}
