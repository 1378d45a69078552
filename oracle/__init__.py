"""oracle -- ctypes binding of liboracle.so, the FP64 CPU restatement of the
reference FMM-BEM algorithm (fmb_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects the sanitizer build (make -C oracle san)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

RULE_COLLAPSED, RULE_GAUSS7 = 0, 1
COAX_RR, COAX_SR = 0, 1

_lib = None
_vp = ctypes.c_void_p
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C")
_ci, _cd = ctypes.c_int, ctypes.c_double

_SIGS = {
    "fo_acoef": (_cd, [_ci, _ci]),
    "fo_bcoef": (_cd, [_ci, _ci]),
    "fo_sph_jn": (None, [_ci, _cd, _dp]),
    "fo_sph_yn": (None, [_ci, _cd, _dp]),
    "fo_legendre_table": (None, [_ci, _cd, _cd, _dp]),
    "fo_ylm": (None, [_ci, _cd, _cd, _dp, _dp]),
    "fo_rbasis": (None, [_ci, _cd, _dp, _dp, _dp]),
    "fo_sbasis": (None, [_ci, _cd, _dp, _dp, _dp]),
    "fo_rbasis_dn": (None, [_ci, _cd, _dp, _dp, _dp, _dp]),
    "fo_gauss_legendre": (_ci, [_ci, _dp, _dp]),
    "fo_tri_rule": (_ci, [_ci, _ci, _dp, _dp]),
    "fo_kernels": (None, [_cd, _dp, _dp, _dp, _dp, _dp, _dp]),
    "fo_panel_regular": (None, [_cd, _dp, _dp, _dp, _ci, _ci, _dp, _dp]),
    "fo_panel_self": (None, [_cd, _dp, _ci, _dp, _dp]),
    "fo_panel_hyper_ref": (None, [_cd, _dp, _ci, ctypes.POINTER(_cd), ctypes.POINTER(_cd)]),
    "fo_trunc_num": (_ci, [_cd, _cd, _cd, _cd]),
    "fo_morton_key": (_ci, [_dp, _ci]),
    "fo_box_center": (None, [_ci, _ci, _dp]),
    "fo_neighbors": (_ci, [_ci, _ci, _ip]),
    "fo_I4": (_ci, [_ci, _ci, _ip]),
    "fo_deter_lmax": (_ci, [_dp, _ci, _ci]),
    "fo_bounding_cube": (_ci, [_dp, _ci, _cd, _dp, ctypes.POINTER(_cd)]),
    "fo_tree_build": (_ci, [_dp, _ci, _ip, _ci, _ci, _ci, ctypes.POINTER(_vp)]),
    "fo_tree_free": (None, [_vp]),
    "fo_tree_lmax": (_ci, [_vp]),
    "fo_tree_nbox": (_ci, [_vp, _ci]),
    "fo_tree_keys": (None, [_vp, _ci, _ip]),
    "fo_tree_geom": (None, [_vp, ctypes.POINTER(_cd), _dp]),
    "fo_tree_perm": (None, [_vp, _ip]),
    "fo_tree_num_sr": (_ci, [_vp, _ci, _ip]),
    "fo_rot_size": (_ci, [_ci]),
    "fo_coax_size": (_ci, [_ci]),
    "fo_coax_block": (_ci, [_ci, _ci]),
    "fo_rot_H": (None, [_ci, _cd, _dp]),
    "fo_rot_H_ref": (None, [_ci, _cd, _dp]),
    "fo_rot_sparse": (None, [_ci, _cd, _cd, _cd, _dp, _dp]),
    "fo_coax_sparse": (None, [_ci, _ci, _cd, _cd, _dp, _dp]),
    "fo_rot_apply": (None, [_ci, _dp, _dp, _dp, _dp, _dp, _dp]),
    "fo_coax_apply": (None, [_ci, _ci, _ci, _dp, _dp, _dp, _dp, _dp, _dp]),
    "fo_translate": (None, [_ci, _ci, _ci, _cd, _dp, _dp, _dp, _dp, _dp]),
    "fo_bem_create": (_ci, [_dp, _ci, _ip, _ci, _cd, _dp, _vp, _vp, _ci, _ci, _ci,
                            ctypes.POINTER(_vp)]),
    "fo_bem_free": (None, [_vp]),
    "fo_bem_set_near": (None, [_vp, _ci, _ci, _cd]),
    "fo_bem_n": (_ci, [_vp]),
    "fo_bem_dense_row": (None, [_vp, _ci, _dp, _dp]),
    "fo_bem_dense_matvec": (None, [_vp, _ci, _dp, _dp, _dp, _dp]),
    "fo_bem_rhs_plane": (None, [_vp, _dp, _dp, _dp]),
    "fo_bem_centroids": (None, [_vp, _dp, _dp, _dp]),
    "fo_bem_fmm_setup": (_ci, [_vp, _ci, _ci, _vp]),
    "fo_bem_fmm_orders": (_ci, [_vp, _ip]),
    "fo_bem_fmm_matvec": (None, [_vp, _ci, _dp, _dp, _dp, _dp]),
    "fo_bem_field": (None, [_vp, _ci, _dp, _dp, _dp, _ci, _vp, _dp, _dp]),
    "fo_bem_field_fmm": (None, [_vp, _ci, _dp, _dp, _dp, _ci, _vp, _dp, _dp]),
    "fo_bem_near_matvec": (None, [_vp, _ci, _dp, _dp, _dp, _dp]),
    "fo_gmres": (_ci, [_vp, _ci, _ci, _cd, _ci, _dp, _dp, _dp, _dp, ctypes.POINTER(_ci),
                       ctypes.POINTER(_cd)]),
    "fo_mie_rigid": (None, [_cd, _ci, _dp, _ci, _dp, _dp]),
    "fo_num_threads": (_ci, []),
}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def sph_jn(nmax, x):
    j = np.zeros(nmax + 1)
    lib().fo_sph_jn(nmax, x, j)
    return j


def sph_yn(nmax, x):
    y = np.zeros(nmax + 1)
    lib().fo_sph_yn(nmax, x, y)
    return y


def ylm(nmax, theta, phi):
    re = np.zeros((nmax + 1) ** 2)
    im = np.zeros_like(re)
    lib().fo_ylm(nmax, theta, phi, re, im)
    return re + 1j * im


def rbasis(nmax, k, v):
    re = np.zeros((nmax + 1) ** 2)
    im = np.zeros_like(re)
    lib().fo_rbasis(nmax, k, _c(v), re, im)
    return re + 1j * im


def sbasis(nmax, k, v):
    re = np.zeros((nmax + 1) ** 2)
    im = np.zeros_like(re)
    lib().fo_sbasis(nmax, k, _c(v), re, im)
    return re + 1j * im


def rbasis_dn(nmax, k, v, nrm):
    re = np.zeros((nmax + 1) ** 2)
    im = np.zeros_like(re)
    lib().fo_rbasis_dn(nmax, k, _c(v), _c(nrm), re, im)
    return re + 1j * im


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().fo_gauss_legendre(n, x, w)
    return x, w


def tri_rule(rule, order=3):
    n = 7 if rule == RULE_GAUSS7 else order * order
    b = np.zeros(3 * n)
    w = np.zeros(n)
    q = lib().fo_tri_rule(rule, order, b, w)
    return b.reshape(-1, 3)[:q], w[:q]


def kernels(k, x, nx, y, ny):
    re, im = np.zeros(4), np.zeros(4)
    lib().fo_kernels(k, _c(x), _c(nx), _c(y), _c(ny), re, im)
    return re + 1j * im


def panel_regular(k, tri, x, nx, rule=RULE_GAUSS7, order=3):
    re, im = np.zeros(4), np.zeros(4)
    lib().fo_panel_regular(k, _c(tri).ravel(), _c(x), _c(nx), rule, order, re, im)
    return re + 1j * im


def panel_self(k, tri, nedge=24):
    re, im = np.zeros(4), np.zeros(4)
    lib().fo_panel_self(k, _c(tri).ravel(), nedge, re, im)
    return re + 1j * im


def panel_hyper_ref(k, tri, order):
    re, im = _cd(), _cd()
    lib().fo_panel_hyper_ref(k, _c(tri).ravel(), order, ctypes.byref(re), ctypes.byref(im))
    return re.value + 1j * im.value


def trunc_num(k, eps, sigma, a):
    return lib().fo_trunc_num(k, eps, sigma, a)


def morton_key(s, l):
    return lib().fo_morton_key(_c(s), l)


def box_center(key, l):
    c = np.zeros(3)
    lib().fo_box_center(key, l, c)
    return c


def neighbors(key, l):
    out = np.zeros(27, dtype=np.int32)
    n = lib().fo_neighbors(key, l, out)
    return out[:n].copy()


def I4(key, l):
    out = np.zeros(216, dtype=np.int32)
    n = lib().fo_I4(key, l, out)
    return out[:n].copy()


def deter_lmax(pts, s):
    pts = _c(pts).reshape(-1, 3)
    return lib().fo_deter_lmax(pts.ravel(), pts.shape[0], s)


class Tree:
    def __init__(self, V, F, s=5, depth=0):
        V = _c(V).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        self.h = _vp()
        rc = lib().fo_tree_build(V.ravel(), V.shape[0], F.ravel(), F.shape[0], s, depth,
                                 ctypes.byref(self.h))
        if rc:
            raise RuntimeError("fo_tree_build failed")
        self.n = F.shape[0]

    @property
    def lmax(self):
        return lib().fo_tree_lmax(self.h)

    def keys(self, l):
        n = lib().fo_tree_nbox(self.h, l)
        k = np.zeros(n, dtype=np.int32)
        lib().fo_tree_keys(self.h, l, k)
        return k

    def geom(self):
        D = _cd()
        p = np.zeros(3)
        lib().fo_tree_geom(self.h, ctypes.byref(D), p)
        return D.value, p

    def perm(self):
        p = np.zeros(self.n, dtype=np.int32)
        lib().fo_tree_perm(self.h, p)
        return p

    def num_sr(self, l):
        n = np.zeros(lib().fo_tree_nbox(self.h, l), dtype=np.int32)
        lib().fo_tree_num_sr(self.h, l, n)
        return n

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fo_tree_free(self.h)
            self.h = None


def rot_H(p, beta):
    H = np.zeros(lib().fo_rot_size(p))
    lib().fo_rot_H(p, beta, H)
    return H


def rot_H_ref(p, beta):
    H = np.zeros(lib().fo_rot_size(p))
    lib().fo_rot_H_ref(p, beta, H)
    return H


def rot_sparse(p, alpha, beta, gamma):
    n = lib().fo_rot_size(p)
    re, im = np.zeros(n), np.zeros(n)
    lib().fo_rot_sparse(p, alpha, beta, gamma, re, im)
    return re + 1j * im


def coax_sparse(kind, p, k, t):
    n = lib().fo_coax_size(p)
    re, im = np.zeros(n), np.zeros(n)
    lib().fo_coax_sparse(kind, p, k, t, re, im)
    return re + 1j * im


def translate(kind, pin, pout, k, t, x):
    x = np.asarray(x, dtype=np.complex128)
    yr, yi = np.zeros(pout * pout), np.zeros(pout * pout)
    lib().fo_translate(kind, pin, pout, k, _c(t), _c(x.real), _c(x.imag), yr, yi)
    return yr + 1j * yi


def mie_rigid(ka, cos_theta, nterms=80):
    ct = _c(cos_theta)
    re, im = np.zeros(ct.size), np.zeros(ct.size)
    lib().fo_mie_rigid(ka, ct.size, ct, nterms, re, im)
    return re + 1j * im


def mie_rigid_field(k, a, pts, d, nterms=None):
    """Total pressure at points pts (M x 3) off a rigid sphere of radius a at
    the origin, for the unit plane wave e^{ik d.x}: the exterior series
    e^{ik d.x} - sum_n (2n+1) i^n j_n'(ka)/h_n'(ka) h_n(kr) P_n(d.x/r) with
    h_n = j_n + i y_n (scipy's spherical Bessel functions).  At r = a it is
    mie_rigid, which sums the Wronskian form of the same series."""
    import scipy.special as sp
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    d = np.asarray(d, dtype=np.float64)
    r = np.linalg.norm(pts, axis=1)
    ct = (pts @ d) / r
    if nterms is None:
        nterms = int(k * a) + 40
    n = np.arange(nterms + 1)
    jp = sp.spherical_jn(n, k * a, derivative=True)
    hp = jp + 1j * sp.spherical_yn(n, k * a, derivative=True)
    c = -(2 * n + 1) * (1j ** n) * jp / hp
    s = np.zeros(len(r), dtype=np.complex128)
    P0, P1 = np.ones_like(ct), ct
    for q in range(nterms + 1):
        if q == 0:
            P = P0
        elif q == 1:
            P = P1
        else:
            P = ((2 * q - 1) * ct * P1 - (q - 1) * P0) / q
            P0, P1 = P1, P
        h = sp.spherical_jn(q, k * r) + 1j * sp.spherical_yn(q, k * r)
        s += c[q] * h * P
    return np.exp(1j * k * (pts @ d)) + s


class Bem:
    """Burton-Miller collocation BEM (dense and FMM operators, GMRES)."""

    def __init__(self, V, F, k, alpha=None, sigma=None, rule=RULE_GAUSS7, order=3,
                 nedge=24, near=(RULE_COLLAPSED, 8, 2.0)):
        V = _c(V).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        if alpha is None:
            alpha = 1j / k
        self._sig = None
        sr = si = None
        if sigma is not None:
            s = np.asarray(sigma, dtype=np.complex128)
            self._sig = (_c(s.real), _c(s.imag))
            sr = self._sig[0].ctypes.data_as(_vp)
            si = self._sig[1].ctypes.data_as(_vp)
        self.h = _vp()
        rc = lib().fo_bem_create(V.ravel(), V.shape[0], F.ravel(), F.shape[0], k,
                                 np.array([alpha.real, alpha.imag]), sr, si, rule, order,
                                 nedge, ctypes.byref(self.h))
        if rc:
            raise RuntimeError("fo_bem_create failed")
        if near is not None:
            lib().fo_bem_set_near(self.h, near[0], near[1], near[2])
        self.n = F.shape[0]
        self.k = k

    def fmm_setup(self, s=5, depth=0, plev=None):
        pl = None
        if plev is not None:
            self._plev = np.zeros(16, dtype=np.int32)
            for l, p in plev.items():
                self._plev[l] = p
            pl = self._plev.ctypes.data_as(_vp)
        rc = lib().fo_bem_fmm_setup(self.h, s, depth, pl)
        if rc:
            raise RuntimeError("fo_bem_fmm_setup failed")

    def orders(self):
        p = np.zeros(16, dtype=np.int32)
        L = lib().fo_bem_fmm_orders(self.h, p)
        return L, p

    def _mv(self, fn, x):
        x = np.atleast_2d(np.asarray(x, dtype=np.complex128))
        nr = x.shape[0]
        yr, yi = np.zeros(x.shape), np.zeros(x.shape)
        fn(self.h, nr, _c(x.real), _c(x.imag), yr, yi)
        return yr + 1j * yi

    def dense_matvec(self, x):
        return self._mv(lib().fo_bem_dense_matvec, x)

    def fmm_matvec(self, x):
        return self._mv(lib().fo_bem_fmm_matvec, x)

    def near_matvec(self, x):
        return self._mv(lib().fo_bem_near_matvec, x)

    def field(self, x, pts, dirs=None, fmm=False):
        """Pressure at field points pts (M x 3) from the surface solution x
        (nrhs x N or N); dirs: incident plane-wave directions (nrhs x 3) or
        None for the scattered field only.  fmm: through the tree (after
        fmm_setup) instead of the direct sum."""
        x = np.atleast_2d(np.asarray(x, dtype=np.complex128))
        nrhs = x.shape[0]
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
        ore = np.zeros((nrhs, pts.shape[0]))
        oim = np.zeros((nrhs, pts.shape[0]))
        d = None
        if dirs is not None:
            d = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
        fn = lib().fo_bem_field_fmm if fmm else lib().fo_bem_field
        fn(self.h, nrhs, xr, xi, pts, pts.shape[0],
           d.ctypes.data_as(ctypes.c_void_p) if d is not None else None, ore, oim)
        return ore + 1j * oim

    def dense_row(self, i):
        re, im = np.zeros(self.n), np.zeros(self.n)
        lib().fo_bem_dense_row(self.h, i, re, im)
        return re + 1j * im

    def dense_matrix(self):
        return np.array([self.dense_row(i) for i in range(self.n)])

    def rhs_plane(self, d):
        re, im = np.zeros(self.n), np.zeros(self.n)
        lib().fo_bem_rhs_plane(self.h, _c(d), re, im)
        return re + 1j * im

    def centroids(self):
        c, n, a = np.zeros(3 * self.n), np.zeros(3 * self.n), np.zeros(self.n)
        lib().fo_bem_centroids(self.h, c, n, a)
        return c.reshape(-1, 3), n.reshape(-1, 3), a

    def gmres(self, b, use_fmm=True, m=30, tol=1e-6, maxit=2000):
        b = np.asarray(b, dtype=np.complex128)
        xr, xi = np.zeros(self.n), np.zeros(self.n)
        it, rr = _ci(), _cd()
        rc = lib().fo_gmres(self.h, int(use_fmm), m, tol, maxit, _c(b.real), _c(b.imag), xr,
                            xi, ctypes.byref(it), ctypes.byref(rr))
        return xr + 1j * xi, it.value, rr.value, rc

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fo_bem_free(self.h)
            self.h = None
