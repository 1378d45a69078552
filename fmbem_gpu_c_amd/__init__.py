"""fmbem_gpu_c_amd -- Python shim over libfmbem.so, the MI355X (gfx950) FMM-BEM
solver for 3-D Helmholtz acoustics.

The C ABI (include/fmbem.h) is the product; this module binds it with ctypes
so tests, benchmarks and torch code can drive it.  Device vectors are torch
complex128 CUDA tensors (interleaved re/im, the layout of fmb_c64).  There
is no fallback: if libfmbem.so is missing or a call fails, FmbError is
raised.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FMBEM_LIB selects another build of the same library (the sanitizer build)
LIB_PATH = os.environ.get("FMBEM_LIB") or os.path.join(_HERE, "libfmbem.so")

FMB_FP64, FMB_MIXED = 0, 1
PRULE_TRUNCNUM, PRULE_EXCESS, PRULE_EXPLICIT = 0, 1, 2
RULE_COLLAPSED, RULE_GAUSS7 = 0, 1
COAX_RR, COAX_SR = 0, 1
RCR_AUTO, RCR_VALU = 0, 1
RCR_MODE_RULE, RCR_MODE_LDS, RCR_MODE_GWS, RCR_MODE_LDSP, RCR_MODE_LDSO = 0, 1, 2, 3, 4


class FmbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fmbem error %d: %s" % (code, msg))
        self.code = code


class PlanParams(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_double),
        ("alpha_re", ctypes.c_double),
        ("alpha_im", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("p_rule", ctypes.c_int),
        ("eps", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("p_levels", ctypes.c_int * 16),
        ("p_min", ctypes.c_int),
        ("quad_rule", ctypes.c_int),
        ("quad_order", ctypes.c_int),
        ("near_rule", ctypes.c_int),
        ("near_order", ctypes.c_int),
        ("near_ratio", ctypes.c_double),
        ("nedge", ctypes.c_int),
        ("nrhs_max", ctypes.c_int),
        ("rcr_path", ctypes.c_int),
        ("k_max", ctypes.c_double),
        ("autotune", ctypes.c_int),
        ("serial_levels", ctypes.c_int),
        ("waves", ctypes.c_int),
        ("split_bf16", ctypes.c_int),
        ("ldso", ctypes.c_int),
        ("m2l_tile", ctypes.c_int),
        ("m2l_tile_min_mb", ctypes.c_int),
        ("rcr_mode", ctypes.c_int),
    ]


class GmresInfo(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("matvecs_fp64", ctypes.c_int),
        ("matvecs_work", ctypes.c_int),
        ("restarts", ctypes.c_int),
        ("relres", ctypes.c_double * 64),
        ("seconds", ctypes.c_double),
    ]


_lib = None
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every entry is declared in include/fmbem.h
_SIGS = {
    "fmb_last_error": (ctypes.c_char_p, []),
    "fmb_version": (ctypes.c_int, []),
    "fmb_obj_count": (ctypes.c_int, [ctypes.c_char_p, _ip, _ip]),
    "fmb_mesh_load_obj": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "fmb_mesh_from_arrays": (ctypes.c_int, [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.POINTER(_vp)]),
    "fmb_mesh_icosphere": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)]),
    "fmb_mesh_ellipsoid": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_ulonglong, ctypes.POINTER(_vp)]),
    "fmb_mesh_orient_outward": (ctypes.c_int, [_vp, _ip]),
    "fmb_mesh_sizes": (ctypes.c_int, [_vp, _ip, _ip]),
    "fmb_mesh_export": (ctypes.c_int, [_vp, _dp, _ip]),
    "fmb_mesh_destroy": (None, [_vp]),
    "fmb_tree_build": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "fmb_tree_info": (ctypes.c_int, [_vp, _ip, _ip, _dp, _dp]),
    "fmb_tree_level": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "fmb_tree_num_sr": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "fmb_tree_elem_box": (ctypes.c_int, [_vp, _ip, _ip]),
    "fmb_tree_pairs": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _ip]),
    "fmb_tree_destroy": (None, [_vp]),
    "fmb_trunc_num": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "fmb_rot_size": (ctypes.c_int, [ctypes.c_int]),
    "fmb_coax_size": (ctypes.c_int, [ctypes.c_int]),
    "fmb_gen_rot_mats": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _vp]),
    "fmb_gen_coax_mats": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                                         ctypes.c_int, _vp]),
    "fmb_gen_coax_mats_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _vp,
                                                 ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "fmb_coax_scratch_bytes": (ctypes.c_size_t, []),
    "fmb_reduce_rot_mat": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int]),
    "fmb_reduce_coax_mat": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int]),
    "fmb_plan_params_default": (None, [ctypes.POINTER(PlanParams)]),
    "fmb_plan_create": (ctypes.c_int, [_vp, ctypes.POINTER(PlanParams), _vp, _vp, ctypes.POINTER(_vp)]),
    "fmb_plan_orders": (ctypes.c_int, [_vp, _ip]),
    "fmb_plan_memory": (ctypes.c_int, [_vp, _dp]),
    "fmb_matvec": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "fmb_matvec_parts": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]),
    "fmb_rhs_plane_wave": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _vp, _vp]),
    "fmb_field_eval": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "fmb_field_eval_fmm": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "fmb_plan_retune": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp]),
    "fmb_matvec_transpose": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                            ctypes.c_int, _vp]),
    "fmb_gmres_transpose": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, _vp, _vp]),
    "fmb_adjoint_grad": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp,
                                        _vp, _vp, _vp, _vp, _vp]),
    "fmb_tree_vertex_normals": (ctypes.c_int, [_vp, _vp]),
    "fmb_matvec_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _vp, _vp, ctypes.c_int, _vp]),
    "fmb_gmres_batch": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp]),
    "fmb_plan_tile_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           _vp, _vp, _vp, _vp, _vp, _vp]),
    "fmb_plan_wavenumber": (ctypes.c_int, [_vp, _vp]),
    "fmb_sweep_guess": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "fmb_gmres": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.POINTER(GmresInfo), _vp]),
    "fmb_gmres_refined": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(GmresInfo), _vp]),
    "fmb_gmres_refined_transpose": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int,
                                                   ctypes.POINTER(GmresInfo), _vp]),
    "fmb_gmres_batch_refined": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                               _vp, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, _vp, _vp]),
    "fmb_adjoint_grad_refined": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_int,
                                                _vp, _vp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_int, _vp, _vp, _vp,
                                                _vp, _vp, _vp]),
    "fmb_plan_set_work_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fmb_plan_destroy": (None, [_vp]),
    "fmb_last_timing": (ctypes.c_int, [_vp, _dp, _dp]),
    "fmb_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
}


def lib():
    """Load libfmbem.so (built in-tree by fmbem_gpu_c_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FmbError(-3, "libfmbem.so not built: run python -m fmbem_gpu_c_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise FmbError(rc, lib().fmb_last_error().decode())
    return rc


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _dev(t, dtype=None):
    import torch
    dtype = torch.complex128 if dtype is None else dtype
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype
            and t.is_contiguous()):
        raise FmbError(-1, "expected a contiguous %s CUDA tensor" % str(dtype).split(".")[-1])
    return ctypes.c_void_p(t.data_ptr())


def _vecs(t, n, name, rows=None, like=None):
    """Checks a tensor of right-hand-side vectors before its pointer reaches
    the C ABI, which cannot see sizes: shape (n,) or (r, n) with r >= 1 (r
    equal to rows when given, and the shape of `like` when given).  Returns
    the number of vectors r.  Raised as FmbError(-1), before any device
    access, so a mis-sized tensor never becomes an out-of-bounds kernel
    access."""
    shape = tuple(getattr(t, "shape", ()))
    if len(shape) not in (1, 2) or shape[-1] != n or (len(shape) == 2 and shape[0] < 1):
        raise FmbError(-1, "%s: expected shape (%d,) or (r, %d), got %s" % (name, n, n, shape))
    r = 1 if len(shape) == 1 else shape[0]
    if rows is not None and r != rows:
        raise FmbError(-1, "%s: expected %d vectors, got %d" % (name, rows, r))
    if like is not None and shape != tuple(like.shape):
        raise FmbError(-1, "%s: shape %s differs from %s" % (name, shape, tuple(like.shape)))
    return r


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Mesh:
    """Triangle mesh (fmb_mesh)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_obj(cls, path):
        h = _vp()
        _check(lib().fmb_mesh_load_obj(path.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, V, F):
        V = np.ascontiguousarray(V, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.int32)
        h = _vp()
        _check(lib().fmb_mesh_from_arrays(_ptr(V, ctypes.c_double), V.shape[0],
                                          _ptr(F, ctypes.c_int), F.shape[0], ctypes.byref(h)))
        return cls(h)

    @classmethod
    def icosphere(cls, level, radius):
        h = _vp()
        _check(lib().fmb_mesh_icosphere(level, radius, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def ellipsoid(cls, nu, a, b, c, lmin=2, lmax=8, rms=0.05, seed=0):
        h = _vp()
        _check(lib().fmb_mesh_ellipsoid(nu, a, b, c, lmin, lmax, rms, seed, ctypes.byref(h)))
        return cls(h)

    def orient_outward(self):
        n = ctypes.c_int()
        _check(lib().fmb_mesh_orient_outward(self.h, ctypes.byref(n)))
        return n.value

    def arrays(self):
        nv, nf = ctypes.c_int(), ctypes.c_int()
        _check(lib().fmb_mesh_sizes(self.h, ctypes.byref(nv), ctypes.byref(nf)))
        V = np.zeros((nv.value, 3))
        F = np.zeros((nf.value, 3), dtype=np.int32)
        _check(lib().fmb_mesh_export(self.h, _ptr(V, ctypes.c_double), _ptr(F, ctypes.c_int)))
        return V, F

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fmb_mesh_destroy(self.h)
            self.h = None


class Tree:
    """Device Morton octree (fmb_tree)."""

    def __init__(self, mesh, s=5, depth=0):
        self.mesh = mesh
        nv, nf = ctypes.c_int(), ctypes.c_int()
        _check(lib().fmb_mesh_sizes(mesh.h, ctypes.byref(nv), ctypes.byref(nf)))
        self.n, self.nvert = nf.value, nv.value
        self.h = _vp()
        _check(lib().fmb_tree_build(mesh.h, s, depth, _stream(), ctypes.byref(self.h)))

    def info(self):
        lmin, lmax = ctypes.c_int(), ctypes.c_int()
        D = ctypes.c_double()
        pmin = np.zeros(3)
        _check(lib().fmb_tree_info(self.h, ctypes.byref(lmin), ctypes.byref(lmax),
                                   ctypes.byref(D), _ptr(pmin, ctypes.c_double)))
        return lmin.value, lmax.value, D.value, pmin

    def level(self, l):
        n = ctypes.c_int()
        _check(lib().fmb_tree_level(self.h, l, ctypes.byref(n), None))
        keys = np.zeros(n.value, dtype=np.int32)
        _check(lib().fmb_tree_level(self.h, l, ctypes.byref(n), _ptr(keys, ctypes.c_int)))
        return keys

    def num_sr(self, l):
        num = np.zeros(len(self.level(l)), dtype=np.int32)
        _check(lib().fmb_tree_num_sr(self.h, l, _ptr(num, ctypes.c_int)))
        return num

    def vertex_normals(self):
        """Unit area-weighted vertex normals (nvert x 3)."""
        nv = self.nvert
        out = np.zeros((nv, 3), dtype=np.float64)
        _check(lib().fmb_tree_vertex_normals(self.h, _ptr(out, ctypes.c_double)))
        return out

    def elem_box(self):
        n = self.n
        box = np.zeros(n, dtype=np.int32)
        perm = np.zeros(n, dtype=np.int32)
        _check(lib().fmb_tree_elem_box(self.h, _ptr(box, ctypes.c_int), _ptr(perm, ctypes.c_int)))
        return box, perm

    def pairs(self, l):
        n = ctypes.c_int()
        _check(lib().fmb_tree_pairs(self.h, l, ctypes.byref(n), None, None, None))
        t = np.zeros(n.value, dtype=np.int32)
        s = np.zeros(n.value, dtype=np.int32)
        o = np.zeros(n.value, dtype=np.int32)
        _check(lib().fmb_tree_pairs(self.h, l, ctypes.byref(n), _ptr(t, ctypes.c_int),
                                    _ptr(s, ctypes.c_int), _ptr(o, ctypes.c_int)))
        return t, s, o

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fmb_tree_destroy(self.h)
            self.h = None


def plan_params(k, **kw):
    p = PlanParams()
    lib().fmb_plan_params_default(ctypes.byref(p))
    p.k = k
    names = {f[0] for f in PlanParams._fields_}
    for key, v in kw.items():
        if key not in names:
            raise FmbError(-1, "unknown plan parameter %r" % key)
        if key == "p_levels":
            for i, q in enumerate(v):
                p.p_levels[i] = int(q)
        else:
            setattr(p, key, v)
    return p


class Plan:
    """Per-frequency FMM-BEM plan (fmb_plan)."""

    def __init__(self, tree, k, robin=None, **kw):
        self.tree = tree
        self.n = tree.n
        self.params = plan_params(k, **kw)
        self.nrhs_max = self.params.nrhs_max
        self.h = _vp()
        if robin is not None:
            _vecs(robin, self.n, "robin", rows=1)
        rb = _dev(robin) if robin is not None else None
        _check(lib().fmb_plan_create(tree.h, ctypes.byref(self.params), rb, _stream(),
                                     ctypes.byref(self.h)))

    def retune(self, k, robin=None):
        """Re-targets the plan at wavenumber k (fmb_plan_retune); robin as
        for the constructor, at the new frequency."""
        if robin is not None:
            _vecs(robin, self.n, "robin", rows=1)
        rb = _dev(robin) if robin is not None else None
        _check(lib().fmb_plan_retune(self.h, float(k), rb, _stream()))

    @property
    def k(self):
        v = ctypes.c_double()
        _check(lib().fmb_plan_wavenumber(self.h, ctypes.byref(v)))
        return v.value

    def sweep_guess(self, k_from, x_from, dirs=None):
        """Initial guess at the plan's wavenumber from the solution x_from at
        k_from, carrying each right-hand side's incident phase."""
        import torch
        nrhs = self._nrhs(x_from, "x_from")
        d = None
        if dirs is not None:
            d = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
            if d.shape[0] != nrhs:
                raise FmbError(-1, "one incident direction per right-hand side")
        x0 = torch.empty_like(x_from)
        _check(lib().fmb_sweep_guess(self.h, float(k_from), _dev(x_from),
                                     ctypes.c_void_p(d.ctypes.data) if d is not None else None,
                                     nrhs, _dev(x0), _stream()))
        return x0

    def tile_shape(self, kind, level, prec=FMB_FP64):
        """(mode, nt, mt, waves per workgroup, tuned ms) of a matrix-core
        translation (fmb_plan_tile_shape; kind 0 M2L, 1 M2M, 2 L2L)."""
        r = self.tile_shape_bf(kind, level, prec)
        return r[:4] + (r[5],)

    def tile_shape_bf(self, kind, level, prec=FMB_FP64):
        """(mode, nt, mt, waves per workgroup, split-BF16 rotations, tuned ms)."""
        v = [ctypes.c_int() for _ in range(5)]
        ms = ctypes.c_double()
        _check(lib().fmb_plan_tile_shape(self.h, kind, level, prec, ctypes.byref(v[0]),
                                         ctypes.byref(v[1]), ctypes.byref(v[2]),
                                         ctypes.byref(v[3]), ctypes.byref(v[4]),
                                         ctypes.byref(ms)))
        return tuple(x.value for x in v) + (ms.value,)

    def _nrhs(self, t, name, like=None):
        """Vectors in t (checked against the plan's element count and
        nrhs_max)."""
        r = _vecs(t, self.n, name, like=like)
        if r > self.nrhs_max:
            raise FmbError(-1, "%s: %d vectors above the plan's nrhs_max %d"
                           % (name, r, self.nrhs_max))
        return r

    def orders(self):
        p = np.zeros(16, dtype=np.int32)
        _check(lib().fmb_plan_orders(self.h, _ptr(p, ctypes.c_int)))
        return p

    def memory(self):
        b = ctypes.c_double()
        _check(lib().fmb_plan_memory(self.h, ctypes.byref(b)))
        return b.value

    def matvec(self, x, y=None, prec=FMB_FP64):
        import torch
        nrhs = self._nrhs(x, "x")
        if y is None:
            y = torch.empty_like(x)
        _vecs(y, self.n, "y", like=x)
        _check(lib().fmb_matvec(self.h, prec, _dev(x), _dev(y), nrhs, _stream()))
        return y

    def matvec_parts(self, x, parts, prec=FMB_FP64):
        import torch
        nrhs = self._nrhs(x, "x")
        y = torch.empty_like(x)
        _check(lib().fmb_matvec_parts(self.h, prec, parts, _dev(x), _dev(y), nrhs, _stream()))
        return y

    def rhs_plane_wave(self, dirs):
        import torch
        dirs = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
        if dirs.ndim != 2 or dirs.shape[1] != 3 or not 1 <= dirs.shape[0] <= self.nrhs_max:
            raise FmbError(-1, "dirs: expected 1..%d unit directions (r, 3), got %s"
                           % (self.nrhs_max, dirs.shape))
        n = self.n
        b = torch.empty((dirs.shape[0], n), dtype=torch.complex128, device="cuda")
        _check(lib().fmb_rhs_plane_wave(self.h, _ptr(dirs, ctypes.c_double), dirs.shape[0],
                                        _dev(b), _stream()))
        return b

    def field(self, x, pts, dirs=None, method="direct"):
        """Pressure at field points pts (M x 3) from the surface solution x
        (complex128 CUDA tensor, [nrhs][N] or [N]); dirs: incident unit
        plane waves (nrhs x 3) added to the scattered field, or None.
        method: "direct" (all panels, fmb_field_eval) or "fmm" (multipole
        expansions of the well-separated boxes, fmb_field_eval_fmm).
        Returns a [nrhs][M] complex128 CUDA tensor."""
        import torch
        if method not in ("direct", "fmm"):
            raise FmbError(-1, "field method is 'direct' or 'fmm'")
        nrhs = _vecs(x, self.n, "x")
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        d = None
        if dirs is not None:
            d = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
            if d.shape[0] != nrhs:
                raise FmbError(-1, "one incident direction per right-hand side")
        out = torch.empty((nrhs, pts.shape[0]), dtype=torch.complex128, device="cuda")
        fn = lib().fmb_field_eval if method == "direct" else lib().fmb_field_eval_fmm
        _check(fn(self.h, _dev(x), nrhs, ctypes.c_void_p(pts.ctypes.data), pts.shape[0],
                  ctypes.c_void_p(d.ctypes.data) if d is not None else None,
                  _dev(out), _stream()))
        return out

    def gmres(self, b, x0=None, restart=30, tol=1e-6, maxit=1000, check=True):
        import torch
        nrhs = self._nrhs(b, "b")
        if x0 is not None:
            _vecs(x0, self.n, "x0", like=b)
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        info = GmresInfo()
        rc = lib().fmb_gmres(self.h, _dev(b), _dev(x), nrhs, restart, tol, maxit,
                             ctypes.byref(info), _stream())
        if check:
            _check(rc)
        return x, info

    def matvec_t(self, x, parts=3, prec=FMB_FP64):
        """y = A^T x (fmb_matvec_transpose)."""
        import torch
        nrhs = self._nrhs(x, "x")
        y = torch.empty_like(x)
        _check(lib().fmb_matvec_transpose(self.h, prec, parts, _dev(x), _dev(y), nrhs,
                                          _stream()))
        return y

    def gmres_refined(self, inner, b, x0=None, restart=30, tol=1e-6, inner_tol=1e-3,
                      maxit=1000, check=True):
        """Solves A x = b to tol on this plan's FP64 pass by iterative
        refinement around GMRES on `inner`, a plan of the same tree,
        wavenumber and boundary condition, typically with a looser
        truncation (fmb_gmres_refined)."""
        return self._refined(lib().fmb_gmres_refined, inner, b, x0, restart, tol, inner_tol,
                             maxit, check)

    def gmres_refined_t(self, inner, b, x0=None, restart=30, tol=1e-6, inner_tol=1e-3,
                        maxit=1000, check=True):
        """gmres_refined on A^T x = b (fmb_gmres_refined_transpose)."""
        return self._refined(lib().fmb_gmres_refined_transpose, inner, b, x0, restart, tol,
                             inner_tol, maxit, check)

    def _refined(self, fn, inner, b, x0, restart, tol, inner_tol, maxit, check):
        import torch
        if inner.n != self.n:
            raise FmbError(-1, "inner plan of %d elements, outer of %d" % (inner.n, self.n))
        nrhs = self._nrhs(b, "b")
        inner._nrhs(b, "b")
        if x0 is not None:
            _vecs(x0, self.n, "x0", like=b)
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        info = GmresInfo()
        rc = fn(self.h, inner.h, _dev(b), _dev(x), nrhs, restart, tol, inner_tol, maxit,
                ctypes.byref(info), _stream())
        if check:
            _check(rc)
        return x, info

    def gmres_t(self, b, x0=None, restart=30, tol=1e-6, maxit=1000, check=True):
        """Solves A^T x = b (fmb_gmres_transpose)."""
        import torch
        nrhs = self._nrhs(b, "b")
        if x0 is not None:
            _vecs(x0, self.n, "x0", like=b)
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        info = GmresInfo()
        rc = lib().fmb_gmres_transpose(self.h, _dev(b), _dev(x), nrhs, restart, tol, maxit,
                                       ctypes.byref(info), _stream())
        if check:
            _check(rc)
        return x, info

    def adjoint_grad(self, x, pts, dirs=None, restart=30, tol=1e-10, maxit=1000,
                     vertices=False, inner=None, inner_tol=1e-3):
        """J = sum |p|^2 at pts (M x 3) for the surface solution x and its
        gradient dJ/dRe sigma + i dJ/dIm sigma per element (complex128 CUDA
        tensor, caller order) with respect to the Robin coefficients
        (fmb_adjoint_grad).  Returns (J, grad, info of the adjoint solve), and
        with vertices=True also dJ/dV (nvert x 3) and dJ/d(vertex normal
        displacement) (nvert) as float64 CUDA tensors.  With an inner plan the
        adjoint solve refines around GMRES on it (fmb_adjoint_grad_refined)."""
        import torch
        nrhs = self._nrhs(x, "x")
        if inner is not None and inner.n != self.n:
            raise FmbError(-1, "inner plan of %d elements, outer of %d" % (inner.n, self.n))
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        d = None
        if dirs is not None:
            d = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
            if d.shape[0] != nrhs:
                raise FmbError(-1, "one incident direction per right-hand side")
        n = self.n
        g = torch.empty(n, dtype=torch.complex128, device=x.device)
        dV = dvn = None
        if vertices:
            nv = self.tree.nvert
            dV = torch.empty((nv, 3), dtype=torch.float64, device=x.device)
            dvn = torch.empty(nv, dtype=torch.float64, device=x.device)
        J = ctypes.c_double()
        info = GmresInfo()
        rest = (_dev(x), nrhs, ctypes.c_void_p(d.ctypes.data) if d is not None else None,
                ctypes.c_void_p(pts.ctypes.data), pts.shape[0], restart, tol, maxit,
                ctypes.byref(J), _dev(g), _dev(dV, torch.float64) if vertices else None,
                _dev(dvn, torch.float64) if vertices else None, ctypes.byref(info), _stream())
        if inner is None:
            _check(lib().fmb_adjoint_grad(self.h, *rest))
        else:
            _check(lib().fmb_adjoint_grad_refined(self.h, inner.h, inner_tol, *rest))
        if vertices:
            return J.value, g, info, dV, dvn
        return J.value, g, info

    def set_work_precision(self, prec):
        _check(lib().fmb_plan_set_work_precision(self.h, prec))

    def set_timing(self, on=True):
        _check(lib().fmb_set_timing(self.h, int(on)))

    def timing(self):
        ph = np.zeros(7)
        m2l = np.zeros(16)
        _check(lib().fmb_last_timing(self.h, _ptr(ph, ctypes.c_double), _ptr(m2l, ctypes.c_double)))
        return ph, m2l

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fmb_plan_destroy(self.h)
            self.h = None


def trunc_num(k, eps, sigma, a):
    return lib().fmb_trunc_num(k, eps, sigma, a)


def gen_rot_mats(angles, p):
    angles = np.ascontiguousarray(np.atleast_2d(angles), dtype=np.float64)
    n = angles.shape[0]
    out = np.zeros((n, lib().fmb_rot_size(p)), dtype=np.complex128)
    _check(lib().fmb_gen_rot_mats(_ptr(angles, ctypes.c_double), n, p,
                                  out.ctypes.data_as(ctypes.c_void_p)))
    return out


def gen_coax_mats(kind, k, t, p):
    t = np.ascontiguousarray(np.atleast_1d(t), dtype=np.float64)
    out = np.zeros((t.shape[0], lib().fmb_coax_size(p)), dtype=np.complex128)
    _check(lib().fmb_gen_coax_mats(kind, k, _ptr(t, ctypes.c_double), t.shape[0], p,
                                   out.ctypes.data_as(ctypes.c_void_p)))
    return out


def _batch_rows(plans, nrhs, t, name):
    """Checks the columns of a batch: one count per plan within its
    nrhs_max, every plan on a mesh of the same element count, and t of
    shape (sum(nrhs), N)."""
    nrhs = [int(r) for r in nrhs]
    if len(plans) < 1 or len(nrhs) != len(plans):
        raise FmbError(-1, "one right-hand-side count per plan")
    n = plans[0].n
    for p, r in zip(plans, nrhs):
        if p.n != n:
            raise FmbError(-1, "batch plans on meshes of %d and %d elements" % (n, p.n))
        if not 1 <= r <= p.nrhs_max:
            raise FmbError(-1, "%d right-hand sides for a plan with nrhs_max %d" % (r, p.nrhs_max))
    tot = sum(nrhs)
    shape = tuple(getattr(t, "shape", ()))
    if shape != (tot, n) and not (tot == 1 and shape == (n,)):
        raise FmbError(-1, "%s: expected shape (%d, %d), got %s" % (name, tot, n, shape))
    return nrhs


def gen_coax_mats_noalloc(kind, k, t, p, out=None):
    """fmb_gen_coax_mats_noalloc: coaxial matrices of order p for the
    distances t (a float64 CUDA tensor) into out (complex128 CUDA tensor
    [len(t)][coax_size(p)], created when None) on the current stream, with
    no device allocation inside the library."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.dim() == 1 and t.numel() >= 1):
        raise FmbError(-1, "t: expected a 1-D tensor of distances")
    n, cs = t.numel(), lib().fmb_coax_size(p)
    if out is None:
        out = torch.empty((n, cs), dtype=torch.complex128, device=t.device)
    if tuple(out.shape) != (n, cs):
        raise FmbError(-1, "out: expected shape (%d, %d), got %s" % (n, cs, tuple(out.shape)))
    _check(lib().fmb_gen_coax_mats_noalloc(kind, float(k), _dev(t, torch.float64), n, p,
                                           _dev(out), _stream()))
    return out


def reduce_rot_mat(mat, p1, p2):
    """A rotation matrix of order p1 (reference layout) truncated to order p2
    (fmb_reduce_rot_mat)."""
    a = np.ascontiguousarray(mat, dtype=np.complex128)
    if a.shape != (lib().fmb_rot_size(p1),):
        raise FmbError(-1, "expected %d values of an order-%d rotation matrix"
                       % (lib().fmb_rot_size(p1), p1))
    out = np.zeros(lib().fmb_rot_size(max(p2, 1)), dtype=np.complex128)
    _check(lib().fmb_reduce_rot_mat(a.ctypes.data_as(ctypes.c_void_p), p1,
                                    out.ctypes.data_as(ctypes.c_void_p), p2))
    return out


def reduce_coax_mat(mat, p1, p2):
    """A coaxial matrix of order p1 (|m|-block layout) truncated to order p2
    (fmb_reduce_coax_mat)."""
    a = np.ascontiguousarray(mat, dtype=np.complex128)
    if a.shape != (lib().fmb_coax_size(p1),):
        raise FmbError(-1, "expected %d values of an order-%d coaxial matrix"
                       % (lib().fmb_coax_size(p1), p1))
    out = np.zeros(lib().fmb_coax_size(max(p2, 1)), dtype=np.complex128)
    _check(lib().fmb_reduce_coax_mat(a.ctypes.data_as(ctypes.c_void_p), p1,
                                     out.ctypes.data_as(ctypes.c_void_p), p2))
    return out


def matvec_batch(plans, nrhs, x, graph=True, prec=FMB_FP64, reps=1):
    """fmb_matvec_batch: every plan applied to its columns of x
    ([sum(nrhs)][N]); returns [reps][sum(nrhs)][N] (one slice per replay)."""
    import torch
    nrhs = _batch_rows(plans, nrhs, x, "x")
    if reps < 1:
        raise FmbError(-1, "reps must be at least 1")
    arr = (_vp * len(plans))(*[p.h for p in plans])
    nr = (ctypes.c_int * len(plans))(*nrhs)
    y = torch.empty((reps,) + tuple(x.shape), dtype=x.dtype, device=x.device)
    _check(lib().fmb_matvec_batch(arr, nr, len(plans), int(graph), prec, _dev(x), _dev(y),
                                  reps, _stream()))
    return y


def gmres_batch(plans, nrhs, b, x0=None, graph=True, restart=30, tol=1e-6, maxit=1000,
                check=True):
    """fmb_gmres_batch: GMRES over several plans (frequencies) in lock step.
    b: complex128 CUDA tensor [sum(nrhs)][N], plan 0's columns first.
    Returns (x, list of GmresInfo, one per plan)."""
    import torch
    nrhs = _batch_rows(plans, nrhs, b, "b")
    if x0 is not None:
        _batch_rows(plans, nrhs, x0, "x0")
    arr = (_vp * len(plans))(*[p.h for p in plans])
    nr = (ctypes.c_int * len(plans))(*nrhs)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    infos = (GmresInfo * len(plans))()
    rc = lib().fmb_gmres_batch(arr, nr, len(plans), int(graph), _dev(b), _dev(x), restart,
                               tol, maxit, infos, _stream())
    if check:
        _check(rc)
    return x, list(infos)


def gmres_batch_refined(outers, inners, nrhs, b, x0=None, graph=True, restart=30, tol=1e-6,
                        inner_tol=1e-3, maxit=1000, check=True):
    """fmb_gmres_batch_refined: refinement on each outer plan's FP64 residual
    around GMRES on the inner plans in lock step (inner p on outer p's tree,
    wavenumber and boundary condition).  b and the result as gmres_batch."""
    import torch
    if len(inners) != len(outers):
        raise FmbError(-1, "one inner plan per outer plan")
    nrhs = _batch_rows(outers, nrhs, b, "b")
    _batch_rows(inners, nrhs, b, "b")
    if x0 is not None:
        _batch_rows(outers, nrhs, x0, "x0")
    no = (_vp * len(outers))(*[p.h for p in outers])
    ni = (_vp * len(inners))(*[p.h for p in inners])
    nr = (ctypes.c_int * len(outers))(*nrhs)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    infos = (GmresInfo * len(outers))()
    rc = lib().fmb_gmres_batch_refined(no, ni, nr, len(outers), int(graph), _dev(b), _dev(x),
                                       restart, tol, inner_tol, maxit, infos, _stream())
    if check:
        _check(rc)
    return x, list(infos)


def _admittance_function():
    import torch

    class AdmittanceObjective(torch.autograd.Function):
        """J(Y) = sum_r sum_q |p_r(y_q)|^2 for plane waves dirs on a surface with
        per-element admittance Y (Robin coefficient sigma = i k Y), solved on
        `plan` (re-targeted at its own wavenumber with the new sigma).
        backward returns dJ/dRe Y + i dJ/dIm Y from fmb_adjoint_grad, the
        convention torch uses for a real loss of a complex tensor."""

        @staticmethod
        def forward(ctx, Y, plan, pts, dirs, tol):
            k = plan.k
            plan.retune(k, robin=(1j * k) * Y.detach())
            b = plan.rhs_plane_wave(dirs)
            x, _ = plan.gmres(b, tol=tol)
            J, gs, _ = plan.adjoint_grad(x, pts, dirs, tol=tol)
            ctx.save_for_backward((-1j * k) * gs)
            return torch.tensor(J, dtype=torch.float64, device=Y.device)

        @staticmethod
        def backward(ctx, gJ):
            (gY,) = ctx.saved_tensors
            return gJ * gY, None, None, None, None

    return AdmittanceObjective


def admittance_objective(Y, plan, pts, dirs, tol=1e-10):
    """Differentiable J(Y) (torch.autograd) through the adjoint FMM-BEM."""
    _vecs(Y, plan.n, "Y", rows=1)
    return _admittance_function().apply(Y, plan, pts, dirs, tol)
