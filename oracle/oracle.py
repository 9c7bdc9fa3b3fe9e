"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP path and never
as the thing measured or shipped.  The C sources state which reference lines
each function restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

BC_PERIODIC = 0
BC_DIRICHLET = 1
INIT_SIN = 0
INIT_REF_QUIRK = 1
TREE_GPU = 0
TREE_SEQUENTIAL = 1

OK, EINVAL, ENOMEM, ENOTCONV, EBREAKDOWN = 0, 1, 2, 5, 6

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class _Csr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("rowptr", _ip), ("col", _ip), ("val", _dp),
                ("base", ctypes.c_int)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_double, c_u64, c_i64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64
        L.orc_jacobi_init2d_f64.argtypes = [_dp, c_int, c_int, c_int]
        L.orc_jacobi_init3d_f64.argtypes = [_dp, c_int, c_int, c_int, c_int]
        L.orc_jacobi_init2d_f32.argtypes = [_fp, c_int, c_int, c_int]
        L.orc_set_threads.argtypes = [c_int]
        L.orc_set_threads.restype = None
        L.orc_get_threads.argtypes = []
        L.orc_get_threads.restype = c_int
        L.orc_jacobi_init3d_f32.argtypes = [_fp, c_int, c_int, c_int, c_int]
        for nm, p in (("orc_jacobi2d_f64", _dp), ("orc_jacobi2d_f32", _fp)):
            getattr(L, nm).argtypes = [p, p, c_int, c_int, c_int, c_int, c_int, _dp, c_int]
            getattr(L, nm).restype = c_int
        for nm, p in (("orc_jacobi3d_f64", _dp), ("orc_jacobi3d_f32", _fp)):
            getattr(L, nm).argtypes = [p, p, c_int, c_int, c_int, c_int, c_int, c_int, _dp, c_int]
            getattr(L, nm).restype = c_int
        L.orc_dot.argtypes = [_dp, _dp, c_int, c_int]
        L.orc_dot.restype = c_double
        L.orc_dot_f32.argtypes = [_fp, _fp, c_int, c_int]
        L.orc_dot_f32.restype = c_double
        L.orc_spmv.argtypes = [c_int, _ip, _ip, _dp, c_int, _dp, _dp]
        L.orc_cg_standard.argtypes = [ctypes.POINTER(_Csr), _dp, _dp, c_double, c_int, c_int, _dp, _ip]
        L.orc_cg_pipelined.argtypes = [ctypes.POINTER(_Csr), _dp, _dp, c_double, c_int, c_int, c_int,
                                       _dp, _ip]
        L.orc_cg_ir.argtypes = [ctypes.POINTER(_Csr), _dp, _dp, c_double, c_int, c_double, c_int, _dp,
                                _ip, _ip]
        L.orc_jacobi_smooth.argtypes = [ctypes.POINTER(_Csr), _dp, _dp, c_int, c_double]
        L.orc_splitmix64_next.argtypes = [ctypes.POINTER(c_u64)]
        L.orc_splitmix64_next.restype = c_u64
        L.orc_stream_state.argtypes = [c_u64, c_u64]
        L.orc_stream_state.restype = c_u64
        L.orc_uniform.argtypes = [c_u64, c_u64, _dp, c_i64]
        L.orc_poisson3d.argtypes = [c_int, c_int, c_int, _ip, _ip, _dp]
        L.orc_poisson3d.restype = c_i64
        L.orc_varcoef3d.argtypes = [c_u64, c_int, c_int, c_int, _ip, _ip, _ip, _dp, _dp]
        L.orc_varcoef3d.restype = c_i64
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def set_threads(n):
    """Threads of the oracle's OpenMP loops (returns the previous count)."""
    prev = lib().orc_get_threads()
    lib().orc_set_threads(int(n))
    return prev


# ------------------------------------------------------------------ Jacobi --

def init2d(nx, ny, kind=INIT_SIN, dtype=np.float64):
    """Initial 2D field; float32 uses the float-rounded pi of the reference's `real` PI."""
    a = np.empty((ny, nx), dtype=dtype)
    if a.dtype == np.float32:
        lib().orc_jacobi_init2d_f32(_ptr(a, _fp), nx, ny, kind)
    else:
        lib().orc_jacobi_init2d_f64(_ptr(a, _dp), nx, ny, kind)
    return a


def init3d(nx, ny, nz, kind=INIT_SIN, dtype=np.float64):
    a = np.empty((nz, ny, nx), dtype=dtype)
    if a.dtype == np.float32:
        lib().orc_jacobi_init3d_f32(_ptr(a, _fp), nx, ny, nz, kind)
    else:
        lib().orc_jacobi_init3d_f64(_ptr(a, _dp), nx, ny, nz, kind)
    return a


def jacobi(init, niter, bc=BC_DIRICHLET, norm_every=0):
    """Run the oracle sweep loop on a 2D (ny, nx) or 3D (nz, ny, nx) array of
    float64 or float32.  Returns (field, norms)."""
    init = np.ascontiguousarray(init)
    out = np.empty_like(init)
    cap = (niter + norm_every - 1) // norm_every if norm_every > 0 else 0
    norms = np.zeros(max(cap, 1), dtype=np.float64)
    f64 = init.dtype == np.float64
    if init.dtype not in (np.float64, np.float32):
        raise TypeError("jacobi oracle takes float64 or float32")
    p = _dp if f64 else _fp
    L = lib()
    if init.ndim == 2:
        ny, nx = init.shape
        fn = L.orc_jacobi2d_f64 if f64 else L.orc_jacobi2d_f32
        rc = fn(_ptr(init, p), _ptr(out, p), nx, ny, niter, bc, norm_every, _ptr(norms, _dp), cap)
    elif init.ndim == 3:
        nz, ny, nx = init.shape
        fn = L.orc_jacobi3d_f64 if f64 else L.orc_jacobi3d_f32
        rc = fn(_ptr(init, p), _ptr(out, p), nx, ny, nz, niter, bc, norm_every, _ptr(norms, _dp), cap)
    else:
        raise ValueError("2D or 3D field expected")
    if rc < 0:
        raise ValueError("oracle rejected the arguments")
    return out, norms[:cap]


# ---------------------------------------------------------------------- CG --

def dot(a, b, tree=TREE_GPU):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return lib().orc_dot_f32(_ptr(a, _fp), _ptr(b, _fp), a.size, tree)
    a = a.astype(np.float64, copy=False)
    b = b.astype(np.float64, copy=False)
    return lib().orc_dot(_ptr(a, _dp), _ptr(b, _dp), a.size, tree)


class Csr:
    """Host CSR matrix (int32 indices, float64 values)."""

    def __init__(self, rowptr, col, val, base=0):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.base = int(base)
        self.n = self.rowptr.size - 1

    def _c(self):
        return _Csr(self.n, _ptr(self.rowptr, _ip), _ptr(self.col, _ip), _ptr(self.val, _dp), self.base)

    @property
    def nnz(self):
        return int(self.rowptr[-1] - self.base)


def spmv(A, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(A.n, dtype=np.float64)
    lib().orc_spmv(A.n, _ptr(A.rowptr, _ip), _ptr(A.col, _ip), _ptr(A.val, _dp), A.base,
                   _ptr(x, _dp), _ptr(y, _dp))
    return y


def cg(A, b, x0=None, tol=1e-8, max_iter=10000, variant="standard", tree=TREE_GPU, quirk=True):
    """Returns (status, x, hist[:iters+1], iters)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max_iter + 1, dtype=np.float64)
    it = ctypes.c_int(0)
    c = A._c()
    if variant == "standard":
        st = lib().orc_cg_standard(ctypes.byref(c), _ptr(b, _dp), _ptr(x, _dp), tol, max_iter, tree,
                                   _ptr(hist, _dp), ctypes.byref(it))
    elif variant == "pipelined":
        st = lib().orc_cg_pipelined(ctypes.byref(c), _ptr(b, _dp), _ptr(x, _dp), tol, max_iter, tree,
                                    1 if quirk else 0, _ptr(hist, _dp), ctypes.byref(it))
    else:
        raise ValueError(variant)
    return st, x, hist[: it.value + 1], it.value


def cg_ir(A, b, x0=None, tol=1e-8, max_outer=50, inner_tol=1e-5, inner_max=10000):
    """FP32 inner CG inside an FP64 refinement loop.
    Returns (status, x, outer_hist, outer_iters, inner_total)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max_outer + 1, dtype=np.float64)
    oi, ii = ctypes.c_int(0), ctypes.c_int(0)
    c = A._c()
    st = lib().orc_cg_ir(ctypes.byref(c), _ptr(b, _dp), _ptr(x, _dp), tol, max_outer, inner_tol,
                         inner_max, _ptr(hist, _dp), ctypes.byref(oi), ctypes.byref(ii))
    return st, x, hist[: oi.value + 1], oi.value, ii.value


def jacobi_smooth(A, b, x, sweeps, omega=0.8):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x, dtype=np.float64, copy=True)
    c = A._c()
    lib().orc_jacobi_smooth(ctypes.byref(c), _ptr(b, _dp), _ptr(x, _dp), sweeps, omega)
    return x


# -------------------------------------------------------------- generators --

def splitmix64(state, count):
    st = ctypes.c_uint64(state)
    return [lib().orc_splitmix64_next(ctypes.byref(st)) for _ in range(count)]


def uniform(seed, stream_id, count):
    out = np.empty(count, dtype=np.float64)
    lib().orc_uniform(seed, stream_id, _ptr(out, _dp), count)
    return out


def poisson3d(nx, ny, nz):
    L = lib()
    nnz = L.orc_poisson3d(nx, ny, nz, None, None, None)
    n = nx * ny * nz
    rowptr = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    L.orc_poisson3d(nx, ny, nz, _ptr(rowptr, _ip), _ptr(col, _ip), _ptr(val, _dp))
    return Csr(rowptr, col, val)


def varcoef3d(seed, system_id, lmin, lmax):
    """Returns (Csr, f, dims)."""
    L = lib()
    dims = np.zeros(3, dtype=np.int32)
    nnz = L.orc_varcoef3d(seed, system_id, lmin, lmax, _ptr(dims, _ip), None, None, None, None)
    n = int(dims.prod())
    rowptr = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    f = np.empty(n, dtype=np.float64)
    L.orc_varcoef3d(seed, system_id, lmin, lmax, _ptr(dims, _ip), _ptr(rowptr, _ip), _ptr(col, _ip),
                    _ptr(val, _dp), _ptr(f, _dp))
    return Csr(rowptr, col, val), f, tuple(int(d) for d in dims)
