"""ctypes binding of libcfm.so, the MI355X-native CPU-free solver library.

Mirrors the reference's per-version entry points for the single-GPU path:

* ``jacobi(...)``   -- the stencil versions' ``init(argc, argv)``
  (Stencil/jacobi2D/src/main.cu:25-68, Stencil/jacobi3D/src/main.cu:31-72):
  allocate, initialise both arrays, run ``niter`` sweeps, copy back.
* ``cg(...)``       -- the CG versions' ``init(rowptr_d, col_d, val_d, n, nnz,
  zero_indexed, iter_max, x_final, ...)`` (CG/include/common.h:122-133).

Every call goes through the C ABI declared in include/cfm.h; there is no CPU
fallback: if the shared library is missing, importing the module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfm.so")

OK, EINVAL, ENOMEM, EDEVICE, ETIMEOUT, ENOTCONV, EBREAKDOWN, ENOTSUP = range(8)
F64, F32 = 0, 1
BC_PERIODIC, BC_DIRICHLET = 0, 1
LAUNCH_PER_ITER, PERSISTENT, NO_COMPUTE = 0, 1, 2
INIT_SIN, INIT_REF_QUIRK = 0, 1
CG_STANDARD, CG_PIPELINED, CG_IR_F32 = 0, 1, 2
CG_PERSISTENT, CG_DISCRETE = 0, 1

# Every symbol include/cfm.h declares.
EXPORTS = (
    "cfm_version", "cfm_status_str", "cfm_last_error", "cfm_device_count", "cfm_device_info", "cfm_device_synchronize",
    "cfm_bandwidth_probe", "cfm_bandwidth_probe_shape",
    "cfm_jacobi_plan_create", "cfm_jacobi_plan_load", "cfm_jacobi_plan_run", "cfm_jacobi_plan_download",
    "cfm_jacobi_plan_field", "cfm_jacobi_plan_destroy", "cfm_jacobi_run", "cfm_jacobi_init_field",
    "cfm_cg_plan_create", "cfm_cg_plan_set_rhs", "cfm_cg_plan_set_x0", "cfm_cg_plan_solve",
    "cfm_cg_plan_get_x", "cfm_cg_plan_destroy", "cfm_cg_solve", "cfm_gen_poisson3d", "cfm_gen_uniform",
    "cfm_gen_varcoef3d", "cfm_cg_batch_plan_create", "cfm_cg_batch_set_rhs", "cfm_cg_batch_set_x0",
    "cfm_cg_batch_solve", "cfm_cg_batch_get_x", "cfm_cg_batch_tune", "cfm_cg_batch_matrix_format", "cfm_cg_batch_plan_destroy",
    "cfm_unet_param_count", "cfm_unet_create", "cfm_unet_destroy", "cfm_unet_init_params", "cfm_unet_set_params",
    "cfm_unet_get_params", "cfm_unet_forward", "cfm_unet_gradient", "cfm_unet_train_step",
    "cfm_cg_batch_x0_rejected", "cfm_cg_batch_load_features", "cfm_cg_batch_warm_start", "cfm_gen_varcoef3d_coef",
    "cfm_cg_batch_smooth_x0", "cfm_gen_tridiag", "cfm_mm_read", "cfm_host_free", "cfm_jacobi_plan_phase_times",
    "cfm_cg_plan_phase_times", "cfm_jacobi_run_noalloc",
)
PHASES = ("spmv", "update", "reduce", "barrier", "other")


class CfmError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = last_error()
        super().__init__(f"{where}: {status_str(status)} ({status})" + (f": {msg}" if msg else ""))


class JacobiParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nx", "ny", "nz", "niter", "norm_every", "prec", "bc", "variant", "tile_rows", "grid_blocks",
        "chunks", "spin_limit_us", "block_threads", "time_block", "phase_timers")]


class Timing(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("grid_blocks", ctypes.c_int), ("block_threads", ctypes.c_int),
                ("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("cooperative", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CgParams(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("mode", ctypes.c_int), ("tol", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("restart_quirk", ctypes.c_int), ("max_outer", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("grid_blocks", ctypes.c_int),
                ("spin_limit_us", ctypes.c_int), ("block_threads", ctypes.c_int), ("x0_guard", ctypes.c_int),
                ("matrix_format", ctypes.c_int), ("phase_timers", ctypes.c_int)]


class PhaseTimes(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 5), ("max_ms", ctypes.c_double * 5), ("total_ms", ctypes.c_double),
                ("workgroups", ctypes.c_int), ("host_timed", ctypes.c_int)]

    def as_dict(self):
        return {"ms": dict(zip(PHASES, self.ms)), "max_ms": dict(zip(PHASES, self.max_ms)),
                "total_ms": self.total_ms, "workgroups": self.workgroups, "host_timed": bool(self.host_timed)}


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def _loaded_hip_runtimes():
    try:
        with open("/proc/self/maps") as f:
            return {ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln}
    except OSError:
        return set()


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C cpu-free-model_amd)")
    # libcfm binds to whichever libamdhip64.so.7 the process already holds (the
    # soname is shared).  A different runtime loaded first -- PyTorch wheels
    # bundle their own -- would run these kernels on a runtime they were not
    # built for, so refuse instead of running on it.
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    want = os.path.realpath(os.path.join(rocm, "lib", "libamdhip64.so.7"))
    other = {p for p in _loaded_hip_runtimes() if os.path.realpath(p) != want}
    if other and os.path.exists(want):
        raise ImportError(f"cfm: this process already loaded the HIP runtime {sorted(other)[0]} (for example by "
                          f"importing torch first); libcfm needs {want}.  Import cfm before torch.")
    L = ctypes.CDLL(LIB_PATH)
    L.cfm_version.restype = ctypes.c_char_p
    L.cfm_status_str.restype = ctypes.c_char_p
    L.cfm_status_str.argtypes = [ctypes.c_int]
    L.cfm_last_error.argtypes = []
    L.cfm_last_error.restype = ctypes.c_char_p
    L.cfm_device_count.argtypes = [_ip]
    L.cfm_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _ip,
                                  ctypes.POINTER(ctypes.c_size_t)]
    L.cfm_bandwidth_probe.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp, _dp]
    L.cfm_bandwidth_probe_shape.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _dp, _dp]
    L.cfm_jacobi_plan_create.argtypes = [ctypes.POINTER(JacobiParams), ctypes.POINTER(_vp)]
    L.cfm_jacobi_plan_load.argtypes = [_vp, _vp]
    L.cfm_jacobi_plan_run.argtypes = [_vp, _dp, ctypes.c_int, _ip, ctypes.POINTER(Timing)]
    L.cfm_jacobi_plan_download.argtypes = [_vp, _vp]
    L.cfm_jacobi_plan_field.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.cfm_jacobi_plan_destroy.argtypes = [_vp]
    L.cfm_jacobi_plan_destroy.restype = None
    L.cfm_jacobi_run.argtypes = [ctypes.POINTER(JacobiParams), _vp, _vp, _dp, ctypes.c_int, _ip,
                                 ctypes.POINTER(Timing)]
    L.cfm_jacobi_init_field.argtypes = [ctypes.c_int] * 5 + [_vp]
    L.cfm_jacobi_run_noalloc.argtypes = [ctypes.POINTER(JacobiParams), _vp, _vp, _vp, _dp, ctypes.c_int, _ip,
                                         ctypes.POINTER(Timing)]
    L.cfm_cg_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(CgParams), ctypes.POINTER(_vp)]
    L.cfm_cg_plan_set_rhs.argtypes = [_vp, _dp]
    L.cfm_cg_plan_set_x0.argtypes = [_vp, _dp]
    L.cfm_cg_plan_solve.argtypes = [_vp, _dp, ctypes.c_int, _ip, ctypes.POINTER(Timing)]
    L.cfm_cg_plan_get_x.argtypes = [_vp, _dp]
    L.cfm_cg_plan_destroy.argtypes = [_vp]
    L.cfm_cg_plan_destroy.restype = None
    L.cfm_cg_solve.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _dp, _dp,
                               ctypes.POINTER(CgParams), _dp, ctypes.c_int, _ip, ctypes.POINTER(Timing)]
    L.cfm_gen_poisson3d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp]
    L.cfm_gen_poisson3d.restype = ctypes.c_int64
    L.cfm_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _dp, ctypes.c_int64]
    L.cfm_gen_varcoef3d.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp,
                                    _dp]
    L.cfm_gen_varcoef3d.restype = ctypes.c_int64
    L.cfm_cg_batch_plan_create.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.POINTER(CgParams),
                                           ctypes.POINTER(_vp)]
    L.cfm_cg_batch_set_rhs.argtypes = [_vp, _dp]
    L.cfm_cg_batch_set_x0.argtypes = [_vp, _dp]
    L.cfm_cg_batch_solve.argtypes = [_vp, _ip, _ip, _dp, ctypes.c_int, ctypes.POINTER(Timing)]
    L.cfm_cg_batch_get_x.argtypes = [_vp, _dp]
    L.cfm_cg_batch_tune.argtypes = [_vp, ctypes.c_int, ctypes.c_char_p, _ip, _ip]
    L.cfm_cg_batch_matrix_format.argtypes = [_vp]
    L.cfm_cg_batch_matrix_format.restype = ctypes.c_int
    L.cfm_cg_batch_plan_destroy.argtypes = [_vp]
    L.cfm_cg_batch_plan_destroy.restype = None
    _fp = ctypes.POINTER(ctypes.c_float)
    L.cfm_unet_param_count.argtypes = []
    L.cfm_unet_param_count.restype = ctypes.c_size_t
    L.cfm_unet_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.cfm_unet_destroy.argtypes = [_vp]
    L.cfm_unet_destroy.restype = None
    L.cfm_unet_init_params.argtypes = [ctypes.c_uint64, _fp]
    L.cfm_unet_set_params.argtypes = [_vp, _fp]
    L.cfm_unet_get_params.argtypes = [_vp, _fp]
    L.cfm_unet_forward.argtypes = [_vp, _fp, _fp]
    L.cfm_unet_gradient.argtypes = [_vp, _fp, _fp, _fp, _dp]
    L.cfm_unet_train_step.argtypes = [_vp, _fp, _fp, ctypes.c_float, _dp]
    L.cfm_cg_batch_x0_rejected.argtypes = [_vp, _ip]
    L.cfm_cg_batch_load_features.argtypes = [_vp, _fp]
    L.cfm_cg_batch_warm_start.argtypes = [_vp, _vp, ctypes.c_double, ctypes.POINTER(Timing)]
    L.cfm_gen_varcoef3d_coef.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp]
    L.cfm_cg_batch_smooth_x0.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Timing)]
    L.cfm_jacobi_plan_phase_times.argtypes = [_vp, ctypes.POINTER(PhaseTimes)]
    L.cfm_cg_plan_phase_times.argtypes = [_vp, ctypes.POINTER(PhaseTimes)]
    L.cfm_gen_tridiag.argtypes = [ctypes.c_int, ctypes.c_uint, _vp, _vp, _vp]
    L.cfm_gen_tridiag.restype = ctypes.c_int64
    L.cfm_mm_read.argtypes = [ctypes.c_char_p, _ip, _ip, _ip, ctypes.POINTER(_ip), ctypes.POINTER(_ip),
                              ctypes.POINTER(_dp)]
    L.cfm_host_free.argtypes = [_vp]
    L.cfm_host_free.restype = None
    return L


lib = _load()


def status_str(s):
    return lib.cfm_status_str(int(s)).decode()


def last_error():
    m = lib.cfm_last_error()
    return m.decode() if m else ""


def version():
    return lib.cfm_version().decode()


def _check(st, where):
    if st != OK:
        raise CfmError(st, where)


def device_count():
    n = ctypes.c_int(0)
    st = lib.cfm_device_count(ctypes.byref(n))
    return n.value if st == OK else 0


def device_info(dev=0):
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int(0)
    mem = ctypes.c_size_t(0)
    _check(lib.cfm_device_info(dev, name, 256, ctypes.byref(cus), ctypes.byref(mem)), "cfm_device_info")
    return {"name": name.value.decode(), "compute_units": cus.value, "mem_bytes": mem.value}


def bandwidth_probe(nbytes=1 << 30, reps=5):
    """Best copy bandwidth (GB/s, read + write) of a plain streaming copy."""
    gbs, ms = ctypes.c_double(0), ctypes.c_double(0)
    _check(lib.cfm_bandwidth_probe(nbytes, reps, ctypes.byref(gbs), ctypes.byref(ms)), "cfm_bandwidth_probe")
    return gbs.value, ms.value


def bandwidth_probe_shape(kind, grid, unroll=1, nt=False, nbytes=1 << 30, reps=5):
    """Best bandwidth (GB/s) of one probe shape: kind "copy" (read + write bytes), "read" or
    "write"; grid workgroups of 256 lanes, unroll 16-byte vectors in flight per lane."""
    gbs, ms = ctypes.c_double(0), ctypes.c_double(0)
    k = {"copy": 0, "read": 1, "write": 2}[kind]
    _check(lib.cfm_bandwidth_probe_shape(nbytes, reps, k, grid, unroll, 1 if nt else 0, ctypes.byref(gbs),
                                         ctypes.byref(ms)), "cfm_bandwidth_probe_shape")
    return gbs.value, ms.value


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------- device plumbing --
# Device memory for callers that hand the library device pointers (the
# borrowed-CSR path, the no-allocation entry point): allocated through the same
# HIP runtime libcfm is linked against, never through another one.

_hip_rt = None


def hip_runtime():
    global _hip_rt
    if _hip_rt is None:
        rt = ctypes.CDLL("libamdhip64.so.7")  # the soname libcfm resolved: the same loaded runtime
        rt.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        rt.hipFree.argtypes = [_vp]
        rt.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        rt.hipDeviceSynchronize.argtypes = []
        for fn in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipDeviceSynchronize):
            fn.restype = ctypes.c_int
        _hip_rt = rt
    return _hip_rt


_H2D, _D2H = 1, 2


class DeviceBuffer:
    """A plain hipMalloc allocation of nbytes (ptr: its device address)."""

    def __init__(self, nbytes):
        rt = hip_runtime()
        self.nbytes = int(nbytes)
        self.ptr = _vp()
        if rt.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed")

    @classmethod
    def of(cls, array):
        array = np.ascontiguousarray(array)
        buf = cls(array.nbytes)
        buf.upload(array)
        return buf

    def upload(self, array):
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if hip_runtime().hipMemcpy(self.ptr, _ptr(array), array.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy host -> device failed")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("request larger than the buffer")
        if hip_runtime().hipMemcpy(_ptr(out), self.ptr, out.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy device -> host failed")
        return out

    def free(self):
        if self.ptr:
            hip_runtime().hipFree(self.ptr)
            self.ptr = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ Jacobi --

def init_field(nx, ny, nz=1, kind=INIT_SIN, prec=F64):
    shape = (ny, nx) if nz == 1 else (nz, ny, nx)
    a = np.empty(shape, dtype=np.float64 if prec == F64 else np.float32)
    _check(lib.cfm_jacobi_init_field(nx, ny, nz, kind, prec, _ptr(a)), "cfm_jacobi_init_field")
    return a


def jacobi_params(nx, ny, nz=1, niter=1000, norm_every=0, prec=F64, bc=BC_DIRICHLET,
                  variant=PERSISTENT, tile_rows=0, grid_blocks=0, chunks=0, spin_limit_us=0, block_threads=0,
                  time_block=0, phase_timers=0):
    return JacobiParams(nx, ny, nz, niter, norm_every, prec, bc, variant, tile_rows, grid_blocks, chunks,
                        spin_limit_us, block_threads, time_block, phase_timers)


class JacobiPlan:
    """Device-resident Jacobi state (two working arrays) for repeated runs."""

    def __init__(self, params):
        self.params = params
        self.dtype = np.float64 if params.prec == F64 else np.float32
        self.shape = (params.ny, params.nx) if params.nz == 1 else (params.nz, params.ny, params.nx)
        h = _vp()
        _check(lib.cfm_jacobi_plan_create(ctypes.byref(params), ctypes.byref(h)), "cfm_jacobi_plan_create")
        self._h = h

    def load(self, field):
        field = np.ascontiguousarray(field, dtype=self.dtype)
        if field.shape != self.shape:
            raise ValueError(f"field shape {field.shape} != {self.shape}")
        _check(lib.cfm_jacobi_plan_load(self._h, _ptr(field)), "cfm_jacobi_plan_load")

    def run(self):
        p = self.params
        cap = (p.niter + p.norm_every - 1) // p.norm_every if p.norm_every > 0 else 0
        norms = np.zeros(max(cap, 1))
        cnt = ctypes.c_int(0)
        t = Timing()
        _check(lib.cfm_jacobi_plan_run(self._h, norms.ctypes.data_as(_dp), cap, ctypes.byref(cnt),
                                       ctypes.byref(t)), "cfm_jacobi_plan_run")
        return norms[: cnt.value], t

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib.cfm_jacobi_plan_download(self._h, _ptr(out)), "cfm_jacobi_plan_download")
        return out

    def phase_times(self):
        pt = PhaseTimes()
        _check(lib.cfm_jacobi_plan_phase_times(self._h, ctypes.byref(pt)), "cfm_jacobi_plan_phase_times")
        return pt.as_dict()

    def close(self):
        if self._h:
            lib.cfm_jacobi_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def jacobi(init, niter, bc=BC_DIRICHLET, norm_every=0, variant=PERSISTENT, **tuning):
    """Run niter sweeps on a (ny, nx) or (nz, ny, nx) float64/float32 field.
    Returns (field, norms, timing_dict)."""
    init = np.ascontiguousarray(init)
    prec = F64 if init.dtype == np.float64 else F32
    if init.dtype not in (np.float64, np.float32):
        raise TypeError("float64 or float32 field expected")
    if init.ndim == 2:
        ny, nx = init.shape
        nz = 1
    else:
        nz, ny, nx = init.shape
    p = jacobi_params(nx, ny, nz, niter, norm_every, prec, bc, variant, **tuning)
    with JacobiPlan(p) as plan:
        plan.load(init)
        norms, t = plan.run()
        out = plan.download()
    return out, norms, t.as_dict()


def jacobi_noalloc(params, field, work, stream=None):
    """cfm_jacobi_run_noalloc on two device buffers (DeviceBuffer or raw
    device addresses) of params' field size: field holds the initial field
    and receives the result, work is overwritten.  Returns (norms, timing_dict)."""
    p = params
    cap = (p.niter + p.norm_every - 1) // p.norm_every if p.norm_every > 0 else 0
    norms = np.zeros(max(cap, 1))
    cnt = ctypes.c_int(0)
    t = Timing()
    fptr = field.ptr if isinstance(field, DeviceBuffer) else _vp(field)
    wptr = work.ptr if isinstance(work, DeviceBuffer) else _vp(work)
    _check(lib.cfm_jacobi_run_noalloc(ctypes.byref(p), fptr, wptr, _vp(stream), norms.ctypes.data_as(_dp), cap,
                                      ctypes.byref(cnt), ctypes.byref(t)), "cfm_jacobi_run_noalloc")
    return norms[: cnt.value], t.as_dict()


# ---------------------------------------------------------------------- CG --

def poisson3d(nx, ny, nz):
    nnz = lib.cfm_gen_poisson3d(nx, ny, nz, None, None, None)
    if nnz < 0:
        raise ValueError("grid too large")
    n = nx * ny * nz
    rowptr = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    lib.cfm_gen_poisson3d(nx, ny, nz, rowptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                          val.ctypes.data_as(_dp))
    return rowptr, col, val


def tridiag(n, seed=1):
    """The reference's default CG matrix (genTridiag): returns (rowptr, col, val), 0-based."""
    nnz = lib.cfm_gen_tridiag(n, seed, None, None, None)
    if nnz < 0:
        raise ValueError("n must be >= 2")
    rowptr = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    if lib.cfm_gen_tridiag(n, seed, _ptr(rowptr), _ptr(col), _ptr(val)) != nnz:
        raise RuntimeError("cfm_gen_tridiag failed")
    return rowptr, col, val


def mm_read(path):
    """The reference's Matrix Market reader: returns (rowptr, col, val, index_base).
    Raises CfmError with the reader's message for files the reference refuses."""
    n, nnz, base = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    rp, cp, vp = _ip(), _ip(), _dp()
    _check(lib.cfm_mm_read(os.fsencode(path), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(base),
                           ctypes.byref(rp), ctypes.byref(cp), ctypes.byref(vp)), "cfm_mm_read")
    try:
        rowptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
        col = np.ctypeslib.as_array(cp, shape=(max(nnz.value, 1),))[: nnz.value].copy()
        val = np.ctypeslib.as_array(vp, shape=(max(nnz.value, 1),))[: nnz.value].copy()
    finally:
        for q in (rp, cp, vp):
            lib.cfm_host_free(ctypes.cast(q, _vp))
    return rowptr, col, val, base.value


def uniform(seed, stream_id, count):
    out = np.empty(count, dtype=np.float64)
    _check(lib.cfm_gen_uniform(seed, stream_id, out.ctypes.data_as(_dp), count), "cfm_gen_uniform")
    return out


def cg_params(method=CG_STANDARD, mode=CG_PERSISTENT, tol=1e-8, max_iter=10000, restart_quirk=1,
              max_outer=0, inner_tol=0.0, grid_blocks=0, spin_limit_us=0, block_threads=0, x0_guard=0,
              matrix_format=0, phase_timers=0):
    return CgParams(method, mode, tol, max_iter, restart_quirk, max_outer, inner_tol, grid_blocks,
                    spin_limit_us, block_threads, x0_guard, matrix_format, phase_timers)


class CgPlan:
    """Device-resident CSR (stored as SELL-64 or DIA) plus CG work vectors.

    rowptr/col/val are host arrays, or DeviceBuffer objects holding a CSR the
    caller already uploaded (the reference's calling convention: main copies
    the CSR to the device and hands init() the device pointers,
    CG/src/main.cu:188-197); the library reads a borrowed CSR once."""

    def __init__(self, rowptr, col, val, params, index_base=0, n=None, nnz=None):
        self.params = params
        h = _vp()
        if isinstance(rowptr, DeviceBuffer):
            if n is None or nnz is None:
                raise ValueError("a device CSR needs n and nnz")
            self.n = int(n)
            st = lib.cfm_cg_plan_create(self.n, int(nnz), rowptr.ptr, col.ptr, val.ptr, index_base, 1,
                                        ctypes.byref(params), ctypes.byref(h))
        else:
            self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
            self.col = np.ascontiguousarray(col, dtype=np.int32)
            self.val = np.ascontiguousarray(val, dtype=np.float64)
            self.n = self.rowptr.size - 1
            st = lib.cfm_cg_plan_create(self.n, self.col.size, _ptr(self.rowptr), _ptr(self.col), _ptr(self.val),
                                        index_base, 0, ctypes.byref(params), ctypes.byref(h))
        _check(st, "cfm_cg_plan_create")
        self._h = h

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib.cfm_cg_plan_set_rhs(self._h, b.ctypes.data_as(_dp)), "cfm_cg_plan_set_rhs")

    def set_x0(self, x0=None):
        if x0 is None:
            _check(lib.cfm_cg_plan_set_x0(self._h, None), "cfm_cg_plan_set_x0")
        else:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            _check(lib.cfm_cg_plan_set_x0(self._h, x0.ctypes.data_as(_dp)), "cfm_cg_plan_set_x0")

    def solve(self, hist_cap=None):
        """Returns (status, hist, iters, timing). ENOTCONV/EBREAKDOWN are statuses, not exceptions."""
        p = self.params
        if hist_cap is None:
            hist_cap = (p.max_outer if p.method == CG_IR_F32 and p.max_outer > 0 else
                        50 if p.method == CG_IR_F32 else p.max_iter) + 1
        hist = np.zeros(max(hist_cap, 1))
        it = ctypes.c_int(0)
        t = Timing()
        st = lib.cfm_cg_plan_solve(self._h, hist.ctypes.data_as(_dp), hist_cap, ctypes.byref(it),
                                   ctypes.byref(t))
        if st not in (OK, ENOTCONV, EBREAKDOWN):
            raise CfmError(st, "cfm_cg_plan_solve")
        nh = (t.outer_iterations if p.method == CG_IR_F32 else it.value) + 1
        return st, hist[: min(nh, hist_cap)], it.value, t.as_dict()

    def x(self):
        out = np.empty(self.n)
        _check(lib.cfm_cg_plan_get_x(self._h, out.ctypes.data_as(_dp)), "cfm_cg_plan_get_x")
        return out

    def phase_times(self):
        pt = PhaseTimes()
        _check(lib.cfm_cg_plan_phase_times(self._h, ctypes.byref(pt)), "cfm_cg_plan_phase_times")
        return pt.as_dict()

    def close(self):
        if self._h:
            lib.cfm_cg_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cg(rowptr, col, val, b, x0=None, index_base=0, **kw):
    """One-shot solve. Returns (status, x, hist, iters, timing)."""
    p = cg_params(**kw)
    with CgPlan(rowptr, col, val, p, index_base) as plan:
        plan.set_rhs(b)
        plan.set_x0(x0)
        st, hist, it, t = plan.solve()
        x = plan.x()
    return st, x, hist, it, t


# ----------------------------------------------------------------- batched --

def varcoef3d(seed, system_id, lmin, lmax):
    """Variable-coefficient system: returns (rowptr, col, val, f, dims)."""
    dims = np.zeros(3, dtype=np.int32)
    nnz = lib.cfm_gen_varcoef3d(seed, system_id, lmin, lmax, dims.ctypes.data_as(_ip), None, None, None, None)
    if nnz < 0:
        raise ValueError("bad generator arguments")
    n = int(dims.prod())
    rowptr = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    f = np.empty(n, dtype=np.float64)
    lib.cfm_gen_varcoef3d(seed, system_id, lmin, lmax, dims.ctypes.data_as(_ip), rowptr.ctypes.data_as(_ip),
                          col.ctypes.data_as(_ip), val.ctypes.data_as(_dp), f.ctypes.data_as(_dp))
    return rowptr, col, val, f, tuple(int(d) for d in dims)


def varcoef3d_coef(seed, system_id, lmin, lmax):
    """Per-cell coefficients k of the varcoef3d system, shape (Lz, Ly, Lx)."""
    dims = np.zeros(3, dtype=np.int32)
    nnz = lib.cfm_gen_varcoef3d(seed, system_id, lmin, lmax, dims.ctypes.data_as(_ip), None, None, None, None)
    if nnz < 0:
        raise CfmError(EINVAL, "cfm_gen_varcoef3d_coef")
    k = np.empty(int(dims.prod()))
    _check(lib.cfm_gen_varcoef3d_coef(seed, system_id, lmin, lmax, dims.ctypes.data_as(_ip), k.ctypes.data_as(_dp)),
           "cfm_gen_varcoef3d_coef")
    return k.reshape(int(dims[2]), int(dims[1]), int(dims[0]))


def block_diag(systems):
    """Concatenate [(rowptr, col, val), ...] into one block-diagonal CSR.
    Returns (sys_row_ptr int64, rowptr, col, val)."""
    n = [s[0].size - 1 for s in systems]
    sys_row_ptr = np.zeros(len(systems) + 1, dtype=np.int64)
    sys_row_ptr[1:] = np.cumsum(n)
    nnz_off = np.zeros(len(systems) + 1, dtype=np.int64)
    nnz_off[1:] = np.cumsum([s[1].size for s in systems])
    rowptr = np.empty(sys_row_ptr[-1] + 1, dtype=np.int32)
    col = np.empty(nnz_off[-1], dtype=np.int32)
    val = np.empty(nnz_off[-1], dtype=np.float64)
    for i, (rp, c, v) in enumerate(systems):
        r0, e0 = sys_row_ptr[i], nnz_off[i]
        rowptr[r0:r0 + n[i]] = rp[:-1] + e0
        col[e0:e0 + c.size] = c + r0
        val[e0:e0 + v.size] = v
    rowptr[-1] = nnz_off[-1]
    return sys_row_ptr, rowptr, col, val


class CgBatchPlan:
    """Device-resident block-diagonal batch solved by one persistent launch."""

    def __init__(self, sys_row_ptr, rowptr, col, val, params=None):
        self.sys_row_ptr = np.ascontiguousarray(sys_row_ptr, dtype=np.int64)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.nsys = self.sys_row_ptr.size - 1
        self.n = int(self.sys_row_ptr[-1])
        self.params = params or cg_params(method=CG_PIPELINED)
        h = _vp()
        _check(lib.cfm_cg_batch_plan_create(self.nsys, _ptr(self.sys_row_ptr), _ptr(self.rowptr), _ptr(self.col),
                                            _ptr(self.val), ctypes.byref(self.params), ctypes.byref(h)),
               "cfm_cg_batch_plan_create")
        self._h = h

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib.cfm_cg_batch_set_rhs(self._h, b.ctypes.data_as(_dp)), "cfm_cg_batch_set_rhs")

    def set_x0(self, x0=None):
        if x0 is None:
            _check(lib.cfm_cg_batch_set_x0(self._h, None), "cfm_cg_batch_set_x0")
        else:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            _check(lib.cfm_cg_batch_set_x0(self._h, x0.ctypes.data_as(_dp)), "cfm_cg_batch_set_x0")

    def solve(self, hist_cap=0):
        """Returns (status, per-system iters, per-system status, hist[nsys, hist_cap] or None, timing)."""
        iters = np.zeros(self.nsys, dtype=np.int32)
        stat = np.zeros(self.nsys, dtype=np.int32)
        hist = np.zeros((self.nsys, max(hist_cap, 1)))
        t = Timing()
        st = lib.cfm_cg_batch_solve(self._h, iters.ctypes.data_as(_ip), stat.ctypes.data_as(_ip),
                                    hist.ctypes.data_as(_dp) if hist_cap > 0 else None, hist_cap, ctypes.byref(t))
        if st not in (OK, ENOTCONV, EBREAKDOWN):
            raise CfmError(st, "cfm_cg_batch_solve")
        return st, iters, stat, (hist if hist_cap > 0 else None), t.as_dict()

    def x(self):
        out = np.empty(self.n)
        _check(lib.cfm_cg_batch_get_x(self._h, out.ctypes.data_as(_dp)), "cfm_cg_batch_get_x")
        return out

    def x0_rejected(self):
        out = np.zeros(self.nsys, dtype=np.int32)
        _check(lib.cfm_cg_batch_x0_rejected(self._h, out.ctypes.data_as(_ip)), "cfm_cg_batch_x0_rejected")
        return out

    def load_features(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32)
        if f.size != 2 * self.n:
            raise ValueError("features must hold 2 values per row")
        _check(lib.cfm_cg_batch_load_features(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "cfm_cg_batch_load_features")

    def warm_start(self, net, scale):
        """x0 := scale * net(features) on the device; returns the timing dict."""
        t = Timing()
        _check(lib.cfm_cg_batch_warm_start(self._h, net._h, float(scale), ctypes.byref(t)), "cfm_cg_batch_warm_start")
        return t.as_dict()

    def smooth_x0(self, sweeps, omega=0.8):
        """Damped Jacobi sweeps on x0 (device); returns the timing dict."""
        t = Timing()
        _check(lib.cfm_cg_batch_smooth_x0(self._h, int(sweeps), float(omega), ctypes.byref(t)), "cfm_cg_batch_smooth_x0")
        return t.as_dict()

    def tune(self, trial_iters=20, cache_path=None):
        bb, bg = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.cfm_cg_batch_tune(self._h, trial_iters, cache_path.encode() if cache_path else None,
                                     ctypes.byref(bb), ctypes.byref(bg)), "cfm_cg_batch_tune")
        return bb.value, bg.value

    def matrix_format(self):
        """Storage of the next solve: "sell" or "dia"."""
        return {1: "sell", 2: "dia"}[lib.cfm_cg_batch_matrix_format(self._h)]

    def close(self):
        if self._h:
            lib.cfm_cg_batch_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- learned warm start --

# (Cin, Cout) of the ten 3x3x3 convolutions, in parameter order.
UNET_CONVS = ((2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16))
UNET_HEAD_IN = 16


def unet_param_count():
    return int(lib.cfm_unet_param_count())


def unet_init_params(seed):
    out = np.empty(unet_param_count(), dtype=np.float32)
    _check(lib.cfm_unet_init_params(seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
           "cfm_unet_init_params")
    return out


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class UNet:
    """The warm-start network on the device: forward, gradient and Adam steps
    on batches of n^3 grids (input [batch, 2, n, n, n], output [batch, n, n, n])."""

    def __init__(self, n, batch, params=None, seed=0):
        self.n, self.batch = int(n), int(batch)
        h = _vp()
        _check(lib.cfm_unet_create(self.n, self.batch, ctypes.byref(h)), "cfm_unet_create")
        self._h = h
        self.set_params(unet_init_params(seed) if params is None else params)

    def _in(self, x):
        x = _f32(x)
        if x.size != self.batch * 2 * self.n ** 3:
            raise ValueError("input must hold batch * 2 * n^3 values")
        return x

    def _tgt(self, y):
        y = _f32(y)
        if y.size != self.batch * self.n ** 3:
            raise ValueError("target must hold batch * n^3 values")
        return y

    def set_params(self, params):
        p = _f32(params)
        if p.size != unet_param_count():
            raise ValueError("parameter vector has the wrong length")
        _check(lib.cfm_unet_set_params(self._h, _fptr(p)), "cfm_unet_set_params")

    def params(self):
        out = np.empty(unet_param_count(), dtype=np.float32)
        _check(lib.cfm_unet_get_params(self._h, _fptr(out)), "cfm_unet_get_params")
        return out

    def forward(self, x):
        x = self._in(x)
        out = np.empty((self.batch, self.n, self.n, self.n), dtype=np.float32)
        _check(lib.cfm_unet_forward(self._h, _fptr(x), _fptr(out)), "cfm_unet_forward")
        return out

    def gradient(self, x, target):
        """Returns (loss, flat gradient)."""
        x, y = self._in(x), self._tgt(target)
        g = np.empty(unet_param_count(), dtype=np.float32)
        loss = ctypes.c_double(0)
        _check(lib.cfm_unet_gradient(self._h, _fptr(x), _fptr(y), _fptr(g), ctypes.byref(loss)),
               "cfm_unet_gradient")
        return loss.value, g

    def train_step(self, x, target, lr):
        x, y = self._in(x), self._tgt(target)
        loss = ctypes.c_double(0)
        _check(lib.cfm_unet_train_step(self._h, _fptr(x), _fptr(y), float(lr), ctypes.byref(loss)),
               "cfm_unet_train_step")
        return loss.value

    def close(self):
        if self._h:
            lib.cfm_unet_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
