"""ctypes binding of libparcel_hip.so (C ABI declared in include/parcel_hip.h).

Host-side mirror of the reference's solver interface: `dcg` keeps the
argument meaning of `parcel_dcg` (src/wrapper.f90:1995-2083) -- 1-based CRS,
`itrmax` in/out, residual history, `iret` -- and `Plan` exposes the
setup/solve split of KSP_CRS (src/krylov.f90:447-735).

Arrays may be numpy arrays (host memory: the library copies them) or
DeviceArray buffers (HBM, allocated through the library's own HIP runtime).  The library is loaded from lib/ next to this file;
if it is missing, loading raises instead of falling back to anything.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libparcel_hip.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_ZERO_DIAG = -3
ERR_UNSUPPORTED = -4

SOLVER_CG, SOLVER_BICG, SOLVER_GMRES, SOLVER_CAGMRES, SOLVER_CBCG = 1, 2, 3, 4, 5
PRECON_NONE, PRECON_PJ, PRECON_BJ, PRECON_AS = 0, 1, 2, 3
ILU0, ILUD, MLUD = 0, 1, 2

_PARAM_FIELDS = [
    "solver", "ipreflag", "ILU_method", "addL", "iflagAS", "iovlflag", "precon_thblock",
    "Neuman_itr", "independ_nvec", "nBlock", "GMRES_m", "GMRES_GSflag", "CAGMRES_sstep",
    "CAGMRES_tstep", "CAGMRES_basis", "CAGMRES_QRflag", "CBCG_kstep", "CBCG_Eigenflag",
    "power_method_itrmax", "CAARNOLDI_sstep", "CAARNOLDI_tstep", "CAARNOLDI_basis",
    "CAARNOLDI_QRflag", "profile",
]


class Param(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in _PARAM_FIELDS]


class Stats(ctypes.Structure):
    _fields_ = [
        ("setup_s", ctypes.c_double),
        ("precon_setup_s", ctypes.c_double),
        ("solve_s", ctypes.c_double),
        ("solve_device_ms", ctypes.c_double),
        ("niter", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("spmv_ms", ctypes.c_double),
        ("spmv_calls", ctypes.c_int),
        ("precon_ms", ctypes.c_double),
        ("precon_calls", ctypes.c_int),
        ("vec_ms", ctypes.c_double),
        ("vec_calls", ctypes.c_int),
        ("bj_blocks", ctypes.c_int),
        ("bj_levels_fwd", ctypes.c_int),
        ("bj_levels_bwd", ctypes.c_int),
        ("bj_entries", ctypes.c_long),
        ("bj_nnz_l", ctypes.c_long),
        ("bj_nnz_u", ctypes.c_long),
        ("inner_niter", ctypes.c_int),
        ("final_rel_res", ctypes.c_double),
        ("bj_lscan", ctypes.c_int),
        ("spmv_stencil", ctypes.c_int),
        ("sell_entries", ctypes.c_long),
        ("sell_overflow", ctypes.c_long),
        ("bj_lines", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol the header declares, with (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int
_IP = ctypes.POINTER(ctypes.c_int)
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
SYMBOLS = {
    "parcel_hip_init": (_I, [_I]),
    "parcel_hip_finalize": (_I, []),
    "parcel_hip_last_error": (ctypes.c_char_p, []),
    "parcel_hip_default_param": (None, [ctypes.POINTER(Param)]),
    "parcel_hip_dcg": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                            _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dbicgstab": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                                  _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dgmres": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                               _I, _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dcagmres": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                                 _I, _I, _I, _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_tsqr_r": (_I, [_I, _I, _P, _I, _P]),
    "parcel_hip_dcbcg": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                              _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dcg_mixed": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                                  _I, _I, _I, _I, _I, _D, _I, _IP]),
    "parcel_hip_plan_set_inner": (_I, [_P, _D, _I]),
    "parcel_hip_ddcg": (_I, [_I, _P, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _I, _I, _I, _I, _IP, _D, _P,
                             _I, _I, _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_plan_create_dd": (_I, [ctypes.POINTER(_P), _I, _I, _I, _I, _P, _P, _P, _P, ctypes.POINTER(Param),
                                       _I, _I]),
    "parcel_hip_plan_solve_dd": (_I, [_P, _P, _P, _P, _P, _IP, _D, _P, _I, _I, _IP]),
    "parcel_hip_plan_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _I, _P, _P, _P,
                                    ctypes.POINTER(Param)]),
    "parcel_hip_plan_solve": (_I, [_P, _P, _P, _IP, _D, _P, _IP]),
    "parcel_hip_plan_stats": (_I, [_P, ctypes.POINTER(Stats)]),
    "parcel_hip_plan_destroy": (_I, [_P]),
    "parcel_hip_plan_eigen": (_I, [_P, _P, _I, _P, _P, _P, _P, _IP, _IP]),
    "parcel_hip_plan_set_profile": (_I, [_P, _I]),
    "parcel_hip_plan_spmv": (_I, [_P, _P, _P]),
    "parcel_hip_plan_precon": (_I, [_P, _P, _P]),
    "parcel_hip_dot": (_I, [_I, _P, _P, _DP]),
    "parcel_hip_dot_noalloc": (_I, [_I, _P, _P, _P]),
    "parcel_hip_plan_time_kernel": (_I, [_P, _I, _I, _DP]),
    "parcel_hip_malloc": (_I, [ctypes.POINTER(_P), ctypes.c_ulong]),
    "parcel_hip_free": (_I, [_P]),
    "parcel_hip_memcpy": (_I, [_P, _P, ctypes.c_ulong, _I]),
    "parcel_hip_memset": (_I, [_P, _I, ctypes.c_ulong]),
    "parcel_hip_device_synchronize": (_I, []),
    "parcel_gen_poisson3d_parcel": (_I, [_I, _I, _I, _IP, _P, _P, _P]),
    "parcel_gen_poisson3d_rhs_parcel": (_I, [_I, _I, _I, _P, _P]),
    "parcel_gen_poisson2d": (_I, [_I, _I, _D, _IP, _P, _P, _P]),
    "parcel_gen_poisson3d_spd": (_I, [_I, _I, _I, _IP, _P, _P, _P]),
    "parcel_gen_uniform": (_I, [ctypes.c_long, ctypes.c_ulonglong, _D, _D, _P]),
    "parcel_gen_normal": (_I, [ctypes.c_long, ctypes.c_ulonglong, _P]),
    "parcel_gen_convdiff3d": (_I, [_I, _I, _I, _D, _D, _D, _D, _IP, _P, _P, _P]),
    "parcel_gen_loguniform": (_I, [ctypes.c_long, ctypes.c_ulonglong, _D, _D, _P]),
    "parcel_gen_diffusion3d": (_I, [_I, _I, _I, _P, _IP, _P, _P, _P]),
    "parcel_gen_spmv_dd": (_I, [_I, _P, _P, _P, _P, _P, _P]),
    "parcel_gen_crs_to_dia": (_I, [_I, _P, _P, _P, _IP, _P, _P]),
    "parcel_hip_plan_create_dia": (_I, [ctypes.POINTER(_P), _I, _I, _I, _P, _P, _I, ctypes.POINTER(Param)]),
    "parcel_hip_dcg_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _D, _P, _I, _I, _I, _I,
                                _I, _IP]),
    "parcel_hip_dbicgstab_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _D, _P, _I, _I, _I,
                                      _I, _I, _IP]),
    "parcel_hip_dgmres_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _D, _P, _I, _I, _I, _I,
                                   _I, _I, _I, _IP]),
    "parcel_hip_dcagmres_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _D, _P, _I, _I, _I,
                                     _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dcbcg_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _D, _P, _I, _I, _I, _I,
                                  _I, _I, _I, _I, _I, _I, _I, _I, _IP]),
    "parcel_hip_dpower": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _I, _I, _I, _I, _I,
                               _I, _P, _P, _P, _P, _IP, _IP, _IP]),
    "parcel_hip_dpower_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _IP, _I, _I, _I, _I, _I,
                                   _I, _P, _P, _P, _P, _IP, _IP, _IP]),
    "parcel_hip_dcaarnoldi": (_I, [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I,
                                   _I, _I, _I, _I, _P, _P, _P, _P, _IP, _IP, _IP]),
    "parcel_hip_dcaarnoldi_dia": (_I, [_I, _P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I,
                                       _I, _I, _I, _I, _P, _P, _P, _P, _IP, _IP, _IP]),
    "parcel_gen_poisson3d_ddm": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "parcel_hip_dcg_ddm": (_I, [_I, _P, _P, _I, _I, _P, _P, _I, _P, _I, _I, _P, _P, _P, _I, _I, _I, _I, _IP, _D,
                                _P, _I, _I, _I, _I, _I, _IP]),
    "parcel_gen_grf2d": (_I, [_I, _I, _D, _D, ctypes.c_ulonglong, _P]),
    "parcel_gen_diffusion2d": (_I, [_I, _I, _P, _IP, _P, _P, _P]),
    "parcel_gen_grf_diffusion2d": (_I, [_I, _I, _D, _D, ctypes.c_ulonglong, _I, _IP, _P, _P, _P, _P]),
    "parcel_hip_batch_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _P, _P, _P, _I]),
    "parcel_hip_batch_set_factor": (_I, [_P, _P, _P, _P]),
    "parcel_hip_batch_solve": (_I, [_P, _P, _P, _I, _D, _I, _P, _P, _P]),
    "parcel_hip_batch_set_inner": (_I, [_P, _D, _I]),
    "parcel_hip_batch_set_strips": (_I, [_P, _I]),
    "parcel_hip_batch_set_fallback": (_I, [_P, _I, _I, _I, _I]),
    "parcel_hip_batch_set_ic0_blocks": (_I, [_P, _I]),
    "parcel_hip_batch_get_fallback": (_I, [_P, _P]),
    "parcel_hip_batch_set_profile": (_I, [_P, _I]),
    "parcel_hip_batch_get_stats": (_I, [_P, _P]),
    "parcel_hip_batch_spmv": (_I, [_P, _P, _P]),
    "parcel_hip_batch_precon_apply": (_I, [_P, _P, _P]),
    "parcel_hip_batch_precon_apply32": (_I, [_P, _P, _P]),
    "parcel_hip_batch_destroy": (_I, [_P]),
    "parcel_hip_batch_get_factor": (_I, [_P, _P, _P, _P]),
    "parcel_hip_mpnn_nparams": (_I, []),
    "parcel_hip_mpnn_create": (_I, [ctypes.POINTER(_P), _P, _I]),
    "parcel_hip_mpnn_factor": (_I, [_P, _P]),
    "parcel_hip_mpnn_inspect": (_I, [_P, _P, _P, _P, _P, _P]),
    "parcel_hip_mpnn_destroy": (_I, [_P]),
    "parcel_hip_mpnn_trainer_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _P]),
    "parcel_hip_mpnn_trainer_add": (_I, [_P, _P, _P, _P]),
    "parcel_hip_mpnn_trainer_grad": (_I, [_P, _P, _I, _P, _DP]),
    "parcel_hip_mpnn_trainer_step": (_I, [_P, _P, _I, ctypes.c_float, _DP]),
    "parcel_hip_mpnn_trainer_params": (_I, [_P, _P]),
    "parcel_hip_mpnn_trainer_destroy": (_I, [_P]),
}

BATCH_NONE, BATCH_JACOBI, BATCH_IC0, BATCH_FACTOR = 0, 1, 2, 3
BATCH_FP64, BATCH_REFINE32, BATCH_APPLY32 = 64, 32, 48


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("setup_s", ctypes.c_double),
        ("solve_s", ctypes.c_double),
        ("solve_device_ms", ctypes.c_double),
        ("max_niter", ctypes.c_int),
        ("outer_steps", ctypes.c_int),
        ("inner_total", ctypes.c_long),
        ("inner_steps", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("apply_ms", ctypes.c_double),
        ("apply_calls", ctypes.c_int),
        ("spmv_ms", ctypes.c_double),
        ("spmv_calls", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("stencil5", ctypes.c_int),
        ("fallback_systems", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

_lib = None


class ParcelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("parcel_hip error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libparcel_hip.so; raises if the library was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError("libparcel_hip.so not built: %s (run __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != OK:
        msg = load().parcel_hip_last_error()
        raise ParcelError(rc, msg.decode() if msg else "")


class DeviceArray:
    """A buffer in HBM owned through the library's own runtime
    (parcel_hip_malloc); `shape`/`dtype` mirror the numpy array it copies."""

    def __init__(self, host=None, n=None, dtype=np.float64):
        lib = load()
        if host is not None:
            host = np.ascontiguousarray(host)
            self.dtype, self.shape = host.dtype, host.shape
        else:
            self.dtype, self.shape = np.dtype(dtype), (int(n),)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib.parcel_hip_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p
        if host is not None:
            _check(lib.parcel_hip_memcpy(self.ptr, ctypes.c_void_p(host.ctypes.data), self.nbytes, 1))
        else:
            _check(lib.parcel_hip_memset(self.ptr, 0, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(load().parcel_hip_memcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, self.nbytes, 2))
        return out

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        _check(load().parcel_hip_memcpy(self.ptr, ctypes.c_void_p(host.ctypes.data), self.nbytes, 1))

    def zero(self):
        _check(load().parcel_hip_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            load().parcel_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    """Pointer of a numpy array (host) or a DeviceArray (HBM)."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def _size(a):
    return int(np.prod(a.shape))


def _as(a, dtype, size=None, name="array"):
    """An input the library only reads: a DeviceArray of exactly `dtype` (and
    `size` elements), or anything numpy converts, copied to a C-contiguous
    array of `dtype` when it is not one already."""
    dtype = np.dtype(dtype)
    if isinstance(a, DeviceArray):
        if a.dtype != dtype:
            raise ValueError("%s: device array of %s, expected %s" % (name, a.dtype, dtype))
    else:
        a = np.ascontiguousarray(a, dtype=dtype)
    if size is not None and _size(a) != size:
        raise ValueError("%s: %d elements, expected %d" % (name, _size(a), size))
    return a


def _io(a, dtype, size=None, name="array"):
    """An array the library writes in place: it must already be a C-contiguous
    numpy array (or DeviceArray) of exactly `dtype` and `size` elements, since
    a converted copy would not return the result."""
    dtype = np.dtype(dtype)
    if isinstance(a, DeviceArray):
        ok = a.dtype == dtype
    else:
        ok = isinstance(a, np.ndarray) and a.dtype == dtype and a.flags["C_CONTIGUOUS"] and a.flags["WRITEABLE"]
    if not ok:
        raise ValueError("%s: expected a writeable C-contiguous %s array" % (name, dtype))
    if size is not None and _size(a) != size:
        raise ValueError("%s: %d elements, expected %d" % (name, _size(a), size))
    return a


def _csr(rowptr, col, val):
    """Validated 1-based CSR inputs: (rowptr int32, col int32, val float64, n, nnz)."""
    rowptr = _as(rowptr, np.int32, name="rowptr")
    n = _size(rowptr) - 1
    if n < 0:
        raise ValueError("rowptr: needs n+1 >= 1 entries")
    col = _as(col, np.int32, name="col")
    nnz = _size(col)
    val = _as(val, np.float64, nnz, name="val")
    return rowptr, col, val, n, nnz


def default_param(**kw):
    p = Param()
    load().parcel_hip_default_param(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, int(v))
    return p


# --- generators ---------------------------------------------------------------

def _gen_csr(fn, *args):
    lib = load()
    nnz = ctypes.c_int(0)
    _check(fn(*args, ctypes.byref(nnz), None, None, None))
    return nnz.value


def poisson3d_parcel(gnx, gny, gnz):
    """make_poisson3D_CRS (src/example/problem.f90:448-618): diag -6, off +1."""
    lib = load()
    nnz = _gen_csr(lib.parcel_gen_poisson3d_parcel, gnx, gny, gnz)
    n = gnx * gny * gnz
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    c = ctypes.c_int(0)
    _check(lib.parcel_gen_poisson3d_parcel(gnx, gny, gnz, ctypes.byref(c), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def poisson3d_spd(gnx, gny, gnz):
    """3D 7-point constant-coefficient Poisson, diag 6, off -1."""
    lib = load()
    nnz = _gen_csr(lib.parcel_gen_poisson3d_spd, gnx, gny, gnz)
    n = gnx * gny * gnz
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    c = ctypes.c_int(0)
    _check(lib.parcel_gen_poisson3d_spd(gnx, gny, gnz, ctypes.byref(c), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def convdiff3d(gnx, gny, gnz, peclet=0.5, vel=(1.0, 0.5, 0.25)):
    """3D 7-point first-order upwind convection-diffusion, cell Peclet
    max|v| h / (2D) = peclet, divided by D*h (1-based CSR)."""
    lib = load()
    nnz = _gen_csr(lib.parcel_gen_convdiff3d, gnx, gny, gnz, peclet, *vel)
    n = gnx * gny * gnz
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    c = ctypes.c_int(0)
    _check(lib.parcel_gen_convdiff3d(gnx, gny, gnz, peclet, *vel, ctypes.byref(c), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def loguniform(n, seed, kmin=1e-4, kmax=1e4):
    k = np.empty(n, np.float64)
    _check(load().parcel_gen_loguniform(n, seed, kmin, kmax, _ptr(k)))
    return k


def diffusion3d(gnx, gny, gnz, kappa):
    """make_poisson3D_adv_CRS (src/example/problem.f90:620-770), harmonic-mean
    faces of the per-cell conductivity kappa (1-based CSR)."""
    lib = load()
    kappa = np.ascontiguousarray(kappa, dtype=np.float64)
    if kappa.shape != (gnx * gny * gnz,):
        raise ValueError("kappa must hold gnx*gny*gnz values")
    nnz = ctypes.c_int(0)
    _check(lib.parcel_gen_diffusion3d(gnx, gny, gnz, None, ctypes.byref(nnz), None, None, None))
    n = gnx * gny * gnz
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    _check(lib.parcel_gen_diffusion3d(gnx, gny, gnz, _ptr(kappa), ctypes.byref(nnz), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def spmv_dd(rowptr, col, val, x):
    """b = A x in double-double on the host: (b_hi, b_lo)."""
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    x = _as(x, np.float64, n, "x")
    bh, bl = np.empty(n), np.empty(n)
    _check(load().parcel_gen_spmv_dd(n, _ptr(rowptr), _ptr(col), _ptr(val), _ptr(x), _ptr(bh), _ptr(bl)))
    return bh, bl


def crs_to_dia(rowptr, col, val):
    """(offset int32[nnd] ascending, diaA float64[nnd*n]) of a 1-based CSR matrix
    (make_poisson2D_matrix_DIA_get_nnd / _from_CRS)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    nnd = ctypes.c_int(0)
    _check(lib.parcel_gen_crs_to_dia(n, _ptr(rowptr), _ptr(col), _ptr(val), ctypes.byref(nnd), None, None))
    off = np.zeros(nnd.value, np.int32)
    dia = np.zeros(nnd.value * n)
    _check(lib.parcel_gen_crs_to_dia(n, _ptr(rowptr), _ptr(col), _ptr(val), ctypes.byref(nnd), _ptr(off), _ptr(dia)))
    return off, dia


_DIA_SOLVERS = {"cg": "parcel_hip_dcg_dia", "bicgstab": "parcel_hip_dbicgstab_dia", "gmres": "parcel_hip_dgmres_dia",
                "cagmres": "parcel_hip_dcagmres_dia", "cbcg": "parcel_hip_dcbcg_dia"}


def solve_dia(solver, x, b, offset, diaA, ipreflag=PRECON_NONE, ILU_method=ILUD, itrmax=1000, rtolmax=1e-8,
              precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0, Neuman_itr=1, nBlock=2000,
              extra=()):
    """parcel_d<solver>_dia: DIA matrix (offset, diaA diagonal-major); extra:
    the solver's own trailing parameters (GMRES_m, GSflag / CA-GMRES s, t,
    basis, QR flag / CBCG k, eigenflag, pm itrmax, CA-Arnoldi s, t, basis, QR);
    solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    offset = _as(offset, np.int32, name="offset")
    nnd = _size(offset)
    b = _as(b, np.float64, name="b")
    n = _size(b)
    x = _io(x, np.float64, n, "x")
    diaA = _as(diaA, np.float64, nnd * n, "diaA")
    hist = np.zeros(max(itrmax, 1) + 256, np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(getattr(lib, _DIA_SOLVERS[solver])(0, _ptr(x), _ptr(b), n, n, 1, _ptr(diaA), _ptr(offset), nnd, ipreflag,
                                              ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                                              iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock,
                                              *[int(v) for v in extra], ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


def dpower(x, rowptr=None, col=None, val=None, dia=None, power_method_itrmax=100, ipreflag=PRECON_NONE,
           ILU_method=ILUD, precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0, Neuman_itr=1,
           nBlock=2000, caarnoldi=None, fill=0.0):
    """parcel_dpower (or, with caarnoldi=(s, t, basis, QRflag), parcel_dcaarnoldi,
    which runs the same power method with one iteration) from the start
    vector x (all ones when x is 0), on a 1-based CSR matrix or dia=(offset,
    diaA).  Returns a dict: EigenValueReal, EigenValueImag (lists of one),
    EigenVector (n; the CRS forms write its first entry only, the rest keep
    `fill`), EigenErrorNorm (list of one), MaxEigenID, MinEigenID, itrmax
    (parcel_dpower only), iret; x is left unchanged."""
    lib = load()
    if dia is not None:
        offset, diaA = dia
        offset = _as(offset, np.int32, name="offset")
        nnd = _size(offset)
        x = _as(x, np.float64, name="x")
        n = _size(x)
        diaA = _as(diaA, np.float64, nnd * n, "diaA")
        mat = (1, _ptr(diaA), _ptr(offset), nnd)
    else:
        rowptr, col, val, n, nnz = _csr(rowptr, col, val)
        x = _as(x, np.float64, n, "x")
        mat = (nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col))
    er, ei, en = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    vec = np.full(n, fill)
    maxid, minid, iret, it = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(-1), ctypes.c_int(-1)
    pre = (ipreflag, ILU_method, addL, iflagAS)
    common = (iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock)
    outs = (ctypes.byref(er), ctypes.byref(ei), _ptr(vec), ctypes.byref(en), ctypes.byref(maxid),
            ctypes.byref(minid), ctypes.byref(iret))
    if caarnoldi is None:
        fn = lib.parcel_hip_dpower_dia if dia is not None else lib.parcel_hip_dpower
        _check(fn(0, _ptr(x), None, n, n, *mat, *pre, ctypes.byref(it), *common, power_method_itrmax, *outs))
    else:
        fn = lib.parcel_hip_dcaarnoldi_dia if dia is not None else lib.parcel_hip_dcaarnoldi
        _check(fn(0, _ptr(x), None, n, n, *mat, *pre, *common, *[int(v) for v in caarnoldi], *outs))
    return {"EigenValueReal": [er.value], "EigenValueImag": [ei.value], "EigenVector": vec,
            "EigenErrorNorm": [en.value], "MaxEigenID": maxid.value, "MinEigenID": minid.value,
            "itrmax": it.value if caarnoldi is None else None, "iret": iret.value}


def poisson3d_rhs_parcel(gnx, gny, gnz):
    """make_poisson3D_right_vector (src/example/problem.f90:772-875): (b, x_exact)."""
    n = gnx * gny * gnz
    b = np.empty(n, np.float64)
    x = np.empty(n, np.float64)
    _check(load().parcel_gen_poisson3d_rhs_parcel(gnx, gny, gnz, _ptr(b), _ptr(x)))
    return b, x


def poisson2d(gnx, gny, scale=1.0):
    """make_poisson2D_CRS (src/example/problem.f90:210-340) times `scale`."""
    lib = load()
    nnz = ctypes.c_int(0)
    _check(lib.parcel_gen_poisson2d(gnx, gny, scale, ctypes.byref(nnz), None, None, None))
    n = gnx * gny
    rp = np.empty(n + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    _check(lib.parcel_gen_poisson2d(gnx, gny, scale, ctypes.byref(nnz), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def uniform(n, seed, lo=-1.0, hi=1.0):
    x = np.empty(n, np.float64)
    _check(load().parcel_gen_uniform(n, seed, lo, hi, _ptr(x)))
    return x


def poisson3d_ddm(gnx, gny, gnz, margin=1):
    """make_poisson_3D_DDM (src/example/ddm_setup.f90:133-349), one rank:
    dict with val_dia (7*m), ioff_dia (7), ioff_grids (7*3, column-major), b,
    x_exact (padded layout), m and the grid."""
    m = (gnx + 2 * margin) * (gny + 2 * margin) * (gnz + 2 * margin)
    val = np.zeros(7 * m)
    ioff = np.zeros(7, np.int32)
    iog = np.zeros(21, np.int32)
    b = np.zeros(m)
    xe = np.zeros(m)
    _check(load().parcel_gen_poisson3d_ddm(gnx, gny, gnz, margin, _ptr(val), _ptr(ioff), _ptr(iog), _ptr(b),
                                           _ptr(xe)))
    return dict(val_dia=val, ioff_dia=ioff, ioff_grids=iog, b=b, x_exact=xe, m=m, grid=(gnx, gny, gnz),
                margin=margin)


def dcg_ddm(x, b, grid, margin, val_dia, ioff_dia, ioff_grids, ipreflag=PRECON_BJ, itrmax=1000, rtolmax=1e-8,
            independ_nvec=1, addL=0, iflagAS=0):
    """parcel_dcg_ddm (src/ddm_kernel.f90:2893-2986): x, b in the padded
    layout; returns (hist[:niter], niter, iret)."""
    ioff_dia = _as(ioff_dia, np.int32, name="ioff_dia")
    nnd = _size(ioff_dia)
    b = _as(b, np.float64, name="b")
    m = _size(b)
    x = _io(x, np.float64, m, "x")
    val_dia = _as(val_dia, np.float64, nnd * m, "val_dia")
    ioff_grids = _as(ioff_grids, np.int32, nnd * len(grid), "ioff_grids")
    ng = np.ascontiguousarray(grid, dtype=np.int32)
    nd = np.ones(3, np.int32)
    hist = np.zeros(max(itrmax, 1))
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(load().parcel_hip_dcg_ddm(0, _ptr(x), _ptr(b), m, nnd, _ptr(ioff_dia), _ptr(val_dia), 0, None, margin,
                                     int(ng.shape[0]), _ptr(nd), _ptr(ng), _ptr(ioff_grids), 1, ipreflag, addL, iflagAS,
                                     ctypes.byref(it), rtolmax, _ptr(hist), 0, -1, 1, independ_nvec, 2000,
                                     ctypes.byref(iret)))
    return hist[: it.value].copy(), it.value, iret.value


def normal(n, seed):
    """x ~ N(0,1) from the counter-based stream `seed` (Box-Muller)."""
    x = np.empty(n, np.float64)
    _check(load().parcel_gen_normal(n, seed, _ptr(x)))
    return x


def grf2d(nx, ny, ell, sigma, seed):
    """Squared-exponential Gaussian random field on an nx*ny torus."""
    g = np.empty(nx * ny, np.float64)
    _check(load().parcel_gen_grf2d(nx, ny, ell, sigma, seed, _ptr(g)))
    return g


def diffusion2d(nx, ny, kappa):
    """Five-point harmonic-mean diffusion with per-cell kappa (1-based CSR)."""
    lib = load()
    kappa = np.ascontiguousarray(kappa, dtype=np.float64)
    nnz = ctypes.c_int(0)
    _check(lib.parcel_gen_diffusion2d(nx, ny, _ptr(kappa), ctypes.byref(nnz), None, None, None))
    rp = np.empty(nx * ny + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    _check(lib.parcel_gen_diffusion2d(nx, ny, _ptr(kappa), ctypes.byref(nnz), _ptr(rp), _ptr(col), _ptr(val)))
    return rp, col, val


def grf_diffusion2d(nx, ny, seed0, nsys, ell=16.0, sigma=2.0, fields=False):
    """nsys systems kappa = exp(grf(seed0 + s)): (rowptr, col, vals[nsys, nnz][, g[nsys, n]])."""
    lib = load()
    nnz = ctypes.c_int(0)
    _check(lib.parcel_gen_grf_diffusion2d(nx, ny, ell, sigma, seed0, nsys, ctypes.byref(nnz), None, None, None,
                                          None))
    rp = np.empty(nx * ny + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    vals = np.empty((nsys, nnz.value), np.float64)
    g = np.empty((nsys, nx * ny), np.float64) if fields else None
    _check(lib.parcel_gen_grf_diffusion2d(nx, ny, ell, sigma, seed0, nsys, ctypes.byref(nnz), _ptr(rp), _ptr(col),
                                          _ptr(vals), _ptr(g)))
    return (rp, col, vals, g) if fields else (rp, col, vals)


# --- solvers -------------------------------------------------------------------

def init(myrank=0):
    _check(load().parcel_hip_init(myrank))


def dcg(x, b, rowptr, col, val, ipreflag=PRECON_NONE, ILU_method=ILUD, itrmax=1000, rtolmax=1e-8,
        precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0, Neuman_itr=1, nBlock=2000,
        _entry="parcel_hip_dcg"):
    """parcel_dcg: solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    b, x = _as(b, np.float64, n, "b"), _io(x, np.float64, n, "x")
    hist = np.zeros(max(itrmax, 1), np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    rc = getattr(lib, _entry)(0, _ptr(x), _ptr(b), n, n, nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col),
                            ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                            iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock,
                            ctypes.byref(iret))
    _check(rc)
    return hist[: it.value].copy(), it.value, iret.value


def dbicgstab(x, b, rowptr, col, val, **kw):
    """parcel_dbicgstab: same arguments and results as dcg."""
    return dcg(x, b, rowptr, col, val, _entry="parcel_hip_dbicgstab", **kw)


def dgmres(x, b, rowptr, col, val, GMRES_m=4, GMRES_GSflag=1, ipreflag=PRECON_NONE, ILU_method=ILUD,
           itrmax=1000, rtolmax=1e-8, precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0,
           Neuman_itr=1, nBlock=2000):
    """parcel_dgmres: restarted GMRES(m); solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    b, x = _as(b, np.float64, n, "b"), _io(x, np.float64, n, "x")
    hist = np.zeros(max(itrmax, 1), np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(lib.parcel_hip_dgmres(0, _ptr(x), _ptr(b), n, n, nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col),
                                 ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                                 iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, GMRES_m,
                                 GMRES_GSflag, ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


def dcagmres(x, b, rowptr, col, val, CAGMRES_sstep=4, CAGMRES_tstep=1, CAGMRES_basis=0, CAGMRES_QRflag=4,
             ipreflag=PRECON_NONE, ILU_method=ILUD, itrmax=1000, rtolmax=1e-8, precon_thblock=-1, independ_nvec=1,
             addL=0, iflagAS=4, iovlflag=0, Neuman_itr=1, nBlock=2000):
    """parcel_dcagmres: CA-GMRES(s,t); solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    b, x = _as(b, np.float64, n, "b"), _io(x, np.float64, n, "x")
    hist = np.zeros(max(itrmax, 1), np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(lib.parcel_hip_dcagmres(0, _ptr(x), _ptr(b), n, n, nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col),
                                   ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                                   iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep,
                                   CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag, ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


def tsqr_r(V):
    """R factor (non-negative diagonal) of the n x nc matrix V by the library's TSQR."""
    lib = load()
    V = np.asarray(V, dtype=np.float64)
    n, nc = V.shape
    vt = np.ascontiguousarray(V.T)  # column-major n x nc
    rt = np.zeros((nc, nc))
    _check(lib.parcel_hip_tsqr_r(n, nc, _ptr(vt), n, _ptr(rt)))
    return rt.T.copy()


def dcbcg(x, b, rowptr, col, val, CBCG_kstep=4, CBCG_Eigenflag=2, power_method_itrmax=100, CAARNOLDI_sstep=4,
          CAARNOLDI_tstep=10, CAARNOLDI_basis=1, CAARNOLDI_QRflag=3, ipreflag=PRECON_NONE, ILU_method=ILUD,
          itrmax=1000, rtolmax=1e-8, precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0,
          Neuman_itr=1, nBlock=2000):
    """parcel_dcbcg: k-step Chebyshev-basis CG; solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    b, x = _as(b, np.float64, n, "b"), _io(x, np.float64, n, "x")
    hist = np.zeros(max(itrmax, 1) + 1, np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(lib.parcel_hip_dcbcg(0, _ptr(x), _ptr(b), n, n, nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col),
                                ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                                iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, CBCG_kstep,
                                CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep,
                                CAARNOLDI_basis, CAARNOLDI_QRflag, ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


def dcg_mixed(x, b, rowptr, col, val, inner_rtol=1e-4, inner_itrmax=10000, ipreflag=PRECON_NONE,
              ILU_method=ILUD, itrmax=50, rtolmax=1e-10, precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4,
              iovlflag=0, Neuman_itr=1, nBlock=2000):
    """FP64 refinement around FP32 PCG; solves in place into x; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, val, n, nnz = _csr(rowptr, col, val)
    b, x = _as(b, np.float64, n, "b"), _io(x, np.float64, n, "x")
    hist = np.zeros(max(itrmax, 1), np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    _check(lib.parcel_hip_dcg_mixed(0, _ptr(x), _ptr(b), n, n, nnz, 1, _ptr(val), _ptr(rowptr), _ptr(col),
                                    ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it), rtolmax, _ptr(hist),
                                    iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, inner_rtol,
                                    inner_itrmax, ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


def ddcg(xh, xl, bh, bl, rowptr, col, vh, vl, ipreflag=PRECON_NONE, ILU_method=ILUD, itrmax=1000, rtolmax=1e-20,
         precon_thblock=-1, independ_nvec=1, addL=0, iflagAS=4, iovlflag=0, nBlock=2000, PRECISION_A=2,
         PRECISION_b=2, PRECISION_x=2, PRECISION_PRECONDITION=2):
    """parcel_ddcg: double-double PCG, x (hi, lo) in place; returns (hist[:niter], niter, iret)."""
    lib = load()
    rowptr, col, vh, n, nnz = _csr(rowptr, col, vh)
    vl = None if vl is None else _as(vl, np.float64, nnz, "vl")
    bh, bl = _as(bh, np.float64, n, "bh"), None if bl is None else _as(bl, np.float64, n, "bl")
    xh, xl = _io(xh, np.float64, n, "xh"), None if xl is None else _io(xl, np.float64, n, "xl")
    hist = np.zeros(max(itrmax, 1), np.float64)
    it = ctypes.c_int(itrmax)
    iret = ctypes.c_int(-1)
    opt = lambda a: None if a is None else _ptr(a)
    _check(lib.parcel_hip_ddcg(0, _ptr(xh), opt(xl), _ptr(bh), opt(bl), n, n, nnz, 1, _ptr(vh), opt(vl),
                               _ptr(rowptr), _ptr(col), ipreflag, ILU_method, addL, iflagAS, ctypes.byref(it),
                               rtolmax, _ptr(hist), iovlflag, precon_thblock, independ_nvec, nBlock, PRECISION_A,
                               PRECISION_b, PRECISION_x, PRECISION_PRECONDITION, ctypes.byref(iret)))
    return hist[: max(it.value, 0)].copy(), it.value, iret.value


class Plan:
    """Setup once, solve many (parcel_hip_plan_*)."""

    def __init__(self, rowptr, col, val, param=None, **kw):
        lib = load()
        self.lib = lib
        rowptr, col, val, self.n, nnz = _csr(rowptr, col, val)
        p = param if param is not None else default_param(**kw)
        h = ctypes.c_void_p()
        _check(lib.parcel_hip_plan_create(ctypes.byref(h), self.n, self.n, nnz, 1, _ptr(val), _ptr(rowptr),
                                          _ptr(col), ctypes.byref(p)))
        self.h = h

    def solve(self, x, b, itrmax, rtolmax):
        b, x = _as(b, np.float64, self.n, "b"), _io(x, np.float64, self.n, "x")
        hist = np.zeros(max(itrmax, 1), np.float64)
        it = ctypes.c_int(itrmax)
        iret = ctypes.c_int(-1)
        _check(self.lib.parcel_hip_plan_solve(self.h, _ptr(x), _ptr(b), ctypes.byref(it), rtolmax, _ptr(hist),
                                              ctypes.byref(iret)))
        return hist[: it.value].copy(), it.value, iret.value

    def set_inner(self, inner_rtol, inner_itrmax):
        _check(self.lib.parcel_hip_plan_set_inner(self.h, inner_rtol, inner_itrmax))

    def set_profile(self, on):
        _check(self.lib.parcel_hip_plan_set_profile(self.h, 1 if on else 0))

    def spmv(self, x, y):
        x, y = _as(x, np.float64, self.n, "x"), _io(y, np.float64, self.n, "y")
        _check(self.lib.parcel_hip_plan_spmv(self.h, _ptr(x), _ptr(y)))

    def precon(self, r, s):
        r, s = _as(r, np.float64, self.n, "r"), _io(s, np.float64, self.n, "s")
        _check(self.lib.parcel_hip_plan_precon(self.h, _ptr(r), _ptr(s)))

    def eigen(self, x, power_method_itrmax=100, vec=None):
        """The power method of parcel_dpower on this plan, from x (all ones
        when x is 0): (Emax, EigenErrorNorm, eigenvector); vec (numpy or
        DeviceArray of n) receives the eigenvector when given."""
        x = _as(x, np.float64, self.n, "x")
        vec = np.zeros(self.n) if vec is None else _io(vec, np.float64, self.n, "vec")
        er, ei, en = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(0.0)
        mx, mn = ctypes.c_int(0), ctypes.c_int(0)
        _check(self.lib.parcel_hip_plan_eigen(self.h, _ptr(x), power_method_itrmax, ctypes.byref(er),
                                              ctypes.byref(ei), _ptr(vec), ctypes.byref(en), ctypes.byref(mx),
                                              ctypes.byref(mn)))
        return er.value, en.value, vec

    def stats(self):
        st = Stats()
        _check(self.lib.parcel_hip_plan_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def time_kernel(self, which, reps):
        ms = ctypes.c_double(0.0)
        _check(self.lib.parcel_hip_plan_time_kernel(self.h, which, reps, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            self.lib.parcel_hip_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dot(x, y):
    n = _size(x)
    x, y = _as(x, np.float64, n, "x"), _as(y, np.float64, n, "y")
    r = ctypes.c_double(0.0)
    _check(load().parcel_hip_dot(n, _ptr(x), _ptr(y), ctypes.byref(r)))
    return r.value


class Batch:
    """nsys systems with one sparsity pattern (parcel_hip_batch_*).  vals is
    (nsys, nnz) in CSR order (host array or DeviceArray)."""

    def __init__(self, rowptr, col, vals, precon=BATCH_JACOBI, nsys=None):
        lib = load()
        self.lib = lib
        rowptr = _as(rowptr, np.int32, name="rowptr")
        col = _as(col, np.int32, name="col")
        self.n, nnz = _size(rowptr) - 1, _size(col)
        if nsys is None:
            nsys = int(vals.shape[0]) if len(vals.shape) == 2 else _size(vals) // max(nnz, 1)
        vals = _as(vals, np.float64, nsys * nnz, "vals")
        self.nsys = nsys
        h = ctypes.c_void_p()
        _check(lib.parcel_hip_batch_create(ctypes.byref(h), nsys, self.n, nnz, _ptr(rowptr), _ptr(col), _ptr(vals),
                                           precon))
        self.h = h

    def set_factor(self, Ls, Lw, Ld):
        """bf16 bit patterns (uint16 arrays of nsys*n, host or DeviceArray)."""
        nv = self.nsys * self.n
        Ls, Lw, Ld = (_as(a, np.uint16, nv, nm) for a, nm in ((Ls, "Ls"), (Lw, "Lw"), (Ld, "Ld")))
        _check(self.lib.parcel_hip_batch_set_factor(self.h, _ptr(Ls), _ptr(Lw), _ptr(Ld)))

    def solve(self, x, b, itrmax, rtol, precision=64, history=True):
        """x, b: (nsys, n) host arrays or DeviceArrays; returns (hist, niter, iret)."""
        nv = self.nsys * self.n
        b, x = _as(b, np.float64, nv, "b"), _io(x, np.float64, nv, "x")
        hist = np.zeros((self.nsys, max(itrmax, 1))) if history else None
        niter = np.zeros(self.nsys, np.int32)
        iret = np.zeros(self.nsys, np.int32)
        _check(self.lib.parcel_hip_batch_solve(self.h, _ptr(x), _ptr(b), itrmax, rtol, precision, _ptr(hist),
                                               _ptr(niter), _ptr(iret)))
        return hist, niter, iret

    def set_inner(self, inner_rtol, inner_itrmax):
        _check(self.lib.parcel_hip_batch_set_inner(self.h, inner_rtol, inner_itrmax))

    def set_strips(self, nstrip):
        _check(self.lib.parcel_hip_batch_set_strips(self.h, nstrip))

    def set_profile(self, on):
        _check(self.lib.parcel_hip_batch_set_profile(self.h, 1 if on else 0))

    def set_ic0_blocks(self, nblocks):
        """IC0 batches: block-Jacobi IC(0) over nblocks strips of grid lines."""
        _check(self.lib.parcel_hip_batch_set_ic0_blocks(self.h, nblocks))

    def set_fallback(self, ipreflag=PRECON_BJ, ILU_method=ILU0, precon_thblock=32, itrmax=20000):
        """Re-solve systems that end with iret != 0 by parcel_dcg's FP64 PCG
        (itrmax <= 0: off; FACTOR batches start with these defaults)."""
        _check(self.lib.parcel_hip_batch_set_fallback(self.h, ipreflag, ILU_method, precon_thblock, itrmax))

    def fell_back(self):
        """Per system: True if the last solve fell back."""
        f = np.zeros(self.nsys, np.int32)
        _check(self.lib.parcel_hip_batch_get_fallback(self.h, _ptr(f)))
        return f.astype(bool)

    def stats(self):
        st = BatchStats()
        _check(self.lib.parcel_hip_batch_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def spmv(self, x, y):
        nv = self.nsys * self.n
        x, y = _as(x, np.float64, nv, "x"), _io(y, np.float64, nv, "y")
        _check(self.lib.parcel_hip_batch_spmv(self.h, _ptr(x), _ptr(y)))

    def factor(self):
        """BF16 factor bits (Ls, Lw, Ld), each (nsys*n,) uint16."""
        out = [np.zeros(self.nsys * self.n, np.uint16) for _ in range(3)]
        _check(self.lib.parcel_hip_batch_get_factor(self.h, _ptr(out[0]), _ptr(out[1]), _ptr(out[2])))
        return tuple(out)

    def precon(self, r, z):
        f32 = np.dtype(r.dtype) == np.float32
        fn = self.lib.parcel_hip_batch_precon_apply32 if f32 else self.lib.parcel_hip_batch_precon_apply
        dt, nv = (np.float32 if f32 else np.float64), self.nsys * self.n
        r, z = _as(r, dt, nv, "r"), _io(z, dt, nv, "z")
        _check(fn(self.h, _ptr(r), _ptr(z)))

    def close(self):
        if self.h:
            self.lib.parcel_hip_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mpnn_nparams():
    return load().parcel_hip_mpnn_nparams()


class Mpnn:
    """The learned-factor network (inference) with FP32 parameters."""

    def __init__(self, params):
        self.lib = load()
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        _check(self.lib.parcel_hip_mpnn_create(ctypes.byref(h), _ptr(p), int(p.shape[0])))
        self.h = h

    def factor(self, batch):
        """Predict and install the factor of every system of a FACTOR batch."""
        _check(self.lib.parcel_hip_mpnn_factor(self.h, batch.h))

    def inspect(self, batch):
        nv = batch.nsys * batch.n
        o = np.zeros(nv * 8, np.float32)
        hl = np.zeros(nv * 32, np.uint16)
        ah = np.zeros(nv * 4, np.float32)
        sd = np.zeros(nv, np.float32)
        _check(self.lib.parcel_hip_mpnn_inspect(self.h, batch.h, _ptr(o), _ptr(hl), _ptr(ah), _ptr(sd)))
        return dict(o=o.reshape(nv, 8), hlast=hl.reshape(nv, 32), ahat=ah.reshape(nv, 4), sd=sd)

    def close(self):
        if self.h:
            self.lib.parcel_hip_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MpnnTrainer:
    """Device-side training pool, minibatch gradients and Adam steps."""

    def __init__(self, n, nx, capacity, params):
        self.lib = load()
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        _check(self.lib.parcel_hip_mpnn_trainer_create(ctypes.byref(h), n, nx, capacity, _ptr(p)))
        self.h = h
        self.count = 0

    def add(self, batch, x, b):
        nv = batch.nsys * batch.n
        x, b = _as(x, np.float64, nv, "x"), _as(b, np.float64, nv, "b")
        _check(self.lib.parcel_hip_mpnn_trainer_add(self.h, batch.h, _ptr(x), _ptr(b)))
        self.count += batch.nsys

    def grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(mpnn_nparams(), np.float32)
        loss = ctypes.c_double(0.0)
        _check(self.lib.parcel_hip_mpnn_trainer_grad(self.h, _ptr(idx), int(idx.shape[0]), _ptr(g),
                                                     ctypes.byref(loss)))
        return loss.value, g

    def step(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ctypes.c_double(0.0)
        _check(self.lib.parcel_hip_mpnn_trainer_step(self.h, _ptr(idx), int(idx.shape[0]), lr, ctypes.byref(loss)))
        return loss.value

    def params(self):
        p = np.zeros(mpnn_nparams(), np.float32)
        _check(self.lib.parcel_hip_mpnn_trainer_params(self.h, _ptr(p)))
        return p

    def close(self):
        if self.h:
            self.lib.parcel_hip_mpnn_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
