"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

RED_SEQ, RED_TILED = 0, 1
# the HIP library's reduction layout (parcel-legacy_amd/csrc/common.hpp)
GPU_NT, GPU_NBMAX = 256, 2048


class Red(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("nt", ctypes.c_int), ("nbmax", ctypes.c_int)]


TILED = Red(RED_TILED, GPU_NT, GPU_NBMAX)
SEQ = Red(RED_SEQ, GPU_NT, GPU_NBMAX)

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_RP = ctypes.POINTER(Red)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError("oracle not built: %s" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "oracle_dot": (_D, [_RP, _I, _P, _P]),
        "oracle_reduce_partials": (_D, [_P, _I]),
        "oracle_divide_region": (None, [_I, _I, _I, _IP, _IP]),
        "oracle_spmv": (None, [_I, _P, _P, _P, _P, _P]),
        "oracle_precon_create": (_P, [_I, _I, _I, _I, _I, _P, _P, _P, _IP]),
        "oracle_precon_create_blocks": (_P, [_I, _I, _I, _P, _I, _I, _P, _P, _P, _IP]),
        "oracle_pcg_with": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _D, _RP, _P, _IP, _IP]),
        "oracle_precon_apply": (None, [_P, _P, _P]),
        "oracle_precon_free": (None, [_P]),
        "oracle_precon_nblocks": (_I, [_P]),
        "oracle_pcg": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _RP, _P, _IP, _IP]),
        "oracle_last_times": (None, [_P]),
        "oracle_bicgstab": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _RP, _P, _IP, _IP]),
        "oracle_gmres": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _I, _I, _RP, _P, _IP, _IP]),
        "oracle_cagmres": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _I, _I, _I, _I, _RP, _P, _IP,
                                _IP]),
        "oracle_pcg_mixed": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _I, _D, _RP, _P, _IP, _IP, _IP]),
        "oracle_dd_op": (None, [_I, _I, _P, _P, _P, _P, _P, _P]),
        "oracle_dd_dot": (None, [_RP, _I, _P, _P, _P, _P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
        "oracle_dd_spmv": (None, [_I, _P, _P, _P, _P, _P, _P, _P, _P]),
        "oracle_ddprecon_create": (_P, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _IP]),
        "oracle_ddprecon_apply": (None, [_P, _P, _P, _P, _P]),
        "oracle_ddprecon_free": (None, [_P]),
        "oracle_ddcg": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _RP, _P, _IP, _IP]),
        "oracle_gemm_nn": (None, [_I, _I, _I, _P, _I, _P, _I, _P, _I]),
        "oracle_trmm_lunn": (None, [_I, _P, _I, _P, _I]),
        "oracle_trinv_upper": (None, [_I, _P, _I]),
        "oracle_chol_upper": (_I, [_I, _P, _I]),
        "oracle_hqr_eigs": (_I, [_I, _P, _I, _P, _P]),
        "oracle_leja": (None, [_I, _P, _P]),
        "oracle_sym_pinv": (None, [_I, _P, _I, _P, _I]),
        "oracle_precon_create_ras": (_P, [_I, _I, _P, _P, _P, _P, _P, _I, _P, _IP]),
        "oracle_precon_create_as": (_P, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _IP]),
        "oracle_tsqr_r": (_I, [_I, _I, _P, _I, _P]),
        "oracle_house_r": (None, [_I, _I, _P, _I, _P]),
        "oracle_tsqr_waves": (_I, [_I]),
        "oracle_power_emax": (_D, [_P, _I, _P, _P, _P, _P, _I, _RP]),
        "oracle_power_eigen": (None, [_P, _I, _P, _P, _P, _P, _I, _RP, _P, _P, _P]),
        "oracle_cbcg": (_I, [_I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _I, _I, _I, _RP, _P, _IP, _IP,
                             ctypes.POINTER(_D), ctypes.POINTER(_D)]),
        "oracle_cbcg_with": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _D, _I, _D, _D, _RP, _P, _IP, _IP]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def dot(x, y, red=TILED):
    return load().oracle_dot(ctypes.byref(red), int(x.shape[0]), _p(x), _p(y))


def reduce_partials(part):
    return load().oracle_reduce_partials(_p(part), int(part.shape[0]))


def divide_region(rank, nproc, n):
    lo, hi = ctypes.c_int(), ctypes.c_int()
    load().oracle_divide_region(rank, nproc, n, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def spmv(rowptr, col, val, x):
    n = rowptr.shape[0] - 1
    y = np.empty(n, np.float64)
    load().oracle_spmv(n, _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


class RasPrecon:
    """Restricted additive Schwarz: gather src, block D-ILU of the expanded
    matrix (rp_e, col_e, val_e; 1-based CSR, blocks at starts), scatter to dst."""

    def __init__(self, n, src, dst, rp_e, col_e, val_e, starts):
        self._keep = [np.ascontiguousarray(a, dtype=t) for a, t in
                      ((src, np.int32), (dst, np.int32), (rp_e, np.int32), (col_e, np.int32), (val_e, np.float64),
                       (starts, np.int32))]
        s, d, r, c, v, st = self._keep
        self.n = n
        err = ctypes.c_int(0)
        self.h = load().oracle_precon_create_ras(n, int(s.shape[0]), _p(s), _p(d), _p(r), _p(c), _p(v),
                                                 int(st.shape[0]) - 1, _p(st), ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle RAS setup failed: %d" % err.value)

    def apply(self, r):
        s = np.empty(self.n, np.float64)
        load().oracle_precon_apply(self.h, _p(np.ascontiguousarray(r, dtype=np.float64)), _p(s))
        return s

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_precon_free(self.h)


class AsPrecon:
    """Additive Schwarz of the CRS path at one rank (AdditiveSchwarz_crs):
    divide_region blocks extended by addL rows, D-ILU per extended block,
    iflagAS 1 BASIC, 2 RESTRICT, 3 INTERPOLATE, 4 NONE."""

    def __init__(self, rowptr, col, val, ilu_method=1, nblocks=1, addL=1, iflagAS=2, independ_nvec=1):
        self._keep = (rowptr, col, val)
        self.n = rowptr.shape[0] - 1
        err = ctypes.c_int(0)
        self.h = load().oracle_precon_create_as(ilu_method, nblocks, addL, iflagAS, independ_nvec, self.n,
                                                _p(rowptr), _p(col), _p(val), ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle additive Schwarz setup failed: %d" % err.value)

    def apply(self, r):
        s = np.empty(self.n, np.float64)
        load().oracle_precon_apply(self.h, _p(np.ascontiguousarray(r, dtype=np.float64)), _p(s))
        return s

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_precon_free(self.h)
            self.h = None


class Precon:
    def __init__(self, rowptr, col, val, ipreflag, ilu_method=1, nblocks=1, independ_nvec=1, starts=None):
        """starts: explicit block-Jacobi blocks (0-based starts, nblocks+1 entries)."""
        self._keep = (rowptr, col, val, starts)
        self.n = rowptr.shape[0] - 1
        err = ctypes.c_int(0)
        if starts is None:
            self.h = load().oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec, self.n, _p(rowptr),
                                                 _p(col), _p(val), ctypes.byref(err))
        else:
            starts = np.ascontiguousarray(starts, dtype=np.int32)
            self._keep = (rowptr, col, val, starts)
            self.h = load().oracle_precon_create_blocks(ipreflag, ilu_method, int(starts.shape[0]) - 1, _p(starts),
                                                        independ_nvec, self.n, _p(rowptr), _p(col), _p(val),
                                                        ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle precon setup failed: %d" % err.value)

    def apply(self, r):
        s = np.empty(self.n, np.float64)
        load().oracle_precon_apply(self.h, _p(r), _p(s))
        return s

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_precon_free(self.h)
            self.h = None


def bicgstab(rowptr, col, val, b, x0=None, **kw):
    return pcg(rowptr, col, val, b, x0, _fn="oracle_bicgstab", **kw)


def pcg(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=1000,
        rtol=1e-8, red=TILED, _fn="oracle_pcg"):
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 1))
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    rc = getattr(load(), _fn)(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), ipreflag, ilu_method, nblocks,
                           independ_nvec, itrmax, rtol, ctypes.byref(red), _p(hist), ctypes.byref(niter),
                           ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_pcg failed: %d" % rc)
    return x, hist[: niter.value].copy(), niter.value, iret.value


def gmres(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=1000,
          rtol=1e-8, m=4, gsflag=1, red=TILED):
    """Restarted GMRES(m); returns (x, hist[:niter], niter, iret)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, m + 1, 1))
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    rc = load().oracle_gmres(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), ipreflag, ilu_method, nblocks,
                             independ_nvec, itrmax, rtol, m, gsflag, ctypes.byref(red), _p(hist),
                             ctypes.byref(niter), ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_gmres failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value


def cagmres(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=1000,
            rtol=1e-8, s=4, t=1, basis=0, qrflag=4, red=TILED):
    """CA-GMRES(s,t); returns (x, hist[:niter], niter, iret)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 0) + s * t + 1)
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    rc = load().oracle_cagmres(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), ipreflag, ilu_method, nblocks,
                               independ_nvec, itrmax, rtol, s, t, basis, qrflag, ctypes.byref(red), _p(hist),
                               ctypes.byref(niter), ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_cagmres failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value


def cbcg(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=1000,
         rtol=1e-8, k=4, eigflag=1, pm_itrmax=100, red=TILED):
    """Chebyshev-basis CG; returns (x, hist[:niter], niter, iret, (Emax, Emin))."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 0) + 2)
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    emax, emin = ctypes.c_double(0), ctypes.c_double(0)
    rc = load().oracle_cbcg(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), ipreflag, ilu_method, nblocks,
                            independ_nvec, itrmax, rtol, k, eigflag, pm_itrmax, ctypes.byref(red), _p(hist),
                            ctypes.byref(niter), ctypes.byref(iret), ctypes.byref(emax), ctypes.byref(emin))
    if rc != 0:
        raise RuntimeError("oracle_cbcg failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value, (emax.value, emin.value)


def cbcg_bounds(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1,
                itrmax=1000, rtol=1e-8, k=4, emax=1.0, emin=0.0, red=TILED):
    """CBCG with given eigenvalue bounds; returns (x, hist[:niter], niter, iret)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 0) + 2)
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    pre = Precon(rowptr, col, val, ipreflag, ilu_method, nblocks, independ_nvec)
    rc = load().oracle_cbcg_with(pre.h, n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), itrmax, rtol, k, emax,
                                 emin, ctypes.byref(red), _p(hist), ctypes.byref(niter), ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_cbcg_with failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value


def power_emax(rowptr, col, val, x0, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=100, red=TILED):
    """Power-method estimate of the largest eigenvalue of A M^-1."""
    n = rowptr.shape[0] - 1
    pre = Precon(rowptr, col, val, ipreflag, ilu_method, nblocks, independ_nvec)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    return load().oracle_power_emax(pre.h, n, _p(rowptr), _p(col), _p(val), _p(x0), itrmax, ctypes.byref(red))


def power_eigen(rowptr, col, val, x0, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=100,
                red=TILED):
    """power_method_symMatrix in full: (Emax, ||A x - Emax x|| / ||Emax x||, x)."""
    n = rowptr.shape[0] - 1
    pre = Precon(rowptr, col, val, ipreflag, ilu_method, nblocks, independ_nvec)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    e, en, vec = ctypes.c_double(0.0), ctypes.c_double(0.0), np.zeros(n)
    load().oracle_power_eigen(pre.h, n, _p(rowptr), _p(col), _p(val), _p(x0), itrmax, ctypes.byref(red),
                              ctypes.byref(e), ctypes.byref(en), _p(vec))
    return e.value, en.value, vec


def tsqr_r(V):
    """TSQR R factor (non-negative diagonal) of the n x nc matrix V."""
    V = np.asarray(V, dtype=np.float64)
    n, nc = V.shape
    vt = np.ascontiguousarray(V.T)  # column-major n x nc
    rt = np.zeros((nc, nc))
    rc = load().oracle_tsqr_r(n, nc, _p(vt), n, _p(rt))
    if rc != 0:
        raise RuntimeError("oracle_tsqr_r failed")
    return rt.T.copy()


def house_r(A):
    """House_local_Ronly_omp R factor of the m x s matrix A (diagonal sign as produced)."""
    A = np.asarray(A, dtype=np.float64)
    m, s = A.shape
    at = np.ascontiguousarray(A.T)
    rt = np.zeros((s, s))
    load().oracle_house_r(m, s, _p(at), m, _p(rt))
    return rt.T.copy()


def sym_pinv(a):
    """Pseudo-inverse of a symmetric matrix (upper triangle read)."""
    at = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)  # column-major view
    k = at.shape[0]
    pt = np.zeros((k, k))
    load().oracle_sym_pinv(k, _p(at), k, _p(pt), k)
    return pt.T.copy()


def pcg_mixed(rowptr, col, val, b, x0=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1, itrmax=50,
              rtol=1e-10, inner_itrmax=5000, inner_rtol=1e-4, red=TILED):
    """FP64 iterative refinement around FP32 PCG; returns (x, hist[:niter], niter, iret, inner_total)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 1))
    niter, iret, tot = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    rc = load().oracle_pcg_mixed(n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), ipreflag, ilu_method, nblocks,
                                 independ_nvec, itrmax, rtol, inner_itrmax, inner_rtol, ctypes.byref(red),
                                 _p(hist), ctypes.byref(niter), ctypes.byref(iret), ctypes.byref(tot))
    if rc != 0:
        raise RuntimeError("oracle_pcg_mixed failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value, tot.value


DD_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4}


def dd_op(op, ah, al, bh=None, bl=None):
    ah = np.ascontiguousarray(ah, dtype=np.float64)
    al = np.ascontiguousarray(al, dtype=np.float64)
    bh = np.zeros_like(ah) if bh is None else np.ascontiguousarray(bh, dtype=np.float64)
    bl = np.zeros_like(ah) if bl is None else np.ascontiguousarray(bl, dtype=np.float64)
    ch, cl = np.empty_like(ah), np.empty_like(ah)
    load().oracle_dd_op(DD_OPS[op], int(ah.shape[0]), _p(ah), _p(al), _p(bh), _p(bl), _p(ch), _p(cl))
    return ch, cl


def dd_dot(xh, xl, yh, yl, red=TILED):
    rh, rl = ctypes.c_double(), ctypes.c_double()
    load().oracle_dd_dot(ctypes.byref(red), int(xh.shape[0]), _p(xh), _p(xl), _p(yh), _p(yl), ctypes.byref(rh),
                         ctypes.byref(rl))
    return rh.value, rl.value


def dd_spmv(rowptr, col, vh, vl, xh, xl):
    n = rowptr.shape[0] - 1
    yh, yl = np.empty(n), np.empty(n)
    load().oracle_dd_spmv(n, _p(rowptr), _p(col), _p(vh), None if vl is None else _p(vl), _p(xh), _p(xl), _p(yh),
                          _p(yl))
    return yh, yl


class DDPrecon:
    def __init__(self, rowptr, col, vh, vl, ipreflag, ilu_method=1, nblocks=1, independ_nvec=1, prec=2):
        self._keep = (rowptr, col, vh, vl)
        self.n = rowptr.shape[0] - 1
        err = ctypes.c_int(0)
        self.h = load().oracle_ddprecon_create(ipreflag, ilu_method, nblocks, independ_nvec, prec, self.n,
                                               _p(rowptr), _p(col), _p(vh), None if vl is None else _p(vl),
                                               ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle dd precon setup failed: %d" % err.value)

    def apply(self, rh, rl):
        sh, sl = np.zeros(self.n), np.zeros(self.n)
        load().oracle_ddprecon_apply(self.h, _p(rh), _p(rl), _p(sh), _p(sl))
        return sh, sl

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_ddprecon_free(self.h)
            self.h = None


def ddcg(rowptr, col, vh, vl, bh, bl, x0h=None, x0l=None, ipreflag=0, ilu_method=1, nblocks=1, independ_nvec=1,
         prec=2, itrmax=1000, rtol=1e-20, red=TILED):
    """dd PCG; returns (xh, xl, hist[:niter], niter, iret)."""
    n = rowptr.shape[0] - 1
    xh = np.zeros(n) if x0h is None else np.array(x0h, dtype=np.float64, copy=True)
    xl = np.zeros(n) if x0l is None else np.array(x0l, dtype=np.float64, copy=True)
    bl = np.zeros(n) if bl is None else bl
    hist = np.zeros(max(itrmax, 1))
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    rc = load().oracle_ddcg(n, _p(rowptr), _p(col), _p(vh), None if vl is None else _p(vl), _p(bh), _p(bl), _p(xh),
                            _p(xl), ipreflag, ilu_method, nblocks, independ_nvec, prec, itrmax, rtol,
                            ctypes.byref(red), _p(hist), ctypes.byref(niter), ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_ddcg failed: %d" % rc)
    return xh, xl, hist[: max(niter.value, 0)].copy(), niter.value, iret.value


def _fortran(a):
    return np.asfortranarray(np.array(a, dtype=np.float64))


def hqr_eigs(H):
    """Eigenvalues (wr, wi) of an upper Hessenberg matrix (DLAHQR order)."""
    A = _fortran(H)
    n = A.shape[0]
    wr, wi = np.zeros(n), np.zeros(n)
    rc = load().oracle_hqr_eigs(n, ctypes.c_void_p(A.ctypes.data), n, _p(wr), _p(wi))
    if rc != 0:
        raise RuntimeError("hqr did not converge: %d" % rc)
    return wr, wi


def leja(re, im):
    re = np.array(re, dtype=np.float64)
    im = np.array(im, dtype=np.float64)
    load().oracle_leja(int(re.shape[0]), _p(re), _p(im))
    return re, im


def chol_upper(A):
    R = _fortran(A)
    info = load().oracle_chol_upper(R.shape[0], ctypes.c_void_p(R.ctypes.data), R.shape[0])
    return np.triu(R), info


def trinv_upper(R):
    A = _fortran(R)
    load().oracle_trinv_upper(A.shape[0], ctypes.c_void_p(A.ctypes.data), A.shape[0])
    return np.triu(A)


def gemm_nn(A, B):
    A, B = _fortran(A), _fortran(B)
    C = np.zeros((A.shape[0], B.shape[1]), order="F")
    load().oracle_gemm_nn(A.shape[0], B.shape[1], A.shape[1], ctypes.c_void_p(A.ctypes.data), A.shape[0],
                          ctypes.c_void_p(B.ctypes.data), B.shape[0], ctypes.c_void_p(C.ctypes.data), A.shape[0])
    return C


def trmm_lunn(A, B):
    A, B = _fortran(A), _fortran(B)
    load().oracle_trmm_lunn(A.shape[0], ctypes.c_void_p(A.ctypes.data), A.shape[0], ctypes.c_void_p(B.ctypes.data),
                            B.shape[0])
    return B


def last_times():
    """(setup_s, iterations_s) of the last pcg() call."""
    t = np.zeros(2)
    load().oracle_last_times(_p(t))
    return float(t[0]), float(t[1])


# --- batched solver's factor preconditioners (batched.c) -----------------------

def _sig_batched(lib):
    sig = {
        "oracle_pcg_with": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _D, _RP, _P, _IP, _IP]),
        "oracle_pcg_mixed_with": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _D, _I, _D, _RP, _P, _IP, _IP, _IP]),
        "oracle_ic0_5pt": (_I, [_I, _I, _P, _P, _P]),
        "oracle_ic0_5pt_blocks": (_I, [_I, _I, _I, _P, _P, _P]),
        "oracle_ic0_grid_blocks": (_I, [_I, _I, _I, _P, _P, _P, _P]),
        "oracle_derive": (None, [_I, _I, _I, _P, _P, _P]),
        "oracle_precon_create_factor": (_P, [_I, _I, _P, _P]),
        "oracle_lscan_apply": (None, [_I, _I, _P, _P, _P]),
        "oracle_lscan_apply32": (None, [_I, _I, _P, _P, _P]),
        "oracle_lscan_apply_strips": (None, [_I, _I, _I, _P, _P, _P]),
        "oracle_f32_to_bf16": (ctypes.c_ushort, [ctypes.c_float]),
        "oracle_precon_factor_apply32": (None, [_P, _I]),
        "oracle_precon_factor_strips": (None, [_P, _I]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a


def _blib():
    lib = load()
    if not getattr(lib, "_batched_sig", False):
        _sig_batched(lib)
        lib._batched_sig = True
    return lib


def ic0_5pt(rowptr, val, nx, nblocks=1):
    """IC(0) factor (Ls, Lw, Ld) of a five-point matrix in natural order;
    nblocks > 1: block-Jacobi IC(0) over strips of grid lines."""
    n = rowptr.shape[0] - 1
    L = np.zeros(3 * n)
    if nblocks > 1:
        rc = _blib().oracle_ic0_5pt_blocks(n, nx, nblocks, _p(rowptr), _p(val), _p(L))
    else:
        rc = _blib().oracle_ic0_5pt(n, nx, _p(rowptr), _p(val), _p(L))
    if rc != 0:
        raise RuntimeError("IC(0) breakdown")
    return L[:n].copy(), L[n:2 * n].copy(), L[2 * n:].copy()


def ic0_grid(rowptr, col, val, nx, nblocks=1):
    """Block-Jacobi IC(0) factor (Ls, Lw, Ld) of a CRS matrix over strips of
    whole grid lines (nx cells per line), couplings found by column: the
    factor of the mixed solver's FP32 line-scan apply."""
    n = rowptr.shape[0] - 1
    L = np.zeros(3 * n)
    rc = _blib().oracle_ic0_grid_blocks(n, nx, nblocks, _p(rowptr), _p(col), _p(val), _p(L))
    if rc != 0:
        raise RuntimeError("IC(0) breakdown")
    return L[:n].copy(), L[n:2 * n].copy(), L[2 * n:].copy()


def derive(Ls, Lw, Ld, nx):
    """Apply arrays (F64, F32), each (4, n): ls, 1/Ld, a_fwd, a_bwd.  The
    factor is double (float64 arrays) or bf16 (uint16 bit patterns)."""
    n = Ls.shape[0]
    src = 1 if Ls.dtype == np.uint16 else 0
    L = np.ascontiguousarray(np.concatenate([Ls, Lw, Ld]))
    F64 = np.zeros(4 * n)
    F32 = np.zeros(4 * n, np.float32)
    _blib().oracle_derive(n, nx, src, _p(L), _p(F64), _p(F32))
    return F64, F32


def lscan_apply(F, r, nx, nstrip=1):
    """z = (L L^T)^-1 r with the GPU's lane-scan order (F float64 or float32)."""
    n = r.shape[0]
    z = np.zeros_like(r)
    if nstrip != 1:
        assert r.dtype == np.float64
        _blib().oracle_lscan_apply_strips(nx, n // nx, nstrip, _p(np.ascontiguousarray(F)),
                                          _p(np.ascontiguousarray(r)), _p(z))
        return z
    fn = _blib().oracle_lscan_apply32 if r.dtype == np.float32 else _blib().oracle_lscan_apply
    fn(nx, n // nx, _p(np.ascontiguousarray(F)), _p(np.ascontiguousarray(r)), _p(z))
    return z


class FactorPrecon:
    def __init__(self, F64, F32, nx, apply32=False, nstrip=1):
        """apply32: FP64 vectors through the FP32 apply (r rounded, z widened);
        nstrip: independent strips of grid lines."""
        self.n = F64.shape[0] // 4
        self.h = _blib().oracle_precon_create_factor(self.n, nx, _p(F64), _p(F32))
        if apply32:
            _blib().oracle_precon_factor_apply32(self.h, 1)
        if nstrip != 1:
            _blib().oracle_precon_factor_strips(self.h, nstrip)

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_precon_free(self.h)
            self.h = None


def pcg_factor(rowptr, col, val, b, M, x0=None, itrmax=1000, rtol=1e-8, red=TILED):
    """oracle PCG with a prebuilt preconditioner (FactorPrecon or Precon);
    returns (x, hist[:niter], niter, iret)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 1))
    niter, iret = ctypes.c_int(0), ctypes.c_int(0)
    rc = load().oracle_pcg_with(M.h, n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), itrmax, rtol,
                                 ctypes.byref(red), _p(hist), ctypes.byref(niter), ctypes.byref(iret))
    if rc != 0:
        raise RuntimeError("oracle_pcg_with failed: %d" % rc)
    return x, hist[: niter.value].copy(), niter.value, iret.value


def pcg_mixed_factor(rowptr, col, val, b, M, x0=None, itrmax=50, rtol=1e-10, inner_itrmax=5000, inner_rtol=1e-4,
                     red=TILED):
    """Mixed refinement with a FactorPrecon; returns (x, hist[:niter], niter, iret, inner_total)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    hist = np.zeros(max(itrmax, 1))
    niter, iret, tot = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    rc = _blib().oracle_pcg_mixed_with(M.h, n, _p(rowptr), _p(col), _p(val), _p(b), _p(x), itrmax, rtol,
                                       inner_itrmax, inner_rtol, ctypes.byref(red), _p(hist), ctypes.byref(niter),
                                       ctypes.byref(iret), ctypes.byref(tot))
    if rc != 0:
        raise RuntimeError("oracle_pcg_mixed_with failed: %d" % rc)
    return x, hist[: max(niter.value, 0)].copy(), niter.value, iret.value, tot.value


def f32_to_bf16(x):
    """Round float32 to bf16 bit patterns, nearest even (NaN -> quiet NaN)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def f32_to_bf16_c(v):
    """The C oracle's scalar rounding (bit pattern)."""
    return int(_blib().oracle_f32_to_bf16(float(v)))


def bf16_to_f32(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


# --- random-coefficient diffusion (the generator's restatement in numpy) -------

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix_u01(seed, idx):
    """u_i in [0,1): top 53 bits of SplitMix64 of seed + (i+1)*golden."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(n, seed):
    k = np.arange(n, dtype=np.uint64)
    p = k >> np.uint64(1)
    u1 = splitmix_u01(seed, 2 * p)
    u2 = splitmix_u01(seed, 2 * p + np.uint64(1))
    rad = np.sqrt(-2.0 * np.log(1.0 - u1))
    ang = 6.283185307179586 * u2
    return np.where((k & np.uint64(1)) == 1, rad * np.sin(ang), rad * np.cos(ang))


def grf_spectrum1d(N, ell):
    """S_N(k) = sum_i exp(-i^2/(2 ell^2)) e^{-2 pi i i k/N} by Poisson summation."""
    f = np.arange(N) / N
    m = np.arange(-1, 3)[:, None]
    return np.sqrt(2 * np.pi) * ell * np.exp(-2 * np.pi ** 2 * ell ** 2 * (f[None, :] - m) ** 2).sum(axis=0)


def grf_eigenvalues(nx, ny, ell, sigma):
    """Eigenvalues of the circulant SE covariance, (ny, nx)."""
    return sigma * sigma * np.outer(grf_spectrum1d(ny, ell), grf_spectrum1d(nx, ell))


def grf2d(nx, ny, ell, sigma, seed):
    """Circulant-embedding GRF with squared-exponential covariance on the torus."""
    lam = grf_eigenvalues(nx, ny, ell, sigma)
    sq = np.sqrt(lam)
    w = normal(nx * ny, seed).reshape(ny, nx)
    return np.fft.ifft2(sq * np.fft.fft2(w)).real.reshape(-1)


def diffusion2d(nx, ny, kappa):
    """Five-point harmonic-mean diffusion (1-based CSR), ghost kappa = cell kappa."""
    k = np.asarray(kappa, dtype=np.float64).reshape(ny, nx)

    def face(a, b):
        return 1.0 / (0.5 * (1.0 / a + 1.0 / b))

    ks = np.vstack([k[:1], k[:-1]])
    kw = np.hstack([k[:, :1], k[:, :-1]])
    ke = np.hstack([k[:, 1:], k[:, -1:]])
    kn = np.vstack([k[1:], k[-1:]])
    fs, fw, fe, fn = face(k, ks), face(k, kw), face(k, ke), face(k, kn)
    diag = ((fs + fw) + fe) + fn
    n = nx * ny
    rp = np.zeros(n + 1, np.int32)
    col = []
    val = []
    e = 0
    for y in range(ny):
        for x in range(nx):
            r = x + nx * y
            rp[r] = e + 1
            if y > 0:
                col.append(r - nx + 1); val.append(-fs[y, x]); e += 1
            if x > 0:
                col.append(r - 1 + 1); val.append(-fw[y, x]); e += 1
            col.append(r + 1); val.append(diag[y, x]); e += 1
            if x < nx - 1:
                col.append(r + 1 + 1); val.append(-fe[y, x]); e += 1
            if y < ny - 1:
                col.append(r + nx + 1); val.append(-fn[y, x]); e += 1
    rp[n] = e + 1
    return rp, np.array(col, np.int32), np.array(val, np.float64)
