"""CPU tests of the C ABI: the library loads, exports every symbol the public
header declares, its host-side generators work, and the entry points report
errors (instead of crashing) when no device is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import parcel_hip as ph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parcel_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(parcel_(?:hip|gen)_\w+)\s*\(", text)))


FORTRAN_HEADER = os.path.join(ROOT, "include", "parcel_fortran.h")


def test_by_reference_symbols_exported():
    text = re.sub(r"/\*.*?\*/", "", open(FORTRAN_HEADER).read(), flags=re.S)
    syms = sorted(set(re.findall(r"\b(parcel_\w+_c_)\s*\(", text)))
    assert syms == ["parcel_dbicgstab_c_", "parcel_dbicgstab_dia_c_", "parcel_dcaarnoldi_c_",
                    "parcel_dcaarnoldi_dia_c_", "parcel_dcagmres_c_", "parcel_dcagmres_dia_c_", "parcel_dcbcg_c_",
                    "parcel_dcbcg_dia_c_", "parcel_dcg_c_", "parcel_dcg_ddm_c_", "parcel_dcg_dia_c_", "parcel_ddcg_c_",
                    "parcel_dgmres_c_", "parcel_dgmres_dia_c_", "parcel_free_c_", "parcel_init_c_"]
    lib = ph.load()
    for s in syms:
        assert hasattr(lib, s), s


def test_devicemem_symbols_exported():
    # the GPU-build drivers' convention: the same names, in their own library
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "parcel_devicemem.h")).read(), flags=re.S)
    syms = sorted(set(re.findall(r"\b(parcel_\w+_c_)\s*\(", text)) - {"parcel_init_c_", "parcel_free_c_"})
    assert len(syms) == 7
    dev = ctypes.CDLL(os.path.join(ROOT, "parcel-legacy_amd", "lib", "libparcel_devicemem.so"))
    for s in syms:
        assert hasattr(dev, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "parcel-legacy_amd", "lib",
                                                                      "libparcel_devicemem.so")],
                         capture_output=True, text=True)
    defined = sorted(set(re.findall(r" T (parcel_\w+_c_)", out.stdout)))
    assert defined == syms


def test_fortran_module_built():
    lib_dir = os.path.join(ROOT, "parcel-legacy_amd", "lib")
    assert os.path.exists(os.path.join(lib_dir, "libparcel_fortran.so"))
    mod = open(os.path.join(lib_dir, "parcel_wrapper.mod"), "rb").read().lower()
    for name in (b"parcel_dcg", b"parcel_dbicgstab", b"parcel_dgmres", b"parcel_dcagmres", b"parcel_dcbcg",
                 b"parcel_ddcg", b"parcel_dcg_ddm", b"parcel_dpower", b"parcel_dcaarnoldi", b"parcel_init",
                 b"parcel_free", b"parcel_dcg_dia", b"parcel_dbicgstab_dia", b"parcel_dgmres_dia",
                 b"parcel_dcagmres_dia", b"parcel_dcbcg_dia"):
        assert name in mod


def test_header_symbols_exported_and_bound():
    lib = ph.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
        assert s in ph.SYMBOLS, "binding missing for %s" % s


def test_generators_poisson3d_parcel():
    rp, col, val = ph.poisson3d_parcel(5, 4, 3)
    n = 60
    assert rp[0] == 1 and rp[-1] - 1 == col.shape[0] == 7 * n - 2 * (4 * 3 + 5 * 3 + 5 * 4)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))
    assert np.all(A.diagonal() == -6.0)
    assert (A != A.T).nnz == 0
    # columns sorted in every row, entry order b, s, w, c, e, n, t
    for i in range(n):
        c = col[rp[i] - 1: rp[i + 1] - 1]
        assert np.all(np.diff(c) > 0)


def test_generator_rhs_matches_exact_solution():
    # b = A x_exact up to the boundary terms folded into b: A x_exact - b is
    # O(h^8) small relative to b (problem.f90:772-875 builds b for x = h^6 x^2 y^2 z^2)
    g = 12
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, xe = ph.poisson3d_rhs_parcel(g, g, g)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(g ** 3, g ** 3))
    assert np.abs(A @ xe - b).max() <= 1e-12 * np.abs(b).max()


def test_generator_poisson2d_scaled():
    rp, col, val = ph.poisson2d(8, 6, -1.0)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(48, 48))
    assert np.all(A.diagonal() == 4.0)
    assert set(np.unique(A.data)) == {-1.0, 4.0}
    assert A.nnz == 5 * 48 - 2 * 8 - 2 * 6


def test_uniform_stream_is_counter_based():
    a = ph.uniform(1000, 3)
    b = ph.uniform(10, 3)
    assert np.array_equal(a[:10], b)
    assert a.min() >= -1.0 and a.max() < 1.0
    assert abs(a.mean()) < 0.1
    assert not np.array_equal(ph.uniform(10, 4), b)
    # SplitMix64 known answer: seed 0, first output 0xE220A8397B1DCDAF
    z = (0 + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    z ^= z >> 31
    assert z == 0xE220A8397B1DCDAF
    assert ph.uniform(1, 0, 0.0, 1.0)[0] == (z >> 11) / 2.0 ** 53


def test_entry_points_fail_cleanly_without_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    lib = ph.load()
    assert lib.parcel_hip_init(0) != 0
    assert lib.parcel_hip_last_error()
    rp, col, val = ph.poisson2d(4, 4, -1.0)
    with pytest.raises(ph.ParcelError):
        ph.Plan(rp, col, val, ipreflag=1)
