"""GPU parity tests of the eigen entry points through the C ABI:
parcel_dpower (src/wrapper.f90:913-1022) and parcel_dcaarnoldi (:1561-1680),
their DIA forms (:1025, :1683) and the plan-level power method, all running
power_method_symMatrix (src/power_method.f90:188-437).  Emax, the eigenpair
error and the eigenvector are bit-exact against the oracle (device dots in
the tiled order).  The CRS forms return the eigenvector's first component
only (src/krylov.f90:360-369), the DIA forms all n (get_EigenPair_KSP,
src/EigenProblem.f90:18-46); parcel_dcaarnoldi runs one iteration, its
power_method_itrmax being left at -1 and raised to 1 (src/wrapper.f90:24,
src/krylov.f90:32, src/type.f90:1331-1336)."""
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


PROBLEMS = {
    "parcel3d_14": lambda: ph.poisson3d_parcel(14, 12, 10),
    "spd3d_20": lambda: ph.poisson3d_spd(20, 20, 20),
    "convdiff_10": lambda: convdiff3d(10, peclet=0.3)[:3],
    "poisson2d_760": lambda: ph.poisson2d(760, 760, -1.0),  # n = 577600, many tiles per block
}

PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_4": dict(ipreflag=2, ILU_method=0, precon_thblock=4),
    "ilud_8": dict(ipreflag=2, ILU_method=1, precon_thblock=8),
}

SENTINEL = -7.25


def _oracle(rp, col, val, x0, pkw, itr):
    return orc.power_eigen(rp, col, val, x0, ipreflag=pkw["ipreflag"], ilu_method=pkw.get("ILU_method", 1),
                           nblocks=pkw.get("precon_thblock", 1), itrmax=itr)


def _kw(pkw):
    return dict(ipreflag=pkw["ipreflag"], ILU_method=pkw.get("ILU_method", 1),
                precon_thblock=pkw.get("precon_thblock", 1))


def _check_crs(out, e, en, vec):
    assert out["EigenValueReal"] == [e] and out["EigenErrorNorm"] == [en]
    assert out["EigenValueImag"] == [0.0]
    assert (out["MaxEigenID"], out["MinEigenID"], out["iret"]) == (1, 1, 0)
    v = out["EigenVector"]
    assert v[0] == vec[0] and np.all(v[1:] == SENTINEL)


@pytest.mark.parametrize("prob", sorted(PROBLEMS))
@pytest.mark.parametrize("pre", sorted(PRECONS))
def test_dpower_crs_and_dia_match_oracle(prob, pre):
    rp, col, val = PROBLEMS[prob]()
    pkw = PRECONS[pre]
    n = rp.shape[0] - 1
    rng = np.random.default_rng(21)
    x0 = rng.standard_normal(n)
    itr = 40
    say("dpower", prob, pre, "n", n, "itr", itr)
    e, en, vec = _oracle(rp, col, val, x0, pkw, itr)
    x = x0.copy()
    out = ph.dpower(x, rp, col, val, power_method_itrmax=itr, fill=SENTINEL, **_kw(pkw))
    assert np.array_equal(x, x0)
    assert out["itrmax"] == 0
    _check_crs(out, e, en, vec)
    off, dia = ph.crs_to_dia(rp, col, val)
    say("dpower_dia", prob, pre, "nnd", off.size)
    out = ph.dpower(x, dia=(off, dia), power_method_itrmax=itr, fill=SENTINEL, **_kw(pkw))
    assert out["EigenValueReal"] == [e] and out["EigenErrorNorm"] == [en]
    assert np.array_equal(out["EigenVector"], vec)
    assert (out["MaxEigenID"], out["MinEigenID"], out["itrmax"], out["iret"]) == (1, 1, 0, 0)


def test_dcaarnoldi_is_one_power_step():
    rp, col, val = ph.poisson3d_parcel(16, 16, 16)
    n = rp.shape[0] - 1
    x0 = np.random.default_rng(3).standard_normal(n)
    for pre in ("none", "ilu0_4"):
        pkw = PRECONS[pre]
        say("dcaarnoldi", pre)
        e, en, vec = _oracle(rp, col, val, x0, pkw, 1)
        out = ph.dpower(x0, rp, col, val, caarnoldi=(4, 10, 1, 3), fill=SENTINEL, **_kw(pkw))
        _check_crs(out, e, en, vec)
        assert out["itrmax"] is None
        off, dia = ph.crs_to_dia(rp, col, val)
        out = ph.dpower(x0, dia=(off, dia), caarnoldi=(4, 10, 1, 3), **_kw(pkw))
        assert out["EigenValueReal"] == [e] and out["EigenErrorNorm"] == [en]
        assert np.array_equal(out["EigenVector"], vec)


def test_power_zero_start_and_itrmax_floor():
    rp, col, val = ph.poisson3d_spd(12, 12, 12)
    n = rp.shape[0] - 1
    say("dpower zero start")
    e, en, vec = _oracle(rp, col, val, np.zeros(n), PRECONS["pj"], 25)
    out = ph.dpower(np.zeros(n), rp, col, val, power_method_itrmax=25, fill=SENTINEL, ipreflag=1)
    _check_crs(out, e, en, vec)
    # power_method_itrmax < 1 runs one iteration
    e1, en1, vec1 = _oracle(rp, col, val, np.ones(n), PRECONS["none"], 1)
    for itr in (0, -5):
        out = ph.dpower(np.ones(n), rp, col, val, power_method_itrmax=itr, fill=SENTINEL)
        _check_crs(out, e1, en1, vec1)


def test_plan_eigen_full_vector_host_and_device():
    rp, col, val = ph.poisson3d_parcel(24, 20, 16)
    n = rp.shape[0] - 1
    x0 = np.random.default_rng(8).standard_normal(n)
    pkw = PRECONS["ilud_8"]
    e, en, vec = _oracle(rp, col, val, x0, pkw, 60)
    say("plan eigen host")
    with ph.Plan(rp, col, val, ipreflag=2, ILU_method=1, precon_thblock=8) as plan:
        ge, gen, gvec = plan.eigen(x0, 60)
        assert (ge, gen) == (e, en) and np.array_equal(gvec, vec)
        say("plan eigen device")
        xd = ph.DeviceArray(x0)
        vd = ph.DeviceArray(n=n)
        ge, gen, _ = plan.eigen(xd, 60, vec=vd)
        assert (ge, gen) == (e, en)
        assert np.array_equal(vd.numpy(), vec)
        assert np.array_equal(xd.numpy(), x0)
        # the plan still solves after the power method used its work vectors
        b = np.ones(n)
        x = np.zeros(n)
        hist, niter, iret = plan.solve(x, b, 500, 1e-8)
        xo, ho, no, io = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=8, itrmax=500, rtol=1e-8)[:4]
        assert (niter, iret) == (no, io) and np.array_equal(hist, ho) and np.array_equal(x, xo)
