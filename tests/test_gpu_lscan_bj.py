"""GPU parity of the mixed solver's FP32 block-Jacobi apply in line-scan form
(precon.hip: lscan_setup; batched.hip: k_lscan) through the C ABI.  When the
blocks are strips of whole grid lines coupled only within a line and to the
line below, symmetric bit for bit, the FP32 inner apply is the block IC(0)
factor solved by line scans (block ILU(0) of src/matrix.f90:8249-8392 in
exact arithmetic); the whole mixed solve must equal the oracle's mixed
refinement with that factor (oracle ic0_grid + FactorPrecon strips) bit for
bit.  Everything else keeps the level-scheduled FP32 ILU(0) and must equal
the oracle's sequential one: nonsymmetric couplings, blocks of several
planes, blocks that cut a line, independ_nvec > 1, D-ILU."""
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def _kappa(nx, ny, seed):
    return np.exp(ph.grf2d(512, 512, 4.0, 1.5, seed).reshape(512, 512)[:ny, :nx].ravel())


def _nonsym():
    # a diffusion matrix with one coupling changed in its last bit: not
    # symmetric bit for bit, so the line-scan form does not apply
    rp, col, val = ph.diffusion2d(40, 40, _kappa(40, 40, 9))
    r = 500
    for e in range(rp[r] - 1, rp[r + 1] - 1):
        if col[e] - 1 == r - 1:
            val[e] = np.nextafter(val[e], 0.0)
    return rp, col, val


# (generator, nx, blocks)
LSCAN = {
    "p2d_64x200_b2": (lambda: ph.poisson2d(64, 200, -1.0), 64, 2),
    "p2d_8x300_b1": (lambda: ph.poisson2d(8, 300, -1.0), 8, 1),
    "diff2d_100x60_b4": (lambda: ph.diffusion2d(100, 60, _kappa(100, 60, 5)), 100, 4),
    "diff2d_512x24_b3": (lambda: ph.diffusion2d(512, 24, _kappa(512, 24, 6)), 512, 3),
    "p3d_16_b16": (lambda: ph.poisson3d_spd(16, 16, 16), 16, 16),
    "p3d_64x64x12_b12": (lambda: ph.poisson3d_spd(64, 64, 12), 64, 12),
    "p3d_256x32x4_b4": (lambda: ph.poisson3d_spd(256, 32, 4), 256, 4),
}

# (generator, blocks, ILU_method, independ_nvec): the level-scheduled FP32 apply
SEQUENTIAL = {
    "nonsym_40x40_b4": (_nonsym, 4, 0, 1),
    "p3d_24_b8_planes3": (lambda: ph.poisson3d_spd(24, 24, 24), 8, 0, 1),
    "p2d_64x64_b7_cut": (lambda: ph.poisson2d(64, 64, -1.0), 7, 0, 1),
    "p2d_64x64_nvec2": (lambda: ph.poisson2d(64, 64, -1.0), 4, 0, 2),
    "p2d_64x64_ilud": (lambda: ph.poisson2d(64, 64, -1.0), 4, 1, 1),
}


def _mixed(rp, col, val, b, nb, ilu=0, nvec=1):
    p = ph.default_param(solver=11, ipreflag=2, ILU_method=ilu, precon_thblock=nb, independ_nvec=nvec)
    with ph.Plan(rp, col, val, param=p) as P:
        P.set_inner(1e-4, 3000)
        x = np.zeros_like(b)
        hist, niter, iret = P.solve(x, b, 30, 1e-10)
        st = P.stats()
    return x, hist, niter, iret, st


@pytest.mark.parametrize("case", sorted(LSCAN))
def test_lscan_mixed_bit_exact(case):
    gen, nx, nb = LSCAN[case]
    say("lscan mixed", case)
    rp, col, val = gen()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 61)
    x, hist, niter, iret, st = _mixed(rp, col, val, b, nb)
    assert st["bj_lscan"] == 1
    F64, F32 = orc.derive(*orc.ic0_grid(rp, col, val, nx, nblocks=nb), nx)
    M = orc.FactorPrecon(F64, F32, nx, nstrip=nb)
    xo, ho, no, io, inner = orc.pcg_mixed_factor(rp, col, val, b, M, itrmax=30, rtol=1e-10, inner_itrmax=3000,
                                                 inner_rtol=1e-4)
    assert (niter, iret) == (no, io) and iret == 0
    assert st["inner_niter"] == inner
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


@pytest.mark.parametrize("case", sorted(SEQUENTIAL))
def test_level_scheduled_fp32_kept_elsewhere(case):
    gen, nb, ilu, nvec = SEQUENTIAL[case]
    say("sequential mixed", case)
    rp, col, val = gen()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 67)
    x, hist, niter, iret, st = _mixed(rp, col, val, b, nb, ilu, nvec)
    assert st["bj_lscan"] == 0
    xo, ho, no, io, inner = orc.pcg_mixed(rp, col, val, b, ipreflag=2, ilu_method=ilu, nblocks=nb,
                                          independ_nvec=nvec, itrmax=30, rtol=1e-10, inner_itrmax=3000,
                                          inner_rtol=1e-4)
    assert (niter, iret) == (no, io)
    assert st["inner_niter"] == inner
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


def test_fp64_solver_keeps_the_sequential_factor():
    # the same plan for FP64 PCG: the line-scan form is the mixed solver's only
    rp, col, val = ph.poisson3d_spd(16, 16, 16)
    with ph.Plan(rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=16) as P:
        assert P.stats()["bj_lscan"] == 0
