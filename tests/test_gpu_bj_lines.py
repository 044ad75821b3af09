"""GPU parity of the line-wavefront form of the FP64 block-Jacobi apply
(precon.hip: line_setup, k_bj_lines) through the C ABI.

When every block is whole grid lines (strips of one plane, or whole planes)
and its ILU(0) / D-ILU factor rows couple only to the neighbouring cell of
the line and the same cell of the neighbouring line and plane (a 3D 7-point
operator, a 2D 5-point one), the apply runs as a wavefront over the block's
lines instead of level by level.  Each row still sums its couplings in the stored order of the
reference sweeps (solve_LU_crs src/matrix.f90:8548-8746, solve_ILU_crs
:4987-5163), so the solves equal the oracle's sequential ones bit for bit,
and one apply equals the level-scheduled kernel's
(PARCEL_HIP_BJ_LEVELS=1) bit for bit.  The cases cover one to eight
wavefronts per block (16 to 512 lines), symmetric and nonsymmetric
couplings, blocks of one plane or strip and of two to four whole planes, the
fused CG update (dcg) and the plain apply (BiCGSTAB, GMRES)."""
import os
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


# (generator, blocks, wavefronts per block): one plane (or strip) per block,
# then blocks of 2 to 4 whole planes (the plane neighbour through the LDS ring)
CASES = {
    "p3d_16_b16": (lambda: ph.poisson3d_spd(16, 16, 16), 16, 1),
    "p3d_48x100x6_b6": (lambda: ph.poisson3d_spd(48, 100, 6), 6, 2),
    "p2d_100x300_b2": (lambda: ph.poisson2d(100, 300, -1.0), 2, 4),
    "p2d_40x512_b1": (lambda: ph.poisson2d(40, 512, -1.0), 1, 8),
    "cd3d_20_b20": (lambda: ph.convdiff3d(20, 20, 20), 20, 1),
    "cd3d_12x130x3_b3": (lambda: ph.convdiff3d(12, 130, 3), 3, 4),
    "p3d_24_b8_planes3": (lambda: ph.poisson3d_spd(24, 24, 24), 8, 4),
    "p3d_16x100x4_b2_planes2": (lambda: ph.poisson3d_spd(16, 100, 4), 2, 4),
    "p3d_10x30x16_b4_planes4": (lambda: ph.poisson3d_spd(10, 30, 16), 4, 4),
    "cd3d_20_b10_planes2": (lambda: ph.convdiff3d(20, 20, 20), 10, 2),
    "cd3d_12x70x4_b2_planes2": (lambda: ph.convdiff3d(12, 70, 4), 2, 4),
}
SYM = ("p3d_16_b16", "p3d_48x100x6_b6", "p2d_100x300_b2", "p2d_40x512_b1", "p3d_24_b8_planes3",
       "p3d_16x100x4_b2_planes2", "p3d_10x30x16_b4_planes4")


def _levels_plan(rp, col, val, **kw):
    os.environ["PARCEL_HIP_BJ_LEVELS"] = "1"
    try:
        return ph.Plan(rp, col, val, **kw)
    finally:
        del os.environ["PARCEL_HIP_BJ_LEVELS"]


@pytest.mark.parametrize("ilu", [0, 1, 2])
@pytest.mark.parametrize("case", sorted(CASES))
def test_apply_equals_level_scheduled(case, ilu):
    gen, nb, waves = CASES[case]
    say("bj lines apply", case, "ILU_method", ilu)
    rp, col, val = gen()
    n = rp.shape[0] - 1
    r = ph.uniform(n, 71)
    kw = dict(ipreflag=2, ILU_method=ilu, precon_thblock=nb)
    with ph.Plan(rp, col, val, **kw) as P:
        st = P.stats()
        assert st["bj_lines"] == 1 and st["bj_blocks"] == nb
        s = np.zeros(n)
        P.precon(r, s)
        s2 = np.full(n, np.nan)
        P.precon(r, s2)  # a second apply on the same plan gives the same bits
    with _levels_plan(rp, col, val, **kw) as Q:
        assert Q.stats()["bj_lines"] == 0
        sl = np.zeros(n)
        Q.precon(r, sl)
    # the same bits, signed zeros included
    assert np.array_equal(s.view(np.uint64), sl.view(np.uint64)) and np.array_equal(s2.view(np.uint64), s.view(np.uint64))


@pytest.mark.parametrize("ilu", [0, 1, 2])
@pytest.mark.parametrize("case", SYM)
def test_pcg_bit_exact(case, ilu):
    gen, nb, _ = CASES[case]
    say("bj lines pcg", case, "ILU_method", ilu)
    rp, col, val = gen()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 73)
    x = np.zeros(n)
    hist, niter, iret = ph.dcg(x, b, rp, col, val, ipreflag=2, ILU_method=ilu, precon_thblock=nb, itrmax=3000,
                               rtolmax=1e-10)
    xo, ho, no, io = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=ilu, nblocks=nb, itrmax=3000, rtol=1e-10)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


@pytest.mark.parametrize("case", ["cd3d_20_b20", "cd3d_12x130x3_b3", "p2d_100x300_b2", "cd3d_20_b10_planes2",
                                  "cd3d_12x70x4_b2_planes2"])
def test_nonsymmetric_solvers_bit_exact(case):
    gen, nb, _ = CASES[case]
    rp, col, val = gen()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 79)
    say("bj lines bicgstab", case)
    x = np.zeros(n)
    hist, niter, iret = ph.dbicgstab(x, b, rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=nb,
                                     itrmax=2000, rtolmax=1e-10)
    xo, ho, no, io = orc.bicgstab(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=nb, itrmax=2000, rtol=1e-10)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)
    say("bj lines gmres", case)
    x = np.zeros(n)
    hist, niter, iret = ph.dgmres(x, b, rp, col, val, GMRES_m=20, GMRES_GSflag=1, ipreflag=2, ILU_method=1,
                                  precon_thblock=nb, itrmax=2000, rtolmax=1e-10)
    xo, ho, no, io = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=nb, itrmax=2000, rtol=1e-10,
                               m=20, gsflag=1)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


@pytest.mark.parametrize("gen,nb", [
    (lambda: ph.poisson3d_spd(16, 16, 16), 3),    # blocks that cut a plane
    (lambda: ph.poisson2d(64, 64, -1.0), 7),      # blocks that cut a line
], ids=["cut_plane", "cut_line"])
def test_other_blocks_keep_level_schedule(gen, nb):
    rp, col, val = gen()
    with ph.Plan(rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=nb) as P:
        assert P.stats()["bj_lines"] == 0
    with ph.Plan(rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=nb, independ_nvec=2) as P:
        assert P.stats()["bj_lines"] == 0
