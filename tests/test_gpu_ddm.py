"""GPU tests of the stencil (DDM) entry point parcel_hip_dcg_ddm: bit for bit
against the oracle PCG on the box-major CRS of the same problem (block Jacobi
= D-ILU on the GPU build's boxes, point Jacobi, none), and the reference's
own GPU-build golden history (src/example/template_gpu/STDOUT_run_ddm: 160^3,
block Jacobi, rtol 1e-9, 345 iterations)."""
import json
import os

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import ddm_box_crs, ddm_ras

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


@pytest.mark.parametrize("grid,ipre", [((16, 16, 16), 2), ((13, 10, 9), 2), ((12, 12, 12), 1), ((10, 9, 8), 0)])
def test_dcg_ddm_matches_oracle(grid, ipre):
    print("dcg_ddm", grid, ipre, flush=True)
    d = ph.poisson3d_ddm(*grid)
    x = np.zeros(d["m"])
    hist, niter, iret = ph.dcg_ddm(x, d["b"], grid, 1, d["val_dia"], d["ioff_dia"], d["ioff_grids"], ipreflag=ipre,
                                   itrmax=2000, rtolmax=1e-9)
    rp, col, val, perm, starts = ddm_box_crs(d)
    M = orc.Precon(rp, col, val, ipre, ilu_method=1, starts=starts if ipre == 2 else None)
    xo, ho, no, io = orc.pcg_factor(rp, col, val, d["b"][perm].copy(), M, itrmax=2000, rtol=1e-9)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho)
    assert np.array_equal(x[perm], xo)
    mask = np.ones(d["m"], bool)
    mask[perm] = False
    assert np.all(x[mask] == 0)  # ghosts untouched


def test_dcg_ddm_reference_gpu_golden():
    print("dcg_ddm golden 160^3", flush=True)
    gold = json.load(open(os.path.join(HERE, "golden", "templates.json")))["template_gpu/STDOUT_run_ddm"]
    ref = np.array([v[1] for v in gold["histories"][0]])
    d = ph.poisson3d_ddm(160, 160, 160)
    x = np.zeros(d["m"])
    hist, niter, iret = ph.dcg_ddm(x, d["b"], (160, 160, 160), 1, d["val_dia"], d["ioff_dia"], d["ioff_grids"],
                                   ipreflag=2, itrmax=2000, rtolmax=1e-9)
    print("niter", niter, "reference", ref.size, flush=True)
    assert iret == 0 and niter == ref.size == 345
    rel = np.abs(hist - ref) / ref
    print("max rel dev", rel.max(), flush=True)
    assert rel.max() <= 1e-12


@pytest.mark.parametrize("grid,addL", [((16, 16, 16), 1), ((16, 16, 16), 2), ((13, 10, 9), 1), ((24, 8, 1), 3)])
def test_dcg_ddm_additive_schwarz_matches_oracle(grid, addL):
    # ipreflag 3: restricted additive Schwarz over the boxes extended by addL
    print("dcg_ddm AS", grid, addL, flush=True)
    d = ph.poisson3d_ddm(*grid)
    x = np.zeros(d["m"])
    hist, niter, iret = ph.dcg_ddm(x, d["b"], grid, 1, d["val_dia"], d["ioff_dia"], d["ioff_grids"], ipreflag=3,
                                   itrmax=2000, rtolmax=1e-9, addL=addL)
    rp, col, val, perm, starts = ddm_box_crs(d)
    src, dst, rpe, cole, vale, st = ddm_ras(d, addL)
    M = orc.RasPrecon(perm.size, src, dst, rpe, cole, vale, st)
    xo, ho, no, io = orc.pcg_factor(rp, col, val, d["b"][perm].copy(), M, itrmax=2000, rtol=1e-9)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho)
    assert np.array_equal(x[perm], xo)


def test_dcg_ddm_rejects_bad_overlap():
    d = ph.poisson3d_ddm(8, 8, 8)
    x = np.zeros(d["m"])
    # 8 cells make two boxes of 4: an overlap of 5 reaches past the grid
    with pytest.raises(ph.ParcelError) as e:
        ph.dcg_ddm(x, d["b"], (8, 8, 8), 1, d["val_dia"], d["ioff_dia"], d["ioff_grids"], ipreflag=3, addL=5)
    assert e.value.code == ph.ERR_UNSUPPORTED
    with pytest.raises(ph.ParcelError) as e:
        ph.dcg_ddm(x, d["b"], (8, 8, 8), 1, d["val_dia"], d["ioff_dia"], d["ioff_grids"], ipreflag=3, addL=-1)
    assert e.value.code == ph.ERR_ARG
