"""CPU check that tests/golden/configs.json (the oracle's results on the
full-size benchmark configurations, which the GPU tests compare against)
still is what the oracle computes: a prefix of every history is recomputed
here at the full size.  For CG the history entries before the limit do not
depend on itrmax; for GMRES and CA-GMRES the limit cuts the last cycle short
and the entries before it stay the same."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_config_fixtures as F  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "configs.json")))


def test_fixture_entries_complete():
    for k in ("c2_gmres30", "c2_cagmres_s8t4_basis1", "c2_cagmres_s8t4_basis2", "c3_ddcg", "c4_bj", "c4_mixed_bj"):
        g = GOLD[k]
        assert len(g["hist"]) == g["niter"] and g["iret"] == 0, k
    # the reference's Newton step with complex Ritz shifts stops on an estimate
    # the true residual does not follow; the consistent step reaches rtol
    assert GOLD["c2_cagmres_s8t4_basis1"]["true_rel_res"] > 1e-6
    assert GOLD["c2_cagmres_s8t4_basis2"]["true_rel_res"] < 1e-9
    # CA-GMRES(8, 4) with a consistent basis is GMRES(32) in exact arithmetic:
    # its history follows GMRES(30)'s over the first restart cycle
    h30, h32 = np.array(GOLD["c2_gmres30"]["hist"][:30]), np.array(GOLD["c2_cagmres_s8t4_basis2"]["hist"][:30])
    assert np.allclose(h32, h30, rtol=1e-6)


@pytest.fixture(scope="module")
def c2():
    return F.c2_problem()


def test_c2_gmres_prefix(c2):
    rp, col, val, b = c2
    c = F.C2_GMRES
    _, h, ni, ir = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], itrmax=35,
                             rtol=c["rtol"], m=c["m"], gsflag=c["gsflag"])
    assert ni == 35 and np.array_equal(h, np.array(GOLD["c2_gmres30"]["hist"][:35]))


@pytest.mark.parametrize("basis", [1, 2])
def test_c2_cagmres_prefix(c2, basis):
    rp, col, val, b = c2
    c = F.C2_CAGMRES
    _, h, ni, ir = orc.cagmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], itrmax=40,
                               rtol=c["rtol"], s=c["s"], t=c["t"], basis=basis, qrflag=c["qrflag"])
    g = np.array(GOLD["c2_cagmres_s8t4_basis%d" % basis]["hist"])
    assert np.array_equal(h[:40], g[:40])


def test_c3_ddcg_prefix():
    rp, col, val, bh, bl = F.c3_problem()
    c = F.C3_DDCG
    _, _, h, ni, ir = orc.ddcg(rp, col, val, None, bh, bl, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], prec=2,
                               itrmax=25, rtol=c["rtol"])
    assert ni == 25 and np.array_equal(h, np.array(GOLD["c3_ddcg"]["hist"][:25]))


def test_c4_bj_prefix():
    rp, col, val, b = F.c4_problem()
    c = F.C4_BJ
    _, h, ni, ir = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], itrmax=4, rtol=c["rtol"])
    assert ni == 4 and np.array_equal(h, np.array(GOLD["c4_bj"]["hist"][:4]))
