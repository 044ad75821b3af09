"""CPU tests of the learned-factor network's oracle: its backward pass is
pinned by central finite differences of its own FP64 forward (no BF16
rounding), its factor by closed forms at the untrained state."""
import numpy as np

import mpnn
import oracle as orc
import parcel_hip as ph


def small(nx=8, ny=6, seed=3):
    k = np.exp(ph.grf2d(8, 8, 2.0, 1.0, seed)[: nx * ny])
    rp, col, val = ph.diffusion2d(nx, ny, k)
    return rp, col, val


def random_params(seed=1):
    rng = np.random.default_rng(seed)
    p = mpnn.init_params(seed, scale=0.3).astype(np.float64)
    p[mpnn.OFF_WO: mpnn.OFF_WO + mpnn.NO * mpnn.C] = rng.standard_normal(mpnn.NO * mpnn.C) * 0.01
    p[mpnn.OFF_BO:] = rng.standard_normal(mpnn.NO) * 0.1
    for l in range(mpnn.NL):
        p[l * mpnn.LAYER + 2 * mpnn.C * mpnn.C: (l + 1) * mpnn.LAYER] = rng.standard_normal(mpnn.C) * 0.1
    return p


def test_scaled_couplings_closed_form():
    nx, ny = 8, 6
    rp, col, val = ph.diffusion2d(nx, ny, np.full(nx * ny, 2.0))
    ah, sd = mpnn.scaled_couplings(rp, col, val, nx)
    # constant kappa: interior row diag 8, off -2 -> -1/4; boundary diag larger
    i = 3 + nx * 2
    assert np.allclose(ah[i], -0.25) and np.isclose(sd[i], np.sqrt(8.0))
    assert ah[0, 0] == 0 and ah[0, 1] == 0 and ah[0, 2] < 0


def test_untrained_factor_is_unit_lower_plus_couplings():
    nx, ny = 8, 6
    rp, col, val = small(nx, ny)
    ah, sd = mpnn.scaled_couplings(rp, col, val, nx)
    p = mpnn.init_params(0)  # zero head
    fw = mpnn.forward(p, ah, nx)
    Ls, Lw, Ld = fw["Lh"]
    assert np.allclose(Ld, np.sqrt(1.0 + 1e-6))
    assert np.array_equal(Lw, ah[:, 1].astype(np.float64)) and np.array_equal(Ls, ah[:, 0].astype(np.float64))


def test_backward_matches_finite_differences():
    nx, ny = 8, 6
    rp, col, val = small(nx, ny)
    ah, _ = mpnn.scaled_couplings(rp, col, val, nx)
    rng = np.random.default_rng(0)
    p = random_params()
    xh, bh = rng.standard_normal(nx * ny), rng.standard_normal(nx * ny)

    def loss(q):
        return mpnn.loss_and_grad_factor(mpnn.forward(q, ah, nx, bf=False)["Lh"], xh, bh, nx)[0]

    fw = mpnn.forward(p, ah, nx, bf=False)
    _, G = mpnn.loss_and_grad_factor(fw["Lh"], xh, bh, nx)
    g = mpnn.backward(p, ah, nx, fw, G, bf=False)
    idx = list(rng.choice(mpnn.NPARAM, 150, replace=False)) + list(range(mpnn.OFF_WO, mpnn.NPARAM))
    checked = 0
    for j in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[j] += e
        pm[j] -= e
        fd = (loss(pp) - loss(pm)) / (2 * e)
        if abs(fd) > 1e-6:
            checked += 1
            assert abs(fd - g[j]) <= 1e-4 * abs(fd), (j, fd, g[j])
    assert checked > 100


def test_backward_plants_a_wrong_head_derivative_and_fails():
    # the finite-difference check above catches a dropped factor: doubling
    # the diagonal output's derivative (a mistake this oracle once had)
    nx, ny = 8, 6
    rp, col, val = small(nx, ny)
    ah, _ = mpnn.scaled_couplings(rp, col, val, nx)
    rng = np.random.default_rng(0)
    p = random_params()
    xh, bh = rng.standard_normal(nx * ny), rng.standard_normal(nx * ny)
    fw = mpnn.forward(p, ah, nx, bf=False)
    _, G = mpnn.loss_and_grad_factor(fw["Lh"], xh, bh, nx)
    G2 = G.copy()
    G2[2] *= 2.0
    g = mpnn.backward(p, ah, nx, fw, G, bf=False)
    gw = mpnn.backward(p, ah, nx, fw, G2, bf=False)
    j = mpnn.OFF_BO  # bias of the diagonal output
    assert abs(gw[j] - g[j]) > 1e-3 * abs(g[j])


def test_bf16_forward_close_to_fp64():
    nx, ny = 16, 8
    rp, col, val = ph.diffusion2d(nx, ny, np.exp(ph.grf2d(16, 8, 3.0, 1.0, 5)))
    ah, _ = mpnn.scaled_couplings(rp, col, val, nx)
    p = random_params(2)
    a = mpnn.forward(p, ah, nx, bf=True)["Lh"]
    b = mpnn.forward(p, ah, nx, bf=False)["Lh"]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 2e-2


# --- pinned by an independent dense restatement (PyTorch autograd) -----------
# tests/golden/make_mpnn_fixtures.py writes the fixture: forward, loss and
# parameter gradient of the network on a 16 x 8 grid, computed with dense
# matrices (A_hat = D^-1/2 A D^-1/2 off its diagonal, Lh as a dense lower
# triangle) instead of the oracle's stencil loops.

import os  # noqa: E402

import scipy.sparse as sp  # noqa: E402

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mpnn_16x8.npz")


def _fix():
    return dict(np.load(FIX, allow_pickle=False))


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _oracle_all(f, bf):
    nx = int(f["nx"])
    ah, sd = mpnn.scaled_couplings(f["rowptr"], f["col"], f["val"], nx)
    fw = mpnn.forward(f["params"], ah, nx, bf=bf)
    loss, G = mpnn.loss_and_grad_factor(fw["Lh"], f["xh"], f["bh"], nx)
    return ah, fw, loss, G


def test_oracle_forward_loss_grad_match_dense_autograd():
    f = _fix()
    nx = int(f["nx"])
    ah, fw, loss, G = _oracle_all(f, bf=False)
    Ls, Lw, Ld = fw["Lh"]
    assert _rel(Ld, f["Ld"]) < 1e-12 and _rel(Lw, f["Lw"]) < 1e-12 and _rel(Ls, f["Ls"]) < 1e-12
    assert _rel(fw["o"], f["o"]) < 1e-12 and _rel(fw["hs"][-1], f["hlast"]) < 1e-12
    assert abs(loss - float(f["loss"])) <= 1e-12 * float(f["loss"])
    g = mpnn.backward(f["params"], ah, nx, fw, G, bf=False)
    gt = f["grad"]
    for l in range(mpnn.NL):
        sl = slice(l * mpnn.LAYER, (l + 1) * mpnn.LAYER)
        assert _rel(g[sl], gt[sl]) < 1e-10, l
    sl = slice(mpnn.OFF_WO, mpnn.NPARAM)
    assert _rel(g[sl], gt[sl]) < 1e-10


def test_oracle_bf16_forward_matches_dense_autograd():
    # the same rounding points (FP32 then nearest-even BF16) in torch's
    # bfloat16; float64 sums in another order may flip a rare rounding
    f = _fix()
    _, fw, loss, _ = _oracle_all(f, bf=True)
    for k, name in enumerate(("Ls_bf", "Lw_bf", "Ld_bf")):
        assert _rel(fw["Lh"][k], f[name]) < 1e-6, name
    assert _rel(fw["hs"][-1], f["hlast_bf"]) < 1e-6
    assert abs(loss - float(f["loss_bf"])) <= 1e-6 * float(f["loss_bf"])


def test_aggregation_and_loss_against_scipy():
    f = _fix()
    nx = int(f["nx"])
    rp, col, val = f["rowptr"], f["col"], f["val"]
    n = rp.shape[0] - 1
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))
    d = A.diagonal()
    Ahat = sp.diags(1 / np.sqrt(d)) @ (A - sp.diags(d)) @ sp.diags(1 / np.sqrt(d))
    ah, fw, loss, _ = _oracle_all(f, bf=False)
    # m = A_hat h for every layer input (couplings rounded to FP32 as handed over)
    Ahat32 = Ahat.astype(np.float32).astype(np.float64)
    for h, m in zip(fw["hs"][:-1], fw["ms"]):
        assert np.allclose(m, Ahat32 @ h, rtol=0, atol=1e-13 * np.abs(h).max())
    Ls, Lw, Ld = fw["Lh"]
    i = np.arange(n)
    w, s = i % nx > 0, i >= nx
    Lh = sp.csr_matrix((np.concatenate([Ld, Lw[w], Ls[s]]),
                        (np.concatenate([i, i[w], i[s]]), np.concatenate([i, i[w] - 1, i[s] - nx]))), shape=(n, n))
    r = Lh @ (Lh.T @ f["xh"]) - f["bh"]
    assert abs(loss - r @ r / (f["bh"] @ f["bh"])) <= 1e-12 * loss


def test_planted_aggregation_mistake_fails(monkeypatch):
    # swapping the west and east neighbours in the message sums must break
    # agreement with the dense restatement
    f = _fix()
    good = mpnn.neighbours

    def swapped(n, nx):
        idx = good(n, nx)
        return idx[:, [0, 2, 1, 3]]

    monkeypatch.setattr(mpnn, "neighbours", swapped)
    _, fw, loss, _ = _oracle_all(f, bf=False)
    assert _rel(fw["hs"][-1], f["hlast"]) > 1e-3 or _rel(fw["Lh"][2], f["Ld"]) > 1e-3


def test_planted_transposed_loss_fails(monkeypatch):
    # Lh^T Lh instead of Lh Lh^T in the loss must be caught
    f = _fix()
    lt, la = mpnn.lt_apply, mpnn.l_apply
    monkeypatch.setattr(mpnn, "lt_apply", la)
    monkeypatch.setattr(mpnn, "l_apply", lt)
    _, _, loss, _ = _oracle_all(f, bf=False)
    assert abs(loss - float(f["loss"])) > 1e-6 * float(f["loss"])
