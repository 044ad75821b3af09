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
