"""GPU tests of the learned factor: the network's BF16 MFMA forward and
backward kernels against the oracle network (tolerances from FP32
accumulation order and BF16 rounding flips, stated per check), the PCG with
the predicted factor bit for bit against the oracle PCG on the same factor,
and training that lowers the loss."""
import numpy as np
import pytest

import mpnn
import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def problem(nx, ny, nsys, seed0=1000, ell=4.0):
    nxp = 1 << max(nx - 1, 1).bit_length()
    nyp = 1 << max(ny - 1, 1).bit_length()
    vals = []
    for s in range(nsys):
        g = ph.grf2d(nxp, nyp, ell, 2.0, seed0 + s).reshape(nyp, nxp)[:ny, :nx].ravel()
        rp, col, v = ph.diffusion2d(nx, ny, np.exp(g))
        vals.append(v)
    b = np.stack([ph.normal(nx * ny, 77 + s) for s in range(nsys)])
    return rp, col, np.ascontiguousarray(np.stack(vals)), np.ascontiguousarray(b)


def random_params(seed=1):
    rng = np.random.default_rng(seed)
    p = mpnn.init_params(seed, scale=0.2)
    p[mpnn.OFF_WO: mpnn.OFF_WO + mpnn.NO * mpnn.C] = rng.standard_normal(mpnn.NO * mpnn.C) * 0.02
    p[mpnn.OFF_BO:] = rng.standard_normal(mpnn.NO) * 0.1
    for l in range(mpnn.NL):
        p[l * mpnn.LAYER + 2 * mpnn.C * mpnn.C: (l + 1) * mpnn.LAYER] = rng.standard_normal(mpnn.C) * 0.1
    return p.astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("nx,ny", [(64, 24), (100, 16)])
def test_mpnn_forward_matches_oracle(nx, ny):
    print("mpnn forward", nx, ny, flush=True)
    nsys = 2
    rp, col, vals, b = problem(nx, ny, nsys)
    n = nx * ny
    p = random_params()
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        net = ph.Mpnn(p)
        out = net.inspect(B)
        net.factor(B)
        Ls, Lw, Ld = B.factor()
        net.close()
    for s in range(nsys):
        sl = slice(s * n, (s + 1) * n)
        ah, sd = mpnn.scaled_couplings(rp, col, vals[s], nx)
        # couplings and sqrt(diag): IEEE division and square root, bit for bit
        assert np.array_equal(out["ahat"][sl], ah) and np.array_equal(out["sd"][sl], sd)
        fw = mpnn.forward(p, ah, nx, bf=True)
        # six BF16-rounded layers: FP32 accumulation order flips a few roundings
        h_gpu = orc.bf16_to_f32(out["hlast"][sl].ravel()).reshape(n, 32)
        assert rel(h_gpu, fw["hs"][-1]) < 1e-2
        assert rel(out["o"][sl, :5], fw["o"][:, :5]) < 1e-2
        Lref = mpnn.factor_of_A(fw["Lh"], sd)
        for k, L in enumerate((Ls, Lw, Ld)):
            assert rel(orc.bf16_to_f32(L[sl]), orc.bf16_to_f32(Lref[k])) < 2e-2, k


def test_mpnn_factor_pcg_bit_exact():
    print("mpnn factor pcg", flush=True)
    nx, ny, nsys = 128, 32, 2
    rp, col, vals, b = problem(nx, ny, nsys, seed0=2000)
    n = nx * ny
    p = random_params(3)
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        ph.Mpnn(p).factor(B)
        Lb = B.factor()
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 3000, 1e-10)
        B.set_inner(1e-4, 5000)
        x32 = np.zeros_like(b)
        h32, n32, i32 = B.solve(x32, b, 30, 1e-10, precision=32)
    for s in range(nsys):
        sl = slice(s * n, (s + 1) * n)
        F64, F32 = orc.derive(Lb[0][sl], Lb[1][sl], Lb[2][sl], nx)
        M = orc.FactorPrecon(F64, F32, nx)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=3000, rtol=1e-10)
        assert iret[s] == 0 and niter[s] == no and np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo)
        xo, ho, no, io, _ = orc.pcg_mixed_factor(rp, col, vals[s], b[s], M, itrmax=30, rtol=1e-10,
                                                 inner_itrmax=5000, inner_rtol=1e-4)
        assert i32[s] == 0 and n32[s] == no and np.array_equal(h32[s, :no], ho) and np.array_equal(x32[s], xo)


def oracle_loss_grad(p, rp, col, vals, x, b, nx):
    S = vals.shape[0]
    loss, g = 0.0, np.zeros(mpnn.NPARAM)
    for s in range(S):
        ah, sd = mpnn.scaled_couplings(rp, col, vals[s], nx)
        sd64 = sd.astype(np.float64)
        xh = (sd64 * x[s]).astype(np.float32).astype(np.float64)
        bh = (b[s] / sd64).astype(np.float32).astype(np.float64)
        fw = mpnn.forward(p, ah, nx, bf=True)
        l, G = mpnn.loss_and_grad_factor(fw["Lh"], xh, bh, nx, weight=1.0 / S)
        loss += l
        g += mpnn.backward(p, ah, nx, fw, G, bf=True)
    return loss, g


def blocks():
    out = []
    for l in range(mpnn.NL):
        out.append(("W%d" % l, slice(l * mpnn.LAYER, l * mpnn.LAYER + 2 * mpnn.C * mpnn.C)))
        out.append(("b%d" % l, slice(l * mpnn.LAYER + 2 * mpnn.C * mpnn.C, (l + 1) * mpnn.LAYER)))
    out.append(("Wo", slice(mpnn.OFF_WO, mpnn.OFF_WO + 5 * mpnn.C)))
    out.append(("bo", slice(mpnn.OFF_BO, mpnn.OFF_BO + 5)))
    return out


@pytest.mark.parametrize("nx,ny", [(64, 16), (96, 12)])
def test_mpnn_gradients_match_oracle(nx, ny):
    print("mpnn gradients", nx, ny, flush=True)
    nsys = 2
    rp, col, vals, b = problem(nx, ny, nsys, seed0=3000)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(b.shape)
    p = random_params(4)
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        T = ph.MpnnTrainer(nx * ny, nx, nsys, p)
        T.add(B, x, b)
        loss, g = T.grad([0, 1])
        T.close()
    lo, go = oracle_loss_grad(p.astype(np.float64), rp, col, vals, x, b, nx)
    assert abs(loss - lo) <= 1e-3 * lo
    # BF16 dz operands and FP32 sums: each block within 3% of the oracle's
    for name, sl in blocks():
        assert rel(g[sl], go[sl]) < 3e-2, name
    assert np.all(g[mpnn.OFF_WO + 5 * mpnn.C: mpnn.OFF_BO] == 0) and np.all(g[mpnn.OFF_BO + 5:] == 0)


def test_mpnn_training_lowers_loss():
    print("mpnn training", flush=True)
    nx, ny, nsys = 64, 32, 8
    rp, col, vals, b = problem(nx, ny, nsys, seed0=4000)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as Bs:
        x = np.zeros_like(b)
        _, _, iret = Bs.solve(x, b, 2000, 1e-8)
        assert np.all(iret == 0)
    p = mpnn.init_params(0)
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        T = ph.MpnnTrainer(nx * ny, nx, nsys, p)
        T.add(B, x, b)
        l0, _ = T.grad(np.arange(nsys))
        rng = np.random.default_rng(0)
        for it in range(60):
            T.step(rng.choice(nsys, 2, replace=False), 2e-3)
        l1, _ = T.grad(np.arange(nsys))
        p1 = T.params()
        T.close()
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.7 * l0
    # the trained factor through the solver: fewer iterations than untrained
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        its = []
        for q in (p, p1):
            ph.Mpnn(q).factor(B)
            xx = np.zeros_like(b)
            _, niter, iret = B.solve(xx, b, 3000, 1e-10)
            assert np.all(iret == 0)
            its.append(int(niter.sum()))
    print("iterations untrained", its[0], "trained", its[1], flush=True)
    assert its[1] < its[0]
