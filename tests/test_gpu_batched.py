"""GPU parity of the batched solver: every system of a batch against the
oracle's single-system solve on the same inputs, bit for bit (histories,
iteration counts, solutions), for point Jacobi, IC(0) and a BF16 factor, in
FP64 and in the mixed FP64/FP32 refinement."""
import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def batch_problem(nx, ny, nsys, seed0=1000, ell=4.0):
    nxp = 1 << max(nx - 1, 1).bit_length()
    nyp = 1 << max(ny - 1, 1).bit_length()
    vals = []
    for s in range(nsys):
        g = ph.grf2d(nxp, nyp, ell, 2.0, seed0 + s).reshape(nyp, nxp)[:ny, :nx].ravel()
        rp, col, v = ph.diffusion2d(nx, ny, np.exp(g))
        vals.append(v)
    vals = np.ascontiguousarray(np.stack(vals))
    b = np.stack([ph.normal(nx * ny, 77 + s) for s in range(nsys)])
    return rp, col, vals, np.ascontiguousarray(b)


def ic0_factor(rp, val, nx):
    return orc.ic0_5pt(rp, val, nx)


@pytest.mark.parametrize("nx,ny,nsys", [(100, 37, 3), (64, 20, 2)])
def test_batch_jacobi_fp64_bit_exact(nx, ny, nsys):
    print("batch jacobi", nx, ny, nsys, flush=True)
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_JACOBI) as B:
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 3000, 1e-10)
    for s in range(nsys):
        xo, ho, no, io = orc.pcg(rp, col, vals[s], b[s], ipreflag=1, itrmax=3000, rtol=1e-10)
        assert niter[s] == no and iret[s] == io, s
        assert np.array_equal(hist[s, :no], ho), s
        assert np.array_equal(x[s], xo), s


@pytest.mark.parametrize("nx,ny,nsys", [(100, 37, 3), (512, 6, 2), (64, 33, 2), (1000, 3, 1), (1, 40, 1)])
def test_batch_ic0_fp64_bit_exact(nx, ny, nsys):
    print("batch ic0", nx, ny, nsys, flush=True)
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
        st = B.stats()
        # a single column (nx = 1) is a tridiagonal: one grid line of ny cells
        gx = nx if nx > 1 else ny
        assert (st["nx"], st["ny"]) == (gx, nx * ny // gx)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 2000, 1e-10)
    for s in range(nsys):
        F64, F32 = orc.derive(*ic0_factor(rp, vals[s], gx), gx)
        M = orc.FactorPrecon(F64, F32, gx)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=2000, rtol=1e-10)
        assert niter[s] == no and iret[s] == io, s
        assert np.array_equal(hist[s, :no], ho), s
        assert np.array_equal(x[s], xo), s


# line layouts: one wavefront (64, 100), two (200 padded, 256), four (300 with
# a whole wavefront of padding, 512, 1000); strips shorter and longer than the
# 8-line pipeline step
@pytest.mark.parametrize("nx,ny", [(512, 5), (100, 9), (64, 4), (1000, 2), (200, 7), (256, 19), (300, 11),
                                   (512, 37)])
def test_batch_precon_apply_bit_exact(nx, ny):
    print("batch apply", nx, ny, flush=True)
    nsys = 2
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
        z = np.zeros_like(b)
        B.precon(b, z)
        b32 = b.astype(np.float32)
        z32 = np.zeros_like(b32)
        B.precon(b32, z32)
    for s in range(nsys):
        F64, F32 = orc.derive(*ic0_factor(rp, vals[s], nx), nx)
        assert np.array_equal(z[s], orc.lscan_apply(F64, b[s], nx)), s
        assert np.array_equal(z32[s], orc.lscan_apply(F32, b32[s], nx)), s


def test_batch_mixed_ic0_bit_exact():
    print("batch mixed ic0", flush=True)
    nx, ny, nsys = 128, 40, 3
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
        B.set_inner(1e-4, 5000)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 30, 1e-10, precision=32)
        st = B.stats()
    tot = 0
    for s in range(nsys):
        F64, F32 = orc.derive(*ic0_factor(rp, vals[s], nx), nx)
        M = orc.FactorPrecon(F64, F32, nx)
        xo, ho, no, io, ti = orc.pcg_mixed_factor(rp, col, vals[s], b[s], M, itrmax=30, rtol=1e-10,
                                                   inner_itrmax=5000, inner_rtol=1e-4)
        tot += ti
        assert niter[s] == no and iret[s] == io == 0, s
        assert np.array_equal(hist[s, :no], ho), s
        assert np.array_equal(x[s], xo), s
    assert st["inner_total"] == tot


def test_batch_mixed_jacobi_bit_exact():
    print("batch mixed jacobi", flush=True)
    nx, ny, nsys = 64, 30, 2
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_JACOBI) as B:
        B.set_inner(1e-4, 5000)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 30, 1e-10, precision=32)
    for s in range(nsys):
        xo, ho, no, io, ti = orc.pcg_mixed(rp, col, vals[s], b[s], ipreflag=1, itrmax=30, rtol=1e-10,
                                           inner_itrmax=5000, inner_rtol=1e-4)
        assert niter[s] == no and iret[s] == io, s
        assert np.array_equal(hist[s, :no], ho), s
        assert np.array_equal(x[s], xo), s


def test_batch_bf16_factor_bit_exact():
    print("batch bf16 factor", flush=True)
    nx, ny, nsys = 128, 24, 2
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    Lb = [[], [], []]
    for s in range(nsys):
        for k, a in enumerate(ic0_factor(rp, vals[s], nx)):
            Lb[k].append(orc.f32_to_bf16(a.astype(np.float32)))
    Lb = [np.ascontiguousarray(np.concatenate(v)) for v in Lb]
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        B.set_factor(*Lb)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 2000, 1e-10)
        x32 = np.zeros_like(b)
        B.set_inner(1e-4, 5000)
        h32, n32, i32 = B.solve(x32, b, 30, 1e-10, precision=32)
    n = nx * ny
    for s in range(nsys):
        F64, F32 = orc.derive(Lb[0][s * n:(s + 1) * n], Lb[1][s * n:(s + 1) * n], Lb[2][s * n:(s + 1) * n], nx)
        M = orc.FactorPrecon(F64, F32, nx)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=2000, rtol=1e-10)
        assert niter[s] == no and np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo), s
        xo, ho, no, io, _ = orc.pcg_mixed_factor(rp, col, vals[s], b[s], M, itrmax=30, rtol=1e-10,
                                                 inner_itrmax=5000, inner_rtol=1e-4)
        assert n32[s] == no and np.array_equal(h32[s, :no], ho) and np.array_equal(x32[s], xo), s


def test_batch_device_resident_and_spmv():
    print("batch device resident", flush=True)
    nx, ny, nsys = 64, 16, 4
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    vd = ph.DeviceArray(vals)
    with ph.Batch(rp, col, vd, ph.BATCH_JACOBI, nsys=nsys) as B:
        xd = ph.DeviceArray(np.zeros_like(b))
        bd = ph.DeviceArray(b)
        hist, niter, iret = B.solve(xd, bd, 3000, 1e-10)
        x = xd.numpy()
        y = np.zeros_like(b)
        B.spmv(x, y)
    for s in range(nsys):
        assert np.array_equal(y[s], orc.spmv(rp, col, vals[s], x[s])), s
        assert np.linalg.norm(b[s] - y[s]) / np.linalg.norm(b[s]) < 1e-9


def test_batch_rejects_non_five_point_for_factor():
    rp, col, val = ph.poisson3d_spd(4, 4, 4)
    with pytest.raises(ph.ParcelError) as e:
        ph.Batch(rp, col, val[None, :].copy(), ph.BATCH_IC0)
    assert e.value.code == ph.ERR_UNSUPPORTED
    # Jacobi works on any pattern
    with ph.Batch(rp, col, val[None, :].copy(), ph.BATCH_JACOBI) as B:
        b = np.ones((1, 64))
        x = np.zeros_like(b)
        _, niter, iret = B.solve(x, b, 500, 1e-10)
        assert iret[0] == 0


# the headline path (FP64 PCG, factor applied in FP32, z kept in FP32) on
# full four-wavefront lines and on padded two- and four-wavefront lines
@pytest.mark.parametrize("kind,nx,ny", [("ic0", 512, 12), ("bf16", 512, 12), ("ic0", 300, 10), ("bf16", 200, 9)])
def test_batch_fp32_apply_in_fp64_pcg_bit_exact(kind, nx, ny):
    print("batch fp32 apply", kind, nx, ny, flush=True)
    nsys = 2
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    n = nx * ny
    if kind == "ic0":
        B = ph.Batch(rp, col, vals, ph.BATCH_IC0)
        facs = [orc.derive(*ic0_factor(rp, vals[s], nx), nx) for s in range(nsys)]
    else:
        Lb = [[], [], []]
        for s in range(nsys):
            for k, a in enumerate(ic0_factor(rp, vals[s], nx)):
                Lb[k].append(orc.f32_to_bf16(a.astype(np.float32)))
        Lb = [np.ascontiguousarray(np.concatenate(v)) for v in Lb]
        B = ph.Batch(rp, col, vals, ph.BATCH_FACTOR)
        B.set_factor(*Lb)
        facs = [orc.derive(Lb[0][s * n:(s + 1) * n], Lb[1][s * n:(s + 1) * n], Lb[2][s * n:(s + 1) * n], nx)
                for s in range(nsys)]
    with B:
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 3000, 1e-10, precision=ph.BATCH_APPLY32)
    for s in range(nsys):
        M = orc.FactorPrecon(facs[s][0], facs[s][1], nx, apply32=True)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=3000, rtol=1e-10)
        assert iret[s] == 0 and niter[s] == no, s
        assert np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo), s


@pytest.mark.parametrize("nstrip,precision", [(4, 64), (8, 48), (3, 32)])
def test_batch_strips_bit_exact(nstrip, precision):
    print("batch strips", nstrip, precision, flush=True)
    nx, ny, nsys = 512, 16, 2
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
        B.set_strips(nstrip)
        B.set_inner(1e-4, 5000)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 30 if precision == 32 else 3000, 1e-10, precision=precision)
    for s in range(nsys):
        F64, F32 = orc.derive(*ic0_factor(rp, vals[s], nx), nx)
        M = orc.FactorPrecon(F64, F32, nx, apply32=precision == 48, nstrip=nstrip)
        if precision == 32:
            xo, ho, no, io, _ = orc.pcg_mixed_factor(rp, col, vals[s], b[s], M, itrmax=30, rtol=1e-10,
                                                     inner_itrmax=5000, inner_rtol=1e-4)
        else:
            xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=3000, rtol=1e-10)
        assert iret[s] == 0 and niter[s] == no, s
        assert np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo), s


@pytest.mark.parametrize("nx,ny,nsys,sym", [(128, 128, 3, True), (256, 96, 2, True), (128, 128, 2, False),
                                            (100, 37, 2, True)])
def test_batch_stencil_spmv_bit_exact(nx, ny, nsys, sym):
    # exactly symmetric five-point systems take the stencil SpMV (three
    # coefficient arrays, several tile owners per workgroup when the tile
    # count allows); a nonsymmetric one keeps the SELL product; both must
    # give the oracle's histories bit for bit
    print("batch stencil", nx, ny, nsys, sym, flush=True)
    rp, col, vals, b = batch_problem(nx, ny, nsys)
    if not sym:
        rng = np.random.default_rng(5)
        vals = vals * (1.0 + 0.01 * (rng.random(vals.shape) - 0.5) * (col - 1 != np.repeat(
            np.arange(nx * ny), np.diff(rp)))[None, :])
        vals = np.ascontiguousarray(vals)
    with ph.Batch(rp, col, vals, ph.BATCH_IC0 if sym else ph.BATCH_JACOBI) as B:
        assert B.stats()["stencil5"] == (1 if sym else 0)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 4000, 1e-10)
    for s in range(nsys):
        if sym:
            F64, F32 = orc.derive(*ic0_factor(rp, vals[s], nx), nx)
            M = orc.FactorPrecon(F64, F32, nx)
            xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=4000, rtol=1e-10)
        else:
            xo, ho, no, io = orc.pcg(rp, col, vals[s], b[s], ipreflag=1, itrmax=4000, rtol=1e-10)
        assert niter[s] == no and iret[s] == io, s
        assert np.array_equal(hist[s, :no], ho), s
        assert np.array_equal(x[s], xo), s



def _fallback_problem():
    nx, ny, nsys = 64, 32, 3
    vals = []
    for s in range(nsys):
        rp, col, v = ph.diffusion2d(nx, ny, np.exp(ph.grf2d(64, 32, 8.0, 2.0, 900 + s)))
        vals.append(v)
    b = np.stack([ph.normal(nx * ny, 950 + s) for s in range(nsys)])
    return nx, rp, col, np.ascontiguousarray(np.stack(vals)), np.ascontiguousarray(b)


def test_learned_fallback_bit_exact():
    # systems 0 and 2 get a useless factor (M = I) and a budget of 150
    # iterations, so they end with iret = 1 and fall back to FP64 PCG +
    # block-Jacobi ILU(0); system 1 gets its IC(0) factor and converges
    print("learned fallback", flush=True)
    nx, rp, col, vals, b = _fallback_problem()
    nsys, n = b.shape
    one = orc.f32_to_bf16(np.ones(n, np.float32))
    zero = np.zeros(n, np.uint16)
    fac = [(zero, zero, one)] * nsys
    ic = [orc.f32_to_bf16(a.astype(np.float32)) for a in orc.ic0_5pt(rp, vals[1], nx)]
    fac[1] = tuple(ic)
    Ls, Lw, Ld = (np.ascontiguousarray(np.concatenate([f[k] for f in fac])) for k in range(3))
    itrmax = 150
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        B.set_factor(Ls, Lw, Ld)
        B.set_fallback(ph.PRECON_BJ, ph.ILU0, 4, 5000)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, itrmax, 1e-10, precision=ph.BATCH_APPLY32)
        fb = B.fell_back()
        assert B.stats()["fallback_systems"] == 2
        # the same batch with the fallback off reports the failures as they are
        B.set_fallback(itrmax=0)
        x2 = np.zeros_like(b)
        _, n2, i2 = B.solve(x2, b, itrmax, 1e-10, precision=ph.BATCH_APPLY32)
        assert not B.fell_back().any() and list(i2) == [1, 0, 1] and list(n2[[0, 2]]) == [itrmax, itrmax]
    assert list(fb) == [True, False, True]
    for s in (0, 2):
        xo, ho, no, io = orc.pcg(rp, col, vals[s], b[s], ipreflag=2, ilu_method=0, nblocks=4, itrmax=5000,
                                 rtol=1e-10)
        assert (niter[s], iret[s]) == (no, io) and io == 0, s
        k = min(no, itrmax)
        assert np.array_equal(hist[s, :k], ho[:k]) and np.all(hist[s, k:] == 0)
        assert np.array_equal(x[s], xo), s
    F64, F32 = orc.derive(*fac[1], nx)
    M = orc.FactorPrecon(F64, F32, nx, apply32=True)
    xo, ho, no, io = orc.pcg_factor(rp, col, vals[1], b[1], M, itrmax=itrmax, rtol=1e-10)
    assert (niter[1], iret[1]) == (no, io) and io == 0
    assert np.array_equal(hist[1, :no], ho) and np.array_equal(x[1], xo)
    assert np.array_equal(x2[1], xo)


@pytest.mark.parametrize("nb", [1, 4, 8])
def test_block_ic0_batch_bit_exact(nb):
    # block-Jacobi IC(0): the batched form of parcel_dcg's block-Jacobi
    # ILU(0) for these symmetric systems (exact arithmetic); every system bit
    # for bit against the oracle PCG with the same strip factor, and its
    # iteration count next to parcel_dcg's block-Jacobi ILU(0) on the same blocks
    print("block ic0 batch", nb, flush=True)
    rp, col, vals, b = batch_problem(96, 32, 3, seed0=1200)
    nx = 96
    with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
        B.set_ic0_blocks(nb)
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 3000, 1e-10)
    for s in range(3):
        F64, F32 = orc.derive(*orc.ic0_5pt(rp, vals[s], nx, nblocks=nb), nx)
        M = orc.FactorPrecon(F64, F32, nx, nstrip=nb)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=3000, rtol=1e-10)
        assert (niter[s], iret[s]) == (no, io) and io == 0, s
        assert np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo), s
        _, hb, nbj, _ = orc.pcg(rp, col, vals[s], b[s], ipreflag=2, ilu_method=0, nblocks=nb, itrmax=3000,
                                rtol=1e-10)
        assert abs(int(niter[s]) - nbj) <= 2, (s, niter[s], nbj)
