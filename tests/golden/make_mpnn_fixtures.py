"""Writes tests/golden/mpnn_16x8.npz: forward, loss and gradient of the
learned-factor network on a 16 x 8 grid for fixed parameters, computed by an
independent dense restatement of the network in PyTorch (CPU) with autograd.

The restatement works on matrices, not on the stencil loops of the oracle:
  A_hat = D^-1/2 A D^-1/2 with its diagonal removed (dense n x n),
  m = A_hat h (one matrix product per layer),
  Lh = diag(Ld) + the w and s couplings placed as dense entries,
  loss = ||Lh Lh^T xh - bh||^2 / ||bh||^2,
so a wrong neighbour, a transposed operand or a dropped term in the oracle's
stencil form shows up as a mismatch.  Written once here (torch is present in
the build container); the CPU tests compare the oracle with the file.

    python tests/golden/make_mpnn_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "parcel-legacy_amd")]

OUT = os.path.join(HERE, "mpnn_16x8.npz")
NX, NY = 16, 8
C, NL, NO = 32, 6, 8
LAYER = C * 2 * C + C
NPARAM = NL * LAYER + NO * C + NO


def problem():
    import parcel_hip as ph
    k = np.exp(ph.grf2d(16, 8, 3.0, 1.5, 11))
    rp, col, val = ph.diffusion2d(NX, NY, k)
    n = NX * NY
    A = np.zeros((n, n))
    for i in range(n):
        for e in range(rp[i] - 1, rp[i + 1] - 1):
            A[i, col[e] - 1] = val[e]
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    b = A @ x + 0.05 * rng.standard_normal(n)  # not an exact solution: the loss is not zero at Lh = chol
    return rp, col, val, A, x, b


def params():
    rng = np.random.default_rng(21)
    p = np.zeros(NPARAM)
    for l in range(NL):
        p[l * LAYER: l * LAYER + 2 * C * C] = rng.standard_normal(2 * C * C) * 0.2
        p[l * LAYER + 2 * C * C: (l + 1) * LAYER] = rng.standard_normal(C) * 0.1
    p[NL * LAYER: NL * LAYER + NO * C] = rng.standard_normal(NO * C) * 0.05
    p[NL * LAYER + NO * C:] = rng.standard_normal(NO) * 0.1
    return p.astype(np.float32).astype(np.float64)  # parameters are FP32 in the product


def torch_forward(torch, A, p, x, b, bf):
    """Dense restatement; returns (Lh dense, loss, intermediates)."""
    n = A.shape[0]
    At = torch.tensor(A, dtype=torch.float64)
    d = torch.diagonal(At)
    sq = torch.sqrt(d)
    Ahat = At / sq[:, None] / sq[None, :]
    Ahat = Ahat - torch.diag(torch.diagonal(Ahat))
    # couplings are handed over in FP32
    Ahat = Ahat.to(torch.float32).to(torch.float64)
    sd = sq.to(torch.float32).to(torch.float64)

    def q(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64) if bf else t

    idx = torch.arange(n)
    col = idx % NX
    west = torch.where(col > 0, idx - 1, idx)
    south = torch.where(idx >= NX, idx - NX, idx)
    east = torch.where(col < NX - 1, idx + 1, idx)
    north = torch.where(idx + NX < n, idx + NX, idx)
    a_s = torch.where(idx >= NX, Ahat[idx, south], torch.zeros(n, dtype=torch.float64))
    a_w = torch.where(col > 0, Ahat[idx, west], torch.zeros(n, dtype=torch.float64))
    a_e = torch.where(col < NX - 1, Ahat[idx, east], torch.zeros(n, dtype=torch.float64))
    a_n = torch.where(idx + NX < n, Ahat[idx, north], torch.zeros(n, dtype=torch.float64))
    # the row sum feature is formed in FP32 in the product
    f32 = lambda t: t.to(torch.float32)
    rho = (((1.0 + f32(a_s)) + f32(a_w)) + f32(a_e)) + f32(a_n)
    h = torch.zeros((n, C), dtype=torch.float64)
    h[:, 0], h[:, 1], h[:, 2], h[:, 3] = a_s, a_w, a_e, a_n
    h[:, 4] = rho.to(torch.float64)
    h[:, 5] = 1.0
    h = q(h)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    hs = [h]
    for l in range(NL):
        W = P[l * LAYER: l * LAYER + 2 * C * C].reshape(C, 2 * C)
        bl = P[l * LAYER + 2 * C * C: (l + 1) * LAYER]
        m = q(Ahat @ h)
        z = torch.cat([h, m], dim=1) @ q(W).T + bl
        h = q(h + torch.relu(z))
        hs.append(h)
    Wo = P[NL * LAYER: NL * LAYER + NO * C].reshape(NO, C)
    bo = P[NL * LAYER + NO * C:]
    o = h @ Wo.T + bo
    clamp = lambda t, c: torch.clamp(t, -c, c)
    Ld = torch.sqrt(torch.exp(clamp(2.0 * o[:, 0], 20.0)) + 1e-6)
    zero = torch.zeros(n, dtype=torch.float64)
    tw = o[:, 1] + torch.where(col > 0, o[west, 3], zero)
    ts = o[:, 2] + torch.where(idx >= NX, o[south, 4], zero)
    Lw = a_w * torch.exp(clamp(tw, 10.0))
    Ls = a_s * torch.exp(clamp(ts, 10.0))
    L = torch.diag(Ld)
    L = L + torch.zeros((n, n), dtype=torch.float64).index_put((idx[col > 0], west[col > 0]), Lw[col > 0])
    L = L + torch.zeros((n, n), dtype=torch.float64).index_put((idx[idx >= NX], south[idx >= NX]), Ls[idx >= NX])
    xh = sd * torch.tensor(x, dtype=torch.float64)
    bh = torch.tensor(b, dtype=torch.float64) / sd
    r = L @ (L.T @ xh) - bh
    loss = torch.dot(r, r) / torch.dot(bh, bh)
    return P, L, Ls, Lw, Ld, o, hs, loss, xh, bh


def main():
    import torch
    torch.set_num_threads(1)
    rp, col, val, A, x, b = problem()
    p = params()
    out = dict(rowptr=rp, col=col, val=val, x=x, b=b, params=p, nx=np.int64(NX))
    for bf in (False, True):
        P, L, Ls, Lw, Ld, o, hs, loss, xh, bh = torch_forward(torch, A, p, x, b, bf)
        loss.backward()
        sfx = "_bf" if bf else ""
        out["Ls" + sfx] = Ls.detach().numpy()
        out["Lw" + sfx] = Lw.detach().numpy()
        out["Ld" + sfx] = Ld.detach().numpy()
        out["o" + sfx] = o.detach().numpy()
        out["hlast" + sfx] = hs[-1].detach().numpy()
        out["loss" + sfx] = np.float64(loss.item())
        if not bf:
            out["grad"] = P.grad.numpy().copy()
            out["xh"] = xh.numpy()
            out["bh"] = bh.numpy()
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, "loss", out["loss"], "loss_bf", out["loss_bf"])


if __name__ == "__main__":
    main()
