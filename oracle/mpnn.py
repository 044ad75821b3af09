"""Oracle of the learned preconditioner's message-passing network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's cpu_baseline
leg), never by the product library.  No reference counterpart: the network,
its factor parameterisation and the training loss are this build's design,
specified here in float64 with the product's BF16 rounding points emulated
(`bf=True`), so the GPU kernels can be compared against it with tolerances
that follow from FP32 accumulation order alone.

Network over the graph of a 2D five-point matrix A (n cells, grid line nx):
  scaled couplings  a_ij = A_ij / sqrt(A_ii A_jj) for the neighbours
                    s (i-nx), w (i-1), e (i+1), n (i+nx); 0 where absent;
  input             h0_i = [a_s, a_w, a_e, a_n, 1 + a_s + a_w + a_e + a_n, 1,
                    0 ... 0]  (32 channels, BF16);
  6 layers          m_i = sum_j a_ij h_j  (order s, w, e, n, FP32, then BF16)
                    z_i = W_l [h_i; m_i] + b_l   (W_l 32 x 64, BF16 operands,
                    FP32 accumulation)
                    h_i <- BF16(h_i + relu(z_i));
  head              o_i = Wo h_i + bo  (5 outputs used, Wo 8 x 32, FP32);
  factor of the Jacobi-scaled matrix (clamp(t) = min(max(t, -10), 10)):
                    Lh_ii = sqrt(exp(2 clamp(o0_i)) + 1e-6)
                    Lh_iw = a_iw exp(clamp(o1_i + o3_{i-1}))
                    Lh_is = a_is exp(clamp(o2_i + o4_{i-nx}))
  and of A          L = diag(sqrt(A_ii)) Lh, stored in BF16; M = L L^T.
Loss of one system with solution x of A x = b (xh = sqrt(A_ii) x,
bh = b / sqrt(A_ii)):  ||Lh Lh^T xh - bh||^2 / ||bh||^2, averaged over the
systems of a minibatch.
"""
import numpy as np

C = 32
NL = 6
NO = 8          # head rows (5 used)
LAYER = C * 2 * C + C
OFF_WO = NL * LAYER
OFF_BO = OFF_WO + NO * C
NPARAM = OFF_BO + NO
CLAMP = 10.0
EPS = 1e-6


def bf16(x):
    """Round to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    W = [p[l * LAYER: l * LAYER + 2 * C * C].reshape(C, 2 * C) for l in range(NL)]
    b = [p[l * LAYER + 2 * C * C: (l + 1) * LAYER] for l in range(NL)]
    Wo = p[OFF_WO: OFF_WO + NO * C].reshape(NO, C)
    bo = p[OFF_BO: OFF_BO + NO]
    return W, b, Wo, bo


def init_params(seed=0, scale=0.05):
    """Small random layers, zero head: the untrained factor is I + strictly
    lower part of the scaled matrix (Lh_ii = sqrt(1 + 1e-6))."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for l in range(NL):
        p[l * LAYER: l * LAYER + 2 * C * C] = rng.standard_normal(2 * C * C) * scale
    return p


def scaled_couplings(rowptr, col, val, nx):
    """(ahat (n,4) float32 in order s, w, e, n; sd = sqrt(A_ii) float32)."""
    n = rowptr.shape[0] - 1
    d = np.empty(n)
    ah = np.zeros((n, 4))
    for i in range(n):
        for e in range(rowptr[i] - 1, rowptr[i + 1] - 1):
            if col[e] - 1 == i:
                d[i] = val[e]
    for i in range(n):
        for e in range(rowptr[i] - 1, rowptr[i + 1] - 1):
            j = col[e] - 1
            if j == i:
                continue
            k = {-nx: 0, -1: 1, 1: 2, nx: 3}[j - i]
            ah[i, k] = val[e] / np.sqrt(d[i] * d[j])
    return ah.astype(np.float32), np.sqrt(d).astype(np.float32)


def neighbours(n, nx):
    i = np.arange(n)
    x = i % nx
    idx = np.stack([np.where(i >= nx, i - nx, i), np.where(x > 0, i - 1, i),
                    np.where(x < nx - 1, i + 1, i), np.where(i + nx < n, i + nx, i)], axis=1)
    return idx


def aggregate(ah, h, nbr, bf=True):
    """m_i = sum_j a_ij h_j in order s, w, e, n (float64; FP32 on the GPU)."""
    m = np.zeros_like(h)
    for k in range(4):
        m = m + ah[:, k:k + 1].astype(np.float64) * h[nbr[:, k]]
    return bf16(m) if bf else m


def input_features(ah, bf=True):
    a = ah.astype(np.float32)
    h = np.zeros((a.shape[0], C))
    h[:, :4] = a
    rho = (((np.float32(1.0) + a[:, 0]) + a[:, 1]) + a[:, 2]) + a[:, 3]
    h[:, 4] = rho
    h[:, 5] = 1.0
    return bf16(h) if bf else h


def forward(params, ah, nx, bf=True):
    """Returns dict with hs (list of 7 activations), ms, zs, o, Lh (3, n)."""
    W, b, Wo, bo = unpack(params)
    n = ah.shape[0]
    nbr = neighbours(n, nx)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    h = input_features(ah, bf)
    hs, ms, zs = [h], [], []
    for l in range(NL):
        m = aggregate(ah, h, nbr, bf)
        z = np.concatenate([h, m], axis=1) @ q(W[l]).T + b[l]
        h = q(h + np.maximum(z, 0.0))
        hs.append(h)
        ms.append(m)
        zs.append(z)
    o = h @ Wo.T + bo
    Lh, aux = factor_from_head(o, ah, nx)
    return dict(hs=hs, ms=ms, zs=zs, o=o, Lh=Lh, aux=aux, nbr=nbr)


def factor_from_head(o, ah, nx):
    n = o.shape[0]
    i = np.arange(n)
    x = i % nx
    ahd = ah.astype(np.float64)
    t0 = np.clip(2.0 * o[:, 0], -2 * CLAMP, 2 * CLAMP)
    e0 = np.exp(t0)
    Ld = np.sqrt(e0 + EPS)
    tw = o[:, 1] + np.where(x > 0, o[np.maximum(i - 1, 0), 3], 0.0)
    ts = o[:, 2] + np.where(i >= nx, o[np.maximum(i - nx, 0), 4], 0.0)
    cw = np.clip(tw, -CLAMP, CLAMP)
    cs = np.clip(ts, -CLAMP, CLAMP)
    Lw = ahd[:, 1] * np.exp(cw)
    Ls = ahd[:, 0] * np.exp(cs)
    aux = dict(e0=e0, t0=t0, tw=tw, ts=ts, x=x)
    return np.stack([Ls, Lw, Ld]), aux


def factor_of_A(Lh, sd):
    """L = diag(sd) Lh in BF16 bits (uint16 Ls, Lw, Ld)."""
    import oracle as orc
    return tuple(orc.f32_to_bf16((Lh[k] * sd.astype(np.float64)).astype(np.float32)) for k in range(3))


def lt_apply(Lh, v, nx):
    """Lh^T v: (Lh^T v)_j = Ld_j v_j + Lw_{j+1} v_{j+1} + Ls_{j+nx} v_{j+nx}."""
    Ls, Lw, Ld = Lh
    n = v.shape[0]
    out = Ld * v
    j = np.arange(n)
    nxt = (j % nx < nx - 1) & (j + 1 < n)
    out[nxt] += Lw[j[nxt] + 1] * v[j[nxt] + 1]
    up = j + nx < n
    out[up] += Ls[j[up] + nx] * v[j[up] + nx]
    return out


def l_apply(Lh, v, nx):
    Ls, Lw, Ld = Lh
    n = v.shape[0]
    out = Ld * v
    i = np.arange(n)
    w = i % nx > 0
    out[w] += Lw[w] * v[i[w] - 1]
    s = i >= nx
    out[s] += Ls[s] * v[i[s] - nx]
    return out


def loss_and_grad_factor(Lh, xh, bh, nx, weight=1.0):
    """loss = weight ||Lh Lh^T xh - bh||^2 / ||bh||^2 and dloss/dLh (3, n)."""
    u = lt_apply(Lh, xh, nx)
    v = l_apply(Lh, u, nx)
    r = v - bh
    bb = float(np.dot(bh, bh))
    loss = weight * float(np.dot(r, r)) / bb
    g = 2.0 * weight * r / bb
    w = lt_apply(Lh, g, nx)
    n = xh.shape[0]
    i = np.arange(n)
    G = np.zeros((3, n))
    G[2] = g * u + xh * w
    wm = i % nx > 0
    G[1, wm] = g[wm] * u[i[wm] - 1] + xh[wm] * w[i[wm] - 1]
    sm = i >= nx
    G[0, sm] = g[sm] * u[i[sm] - nx] + xh[sm] * w[i[sm] - nx]
    return loss, G


def backward(params, ah, nx, fw, G, bf=True):
    """Gradient of sum_ij G_ij Lh_ij with respect to the parameters."""
    W, b, Wo, bo = unpack(params)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    o, Lh, aux = fw["o"], fw["Lh"], fw["aux"]
    n = o.shape[0]
    i = np.arange(n)
    x = aux["x"]
    do = np.zeros((n, NO))
    # Ld = sqrt(exp(t0) + eps), t0 = clamp(2 o0)
    inside0 = np.abs(2.0 * o[:, 0]) < 2 * CLAMP
    do[:, 0] = G[2] * aux["e0"] / Lh[2] * inside0
    gw = G[1] * Lh[1] * (np.abs(aux["tw"]) < CLAMP)
    gs = G[0] * Lh[0] * (np.abs(aux["ts"]) < CLAMP)
    do[:, 1] = gw
    do[:, 2] = gs
    wm = x > 0
    do[i[wm] - 1, 3] += gw[wm]
    sm = i >= nx
    do[i[sm] - nx, 4] += gs[sm]
    grad = np.zeros(NPARAM)
    h = fw["hs"][-1]
    grad[OFF_WO: OFF_WO + NO * C] = (do.T @ h).ravel()
    grad[OFF_BO: OFF_BO + NO] = do.sum(axis=0)
    dh = do @ Wo
    nbr = fw["nbr"]
    for l in reversed(range(NL)):
        hin, m, z = fw["hs"][l], fw["ms"][l], fw["zs"][l]
        dz = dh * (z > 0)
        dzq = q(dz)
        hm = np.concatenate([hin, m], axis=1)
        grad[l * LAYER: l * LAYER + 2 * C * C] = (dzq.T @ hm).ravel()
        grad[l * LAYER + 2 * C * C: (l + 1) * LAYER] = dz.sum(axis=0)
        dhm = dzq @ q(W[l])
        dm = dhm[:, C:]
        # m_i = sum_j a_ij h_j with a symmetric: dh_j += sum_i a_ij dm_i
        agg = np.zeros_like(dm)
        for k in range(4):
            agg = agg + ah[:, k:k + 1].astype(np.float64) * dm[nbr[:, k]]
        dh = dh + dhm[:, :C] + agg
    return grad
