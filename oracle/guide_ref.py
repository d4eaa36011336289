"""CPU FP64 restatement of the learned proposal's network, mixture density
and training gradient (test infrastructure: used only by tests/ to check
the GPU forward and backward kernels).

The network is defined by the library (include/snoggpu.h, "learned
directional proposal"): 6 inputs, 78-feature sin/cos encoding padded to
80, three 64-unit ReLU layers, 40 outputs = 8 lobes x (mean direction
3, log concentration 1, logit 1).  With ``emulate=True`` the bf16
roundings of the GPU path are applied at the same points (weights,
encoded features, hidden activations, output-side gradients); products
are then summed in FP64 instead of the GPU's FP32.  The gradient is
pinned against central finite differences of the FP64 loss.
"""
import math

import numpy as np

IN, FEAT, H, OUT, OUTP, LOBES, NFREQ = 6, 80, 64, 40, 48, 8, 6
KLO, KHI = -3.0, 6.0
SHAPES = [("W1", (H, FEAT)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, H)), ("b3", (H,)), ("W4", (OUTP, H)), ("b4", (OUTP,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    out, o = {}, 0
    for k, s in SHAPES:
        n = int(np.prod(s))
        out[k] = np.asarray(w[o:o + n], np.float64).reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([d[k].ravel() for k, _ in SHAPES])


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(in6, emulate=False):
    """(n, 80) encoding: raw inputs, then sin / cos of 2^k pi u_i."""
    u = np.asarray(in6, np.float64).reshape(-1, IN)
    n = len(u)
    f = np.zeros((n, FEAT))
    f[:, :IN] = u
    for i in range(IN):
        for k in range(NFREQ):
            if emulate:  # the GPU forms the argument in float32
                x = (np.float32(u[:, i].astype(np.float32)) * np.float32(1 << k)
                     * np.float32(math.pi)).astype(np.float64)
            else:
                x = u[:, i] * (1 << k) * math.pi
            c = IN + (i * NFREQ + k) * 2
            f[:, c] = np.sin(x)
            f[:, c + 1] = np.cos(x)
    return bf16(f) if emulate else f


def forward(w, in6, emulate=False):
    """Raw outputs (n, 40) and the activations (for backward)."""
    P = unpack(w)
    if emulate:
        for k in ("W1", "W2", "W3", "W4"):
            P[k] = bf16(P[k])
        for k in ("b1", "b2", "b3", "b4"):
            P[k] = np.asarray(P[k], np.float32).astype(np.float64)
    a0 = features(in6, emulate)
    acts = [a0]
    a = a0
    for l in (1, 2, 3):
        z = a @ P[f"W{l}"].T + P[f"b{l}"]
        a = np.maximum(z, 0.0)
        if emulate:
            a = bf16(a)
        acts.append(a)
    out = a @ P["W4"].T + P["b4"]
    return out[:, :OUT], acts, P


def lobes(raw):
    """(n, 8, 5): unit mean, concentration, softmax weight."""
    r = np.asarray(raw, np.float64).reshape(-1, LOBES, 5)
    u = r[..., :3]
    nn = np.linalg.norm(u, axis=-1, keepdims=True)
    mu = np.where(nn > 1e-12, u / np.where(nn > 1e-12, nn, 1), np.array([0.0, 0.0, 1.0]))
    kap = np.exp(np.clip(r[..., 3], KLO, KHI))
    lg = r[..., 4]
    e = np.exp(lg - lg.max(-1, keepdims=True))
    wt = e / e.sum(-1, keepdims=True)
    return np.concatenate([mu, kap[..., None], wt[..., None]], -1)


def vmf_pdf(kappa, c):
    """kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (c - 1)}; 1 / (4 pi) as kappa -> 0."""
    kappa = np.asarray(kappa, np.float64)
    c = np.asarray(c, np.float64)
    small = kappa < 1e-4
    ks = np.where(small, 1.0, kappa)
    p = ks / (2 * math.pi * (-np.expm1(-2 * ks))) * np.exp(ks * (c - 1))
    return np.where(small, 1 / (4 * math.pi), p)


def mix_pdf(raw, wi):
    L = lobes(raw)
    wi = np.asarray(wi, np.float64).reshape(-1, 1, 3)
    c = (L[..., :3] * wi).sum(-1)
    return (L[..., 4] * vmf_pdf(L[..., 3], c)).sum(-1)


def vmf_sample(mu, kappa, u1, u2):
    """Inversion sampling of one lobe (the same formula as the device)."""
    mu = np.asarray(mu, np.float64)
    w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    w = np.clip(w, -1, 1)
    sr = np.sqrt(np.maximum(0, 1 - w * w))
    ph = 2 * math.pi * u2
    lx, ly = sr * np.cos(ph), sr * np.sin(ph)
    nx, ny, nz = mu
    sg = 1.0 if nz >= 0 else -1.0
    a = -1 / (sg + nz)
    b = nx * ny * a
    t = np.array([1 + sg * nx * nx * a, sg * b, -sg * nx])
    bb = np.array([b, sg + ny * ny * a, -ny])
    return lx[:, None] * t + ly[:, None] * bb + w[:, None] * mu


def loss_sum(w, rec, emulate=False):
    rec = np.asarray(rec, np.float64).reshape(-1, 16)
    raw, _, _ = forward(w, rec[:, :6], emulate)
    q = mix_pdf(raw, rec[:, 6:9])
    wt = np.maximum(rec[:, 9], 0)
    return float((wt * -np.log(np.maximum(q, 1e-300))).sum()), float(wt.sum())


def raw_grad(raw, wi):
    """d(-log q)/d(raw) per row (n, 40), analytically."""
    r = np.asarray(raw, np.float64).reshape(-1, LOBES, 5)
    L = lobes(raw)
    wi = np.asarray(wi, np.float64).reshape(-1, 1, 3)
    c = (L[..., :3] * wi).sum(-1)
    v = vmf_pdf(L[..., 3], c)
    q = np.maximum((L[..., 4] * v).sum(-1, keepdims=True), 1e-300)
    gam = L[..., 4] * v / q
    kap = L[..., 3]
    nn = np.linalg.norm(r[..., :3], axis=-1)
    s = np.where(nn > 1e-12, gam * kap / np.where(nn > 1e-12, nn, 1), 0)
    g = np.zeros_like(r)
    g[..., :3] = -s[..., None] * (wi - c[..., None] * L[..., :3])
    inside = (r[..., 3] > KLO) & (r[..., 3] < KHI)
    e2 = np.exp(-2 * kap)
    dk = gam * kap * (1 / kap - 2 * e2 / (1 - e2) + (c - 1))
    g[..., 3] = np.where(inside, -dk, 0)
    g[..., 4] = -(gam - L[..., 4])
    return g.reshape(-1, OUT)


def gradient(w, rec, emulate=False):
    """(grad of sum(w * -log q) / sum(w) wrt the parameters, loss sum, weight sum)."""
    rec = np.asarray(rec, np.float64).reshape(-1, 16)
    raw, acts, P = forward(w, rec[:, :6], emulate)
    wt = np.maximum(rec[:, 9], 0)
    g4 = np.zeros((len(rec), OUTP))
    g4[:, :OUT] = wt[:, None] * raw_grad(raw, rec[:, 6:9])
    if emulate:
        g4 = bf16(g4)
    grads = {"W4": g4.T @ acts[3], "b4": g4.sum(0)}
    d = g4
    for l in (3, 2, 1):
        Wn = P[f"W{l + 1}"]
        da = d @ Wn
        d = np.where(acts[l] > 0, da, 0.0)
        if emulate:
            d = bf16(d)
        grads[f"W{l}"] = d.T @ acts[l - 1]
        grads[f"b{l}"] = d.sum(0)
    ls, ws = loss_sum(w, rec, emulate)
    g = pack({k: grads[k] for k, _ in SHAPES})
    return g / ws if ws > 0 else g * 0, ls, ws
