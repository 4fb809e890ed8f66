"""Float64 CPU restatement of the learned directional proposal.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing shipped).  The
reference renderer has no learned guiding, so nothing here restates a
reference file; the definitions follow the renderer's own specification
(colvillea_amd/csrc/clpt_guide.h, guide.hip):

  encoding   9 inputs -> 256 features: group g < 27 holds, for input g // 3
             and frequencies f = 4 (g % 3) + 0..3, the pairs
             sin(2^f pi v), cos(2^f pi v); group 27 the raw inputs 0..7;
             group 28 raw input 8, then 1, then zeros; groups 29..31 zeros.
  network    256 -> 64 -> 64 -> 64 -> 64 -> 64, ReLU between, no biases;
             weights W_l[in][out], outputs 0..39 used.
  mixture    lobe j: logit out[5j], mean out[5j+1..3] (normalised),
             kappa = exp(clamp(out[5j+4], -4, 8)); weights softmax(logits);
             vMF_j(w) = kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}.
  loss       L = -sum_n (w_n / sum w) log p(wi_n).
  Adam       beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected.

Pins (tests/test_guide_oracle.py): central finite differences of the loss
gradient; vMF normalisation by quadrature; the mixture integrates to 1;
chi-square of mixture sampling against the pdf; Adam's first step in closed
form; encoding values at known arguments; bfloat16 rounding of known bit
patterns.
"""
from __future__ import annotations

import numpy as np

IN, HID, NOUT, LOBES = 256, 64, 40, 8
PARAMS = IN * HID + 4 * HID * HID
KMIN, KMAX = -4.0, 8.0


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) via float32, as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def e4m3(x):
    """Round to OCP FP8 e4m3 (nearest, ties to even; saturated at +-448;
    subnormal step 2^-9), as float64."""
    x = np.clip(np.asarray(x, np.float64), -448.0, 448.0)
    a = np.abs(x)
    _, e = np.frexp(a)
    step = np.where(a >= 2.0 ** -6, np.ldexp(1.0, e - 4), 2.0 ** -9)
    return np.sign(x) * (np.round(a / step) * step)


def forward_fp8(Ws, X, scl):
    """FP8 inference as the renderer's ablation computes it: the input of
    layer l enters as e4m3(x * as_l) and the weights as e4m3(W * ws_l) with
    ws_l = 448 / max|W_l| (float32), the accumulator is scaled by scl[2l+1]
    (= 1 / (as_l ws_l)); hidden inputs are relu(acc) * (scl[2l-1] scl[2l])."""
    f32 = np.float32
    scl = np.asarray(scl, np.float32)
    h = e4m3(np.asarray(X, np.float32).astype(np.float64) * float(scl[0]))
    acc = None
    for l in range(5):
        W = np.asarray(Ws[l], np.float32)
        ws = f32(448.0) / f32(np.abs(W).max())
        Wq = e4m3(W.astype(np.float64) * float(ws)).astype(np.float64)
        if l > 0:
            h = e4m3(np.maximum(acc, 0.0) * float(f32(scl[2 * l - 1]) * f32(scl[2 * l])))
        acc = h @ Wq
    return acc * float(scl[9])


def split_params(p):
    """Flat parameter vector -> [W0 (256x64), W1..W4 (64x64)]."""
    p = np.asarray(p, np.float64)
    Ws = [p[:IN * HID].reshape(IN, HID)]
    off = IN * HID
    for _ in range(4):
        Ws.append(p[off:off + HID * HID].reshape(HID, HID))
        off += HID * HID
    return Ws


def join_params(Ws):
    return np.concatenate([w.reshape(-1) for w in Ws])


def encode(q):
    q = np.asarray(q, np.float64)
    N = q.shape[0]
    X = np.zeros((N, IN))
    for g in range(27):
        d, fb = divmod(g, 3)
        for k in range(4):
            a = np.pi * (2.0 ** (4 * fb + k)) * q[:, d]
            X[:, 8 * g + 2 * k] = np.sin(a)
            X[:, 8 * g + 2 * k + 1] = np.cos(a)
    X[:, 216:224] = q[:, :8]
    X[:, 224] = q[:, 8]
    X[:, 225] = 1.0
    return X


def forward(Ws, X, bf=False):
    """Outputs (N x 64) and the per-layer inputs (post-ReLU, as used)."""
    r = bf16 if bf else (lambda a: a)
    A = [r(X)]
    h = A[0]
    for l in range(5):
        z = h @ r(Ws[l])
        if l < 4:
            h = r(np.maximum(z, 0.0))
            A.append(h)
        else:
            return z, A


def mixture(out):
    out = np.asarray(out, np.float64)[:, :NOUT].reshape(-1, LOBES, 5)
    a = out[:, :, 0]
    a = a - a.max(axis=1, keepdims=True)
    w = np.exp(a)
    w /= w.sum(axis=1, keepdims=True)
    raw = out[:, :, 1:4]
    nr = np.linalg.norm(raw, axis=2, keepdims=True)
    mu = np.where(nr ** 2 > 1e-12, raw / np.maximum(nr, 1e-300), np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(out[:, :, 4], KMIN, KMAX))
    norm = w * kappa / (2.0 * np.pi * (-np.expm1(-2.0 * kappa)))
    return w, mu, kappa, norm


def vmf_pdf(mu, kappa, d):
    return kappa / (2.0 * np.pi * (-np.expm1(-2.0 * kappa))) * np.exp(kappa * (np.sum(mu * d, -1) - 1.0))


def mix_pdf(out, d):
    """Mixture density of each row's mixture at direction d (N x 3)."""
    w, mu, kappa, norm = mixture(out)
    cosv = np.einsum("njk,nk->nj", mu, np.asarray(d, np.float64))
    return np.sum(norm * np.exp(kappa * (cosv - 1.0)), axis=1)


def frame(n):
    """Orthonormal t, b around unit n, completed as makeFrame
    (libkernel/base/math.h:33-48) completes it -- c in the plane of n and
    the larger of its x, y axes, t = c x n, b = c -- so that a sample drawn
    here and one drawn by the renderer from the same numbers agree."""
    n = np.asarray(n, np.float64)
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    use_x = np.abs(x) > np.abs(y)
    with np.errstate(divide="ignore", invalid="ignore"):
        ix = 1.0 / np.sqrt(x * x + z * z)
        iy = 1.0 / np.sqrt(y * y + z * z)
        zero = np.zeros_like(x)
        c = np.where(use_x[..., None], np.stack([z * ix, zero, -x * ix], -1), np.stack([zero, z * iy, -y * iy], -1))
    return np.cross(c, n), c


def mix_sample(out, usel, u1, u2):
    """Lobe by usel over the weights, then the Jakob (2012) inversion."""
    w, mu, kappa, _ = mixture(out)
    c = np.cumsum(w, axis=1)
    j = np.minimum((usel[:, None] >= c[:, :-1]).sum(axis=1), LOBES - 1)
    idx = np.arange(len(j))
    m, k = mu[idx, j], kappa[idx, j]
    ct = np.clip(1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * k)) / k, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    t, b = frame(m)
    ph = 2.0 * np.pi * u2
    return st[:, None] * (np.cos(ph)[:, None] * t + np.sin(ph)[:, None] * b) + ct[:, None] * m


def loss_and_grad(params, q, wi, weight, bf=False):
    """L = -sum_n w_n / W log p(wi_n) and dL/dparams (flat), float64.
    bf=True mirrors every bfloat16 rounding point of the GPU (weights,
    features, activations, backward gradients); bf="forward" is the exact
    float64 gradient of the bfloat16 network (rounded forward, unrounded
    backward)."""
    Ws = split_params(params)
    r = bf16 if bf is True else (lambda a: a)
    X = encode(q)
    out, A = forward(Ws, X, bool(bf))
    wn = np.asarray(weight, np.float64) / np.sum(weight)
    wi = np.asarray(wi, np.float64)
    N = out.shape[0]
    o = out[:, :NOUT].reshape(N, LOBES, 5)
    w, mu, kappa, norm = mixture(out)
    cosv = np.einsum("njk,nk->nj", mu, wi)
    v = norm * np.exp(kappa * (cosv - 1.0))
    p = v.sum(axis=1)
    loss = float(-np.sum(wn * np.log(np.maximum(p, 1e-300))))
    rr = v / p[:, None]
    G = np.zeros((N, LOBES, 5))
    G[:, :, 0] = -wn[:, None] * (rr - w)
    raw = o[:, :, 1:4]
    nr = np.linalg.norm(raw, axis=2)
    ok = nr ** 2 > 1e-12
    s1 = np.where(ok, -wn[:, None] * rr * kappa / np.maximum(nr, 1e-300), 0.0)
    G[:, :, 1:4] = s1[:, :, None] * (wi[:, None, :] - cosv[:, :, None] * mu)
    sr = o[:, :, 4]
    inside = (sr > KMIN) & (sr < KMAX)
    e = np.exp(-2.0 * kappa)
    dl = 1.0 / kappa - 2.0 * e / (-np.expm1(-2.0 * kappa)) + (cosv - 1.0)
    G[:, :, 4] = np.where(inside, -wn[:, None] * rr * dl * kappa, 0.0)
    g = np.zeros((N, HID))
    g[:, :NOUT] = G.reshape(N, NOUT)
    g = r(g)
    grads = [None] * 5
    for l in range(4, -1, -1):
        grads[l] = A[l].T @ g
        if l > 0:
            Wl = bf16(Ws[l]) if bf else Ws[l]
            g = r((g @ Wl.T) * (A[l] > 0))
    return loss, join_params(grads)


def adam(params, m, v, t, g, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=1):
    """He-uniform weights from a SplitMix64 stream (the library's
    clhip_guide_reset draws the same stream in the same order)."""
    st = np.uint64(seed)
    out = np.zeros(PARAMS)
    vals = []
    M = (1 << 64) - 1
    s = int(seed)
    n = IN * HID + 4 * HID * HID
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        vals.append((z >> 11) * (1.0 / 9007199254740992.0))
    u = np.array(vals)
    off = 0
    for l in range(5):
        fan = IN if l == 0 else HID
        a = np.sqrt(6.0 / fan) * (0.1 if l == 4 else 1.0)
        blk = (2.0 * u[off:off + fan * HID] - 1.0) * a
        blk = blk.reshape(fan, HID)
        if l == 4:
            blk[:, NOUT:] = 0.0
        out[off:off + fan * HID] = blk.reshape(-1)
        off += fan * HID
    del st
    return out.astype(np.float32).astype(np.float64)


def gather(V, E, maxd):
    """Training samples of recorded paths (the renderer's k_guide_gather_*,
    restated in float32 with its operation order): for each path slot in
    order, vertices from depth maxd up to 1, with S the float32 running sum
    of the contributions E at the deeper vertices, weight
    w = (0.2126 S.r R.r + 0.7152 S.g R.g) + 0.0722 S.b R.b, divided by p,
    where plane 3 of V holds (p, R.r, R.g, R.b); a sample is kept when
    0 < w < 3e38.  V: [4, D, S, 4], E: [D, S, 4].  Returns [n, 13]:
    planes 0-2 of the vertex (query 9, direction 3) and w."""
    f = np.float32
    V = np.asarray(V, f)
    E = np.asarray(E, f)
    S = V.shape[2]
    sr = np.zeros(S, f)
    sg = np.zeros(S, f)
    sb = np.zeros(S, f)
    keys, rows = [], []
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        for j in range(maxd, 0, -1):
            v3 = V[3, j]
            lum = (f(0.2126) * (sr * v3[:, 1]) + f(0.7152) * (sg * v3[:, 2])) + f(0.0722) * (sb * v3[:, 3])
            w = np.where(v3[:, 0] > 0, lum / np.where(v3[:, 0] > 0, v3[:, 0], f(1)), f(0)).astype(f)
            keep = np.nonzero((w > 0) & (w < f(3.0e38)))[0]
            if len(keep):
                q = np.concatenate([V[0, j, keep], V[1, j, keep], V[2, j, keep], w[keep, None]], 1)
                rows.append(q)
                keys.append(np.stack([keep, np.full(len(keep), maxd - j)], 1))
            sr = (sr + E[j, :, 0]).astype(f)
            sg = (sg + E[j, :, 1]).astype(f)
            sb = (sb + E[j, :, 2]).astype(f)
    if not rows:
        return np.zeros((0, 13), f)
    rows = np.concatenate(rows)
    keys = np.concatenate(keys)
    order = np.lexsort((keys[:, 1], keys[:, 0]))
    return rows[order]

