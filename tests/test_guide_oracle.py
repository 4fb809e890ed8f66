"""Pins of the float64 guiding oracle (oracle/guide_ref.py) by mathematics
that does not restate its formulas: finite differences, quadrature,
chi-square, closed forms."""
import numpy as np
import pytest
from scipy import stats

import guide_ref as G


def rand_queries(rng, n):
    q = np.empty((n, 9))
    q[:, :3] = rng.uniform(0, 1, (n, 3))
    for s in (3, 6):
        v = rng.normal(size=(n, 3))
        q[:, s:s + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return q


def rand_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_bf16_rounding_known_patterns():
    one = 1.0
    assert G.bf16(one) == 1.0
    assert G.bf16(1.0 + 2.0 ** -8) == 1.0                     # tie -> even (mantissa 0)
    assert G.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6     # tie -> even (mantissa 2)
    assert G.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert G.bf16(-2.5) == -2.5
    assert G.bf16(3.14159265) == 3.140625


def test_e4m3_rounding_known_values():
    assert G.e4m3(1.0) == 1.0 and G.e4m3(448.0) == 448.0 and G.e4m3(1e6) == 448.0
    assert G.e4m3(1.0625) == 1.0                      # tie (step 1/8) -> even
    assert G.e4m3(1.1875) == 1.25                     # tie -> even (mantissa 2)
    assert G.e4m3(2.0 ** -9) == 2.0 ** -9             # smallest subnormal
    assert G.e4m3(2.0 ** -10) == 0.0                  # tie with 0 -> even (0)
    assert G.e4m3(-3.3) == -3.25
    # brute force against the format: zero, 7 subnormals m 2^-9, normals
    # (1 + m/8) 2^(e-7) for e = 1..15 except e = 15, m = 7 (NaN): 127 values
    vals = [m * 2.0 ** -9 for m in range(8)]
    vals += [(1 + m / 8) * 2.0 ** (e - 7) for e in range(1, 16) for m in range(8) if not (e == 15 and m == 7)]
    vals = np.array(sorted(vals))
    assert len(vals) == 127 and vals[-1] == 448.0
    assert np.all(G.e4m3(vals) == vals)
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 448, 20000), rng.uniform(0, 0.02, 5000), (vals[1:] + vals[:-1]) / 2])
    got = G.e4m3(x)
    d = np.abs(x[:, None] - vals[None, :])
    best = d.min(axis=1)
    # nearest value; at a tie, the one with an even mantissa (even index in the sorted list)
    for xi, gi, bi, row in zip(x, got, best, d):
        cand = np.nonzero(row == bi)[0]
        want = vals[cand[0]] if len(cand) == 1 else vals[[c for c in cand if c % 2 == 0][0]]
        assert gi == want, (xi, gi, want)


def test_encoding_known_values():
    q = np.zeros((1, 9))
    q[0, :] = [0.25, 0.5, 0.0, 1.0, -0.5, 0.125, 0.75, -1.0, 0.3]
    X = G.encode(q)[0]
    # input 0, frequencies 0..3: sin/cos of pi/4, pi/2, pi, 2pi
    assert np.allclose(X[0:8], [np.sqrt(0.5), np.sqrt(0.5), 1, 0, 0, -1, 0, 1], atol=1e-12)
    # input 1 = 0.5, frequency 4 (group 4 = input 1, block 1): 16 * pi/2 = 8 pi
    assert np.allclose(X[8 * 4:8 * 4 + 2], [0, 1], atol=1e-12)
    assert np.allclose(X[216:224], q[0, :8])
    assert X[224] == 0.3 and X[225] == 1.0 and np.all(X[226:] == 0)


@pytest.mark.parametrize("kappa", [np.exp(-4.0), 1.0, 50.0, np.exp(8.0)])
def test_vmf_normalisation_quadrature(kappa):
    # integrate over the polar angle from the mean (substitution c = cos t)
    mu = np.array([[0.0, 0.0, 1.0]])
    # geometric grid of 1 - c resolves the peak at large kappa
    x = np.concatenate([[0.0], np.geomspace(1e-9, 2.0, 200001)])
    c = 1.0 - x
    d = np.stack([np.sqrt(np.maximum(0, 1 - c * c)), np.zeros_like(c), c], 1)
    f = G.vmf_pdf(mu, kappa, d)
    integral = 2 * np.pi * np.trapezoid(f, x)
    assert abs(integral - 1.0) < 1e-6, integral


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(3)
    out = rng.normal(scale=1.5, size=(4, 64))
    out[:, 4::5][:, :8] = rng.uniform(-3, 4, (4, 8))  # log kappa
    # Fibonacci sphere with equal-area cells
    n = 400000
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    ph = np.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    for row in out:
        p = G.mix_pdf(np.repeat(row[None], n, 0), d)
        assert abs(4 * np.pi * p.mean() - 1.0) < 2e-3


def test_mixture_sampling_chi_square():
    rng = np.random.default_rng(5)
    row = rng.normal(scale=1.0, size=64)
    row[4:40:5] = rng.uniform(0.0, 2.5, 8)
    n = 400000
    outs = np.repeat(row[None], n, 0)
    d = G.mix_sample(outs, rng.uniform(size=n), rng.uniform(size=n), rng.uniform(size=n))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-9)
    # bins in (z, phi); expected counts by midpoint-refined quadrature
    nz, nph = 24, 48
    zi = np.minimum(((d[:, 2] + 1) / 2 * nz).astype(int), nz - 1)
    pi_ = np.minimum(((np.arctan2(d[:, 1], d[:, 0]) + np.pi) / (2 * np.pi) * nph).astype(int), nph - 1)
    obs = np.bincount(zi * nph + pi_, minlength=nz * nph)
    sub = 8
    zz = -1 + (np.arange(nz * sub) + 0.5) * (2 / (nz * sub))
    pp = -np.pi + (np.arange(nph * sub) + 0.5) * (2 * np.pi / (nph * sub))
    Z, P = np.meshgrid(zz, pp, indexing="ij")
    R = np.sqrt(1 - Z ** 2)
    dirs = np.stack([R * np.cos(P), R * np.sin(P), Z], -1).reshape(-1, 3)
    pdf = G.mix_pdf(np.repeat(row[None], len(dirs), 0), dirs).reshape(nz * sub, nph * sub)
    cell = (2 / (nz * sub)) * (2 * np.pi / (nph * sub))
    exp = (pdf * cell).reshape(nz, sub, nph, sub).sum(axis=(1, 3)).reshape(-1) * n
    m = exp > 5
    chi2 = np.sum((obs[m] - exp[m]) ** 2 / exp[m])
    dof = m.sum() - 1
    assert stats.chi2.sf(chi2, dof) > 1e-3, (chi2, dof)


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(7)
    params = G.init_params(11)
    n = 12
    q = rand_queries(rng, n)
    wi = rand_dirs(rng, n)
    w = rng.uniform(0.1, 2.0, n)
    loss, g = G.loss_and_grad(params, q, wi, w)
    idx = np.concatenate([rng.choice(256 * 64, 40, replace=False),
                          256 * 64 + rng.choice(4 * 4096, 60, replace=False)])
    # output-layer columns that are used (0..39) must be among those checked
    idx = np.concatenate([idx, [256 * 64 + 3 * 4096 + k * 64 + m for k in (0, 17, 40) for m in (0, 4, 21, 39)]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (G.loss_and_grad(pp, q, wi, w)[0] - G.loss_and_grad(pm, q, wi, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])
    # unused output columns get no gradient
    W4 = g[256 * 64 + 3 * 4096:].reshape(64, 64)
    assert np.all(W4[:, 40:] == 0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(9)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = G.adam(p, np.zeros(100), np.zeros(100), 1, g, lr=1e-3)
    # first bias-corrected step: m_hat = g, v_hat = g^2
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def test_init_params_matches_he_uniform_bounds():
    p = G.init_params(1)
    Ws = G.split_params(p)
    assert np.all(np.abs(Ws[0]) <= np.sqrt(6 / 256) + 1e-7)
    assert np.all(np.abs(Ws[1]) <= np.sqrt(6 / 64) + 1e-7)
    assert np.all(np.abs(Ws[4][:, :40]) <= 0.1 * np.sqrt(6 / 64) + 1e-7)
    assert np.all(Ws[4][:, 40:] == 0)
    # a uniform draw fills its range
    assert np.abs(Ws[0]).max() > 0.95 * np.sqrt(6 / 256)


def test_gather_known_paths_and_loop_form():
    """The gather restatement: hand-made paths with known weights, and
    random records against a plain per-path loop in float64."""
    D, S = 5, 3
    V = np.zeros((4, D, S, 4), np.float32)
    E = np.zeros((D, S, 4), np.float32)
    # path 0: light arrives at depth 3; vertex 2 sampled with p = 2, R = 1
    E[3, 0, :3] = 1.0
    V[3, 2, 0] = [2.0, 1.0, 1.0, 1.0]
    V[0, 2, 0] = [0.1, 0.2, 0.3, 0.4]
    # vertex 1 of path 0 sees the same light through R = (2, 0, 0), p = 4
    V[3, 1, 0] = [4.0, 2.0, 0.0, 0.0]
    # path 1: a vertex below its only contribution (no light after it)
    E[1, 1, :3] = 5.0
    V[3, 2, 1] = [1.0, 1.0, 1.0, 1.0]
    # path 2: p = 0 (never sampled) is skipped
    E[4, 2, :3] = 1.0
    V[3, 3, 2] = [0.0, 1.0, 1.0, 1.0]
    out = G.gather(V, E, D - 1)
    assert out.shape == (2, 13)
    assert np.isclose(out[0, 12], (0.2126 + 0.7152 + 0.0722) / 2.0, rtol=1e-6)  # deepest first
    assert np.allclose(out[0, :4], [0.1, 0.2, 0.3, 0.4])
    assert np.isclose(out[1, 12], 0.2126 * 2.0 / 4.0, rtol=1e-6)

    rng = np.random.default_rng(4)
    D, S = 7, 200
    V = rng.uniform(0, 1, (4, D, S, 4)).astype(np.float32)
    V[3, :, :, 0] *= rng.uniform(size=(D, S)) > 0.3  # some vertices unsampled
    E = (rng.uniform(0, 1, (D, S, 4)) * (rng.uniform(size=(D, S, 1)) > 0.6)).astype(np.float32)
    out = G.gather(V, E, D - 1)
    ref = []
    for s in range(S):
        acc = np.zeros(3)
        for j in range(D - 1, 0, -1):
            p, r = float(V[3, j, s, 0]), V[3, j, s, 1:].astype(np.float64)
            if p > 0:
                w = float(np.dot([0.2126, 0.7152, 0.0722], acc * r)) / p
                if w > 0:
                    ref.append(np.concatenate([V[0, j, s], V[1, j, s], V[2, j, s], [w]]))
            acc += E[j, s, :3]
    ref = np.array(ref)
    assert out.shape == ref.shape
    assert np.allclose(out[:, :12], ref[:, :12])
    assert np.allclose(out[:, 12], ref[:, 12], rtol=1e-5)


def test_frame_is_orthonormal_and_right_handed():
    rng = np.random.default_rng(9)
    n = rng.normal(size=(20000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n[:6] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]]
    t, b = G.frame(n)
    for v in (t, b):
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-12)
    assert np.allclose(np.sum(t * b, 1), 0, atol=1e-12)
    assert np.allclose(np.sum(t * n, 1), 0, atol=1e-12) and np.allclose(np.sum(b * n, 1), 0, atol=1e-12)
    assert np.allclose(np.cross(t, b), n, atol=1e-12)
