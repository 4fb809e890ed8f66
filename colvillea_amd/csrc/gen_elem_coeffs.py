"""Generate the minimax-style polynomial coefficients used by clpt_elem.h.

Each reduced-range kernel is fitted by iteratively reweighted least squares
(Lawson's algorithm) on a dense grid in float64, with relative-error
weighting, then rounded to float32.  Run it to regenerate the numbers that
clpt_elem.h carries; tests check the resulting float functions against
float64 libm to a stated ulp bound.
"""
import numpy as np


def lawson(x, target, basis, iters=200):
    """Approximate minimax fit: minimise max |basis@c - target| over x."""
    A = np.stack([b(x) for b in basis], axis=1)
    w = np.ones_like(x) / len(x)
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], target * sw, rcond=None)
        err = np.abs(A @ c - target)
        w = w * (err + 1e-300)
        w /= w.sum()
    return c, np.max(np.abs(A @ c - target))


def series(coef, z, nterms=80):
    """Sum coef(k) * z**k for k = 0..nterms-1 (well-conditioned targets)."""
    out = np.zeros_like(z)
    for k in reversed(range(nterms)):
        out = out * z + coef(k)
    return out


def fit(name, zmax, g, deg, wpow=1):
    """Fit P(z) ~ g(z); the error is weighted by z**wpow, which is how an
    error in P reaches the final function value."""
    z = np.linspace(0.0, zmax, 20001)
    wgt = np.maximum(z, 1e-300) ** wpow
    basis = [lambda z, k=k: z ** k for k in range(deg + 1)]
    A = np.stack([b(z) for b in basis], axis=1) * wgt[:, None]
    c, e = lawson(z, g(z) * wgt, [lambda _z, j=j: A[:, j] for j in range(deg + 1)])
    c32 = c.astype(np.float32)
    print(f"// {name}: deg {deg} on z in [0, {zmax:.6g}], fit max abs err {e:.3e}")
    for k, v in enumerate(c32):
        print(f"#define {name}{k} {float(v)!r:>26}f")
    return c32


if __name__ == "__main__":
    q = (np.pi / 4) ** 2
    # sin(r) = r + r*z*S(z),  S(z) ~ (sin(sqrt z) - sqrt z) / (z sqrt z)
    from math import factorial, comb
    fit("CLPT_SIN_S", q, lambda z: series(lambda k: (-1) ** (k + 1) / factorial(2 * k + 3), z, 30), 3)
    # cos(r) = 1 - z/2 + z*z*C(z)
    fit("CLPT_COS_C", q, lambda z: series(lambda k: (-1) ** k / factorial(2 * k + 4), z, 30), 3, 2)
    # atan(t) = t + t*z*A(z),  |t| <= tan(pi/8)
    t8 = np.tan(np.pi / 8) ** 2
    fit("CLPT_ATAN_A", t8, lambda z: series(lambda k: (-1) ** (k + 1) / (2 * k + 3), z, 120), 6)
    # asin(x) = x + x*z*P(z),  |x| <= 0.5
    fit("CLPT_ASIN_P", 0.25, lambda z: series(
        lambda k: comb(2 * k + 2, k + 1) / (4 ** (k + 1) * (2 * k + 3)), z, 120), 5)
    # exp(r) = 1 + r + r*r*E(r), |r| <= ln2/2 (polynomial in r, not r^2)
    h = np.log(2) / 2
    r = np.linspace(-h, h, 20001)
    A = np.stack([r ** k for k in range(5)], axis=1) * (r * r)[:, None]
    tgt = series(lambda k: 1.0 / factorial(k + 2), r, 30) * r * r
    c, e = lawson(r, tgt, [lambda _r, j=j: A[:, j] for j in range(5)])
    print(f"// CLPT_EXP_E: deg 4 in r on [-ln2/2, ln2/2], fit max abs err {e:.3e}")
    for k, v in enumerate(c.astype(np.float32)):
        print(f"#define CLPT_EXP_E{k} {float(v)!r:>26}f")
    # log(1+f) = 2s + 2s*w*L(w), s = f/(2+f), w = s^2 <= (3-2*sqrt2)^2
    wmax = (3 - 2 * np.sqrt(2)) ** 2
    fit("CLPT_LOG_L", wmax, lambda w: series(lambda k: 1.0 / (2 * k + 3), w, 40), 4)
    # Cody-Waite split of pi/2 into a short head (exact k*C1 for |k| < 2^16),
    # a 12-bit middle part (exact k*C2 for |k| < 2^12) and a tail.
    import struct
    def trunc_bits(v, bits):
        m, e = np.frexp(v)
        return float(np.ldexp(np.floor(m * 2 ** bits) / 2 ** bits, e))
    hp = np.pi / 2
    c1 = trunc_bits(hp, 8)
    c2 = trunc_bits(hp - c1, 12)
    c3 = float(np.float32(hp - c1 - c2))
    print(f"#define CLPT_PIO2_C1 {c1!r}f\n#define CLPT_PIO2_C2 {c2!r}f\n#define CLPT_PIO2_C3 {c3!r}f")
