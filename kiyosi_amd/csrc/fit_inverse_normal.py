"""Fits the single-precision inverse normal used by the FP32 path kernels.

Form (after M. Giles, "Approximating the erfinv function", GPU Computing Gems
Jade Edition, 2011, which writes erfinv(x) = x * P(w) with
w = -ln((1 - x)(1 + x))): for p in (0, 1) let x = 2p - 1 and
w = -ln(4 p (1 - p)) and l = -w / ln 2 = log2(4 p (1 - p)), the value the
hardware's log2 returns; then

    Phi^-1(p) = x * P_c(l)                if w < 6
    Phi^-1(p) = x * P_t(sqrt(w) - 3)      otherwise

with polynomials of degree 7 (central) and 8 (tail) fitted here, by least
squares on Chebyshev nodes, to SciPy's double-precision ndtri. The central
polynomial runs in l itself, so the hot path spends no instruction on a
change of variable. The split at w = 6 (p below 6.2e-4 or above 1 - 6.2e-4)
keeps the tail branch rare: in float32 the degree-7 central polynomial
reaches 1.2e-6 up to w = 6, and moving the split to w = 5 would only make
the tail, 11 more instructions for a whole wave, run 2.5 times as often.
The worst case over all inputs is 1.2e-6. The coefficients
are this script's own fit, printed as C++ for fp32_math.hpp.
Run: python fit_inverse_normal.py
"""

import numpy as np
from scipy.special import ndtri

DEGREE_CENTRAL = 7
DEGREE_TAIL = 8
W_SPLIT = 6.0
P_MIN = 2.0 ** -33  # smallest midpoint of a 32-bit uniform


def target(w):
    # p < 1/2 branch: 1 - x^2 = exp(-w), x = -sqrt(-expm1(-w)).
    x = -np.sqrt(-np.expm1(-w))
    p = 0.5 * (1.0 + x)
    return ndtri(p) / x


def cheb_nodes(a, b, n):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * n))


def fit(u, values, degree):
    coeffs = np.polynomial.polynomial.polyfit(u, values, degree)
    return coeffs[::-1]  # highest degree first, for Horner


def horner(coeffs, x):
    r = np.full_like(x, coeffs[0])
    for c in coeffs[1:]:
        r = (r.astype(np.float64) * x + c).astype(np.float32)
    return r


def main():
    w_max = -np.log(4.0 * P_MIN * (1.0 - P_MIN))
    wc = cheb_nodes(1e-12, W_SPLIT, 4000)
    central = fit(-wc / np.log(2.0), target(wc), DEGREE_CENTRAL)
    wt = cheb_nodes(W_SPLIT, w_max * 1.01, 4000)
    tail = fit(np.sqrt(wt) - 3.0, target(wt), DEGREE_TAIL)

    # Float32 evaluation check against ndtri on a dense grid of 32-bit inputs.
    bits = np.unique(np.concatenate([
        np.arange(0, 1 << 16, dtype=np.uint64),
        np.linspace(0, 2**32 - 1, 2_000_001).astype(np.uint64),
        (2**32 - 1 - np.arange(0, 1 << 16, dtype=np.uint64)),
    ]))
    # As the kernel evaluates it: s = bits read as int32,
    # q = fl(fl(s) * -2^-32 - 2^-33) (one fma), ps = |q| = min(p, 1 - p),
    # l = log2(4 ps (1 - ps)), central variable l and tail variable
    # sqrt(ln2) sqrt(-l) - 3, Horner in float32 with one rounding per step as
    # fmaf gives; the sign of the normal is the sign of q.
    s = bits.astype(np.int64)
    s = np.where(s >= 2**31, s - 2**32, s)
    q = (s.astype(np.float32).astype(np.float64) * -2.0 ** -32 - 2.0 ** -33).astype(np.float32)
    ps = np.abs(q)
    ax = (1.0 - 2.0 * ps.astype(np.float64)).astype(np.float32)
    four = (4.0 - 4.0 * ps.astype(np.float64)).astype(np.float32)
    l = np.log2((ps * four).astype(np.float32)).astype(np.float32)
    t = (np.sqrt(np.log(2.0)) * np.sqrt(-l.astype(np.float32)).astype(np.float64) - 3.0).astype(np.float32)
    c32 = central.astype(np.float32)
    t32 = tail.astype(np.float32)
    pc = horner(c32, l)
    pt = horner(t32, t)
    z = np.copysign(np.where(l <= np.float32(-W_SPLIT / np.log(2.0)), pt, pc).astype(np.float32) * ax, q)
    exact = ndtri((bits.astype(np.float64) + 0.5) * 2.0 ** -32)
    err = np.abs(z.astype(np.float64) - exact)
    rel = err / np.maximum(np.abs(exact), 1e-3)
    print(f"// max abs error {err.max():.3e}, max rel error {rel.max():.3e} over {bits.size} inputs")
    for name, c in (("CENTRAL", central), ("TAIL", tail)):
        body = ", ".join(f"{v:.9e}f" for v in c.astype(np.float32))
        print(f"constexpr float {name}[{c.size}] = {{{body}}};")


if __name__ == "__main__":
    main()
