"""Provenance of the exp2 polynomial used by the blend (gr::exp2_poly in
3dgaussianrenderer_amd/csrc/gr_math.hpp and its host restatement in
oracle/mirror.cpp).

The blend needs density = 2^e for e in [-26, 0].  The build evaluates it as
n = floor(e), f = min(e - n, 1 - 2^-24) in [0, 1), 2^f ~ p(f),
result = ldexp(p(f), n), so that the GPU and the host oracle compute the same
bits (CUDA's __expf, which the reference calls at core/gaussian_render.cu:733,
is closed code).  On gfx950 f is one v_fract_f32 (x - floor(x), rounded to
nearest even and clamped below 1) and n one v_cvt_flr_i32_f32.

p is a degree-4 polynomial with p(0) = 1 exactly and a near-minimax relative
error on [0, 1]: Lawson-weighted least squares (iteratively reweighted by the
current error) on 4000 Chebyshev nodes.  The script prints the fitted FP32
coefficients (c0 .. c4, the constants kExp2C0 .. kExp2C4), the maximum
relative error of the FP32 Horner/FMA evaluation on a dense grid of [0, 1)
(about 3.0e-6) and p at the largest f below 1, which must stay below 2 so that
2^e <= 1 for every e <= 0.  The committed constants were produced by this
script; the oracle's tests pin the polynomial's accuracy against numpy's exp2,
its exact powers of two and its bound by 1 (tests/test_oracle.py).

    python3 tests/golden/fit_exp2.py
"""
import numpy as np

DEGREE = 4


def fit(deg: int = DEGREE, nodes: int = 4000, iters: int = 300) -> np.ndarray:
    """FP32 coefficients c0..c_deg of p(f) = 1 + sum_{k>=1} c_k f^k, fitted on [0, 1]."""
    x = (np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes) + 1.0) * 0.5
    y = 2.0 ** x
    w = np.ones_like(x)
    c = None
    for _ in range(iters):
        # minimise the relative error (p - y) / y with c0 fixed at 1
        A = np.vander(x, deg + 1, increasing=True)[:, 1:] / y[:, None]
        b = (y - 1) / y
        Aw = A * np.sqrt(w)[:, None]
        c, _, _, _ = np.linalg.lstsq(Aw, np.sqrt(w) * b, rcond=None)
        err = np.abs(A @ c - b)
        w = w * err
        w /= w.sum()
    return np.concatenate([[1.0], c]).astype(np.float32)


def horner_fp32(c32: np.ndarray, xx: np.ndarray) -> np.ndarray:
    """p(xx) evaluated as the kernels do: FP32 Horner steps, each a fused multiply-add."""
    def fma(a, b, c):
        return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)

    p = np.full_like(xx, c32[-1])
    for k in range(len(c32) - 2, -1, -1):
        p = fma(p, xx, np.float32(c32[k]) * np.ones_like(xx))
    return p


def main() -> None:
    c32 = fit()
    xx = np.linspace(0.0, 1.0, 400001).astype(np.float32)
    xx[-1] = np.nextafter(np.float32(1.0), np.float32(0.0))   # the largest f
    p = horner_fp32(c32, xx)
    rel = np.abs(p.astype(np.float64) / 2.0 ** xx.astype(np.float64) - 1)
    print("max rel err", rel.max(), "p(0) =", p[0], "p(1 - 2^-24) =", p[-1], "(< 2:", bool(p.max() < 2.0), ")")
    for v in c32:
        print("%.9e" % v)


if __name__ == "__main__":
    main()
