"""Provenance of the exp2 polynomial used by the blend (gr::exp2_poly in
3dgaussianrenderer_amd/csrc/gr_math.hpp and its host restatement in
oracle/mirror.cpp).

The blend needs density = 2^e for e in [-26, 0].  The build evaluates it as
n = rint(e), f = e - n in [-1/2, 1/2], 2^f ~ p(f), result = ldexp(p(f), n), so
that the GPU and the host oracle compute the same bits (CUDA's __expf, which
the reference calls at core/gaussian_render.cu:733, is closed code).

p is a degree-5 polynomial with p(0) = 1 exactly and a near-minimax relative
error on [-1/2, 1/2]: Lawson-weighted least squares (iteratively reweighted
by the current error) on 4000 Chebyshev nodes.  The script prints the fitted
FP32 coefficients (c0 .. c5, the constants kExp2C0 .. kExp2C5) and the
maximum relative error of the FP32 Horner/FMA evaluation on a dense grid
(about 1.7e-7).  The committed constants were produced by this script; the
oracle's tests pin the polynomial's accuracy against numpy's exp2
(tests/test_oracle.py).

    python3 tests/golden/fit_exp2.py
"""
import numpy as np


def fit(deg: int = 5, nodes: int = 4000, iters: int = 300) -> np.ndarray:
    """FP32 coefficients c0..c_deg of p(f) = 1 + sum_{k>=1} c_k f^k."""
    x = np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes) * 0.5
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
    xx = np.linspace(-0.5, 0.5, 400001).astype(np.float32)
    p = horner_fp32(c32, xx)
    rel = np.abs(p.astype(np.float64) / 2.0 ** xx.astype(np.float64) - 1)
    print("max rel err", rel.max(), "p(0)=", p[200000], "max p on [-.5,0]", p[:200001].max())
    for v in c32:
        print("%.9e" % v)


if __name__ == "__main__":
    main()
