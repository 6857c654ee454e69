"""The fit behind the erfc approximation of the pair kernels
(mdir_amd/csrc/pairs.hip, erfc_from_exp; mdir_amd/csrc/adjoint.hip,
erfc_from_exp_t; the oracle's f32 rows, md_row):

    erfc(x) ~= t exp(-x^2) P(t),  t = 1 / (1 + 0.475 x),  P of degree 6,

on x in [0, 3.6] (beta r up to beta rc = 3.12 at rc 0.9 nm and a
tolerance of 1e-5, with room).  Method: P fits erfc(x) exp(x^2) / t by
least squares in t on 4,001 Chebyshev-spaced points of x, weighted by
1 / (erfc(x) exp(x^2) / t) so that the relative error of erfc is what is
minimised; then two passes of iterative reweighting by the current error
magnitude (a Lawson-style step towards the minimax fit).

    python tests/golden/erfc_fit.py      # prints the coefficients
"""
import numpy as np
from scipy.special import erfc

# the coefficients the kernels carry, highest degree first
KERNEL = [0.09289151430130005, -0.4289827346801758, 0.6025476455688477,
          -0.1264616847038269, 0.3417808711528778, 0.24873463809490204,
          0.26948970556259155]


def approx(coeffs, x):
    t = 1.0 / (1.0 + 0.475 * x)
    return t * np.exp(-x * x) * np.polyval(coeffs, t)


def relative_error(coeffs, x):
    return approx(coeffs, x) / erfc(x) - 1.0


def fit(degree=6, lo=0.0, hi=3.6, npts=4001, passes=2):
    k = np.arange(npts)
    x = 0.5 * (lo + hi) + 0.5 * (hi - lo) * np.cos(np.pi * (k + 0.5) / npts)
    t = 1.0 / (1.0 + 0.475 * x)
    target = erfc(x) * np.exp(x * x) / t
    weight = 1.0 / target
    A = np.vander(t, degree + 1)
    for _ in range(passes + 1):
        coeffs, *_ = np.linalg.lstsq(A * weight[:, None], target * weight, rcond=None)
        err = np.abs(np.polyval(coeffs, t) / target - 1.0)
        weight = weight * np.sqrt(err / err.max() + 1e-3)
    return coeffs


if __name__ == "__main__":
    c = fit()
    x = np.linspace(0.0, 3.6, 200001)
    print("coefficients (highest degree first):", [float(v) for v in c])
    print(f"max relative error of the fit {np.abs(relative_error(c, x)).max():.3e}")
    print(f"max relative error of the kernels' {np.abs(relative_error(KERNEL, x)).max():.3e}")
