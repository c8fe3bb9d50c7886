"""Derive the polynomial coefficients used by tr_fmath.h.

Each polynomial is a weighted least-squares fit in float64 on 4097
Chebyshev nodes of the reduced interval, weighted for relative error, then
rounded to float32.  Run it to reproduce the constants in tr_fmath.h:

    python tempestrenderer_amd/csrc/fit_fmath.py
"""
import numpy as np


def cheb_nodes(a, b, n=4097):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (k + 0.5) / n)


def fit(xs, target, basis, weight):
    A = np.stack([f(xs) for f in basis], axis=1) * weight[:, None]
    coef, *_ = np.linalg.lstsq(A, target * weight, rcond=None)
    return coef


def main():
    # sin(2*pi*x) = x * P(x^2),  |x| <= 1/8, P of degree 4 in x^2
    x = cheb_nodes(1e-6, 0.125)
    y = x * x
    sin_c = fit(x, np.sin(2 * np.pi * x) / x, [lambda _, k=k: y ** k for k in range(5)],
                1.0 / (np.sin(2 * np.pi * x) / x))
    # cos(2*pi*x) = 1 + y * Q(y), Q of degree 4
    cos_c = fit(x, (np.cos(2 * np.pi * x) - 1.0) / y, [lambda _, k=k: y ** k for k in range(5)],
                y / np.cos(2 * np.pi * x))
    # log(m) = 2s + s*z*R(z), s = (m-1)/(m+1), z = s^2, m in [sqrt(1/2), sqrt(2)]
    s = cheb_nodes(1e-6, (np.sqrt(2) - 1) / (np.sqrt(2) + 1))
    z = s * s
    logm = np.log((1 + s) / (1 - s))
    log_c = fit(s, (logm - 2 * s) / (s * z), [lambda _, k=k: z ** k for k in range(4)],
                s * z / np.abs(logm))
    # exp(r) = 1 + r + r^2 * E(r), |r| <= ln2/2, E of degree 5
    r = cheb_nodes(-0.5 * np.log(2), 0.5 * np.log(2))
    r = r[np.abs(r) > 1e-6]
    exp_c = fit(r, (np.exp(r) - 1 - r) / (r * r), [lambda _, k=k: r ** k for k in range(6)],
                r * r / np.exp(r))
    for name, c in (("sin", sin_c), ("cos", cos_c), ("log", log_c), ("exp", exp_c)):
        print(name, ", ".join("%.9ef" % np.float32(v) for v in c))


if __name__ == "__main__":
    main()
