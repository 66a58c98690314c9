"""Tracy's analytic solutions and the final-time error metrics of the reference CLI.

Restates src/utils/exact_solutions.py:4-79 (Fourier series, nt = 200 terms)
and src/utils/error_calculation.py:9-56 (RMS, max and relative L2 of h and of
Se = (theta - theta_r)/(theta_s - theta_r) at the last recorded time, with
hd = -15.24, eps0 = exp(alpha0 hd), d = alpha0 (theta_s - theta_r) / Ks).
The series is evaluated with the terms as an array axis.
"""
from __future__ import annotations

import numpy as np


def exact_solution(x, y, t, test_number, L, alpha0, eps0, d, nt=200):
    x = np.asarray(x, dtype=np.float64)[:, None]
    y = np.asarray(y, dtype=np.float64)[:, None]
    c = 2.0 / (L * d)
    if test_number in (1, 2):
        k = np.arange(1, nt + 1, dtype=np.float64)[None, :]
        lam = np.pi * k / L
        sign = np.where(k % 2 == 0, 1.0, -1.0)

        def mode(beta, freq):
            mu = (beta ** 2 + lam ** 2) / d
            s = np.sum(sign * lam / mu * np.sin(lam * y) * np.exp(-mu * t), axis=1)
            return np.sinh(beta * y[:, 0]) / np.sinh(beta * L) + c * s

        b1 = np.sqrt(0.25 * alpha0 ** 2 + (np.pi / L) ** 2)
        if test_number == 1:
            b3 = np.sqrt(0.25 * alpha0 ** 2 + (3 * np.pi / L) ** 2)
            psi = (0.75 * np.sin(np.pi * x[:, 0] / L) * mode(b1, 1)
                   - 0.25 * np.sin(3 * np.pi * x[:, 0] / L) * mode(b3, 3))
        else:
            psi = mode(b1, 1) * np.sin(np.pi * x[:, 0] / L)
        hb = (1 - eps0) * np.exp(0.5 * alpha0 * (L - y[:, 0])) * psi
    elif test_number == 3:
        k = np.arange(0, nt + 1, dtype=np.float64)[None, :]
        lam = np.pi * k / L
        sign = np.where(k % 2 == 0, 1.0, -1.0)
        w2 = (2 * np.pi / L) ** 2
        mu1 = (0.25 * alpha0 ** 2 + lam ** 2) / d
        mu2 = (0.25 * alpha0 ** 2 + lam ** 2 + w2) / d
        cosx = np.cos(2 * np.pi * x / L)
        s = np.sum(sign * lam * (np.exp(-mu1 * t) / mu1 - cosx * np.exp(-mu2 * t) / mu2) * np.sin(lam * y), axis=1)
        b1 = np.sqrt(0.25 * alpha0 ** 2 + w2)
        psi = (np.sinh(0.5 * alpha0 * y[:, 0]) / np.sinh(0.5 * alpha0 * L)
               - cosx[:, 0] * np.sinh(b1 * y[:, 0]) / np.sinh(b1 * L) + c * s)
        hb = 0.5 * (1 - eps0) * np.exp(0.5 * alpha0 * (L - y[:, 0])) * psi
    else:
        raise ValueError("Invalid test number. Please choose 1, 2, or 3.")
    h = np.log(np.maximum(eps0 + hb, 1e-10)) / alpha0
    return h, np.exp(alpha0 * h)


def final_errors(points, h, theta, t, test_number, alpha0=0.164, thetas=0.45, thetar=0.15, Ks=0.1, L=15.24):
    eps0 = np.exp(alpha0 * -15.24)
    d = alpha0 * (thetas - thetar) / Ks
    he, se_e = exact_solution(points[:, 0], points[:, 1], t, test_number, L, alpha0, eps0, d)
    se = (np.asarray(theta) - thetar) / (thetas - thetar)
    dh, ds = np.asarray(h) - he, se - se_e
    return {"l2_pressure": float(np.sqrt(np.mean(dh ** 2))), "linf_pressure": float(np.max(np.abs(dh))),
            "l2_saturation": float(np.sqrt(np.mean(ds ** 2))), "linf_saturation": float(np.max(np.abs(ds))),
            "l2_relative_pressure": float(np.linalg.norm(dh) / np.linalg.norm(he)),
            "l2_relative_saturation": float(np.linalg.norm(ds) / np.linalg.norm(se_e))}
