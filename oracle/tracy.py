"""Oracle Tracy analytic solutions and final-time error metrics.

Restates src/utils/exact_solutions.py:4-79 (series with nt=200 terms) and
src/utils/error_calculation.py:9-56 (RMS, max and relative L2 of h and Se at
the last recorded time; hd = -15.24, eps0 = exp(alpha0*hd),
d = alpha0*(theta_s-theta_r)/Ks).
"""
from __future__ import annotations

import numpy as np


def exact_solution(x, y, t, test_number, L, alpha0, eps0, d, nt=200):
    def series(beta):
        s = 0.0
        for k in range(1, nt + 1):
            lind = np.pi * k / L
            mu = (beta ** 2 + lind ** 2) / d
            s = s + ((-1) ** k) * lind / mu * np.sin(lind * y) * np.exp(-mu * t)
        return s

    if test_number == 1:
        b1 = np.sqrt(0.25 * alpha0 ** 2 + (np.pi / L) ** 2)
        b3 = np.sqrt(0.25 * alpha0 ** 2 + (3 * np.pi / L) ** 2)
        s1, s2 = series(b1), series(b3)
        psi = (0.75 * np.sin(np.pi * x / L) * (np.sinh(b1 * y) / np.sinh(b1 * L) + (2 / (L * d)) * s1)
               - 0.25 * np.sin(3 * np.pi * x / L) * (np.sinh(b3 * y) / np.sinh(b3 * L) + (2 / (L * d)) * s2))
    elif test_number == 2:
        b1 = np.sqrt(0.25 * alpha0 ** 2 + (np.pi / L) ** 2)
        s1 = series(b1)
        psi = (np.sinh(b1 * y) / np.sinh(b1 * L) + (2 / (L * d)) * s1) * np.sin(np.pi * x / L)
    elif test_number == 3:
        b1 = np.sqrt(0.25 * alpha0 ** 2 + (2 * np.pi / L) ** 2)
        s1 = 0.0
        for k1 in range(nt + 1):
            lind = np.pi * k1 / L
            mu1 = (0.25 * alpha0 ** 2 + lind ** 2) / d
            mu2 = (0.25 * alpha0 ** 2 + lind ** 2 + (2 * np.pi / L) ** 2) / d
            s1 = s1 + ((-1) ** k1) * lind * (1 / mu1 * np.exp(-mu1 * t)
                                             - 1 / mu2 * np.cos(2 * np.pi * x / L) * np.exp(-mu2 * t)) * np.sin(lind * y)
        psi = (np.sinh(0.5 * alpha0 * y) / np.sinh(0.5 * alpha0 * L)
               - np.cos(2 * np.pi * x / L) * (np.sinh(b1 * y) / np.sinh(b1 * L)) + (2 / (L * d)) * s1)
    else:
        raise ValueError("Invalid test number. Please choose 1, 2, or 3.")
    if test_number in (1, 2):
        hb = (1 - eps0) * np.exp(0.5 * alpha0 * (L - y)) * psi
    else:
        hb = 0.5 * (1 - eps0) * np.exp(0.5 * alpha0 * (L - y)) * psi
    hex_ = np.log(np.maximum(eps0 + hb, 1e-10)) / alpha0
    return hex_, np.exp(alpha0 * hex_)


def final_errors(points, h_final, theta_final, t_final, test_number,
                 alpha0=0.164, thetas=0.45, thetar=0.15, Ks=0.1, L=15.24):
    hd = -15.24
    eps0 = np.exp(alpha0 * hd)
    d = alpha0 * (thetas - thetar) / Ks
    h_ex, se_ex = exact_solution(points[:, 0], points[:, 1], t_final, test_number, L, alpha0, eps0, d)
    se = (theta_final - thetar) / (thetas - thetar)
    return {
        "l2_pressure": float(np.sqrt(np.mean((h_final - h_ex) ** 2))),
        "linf_pressure": float(np.max(np.abs(h_final - h_ex))),
        "l2_saturation": float(np.sqrt(np.mean((se - se_ex) ** 2))),
        "linf_saturation": float(np.max(np.abs(se - se_ex))),
        "l2_relative_pressure": float(np.linalg.norm(h_final - h_ex) / np.linalg.norm(h_ex)),
        "l2_relative_saturation": float(np.linalg.norm(se - se_ex) / np.linalg.norm(se_ex)),
    }
