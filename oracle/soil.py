"""Oracle soil models (vectorised over nodes).

van_genuchten follows src/models/van_genuchten.py:9-47 (phi = [alpha, theta_s,
theta_r, n, m, Ks], config/settings.py:19-22); exponential follows
src/models/exponential.py:10-42 (phi = [alpha0, theta_s, theta_r, Ks]).
Both return (C, K, theta) like the reference.
"""
from __future__ import annotations

import numpy as np


def van_genuchten(h, phi):
    alpha, ts, tr, n, m, ks = [float(v) for v in phi]
    h = np.asarray(h)
    neg = h < 0
    ah = alpha * np.abs(h)
    with np.errstate(all="ignore"):
        theta = np.where(neg, (ts - tr) * ((1 + ah ** n) ** (-m)) + tr, ts)
        se = np.where(neg, (theta - tr) / (ts - tr), 1.0)
        K = np.where(neg, ks * ((se ** 0.5) * (1 - (1 - se ** (1 / m)) ** m) ** 2), ks)
        C = np.where(neg, (ts - tr) * alpha * m * n * ah ** (n - 1) * (1 + ah ** n) ** (-m - 1), 0.0)
    return C, K, theta


def exponential(h, phi):
    a0, ts, tr, ks = [float(v) for v in phi]
    h = np.asarray(h)
    le = h <= 0
    e = np.exp(a0 * h)
    theta = np.where(le, (ts - tr) * e + tr, ts)
    K = np.where(le, ks * e, ks)
    C = np.where(le, a0 * (ts - tr) * e, 0.0)
    return C, K, theta


def soil(model: str, h, phi):
    if model in ("vg", "van_genuchten"):
        return van_genuchten(h, phi)
    if model in ("gardner", "exponential"):
        return exponential(h, phi)
    raise ValueError(f"unknown soil model {model!r}")
