"""Comparing two chi(k): the R-space R-factor with a Hanning-tapered k window.

Same definition as feffgpu's parity metric (src/feffgpu/compare.py:6-27): Hanning sills of width ``dk`` centred
on ``kmin`` and ``kmax``; FT of k^kweight chi over ``nfft`` points scaled by dk/sqrt(pi); R-factor
sum |FT(fit) - FT(data)|^2 / sum |FT(data)|^2 over rmin <= r <= rmax.
"""
from __future__ import annotations

import math

import numpy as np


def hanning(k, kmin, kmax, dk=1.0):
    k = np.asarray(k, float)
    if dk <= 0:
        return ((k >= kmin) & (k <= kmax)).astype(float)
    rise = 0.5 * (1.0 - np.cos(math.pi * np.clip((k - kmin + dk / 2) / dk, 0.0, 1.0)))
    fall = 0.5 * (1.0 + np.cos(math.pi * np.clip((k - kmax + dk / 2) / dk, 0.0, 1.0)))
    return rise * fall


def r_factor(k, data, fit, kmin=3.0, kmax=23.0, rmin=1.0, rmax=6.0, kweight=2, nfft=2048, dk=1.0):
    k, data, fit = (np.asarray(v, float) for v in (k, data, fit))
    step = k[1] - k[0]
    w = hanning(k, kmin, kmax, dk) * k ** kweight

    def ft(c):
        return np.fft.fft(c * w, n=nfft)[: nfft // 2] * step / math.sqrt(math.pi)

    r = np.arange(nfft // 2) * math.pi / (nfft * step)
    sel = (r >= rmin) & (r <= rmax)
    d, m = ft(data)[sel], ft(fit)[sel]
    return float(np.sum(np.abs(m - d) ** 2) / np.sum(np.abs(d) ** 2))


def max_rel(ref, got):
    """max |got - ref| / max |ref|."""
    ref, got = np.asarray(ref), np.asarray(got)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))
