"""Error metrics used by the parity tests.

rel_mean(a, ref)  = sum|a - ref| / sum|ref|                      (bound 1e-3)
max_pixel(a, ref) = max over pixels of max_channel|a - ref| /
                    (max_channel|ref| + 1e-2 * mean|ref|)        (bound 1e-2)
The 1e-2*mean floor keeps near-black pixels from dividing by ~0.
"""
from __future__ import annotations

import numpy as np

REL_MEAN_TOL = 1e-3
MAX_PIXEL_TOL = 1e-2


def rel_mean(a, ref) -> float:
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.abs(ref).sum()
    return float(np.abs(a - ref).sum() / den) if den > 0 else float(np.abs(a - ref).sum())


def max_pixel(a, ref) -> float:
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    if a.ndim == 2:
        a = a[..., None]
        ref = ref[..., None]
    floor = 1e-2 * np.abs(ref).mean() + 1e-30
    err = np.abs(a - ref).max(axis=-1)
    mag = np.abs(ref).max(axis=-1)
    return float((err / (mag + floor)).max())


def assert_close(a, ref, what: str, rel_tol=REL_MEAN_TOL, pix_tol=MAX_PIXEL_TOL):
    rm, mp = rel_mean(a, ref), max_pixel(a, ref)
    assert np.isfinite(a).all(), f"{what}: non-finite values"
    assert rm <= rel_tol and mp <= pix_tol, f"{what}: rel_mean={rm:.3g} (<= {rel_tol}) max_pixel={mp:.3g} (<= {pix_tol})"
    return rm, mp


def pixel_errors(a, ref) -> np.ndarray:
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    if a.ndim == 1 or (a.ndim == 2 and a.shape[-1] not in (2, 3, 4)):
        a = a[..., None]
        ref = ref[..., None]
    floor = 1e-2 * np.abs(ref).mean() + 1e-30
    return np.abs(a - ref).max(axis=-1) / (np.abs(ref).max(axis=-1) + floor)


def assert_close_frac(a, ref, what: str, rel_tol=REL_MEAN_TOL, pix_tol=MAX_PIXEL_TOL, frac=0.01):
    """For outputs downstream of RIS selections: the FP64 oracle can take a
    different branch at a selection whose float comparison is within rounding
    of the random number, after which the pixel's path differs.  Bar:
    rel_mean <= rel_tol and at most `frac` of the pixels above pix_tol."""
    rm = rel_mean(a, ref)
    e = pixel_errors(a, ref)
    bad = float(np.mean(e > pix_tol))
    assert np.isfinite(a).all(), f"{what}: non-finite values"
    assert rm <= rel_tol and bad <= frac, f"{what}: rel_mean={rm:.3g} (<= {rel_tol}) pixels>{pix_tol}: {bad:.3%} (<= {frac:.1%})"
    return rm, bad
