"""Shared checks of the tests."""
from __future__ import annotations

import numpy as np


def rel_err_vec(a: np.ndarray, b: np.ndarray, floor: float = 1.0) -> float:
    """Max over rows of |a-b|_inf / max(|b|_inf, floor): relative error of
    each vector against its magnitude (floor: the grid-spacing scale)."""
    if np.asarray(a).size == 0 and np.asarray(b).size == 0:
        return 0.0
    a = np.asarray(a, np.float64).reshape(len(a), -1)
    b = np.asarray(b, np.float64).reshape(len(b), -1)
    num = np.abs(a - b).max(axis=1)
    den = np.maximum(np.abs(b).max(axis=1), floor)
    return float((num / den).max())


def dax_test_equal(a, b, tol: float = 1e-4) -> bool:
    """The reference's test_equal (dax/testing/Testing.h:405-431): per
    component, both below 2*tol in magnitude, or ratio within 1 +- tol."""
    a = np.atleast_1d(np.asarray(a, np.float64)).ravel()
    b = np.atleast_1d(np.asarray(b, np.float64)).ravel()
    small = (np.abs(a) < 2 * tol) & (np.abs(b) < 2 * tol)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = a / b
    ok = small | ((ratio > 1.0 - tol) & (ratio < 1.0 + tol))
    return bool(ok.all())
