"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

# Fields whose values are strictly positive compare purely relatively; winds
# (which cross zero) use a floor of 1e-3 of the field's largest magnitude.
POSITIVE = ("nds", "temp", "rho", "nden")
SIGNED = ("vvel", "vel")
ALL = ("nds", "vvel", "vel", "temp", "rho", "logns", "nden")


def rel_errors(out: dict, ref: dict, fields=ALL) -> dict:
    errs = {}
    for k in fields:
        if k not in out or k not in ref:
            continue
        a = np.asarray(out[k], dtype=np.float64)
        b = np.asarray(ref[k], dtype=np.float64)
        if k in SIGNED:
            floor = 1e-3 * np.max(np.abs(b))
        else:
            floor = np.finfo(np.float64).tiny
        d = np.abs(a - b) / np.maximum(np.abs(b), floor)
        errs[k] = float(np.max(d)) if d.size else 0.0
    return errs


def assert_parity(out: dict, ref: dict, rtol: float, fields=ALL, what: str = "") -> dict:
    errs = rel_errors(out, ref, fields)
    bad = {k: v for k, v in errs.items() if not (v <= rtol)}
    assert not bad, f"{what}: relative error above {rtol:g}: {bad} (all: {errs})"
    for k in fields:
        if k in out:
            assert np.all(np.isfinite(out[k])), f"{what}: non-finite values in {k}"
    return errs


def subset(cols: dict, idx) -> dict:
    return {k: np.ascontiguousarray(v[idx]) for k, v in cols.items()}
