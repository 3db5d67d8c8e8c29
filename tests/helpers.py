"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

# Strictly positive fields compare elementwise: |a - b| / |b| at every point.
# Winds cross zero, where an elementwise ratio is undefined; they compare
# normwise per component (each species' vertical wind, each of the three
# wind components): max|a - b| / max|b| over all columns and levels.
# ``floored_errors`` gives the stricter elementwise ratio with a floor of 1e-3
# of the component's largest magnitude, for reporting.
POSITIVE = ("nds", "temp", "rho", "nden", "ions")
SIGNED = ("vvel", "vel")
ALL = ("nds", "vvel", "vel", "temp", "rho", "logns", "nden", "ions")


def _components(x: np.ndarray):
    """Split a [column][component][level] field into its components."""
    if x.ndim == 3:
        return [x[:, c, :] for c in range(x.shape[1])]
    return [x]


def rel_errors(out: dict, ref: dict, fields=ALL) -> dict:
    errs = {}
    for k in fields:
        if k not in out or k not in ref:
            continue
        a = np.asarray(out[k], dtype=np.float64)
        b = np.asarray(ref[k], dtype=np.float64)
        if a.size == 0:
            errs[k] = 0.0
        elif k in SIGNED:
            e = 0.0
            for ac, bc in zip(_components(a), _components(b)):
                scale = np.max(np.abs(bc))
                d = np.max(np.abs(ac - bc))
                e = max(e, float(d / scale) if scale > 0 else float(d))
            errs[k] = e
        else:
            d = np.abs(a - b) / np.maximum(np.abs(b), np.finfo(np.float64).tiny)
            errs[k] = float(np.max(d))
    return errs


def floored_errors(out: dict, ref: dict, fields=SIGNED) -> dict:
    """Elementwise relative error of the winds with a floor of 1e-3 of each
    component's largest magnitude."""
    errs = {}
    for k in fields:
        if k not in out or k not in ref:
            continue
        e = 0.0
        a = np.asarray(out[k], dtype=np.float64)
        b = np.asarray(ref[k], dtype=np.float64)
        for ac, bc in zip(_components(a), _components(b)):
            floor = 1e-3 * np.max(np.abs(bc))
            if ac.size:
                e = max(e, float(np.max(np.abs(ac - bc) / np.maximum(np.abs(bc), floor))))
        errs[k] = e
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


def ulp_perturbed(cols: dict, seed: int = 0, fields=("nds", "vvel", "vel", "temp", "rho")) -> dict:
    """A copy of ``cols`` with every value of ``fields`` moved one ulp up or
    down at random."""
    rng = np.random.default_rng(seed)
    out = {k: np.array(v) for k, v in cols.items()}
    for k in fields:
        direction = rng.choice([-np.inf, np.inf], size=out[k].shape)
        out[k] = np.nextafter(out[k], direction)
    return out


def ulp_envelope(cols: dict, params, solver_id: int, n_steps: int, ref: dict,
                 threads: int = 8) -> dict:
    """How far the reference algorithm itself moves after ``n_steps`` when its
    input is perturbed by one ulp: the oracle (bit-identical to the reference
    build) from two independently 1-ulp-perturbed inputs, compared with the
    unperturbed result ``ref``; the larger error per field."""
    from oracle import cpu

    env = {}
    for seed in (1, 2):
        pert = cpu.advance(ulp_perturbed(cols, seed), params, solver_id, n_steps,
                           threads=threads)
        for k, v in rel_errors(pert, ref).items():
            env[k] = max(env.get(k, 0.0), v)
    return env


def assert_within_envelope(out: dict, ref: dict, env: dict, rtol: float,
                           what: str = "") -> dict:
    """Multi-step trajectories: every field within max(rtol, 2 x envelope)."""
    errs = rel_errors(out, ref)
    bad = {k: (v, env[k]) for k, v in errs.items() if not (v <= max(rtol, 2.0 * env[k]))}
    assert not bad, f"{what}: outside the 1-ulp envelope: {bad} (all: {errs}, env: {env})"
    for k in ALL:
        if k in out:
            assert np.all(np.isfinite(out[k])), f"{what}: non-finite values in {k}"
    return errs
