"""Shared comparison helpers for the parity tests."""
import numpy as np

FIELDS = ("rho", "ux", "uy", "cpos", "cneg", "psi")


def rel_l2(a, b, rows=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if rows is not None:
        a, b = a[rows], b[rows]
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / scale if scale > 0 else np.max(np.abs(a - b))


def field_errors(solver, ch, oracle, fields=FIELDS, rows=None):
    return {f: rel_l2(solver.get_field(ch, f), oracle.get_field(f), rows) for f in fields}


# Absolute floors (lattice units) below which a difference is rounding of a
# cancellation, not a discrepancy: velocities and E fields start near zero
# and grow from differences of O(1) populations / potentials.
ABS_FLOOR = {"rho": 1e-14, "ux": 1e-15, "uy": 1e-15, "cpos": 1e-14, "cneg": 1e-14, "psi": 1e-13,
             "ex": 1e-14, "ey": 1e-14}


def parity_error(a, b, field, rtol):
    """max|a-b| / (rtol * max|b| + floor): pass when <= 1."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / (rtol * np.max(np.abs(b)) + ABS_FLOOR[field]))


def assert_parity(solver, ch, oracle, rtol, fields=FIELDS, tag=""):
    worst = {}
    for f in fields:
        a, b = solver.get_field(ch, f), oracle.get_field(f)
        worst[f] = (parity_error(a, b, f, rtol), rel_l2(a, b), float(np.max(np.abs(a - b))))
    print(tag, " ".join(f"{k}: rel={v[1]:.2e} abs={v[2]:.2e}" for k, v in worst.items()), flush=True)
    bad = {k: v for k, v in worst.items() if not v[0] <= 1.0}
    assert not bad, f"{tag}: {bad}"
