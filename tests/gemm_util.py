"""Shared helpers for the tests: golden fixtures, operand views, error metrics,
and host<->device copies of column-major buffers."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U64 = 2.0 ** -53
U32 = 2.0 ** -24


def notrans(c: str) -> bool:
    return c in ("N", "n")


def stored_shape(trans: str, rows: int, cols: int):
    """Stored (rows, cols) of an operand whose op() is rows x cols."""
    return (rows, cols) if notrans(trans) else (cols, rows)


def op(buf, trans, rows, cols):
    """Logical op(X) (rows x cols) of a Fortran-ordered ld x * buffer."""
    r, c = stored_shape(trans, rows, cols)
    M = buf[:r, :c]
    return M if notrans(trans) else M.T


def load_fixtures():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        cases = json.load(f)
    vals = np.fromfile(os.path.join(GOLDEN, "values.f64"), dtype="<f8")
    idx = np.fromfile(os.path.join(GOLDEN, "index.i32"), dtype="<i4")
    for c in cases:
        c["values"] = vals[c["voff"]:c["voff"] + c["count"]]
        if c["kind"] == "sampled":
            ij = idx[c["ioff"]:c["ioff"] + 2 * c["count"]].reshape(-1, 2)
            c["is"], c["js"] = ij[:, 0].copy(), ij[:, 1].copy()
    return cases


def case_inputs(case, oracle_mod):
    """Host buffers (A, B, C) of a fixture case, from the generator."""
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    m, n, k = case["m"], case["n"], case["k"]
    ra, ca = stored_shape(case["transa"], m, k)
    rb, cb = stored_shape(case["transb"], k, n)
    s, d, p = case["seed"], case["dist"], case["param"]
    A = oracle_mod.fill(ra, ca, case["lda"], s, 0, dt, d, p)
    B = oracle_mod.fill(rb, cb, case["ldb"], s, 1, dt, d, p)
    C = oracle_mod.fill(m, n, case["ldc"], s, 2, dt, d, p)
    return A, B, C


def abs_bound(A, B, C0, ta, tb, m, n, k, alpha, beta):
    """|alpha| |op(A)| |op(B)| + |beta C0|, the scale of componentwise errors."""
    oa = np.abs(op(A, ta, m, k).astype(np.float64))
    ob = np.abs(op(B, tb, k, n).astype(np.float64))
    bound = abs(alpha) * (oa @ ob) if k > 0 else np.zeros((m, n))
    c0 = C0[:m, :n].astype(np.float64)
    return bound + (abs(beta) * np.abs(c0) if beta != 1.0 else np.abs(c0))


def rel_fro(X, R):
    X = np.asarray(X, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    den = np.linalg.norm(R)
    return float(np.linalg.norm(X - R) / (den if den > 0 else 1.0))


def max_comp(X, R, bound):
    X = np.asarray(X, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    return float(np.max(np.abs(X - R) / (bound + 1e-300))) if X.size else 0.0


def to_device(buf, torch_mod, device="cuda"):
    """1-D device copy of a Fortran-ordered buffer (column-major memory)."""
    flat = np.ravel(buf, order="F")
    return torch_mod.from_numpy(flat.copy()).to(device)


def from_device(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")
