"""Shared helpers of the parity tests: stage runners on the oracle and
comparison metrics."""
import ctypes as C

import numpy as np


def relerr(a, b, floor=1e-9):
    """max_i |a_i - b_i| / max(|b_i|, floor * max|b|): elementwise relative
    error, with entries far below the field's scale measured against it."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    if np.isnan(a).any() or np.isnan(b).any():
        if not np.array_equal(np.isnan(a), np.isnan(b)):
            return np.inf
        m = ~np.isnan(b)
        a, b = a[m], b[m]
        if b.size == 0:
            return 0.0
    inf_a, inf_b = np.isinf(a), np.isinf(b)
    if inf_a.any() or inf_b.any():
        if not (np.array_equal(inf_a, inf_b) and np.array_equal(a[inf_a], b[inf_b])):
            return np.inf
        a, b = a[~inf_a], b[~inf_b]
    scale = np.max(np.abs(b)) if b.size else 0.0
    den = np.maximum(np.abs(b), floor * scale)
    den[den == 0] = 1.0
    return float(np.max(np.abs(a - b) / den)) if b.size else 0.0


def row_relerr(Ic, A, B):
    """CSR values: max over rows of max_j |A_ij - B_ij| / max_j |B_ij|."""
    worst = 0.0
    d = np.abs(A - B)
    for r in range(len(Ic) - 1):
        s, e = Ic[r], Ic[r + 1]
        if e > s:
            m = np.max(np.abs(B[s:e]))
            if m > 0:
                worst = max(worst, float(np.max(d[s:e]) / m))
            elif np.max(d[s:e]) > 0:
                worst = np.inf
    return worst


def salloc(orc, S, neq):
    orc.orc_salloc(C.byref(S), neq)


def numbering(orc, mdz, m):
    S = mdz.SparseMat()
    orc.orc_EvalNumberOfEquations(m.m.mesh, C.byref(S))
    return S


def blocks(orc, mdz, S, n=4):
    out = [mdz.SparseMat() for _ in range(n)]
    for i, b in enumerate(out):
        salloc(orc, b, S.neq_mom if i % 4 < 2 else S.neq_cont)
    return out


def oracle_rheology_inputs(orc, m):
    """Strain rates of the current iterate (stage before the rheology)."""
    mm = m.m
    orc.orc_StrainRateComponents(mm.mesh, mm.scaling, C.byref(mm.model))


def oracle_update(orc, m):
    mm = m.m
    orc.orc_UpdateNonLinearity(mm.mesh, mm.materials, C.byref(mm.model), C.byref(mm.Nmodel), mm.scaling)
