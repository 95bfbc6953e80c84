"""GPU tests of deferred diagnostics (mdz_ctx_defer_diagnostics): units that
write only the rheology outputs the unit reads give the same residual norms
and CSR blocks as units with every output, bit for bit, and the grid that
mdz_ctx_download_grid hands back (every output field, the diagnostics
recomputed at the download) is the same bit for bit, also after a line
search and a new iterate uploaded since the last unit, and in the mixed and
learned modes."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu
DP, FP = C.POINTER(C.c_double), C.POINTER(C.c_float)


def _network(mdz, hot, m):
    """an untrained network (seeded initial weights): any guess serves, the
    point is that both contexts start from the same one"""
    h = C.c_void_p()
    mdz.check(hot.mdz_mlp_create(C.byref(h), C.byref(m.m.materials), C.byref(m.m.model), C.byref(m.m.scaling)))
    try:
        mdz.check(hot.mdz_mlp_init(h, 11))
        w = np.zeros(mo.NPARAMS, np.float32)
        mdz.check(hot.mdz_mlp_get_weights(h, w.ctypes.data_as(FP), mo.NPARAMS))
        return w
    finally:
        hot.mdz_mlp_destroy(h)


def _blocks(mdz, hot, ctx):
    out = []
    for i in range(8):
        neq, nnz = C.c_int(), C.c_int()
        mdz.check(hot.mdz_ctx_block_info(ctx, i, C.byref(neq), C.byref(nnz)))
        A, b = np.zeros(max(nnz.value, 1)), np.zeros(max(neq.value, 1))
        mdz.check(hot.mdz_ctx_download_block(ctx, i, None, None, A.ctypes.data_as(DP), b.ctypes.data_as(DP), None))
        out.append((A, b))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fp64", "mixed", "learned"])
@pytest.mark.parametrize("model", ["litho", "pwl"])
def test_deferred_diagnostics_change_nothing_visible(mdz, hot, mode, model):
    base = mdz.Model.lithosphere(96, 48, 3) if model == "litho" else mdz.Model.pwl_inclusion(64, 64)
    w = _network(mdz, hot, base) if mode == 2 and model == "litho" else None
    if mode == 2 and w is None:
        base.free()
        pytest.skip("the network takes the three flow laws of the lithosphere model")
    ms = [base.clone(), base.clone()]
    ctx = [C.c_void_p(), C.c_void_p()]
    try:
        for c, m, defer in zip(ctx, ms, (0, 1)):
            mdz.check(hot.mdz_ctx_create(C.byref(c), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                         C.byref(m.m.scaling)))
            if w is not None:
                mdz.check(hot.mdz_ctx_set_network(c, w.ctypes.data_as(FP), mo.NPARAMS))
            mdz.check(hot.mdz_ctx_set_rheology_mode(c, mode))
            mdz.check(hot.mdz_ctx_defer_diagnostics(c, defer))
        res = []
        for c in ctx:
            r = (C.c_double * 3)()
            mdz.check(hot.mdz_ctx_newton_unit(c, r))
            mdz.check(hot.mdz_ctx_newton_unit(c, r))
            res.append(list(r))
        assert res[0] == res[1]
        for (A0, b0), (A1, b1) in zip(_blocks(mdz, hot, ctx[0]), _blocks(mdz, hot, ctx[1])):
            assert np.array_equal(A0, A1) and np.array_equal(b0, b1)
        # a line search leaves the last trial's fields on the grid
        neq_mom, neq_cont = C.c_int(), C.c_int()
        mdz.check(hot.mdz_ctx_block_info(ctx[0], 0, C.byref(neq_mom), None))
        mdz.check(hot.mdz_ctx_block_info(ctx[0], 3, C.byref(neq_cont), None))
        dx = np.random.default_rng(5).standard_normal(neq_mom.value + neq_cont.value) * 1e-6
        alpha = np.array([0.5, 1.0])
        out = [np.zeros(6), np.zeros(6)]
        for c, o in zip(ctx, out):
            mdz.check(hot.mdz_ctx_line_search(c, dx.ctypes.data_as(DP), 2, alpha.ctypes.data_as(DP),
                                              o.ctypes.data_as(DP)))
        assert np.array_equal(out[0], out[1])
        # a new iterate uploaded after the last rheology pass: the grid still
        # shows that pass's fields (the deferred ones come from a snapshot of p)
        nu, nw, nc = base.Nx * (base.Nz + 1), (base.Nx + 1) * base.Nz, (base.Nx - 1) * (base.Nz - 1)
        rng = np.random.default_rng(9)
        u1, v1, p1 = rng.standard_normal(nu), rng.standard_normal(nw), rng.standard_normal(nc)
        for c in ctx:
            mdz.check(hot.mdz_ctx_upload_solution(c, u1.ctypes.data_as(DP), v1.ctypes.data_as(DP),
                                                  p1.ctypes.data_as(DP)))
        for c, m in zip(ctx, ms):
            mdz.check(hot.mdz_ctx_download_grid(c, m.m.mesh))
        names = mdz.field_names(2)
        assert "Wtot" in names and "eII_pl" in names and "sxz" in names
        for n in names:
            assert np.array_equal(ms[0].field(n), ms[1].field(n), equal_nan=True), n
    finally:
        for c in ctx:
            if c:
                hot.mdz_ctx_destroy(c)
        for m in ms:
            m.free()
        base.free()


def test_defer_diagnostics_argument_errors(hot):
    assert hot.mdz_ctx_defer_diagnostics(None, 1) != 0
