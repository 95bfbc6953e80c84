"""GPU tests of ensembles: independent lithosphere models of uneven grids
taken through one Newton unit each by one captured graph launch.

The graph runs the same kernels as the eager unit, so residual norms and CSR
values must equal the eager ones bit for bit on every replay, whatever block
shapes the autotuner picked.  Against the oracle, per-model residual norms
are held to 1e-12 relative in FP64 mode and 1e-10 in the learned and mixed
modes (the FP32 start changes eta within the local tolerance)."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as mo
from helpers import free_blocks, newton_unit, numbering

pytestmark = pytest.mark.gpu

# (ncx, ncz, seed, friction [deg], cohesion [Pa], extension rate [1/s])
MODELS = [(48, 24, 1, 15.0, 2.0e7, 3e-15), (96, 48, 2, 30.0, 1.0e7, 1e-16), (33, 16, 3, 10.0, 5.0e6, 1e-14),
          (130, 65, 4, 35.0, 5.0e7, 5e-16), (64, 32, 5, 22.0, 3.0e7, 2e-15)]
IP, DP = C.POINTER(C.c_int), C.POINTER(C.c_double)


class Ensemble:
    def __init__(self, mdz, hot):
        self.mdz, self.hot = mdz, hot
        self.models = [mdz.Model.lithosphere(*p) for p in MODELS]
        self.ctx = []
        for m in self.models:
            c = C.c_void_p()
            mdz.check(hot.mdz_ctx_create(C.byref(c), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                         C.byref(m.m.scaling)))
            self.ctx.append(c)
        self.ens = None

    def create(self):
        arr = (C.c_void_p * len(self.ctx))(*[c.value for c in self.ctx])
        self.ens = C.c_void_p()
        self.mdz.check(self.hot.mdz_ensemble_create(C.byref(self.ens), arr, len(self.ctx)))

    def residuals(self):
        r = np.zeros((len(self.ctx), 3))
        self.mdz.check(self.hot.mdz_ensemble_residuals(self.ens, r.ctypes.data_as(DP)))
        return r

    def block_values(self, i, which):
        neq, nnz = C.c_int(), C.c_int()
        self.mdz.check(self.hot.mdz_ctx_block_info(self.ctx[i], which, C.byref(neq), C.byref(nnz)))
        A = np.zeros(max(nnz.value, 1))
        b = np.zeros(max(neq.value, 1))
        self.mdz.check(self.hot.mdz_ctx_download_block(self.ctx[i], which, None, None, A.ctypes.data_as(DP),
                                                       b.ctypes.data_as(DP), None))
        return A, b

    def close(self):
        if self.ens:
            self.hot.mdz_ensemble_destroy(self.ens)
        for c in self.ctx:
            self.hot.mdz_ctx_destroy(c)
        for m in self.models:
            m.free()


@pytest.fixture
def ens(mdz, hot):
    e = Ensemble(mdz, hot)
    yield e
    e.close()


def _eager(e):
    res, vals = [], []
    for i, c in enumerate(e.ctx):
        r = (C.c_double * 3)()
        e.mdz.check(e.hot.mdz_ctx_newton_unit(c, r))
        res.append(list(r))
        vals.append([e.block_values(i, w) for w in (0, 4, 7)])
    return np.array(res), vals


def test_graph_replays_match_eager_units(mdz, hot, ens):
    """Autotuned shapes are valid; three graph replays, with eager units of
    the program in between, reproduce the eager norms and the Picard A,
    Jacobian A and Jacobian D values bit for bit."""
    ref_res, ref_vals = _eager(ens)
    ens.create()
    shapes = np.zeros((len(MODELS), 3), np.int32)
    br = C.c_int()
    mdz.check(hot.mdz_ensemble_autotune(ens.ens, 2, shapes.ctypes.data_as(IP), C.byref(br)))
    print("autotuned shapes", shapes.tolist(), "branches", br.value)
    assert br.value in (1, 2, 4, 8)
    for rh, pi, ja in shapes:
        assert rh in (64, 128, 256) and pi in (0, 64, 128, 256) and ja in (0, 64, 128)
    for replay in range(3):
        mdz.check(hot.mdz_ensemble_launch(ens.ens))
        mdz.check(hot.mdz_ensemble_synchronize(ens.ens))
        assert np.array_equal(ens.residuals(), ref_res), replay
        for i in range(len(MODELS)):
            for (A, b), (Ar, br_) in zip([ens.block_values(i, w) for w in (0, 4, 7)], ref_vals[i]):
                assert np.array_equal(A, Ar) and np.array_equal(b, br_), (replay, i)
        # the program's own eager work between replays
        k = replay % len(MODELS)
        r = (C.c_double * 3)()
        mdz.check(hot.mdz_ctx_newton_unit(ens.ctx[k], r))
        assert np.array_equal(np.array(list(r)), ref_res[k])


def _trained_network(mdz, hot, model):
    """A network trained briefly on the GPU (2^20 samples, 4 epochs): the
    FP64 polish makes every result independent of its quality."""
    h = C.c_void_p()
    mdz.check(hot.mdz_mlp_create(C.byref(h), C.byref(model.m.materials), C.byref(model.m.model),
                                 C.byref(model.m.scaling)))
    try:
        mdz.check(hot.mdz_mlp_generate(h, 1 << 20, 17))
        mdz.check(hot.mdz_mlp_init(h, 2))
        loss = (C.c_double * 4)()
        mdz.check(hot.mdz_mlp_train(h, 4, 16384, 1e-3, 1, loss))
        w = np.zeros(mo.NPARAMS, np.float32)
        mdz.check(hot.mdz_mlp_get_weights(h, w.ctypes.data_as(C.POINTER(C.c_float)), mo.NPARAMS))
        return w
    finally:
        hot.mdz_mlp_destroy(h)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["fp64", "mixed", "learned"])
def test_per_model_residuals_against_oracle(mdz, orc, hot, ens, mode):
    """Per-model residual error of the ensemble graph against the oracle's
    Newton unit on the same model."""
    ref = []
    for p in MODELS:
        m = mdz.Model.lithosphere(*p)
        S = numbering(orc, mdz, m)
        bo = newton_unit(orc, mdz, m, S)
        ref.append([m.m.Nmodel.resx, m.m.Nmodel.resz, m.m.Nmodel.resp])
        free_blocks(orc, bo)
        m.free()
    ref = np.array(ref)
    if mode == 2:
        w = _trained_network(mdz, hot, ens.models[0])
        for c in ens.ctx:
            mdz.check(hot.mdz_ctx_set_network(c, w.ctypes.data_as(C.POINTER(C.c_float)), mo.NPARAMS))
    for c in ens.ctx:
        mdz.check(hot.mdz_ctx_set_rheology_mode(c, mode))
    ens.create()
    mdz.check(hot.mdz_ensemble_autotune(ens.ens, 1, None, None))
    mdz.check(hot.mdz_ensemble_launch(ens.ens))
    res = ens.residuals()
    err = np.abs(res - ref) / np.abs(ref)
    for i, p in enumerate(MODELS):
        print("model %d (%dx%d): residual rel. error x %.1e z %.1e p %.1e" % ((i, p[0], p[1]) + tuple(err[i])))
    assert err.max() <= (1e-12 if mode == 0 else 1e-10)


def test_setting_change_recaptures(mdz, hot, ens):
    """A context setting changed after the capture (here: histogram
    collection switched on for one model) takes effect at the next launch."""
    ens.create()
    mdz.check(hot.mdz_ensemble_set_branches(ens.ens, 2))
    mdz.check(hot.mdz_ensemble_launch(ens.ens))
    mdz.check(hot.mdz_ensemble_synchronize(ens.ens))
    h = (C.c_longlong * 192)()
    mdz.check(hot.mdz_ctx_iteration_histogram(ens.ctx[1], h))
    assert sum(h) == 0
    mdz.check(hot.mdz_ctx_collect_histogram(ens.ctx[1], 1))
    mdz.check(hot.mdz_ensemble_launch(ens.ens))
    mdz.check(hot.mdz_ensemble_synchronize(ens.ens))
    mdz.check(hot.mdz_ctx_iteration_histogram(ens.ctx[1], h))
    assert sum(h[0:64]) > 0
    mdz.check(hot.mdz_ctx_iteration_histogram(ens.ctx[0], h))
    assert sum(h) == 0


def test_ensemble_argument_errors(mdz, hot, ens):
    e = C.c_void_p()
    assert hot.mdz_ensemble_create(C.byref(e), None, 0) != 0
    arr = (C.c_void_p * 2)(ens.ctx[0].value, ens.ctx[0].value)
    assert hot.mdz_ensemble_create(C.byref(e), arr, 2) != 0  # repeated context
    assert hot.mdz_ctx_set_launch_shape(ens.ctx[0], 96, 64, 64) != 0
    assert hot.mdz_ctx_set_launch_shape(ens.ctx[0], 128, 64, 256) != 0
    mdz.check(hot.mdz_ctx_set_launch_shape(ens.ctx[0], 64, 0, 0))
    sh = (C.c_int * 3)()
    mdz.check(hot.mdz_ctx_get_launch_shape(ens.ctx[0], sh))
    assert list(sh) == [64, 0, 0]
    ens.create()
    assert hot.mdz_ensemble_set_branches(ens.ens, 9) != 0
    assert hot.mdz_ensemble_set_branches(ens.ens, 0) != 0


@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_every_launch_shape_gives_identical_results(mdz, hot, ens, mode):
    """Each candidate shape of the autotuner, one stage at a time, against the
    default shape: norms and CSR values bit for bit."""
    c = ens.ctx[3]
    mdz.check(hot.mdz_ctx_set_rheology_mode(c, mode))
    r0 = (C.c_double * 3)()
    mdz.check(hot.mdz_ctx_newton_unit(c, r0))
    ref = [ens.block_values(3, w) for w in (0, 1, 4, 5)]
    shapes = [(r, 256, 128) for r in (64, 128)] + [(256, p, 128) for p in (0, 64, 128)] + \
             [(256, 256, j) for j in (0, 64)]
    for sh in shapes:
        mdz.check(hot.mdz_ctx_set_launch_shape(c, *sh))
        r = (C.c_double * 3)()
        mdz.check(hot.mdz_ctx_newton_unit(c, r))
        assert list(r) == list(r0), sh
        for (A, b), (Ar, br_) in zip([ens.block_values(3, w) for w in (0, 1, 4, 5)], ref):
            assert np.array_equal(A, Ar) and np.array_equal(b, br_), sh


def test_deferred_diagnostics_in_the_graph(mdz, hot, ens):
    """Contexts with deferred diagnostics in a captured ensemble: the replayed
    units give the eager full-output norms, and the grid downloaded after a
    replay (diagnostics recomputed then) equals the one after an eager
    full-output unit, field for field."""
    names = mdz.field_names(2)
    ref_res, ref_grid = [], []
    for i, c in enumerate(ens.ctx):
        r = (C.c_double * 3)()
        mdz.check(hot.mdz_ctx_newton_unit(c, r))
        ref_res.append(list(r))
        m = ens.models[i].clone()
        mdz.check(hot.mdz_ctx_download_grid(c, m.m.mesh))
        ref_grid.append({n: m.field(n).copy() for n in names})
        m.free()
    for c in ens.ctx:
        mdz.check(hot.mdz_ctx_defer_diagnostics(c, 1))
    ens.create()
    for replay in range(2):
        mdz.check(hot.mdz_ensemble_launch(ens.ens))
        mdz.check(hot.mdz_ensemble_synchronize(ens.ens))
        assert np.array_equal(ens.residuals(), np.array(ref_res)), replay
    for i, c in enumerate(ens.ctx):
        m = ens.models[i].clone()
        mdz.check(hot.mdz_ctx_download_grid(c, m.m.mesh))
        for n in names:
            assert np.array_equal(m.field(n), ref_grid[i][n], equal_nan=True), (i, n)
        m.free()
