"""GPU test of the network evaluated inside the log-space rheology kernels
(MDZ_RHEO_LEARNED with mdz_ctx_fuse_guess) against the separate inference
pass (the default): both compute the same features and the same
forward pass, so the guess, the local iterations and everything after them
are the same bit for bit -- residual norms, the eight CSR blocks, every
downloaded grid field and the iteration histograms -- for each rheology
block size, with and without deferred diagnostics, over two units."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as mo
from test_gpu_deferred_diagnostics import _blocks, _network

pytestmark = pytest.mark.gpu
FP = C.POINTER(C.c_float)


@pytest.mark.parametrize("defer", [0, 1], ids=["full", "deferred"])
@pytest.mark.parametrize("rheo_block", [64, 128, 256])
def test_fused_guess_equals_separate_pass(mdz, hot, rheo_block, defer):
    # 97 x 47 cells: neither node count is a multiple of 64, so the last
    # wave of each kernel has lanes past the grid in the matrix instructions
    base = mdz.Model.lithosphere(97, 47, 9)
    w = _network(mdz, hot, base)
    ms = [base.clone(), base.clone()]
    ctx = [C.c_void_p(), C.c_void_p()]
    try:
        for c, m, fused in zip(ctx, ms, (1, 0)):
            mdz.check(hot.mdz_ctx_create(C.byref(c), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                         C.byref(m.m.scaling)))
            mdz.check(hot.mdz_ctx_set_network(c, w.ctypes.data_as(FP), mo.NPARAMS))
            mdz.check(hot.mdz_ctx_set_rheology_mode(c, 2))
            mdz.check(hot.mdz_ctx_fuse_guess(c, fused))
            mdz.check(hot.mdz_ctx_defer_diagnostics(c, defer))
            mdz.check(hot.mdz_ctx_collect_histogram(c, 1))
            shape = (C.c_int * 3)()
            mdz.check(hot.mdz_ctx_get_launch_shape(c, shape))
            mdz.check(hot.mdz_ctx_set_launch_shape(c, rheo_block, shape[1], shape[2]))
        res, hists = [], []
        for c in ctx:
            r = (C.c_double * 3)()
            mdz.check(hot.mdz_ctx_newton_unit(c, r))
            mdz.check(hot.mdz_ctx_newton_unit(c, r))
            res.append(list(r))
            h = (C.c_longlong * 192)()
            mdz.check(hot.mdz_ctx_iteration_histogram(c, h))
            hists.append(list(h))
        print("residual norms", res[0])
        assert res[0] == res[1]
        assert hists[0] == hists[1]
        # the FP32 iterations ran from the network's guess at every node
        assert sum(hists[0][128:192]) > 0
        for (A0, b0), (A1, b1) in zip(_blocks(mdz, hot, ctx[0]), _blocks(mdz, hot, ctx[1])):
            assert np.array_equal(A0, A1) and np.array_equal(b0, b1)
        for c, m in zip(ctx, ms):
            mdz.check(hot.mdz_ctx_download_grid(c, m.m.mesh))
        for name in mdz.field_names(2):
            a, b = ms[0].field(name), ms[1].field(name)
            assert np.array_equal(a, b, equal_nan=True), name
    finally:
        for c in ctx:
            if c:
                hot.mdz_ctx_destroy(c)
        for m in ms:
            m.free()
        base.free()
