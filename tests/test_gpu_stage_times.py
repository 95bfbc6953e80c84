"""GPU tests of the per-stage timing of the device-resident unit: it runs the
unit's kernels (the residual vectors afterwards equal those of
mdz_ctx_newton_unit bit for bit), an isotropic model reports its fused D pass
as the J = 1 stage and nothing in the two J = 0 stages, and the stages add up
to the whole unit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DP = C.POINTER(C.c_double)


def test_stage_times_of_an_isotropic_model(mdz, hot):
    m = mdz.Model.lithosphere(256, 128, 3)
    ctx = C.c_void_p()
    mdz.check(hot.mdz_ctx_create(C.byref(ctx), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                 C.byref(m.m.scaling)))
    try:
        neq_mom, neq_cont = C.c_int(), C.c_int()
        mdz.check(hot.mdz_ctx_block_info(ctx, 0, C.byref(neq_mom), None))
        mdz.check(hot.mdz_ctx_block_info(ctx, 3, C.byref(neq_cont), None))

        def residual():
            fm, fc = np.zeros(neq_mom.value), np.zeros(neq_cont.value)
            mdz.check(hot.mdz_ctx_download_residual(ctx, fm.ctypes.data_as(DP), fc.ctypes.data_as(DP)))
            return fm, fc

        mdz.check(hot.mdz_ctx_newton_unit(ctx, None))
        f0 = residual()
        ms = (C.c_double * 11)()
        mdz.check(hot.mdz_ctx_stage_times(ctx, 3, ms))
        print("stage ms", [round(x, 4) for x in ms])
        # events recorded back to back: no work between them
        assert ms[4] < 0.01 and ms[8] < 0.01
        for i in (0, 1, 2, 3, 5, 6, 7, 9):
            assert ms[i] > 0.0, i
        assert sum(ms[:10]) <= ms[10] * 1.001 + 1e-6
        assert sum(ms[:10]) >= 0.5 * ms[10]
        # the timed units computed what a unit computes
        f1 = residual()
        assert np.array_equal(f0[0], f1[0]) and np.array_equal(f0[1], f1[1])
    finally:
        hot.mdz_ctx_destroy(ctx)
        m.free()


def test_stage_times_argument_errors(mdz, hot):
    assert hot.mdz_ctx_stage_times(None, 1, (C.c_double * 11)()) != 0
    m = mdz.Model.lithosphere(32, 16, 1)
    ctx = C.c_void_p()
    mdz.check(hot.mdz_ctx_create(C.byref(ctx), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                 C.byref(m.m.scaling)))
    try:
        assert hot.mdz_ctx_stage_times(ctx, 0, (C.c_double * 11)()) != 0
        assert hot.mdz_ctx_stage_times(ctx, 1, None) != 0
    finally:
        hot.mdz_ctx_destroy(ctx)
        m.free()
