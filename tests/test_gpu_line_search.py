"""GPU tests of the line-search batch (LineSearchDecoupled,
StokesRoutines.c:396-622, as one graph launch): the norms of every trial equal
those of an eager unit at the same trial iterate bit for bit, and the
oracle's within 1e-12; the iterate is restored; a second launch with other
step lengths reuses the graph and still matches."""
import ctypes as C

import numpy as np
import pytest

import newton_driver as nd

pytestmark = pytest.mark.gpu
DP = C.POINTER(C.c_double)


def _trial_iterate(it, st, dx, alpha):
    it.set_state(st)
    it.add(dx, alpha)
    return it.state()


@pytest.mark.parametrize("model", ["pwl", "litho"])
def test_line_search_batch(mdz, orc, hot, model):
    m = mdz.Model.pwl_inclusion(100, 100) if model == "pwl" else mdz.Model.lithosphere(96, 48, 3)
    mo = m.clone()
    gpu = None
    try:
        it = nd.Iterate(mdz, orc, m)
        gpu = nd.GpuAssembler(mdz, hot, m, it)
        csr4, F, res = gpu()
        K = nd.coupled(it.neq_mom, it.neq_cont, csr4)
        import scipy.sparse.linalg as spla
        dx = spla.spsolve(K, -F)
        st = it.state()
        for alphas in ([0.2, 0.4, 0.6, 0.8, 1.0], [1.0, 0.5, 0.25]):
            a = np.array(alphas)
            out = np.zeros((len(a), 3))
            mdz.check(hot.mdz_ctx_line_search(gpu.ctx, dx.ctypes.data_as(DP), len(a), a.ctypes.data_as(DP),
                                              out.ctypes.data_as(DP)))
            # the iterate is restored
            u, v, p = np.zeros_like(it.u), np.zeros_like(it.v), np.zeros_like(it.p)
            mdz.check(hot.mdz_ctx_download_solution(gpu.ctx, u.ctypes.data_as(DP), v.ctypes.data_as(DP),
                                                    p.ctypes.data_as(DP)))
            assert np.array_equal(u, st[0]) and np.array_equal(v, st[1]) and np.array_equal(p, st[2])
            ito = nd.Iterate(mdz, orc, mo)
            for k, alpha in enumerate(alphas):
                _trial_iterate(it, st, dx, alpha)
                _, _, r_eager = gpu()
                assert tuple(out[k]) == r_eager, (alpha, out[k], r_eager)
                ito.set_state(_trial_iterate(it, st, dx, alpha))
                _, _, r_orc = nd.OracleAssembler(mdz, orc, mo, ito)()
                assert np.allclose(out[k], r_orc, rtol=1e-12, atol=0), (alpha, out[k], r_orc)
            it.set_state(st)
            gpu()  # back to the unit at the saved iterate
            print(model, "alphas", alphas, "resx", ["%.3e" % x for x in out[:, 0]])
    finally:
        if gpu:
            gpu.close()
        m.free()
        mo.free()


def test_line_search_argument_errors(mdz, hot):
    m = mdz.Model.lithosphere(32, 16, 1)
    ctx = C.c_void_p()
    mdz.check(hot.mdz_ctx_create(C.byref(ctx), m.m.mesh, C.byref(m.m.materials), C.byref(m.m.model),
                                 C.byref(m.m.scaling)))
    try:
        dx = np.zeros(10)
        a = np.zeros(17)
        out = np.zeros(51)
        assert hot.mdz_ctx_line_search(ctx, dx.ctypes.data_as(DP), 17, a.ctypes.data_as(DP),
                                       out.ctypes.data_as(DP)) != 0
        assert hot.mdz_ctx_line_search(ctx, dx.ctypes.data_as(DP), 0, a.ctypes.data_as(DP),
                                       out.ctypes.data_as(DP)) != 0
    finally:
        hot.mdz_ctx_destroy(ctx)
        m.free()
