"""Outer Newton iterations with the GPU path assembling the Jacobian and the
residual, against the same driver on the oracle, on the reference's
power-law pure-shear benchmark (PureShear_pwl, 101 x 101 nodes): iteration
counts equal within +-1 (and both within +-1 of the 6 the reference
documents, Documentation/MDOODZ_docu.tex:293); converged iterates equal to
1e-6 relative (the stop rule's tolerance)."""
import numpy as np
import pytest

import newton_driver as nd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_gpu_newton_count_matches_oracle(mdz, orc, hot, mode):
    mo, mg = mdz.Model.pwl_inclusion(100, 100), mdz.Model.pwl_inclusion(100, 100)
    gpu = None
    try:
        ito = nd.Iterate(mdz, orc, mo)
        no, ho = nd.newton(ito, nd.OracleAssembler(mdz, orc, mo, ito))
        itg = nd.Iterate(mdz, orc, mg)
        gpu = nd.GpuAssembler(mdz, hot, mg, itg, mode)
        ng, hg = nd.newton(itg, gpu)
        print("Newton iterations: oracle %d, GPU %d" % (no, ng))
        print("oracle %s" % ["%.2e" % max(r) for r in ho])
        print("GPU    %s" % ["%.2e" % max(r) for r in hg])
        assert abs(ng - no) <= 1
        assert abs(ng - 6) <= 1
        for a, b in ((itg.u, ito.u), (itg.v, ito.v)):
            assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
    finally:
        if gpu:
            gpu.close()
        mo.free()
        mg.free()
