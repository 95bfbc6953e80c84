"""Outer Newton iterations with the oracle assembling the path, on the
reference's power-law pure-shear benchmark (BENCHMARKS_old_input/
PureShear_pwl.txt at its 101 x 101 nodes): the reference documents that its
Newton solver converges in 6 nonlinear iterations
(Documentation/MDOODZ_docu.tex:293).  The driver (tests/newton_driver.py)
takes the reference's line search and stop rule, with scipy's sparse LU in
place of CHOLMOD; the count is held to the documented one within +-1."""
import numpy as np

import newton_driver as nd


def test_oracle_newton_count_on_powerlaw_benchmark(mdz, orc):
    m = mdz.Model.pwl_inclusion(100, 100)
    try:
        it = nd.Iterate(mdz, orc, m)
        n, hist = nd.newton(it, nd.OracleAssembler(mdz, orc, m, it))
        print("Newton iterations %d, max residual per iteration %s" % (n, ["%.2e" % max(r) for r in hist]))
        assert abs(n - 6) <= 1
        assert nd.converged(hist[-1], hist[0], 1e-6, 1e-6)
        # fast (Newton) convergence over the last steps
        r = [max(x) for x in hist]
        assert r[-1] < 0.1 * r[-2] and r[-2] < 0.1 * r[-3]
        assert np.isfinite(it.u).all() and np.isfinite(it.v).all() and np.isfinite(it.p).all()
    finally:
        m.free()
