"""The reference's own example input (NADIA/examples/image_files/planar_data.tif,
a measured 1024^2 diffraction pattern, kept as tests/golden/planar_data.npz)
reconstructed with the schedule of NADIA/examples/PlanarCDI_example.c
(initialise_estimate(0), ER with shrinkwrap, relaxation 0.5, HIO, 4 best
estimates kept) on the device (FP64) and by the oracle: per-iteration errors,
the estimate, the shrinkwrapped support and the best estimates agree."""
import importlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

synth = importlib.import_module("nadia-coecxs_amd.synth")
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_reference_example_input_fp64(nd, oracle):
    d = synth.reference_planar_example(os.path.join(HERE, "golden", "planar_data.npz"))
    n = d["intensity"].shape[0]
    est = nd.Complex2D(n, n)
    g = nd.PlanarCDI(est, n_best=4)
    xo = np.zeros((n, n), np.complex128)
    o = oracle.OracleCDI(xo, oracle.PLANAR, n_best=4)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.initialise_estimate(0)
    assert rel(est.get(), xo) <= 1e-15
    # the first cycle of the example, HIO and the second ER shortened
    sched = dict(synth.REFERENCE_EXAMPLE_SCHEDULE, hio=12, er2=8)
    algs = {"ER": (nd.ER, oracle.ER), "HIO": (nd.HIO, oracle.HIO)}
    eg = synth.run_reference_example(
        g, alg=lambda a: g.set_algorithm(algs[a][0]), iterate=lambda: float(g.iterate(1)[0]),
        shrinkwrap=g.apply_shrinkwrap, set_beta=g.set_relaxation_parameter, cycles=1, sched=sched)

    def o_iter():
        o.iterate()
        return o.get_error()
    eo = synth.run_reference_example(
        o, alg=lambda a: o.set_algorithm(algs[a][1]), iterate=o_iter, shrinkwrap=o.apply_shrinkwrap,
        set_beta=o.set_relaxation_parameter, cycles=1, sched=sched)
    eg, eo = np.array(eg), np.array(eo)
    assert len(eg) == 70
    assert (np.abs(eg - eo) / eo).max() <= 1e-9
    assert rel(est.get(), xo) <= 1e-9
    assert np.array_equal(g.get_support(), o.get_support())
    for k in range(4):
        bg, ebg = g.get_best_result(k)
        bo, ebo = o.get_best_result(k)
        assert abs(ebg - ebo) <= 1e-9 * ebo
        assert rel(bg, bo) <= 1e-9
