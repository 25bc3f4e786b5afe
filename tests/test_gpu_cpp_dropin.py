"""The C++ drop-in classes (include/nadia/*.h, libNADIA[d]_amd.so) driven by
an example program written like NADIA's examples/PlanarCDI_example.c,
checked against the CPU oracle running the same loop."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "nadia-coecxs_amd", "cpp")


def run_example(binary, tmp_path, d, er1, hio, er2, sw):
    n = d["x0"].shape[0]
    d["intensity"].astype(np.float64).tofile(tmp_path / "intensity.raw")
    d["support"].astype(np.float64).tofile(tmp_path / "support.raw")
    d["x0"].astype(np.complex128).tofile(tmp_path / "x0.raw")
    out = subprocess.run([os.path.join(CPP, binary), str(tmp_path), str(n), str(n), str(er1), str(hio), str(er2),
                          str(sw)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    errs = [float(l.split()[-1]) for l in out.stdout.splitlines() if l.startswith("iteration ")]
    best = [float(l.split()[-1]) for l in out.stdout.splitlines() if l.startswith("best error")][0]
    x = np.fromfile(tmp_path / "out.raw", dtype=np.complex128).reshape(n, n)
    return np.array(errs), x, best


def oracle_loop(oracle, d, er1, hio, er2, sw):
    x = d["x0"].astype(np.complex128).copy()
    o = oracle.OracleCDI(x, oracle.PLANAR, n_best=4)
    o.set_support(d["support"])
    o.set_intensity(d["intensity"])
    errs = []
    it = 0
    for alg, cnt in ((oracle.ER, er1), (oracle.HIO, hio), (oracle.ER, er2)):
        o.set_algorithm(alg)
        for _ in range(cnt):
            o.iterate()
            errs.append(o.get_error())
            if sw > 0 and it % sw == sw - 1:
                o.apply_shrinkwrap(2.0, 0.1)
            it += 1
    return np.array(errs), x, o.get_best_result(0)[1]


def test_cpp_planar_example_fp64(oracle, synth, tmp_path):
    d = synth.planar_block(n=128, block=48, support=56)
    eg, xg, bg = run_example("planar_exampled", tmp_path, d, 20, 20, 20, 25)
    eo, xo, bo = oracle_loop(oracle, d, 20, 20, 20, 25)
    assert len(eg) == 60
    assert (np.abs(eg - eo) / eo).max() <= 1e-9
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-9
    assert abs(bg - bo) <= 1e-9 * bo


def test_cpp_planar_example_fp32(oracle, synth, tmp_path):
    d = synth.planar_block(n=128, block=48, support=56)
    eg, xg, bg = run_example("planar_example", tmp_path, d, 20, 0, 0, 0)
    eo, xo, bo = oracle_loop(oracle, d, 20, 0, 0, 0)
    assert (np.abs(eg - eo) / eo).max() <= 1e-4
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-4


def test_cpp_phase_diverse_with_constraint(oracle, synth, tmp_path):
    """C++ drop-in PhaseDiverseCDI + TransmissionConstraint + .dbin/.cplx
    files (multi_example) against the oracle running the same set-up."""
    import importlib
    files = importlib.import_module("nadia-coecxs_amd.files")
    n, x1, y1, iters = 64, 10.0, -6.0, 3
    frames = []
    for k in range(2):
        d = synth.planar_block(n=n, block=24, support=30, seed_amp=30 + k, seed_poisson=30 + k, seed_phase=40 + k)
        files.write_dbin(tmp_path / ("intensity_%d.dbin" % k), d["intensity"])
        files.write_dbin(tmp_path / ("support_%d.dbin" % k), d["support"])
        files.write_cplx(tmp_path / ("x0_%d.cplx" % k), d["x0"])
        frames.append(d)
    region = np.zeros((n, n))
    region[n // 2 - 12:n // 2 + 12, n // 2 - 12:n // 2 + 12] = 1
    files.write_dbin(tmp_path / "region.dbin", region)
    out = subprocess.run([os.path.join(CPP, "multi_exampled"), str(tmp_path), str(n), str(x1), str(y1), str(iters)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Added position 1" in out.stdout and "Error for frame 1 is" in out.stdout
    c_gpu = [float(l.split()[1]) for l in out.stdout.splitlines() if l.startswith("c_mean")][0]

    xs, objs = [], []
    for d in frames:
        x = d["x0"].astype(np.complex128).copy()
        o = oracle.OracleCDI(x, oracle.PLANAR)
        o.set_support(d["support"])
        o.set_intensity(d["intensity"])
        o.set_algorithm(oracle.HIO)
        xs.append(x)
        objs.append(o)
    tc = oracle.OracleConstraint(n, n)
    k = tc.add_complex_constraint(region, 0.5, 0.5)
    tc.set_enforce_unity(False)
    objs[0].set_complex_constraint(tc)
    pd = oracle.OraclePhaseDiverse(beta=0.8, gamma=1.0, parallel=True)
    pd.add_new_position(objs[0], 0, 0)
    pd.add_new_position(objs[1], x1, y1)
    pd.initialise_estimate()
    for _ in range(iters):
        pd.iterate()
    t = pd.get_transmission()
    nx, ny = t.shape
    got = files.read_cplx(tmp_path / "object.cplx", nx, ny)
    assert np.linalg.norm(got - t) / np.linalg.norm(t) <= 1e-10
    f0 = files.read_cplx(tmp_path / "frame0.cplx", n, n)
    assert np.linalg.norm(f0 - xs[0]) / np.linalg.norm(xs[0]) <= 1e-10
    assert abs(c_gpu - tc.c_mean(k)) <= 1e-10 * abs(tc.c_mean(k))


@pytest.mark.parametrize("kind,forward", [(0, 1), (1, 0)], ids=["cross_correlation", "minimum_error_backward"])
def test_cpp_adjust_positions(oracle, tmp_path, kind, forward):
    """PhaseDiverseCDI::adjust_positions through the C++ drop-in (align_example)
    against the oracle and the known frame offsets."""
    import importlib
    files = importlib.import_module("nadia-coecxs_amd.files")
    synth = importlib.import_module("nadia-coecxs_amd.synth")
    d = synth.ptycho_frames()
    truth = d["positions"]
    ref = 0 if forward else len(truth) - 1
    starts = [(x + (0 if k == ref else 2.0 * (-1) ** k), y + (0 if k == ref else -3.0 + k))
              for k, (x, y) in enumerate(truth)]
    m = d["frames"][0].shape[0]
    for k, (f, inten, sup) in enumerate(zip(d["frames"], d["intensities"], d["supports"])):
        files.write_cplx(tmp_path / ("x0_%d.cplx" % k), f)
        files.write_dbin(tmp_path / ("intensity_%d.dbin" % k), inten)
        files.write_dbin(tmp_path / ("support_%d.dbin" % k), sup)
    args = [str(v) for xy in starts for v in xy]
    out = subprocess.run([os.path.join(CPP, "align_exampled"), str(tmp_path), str(m), str(kind), str(forward)] + args,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = {}
    for line in out.stdout.splitlines():
        if line.startswith("position "):
            _, k, x, y = line.split()
            got[int(k)] = (float(x), float(y))
    assert out.stdout.count("moving prob") == len(truth) - 1
    objs = []
    pd = oracle.OraclePhaseDiverse(beta=1.0)
    for f, inten, sup, (x0, y0) in zip(d["frames"], d["intensities"], d["supports"], starts):
        x = f.astype(np.complex128).copy()
        o = oracle.OracleCDI(x, oracle.PLANAR)
        o.set_support(sup)
        o.set_intensity(inten)
        o.set_algorithm(oracle.ER)
        objs.append((x, o))
        pd.add_new_position(o, x0, y0)
    pd.initialise_estimate()
    pd.adjust_positions(kind, bool(forward))
    for k in range(len(truth)):
        assert got[k] == pd.get_position(k) == truth[k]
