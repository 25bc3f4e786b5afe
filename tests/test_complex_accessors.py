"""Element accessors of the drop-in Complex_2D (include/nadia/Complex_2D.h)
against NADIA's semantics (NADIA/include/Complex_2D.h:167-305), in both
precisions, through a small C++ program built with g++ (host code; no device
call in this mode):

* set_mag on a zero element gives (value, 0); otherwise it scales the
  element, keeping the phase (:205-215);
* set_phase keeps the magnitude (:225-229);
* get_phase is atan2 with values above M_PI wrapped by -2 pi (:282-290) --
  which turns the single-precision atan2 of a negative real number (pi
  rounded up in float) into about -pi, while double precision returns +pi;
* get_value(MAG_SQ) is pow(get_mag, 2)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nadia-coecxs_amd")


def build_probe(out_dir, double):
    exe = os.path.join(str(out_dir), "complex_probe" + ("d" if double else ""))
    cmd = ["g++", "-O1", "-std=c++11", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "include", "nadia"), os.path.join(ROOT, "tests", "cpp", "complex_probe.cpp"),
           "-o", exe, "-L", PKG, "-lnadia_hip", "-Wl,-rpath," + PKG]
    if double:
        cmd.insert(1, "-DDOUBLE_PRECISION")
    subprocess.run(cmd, check=True)
    return exe


def run_probe(exe, *args):
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = {}
    for line in out.stdout.splitlines():
        f = line.split()
        rows.setdefault(f[0], []).append([float(v) for v in f[1:]])
    return {k: np.array(v) for k, v in rows.items()}


def expected_elements():
    """(re, im) after the probe's accessor calls, computed in double."""
    e = np.zeros((4, 6), complex)
    e[0, 0] = 2.5
    e[1, 2] = (3 - 4j) * (10.0 / 5.0)
    e[2, 3] = np.sqrt(2.0) * np.exp(0.75j)
    e[3, 5] = -2.0
    e[3, 4] = 1.5 * np.exp(-2.0j)
    return e


@pytest.fixture(scope="module", params=[False, True], ids=["fp32", "fp64"])
def probe(request, tmp_path_factory):
    if not os.path.exists(os.path.join(PKG, "libnadia_hip.so")):
        pytest.skip("libnadia_hip.so not built")
    return request.param, build_probe(tmp_path_factory.mktemp("cprobe"), request.param)


def test_host_accessors(probe):
    double, exe = probe
    tol = 1e-15 if double else 1e-6
    r = run_probe(exe)["E"]
    e = expected_elements()
    for i, j, re, im, mag, ph, msq in r:
        i, j = int(i), int(j)
        want = e[i, j]
        assert abs(re - want.real) <= tol * 10 and abs(im - want.imag) <= tol * 10, (i, j)
        assert abs(mag - abs(want)) <= tol * 10, (i, j)
        assert abs(msq - mag * mag) <= tol * 100, (i, j)
        if (i, j) == (3, 5):
            # atan2 of a negative real: +pi in double; float pi rounds above
            # M_PI and the wrap gives about -pi
            assert ph == (np.pi if double else float(np.float32(float(np.float32(np.pi)) - 2 * np.pi)))
        elif want != 0:
            assert abs(ph - np.angle(want)) <= tol * 10, (i, j)
        else:
            assert ph == 0.0
    # the |c| = 0 branch of set_mag: exactly (value, 0)
    z = r[(r[:, 0] == 0) & (r[:, 1] == 0)][0]
    assert z[2] == 2.5 and z[3] == 0.0
