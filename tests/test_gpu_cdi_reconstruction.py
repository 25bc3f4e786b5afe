"""The configuration-file reconstruction tool (nadia-coecxs_amd/cpp/
cdi_reconstruction.cpp, the counterpart of NADIA/tools/CDI_reconstruction.c)
run on .dbin inputs and checked against the oracle running the same
schedule: planar (HIO then ER with shrinkwrap, intermediate magnitude and
diffraction images) and the two-step Fresnel pipeline (fresnel_wf, then
fresnel with the reconstructed white field, transmission function out)."""
import importlib
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "nadia-coecxs_amd", "cpp")
files = importlib.import_module("nadia-coecxs_amd.files")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_tool(binary, cwd, cfg_text, reco_type, seed):
    cfg = cwd / ("%s.config" % reco_type)
    cfg.write_text(cfg_text)
    out = subprocess.run([os.path.join(CPP, binary), str(cfg), reco_type, str(seed)], cwd=str(cwd),
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def printed_errors(stdout):
    return {int(m.group(1)): float(m.group(2))
            for m in re.finditer(r"^Error for iteration (\d+) is (\S+)$", stdout, re.M)}


def oracle_schedule(o, x, schedule, sw, gauss, thr, out_every=0):
    """The tool's loop (NADIA/tools/CDI_reconstruction.c:383-441) on the oracle."""
    errs, images, diffs, i = {}, {}, {}, 0
    for alg, n in schedule:
        o.set_algorithm(alg)
        for _ in range(n):
            o.iterate()
            errs[i] = o.get_error()
            if out_every and i % out_every == 0:
                images[i] = np.abs(x)
                t = x.copy()
                o.propagate_to_detector(t)
                diffs[i] = np.abs(t) ** 2
            if sw and i % sw == sw - 1:
                o.apply_shrinkwrap(gauss, thr)
            i += 1
    return errs, images, diffs


PLANAR_CFG = """# planar reconstruction
data_file_name = {d}/intensity.dbin
pixels_x = {n}
pixels_y = {n}
support_file_name = {d}/support.dbin
output_file_name_prefix = {d}/planar
output_file_type = dbin
output_iterations = 10
algorithm = {algs}
iterations = {iters}
info_level = 2
output_diffraction_estimate = 1
use_log_scale_for_diffraction = 0
use_log_scale_for_object = 0
shrinkwrap_iterations = {sw}
shrinkwrap_gauss_width = 1.5
shrinkwrap_threshold = 0.1
"""


@pytest.mark.parametrize("binary,double,algs,iters,sw,tol,n", [
    ("cdi_reconstructiond", True, "HIO ER", "20 10", 15, 1e-9, 128),
    ("cdi_reconstruction", False, "ER", "20", 0, 1e-4, 128),
    # sizes without compiled passes: 120 = 8*3*5 (direct plan), 106 = 2*53 (Bluestein)
    ("cdi_reconstructiond", True, "HIO ER", "20 10", 15, 1e-9, 120),
    ("cdi_reconstructiond", True, "HIO ER", "20 10", 15, 1e-9, 106),
], ids=["fp64", "fp32", "fp64-120", "fp64-106"])
def test_planar_config_run(oracle, synth, tmp_path, binary, double, algs, iters, sw, tol, n):
    seed = 3
    d = synth.planar_block(n=n, block=48, support=56)
    files.write_dbin(tmp_path / "intensity.dbin", d["intensity"])
    files.write_dbin(tmp_path / "support.dbin", d["support"])
    stdout = run_tool(binary, tmp_path, PLANAR_CFG.format(d=tmp_path, n=n, algs=algs, iters=iters, sw=sw),
                      "planar", seed)
    assert "Switching to the HIO algorithm" in stdout or algs == "ER"

    x = np.zeros((n, n), np.complex128)
    o = oracle.OracleCDI(x, oracle.PLANAR)
    o.set_support(d["support"])
    o.set_intensity(d["intensity"])
    o.initialise_estimate(seed)
    schedule = [(getattr(oracle, a), int(k)) for a, k in zip(algs.split(), iters.split())]
    errs, images, diffs = oracle_schedule(o, x, schedule, sw, 1.5, 0.1, out_every=10)

    got = files.read_cplx(tmp_path / "planar.cplx", n, n)
    assert rel(got, x) <= tol
    total = sum(k for _, k in schedule)
    if sw:
        assert stdout.count("Applying shrink-wrap at iteration") == total // sw
    # "Error for iteration i-1" is printed after iteration i (6 significant digits)
    pe = printed_errors(stdout)
    assert sorted(pe) == list(range(total - 1))
    for i, e in pe.items():
        assert abs(e - errs[i + 1]) <= max(2e-5, tol) * errs[i + 1] + 1e-12
    for i in images:
        img = files.read_dbin(tmp_path / ("planar_%d.dbin" % i), n, n)
        assert rel(img, images[i]) <= tol
        dif = files.read_dbin(tmp_path / ("planar_diffraction_%d.dbin" % i), n, n)
        assert rel(dif, diffs[i]) <= tol


FRESNEL_CFG = """# Fresnel white field, then object
beam_wavelength = {wavelength!r}
zone_focal_length = {z12!r}
focal_detector_length = {zfd!r}
focal_sample_length = {zfs!r}
pixel_size = {pixel!r}
normalisation = 1.0
pixels_x = {n}
pixels_y = {n}
white_field_reco_file_name = {d}/wf.cplx
white_field_data_file_name = {d}/wf_intensity.dbin
white_field_support_file_name = {d}/wf_support.dbin
wf_iterations = 20
data_file_name = {d}/intensity.dbin
support_file_name = {d}/support.dbin
output_file_name_prefix = {d}/fresnel
output_file_type = dbin
output_iterations = 100
algorithm = ER HIO ER
iterations = 8 8 8
info_level = 1
output_diffraction_estimate = 0
shrinkwrap_iterations = 12
shrinkwrap_gauss_width = 1.5
shrinkwrap_threshold = 0.1
"""


def test_fresnel_pipeline_config_run(oracle, synth, tmp_path):
    n, seed = 128, 2
    d = synth.fresnel_config(n=n, pupil_diameter=160e-6 * n / 1024, disk_radius=20 * n / 1024 + 2,
                             ndisks=4, support=500)
    # white-field support: the softened pupil disk of FresnelCDI_WF::set_support
    xw = np.zeros((n, n), np.complex128)
    ow = oracle.OracleCDI(xw, oracle.FRESNEL_WF)
    ow.wf_init(d["wavelength"], d["z12"], d["z23"], d["pixel"])
    ow.wf_set_support(d["pupil_diameter"])
    wf_support = ow.get_support()
    files.write_dbin(tmp_path / "wf_intensity.dbin", d["white_field_intensity"])
    files.write_dbin(tmp_path / "wf_support.dbin", wf_support)
    files.write_dbin(tmp_path / "intensity.dbin", d["intensity"])
    files.write_dbin(tmp_path / "support.dbin", d["support"])
    cfg = FRESNEL_CFG.format(d=tmp_path, n=n, **{k: d[k] for k in ("wavelength", "z12", "zfd", "zfs", "pixel")})

    # step 1: white field (the tool sets the file's support through BaseCDI)
    run_tool("cdi_reconstructiond", tmp_path, cfg, "fresnel_wf", seed)
    ow.set_support(wf_support)
    ow.set_intensity(d["white_field_intensity"])
    ow.initialise_estimate(seed)
    oracle_schedule(ow, xw, [(oracle.ER, 20)], 0, 0, 0)
    wf = files.read_cplx(tmp_path / "wf.cplx", n, n)
    assert rel(wf, xw) <= 1e-9

    # step 2: object with the reconstructed white field
    stdout = run_tool("cdi_reconstructiond", tmp_path, cfg, "fresnel", seed)
    # the tool prints |white field(0, 0)|^2 after reading it
    printed = [l for l in stdout.splitlines() if re.fullmatch(r"[-+0-9.eE]+", l.strip())]
    assert float(printed[0]) == pytest.approx(abs(wf[0, 0]) ** 2, rel=1e-5)
    xf = np.zeros((n, n), np.complex128)
    of = oracle.OracleCDI(xf, oracle.FRESNEL)
    of.fresnel_init(xw, d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
    of.set_support(d["support"])
    of.set_intensity(d["intensity"])
    of.initialise_estimate(seed)
    oracle_schedule(of, xf, [(oracle.ER, 8), (oracle.HIO, 8), (oracle.ER, 8)], 12, 1.5, 0.1)
    assert stdout.count("Applying shrink-wrap at iteration") == 2
    esw = files.read_cplx(tmp_path / "fresnel.cplx", n, n)
    assert rel(esw, xf) <= 1e-9
    T = files.read_cplx(tmp_path / "transmission_function.cplx", n, n)
    assert rel(T, of.get_transmission_function()) <= 1e-9
    img = files.read_dbin(tmp_path / "fresnel_0.dbin", n, n)
    assert img.shape == (n, n)
