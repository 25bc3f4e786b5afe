"""The Config reader of the drop-in API (include/nadia/Config.h,
nadia-coecxs_amd/cpp/Config.cpp), built with g++ alone (no device code) and
probed on a file holding every parsing rule of NADIA/src/Config.c++:22-161:
comments, '=' or space separators, indented keys, first occurrence winning,
first-word strings, word lists, atof / atoi conversion, and the messages and
FAILURE status of missing keys, empty lists and unreadable files."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CONFIG = """#set the parameters
data_file_name=./image_files/planar_data.dbin   # trailing comment
pixels_x = 1024
pixels_y 512
  indented_key = 7
output_file_name_prefix = planar extra words
algorithm = HIO ER   RAAR
iterations= 250 150 25
beam_wavelength = 1.55e-10
shrinkwrap_threshold = 0.1 # threshold at 10%
dup = first
dup = second
int_with_text = 12abc
empty_list =
#commented_out = 3
no_separator_line
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("config")
    exe = d / "config_probe"
    subprocess.run(["g++", "-O1", "-std=c++11", "-I", os.path.join(ROOT, "include", "nadia"),
                    os.path.join(ROOT, "tests", "cpp", "config_probe.cpp"),
                    os.path.join(ROOT, "nadia-coecxs_amd", "cpp", "Config.cpp"), "-o", str(exe)], check=True)
    cfg = d / "example.config"
    cfg.write_text(CONFIG)

    def run(*queries, path=cfg):
        out = subprocess.run([str(exe), str(path), *queries], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
        return out.stdout.splitlines()
    run.path = cfg
    return run


def answers(lines):
    return {l.split(" ->")[0][2:]: l.split(" ->")[1].strip() for l in lines if l.startswith("= ")}


def test_values_and_separators(probe):
    lines = probe("s:data_file_name", "i:pixels_x", "i:pixels_y", "i:indented_key", "s:output_file_name_prefix",
                  "d:beam_wavelength", "d:shrinkwrap_threshold", "s:dup", "i:int_with_text", "status")
    a = answers(lines)
    assert a["s data_file_name"] == "[./image_files/planar_data.dbin]"
    assert a["i pixels_x"] == "1024"
    assert a["i pixels_y"] == "512"          # space as the separator
    assert a["i indented_key"] == "7"
    assert a["s output_file_name_prefix"] == "[planar]"   # first word only
    assert float(a["d beam_wavelength"]) == 1.55e-10
    assert float(a["d shrinkwrap_threshold"]) == 0.1
    assert a["s dup"] == "[first]"           # the first occurrence wins
    assert a["i int_with_text"] == "12"      # atoi
    assert lines[-1] == "status 1"


def test_lists(probe):
    a = answers(probe("S:algorithm", "I:iterations"))
    assert a["S algorithm"] == "[HIO] [ER] [RAAR]"
    assert a["I iterations"] == "250 150 25"


def test_missing_and_empty_keys_fail(probe):
    lines = probe("status", "s:starting_point_file_name", "status")
    assert lines[0] == "status 1"
    assert "Key:starting_point_file_name not found in config file" in lines
    assert answers(lines)["s starting_point_file_name"] == "[]"
    assert lines[-1] == "status 0"
    for q, msg in (("i:commented_out", "not found"), ("d:no_separator_line", "not found"),
                   ("S:empty_list", "has no entries"), ("I:empty_list", "has no entries")):
        lines = probe(q, "status")
        assert any(l.startswith("Key:" + q[2:]) and msg in l for l in lines), lines
        assert lines[-1] == "status 0"
        if q[0] in "SI":
            assert answers(lines)[q[0] + " " + q[2:]] == "null"
        else:
            assert float(answers(lines)[q[0] + " " + q[2:]]) == 0


def test_unreadable_file(probe, tmp_path):
    lines = probe("status", path=tmp_path / "absent.config")
    assert lines[0] == "Could not open the file " + str(tmp_path / "absent.config")
    assert lines[-1] == "status 0"


def test_python_config_matches_cpp(probe, capsys):
    """The request loop's front end (nadia-coecxs_amd/frames.py) reads
    configuration files with a Python Config: every answer, message and
    status equals the C++ Config's on the same file."""
    import importlib

    frames = importlib.import_module("nadia-coecxs_amd.frames")
    queries = ["s:data_file_name", "i:pixels_x", "i:pixels_y", "i:indented_key", "s:output_file_name_prefix",
               "d:beam_wavelength", "d:shrinkwrap_threshold", "s:dup", "i:int_with_text", "S:algorithm",
               "I:iterations", "s:starting_point_file_name", "i:commented_out", "S:empty_list", "I:empty_list",
               "d:no_separator_line"]
    cpp = probe(*queries, "status")
    capsys.readouterr()
    c = frames.Config(str(probe.path))
    lines = []
    for q in queries:
        kind, key = q[0], q[2:]
        if kind == "s":
            r = " [%s]" % c.get_string(key)
        elif kind == "d":
            r = " %.17g" % c.get_double(key)
        elif kind == "i":
            r = " %d" % c.get_int(key)
        elif kind == "S":
            v = c.get_string_list(key)
            r = " null" if v is None else "".join(" [%s]" % w for w in v)
        else:
            v = c.get_int_list(key)
            r = " null" if v is None else "".join(" %d" % w for w in v)
        lines += capsys.readouterr().out.splitlines()
        lines.append("= %s %s ->%s" % (kind, key, r))
    lines.append("status %d" % c.status)
    assert lines == cpp
