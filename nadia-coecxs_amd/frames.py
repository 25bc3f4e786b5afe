"""Configuration files as frames of the serving loop.

NADIA's reconstruction tool (NADIA/tools/CDI_reconstruction.c:94-466) runs one
reconstruction described by a configuration file (NADIA/src/Config.c++:22-161).
Here each configuration file is one frame of a beamline stream: its data and
support (.dbin), starting point (.cplx, optional; else initialise_estimate with
the frame's seed), algorithm schedule ("algorithm = HIO ER", "iterations =
250 150") and shrinkwrap settings (shrinkwrap_iterations / _gauss_width /
_threshold) become one request of serve.Server, so many frames are
reconstructed together with continuous batching, and each result is written
like the tool's, to <output_file_name_prefix>.cplx.  A frame's result is the
one the tool computes for that file (same operations in the same order:
set_support, set_intensity, initialise_estimate, set_algorithm per block,
shrinkwrap after iteration i when i % shrinkwrap_iterations ==
shrinkwrap_iterations - 1).

    python -m nadia-coecxs_amd.frames [--double] [--rate R] [--seed S] a.config b.config ...
"""
import argparse
import importlib
import os
import re
import sys

import numpy as np

_NUM = re.compile(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)")
_INT = re.compile(r"\s*([+-]?\d+)")


def _atof(s):
    m = _NUM.match(s)
    return float(m.group(1)) if m else 0.0


def _atoi(s):
    m = _INT.match(s)
    return int(m.group(1)) if m else 0


class Config:
    """NADIA's Config (NADIA/src/Config.c++:22-161): '#' comments, "key = value"
    or "key value", the first occurrence of a key wins; get_string is the
    value's first word, lists are every word, numbers use atof / atoi; a
    missing key or an empty list reports the reference's message and sets
    the status to failure (0)."""

    def __init__(self, path):
        self.mapping = {}
        self.status = 0
        try:
            with open(path) as f:
                lines = f.read().split("\n")
        except OSError:
            print("Could not open the file %s" % path)
            return
        for line in lines:
            line = line.split("#", 1)[0]
            sep = line.find("=")
            if sep < 0:
                sep = line.find(" ")
            if sep < 0:
                continue
            k = line[:sep].split()
            self.mapping.setdefault(k[0] if k else "", line[sep + 1:])
        self.status = 1

    def _lookup(self, key):
        if key not in self.mapping:
            print("Key:%s not found in config file" % key)
            self.status = 0
            return None
        return self.mapping[key]

    def get_string(self, key):
        v = self._lookup(key)
        w = v.split() if v is not None else []
        return w[0] if w else ""

    def get_double(self, key):
        v = self._lookup(key)
        return _atof(v) if v is not None else 0.0

    def get_int(self, key):
        v = self._lookup(key)
        return _atoi(v) if v is not None else 0

    def _list(self, key):
        v = self._lookup(key)
        if v is None:
            return None
        w = v.split()
        if not w:
            print("Key:%s has no entries in the config file" % key)
            self.status = 0
            return None
        return w

    def get_string_list(self, key):
        return self._list(key)

    def get_int_list(self, key):
        w = self._list(key)
        return None if w is None else [_atoi(x) for x in w]


class Frame:
    """One planar reconstruction described by a configuration file."""

    def __init__(self, path, seed=0):
        nd = importlib.import_module("nadia-coecxs_amd")
        c = Config(path)
        if not c.status:
            raise ValueError("could not open the configuration file %s" % path)
        self.path = path
        self.seed = seed
        self.nx, self.ny = c.get_int("pixels_x"), c.get_int("pixels_y")
        algs = c.get_string_list("algorithm")
        its = c.get_int_list("iterations")
        if not c.status:
            raise ValueError("problem reading the configuration file %s" % path)
        if len(algs) != len(its):
            raise ValueError("%s: the number of algorithms and the number of iterations do not match" % path)
        self.schedule = []
        for a, n in zip(algs, its):
            alg = nd.get_alg_from_name(a)
            if alg < 0:
                raise ValueError("%s: could not find reconstruction algorithm with the name %s" % (path, a))
            self.schedule.append((alg, int(n)))
        # optional keys, 0 / "" when absent (as the tool reads them)
        opt = Config(path)
        self.sw_every = opt.get_int("shrinkwrap_iterations")
        self.sw_sigma = opt.get_double("shrinkwrap_gauss_width")
        self.sw_threshold = opt.get_double("shrinkwrap_threshold")
        base = os.path.dirname(os.path.abspath(path))
        rel = lambda p: p if not p or os.path.isabs(p) else os.path.join(base, p)
        self.data_file = rel(c.get_string("data_file_name"))
        self.support_file = rel(c.get_string("support_file_name"))
        self.start_file = rel(opt.get_string("starting_point_file_name"))
        self.output = rel(c.get_string("output_file_name_prefix") + ".cplx")
        if not c.status:
            raise ValueError("problem reading the configuration file %s" % path)

    def load(self, double=False):
        """The frame's arrays in the serving precision."""
        files = importlib.import_module("nadia-coecxs_amd.files")
        rdt, cdt = (np.float64, np.complex128) if double else (np.float32, np.complex64)
        d = {"intensity": files.read_dbin(self.data_file, self.nx, self.ny).astype(rdt),
             "support": files.read_dbin(self.support_file, self.nx, self.ny).astype(rdt)}
        if self.start_file:
            d["x0"] = files.read_cplx(self.start_file, self.nx, self.ny).astype(cdt)
        return d

    def request(self, rid, arrival=0.0):
        serve = importlib.import_module("nadia-coecxs_amd.serve")
        if self.nx != self.ny:
            raise ValueError("%s: the request loop takes square frames" % self.path)
        total = sum(n for _, n in self.schedule)
        return serve.Request(rid, arrival, self.nx, self.schedule[0][0], 0.9, total, self.sw_every > 0, self.seed,
                             schedule=tuple(self.schedule), sw_every=self.sw_every, sw_sigma=self.sw_sigma,
                             sw_threshold=self.sw_threshold)


def serve_configs(nd, paths, seed=0, rate=None, double=False, hbm_budget=200e9, write=True, tick=50):
    """Reconstruct the frames of `paths` through the request loop (arrivals at
    1/rate intervals, or all at once); writes each frame's result to its
    output file when `write`.  Returns the serve.Completed records (results
    included) in request order."""
    serve = importlib.import_module("nadia-coecxs_amd.serve")
    files = importlib.import_module("nadia-coecxs_amd.files")
    frames = [Frame(p, seed) for p in paths]
    reqs = [f.request(i, 0.0 if not rate else i / rate) for i, f in enumerate(frames)]
    data = {i: f.load(double) for i, f in enumerate(frames)}
    be = serve.DeviceBackend(nd, lambda r: data[r.rid], double=double)
    srv = serve.Server(be, hbm_budget=hbm_budget, tick=tick, keep_results=True)
    done = sorted(srv.run(reqs), key=lambda c: c.rid)
    if write:
        for c in done:
            files.write_cplx(frames[c.rid].output, c.result)
    return done


def main(argv=None):
    ap = argparse.ArgumentParser(description="reconstruct configuration-file frames through the request loop")
    ap.add_argument("configs", nargs="+")
    ap.add_argument("--double", action="store_true", help="FP64 throughout (default: FP32 bulk, FP64 sums)")
    ap.add_argument("--rate", type=float, default=None, help="frames per second (default: all at once)")
    ap.add_argument("--seed", type=int, default=0, help="initialise_estimate seed of frames without a start")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    a = ap.parse_args(argv)
    nd = importlib.import_module("nadia-coecxs_amd")
    done = serve_configs(nd, a.configs, a.seed, a.rate, a.double, a.budget_gb * 1e9)
    for c in done:
        print("frame %d (%s): %d iterations, error %.6g, %.1f ms" % (c.rid, a.configs[c.rid], c.iterations, c.error,
                                                                    c.latency * 1e3))
    return 0


if __name__ == "__main__":
    sys.exit(main())
