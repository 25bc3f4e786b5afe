"""Request loop for reconstruction jobs with continuous batching.

Frames of a beamline stream arrive as requests, each a PlanarCDI
reconstruction with its own size, algorithm, iteration count and shrinkwrap
schedule.  The server

* admits arrived requests in arrival order while their device memory fits
  the HBM budget (and a slot limit, if set); a request whose memory does not
  fit waits, and so do the ones behind it;
* keeps every admitted job's estimate, support, intensity and work fields
  on the device between iterations;
* iterates all admitted jobs together in ticks: one tick gives each job up
  to `tick` iterations, cut at its next shrinkwrap point or its end, and
  jobs of one precision and size share every launch of the fused passes
  (nadia_batch_iterate_each);
* retires a job as soon as its last iteration is done, and admits waiting
  requests in the memory and slots it frees, while the others keep running.
* With a backend that can mark its queue (mark / done / wait_mark, retire /
  collect), the device never waits for the host: tick k+1 is queued before
  the results of tick k are read (asynchronously, into page-locked memory),
  so retiring, admitting and scheduling overlap the device's work.

A job's result does not depend on what it ran beside: batched launches
compute every job independently, with the same kernels and reductions as a
job iterated alone.

The device side is a backend object, so the scheduling logic runs (and is
tested) with a simulated backend on the CPU; `DeviceBackend` is the one that
drives the HIP library.
"""
import collections
import dataclasses
import time

import numpy as np

# algorithm ids as in NADIA's BaseCDI enum (NADIA/include/BaseCDI.h)
ER, HIO, RAAR = 0, 3, 8


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float      # seconds after the stream starts
    n: int              # n x n pixels
    algorithm: int
    beta: float
    iterations: int
    shrinkwrap: bool    # shrinkwrap every `every` iterations
    seed: int
    # configuration-file frames (frames.py): an algorithm schedule of
    # (algorithm, iterations) blocks (algorithm is its first), and shrinkwrap
    # after every sw_every iterations with a fixed width and threshold; the
    # synthetic stream leaves them unset (one algorithm, shrinkwrap every
    # `every` iterations with the width decreasing 3 -> 1.5)
    schedule: tuple = None
    sw_every: int = 0
    sw_sigma: float = 1.5
    sw_threshold: float = 0.1


@dataclasses.dataclass
class Completed:
    rid: int
    n: int
    arrival: float
    admitted: float
    finished: float
    iterations: int
    error: float
    result: object = None

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_stream(count=2000, rate=40.0, sizes=(256, 512, 1024, 2048), probs=(0.4, 0.3, 0.2, 0.1),
                   iterations=(100, 1000), shrinkwrap_prob=0.5, seed=2024):
    """Synthetic request stream: exponential inter-arrival times of mean
    1/rate, size drawn from `sizes` with `probs`, algorithm uniform over
    {ER, HIO beta 0.9, RAAR beta 0.85}, iteration count uniform on the closed
    range `iterations`, shrinkwrap with probability `shrinkwrap_prob`.  All
    draws from numpy's PCG64 generator seeded with `seed`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    n = rng.choice(np.asarray(sizes), size=count, p=np.asarray(probs, dtype=float))
    alg = rng.integers(0, 3, count)
    its = rng.integers(iterations[0], iterations[1] + 1, count)
    sw = rng.random(count) < shrinkwrap_prob
    table = ((ER, 0.9), (HIO, 0.9), (RAAR, 0.85))
    return [Request(i, float(t[i]), int(n[i]), table[alg[i]][0], table[alg[i]][1], int(its[i]), bool(sw[i]),
                    int(seed * 100003 + i)) for i in range(count)]


def shrinkwrap_sigma(call, ncalls, start=3.0, end=1.5):
    """Gaussian width of shrinkwrap call `call` of `ncalls`: linear start->end."""
    if ncalls <= 1:
        return end
    return start + (end - start) * min(call, ncalls - 1) / (ncalls - 1)


def percentile(values, q):
    """q-th percentile (linear interpolation between order statistics)."""
    if not len(values):
        return None
    return float(np.percentile(np.asarray(values, dtype=float), q))


class _Job:
    """A running request: its iteration count and the events of its schedule
    -- (after iteration count, kind, arguments) with kind "sw" (shrinkwrap:
    width, threshold) or "alg" (switch algorithm) -- in order; a tick is cut
    at the next event, and events at the same count run shrinkwrap first
    (the reference tool shrinkwraps at the end of an iteration and switches
    algorithm before the next block)."""
    __slots__ = ("req", "handle", "bytes", "done", "admitted", "pending", "events")

    def __init__(self, req, handle, nbytes, admitted, every, threshold):
        self.req = req
        self.handle = handle
        self.bytes = nbytes
        self.done = 0
        self.admitted = admitted
        ev = []
        if req.sw_every > 0:
            ev += [(k, 0, "sw", (req.sw_sigma, req.sw_threshold))
                   for k in range(req.sw_every, req.iterations, req.sw_every)]
        elif req.shrinkwrap:
            ncalls = (req.iterations - 1) // every
            ev += [((c + 1) * every, 0, "sw", (shrinkwrap_sigma(c, ncalls), threshold)) for c in range(ncalls)]
        if req.schedule:
            at = 0
            for b in range(len(req.schedule) - 1):
                at += req.schedule[b][1]
                ev.append((at, 1, "alg", (req.schedule[b + 1][0],)))
        ev.sort(key=lambda e: (e[0], e[1]))
        self.events = collections.deque(e for e in ev if 0 < e[0] < req.iterations)

    def count(self, tick):
        """Iterations of the next tick: up to `tick`, cut at the next event."""
        c = min(tick, self.req.iterations - self.done)
        if self.events:
            c = min(c, self.events[0][0] - self.done)
        return c

    def advance(self, c, be):
        """Count c iterations as done and run the events that fall due."""
        self.done += c
        while self.events and self.events[0][0] == self.done:
            _, _, kind, arg = self.events.popleft()
            if kind == "sw":
                be.shrinkwrap(self.handle, *arg)
            else:
                be.set_algorithm(self.handle, *arg)


class Server:
    """Continuous-batching request loop (see the module docstring).

    backend: object with bytes_for(req), start(req) -> handle,
    run(handles, counts), shrinkwrap(handle, sigma, threshold), wait(),
    finish(handle, keep) -> (error, result), release_idle(nbytes) -> freed
    bytes (optional), held_bytes() -> bytes of idle cached jobs (optional).
    Asynchronous backends add mark() -> token (after the queued work),
    done(token), wait_mark(token), retire(handle, keep) -> pending (the job
    leaves the running set; its results are read asynchronously) and
    collect(pending) -> (error, result) once a later mark has completed.
    """

    def __init__(self, backend, hbm_budget=200e9, max_active=None, tick=50, every=50, threshold=0.1,
                 keep_results=False, admit_per_tick=16):
        self.backend = backend
        # asynchronous loop: requests admitted per pass, so that a burst of
        # arrivals is uploaded while the device runs the jobs already in
        self.admit_per_tick = admit_per_tick
        self.budget = float(hbm_budget)
        self.max_active = max_active
        self.tick = int(tick)
        self.every = int(every)
        self.threshold = threshold
        self.keep = keep_results
        self.peak_bytes = 0
        self.peak_active = 0
        self.ticks = 0
        self.iterations = 0

    def _held(self):
        f = getattr(self.backend, "held_bytes", None)
        return f() if f else 0

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = collections.deque()
        active = []
        used = 0
        done = []
        t0 = clock()
        be = self.backend
        if hasattr(be, "mark"):
            return self._run_async(pending, waiting, clock, sleep)
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: arrival order, HBM budget, slot limit
            while waiting and (self.max_active is None or len(active) < self.max_active):
                req = waiting[0]
                need = be.bytes_for(req)
                if need > self.budget:
                    raise ValueError("request %d needs %d bytes, more than the whole budget" % (req.rid, need))
                if used + self._held() + need > self.budget and hasattr(be, "release_idle"):
                    be.release_idle(used + self._held() + need - self.budget)
                if used + self._held() + need > self.budget:
                    break
                waiting.popleft()
                h = be.start(req)
                active.append(_Job(req, h, need, clock() - t0, self.every, self.threshold))
                used += need
            self.peak_bytes = max(self.peak_bytes, used + self._held())
            self.peak_active = max(self.peak_active, len(active))
            if not active:
                if pending:
                    dt = pending[0].arrival - (clock() - t0)
                    if dt > 0:
                        sleep(dt)
                continue
            # one tick: up to `tick` iterations per job, cut at its next
            # shrinkwrap point or its end
            counts = [j.count(self.tick) for j in active]
            be.run([j.handle for j in active], counts)
            self.ticks += 1
            self.iterations += sum(counts)
            for j, c in zip(active, counts):
                j.advance(c, be)
            finished = [j for j in active if j.done >= j.req.iterations]
            if finished:
                be.wait()
                tf = clock() - t0
                for j in finished:
                    err, res = be.finish(j.handle, self.keep)
                    done.append(Completed(j.req.rid, j.req.n, j.req.arrival, j.admitted, tf, j.done, err, res))
                    used -= j.bytes
                active = [j for j in active if j.done < j.req.iterations]
            else:
                be.wait()
        self.elapsed = clock() - t0
        return done

    def _admit(self, waiting, active, used, t0, clock, limit=None):
        be = self.backend
        n = 0
        while waiting and (self.max_active is None or len(active) < self.max_active) and (limit is None or n < limit):
            n += 1
            req = waiting[0]
            need = be.bytes_for(req)
            if need > self.budget:
                raise ValueError("request %d needs %d bytes, more than the whole budget" % (req.rid, need))
            if used + self._held() + need > self.budget and hasattr(be, "release_idle"):
                be.release_idle(used + self._held() + need - self.budget)
            if used + self._held() + need > self.budget:
                break
            waiting.popleft()
            h = be.start(req)
            active.append(_Job(req, h, need, clock() - t0, self.every, self.threshold))
            used += need
        return used

    def _collect(self, entry, done, t0, clock):
        """Wait for a queued tick and record its finished jobs; returns the
        bytes they free."""
        tok, fin = entry
        self.backend.wait_mark(tok)
        tf = clock() - t0
        freed = 0
        for j in fin:
            err, res = self.backend.collect(j.pending)
            done.append(Completed(j.req.rid, j.req.n, j.req.arrival, j.admitted, tf, j.done, err, res))
            freed += j.bytes
        return freed

    def _run_async(self, pending, waiting, clock, sleep):
        """The request loop over an asynchronous backend: at most two ticks
        are queued; the older one's finished jobs are read while the device
        runs the newer one."""
        be = self.backend
        active, done = [], []
        inflight = collections.deque()  # (token, jobs that finished in that tick)
        used = 0
        t0 = clock()
        while pending or waiting or active or inflight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            before = len(active)
            used = self._admit(waiting, active, used, t0, clock, self.admit_per_tick)
            if waiting and len(active) - before < self.admit_per_tick and any(fin for _, fin in inflight):
                # admission is blocked (memory or slots) and queued ticks
                # hold finished jobs: free them first, then admit again
                while inflight:
                    used -= self._collect(inflight.popleft(), done, t0, clock)
                used = self._admit(waiting, active, used, t0, clock, self.admit_per_tick - (len(active) - before))
            self.peak_bytes = max(self.peak_bytes, used + self._held())
            self.peak_active = max(self.peak_active, len(active))
            if active:
                counts = [j.count(self.tick) for j in active]
                be.run([j.handle for j in active], counts)
                self.ticks += 1
                self.iterations += sum(counts)
                for j, c in zip(active, counts):
                    j.advance(c, be)
                finished = [j for j in active if j.done >= j.req.iterations]
                for j in finished:
                    j.pending = be.retire(j.handle, self.keep)
                inflight.append((be.mark(), finished))
                active = [j for j in active if j.done < j.req.iterations]
            # read the older tick while the newer one runs; wait for the last
            # one only when nothing else is queued behind it
            while inflight and (len(inflight) > 1 or not active or be.done(inflight[0][0])):
                used -= self._collect(inflight.popleft(), done, t0, clock)
                if active:
                    break
            if not active and not inflight and not waiting and pending:
                dt = pending[0].arrival - (clock() - t0)
                if dt > 0:
                    sleep(dt)
        self.elapsed = clock() - t0
        return done

    @staticmethod
    def summary(done, elapsed):
        lat = [c.latency * 1e3 for c in done]
        its = sum(c.iterations for c in done)
        return {"requests": len(done), "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "mean_ms": float(np.mean(lat)) if lat else None, "elapsed_s": elapsed,
                "aggregate_iter_s": its / elapsed if elapsed > 0 else None}


class DeviceBackend:
    """Backend over the HIP library: one nadia_batch, PlanarCDI jobs in FP32
    bulk (FP64 reductions) or FP64.  Finished jobs' device objects are kept
    per size and reused by later requests of that size; they count against
    the budget (held_bytes) and are freed when admission needs the memory.

    inputs(req) -> dict(intensity, support, x0[, beam_stop]) supplies each
    request's data (host arrays)."""

    def __init__(self, nd, inputs, double=False):
        self.nd = nd
        self.inputs = inputs
        self.double = double
        self.batch = nd.Batch()
        # (n, has beam stop) -> [(est, recon)]: a reconstruction keeps its
        # beam stop once set (BaseCDI has no way to clear it), so it is only
        # reused by requests with the same beam-stop state
        self.idle = collections.defaultdict(list)
        self.idle_bytes = 0
        self._bytes = {}
        self._events = []       # recorded-and-waited events, reused
        self._pinned = []       # PinnedDoubles blocks of error slots
        self._free_slots = []

    def bytes_for(self, req):
        if req.n not in self._bytes:
            self._bytes[req.n] = self.nd.bytes_estimate(req.n, req.n, double=self.double)
        return self._bytes[req.n]

    def held_bytes(self):
        return self.idle_bytes

    def release_idle(self, nbytes):
        freed = 0
        for key in sorted(self.idle, reverse=True):
            while self.idle[key] and freed < nbytes:
                est, p = self.idle[key].pop()
                del p, est
                freed += self._bytes[key[0]]
        self.idle_bytes -= freed
        return freed

    def start(self, req):
        nd = self.nd
        d = self.inputs(req)
        cdt = np.complex128 if self.double else np.complex64
        key = (req.n, "beam_stop" in d)
        # no copy when already in the serving precision; without a starting
        # point the estimate comes from initialise_estimate(seed), as in the
        # reference tool (NADIA/tools/CDI_reconstruction.c)
        x0 = np.ascontiguousarray(d["x0"], dtype=cdt) if "x0" in d else None
        if self.idle[key]:
            est, p = self.idle[key].pop()
            self.idle_bytes -= self._bytes[req.n]
            if x0 is not None:
                est.set(x0)
            # back to a new object's algorithm state (ER, beta 0.9,
            # NADIA/src/BaseCDI.c++:33-66) so the request below selects its
            # algorithm exactly as on a fresh object
            p.set_relaxation_parameter(0.9)
            p.set_algorithm(ER)
        else:
            est = nd.Complex2D(req.n, req.n, double=self.double, data=x0)
            p = nd.PlanarCDI(est)
        p.set_support(d["support"])
        p.set_intensity(d["intensity"])
        if key[1]:
            p.set_beam_stop(d["beam_stop"])
        if x0 is None:
            p.initialise_estimate(req.seed)
        # set_relaxation_parameter before set_algorithm: selecting the
        # algorithm a job already has does nothing (NADIA/src/BaseCDI.c++:194-195)
        p.set_relaxation_parameter(req.beta)
        p.set_algorithm(req.algorithm)
        self.batch.add(p)
        return (req, est, p, key[1])

    def run(self, handles, counts):
        # the batch holds exactly the active jobs, in admission order
        order = {id(h[2]): c for h, c in zip(handles, counts)}
        self.batch.iterate_each([order[id(p)] for p in self.batch.jobs])

    def shrinkwrap(self, handle, sigma, threshold):
        handle[2].apply_shrinkwrap(sigma, threshold)

    def set_algorithm(self, handle, alg):
        handle[2].set_algorithm(alg)

    def wait(self):
        self.nd.synchronize()

    # ---- asynchronous retirement ----
    def mark(self):
        ev = self._events.pop() if self._events else self.nd.Event()
        ev.record()
        return ev

    def done(self, ev):
        return ev.query()

    def wait_mark(self, ev):
        ev.synchronize()
        self._events.append(ev)

    def retire(self, handle, keep):
        """The job leaves the batch now (the next tick does not run it); its
        error is copied out asynchronously into a page-locked slot."""
        if not self._free_slots:
            blk = self.nd.PinnedDoubles(256)
            self._pinned.append(blk)
            self._free_slots.extend((blk, i) for i in range(256))
        slot = self._free_slots.pop()
        handle[2].get_error_async(slot[0], slot[1])
        self.batch.remove(handle[2])
        return (handle, slot, keep)

    def collect(self, pending):
        handle, slot, keep = pending
        req, est, p, beam = handle
        err = float(slot[0].a[slot[1]])
        self._free_slots.append(slot)
        res = est.get() if keep else None
        self.idle[(req.n, beam)].append((est, p))
        self.idle_bytes += self._bytes[req.n]
        return err, res

    def finish(self, handle, keep):
        req, est, p, beam = handle
        err = p.get_error()
        res = est.get() if keep else None
        self.batch.remove(p)
        self.idle[(req.n, beam)].append((est, p))
        self.idle_bytes += self._bytes[req.n]
        return err, res


def input_pool(synth, sizes, per_size=2, seed=7, double=False):
    """Inputs as the 2048^2 star-object workload with radius and support
    scaled by n/2048, `per_size` datasets per size (seeds seed + 10*k); a
    request uses dataset req.seed % per_size of its size.  The frames are
    stored in the serving precision (float32 / complex64 unless double), as
    a detector stream would deliver them, so admission only uploads."""
    rdt, cdt = (np.float64, np.complex128) if double else (np.float32, np.complex64)

    def frame(n, k):
        d = synth.planar_star(n=n, seed=seed + 10 * k)
        return {key: np.ascontiguousarray(v, dtype=cdt if key == "x0" else rdt)
                for key, v in d.items() if key in ("intensity", "support", "x0", "beam_stop")}

    pool = {n: [frame(n, k) for k in range(per_size)] for n in sizes}
    return lambda req: pool[req.n][req.seed % per_size]
