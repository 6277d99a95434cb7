"""Serving forward-model requests with continuous batching.

A request is one RELAX forward model: a grid, a rheology, coseismic slip
patches, optionally an afterslip plane, and a number of RK2 steps -- a
workload description, or a RELAX input deck (deck_request).  The
server keeps every admitted model's state (u, sigma, tau, gamma, fault
patches) on the device from its coseismic set-up to its last step, and steps
the models of one grid class together as a batch (relax_batch_*: shared
launches replayed from a hipGraph).  Each grid class has its own batch and
stream, and a batch takes its next step as soon as its previous one is done,
so the classes overlap on the device.  Requests are admitted in arrival order
while their HBM footprint (relax_model_footprint) fits the budget, joining
their class's batch between two of its steps; finished requests leave, return
their surface displacement, and free their memory for the queue.

With mixed=True the server steps every admitted request, whatever its grid,
in one batch: the divergence and strain passes are one work-list launch over
all members, the transforms and Green passes one launch per grid class.

The baseline runs the same requests back to back through the single-model
path: set-up, steps, surface displacement, release -- one at a time.
"""
import ctypes
import dataclasses
import math
import threading
import time

import numpy as np

from . import workloads
from ._lib import RelaxError, check, load
from .model import Batch, footprint

# grid classes of the request stream: (sx1, sx2, sx3, dx) and their weights
GRIDS = ((128, 128, 64, 0.5), (256, 256, 128, 1.0), (512, 512, 256, 0.5))
WEIGHTS = (6, 3, 1)
# A class's batch holds at most this many grid points (over its members):
# beyond it one batched step already keeps HBM busy, and more members only
# lengthen every member's step.
MAX_BATCH_POINTS = 1 << 25
# A mixed batch (Server(mixed=True): every class in one batch) holds at most
# this many grid points over its members: one 512x512x256 request and the
# smaller ones beside it.
MIXED_BATCH_POINTS = 1 << 27
MIXED = "mixed"  # the worker key of the mixed batch


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float       # seconds after the start of the stream
    spec: dict           # a workload description (relax_amd.workloads format)
    nsteps: int
    grid: int            # index into GRIDS (-1: a grid of its own, e.g. a deck's)

    @property
    def cls(self):
        return grid_class(self.spec)


def grid_class(s):
    """The batch class of a spec: its grid, spacing and elastic constants.
    The per-class server steps the models of one class as one batch (its
    transforms and Green passes share launches); a mixed batch holds
    several classes (relax_batch_add)."""
    return (s["sx1"], s["sx2"], s["sx3"]) + workloads.spacing(s) + workloads.constants(s)[:3]


def deck_request(rid, source, nsteps, arrival=0.0, **options):
    """A request from a RELAX input deck (its text or a file name; options
    are relax_amd.deck.read's section switches)."""
    from . import deck
    s = deck.read(source, **options).spec(name="deck %d" % rid, nsteps=nsteps)
    return Request(rid, arrival, s, nsteps, -1)


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    start: float
    finish: float
    nsteps: int
    t: float             # model time after the last step
    status: int
    surface: tuple       # (u1, u2, u3) at x3 = 0, (sx2, sx1) each

    @property
    def latency(self):
        return self.finish - self.arrival


def request_spec(rng, gi, afterslip):
    """One request: rheology as the batched-model workload (log-uniform eta in
    [1e17, 1e20] Pa s, n in {1, 2, 3, 3.5}, a 15-40 km lid), 1-10 patches of
    0.5-5 m slip, and on request an afterslip plane with (a-b)sigma = 1 MPa,
    v0 = 1 m/yr, dip 20, rake 90, away from the slip."""
    sx1, sx2, sx3, dx = GRIDS[gi]
    half = 0.5 * sx1 * dx
    n = [1.0, 2.0, 3.0, 3.5][rng.integers(0, 4)]
    eta = 10 ** rng.uniform(17, 20)
    lid = rng.uniform(15, 40)
    lay = [(lid, workloads.gdot0(eta, n), n, 0.0, 0.6, 0.0)]
    deck = []
    for _ in range(int(rng.integers(1, 11))):
        p = list(workloads._deck_patch(rng, -0.4 * half, 0.0, min(10.0, 0.25 * (sx3 // 2) * dx), 5.0, 5e-3))
        p[0] = rng.uniform(0.5e-3, 5e-3)
        deck.append(tuple(p))
    spec = dict(name="request", sx1=sx1, sx2=sx2, sx3=sx3, dx=dx, linear=lay if n == 1.0 else None,
                nonlinear=None if n == 1.0 else lay, friction=None, planes=None, deck=deck, interval=1e9, odt=1.0,
                skip=1, tscale=1.0)
    if afterslip:
        L = 0.25 * sx1 * dx
        W = 0.5 * L
        dip = 20.0
        spec["friction"] = [(0.0, 1e-3, 1.0, 0.0, 0.6, 0.0)]
        spec["planes"] = [dict(x=0.3 * half, y=0.3 * half, z=1.0, length=L, width=W, strike=0.0, dip=dip,
                               rake=90.0)]
    return spec


def request_stream(n, rate=20.0, seed=20261206, afterslip_fraction=0.3, steps=(10, 30)):
    """Seeded Poisson arrivals (mean `rate` per second) of n requests over the
    grid classes in the ratio 6:3:1, with 10-30 RK2 steps each."""
    rng = np.random.default_rng(seed)
    p = np.asarray(WEIGHTS, float) / sum(WEIGHTS)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        gi = int(rng.choice(len(GRIDS), p=p))
        spec = request_spec(rng, gi, rng.uniform() < afterslip_fraction)
        out.append(Request(i, t, spec, int(rng.integers(steps[0], steps[1] + 1)), gi))
    return out


def meminfo():
    """(free, total) bytes of device memory."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(load().relax_hip_meminfo(ctypes.byref(f), ctypes.byref(t)), "meminfo")
    return f.value, t.value


def _surface(m, s):
    sx1, sx2 = s["sx1"], s["sx2"]
    u = [np.zeros((sx2, sx1), np.float32) for _ in range(3)]
    m.surface(*u)
    return tuple(u)


def _footprint(s):
    return footprint(s["sx1"], s["sx2"], s["sx3"])


def _points(s):
    return s["sx1"] * s["sx2"] * s["sx3"]


def batch_cap(cls):
    """Members of one batch class (grid_class) at most."""
    return max(1, MAX_BATCH_POINTS // (cls[0] * cls[1] * cls[2]))


def tune(grids=None, seed=7):
    """Choose launch shapes for the request stream's grid classes before
    serving: per grid, the single-model autotuner on one model
    (relax_model_autotune, used by the single-model path), and the batch
    autotuner on a full batch of throw-away models, once without and once
    with an afterslip member (the two rheology classes a batch of the stream
    runs, relax_batch_autotune).  Later models and batches of these classes
    in the process start from the choices.  Returns {(grid index,
    "single" | "batch" | "batch afterslip"): choice}."""
    rng = np.random.default_rng(seed)
    out = {}
    for gi in range(len(GRIDS)) if grids is None else grids:
        m = workloads.build(request_spec(rng, gi, False))
        try:
            out[(gi, "single")] = m.autotune()
        finally:
            m.close()
        for afterslip in (False, True):
            first = request_spec(rng, gi, afterslip)
            specs = [first] + [request_spec(rng, gi, False) for _ in range(batch_cap(grid_class(first)) - 1)]
            models = [workloads.build(s) for s in specs]
            b = Batch(models)
            try:
                out[(gi, "batch afterslip" if afterslip else "batch")] = b.autotune()
            finally:
                b.close()
                for m in models:
                    m.close()
    return out


def kernel_time(requests):
    """Device time of the stream's RK2 steps on the single-model path: each
    request's model, set up as for serving, runs its steps under the
    HIP-event phase timers (relax_model_profile).  Returns the summed kernel
    time of the steps in seconds, by grid index; the set-up (the coseismic
    event) is not counted."""
    out = {}
    for r in requests:
        m = workloads.build(r.spec)
        try:
            prof = m.profile(r.nsteps)
        finally:
            m.close()
        out[r.grid] = out.get(r.grid, 0.0) + sum(ms for ms, _ in prof.values()) / 1e3
    return out


def run_back_to_back(requests, realtime=True):
    """Each request alone through the single-model path, in arrival order."""
    t0 = time.perf_counter()
    results = []
    for r in requests:
        now = time.perf_counter() - t0
        if realtime and now < r.arrival:
            time.sleep(r.arrival - now)
        start = time.perf_counter() - t0
        m = workloads.build(r.spec)
        status = 0
        info = None
        for _ in range(r.nsteps - 1):
            m.step(info=False)
        try:
            info = m.step()
        except RelaxError as e:  # a step's error status (e.g. fault creep overflow)
            status = e.status
        surf = _surface(m, r.spec)
        m.close()
        finish = time.perf_counter() - t0
        results.append(Result(r.rid, r.arrival if realtime else 0.0, start, finish, r.nsteps,
                              info["t"] if info else math.nan, info["status"] if info else status, surf))
    return results


class Server:
    """Continuous batching over batch classes (see the module docstring).

    Every class has a worker thread of its own that admits the class's
    requests into its batch, steps the batch and retires finished members;
    the library releases the interpreter lock inside its calls, so a class's
    next steps are enqueued while another class's thread sets up a model,
    and no class waits for the host work of another.  The calling thread
    hands each request to its class at its arrival time.

    Finished models go back to a pool of their class and are reset for the
    next request: releasing device memory waits for the whole device, which
    would stall every class's queued steps.  `used` counts the footprint of
    every model held, pooled ones included; pooled models of other classes
    are released when an admission needs their memory, and a request that
    does not fit waits until memory is returned."""

    def __init__(self, budget_bytes=None, reserve=0.1, admit_per_pass=2, chunk=4, mixed=False):
        self.admit_per_pass = admit_per_pass
        self.chunk = chunk
        # mixed: one worker steps every admitted request in one batch of
        # uneven grids (relax_batch_* work-list launches), up to
        # MIXED_BATCH_POINTS grid points over its members
        self.mixed = mixed
        if budget_bytes is None:
            free, _ = meminfo()
            budget_bytes = int(free * (1.0 - reserve))
        self.budget = budget_bytes
        self.used = 0
        self.pool = {}      # batch class -> idle models
        self._mem = threading.Condition()  # guards used and pool

    def _reserve(self, r):
        """Claim the memory of request r (an idle model of its class, or new
        footprint, releasing idle models of other classes if needed)."""
        with self._mem:
            idle = self.pool.get(r.cls)
            if idle:
                r._model = idle.pop()  # taken here: other classes release only idle models
                return True
            fp = _footprint(r.spec)
            need = self.used + fp - self.budget
            if need > 0:
                for cls, idle in self.pool.items():
                    while idle and need > 0 and cls != r.cls:
                        m = idle.pop()
                        m.close()
                        f = footprint(*cls[:3])
                        self.used -= f
                        need -= f
            if need > 0:
                return False
            self.used += fp
            r._model = None
            return True

    def _admit(self, r, now):
        """Set up request r's model (the pooled one _reserve took, or a new
        one) and return it."""
        m = getattr(r, "_model", None)
        r._model = None
        if m is not None:
            m.reset()
            workloads.configure(m, r.spec)
        else:
            m = workloads.build(r.spec)
        return m

    def _release(self, cls, m):
        with self._mem:
            self.pool.setdefault(cls, []).append(m)
            self._mem.notify_all()

    def _room(self, cls, members, r):
        """Whether request r may join the batch of `members`."""
        if cls != MIXED:
            return len(members) < batch_cap(cls)
        pts = sum(_points(e[0].spec) for e in members)
        return not members or pts + _points(r.spec) <= MIXED_BATCH_POINTS

    def _worker(self, cls, inbox, results, t0, realtime):
        batch = Batch()
        members = []  # [request, model, steps taken, start]
        waiting = inbox["queue"]
        try:
            while True:
                # admission in arrival order, a few per pass, while the
                # class has room and the memory fits
                admitted = 0
                while admitted < self.admit_per_pass:
                    with inbox["cv"]:
                        r = waiting[0] if waiting else None
                    if r is None or not self._room(cls, members, r) or not self._reserve(r):
                        break
                    with inbox["cv"]:
                        waiting.pop(0)
                    start = time.perf_counter() - t0
                    m = self._admit(r, start)
                    batch.add(m)
                    members.append([r, m, 0, start])
                    admitted += 1
                if not members:
                    with inbox["cv"]:
                        if not waiting and inbox["closed"]:
                            return
                        if not waiting:
                            inbox["cv"].wait(0.05)
                            continue
                    # a request waits for memory held by other classes
                    with self._mem:
                        self._mem.wait(0.005)
                    continue
                # up to `chunk` steps at once (no member past its last
                # step), one while requests wait that the class could admit
                with inbox["cv"]:
                    more = bool(waiting) and self._room(cls, members, waiting[0])
                nst = 1 if more else min(self.chunk, min(r.nsteps - k for r, _, k, _ in members))
                for _ in range(nst):
                    batch.step(info=False)
                for e in members:
                    e[2] += nst
                recs = batch.scalars()  # waits for the steps
                keep = []
                for (r, m, k, start), rec in zip(members, recs):
                    if k >= r.nsteps or rec["status"]:
                        batch.remove(m)
                        surf = _surface(m, r.spec)
                        self._release(r.cls, m)
                        res = Result(r.rid, r.arrival if realtime else 0.0, start, time.perf_counter() - t0, k,
                                     rec["t"], rec["status"], surf)
                        with inbox["cv"]:
                            results.append(res)
                    else:
                        keep.append([r, m, k, start])
                members = keep
        except BaseException as e:  # surfaces in run()
            inbox["error"] = e
        finally:
            batch.close()

    def run(self, requests, realtime=True, poll=2e-4):
        pending = sorted(requests, key=lambda r: r.arrival)
        inboxes = {}
        threads = []
        results = []
        t0 = time.perf_counter()
        for r in pending:
            now = time.perf_counter() - t0
            if realtime and r.arrival > now:
                time.sleep(r.arrival - now)
            key = MIXED if self.mixed else r.cls
            box = inboxes.get(key)
            if box is None:
                box = inboxes[key] = {"queue": [], "cv": threading.Condition(), "closed": False, "error": None}
                th = threading.Thread(target=self._worker, args=(key, box, results, t0, realtime), daemon=True)
                threads.append(th)
                th.start()
            with box["cv"]:
                box["queue"].append(r)
                box["cv"].notify()
        for box in inboxes.values():
            with box["cv"]:
                box["closed"] = True
                box["cv"].notify()
        for th in threads:
            th.join()
        with self._mem:
            for idle in self.pool.values():
                for m in idle:
                    m.close()
            self.pool.clear()
            self.used = 0
        for box in inboxes.values():
            if box["error"] is not None:
                raise box["error"]
        return sorted(results, key=lambda x: x.finish)


def summarize(results, label):
    lat = np.array([r.latency for r in results])
    span = max(r.finish for r in results) - min(r.arrival for r in results)
    ok = sum(1 for r in results if r.status == 0)
    return {"path": label, "requests": len(results), "ok": ok, "solves_per_s": round(len(results) / span, 3),
            "makespan_s": round(span, 3), "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4)}
