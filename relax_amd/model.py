"""Device-resident RK2 loop of RELAX (relax.f90:555-890) through the C ABI."""
import ctypes

import numpy as np

from ._lib import LAYER, PATCH, PLANE, SOURCE, TENSOR_LAYER, TETRA, WEAK, EventDesc, ModelDesc, StepInfo, check, load, ptr

# mechanisms of set_structure (relax.f90:565-650): linear and power-law
# viscosity, fault creep, linear and power-law transient creep
LINEAR, POWERLAW, FAULTCREEP, LTRANSIENT, NLTRANSIENT = 0, 1, 2, 3, 4


def footprint(sx1, sx2, sx3):
    d = ModelDesc(sx1, sx2, sx3, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.2, 1.0, 1.0, 1.0, 1, 0, None)
    return int(load().relax_model_footprint(ctypes.byref(d)))


class Model:
    """One RELAX model held in HBM.

    State is exchanged in the reference's layout: u1..u3 and gamma as
    (sx3/2, sx2, sx1+2) float32 (align_data=1) or (sx3/2, sx2, sx1), tau and
    sig as (sx3/2, sx2, sx1, 6) float32.
    """

    def __init__(self, sx1, sx2, sx3, dx1, dx2, dx3, lam, mu, gamma=0.0, beta=0.2, interval=1e30, odt=1.0,
                 tscale=1.0, skip=1, event_times=(0.0,)):
        self._lib = load()
        self.shape = (sx1, sx2, sx3)
        self.dx = (dx1, dx2, dx3)
        self._events = np.ascontiguousarray(event_times, np.float64)
        self._desc = ModelDesc(sx1, sx2, sx3, dx1, dx2, dx3, lam, mu, gamma, beta, interval, odt, tscale, skip,
                               len(self._events), self._events.ctypes.data)
        h = ctypes.c_void_p()
        check(self._lib.relax_model_create(ctypes.byref(self._desc), ctypes.byref(h)), "relax_model_create")
        self._h = h
        self.planes = None

    def close(self):
        if self._h:
            self._lib.relax_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def reset(self):
        """Back to a new model's state, keeping the device memory."""
        check(self._lib.relax_model_reset(self._h), "reset")

    def set_structure(self, which, layers):
        if layers is None:
            check(self._lib.relax_model_set_structure(self._h, which, None, 0), "set_structure")
            return
        layers = np.ascontiguousarray(layers, LAYER)
        check(self._lib.relax_model_set_structure(self._h, which, ptr(layers), len(layers)), "set_structure")

    def set_weakzones(self, which, zones):
        """Weak zones (WEAK array) of mechanism LINEAR, POWERLAW, LTRANSIENT
        or NLTRANSIENT; an empty array removes them."""
        zones = np.ascontiguousarray(zones, WEAK)
        check(self._lib.relax_model_set_weakzones(self._h, which, ptr(zones) if len(zones) else None, len(zones)),
              "set_weakzones")

    def set_transient_strain(self, epsilonik):
        """epsilonik as (sx3/2, sx2, sx1, 6) float32 (needs a transient structure)."""
        e = np.ascontiguousarray(epsilonik, np.float32)
        sx1, sx2, sx3 = self.shape
        if e.shape != (sx3 // 2, sx2, sx1, 6):
            raise ValueError("epsilonik must be (sx3/2, sx2, sx1, 6)")
        check(self._lib.relax_model_set_transient_strain(self._h, ptr(e)), "set_transient_strain")

    def transient_strain(self):
        sx1, sx2, sx3 = self.shape
        e = np.zeros((sx3 // 2, sx2, sx1, 6), np.float32)
        check(self._lib.relax_model_get_transient_strain(self._h, ptr(e)), "get_transient_strain")
        return e

    def set_prestress(self, prestress):
        prestress = np.ascontiguousarray(prestress, TENSOR_LAYER)
        check(self._lib.relax_model_set_prestress(self._h, ptr(prestress), len(prestress)), "set_prestress")

    def set_planes(self, planes, patches=None):
        planes = np.ascontiguousarray(planes, PLANE)
        self.planes = planes
        pa = None if patches is None else np.ascontiguousarray(patches, PATCH)
        check(self._lib.relax_model_set_planes(self._h, ptr(planes), len(planes), ptr(pa)), "set_planes")

    def get_patches(self):
        n = int(sum(int(p["px2"]) * int(p["px3"]) for p in self.planes)) if self.planes is not None else 0
        out = np.zeros(n, PATCH)
        if n:
            check(self._lib.relax_model_get_patches(self._h, ptr(out)), "get_patches")
        return out

    def upload(self, u1, u2, u3, tau, sig=None, gamma=None, t=0.0, oi=1, i=0, align_data=1):
        arrs = [np.ascontiguousarray(a, np.float32) if a is not None else None for a in (u1, u2, u3, tau, sig, gamma)]
        check(self._lib.relax_model_upload(self._h, *[ptr(a) for a in arrs], t, oi, i, align_data), "upload")

    def download(self, align_data=1):
        sx1, sx2, sx3 = self.shape
        h3 = sx3 // 2
        p = sx1 + 2 if align_data else sx1
        u = [np.zeros((h3, sx2, p), np.float32) for _ in range(3)]
        tau = np.zeros((h3, sx2, sx1, 6), np.float32)
        sig = np.zeros_like(tau)
        gam = np.zeros((h3, sx2, p), np.float32)
        t = ctypes.c_double()
        check(self._lib.relax_model_download(self._h, ptr(u[0]), ptr(u[1]), ptr(u[2]), ptr(tau), ptr(sig), ptr(gam),
                                             ctypes.byref(t), align_data), "download")
        return dict(u1=u[0], u2=u[1], u3=u[2], tau=tau, sig=sig, gamma=gam, t=t.value)

    def step(self, dt=None, info=True):
        """One predictor/corrector step; returns the step record as a dict."""
        rec = StepInfo() if info else None
        d = None if dt is None else ctypes.byref(ctypes.c_double(dt))
        check(self._lib.relax_model_step(self._h, d, None if rec is None else ctypes.byref(rec)), "step")
        return None if rec is None else _info(rec)

    def run(self, nsteps, dt_replay=None):
        recs = (StepInfo * max(nsteps, 1))()
        done = ctypes.c_int()
        dts = None if dt_replay is None else np.ascontiguousarray(dt_replay, np.float64)
        check(self._lib.relax_model_run(self._h, nsteps, ptr(dts), recs, ctypes.byref(done)), "run")
        return [_info(recs[k]) for k in range(done.value)]

    def set_interseismic(self, sources=None, tensile=None):
        """Inter-seismic loading (relax.f90:669-674, :795-799): SOURCE arrays
        of shear segments (slip rate) and tensile segments (opening rate);
        none removes it."""
        arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE) for a in (sources, tensile)]
        ev = EventDesc(ptr(arr[0]), len(arr[0]), ptr(arr[1]), len(arr[1]), None, 0, None, 0)
        check(self._lib.relax_model_set_interseismic(self._h, ctypes.byref(ev)), "set_interseismic")

    def okada(self, sources=None, tensile=None, nyquist=1.0):
        """dislocations_disp (relax.f90:357-360, :1300-1349): Okada's
        displacement of the shear patches and tensile cracks at or below the
        nyquist size added to u, then sig = -tau + C:eps(u).  SOURCE arrays
        in the deck's convention (structure.deck_copies)."""
        arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE) for a in (sources, tensile)]
        check(self._lib.relax_model_okada(self._h, ptr(arr[0]), len(arr[0]), ptr(arr[1]), len(arr[1]),
                                          float(nyquist)), "okada")

    def add_shear(self, patches, slip):
        """momentdensityshear of each patch (PLANE array, internal convention) into tau."""
        patches = np.ascontiguousarray(patches, PLANE)
        slip = np.ascontiguousarray(slip, np.float64)
        check(self._lib.relax_model_add_shear(self._h, ptr(patches), ptr(slip), len(patches)), "add_shear")

    def coseismic(self):
        """u = Green(body force of tau), sig = -tau + C:eps(u)."""
        check(self._lib.relax_model_coseismic(self._h), "coseismic")

    def event(self, sources=None, loads=None, nyquist=1.0, tensile=None, mogi=None, cuboids=None, tetrahedra=None,
              vertices=None):
        """A coseismic event (dislocations, relax.f90:1116-1289, and traction):
        the first at t = 0 on a new model, a later one once a step has reached
        its time (the step's "e" advanced).  SOURCE arrays (or None) of shear
        patches, surface loads, tensile cracks (opening) and point dilatation
        sources (slip = volume change); eigenstrain cuboids (WEAK array, the
        eigenstrain in e) and tetrahedra (TETRA array, 1-based indices into
        vertices, (n, 3) float64).  u += Green(the body forces of the sources
        larger than nyquist*min(dx), of every point source, of the eigenstrain
        volumes, and the loads), tau += their moment densities,
        sig = -tau + C:eps(u)."""
        arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE)
               for a in (sources, tensile, mogi, loads)]
        cb = np.ascontiguousarray(np.zeros(0, WEAK) if cuboids is None else cuboids, WEAK)
        th = np.ascontiguousarray(np.zeros(0, TETRA) if tetrahedra is None else tetrahedra, TETRA)
        vx = np.ascontiguousarray(np.zeros((0, 3)) if vertices is None else vertices, np.float64).reshape(-1, 3)
        ev = EventDesc(ptr(arr[0]), len(arr[0]), ptr(arr[1]), len(arr[1]), ptr(arr[2]), len(arr[2]), ptr(arr[3]),
                       len(arr[3]), ptr(cb), len(cb), ptr(th), len(th), ptr(vx), len(vx))
        check(self._lib.relax_model_event(self._h, ctypes.byref(ev), float(nyquist)), "event")

    def profile(self, nsteps):
        """Run nsteps steps with HIP-event phase timers; returns
        {kind: (total_ms, launches)} for the kernel kinds of PROFILE_KINDS."""
        ms = np.zeros(len(PROFILE_KINDS), np.float64)
        n = np.zeros(len(PROFILE_KINDS), np.int32)
        check(self._lib.relax_model_profile(self._h, nsteps, ptr(ms), ptr(n)), "profile")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(PROFILE_KINDS)}

    def synchronize(self):
        check(self._lib.relax_model_synchronize(self._h), "synchronize")

    TUNE_VARIANTS = 6

    def autotune(self, times=False):
        """Choose the launch shapes of the tunable kernels for this grid by
        timing them (relax_model_autotune); returns the chosen variants, and
        with times=True also the ms per launch of every variant (None where
        a variant does not apply)."""
        c = np.zeros(3, np.int32)
        ms = np.zeros((3, self.TUNE_VARIANTS), np.float32)
        check(self._lib.relax_model_autotune(self._h, ptr(c), ptr(ms)), "autotune")
        out = dict(fft_cols=int(c[0]), green=int(c[1]), bodyforce=int(c[2]))
        if times:
            out["ms"] = {k: [None if v < 0 else round(float(v), 4) for v in ms[i]]
                         for i, k in enumerate(("fft_cols", "green", "bodyforce"))}
        return out

    def set_tuning(self, fft_cols=0, green=0, bodyforce=0):
        """Use the given launch-shape variants (relax_model_set_tuning)."""
        c = np.array([fft_cols, green, bodyforce], np.int32)
        check(self._lib.relax_model_set_tuning(self._h, ptr(c)), "set_tuning")

    def points(self, xyz, rot=0.0):
        """u and sig at observation points (exportpoints, util.f90:59-139):
        xyz (n, 3) in the computational frame; returns u (n, 3) float64 in
        the user frame (rotation by -rot degrees) and sig (n, 6) float32."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        u = np.zeros((len(x), 3), np.float64)
        sig = np.zeros((len(x), 6), np.float32)
        check(self._lib.relax_model_points(self._h, ptr(x), len(x), float(rot), ptr(u), ptr(sig)), "points")
        return u, sig

    def surface(self, u1, u2, u3):
        """Copy the surface plane (x3 = 0) of u1..u3 into (sx2, sx1) float32
        arrays."""
        for a in (u1, u2, u3):
            if a.dtype != np.float32 or not a.flags.c_contiguous:
                raise ValueError("surface arrays must be C-contiguous float32")
        check(self._lib.relax_model_surface(self._h, ptr(u1), ptr(u2), ptr(u3), 0), "surface")


class Batch:
    """Models of one grid stepped together: every kernel of the RK2 step is
    one launch over all members, replayed from a captured hipGraph while the
    membership is unchanged (relax_batch_* of relax_hip.h)."""

    def __init__(self, models=()):
        self._lib = load()
        h = ctypes.c_void_p()
        check(self._lib.relax_batch_create(ctypes.byref(h)), "batch create")
        self._h = h
        self.models = []
        for m in models:
            self.add(m)

    def add(self, model):
        check(self._lib.relax_batch_add(self._h, model.handle), "batch add")
        self.models.append(model)

    def remove(self, model):
        check(self._lib.relax_batch_remove(self._h, model.handle), "batch remove")
        self.models.remove(model)

    def __len__(self):
        return int(self._lib.relax_batch_size(self._h))

    def step(self, info=True):
        """One RK2 step of every member; returns the members' step records
        (in the order of addition) or None."""
        n = len(self.models)
        if not info:
            check(self._lib.relax_batch_step(self._h, None), "batch step")
            return None
        recs = (StepInfo * max(n, 1))()
        check(self._lib.relax_batch_step(self._h, recs), "batch step")
        return [_info(recs[k]) for k in range(n)]

    def synchronize(self):
        check(self._lib.relax_batch_synchronize(self._h), "batch synchronize")

    def busy(self):
        """True while the batch has queued work (does not wait)."""
        return bool(self._lib.relax_batch_busy(self._h))

    def autotune(self, times=False):
        """Choose the launch shapes of the tunable kernels for this batch by
        timing one launch over all current members per variant
        (relax_batch_autotune); returns the choice as Model.autotune does.
        Later batches of the grid and rheology class start from it."""
        c = np.zeros(3, np.int32)
        ms = np.zeros((3, Model.TUNE_VARIANTS), np.float32)
        check(self._lib.relax_batch_autotune(self._h, ptr(c), ptr(ms)), "batch autotune")
        out = dict(fft_cols=int(c[0]), green=int(c[1]), bodyforce=int(c[2]))
        if times:
            out["ms"] = {k: [None if v < 0 else round(float(v), 4) for v in ms[i]]
                         for i, k in enumerate(("fft_cols", "green", "bodyforce"))}
        return out

    def scalars(self):
        """The members' current step records (waits for queued work)."""
        n = len(self.models)
        recs = (StepInfo * max(n, 1))()
        if n:
            check(self._lib.relax_batch_scalars(self._h, recs), "batch scalars")
        return [_info(recs[k]) for k in range(n)]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.relax_batch_destroy(self._h)
            self._h = None
            self.models = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# kernel kinds of relax_model_profile, in the order of RELAX_PROF_* (relax_hip.h)
PROFILE_KINDS = ("bodyforce", "fft_x_fwd", "fft_y", "green_z", "fft_x_inv", "stress_pred", "stress_corr", "small")


def _info(r):
    return dict(i=r.i, Dt=r.Dt, tm=r.tm, maxwell=list(r.maxwell), power=r.power, cei=r.cei, t=r.t, oi=r.oi,
                status=r.status, e=r.e)
