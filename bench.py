#!/usr/bin/env python3
"""Throughput of RELAX's RK2 elastic + relaxation step on one MI355X.

A step is one predictor/corrector pass of the RK2 loop (relax.f90:555-890)
over the 512x512x256 workload (40 slip patches, power law n = 3 below a 25 km
lid; relax_amd.workloads "c1"), with the whole state resident in HBM.

  python bench.py [--steps K] [--warmup W] [--workload c1] [--no-cpu-baseline]
                  [--no-batched] [--no-serving] [--only batched|serving|dropin]

Prints one JSON line.  The timed region is K back-to-back steps on the
model's stream with HIP-event phase timers (one synchronisation at the end),
bracketed by device synchronisations; `roofline` reports the dominant
kernel's algorithmic bytes per launch over its event-measured average
duration.  `cpu_baseline` times the oracle (the C++ restatement of the
reference, OpenMP on the host cores) on two steps of the same state.
`batched` (64 models of 128^3) and `serving` (the request stream) ride on
the same line; --only runs one of them alone (dropin: the per-routine Green's
function entry with host arrays, against the same solve on a resident model).
"""
import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "RK2 steps/s at 512x512x256 + % HBM roofline; batched solves/s, p99 latency"


def kernel_bytes(sx1, sx2, sx3):
    """Algorithmic bytes per launch of each kernel kind: every input read
    once, every output written once (FP32 fields, H = sx1*sx2*sx3/2)."""
    H = sx1 * sx2 * (sx3 // 2)
    nq = 8 * 3 * sx2 * (sx3 // 2)  # Nyquist planes of the work spectrum
    tp = 4 * 3 * sx1 * sx2         # the three traction planes (real), or their half spectra
    tn = 8 * 3 * sx2               # the traction planes' Nyquist rows
    return {
        "bodyforce": 4 * (6 * H + 3 * H) + 12 * sx1 * sx2,   # read sig, write body force + traction
        # read real rows, write half spectrum; the traction planes ride along
        "fft_x_fwd": 4 * (3 * H + 3 * H) + nq + 2 * tp + tn,
        # read + write the half spectrum, with the Nyquist rows riding along
        # (and, forward, the traction planes): the mean of the forward and
        # inverse launches
        "fft_y": 4 * (3 * H + 3 * H) + 2 * nq + (2 * tp + 2 * tn) // 2,
        "green_z": 4 * (3 * H + 3 * H) + 2 * nq,              # read + write spectrum and Nyquist planes
        "fft_x_inv": 4 * (3 * H + 3 * H + 3 * H) + nq,        # read spectrum + u, write v or u
        "stress_pred": 4 * (3 * H + 6 * H + 6 * H + 6 * H),   # read v, sig, tau; write sig
        "stress_corr": 4 * (3 * H + 6 * H + 6 * H + H + 6 * H + 6 * H + H),  # + tau, gamma in/out
    }


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "rk2_step_c1_512x512x256_hbm_traffic.md")


# the tunable kernel kinds and their autotuner field
TUNED_KIND = {"fft_y": "fft_cols", "green_z": "green", "bodyforce": "bodyforce"}


def measured_traffic(tuning):
    """HBM bytes per launch of each kernel kind at 512x512x256, read from the
    committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE, each scaled by its
    reading for a device copy of known size).  The summary gives the tunable
    kinds per launch-shape variant; a kind whose variant this run uses was
    not measured is left out."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    rec = json.loads(line.split(":", 1)[1])
                    break
            else:
                return None, "no traffic line in the profile"
    except OSError as e:
        return None, "profile not readable: %s" % e
    out, missing = {}, []
    for kind, b in rec["bytes"].items():
        if kind in TUNED_KIND:
            field = TUNED_KIND[kind]
            v = str(tuning[field])
            # {variant: bytes}, or the bytes of the variant the profile's
            # "tuning" names
            if not isinstance(b, dict):
                b = {str(rec.get("tuning", {}).get(field)): b}
            if v not in b:
                missing.append("%s variant %s" % (kind, v))
                continue
            b = b[v]
        out[kind] = b
    return out, ("not profiled: " + ", ".join(missing)) if missing else None


def step_bytes_model(sx1, sx2, sx3):
    """Compulsory bytes of one RK2 step in the unfused accounting of the
    path (404 N + 192 N/sx1 + 96 sx1 sx2, N = sx1 sx2 sx3)."""
    N = sx1 * sx2 * sx3
    return 404 * N + 192 * N / sx1 + 96 * sx1 * sx2


def cpu_baseline(spec, state, nthreads, nsteps=2):
    """The oracle (oracle/, FP32 storage, OpenMP over nthreads host threads)
    on nsteps RK2 steps from the GPU's state; returns the seconds taken."""
    ROOT = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, ROOT)
    os.environ["OMP_NUM_THREADS"] = str(nthreads)
    import oracle as O
    from relax_amd import workloads

    t = workloads.tables(spec)
    dx1, dx2, dx3 = workloads.spacing(spec)
    m = dict(sx1=spec["sx1"], sx2=spec["sx2"], sx3=spec["sx3"], dx1=dx1, dx2=dx2, dx3=dx3,
             lam=workloads.LAM, mu=workloads.MU, gam=workloads.GAM, beta=workloads.BETA, linear=t["linear"],
             nonlinear=t["nonlinear"], faultcreep=t["friction"], prestress=t["prestress"], interval=1e30,
             odt=1.0, tscale=1.0, skip=1, event_times=[0.0])
    st = {k: state[k] for k in ("u1", "u2", "u3", "tau", "sig", "gamma")}
    st.update(t=0.0, oi=1, i=0)
    t0 = time.perf_counter()
    O.rk2_run(m, st, nsteps)
    return time.perf_counter() - t0


def batched_c3(nmodels=64, nsteps=10):
    """64 independent 128^3 models (relax_amd.workloads c3<k>: seeded patches,
    viscosity and stress exponent), 10 RK2 steps each: one batch of shared
    launches replayed from a hipGraph, against the same models stepped back to
    back through the single-model path (no host synchronisation between
    steps), each path with the launch shapes its own autotuner picks.  A solve
    is the 10 steps of one model from its coseismic state."""
    import numpy as np

    from relax_amd import workloads
    from relax_amd.model import Batch

    specs = [workloads.spec("c3%d" % k) for k in range(nmodels)]
    t0 = time.perf_counter()
    single = [workloads.build(s) for s in specs]
    setup = (time.perf_counter() - t0) / nmodels
    batched = [workloads.build(s) for s in specs]
    # warm-up of both paths on throw-away models of the same grid; each path
    # gets the launch shapes its autotuner picks (one model, or one launch
    # over all 64 members)
    w = workloads.build(specs[0])
    tune_single = w.autotune()
    w.step(info=False)
    w.synchronize()
    wb = Batch([workloads.build(specs[1]), workloads.build(specs[2])])
    wb.step(info=False)
    wb.synchronize()
    for m in wb.models + [w]:
        m.close()
    wb.close()
    for m in single:
        m.set_tuning(**tune_single)
    t0 = time.perf_counter()
    for m in single:
        for _ in range(nsteps):
            m.step(info=False)
    for m in single:
        m.synchronize()
    t_single = time.perf_counter() - t0
    b = Batch(batched)
    tune_batch = b.autotune()
    b.step(info=False)  # capture (not timed); one step more than the single path
    b.synchronize()
    for m in single:  # bring the single path level
        m.step(info=False)
    for m in single:
        m.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        b.step(info=False)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    same, diff = True, 0.0
    for k in (0, nmodels // 3, nmodels - 1):
        a, c = single[k].download(), batched[k].download()
        for f in ("u1", "u2", "u3", "sig", "tau", "gamma"):
            same = same and np.array_equal(a[f], c[f])
            ref = np.linalg.norm(a[f].astype(np.float64))
            if ref > 0:
                diff = max(diff, float(np.linalg.norm(c[f].astype(np.float64) - a[f]) / ref))
    b.close()
    for m in single + batched:
        m.close()
    return {"workload": "c3: %d models of 128^3, dx=1 km, %d RK2 steps each" % (nmodels, nsteps),
            "solves_per_s": round(nmodels / t_batch, 2),
            "single_path_solves_per_s": round(nmodels / t_single, 2),
            "speedup": round(t_single / t_batch, 2),
            "setup_s_per_model": round(setup, 4),
            "launch_shapes": {"single_path": tune_single, "batch": tune_batch},
            # bitwise when both paths use the same shapes; the Green variants
            # differ from each other in FP32 rounding only
            "identical_to_single_path": bool(same),
            "max_rel_l2_to_single_path": diff}


def serving_c4(nreq=60, rate=20.0):
    """The serving stream (relax_amd.serve.request_stream: Poisson arrivals,
    grids 128x128x64 / 256x256x128 / 512x512x256 in the ratio 6:3:1, 10-30
    steps, 30 % with an afterslip plane).  Capacity: the whole stream queued
    at once, continuous batching against the single-model path back to back.
    Latency: arrivals replayed at their times, p50/p99 time-to-solution
    (set-up, steps and the surface displacement of each request).  Both
    also with every request in one batch of uneven grids (mixed=True).  Also the
    device time of the stream's step kernels on the single-model path
    (relax_model_profile): the single path's makespan over it bounds the
    speedup of any scheme that only removes host work and launch gaps."""
    from relax_amd import serve

    warm = serve.request_stream(6, seed=1, steps=(2, 3))
    serve.run_back_to_back(warm, realtime=False)
    serve.Server().run(warm, realtime=False)
    tuned = serve.tune()  # launch shapes per grid class: single models and full batches
    reqs = serve.request_stream(nreq, rate=rate)
    cap_single = serve.summarize(serve.run_back_to_back(reqs, realtime=False), "single-model path, back to back")
    cap_served = serve.summarize(serve.Server().run(reqs, realtime=False), "continuous batching")
    lat_single = serve.summarize(serve.run_back_to_back(reqs, realtime=True), "single-model path, back to back")
    lat_served = serve.summarize(serve.Server().run(reqs, realtime=True), "continuous batching")
    # every request in one batch of uneven grids (work-list launches)
    cap_mixed = serve.summarize(serve.Server(mixed=True).run(reqs, realtime=False), "one mixed batch")
    lat_mixed = serve.summarize(serve.Server(mixed=True).run(reqs, realtime=True), "one mixed batch")
    # the stream's step kernels alone (single-model launches, no gaps): a
    # scheme that only removed host work and launch gaps could not finish
    # the queued stream faster than this
    work = serve.kernel_time(reqs)
    kernel_s = sum(work.values())
    return {"workload": "c4: %d requests, Poisson %g/s, grids 128x128x64:256x256x128:512x512x256 = 6:3:1, "
                        "10-30 steps, 30%% afterslip" % (nreq, rate),
            "solves_per_s": cap_served["solves_per_s"],
            "single_path_solves_per_s": cap_single["solves_per_s"],
            "speedup": round(cap_served["solves_per_s"] / cap_single["solves_per_s"], 2),
            "p50_s": lat_served["p50_s"], "p99_s": lat_served["p99_s"],
            "single_path_p50_s": lat_single["p50_s"], "single_path_p99_s": lat_single["p99_s"],
            "mixed_batch_solves_per_s": cap_mixed["solves_per_s"],
            "mixed_batch_p50_s": lat_mixed["p50_s"], "mixed_batch_p99_s": lat_mixed["p99_s"],
            "step_kernel_s": round(kernel_s, 4),
            "step_kernel_s_by_grid": {"x".join(map(str, serve.GRIDS[g][:3])): round(v, 4) for g, v in work.items()},
            "step_kernel_s_over_served_makespan": round(kernel_s / cap_served["makespan_s"], 3),
            "speedup_bound_from_step_kernels": round(cap_single["makespan_s"] / kernel_s, 2),
            "launch_shapes": {"%s %s" % ("x".join(map(str, serve.GRIDS[gi][:3])), kind): [
                c["fft_cols"], c["green"], c["bodyforce"]] for (gi, kind), c in tuned.items()},
            "capacity": [cap_single, cap_served, cap_mixed], "at_arrival_rate": [lat_single, lat_served, lat_mixed]}


def dropin_green(calls=3):
    """The literal drop-in's cost: relax_greenfunctioncowling, the
    per-routine entry a relinked relax.f90 calls with host arrays, on the
    512x512x256 grid (copies of c1..c3 and t1..t3 to the device and back on
    every call, as the reference's CUDA port does), against the same Green
    solve on the device-resident model (its transforms and Green pass, from
    the HIP-event phase timers of one step, divided by the step's two
    solves)."""
    import numpy as np

    from relax_amd import green, workloads

    sx1, sx2, sx3 = 512, 512, 256
    rng = np.random.default_rng(5)
    c = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2), np.float32) for _ in range(3)]
    for a in c:
        a[: sx3 // 2, :, :sx1] = rng.standard_normal((sx3 // 2, sx2, sx1))
    for a in t:
        a[:, :sx1] = rng.standard_normal((sx2, sx1))
    green.greenfunctioncowling(*c, *t, 0.5, 0.5, 0.5, workloads.LAM, workloads.MU, 0.0)  # warm-up
    t0 = time.perf_counter()
    for _ in range(calls):
        green.greenfunctioncowling(*c, *t, 0.5, 0.5, 0.5, workloads.LAM, workloads.MU, 0.0)
    per_call = (time.perf_counter() - t0) / calls
    m = workloads.build(workloads.spec("c1"))
    m.autotune()
    m.step(info=False)
    prof = m.profile(4)
    m.close()
    solve_ms = sum(prof[k][0] for k in ("fft_x_fwd", "fft_y", "green_z", "fft_x_inv") if k in prof) / (2 * 4)
    return {"workload": "greenfunctioncowling at 512x512x256, host arrays in and out",
            "per_routine_ms_per_call": round(per_call * 1e3, 2),
            "device_resident_solve_ms": round(solve_ms, 3),
            "ratio": round(per_call * 1e3 / solve_ms, 1),
            "bytes_copied_per_call": 2 * 4 * (3 * sx3 * sx2 * (sx1 + 2) + 3 * sx2 * (sx1 + 2))}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 64 x 128^3 measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--only", choices=("batched", "serving", "dropin"),
                    help="run only this measurement and print its JSON (for profiling it alone); dropin: the "
                         "per-routine Green entry with host arrays against the device-resident solve")
    args = ap.parse_args()
    if args.only:
        out = {"batched": batched_c3, "serving": serving_c4, "dropin": dropin_green}[args.only]()
        print(json.dumps(out), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the RK2 step of one model per GPU "
              "(--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from relax_amd import workloads

    spec = workloads.spec(args.workload)
    sx1, sx2, sx3 = spec["sx1"], spec["sx2"], spec["sx3"]
    print("# building %s (%dx%dx%d) on the device" % (args.workload, sx1, sx2, sx3), file=sys.stderr, flush=True)
    model = workloads.build(spec)
    tuned = model.autotune(times=True)  # launch shapes of the tunable kernels, timed on this grid
    for _ in range(args.warmup):
        model.step(info=False)
    model.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prof = model.profile(args.steps)
    model.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    info = model.step(info=True)  # one more step to read the scalars (not timed)
    steps_per_s = args.steps / elapsed

    kb = kernel_bytes(sx1, sx2, sx3)
    kernels = {}
    for k, (ms, n) in prof.items():
        if n == 0:
            continue
        avg = ms / n
        e = {"launches_per_step": n / args.steps, "avg_ms": round(avg, 5), "ms_per_step": round(ms / args.steps, 4)}
        if k in kb:
            e["algorithmic_bytes"] = kb[k]
            e["achieved_GBs"] = round(kb[k] / (avg * 1e-3) / 1e9, 1)
        kernels[k] = e
    dom = max((k for k in kernels if k in kb), key=lambda k: prof[k][0])
    achieved = kernels[dom]["achieved_GBs"]
    B_step = step_bytes_model(sx1, sx2, sx3)
    traffic, why = measured_traffic(tuned) if args.workload == "c1" else (None, "profiled at c1 only")
    for k, e in kernels.items():
        if traffic and k in traffic:
            e["hbm_bytes_profiled"] = traffic[k]
    fused_bytes = sum(kb[k] * kernels[k]["launches_per_step"] for k in kb if k in kernels)
    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 3),
        "unit": "RK2 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded slip patches, relax_amd.workloads.%s)" % args.workload,
        "config": {"workload": "%s: %dx%dx%d FP32, dx=%g km, 40 rectangular 5x5 km slip patches, power law n=3 "
                               "below a 25 km lid, adaptive RK2" % (args.workload, sx1, sx2, sx3,
                                                                    workloads.spacing(spec)[0]),
                   "grid": [sx1, sx2, sx3], "dt_last": info["Dt"], "t_last": info["t"], "autotune": tuned},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get(dom) if traffic else None},
        # frac: the unfused model bytes of the path (404 N + ...) over the step
        # time, bytes this implementation never moves; fused_frac: the bytes
        # the fused kernels must move
        "step_roofline": {"bytes_per_step_model": B_step, "frac": round(B_step * steps_per_s / 8e12, 4),
                          "fused_bytes_per_step": fused_bytes,
                          "fused_frac": round(fused_bytes * steps_per_s / 8e12, 4)},
        "kernels": kernels,
    }
    if why:
        out["roofline"]["traffic_note"] = why
    state = model.download() if not args.no_cpu_baseline else None
    model.close()
    if not args.no_batched:
        print("# batched: 64 x 128^3 models, 10 steps each", file=sys.stderr, flush=True)
        try:
            out["batched"] = batched_c3()
        except Exception as e:  # keep the headline line
            out["batched"] = {"error": repr(e)}
    if not args.no_serving:
        print("# serving stream", file=sys.stderr, flush=True)
        try:
            out["serving"] = serving_c4()
        except Exception as e:
            out["serving"] = {"error": repr(e)}
    if state is not None:
        # the host threads this process may use: the box's CPU share for
        # its GPU (OMP_NUM_THREADS), not every core the host reports
        nproc = os.cpu_count() or 1
        nthreads = int(os.environ.get("OMP_NUM_THREADS", nproc))
        nsteps_cpu = 2
        print("# cpu baseline: oracle, %d RK2 steps of the same state, %d threads" % (nsteps_cpu, nthreads),
              file=sys.stderr, flush=True)
        dt = cpu_baseline(spec, state, nthreads, nsteps_cpu)
        out["cpu_baseline"] = {"value": round(nsteps_cpu / dt, 5), "unit": "RK2 steps/s", "cores": nthreads,
                               "kind": "port",
                               "sample": "%d RK2 steps of the %s state (%dx%dx%d) in %.1f s on %d OpenMP threads "
                                         "(host reports %d cores; OMP_NUM_THREADS is this GPU's CPU share)" % (
                                             nsteps_cpu, args.workload, sx1, sx2, sx3, dt, nthreads, nproc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
