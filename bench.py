#!/usr/bin/env python3
"""Throughput of RELAX's RK2 elastic + relaxation step on one MI355X.

A step is one predictor/corrector pass of the RK2 loop (relax.f90:555-890)
over the 512x512x256 workload (40 slip patches, power law n = 3 below a 25 km
lid; relax_amd.workloads "c1"), with the whole state resident in HBM.

  python bench.py [--steps K] [--warmup W] [--workload c1] [--no-cpu-baseline]

Prints one JSON line.  The timed region is K back-to-back steps on the
model's stream with HIP-event phase timers (one synchronisation at the end),
bracketed by device synchronisations; `roofline` reports the dominant
kernel's algorithmic bytes per launch over its event-measured average
duration.  `cpu_baseline` times the oracle (the C++ restatement of the
reference, OpenMP on the host cores) on one step of the same state.
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
    return {
        "bodyforce": 4 * (6 * H + 3 * H) + 12 * sx1 * sx2,   # read sig, write body force + traction
        "fft_x_fwd": 4 * (3 * H + 3 * H) + nq,                # read real rows, write half spectrum
        "fft_y": 4 * (3 * H + 3 * H),                         # read + write the half spectrum
        "green_z": 4 * (3 * H + 3 * H) + 2 * nq,              # read + write spectrum and Nyquist planes
        "fft_x_inv": 4 * (3 * H + 3 * H + 3 * H) + nq,        # read spectrum + u, write v or u
        "stress_pred": 4 * (3 * H + 6 * H + 6 * H + 6 * H),   # read v, sig, tau; write sig
        "stress_corr": 4 * (3 * H + 6 * H + 6 * H + H + 6 * H + 6 * H + H),  # + tau, gamma in/out
    }


def step_bytes_model(sx1, sx2, sx3):
    """Compulsory bytes of one RK2 step in the unfused accounting of the
    path (404 N + 192 N/sx1 + 96 sx1 sx2, N = sx1 sx2 sx3)."""
    N = sx1 * sx2 * sx3
    return 404 * N + 192 * N / sx1 + 96 * sx1 * sx2


def cpu_baseline(spec, state, nthreads):
    """The oracle (oracle/, FP32 storage) on one RK2 step of the GPU's state."""
    import numpy as np

    ROOT = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, ROOT)
    os.environ.setdefault("OMP_NUM_THREADS", str(nthreads))
    import oracle as O
    from relax_amd import workloads

    t = workloads.tables(spec)
    m = dict(sx1=spec["sx1"], sx2=spec["sx2"], sx3=spec["sx3"], dx1=spec["dx"], dx2=spec["dx"], dx3=spec["dx"],
             lam=workloads.LAM, mu=workloads.MU, gam=workloads.GAM, beta=workloads.BETA, linear=t["linear"],
             nonlinear=t["nonlinear"], faultcreep=t["friction"], prestress=t["prestress"], interval=1e30,
             odt=1.0, tscale=1.0, skip=1, event_times=[0.0])
    st = {k: state[k] for k in ("u1", "u2", "u3", "tau", "sig", "gamma")}
    st.update(t=0.0, oi=1, i=0)
    t0 = time.perf_counter()
    O.rk2_run(m, st, 1)
    dt = time.perf_counter() - t0
    return dt


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
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
                               "below a 25 km lid, adaptive RK2" % (args.workload, sx1, sx2, sx3, spec["dx"]),
                   "grid": [sx1, sx2, sx3], "dt_last": info["Dt"], "t_last": info["t"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "step_roofline": {"bytes_per_step_model": B_step, "frac": round(B_step * steps_per_s / 8e12, 4),
                          "fused_bytes_per_step": fused_bytes,
                          "fused_frac": round(fused_bytes * steps_per_s / 8e12, 4)},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        nthreads = min(16, os.cpu_count() or 1)
        print("# cpu baseline: oracle, one RK2 step of the same state, %d threads" % nthreads, file=sys.stderr,
              flush=True)
        state = model.download()
        model.close()
        dt = cpu_baseline(spec, state, nthreads)
        out["cpu_baseline"] = {"value": round(1.0 / dt, 5), "unit": "RK2 steps/s", "cores": nthreads, "kind": "port",
                               "sample": "one RK2 step of the %s state (%dx%dx%d), %.1f s" % (
                                   args.workload, sx1, sx2, sx3, dt)}
    else:
        model.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
