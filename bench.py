"""Benchmark of the spectral path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (the metric's configuration): seeded 200-knot deck with attenuation
and full self-gravitation, lmax = 600, fmin = 0.1 mHz, fmax = 40 mHz, 360 min
series at dt = 1 s (nt = 32768), filters 0.2/0.5/35/38 mHz, one source at
600 km depth with a seeded moment tensor, 200 seeded receivers at the
surface: 787,310 (l, w) pairs.

One step = one full forward: every (l, w) solve, the Legendre receiver sum
and the filtered inverse FFT to seismograms, with the model, run tables and
all buffers resident in HBM before timing starts.  Every library call is
synchronous (the context's HIP stream is synchronised before it returns), so
the wall clock around K steps brackets exactly K steps.

Prints one JSON line.  value = (l, w) solves/s of the full forward (pairs per
step / seconds per step); seismograms/s = receivers per step / seconds per
step.  "batched": the 64-event workload (depth U[0,700] km, fmax U[10,40]
mHz, lmax = 15 fmax, 16 receivers each, 180 min at 1 s; 10.3 M pairs) as
full forwards through the plain path (events one after another) and the
batched path (all events per launch, cost-sorted work, HIP graph), after
yspec_amd/tune.py has picked the batched launch shape.  roofline: the dominant kernel's algorithmic FP64 FLOPs (exact step
counts x the per-step constants of yspec_amd/flops.py) over its HIP-event
time.  cpu_baseline: the reference Fortran (compiled from the reference's
own sources into oracle/_ref, run through its FP64 driver) timed on this
host on a stride sample of degrees, all frequencies and receivers, one core;
cpu_baseline_all_cores: the same program as one process per core (at most
16) on a finer degree sample split between them, pairs over wall time.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from yspec_amd import api  # noqa: E402
from yspec_amd.flops import ode_flops, sum_flops, fft_bytes  # noqa: E402
from yspec_amd.synth import (batch_events, bump_deck, inversion_events, synth_deck,  # noqa: E402
                             workload)
from yspec_amd.tune import tune  # noqa: E402

METRIC = "(l,ω) solves/s + seismograms/s, batched vs plain GPU; gradient/forward time"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the ODE kernels at this workload: rocprofv3
# FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (factors from a 1 GiB device copy),
# profiles/forward_cfg2_hbm_traffic.md
TRAFFIC_BYTES = {"bvec": 1429.7e6, "minors": 1560.2e6, "tor": 200.8e6, "lsum": 483.2e6}
PHASES = ("source", "tor", "minors", "bvec", "rad", "assemble", "lsum", "ifft")


def brow_rows(p):
    """receiver-side rows of the b-vector row march for parameters p"""
    return 3 if (p.out == 2 and p.cor != 0) else 2


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ref_dump")


def _ref_inputs(td, deck):
    with open(os.path.join(td, "synth.200"), "w") as f:
        f.write(deck.to_text())
    q = workload("cfg2", model_file="synth.200")
    with open(os.path.join(td, "in.txt"), "w") as f:
        f.write(q.to_text())


def _parse_times(out):
    rows = [line.split() for line in out.splitlines() if line.startswith("l ")]
    return int(sum(int(r[3]) for r in rows)), sum(float(r[-1]) for r in rows)


def cpu_baseline(deck, p, sample):
    """Reference yspec (oracle/_ref/ref_dump --time) on degrees `sample`,
    every frequency and receiver, one core.  Returns (pairs/s, pairs, s)."""
    if not os.path.exists(REF_EXE):
        return None
    with tempfile.TemporaryDirectory() as td:
        _ref_inputs(td, deck)
        t0 = time.perf_counter()
        out = subprocess.run([REF_EXE, "in.txt", "--time", *map(str, sample)], cwd=td, check=True,
                             capture_output=True, text=True).stdout
        wall = time.perf_counter() - t0
    pairs, secs = _parse_times(out)
    return pairs / secs, pairs, secs, wall


def cpu_baseline_all_cores(deck, sample, nproc):
    """The same program as `nproc` processes at once, degrees of `sample`
    dealt out in turn; pairs over the wall time of the whole run."""
    if not os.path.exists(REF_EXE):
        return None
    with tempfile.TemporaryDirectory() as td:
        _ref_inputs(td, deck)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([REF_EXE, "in.txt", "--time", *map(str, sample[k::nproc])],
                                  cwd=td, stdout=subprocess.PIPE, text=True)
                 for k in range(nproc) if sample[k::nproc]]
        outs = [pr.communicate()[0] for pr in procs]
        wall = time.perf_counter() - t0
        if any(pr.returncode for pr in procs):
            return None
    pairs = sum(_parse_times(o)[0] for o in outs)
    return pairs / wall, pairs, wall, len(procs)


def cpu_model():
    """the host CPU's model name"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


REF_YSPEC = os.path.join(ROOT, "oracle", "_ref", "yspec")


def cfg1_compare(deck, steps, warmup, cpu=True):
    """The lmax = 150, fmax = 10 mHz workload (60 min at 10 s, 100 km
    source, 10 receivers, no attenuation) in full: the reference program
    (oracle/_ref/yspec, one core, its own parameter file and deck, output
    files included) against full GPU forwards."""
    p = workload("cfg1")
    g = api.YSpec(deck, p, device=0)
    for _ in range(warmup):
        g.forward(copy=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.forward(copy=False)
    sec = (time.perf_counter() - t0) / steps
    out = {"workload": "lmax=150, fmax=10 mHz, 60 min, dt=10 s, 100 km source, 10 receivers, "
                       "self-gravitation, no attenuation",
           "pairs": g.npair, "gpu_ms_per_forward": sec * 1e3,
           "gpu_solves_per_s": g.npair / sec, "gpu_seismograms_per_s": p.nr / sec}
    g.close()
    if cpu and os.path.exists(REF_YSPEC):
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "synth.200"), "w") as f:
                f.write(deck.to_text())
            with open(os.path.join(td, "in.txt"), "w") as f:
                f.write(workload("cfg1", model_file="synth.200").to_text())
            t0 = time.perf_counter()
            subprocess.run([REF_YSPEC, "in.txt"], cwd=td, check=True, capture_output=True)
            wall = time.perf_counter() - t0
        out["cpu_reference_s"] = wall
        out["cpu_solves_per_s"] = out["pairs"] / wall
        out["cpu_seismograms_per_s"] = p.nr / wall
        out["cpu_cores"] = 1
        out["cpu_note"] = ("reference yspec (amdflang -O3) run once in full on one core, wall "
                           "time including its parameter-file, deck and output-file I/O")
    return out


def batched_compare(deck, steps, warmup, do_tune=True):
    """The 64-event workload through the plain path (events one after
    another, each with its own launches) and the batched path (all events
    per launch from cost-sorted work lists, replayed from a HIP graph),
    after the autotuner has picked the batched launch shape."""
    ev = batch_events()
    g = api.YSpec(deck)
    g.set_events(ev)
    P, nrec = g.npair, sum(p.nr for p in ev)
    print(f"# workload 64 events: {P} (l,w) pairs, {nrec} receivers", file=sys.stderr, flush=True)
    tuning = tune(g, log=lambda m: print("# " + m, file=sys.stderr, flush=True)) if do_tune \
        else None
    res = {}
    for mode, name in ((0, "plain"), (1, "batched")):
        # the plain path with its default lanes per radial solve, the
        # batched path with the tuned ones
        g.set_lanes(tuning["best"]["rad_lanes"] if (tuning and mode == 1) else 0)
        for _ in range(warmup):
            g.forward_events(mode)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.forward_events(mode)
        sec = (time.perf_counter() - t0) / steps
        print(f"# {name}: {sec * 1e3:.1f} ms per forward of all events", file=sys.stderr,
              flush=True)
        res[name] = {"ms_per_step": sec * 1e3, "solves_per_s": P / sec,
                     "seismograms_per_s": nrec / sec}
    # per-event moment-tensor gradients after the batched forward (the
    # records are that forward's): kernel time of the three adjoint stages
    rng = np.random.default_rng(11)
    sbs = [rng.standard_normal((g.info[e]["nt"], 3, p.nr)) for e, p in enumerate(ev)]
    g.forward_events(1)
    g.events_gradient_mt(sbs)
    grad_ms = float(g.vjp_timings().sum())
    print(f"# moment-tensor gradients of all events: {grad_ms:.2f} ms of kernels",
          file=sys.stderr, flush=True)
    g.close()
    return {
        "workload": "64 events on the same deck: depth U[0,700] km, fmax U[10,40] mHz, "
                    "lmax = 15 fmax, 16 receivers each, 180 min at dt = 1 s, attenuation + "
                    "self-gravitation",
        "events": len(ev), "pairs": P, "receivers": nrec,
        "plain": res["plain"], "batched": res["batched"],
        "speedup": res["plain"]["ms_per_step"] / res["batched"]["ms_per_step"],
        "tuning": tuning["best"] if tuning else None,
        "moment_tensor_gradients": {
            "kernel_ms": grad_ms,
            "over_batched_forward_ms": grad_ms / res["batched"]["ms_per_step"],
            "note": "per-event d/dM of sum(seis_bar_e * seismograms_e) from the batched "
                    "forward's records; seis_bar upload excluded",
        },
    }


def gradient_compare(deck, steps, warmup):
    """The misfit-gradient workload (lmax 300, 20 mHz, 120 min at 2 s, 50 km
    source, 50 receivers, attenuation + self-gravitation; data from the deck
    with 2% Gaussian bumps in vp and vs): wall time of one forward against
    one ysx_misfit_grad (its own forward, the residual, every adjoint stage
    and the deck -> table adjoint, the 800 knot values' gradient back on the
    host), both from inputs resident on the device."""
    p = workload("cfg4")
    g = api.YSpec(bump_deck(deck), p, device=0)
    obs = g.forward()
    g.close()
    g = api.YSpec(deck, p, device=0)
    # the observed data in the library's layout, packed once as an
    # inversion's iterations would reuse them
    obs = g.pack_obs(obs)
    for _ in range(warmup):
        g.forward(copy=False)
        g.misfit_grad(obs)
    tf, tg = [], []
    fwd_k = np.zeros(8)
    vjp_k = np.zeros(3)
    grad_k = np.zeros(7)
    for k in range(steps):
        t0 = time.perf_counter()
        g.forward(copy=False)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        g.misfit_grad(obs)
        tg.append(time.perf_counter() - t0)
        fwd_k += g.timings()
        vjp_k += g.vjp_timings()
        grad_k += g.grad_timings()
        print(f"# gradient step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    fwd_k /= steps
    vjp_k /= steps
    grad_k /= steps
    out = {
        "workload": "lmax=300, fmax=20 mHz, 120 min, dt=2 s, 50 km source, 50 receivers, "
                    "attenuation + self-gravitation; 800 knot parameters",
        "pairs": g.npair,
        "forward_ms": 1e3 * float(np.median(tf)),
        "gradient_ms": 1e3 * float(np.median(tg)),
        "gradient_over_forward": float(np.median(tg) / np.median(tf)),
        "forward_kernel_ms": {k: round(float(v), 4) for k, v in zip(PHASES, fwd_k)},
        "adjoint_kernel_ms": {
            "seismograms": float(vjp_k[0]), "receiver_sum_and_assembly": float(vjp_k[1]),
            "bvec_sweep": float(grad_k[0]), "minor_sweep": float(grad_k[1]),
            "tor_sweep": float(grad_k[2]), "starting_solutions": float(grad_k[3]),
            "radial": float(grad_k[4]), "source": float(grad_k[5])},
        "gradient_device_bytes": dict(zip(("per_step_states", "step_scratch_and_radial_tape",
                                           "per_pair_adjoints"),
                                          map(float, g.grad_memory()))),
        "note": "gradient_ms: one ysx_misfit_grad (forward included; the observed data "
                "packed once beforehand, as an inversion reuses them) over one forward's "
                "wall time, medians of the steps; per_step_states include the minor and "
                "toroidal states the single-event forward stores for the sweeps",
    }
    g.close()
    return out


def inversion_compare(deck, steps, warmup):
    """The batched inversion step: 32 events (depths U[0,700] km, 100
    receivers each, lmax 600, 40 mHz, 180 min at 1 s, attenuation +
    self-gravitation; data from the deck with 2% Gaussian bumps in vp and
    vs): one batched forward launch against one ysx_misfit_grad over all
    events (its own batched forward, the residuals, the adjoint sweeps as
    one persistent launch each over every event's pairs, the per-event
    gradients summed on the device)."""
    events = inversion_events(32)
    g = api.YSpec(bump_deck(deck), device=0)
    g.set_events(events)
    g.forward_events(1)
    obs = [g.event_seismograms(e)[:g.info[e]["nout"]] for e in range(len(events))]
    g.close()
    g = api.YSpec(deck, device=0)
    g.set_events(events)
    obs = g.pack_obs(obs)
    for _ in range(warmup):
        g.forward_events(1)
        g.misfit_grad(obs)
    tf, tg, gk = [], [], np.zeros(7)
    for k in range(steps):
        t0 = time.perf_counter()
        g.forward_events(1)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        g.misfit_grad(obs)
        tg.append(time.perf_counter() - t0)
        gk += g.grad_timings()
        print(f"# inversion step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    gk /= steps
    mem = g.grad_memory()
    out = {
        "workload": "32 events, depths U[0,700] km, 100 receivers each, lmax=600, "
                    "fmax=40 mHz, 180 min, dt=1 s, attenuation + self-gravitation",
        "pairs": g.npair,
        "forward_batched_ms": 1e3 * float(np.median(tf)),
        "gradient_ms": 1e3 * float(np.median(tg)),
        "gradient_over_forward": float(np.median(tg) / np.median(tf)),
        "adjoint_kernel_ms": {"bvec_sweep": float(gk[0]), "minor_sweep": float(gk[1]),
                              "tor_sweep": float(gk[2]), "starting_solutions": float(gk[3]),
                              "radial": float(gk[4]), "source": float(gk[5])},
        "gradient_device_bytes": {"per_step_states": float(mem[0]),
                                  "step_scratch_and_radial_tape": float(mem[1]),
                                  "per_pair_adjoints": float(mem[2])},
        "note": "gradient_ms: one ysx_misfit_grad over all events (batched forward "
                "included; the observed data packed once beforehand); per-step states and "
                "step scratch are per persistent block, so they do not grow with the number "
                "of pairs",
    }
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--lmax", type=int, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-event plain vs batched comparison")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the misfit-gradient workload")
    ap.add_argument("--gradient-only", action="store_true",
                    help="only the misfit-gradient workload, for profiling")
    ap.add_argument("--no-inversion", action="store_true",
                    help="skip the 32-event batched forward + adjoint workload")
    ap.add_argument("--inversion-only", action="store_true",
                    help="only the 32-event batched forward + adjoint workload")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the 64-event comparison (no tuning), for profiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the work would shard by frequency band; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    deck = synth_deck(0)
    if args.gradient_only:
        print(json.dumps(gradient_compare(deck, args.steps, args.warmup)), flush=True)
        return
    if args.inversion_only:
        print(json.dumps(inversion_compare(deck, args.steps, args.warmup)), flush=True)
        return
    if args.batched_only:
        print(json.dumps(batched_compare(deck, args.steps, args.warmup, do_tune=False)),
              flush=True)
        return
    p = workload("cfg2") if args.lmax is None else workload("cfg2", lmax=args.lmax)
    g = api.YSpec(deck, p, device=0)
    P, nr = g.npair, p.nr
    print(f"# workload cfg2: {P} (l,w) pairs, {nr} receivers, nt={g.nt}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        g.forward(copy=False)
    phase = np.zeros(8)
    t0 = time.perf_counter()
    for k in range(args.steps):
        g.forward(copy=False)
        phase += g.timings()
        print(f"# step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    # the same forward handing the (nt, 3, nr) seismograms back to host
    # memory, as the Fortran host receives them (outside the timed steps):
    # the copy alone into an array the host keeps (touched once, as a
    # program reusing its output array), and a forward with the copy into a
    # new array
    host = np.zeros((g.nt, 3, nr), order="F")
    g.event_seismograms(0, out=host)
    t2 = time.perf_counter()
    g.event_seismograms(0, out=host)
    copy_ms = (time.perf_counter() - t2) * 1e3
    t2 = time.perf_counter()
    g.forward(copy=True)
    pcie_ms = (time.perf_counter() - t2) * 1e3
    # gradient of sum(seis_bar * seismograms) with respect to the moment
    # tensor through the three adjoint stages, after one more forward
    # (outside the timed steps)
    sb = np.asfortranarray(np.random.default_rng(7).standard_normal((g.nt, 3, nr)))
    g.forward(copy=False)
    t3 = time.perf_counter()
    g.misfit_gradient_mt(sb)
    mt_wall = (time.perf_counter() - t3) * 1e3
    mt_k = g.vjp_timings()
    phase /= args.steps
    counts = g.step_counts()
    fl = ode_flops(counts, brows=brow_rows(p))
    work = {"tor": fl["tor"], "minors": fl["minors"], "bvec": fl["bvec"], "rad": fl["rad"],
            "lsum": sum_flops(g.nw, p.lmax, nr)}
    ms = dict(zip(PHASES, phase))
    # the dominant kernel of the forward's own stream (the radial solves run
    # beside it on a second stream)
    dom = max(("tor", "minors", "bvec", "lsum"), key=lambda k: ms[k])
    achieved = work[dom] / (ms[dom] * 1e-3) / 1e12
    ode_ms = ms["tor"] + ms["minors"] + ms["bvec"] + ms["rad"]
    ode_flop = fl["tor"] + fl["minors"] + fl["bvec"] + fl["rad"]
    line = {
        "metric": METRIC,
        "value": P / sec,
        "unit": "(l,ω) solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 200-knot deck, seeded moment tensor and receivers",
        "config": {"workload": "cfg2 plain path: lmax=600, fmax=40 mHz, 360 min, dt=1 s, "
                               "600 km source, 200 receivers, attenuation + self-gravitation",
                   "pairs": P, "receivers": nr, "nt": g.nt, "nw": g.nw},
        "seismograms_per_s": nr / sec,
        "ode_solves_per_s": P / (ode_ms * 1e-3),
        "phase_ms": {k: round(float(v), 4) for k, v in ms.items()},
        "ode_tflops": ode_flop / (ode_ms * 1e-3) / 1e12,
        # solves/s as a fraction of the FP64 roofline: ODE FLOP rate / peak
        "ode_roofline_frac": ode_flop / (ode_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        "step_counts": {k: int(v) for k, v in zip(("tor", "solid", "fluid", "bvec", "rad"),
                                                  counts)},
        "roofline": {"bound": "fp64_vector", "kernel": dom, "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": TRAFFIC_BYTES.get(dom) if args.lmax is None else None,
                     "note": "FP64 vector work; MI355X FP64 vector and matrix peaks are both "
                             "78.6 TFLOP/s; traffic = HBM bytes per launch from rocprofv3 "
                             "FETCH_SIZE/WRITE_SIZE scaled by a 1 GiB copy"},
        "host_copy": {"copy_ms": copy_ms, "forward_with_copy_ms": pcie_ms,
                      "bytes": g.nt * 3 * nr * 8,
                      "note": "device-to-host copy of the (nt, 3, nr) seismograms into an array "
                              "the host keeps; forward_with_copy_ms: one forward plus the copy "
                              "into a newly allocated array"},
        # the metric's third part: the misfit gradient with respect to the
        # knot values at its own workload (filled in below)
        "gradient_over_forward": None,
        "moment_tensor_gradient": {
            "kernel_ms": float(mt_k.sum()),
            "stage_kernel_ms": {"seismograms": float(mt_k[0]),
                                "receiver_sum_and_assembly": float(mt_k[1]),
                                "source": float(mt_k[2])},
            "kernel_ms_over_forward_ms": float(mt_k.sum()) / (sec * 1e3),
            "wall_ms": mt_wall,
            "note": "d/dM of sum(seis_bar * seismograms) given the forward's records; "
                    "wall_ms includes the upload of the (nt, 3, nr) seis_bar from host memory",
        },
    }
    g.close()
    if not args.no_gradient and args.lmax is None:
        gr = gradient_compare(deck, args.steps, args.warmup)
        line["gradient_over_forward"] = gr["gradient_over_forward"]
        line["gradient"] = gr
    if not args.no_inversion and args.lmax is None:
        line["inversion"] = inversion_compare(deck, 1, 1)
    if not args.no_batched and args.lmax is None:
        line["batched"] = batched_compare(deck, args.steps, args.warmup)
    if args.lmax is None:
        line["cfg1"] = cfg1_compare(deck, args.steps, args.warmup, cpu=not args.no_cpu)
    if not args.no_cpu:
        sample = list(range(0, p.lmax + 1, 30))
        cb = cpu_baseline(deck, p, sample)
        if cb:
            rate, pairs, secs, wall = cb
            line["cpu_baseline"] = {
                "value": rate, "unit": "(l,ω) solves/s", "cores": 1, "kind": "reference",
                "sample": f"reference yspec (amdflang -O3) on l in {sample}, all {g.nw} w, "
                          f"{nr} receivers: {pairs} pairs in {secs:.1f} s",
                "cpu_model": cpu_model(),
                # the full forward's solves at this rate (the sum over l and
                # the IFFT take well under 1% of the reference's time)
                "seismograms_per_s_extrapolated": nr / (P / rate),
            }
        nproc = max(1, min(16, os.cpu_count() or 1))
        sample_all = list(range(0, p.lmax + 1, 5))
        ca = cpu_baseline_all_cores(deck, sample_all, nproc)
        if ca:
            rate, pairs, wall, used = ca
            line["cpu_baseline_all_cores"] = {
                "value": rate, "unit": "(l,ω) solves/s", "cores": used, "kind": "reference",
                "sample": f"reference yspec (amdflang -O3), {used} processes, l = 0, 5, ..., "
                          f"{p.lmax} dealt out in turn, all {g.nw} w, {nr} receivers: "
                          f"{pairs} pairs in {wall:.1f} s wall",
            }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
