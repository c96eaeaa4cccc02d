#!/usr/bin/env python3
"""Throughput of the LISA time-domain response on one MI355X: mixed precision vs FP64.

Workload: one galactic-binary source over one year, N = 3,153,600 samples at dt = 10 s
(f0 = 5 mHz, fdot = 1e-16), order-25 Lagrange stencils, 2nd-generation TDI, A/E/T output,
analytic equal-arm cartwheel orbits tabulated every 1e4 s. A step is one pass of the response
over the source with its input already resident in HBM; `value` is AET samples/s of the
mixed-precision form (--precision, default mixed-c128): one fused projection + TDI launch (the
six projections stay in LDS), times, delays and stencil arithmetic in FP64, composite filters
for the equal-arm TDI chains, float32 A, E, T out. The default reads the complex128 input, the
one format whose error stays within 1e-6 (rel-L2) on all three channels; the complex64
(h / max|h|) and 32-bit fixed-point inputs are timed beside it with their errors. The launch
reads a few percent of HBM bandwidth, so the input format hardly changes the speed.

Also on the same JSON line:
  fp64          the plain FP64 drop-in (separate projection and TDI launches, exact stencils,
                bitwise equal to the CPU oracle) timed in the same run, and mixed / fp64;
                fp64_fast and fp64_fused, the FP64 forms with division-free stencils (drop-in,
                and one fused launch), and mixed / each
  mixed_q31,    the same fused launch with a 32-bit fixed-point or a complex128 input (float32
  mixed_c128    outputs): less or no storage error at the same arithmetic
  roofline      the fused mixed-precision kernel's algorithmic FP64 rate (stencil
                multiply-adds, 3744 FLOP per 2nd-generation sample) against the 78.6 TFLOP/s
                FP64 vector peak, its launch time from HIP events in the timed region, and HBM
                bytes per launch from the committed rocprofv3 counter summary
  batch         the 512-source batch (durations U[0.05, 1] yr): FP64 exact vs autotuned
                mixed-precision samples/s (complex64, fixed-point and complex128 inputs), and
                each form's per-source rel-L2 against the FP64 exact batch
  latency       p50/p99 of a short open-loop serving run: Poisson arrivals at 200 req/s of
                1..64-source requests (2^18/2^20/2^22 samples, dt 5/10 s, 30/70 1st/2nd gen);
                latency_curve: the same traffic at 100, 150 and 200 req/s
  parity        rel-L2 / max-abs of A, E, T against the FP64 oracle on the full workload, and
                every form's rel-L2 against the oracle built with extended-precision values
                (rel_l2_AET_vs_wide; fp64_oracle: the FP64 oracle's own, its rounding floor)
  cpu_baseline  the FP64 CPU oracle (1 thread, the reference CPU path's serial arithmetic) on
                a bounded prefix of the same workload; `all_cores`: the same oracle with OpenMP
                over samples (up to 16 threads, the host share of one GPU) on the whole source
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "lisa-on-gpu_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

YRSID_SI = 31558149.763545603
METRIC = "TDI AET samples/s, 1-yr order-25 2nd-gen (mixed vs FP64 GPU); p50/p99 latency"
N_SAMPLES = 3_153_600
DT = 10.0
ORDER = 25
T_BUFFER = 10000.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
TAPS = ORDER + 1
# Algorithmic FLOPs per output sample: 2 per stencil tap. Projection: 6 links x (receiver,
# emitter) x (h+, hx); TDI: one real stencil per unit (24 or 48 units).
FLOP_PROJ = 2 * TAPS * 6 * 2 * 2
FLOP_TDI = {"1st generation": 2 * TAPS * 24, "2nd generation": 2 * TAPS * 48}
# Algorithmic HBM bytes per sample: h in, A, E, T out (y stays on chip in the fused kernel);
# the FP64 drop-in also writes and reads the six projections.
BYTES_IO = {"mixed": 8 + 12, "mixed-q31": 8 + 12, "mixed-c128": 16 + 12}
FORMS = {"mixed": "complex64 input h / max|h|, float32 A, E, T",
         "mixed-q31": "32-bit fixed-point input h / max(|Re h|, |Im h|), float32 A, E, T",
         "mixed-c128": "complex128 input (no storage error), float32 A, E, T"}
BYTES_FP64 = 16 + 48 + 48 + 24


def gb_waveform(n, dt, A=1e-21, f=5e-3, fdot=1e-16, iota=0.8, phi0=0.3, psi=0.5):
    import numpy as np
    t = np.arange(n) * dt
    fddot = 11.0 / 3.0 * fdot ** 2 / f
    phase = 2 * np.pi * (f * t + 0.5 * fdot * t ** 2 + fddot * t ** 3 / 6.0) - phi0
    ci = np.cos(iota)
    hSp = -np.cos(phase) * A * (1.0 + ci * ci)
    hSc = -np.sin(phase) * 2.0 * A * ci
    hp = hSp * np.cos(2 * psi) - hSc * np.sin(2 * psi)
    hc = hSp * np.sin(2 * psi) + hSc * np.cos(2 * psi)
    return hp + 1j * hc


def pmc_traffic(name: str, variant: str = ""):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (FETCH_SIZE and
    WRITE_SIZE, calibrated against a copy kernel): the HBM_BYTES_PER_LAUNCH line (of
    `variant`, when the summary has one per variant), or None if there is none."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    key = r"HBM_BYTES_PER_LAUNCH" + (r"\[" + re.escape(variant) + r"\]" if variant else "")
    m = re.search(r"^" + key + r"\s+([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(orbits, h, lam, beta, tdi, n_sample, max_cores=16):
    """Oracle projection + TDI + AET on the first n_sample samples, one thread (the reference
    CPU path is serial); beside it, the same oracle over samples on up to max_cores threads
    on the whole source."""
    from oracle import oracle as orc
    oo = orc.OracleOrbits(*orbits.pycppdetector_args)
    import numpy as np
    hs = np.ascontiguousarray(h[:n_sample])
    t0 = time.perf_counter()
    orc.response_pipeline(oo, hs, lam, beta, DT, ORDER, tdi, x_base=orbits.x_base, nthreads=1)
    el = time.perf_counter() - t0
    out = {"value": n_sample / el, "unit": "samples/s", "cores": 1, "kind": "port",
           "sample": f"first {n_sample} samples of the same source, projection + {tdi} TDI + AET "
                     f"({el:.1f} s, FP64 oracle restating the reference CPU path)"}
    threads = max(1, min(max_cores, orc.max_threads(), os.cpu_count() or 1))
    if threads > 1:
        hf = np.ascontiguousarray(h)
        t0 = time.perf_counter()
        orc.response_pipeline(oo, hf, lam, beta, DT, ORDER, tdi, x_base=orbits.x_base,
                              nthreads=threads)
        el = time.perf_counter() - t0
        out["all_cores"] = {"value": len(hf) / el, "threads": threads,
                            "sample": f"the whole {len(hf)}-sample source, OpenMP over samples "
                                      f"({el:.1f} s)"}
    return out


def timed(fn, steps, warmup):
    """Warm up, then time `steps` calls with HIP events on the current stream and the wall
    clock (device synchronised on both sides). Returns (wall s, mean event ms per call)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, sum(a.elapsed_time(b) for a, b in ev) / steps


def batch_run(num_sources, steps, seed=3):
    """The 512-source batch: durations U[0.05, 1] yr at dt = 10 s, sky uniform on the sphere,
    psi, phi0 U[0, 2 pi), cos i U[-1, 1], f0 log-U[1e-4, 1e-2] Hz, fdot log-U[1e-19, 1e-15] Hz/s,
    log10 A U[-23, -21] (numpy default_rng(3)); 2nd-gen AET, order 25. Times the FP64 exact
    batch (two persistent launches) and the autotuned mixed-precision fused batch; reports the
    mixed form's rel-L2 per source against the FP64 exact result (bitwise the oracle's)."""
    import numpy as np
    import torch
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    rng = np.random.default_rng(seed)
    orb = EqualArmlengthOrbits(t_max=1.2 * YRSID_SI)
    orb.configure(linear_interp_setup=True)
    lengths = (rng.uniform(0.05, 1.0, num_sources) * YRSID_SI / DT).astype(np.int64)
    lam = rng.uniform(0, 2 * np.pi, num_sources)
    beta = np.arcsin(rng.uniform(-1, 1, num_sources))
    psi = rng.uniform(0, 2 * np.pi, num_sources)
    phi0 = rng.uniform(0, 2 * np.pi, num_sources)
    ci = rng.uniform(-1, 1, num_sources)
    f0 = 10 ** rng.uniform(-4, -2, num_sources)
    fd = 10 ** rng.uniform(-19, -15, num_sources)
    amp = 10 ** rng.uniform(-23, -21, num_sources)
    total = int(lengths.sum())
    # the reference test's galactic-binary generator, on the device (waveforms.generate_gb)
    from fastlisaresponse_amd.waveforms import generate_gb
    gb = np.stack([amp, f0, fd, np.arccos(ci), phi0, psi], axis=1)
    h, _ = generate_gb(gb, lengths, DT, fmt="complex128")
    torch.cuda.synchronize()
    res = {"sources": num_sources, "samples": total, "seed": seed,
           "workload": "512 galactic binaries, durations U[0.05,1] yr at dt=10 s, 2nd-gen AET"}

    def time_plan(plan):
        return timed(plan.run, steps, 1)

    ref = BatchedResponse(1 / DT, tdi="2nd generation", orbits=orb, precision="fp64")
    p64 = ref.plan(h, lengths, lam, beta)
    res["fp64_tune"] = p64.tune(use_cache=False, reps=1)["best"]
    wall, ms = time_plan(p64)
    p64.check_errors()
    res["fp64_exact_samples_per_s"] = total * steps / wall
    out64 = p64.out_all.view(-1)
    del p64

    T = int(T_BUFFER / DT)

    def rel_l2(out):
        """Per-source rel-L2 of a float32 batch output against the FP64 exact batch."""
        rel = np.zeros((num_sources, 3))
        o = 0
        for s in range(num_sources):
            n = int(lengths[s])
            a = out64[3 * o: 3 * (o + n)].view(3, n)[:, T:n - T]
            b = out[3 * o: 3 * (o + n)].view(3, n)[:, T:n - T].double()
            rel[s] = ((b - a).norm(dim=1) / a.norm(dim=1)).cpu().numpy()
            o += n
        return rel

    def summary(rel):
        return {ch: {"median": float(np.median(rel[:, c])),
                     "p90": float(np.percentile(rel[:, c], 90)), "max": float(rel[:, c].max()),
                     "sources_within_1e-6": int((rel[:, c] <= 1e-6).sum())}
                for c, ch in enumerate("AET")}

    from fastlisaresponse_amd.batch import pack_q31
    h32, scales = pack_mixed(h, lengths)
    hq, qscales = pack_q31(h, lengths)  # both forms quantise the same FP64 waveforms
    mb = BatchedResponse(1 / DT, tdi="2nd generation", orbits=orb, precision="mixed",
                         storage_check="off")
    pm = mb.plan(h32, lengths, lam, beta, scales=scales)
    tune = pm.tune(use_cache=False, reps=2)
    res["mixed_tune"] = {"best": tune["best"], "ms": tune["ms"]}
    wall, ms = time_plan(pm)
    pm.check_errors()
    res["mixed_samples_per_s"] = total * steps / wall
    res["mixed_vs_fp64"] = res["mixed_samples_per_s"] / res["fp64_exact_samples_per_s"]
    rel = rel_l2(pm.out_all.view(-1))
    res["mixed_rel_l2_vs_fp64"] = summary(rel)
    by_f = {}
    for lo in (-4, -3):
        sel = (np.log10(f0) >= lo) & (np.log10(f0) < lo + 1)
        by_f[f"f0 in [1e{lo}, 1e{lo + 1})"] = float(np.median(rel[sel, 0])) if sel.any() else None
    res["mixed_rel_l2_A_median_by_f0"] = by_f
    del pm, h32
    # the same batch with 32-bit fixed-point input
    qb = BatchedResponse(1 / DT, tdi="2nd generation", orbits=orb, precision="mixed-q31",
                         storage_check="off")
    pq = qb.plan(hq, lengths, lam, beta, scales=qscales)
    pq.tune(use_cache=False, reps=1)
    wall, ms = time_plan(pq)
    pq.check_errors()
    res["q31_samples_per_s"] = total * steps / wall
    res["q31_rel_l2_vs_fp64"] = summary(rel_l2(pq.out_all.view(-1)))
    del pq, hq
    # and reading the complex128 input itself (no storage error)
    cb = BatchedResponse(1 / DT, tdi="2nd generation", orbits=orb, precision="mixed-c128")
    pc = cb.plan(h, lengths, lam, beta)
    pc.tune(use_cache=False, reps=1)
    wall, ms = time_plan(pc)
    pc.check_errors()
    res["c128_samples_per_s"] = total * steps / wall
    res["c128_vs_fp64"] = res["c128_samples_per_s"] / res["fp64_exact_samples_per_s"]
    res["c128_rel_l2_vs_fp64"] = summary(rel_l2(pc.out_all.view(-1)))
    return res


def serving_run(seconds, rate, seed=5):
    """Poisson arrivals at `rate` req/s for `seconds` s (numpy default_rng(seed)), served with
    the default admission budget (90% of the device's memory), drained; latency and
    throughput."""
    import numpy as np
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.serving import ResponseServer, poisson_requests, run_open_loop
    orb = EqualArmlengthOrbits(t_max=1.5 * YRSID_SI)
    orb.configure(linear_interp_setup=True)
    keys = [(dt, g) for dt in (5.0, 10.0) for g in ("1st generation", "2nd generation")]
    server = ResponseServer(orb, keys, num_streams=2)
    # warm-up: a few small requests, outside the measurement
    wt, wr = poisson_requests(np.random.default_rng(seed + 1), 50.0, 0.2,
                              lengths=(1 << 18,), max_sources=4)
    run_open_loop(server, wt, wr)
    server.done.clear()
    server.busy_intervals.clear()
    server.launches = 0
    # requests carry galactic-binary parameters: the server generates each input on the device
    times, reqs = poisson_requests(np.random.default_rng(seed), rate, seconds)
    stats = run_open_loop(server, times, reqs, drain_timeout=300.0)
    server.close()
    stats.pop("queue_trace", None)
    stats.update({"arrival_rate_per_s": rate, "arrival_seconds": seconds,
                  "hbm_budget_bytes": server.hbm_budget,
                  "workload": "Poisson arrivals, 1-64 sources/request, N in {2^18, 2^20, 2^22}, "
                              "dt in {5, 10} s, 30/70 1st/2nd gen, galactic-binary parameters "
                              "generated on the device as complex128, fused kernel with "
                              "float32 outputs, 2 streams, admission budget 90% of HBM"})
    return stats


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tdi", default="2nd generation", choices=list(FLOP_TDI))
    ap.add_argument("--orbits", default="equal", choices=["equal", "unequal"],
                    help="equal-arm cartwheel (default) or the unequal-arm cartwheel")
    ap.add_argument("--cpu-sample", type=int, default=1 << 21,
                    help="samples in the CPU-baseline sample (0 skips it)")
    ap.add_argument("--no-parity", action="store_true", help="skip the full-size oracle check")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="seconds of Poisson arrivals in the serving run (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="requests per second")
    ap.add_argument("--batch-sources", type=int, default=512,
                    help="sources of the batched workload (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--launch", default=None, metavar="TILExGRID",
                    help="fixed launch shape of the mixed plan (e.g. 512x32) instead of autotuning")
    ap.add_argument("--warm-seconds", type=float, default=1.5,
                    help="seconds of untimed launches before the first measurement")
    ap.add_argument("--precision", choices=list(FORMS), default="mixed-c128",
                    help="input format of the headline mixed-precision form")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="time one form only (for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also run one elementwise FP64 kernel reading and writing BYTES each "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE against a known size)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sources are "
              "independent and a long source shards by time range with halo recompute, but the "
              "multi-GPU path is not implemented.", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from fastlisaresponse_amd import pyResponseTDI
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits

    torch.cuda.set_device(0)
    orbit_cls = EqualArmlengthOrbits if args.orbits == "equal" else CartwheelOrbits
    orbits = orbit_cls(t_max=2.0 * YRSID_SI)
    orbits.configure(linear_interp_setup=True)
    lam, beta = 1.2, 0.3
    h_host = gb_waveform(N_SAMPLES, DT)
    h = torch.as_tensor(h_host, device="cuda")
    flop = FLOP_PROJ + FLOP_TDI[args.tdi]
    line = {"metric": METRIC, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "storage": FORMS[args.precision],
            "data": "synthetic (one galactic binary, f0=5 mHz, fdot=1e-16; analytic "
                    + ("equal-arm" if args.orbits == "equal" else "unequal-arm") + " orbits)",
            "config": {"workload": f"1-yr order-25 {args.tdi} TDI AET, mixed precision fused "
                                   "launch vs plain FP64 drop-in",
                       "num_pts": N_SAMPLES, "dt": DT, "order": ORDER, "tdi": args.tdi,
                       "tdi_chan": "AET", "orbits": args.orbits + "-arm cartwheel"}}

    from fastlisaresponse_amd.autotune import apply
    from fastlisaresponse_amd.batch import pack_q31

    warmed = []

    def warm_up(fn):
        """Bring the GPU to its working clocks with untimed calls of the leg about to be timed
        (--warm-seconds before the first leg, a fifth of that before each later one: the clocks
        drop while the host prepares a leg)."""
        end = time.perf_counter() + (args.warm_seconds / 5 if warmed else args.warm_seconds)
        while time.perf_counter() < end:
            for _ in range(20):
                fn()
            torch.cuda.synchronize()
        warmed.append(True)

    def mixed_leg(precision, launch=None, reps=3):
        """One fused mixed-precision plan over the workload: autotuned (or at `launch`), then
        timed. Returns (result dict, output [3, N])."""
        br = BatchedResponse(1 / DT, tdi=args.tdi, orbits=orbits, tdi_chan="AET",
                             precision=precision)
        if precision == "mixed":
            hin, sc = pack_mixed(h, [N_SAMPLES])
        elif precision == "mixed-q31":
            hin, sc = pack_q31(h, [N_SAMPLES])
        else:
            hin, sc = h, None
        plan = br.plan(hin, [N_SAMPLES], [lam], [beta], scales=sc)
        warm_up(plan.run)
        if launch:
            apply(plan, *launch)
            shape = {"tile_window": launch[0], "workgroups_per_cu": launch[1]}
        else:
            tune = plan.tune(use_cache=False, reps=reps)
            shape = {"tile_window": tune["best"][0], "workgroups_per_cu": tune["best"][1],
                     "autotune_ms": tune["ms"]}
        wall, ms = timed(plan.run, args.steps, args.warmup)
        plan.check_errors()
        plans[precision] = plan
        return ({"value": N_SAMPLES * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
                 "kernel_ms": ms, "launch_shape": shape, "form": FORMS[precision]},
                plan.out_all.view(3, -1))

    plans = {}
    outs = {}
    out_fp64 = None
    if args.only in (None, "mixed"):
        launch = tuple(int(x) for x in args.launch.lower().split("x")) if args.launch else None
        head, outs[args.precision] = mixed_leg(args.precision, launch)
        ms = head["kernel_ms"]
        line["config"]["precision"] = args.precision
        line["config"]["form"] = head["form"]
        line["config"]["launch_shape"] = head["launch_shape"]
        line.update({"value": head["value"], "ms_per_step": head["ms_per_step"]})
        achieved = flop * N_SAMPLES / (ms * 1e-3) / 1e12
        tile = head["launch_shape"]["tile_window"]
        traffic = pmc_traffic(f"fused_{args.precision}_1yr_{'2nd' if args.tdi.startswith('2nd') else '1st'}"
                              f"_gen_{args.orbits}_pmc.txt", f"tile{tile}")
        nbytes = BYTES_IO[args.precision]
        line["roofline"] = {
            "bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "kernel": ("fused_lti_kernel" if args.orbits == "equal" else "fused_batch_kernel")
                      + f" ({args.precision})", "kernel_ms": ms,
            "algorithmic_flop_per_sample": flop, "algorithmic_bytes_per_sample": nbytes,
            "hbm_achieved_GBs": nbytes * N_SAMPLES / (ms * 1e-3) / 1e9,
            "hbm_frac": nbytes * N_SAMPLES / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    if args.only is None:
        for prec in FORMS:  # the other input formats, each autotuned
            if prec != args.precision:
                line[prec.replace("-", "_")], outs[prec] = mixed_leg(prec, reps=2)

    if args.only in (None, "fp64"):
        r = pyResponseTDI(1 / DT, N_SAMPLES, order=ORDER, tdi=args.tdi, orbits=orbits,
                          tdi_chan="AET")
        res = {}

        def fp64_step():
            r.get_projections(h, lam, beta, t0=0.0, t_buffer=T_BUFFER)
            res["out"] = r.get_tdi_delays()

        warm_up(fp64_step)
        wall64, ms64 = timed(fp64_step, args.steps, args.warmup)
        r.check_errors()
        out_fp64 = torch.stack(res["out"])
        fp64 = {"value": N_SAMPLES * args.steps / wall64, "ms_per_step": wall64 * 1e3 / args.steps,
                "form": "plain FP64 drop-in: projection + TDI launches, exact stencils",
                "algorithmic_bytes_per_sample": BYTES_FP64,
                "achieved_TFLOPs": flop * N_SAMPLES / (ms64 * 1e-3) / 1e12}
        line["fp64"] = fp64
        if args.only == "fp64":
            line.update({"value": fp64["value"], "ms_per_step": fp64["ms_per_step"]})
        elif "value" in line:
            line["mixed_vs_fp64"] = line["value"] / fp64["value"]

    if args.only is None and args.precision == "mixed-c128":
        # the same form through the drop-in API (pyResponseTDI, one reserved batch rebound per
        # call): what a reference user switching precision= gets, host work included
        rd = pyResponseTDI(1 / DT, N_SAMPLES, order=ORDER, tdi=args.tdi, orbits=orbits,
                           tdi_chan="AET", precision="mixed-c128")

        def dropin_step():
            rd.get_projections(h, lam, beta, t0=0.0, t_buffer=T_BUFFER)
            rd.get_tdi_delays()

        warm_up(dropin_step)
        wd, _ = timed(dropin_step, args.steps, args.warmup)
        rd.check_errors()
        line["mixed_c128_dropin"] = {"value": N_SAMPLES * args.steps / wd,
                                     "ms_per_step": wd * 1e3 / args.steps,
                                     "form": "pyResponseTDI(precision='mixed-c128'): "
                                             "get_projections + get_tdi_delays per step"}

    out_fast = out_fused64 = None
    if args.only is None:
        # the other FP64 GPU forms (rel-L2 <= 1e-12 against the oracle rather than bitwise):
        # the drop-in with division-free stencils, and one fused FP64 launch (y kept in LDS)
        rf = pyResponseTDI(1 / DT, N_SAMPLES, order=ORDER, tdi=args.tdi, orbits=orbits,
                           tdi_chan="AET", exact=False)
        resf = {}

        def fast_step():
            rf.get_projections(h, lam, beta, t0=0.0, t_buffer=T_BUFFER)
            resf["out"] = rf.get_tdi_delays()

        warm_up(fast_step)
        wf, _ = timed(fast_step, args.steps, args.warmup)
        rf.check_errors()
        out_fast = torch.stack(resf["out"])
        bf = BatchedResponse(1 / DT, tdi=args.tdi, orbits=orbits, tdi_chan="AET",
                             precision="fp64", exact=False, fused=True)
        pf = bf.plan(h, [N_SAMPLES], [lam], [beta])
        tf = pf.tune(use_cache=False, reps=2)
        warm_up(pf.run)
        wfu, msfu = timed(pf.run, args.steps, args.warmup)
        pf.check_errors()
        out_fused64 = pf.out_all.view(3, -1)
        line["fp64_fast"] = {"value": N_SAMPLES * args.steps / wf, "ms_per_step": wf * 1e3 / args.steps,
                             "form": "plain FP64 drop-in, division-free stencils"}
        line["fp64_fused"] = {"value": N_SAMPLES * args.steps / wfu,
                              "ms_per_step": wfu * 1e3 / args.steps, "kernel_ms": msfu,
                              "launch_shape": list(tf["best"]),
                              "achieved_TFLOPs": flop * N_SAMPLES / (msfu * 1e-3) / 1e12,
                              "form": "one fused FP64 launch, division-free stencils"}
        if "value" in line:
            line["mixed_vs_fp64_fast"] = line["value"] / line["fp64_fast"]["value"]
            line["mixed_vs_fp64_fused"] = line["value"] / line["fp64_fused"]["value"]

    if args.only is None and args.precision in plans:
        # fused inner products against a resident data series (here the response itself,
        # stored as float32): the same launch summing d h, h^2, h and the alternating sum per
        # channel instead of writing A, E, T; beside it the store-then-reduce alternative
        plan = plans[args.precision]
        data = plan.out_all.clone()
        sums = torch.empty((1, 3, 4), dtype=torch.float64, device="cuda")
        warm_up(lambda: plan.inner_products(data, sums))
        wi, msi = timed(lambda: plan.inner_products(data, sums), args.steps, args.warmup)
        plan.check_errors()
        dd = data.view(3, -1).double()

        def store_then_dot():
            plan.run()
            o = plan.out_all.view(3, -1).double()
            return (o * dd).sum(1), (o * o).sum(1)

        warm_up(store_then_dot)
        ws, _ = timed(store_then_dot, args.steps, args.warmup)
        T0 = int(T_BUFFER / DT)
        hh = dd[:, T0:N_SAMPLES - T0]
        got = sums[0].cpu().numpy()
        line["inner_products"] = {
            "value": N_SAMPLES * args.steps / wi, "ms_per_step": wi * 1e3 / args.steps,
            "kernel_ms": msi, "store_then_torch_dot_ms_per_step": ws * 1e3 / args.steps,
            "rel_err_sum_hh_vs_float32_outputs": [
                float(abs(got[c, 1] - float((hh[c] * hh[c]).sum())) / got[c, 1]) for c in range(3)],
            "form": "BatchPlan.inner_products (flr_fused_inner): sums of d h, h^2, h and "
                    "(-1)^j h per channel in the fused launch, no A, E, T written"}

    if args.only is None:
        # unequal, time-varying arms (CartwheelOrbits): the mixed-c128 fused launch (composite
        # filters per tile with a drift term, on the matrix cores) against the FP64 drop-in
        orb_u = CartwheelOrbits(t_max=2.0 * YRSID_SI)
        orb_u.configure(linear_interp_setup=True)
        bu = BatchedResponse(1 / DT, tdi=args.tdi, orbits=orb_u, tdi_chan="AET",
                             precision="mixed-c128")
        pu = bu.plan(h, [N_SAMPLES], [lam], [beta])
        warm_up(pu.run)
        tu = pu.tune(use_cache=False, reps=2)
        wu, msu = timed(pu.run, args.steps, args.warmup)
        pu.check_errors()
        ru = pyResponseTDI(1 / DT, N_SAMPLES, order=ORDER, tdi=args.tdi, orbits=orb_u,
                           tdi_chan="AET")
        resu = {}

        def fp64u_step():
            ru.get_projections(h, lam, beta, t0=0.0, t_buffer=T_BUFFER)
            resu["out"] = ru.get_tdi_delays()

        warm_up(fp64u_step)
        wu64, _ = timed(fp64u_step, args.steps, args.warmup)
        ru.check_errors()
        un = {"mixed_c128": {"value": N_SAMPLES * args.steps / wu,
                             "ms_per_step": wu * 1e3 / args.steps, "kernel_ms": msu,
                             "launch_shape": list(tu["best"]),
                             "achieved_TFLOPs": flop * N_SAMPLES / (msu * 1e-3) / 1e12},
              "fp64": {"value": N_SAMPLES * args.steps / wu64,
                       "ms_per_step": wu64 * 1e3 / args.steps,
                       "form": "plain FP64 drop-in, exact stencils"},
              "orbits": "CartwheelOrbits: six different light travel times drifting ~1e-10 s/s"}
        un["mixed_vs_fp64"] = un["mixed_c128"]["value"] / un["fp64"]["value"]
        if not args.no_parity:
            from oracle import oracle as orc
            oou = orc.OracleOrbits(*orb_u.pycppdetector_args)
            _, ref_u = orc.response_pipeline(oou, h_host, lam, beta, DT, ORDER, args.tdi,
                                             x_base=orb_u.x_base,
                                             nthreads=max(1, min(16, os.cpu_count() or 1)))
            T0 = int(T_BUFFER / DT)
            for name, out in (("mixed_c128", pu.out_all.view(3, -1)),
                              ("fp64", torch.stack(resu["out"]))):
                got = out.double().cpu().numpy()
                un[name]["rel_l2_AET_vs_oracle"] = [
                    float(np.linalg.norm(got[c, T0:-T0] - ref_u[c, T0:-T0]) /
                          np.linalg.norm(ref_u[c, T0:-T0])) for c in range(3)]
                un[name]["bitwise_equal_oracle"] = bool(np.array_equal(got, ref_u))
        line["unequal_arms"] = un

    if not args.no_parity:
        from oracle import oracle as orc
        oo = orc.OracleOrbits(*orbits.pycppdetector_args)
        nth = max(1, min(16, os.cpu_count() or 1))
        _, aet_ref = orc.response_pipeline(oo, h_host, lam, beta, DT, ORDER, args.tdi,
                                           x_base=orbits.x_base, nthreads=nth)
        # the same arithmetic with extended-precision values (time/delay/index arithmetic kept):
        # measures how far every FP64 form, the oracle included, is from its exact value
        _, aet_wide = orc.response_pipeline(oo, h_host, lam, beta, DT, ORDER, args.tdi,
                                            x_base=orbits.x_base, nthreads=nth, wide=True)
        aet_wide = aet_wide.astype(np.float64)
        par = {"vs": "FP64 CPU oracle, full workload; rel_l2_AET_vs_wide: against the same "
                     "arithmetic in x87 extended precision (the FP64 rounding floor)"}
        T = int(T_BUFFER / DT)

        def vs_wide(got):
            return [float(np.linalg.norm(got[c, T:-T] - aet_wide[c, T:-T]) /
                          np.linalg.norm(aet_wide[c, T:-T])) for c in range(3)]

        par["fp64_oracle"] = {"rel_l2_AET_vs_wide": vs_wide(aet_ref)}
        for name, out in (*((p.replace("-", "_"), o) for p, o in outs.items()),
                          ("fp64", out_fp64),
                          ("fp64_fast", out_fast), ("fp64_fused", out_fused64)):
            if out is None:
                continue
            got = out.double().cpu().numpy()
            par[name] = {
                "rel_l2_AET": [float(np.linalg.norm(got[c, T:-T] - aet_ref[c, T:-T]) /
                                     np.linalg.norm(aet_ref[c, T:-T])) for c in range(3)],
                "max_abs_AET": [float(np.abs(got[c] - aet_ref[c]).max()) for c in range(3)],
                "bitwise_equal": bool(np.array_equal(got, aet_ref)),
                "rel_l2_AET_vs_wide": vs_wide(got)}
        line["parity"] = par

    if args.batch_sources > 0 and args.only is None:
        line["batch"] = batch_run(args.batch_sources, args.batch_steps)

    if args.serve_seconds > 0 and args.only is None:
        s = serving_run(args.serve_seconds, args.serve_rate)
        line["latency"] = {"p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"]}
        line["serving"] = s
        # the same traffic at lower arrival rates: latency below saturation
        curve = []
        for rate in (r for r in (0.5 * args.serve_rate, 0.75 * args.serve_rate) if r > 0):
            c = serving_run(args.serve_seconds, rate)
            curve.append({k: c[k] for k in ("arrival_rate_per_s", "p50_ms", "p99_ms",
                                             "samples_per_s", "offered_samples_per_s",
                                             "gpu_busy_frac", "queue_max")})
        curve.append({k: s[k] for k in ("arrival_rate_per_s", "p50_ms", "p99_ms",
                                        "samples_per_s", "offered_samples_per_s",
                                        "gpu_busy_frac", "queue_max")})
        line["latency_curve"] = curve

    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        line["copy_probe_bytes"] = args.copy_probe
    if args.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(orbits, h_host, lam, beta, args.tdi, args.cpu_sample)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
