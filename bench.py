#!/usr/bin/env python3
"""Throughput of the LISA time-domain response on one MI355X.

Workload: one galactic-binary source over one year, N = 3,153,600 samples at dt = 10 s,
order-25 Lagrange stencils, 2nd-generation TDI, A/E/T output, FP64 throughout, equal-arm
analytic cartwheel orbits tabulated every 1e4 s. A step is one projection + TDI(+AET) pass over
the source with its inputs already resident in HBM; `value` is AET samples/s.

Also reported on the same JSON line:
  roofline      dominant kernel's algorithmic FP64 rate (stencil multiply-adds) against the
                78.6 TFLOP/s FP64 peak, its launch time from HIP events in the timed region,
                and HBM bytes per launch from the committed rocprofv3 counter summary
  cpu_baseline  the FP64 CPU oracle (1 thread, the reference CPU path's serial arithmetic) on
                a bounded prefix of the same workload
  parity        rel-L2 / max-abs of A, E, T against the FP64 oracle on the full workload
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
# Algorithmic FLOPs per output sample: 2 * (stencil taps) per stencil evaluation.
TAPS = ORDER + 1
FLOP_TDI = {"1st generation": 2 * TAPS * 24, "2nd generation": 2 * TAPS * 48}
FLOP_PROJ = 2 * TAPS * 2 * 2 * 6  # 6 links x (rec, em) x (h+, hx)


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


def pmc_traffic(form: str, tdi: str, orbits: str):
    """HBM bytes per TDI launch from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE, calibrated against a copy kernel), or None if none exists."""
    gen = "2nd_gen" if tdi.startswith("2nd") else "1st_gen"
    arms = "" if orbits == "equal" else "_unequal"
    path = os.path.join(ROOT, "profiles", f"kernels_1yr_{gen}_fp64_{form}{arms}_pmc.txt")
    try:
        with open(path) as f:
            m = re.search(r"^HBM_BYTES_PER_LAUNCH\s+([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(orbits, h, lam, beta, tdi, n_sample):
    """Oracle projection + TDI + AET on the first n_sample samples, one thread."""
    import numpy as np
    from oracle import oracle as orc
    oo = orc.OracleOrbits(*orbits.pycppdetector_args)
    hs = np.ascontiguousarray(h[:n_sample])
    t0 = time.perf_counter()
    orc.response_pipeline(oo, hs, lam, beta, DT, ORDER, tdi, x_base=orbits.x_base, nthreads=1)
    el = time.perf_counter() - t0
    return {"value": n_sample / el, "unit": "samples/s", "cores": 1, "kind": "port",
            "sample": f"first {n_sample} samples of the same source, projection + {tdi} TDI + AET "
                      f"({el:.1f} s, FP64 oracle restating the reference CPU path)"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tdi", default="2nd generation", choices=list(FLOP_TDI))
    ap.add_argument("--fast", action="store_true",
                    help="reciprocal-free stencils instead of the (default) bitwise-exact form")
    ap.add_argument("--orbits", default="equal", choices=["equal", "unequal"],
                    help="equal-arm cartwheel (default) or the unequal-arm cartwheel")
    ap.add_argument("--cpu-sample", type=int, default=1 << 21,
                    help="samples in the CPU-baseline sample (0 skips it)")
    ap.add_argument("--no-parity", action="store_true", help="skip the full-size oracle check")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also run one elementwise FP64 kernel reading and writing BYTES each "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE against a known size)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sources are "
              "independent and a long source shards by time range with halo recompute, but no "
              "multi-GPU run is part of this build.", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from fastlisaresponse_amd import pyResponseTDI
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits

    torch.cuda.set_device(0)
    orbit_cls = EqualArmlengthOrbits if args.orbits == "equal" else CartwheelOrbits
    orbits = orbit_cls(t_max=2.0 * YRSID_SI)
    orbits.configure(linear_interp_setup=True)
    lam, beta = 1.2, 0.3
    r = pyResponseTDI(1 / DT, N_SAMPLES, order=ORDER, tdi=args.tdi, orbits=orbits, tdi_chan="AET",
                      exact=not args.fast)
    h_host = gb_waveform(N_SAMPLES, DT)
    h = torch.as_tensor(h_host, device="cuda")

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        r.get_projections(h, lam, beta, t0=0.0, t_buffer=T_BUFFER)
        if ev is not None:
            ev[1].record()
        out = r.get_tdi_delays()
        if ev is not None:
            ev[2].record()
        return out

    for _ in range(args.warmup):
        step()
    r.check_errors()
    torch.cuda.synchronize()

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        out = step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    r.check_errors()

    proj_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    tdi_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    value = N_SAMPLES * args.steps / elapsed
    if tdi_ms >= proj_ms:
        kname, kms, flop = "tdi_kernel", tdi_ms, FLOP_TDI[args.tdi]
    else:
        kname, kms, flop = "projection_kernel", proj_ms, FLOP_PROJ
    achieved = flop * N_SAMPLES / (kms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": (pmc_traffic("fast" if args.fast else "exact", args.tdi, args.orbits)
                            if kname == "tdi_kernel" else None),
                "kernel": kname, "kernel_ms": kms,
                "algorithmic_flop_per_sample": flop}

    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (one galactic binary, f0=5 mHz, fdot=1e-16; analytic equal-arm orbits)",
        "config": {"workload": f"1-yr order-25 {args.tdi} TDI AET, FP64 plain GPU form",
                   "num_pts": N_SAMPLES, "dt": DT, "order": ORDER, "tdi": args.tdi,
                   "tdi_chan": "AET", "stencil_form": "fast" if args.fast else "exact",
                   "orbits": "equal-arm cartwheel" if args.orbits == "equal"
                   else "unequal-arm cartwheel"},
        "kernels_ms": {"projection_kernel": proj_ms, "tdi_kernel": tdi_ms},
        "roofline": roofline,
    }

    if not args.no_parity:
        from oracle import oracle as orc
        oo = orc.OracleOrbits(*orbits.pycppdetector_args)
        nth = max(1, min(16, os.cpu_count() or 1))
        _, aet_ref = orc.response_pipeline(oo, h_host, lam, beta, DT, ORDER, args.tdi,
                                           x_base=orbits.x_base, nthreads=nth)
        got = torch.stack(out).cpu().numpy()
        rel = [float(np.linalg.norm(got[c] - aet_ref[c]) / np.linalg.norm(aet_ref[c])) for c in range(3)]
        mabs = [float(np.abs(got[c] - aet_ref[c]).max()) for c in range(3)]
        line["parity"] = {"vs": "FP64 CPU oracle, full workload", "rel_l2_AET": rel,
                          "max_abs_AET": mabs, "bitwise_equal": bool(np.array_equal(got, aet_ref))}
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
