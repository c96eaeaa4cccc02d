"""Continuous-batching request loop for the LISA response on one GPU.

A request is a batch of sources (e.g. the walkers of one MCMC step) that share a sampling
interval and a TDI generation; each source has its own length, sky position, amplitude scale
and input h / scale (complex128 by default, or complex64 with precision="mixed"). The server keeps resident, per (dt, generation), one response
handle (orbit, Lagrange and TDI tables on the device) and, per stream, one reserved fused batch
(device source table + pinned staging), so the loop allocates no device tables and never
synchronises the device:

* arrival   -- `submit` queues the request (FIFO). Inputs the client passes as device tensors
  may still be in production on the client's current stream: submit records an event there,
  the launch's stream waits for it before copying, and the copy keeps the tensor's memory from
  being reused until it has been read, so the client needs no synchronisation.
* admission -- the head request is admitted when the bytes in flight plus its own (packed input
  and float32 outputs) fit the HBM budget and a stream is free; queued requests with the same
  key are merged into the same launch up to `max_launch_samples` / `max_sources`.
* launch    -- the admitted inputs are packed into one arena (device copies of the clients'
  waveforms, or waveforms generated on the device from galactic-binary parameters), the
  sources are bound (asynchronous table upload) and one fused projection + TDI launch is
  enqueued on the stream, followed by an event.
* completion -- `poll` checks the events; a finished launch completes all its requests
  (latency = completion - arrival) and releases its memory.

Launches on different streams overlap on the device, so requests that arrive while others run
are admitted as soon as a stream and memory free up. Every source's result is the same bits
whichever requests it shares a launch with (tiles never mix sources).
"""
from __future__ import annotations

import ctypes
import os
import time
import warnings
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _capi
from ._capi import check, lib
from .batch import (C_SI, STORAGE_K, STORAGE_TOL, BatchedResponse, StorageErrorWarning,
                    storage_error_estimate)
from .tables import sky_basis


@dataclass
class SourceSpec:
    """One source of a request: either its input h (h / scale on the device, in the server's
    input format: complex128, or complex64 for precision="mixed") or the
    galactic-binary parameters gb = (A, f0, fdot, iota, phi0, psi) and length n, from which the
    server generates the input on the device."""
    h: Optional[torch.Tensor]
    lam: float
    beta: float
    scale: float = 1.0
    t0: float = 0.0
    gb: Optional[Tuple[float, ...]] = None
    n: int = 0

    @property
    def num_pts(self) -> int:
        return int(self.h.numel()) if self.h is not None else int(self.n)


@dataclass
class Request:
    rid: int
    dt: float
    tdi: str
    sources: List[SourceSpec]
    arrival: float = 0.0          # perf_counter time the request arrived
    admitted: float = 0.0
    completed: float = 0.0
    result: Optional[torch.Tensor] = None  # [3 * samples] float32 (A, E, T per source) if kept
    offsets: List[int] = field(default_factory=list)
    keep: Optional[bool] = None  # keep this request's result (None: the server's keep_results)
    ready: Optional[torch.cuda.Event] = None  # recorded at submit on the caller's stream

    @property
    def key(self) -> Tuple[float, str]:
        return (self.dt, self.tdi)

    @property
    def samples(self) -> int:
        return sum(s.num_pts for s in self.sources)

    def channels(self, s: int):
        """(A, E, T) of source s from the kept result."""
        o, n = self.offsets[s], self.sources[s].num_pts
        return self.result[3 * o: 3 * (o + n)].view(3, n)

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


class _Launch:
    def __init__(self, stream, requests, arena, out, nbytes, start, end):
        self.stream = stream
        self.requests = requests
        self.arena = arena
        self.out = out
        self.nbytes = nbytes
        self.start = start
        self.end = end


class ResponseServer:
    """Serve response requests with continuous batching (mixed precision, fused kernels).

    Args:
        orbits: orbit provider shared by all requests (tables must cover the longest request).
        keys: (dt, tdi) pairs to keep resident.
        hbm_budget: bytes of request inputs + outputs allowed in flight; None (default) is
            90% of the device's memory (259 GB of an MI355X's 288 GB), the rest left to the
            resident tables, torch's allocator cache and other users.
        num_streams: launches that may run concurrently.
        max_launch_samples, max_sources: caps on one merged launch.
        keep_results: keep each request's outputs (tests); otherwise they are dropped. A
            request's own `keep` (True or False) overrides it.
        t_buffer: edge length (s) left out of TDI, as pyResponseTDI.get_projections.
        precision: "mixed-c128" (complex128 inputs, float32 outputs; default) or "mixed"
            (complex64 inputs h / scale: half the input memory, FP32 storage error).
        launch: (tile window samples, workgroups per CU) of every launch; the default 1024-sample
            window is the autotuner's pick for long sources. A key whose delay windows do not
            fit that window in LDS (very fine sampling) keeps the library's default shape.
    """

    INPUT_BYTES = {"mixed": 8, "mixed-c128": 16}
    BYTES_PER_SAMPLE = 16 + 12  # per sample in flight: input + float32 A, E, T (default format)

    def __init__(self, orbits, keys: Sequence[Tuple[float, str]], hbm_budget: Optional[int] = None,
                 num_streams: int = 2, max_launch_samples: int = 1 << 27, max_sources: int = 512,
                 keep_results: bool = False, t_buffer: float = 10000.0, tdi_chan: str = "AET",
                 precision: str = "mixed-c128", launch: Optional[Tuple[int, int]] = (1024, 32)):
        if precision not in self.INPUT_BYTES:
            raise ValueError(f"precision must be one of {sorted(self.INPUT_BYTES)}")
        queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        if not 1 <= num_streams <= queues:
            # more streams than hardware queues serialise behind one another (measured: 8
            # streams on 4 queues served a quarter of the throughput)
            raise ValueError(f"num_streams must be in [1, {queues}] (hardware queues per process)")
        self.precision = precision
        if precision == "mixed":
            # the clients' complex64 inputs are not inspected in the loop (that would synchronise
            # per request); parameter requests are checked at submit from their frequency
            f_lim = (STORAGE_K[precision] / STORAGE_TOL) ** (1.0 / 3.0) * C_SI / (
                2.0 * np.pi * orbits.armlength)
            warnings.warn(f"precision='mixed': the complex64 storage error on A and E exceeds "
                          f"{STORAGE_TOL:g} (rel-L2) for sources below ~{f_lim * 1e3:.1f} mHz; "
                          "'mixed-c128' has none", StorageErrorWarning, stacklevel=2)
        self.armlength = float(orbits.armlength)
        self.launch = None if launch is None else (int(launch[0]), int(launch[1]))
        self.BYTES_PER_SAMPLE = self.INPUT_BYTES[precision] + 12
        self._in_dtype = torch.complex128 if precision == "mixed-c128" else torch.complex64
        self.models: Dict[Tuple[float, str], BatchedResponse] = {}
        for dt, tdi in keys:
            self.models[(dt, tdi)] = BatchedResponse(1.0 / dt, tdi=tdi, orbits=orbits,
                                                     tdi_chan=tdi_chan, precision=precision,
                                                     t_buffer=t_buffer)
        if hbm_budget is None:
            hbm_budget = int(0.9 * torch.cuda.get_device_properties(torch.cuda.current_device())
                             .total_memory)
        self.hbm_budget = int(hbm_budget)
        self.max_launch_samples = int(max_launch_samples)
        self.max_sources = int(max_sources)
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream() for _ in range(num_streams)]
        self.busy: List[Optional[_Launch]] = [None] * num_streams
        self._batches: Dict[Tuple[int, Tuple[float, str]], ctypes.c_void_p] = {}
        self.pending: deque = deque()
        self.in_flight_bytes = 0
        self.done: List[Request] = []
        self.launches = 0
        self.busy_intervals: List[Tuple[torch.cuda.Event, torch.cuda.Event]] = []

    # ------------------------------------------------------------ resident state
    def _batch(self, slot: int, key) -> ctypes.c_void_p:
        b = self._batches.get((slot, key))
        if b is None:
            br = self.models[key]
            flags = (_capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE | _capi.FLR_FLAG_MIXED |
                     (_capi.FLR_FLAG_AET if br.tdi_chan != "XYZ" else 0) |
                     (_capi.FLR_FLAG_C128 if self.precision == "mixed-c128" else 0))
            b = ctypes.c_void_p()
            check(lib().flr_fused_reserve(br.model.cpp_response._h, self.max_sources, br.dt,
                                          br.sampling_frequency, flags, ctypes.byref(b)))
            if self.launch is not None:
                # a shape the key's windows do not fit is refused: keep the default one
                lib().flr_batch_set_launch(b, self.launch[0], self.launch[1])
            self._batches[(slot, key)] = b
        return b

    def close(self) -> None:
        for s in self.streams:
            s.synchronize()
        for b in self._batches.values():
            lib().flr_batch_destroy(b)
        self._batches.clear()

    def __del__(self):
        try:
            if self._batches and _capi._lib is not None:
                self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ loop
    def submit(self, req: Request, now: Optional[float] = None) -> None:
        if req.key not in self.models:
            raise ValueError(f"no resident response for dt={req.dt}, tdi={req.tdi!r}")
        if not req.sources or len(req.sources) > self.max_sources:
            raise ValueError("a request needs 1..max_sources sources")
        if any(sp.h is not None and sp.h.dtype != self._in_dtype for sp in req.sources):
            raise TypeError(f"source inputs must be {self._in_dtype} for precision={self.precision!r}")
        need = req.samples * self.BYTES_PER_SAMPLE
        if need > self.hbm_budget:
            raise ValueError("request larger than the HBM budget")
        req.arrival = time.perf_counter() if now is None else now
        if self.precision == "mixed":
            f0 = [sp.gb[1] for sp in req.sources if sp.h is None]
            if f0 and max(np.atleast_1d(storage_error_estimate(f0, self.precision,
                                                                self.armlength))) > STORAGE_TOL:
                warnings.warn(f"request {req.rid}: a source's estimated complex64 storage error "
                              f"on A and E exceeds {STORAGE_TOL:g}; use precision='mixed-c128'",
                              StorageErrorWarning, stacklevel=2)
        if any(sp.h is not None for sp in req.sources):
            # the inputs may still be being produced on the caller's stream: the launch that
            # copies them waits for this point of it (no host synchronisation)
            req.ready = torch.cuda.Event()
            req.ready.record(torch.cuda.current_stream())
        self.pending.append(req)

    def poll(self) -> int:
        """Complete finished launches; returns the number of requests completed."""
        n = 0
        for slot, ln in enumerate(self.busy):
            if ln is None or not ln.end.query():
                continue
            now = time.perf_counter()
            for req in ln.requests:
                req.completed = now
                if not (self.keep_results if req.keep is None else req.keep):
                    req.result = None
                self.done.append(req)
                n += 1
            self.in_flight_bytes -= ln.nbytes
            self.busy_intervals.append((ln.start, ln.end))
            self.busy[slot] = None
        return n

    def _take(self) -> List[Request]:
        """Admit the head request and the queued requests of its key that fit."""
        head = self.pending[0]
        key = head.key
        taken: List[Request] = []
        samples = sources = 0
        budget = self.hbm_budget - self.in_flight_bytes
        keep = deque()
        while self.pending:
            req = self.pending.popleft()
            fits = (req.key == key and sources + len(req.sources) <= self.max_sources and
                    (not taken or samples + req.samples <= self.max_launch_samples) and
                    (samples + req.samples) * self.BYTES_PER_SAMPLE <= budget)
            if fits:
                taken.append(req)
                samples += req.samples
                sources += len(req.sources)
            else:
                keep.append(req)
                if req is head:  # head-of-line does not fit: admit nothing
                    break
        keep.extend(self.pending)
        self.pending = keep
        return taken

    def admit(self) -> int:
        """Launch queued work on free streams; returns the number of launches started."""
        started = 0
        while self.pending:
            slot = next((k for k, ln in enumerate(self.busy) if ln is None), None)
            if slot is None:
                break
            head = self.pending[0]
            if self.in_flight_bytes + head.samples * self.BYTES_PER_SAMPLE > self.hbm_budget:
                break
            reqs = self._take()
            if not reqs:
                break
            self._launch(slot, reqs)
            started += 1
        return started

    def _launch(self, slot: int, reqs: List[Request]) -> None:
        key = reqs[0].key
        br = self.models[key]
        stream = self.streams[slot]
        total = sum(r.samples for r in reqs)
        nsrc = sum(len(r.sources) for r in reqs)
        now = time.perf_counter()
        srcs = (_capi.FlrSource * nsrc)()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        gen_params, gen_len, gen_off, gen_t0, gen_idx = [], [], [], [], []
        with torch.cuda.stream(stream):
            start.record(stream)
            arena = torch.empty(total, dtype=self._in_dtype, device="cuda")
            out = torch.empty(3 * total, dtype=torch.float32, device="cuda")
            o = q = 0
            for req in reqs:
                req.admitted = now
                req.offsets = []
                if req.ready is not None:
                    stream.wait_event(req.ready)
                for sp in req.sources:
                    n = sp.num_pts
                    s = srcs[q]
                    if sp.h is not None:
                        arena[o:o + n].copy_(sp.h, non_blocking=True)
                        # the caller may drop its tensor now: its memory is not reused before
                        # this stream's copy has read it
                        sp.h.record_stream(stream)
                        s.scale = sp.scale
                    else:  # generated below, all of the launch's parameter sources at once
                        gen_params.append(sp.gb)
                        gen_len.append(n)
                        gen_off.append(o)
                        gen_t0.append(sp.t0)
                        gen_idx.append(q)
                    s.offset = o
                    s.num_pts = n
                    s.projections_start_ind = br.projections_start_ind
                    s.tdi_start_ind = br.tdi_start_ind
                    s.t0 = sp.t0
                    k, u, v = sky_basis(sp.lam, sp.beta)
                    for c in range(3):
                        s.k[c], s.u[c], s.v[c] = k[c], u[c], v[c]
                    req.offsets.append(o)
                    o += n
                    q += 1
            if gen_params:
                from .waveforms import generate_gb
                fmt = "complex128" if self.precision == "mixed-c128" else "complex64"
                _, gscales = generate_gb(np.asarray(gen_params), gen_len, key[0], fmt=fmt,
                                         t0s=gen_t0, out=arena, stream=stream, offsets=gen_off)
                for q, sc in zip(gen_idx, gscales):
                    srcs[q].scale = float(sc)
            b = self._batch(slot, key)
            check(lib().flr_fused_bind(b, srcs, nsrc, arena.data_ptr(), out.data_ptr(),
                                       ctypes.c_void_p(stream.cuda_stream)))
            check(lib().flr_batch_run(b, ctypes.c_void_p(stream.cuda_stream)))
            end.record(stream)
        base = 0
        for req in reqs:
            req.result = out[3 * base: 3 * (base + req.samples)]
            req.offsets = [x - base for x in req.offsets]
            base += req.samples
        nbytes = total * self.BYTES_PER_SAMPLE
        self.in_flight_bytes += nbytes
        self.busy[slot] = _Launch(stream, reqs, arena, out, nbytes, start, end)
        self.launches += 1

    def idle(self) -> bool:
        return not self.pending and all(ln is None for ln in self.busy)

    def step(self) -> None:
        self.poll()
        self.admit()

    def check_errors(self) -> None:
        for br in self.models.values():
            br.model.cpp_response.check_errors()

    def gpu_busy_seconds(self) -> float:
        """Union of the launches' device intervals (start/end events), in seconds."""
        if not self.busy_intervals:
            return 0.0
        ref = self.busy_intervals[0][0]
        iv = sorted((ref.elapsed_time(a), ref.elapsed_time(b)) for a, b in self.busy_intervals)
        tot, cur_s, cur_e = 0.0, iv[0][0], iv[0][1]
        for s, e in iv[1:]:
            if s > cur_e:
                tot += cur_e - cur_s
                cur_s, cur_e = s, e
            else:
                cur_e = max(cur_e, e)
        tot += cur_e - cur_s
        return tot * 1e-3


# ------------------------------------------------------------------ synthetic traffic

def gb_pool(num: int, n: int, dt: float, seed: int = 0,
            dtype: torch.dtype = torch.complex64) -> List[torch.Tensor]:
    """`num` normalised galactic-binary waveforms (max |h| = 1, `dtype`) of n samples, made on the device
    (f0 log-uniform in [1e-4, 1e-2] Hz, fdot log-uniform in [1e-19, 1e-15] Hz/s, cos iota,
    phi0, psi uniform)."""
    rng = np.random.default_rng(seed)
    t = torch.arange(n, dtype=torch.float64, device="cuda") * dt
    out = []
    for _ in range(num):
        f = 10 ** rng.uniform(-4, -2)
        fd = 10 ** rng.uniform(-19, -15)
        ci = rng.uniform(-1, 1)
        phi0, psi = rng.uniform(0, 2 * np.pi, 2)
        fdd = 11.0 / 3.0 * fd ** 2 / f
        ph = 2 * np.pi * (f * t + 0.5 * fd * t * t + fdd * t ** 3 / 6.0) - phi0
        hp0 = -torch.cos(ph) * (1.0 + ci * ci)
        hc0 = -torch.sin(ph) * 2.0 * ci
        c2, s2 = np.cos(2 * psi), np.sin(2 * psi)
        h = torch.complex(hp0 * c2 - hc0 * s2, hp0 * s2 + hc0 * c2)
        h = h / h.abs().max()
        out.append(h.to(dtype))
    torch.cuda.synchronize()  # the server's streams read these
    return out


def poisson_requests(rng: np.random.Generator, rate: float, duration: float, pool=None,
                     lengths=(1 << 18, 1 << 20, 1 << 22), dts=(5.0, 10.0),
                     gens=(("1st generation", 0.3), ("2nd generation", 0.7)), max_sources=64):
    """Arrival offsets (s) and requests: Poisson arrivals at `rate`/s over `duration` s, each a
    batch of 1..max_sources sources sharing one length, dt and TDI generation (probabilities
    in `gens`); sky uniform on the sphere, log10 amplitude uniform in [-23, -21]. With a `pool`
    of waveforms each source takes a slice of one; without, each source carries
    galactic-binary parameters (f0 log-uniform in [1e-4, 1e-2] Hz, fdot log-uniform in
    [1e-19, 1e-15] Hz/s, cos i, phi0, psi uniform) and the server generates it."""
    times, reqs = [], []
    t = rng.exponential(1.0 / rate)
    rid = 0
    names = [g for g, _ in gens]
    probs = np.array([p for _, p in gens], dtype=float)
    probs /= probs.sum()
    while t < duration:
        n = int(rng.choice(lengths))
        dt = float(rng.choice(dts))
        tdi = names[int(rng.choice(len(names), p=probs))]
        ns = int(rng.integers(1, max_sources + 1))
        srcs = []
        for _ in range(ns):
            lam = float(rng.uniform(0, 2 * np.pi))
            beta = float(np.arcsin(rng.uniform(-1, 1)))
            amp = float(10 ** rng.uniform(-23, -21))
            if pool is not None:
                h = pool[int(rng.integers(len(pool)))][:n]
                srcs.append(SourceSpec(h=h, lam=lam, beta=beta, scale=amp))
            else:
                gb = (amp, float(10 ** rng.uniform(-4, -2)), float(10 ** rng.uniform(-19, -15)),
                      float(np.arccos(rng.uniform(-1, 1))), float(rng.uniform(0, 2 * np.pi)),
                      float(rng.uniform(0, 2 * np.pi)))
                srcs.append(SourceSpec(h=None, lam=lam, beta=beta, gb=gb, n=n))
        reqs.append(Request(rid, dt, tdi, srcs))
        times.append(t)
        rid += 1
        t += rng.exponential(1.0 / rate)
    return times, reqs


def run_open_loop(server: ResponseServer, times: Sequence[float], reqs: Sequence[Request],
                  drain_timeout: float = 120.0, sample_every: float = 0.1) -> dict:
    """Replay arrivals on the wall clock, serve until drained; report latency percentiles,
    throughput, GPU utilisation and the queue length over time."""
    t_start = time.perf_counter()
    nxt = 0
    queue_trace = []
    next_sample = 0.0
    while True:
        now = time.perf_counter() - t_start
        while nxt < len(reqs) and times[nxt] <= now:
            server.submit(reqs[nxt], now=t_start + times[nxt])
            nxt += 1
        server.step()
        if now >= next_sample:
            queue_trace.append((round(now, 3), len(server.pending)))
            next_sample += sample_every
        if nxt == len(reqs) and server.idle():
            break
        if now > (times[-1] if times else 0.0) + drain_timeout:
            raise TimeoutError("requests did not drain")
        if not any(ln is not None and not ln.end.query() for ln in server.busy) and nxt < len(reqs):
            time.sleep(min(1e-4, max(0.0, times[nxt] - now)))
        else:
            time.sleep(5e-5)
    wall = time.perf_counter() - t_start
    server.check_errors()
    lat = np.array([r.latency for r in server.done])
    samples = sum(r.samples for r in server.done)
    last_done = max(r.completed for r in server.done) - t_start
    busy = server.gpu_busy_seconds()
    offered = sum(r.samples for r in reqs) / max(times[-1], 1e-9) if times else 0.0
    return {
        "requests": len(server.done), "launches": server.launches,
        "samples": samples, "wall_s": wall,
        "samples_per_s": samples / last_done,
        "offered_samples_per_s": offered,
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3),
        "gpu_busy_frac": busy / last_done,
        "queue_max": max(q for _, q in queue_trace) if queue_trace else 0,
        "queue_trace": queue_trace[:: max(1, len(queue_trace) // 20)],
    }
