"""Many sources of uneven length in one persistent launch per stage.

The reference processes one source per `get_projections` / `get_tdi_delays` call. Here a
batch of sources (MCMC walkers, a catalogue of binaries) is packed back to back in HBM and
runs as one persistent projection launch plus one persistent TDI launch over a flattened
(source, tile) work list; the launches can be captured into a HIP graph and replayed. Each
source's result is bitwise the same as running it alone (the kernels are deterministic and
tile results do not depend on their neighbours).

    br = BatchedResponse(0.1, order=25, tdi="2nd generation", orbits=orbits, tdi_chan="AET")
    plan = br.plan(h_all, lengths, lams, betas)     # h_all: packed complex128 device tensor
    plan.run()                                       # enqueues the two launches
    A, E, T = plan.channels(7)                       # views of source 7's output

Fused plans (fused=True, order 25) run projection and TDI as one launch with the projections
kept in LDS; precision="mixed" additionally stores the input as complex64 h / scale (one scale
per source, see pack_mixed) and the outputs as float32, with FP64 time/delay arithmetic and
FP64 stencil arithmetic inside the kernel; precision="mixed-q31" stores the input as 32-bit
fixed point instead (pack_q31), ~2^7 times less quantisation error at the same 8 bytes;
precision="mixed-c128" reads the complex128 input itself (16 bytes, no storage error) and keeps
the float32 outputs and the mixed arithmetic.
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Optional, Sequence

import numpy as np
import torch

from . import _capi
from ._capi import check, lib
from .backend import current_stream
from .response import pyResponseTDI
from .tables import buffer_indices, sky_basis

PRECISIONS = ("fp64", "mixed", "mixed-q31", "mixed-c128")

C_SI = 299792458.0

# ------------------------------------------------------------------ storage-error guard
# The 8-byte input formats round h to 2^-24 of each value ("mixed", complex64) or 2^-31 of full
# scale ("mixed-q31"). That rounding is white noise, while the response's signal cancels: the
# projection differences h at two times ~L apart (a factor ~wL) and TDI differences the links
# again, so the relative error of A and E grows about as (wL)^-3 towards low frequencies
# (w = 2 pi f, L = armlength / c), and faster still on the null channel T. The model
# K (wL)^-3 is calibrated on the 1-yr 5 mHz source (wL = 0.262): complex64 A, E rel-L2 2.5e-7
# (measured against the FP64 oracle), q31 2.8e-9 (the oracle on the quantised input); it
# reproduces the 512-source batch's medians (1.2e-6 for f0 in [1, 10] mHz, 9e-4 below).
STORAGE_K = {"fp64": 0.0, "mixed-c128": 0.0, "mixed": 4.5e-9, "mixed-q31": 5.0e-11}
STORAGE_TOL = 1e-6


class StorageErrorWarning(UserWarning):
    """An 8-byte input format whose rounding error on the A and E channels, amplified by the
    response's cancellations, is estimated above the tolerance; precision="mixed-c128" (the
    complex128 input read directly) has no storage error."""


def storage_error_estimate(f, precision: str, armlength: float = 2.5e9):
    """Estimated rel-L2 error of A and E that the storage format of `precision` causes for a
    source of (RMS) frequency f Hz: K (2 pi f L / c)^-3; 0 for the formats without storage
    error, and for f <= 0 (no signal)."""
    k = STORAGE_K[precision]
    f = np.asarray(f, dtype=np.float64)
    wl = 2.0 * np.pi * f * armlength / C_SI
    with np.errstate(divide="ignore", invalid="ignore"):
        est = np.where(wl > 0.0, k / np.maximum(wl, 1e-300) ** 3, 0.0)
    return est if est.ndim else float(est)


def rms_frequencies(h: torch.Tensor, lengths: Sequence[int], dt: float) -> np.ndarray:
    """RMS frequency (Hz) of each source of a packed complex series: with r^2 = mean |h[i+1] -
    h[i]|^2 / mean |h[i]|^2 over the source, f = 2 arcsin(r / 2) / (2 pi dt), which is exact for a
    single complex sinusoid. One pass over the samples (per-source sums by index), one host
    sync."""
    lengths = np.asarray([int(n) for n in lengths], dtype=np.int64)
    if h.dtype == torch.int32:  # q31 pairs
        h = torch.view_as_complex(h.view(-1, 2).to(torch.float64).contiguous())
    h = h.to(torch.complex128) if h.dtype != torch.complex128 else h
    nsrc = len(lengths)
    ids = torch.repeat_interleave(torch.arange(nsrc, device=h.device),
                                  torch.as_tensor(lengths, device=h.device))
    p = torch.zeros(nsrc, dtype=torch.float64, device=h.device)
    d = torch.zeros(nsrc, dtype=torch.float64, device=h.device)
    p.index_add_(0, ids, h.abs().square())
    if h.numel() > 1:
        diff = (h[1:] - h[:-1]).abs().square()
        same = ids[1:] == ids[:-1]  # no differences across source boundaries
        d.index_add_(0, ids[1:][same], diff[same])
    p, d = p.cpu().numpy() / np.maximum(lengths, 1), d.cpu().numpy() / np.maximum(lengths - 1, 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(np.where(p > 0.0, d / np.maximum(p, 1e-300), 0.0))
    return 2.0 * np.arcsin(np.minimum(r, 2.0) / 2.0) / (2.0 * np.pi * dt)


def check_storage(h: torch.Tensor, lengths: Sequence[int], dt: float, precision: str,
                  armlength: float = 2.5e9, tol: float = STORAGE_TOL, action: str = "warn"):
    """Estimate each source's storage error (storage_error_estimate at its RMS frequency) and
    warn (StorageErrorWarning) or raise (ValueError, action="raise") when one exceeds tol;
    returns the estimates. action="off" or a format without storage error skips the pass."""
    if action not in ("warn", "raise", "off"):
        raise ValueError("action must be 'warn', 'raise' or 'off'")
    if action == "off" or STORAGE_K[precision] == 0.0:
        return np.zeros(len(lengths))
    est = np.atleast_1d(storage_error_estimate(rms_frequencies(h, lengths, dt), precision,
                                               armlength))
    bad = np.flatnonzero(est > tol)
    if bad.size:
        msg = (f"precision={precision!r}: {bad.size} of {len(est)} source(s) have an estimated "
               f"storage error on A and E up to {est.max():.1e} (rel-L2, > {tol:g}; the null "
               "channel T is worse) -- the input rounding is amplified ~(2 pi f L/c)^-3 at low "
               "frequencies; precision='mixed-c128' reads the complex128 input with no storage "
               "error at the same speed")
        if action == "raise":
            raise ValueError(msg)
        warnings.warn(msg, StorageErrorWarning, stacklevel=3)
    return est


def pack_mixed(h: torch.Tensor, lengths: Sequence[int]):
    """complex128 packed sources -> (complex64 h / scale, scales): scale = max |h| per source,
    so the stored samples are O(1) whatever the amplitude."""
    if not (isinstance(h, torch.Tensor) and h.dtype == torch.complex128):
        raise TypeError("h must be a complex128 tensor")
    lengths = [int(n) for n in lengths]
    if sum(lengths) != h.numel():
        raise ValueError("lengths do not add up to the number of samples")
    out = torch.empty(h.numel(), dtype=torch.complex64, device=h.device)
    scales = []
    o = 0
    for n in lengths:
        seg = h[o:o + n]
        sc = float(seg.abs().max()) if n else 1.0
        sc = sc if sc > 0.0 else 1.0
        out[o:o + n] = (seg / sc).to(torch.complex64)
        scales.append(sc)
        o += n
    return out, scales


def pack_q31(h: torch.Tensor, lengths: Sequence[int]):
    """complex128 packed sources -> (int32 [2 * samples] fixed point, scales): per source,
    s = max(|Re h|, |Im h|) and each component is stored as round(x / s * (2^31 - 1)); the
    returned scale is s / (2^31 - 1), so h = stored * scale. Same 8 bytes per sample as
    complex64, with steps of 2^-31 of full scale instead of 2^-24 of the value."""
    if not (isinstance(h, torch.Tensor) and h.dtype == torch.complex128):
        raise TypeError("h must be a complex128 tensor")
    lengths = [int(n) for n in lengths]
    if sum(lengths) != h.numel():
        raise ValueError("lengths do not add up to the number of samples")
    full = float(2 ** 31 - 1)
    out = torch.empty(2 * h.numel(), dtype=torch.int32, device=h.device)
    scales = []
    o = 0
    for n in lengths:
        seg = torch.view_as_real(h[o:o + n])
        s = float(seg.abs().max()) if n else 1.0
        s = s if s > 0.0 else 1.0
        out[2 * o:2 * (o + n)] = torch.round(seg.reshape(-1) * (full / s)).to(torch.int32)
        scales.append(s / full)
        o += n
    return out, scales


class BatchedResponse:
    """Response configuration shared by a batch: sampling, order, TDI, orbits, channels.

    Args follow pyResponseTDI (sampling_frequency, order, tdi, orbits, tdi_chan, exact) plus
    t_buffer (s), the edge length left out of TDI for every source; fused (one launch, y kept
    on chip, order 25) and precision ("fp64", "mixed" with complex64 input, "mixed-q31"
    with 32-bit fixed-point input, see pack_q31, or "mixed-c128" with complex128 input; the
    mixed forms imply fused and the fast stencils and write float32 outputs). exact defaults to
    True for "fp64" and False otherwise. storage_check ("warn", "raise" or "off"): for "mixed"
    and "mixed-q31", each plan estimates every source's storage error on A and E
    (check_storage) and warns (StorageErrorWarning) or raises above 1e-6."""

    def __init__(self, sampling_frequency: float, order: int = 25, tdi="2nd generation",
                 orbits=None, tdi_chan: str = "AET", exact: Optional[bool] = None,
                 t_buffer: float = 10000.0, fused: bool = False, precision: str = "fp64",
                 storage_check: str = "warn"):
        if tdi_chan not in ("XYZ", "AET", "AE"):
            raise ValueError("tdi_chan must be 'XYZ', 'AET' or 'AE'.")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        if exact is None:
            exact = precision == "fp64"
        if precision != "fp64":
            if exact:
                raise ValueError("exact stencils need FP64 storage (precision='fp64')")
            fused = True
        if fused and order != 25:
            raise ValueError("fused batches are built for order 25")
        self.model = pyResponseTDI(sampling_frequency, 1, order=order, tdi=tdi, orbits=orbits,
                                   tdi_chan=tdi_chan, exact=exact)
        self.sampling_frequency = sampling_frequency
        self.dt = self.model.dt
        self.order = order
        self.tdi_chan = tdi_chan
        self.exact = bool(exact)
        self.fused = bool(fused)
        self.precision = precision
        if storage_check not in ("warn", "raise", "off"):
            raise ValueError("storage_check must be 'warn', 'raise' or 'off'")
        self.storage_check = storage_check
        self.t_buffer = t_buffer
        (self.tdi_start_ind, _, self.projection_buffer,
         self.projections_start_ind) = buffer_indices(sampling_frequency, order, t_buffer,
                                                      self.model.response_orbits.x_base)
        if self.projections_start_ind < self.projection_buffer:
            raise ValueError("Need to increase t_buffer. The initial buffer is not large enough.")

    def plan(self, h_all: torch.Tensor, lengths: Sequence[int], lams: Sequence[float],
             betas: Sequence[float], t0s: Optional[Sequence[float]] = None,
             scales: Optional[Sequence[float]] = None) -> "BatchPlan":
        """Bind a packed input (sources back to back in h_all: complex128 for "fp64" and
        "mixed-c128", complex64 h / scale for "mixed", int32 pairs for "mixed-q31") and allocate
        the outputs."""
        return BatchPlan(self, h_all, lengths, lams, betas, t0s, scales)


class BatchPlan:
    """A bound batch: device source table, packed projections y_all and outputs out_all."""

    def __init__(self, br: BatchedResponse, h_all, lengths, lams, betas, t0s=None, scales=None):
        lengths = np.asarray(lengths, dtype=np.int64)
        n = len(lengths)
        if (n < 1 or len(lams) != n or len(betas) != n or (t0s is not None and len(t0s) != n)
                or (scales is not None and len(scales) != n)):
            raise ValueError("lengths, lams, betas (and t0s, scales) must have one entry per source")
        if np.any(lengths < 1):
            raise ValueError("every source needs at least one sample")
        mixed = br.precision != "fp64"
        q31 = br.precision == "mixed-q31"
        want = {"fp64": torch.complex128, "mixed": torch.complex64,
                "mixed-q31": torch.int32, "mixed-c128": torch.complex128}[br.precision]
        if not (isinstance(h_all, torch.Tensor) and h_all.is_cuda and h_all.dtype == want
                and h_all.is_contiguous()):
            raise TypeError(f"h_all must be a contiguous {want} tensor on the GPU")
        if scales is not None and not br.fused:
            raise ValueError("scales apply to fused plans only")
        total = int(lengths.sum())
        held = h_all.numel() // 2 if q31 else h_all.numel()
        if held != total or (q31 and h_all.numel() % 2):
            raise ValueError(f"h_all holds {held} samples, the lengths sum to {total}")
        self.br = br
        self.h_all = h_all
        self.lengths = lengths
        self.storage_error = check_storage(h_all, lengths, br.dt, br.precision,
                                           br.model.response_orbits.armlength, action=br.storage_check)
        self.offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
        dev = h_all.device
        self.y_all = None if br.fused else torch.empty(6 * total, dtype=torch.float64, device=dev)
        self.out_all = torch.empty(3 * total, dtype=torch.float32 if mixed else torch.float64,
                                   device=dev)
        srcs = self._records(lams, betas, t0s, scales)
        flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE
        if br.tdi_chan != "XYZ":
            flags |= _capi.FLR_FLAG_AET
        if br.exact:
            flags |= _capi.FLR_FLAG_EXACT
        if mixed:
            flags |= _capi.FLR_FLAG_MIXED
        if q31:
            flags |= _capi.FLR_FLAG_Q31
        if br.precision == "mixed-c128":
            flags |= _capi.FLR_FLAG_C128
        h = ctypes.c_void_p()
        if br.fused:
            check(lib().flr_fused_create(br.model.cpp_response._h, srcs, n, h_all.data_ptr(),
                                         self.out_all.data_ptr(), br.dt, br.sampling_frequency,
                                         flags, ctypes.byref(h)))
        else:
            check(lib().flr_batch_create(br.model.cpp_response._h, srcs, n, h_all.data_ptr(),
                                         self.y_all.data_ptr(), self.out_all.data_ptr(), br.dt,
                                         br.sampling_frequency, flags, ctypes.byref(h)))
        self._h = h
        self.num_sources = n
        self.samples = int(lib().flr_batch_samples(h))
        self.launch_shape = (512 if br.fused else 256, 32)  # (tile window samples, WGs per CU)

    def _records(self, lams, betas, t0s=None, scales=None):
        """flr_source records of the plan's sources (packed offsets, the lengths of the plan)."""
        n = len(self.lengths)
        if (len(lams) != n or len(betas) != n or (t0s is not None and len(t0s) != n)
                or (scales is not None and len(scales) != n)):
            raise ValueError("lams, betas (and t0s, scales) must have one entry per source")
        srcs = (_capi.FlrSource * n)()
        for s in range(n):
            k, u, v = sky_basis(lams[s], betas[s])
            q = srcs[s]
            q.offset = int(self.offsets[s])
            q.num_pts = int(self.lengths[s])
            q.projections_start_ind = self.br.projections_start_ind
            q.tdi_start_ind = self.br.tdi_start_ind
            q.t0 = 0.0 if t0s is None else float(t0s[s])
            q.scale = 1.0 if scales is None else float(scales[s])
            for c in range(3):
                q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
        return srcs

    def rebind(self, lams, betas, t0s=None, scales=None, h_all: Optional[torch.Tensor] = None,
               stream: Optional[int] = None) -> None:
        """Fused plans: bind new source parameters (sky positions, start times, scales) and
        optionally a new packed input of the same format and size, keeping the lengths, the
        outputs and the launch shape -- the MCMC pattern of many steps over fixed-length
        sources. Enqueued on `stream` (default: the current one) without allocating; the
        composite TDI filters are rebuilt only if a start time changed. A graph captured from
        run() keeps replaying correctly after a rebind, since the kernel reads the rebound
        tables at run time."""
        if not self.br.fused:
            raise ValueError("rebind needs a fused plan (fused=True or a mixed precision)")
        if h_all is not None:
            if (not isinstance(h_all, torch.Tensor) or h_all.dtype != self.h_all.dtype or
                    h_all.numel() != self.h_all.numel() or not h_all.is_cuda or
                    not h_all.is_contiguous()):
                raise TypeError("h_all must be a contiguous device tensor like the plan's input")
            self.storage_error = check_storage(h_all, self.lengths, self.br.dt, self.br.precision,
                                               self.br.model.response_orbits.armlength,
                                               action=self.br.storage_check)
            self.h_all = h_all
        srcs = self._records(lams, betas, t0s, scales)
        check(lib().flr_fused_bind(self._h, srcs, len(self.lengths), self.h_all.data_ptr(),
                                   self.out_all.data_ptr(),
                                   current_stream() if stream is None else stream))

    def inner_products(self, data: torch.Tensor, sums: Optional[torch.Tensor] = None,
                       stream: Optional[int] = None) -> torch.Tensor:
        """Fused inner products of the plan's responses with a resident data series, without
        writing the responses (flr_fused_inner): `data` is laid out as out_all and has its dtype
        (float32 for the mixed forms). Returns sums [num_sources, 3, 4] (float64, device): per
        source and output channel, over [tdi_start, n - tdi_start), sum d h, sum h^2, sum h and
        sum (-1)^(i - tdi_start) h, with h the response in FP64 before its float32 rounding
        (utility.inner_from_sums, overlap_from_sums and log_likelihood_from_sums turn them into
        inner products, get_overlap's overlap and a white-noise log-likelihood). Mixed plans
        with the stock TDI tables. Enqueued on `stream` (default: the current one); the first
        call allocates the per-tile buffer, later ones nothing (graph-capturable after it)."""
        if self.br.precision == "fp64":
            raise ValueError("fused inner products run in the mixed-precision kernels "
                             "(precision='mixed-c128', 'mixed' or 'mixed-q31')")
        if self.br.tdi_chan == "AE":
            raise ValueError("inner products need all three channels (tdi_chan 'AET' or 'XYZ')")
        if (not isinstance(data, torch.Tensor) or not data.is_cuda or data.dtype != self.out_all.dtype
                or data.numel() != self.out_all.numel() or not data.is_contiguous()):
            raise TypeError("data must be a contiguous device tensor laid out as out_all")
        n = len(self.lengths)
        if sums is None:
            sums = torch.empty((n, 3, 4), dtype=torch.float64, device=self.out_all.device)
        elif sums.shape != (n, 3, 4) or sums.dtype != torch.float64 or not sums.is_contiguous():
            raise TypeError("sums must be a contiguous float64 tensor of shape (sources, 3, 4)")
        check(lib().flr_fused_inner(self._h, data.data_ptr(), sums.data_ptr(),
                                    current_stream() if stream is None else stream))
        return sums

    def tune(self, **kw) -> dict:
        """Pick the fastest launch shape for this plan (fastlisaresponse_amd.autotune.tune)."""
        from .autotune import tune
        return tune(self, **kw)

    def run(self, stream: Optional[int] = None) -> None:
        """Enqueue the batch's projection and TDI launches (no allocation, no sync)."""
        check(lib().flr_batch_run(self._h, current_stream() if stream is None else stream))

    def check_errors(self) -> None:
        self.br.model.cpp_response.check_errors()

    def channels(self, s: int):
        """Output channels of source s: (X, Y, Z), (A, E, T) or (A, E), as views."""
        o, n = int(self.offsets[s]), int(self.lengths[s])
        out = self.out_all[3 * o: 3 * (o + n)].view(3, n)
        return (out[0], out[1]) if self.br.tdi_chan == "AE" else (out[0], out[1], out[2])

    def projections(self, s: int) -> torch.Tensor:
        if self.y_all is None:
            raise ValueError("fused plans keep the projections on chip; use fused=False for y")
        o, n = int(self.offsets[s]), int(self.lengths[s])
        return self.y_all[6 * o: 6 * (o + n)].view(6, n)

    def graph(self) -> torch.cuda.CUDAGraph:
        """Capture run() into a graph (replay with .replay())."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run()  # warm-up outside capture (kernel attributes are set once)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            self.run()
        return g

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _capi._lib is not None:
            _capi._lib.flr_batch_destroy(h)
            self._h = None
