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
from typing import Optional, Sequence

import numpy as np
import torch

from . import _capi
from ._capi import check, lib
from .backend import current_stream
from .response import pyResponseTDI
from .tables import buffer_indices, sky_basis

PRECISIONS = ("fp64", "mixed", "mixed-q31", "mixed-c128")


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
    True for "fp64" and False otherwise."""

    def __init__(self, sampling_frequency: float, order: int = 25, tdi="2nd generation",
                 orbits=None, tdi_chan: str = "AET", exact: Optional[bool] = None,
                 t_buffer: float = 10000.0, fused: bool = False, precision: str = "fp64"):
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
        self.offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
        dev = h_all.device
        self.y_all = None if br.fused else torch.empty(6 * total, dtype=torch.float64, device=dev)
        self.out_all = torch.empty(3 * total, dtype=torch.float32 if mixed else torch.float64,
                                   device=dev)
        srcs = (_capi.FlrSource * n)()
        for s in range(n):
            k, u, v = sky_basis(lams[s], betas[s])
            q = srcs[s]
            q.offset = int(self.offsets[s])
            q.num_pts = int(lengths[s])
            q.projections_start_ind = br.projections_start_ind
            q.tdi_start_ind = br.tdi_start_ind
            q.t0 = 0.0 if t0s is None else float(t0s[s])
            q.scale = 1.0 if scales is None else float(scales[s])
            for c in range(3):
                q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
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
