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


class BatchedResponse:
    """Response configuration shared by a batch: sampling, order, TDI, orbits, channels.

    Args follow pyResponseTDI (sampling_frequency, order, tdi, orbits, tdi_chan, exact) plus
    t_buffer (s), the edge length left out of TDI for every source."""

    def __init__(self, sampling_frequency: float, order: int = 25, tdi="2nd generation",
                 orbits=None, tdi_chan: str = "AET", exact: bool = True, t_buffer: float = 10000.0):
        if tdi_chan not in ("XYZ", "AET", "AE"):
            raise ValueError("tdi_chan must be 'XYZ', 'AET' or 'AE'.")
        self.model = pyResponseTDI(sampling_frequency, 1, order=order, tdi=tdi, orbits=orbits,
                                   tdi_chan=tdi_chan, exact=exact)
        self.sampling_frequency = sampling_frequency
        self.dt = self.model.dt
        self.order = order
        self.tdi_chan = tdi_chan
        self.exact = exact
        self.t_buffer = t_buffer
        (self.tdi_start_ind, _, self.projection_buffer,
         self.projections_start_ind) = buffer_indices(sampling_frequency, order, t_buffer,
                                                      self.model.response_orbits.x_base)
        if self.projections_start_ind < self.projection_buffer:
            raise ValueError("Need to increase t_buffer. The initial buffer is not large enough.")

    def plan(self, h_all: torch.Tensor, lengths: Sequence[int], lams: Sequence[float],
             betas: Sequence[float], t0s: Optional[Sequence[float]] = None) -> "BatchPlan":
        """Bind a packed input (sources back to back in h_all) and allocate the outputs."""
        return BatchPlan(self, h_all, lengths, lams, betas, t0s)


class BatchPlan:
    """A bound batch: device source table, packed projections y_all and outputs out_all."""

    def __init__(self, br: BatchedResponse, h_all, lengths, lams, betas, t0s=None):
        lengths = np.asarray(lengths, dtype=np.int64)
        n = len(lengths)
        if n < 1 or len(lams) != n or len(betas) != n or (t0s is not None and len(t0s) != n):
            raise ValueError("lengths, lams, betas (and t0s) must have one entry per source")
        if np.any(lengths < 1):
            raise ValueError("every source needs at least one sample")
        if not (isinstance(h_all, torch.Tensor) and h_all.is_cuda and h_all.dtype == torch.complex128
                and h_all.is_contiguous()):
            raise TypeError("h_all must be a contiguous complex128 tensor on the GPU")
        total = int(lengths.sum())
        if h_all.numel() != total:
            raise ValueError(f"h_all holds {h_all.numel()} samples, the lengths sum to {total}")
        self.br = br
        self.h_all = h_all
        self.lengths = lengths
        self.offsets = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
        dev = h_all.device
        self.y_all = torch.empty(6 * total, dtype=torch.float64, device=dev)
        self.out_all = torch.empty(3 * total, dtype=torch.float64, device=dev)
        srcs = (_capi.FlrSource * n)()
        for s in range(n):
            k, u, v = sky_basis(lams[s], betas[s])
            q = srcs[s]
            q.offset = int(self.offsets[s])
            q.num_pts = int(lengths[s])
            q.projections_start_ind = br.projections_start_ind
            q.tdi_start_ind = br.tdi_start_ind
            q.t0 = 0.0 if t0s is None else float(t0s[s])
            for c in range(3):
                q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
        flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE
        if br.tdi_chan != "XYZ":
            flags |= _capi.FLR_FLAG_AET
        if br.exact:
            flags |= _capi.FLR_FLAG_EXACT
        h = ctypes.c_void_p()
        check(lib().flr_batch_create(br.model.cpp_response._h, srcs, n, h_all.data_ptr(),
                                     self.y_all.data_ptr(), self.out_all.data_ptr(), br.dt,
                                     br.sampling_frequency, flags, ctypes.byref(h)))
        self._h = h
        self.num_sources = n
        self.samples = int(lib().flr_batch_samples(h))

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
