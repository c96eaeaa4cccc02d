"""Galactic-binary waveforms generated on the GPU (flr_gb_generate).

`GBWave` has the call signature of the generator in the reference's tests
(tests/test_fastlisaresponse.py:34-63) and returns h+ + i hx as a complex128 device tensor, so it
can drive `ResponseWrapper` with no host-side waveform. `generate_gb` writes many sources into one
packed response input in any storage format the batch / fused kernels take (complex128,
complex64 h / scale, or 32-bit fixed point), so that sources can be described by their
parameters alone (MCMC walkers, catalogues, served requests).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _capi
from ._capi import check, lib

YRSID_SI = 31558149.763545603

# flr_gb_source of include/flr.h
GB_SOURCE_DTYPE = np.dtype([
    ("offset", np.int64), ("num_pts", np.int32), ("reserved", np.int32),
    ("amplitude", np.float64), ("f0", np.float64), ("fdot", np.float64), ("fddot", np.float64),
    ("cos_iota", np.float64), ("cos2psi", np.float64), ("sin2psi", np.float64),
    ("phi0", np.float64), ("t0", np.float64), ("scale", np.float64),
])
assert GB_SOURCE_DTYPE.itemsize == 96

FORMATS = {"complex128": _capi.FLR_GB_COMPLEX128, "complex64": _capi.FLR_GB_COMPLEX64,
           "q31": _capi.FLR_GB_Q31}
Q31_FULL = float(2 ** 31 - 1)


def gb_table(params: np.ndarray, lengths: Sequence[int], t0s: Optional[Sequence[float]] = None,
             fmt: str = "complex128",
             offsets: Optional[Sequence[int]] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Source records for `params` rows (A, f0, fdot, iota, phi0, psi) at `offsets` (default:
    packed back to back), and the response scales of the stored format: 1 for complex128;
    s = A (1 + cos^2 i) -- a bound on |h+| and |hx| -- for complex64 (h / s stored) and
    s / (2^31 - 1) for q31."""
    params = np.asarray(params, dtype=np.float64).reshape(-1, 6)
    n = len(params)
    lengths = np.asarray(lengths, dtype=np.int64)
    if len(lengths) != n:
        raise ValueError("one length per source")
    rec = np.zeros(n, dtype=GB_SOURCE_DTYPE)
    A, f0, fdot, iota, phi0, psi = params.T
    ci = np.cos(iota)
    s = A * (1.0 + ci * ci)
    if offsets is None:
        rec["offset"] = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    else:
        offsets = np.asarray(offsets, dtype=np.int64)
        if len(offsets) != n or np.any(np.diff(offsets) < lengths[:-1]) or np.any(offsets < 0):
            raise ValueError("offsets must be increasing and leave room for each source")
        rec["offset"] = offsets
    rec["num_pts"] = lengths
    rec["amplitude"], rec["f0"], rec["fdot"] = A, f0, fdot
    rec["fddot"] = 11.0 / 3.0 * fdot ** 2 / f0
    rec["cos_iota"] = ci
    rec["cos2psi"], rec["sin2psi"] = np.cos(2.0 * psi), np.sin(2.0 * psi)
    rec["phi0"] = phi0
    rec["t0"] = 0.0 if t0s is None else np.asarray(t0s, dtype=np.float64)
    rec["scale"] = s
    scales = {"complex128": np.ones(n), "complex64": s, "q31": s / Q31_FULL}[fmt]
    return rec, scales


def generate_gb(params, lengths: Sequence[int], dt: float, fmt: str = "complex128",
                t0s: Optional[Sequence[float]] = None, out: Optional[torch.Tensor] = None,
                stream: Optional[torch.cuda.Stream] = None, device="cuda",
                offsets: Optional[Sequence[int]] = None):
    """Waveforms of the sources (rows of params: A, f0, fdot, iota, phi0, psi) in `fmt`, packed
    back to back (or at `offsets` in `out`, other samples untouched); returns (tensor, scales).
    `out`, if given, must have the format's dtype. The records are uploaded and the generator
    enqueued on `stream` (default: the current one)."""
    if fmt not in FORMATS:
        raise ValueError(f"fmt must be one of {sorted(FORMATS)}")
    rec, scales = gb_table(params, lengths, t0s, fmt, offsets)
    per = 2 if fmt == "q31" else 1
    dtype = {"complex128": torch.complex128, "complex64": torch.complex64, "q31": torch.int32}[fmt]
    if offsets is None:
        total = int(np.sum(lengths))
    else:
        if out is None:
            raise ValueError("offsets need an output tensor")
        total = out.numel() // per
    end = int(rec["offset"][-1] + rec["num_pts"][-1]) if len(rec) else 0
    if out is not None and (out.dtype != dtype or not out.is_contiguous() or
                            out.numel() != per * total or end > total):
        raise ValueError(f"out must be a contiguous {dtype} tensor holding every source")
    numel = per * total
    stream = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(stream):  # table allocated, filled and read in this stream's order
        if out is None:
            out = torch.empty(numel, dtype=dtype, device=device)
        table = torch.from_numpy(rec.view(np.uint8).copy()).to(out.device, non_blocking=False)
        check(lib().flr_gb_generate(ctypes.c_void_p(table.data_ptr()), len(rec), total, float(dt),
                                    FORMATS[fmt], ctypes.c_void_p(out.data_ptr()),
                                    ctypes.c_void_p(stream.cuda_stream)))
    return out, scales


class GBWave:
    """h+ + i hx of a galactic binary on the GPU, with the reference test generator's
    signature: GBWave()(A, f, fdot, iota, phi0, psi, T=1.0, dt=10.0) over t = arange(0, T yr,
    dt)."""

    def __call__(self, A, f, fdot, iota, phi0, psi, T=1.0, dt=10.0):
        n = len(np.arange(0.0, T * YRSID_SI, dt))
        h, _ = generate_gb(np.array([[A, f, fdot, iota, phi0, psi]]), [n], dt)
        return h
