"""Channel helpers: XYZ -> AET and the FFT overlap used by the reference's tests."""
from __future__ import annotations

import numpy as np
import torch


def AET(X, Y, Z):
    """A = (Z - X)/sqrt(2), E = (X - 2Y + Z)/sqrt(6), T = (X + Y + Z)/sqrt(3).

    The convention of lisatools.utils.utility.AET (called at response.py:665); the TDI kernel
    applies the same arithmetic in its epilogue when tdi_chan is "AET" or "AE"."""
    return ((Z - X) / np.sqrt(2.0), (X - 2.0 * Y + Z) / np.sqrt(6.0), (X + Y + Z) / np.sqrt(3.0))


def get_overlap(sig1, sig2, phase_maximize: bool = False) -> float:
    """Normalised overlap of two channel sets, averaged over channels (utils/utility.py:16-82).

    Accepts lists (or stacked arrays) of torch tensors or numpy arrays."""

    def chans(s):
        if isinstance(s, (list, tuple)):
            return [np.asarray(c.detach().cpu().numpy() if isinstance(c, torch.Tensor) else c) for c in s]
        a = s.detach().cpu().numpy() if isinstance(s, torch.Tensor) else np.asarray(s)
        return [a] if a.ndim < 2 else list(a)

    s1, s2 = chans(sig1), chans(sig2)
    if len(s1) != len(s2) or len(s1[0]) != len(s2[0]):
        raise ValueError("sig1 and sig2 must have the same channels and lengths")
    overlap = 0.0 + 0.0j
    for a, b in zip(s1, s2):
        fa, fb = np.fft.rfft(a), np.fft.rfft(b)
        overlap += np.dot(fa.conj(), fb) / np.sqrt(np.dot(fa.conj(), fa) * np.dot(fb.conj(), fb))
    overlap /= len(s1)
    return float(np.abs(overlap)) if phase_maximize else float(overlap.real)
