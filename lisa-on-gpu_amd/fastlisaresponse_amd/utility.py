"""Channel helpers: XYZ -> AET and the FFT overlap used by the reference's tests."""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch


def AET(X, Y, Z):
    """A = (Z - X)/sqrt(2), E = (X - 2Y + Z)/sqrt(6), T = (X + Y + Z)/sqrt(3).

    The convention of lisatools.utils.utility.AET (called at response.py:665); the TDI kernel
    applies the same arithmetic in its epilogue when tdi_chan is "AET" or "AE"."""
    return ((Z - X) / np.sqrt(2.0), (X - 2.0 * Y + Z) / np.sqrt(6.0), (X + Y + Z) / np.sqrt(3.0))


def _channels(sig, device: bool):
    """A list of 1-D channels: a list/tuple of channels, or one array of shape (n,) or
    (channels, n)."""
    if isinstance(sig, (list, tuple)):
        items = list(sig)
    else:
        if not isinstance(sig, (torch.Tensor, np.ndarray)):
            raise ValueError("signals must be a list of arrays or a single array")
        items = [sig] if sig.ndim < 2 else list(sig)
    if device:
        return [c.to(torch.float64) if isinstance(c, torch.Tensor) and c.is_cuda
                else torch.as_tensor(np.asarray(c), dtype=torch.float64, device="cuda")
                for c in items]
    return [np.asarray(c.detach().cpu().numpy() if isinstance(c, torch.Tensor) else c,
                       dtype=np.float64) for c in items]


def get_overlap(sig1, sig2, phase_maximize: bool = False, use_gpu: Optional[bool] = None) -> float:
    """Normalised overlap of two channel sets, averaged over channels (utils/utility.py:16-82):
    sum_c <rfft s1_c, rfft s2_c> / sqrt(<s1_c, s1_c> <s2_c, s2_c>) / channels, its real part, or
    its modulus when phase_maximize.

    Accepts lists (or stacked arrays) of torch tensors or numpy arrays. use_gpu=None runs on the
    device when any channel is a GPU tensor; True moves host arrays to the GPU; False runs numpy
    on the host. The device form is the same arithmetic in FP64 through torch.fft (rocFFT)."""
    if use_gpu is None:
        def on_gpu(s):
            if isinstance(s, (list, tuple)):
                return any(isinstance(c, torch.Tensor) and c.is_cuda for c in s)
            return isinstance(s, torch.Tensor) and s.is_cuda
        use_gpu = on_gpu(sig1) or on_gpu(sig2)
    s1, s2 = _channels(sig1, use_gpu), _channels(sig2, use_gpu)
    if len(s1) != len(s2) or len(s1[0]) != len(s2[0]):
        raise ValueError("sig1 and sig2 must have the same channels and lengths")
    if use_gpu:
        a, b = torch.stack(s1), torch.stack(s2)
        fa, fb = torch.fft.rfft(a, dim=1), torch.fft.rfft(b, dim=1)
        ab = (fa.conj() * fb).sum(dim=1)
        aa = (fa.conj() * fa).sum(dim=1)
        bb = (fb.conj() * fb).sum(dim=1)
        overlap = complex((ab / torch.sqrt(aa * bb)).sum().item()) / len(s1)
    else:
        overlap = 0.0 + 0.0j
        for a, b in zip(s1, s2):
            fa, fb = np.fft.rfft(a), np.fft.rfft(b)
            overlap += np.dot(fa.conj(), fb) / np.sqrt(np.dot(fa.conj(), fa) * np.dot(fb.conj(), fb))
        overlap /= len(s1)
    return float(abs(overlap)) if phase_maximize else float(overlap.real)
