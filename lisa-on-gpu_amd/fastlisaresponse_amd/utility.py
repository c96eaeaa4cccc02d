"""Channel helpers: XYZ -> AET, the FFT overlap used by the reference's tests, and the
quantities that follow from the fused inner-product sums (BatchPlan.inner_products)."""
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


# ------------------------------------------------------------------ fused inner products
# BatchPlan.inner_products returns, per source and channel, over the N = n - 2*tdi_start
# samples of [tdi_start, n - tdi_start): S0 = sum d h, S1 = sum h^2, S2 = sum h and
# S3 = sum (-1)^j h (j counted from tdi_start). For real series a, b of length N,
#   sum_{k=0}^{floor(N/2)} conj(rfft(a)_k) rfft(b)_k
# has real part (N sum a b + a_0 b_0 + [N even] a_N b_N) / 2 (Parseval, with the DC bin
# a_0 = sum a and the Nyquist bin a_N = sum (-1)^j a counted once), which is how get_overlap's
# rfft inner products follow from time-domain sums without the series.


def data_sums(data, lengths, offsets, tdi_start: int) -> torch.Tensor:
    """sum d^2, sum d, sum (-1)^j d per source and channel of a data series laid out as a plan's
    out_all ([sources, 3, 3], float64), over the same samples as the fused sums."""
    out = []
    for o, n in zip(offsets, lengths):
        d = data[3 * int(o): 3 * (int(o) + int(n))].view(3, int(n))[:, tdi_start:int(n) - tdi_start]
        d = d.to(torch.float64)
        alt = torch.ones(d.shape[1], dtype=torch.float64, device=d.device)
        alt[1::2] = -1.0
        out.append(torch.stack([(d * d).sum(1), d.sum(1), (d * alt).sum(1)], dim=1))
    return torch.stack(out)


def _rfft_products(N, S_ab, a0, b0, aN, bN):
    """Real part of the rfft inner product of two length-N series from their time-domain
    product sum and DC / Nyquist bins."""
    nyq = (aN * bN) if (N % 2 == 0) else 0.0
    return 0.5 * (N * S_ab + a0 * b0 + nyq)


def overlap_from_sums(sums, dsums, N: int) -> torch.Tensor:
    """get_overlap(data, response) (utils/utility.py:16-82, phase_maximize=False) per source,
    from the fused sums and data_sums over N samples: the channel average of
    Re<d, h> / sqrt(<d, d> <h, h>) with <., .> the rfft inner product."""
    S0, S1, S2, S3 = sums.unbind(-1)
    D1, D2, D3 = dsums.unbind(-1)
    dh = _rfft_products(N, S0, D2, S2, D3, S3)
    hh = _rfft_products(N, S1, S2, S2, S3, S3)
    dd = _rfft_products(N, D1, D2, D2, D3, D3)
    return (dh / torch.sqrt(dd * hh)).mean(-1)


def inner_from_sums(sums, noise_var=1.0) -> torch.Tensor:
    """White-noise inner products (d|h) = sum_c sum_i d h / sigma_c^2 and (h|h) per source
    ([sources, 2]); noise_var: sigma^2, a scalar or one value per channel."""
    w = torch.as_tensor(noise_var, dtype=torch.float64, device=sums.device).reshape(-1)
    w = w.expand(3) if w.numel() == 1 else w
    return torch.stack([(sums[..., 0] / w).sum(-1), (sums[..., 1] / w).sum(-1)], dim=-1)


def log_likelihood_from_sums(sums, dsums, noise_var=1.0) -> torch.Tensor:
    """Gaussian white-noise log-likelihood -1/2 sum_c sum_i (d - h)^2 / sigma_c^2 per source
    (up to the normalisation constant), from the fused sums and data_sums."""
    w = torch.as_tensor(noise_var, dtype=torch.float64, device=sums.device).reshape(-1)
    w = w.expand(3) if w.numel() == 1 else w
    chi2 = (dsums[..., 0] - 2.0 * sums[..., 0] + sums[..., 1]) / w
    return -0.5 * chi2.sum(-1)
