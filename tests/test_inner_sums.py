"""The host side of the fused inner products (no GPU): the overlap, inner products and
log-likelihood that utility.py derives from the time-domain sums agree with get_overlap's rfft
arithmetic (utils/utility.py:16-82) and the direct chi^2, for even and odd lengths."""
import numpy as np
import pytest
import torch

from fastlisaresponse_amd.utility import (get_overlap, inner_from_sums, log_likelihood_from_sums,
                                          overlap_from_sums)


def _sums(d, h):
    alt = np.where(np.arange(h.shape[1]) % 2 == 0, 1.0, -1.0)
    s = np.stack([(d * h).sum(1), (h * h).sum(1), h.sum(1), (h * alt).sum(1)], axis=1)
    ds = np.stack([(d * d).sum(1), d.sum(1), (d * alt).sum(1)], axis=1)
    return torch.as_tensor(s), torch.as_tensor(ds)


@pytest.mark.parametrize("N", [4096, 4097, 1000, 999])
def test_overlap_from_sums_is_get_overlap(N):
    rng = np.random.default_rng(N)
    t = np.arange(N)
    h = np.stack([np.sin(0.03 * t + p) + 0.2 for p in (0.1, 0.7, 1.9)])
    h[:, ::2] += 0.05  # a Nyquist component
    d = h + 0.5 * rng.standard_normal(h.shape) + 0.1
    s, ds = _sums(d, h)
    got = float(overlap_from_sums(s, ds, N))
    want = get_overlap(list(d), list(h), use_gpu=False)
    assert abs(got - want) <= 1e-12, (got, want)


def test_inner_products_and_likelihood():
    rng = np.random.default_rng(3)
    h = rng.standard_normal((3, 777))
    d = h + rng.standard_normal((3, 777))
    s, ds = _sums(d, h)
    w = np.array([1.0, 0.5, 3.0])
    ip = inner_from_sums(s, torch.as_tensor(w)).numpy()
    assert np.isclose(ip[0], ((d * h).sum(1) / w).sum(), rtol=1e-13)
    assert np.isclose(ip[1], ((h * h).sum(1) / w).sum(), rtol=1e-13)
    ll = float(log_likelihood_from_sums(s, ds, torch.as_tensor(w)))
    assert np.isclose(ll, -0.5 * (((d - h) ** 2).sum(1) / w).sum(), rtol=1e-12)
    assert np.isclose(float(log_likelihood_from_sums(s, ds)), -0.5 * ((d - h) ** 2).sum(),
                      rtol=1e-12)
