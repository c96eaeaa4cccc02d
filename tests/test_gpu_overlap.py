"""get_overlap on the device (torch.fft) against the host form (numpy), on the GPU.

The host form is pinned on the CPU (tests/test_host_layer.py: closed forms and a brute-force
DFT); here the device form must agree with it to rounding on long channels, float32 channels
included (they are widened to FP64 before the transform, as the host form does)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_device_overlap_matches_host(torch_gpu):
    torch = torch_gpu
    from fastlisaresponse_amd.utility import get_overlap
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(3, 100003, dtype=torch.float64, device="cuda", generator=g)
    b = a + 0.3 * torch.randn(3, 100003, dtype=torch.float64, device="cuda", generator=g)
    b = torch.roll(b, 7, dims=1)
    an, bn = a.cpu().numpy(), b.cpu().numpy()
    for pm in (False, True):
        host = get_overlap(an, bn, phase_maximize=pm)
        assert abs(get_overlap(a, b, phase_maximize=pm) - host) < 1e-12, pm
        assert abs(get_overlap(list(a), list(b), phase_maximize=pm) - host) < 1e-12, pm
        assert abs(get_overlap(an, bn, phase_maximize=pm, use_gpu=True) - host) < 1e-12, pm
    a32, b32 = a.float(), b.float()
    host32 = get_overlap(a32.cpu().numpy(), b32.cpu().numpy())
    assert abs(get_overlap(a32, b32) - host32) < 1e-12
