"""On-device galactic-binary generator (fastlisaresponse_amd.waveforms), on the GPU.

The generator follows the reference test's formula (tests/test_fastlisaresponse.py:41-63) in
FP64; against numpy the phases agree to an ulp or two and sin/cos to an ulp, so h agrees to
~1e-10 of full scale at 1-yr phases (pinned at 1e-9). The complex64 and q31 outputs are the
FP64 values divided by the scale and rounded -- checked bit for bit against that conversion of
the complex128 output.
"""
import numpy as np
import pytest

from conftest import YRSID_SI, gb_waveform, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _params(rng, n):
    return np.stack([10 ** rng.uniform(-23, -21, n), 10 ** rng.uniform(-4, -2, n),
                     10 ** rng.uniform(-19, -15, n), np.arccos(rng.uniform(-1, 1, n)),
                     rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)], axis=1)


def test_gb_generator_matches_numpy(torch_gpu):
    from fastlisaresponse_amd.waveforms import generate_gb
    rng = np.random.default_rng(61)
    params = _params(rng, 5)
    lengths = [100000, 3000, 250000, 1, 77777]
    t0s = [0.0, 1.0e6, 2.5e7, 0.0, 3.0e7]
    h, scales = generate_gb(params, lengths, 10.0, t0s=t0s)
    got = h.cpu().numpy()
    o = 0
    for s, n in enumerate(lengths):
        A, f, fd, iota, phi0, psi = params[s]
        exp = gb_waveform(n, 10.0, A, f, fd, iota, phi0, psi, t0=t0s[s])
        full = A * (1 + np.cos(iota) ** 2)
        assert np.abs(got[o:o + n] - exp).max() <= 1e-9 * full, s
        o += n
    assert np.all(scales == 1.0)


def test_gb_storage_formats_are_the_scaled_fp64_values(torch_gpu):
    torch = torch_gpu
    from fastlisaresponse_amd.waveforms import Q31_FULL, generate_gb
    rng = np.random.default_rng(67)
    params = _params(rng, 4)
    lengths = [40000, 12345, 1000, 65536]
    h64, _ = generate_gb(params, lengths, 5.0)
    h32, s32 = generate_gb(params, lengths, 5.0, fmt="complex64")
    hq, sq = generate_gb(params, lengths, 5.0, fmt="q31")
    bound = np.repeat(params[:, 0] * (1 + np.cos(params[:, 3]) ** 2), lengths)
    b = torch.as_tensor(bound, device="cuda")
    re, im = h64.real, h64.imag
    assert float((torch.maximum(re.abs(), im.abs()) / b).max()) <= 1.0  # the scale bounds h
    exp32 = torch.complex(re / b, im / b).to(torch.complex64)
    assert torch.equal(h32, exp32)
    expq = torch.stack([torch.round(re / b * Q31_FULL), torch.round(im / b * Q31_FULL)], dim=1)
    assert torch.equal(hq.view(-1, 2), expq.to(torch.int32))
    np.testing.assert_allclose(s32, bound[np.cumsum([0] + lengths[:-1])])
    np.testing.assert_allclose(sq * Q31_FULL, s32)


def test_gbwave_drives_response_wrapper(torch_gpu, oracle_mod):
    """GBWave (device) as ResponseWrapper's generator, against numpy's waveform + the oracle."""
    from fastlisaresponse_amd import ResponseWrapper
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.waveforms import GBWave
    orbits = EqualArmlengthOrbits(t_max=0.2 * YRSID_SI)
    orbits.configure(linear_interp_setup=True)
    n, dt = 1 << 17, 10.0
    T = n * dt / YRSID_SI
    gen = GBWave()

    def gb(*a, **k):  # arange(0, T yr, dt) may round to n + 1 samples
        return gen(*a, **k)[:n]

    wrap = ResponseWrapper(gb, T, dt, 6, 7, remove_sky_coords=True, n_overide=n,
                           orbits=orbits, tdi="1st generation", tdi_chan="AET")
    p = (1e-21, 2e-3, 1e-17, 0.8, 0.3, 0.5)
    chans = np.stack([c.cpu().numpy() for c in wrap(*p, 1.2, 0.3)])
    oo = oracle_mod.OracleOrbits(*orbits.pycppdetector_args)
    _, aet = oracle_mod.response_pipeline(oo, gb_waveform(n, dt, *p), 1.2, 0.3, dt, 25,
                                          "1st generation", x_base=orbits.x_base, nthreads=8)
    for c in range(3):
        assert rel_l2(chans[c], aet[c, 1000:-1000]) <= 1e-8, c
