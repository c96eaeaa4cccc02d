"""Fused inner products against a resident data series (BatchPlan.inner_products,
flr_fused_inner): the mixed kernels sum d h, h^2, h and (-1)^j h per source and channel instead
of storing the response.

Checked against numpy sums over the FP64 fast form's outputs of the same sources (equal arms:
the same FP64 arithmetic up to rounding, so within 1e-9), against numpy sums over the same
mixed plan's stored float32 outputs (both arms: the float32 rounding only), and through
utility.overlap_from_sums against get_overlap's rfft overlap (utils/utility.py:16-82) of the
data and the FP64 outputs. Deterministic: repeated calls, other grid densities and a captured
graph give the same bits.
"""
import numpy as np
import pytest

from conftest import gb_waveform

pytestmark = pytest.mark.gpu

T_BUF = 1000


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module", params=["equal", "unequal"])
def orbits(request):
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits
    cls = EqualArmlengthOrbits if request.param == "equal" else CartwheelOrbits
    o = cls(t_max=4.0e7)
    o.configure(linear_interp_setup=True)
    return o


LENGTHS = [70000, 23456, 41000]
T0S = [0.0, 2.5e7, 1.0e6]
FREQS = [3e-3, 6e-3, 1.5e-3]


def _setup(torch, orbits, tdi_chan="AET", precision="mixed-c128"):
    from fastlisaresponse_amd.batch import BatchedResponse
    hs = [gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.7, 0.2, 0.4, t0=t0)
          for n, f, t0 in zip(LENGTHS, FREQS, T0S)]
    lam, beta = [0.4, 2.2, 4.9], [0.3, -0.9, 1.1]
    h = torch.as_tensor(np.concatenate(hs), device="cuda")
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan=tdi_chan,
                         precision=precision, storage_check="off")
    if precision == "mixed":
        from fastlisaresponse_amd.batch import pack_mixed
        h32, sc = pack_mixed(h, LENGTHS)
        plan = br.plan(h32, LENGTHS, lam, beta, T0S, scales=sc)
    else:
        plan = br.plan(h, LENGTHS, lam, beta, T0S)
    f64 = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan=tdi_chan, fused=True,
                          exact=False).plan(h, LENGTHS, lam, beta, T0S)
    f64.run()
    f64.check_errors()
    # data: a response with a slightly different sky position plus white noise, float32
    alt = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan=tdi_chan, fused=True,
                          exact=False).plan(h, LENGTHS, [x + 0.01 for x in lam], beta, T0S)
    alt.run()
    g = torch.Generator(device="cuda").manual_seed(11)
    noise = torch.randn(alt.out_all.shape, generator=g, device="cuda", dtype=torch.float64)
    data = (alt.out_all + 0.02 * alt.out_all.abs().max() * noise).to(plan.out_all.dtype)
    return plan, f64, data


def _numpy_sums(out, data, plan):
    res = np.zeros((len(LENGTHS), 3, 4))
    for s, n in enumerate(LENGTHS):
        o = int(plan.offsets[s])
        h = out[3 * o:3 * (o + n)].reshape(3, n)[:, T_BUF:n - T_BUF].astype(np.float64)
        d = data[3 * o:3 * (o + n)].reshape(3, n)[:, T_BUF:n - T_BUF].astype(np.float64)
        alt = np.where(np.arange(h.shape[1]) % 2 == 0, 1.0, -1.0)
        res[s] = np.stack([(d * h).sum(1), (h * h).sum(1), h.sum(1), (h * alt).sum(1)], axis=1)
    return res


def _scale(plan, out):
    """Per source and channel: sqrt(N) * rms(h), the size of a sum of N terms of h."""
    sc = np.zeros((len(LENGTHS), 3))
    for s, n in enumerate(LENGTHS):
        o = int(plan.offsets[s])
        h = out[3 * o:3 * (o + n)].reshape(3, n)[:, T_BUF:n - T_BUF].astype(np.float64)
        sc[s] = np.sqrt((h * h).sum(1))
    return sc


@pytest.mark.parametrize("tdi_chan", ["AET", "XYZ"])
def test_inner_sums_match_the_stored_outputs(torch_gpu, orbits, tdi_chan):
    torch = torch_gpu
    plan, f64, data = _setup(torch, orbits, tdi_chan)
    sums = plan.inner_products(data).cpu().numpy()
    plan.check_errors()
    plan.run()
    plan.check_errors()
    dn = data.cpu().numpy()
    stored = _numpy_sums(plan.out_all.cpu().numpy(), dn, plan)  # float32 outputs
    scale = _scale(plan, plan.out_all.cpu().numpy())
    dscale = np.array([[np.sqrt((dn[3 * int(plan.offsets[s]):3 * (int(plan.offsets[s]) + n)]
                                 .astype(np.float64) ** 2).sum())] * 3
                       for s, n in enumerate(LENGTHS)])
    # float32 rounding of the stored outputs (2^-24 relative per sample)
    tol = 1e-7
    assert np.all(np.abs(sums[..., 0] - stored[..., 0]) <= tol * scale * dscale), (sums, stored)
    assert np.all(np.abs(sums[..., 1] - stored[..., 1]) <= tol * scale ** 2)
    assert np.all(np.abs(sums[..., 2:] - stored[..., 2:]) <= tol * scale[..., None] * 10)
    if orbits.ltt.min() == orbits.ltt.max():
        # equal arms: against the FP64 fast form (same arithmetic up to rounding); the null
        # channel T of AET is compared on the scale of A and E
        exact = _numpy_sums(f64.out_all.cpu().numpy(), dn, plan)
        ae = np.maximum(scale, scale[:, :1] if tdi_chan == "XYZ" else scale[:, :2].max(1, keepdims=True))
        assert np.all(np.abs(sums[..., 0] - exact[..., 0]) <= 1e-9 * ae * dscale)
        assert np.all(np.abs(sums[..., 1] - exact[..., 1]) <= 1e-9 * ae ** 2)


def test_overlap_and_likelihood_from_sums(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.utility import (data_sums, get_overlap, log_likelihood_from_sums,
                                              overlap_from_sums)
    plan, f64, data = _setup(torch, orbits)
    sums = plan.inner_products(data)
    ds = data_sums(data, LENGTHS, plan.offsets, T_BUF)
    for s, n in enumerate(LENGTHS):
        N = n - 2 * T_BUF
        got = float(overlap_from_sums(sums[s], ds[s], N))
        o = int(plan.offsets[s])
        d = data[3 * o:3 * (o + n)].view(3, n)[:, T_BUF:n - T_BUF].double().cpu().numpy()
        h = f64.out_all[3 * o:3 * (o + n)].view(3, n)[:, T_BUF:n - T_BUF].cpu().numpy()
        want = get_overlap(list(d), list(h), use_gpu=False)
        assert 0.05 < abs(want) < 1.0
        # equal arms: the same FP64 arithmetic up to rounding; unequal arms: the null channel's
        # delays differ from the FP64 form's by its per-sample time rounding
        tol = 1e-8 if orbits.ltt.min() == orbits.ltt.max() else 1e-6
        assert abs(got - want) <= tol * abs(want), (s, got, want)
        # white-noise log-likelihood against the direct chi^2 of the FP64 outputs
        ll = float(log_likelihood_from_sums(sums[s], ds[s], noise_var=[1.0, 2.0, 4.0]))
        w = np.array([1.0, 2.0, 4.0])[:, None]
        want_ll = -0.5 * float((((d - h) ** 2) / w).sum())
        assert abs(ll - want_ll) <= 1e-6 * abs(want_ll), (s, ll, want_ll)


def test_inner_products_are_deterministic(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.autotune import apply
    plan, _, data = _setup(torch, orbits)
    a = plan.inner_products(data).clone()
    b = plan.inner_products(data).clone()
    assert torch.equal(a, b)
    apply(plan, plan.launch_shape[0], 7)  # another grid density, same tiles
    c = plan.inner_products(data).clone()
    assert torch.equal(a, c)
    # graph capture (after the first call allocated the per-tile buffer), replayed with eager
    # work and synchronisations between replays
    sums = torch.empty_like(a)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.inner_products(data, sums)
    torch.cuda.current_stream().wait_stream(s)
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gr):
        plan.inner_products(data, sums)
    for _ in range(3):
        sums.fill_(float("nan"))
        gr.replay()
        torch.cuda.synchronize()
        assert torch.equal(sums, a)
        plan.run()
        torch.cuda.synchronize()


def test_inner_products_other_inputs_and_errors(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    plan, _, data = _setup(torch, orbits, precision="mixed")
    sums = plan.inner_products(data).cpu().numpy()
    plan.run()
    plan.check_errors()
    stored = _numpy_sums(plan.out_all.cpu().numpy(), data.cpu().numpy(), plan)
    scale = _scale(plan, plan.out_all.cpu().numpy())
    assert np.all(np.abs(sums[..., 1] - stored[..., 1]) <= 1e-7 * scale ** 2)
    with pytest.raises(TypeError):
        plan.inner_products(data.double())
    with pytest.raises(TypeError):
        plan.inner_products(data[:-3])
    h = torch.zeros(20000, dtype=torch.complex128, device="cuda")
    fp = BatchedResponse(0.1, orbits=orbits, fused=True, exact=False).plan(h, [20000], [0.1], [0.2])
    with pytest.raises(ValueError):
        fp.inner_products(torch.zeros_like(fp.out_all))
