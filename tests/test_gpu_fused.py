"""Fused projection + TDI launches (y kept in LDS) and the mixed-precision form, on the GPU.

* Fused exact FP64 is bitwise the unfused exact path (and so the FP64 oracle).
* Fused fast FP64 (one stencil per output and chain length for equal arms) is within the fast
  form's tolerance of the oracle.
* Mixed precision stores h as complex64 h / scale and the outputs as float32; inside, times,
  delays and stencils are FP64. It is pinned two ways: against the oracle run on the same
  complex64-rounded input (isolates the kernel: only the fast form's ~1e-15 and the float32
  output rounding remain, rel-L2 <= 2e-7 per channel), and against the oracle on the FP64
  input (the storage error: rel-L2 <= 1e-6 on A and E at 5 mHz). The 32-bit fixed-point
  input form ("mixed-q31") is pinned the same two ways. The complex128-input form
  ("mixed-c128") has no storage error: against the oracle on the same FP64 input only the
  float32 output rounding (~3e-8 rms relative) and the composite filters' delay tolerance remain.
"""
import numpy as np
import pytest

from conftest import YRSID_SI, check_mixed, gb_waveform, rel_l2, unequal_arms

pytestmark = pytest.mark.gpu

TOL_FAST = 1e-12        # fast form vs oracle, A and E (T: absolute, relative to the A/E scale)
TOL_MIXED_KERNEL = 2e-7  # mixed kernel vs oracle on the same complex64-rounded input
TOL_MIXED_5MHZ = 1e-6    # mixed vs oracle on the FP64 input, A and E, f0 = 5 mHz


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
    o = cls(t_max=1.2 * YRSID_SI)
    o.configure(linear_interp_setup=True)
    return o


@pytest.fixture(scope="module")
def oracle_orbits(orbits, oracle_mod):
    return oracle_mod.OracleOrbits(*orbits.pycppdetector_args)


def _sources(rng, lengths):
    params, hs = [], []
    for n in lengths:
        f = 10 ** rng.uniform(-3, -2)
        fd = 10 ** rng.uniform(-19, -16)
        A = 10 ** rng.uniform(-23, -21)
        iota = np.arccos(rng.uniform(-1, 1))
        phi0, psi = rng.uniform(0, 2 * np.pi, 2)
        lam = rng.uniform(0, 2 * np.pi)
        beta = np.arcsin(rng.uniform(-1, 1))
        t0 = float(rng.choice([0.0, 1.0e6, 2.5e7]))
        hs.append(gb_waveform(n, 10.0, A, f, fd, iota, phi0, psi, t0=t0))
        params.append((lam, beta, t0))
    return hs, params


def _plan(torch, br, hs, params, scales=None, h_all=None):
    if h_all is None:
        h_all = torch.as_tensor(np.concatenate(hs), device="cuda")
    return br.plan(h_all, [len(h) for h in hs], [p[0] for p in params], [p[1] for p in params],
                   [p[2] for p in params], scales=scales)


def _fast_ok(got, exp, aet):
    if not aet:
        for c in range(3):
            assert rel_l2(got[c], exp[c]) <= TOL_FAST, (c, rel_l2(got[c], exp[c]))
        return
    scale = max(np.linalg.norm(exp[0]), np.linalg.norm(exp[1]))
    for c in range(2):
        assert rel_l2(got[c], exp[c]) <= TOL_FAST, (c, rel_l2(got[c], exp[c]))
    assert np.linalg.norm(got[2] - exp[2]) <= TOL_FAST * scale


@pytest.mark.parametrize("tdi,chan", [("2nd generation", "AET"), ("1st generation", "XYZ")])
def test_fused_exact_is_bitwise_the_unfused_path(torch_gpu, orbits, oracle_orbits, oracle_mod,
                                                 tdi, chan):
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    rng = np.random.default_rng(31)
    lengths = [5000, 12345, 2100, 3000, 20000, 7777]  # 2100: only edges
    hs, params = _sources(rng, lengths)
    ref = _plan(torch, BatchedResponse(0.1, tdi=tdi, orbits=orbits, tdi_chan=chan), hs, params)
    fused = _plan(torch, BatchedResponse(0.1, tdi=tdi, orbits=orbits, tdi_chan=chan, fused=True),
                  hs, params)
    ref.run()
    fused.run()
    ref.check_errors()
    fused.check_errors()
    assert torch.equal(fused.out_all, ref.out_all)
    lam, beta, t0 = params[4]
    _, exp = oracle_mod.response_pipeline(oracle_orbits, hs[4], lam, beta, 10.0, 25, tdi, t0=t0,
                                          tdi_chan=chan, x_base=orbits.x_base, nthreads=8)
    np.testing.assert_array_equal(torch.stack(fused.channels(4)).cpu().numpy(), exp)


@pytest.mark.parametrize("tdi,chan", [("2nd generation", "AET"), ("1st generation", "XYZ")])
def test_fused_fast_fp64_vs_oracle(torch_gpu, orbits, oracle_orbits, oracle_mod, tdi, chan):
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    rng = np.random.default_rng(37)
    lengths = [40000, 9000]
    hs, params = _sources(rng, lengths)
    br = BatchedResponse(0.1, tdi=tdi, orbits=orbits, tdi_chan=chan, fused=True, exact=False)
    plan = _plan(torch, br, hs, params)
    plan.run()
    plan.check_errors()
    for s in range(len(lengths)):
        lam, beta, t0 = params[s]
        _, exp = oracle_mod.response_pipeline(oracle_orbits, hs[s], lam, beta, 10.0, 25, tdi,
                                              t0=t0, tdi_chan=chan, x_base=orbits.x_base,
                                              nthreads=8)
        got = torch.stack(plan.channels(s)).cpu().numpy()
        T = br.tdi_start_ind
        assert np.all(got[:, :T] == 0.0) and np.all(got[:, -T:] == 0.0)
        _fast_ok(got[:, T:-T], exp[:, T:-T], chan == "AET")


def test_mixed_kernel_vs_oracle_on_rounded_input(torch_gpu, orbits, oracle_orbits, oracle_mod):
    """The kernel itself: oracle on the same complex64 samples (times scale) in FP64."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    rng = np.random.default_rng(41)
    lengths = [60000, 11000, 25000]
    hs, params = _sources(rng, lengths)
    h64 = torch.as_tensor(np.concatenate(hs), device="cuda")
    h32, scales = pack_mixed(h64, lengths)
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan="AET",
                         precision="mixed")
    plan = _plan(torch, br, hs, params, scales=scales, h_all=h32)
    plan.run()
    plan.check_errors()
    assert plan.out_all.dtype == torch.float32
    h32_host = h32.cpu().numpy()
    o = 0
    for s, n in enumerate(lengths):
        lam, beta, t0 = params[s]
        hq = h32_host[o:o + n].astype(np.complex128) * scales[s]
        o += n
        y, exp = oracle_mod.response_pipeline(oracle_orbits, hq, lam, beta, 10.0, 25,
                                              "2nd generation", t0=t0, x_base=orbits.x_base,
                                              nthreads=8)
        got = torch.stack(plan.channels(s)).cpu().numpy().astype(np.float64)
        check_mixed(oracle_mod, oracle_orbits, got, y, exp, t0, 10.0, "2nd generation", "AET",
                    TOL_MIXED_KERNEL, 1000, (s,), unequal_arms(orbits))


@pytest.mark.parametrize("tdi,chan,dt", [("1st generation", "XYZ", 10.0),
                                         ("2nd generation", "AET", 5.0),
                                         ("1st generation", "AET", 5.0),
                                         ("2nd generation", "XYZ", 2.0)])
def test_mixed_generations_channels_and_sampling(torch_gpu, orbits, oracle_orbits, oracle_mod,
                                                 tdi, chan, dt):
    """Mixed kernels for both generations, both channel sets and sampling intervals where the
    composite filter is nearly full (dt = 5 s: 38 of 40 taps) or does not fit (dt = 2 s: every
    tile takes the per-sample path), against the oracle on the same complex64 input."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    rng = np.random.default_rng(53)
    n = 60000
    f = 10 ** rng.uniform(-3, -2)
    h = gb_waveform(n, dt, 1e-21, f, 1e-17, 0.7, 0.2, 0.4, t0=1.0e6)
    h32, scales = pack_mixed(torch.as_tensor(h, device="cuda"), [n])
    br = BatchedResponse(1.0 / dt, tdi=tdi, orbits=orbits, tdi_chan=chan, precision="mixed")
    plan = br.plan(h32, [n], [0.9], [-0.4], [1.0e6], scales=scales)
    plan.run()
    plan.check_errors()
    hq = h32.cpu().numpy().astype(np.complex128) * scales[0]
    y, exp = oracle_mod.response_pipeline(oracle_orbits, hq, 0.9, -0.4, dt, 25, tdi, t0=1.0e6,
                                          tdi_chan=chan, x_base=orbits.x_base, nthreads=8)
    got = torch.stack(plan.channels(0)).cpu().numpy().astype(np.float64)
    check_mixed(oracle_mod, oracle_orbits, got, y, exp, 1.0e6, dt, tdi, chan, TOL_MIXED_KERNEL,
                br.tdi_start_ind, (), unequal_arms(orbits))


def test_mixed_vs_fp64_oracle_at_5mhz(torch_gpu, orbits, oracle_orbits, oracle_mod):
    """End to end through pyResponseTDI(precision="mixed"): the FP32 storage error."""
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    n = 1 << 18
    h = gb_waveform(n, 10.0, 1e-21, 5e-3, 1e-16, 0.8, 0.3, 0.5)
    r = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                      precision="mixed")
    r.get_projections(h, 1.2, 0.3)
    A, E, T = r.get_tdi_delays()
    assert A.dtype == torch.float32
    got = torch.stack([A, E, T]).cpu().numpy().astype(np.float64)
    _, exp = oracle_mod.response_pipeline(oracle_orbits, h, 1.2, 0.3, 10.0, 25, "2nd generation",
                                          x_base=orbits.x_base, nthreads=8)
    for c in range(2):
        r_c = rel_l2(got[c, 1000:-1000], exp[c, 1000:-1000])
        assert r_c <= TOL_MIXED_5MHZ, (c, r_c)
    # the null channel T is ~1e3 below A, E here: its relative error is ~1e-4
    assert rel_l2(got[2, 1000:-1000], exp[2, 1000:-1000]) <= 1e-3
    # the XYZ view and y_gw are available (FP64, from the original input)
    xyz = r.XYZ.cpu().numpy().astype(np.float64)
    y_exp, xyz_exp = oracle_mod.response_pipeline(oracle_orbits, h, 1.2, 0.3, 10.0, 25,
                                                  "2nd generation", tdi_chan="XYZ",
                                                  x_base=orbits.x_base, nthreads=8)
    for c in range(3):
        assert rel_l2(xyz[c, 1000:-1000], xyz_exp[c, 1000:-1000]) <= TOL_MIXED_5MHZ
    y = r.y_gw.cpu().numpy()
    for l in range(6):
        assert rel_l2(y[l], y_exp[l]) <= TOL_FAST


def test_q31_kernel_vs_oracle_on_dequantised_input(torch_gpu, orbits, oracle_orbits, oracle_mod):
    """32-bit fixed-point input: the kernel against the oracle on the same quantised samples."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse, pack_q31
    rng = np.random.default_rng(47)
    lengths = [50000, 13000]
    hs, params = _sources(rng, lengths)
    hq_all, scales = pack_q31(torch.as_tensor(np.concatenate(hs), device="cuda"), lengths)
    assert hq_all.dtype == torch.int32 and hq_all.numel() == 2 * sum(lengths)
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan="AET",
                         precision="mixed-q31")
    plan = _plan(torch, br, hs, params, scales=scales, h_all=hq_all)
    plan.run()
    plan.check_errors()
    q = hq_all.cpu().numpy().astype(np.float64).reshape(-1, 2)
    o = 0
    for s, n in enumerate(lengths):
        lam, beta, t0 = params[s]
        hq = (q[o:o + n, 0] + 1j * q[o:o + n, 1]) * scales[s]
        o += n
        y, exp = oracle_mod.response_pipeline(oracle_orbits, hq, lam, beta, 10.0, 25,
                                              "2nd generation", t0=t0, x_base=orbits.x_base,
                                              nthreads=8)
        got = torch.stack(plan.channels(s)).cpu().numpy().astype(np.float64)
        check_mixed(oracle_mod, oracle_orbits, got, y, exp, t0, 10.0, "2nd generation", "AET",
                    TOL_MIXED_KERNEL, 1000, (s,), unequal_arms(orbits))


def test_q31_vs_fp64_oracle_at_5mhz(torch_gpu, orbits, oracle_orbits, oracle_mod):
    """pyResponseTDI(precision="mixed-q31"): storage error ~2^7 below complex64's."""
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    n = 1 << 18
    h = gb_waveform(n, 10.0, 1e-21, 5e-3, 1e-16, 0.8, 0.3, 0.5)
    r = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                      precision="mixed-q31")
    r.get_projections(h, 1.2, 0.3)
    got = torch.stack(r.get_tdi_delays()).cpu().numpy().astype(np.float64)
    _, exp = oracle_mod.response_pipeline(oracle_orbits, h, 1.2, 0.3, 10.0, 25, "2nd generation",
                                          x_base=orbits.x_base, nthreads=8)
    # storage error predicted 2.8e-9, 2.4e-9, 1.3e-6 (oracle on the quantised input); on A and E
    # the float32 output rounding (~3e-8 rms relative) dominates it
    for c, tol in ((0, 1e-7), (1, 1e-7), (2, 1e-5)):
        r_c = rel_l2(got[c, 1000:-1000], exp[c, 1000:-1000])
        assert r_c <= tol, (c, r_c)


TOL_C128 = 1e-7  # mixed-c128 vs the oracle on the same FP64 input, every channel


def test_c128_vs_fp64_oracle_at_5mhz(torch_gpu, orbits, oracle_orbits, oracle_mod):
    """pyResponseTDI(precision="mixed-c128"): float32 outputs, no storage error, so the null
    channel T is as accurate (relative) as A and E."""
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    n = 1 << 18
    h = gb_waveform(n, 10.0, 1e-21, 5e-3, 1e-16, 0.8, 0.3, 0.5)
    r = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                      precision="mixed-c128")
    r.get_projections(h, 1.2, 0.3)
    A, E, T = r.get_tdi_delays()
    assert A.dtype == torch.float32
    got = torch.stack([A, E, T]).cpu().numpy().astype(np.float64)
    y, exp = oracle_mod.response_pipeline(oracle_orbits, h, 1.2, 0.3, 10.0, 25, "2nd generation",
                                          x_base=orbits.x_base, nthreads=8)
    check_mixed(oracle_mod, oracle_orbits, got, y, exp, 0.0, 10.0, "2nd generation", "AET",
                TOL_C128, 1000, (), unequal_arms(orbits))


def test_bound_filters_follow_table_changes(torch_gpu, orbits):
    """The composite TDI filters are built when sources are bound; replacing the response's
    TDI table afterwards rebuilds them at the next run (same bits as a plan made for the new
    table)."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.tables import stock_tdi, tdi_tables
    rng = np.random.default_rng(59)
    hs, params = _sources(rng, [20000, 9000])
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, precision="mixed-c128")
    plan = _plan(torch, br, hs, params)
    plan.run()
    out2 = plan.out_all.clone()
    tabs = tdi_tables(stock_tdi("1st generation"))
    br.model.cpp_response.set_tdi(tabs["unit_starts"], tabs["unit_lengths"], tabs["tdi_base_links"],
                                  tabs["tdi_link_combinations"], tabs["tdi_signs"],
                                  tabs["channels"], tabs["num_units"], 3)
    plan.run()
    plan.check_errors()
    b1 = BatchedResponse(0.1, tdi="1st generation", orbits=orbits, precision="mixed-c128")
    p1 = _plan(torch, b1, hs, params)
    p1.run()
    p1.check_errors()
    assert torch.equal(plan.out_all, p1.out_all)
    assert not torch.equal(plan.out_all, out2)


def test_mixed_dropin_reuses_its_batch(torch_gpu, orbits):
    """pyResponseTDI's mixed forms keep one reserved batch per flag set and rebind it per call:
    consecutive calls with different inputs, times and channels give the results of fresh
    instances, and earlier results are not overwritten."""
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    n = 1 << 15
    hs = [gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.8, 0.3, 0.5) for f in (3e-3, 7e-3)]
    for prec in ("mixed-c128", "mixed"):
        r = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                          precision=prec)
        outs = []
        for h, t0 in zip(hs, (0.0, 1.0e6)):
            r.get_projections(h, 1.2, 0.3, t0=t0)
            outs.append([c.clone() for c in r.get_tdi_delays()] + [r.get_tdi_delays()[0]])
        r.tdi_chan = "XYZ"
        xyz = r.get_tdi_delays()[0]
        for (h, t0), got in zip(zip(hs, (0.0, 1.0e6)), outs):
            f = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                              precision=prec)
            f.get_projections(h, 1.2, 0.3, t0=t0)
            exp = f.get_tdi_delays()
            for c in range(3):
                assert torch.equal(got[c], exp[c]), (prec, t0, c)
            assert torch.equal(got[3], exp[0])
        f.tdi_chan = "XYZ"
        assert torch.equal(xyz, f.get_tdi_delays()[0]), prec
        r.check_errors()


def test_mixed_far_times_take_the_per_sample_path(torch_gpu, oracle_mod):
    """At t ~ 1e9 s the rounding of the delays can exceed the composite filters' tolerance, so
    tiles are not cleared as a whole and samples are tested one by one; those that fail take
    the per-sample stencils on the padded window. Against the oracle, and across launch shapes
    (bitwise), for the complex128 and complex64 inputs."""
    torch = torch_gpu
    from fastlisaresponse_amd.autotune import apply, candidates
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    orb = EqualArmlengthOrbits(t_max=1.0e9 + 2.0e6)
    orb.configure(linear_interp_setup=True)
    oo = oracle_mod.OracleOrbits(*orb.pycppdetector_args)
    n, t0 = 60000, 1.0e9
    h = gb_waveform(n, 10.0, 1e-21, 4e-3, 1e-17, 0.7, 0.2, 0.4, t0=t0)
    _, exp = oracle_mod.response_pipeline(oo, h, 0.9, -0.4, 10.0, 25, "2nd generation", t0=t0,
                                          x_base=orb.x_base, nthreads=8)
    hd = torch.as_tensor(h, device="cuda")
    for precision in ("mixed-c128", "mixed"):
        br = BatchedResponse(0.1, tdi="2nd generation", orbits=orb, precision=precision)
        if precision == "mixed":
            hin, sc = pack_mixed(hd, [n])
        else:
            hin, sc = hd, None
        plan = br.plan(hin, [n], [0.9], [-0.4], [t0], scales=sc)
        ref = None
        for shape in candidates(plan):
            apply(plan, *shape)
            plan.out_all.fill_(float("nan"))
            plan.run()
            plan.check_errors()
            if ref is None:
                ref = plan.out_all.clone()
            else:
                assert torch.equal(plan.out_all, ref), (precision, shape)
        got = ref.view(3, n).double().cpu().numpy()
        tol = TOL_C128 if precision == "mixed-c128" else 1e-6
        for c in range(2):
            assert rel_l2(got[c, 1000:-1000], exp[c, 1000:-1000]) <= tol, (precision, c)


@pytest.mark.parametrize("tdi", ["1st generation", "2nd generation"])
def test_c128_batch_vs_oracle(torch_gpu, orbits, oracle_orbits, oracle_mod, tdi):
    """mixed-c128 batch of uneven sources with offsets in time, against the oracle on the same
    complex128 input, XYZ and AET."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    rng = np.random.default_rng(53)
    lengths = [40000, 9000, 25000]
    hs, params = _sources(rng, lengths)
    for chan in ("AET", "XYZ"):
        br = BatchedResponse(0.1, tdi=tdi, orbits=orbits, tdi_chan=chan, precision="mixed-c128")
        plan = _plan(torch, br, hs, params)
        plan.run()
        plan.check_errors()
        assert plan.out_all.dtype == torch.float32
        for s, n in enumerate(lengths):
            lam, beta, t0 = params[s]
            y, exp = oracle_mod.response_pipeline(oracle_orbits, hs[s], lam, beta, 10.0, 25, tdi,
                                                  t0=t0, tdi_chan=chan, x_base=orbits.x_base,
                                                  nthreads=8)
            got = torch.stack(plan.channels(s)).cpu().numpy().astype(np.float64)
            check_mixed(oracle_mod, oracle_orbits, got, y, exp, t0, 10.0, tdi, chan, TOL_C128,
                        1000, (chan, s), unequal_arms(orbits))


def test_fused_argument_checks(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    from fastlisaresponse_amd.batch import BatchedResponse
    with pytest.raises(ValueError):
        BatchedResponse(0.1, orbits=orbits, precision="mixed", exact=True)
    with pytest.raises(ValueError):
        BatchedResponse(0.1, orbits=orbits, order=15, fused=True)
    with pytest.raises(ValueError):
        pyResponseTDI(0.1, 1000, orbits=orbits, precision="mixed", order=15)
    br = BatchedResponse(0.1, orbits=orbits, precision="mixed")
    h = torch.zeros(5000, dtype=torch.complex128, device="cuda")
    with pytest.raises(TypeError):
        br.plan(h, [5000], [0.1], [0.2])  # mixed plans take complex64
    plan = br.plan(h.to(torch.complex64), [5000], [0.1], [0.2])
    with pytest.raises(ValueError):
        plan.projections(0)


@pytest.mark.parametrize("kind", ["mixed", "mixed-q31", "mixed-c128", "fused-fp64", "unfused"])
def test_launch_shapes_give_identical_results(torch_gpu, orbits, kind):
    """Tile length and grid density change speed only: every shape the autotuner may pick gives
    the same bits -- except the mixed forms with unequal arms, whose composite filters are
    built per tile about the tile's middle sample: there the shapes agree within the float32
    outputs' rounding on A and E (the models differ by second-order terms, ~1e-12 of y)."""
    torch = torch_gpu
    from fastlisaresponse_amd.autotune import apply, candidates
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    rng = np.random.default_rng(43)
    lengths = [30000, 4000, 12000]
    hs, params = _sources(rng, lengths)
    kw = {"mixed": dict(precision="mixed"), "mixed-q31": dict(precision="mixed-q31"),
          "mixed-c128": dict(precision="mixed-c128"),
          "unfused": dict(), "fused-fp64": dict(fused=True, exact=False)}[kind]
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan="AET", **kw)
    scales = None
    h_all = None
    if kind in ("mixed", "mixed-q31"):
        from fastlisaresponse_amd.batch import pack_q31
        pack = pack_q31 if kind == "mixed-q31" else pack_mixed
        h_all, scales = pack(torch.as_tensor(np.concatenate(hs), device="cuda"), lengths)
    plan = _plan(torch, br, hs, params, scales=scales, h_all=h_all)
    ref = None
    shapes = list(candidates(plan))
    assert len(shapes) == (12 if kind.startswith("mixed") else 4)
    for shape in shapes:
        apply(plan, *shape)
        plan.out_all.fill_(float("nan"))
        plan.run()
        plan.check_errors()
        if ref is None:
            ref = plan.out_all.clone()
        elif kind.startswith("mixed") and unequal_arms(orbits):
            for s in range(len(lengths)):
                o, n = int(plan.offsets[s]), int(plan.lengths[s])
                a = torch.stack(plan.channels(s)).double()
                b = ref[3 * o:3 * (o + n)].view(3, n).double()
                rel = ((a - b).norm(dim=1) / b.norm(dim=1)).cpu().numpy()
                assert np.all(rel[:2] <= 1e-8) and rel[2] <= 1e-6, (shape, s, rel)
        else:
            assert torch.equal(plan.out_all, ref), shape
    res = plan.tune(use_cache=False, reps=1)
    assert tuple(res["best"]) in shapes and plan.launch_shape == tuple(res["best"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_window_error_is_reported_not_faulted(torch_gpu, orbits, precision):
    """A source whose buffers are too short for the delays: the fused kernel's window checks
    latch FLR_ERR_WINDOW and write NaN for the affected samples; nothing is read out of range
    and the other samples are untouched by the failure."""
    import ctypes
    torch = torch_gpu
    from fastlisaresponse_amd import _capi
    from fastlisaresponse_amd._capi import FLRError, check, lib
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.tables import sky_basis
    br = BatchedResponse(0.1, orbits=orbits, precision=precision, fused=True,
                         exact=False if precision == "fp64" else None)
    n = 20000
    t = np.arange(n) * 10.0
    h = np.exp(2j * np.pi * 3e-3 * t)
    hd = torch.as_tensor(h, device="cuda")
    if precision == "mixed":
        hd = hd.to(torch.complex64)
    out = torch.zeros(3 * n, dtype=torch.float32 if precision == "mixed" else torch.float64,
                      device="cuda")
    srcs = (_capi.FlrSource * 1)()
    q = srcs[0]
    q.offset, q.num_pts, q.projections_start_ind, q.tdi_start_ind = 0, n, 0, 3
    q.scale = 1.0
    k, u, v = sky_basis(1.2, 0.3)
    for c in range(3):
        q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
    flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_AET
    if precision == "mixed":
        flags |= _capi.FLR_FLAG_MIXED
    hb = ctypes.c_void_p()
    check(lib().flr_fused_create(br.model.cpp_response._h, srcs, 1, hd.data_ptr(), out.data_ptr(),
                                 br.dt, br.sampling_frequency, flags, ctypes.byref(hb)))
    try:
        check(lib().flr_batch_run(hb, None))
        with pytest.raises(FLRError):
            br.model.cpp_response.check_errors()
    finally:
        lib().flr_batch_destroy(hb)
    o = out.view(3, n).double().cpu().numpy()
    assert np.isnan(o[:, 3:40]).all()           # the first tile's samples: window error
    assert np.isfinite(o[:, 5000:15000]).all()  # samples far from the ends are computed
    br.model.cpp_response.check_errors()        # the error word was cleared


@pytest.mark.parametrize("kind", ["mixed-c128", "fused-fp64-exact"])
def test_bound_sources_retile_when_chains_lengthen(torch_gpu, orbits, kind):
    """A batch bound under 1st-generation tables whose response switches to 2nd generation: the
    longer delay chains widen the TDI halo and shorten the tile, so the bound sources are
    re-tiled at the next run -- every sample is computed, bitwise as a plan made for the new
    table (a source of 3e5 samples at dt = 10 s spans ~650 tiles, so a stale tiling would leave
    thousands of samples unwritten)."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.tables import stock_tdi, tdi_tables
    rng = np.random.default_rng(61)
    hs, params = _sources(rng, [300000, 50000])
    kw = dict(precision="mixed-c128") if kind == "mixed-c128" else dict(fused=True)
    br = BatchedResponse(0.1, tdi="1st generation", orbits=orbits, **kw)
    plan = _plan(torch, br, hs, params)
    plan.run()
    plan.check_errors()
    tabs = tdi_tables(stock_tdi("2nd generation"))
    br.model.cpp_response.set_tdi(tabs["unit_starts"], tabs["unit_lengths"], tabs["tdi_base_links"],
                                  tabs["tdi_link_combinations"], tabs["tdi_signs"],
                                  tabs["channels"], tabs["num_units"], 3)
    plan.out_all.fill_(float("nan"))
    plan.run()
    plan.check_errors()
    b2 = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, **kw)
    p2 = _plan(torch, b2, hs, params)
    p2.run()
    p2.check_errors()
    assert not bool(torch.isnan(plan.out_all).any())
    assert torch.equal(plan.out_all, p2.out_all)


def test_mixed_dropin_rebinds_across_streams(torch_gpu, orbits):
    """pyResponseTDI's mixed form rebinds its reserved batch per call on the caller's current
    stream: a call on a second stream, made while the first stream's launch may still run, waits
    for that launch before replacing the source table. Both results equal fresh instances'."""
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    n = 1 << 20
    hs = [torch.as_tensor(gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.8, 0.3, 0.5), device="cuda")
          for f in (3e-3, 7e-3)]
    r = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                      precision="mixed-c128")
    outs = []
    torch.cuda.synchronize()
    for h, st in zip(hs, (torch.cuda.Stream(), torch.cuda.Stream())):
        with torch.cuda.stream(st):
            r.get_projections(h, 1.2, 0.3)
            outs.append(torch.stack(r.get_tdi_delays()))
    torch.cuda.synchronize()
    for h, got in zip(hs, outs):
        f = pyResponseTDI(0.1, n, orbits=orbits, tdi="2nd generation", tdi_chan="AET",
                          precision="mixed-c128")
        f.get_projections(h, 1.2, 0.3)
        assert torch.equal(got, torch.stack(f.get_tdi_delays()))
    r.check_errors()


@pytest.mark.parametrize("precision", ["mixed-c128", "mixed"])
def test_fused_mixed_graph_replays_across_rebind(torch_gpu, orbits, precision):
    """The headline form in a HIP graph: run() of a bound mixed-precision fused batch (composite
    filters built at bind) captured once; replays equal the eager run, with eager work and
    synchronisation between them; after a rebind to new parameters and new inputs of the same
    lengths (new start times, so the filters are rebuilt outside the graph) the same graph
    replays the new batch, equal to a fresh plan's eager run."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse, pack_mixed
    rng = np.random.default_rng(67)
    lengths = [40000, 9000, 25000]
    hs, params = _sources(rng, lengths)
    hs2, params2 = _sources(rng, lengths)
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan="AET",
                         precision=precision, storage_check="off")

    def stored(hlist):
        h = torch.as_tensor(np.concatenate(hlist), device="cuda")
        return pack_mixed(h, lengths) if precision == "mixed" else (h, None)

    h1, sc1 = stored(hs)
    plan = _plan(torch, br, hs, params, scales=sc1, h_all=h1)
    plan.run()
    torch.cuda.synchronize()
    eager = plan.out_all.clone()
    g = plan.graph()
    for rep in range(3):
        plan.out_all.fill_(float("nan"))  # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.out_all, eager), rep
    # rebind: new sky positions, start times and input contents in a new tensor of the same size
    h2, sc2 = stored(hs2)
    plan.rebind([p[0] for p in params2], [p[1] for p in params2], [p[2] for p in params2],
                scales=sc2, h_all=h2)
    fresh = _plan(torch, br, hs2, params2, scales=sc2, h_all=h2)
    fresh.run()
    for rep in range(2):
        plan.out_all.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.out_all, fresh.out_all), rep
    assert not torch.equal(eager, fresh.out_all)
    plan.check_errors()


def test_lti_kernel_variants_agree(torch_gpu, oracle_mod):
    """The composite-filter kernel comes in two variants: with the per-sample pass (some tile
    may not clear the filter test as a whole, e.g. times ~1e9 s) and without it (the host proved
    every tile clears). A batch holding a source at t0 = 1e9 s runs the first; its ordinary
    sources must give the same bits as alone, where the second runs."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    orb = EqualArmlengthOrbits(t_max=1.0e9 + 2.0e6)
    orb.configure(linear_interp_setup=True)
    rng = np.random.default_rng(73)
    lengths = [60000, 30000, 45000]
    t0s = [0.0, 1.0e9, 2.5e7]
    hs = [gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.7, 0.2, 0.4, t0=t0)
          for n, f, t0 in zip(lengths, (3e-3, 4e-3, 6e-3), t0s)]
    lams = rng.uniform(0, 6, 3)
    betas = rng.uniform(-1, 1, 3)
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orb, precision="mixed-c128")
    mixed = br.plan(torch.as_tensor(np.concatenate(hs), device="cuda"), lengths, lams, betas, t0s)
    mixed.run()
    mixed.check_errors()
    for s in (0, 2):
        alone = br.plan(torch.as_tensor(hs[s], device="cuda"), [lengths[s]], [lams[s]], [betas[s]],
                        [t0s[s]])
        alone.run()
        alone.check_errors()
        assert torch.equal(torch.stack(mixed.channels(s)), torch.stack(alone.channels(0))), s
    # and the far-time source against the oracle
    oo = oracle_mod.OracleOrbits(*orb.pycppdetector_args)
    _, exp = oracle_mod.response_pipeline(oo, hs[1], lams[1], betas[1], 10.0, 25, "2nd generation",
                                          t0=t0s[1], x_base=orb.x_base, nthreads=8)
    got = torch.stack(mixed.channels(1)).double().cpu().numpy()
    for c in range(3):
        assert rel_l2(got[c, 1000:-1000], exp[c, 1000:-1000]) <= TOL_C128, c


@pytest.mark.parametrize("kind", ["equal", "unequal"])
def test_orbit_tables_finer_than_the_tiles(torch_gpu, oracle_mod, kind):
    """The matrix-core kernels read a tile's orbit nodes from an LDS copy when all of them fit
    the copy (8 nodes), and from the table otherwise. With nodes every 2000 s the 256- and
    512-sample windows (a few nodes per tile) take the first, the 1024-sample window the second:
    equal arms give the same bits on every launch shape, unequal arms agree within the float32
    rounding, and both match the oracle; inner products run on both."""
    torch = torch_gpu
    from fastlisaresponse_amd.autotune import apply, candidates
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits
    cls = EqualArmlengthOrbits if kind == "equal" else CartwheelOrbits
    orb = cls(t_max=4.0e7, dt_base=2000.0)
    orb.configure(linear_interp_setup=True)
    assert orb.dt == 2000.0
    oo = oracle_mod.OracleOrbits(*orb.pycppdetector_args)
    lengths, t0s, freqs = [60000, 45000], [0.0, 2.5e7], [4e-3, 2e-3]
    hs = [gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.7, 0.2, 0.4, t0=t0)
          for n, f, t0 in zip(lengths, freqs, t0s)]
    lams, betas = [0.9, 4.1], [-0.4, 0.8]
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orb, precision="mixed-c128")
    plan = br.plan(torch.as_tensor(np.concatenate(hs), device="cuda"), lengths, lams, betas, t0s)
    ref = None
    windows = set()
    for shape in candidates(plan):
        apply(plan, *shape)
        windows.add(shape[0])
        plan.out_all.fill_(float("nan"))
        plan.run()
        plan.check_errors()
        if ref is None:
            ref = plan.out_all.clone()
        elif kind == "equal":
            assert torch.equal(plan.out_all, ref), shape
        else:
            rel = ((plan.out_all.double() - ref.double()).norm() / ref.double().norm()).item()
            assert rel <= 1e-7, (shape, rel)
    assert windows == {256, 512, 1024}
    for s, n in enumerate(lengths):
        y, exp = oracle_mod.response_pipeline(oo, hs[s], lams[s], betas[s], 10.0, 25,
                                              "2nd generation", t0=t0s[s], x_base=orb.x_base,
                                              nthreads=8)
        got = torch.stack(plan.channels(s)).double().cpu().numpy()
        check_mixed(oracle_mod, oo, got, y, exp, t0s[s], 10.0, "2nd generation", "AET", TOL_C128,
                    1000, (kind, s), kind == "unequal")
    # inner products on the last shape (the 1024-sample window, table reads) and a staged one
    data = plan.out_all.clone()
    sums = plan.inner_products(data).clone()
    apply(plan, 256, 16)
    staged = plan.inner_products(data)
    plan.check_errors()
    for s, n in enumerate(lengths):
        o = int(plan.offsets[s])
        h = ref[3 * o:3 * (o + n)].view(3, n)[:, 1000:n - 1000].double()
        hh = (h * h).sum(dim=1).cpu().numpy()
        for got in (sums, staged):
            g = got[s, :, 1].cpu().numpy()
            assert np.all(np.abs(g - hh) <= 1e-6 * hh), (s, g, hh)
