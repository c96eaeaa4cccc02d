"""Parity of the gfx950 kernels with the FP64 CPU oracle (run on the GPU).

Exact form (the default; reference operation order, correctly rounded divisions): bitwise equal
to the oracle on every output. Fast form (reciprocal-free stencils): relative L2 <= 1e-12 on
X, Y, Z / A, E; the null channel T is checked against the scale of the other channels, since
ulp-level stencil differences do not cancel in it.

Every test runs on two orbit models, which select different TDI kernels:
  equal  -- EqualArmlengthOrbits: all six light travel times equal -> equal-arm stock kernel
  unequal -- CartwheelOrbits: six different, time-varying light travel times -> generic kernel
"""
import numpy as np
import pytest

from conftest import YRSID_SI, gb_waveform, rel_l2

pytestmark = pytest.mark.gpu

TOL_FAST = 1e-12


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastlisaresponse_amd as flr
    from fastlisaresponse_amd import _capi
    _capi.lib()
    return flr


@pytest.fixture(scope="module", params=["equal", "unequal"])
def orbits(request):
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits
    cls = EqualArmlengthOrbits if request.param == "equal" else CartwheelOrbits
    o = cls(t_max=1.2 * YRSID_SI)
    o.configure(linear_interp_setup=True)
    return o


@pytest.fixture(scope="module")
def oracle_orbits(oracle_mod, orbits):
    return oracle_mod.OracleOrbits(*orbits.pycppdetector_args)


def _setup(flr, orbits, order=25, tdi="1st generation"):
    from fastlisaresponse_amd.backend import LISAResponseWrap, OrbitsWrap
    from fastlisaresponse_amd.tables import lagrange_tables, stock_tdi, tdi_tables
    ow = OrbitsWrap(*orbits.pycppdetector_args)
    rw = LISAResponseWrap(ow)
    A, deps, E = lagrange_tables(order)
    rw.set_lagrange(order, A, deps, len(A), E)
    tabs = tdi_tables(stock_tdi(tdi) if isinstance(tdi, str) else tdi)
    rw.set_tdi(tabs["unit_starts"], tabs["unit_lengths"], tabs["tdi_base_links"],
               tabs["tdi_link_combinations"], tabs["tdi_signs"], tabs["channels"],
               tabs["num_units"], 3)
    return ow, rw, (A, deps, E), tabs


def _oracle_tables(tabs):
    return dict(unit_starts=tabs["unit_starts"], unit_lengths=tabs["unit_lengths"],
                tdi_base_link=tabs["tdi_base_links"],
                tdi_link_combinations=tabs["tdi_link_combinations"], tdi_signs=tabs["tdi_signs"],
                channels=tabs["channels"], num_units=tabs["num_units"])


def _smooth_links(n, dt, seed):
    """Projections as sums of sinusoids (white noise would make the tolerance meaningless)."""
    rng = np.random.default_rng(seed)
    tt = np.arange(n) * dt
    y = np.zeros((6, n))
    for l in range(6):
        for _ in range(4):
            y[l] += rng.normal() * np.sin(2 * np.pi * rng.uniform(1e-4, 1e-2) * tt + rng.uniform(0, 6))
    return y


def _assert_fast_channels(got, exp):
    scale = max(np.linalg.norm(exp[0]), np.linalg.norm(exp[1]))
    for c in range(2):
        assert rel_l2(got[c], exp[c]) <= TOL_FAST, (c, rel_l2(got[c], exp[c]))
    assert np.linalg.norm(got[2] - exp[2]) <= TOL_FAST * scale


# ------------------------------------------------------------------ orbits

def test_orbit_interpolation_bitwise(gpu, orbits, oracle_orbits):
    import torch
    ow, _, _, _ = _setup(gpu, orbits)
    rng = np.random.default_rng(11)
    t = np.concatenate([rng.uniform(0, orbits.t[-1], 2000), orbits.t[:50], [orbits.t[-1]],
                        [-5.0, orbits.t[-1] + 7.0]])
    tt = torch.as_tensor(t, device="cuda")
    for link in orbits.LINKS:
        np.testing.assert_array_equal(ow.light_travel_time(tt, link).cpu().numpy(),
                                      oracle_orbits.ltt_at(t, link))
    for sc in (1, 2, 3):
        np.testing.assert_array_equal(ow.position(tt, sc).cpu().numpy(), oracle_orbits.pos_at(t, sc))


# ------------------------------------------------------------------ projection

@pytest.mark.parametrize("t0", [0.0, 3.0e7])
@pytest.mark.parametrize("exact", [True, False])
def test_projection_vs_oracle(gpu, orbits, oracle_orbits, oracle_mod, t0, exact):
    import torch
    from fastlisaresponse_amd import _capi
    from fastlisaresponse_amd.tables import buffer_indices, sky_basis
    _, rw, (A, deps, E), _ = _setup(gpu, orbits)
    n, dt, order = 8192, 10.0, 25
    fs = 1 / dt
    h = gb_waveform(n, dt, f=5e-3, fdot=1e-16, t0=t0)
    k, u, v = sky_basis(1.2, 0.3)
    _, _, _, P = buffer_indices(fs, order, 10000.0, orbits.x_base)
    t_data = t0 + np.arange(n) * dt
    y_exp = oracle_mod.get_response(oracle_orbits, h, t_data, k, u, v, order, fs, A, deps, E, P, 8)
    y = torch.full((6 * n,), 7.0, dtype=torch.float64, device="cuda")
    flags = _capi.FLR_FLAG_ZERO_EDGES | (_capi.FLR_FLAG_EXACT if exact else 0)
    rw.projection(y, torch.as_tensor(h, device="cuda"), n, t0, dt, fs, k, u, v, P, flags)
    rw.check_errors()
    y = y.cpu().numpy().reshape(6, n)
    if exact:
        np.testing.assert_array_equal(y, y_exp)
    else:
        np.testing.assert_array_equal(y[:, :P], 0.0)
        np.testing.assert_array_equal(y[:, n - P:], 0.0)
        for l in range(6):
            assert rel_l2(y[l], y_exp[l]) <= TOL_FAST, (l, rel_l2(y[l], y_exp[l]))


def test_projection_t_data_path_matches_uniform(gpu, orbits):
    """Reading t_data gives the same bits as synthesising t0 + i*dt in-kernel."""
    import torch
    from fastlisaresponse_amd.tables import sky_basis
    _, rw, _, _ = _setup(gpu, orbits)
    n, dt, t0 = 4096, 10.0, 1.234567e7
    h = torch.as_tensor(gb_waveform(n, dt, t0=t0), device="cuda")
    k, u, v = sky_basis(0.4, -0.7)
    t_data = torch.as_tensor(t0 + np.arange(n) * dt, device="cuda")
    y1 = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    y2 = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    rw.projection(y1, h, n, t0, dt, 1 / dt, k, u, v, 780, 0, t_data=t_data)
    rw.projection(y2, h, n, t0, dt, 1 / dt, k, u, v, 780, 0)
    rw.check_errors()
    assert torch.equal(y1, y2)


# ------------------------------------------------------------------ TDI

@pytest.mark.parametrize("tdi", ["1st generation", "2nd generation"])
@pytest.mark.parametrize("exact", [True, False])
def test_tdi_vs_oracle(gpu, orbits, oracle_orbits, oracle_mod, tdi, exact):
    import torch
    from fastlisaresponse_amd import _capi
    _, rw, (A, deps, E), tabs = _setup(gpu, orbits, tdi=tdi)
    n, dt, order, T, t0 = 8192, 10.0, 25, 1000, 2.5e7
    fs = 1 / dt
    y = _smooth_links(n, dt, 5)
    t_arr = t0 + np.arange(n) * dt
    xyz_exp = oracle_mod.get_tdi_delays(oracle_orbits, y, t_arr, _oracle_tables(tabs), order, fs,
                                        A, deps, E, T, 8)
    out = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    flags = _capi.FLR_FLAG_OVERWRITE | _capi.FLR_FLAG_ZERO_EDGES
    flags |= _capi.FLR_FLAG_EXACT if exact else 0
    yd = torch.as_tensor(y.reshape(-1), device="cuda")
    rw.tdi(out, yd, n, t0, dt, fs, T, flags)
    rw.check_errors()
    xyz = out.cpu().numpy().reshape(3, n)
    if exact:
        np.testing.assert_array_equal(xyz, xyz_exp)
    else:
        for c in range(3):
            assert rel_l2(xyz[c], xyz_exp[c]) <= TOL_FAST, (c, rel_l2(xyz[c], xyz_exp[c]))
    # fused AET epilogue: the oracle's AET arithmetic on the same X, Y, Z
    aet = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    rw.tdi(aet, yd, n, t0, dt, fs, T, flags | _capi.FLR_FLAG_AET)
    rw.check_errors()
    np.testing.assert_array_equal(aet.cpu().numpy().reshape(3, n), np.stack(oracle_mod.aet(*xyz)))


def test_tdi_accumulates_into_output(gpu, orbits, oracle_orbits, oracle_mod):
    """Without OVERWRITE the sums are added to what the output holds, unit by unit."""
    import torch
    from fastlisaresponse_amd import _capi
    _, rw, (A, deps, E), tabs = _setup(gpu, orbits, tdi="2nd generation")
    n, dt, T = 6000, 10.0, 1000
    y = _smooth_links(n, dt, 8)
    t_arr = np.arange(n) * dt
    pre = np.random.default_rng(3).normal(size=3 * n)
    exp = oracle_mod.get_tdi_delays(oracle_orbits, y, t_arr, _oracle_tables(tabs), 25, 0.1, A, deps,
                                    E, T, 8, delayed_links=pre)
    out = torch.as_tensor(pre.copy(), device="cuda")
    rw.tdi(out, torch.as_tensor(y.reshape(-1), device="cuda"), n, 0.0, dt, 0.1, T,
           _capi.FLR_FLAG_EXACT)
    rw.check_errors()
    np.testing.assert_array_equal(out.cpu().numpy().reshape(3, n), exp)


def test_tdi_custom_combination(gpu, orbits, oracle_orbits, oracle_mod):
    """A user list of terms (Sagnac-like alpha, one non-unit sign) uses the generic kernel."""
    import torch
    from fastlisaresponse_amd import _capi
    combo = [
        {"link": 12, "links_for_delay": [], "sign": 1},
        {"link": 23, "links_for_delay": [12], "sign": 1},
        {"link": 31, "links_for_delay": [12, 23], "sign": 1},
        {"link": 13, "links_for_delay": [], "sign": -1},
        {"link": 32, "links_for_delay": [13], "sign": -1},
        {"link": 21, "links_for_delay": [13, 32], "sign": -0.5},
    ]
    _, rw, (A, deps, E), tabs = _setup(gpu, orbits, tdi=combo)
    n, dt, T = 6000, 10.0, 1000
    y = _smooth_links(n, dt, 9)
    tt = np.arange(n) * dt
    xyz_exp = oracle_mod.get_tdi_delays(oracle_orbits, y, tt, _oracle_tables(tabs), 25, 0.1, A, deps,
                                        E, T, 8)
    out = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    rw.tdi(out, torch.as_tensor(y.reshape(-1), device="cuda"), n, 0.0, dt, 0.1, T,
           _capi.FLR_FLAG_OVERWRITE | _capi.FLR_FLAG_EXACT)
    rw.check_errors()
    np.testing.assert_array_equal(out.cpu().numpy().reshape(3, n), xyz_exp)


@pytest.mark.parametrize("order", [7, 15, 31])
def test_other_orders(gpu, orbits, oracle_orbits, oracle_mod, order):
    """Orders other than 25 run the runtime-order kernels."""
    import torch
    from fastlisaresponse_amd import _capi
    from fastlisaresponse_amd.tables import buffer_indices, sky_basis
    _, rw, (A, deps, E), tabs = _setup(gpu, orbits, order=order, tdi="2nd generation")
    n, dt = 6000, 10.0
    fs = 1 / dt
    h = gb_waveform(n, dt, f=3e-3)
    k, u, v = sky_basis(2.5, 0.9)
    T, _, _, P = buffer_indices(fs, order, 10000.0, orbits.x_base)
    t_data = np.arange(n) * dt
    y_exp = oracle_mod.get_response(oracle_orbits, h, t_data, k, u, v, order, fs, A, deps, E, P, 8)
    xyz_exp = oracle_mod.get_tdi_delays(oracle_orbits, y_exp, t_data, _oracle_tables(tabs), order,
                                        fs, A, deps, E, T, 8)
    y = torch.empty(6 * n, dtype=torch.float64, device="cuda")
    rw.projection(y, torch.as_tensor(h, device="cuda"), n, 0.0, dt, fs, k, u, v, P,
                  _capi.FLR_FLAG_EXACT | _capi.FLR_FLAG_ZERO_EDGES)
    out = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    rw.tdi(out, y, n, 0.0, dt, fs, T,
           _capi.FLR_FLAG_EXACT | _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE)
    rw.check_errors()
    np.testing.assert_array_equal(y.cpu().numpy().reshape(6, n), y_exp)
    np.testing.assert_array_equal(out.cpu().numpy().reshape(3, n), xyz_exp)


# ------------------------------------------------------------------ drop-in entry points

def test_dropin_wraps_match_oracle(gpu, orbits, oracle_orbits, oracle_mod):
    """get_response_wrap / get_tdi_delays_wrap with the reference's argument lists, accumulate
    semantics and t arrays: bitwise (the drop-in default form)."""
    import torch
    from fastlisaresponse_amd.tables import buffer_indices, sky_basis
    _, rw, (A, deps, E), tabs = _setup(gpu, orbits, tdi="2nd generation")
    n, dt, order, t0 = 5000, 10.0, 25, 4.0e6
    fs = 1 / dt
    h = gb_waveform(n, dt, t0=t0)
    k, u, v = sky_basis(2.0, -0.2)
    T, _, _, P = buffer_indices(fs, order, 10000.0, orbits.x_base)
    t_data = t0 + np.arange(n) * dt
    y = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    rw.get_response_wrap(y, torch.as_tensor(t_data, device="cuda"), k, u, v, dt, n,
                         torch.as_tensor(h, device="cuda"), n, order, fs, 2 * order + 1, A, deps,
                         len(A), E, P)
    y_exp = oracle_mod.get_response(oracle_orbits, h, t_data, k, u, v, order, fs, A, deps, E, P, 8)
    np.testing.assert_array_equal(y.cpu().numpy().reshape(6, n), y_exp)
    pre = np.random.default_rng(1).normal(size=3 * n)
    dl = torch.as_tensor(pre.copy(), device="cuda")
    rw.get_tdi_delays_wrap(dl, y, n, n, torch.as_tensor(t_data, device="cuda"),
                           tabs["unit_starts"], tabs["unit_lengths"], tabs["tdi_base_links"],
                           tabs["tdi_link_combinations"], tabs["tdi_signs"], tabs["channels"],
                           tabs["num_units"], 3, order, fs, 2 * order + 1, A, deps, len(A), E, T)
    exp = oracle_mod.get_tdi_delays(oracle_orbits, y_exp, t_data, _oracle_tables(tabs), order, fs,
                                    A, deps, E, T, 8, delayed_links=pre)
    np.testing.assert_array_equal(dl.cpu().numpy().reshape(3, n), exp)
    # numpy outputs are written back in place
    y_np = np.zeros(6 * n)
    rw.get_response_wrap(y_np, t_data, k, u, v, dt, n, h, n, order, fs, 2 * order + 1, A, deps,
                         len(A), E, P)
    np.testing.assert_array_equal(y_np.reshape(6, n), y_exp)


def test_reference_host_flow_with_numpy_xp(gpu, orbits, oracle_orbits, oracle_mod):
    """The reference's host code with xp = numpy (the backend INTEGRATION.md sketches): its
    get_projections (response.py:526-557) and get_tdi_delays (:592-657) call sequence, arrays
    made with numpy as it makes them, through the classes of backend.backend_methods(); the
    projections, X, Y, Z and A, E, T are bitwise the oracle's."""
    from fastlisaresponse_amd.backend import backend_methods
    from fastlisaresponse_amd.tables import buffer_indices, lagrange_tables, stock_tdi, tdi_tables
    from fastlisaresponse_amd.utility import AET
    xp = np
    m = backend_methods()
    cpp_orbits = m["OrbitsWrap"](*orbits.pycppdetector_args)
    cpp_response = m["LISAResponseWrap"](cpp_orbits)
    order, dt, t0, n = 25, 10.0, 2.0e6, 6000
    fs = 1 / dt
    A_in, deps, E_in = lagrange_tables(order)
    A_in, E_in = xp.asarray(A_in), xp.asarray(E_in)
    tabs = tdi_tables(stock_tdi("2nd generation"))
    tdi_base_links = xp.asarray(tabs["tdi_base_links"]).astype(xp.int32)
    tdi_link_combinations = xp.asarray(tabs["tdi_link_combinations"]).astype(xp.int32)
    tdi_signs = xp.asarray(tabs["tdi_signs"]).astype(xp.float64)
    channels = xp.asarray(tabs["channels"]).astype(xp.int32)
    tdi_operation_index = xp.asarray(tabs["tdi_operation_index"]).astype(xp.int32)
    T, _, _, P = buffer_indices(fs, order, 10000.0, orbits.x_base)
    from fastlisaresponse_amd.tables import sky_basis
    k, u, v = sky_basis(0.7, -0.9)
    k_in, u_in, v_in = xp.asarray(k), xp.asarray(u), xp.asarray(v)
    input_in = xp.asarray(gb_waveform(n, dt, t0=t0))
    t_data = t0 + xp.arange(len(input_in)) * dt
    y_gw = xp.zeros((6 * n,), dtype=xp.float64)
    cpp_response.get_response_wrap(y_gw, t_data, k_in, u_in, v_in, dt, len(input_in), input_in,
                                   len(input_in), order, fs, 2 * order + 1, A_in, deps,
                                   len(A_in), E_in, P)
    delayed_links_flat = xp.zeros((3, n), dtype=xp.float64).flatten()
    t_data = t0 + xp.arange(n) * dt
    num_units = int(tdi_operation_index.max() + 1)
    _, unit_starts, unit_lengths = np.unique(tdi_operation_index, return_index=True,
                                             return_counts=True)
    cpp_response.get_tdi_delays_wrap(delayed_links_flat, y_gw.reshape(6, -1).flatten().copy(), n,
                                     n, t_data, unit_starts.astype(np.int32),
                                     unit_lengths.astype(np.int32), tdi_base_links,
                                     tdi_link_combinations, tdi_signs, channels, num_units, 3,
                                     order, fs, 2 * order + 1, A_in, deps, len(A_in), E_in, T)
    X, Y, Z = delayed_links_flat.reshape(3, -1)
    A, E, Tc = AET(X, Y, Z)
    y_exp = oracle_mod.get_response(oracle_orbits, input_in, t_data, k, u, v, order, fs,
                                    np.asarray(A_in), deps, np.asarray(E_in), P, 8)
    xyz_exp = oracle_mod.get_tdi_delays(oracle_orbits, y_exp, t_data, _oracle_tables(tabs), order,
                                        fs, np.asarray(A_in), deps, np.asarray(E_in), T, 8)
    np.testing.assert_array_equal(y_gw.reshape(6, n), y_exp)
    np.testing.assert_array_equal(np.stack([X, Y, Z]), xyz_exp)
    np.testing.assert_array_equal(np.stack([A, E, Tc]), np.stack(oracle_mod.aet(*xyz_exp)))


def test_dropin_length_checks(gpu, orbits):
    import torch
    _, rw, (A, deps, E), _ = _setup(gpu, orbits)
    n = 3000
    with pytest.raises(ValueError, match="y_gw: input arrays have the incorrect length"):
        rw.get_response_wrap(torch.zeros(6 * n - 1, dtype=torch.float64, device="cuda"),
                             np.arange(n) * 10.0, np.ones(3), np.ones(3), np.ones(3), 10.0, n,
                             np.zeros(n, dtype=complex), n, 25, 0.1, 51, A, deps, len(A), E, 780)


def test_window_error_is_reported_not_faulted(gpu, orbits):
    """A start index too small for the delays: kernels clamp every access and latch an error."""
    import torch
    from fastlisaresponse_amd import _capi
    from fastlisaresponse_amd.tables import sky_basis
    _, rw, _, _ = _setup(gpu, orbits)
    n, dt = 4096, 10.0
    h = torch.as_tensor(gb_waveform(n, dt), device="cuda")
    k, u, v = sky_basis(1.0, 0.5)
    y = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    rw.projection(y, h, n, 0.0, dt, 0.1, k, u, v, 5, 0)
    with pytest.raises(_capi.FLRError, match="window"):
        rw.check_errors()
    out = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    rw.tdi(out, y.nan_to_num(), n, 0.0, dt, 0.1, 3, _capi.FLR_FLAG_OVERWRITE)
    with pytest.raises(_capi.FLRError, match="window"):
        rw.check_errors()
    rw.check_errors()  # the latched word was cleared


# ------------------------------------------------------------------ full API

@pytest.mark.parametrize("tdi", ["1st generation", "2nd generation"])
def test_response_wrapper_vs_oracle(gpu, orbits, oracle_orbits, oracle_mod, tdi):
    """ResponseWrapper on a 2^17-sample galactic binary (the 15-day CPU-reference workload),
    default (exact) form: A, E, T bitwise equal to the oracle."""
    from fastlisaresponse_amd import ResponseWrapper

    n, dt = 2 ** 17, 10.0

    def gb(A, f, fdot, iota, phi0, psi, T=1.0, dt=10.0):
        return gb_waveform(n, dt, A, f, fdot, iota, phi0, psi)

    wrap = ResponseWrapper(gb, n * dt / YRSID_SI, dt, 6, 7, t_buffer=10000.0,
                           remove_sky_coords=True, n_overide=n, orbits=orbits, order=25, tdi=tdi,
                           tdi_chan="AET")
    chans = np.stack([c.cpu().numpy() for c in wrap(1e-21, 2e-3, 1e-17, 0.8, 0.3, 0.5, 1.2, 0.3)])
    _, aet = oracle_mod.response_pipeline(oracle_orbits, gb_waveform(n, dt), 1.2, 0.3, dt, 25, tdi,
                                          x_base=orbits.x_base, nthreads=8)
    T = 1000
    np.testing.assert_array_equal(chans, aet[:, T:-T])
    # the XYZ property recomputes X, Y, Z after a fused AET call
    xyz = wrap.response_model.XYZ.cpu().numpy()
    np.testing.assert_array_equal(np.stack(oracle_mod.aet(*xyz))[:, T:-T], chans)


def test_response_wrapper_fast_form(gpu, orbits, oracle_orbits, oracle_mod):
    from fastlisaresponse_amd import ResponseWrapper
    n, dt = 2 ** 16, 10.0

    def gb(A, f, fdot, iota, phi0, psi, T=1.0, dt=10.0):
        return gb_waveform(n, dt, A, f, fdot, iota, phi0, psi)

    wrap = ResponseWrapper(gb, n * dt / YRSID_SI, dt, 6, 7, remove_sky_coords=True, n_overide=n,
                           orbits=orbits, tdi="2nd generation", tdi_chan="AET", exact=False)
    chans = np.stack([c.cpu().numpy() for c in wrap(1e-21, 5e-3, 1e-16, 0.8, 0.3, 0.5, 1.2, 0.3)])
    _, aet = oracle_mod.response_pipeline(oracle_orbits, gb_waveform(n, dt, f=5e-3, fdot=1e-16), 1.2,
                                          0.3, dt, 25, "2nd generation", x_base=orbits.x_base,
                                          nthreads=8)
    _assert_fast_channels(chans, aet[:, 1000:-1000])


def test_response_wrapper_xyz_and_zero_garbage(gpu, orbits, oracle_orbits, oracle_mod):
    from fastlisaresponse_amd import ResponseWrapper
    n, dt = 40000, 10.0

    def gb(A, f, fdot, iota, phi0, psi, T=1.0, dt=10.0):
        return gb_waveform(n, dt, A, f, fdot, iota, phi0, psi)

    wrap = ResponseWrapper(gb, n * dt / YRSID_SI, dt, 6, 7, remove_sky_coords=True, n_overide=n,
                           orbits=orbits, tdi="2nd generation", tdi_chan="XYZ",
                           remove_garbage="zero")
    chans = np.stack([c.cpu().numpy() for c in wrap(2e-21, 4e-3, 1e-16, 0.3, 1.0, 2.0, 4.0, -1.0)])
    h = gb_waveform(n, dt, 2e-21, 4e-3, 1e-16, 0.3, 1.0, 2.0)
    y_exp, xyz = oracle_mod.response_pipeline(oracle_orbits, h, 4.0, -1.0, dt, 25, "2nd generation",
                                              tdi_chan="XYZ", x_base=orbits.x_base, nthreads=8)
    np.testing.assert_array_equal(wrap.response_model.y_gw.cpu().numpy(), y_exp)
    np.testing.assert_array_equal(chans, xyz)


@pytest.mark.parametrize("precision", ["fp64", "mixed-c128"])
def test_ae_channels(gpu, orbits, oracle_orbits, oracle_mod, precision):
    """tdi_chan="AE" (response.py:663-671): the A and E of "AET", same bits, two channels; the
    exact form bitwise the oracle's A and E."""
    import torch
    from fastlisaresponse_amd import pyResponseTDI
    n, dt = 60000, 10.0
    h = gb_waveform(n, dt, 1e-21, 3e-3, 1e-17, 0.8, 0.3, 0.5)
    out = {}
    for chan in ("AE", "AET"):
        r = pyResponseTDI(1.0 / dt, n, order=25, tdi="2nd generation", orbits=orbits,
                          tdi_chan=chan, precision=precision)
        r.get_projections(torch.as_tensor(h, device="cuda"), 1.2, 0.3)
        got = r.get_tdi_delays()
        r.check_errors()
        out[chan] = np.stack([c.double().cpu().numpy() for c in got])
    assert out["AE"].shape == (2, n) and out["AET"].shape == (3, n)
    np.testing.assert_array_equal(out["AE"], out["AET"][:2])
    if precision == "fp64":
        _, aet = oracle_mod.response_pipeline(oracle_orbits, h, 1.2, 0.3, dt, 25, "2nd generation",
                                              x_base=orbits.x_base, nthreads=8)
        np.testing.assert_array_equal(out["AE"][:, 1000:-1000], aet[:2, 1000:-1000])


def test_get_tdi_delays_with_given_projections(gpu, orbits):
    import torch
    from fastlisaresponse_amd import pyResponseTDI
    n = 20000
    r = pyResponseTDI(0.1, n, orbits=orbits, tdi="1st generation", tdi_chan="XYZ")
    r.get_projections(gb_waveform(n, 10.0), 1.0, 0.2)
    X1 = torch.stack(r.get_tdi_delays()).clone()
    X2 = torch.stack(r.get_tdi_delays(y_gw=r.y_gw.clone()))
    assert torch.equal(X1, X2)
    with pytest.raises(ValueError):
        r.get_tdi_delays(t0=5.0)
