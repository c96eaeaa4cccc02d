"""Host-side behaviour of the Python layer that needs no GPU: argument validation (errors the
reference raises before any kernel runs), orbit providers and their table layout."""
import os

import numpy as np
import pytest

from conftest import YRSID_SI

C_SI = 299792458.0


def test_cpu_backend_is_refused():
    from fastlisaresponse_amd.response import ResponseWrapper, pyResponseTDI
    with pytest.raises(ValueError, match="only the GPU"):
        pyResponseTDI(0.1, 1000, force_backend="cpu")
    with pytest.raises(ValueError, match="only the GPU"):
        ResponseWrapper(lambda *a, **k: None, 0.1, 10.0, 6, 7, force_backend="cpu")


def test_bad_tdi_raises_before_gpu():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.response import pyResponseTDI
    o = EqualArmlengthOrbits(t_max=1e7)
    with pytest.raises(ValueError, match="tdi kwarg"):
        pyResponseTDI(0.1, 1000, tdi="3rd generation", orbits=o)


def test_n_overide_must_be_int():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.response import ResponseWrapper
    with pytest.raises(ValueError, match="n_overide"):
        ResponseWrapper(lambda *a, **k: None, 0.1, 10.0, 6, 7, n_overide=10.5,
                        orbits=EqualArmlengthOrbits(t_max=1e7))


def test_tobs_is_clamped_to_the_orbits():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.response import ResponseWrapper
    o = EqualArmlengthOrbits(t_max=1e7)
    with pytest.warns(UserWarning, match="Tobs is larger"):
        with pytest.raises(Exception):  # the warning precedes the GPU-only construction
            ResponseWrapper(lambda *a, **k: None, 5.0, 10.0, 6, 7, orbits=o, tdi="bogus")


def test_pycppdetector_args_layout():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    o = EqualArmlengthOrbits(t_max=2e6, dt_base=1e4)
    with pytest.raises(ValueError):
        o.pycppdetector_args
    o.configure(linear_interp_setup=True)
    dt, N, n_arr, ltt, x, links, sc_r, sc_e, arm = o.pycppdetector_args
    assert dt == 1e4 and N == len(o.t) == 201 and o.t[0] == 0.0
    assert n_arr.shape == (N * 18,) and ltt.shape == (N * 6,) and x.shape == (N * 9,)
    assert list(links) == [12, 23, 31, 13, 32, 21]
    assert list(sc_r) == [1, 2, 3, 1, 3, 2] and list(sc_e) == [2, 3, 1, 3, 2, 1]
    assert links.dtype == np.int32 and arm == 2.5e9
    # rows are time-major: ltt[t][link], x[t][sc][xyz], n[t][link][xyz]
    assert np.array_equal(ltt.reshape(N, 6), o.ltt)
    assert np.array_equal(x.reshape(N, 3, 3), o.x)
    assert np.allclose(np.linalg.norm(n_arr.reshape(N, 6, 3), axis=-1), 1.0)


def test_equal_arm_cartwheel_geometry():
    from fastlisaresponse_amd.orbits import AU_SI, EqualArmlengthOrbits
    o = EqualArmlengthOrbits(t_max=YRSID_SI)
    o.configure(linear_interp_setup=True)
    assert np.all(o.ltt == 2.5e9 / C_SI)
    for i, j in ((0, 1), (1, 2), (2, 0)):
        arm = np.linalg.norm(o.x[:, i] - o.x[:, j], axis=-1)
        assert np.allclose(arm, 2.5e9, rtol=1e-2)
    centre = o.x.mean(axis=1)
    assert np.allclose(np.linalg.norm(centre, axis=-1), AU_SI, rtol=1e-4)
    # one orbit per year: the guiding centre returns to its start
    assert np.allclose(centre[0], centre[np.argmin(np.abs(o.t - YRSID_SI))], atol=1e-3 * AU_SI)


def test_unequal_arm_cartwheel_light_times():
    from fastlisaresponse_amd.orbits import CartwheelOrbits, cartwheel_positions
    o = CartwheelOrbits(t_max=3e6)
    o.configure(linear_interp_setup=True)
    assert len(np.unique(o.ltt[50])) == 6
    for r, link in enumerate(o.LINKS):
        i, j = divmod(link, 10)
        L = o.ltt[:, r]
        xe = cartwheel_positions(o.t - L)[:, j - 1]
        resid = np.linalg.norm(o.x[:, i - 1] - xe, axis=-1) / C_SI - L
        assert np.abs(resid).max() < 1e-9


def test_static_orbits_are_constant():
    from fastlisaresponse_amd.orbits import StaticOrbits
    pos = np.arange(9.0).reshape(3, 3) * 1e9
    o = StaticOrbits(pos, np.arange(6) + 8.0, t_max=1e5)
    o.configure(linear_interp_setup=True)
    assert np.array_equal(o.x, np.broadcast_to(pos, o.x.shape))
    assert np.array_equal(o.ltt[7], np.arange(6) + 8.0)


def test_overlap_helper():
    from fastlisaresponse_amd.utility import get_overlap
    rng = np.random.default_rng(0)
    a = [rng.normal(size=256) for _ in range(3)]
    assert abs(get_overlap(a, a) - 1.0) < 1e-12
    assert abs(get_overlap(a, [-x for x in a]) + 1.0) < 1e-12
    # cos vs sin of one Fourier bin: rfft overlap -i, so real part 0 and modulus 1
    n, k0 = 256, 17
    t = np.arange(n)
    c, s = np.cos(2 * np.pi * k0 * t / n), np.sin(2 * np.pi * k0 * t / n)
    assert abs(get_overlap(c, s)) < 1e-12
    assert abs(get_overlap(c, s, phase_maximize=True) - 1.0) < 1e-12
    # brute-force DFT over bins 0..n/2 on tiny channels
    b = [rng.normal(size=16) for _ in range(3)]
    a16 = [x[:16] for x in a]
    W = np.exp(-2j * np.pi * np.outer(np.arange(9), np.arange(16)) / 16)
    exp = np.mean([np.vdot(W @ x, W @ y) / np.sqrt(np.vdot(W @ x, W @ x) * np.vdot(W @ y, W @ y))
                   for x, y in zip(a16, b)])
    assert abs(get_overlap(a16, b) - exp.real) < 1e-12
    assert abs(get_overlap(np.stack(a16), np.stack(b), phase_maximize=True) - abs(exp)) < 1e-12
    with pytest.raises(ValueError):
        get_overlap(a16, b[:2])


def test_poisson_traffic_mix():
    """Serving traffic generator: Poisson arrivals, 1..64 sources, lengths, dt and the 30/70
    TDI mix of the serving workload (host-side only: the pool is a list of CPU tensors)."""
    import numpy as np
    import torch
    from fastlisaresponse_amd.serving import poisson_requests
    pool = [torch.zeros(1 << 12, dtype=torch.complex64) for _ in range(2)]
    rng = np.random.default_rng(5)
    times, reqs = poisson_requests(rng, 200.0, 20.0, pool, lengths=(1 << 10, 1 << 11, 1 << 12))
    n = len(reqs)
    assert abs(n - 4000) < 4 * np.sqrt(4000)
    assert np.all(np.diff(times) > 0) and times[-1] < 20.0
    gaps = np.diff([0.0] + list(times))
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0
    second = sum(r.tdi == "2nd generation" for r in reqs) / n
    assert abs(second - 0.7) < 0.03
    assert {r.dt for r in reqs} == {5.0, 10.0}
    ns = [len(r.sources) for r in reqs]
    assert min(ns) == 1 and max(ns) == 64
    for r in reqs[:50]:
        assert len({s.num_pts for s in r.sources}) == 1
        assert all(-23 <= np.log10(s.scale) <= -21 for s in r.sources)


def _shell_server(budget, max_sources=512, max_launch_samples=1 << 27, slots=2):
    """A ResponseServer without device state: admission logic only (_launch recorded)."""
    from collections import deque
    from fastlisaresponse_amd.serving import ResponseServer
    srv = object.__new__(ResponseServer)
    srv.pending = deque()
    srv.hbm_budget = budget
    srv.in_flight_bytes = 0
    srv.max_sources = max_sources
    srv.max_launch_samples = max_launch_samples
    srv.busy = [None] * slots
    srv.launched = []

    def launch(slot, reqs):
        srv.launched.append((slot, [r.rid for r in reqs]))
        srv.in_flight_bytes += sum(r.samples for r in reqs) * ResponseServer.BYTES_PER_SAMPLE
        srv.busy[slot] = object()
    srv._launch = launch
    return srv


def _req(rid, key, n, ns=1):
    import torch
    from fastlisaresponse_amd.serving import Request, SourceSpec
    srcs = [SourceSpec(h=torch.zeros(n, dtype=torch.complex64), lam=0.0, beta=0.0) for _ in range(ns)]
    return Request(rid, key[0], key[1], srcs)


def test_serving_admission_merges_same_key_and_keeps_order():
    from fastlisaresponse_amd.serving import ResponseServer
    B = ResponseServer.BYTES_PER_SAMPLE
    a, b = (10.0, "2nd generation"), (5.0, "1st generation")
    srv = _shell_server(budget=10_000 * B, max_launch_samples=2500)
    for r in [_req(0, a, 1000), _req(1, b, 1000), _req(2, a, 1000), _req(3, a, 1000),
              _req(4, b, 500)]:
        srv.pending.append(r)
    assert srv.admit() == 2
    # slot 0: head key a, merged up to 2500 samples (requests 0, 2; 3 would exceed the cap)
    # slot 1: the next head (1, key b) merged with 4
    assert srv.launched == [(0, [0, 2]), (1, [1, 4])]
    assert [r.rid for r in srv.pending] == [3]
    assert srv.in_flight_bytes == 3500 * B


def test_serving_admission_respects_budget_and_head_of_line():
    from fastlisaresponse_amd.serving import ResponseServer
    B = ResponseServer.BYTES_PER_SAMPLE
    a = (10.0, "2nd generation")
    srv = _shell_server(budget=2500 * B, slots=4)
    for r in [_req(0, a, 2000), _req(1, a, 1000), _req(2, a, 400)]:
        srv.pending.append(r)
    srv.admit()
    # 0 is admitted; 1 does not fit beside it but 2 does, and merges with the head
    assert srv.launched == [(0, [0, 2])]
    assert srv.in_flight_bytes <= srv.hbm_budget
    srv.admit()  # the head (1) no longer fits the remaining budget: nothing is admitted
    assert srv.launched == [(0, [0, 2])] and [r.rid for r in srv.pending] == [1]
    # the source cap splits a launch too
    srv = _shell_server(budget=10 ** 9, max_sources=70, slots=1)
    for r in [_req(0, a, 100, ns=40), _req(1, a, 100, ns=40)]:
        srv.pending.append(r)
    srv.admit()
    assert srv.launched == [(0, [0])]


def test_compile_time_order25_E_table_matches_the_host_table():
    """The fused fast kernels' compile-time order-25 E weights (e25 in flr_device.hpp) are the
    bits of the table pyResponseTDI uploads (the host refuses other tables for those kernels)."""
    import re
    from conftest import ROOT
    from fastlisaresponse_amd.tables import lagrange_tables
    text = open(os.path.join(ROOT, "lisa-on-gpu_amd", "csrc", "flr_device.hpp")).read()
    body = text[text.index("constexpr double e25(int jm1)"):]
    body = body[:body.index("}")]
    consts = [float.fromhex(h) for h in re.findall(r"(-?0x1\.[0-9a-f]+p-?\d+)", body)]
    _, _, E = lagrange_tables(25)
    assert len(consts) == 12
    assert consts == [float(e) for e in E[:12]]


def test_tdi_term_type_is_accepted_and_ignored():
    """A custom TDI term may carry "type" ("delay" or "advance", response.py:66); the reference's
    table builder (response.py:312-433) never reads it, and neither does this one: the same
    tables with or without it."""
    from fastlisaresponse_amd.tables import stock_tdi, tdi_tables
    terms = stock_tdi("2nd generation")
    base = tdi_tables(terms)
    for kind in ("delay", "advance"):
        tabs = tdi_tables([dict(t, type=kind) for t in terms])
        assert set(tabs) == set(base)
        for k in base:
            np.testing.assert_array_equal(np.asarray(tabs[k]), np.asarray(base[k]))
