"""Continuous-batching server (fastlisaresponse_amd.serving), on the GPU.

Requests of different sampling intervals and TDI generations arrive together, are merged into
fused mixed-precision launches on several streams, and complete out of order; every request's
result must be bitwise the one it gets when run alone, and the admitted bytes must never exceed
the HBM budget. Both input formats are served: complex128 (default) and complex64 h / scale.
"""
import numpy as np
import pytest

from conftest import YRSID_SI, rel_l2

pytestmark = pytest.mark.gpu

TOL_MIXED_KERNEL = 2e-7  # mixed kernel vs the oracle on the same complex64-rounded input
TOL_C128 = 1e-7          # complex128-input kernel vs the oracle on the same input
DTYPE = {"mixed": "complex64", "mixed-c128": "complex128"}


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def orbits():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    o = EqualArmlengthOrbits(t_max=0.2 * YRSID_SI)
    o.configure(linear_interp_setup=True)
    return o


KEYS = [(5.0, "1st generation"), (5.0, "2nd generation"), (10.0, "1st generation"),
        (10.0, "2nd generation")]


def _traffic(torch, seed, rate=400.0, duration=0.06, precision="mixed-c128"):
    from fastlisaresponse_amd.serving import gb_pool, poisson_requests
    pool = gb_pool(3, 1 << 16, 10.0, seed=seed, dtype=getattr(torch, DTYPE[precision]))
    rng = np.random.default_rng(seed)
    return poisson_requests(rng, rate, duration, pool, lengths=(1 << 14, 1 << 15, 1 << 16),
                            max_sources=12)


def _alone(torch, orbits, req, precision):
    """The request run by itself through a plain fused mixed-precision plan."""
    from fastlisaresponse_amd.batch import BatchedResponse
    br = BatchedResponse(1.0 / req.dt, tdi=req.tdi, orbits=orbits, precision=precision)
    h = torch.cat([s.h for s in req.sources])
    plan = br.plan(h, [s.num_pts for s in req.sources], [s.lam for s in req.sources],
                   [s.beta for s in req.sources], scales=[s.scale for s in req.sources])
    plan.run()
    plan.check_errors()
    return plan.out_all


@pytest.mark.parametrize("precision", ["mixed-c128", "mixed"])
def test_results_same_alone_and_in_the_mix(torch_gpu, orbits, precision):
    torch = torch_gpu
    from fastlisaresponse_amd.serving import ResponseServer, run_open_loop
    times, reqs = _traffic(torch, 5, precision=precision)
    assert len(reqs) >= 12
    server = ResponseServer(orbits, KEYS, num_streams=3, keep_results=True,
                            max_launch_samples=1 << 20, precision=precision)
    stats = run_open_loop(server, times, reqs)
    assert stats["requests"] == len(reqs)
    assert all(r.completed >= r.arrival for r in reqs)
    for req in reqs:
        assert torch.equal(req.result, _alone(torch, orbits, req, precision)), req.rid
    server.close()


def test_burst_is_merged_and_budget_respected(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.serving import ResponseServer
    times, reqs = _traffic(torch, 9, rate=1000.0, duration=0.03)
    budget = max(r.samples for r in reqs) * ResponseServer.BYTES_PER_SAMPLE * 3
    server = ResponseServer(orbits, KEYS, hbm_budget=budget, num_streams=2, keep_results=True)
    assert server.BYTES_PER_SAMPLE == ResponseServer.BYTES_PER_SAMPLE == 16 + 12
    for r in reqs:  # a burst: everything queued before the first admission
        server.submit(r)
    peak = 0
    while not server.idle():
        server.step()
        peak = max(peak, server.in_flight_bytes)
        assert server.in_flight_bytes <= budget
    server.check_errors()
    assert len(server.done) == len(reqs)
    assert server.launches < len(reqs)  # same-key requests shared launches
    for req in reqs:
        assert torch.equal(req.result, _alone(torch, orbits, req, "mixed-c128")), req.rid
    server.close()


def test_wrong_input_format_is_rejected(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.serving import Request, ResponseServer, SourceSpec
    server = ResponseServer(orbits, KEYS[:1], num_streams=1)
    h = torch.zeros(1 << 14, dtype=torch.complex64, device="cuda")
    with pytest.raises(TypeError):
        server.submit(Request(0, KEYS[0][0], KEYS[0][1], [SourceSpec(h=h, lam=0.1, beta=0.2)]))
    with pytest.raises(ValueError):
        ResponseServer(orbits, KEYS[:1], precision="mixed-q31")
    server.close()


@pytest.mark.parametrize("precision", ["mixed-c128", "mixed"])
def test_served_source_vs_oracle(torch_gpu, orbits, oracle_mod, precision):
    torch = torch_gpu
    from fastlisaresponse_amd.serving import ResponseServer, run_open_loop
    times, reqs = _traffic(torch, 13, precision=precision)
    tol = TOL_C128 if precision == "mixed-c128" else TOL_MIXED_KERNEL
    server = ResponseServer(orbits, KEYS, num_streams=2, keep_results=True, precision=precision)
    run_open_loop(server, times, reqs)
    oo = oracle_mod.OracleOrbits(*orbits.pycppdetector_args)
    checked = 0
    for req in reqs[:4]:
        sp = req.sources[0]
        hq = sp.h.cpu().numpy().astype(np.complex128) * sp.scale
        _, exp = oracle_mod.response_pipeline(oo, hq, sp.lam, sp.beta, req.dt, 25, req.tdi,
                                              x_base=orbits.x_base, nthreads=8)
        got = req.channels(0).cpu().numpy().astype(np.float64)
        T = int(10000.0 / req.dt)
        for c in range(3):
            assert rel_l2(got[c, T:-T], exp[c, T:-T]) <= tol, (req.rid, c)
        checked += 1
    assert checked == 4
    server.close()


@pytest.mark.parametrize("precision", ["mixed-c128", "mixed"])
def test_parameter_requests_generated_on_device(torch_gpu, orbits, precision):
    """Requests that carry galactic-binary parameters: the server generates their inputs on the
    device; each result equals the same request run alone (generator + plain mixed plan)."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    from fastlisaresponse_amd.serving import ResponseServer, poisson_requests, run_open_loop
    from fastlisaresponse_amd.waveforms import generate_gb
    rng = np.random.default_rng(71)
    times, reqs = poisson_requests(rng, 400.0, 0.05, lengths=(1 << 14, 1 << 15), max_sources=8)
    assert reqs and all(s.h is None for r in reqs for s in r.sources)
    server = ResponseServer(orbits, KEYS, num_streams=3, keep_results=True, precision=precision)
    run_open_loop(server, times, reqs)
    for req in reqs:
        params = np.array([s.gb for s in req.sources])
        lengths = [s.num_pts for s in req.sources]
        h32, scales = generate_gb(params, lengths, req.dt, fmt=DTYPE[precision])
        br = BatchedResponse(1.0 / req.dt, tdi=req.tdi, orbits=orbits, precision=precision)
        plan = br.plan(h32, lengths, [s.lam for s in req.sources], [s.beta for s in req.sources],
                       scales=list(scales))
        plan.run()
        plan.check_errors()
        assert torch.equal(req.result, plan.out_all), req.rid
    server.close()


def test_unsynchronised_producer(torch_gpu, orbits):
    """Clients hand over device inputs still being produced on their own stream, without any
    synchronisation: the server's copy waits for the point of submission on that stream, so
    every result equals the request run alone on the finished input."""
    torch = torch_gpu
    from fastlisaresponse_amd.serving import Request, ResponseServer, SourceSpec, gb_pool
    pool = gb_pool(3, 1 << 16, 10.0, seed=3, dtype=torch.complex128)
    server = ResponseServer(orbits, KEYS, num_streams=2, keep_results=True)
    producer = torch.cuda.Stream()
    reqs = []
    rng = np.random.default_rng(19)
    with torch.cuda.stream(producer):
        for rid in range(6):
            key = KEYS[rid % len(KEYS)]
            h = torch.full((1 << 16,), float("nan"), dtype=torch.complex128, device="cuda")
            x = torch.ones(2048, 2048, dtype=torch.float64, device="cuda")
            for _ in range(6):  # several milliseconds of work ahead of the input's last write
                x = (x @ x) / 2048.0
            h.copy_(pool[rid % 3] + 0.0 * x[0, 0])
            req = Request(rid, key[0], key[1], [SourceSpec(h=h, lam=float(rng.uniform(0, 6)),
                                                           beta=float(rng.uniform(-1, 1)),
                                                           scale=1e-21)])
            server.submit(req)  # no synchronisation: the producer is still running
            reqs.append(req)
            del h, x
            server.step()
    while not server.idle():
        server.step()
    server.check_errors()
    torch.cuda.synchronize()
    for req in reqs:
        assert not bool(torch.isnan(req.result).any()), req.rid
        assert torch.equal(req.result, _alone(torch, orbits, req, "mixed-c128")), req.rid
    server.close()
