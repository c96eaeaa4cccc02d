"""Batched sources (persistent launches over a flattened (source, tile) list), on the GPU.

Each source of a batch must give exactly the bits it gives alone (single-source path and the
FP64 oracle), whatever its neighbours; graph replays must equal the eager run.
"""
import numpy as np
import pytest

from conftest import YRSID_SI, gb_waveform

pytestmark = pytest.mark.gpu


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


def _sources(rng, lengths):
    params = []
    hs = []
    for n in lengths:
        f = 10 ** rng.uniform(-4, -2)
        fd = 10 ** rng.uniform(-19, -15)
        A = 10 ** rng.uniform(-23, -21)
        iota = np.arccos(rng.uniform(-1, 1))
        phi0, psi = rng.uniform(0, 2 * np.pi, 2)
        lam = rng.uniform(0, 2 * np.pi)
        beta = np.arcsin(rng.uniform(-1, 1))
        t0 = float(rng.choice([0.0, 1.0e6, 2.5e7]))
        hs.append(gb_waveform(n, 10.0, A, f, fd, iota, phi0, psi, t0=t0))
        params.append((lam, beta, t0))
    return hs, params


@pytest.mark.parametrize("tdi,chan", [("2nd generation", "AET"), ("1st generation", "XYZ")])
def test_batch_matches_single_source_runs(torch_gpu, orbits, oracle_mod, tdi, chan):
    torch = torch_gpu
    from fastlisaresponse_amd import pyResponseTDI
    from fastlisaresponse_amd.batch import BatchedResponse
    rng = np.random.default_rng(17)
    lengths = [5000, 12345, 2100, 3000, 20000, 7777, 4096]  # 2100: only edges, 3000: short
    hs, params = _sources(rng, lengths)
    br = BatchedResponse(0.1, tdi=tdi, orbits=orbits, tdi_chan=chan)
    h_all = torch.as_tensor(np.concatenate(hs), device="cuda")
    plan = br.plan(h_all, lengths, [p[0] for p in params], [p[1] for p in params],
                   [p[2] for p in params])
    plan.run()
    plan.check_errors()
    assert plan.samples == sum(lengths)
    for s, n in enumerate(lengths):
        r = pyResponseTDI(0.1, n, orbits=orbits, tdi=tdi, tdi_chan=chan)
        lam, beta, t0 = params[s]
        r.get_projections(hs[s], lam, beta, t0=t0)
        alone = torch.stack(r.get_tdi_delays())
        assert torch.equal(torch.stack(plan.channels(s)), alone), s
        assert torch.equal(plan.projections(s), r.y_gw), s
    # two sources against the oracle directly
    oo = oracle_mod.OracleOrbits(*orbits.pycppdetector_args)
    for s in (1, 4):
        lam, beta, t0 = params[s]
        _, ref = oracle_mod.response_pipeline(oo, hs[s], lam, beta, 10.0, 25, tdi, t0=t0,
                                              tdi_chan=chan, x_base=orbits.x_base, nthreads=8)
        np.testing.assert_array_equal(torch.stack(plan.channels(s)).cpu().numpy(), ref)


def test_batch_graph_replays_equal_eager(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    rng = np.random.default_rng(23)
    lengths = [9000, 4000, 15000]
    hs, params = _sources(rng, lengths)
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=orbits, tdi_chan="AET")
    h_all = torch.as_tensor(np.concatenate(hs), device="cuda")
    plan = br.plan(h_all, lengths, [p[0] for p in params], [p[1] for p in params],
                   [p[2] for p in params])
    plan.run()
    torch.cuda.synchronize()
    eager = plan.out_all.clone()
    y_eager = plan.y_all.clone()
    g = plan.graph()
    for rep in range(3):
        plan.out_all.fill_(rep + 1.0)  # eager work between replays
        plan.y_all.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.out_all, eager), rep
        assert torch.equal(plan.y_all, y_eager), rep
    plan.check_errors()


def test_batch_argument_checks(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    br = BatchedResponse(0.1, orbits=orbits)
    h = torch.zeros(5000, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError):
        br.plan(h, [4000], [0.1], [0.2])  # lengths do not add up
    with pytest.raises(TypeError):
        br.plan(h.to(torch.complex64), [5000], [0.1], [0.2])
    with pytest.raises(ValueError):
        BatchedResponse(0.1, orbits=orbits, t_buffer=100.0)
