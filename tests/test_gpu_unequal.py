"""The mixed forms with unequal, time-varying arms (the slowly-varying-arm composite filters).

With six different light travel times the stock TDI's units no longer share delays; the mixed
fused kernel then builds, per tile, composite filters per (channel, base link) from every unit's
position at the tile's middle sample plus a first-order term for the positions' drift across
the tile, and applies them on the matrix cores. Tiles whose positions do not fit that model
(curvature of the arms, a drift across an A-table node) take the per-sample stencils of
LISAResponse.cu:300-333 instead. Checked against the oracle on:

* CartwheelOrbits (all six arms different, drifting by ~1e-10 s/s), 1-yr source, 2nd
  generation: A and E within 1e-7 rel-L2 of the oracle (float32 outputs), T within 1e-6;
* the null channel T against the same projections' TDI with extended-precision delays
  (oracle.tdi_extended_delays): the oracle rounds every sample's delay time to FP64, which T
  amplifies where the arms differ (~2e-7 at t ~ 2.5e7 s and 5 mHz); this form's delays carry no
  such rounding, and late in time it is closer to the extended result than the oracle is;
* 1st generation at 2^20 samples, and X, Y, Z output;
* arms that change fast (the host launches the per-sample kernel), arms within the model's
  tolerances, and arms under which some tiles fail its tests (both paths in one launch);
* a batch of sources from 0.5 to 10 mHz at times up to 2.5e7 s.
Every check of A and E (and X, Y, Z) is against the oracle within 1e-7; T as above.
"""
import os

import numpy as np
import pytest

from conftest import check_mixed, gb_waveform, rel_l2

pytestmark = pytest.mark.gpu

NTH = max(1, min(16, os.cpu_count() or 1))
N_YEAR = 3_153_600
T_BUF = 1000  # tdi_start_ind at dt = 10 s, t_buffer = 1e4 s


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def wobbly_orbits(amp, period):
    """CartwheelOrbits with each light travel time offset by amp * sin(2 pi t / period + r)."""
    from fastlisaresponse_amd.orbits import CartwheelOrbits

    class Wobbly(CartwheelOrbits):
        def _model(self, t):
            x, ltt = super()._model(t)
            ph = 2 * np.pi * np.asarray(t)[:, None] / period + np.arange(6)[None, :]
            return x, ltt + amp * np.sin(ph)

    o = Wobbly(t_max=4.0e7)
    o.configure(linear_interp_setup=True)
    return o


@pytest.fixture(scope="module")
def cartwheel():
    from fastlisaresponse_amd.orbits import CartwheelOrbits
    o = CartwheelOrbits()
    o.configure(linear_interp_setup=True)
    return o


def _mixed_vs_oracle(torch, orbits, oracle_mod, n, f0, tdi, chan="AET", t0=0.0, floor=True):
    """mixed-c128 through pyResponseTDI against the oracle; with floor, the null channel T is
    judged against extended-precision delays (check_mixed). Returns the per-channel errors,
    and the oracle's own T error against extended delays (or None)."""
    from fastlisaresponse_amd import pyResponseTDI
    h = gb_waveform(n, 10.0, 1e-21, f0, 1e-17, 0.8, 0.3, 0.5, t0=t0)
    oo = oracle_mod.OracleOrbits(*orbits.pycppdetector_args)
    y, want = oracle_mod.response_pipeline(oo, h, 1.2, 0.3, 10.0, 25, tdi, t0=t0,
                                           x_base=orbits.x_base, nthreads=NTH, tdi_chan=chan)
    r = pyResponseTDI(0.1, n, order=25, tdi=tdi, orbits=orbits, tdi_chan=chan,
                      precision="mixed-c128")
    r.get_projections(torch.as_tensor(h, device="cuda"), 1.2, 0.3, t0=t0)
    got = torch.stack(r.get_tdi_delays()).double().cpu().numpy()
    r.check_errors()
    assert np.all(np.isfinite(got))
    assert np.all(got[:, :T_BUF] == 0.0) and np.all(got[:, -T_BUF:] == 0.0)
    if not floor:
        return [rel_l2(got[c, T_BUF:-T_BUF], want[c, T_BUF:-T_BUF]) for c in range(3)], None
    errs = check_mixed(oracle_mod, oo, got, y, want, t0, 10.0, tdi, chan, 1e-7, T_BUF,
                       (tdi, chan, t0), unequal=True)
    oracle_floor = None
    if chan == "AET":
        ext = oracle_mod.tdi_extended_delays(oo, y, t0, 10.0, tdi, tdi_start=T_BUF, chan="AET")
        oracle_floor = rel_l2(want[2, T_BUF:-T_BUF], ext[2, T_BUF:-T_BUF])
    return errs, oracle_floor


def test_cartwheel_year_mixed_c128(torch_gpu, cartwheel, oracle_mod):
    """The 1-yr source against the oracle: A and E within 1e-7 (float32 outputs), T within
    1e-6 (the oracle's own T carries the FP64 rounding of its delays, ~2e-7 late in the year)."""
    errs, _ = _mixed_vs_oracle(torch_gpu, cartwheel, oracle_mod, N_YEAR, 5e-3, "2nd generation",
                               floor=False)
    print("cartwheel 1-yr mixed-c128 rel-L2 vs oracle A, E, T:", errs, flush=True)
    assert errs[0] <= 1e-7 and errs[1] <= 1e-7 and errs[2] <= 1e-6, errs


@pytest.mark.parametrize("f0", [5e-3, 1e-3])
def test_cartwheel_late_times_below_the_fp64_floor(torch_gpu, cartwheel, oracle_mod, f0):
    """At t ~ 2.5e7 s the oracle's null channel is ~2e-7 (5 mHz) and ~2e-6 (1 mHz) from the same
    projections' TDI with extended-precision delays; the slowly-varying-arm form, whose delays
    are not rounded per sample, is closer to it than the oracle is."""
    errs, floor = _mixed_vs_oracle(torch_gpu, cartwheel, oracle_mod, 1 << 19, f0,
                                   "2nd generation", t0=2.5e7)
    print("late times", f0, "A, E vs oracle, T vs extended:", errs, "oracle T floor", floor,
          flush=True)
    assert errs[2] < floor, (errs, floor)


@pytest.mark.parametrize("chan", ["AET", "XYZ"])
def test_cartwheel_first_generation(torch_gpu, cartwheel, oracle_mod, chan):
    errs, _ = _mixed_vs_oracle(torch_gpu, cartwheel, oracle_mod, 1 << 20, 3e-3, "1st generation",
                               chan=chan, t0=1.0e6)
    print("cartwheel 1st gen", chan, errs, flush=True)


@pytest.mark.parametrize("amp,period", [(1e-3, 2.0e4), (3e-8, 1.0e5), (2.4e-2, 1.0e8)])
def test_fast_arms_take_the_per_sample_path(torch_gpu, oracle_mod, amp, period):
    """amp 1 ms over 2e4 s: the arms change by ~3e-7 s/s, far past the model's drift bound, so
    the host launches the per-sample kernel (as the oracle); amp 30 ns over 1e5 s: curvature at
    the orbit nodes within the model's tolerance, every tile on the matrix cores; amp 24 ms over
    1e8 s: arm rates ~1.6e-9 s/s, under which ~7% of the tiles see a unit's fraction cross a node
    of the A table and take the filter form's per-sample path (both paths in one launch)."""
    orbits = wobbly_orbits(amp, period)
    errs, _ = _mixed_vs_oracle(torch_gpu, orbits, oracle_mod, 1 << 18, 4e-3, "2nd generation",
                               t0=2.0e6)
    print("wobbly", amp, period, errs, flush=True)


def test_cartwheel_batch_vs_oracle(torch_gpu, cartwheel, oracle_mod):
    """A batch of sources from 0.5 to 10 mHz at times up to 2.5e7 s, each against the oracle
    (T against extended-precision delays, check_mixed)."""
    torch = torch_gpu
    from fastlisaresponse_amd.batch import BatchedResponse
    freqs = np.geomspace(5e-4, 1e-2, 6)
    lengths = [100_000 + 23_000 * s for s in range(len(freqs))]
    t0s = [0.0, 2.5e7, 1.0e6, 2.5e7, 0.0, 1.5e7]
    hs = [gb_waveform(n, 10.0, 1e-21, f, 1e-17, 0.7, 0.2, 0.4, t0=t0)
          for n, f, t0 in zip(lengths, freqs, t0s)]
    lam = np.linspace(0.3, 5.5, len(freqs))
    beta = np.linspace(-1.2, 1.2, len(freqs))
    br = BatchedResponse(0.1, tdi="2nd generation", orbits=cartwheel, tdi_chan="AET",
                         precision="mixed-c128")
    plan = br.plan(torch.as_tensor(np.concatenate(hs), device="cuda"), lengths, lam, beta, t0s)
    plan.run()
    plan.check_errors()
    oo = oracle_mod.OracleOrbits(*cartwheel.pycppdetector_args)
    for s, n in enumerate(lengths):
        y, want = oracle_mod.response_pipeline(oo, hs[s], lam[s], beta[s], 10.0, 25,
                                               "2nd generation", t0=t0s[s],
                                               x_base=cartwheel.x_base, nthreads=NTH)
        got = torch.stack(plan.channels(s)).double().cpu().numpy()
        errs = check_mixed(oracle_mod, oo, got, y, want, t0s[s], 10.0, "2nd generation", "AET",
                           1e-7, T_BUF, (s, freqs[s]), unequal=True)
        print("batch source", s, freqs[s], t0s[s], errs, flush=True)
