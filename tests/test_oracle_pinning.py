"""Pin the FP64 oracle to mathematics (CPU only).

The reference's own tests hold no golden vectors or known answers (they check no-NaN and
CPU/GPU agreement only), and the reference cannot be built or imported here (its orbit class
comes from the absent lisaanalysistools package). So every oracle function is checked against
what mathematics fixes, chosen so that plausible mistakes fail:

* Lagrange tables vs exact rational arithmetic (committed fixture tests/golden/lagrange_o25.npz)
* stencils vs exact Lagrange interpolation (Fractions) at table nodes, and polynomial
  reproduction between nodes within the A-table error
* the delay split against hand-derived sample positions
* orbit interpolation vs numpy.interp
* projection vs the plane-wave formula evaluated analytically on a frozen constellation
* TDI vs the delayed sums evaluated analytically; Michelson cancellation
* the XYZ -> AET transform: orthonormality and the unit-vector images
* mutants: each planted mistake (oracle built with -DORACLE_MUTANT=k) fails one of the above
"""
from fractions import Fraction
from math import factorial

import numpy as np
import pytest

from conftest import ROOT, rel_l2

GOLDEN = f"{ROOT}/tests/golden/lagrange_o25.npz"


# ------------------------------------------------------------------ helpers

def exact_lagrange(x, e, h):
    """sum_m L_m(r) x_m for nodes p = -(h-1)..h and r = 1 - e, in exact rationals."""
    r = 1 - Fraction(e)
    nodes = list(range(-(h - 1), h + 1))
    total = Fraction(0)
    for m, pm in enumerate(nodes):
        w = Fraction(1)
        for pk in nodes:
            if pk != pm:
                w *= (r - pk) / (pm - pk)
        total += w * Fraction(float(x[m]))
    return total


def exact_A(e, h):
    e = Fraction(e)
    prod = Fraction(1)
    for i in range(1, h):
        prod *= (i + e) * (i + 1 - e)
    return prod / (factorial(h - 1) * factorial(h))


def frozen_orbits(L_values, t_max=3.0e6):
    """Static constellation: the cartwheel at t = 0, constant light travel times."""
    from fastlisaresponse_amd.orbits import StaticOrbits, cartwheel_positions
    pos = cartwheel_positions(np.array([0.0]))[0]
    o = StaticOrbits(pos, L_values, t_max=t_max)
    o.configure(linear_interp_setup=True)
    return o, pos


# ------------------------------------------------------------------ tables

def test_lagrange_tables_match_exact_rationals(oracle_mod):
    from fastlisaresponse_amd.tables import lagrange_tables
    g = np.load(GOLDEN)
    A, deps, E = lagrange_tables(25)
    A2, deps2, num_A, E2 = oracle_mod.lagrange_tables(25)
    # the product's host tables and the oracle's restatement agree bitwise ...
    assert np.array_equal(A, A2) and np.array_equal(E, E2) and deps == deps2 and num_A == 1001
    # ... and both sit within a few ulp of the exact rational values
    assert np.all(np.abs(A - g["A_exact"]) <= 8 * np.spacing(g["A_exact"]))
    assert np.all(np.abs(E[:12] - g["E_exact"]) <= 4 * np.spacing(np.abs(g["E_exact"])))
    assert E[12] == 0.0
    assert abs(A[0] - 1.0) < 1e-15 and abs(A[-1] - 1.0) < 1e-15  # A(0) = A(1) = 1


def test_E_is_the_barycentric_weight():
    """E_j (h-1)! h! ... equals (h-1)! h! times the barycentric weight of node 1+j."""
    from fastlisaresponse_amd.tables import lagrange_tables
    h = 13
    _, _, E = lagrange_tables(2 * h - 1)
    nodes = range(-(h - 1), h + 1)
    for j in range(1, h):
        p = 1 + j
        w = Fraction(1)
        for q in nodes:
            if q != p:
                w /= (p - q)
        assert abs(E[j - 1] - float(w * factorial(h - 1) * factorial(h))) <= 4 * np.spacing(abs(E[j - 1]))


def test_tdi_tables_cyclic_structure(oracle_mod):
    from fastlisaresponse_amd.tables import cyclic_permutation, stock_tdi, tdi_tables
    # hand-derived: 1 -> 2 -> 3 -> 1
    assert [cyclic_permutation(13, s) for s in range(3)] == [13, 21, 32]
    assert [cyclic_permutation(31, s) for s in range(3)] == [31, 12, 23]
    for link in (12, 23, 31, 13, 32, 21):
        assert cyclic_permutation(cyclic_permutation(link, 1), 1) == cyclic_permutation(link, 2)
        assert cyclic_permutation(cyclic_permutation(link, 2), 1) == link
        assert cyclic_permutation(link, 1) == oracle_mod.cyclic_permutation(link, 1)
    for gen, nunits, nchain in (("1st generation", 24, 42), ("2nd generation", 48, 174)):
        t = tdi_tables(stock_tdi(gen))
        assert t["num_units"] == nunits and len(t["tdi_link_combinations"]) == nchain
        o = oracle_mod.tdi_tables(oracle_mod.stock_combinations(gen))
        for a, b in (("tdi_base_links", "tdi_base_link"), ("tdi_link_combinations",) * 2,
                     ("tdi_signs",) * 2, ("channels",) * 2, ("unit_starts",) * 2,
                     ("unit_lengths",) * 2):
            assert np.array_equal(t[a], o[b]), a
        # Y's first unit is X's first unit shifted by one: link 13 -> 21, undelayed
        ny = nunits // 3
        assert t["tdi_base_links"][ny] == 21 and t["channels"][ny] == 1
        assert t["tdi_link_combinations"][t["unit_starts"][ny]] == -11
        # every channel's signs sum to zero (a Michelson combination is balanced)
        for c in range(3):
            assert t["tdi_signs"][t["channels"] == c].sum() == 0.0


def test_sky_basis_is_a_right_handed_orthonormal_triad(oracle_mod):
    from fastlisaresponse_amd.tables import sky_basis
    for lam, beta in ((1.2, 0.3), (5.2, -1.1), (0.0, 1.5707)):
        k, u, v = sky_basis(lam, beta)
        k2, u2, v2 = oracle_mod.sky_basis(lam, beta)
        assert np.array_equal(k, k2) and np.array_equal(u, u2) and np.array_equal(v, v2)
        M = np.stack([u, v, k])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-15)
        assert np.allclose(np.cross(u, v), k, atol=1e-15)
        # k propagates away from the source at (lam, beta)
        src = np.array([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam), np.sin(beta)])
        assert np.allclose(k, -src, atol=1e-15)


def test_buffer_indices_known_values():
    from fastlisaresponse_amd.orbits import EqualArmlengthOrbits
    from fastlisaresponse_amd.tables import buffer_indices
    o = EqualArmlengthOrbits(t_max=1e7)
    # dt = 10 s, order 25, t_buffer = 1e4 s: 1000, 10 + 100, int(max|x| / c) + 100, 1000 - 220;
    # the spacecraft stay within 1 AU * (1 +- 1%), i.e. 494..504 light seconds of the Sun
    T, check, pbuf, P = buffer_indices(0.1, 25, 10000.0, o.x_base)
    assert (T, check, P) == (1000, 110, 780) and 494 + 100 <= pbuf <= 504 + 100
    assert buffer_indices(0.2, 25, 10000.0, o.x_base)[::3] == (2000, 1760)
    assert buffer_indices(0.2, 25, 10000.0, o.x_base)[1] == 120


# ------------------------------------------------------------------ stencils

@pytest.mark.parametrize("k", [0, 1, 137, 500, 731, 999])
def test_stencil_at_table_nodes_is_exact_lagrange(oracle_mod, k):
    A, deps, _, E = oracle_mod.lagrange_tables(25)
    e = k * deps
    assert int(e / deps) == k  # the A lookup lands exactly on node k (frac = 0)
    rng = np.random.default_rng(k)
    h, d = 13, 40
    x = rng.normal(size=100)
    got = oracle_mod.interp_single(x, h, d, e, A, deps, E)
    exp = float(exact_lagrange(x[d - (h - 1): d + h + 1], e, h))
    assert abs(got - exp) <= 1e-13 * np.abs(x).max(), (got, exp)
    # complex = real on each part
    xc = x + 1j * rng.normal(size=100)
    gc = oracle_mod.interp_cplx(xc, h, d, e, A, deps, E)
    assert gc.real == got
    assert gc.imag == oracle_mod.interp_single(xc.imag.copy(), h, d, e, A, deps, E)


def test_stencil_reproduces_polynomials_between_nodes(oracle_mod):
    A, deps, _, E = oracle_mod.lagrange_tables(25)
    h, d = 13, 30
    n = np.arange(80, dtype=np.float64)
    poly = np.polynomial.Polynomial([0.3, -1.1, 0.7, 0.05, -0.002, 1e-4])
    x = poly(n / 10.0)
    rng = np.random.default_rng(2)
    for e in rng.uniform(0, 1, 50):
        got = oracle_mod.interp_single(x, h, d, e, A, deps, E)
        exact = poly((d + 1 - e) / 10.0)
        # exact Lagrange up to the linearly interpolated A(e): |A_tab/A - 1| < 3e-7
        ratio = float(exact_A(e, h))
        assert abs(got - exact) <= 3e-7 * abs(exact) + 1e-14, (e, got, exact, ratio)
    # at e = 0 the stencil returns x[d+1] times A(0)
    assert oracle_mod.interp_single(x, h, d, 0.0, A, deps, E) == A[0] * x[d + 1]


def test_split_delay_conventions(oracle_mod):
    # the stencil point is (delay - t0) * fs = d + 1 - e with e in [0, 1)
    assert oracle_mod.split_delay(100.0, 0.0, 0.1) == (9, 0.0)      # exactly on sample 10
    assert oracle_mod.split_delay(102.5, 0.0, 0.1) == (10, 0.75)    # 10.25 = 10 + 1 - 0.75
    assert oracle_mod.split_delay(107.5 + 3.0e7, 3.0e7, 0.1) == (10, 0.25)
    d, e = oracle_mod.split_delay(-4.0, 0.0, 0.5)                    # -2.0: sample -2
    assert (d, e) == (-3, 0.0)


# ------------------------------------------------------------------ orbits

def test_orbit_interpolation_is_linear_interpolation(oracle_mod):
    from fastlisaresponse_amd.orbits import CartwheelOrbits
    o = CartwheelOrbits(t_max=5e6)
    o.configure(linear_interp_setup=True)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    rng = np.random.default_rng(4)
    t = rng.uniform(0, o.t[-1], 300)
    for r, link in enumerate(o.LINKS):
        got = oo.ltt_at(t, link)
        assert np.allclose(got, np.interp(t, o.t, o.ltt[:, r]), rtol=1e-14, atol=0)
        assert np.array_equal(oo.ltt_at(o.t[5:9], link), o.ltt[5:9, r])  # nodes exactly
    for sc in (1, 2, 3):
        got = oo.pos_at(t, sc)
        for c in range(3):
            assert np.allclose(got[:, c], np.interp(t, o.t, o.x[:, sc - 1, c]), rtol=1e-13, atol=1e-3)
    # beyond the table the last interval is extended linearly
    tt = np.array([o.t[-1] + 500.0])
    slope = (o.ltt[-1, 0] - o.ltt[-2, 0]) / o.dt
    assert np.allclose(oo.ltt_at(tt, o.LINKS[0]), o.ltt[-1, 0] + slope * 500.0, rtol=1e-14)


# ------------------------------------------------------------------ closed forms

C_SI = 299792458.0


def _projection_closed_form(oracle_mod, lib=None):
    """Oracle projection vs the plane-wave formula on a frozen constellation."""
    from fastlisaresponse_amd.tables import buffer_indices, lagrange_tables, sky_basis
    L = np.array([8.31, 8.33, 8.35, 8.32, 8.34, 8.36])
    o, pos = frozen_orbits(L)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    n, dt, order = 6000, 10.0, 25
    fs = 1 / dt
    t = np.arange(n) * dt
    w, a, b, p1, p2 = 2 * np.pi * 4e-3, 1.0, 0.6, 0.4, 1.3
    hp = lambda s: a * np.cos(w * s + p1)
    hc = lambda s: b * np.sin(w * s + p2)
    h = hp(t) + 1j * hc(t)
    lam, beta = 2.1, 0.4
    k, u, v = sky_basis(lam, beta)
    A, deps, E = lagrange_tables(order)
    _, _, _, P = buffer_indices(fs, order, 10000.0, o.x_base)
    y = oracle_mod.get_response(oo, h, t, k, u, v, order, fs, A, deps, E, P)
    errs = []
    for r, link in enumerate(o.LINKS):
        i, j = divmod(link, 10)
        xr, xe = pos[i - 1], pos[j - 1]
        nvec = (xr - xe) / np.linalg.norm(xr - xe)
        xi_p = 0.5 * ((u @ nvec) ** 2 - (v @ nvec) ** 2)
        xi_c = (u @ nvec) * (v @ nvec)
        ts = t[P:n - P]
        te = ts - L[r] - (k @ xe) / C_SI
        tr = ts - (k @ xr) / C_SI
        exp = ((hp(te) - hp(tr)) * xi_p + (hc(te) - hc(tr)) * xi_c) / (1 - k @ nvec)
        errs.append(rel_l2(y[r, P:n - P], exp))
    return max(errs)


def _tdi_closed_form(oracle_mod, generation):
    """Oracle TDI vs the delayed sums evaluated analytically (constant, unequal arms)."""
    from fastlisaresponse_amd.tables import lagrange_tables, stock_tdi, tdi_tables
    L = np.array([8.31, 8.33, 8.35, 8.32, 8.34, 8.36])
    o, _ = frozen_orbits(L)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    n, dt, T = 5000, 10.0, 1000
    t = np.arange(n) * dt
    rng = np.random.default_rng(7)
    f = rng.uniform(1e-3, 8e-3, 6)
    ph = rng.uniform(0, 6, 6)
    amp = rng.uniform(0.5, 2, 6)
    ylink = lambda r, s: amp[r] * np.sin(2 * np.pi * f[r] * s + ph[r])
    y = np.stack([ylink(r, t) for r in range(6)])
    combos = stock_tdi(generation)
    tabs = tdi_tables(combos)
    A, deps, E = lagrange_tables(25)
    ot = dict(unit_starts=tabs["unit_starts"], unit_lengths=tabs["unit_lengths"],
              tdi_base_link=tabs["tdi_base_links"], tdi_link_combinations=tabs["tdi_link_combinations"],
              tdi_signs=tabs["tdi_signs"], channels=tabs["channels"], num_units=tabs["num_units"])
    xyz = oracle_mod.get_tdi_delays(oo, y, t, ot, 25, 1 / dt, A, deps, E, T)
    row = {l: r for r, l in enumerate(o.LINKS)}
    ts = t[T:n - T]
    errs = []
    for c in range(3):
        exp = np.zeros_like(ts)
        for u_ in np.where(tabs["channels"] == c)[0]:
            s0, ln = tabs["unit_starts"][u_], tabs["unit_lengths"][u_]
            delay = sum(L[row[l]] for l in tabs["tdi_link_combinations"][s0:s0 + ln] if l != -11)
            exp += tabs["tdi_signs"][u_] * ylink(row[int(tabs["tdi_base_links"][u_])], ts - delay)
        errs.append(rel_l2(xyz[c, T:n - T], exp))
    return max(errs)


def test_projection_matches_plane_wave_formula(oracle_mod):
    assert _projection_closed_form(oracle_mod) < 5e-6


@pytest.mark.parametrize("generation", ["1st generation", "2nd generation"])
def test_tdi_matches_delayed_sums(oracle_mod, generation):
    assert _tdi_closed_form(oracle_mod, generation) < 2e-6


def test_michelson_cancels_for_identical_links(oracle_mod):
    """Equal arms and identical y on every link: each X term is cancelled by its partner."""
    from fastlisaresponse_amd.tables import lagrange_tables, stock_tdi, tdi_tables
    o, _ = frozen_orbits(np.full(6, 8.339))
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    n, dt, T = 4000, 10.0, 1000
    t = np.arange(n) * dt
    y = np.tile(np.sin(2 * np.pi * 3e-3 * t), (6, 1))
    tabs = tdi_tables(stock_tdi("2nd generation"))
    ot = dict(unit_starts=tabs["unit_starts"], unit_lengths=tabs["unit_lengths"],
              tdi_base_link=tabs["tdi_base_links"], tdi_link_combinations=tabs["tdi_link_combinations"],
              tdi_signs=tabs["tdi_signs"], channels=tabs["channels"], num_units=tabs["num_units"])
    A, deps, E = lagrange_tables(25)
    xyz = oracle_mod.get_tdi_delays(oo, y, t, ot, 25, 0.1, A, deps, E, T)
    assert np.abs(xyz[:, T:n - T]).max() < 1e-13


def test_aet_transform(oracle_mod):
    X, Y, Z = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
    A, E, T = oracle_mod.aet(X, Y, Z)
    M = np.stack([A, E, T])  # images of the unit vectors: rows A, E, T
    assert np.allclose(M @ M.T, np.eye(3), atol=1e-15)  # orthonormal
    s2, s6, s3 = np.sqrt(2.0), np.sqrt(6.0), np.sqrt(3.0)
    assert np.allclose(M, [[-1 / s2, 0, 1 / s2], [1 / s6, -2 / s6, 1 / s6], [1 / s3, 1 / s3, 1 / s3]])


def test_oracle_threads_do_not_change_results(oracle_mod):
    from fastlisaresponse_amd.orbits import CartwheelOrbits
    from conftest import gb_waveform
    o = CartwheelOrbits(t_max=2e6)
    o.configure(linear_interp_setup=True)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    h = gb_waveform(6000, 10.0)
    r1 = oracle_mod.response_pipeline(oo, h, 1.0, 0.2, 10.0, 25, "2nd generation", x_base=o.x_base, nthreads=1)
    r4 = oracle_mod.response_pipeline(oo, h, 1.0, 0.2, 10.0, 25, "2nd generation", x_base=o.x_base, nthreads=4)
    assert np.array_equal(r1[0], r4[0]) and np.array_equal(r1[1], r4[1])


# ------------------------------------------------------------------ mutants

@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("mutant", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_detected(oracle_mod, mutant_dir, monkeypatch, mutant):
    """Each planted oracle mistake makes at least one of the checks above fail."""
    monkeypatch.setattr(oracle_mod, "_lib", oracle_mod.load(oracle_mod.build_mutant(mutant, mutant_dir)))
    A_, E_, _ = oracle_mod.aet(np.array([1.0]), np.array([0.0]), np.array([0.0]))
    failures = [
        _projection_closed_form(oracle_mod) >= 5e-6,
        _tdi_closed_form(oracle_mod, "2nd generation") >= 2e-6,
        not np.isclose(A_[0], -1 / np.sqrt(2.0)),
    ]
    assert any(failures), f"mutant {mutant} went undetected"


# ------------------------------------------------------------------ extended-precision delays

def test_extended_delay_tdi_matches_delayed_sums(oracle_mod):
    """tdi_extended_delays against the analytic delayed sums (frozen unequal arms), the closed
    form the FP64 TDI loop is pinned by, to the same bound."""
    from fastlisaresponse_amd.tables import stock_tdi
    L = np.array([8.31, 8.33, 8.35, 8.32, 8.34, 8.36])
    o, _ = frozen_orbits(L)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    n, dt, T = 5000, 10.0, 1000
    t = np.arange(n) * dt
    rng = np.random.default_rng(7)
    f = rng.uniform(1e-3, 8e-3, 6)
    ph = rng.uniform(0, 6, 6)
    amp = rng.uniform(0.5, 2, 6)
    ylink = lambda r, s: amp[r] * np.sin(2 * np.pi * f[r] * s + ph[r])
    y = np.stack([ylink(r, t) for r in range(6)])
    combos = stock_tdi("2nd generation")
    xyz = oracle_mod.tdi_extended_delays(oo, y, 0.0, dt, combos, tdi_start=T)
    tabs = oracle_mod.tdi_tables(combos)
    row = {l: r for r, l in enumerate(o.LINKS)}
    ts = t[T:n - T]
    for c in range(3):
        exp = np.zeros_like(ts)
        for u_ in np.where(tabs["channels"] == c)[0]:
            s0, ln = tabs["unit_starts"][u_], tabs["unit_lengths"][u_]
            delay = sum(L[row[l]] for l in tabs["tdi_link_combinations"][s0:s0 + ln] if l != -11)
            exp += tabs["tdi_signs"][u_] * ylink(row[int(tabs["tdi_base_link"][u_])], ts - delay)
        assert rel_l2(xyz[c, T:n - T], exp) < 2e-6
    assert np.all(xyz[:, :T] == 0.0) and np.all(xyz[:, n - T:] == 0.0)


def test_extended_delay_tdi_differs_from_fp64_by_time_rounding_only(oracle_mod):
    """On the unequal-arm cartwheel the FP64 oracle and the extended-delay TDI of the same y
    differ by the FP64 rounding of each sample's delay, which grows with |t|: XYZ within 1e-11
    at t ~ 1e5 s; at t ~ 2.5e7 s (ulp 3.7e-9 s) about 1e-10 on X, Y, Z and ~1e-7 on the null
    channel T, where it does not cancel."""
    from fastlisaresponse_amd.orbits import CartwheelOrbits
    from conftest import gb_waveform
    o = CartwheelOrbits(t_max=3e7)
    o.configure(linear_interp_setup=True)
    oo = oracle_mod.OracleOrbits(*o.pycppdetector_args)
    n, T = 30000, 1000
    res = {}
    for t0 in (0.0, 2.5e7):
        h = gb_waveform(n, 10.0, 1e-21, 5e-3, 1e-17, 0.7, 0.2, 0.4, t0=t0)
        y, xyz = oracle_mod.response_pipeline(oo, h, 1.0, 0.3, 10.0, 25, "2nd generation", t0=t0,
                                              x_base=o.x_base, tdi_chan="XYZ", nthreads=4)
        ext = oracle_mod.tdi_extended_delays(oo, y, t0, 10.0, "2nd generation", tdi_start=T)
        res[t0] = [rel_l2(xyz[c, T:n - T], ext[c, T:n - T]) for c in range(3)]
        ta, te = oracle_mod.aet(*xyz), oracle_mod.aet(*ext)
        res[t0].append(rel_l2(ta[2][T:n - T], te[2][T:n - T]))
    assert max(res[0.0][:3]) < 1e-11, res
    assert 1e-11 < max(res[2.5e7][:3]) < 1e-9, res
    assert 1e-8 < res[2.5e7][3] < 1e-6, res
