"""FP64 CPU oracle of the fastlisaresponse response path (ctypes wrapper + host tables).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline leg of the benchmark
use this module as the checker of the HIP kernels. The product never imports it.

C part (liboracle_flr.so, built from flr_oracle.c by the Makefile next to this file):
    stencils, projection loop, TDI loop, orbit interpolation, XYZ -> AET.
Python part (this file), restating the host-side arithmetic of
src/fastlisaresponse/response.py:
    lagrange_tables   <- pyResponseTDI._fill_A_E                 response.py:273-310
    tdi_tables        <- _init_TDI_delays + tdi_combinations      response.py:312-433
    cyclic_permutation<- pyResponseTDI._cyclic_permutation        response.py:421-433
    buffers           <- get_projections buffer arithmetic         response.py:474-500
    sky_basis         <- get_projections k, u, v                   response.py:503-523
    unit_ranges       <- get_tdi_delays np.unique                  response.py:619-633

Parity unpinned (third-party lisaanalysistools is not in the image): the lisatools Orbits
interpolation formula (pinned against numpy.interp instead) and the lisatools AET convention
(pinned by orthonormality and the images of unit inputs). Everything else is checked against
closed forms, invariants and brute-force sums in tests/test_oracle_pinning.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from math import factorial

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle_flr.so")
# The same restatement with its value arithmetic in x87 extended precision (time, delay and
# index arithmetic unchanged): the FP64 path's exact values to ~2^-64 instead of ~2^-53.
WIDE_LIB_PATH = os.path.join(ORACLE_DIR, "liboracle_flr_wide.so")
C_INV = 3.3356409519815204e-09  # LISAResponse.hh:7, response.py:36

# The stock TDI combinations, response.py:318-356 (1st generation = first 8 entries).
TDI_X1 = [
    (13, [], +1),
    (31, [13], +1),
    (12, [13, 31], +1),
    (21, [13, 31, 12], +1),
    (12, [], -1),
    (21, [12], -1),
    (13, [12, 21], -1),
    (31, [12, 21, 13], -1),
]
TDI_X2_EXTRA = [
    (12, [13, 31, 12, 21], +1),
    (21, [13, 31, 12, 21, 12], +1),
    (13, [13, 31, 12, 21, 12, 21], +1),
    (31, [13, 31, 12, 21, 12, 21, 13], +1),
    (13, [12, 21, 13, 31], -1),
    (31, [12, 21, 13, 31, 13], -1),
    (12, [12, 21, 13, 31, 13, 31], -1),
    (21, [12, 21, 13, 31, 13, 31, 12], -1),
]


def build() -> str:
    """Compile liboracle_flr.so with the Makefile in this directory."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


class _Orbits(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("N", ctypes.c_int),
        ("ltt", ctypes.c_void_p),
        ("x", ctypes.c_void_p),
        ("links", ctypes.c_int * 6),
        ("sc_r", ctypes.c_int * 6),
        ("sc_e", ctypes.c_int * 6),
        ("armlength", ctypes.c_double),
    ]


_lib = None
_wide_lib = None


def build_mutant(mutant: int, out_dir: str) -> str:
    """Compile the oracle with one planted mistake (-DORACLE_MUTANT=mutant) into out_dir; used
    by the tests that show the oracle's checks detect such mistakes."""
    out = os.path.join(out_dir, f"liboracle_mutant{mutant}.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-shared", f"-DORACLE_MUTANT={mutant}", "-o", out,
                    os.path.join(ORACLE_DIR, "flr_oracle.c"), "-lm"], check=True)
    return out


def lib(wide: bool = False):
    """The FP64 oracle library, or (wide=True) its extended-precision twin."""
    global _lib, _wide_lib
    if wide:
        if _wide_lib is None:
            if not os.path.exists(WIDE_LIB_PATH):
                build()
            _wide_lib = load(WIDE_LIB_PATH)
            assert _wide_lib.oracle_is_wide() == 1
        return _wide_lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def _real(wide: bool):
    """numpy type of the value arithmetic: float64, or the C long double of the wide build."""
    return np.longdouble if wide else np.float64


def load(path: str):
    """Load an oracle library and declare its C signatures."""
    L = ctypes.CDLL(path)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.oracle_ltt.argtypes = [vp, cd, ci]
    L.oracle_ltt.restype = cd
    L.oracle_pos.argtypes = [vp, cd, ci, vp]
    L.oracle_pos.restype = None
    L.oracle_link_row.argtypes = [vp, ci]
    L.oracle_interp_single.argtypes = [vp, ci, ci, cd, vp, cd, vp]
    L.oracle_interp_single.restype = cd
    L.oracle_interp_cplx.argtypes = [vp, vp, ci, ci, cd, vp, cd, vp]
    L.oracle_interp_cplx.restype = None
    L.oracle_split_delay.argtypes = [cd, cd, cd, vp, vp]
    L.oracle_split_delay.restype = None
    L.oracle_get_response.argtypes = [vp, vp, vp, vp, vp, cd, ci, vp, ci, ci, cd, ci, vp, cd,
                                      ci, vp, ci, vp, ci]
    L.oracle_get_response.restype = ci
    L.oracle_get_tdi_delays.argtypes = [vp, vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, ci,
                                        cd, ci, vp, cd, ci, vp, ci, vp, ci]
    L.oracle_get_tdi_delays.restype = ci
    L.oracle_aet.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64]
    L.oracle_aet.restype = None
    L.oracle_max_threads.restype = ci
    L.oracle_is_wide.restype = ci
    return L


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class OracleOrbits:
    """Orbit tables in the lisatools `pycppdetector_args` layout, held for the C oracle."""

    def __init__(self, dt, N, n_arr, ltt_arr, x_arr, links, sc_r, sc_e, armlength):
        self.ltt = np.ascontiguousarray(ltt_arr, dtype=np.float64).reshape(-1)
        self.x = np.ascontiguousarray(x_arr, dtype=np.float64).reshape(-1)
        assert self.ltt.size == N * 6 and self.x.size == N * 9
        self.s = _Orbits()
        self.s.dt = float(dt)
        self.s.N = int(N)
        self.s.ltt = _p(self.ltt)
        self.s.x = _p(self.x)
        for r in range(6):
            self.s.links[r] = int(links[r])
            self.s.sc_r[r] = int(sc_r[r])
            self.s.sc_e[r] = int(sc_e[r])
        self.s.armlength = float(armlength)

    @property
    def ptr(self):
        return ctypes.addressof(self.s)

    def ltt_at(self, t, link) -> np.ndarray:
        t = np.atleast_1d(np.asarray(t, dtype=np.float64))
        return np.array([lib().oracle_ltt(self.ptr, float(ti), int(link)) for ti in t])

    def pos_at(self, t, sc) -> np.ndarray:
        t = np.atleast_1d(np.asarray(t, dtype=np.float64))
        out = np.zeros((t.size, 3))
        buf = np.zeros(3)
        for i, ti in enumerate(t):
            lib().oracle_pos(self.ptr, float(ti), int(sc), _p(buf))
            out[i] = buf
        return out


# ---------------------------------------------------------------- host tables


def lagrange_tables(order: int):
    """A_in [1001], deps, E_in [h] exactly as pyResponseTDI._fill_A_E (response.py:273-310)."""
    fact = np.asarray([float(factorial(n)) for n in range(40)])
    num_A = 1001
    deps = 1.0 / (num_A - 1)
    eps = np.arange(num_A) * deps
    h = int((order + 1) / 2)
    denominator = fact[h - 1] * fact[h]
    A_in = np.zeros_like(eps)
    for j, e in enumerate(eps):
        A = 1.0
        for i in range(1, h):
            A *= (i + e) * (i + 1 - e)
        A /= denominator
        A_in[j] = A
    E_in = np.zeros((h,))
    for j in range(1, h):
        first = fact[h - 1] / fact[h - 1 - j]
        second = fact[h] / fact[h + j]
        value = first * second
        value = value * (-1.0) ** j
        E_in[j - 1] = value
    return A_in, deps, num_A, E_in


def cyclic_permutation(link: int, permutation: int) -> int:
    """Spacecraft digits shifted by `permutation` modulo 3 (response.py:421-433)."""
    out = ""
    for ch in str(link)[:2]:
        temp = int(ch) + permutation
        if temp > 3:
            temp = temp % 3
        out += str(temp)
    return int(out)


def tdi_tables(combinations):
    """Unit tables for three channels by cyclic permutation (response.py:372-419).

    `combinations` is a list of (link, links_for_delay, sign) or of the reference's dicts."""
    norm = []
    for c in combinations:
        if isinstance(c, dict):
            norm.append((c["link"], list(c["links_for_delay"]), c["sign"]))
        else:
            norm.append(c)
    base, comb, signs, opidx, chans = [], [], [], [], []
    idx = 0
    for perm in range(3):
        for link, delays, sign in norm:
            base.append(cyclic_permutation(link, perm))
            signs.append(float(sign))
            chans.append(perm)
            if len(delays) == 0:
                comb.append(-11)
                opidx.append(idx)
            else:
                for ld in delays:
                    comb.append(cyclic_permutation(ld, perm))
                    opidx.append(idx)
            idx += 1
    opidx = np.asarray(opidx, dtype=np.int32)
    _, starts, lengths = np.unique(opidx, return_index=True, return_counts=True)
    return dict(
        tdi_base_link=np.asarray(base, dtype=np.int32),
        tdi_link_combinations=np.asarray(comb, dtype=np.int32),
        tdi_signs=np.asarray(signs, dtype=np.float64),
        channels=np.asarray(chans, dtype=np.int32),
        tdi_operation_index=opidx,
        unit_starts=starts.astype(np.int32),
        unit_lengths=lengths.astype(np.int32),
        num_units=int(opidx.max() + 1),
    )


def stock_combinations(generation: str):
    if generation == "1st generation":
        return list(TDI_X1)
    if generation == "2nd generation":
        return list(TDI_X1) + list(TDI_X2_EXTRA)
    raise ValueError(generation)


def buffers(dt: float, order: int, t_buffer: float, x_base: np.ndarray):
    """(tdi_start_ind, projections_start_ind, projection_buffer), response.py:474-500."""
    fs = 1 / dt
    tdi_start_ind = int(t_buffer / dt)
    check_tdi_buffer = int(100.0 * fs) + 4 * order
    projection_buffer = int((np.sum(x_base * x_base, axis=-1) ** (1 / 2)).max() * C_INV) + 4 * order
    projections_start_ind = tdi_start_ind - 2 * check_tdi_buffer
    return tdi_start_ind, projections_start_ind, projection_buffer


def sky_basis(lam: float, beta: float):
    """(k, u, v) as get_projections builds them, response.py:503-523."""
    k = np.zeros(3)
    u = np.zeros(3)
    v = np.zeros(3)
    cosbeta, sinbeta = np.cos(beta), np.sin(beta)
    coslam, sinlam = np.cos(lam), np.sin(lam)
    v[0] = -sinbeta * coslam
    v[1] = -sinbeta * sinlam
    v[2] = cosbeta
    u[0] = sinlam
    u[1] = -coslam
    u[2] = 0.0
    k[0] = -cosbeta * coslam
    k[1] = -cosbeta * sinlam
    k[2] = -sinbeta
    return k, u, v


# ---------------------------------------------------------------- kernels


def interp_single(x: np.ndarray, h: int, d: int, e: float, A_in, deps, E_in) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_interp_single(_p(x), h, d, e, _p(A_in), deps, _p(E_in))


def interp_cplx(x: np.ndarray, h: int, d: int, e: float, A_in, deps, E_in) -> complex:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(2)
    lib().oracle_interp_cplx(_p(out), _p(x), h, d, e, _p(A_in), deps, _p(E_in))
    return complex(out[0], out[1])


def split_delay(delay: float, t0: float, fs: float):
    d = ctypes.c_int(0)
    e = ctypes.c_double(0.0)
    lib().oracle_split_delay(delay, t0, fs, ctypes.byref(d), ctypes.byref(e))
    return d.value, e.value


def get_response(orbits: OracleOrbits, input_in: np.ndarray, t_data: np.ndarray, k, u, v, order,
                 sampling_frequency, A_in, deps, E_in, projections_start_ind, nthreads=1,
                 wide=False):
    """y_gw [6][N] (zeros outside the computed range), like get_projections + get_response;
    wide=True: the extended-precision build, y as long double."""
    h = np.ascontiguousarray(input_in, dtype=np.complex128)
    t_data = np.ascontiguousarray(t_data, dtype=np.float64)
    n = len(h)
    y = np.zeros(6 * n, dtype=_real(wide))
    k, u, v = (np.ascontiguousarray(a, dtype=np.float64) for a in (k, u, v))
    A_in = np.ascontiguousarray(A_in, dtype=np.float64)
    E_in = np.ascontiguousarray(E_in, dtype=np.float64)
    st = lib(wide).oracle_get_response(_p(y), _p(t_data), _p(k), _p(u), _p(v), 1.0 / sampling_frequency,
                                   n, _p(h), n, order, sampling_frequency, 2 * order + 1, _p(A_in),
                                   deps, len(A_in), _p(E_in), projections_start_ind, orbits.ptr,
                                   nthreads)
    assert st == 0, st
    return y.reshape(6, n)


def get_tdi_delays(orbits: OracleOrbits, y_gw: np.ndarray, t_arr: np.ndarray, tables: dict, order,
                   sampling_frequency, A_in, deps, E_in, tdi_start_ind, nthreads=1,
                   delayed_links=None, wide=False):
    """X, Y, Z [3][N] accumulated into a zeroed (or given) buffer, like get_tdi_delays;
    wide=True: the extended-precision build (y and the outputs as long double)."""
    rt = _real(wide)
    y = np.ascontiguousarray(y_gw, dtype=rt).reshape(-1)
    n = y.size // 6
    t_arr = np.ascontiguousarray(t_arr, dtype=np.float64)
    out = (np.zeros(3 * n, dtype=rt) if delayed_links is None
           else np.ascontiguousarray(delayed_links, dtype=rt).reshape(-1).copy())
    A_in = np.ascontiguousarray(A_in, dtype=np.float64)
    E_in = np.ascontiguousarray(E_in, dtype=np.float64)
    tb = {k_: np.ascontiguousarray(v_) for k_, v_ in tables.items() if isinstance(v_, np.ndarray)}
    st = lib(wide).oracle_get_tdi_delays(
        _p(out), _p(y), n, n, _p(t_arr), _p(tb["unit_starts"]), _p(tb["unit_lengths"]),
        _p(tb["tdi_base_link"]), _p(tb["tdi_link_combinations"]), _p(tb["tdi_signs"]),
        _p(tb["channels"]), tables["num_units"], 3, order, sampling_frequency, 2 * order + 1,
        _p(A_in), deps, len(A_in), _p(E_in), tdi_start_ind, orbits.ptr, nthreads)
    assert st == 0, st
    return out.reshape(3, n)


def aet(X, Y, Z, wide=False):
    X, Y, Z = (np.ascontiguousarray(a, dtype=_real(wide)) for a in (X, Y, Z))
    A, E, T = np.zeros_like(X), np.zeros_like(X), np.zeros_like(X)
    lib(wide).oracle_aet(_p(X), _p(Y), _p(Z), _p(A), _p(E), _p(T), X.size)
    return A, E, T


def response_pipeline(orbits: OracleOrbits, input_in, lam, beta, dt, order, tdi, t0=0.0,
                      t_buffer=10000.0, tdi_chan="AET", x_base=None, nthreads=1, wide=False):
    """Projection -> TDI -> channel selection for one source, as pyResponseTDI runs it
    (wide=True: the extended-precision build; arrays returned as long double)."""
    fs = 1 / dt
    n = len(input_in)
    A_in, deps, num_A, E_in = lagrange_tables(order)
    tdi_start, proj_start, proj_buf = buffers(dt, order, t_buffer, x_base)
    if proj_start < proj_buf:
        raise ValueError("Need to increase t_buffer. The initial buffer is not large enough.")
    k, u, v = sky_basis(lam, beta)
    t_data = t0 + np.arange(n) * dt
    y = get_response(orbits, input_in, t_data, k, u, v, order, fs, A_in, deps, E_in, proj_start,
                     nthreads, wide=wide)
    combos = stock_combinations(tdi) if isinstance(tdi, str) else tdi
    tables = tdi_tables(combos)
    xyz = get_tdi_delays(orbits, y, t_data, tables, order, fs, A_in, deps, E_in, tdi_start, nthreads,
                         wide=wide)
    if tdi_chan == "XYZ":
        return y, xyz
    a, e, t = aet(*xyz, wide=wide)
    return y, np.stack([a, e, t])


def tdi_extended_delays(orbits: OracleOrbits, y: np.ndarray, t0: float, dt: float, tdi,
                        order: int = 25, tdi_start: int = 0, chan: str = "XYZ"):
    """The TDI loop of get_tdi_delays (LISAResponse.cu:300-393) over the FP64 projections y
    [6][N], with every sample's delay arithmetic -- t = t0 + i dt, the orbit tables' linear
    interpolation at t, the chain subtraction, the offset and the scale by fs -- carried in
    numpy.longdouble (x87 extended, 64-bit mantissa) before the integer/fraction split, and the
    stencils in FP64 in interp_single's operation order. The per-sample delays of the FP64 path
    round times of ~1e7 s to ~1e-9 s, which TDI amplifies in T wherever the arms differ; this is
    the same response without that rounding, against which forms that compute their delays
    differently are judged. Returns X, Y, Z (or A, E, T) on [tdi_start, N - tdi_start), zero
    outside, like get_tdi_delays."""
    LD = np.longdouble
    y = np.asarray(y, dtype=np.float64)
    n = y.shape[1]
    fs = 1.0 / dt
    A_in, deps, num_A, E_in = lagrange_tables(order)
    h = int((order + 1) / 2)
    o = orbits
    odt, oN = LD(o.s.dt), int(o.s.N)
    ltt_tab = o.ltt.reshape(oN, 6)
    row = {int(o.s.links[r]): r for r in range(6)}
    ii = np.arange(tdi_start, n - tdi_start)
    t = LD(t0) + ii.astype(LD) * LD(dt)
    q = (t / odt).astype(np.float64)
    w = np.clip(np.where(q >= 0.0, q, 0.0).astype(np.int64), 0, oN - 2)
    tau = t - odt * w.astype(LD)
    L = {}
    for link, r in row.items():
        down = ltt_tab[w, r].astype(LD)
        up = ltt_tab[w + 1, r].astype(LD)
        L[link] = (up - down) / odt * tau + down
    combos = stock_combinations(tdi) if isinstance(tdi, str) else tdi
    tables = tdi_tables(combos)
    out = np.zeros((3, n))
    for u in range(tables["num_units"]):
        c = int(tables["channels"][u])
        delay = t.copy()
        for k in range(tables["unit_starts"][u], tables["unit_starts"][u] + tables["unit_lengths"][u]):
            link = int(tables["tdi_link_combinations"][k])
            if link != -11:
                delay = delay - L[link]
        cf = (delay - LD(t0)) * LD(fs)
        d = np.ceil(cf).astype(np.int64) - 1
        e = (LD(1) + d.astype(LD) - cf).astype(np.float64)
        x = y[row[int(tables["tdi_base_link"][u])]]
        ind = np.minimum((e / deps).astype(np.int64), num_A - 2)  # e = 1 exactly: last segment
        frac = (e - ind * deps) / deps
        A = A_in[ind] * (1.0 - frac) + A_in[ind + 1] * frac
        acc = np.zeros(ii.size)
        for j in range(1, h):
            acc += E_in[j - 1] * (x[d + 1 + j] / (j + e) + x[d - j] / (j + (1 - e)))
        val = A * ((1.0 - e) * x[d + 1] + e * x[d] + (e * (1.0 - e)) * acc)
        out[c, tdi_start:n - tdi_start] += tables["tdi_signs"][u] * val
    if chan == "XYZ":
        return out
    return np.stack(aet(*out))


def max_threads() -> int:
    return lib().oracle_max_threads()
