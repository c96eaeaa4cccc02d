import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PARENT = os.path.join(ROOT, "lisa-on-gpu_amd")
for p in (ROOT, PKG_PARENT):
    if p not in sys.path:
        sys.path.insert(0, p)

YRSID_SI = 31558149.763545603


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); parity tests of the kernels")


def gb_waveform(n, dt, A=1e-21, f=2e-3, fdot=1e-17, iota=0.8, phi0=0.3, psi=0.5, t0=0.0):
    """h+ + i hx of a galactic binary, the generator of tests/test_fastlisaresponse.py:41-63
    (phase 2 pi (f t + fdot t^2/2 + fddot t^3/6) - phi0, rotated by psi) on t = t0 + k dt."""
    t = t0 + np.arange(n) * dt
    cos2psi, sin2psi, cosiota = np.cos(2.0 * psi), np.sin(2.0 * psi), np.cos(iota)
    fddot = 11.0 / 3.0 * fdot ** 2 / f
    phase = 2 * np.pi * (f * t + 0.5 * fdot * t ** 2 + 1.0 / 6.0 * fddot * t ** 3) - phi0
    hSp = -np.cos(phase) * A * (1.0 + cosiota * cosiota)
    hSc = -np.sin(phase) * 2.0 * A * cosiota
    hp = hSp * cos2psi - hSc * sin2psi
    hc = hSp * sin2psi + hSc * cos2psi
    return hp + 1j * hc


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as orc
    orc.lib()
    return orc


def unequal_arms(orbits) -> bool:
    """True when the orbit tables' six light travel times are not all bitwise equal."""
    return not np.all(orbits.ltt == orbits.ltt.reshape(-1)[0])


def check_mixed(oracle_mod, oo, got, y, exp, t0, dt, tdi, chan, tol, T, label=(), unequal=False):
    """Channels of a mixed form (got) against the oracle's (exp, from projections y), rel-L2 over
    [T, N - T) within tol. With unequal arms and A, E, T outputs the null channel T is judged
    against the TDI of the same projections with extended-precision delays instead
    (oracle.tdi_extended_delays): within the oracle's own distance from it plus tol. The oracle
    rounds each sample's delay, a time of up to ~1e7 s, to FP64, and T amplifies that rounding
    where the arms differ; the slowly-varying-arm form computes its delays without it, so its T
    and the oracle's differ by the oracle's floor. Returns the per-channel errors."""
    errs = []
    ext = None
    if unequal and chan == "AET":
        ext = oracle_mod.tdi_extended_delays(oo, y, t0, dt, tdi, tdi_start=T, chan="AET")
    for c in range(3):
        r = rel_l2(got[c, T:-T], exp[c, T:-T])
        if ext is not None and c == 2:
            floor = rel_l2(exp[2, T:-T], ext[2, T:-T])
            r_ext = rel_l2(got[2, T:-T], ext[2, T:-T])
            assert r_ext <= floor + tol, (*label, c, r, r_ext, floor)
            errs.append(r_ext)
        else:
            assert r <= tol, (*label, c, r)
            errs.append(r)
    return errs
