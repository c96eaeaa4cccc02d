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
