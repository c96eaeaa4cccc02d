"""Host-side tables and index arithmetic of the response (no GPU needed).

These reproduce the numbers the reference's host layer hands to its kernels, so that the same
inputs reach the device:
    lagrange_tables    the A(eps) table and E_j weights        response.py:273-310
    stock_tdi          1st/2nd-generation Michelson X terms     response.py:316-356
    tdi_tables         three channels by cyclic permutation    response.py:372-433
    buffer_indices     tdi_start / projection start / check    response.py:474-500
    sky_basis          propagation k and polarisation u, v     response.py:503-523
"""
from __future__ import annotations

from math import factorial
from typing import Iterable, List, Sequence

import numpy as np

C_INV = 3.3356409519815204e-09  # s/m
NUM_A = 1001

# Michelson X combination: (link, links_for_delay, sign); link ij is received by i from j.
_X_FIRST = (
    (13, (), +1), (31, (13,), +1), (12, (13, 31), +1), (21, (13, 31, 12), +1),
    (12, (), -1), (21, (12,), -1), (13, (12, 21), -1), (31, (12, 21, 13), -1),
)
_X_SECOND = (
    (12, (13, 31, 12, 21), +1), (21, (13, 31, 12, 21, 12), +1),
    (13, (13, 31, 12, 21, 12, 21), +1), (31, (13, 31, 12, 21, 12, 21, 13), +1),
    (13, (12, 21, 13, 31), -1), (31, (12, 21, 13, 31, 13), -1),
    (12, (12, 21, 13, 31, 13, 31), -1), (21, (12, 21, 13, 31, 13, 31, 12), -1),
)


def stock_tdi(generation: str) -> List[dict]:
    """The X combination as the list-of-dicts the API accepts."""
    if generation == "1st generation":
        terms = _X_FIRST
    elif generation == "2nd generation":
        terms = _X_FIRST + _X_SECOND
    else:
        raise ValueError(
            "tdi kwarg should be '1st generation', '2nd generation', or a list with a specific "
            "tdi combination.")
    return [{"link": l, "links_for_delay": list(d), "sign": s} for l, d, s in terms]


def lagrange_tables(order: int):
    """(A_in [1001], deps, E_in [h]) for a Lagrange stencil of order `order`, h = (order+1)//2.

    A(eps) = prod_{i=1}^{h-1} (i+eps)(i+1-eps) / ((h-1)! h!) on eps = k/1000, and
    E_j = (-1)^j (h-1)!/(h-1-j)! * h!/(h+j)! for j = 1..h-1 (E_in[h-1] = 0), with the
    factorials rounded to float first, exactly as the reference builds them."""
    if order < 1:
        raise ValueError("order must be >= 1")
    h = (order + 1) // 2
    fact = [float(factorial(n)) for n in range(2 * h + 2)]
    deps = 1.0 / (NUM_A - 1)
    eps = np.arange(NUM_A) * deps
    denom = fact[h - 1] * fact[h]
    A = np.ones(NUM_A)
    for i in range(1, h):
        A = A * ((i + eps) * (i + 1 - eps))
    A = A / denom
    E = np.zeros(h)
    for j in range(1, h):
        val = (fact[h - 1] / fact[h - 1 - j]) * (fact[h] / fact[h + j])
        E[j - 1] = val * (-1.0) ** j
    return A, deps, E


def cyclic_permutation(link: int, shift: int) -> int:
    """Link number with both spacecraft indices advanced by `shift` (1 -> 2 -> 3 -> 1)."""
    a, b = divmod(int(link), 10)
    if not (1 <= a <= 3 and 1 <= b <= 3):
        raise ValueError(f"bad link {link}")
    return 10 * ((a - 1 + shift) % 3 + 1) + ((b - 1 + shift) % 3 + 1)


def tdi_tables(combinations: Sequence[dict]) -> dict:
    """Unit tables for channels X, Y, Z (Y, Z = X under cyclic shifts 1, 2).

    Every combination term becomes one unit per channel; an undelayed term stores the sentinel
    -11 as its single chain entry. Returns int32/float64 arrays plus unit start/length ranges."""
    base, chain, signs, channels, op_index = [], [], [], [], []
    unit = 0
    for ch in range(3):
        for term in combinations:
            base.append(cyclic_permutation(term["link"], ch))
            signs.append(float(term["sign"]))
            channels.append(ch)
            delays = list(term["links_for_delay"])
            if not delays:
                chain.append(-11)
                op_index.append(unit)
            for link in delays:
                chain.append(cyclic_permutation(link, ch))
                op_index.append(unit)
            unit += 1
    op_index = np.asarray(op_index, dtype=np.int32)
    starts = np.searchsorted(op_index, np.arange(unit), side="left").astype(np.int32)
    lengths = np.bincount(op_index, minlength=unit).astype(np.int32)
    return dict(
        tdi_base_links=np.asarray(base, dtype=np.int32),
        tdi_link_combinations=np.asarray(chain, dtype=np.int32),
        tdi_signs=np.asarray(signs, dtype=np.float64),
        channels=np.asarray(channels, dtype=np.int32),
        tdi_operation_index=op_index,
        unit_starts=starts,
        unit_lengths=lengths,
        num_units=unit,
    )


def buffer_indices(sampling_frequency: float, order: int, t_buffer: float, x_base: np.ndarray):
    """(tdi_start_ind, check_tdi_buffer, projection_buffer, projections_start_ind).

    The projection buffer compares max|x| * (1/c) -- seconds -- with a sample index, exactly as
    the reference does; it is a conservative check at dt >= 1 s."""
    dt = 1 / sampling_frequency
    tdi_start_ind = int(t_buffer / dt)
    check_tdi_buffer = int(100.0 * sampling_frequency) + 4 * order
    radius = np.sqrt(np.sum(np.asarray(x_base) * np.asarray(x_base), axis=-1)).max()
    projection_buffer = int(radius * C_INV) + 4 * order
    projections_start_ind = tdi_start_ind - 2 * check_tdi_buffer
    return tdi_start_ind, check_tdi_buffer, projection_buffer, projections_start_ind


def sky_basis(lam: float, beta: float):
    """(k, u, v): propagation direction and polarisation basis for ecliptic (lam, beta)."""
    cb, sb = np.cos(beta), np.sin(beta)
    cl, sl = np.cos(lam), np.sin(lam)
    k = np.array([-cb * cl, -cb * sl, -sb])
    u = np.array([sl, -cl, 0.0])
    v = np.array([-sb * cl, -sb * sl, cb])
    return k, u, v
