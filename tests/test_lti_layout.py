"""CPU replays of the mixed equal-arm TDI's device arithmetic (no GPU needed).

* The matrix-core composite-filter pass (csrc/flr_fused.hpp:lti_mfma): its operand indexing
  -- Toeplitz filter from the zero-padded copy, series from padded window rows, edge blocks
  clamped, the measured v_mfma_f64_16x16x4f64 lane layout (A[l%16][l/16], B[l/16][l%16],
  result element v = D[4v + l/16][l%16]) -- replayed with numpy products and compared with the
  direct 40-tap sums, with every LDS index checked against its array.
* The per-tile test (csrc/flr_fused.hpp:lti_tile_matches): its rounding bound replayed against
  the per-sample chain positions computed in the device's operation order.
"""
import math

import numpy as np

KFIR, KFPAD, KFPADLO = 40, 72, 15


def _col(w):
    return w + (w >> 4)


def _pitch(cap):
    return cap + (cap >> 4) + 1


def test_matrix_core_filter_pass_matches_direct_sums():
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(40):
        passes = int(rng.choice([1, 2, 4]))
        halo_lo, halo_hi = int(rng.integers(19, 24)), int(rng.integers(27, 30))
        win = passes * 256
        tile_len = win - halo_lo - halo_hi
        pitch = _pitch(win)
        n = int(rng.integers(3000, 6000))
        T = int(rng.integers(halo_lo + 5, 900))
        zero_edges = bool(rng.integers(0, 2))
        kmin = -int(rng.integers(halo_lo - 3, halo_lo + 1))
        F = rng.normal(size=(2, KFIR))
        fpad = np.zeros((2, KFPAD))
        for kind in range(2):
            for x in range(KFPAD):
                if 0 <= x - KFPADLO < KFIR:
                    fpad[kind, x] = F[kind, x - KFPADLO]
        ntiles = ((n if zero_edges else n - T) + tile_len - 1) // tile_len
        for tile in range(0, ntiles, 3):
            i0 = (0 if zero_edges else T) + tile * tile_len
            first, last = max(i0, T), min(i0 + tile_len, n - T) - 1
            if first > last:
                continue
            yw_lo = first - halo_lo
            yw_len = last + halo_hi - yw_lo + 1
            S = rng.normal(size=(6, pitch))
            q0, q1 = (first - T) >> 4, (last - T) >> 4
            ngc = (q1 - q0 + 16) >> 4
            lanes = np.arange(64)
            nb, kq = lanes & 15, lanes >> 4
            got = {}
            for g in range(3 * ngc):
                o = g // ngc
                qb = q0 + 16 * (g - o * ngc) + nb
                wb = T + 16 * np.minimum(qb, q1) + kmin - yw_lo
                A = np.zeros((16, 112))
                B = np.zeros((112, 16))
                for kind in range(2):
                    for j in range(14):
                        fidx = kind * KFPAD + KFPADLO + kq - nb + 4 * j
                        assert fidx.min() >= 0 and fidx.max() < 2 * KFPAD
                        w = np.clip(wb + 4 * j + kq, 0, yw_len - 1)
                        lidx = (2 * o + kind) * pitch + _col(w)
                        assert lidx.min() >= 0 and lidx.max() < 6 * pitch
                        A[nb, kind * 56 + 4 * j + kq] = fpad.reshape(-1)[fidx]
                        B[kind * 56 + 4 * j + kq, nb] = S.reshape(-1)[lidx]
                D = A @ B
                for lane in range(64):
                    if qb[lane] > q1:
                        continue
                    for v in range(4):
                        i = T + 16 * qb[lane] + 4 * v + kq[lane]
                        if first <= i <= last:
                            assert 0 <= i - i0 < passes * 256
                            got[(o, i)] = D[4 * v + kq[lane], nb[lane]]
            for i in range(first, last + 1, 7):
                ft = i + kmin - yw_lo
                assert ft >= 0 and ft + KFIR <= yw_len
                for o in range(3):
                    exp = sum(F[kind, k] * S[2 * o + kind, _col(ft + k)]
                              for kind in range(2) for k in range(KFIR))
                    assert abs(got[(o, i)] - exp) <= 1e-12 * (1 + abs(exp)), (o, i)
                    checked += 1
    assert checked > 1000


def _ulp(x):
    return math.ldexp(1.0, math.frexp(x)[1] - 53)


def test_tile_bound_covers_the_delay_rounding():
    """lti_tile_matches: 2 * per_sample + drift bounds |delta_m(i) - delta_m(T)| for every i."""
    L = 2.5e9 / 299792458.0
    for dt, t0, n, M, passes in [(10.0, 0.0, 3153600, 7, True), (5.0, 0.0, 1 << 21, 7, True),
                                 (2.0, 0.0, 1 << 21, 7, True), (10.0, 0.0, 1 << 20, 3, True),
                                 (10.0, 1.0e9, 1 << 18, 7, False)]:
        fs = 1.0 / dt
        T = int(10000.0 / dt)
        i = np.arange(T, n - T, dtype=np.int64)
        t = t0 + i.astype(np.float64) * dt
        tr = t0 + float(T) * dt
        Dm, Dr = t.copy(), tr
        worst = 0.0
        for m in range(M + 1):
            if m > 0:
                Dm, Dr = Dm - L, Dr - L
            cf = (Dm - t0) * fs
            k = np.ceil(cf).astype(np.int64) - 1
            delta = (k - i).astype(np.float64) + 1.0 - ((1.0 + k.astype(np.float64)) - cf)
            cfr = (Dr - t0) * fs
            kr = math.ceil(cfr) - 1
            dr = float(kr - T) + 1.0 - ((1.0 + kr) - cfr)
            worst = max(worst, float(np.abs(delta - dr).max()))
        last = n - T - 1
        tmax = 1.001 * (2 * abs(t0) + last * dt + M * L)
        per_sample = (0.5 * M + 1.5) * _ulp(tmax) * fs + _ulp(tmax * fs)
        from fractions import Fraction
        drift = 1.01 * abs(float(Fraction(dt) * Fraction(fs) - 1)) * (last - T)
        bound = 2 * per_sample + drift + 1e-13
        assert worst <= bound, (dt, t0, worst, bound)
        assert (bound <= 1e-8) == passes, (dt, t0, bound)
