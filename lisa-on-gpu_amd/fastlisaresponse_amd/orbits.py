"""Orbit providers with the surface the response uses from lisatools' `Orbits`.

The reference takes its orbits from the third-party package lisaanalysistools
(`lisatools.detector.Orbits` / `EqualArmlengthOrbits`, imported at response.py:13 and used at
response.py:154,166,227-239,444,481). That package is not part of this framework, so this
module supplies the same surface:

* `t_base`, `x_base` [N,3,3], `ltt_base` [N,6], `n_base` [N,6,3]: the orbit data;
* `configure(t_arr=None, dt=None, linear_interp_setup=False)`: put the data on the uniform grid
  the device interpolates on (node k at t = k*dt); `configured`, `t`, `x`, `ltt`, `n`, `dt`;
* `pycppdetector_args`: (dt, N, n_arr, ltt_arr, x_arr, links, sc_r, sc_e, armlength), the
  argument list of the device orbit object (binding_flr.hpp:35);
* `LINKS` = [12, 23, 31, 13, 32, 21] (link ij: received by spacecraft i, emitted by j), `SC`.

`EqualArmlengthOrbits` is an analytic cartwheel: spacecraft on heliocentric orbits of radius
1 AU and period one (sidereal) year, in the classic first-order-eccentricity constellation
(x = R cos a + e R/2 (cos(2a - b_i) - 3 cos b_i), y = R sin a + e R/2 (sin(2a - b_i) - 3 sin b_i),
z = -sqrt(3) e R cos(a - b_i), a = 2 pi t / yr, b_i = 2 pi (i-1)/3, e = L / (2 sqrt(3) R)),
with every light travel time equal to armlength / c.
"""
from __future__ import annotations

from copy import deepcopy
from typing import List, Optional

import numpy as np

YRSID_SI = 31558149.763545603
C_SI = 299792458.0
AU_SI = 149597870700.0


class Orbits:
    """Tabulated orbits of the three LISA spacecraft.

    Args:
        t_base: sample times of the orbit data (s), increasing, starting at or before 0.
        x_base: spacecraft positions [len(t_base), 3, 3] (m).
        ltt_base: light travel times [len(t_base), 6] (s), columns in `LINKS` order.
        n_base: link unit vectors [len(t_base), 6, 3]; computed from x_base when omitted.
        armlength: nominal armlength (m).
    """

    LINKS: List[int] = [12, 23, 31, 13, 32, 21]
    SC: List[int] = [1, 2, 3]

    def __init__(self, t_base, x_base, ltt_base, n_base=None, armlength: float = 2.5e9):
        self.t_base = np.asarray(t_base, dtype=np.float64)
        self.x_base = np.asarray(x_base, dtype=np.float64)
        self.ltt_base = np.asarray(ltt_base, dtype=np.float64)
        n = len(self.t_base)
        if self.x_base.shape != (n, 3, 3) or self.ltt_base.shape != (n, 6):
            raise ValueError("x_base must be [N,3,3] and ltt_base [N,6] for N = len(t_base)")
        if n < 2 or np.any(np.diff(self.t_base) <= 0):
            raise ValueError("t_base must be increasing with at least two samples")
        self.n_base = (np.asarray(n_base, dtype=np.float64) if n_base is not None
                       else self._link_unit_vectors(self.x_base))
        self.armlength = float(armlength)
        self.configured = False
        self.t = self.x = self.ltt = self.n = None
        self.dt = None

    # ------------------------------------------------------------- link bookkeeping
    @property
    def link_space_craft_r(self) -> List[int]:
        """Receiving spacecraft of each link in `LINKS` order."""
        return [int(str(link)[0]) for link in self.LINKS]

    @property
    def link_space_craft_e(self) -> List[int]:
        """Emitting spacecraft of each link in `LINKS` order."""
        return [int(str(link)[1]) for link in self.LINKS]

    def _link_unit_vectors(self, x: np.ndarray) -> np.ndarray:
        out = np.zeros((x.shape[0], 6, 3))
        for r, (sr, se) in enumerate(zip(self.link_space_craft_r, self.link_space_craft_e)):
            d = x[:, sr - 1] - x[:, se - 1]
            out[:, r] = d / np.linalg.norm(d, axis=-1, keepdims=True)
        return out

    # ------------------------------------------------------------- evaluation
    def _evaluate(self, t: np.ndarray):
        """Positions [n,3,3] and light travel times [n,6] at times t (cubic-spline resampling
        of the base data; subclasses with an analytic model override this)."""
        from scipy.interpolate import CubicSpline

        x = CubicSpline(self.t_base, self.x_base, axis=0)(t)
        ltt = CubicSpline(self.t_base, self.ltt_base, axis=0)(t)
        return x, ltt

    def configure(self, t_arr: Optional[np.ndarray] = None, dt: Optional[float] = None,
                  linear_interp_setup: bool = False) -> None:
        """Tabulate the orbits on the uniform grid t_k = k*dt used by the device interpolation.

        With `linear_interp_setup` the grid spans [0, t_base.max()] with spacing `dt` (default:
        the base spacing). A given `t_arr` must be such a uniform grid starting at 0."""
        if t_arr is not None:
            t = np.asarray(t_arr, dtype=np.float64)
            if len(t) < 2 or t[0] != 0.0 or not np.allclose(np.diff(t), t[1] - t[0], rtol=1e-12, atol=0):
                raise ValueError("t_arr must be a uniform grid starting at 0")
            dt = float(t[1] - t[0])
        else:
            if dt is None:
                dt = float(np.median(np.diff(self.t_base)))
            if not dt > 0:
                raise ValueError("dt must be > 0")
            t = np.arange(0.0, self.t_base.max() + 0.5 * dt, dt)
            t = t[t <= self.t_base.max()]
        x, ltt = self._evaluate(t)
        self.t = t
        self.dt = float(dt)
        self.x = np.ascontiguousarray(x)
        self.ltt = np.ascontiguousarray(ltt)
        self.n = self._link_unit_vectors(self.x)
        self.configured = True

    @property
    def pycppdetector_args(self):
        """Argument list of the device orbit object: (dt, N, n_arr, ltt_arr, x_arr, links,
        sc_r, sc_e, armlength), arrays flattened in row-major order."""
        if not self.configured:
            raise ValueError("Orbits must be configured first (configure(linear_interp_setup=True)).")
        return (
            float(self.dt),
            int(len(self.t)),
            self.n.reshape(-1).copy(),
            self.ltt.reshape(-1).copy(),
            self.x.reshape(-1).copy(),
            np.asarray(self.LINKS, dtype=np.int32),
            np.asarray(self.link_space_craft_r, dtype=np.int32),
            np.asarray(self.link_space_craft_e, dtype=np.int32),
            float(self.armlength),
        )

    def copy(self) -> "Orbits":
        return deepcopy(self)


def cartwheel_positions(t: np.ndarray, armlength: float = 2.5e9, radius: float = AU_SI,
                        period: float = YRSID_SI, kappa: float = 0.0, lambda0: float = 0.0):
    """Positions [len(t), 3, 3] of the first-order-eccentricity cartwheel constellation."""
    t = np.asarray(t, dtype=np.float64)
    alpha = 2.0 * np.pi * t / period + kappa
    ecc = armlength / (2.0 * np.sqrt(3.0) * radius)
    out = np.zeros((len(t), 3, 3))
    for i in range(3):
        beta = 2.0 * np.pi * i / 3.0 + lambda0
        out[:, i, 0] = radius * np.cos(alpha) + 0.5 * ecc * radius * (np.cos(2 * alpha - beta) - 3 * np.cos(beta))
        out[:, i, 1] = radius * np.sin(alpha) + 0.5 * ecc * radius * (np.sin(2 * alpha - beta) - 3 * np.sin(beta))
        out[:, i, 2] = -np.sqrt(3.0) * ecc * radius * np.cos(alpha - beta)
    return out


class EqualArmlengthOrbits(Orbits):
    """Analytic equal-arm cartwheel tabulated every `dt_base` seconds over `t_max` seconds.

    Args:
        armlength: arm length (m); every light travel time is armlength / c.
        t_max: span of the tables (s); 10 years by default.
        dt_base: spacing of the base tables (s); 1e4 s by default.
    """

    def __init__(self, armlength: float = 2.5e9, t_max: float = 10.0 * YRSID_SI,
                 dt_base: float = 1.0e4, **kwargs):
        t_base = np.arange(0.0, t_max + 0.5 * dt_base, dt_base)
        x = cartwheel_positions(t_base, armlength)
        ltt = np.full((len(t_base), 6), armlength / C_SI)
        super().__init__(t_base, x, ltt, None, armlength)
        self._dt_base = float(dt_base)

    def _evaluate(self, t: np.ndarray):
        x = cartwheel_positions(t, self.armlength)
        ltt = np.full((len(t), 6), self.armlength / C_SI)
        return x, ltt

    def configure(self, t_arr=None, dt=None, linear_interp_setup=False) -> None:
        if t_arr is None and dt is None:
            dt = self._dt_base
        super().configure(t_arr=t_arr, dt=dt, linear_interp_setup=linear_interp_setup)


class CartwheelOrbits(Orbits):
    """The analytic cartwheel of EqualArmlengthOrbits with unequal, time-varying light travel
    times: L_ij(t) = |x_i(t) - x_j(t - L_ij)| / c (receiver i, emitter j), solved by fixed-point
    iteration, so L_ij != L_ji and all six differ.

    Args:
        armlength: nominal arm length (m) of the constellation.
        t_max: span of the tables (s); 10 years by default.
        dt_base: spacing of the base tables (s); 1e4 s by default.
    """

    def __init__(self, armlength: float = 2.5e9, t_max: float = 10.0 * YRSID_SI,
                 dt_base: float = 1.0e4):
        t_base = np.arange(0.0, t_max + 0.5 * dt_base, dt_base)
        self.armlength = float(armlength)
        x, ltt = self._model(t_base)
        super().__init__(t_base, x, ltt, None, armlength)
        self._dt_base = float(dt_base)

    def _model(self, t: np.ndarray):
        x = cartwheel_positions(t, self.armlength)
        ltt = np.zeros((len(t), 6))
        for r, link in enumerate(self.LINKS):
            i, j = divmod(link, 10)
            L = np.linalg.norm(x[:, i - 1] - x[:, j - 1], axis=-1) / C_SI
            for _ in range(3):
                xe = cartwheel_positions(t - L, self.armlength)[:, j - 1]
                L = np.linalg.norm(x[:, i - 1] - xe, axis=-1) / C_SI
            ltt[:, r] = L
        return x, ltt

    def _evaluate(self, t: np.ndarray):
        return self._model(np.asarray(t, dtype=np.float64))

    def configure(self, t_arr=None, dt=None, linear_interp_setup=False) -> None:
        if t_arr is None and dt is None:
            dt = self._dt_base
        super().configure(t_arr=t_arr, dt=dt, linear_interp_setup=linear_interp_setup)


class StaticOrbits(Orbits):
    """A frozen constellation (constant positions and light travel times) tabulated every
    `dt_base` seconds; useful for closed-form checks of the response."""

    def __init__(self, positions, ltt, t_max: float = 1.0e7, dt_base: float = 1.0e4,
                 armlength: float = 2.5e9):
        positions = np.asarray(positions, dtype=np.float64).reshape(3, 3)
        ltt = np.broadcast_to(np.asarray(ltt, dtype=np.float64), (6,))
        t_base = np.arange(0.0, t_max + 0.5 * dt_base, dt_base)
        x = np.broadcast_to(positions, (len(t_base), 3, 3)).copy()
        L = np.broadcast_to(ltt, (len(t_base), 6)).copy()
        super().__init__(t_base, x, L, None, armlength)
        self._pos = positions
        self._L = np.array(ltt)
        self._dt_base = float(dt_base)

    def _evaluate(self, t: np.ndarray):
        return (np.broadcast_to(self._pos, (len(t), 3, 3)).copy(),
                np.broadcast_to(self._L, (len(t), 6)).copy())

    def configure(self, t_arr=None, dt=None, linear_interp_setup=False) -> None:
        if t_arr is None and dt is None:
            dt = self._dt_base
        super().configure(t_arr=t_arr, dt=dt, linear_interp_setup=linear_interp_setup)
