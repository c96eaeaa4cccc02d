"""pyResponseTDI and ResponseWrapper: the fastlisaresponse API over the gfx950 kernels.

Constructor arguments, methods, attributes, warnings and errors follow
src/fastlisaresponse/response.py (pyResponseTDI :39-673, ResponseWrapper :676-863). Arrays
returned are torch tensors on the GPU. Differences, all deliberate:

* There is one backend, the HIP one. `force_backend` accepts None or a GPU name; "cpu" raises.
* Tables (Lagrange, TDI, orbits) are uploaded once at construction; each call only enqueues
  kernels on torch's current stream. Sample times t_i = t0 + i*dt are formed in-kernel with
  the same rounding as numpy's `t0 + arange(n)*dt`.
* For tdi_chan "AET"/"AE" the XYZ -> AET transform is fused into the TDI kernel; the `XYZ`
  property still returns X, Y, Z (recomputed on first access).
* `get_tdi_delays(y_gw=...)` works (the reference's check references a missing attribute).
* `exact=True` (the default) evaluates every stencil in the reference's operation order with
  correctly rounded divisions: outputs are bitwise equal to the FP64 CPU path. `exact=False`
  selects the reciprocal-free fast form, equal to ~1e-15 relative on A and E; the null channel
  T (thousands of times smaller than X, Y, Z for low-frequency sources) then differs at the
  1e-12..1e-11 relative level, since ulp-level differences in X, Y, Z do not cancel in it.
* `precision="mixed"` (order 25, fast stencils): get_projections stores the input as
  complex64 h / max|h|; get_tdi_delays runs projection and TDI as one fused launch (the six
  projections stay on chip) and returns float32 channels. Times, delays and the stencil
  arithmetic stay FP64 inside the kernel; the error is that of the FP32 storage of h (rel-L2
  ~1e-7..1e-6 on A, E at mHz frequencies, growing as f^-3 below). `y_gw` is then computed
  on access, in FP64 from the original input. `precision="mixed-q31"` stores h as 32-bit
  fixed point instead (same 8 bytes per sample, ~2^7 times smaller storage error: A, E ~3e-9
  and T ~1e-6 at 5 mHz); `precision="mixed-c128"` keeps the complex128 input itself (no
  storage error: the float32 rounding of the outputs is what remains).
"""
from __future__ import annotations

import warnings
from copy import deepcopy
from typing import Optional

import numpy as np
import torch

from . import _capi
from .backend import LISAResponseWrap, OrbitsWrap, device
from .orbits import EqualArmlengthOrbits, Orbits
from .tables import buffer_indices, lagrange_tables, sky_basis, stock_tdi, tdi_tables

YRSID_SI = 31558149.763545603
C_INV = 3.3356409519815204e-09

_GPU_BACKENDS = (None, "gpu", "cuda", "hip", "rocm", "cuda11x", "cuda12x", "cuda13x")


def _check_backend(force_backend):
    if force_backend not in _GPU_BACKENDS:
        raise ValueError(
            f"force_backend={force_backend!r}: this implementation has only the GPU (HIP gfx950) "
            "backend; use None or 'gpu'")


def _as_orbits(orbits) -> Orbits:
    if orbits is None:
        orbits = EqualArmlengthOrbits()
    elif isinstance(orbits, type) and issubclass(orbits, Orbits):
        orbits = orbits()
    if not isinstance(orbits, Orbits):
        raise TypeError("orbits must be an Orbits instance")
    return orbits


class _FusedBatches:
    """One reserved single-source fused batch per flag set, for pyResponseTDI's mixed
    precisions: reserved on first use, rebound per call. Holds the response handle so that it
    outlives the batches; destroys them after the device has finished with them."""

    def __init__(self, cpp_response):
        self.cpp_response = cpp_response
        self.batches = {}

    def get(self, flags, dt, fs):
        import ctypes
        key = (flags, dt, fs)
        hb = self.batches.get(key)
        if hb is None:
            hb = ctypes.c_void_p()
            _capi.check(_capi.lib().flr_fused_reserve(self.cpp_response._h, 1, dt, fs, flags,
                                                      ctypes.byref(hb)))
            self.batches[key] = hb
        return hb

    def __del__(self):
        try:
            if self.batches and _capi._lib is not None:
                if torch.cuda.is_initialized():
                    torch.cuda.synchronize()
                for hb in self.batches.values():
                    _capi.lib().flr_batch_destroy(hb)
                self.batches.clear()
        except Exception:
            pass


class pyResponseTDI:
    """Generic time-domain LISA response: link projections and TDI combinations.

    Args:
        sampling_frequency (double): sampling rate in Hz.
        num_pts (int): number of points of the output template.
        order (int): Lagrange interpolation order (default 25).
        tdi (str or list): "1st generation", "2nd generation", or a list of
            {"link": 12, "links_for_delay": [21, 13, 31], "sign": 1} terms of channel X.
        orbits (Orbits): orbit provider (default: EqualArmlengthOrbits()).
        tdi_orbits (Orbits): stored for interface parity; the kernels use `orbits`.
        tdi_chan (str): "XYZ", "AET" or "AE".
        force_backend: None or a GPU backend name.
        exact (bool): reference stencil operation order, bitwise FP64 CPU parity (default for
            precision "fp64"); False selects the faster reciprocal-free stencils.
        precision (str): "fp64" (default), "mixed" (complex64 input, float32 outputs, fused
            launch, order 25), "mixed-q31" (the same with 32-bit fixed-point input) or
            "mixed-c128" (the same reading the complex128 input).
        storage_check (str): "warn" (default), "raise" or "off": with "mixed" and "mixed-q31",
            get_projections estimates the input's storage error on A and E from its RMS
            frequency (batch.check_storage) and warns (StorageErrorWarning) or raises when it
            exceeds 1e-6.
    """

    def __init__(self, sampling_frequency, num_pts, order=25, tdi="1st generation",
                 orbits: Optional[Orbits] = None, tdi_orbits: Optional[Orbits] = None,
                 tdi_chan="XYZ", force_backend=None, exact: Optional[bool] = None,
                 precision: str = "fp64", storage_check: str = "warn"):
        _check_backend(force_backend)
        if precision not in ("fp64", "mixed", "mixed-q31", "mixed-c128"):
            raise ValueError("precision must be 'fp64', 'mixed', 'mixed-q31' or 'mixed-c128'")
        if exact is None:
            exact = precision == "fp64"
        if precision != "fp64" and exact:
            raise ValueError("exact stencils need precision='fp64'")
        if precision != "fp64" and order != 25:
            raise ValueError(f"precision={precision!r} is built for order 25")
        self.precision = precision
        if storage_check not in ("warn", "raise", "off"):
            raise ValueError("storage_check must be 'warn', 'raise' or 'off'")
        self.storage_check = storage_check
        self.sampling_frequency = sampling_frequency
        self.dt = 1 / sampling_frequency
        self.tdi_buffer = 200
        self.num_pts = num_pts
        self.order = order
        self.buffer_integer = self.order * 2 + 1
        self.half_order = int((order + 1) / 2)
        self.tdi = tdi
        self.tdi_chan = tdi_chan
        self.exact = bool(exact)
        self._cpp_response = None

        self._fill_A_E()
        self.response_orbits = orbits
        if tdi_orbits is None:
            tdi_orbits = self.response_orbits
        self.tdi_orbits = tdi_orbits

        if self.num_pts * self.dt > self.response_orbits.t_base.max():
            warnings.warn("Input number of points is longer in time than available orbital "
                          "information. Trimming to fit orbital information.")
            self.num_pts = int(self.response_orbits.t_base.max() / self.dt)

        self._init_TDI_delays()

        # host-side configuration is validated above; from here on the GPU is required
        self.device = device()
        self.cpp_orbits = OrbitsWrap(*self.response_orbits.pycppdetector_args)
        self.cpp_response = LISAResponseWrap(self.cpp_orbits)
        self.cpp_response.set_lagrange(self.order, self.A_in, self.deps, self.num_A, self.E_in)
        self.cpp_response.set_tdi(self.unit_starts, self.unit_lengths, self.tdi_base_links,
                                  self.tdi_link_combinations, self.tdi_signs, self.channels,
                                  self.num_units, 3)
        self.cpp_response.set_exact(self.exact)
        self.t0_projection = None
        self.y_gw_flat = None
        self.y_gw_length = None
        self.delayed_links_flat = None
        self._xyz_stale = False
        self._mixed = None  # precision="mixed": (h / scale as complex64, scale, k, u, v)
        self._fused = None  # mixed precisions: reserved one-source fused batches (_FusedBatches)

    # ------------------------------------------------------------ properties
    @property
    def cpp_response(self) -> LISAResponseWrap:
        if self._cpp_response is None:
            raise ValueError("Must add cpp_response and add orbit information.")
        return self._cpp_response

    @cpp_response.setter
    def cpp_response(self, value) -> None:
        self._cpp_response = value

    @property
    def response_gen(self):
        """Drop-in projection entry point (argument list of the reference binding)."""
        return self.cpp_response.get_response_wrap

    @property
    def tdi_gen(self):
        """Drop-in TDI entry point (argument list of the reference binding)."""
        return self.cpp_response.get_tdi_delays_wrap

    @property
    def xp(self):
        return torch

    @property
    def response_orbits(self) -> Orbits:
        return self._response_orbits

    @response_orbits.setter
    def response_orbits(self, orbits) -> None:
        self._response_orbits = deepcopy(_as_orbits(orbits))
        if not self._response_orbits.configured:
            self._response_orbits.configure(linear_interp_setup=True)

    @property
    def tdi_orbits(self) -> Orbits:
        return self._tdi_orbits

    @tdi_orbits.setter
    def tdi_orbits(self, orbits) -> None:
        self._tdi_orbits = deepcopy(_as_orbits(orbits))
        if not self._tdi_orbits.configured:
            self._tdi_orbits.configure(linear_interp_setup=True)

    @property
    def citation(self):
        return "Katz et al. 2022, arXiv:2204.06633 (fastlisaresponse)"

    # ------------------------------------------------------------ tables
    def _fill_A_E(self):
        """Lagrange tables A_in, E_in (response.py:273-310)."""
        A, deps, E = lagrange_tables(self.order)
        self.num_A = len(A)
        self.deps = deps
        self.A_in = A
        self.E_in = E

    def _init_TDI_delays(self):
        if self.tdi in ["1st generation", "2nd generation"]:
            combos = stock_tdi(self.tdi)
        elif isinstance(self.tdi, list):
            combos = self.tdi
        else:
            raise ValueError("tdi kwarg should be '1st generation', '2nd generation', or a list "
                             "with a specific tdi combination.")
        self.tdi_combinations = combos

    @property
    def tdi_combinations(self):
        return self._tdi_combinations

    @tdi_combinations.setter
    def tdi_combinations(self, combos) -> None:
        t = tdi_tables(combos)
        self._tdi_combinations = combos
        self.tdi_base_links = t["tdi_base_links"]
        self.tdi_link_combinations = t["tdi_link_combinations"]
        self.tdi_signs = t["tdi_signs"]
        self.channels = t["channels"]
        self.tdi_operation_index = t["tdi_operation_index"]
        self.unit_starts = t["unit_starts"]
        self.unit_lengths = t["unit_lengths"]
        self.num_units = t["num_units"]
        if getattr(self, "_cpp_response", None) is not None:
            self.cpp_response.set_tdi(self.unit_starts, self.unit_lengths, self.tdi_base_links,
                                      self.tdi_link_combinations, self.tdi_signs, self.channels,
                                      self.num_units, 3)

    def _cyclic_permutation(self, link, permutation):
        from .tables import cyclic_permutation
        return cyclic_permutation(link, permutation)

    # ------------------------------------------------------------ projections
    @property
    def y_gw(self) -> torch.Tensor:
        """Projections along the arms, [6, num_pts], rows in orbits.LINKS order."""
        if self.y_gw_flat is None and self._mixed is not None:
            self._project(self._input_keepalive, self._mixed[2:], self.t0_projection)
        return self.y_gw_flat.reshape(self.nlinks, -1)

    def _project(self, h: torch.Tensor, basis, t0: float) -> None:
        k, u, v = basis
        y_gw = torch.empty(self.nlinks * self.num_pts, dtype=torch.float64, device=self.device)
        flags = _capi.FLR_FLAG_ZERO_EDGES | (_capi.FLR_FLAG_EXACT if self.exact else 0)
        self.cpp_response.projection(y_gw, h, self.num_pts, float(t0), self.dt,
                                     self.sampling_frequency, k, u, v,
                                     self.projections_start_ind, flags)
        self.y_gw_flat = y_gw
        self.y_gw_length = self.num_pts

    def _data_time_check(self, t0: float, n: int, input_in):
        t_last = t0 + (n - 1) * self.dt
        if t_last > self.response_orbits.t.max():
            warnings.warn("Input waveform is longer than available orbital information. "
                          "Trimming to fit orbital information.")
            t_data = t0 + np.arange(n) * self.dt
            max_ind = np.where(t_data <= self.response_orbits.t.max())[0][-1]
            input_in = input_in[:max_ind]
        return input_in

    def get_projections(self, input_in, lam, beta, t0=0.0, t_buffer=10000.0):
        """Project h+ + i hx onto the six links.

        Args:
            input_in: complex time series h+ + i hx (torch tensor or numpy array).
            lam, beta: ecliptic longitude and latitude (rad).
            t0: time of the first sample (s).
            t_buffer: length (s) of the start/end region left out of TDI.

        Raises:
            ValueError: if t_buffer is too short for the delays.
        """
        key = (t_buffer, id(self.response_orbits.x_base))
        if getattr(self, "_buffer_key", None) != key:
            self._buffers = buffer_indices(self.sampling_frequency, self.order, t_buffer,
                                           self.response_orbits.x_base)
            self._buffer_key = key
        self.tdi_start_ind, self.check_tdi_buffer, self.projection_buffer, \
            self.projections_start_ind = self._buffers
        self.t0_projection = t0
        if self.projections_start_ind < self.projection_buffer:
            raise ValueError("Need to increase t_buffer. The initial buffer is not large enough.")

        k, u, v = sky_basis(lam, beta)
        self.nlinks = 6
        h = input_in if isinstance(input_in, torch.Tensor) else torch.as_tensor(np.asarray(input_in))
        self.num_total_points = len(h)
        h = self._data_time_check(t0, len(h), h)
        if len(h) < self.num_pts:
            raise AssertionError("input_in is shorter than num_pts")
        if len(h) != self.num_pts:
            raise ValueError(f"y_gw: input arrays have the incorrect length. Should be "
                             f"{6 * len(h)}. It's length is {6 * self.num_pts}.")
        h = h.to(device=self.device, dtype=torch.complex128).contiguous()
        self._input_keepalive = h
        if self.precision != "fp64":
            from .batch import check_storage, pack_mixed, pack_q31
            if self.precision == "mixed-c128":
                stored, scale = h, 1.0
            else:
                check_storage(h, [len(h)], self.dt, self.precision,
                              self.response_orbits.armlength, action=self.storage_check)
                pack = pack_q31 if self.precision == "mixed-q31" else pack_mixed
                stored, scales = pack(h, [len(h)])
                scale = scales[0]
            self._mixed = (stored, scale, k, u, v)
            self.y_gw_flat = None
            self.y_gw_length = self.num_pts
            return
        self._project(h, (k, u, v), t0)

    # ------------------------------------------------------------ TDI
    @property
    def XYZ(self) -> torch.Tensor:
        """X, Y, Z of the last get_tdi_delays call, [3, num_pts]."""
        if self._xyz_stale:
            self._run_tdi(self._tdi_t0, aet=False)
        return self.delayed_links_flat.reshape(3, -1)

    def _run_fused(self, t0: float, aet: bool) -> torch.Tensor:
        """Mixed precisions: projection + TDI in one launch from the stored input."""
        import ctypes
        h32, scale, k, u, v = self._mixed
        out = torch.empty(3 * self.num_pts, dtype=torch.float32, device=self.device)
        srcs = (_capi.FlrSource * 1)()
        q = srcs[0]
        q.offset = 0
        q.num_pts = self.num_pts
        q.projections_start_ind = self.projections_start_ind
        q.tdi_start_ind = self.tdi_start_ind
        q.t0 = float(t0)
        q.scale = scale
        for c in range(3):
            q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
        flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE | _capi.FLR_FLAG_MIXED
        if self.precision == "mixed-q31":
            flags |= _capi.FLR_FLAG_Q31
        elif self.precision == "mixed-c128":
            flags |= _capi.FLR_FLAG_C128
        if aet:
            flags |= _capi.FLR_FLAG_AET
        from .backend import current_stream
        if self._fused is None:
            self._fused = _FusedBatches(self.cpp_response)
        hb = self._fused.get(flags, self.dt, self.sampling_frequency)
        stream = current_stream()
        # bind (asynchronous table upload) and run on the caller's stream: no allocation and no
        # synchronisation per call; window errors are latched for check_errors
        _capi.check(_capi.lib().flr_fused_bind(hb, srcs, 1, ctypes.c_void_p(h32.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), stream))
        _capi.check(_capi.lib().flr_batch_run(hb, stream))
        if not aet:
            self.delayed_links_flat = out
            self._xyz_stale = False
        return out

    def _run_tdi(self, t0: float, aet: bool) -> torch.Tensor:
        if self._mixed is not None:
            return self._run_fused(t0, aet)
        out = torch.empty(3 * self.y_gw_length, dtype=torch.float64, device=self.device)
        flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_OVERWRITE
        if aet:
            flags |= _capi.FLR_FLAG_AET
        if self.exact:
            flags |= _capi.FLR_FLAG_EXACT
        self.cpp_response.tdi(out, self.y_gw_flat, self.y_gw_length, float(t0), self.dt,
                              self.sampling_frequency, self.tdi_start_ind, flags)
        if not aet:
            self.delayed_links_flat = out
            self._xyz_stale = False
        return out

    def get_tdi_delays(self, t0=None, y_gw=None):
        """TDI combinations from the projections: (X, Y, Z), (A, E, T) or (A, E).

        Args:
            t0: start time; only when projections were not computed here (must then match).
            y_gw: projections [6, num_pts] to use instead of get_projections' result.
        """
        if t0 is None:
            t0 = self.t0_projection
        elif self.t0_projection is not None and self.t0_projection != t0:
            raise ValueError("If running projections with a t0 value, that t0 value needs to be "
                             "the same for TDI. If you want to force otherwise, need to run "
                             "'del class.t0_projections' before TDI computation.")
        if y_gw is not None:
            y = y_gw if isinstance(y_gw, torch.Tensor) else torch.as_tensor(np.asarray(y_gw))
            if tuple(y.shape) != (6, self.num_pts):
                raise ValueError(f"y_gw must have shape (6, {self.num_pts})")
            self.y_gw_flat = y.to(device=self.device, dtype=torch.float64).reshape(-1).clone()
            self.y_gw_length = self.num_pts
            self._mixed = None  # given projections: FP64 TDI from them
            if not hasattr(self, "tdi_start_ind"):
                raise ValueError("get_projections sets tdi_start_ind; run it first or set "
                                 "tdi_start_ind = int(t_buffer / dt)")
        elif self.y_gw_flat is None and self._mixed is None:
            raise ValueError("Need to either enter projection array or have this code determine "
                             "projections.")
        if t0 is None:
            t0 = 0.0
        self._tdi_t0 = t0
        if self.tdi_chan == "XYZ":
            X, Y, Z = self._run_tdi(t0, aet=False).reshape(3, -1)
            return X, Y, Z
        if self.tdi_chan in ("AET", "AE"):
            out = self._run_tdi(t0, aet=True).reshape(3, -1)
            self._xyz_stale = True
            self.delayed_links_flat = None
            A, E, T = out
            return (A, E, T) if self.tdi_chan == "AET" else (A, E)
        raise ValueError("tdi_chan must be 'XYZ', 'AET' or 'AE'.")

    def check_errors(self) -> None:
        """Synchronise and raise if a kernel found a stencil outside its window."""
        self.cpp_response.check_errors()


class ResponseWrapper:
    """Waveform generator -> TDI channels, as the reference's ResponseWrapper.

    Args follow response.py:730-747: waveform_gen, Tobs (yr), dt (s), index_lambda,
    index_beta, t0, t_buffer, flip_hx, remove_sky_coords, is_ecliptic_latitude,
    force_backend, remove_garbage (True, False or "zero"), n_overide, orbits, and **kwargs for
    pyResponseTDI (order, tdi, tdi_chan, tdi_orbits, exact).
    """

    def __init__(self, waveform_gen, Tobs, dt, index_lambda, index_beta, t0=0.0, t_buffer=10000.0,
                 flip_hx=False, remove_sky_coords=False, is_ecliptic_latitude=True,
                 force_backend=None, remove_garbage=True, n_overide=None,
                 orbits: Optional[Orbits] = None, **kwargs):
        _check_backend(force_backend)
        self.waveform_gen = waveform_gen
        self.index_lambda = index_lambda
        self.index_beta = index_beta
        self.dt = dt
        self.t0 = t0
        self.t_buffer = t_buffer
        self.sampling_frequency = 1.0 / dt
        orbits = _as_orbits(orbits)
        if Tobs * YRSID_SI > orbits.t_base.max():
            warnings.warn("Tobs is larger than available orbital information time array. "
                          f"Reducing Tobs to {orbits.t_base.max()}")
            Tobs = orbits.t_base.max() / YRSID_SI
        if n_overide is not None:
            if not isinstance(n_overide, int):
                raise ValueError("n_overide must be an integer if not None.")
            self.n = n_overide
        else:
            self.n = int(Tobs * YRSID_SI / dt)
        self.Tobs = self.n * dt
        self.is_ecliptic_latitude = is_ecliptic_latitude
        self.remove_sky_coords = remove_sky_coords
        self.flip_hx = flip_hx
        self.remove_garbage = remove_garbage
        self.response_model = pyResponseTDI(self.sampling_frequency, self.n, orbits=orbits,
                                            force_backend=force_backend, **kwargs)
        self.Tobs = (self.n * self.response_model.dt) / YRSID_SI

    @property
    def xp(self):
        return torch

    @property
    def citation(self):
        return self.response_model.citation

    def __call__(self, *args, **kwargs):
        args = list(args)
        beta = args[self.index_beta]
        lam = args[self.index_lambda]
        if self.remove_sky_coords:
            args.pop(self.index_beta)
            args.pop(self.index_lambda)
        if not self.is_ecliptic_latitude:
            beta = np.pi / 2.0 - beta
        kwargs["T"] = self.Tobs
        kwargs["dt"] = self.dt
        h = self.waveform_gen(*args, **kwargs)
        if self.flip_hx:
            h = h.real - 1j * h.imag
        self.response_model.get_projections(h, lam, beta, t0=self.t0, t_buffer=self.t_buffer)
        tdi_out = self.response_model.get_tdi_delays()
        out = list(tdi_out)
        T = self.response_model.tdi_start_ind
        if self.remove_garbage is True:
            for i in range(len(out)):
                out[i] = out[i][T:-T]
        elif isinstance(self.remove_garbage, str):
            if self.remove_garbage != "zero":
                raise ValueError("remove_garbage must be True, False, or 'zero'.")
            for i in range(len(out)):
                out[i][:T] = 0.0
                out[i][-T:] = 0.0
        return out
