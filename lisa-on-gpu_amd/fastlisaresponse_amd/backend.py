"""Python classes over the C ABI with the reference binding's names and methods.

`OrbitsWrap` replaces OrbitsWrap{CPU,GPU}_responselisa (binding_flr.hpp:32-75) and
`LISAResponseWrap` replaces LISAResponseWrap{CPU,GPU} (binding_flr.hpp:77-141), with the same
`get_response_wrap` / `get_tdi_delays_wrap` argument lists (binding_flr.cxx:16-58). Arrays may
be torch tensors (device or host) or numpy arrays; they are moved to the GPU for the call and
outputs are written back into the caller's array. The reference's length checks
(binding_flr.hpp:106-112) are applied and raise ValueError with its message.

Extra methods (`set_lagrange`, `set_tdi`, `projection`, `tdi`, `check_errors`) expose the
stream-ordered entry points that keep tables resident between calls.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import check, lib


def device() -> torch.device:
    """The HIP device the library runs on (torch's current device)."""
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU is visible: lisa-on-gpu_amd runs only on a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


_NP_OF = {torch.float64: np.float64, torch.complex128: np.complex128, torch.int32: np.int32}


def _length_error(name: str, expected: int, got: int) -> ValueError:
    return ValueError(f"{name}: input arrays have the incorrect length. Should be {expected}. "
                      f"It's length is {got}.")


def to_device(x, dtype: torch.dtype, name: str, n: int | None = None) -> torch.Tensor:
    """Contiguous device tensor of `dtype` holding x (no copy if it already is one)."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype != dtype:
            if t.is_complex() and not dtype.is_complex:
                raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
            t = t.to(dtype)
        if not t.is_cuda:
            t = t.to(device())
        t = t.contiguous()
    else:
        arr = np.asarray(x)
        if np.iscomplexobj(arr) and not dtype.is_complex:
            raise TypeError(f"{name}: expected real values")
        arr = np.ascontiguousarray(arr, dtype=_NP_OF[dtype])
        t = torch.from_numpy(arr).to(device())
    if n is not None and t.numel() != n:
        raise _length_error(name, n, t.numel())
    return t


def to_host(x, np_dtype, name: str, n: int | None = None) -> np.ndarray:
    """Contiguous numpy copy of a small configuration array."""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    arr = np.ascontiguousarray(np.asarray(x), dtype=np_dtype).reshape(-1)
    if n is not None and arr.size != n:
        raise _length_error(name, n, arr.size)
    return arr


class _Output:
    """Device view of an output array; copies the result back if a staging copy was needed."""

    def __init__(self, x, dtype: torch.dtype, name: str, n: int):
        self.orig = x
        direct = (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == dtype and x.is_contiguous())
        if direct:
            if x.numel() != n:
                raise _length_error(name, n, x.numel())
            self.t = x
        else:
            self.t = to_device(x, dtype, name, n)
        self.copy_back = not direct

    @property
    def ptr(self) -> int:
        return self.t.data_ptr()

    def finish(self) -> None:
        if not self.copy_back:
            return
        if isinstance(self.orig, torch.Tensor):
            self.orig.copy_(self.t.view(self.orig.shape))
        else:
            np.copyto(self.orig, self.t.cpu().numpy().reshape(np.shape(self.orig)))


class OrbitsWrap:
    """Device orbit tables: (dt, N, n_arr, ltt_arr, x_arr, links, sc_r, sc_e, armlength)."""

    def __init__(self, dt, N, n_arr, ltt_arr, x_arr, links, sc_r, sc_e, armlength):
        device()
        N = int(N)
        n_arr = to_host(n_arr, np.float64, "n_arr", N * 6 * 3)
        ltt_arr = to_host(ltt_arr, np.float64, "ltt_arr", N * 6)
        x_arr = to_host(x_arr, np.float64, "x_arr", N * 3 * 3)
        links = to_host(links, np.int32, "links", 6)
        sc_r = to_host(sc_r, np.int32, "sc_r", 6)
        sc_e = to_host(sc_e, np.int32, "sc_e", 6)
        h = ctypes.c_void_p()
        check(lib().flr_orbits_create(float(dt), N, n_arr.ctypes.data, ltt_arr.ctypes.data,
                                      x_arr.ctypes.data, links.ctypes.data, sc_r.ctypes.data,
                                      sc_e.ctypes.data, float(armlength), ctypes.byref(h)))
        self._h = h
        self.dt = float(dt)
        self.N = N
        self.links = [int(v) for v in links]
        self.armlength = float(armlength)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def extent(self):
        """(max |x| over the table in m, max light travel time in s)."""
        r, L = ctypes.c_double(), ctypes.c_double()
        check(lib().flr_orbits_extent(self._h, ctypes.byref(r), ctypes.byref(L)))
        return r.value, L.value

    def light_travel_time(self, t, link: int) -> torch.Tensor:
        """Device interpolation of the light travel time of `link` at times t."""
        tt = to_device(t, torch.float64, "t")
        out = torch.empty_like(tt)
        check(lib().flr_orbits_eval(self._h, tt.data_ptr(), tt.numel(), int(link), 1,
                                    out.data_ptr(), None, current_stream()))
        return out

    def position(self, t, sc: int) -> torch.Tensor:
        """Device interpolation of the position of spacecraft `sc` (1..3) at times t -> [n, 3]."""
        tt = to_device(t, torch.float64, "t")
        out = torch.empty((tt.numel(), 3), dtype=torch.float64, device=tt.device)
        check(lib().flr_orbits_eval(self._h, tt.data_ptr(), tt.numel(), self.links[0], int(sc),
                                    None, out.data_ptr(), current_stream()))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _capi._lib is not None:
            _capi._lib.flr_orbits_destroy(h)
            self._h = None


class LISAResponseWrap:
    """Response handle bound to an OrbitsWrap (kept alive by this object)."""

    def __init__(self, orbits: OrbitsWrap):
        if orbits is None:
            raise ValueError("Must add orbits with add_orbit_information method.")
        self.orbits = orbits
        h = ctypes.c_void_p()
        check(lib().flr_response_create(orbits.handle, ctypes.byref(h)))
        self._h = h

    # ----------------------------------------------------------- drop-in entry points
    def get_response_wrap(self, y_gw, t_data, k_in, u_in, v_in, dt, num_delays, input_in,
                          num_inputs, order, sampling_frequency, buffer_integer, A_in, deps,
                          num_A, E_in, projections_start_ind):
        """Projections y_gw[link * num_delays + i] (binding_flr.cxx:38-58); synchronous."""
        y = _Output(y_gw, torch.float64, "y_gw", int(num_delays) * 6)
        t = to_device(t_data, torch.float64, "t_data", int(num_delays))
        k = to_device(k_in, torch.float64, "k_in", 3)
        u = to_device(u_in, torch.float64, "u_in", 3)
        v = to_device(v_in, torch.float64, "v_in", 3)
        h = to_device(input_in, torch.complex128, "input_in")
        if h.numel() < int(num_inputs):
            raise _length_error("input_in", int(num_inputs), h.numel())
        A = to_device(A_in, torch.float64, "A_in", int(num_A))
        E = to_device(E_in, torch.float64, "E_in")
        check(lib().flr_get_response(self._h, y.ptr, t.data_ptr(), k.data_ptr(), u.data_ptr(),
                                     v.data_ptr(), float(dt), int(num_delays), h.data_ptr(),
                                     int(num_inputs), int(order), float(sampling_frequency),
                                     int(buffer_integer), A.data_ptr(), float(deps), int(num_A),
                                     E.data_ptr(), int(projections_start_ind), current_stream()))
        y.finish()

    def get_tdi_delays_wrap(self, delayed_links, input_links, num_inputs, num_delays, t_arr,
                            unit_starts, unit_lengths, tdi_base_link, tdi_link_combinations,
                            tdi_signs_in, channels, num_units, num_channels, order,
                            sampling_frequency, buffer_integer, A_in, deps, num_A, E_in,
                            tdi_start_ind):
        """TDI sums added into delayed_links[channel * num_delays + i] (binding_flr.cxx:16-36)."""
        out = _Output(delayed_links, torch.float64, "delayed_links", int(num_delays) * 3)
        y = to_device(input_links, torch.float64, "input_links", int(num_inputs) * 6)
        t = to_device(t_arr, torch.float64, "t_arr", int(num_delays))
        us = to_host(unit_starts, np.int32, "unit_starts", int(num_units))
        ul = to_host(unit_lengths, np.int32, "unit_lengths", int(num_units))
        bl = to_host(tdi_base_link, np.int32, "tdi_base_link")
        lc = to_host(tdi_link_combinations, np.int32, "tdi_link_combinations")
        sg = to_host(tdi_signs_in, np.float64, "tdi_signs_in")
        ch = to_host(channels, np.int32, "channels")
        if min(bl.size, sg.size, ch.size) < int(num_units):
            raise ValueError("tdi_base_link, tdi_signs_in and channels need num_units entries")
        if int(num_units) > 0 and lc.size < int((us + ul).max()):
            raise ValueError("tdi_link_combinations is shorter than unit_starts + unit_lengths")
        A = to_host(A_in, np.float64, "A_in", int(num_A))
        E = to_host(E_in, np.float64, "E_in")
        check(lib().flr_get_tdi_delays(self._h, out.ptr, y.data_ptr(), int(num_inputs),
                                       int(num_delays), t.data_ptr(), us.ctypes.data,
                                       ul.ctypes.data, bl.ctypes.data, lc.ctypes.data,
                                       sg.ctypes.data, ch.ctypes.data, int(num_units),
                                       int(num_channels), int(order), float(sampling_frequency),
                                       int(buffer_integer), A.ctypes.data, float(deps), int(num_A),
                                       E.ctypes.data, int(tdi_start_ind), current_stream()))
        out.finish()

    # ----------------------------------------------------------- resident tables
    def set_lagrange(self, order, A_in, deps, num_A, E_in) -> None:
        A = to_host(A_in, np.float64, "A_in", int(num_A))
        E = to_host(E_in, np.float64, "E_in")
        if E.size < (int(order) + 1) // 2 - 1:
            raise ValueError("E_in needs (order+1)/2 - 1 entries")
        check(lib().flr_set_lagrange(self._h, int(order), A.ctypes.data, float(deps), int(num_A),
                                     E.ctypes.data))

    def set_tdi(self, unit_starts, unit_lengths, tdi_base_link, tdi_link_combinations,
                tdi_signs_in, channels, num_units, num_channels=3) -> None:
        us = to_host(unit_starts, np.int32, "unit_starts", int(num_units))
        ul = to_host(unit_lengths, np.int32, "unit_lengths", int(num_units))
        bl = to_host(tdi_base_link, np.int32, "tdi_base_link", int(num_units))
        lc = to_host(tdi_link_combinations, np.int32, "tdi_link_combinations")
        sg = to_host(tdi_signs_in, np.float64, "tdi_signs_in", int(num_units))
        ch = to_host(channels, np.int32, "channels", int(num_units))
        if lc.size < int((us + ul).max()):
            raise ValueError("tdi_link_combinations is shorter than unit_starts + unit_lengths")
        check(lib().flr_set_tdi(self._h, us.ctypes.data, ul.ctypes.data, bl.ctypes.data,
                                lc.ctypes.data, sg.ctypes.data, ch.ctypes.data, int(num_units),
                                int(num_channels)))

    # ----------------------------------------------------------- stream-ordered runs
    def projection(self, y_gw: torch.Tensor, input_in: torch.Tensor, num_pts: int, t0: float,
                   dt: float, sampling_frequency: float, k, u, v, projections_start_ind: int,
                   flags: int, t_data: torch.Tensor | None = None, stream: int | None = None):
        """Enqueue the projection kernel (y_gw and input_in: device tensors)."""
        if y_gw.numel() != 6 * num_pts or input_in.numel() < num_pts:
            raise ValueError("y_gw must hold 6*num_pts values and input_in num_pts")
        k = to_host(k, np.float64, "k", 3)
        u = to_host(u, np.float64, "u", 3)
        v = to_host(v, np.float64, "v", 3)
        if t_data is None:
            flags |= _capi.FLR_FLAG_T_UNIFORM
        check(lib().flr_projection(self._h, y_gw.data_ptr(), input_in.data_ptr(), int(num_pts),
                                   None if t_data is None else t_data.data_ptr(), float(t0),
                                   float(dt), float(sampling_frequency), k.ctypes.data,
                                   u.ctypes.data, v.ctypes.data, int(projections_start_ind),
                                   int(flags), current_stream() if stream is None else stream))

    def tdi(self, out: torch.Tensor, y_gw: torch.Tensor, num_pts: int, t0: float, dt: float,
            sampling_frequency: float, tdi_start_ind: int, flags: int,
            t_arr: torch.Tensor | None = None, stream: int | None = None):
        """Enqueue the TDI kernel (out [3*num_pts] and y_gw [6*num_pts]: device tensors)."""
        if out.numel() != 3 * num_pts or y_gw.numel() != 6 * num_pts:
            raise ValueError("out must hold 3*num_pts values and y_gw 6*num_pts")
        if t_arr is None:
            flags |= _capi.FLR_FLAG_T_UNIFORM
        check(lib().flr_tdi(self._h, out.data_ptr(), y_gw.data_ptr(), int(num_pts),
                            None if t_arr is None else t_arr.data_ptr(), float(t0), float(dt),
                            float(sampling_frequency), int(tdi_start_ind), int(flags),
                            current_stream() if stream is None else stream))

    def check_errors(self, stream: int | None = None) -> None:
        """Synchronise and raise if a kernel latched an error since the last check."""
        check(lib().flr_check_errors(self._h, current_stream() if stream is None else stream))

    def set_exact(self, exact: bool) -> None:
        check(lib().flr_response_set_exact(self._h, 1 if exact else 0))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _capi._lib is not None:
            _capi._lib.flr_response_destroy(h)
            self._h = None


def backend_methods() -> dict:
    """The binding classes a backend of the reference package hands to its host code: the
    fields of FastLISAResponseBackendMethods (src/fastlisaresponse/cutils/__init__.py:13-17)
    except `xp`. The reference's host layer calls numpy-style methods on its arrays (`.astype`,
    `.flatten().copy()`, response.py:403-411,607), so a backend built from these classes takes
    xp=numpy: the classes accept host arrays, stage them to the device for each call and write
    the outputs back in place."""
    return dict(LISAResponseWrap=LISAResponseWrap, LISAResponseBase=LISAResponseWrap,
                OrbitsWrap=OrbitsWrap)
