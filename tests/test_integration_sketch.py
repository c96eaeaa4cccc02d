"""INTEGRATION.md's backend sketch builds (CPU, no GPU needed).

The reference's backend machinery (gpubackendtools, src/fastlisaresponse/cutils/__init__.py)
is not installed here, so the sketch runs against test doubles with the same contract: a
methods dataclass carrying the three binding classes plus the array module `xp`, and a backend
base whose constructor asserts that type and copies the classes (cutils/__init__.py:13-35).
"""
import dataclasses
import inspect
import os
import re

import numpy as np

from conftest import ROOT


@dataclasses.dataclass
class BackendMethods:  # test double: gpubackendtools' base carries the array module
    xp: object


@dataclasses.dataclass
class FastLISAResponseBackendMethods(BackendMethods):  # test double of cutils/__init__.py:13-17
    LISAResponseWrap: object
    LISAResponseBase: object
    OrbitsWrap: object


class FastLISAResponseBackend:  # test double of cutils/__init__.py:20-35
    def __init__(self, methods):
        assert isinstance(methods, FastLISAResponseBackendMethods)
        self.LISAResponseWrap = methods.LISAResponseWrap
        self.LISAResponseBase = methods.LISAResponseBase
        self.OrbitsWrap = methods.OrbitsWrap
        self.xp = methods.xp


class CpuBackend:  # test double: nothing the sketch relies on
    def __init__(self, *args, **kwargs):
        pass


def sketch_source() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```python\n(# src/fastlisaresponse/cutils/__init__.py.*?)```", text, re.S)
    assert m, "INTEGRATION.md lost its backend sketch"
    return m.group(1)


def test_sketch_builds_a_backend():
    ns = dict(CpuBackend=CpuBackend, FastLISAResponseBackend=FastLISAResponseBackend,
              FastLISAResponseBackendMethods=FastLISAResponseBackendMethods)
    exec(compile(sketch_source(), "INTEGRATION.md", "exec"), ns)
    backend = ns["FastLISAResponseRocmBackend"]()
    from fastlisaresponse_amd.backend import LISAResponseWrap, OrbitsWrap
    assert backend.xp is np
    assert backend.OrbitsWrap is OrbitsWrap and backend.LISAResponseWrap is LISAResponseWrap
    assert backend.LISAResponseBase is LISAResponseWrap


def test_binding_methods_take_the_reference_argument_lists():
    """get_response_wrap / get_tdi_delays_wrap keep the argument order of
    binding_flr.cxx:16-58 (the reference's host code passes them positionally)."""
    from fastlisaresponse_amd.backend import LISAResponseWrap
    resp = ["y_gw", "t_data", "k_in", "u_in", "v_in", "dt", "num_delays", "input_in",
            "num_inputs", "order", "sampling_frequency", "buffer_integer", "A_in", "deps",
            "num_A", "E_in", "projections_start_ind"]
    tdi = ["delayed_links", "input_links", "num_inputs", "num_delays", "t_arr", "unit_starts",
           "unit_lengths", "tdi_base_link", "tdi_link_combinations", "tdi_signs_in", "channels",
           "num_units", "num_channels", "order", "sampling_frequency", "buffer_integer", "A_in",
           "deps", "num_A", "E_in", "tdi_start_ind"]
    got_r = list(inspect.signature(LISAResponseWrap.get_response_wrap).parameters)[1:]
    got_t = list(inspect.signature(LISAResponseWrap.get_tdi_delays_wrap).parameters)[1:]
    assert got_r == resp and got_t == tdi


def test_numpy_covers_the_reference_array_calls():
    """The array operations the reference's host layer applies through xp (response.py:
    298-310, 403-411, 532, 537, 592-617) on the types it applies them to."""
    xp = np
    a = xp.asarray([1, 2, 3]).astype(xp.int32)
    assert a.dtype == np.int32
    y = xp.zeros((6 * 4,), dtype=xp.float64)
    assert y.reshape(6, -1).flatten().copy().shape == (24,)
    d = xp.zeros((3, 4), dtype=xp.float64).flatten()
    assert d.shape == (12,) and d.reshape(3, -1).shape == (3, 4)
    t = 5.0 + xp.arange(4) * 10.0
    assert t[-1] == 35.0
