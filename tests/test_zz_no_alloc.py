"""flr_fused_run_static, the allocation-free fused run, through the C ABI on the GPU.

It must give the bits of the allocating form (flr_fused_create + flr_batch_run) on the same
inputs, on consecutive calls that reuse its fixed device tables (the second one smaller), for
the mixed-precision equal-arm form (composite filters in the tables), the unequal-arm mixed form
and the exact FP64 form; and it must refuse more sources than its tables hold.
"""
import ctypes

import numpy as np
import pytest

from conftest import YRSID_SI, gb_waveform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module", params=["equal", "unequal"])
def orbits(request):
    from fastlisaresponse_amd.orbits import CartwheelOrbits, EqualArmlengthOrbits
    cls = EqualArmlengthOrbits if request.param == "equal" else CartwheelOrbits
    o = cls(t_max=1.2 * YRSID_SI)
    o.configure(linear_interp_setup=True)
    return o


def _sources(rng, lengths):
    from fastlisaresponse_amd import _capi
    from fastlisaresponse_amd.tables import sky_basis
    hs = []
    srcs = (_capi.FlrSource * len(lengths))()
    o = 0
    for s, n in enumerate(lengths):
        f = 10 ** rng.uniform(-3.5, -2)
        t0 = float(rng.choice([0.0, 1.0e6, 2.5e7]))
        hs.append(gb_waveform(n, 10.0, 10 ** rng.uniform(-23, -21), f, 1e-17,
                              np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 6), rng.uniform(0, 6),
                              t0=t0))
        k, u, v = sky_basis(rng.uniform(0, 2 * np.pi), np.arcsin(rng.uniform(-1, 1)))
        q = srcs[s]
        q.offset, q.num_pts = o, n
        q.projections_start_ind, q.tdi_start_ind = 780, 1000
        q.t0, q.scale = t0, 1.0
        for c in range(3):
            q.k[c], q.u[c], q.v[c] = k[c], u[c], v[c]
        o += n
    return np.concatenate(hs), srcs


def _allocating(lib, check, resp, srcs, h, out, flags, tile=0, grid=0):
    hb = ctypes.c_void_p()
    check(lib.flr_fused_create(resp, srcs, len(srcs), ctypes.c_void_p(h.data_ptr()),
                               ctypes.c_void_p(out.data_ptr()), 10.0, 0.1, flags, ctypes.byref(hb)))
    try:
        if tile or grid:
            check(lib.flr_batch_set_launch(hb, tile, grid))
        check(lib.flr_batch_run(hb, None))
    finally:
        import torch
        torch.cuda.synchronize()
        lib.flr_batch_destroy(hb)


@pytest.mark.parametrize("form", ["mixed-c128", "fp64-exact"])
def test_static_run_matches_allocating_form(torch_gpu, orbits, form):
    torch = torch_gpu
    from fastlisaresponse_amd import _capi, pyResponseTDI
    from fastlisaresponse_amd._capi import check, lib
    L = lib()
    model = pyResponseTDI(0.1, 1, orbits=orbits, tdi="2nd generation", tdi_chan="AET")
    resp = model.cpp_response._h
    flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_AET
    if form == "mixed-c128":
        flags |= _capi.FLR_FLAG_MIXED | _capi.FLR_FLAG_C128
        odt = torch.float32
    else:
        flags |= _capi.FLR_FLAG_EXACT
        odt = torch.float64
    rng = np.random.default_rng(5)
    # two calls in a row, the second smaller (fewer and shorter sources)
    for lengths in ([60000, 9000, 25000, 4100, 33333], [20000, 7000]):
        h_host, srcs = _sources(rng, lengths)
        h = torch.as_tensor(h_host, device="cuda")
        total = sum(lengths)
        ref = torch.full((3 * total,), float("nan"), dtype=odt, device="cuda")
        _allocating(L, check, resp, srcs, h, ref, flags)
        got = torch.full((3 * total,), float("nan"), dtype=odt, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        check(L.flr_fused_run_static(resp, srcs, len(lengths), ctypes.c_void_p(h.data_ptr()),
                                     ctypes.c_void_p(got.data_ptr()), 10.0, 0.1, flags, 0, 0,
                                     ctypes.c_void_p(stream)))
        model.cpp_response.check_errors()
        assert not bool(torch.isnan(got).any())
        assert torch.equal(got, ref), (form, lengths)


def test_static_run_launch_shapes_and_limits(torch_gpu, orbits):
    torch = torch_gpu
    from fastlisaresponse_amd import _capi, pyResponseTDI
    from fastlisaresponse_amd._capi import check, lib
    L = lib()
    model = pyResponseTDI(0.1, 1, orbits=orbits, tdi="1st generation", tdi_chan="XYZ")
    resp = model.cpp_response._h
    flags = _capi.FLR_FLAG_ZERO_EDGES | _capi.FLR_FLAG_MIXED | _capi.FLR_FLAG_C128
    h_host, srcs = _sources(np.random.default_rng(9), [40000, 12000])
    h = torch.as_tensor(h_host, device="cuda")
    total = 52000
    ref = torch.empty(3 * total, dtype=torch.float32, device="cuda")
    _allocating(L, check, resp, srcs, h, ref, flags)
    for tile, grid in ((256, 8), (1024, 32), (512, 4)):
        got = torch.full((3 * total,), float("nan"), dtype=torch.float32, device="cuda")
        check(L.flr_fused_run_static(resp, srcs, 2, ctypes.c_void_p(h.data_ptr()),
                                     ctypes.c_void_p(got.data_ptr()), 10.0, 0.1, flags, tile, grid,
                                     None))
        torch.cuda.synchronize()
        # the allocating form at the same launch shape: the same bits
        same = torch.empty_like(ref)
        _allocating(L, check, resp, srcs, h, same, flags, tile, grid)
        assert torch.equal(got, same), (tile, grid)
        if orbits.ltt.min() == orbits.ltt.max():
            assert torch.equal(got, ref), (tile, grid)  # equal arms: independent of the shape
        else:
            # unequal arms: composite filters per tile, so shapes agree within float32 rounding
            rel = ((got.double() - ref.double()).norm() / ref.double().norm()).item()
            assert rel <= 1e-7, (tile, grid, rel)
    model.cpp_response.check_errors()
    # more sources than the fixed tables hold: refused before any device work
    many = (_capi.FlrSource * 513)()
    for s in range(513):
        many[s].offset, many[s].num_pts = 0, 10
    st = L.flr_fused_run_static(resp, many, 513, ctypes.c_void_p(h.data_ptr()),
                                ctypes.c_void_p(ref.data_ptr()), 10.0, 0.1, flags, 0, 0, None)
    assert st == _capi.FLR_ERR_UNSUPPORTED
    assert "512" in L.flr_last_error().decode()
    assert L.flr_fused_run_static(resp, srcs, 2, ctypes.c_void_p(h.data_ptr()),
                                  ctypes.c_void_p(ref.data_ptr()), 10.0, 0.1, flags, 768, 0,
                                  None) == _capi.FLR_ERR_UNSUPPORTED
