"""The storage-error guard of the 8-byte input formats (CPU: host tensors, no GPU needed).

The model K (2 pi f L/c)^-3 is checked at its calibration point and for its scaling; the RMS
frequency estimator against sinusoids and a galactic binary of known frequency; and the guard
must warn for a 0.1 mHz source and stay silent for a 5 mHz one (complex64 and q31), never warn
for the formats without storage error, and raise on request.
"""
import warnings

import numpy as np
import pytest
import torch

from conftest import gb_waveform

from fastlisaresponse_amd.batch import (StorageErrorWarning, check_storage, rms_frequencies,
                                        storage_error_estimate)


def test_model_calibration_and_scaling():
    # complex64 at 5 mHz: the 1-yr source's measured A, E error (2.5e-7)
    assert storage_error_estimate(5e-3, "mixed") == pytest.approx(2.5e-7, rel=0.05)
    assert storage_error_estimate(5e-3, "mixed-q31") == pytest.approx(2.8e-9, rel=0.05)
    # (wL)^-3: ten times lower frequency, a thousand times the error
    assert storage_error_estimate(5e-4, "mixed") / storage_error_estimate(5e-3, "mixed") == \
        pytest.approx(1e3, rel=1e-9)
    assert storage_error_estimate(5e-3, "mixed-c128") == 0.0
    assert storage_error_estimate(5e-3, "fp64") == 0.0
    assert storage_error_estimate(0.0, "mixed") == 0.0


@pytest.mark.parametrize("f", [1e-4, 1e-3, 5e-3, 2e-2])
def test_rms_frequency_of_a_binary(f):
    n, dt = 200_000, 10.0
    h = torch.as_tensor(gb_waveform(n, dt, 1e-21, f, 1e-18, 0.7, 0.2, 0.4))
    got = rms_frequencies(h, [n], dt)[0]
    assert got == pytest.approx(f, rel=2e-2)


def test_rms_frequency_per_source_of_a_packed_series():
    dt = 5.0
    lengths = [5000, 12000, 3000]
    fs = [2e-3, 7e-3, 4e-4]
    t = [np.arange(n) * dt for n in lengths]
    h = torch.as_tensor(np.concatenate([np.exp(2j * np.pi * f * tt) for f, tt in zip(fs, t)]))
    got = rms_frequencies(h, lengths, dt)
    np.testing.assert_allclose(got, fs, rtol=1e-9)  # exact for pure tones
    # the same series as complex64 and as q31 pairs
    got32 = rms_frequencies(h.to(torch.complex64), lengths, dt)
    np.testing.assert_allclose(got32, fs, rtol=1e-4)
    q = torch.round(torch.view_as_real(h).reshape(-1) * (2 ** 31 - 1)).to(torch.int32)
    np.testing.assert_allclose(rms_frequencies(q, lengths, dt), fs, rtol=1e-4)
    # silence is not flagged
    assert rms_frequencies(torch.zeros(100, dtype=torch.complex128), [100], dt)[0] == 0.0


@pytest.mark.parametrize("precision", ["mixed", "mixed-q31"])
def test_guard_warns_low_frequency_only(precision):
    n, dt = 100_000, 10.0
    low = torch.as_tensor(gb_waveform(n, dt, 1e-21, 1e-4, 1e-19, 0.8, 0.3, 0.5))
    high = torch.as_tensor(gb_waveform(n, dt, 1e-21, 5e-3, 1e-16, 0.8, 0.3, 0.5))
    with pytest.warns(StorageErrorWarning, match="mixed-c128"):
        est = check_storage(low, [n], dt, precision)
    assert est[0] > 1e-6
    with warnings.catch_warnings():
        warnings.simplefilter("error", StorageErrorWarning)
        est = check_storage(high, [n], dt, precision)
    assert 0.0 < est[0] < 1e-6
    # both in one packed batch: one warning naming one source of two
    with pytest.warns(StorageErrorWarning, match="1 of 2"):
        check_storage(torch.cat([high, low]), [n, n], dt, precision)
    with pytest.raises(ValueError, match="storage error"):
        check_storage(low, [n], dt, precision, action="raise")
    with warnings.catch_warnings():
        warnings.simplefilter("error", StorageErrorWarning)
        check_storage(low, [n], dt, precision, action="off")
        check_storage(low, [n], dt, "mixed-c128")
        check_storage(low, [n], dt, "fp64")
