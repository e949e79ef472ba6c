"""kmc_probe_peaks: the measured roofline denominators are positive and below the
published MI355X peaks (HBM 8 TB/s, FP64 vector 78.6 TFLOP/s) with a small margin."""

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_probe_peaks_within_published_peaks():
    import torch
    from coulomb_kmc_amd import _native
    out = (ctypes.c_double * 2)()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _native.check(_native.lib().kmc_probe_peaks(st, 1 << 30, 3, out), "kmc_probe_peaks")
    copy_gbs, fp64_tflops = out[0], out[1]
    assert 100.0 < copy_gbs < 8000.0 * 1.05
    assert 1.0 < fp64_tflops < 78.6 * 1.05
