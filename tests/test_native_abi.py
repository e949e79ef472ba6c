"""The C-ABI library loads and exports every entry point include/kmc_amd.h declares
(no compute calls: runs without a GPU)."""

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "kmc_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kmc_[a-z_0-9]+)\s*\(", txt)))


def test_header_symbols_exported():
    from coulomb_kmc_amd import _native
    lib = _native.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.EXPORTS)


def test_order_range():
    from coulomb_kmc_amd import _native
    lib = _native.lib()
    assert lib.kmc_min_order() == 2
    assert lib.kmc_max_order() == 20


def test_missing_library_raises(tmp_path, monkeypatch):
    import pytest
    from coulomb_kmc_amd import _native
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        _native.lib()


def test_probe_peaks_rejects_bad_arguments():
    # the argument checks run before any device call
    import ctypes
    from coulomb_kmc_amd import _native
    lib = _native.lib()
    out = (ctypes.c_double * 2)()
    assert lib.kmc_probe_peaks(None, 1024, 1, out) != 0          # below 1 MiB
    assert lib.kmc_probe_peaks(None, 1 << 35, 1, out) != 0       # above 16 GiB
    assert lib.kmc_probe_peaks(None, 1 << 20, 0, out) != 0       # no repetitions
    assert lib.kmc_probe_peaks(None, 1 << 20, 1, None) != 0      # no output
