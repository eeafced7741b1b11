"""The C-ABI library loads and exports every entry point declared in include/*.h (no GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from ctl_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ctlproj_\w+)\s*\(", text):
            names.add(m.group(1))
    names.discard("ctlproj_progress_fn")
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_no_device_is_an_error_not_a_fallback():
    L = _lib.lib()
    n = ctypes.c_int(-1)
    assert L.ctlproj_device_count(ctypes.byref(n)) == _lib.OK
    if n.value == 0:
        h = ctypes.c_void_p()
        assert L.ctlproj_context_create(0, ctypes.byref(h)) == _lib.ERR_NO_DEVICE
        assert "no HIP device" in _lib.last_error()


def test_settings_defaults_match_reference():
    # raycasterprojector.h:31-34
    s = _lib.Settings()
    _lib.lib().ctlproj_settings_default(ctypes.byref(s))
    assert list(s.rays_per_pixel) == [1, 1]
    assert s.ray_sampling == pytest.approx(0.3)
    assert s.volume_upsampling == 1 and s.interpolate == 1
