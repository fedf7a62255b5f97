"""The C-ABI library loads and exports every entry point include/foilbench_amd.h declares
(no compute calls: this runs without a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _declared() -> set[str]:
    text = (ROOT / "include" / "foilbench_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char\*)\s+(fb_\w+)\s*\(", text, flags=re.M))


def test_library_exports_every_declared_symbol():
    from foilbench_amd import _native as N

    assert N.LIB_PATH.exists(), "build the library first: python -m foilbench_amd.build"
    lib = ctypes.CDLL(str(N.LIB_PATH))
    declared = _declared()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes layer binds exactly the declared entry points
    assert set(N.exported_symbols()) == declared


def test_load_without_gpu_fails_loudly():
    import pytest

    from foilbench_amd import _native as N

    lib = N.load(require_gpu=False)
    assert lib.fb_version() == 1
    n = ctypes.c_int(-1)
    rc = lib.fb_device_count(ctypes.byref(n))
    if rc != 0 or n.value < 1:
        with pytest.raises(RuntimeError):
            N.load(require_gpu=True)
