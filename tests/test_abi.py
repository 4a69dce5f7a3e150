"""The C-ABI library: it loads, it exports every symbol include/htrip_gpu.h
declares, and it fails loudly (no CPU fallback) where there is no device.
No compute calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from htrip_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "htrip_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(htg_\w+)\s*\(", txt, re.M)))


def test_library_built_and_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run python -m htrip_amd.build"
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    from htrip_amd import serving

    assert sorted(_lib.EXPORTS + serving.SERVER_EXPORTS) == names


def test_footprint_is_host_only_and_matches_geometry():
    # 1024^2, b=32: leaf D alone is 1024 leaves x 1024^2 x 2 bytes = 2 GiB
    fp = _lib.footprint(32, 2, 1024, 1024)
    assert fp > 2 * 1024 ** 3
    with pytest.raises(_lib.HtgError):
        _lib.footprint(1, 2, 64, 64)


def test_no_silent_cpu_fallback():
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.htg_ctx_create(0, C.byref(h))
    if rc != _lib.HTG_OK:  # no device here: a clear error, not a CPU path
        assert rc in (_lib.HTG_E_HIP, _lib.HTG_E_INVALID)
        assert lib.htg_last_error().decode()
    else:
        lib.htg_ctx_destroy(h)


def test_cpp_dropin_header_compiles(tmp_path):
    out = tmp_path / "dropin"
    from htrip_amd import build

    build.build_dropin_test(str(out))
    assert out.exists()
