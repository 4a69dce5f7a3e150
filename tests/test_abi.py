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


def test_query_argument_checks_need_no_device():
    """Bad query arguments are refused before any device work: missing
    viewpoints for k > 0, negative k, null workspace handles."""
    lib = _lib.load()
    ks = (C.c_int32 * 2)(3, 2)
    out = (C.c_uint16 * 16)()
    assert lib.htg_query_batch(None, None, ks, 2, out) == _lib.HTG_E_INVALID
    assert "null viewpoints" in lib.htg_last_error().decode()
    neg = (C.c_int32 * 1)(-1)
    rc = (C.c_int32 * 2)(0, 0)
    assert lib.htg_query_batch(None, rc, neg, 1, out) == _lib.HTG_E_INVALID
    assert "negative k" in lib.htg_last_error().decode()
    assert lib.htg_query_batch_ws(None, None, None, ks, 2, out) == _lib.HTG_E_INVALID
    assert lib.htg_query_all_pairs_ws(None, None, rc, 1, out) == _lib.HTG_E_INVALID
    h = C.c_void_p()
    assert lib.htg_query_ws_create(None, C.byref(h)) == _lib.HTG_E_INVALID
    assert lib.htg_inject_fault(None, _lib.FAULT_UPDATE, 1) == _lib.HTG_E_INVALID
    n = C.c_uint64()
    assert lib.htg_query_ws_bytes(None, C.byref(n)) == _lib.HTG_E_INVALID
    from htrip_amd import serving

    serving._bind(lib)
    assert lib.htg_server_cancel(None, 1) == _lib.HTG_E_INVALID
    assert lib.htg_server_inject_fault(None, 1, 1, 1) == _lib.HTG_E_INVALID


class _FrontArgs(C.Structure):
    """QFrontArgs (htrip_amd/csrc/kernels.hpp): 12 pointers, then 8 ints."""
    _fields_ = [(f"p{i}", C.c_void_p) for i in range(12)] + [
        (n, C.c_int) for n in ("nq", "ib", "leaf_blocks", "n_slots", "b", "nl", "n_fill", "fill_blocks")]


def _launcher(name):
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    sym = [ln.split()[-1] for ln in out.splitlines() if name in ln]
    assert len(sym) == 1
    return getattr(C.CDLL(_lib.LIB_PATH), sym[0])


def test_query_front_refuses_more_than_64_fill_ranges():
    """The front finds a unit's fill range with one 64-lane ballot, so its
    launcher refuses more ranges (returns false, launches nothing) rather
    than leaving gaps uninitialised.  Called with no device: the guard comes
    before any HIP call."""
    f = _launcher("launch_q_front")
    f.restype = C.c_bool
    f.argtypes = [C.POINTER(_FrontArgs), C.c_void_p]

    def args(n_fill, nq):
        a = _FrontArgs()
        a.nq, a.ib, a.b, a.nl, a.n_fill = nq, 1, 8, 1, n_fill
        return C.byref(a)

    assert f(args(65, 1), None) is False
    assert f(args(-1, 1), None) is False
    assert f(args(64, 0), None) is True  # in range, nothing to launch
