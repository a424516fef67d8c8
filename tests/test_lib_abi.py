"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/drifter_hip/drf.h declares, and its host-only entry points behave
(no GPU is touched here)."""
import ctypes as C
import os
import re

import numpy as np

from drifter_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include", "drifter_hip")):
        if hdr.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", "drifter_hip", hdr)).read()
            decl = r"^\s*(?:const\s+)?(?:drf_status|void|double|int64_t|int|char)\s*\*?\s*(drf_\w+)\s*\("
            names |= set(re.findall(decl, txt, flags=re.M))
    return names


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)


def _declared_arity():
    """Parameter count of every declaration in the headers."""
    out = {}
    for hdr in os.listdir(os.path.join(ROOT, "include", "drifter_hip")):
        if not hdr.endswith(".h"):
            continue
        txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "drifter_hip",
                                                        hdr)).read(), flags=re.S)
        decl = (r"^\s*(?:const\s+)?(?:drf_status|void|double|int64_t|int|char)\s*\*?\s*"
                r"(drf_\w+)\s*\(([^)]*)\)\s*;")
        for name, params in re.findall(decl, txt, flags=re.M):
            p = params.strip()
            out[name] = 0 if p in ("", "void") else p.count(",") + 1
    return out


def test_ctypes_signatures_match_the_header_arity():
    """Every ctypes argument list has as many entries as the C declaration
    (a short list would pass garbage for the last arguments)."""
    ar = _declared_arity()
    assert len(ar) >= 20
    bad = {n: (k, len(_lib.SIGNATURES[n][1])) for n, k in ar.items()
           if n in _lib.SIGNATURES and len(_lib.SIGNATURES[n][1]) != k}
    assert not bad, bad


def test_invalid_order_is_rejected_before_touching_a_device():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.drf_create(0, 0, 0, C.byref(h)) == _lib.DRF_ERR_UNSUPPORTED
    assert b"order" in lib.drf_last_error()
    assert lib.drf_create(9, 0, 0, C.byref(h)) == _lib.DRF_ERR_UNSUPPORTED
    assert lib.drf_create(3, 7, 0, C.byref(h)) == _lib.DRF_ERR_ARG
    assert not h.value


def test_compute_dt_matches_oracle(orc):
    lib = _lib.load()
    for c in (0.0, 1e-15, 3.7, 1e3, 1e12):
        for cfl in (0.5, 0.9):
            a = lib.drf_compute_dt(c, 0.4, cfl, 1e-6, 3600.0)
            b = orc.compute_dt(c, 0.4, cfl, 1e-6, 3600.0)
            assert a == b


def test_basis_export():
    lib = _lib.load()
    x = np.zeros(5)
    w = np.zeros(5)
    D = np.zeros(25)
    assert lib.drf_basis(4, _lib.dptr(x), _lib.dptr(w), _lib.dptr(D)) == 0
    assert x[0] == -1.0 and x[-1] == 1.0 and abs(w.sum() - 2) < 1e-15
    assert lib.drf_basis(0, None, None, None) == _lib.DRF_ERR_ARG


def test_blocked_layout_block_sizes_and_host_conversion():
    """Elements per block of the blocked layout: the volume kernel's widest
    batch (512 lines up to order 4, 128 above, at most 20 elements of order 4
    by the LDS budget); every narrower workgroup's batch divides it.  The
    host conversion round-trips and puts node (e, f, i, L) at
    ((((e / B) 5 + f) n + i) B + e % B) n^2 + L."""
    from drifter_amd import synth

    lib = _lib.load()
    got = [lib.drf_layout_block_elements(o) for o in range(0, 10)]
    assert got == [0, 128, 56, 32, 20, 3, 2, 2, 1, 0]
    for order, B in zip(range(1, 5), got[1:5]):
        n2 = (order + 1) ** 2
        for lines in (128, 256):
            assert B % max(1, min(lines // n2, B)) == 0
    rng = np.random.default_rng(0)
    for order in (1, 4, 5, 8):
        n, B, E = order + 1, got[order], 455
        U = rng.random(E * 5 * n ** 3)
        V = synth.to_blocks(U, E, n, B)
        assert V.size == -(-E // B) * B * 5 * n ** 3
        assert np.array_equal(synth.from_blocks(V, E, n, B), U)
        for e, f, i, L in ((0, 0, 0, 0), (E - 1, 4, n - 1, n * n - 1), (B, 2, 1, 3)):
            off = (((e // B * 5 + f) * n + i) * B + e % B) * n * n + L
            assert V[off] == U[(e * 5 + f) * n ** 3 + i + n * L]


def test_missing_library_fails_loudly(tmp_path):
    """No CPU fallback: without the built HIP library the package raises
    instead of computing anything (checked in a fresh interpreter so the
    loaded library of this process is not touched)."""
    import subprocess
    import sys

    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from drifter_amd import _lib\n"
        "_lib.LIB_PATH = %r\n"
        "try:\n"
        "    _lib.load()\n"
        "except RuntimeError as e:\n"
        "    print('raised:', 'no CPU fallback' in str(e))\n"
        "    sys.exit(0)\n"
        "sys.exit(1)\n"
    ) % (str(ROOT), str(tmp_path / "libdrifter_hip.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "raised: True" in r.stdout
