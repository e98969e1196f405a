"""C ABI checks that need no GPU: the library loads, exports every symbol
include/ggms.h declares, and its host-only entry points agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(ggms_\w+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import ggms
    lib = ggms._lib.load()
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # every symbol the Python mirror binds is declared in the header
    assert set(ggms._lib.SIGNATURES) <= declared


def test_opts_defaults_and_status_names():
    import ggms
    o = ggms.default_opts()
    assert (o.num_buckets, o.num_pivots, o.block_threads, o.blocks_per_cu, o.splitter) == (0, 17, 1024, 0, 0)
    lib = ggms._lib.load()
    assert lib.ggms_status_str(0) == b"GGMS_OK" and lib.ggms_status_str(2) == b"GGMS_ENAN"
    ws = lib.ggms_workspace_bytes(0, 1 << 28, 1001, C.byref(o))
    assert 2 * (1 << 28) * 4 <= ws < 3 * (1 << 28) * 4


def test_host_k_generators_match_oracle():
    import ggms
    for kd in ggms.KDISTS:
        for n, nk, seed in ((1 << 28, 1001, 5), (1000, 101, 7), (1 << 26, 8192, 9)):
            a = ggms.generate_ks(kd, nk, n, seed)
            b = oracle.generate_ks(kd, nk, n, seed)
            assert np.array_equal(a, b), (kd, n, nk)


def test_host_mixture_params_match_oracle():
    import ggms
    for seed in (0, 9, 2**40 + 3):
        m1, mu1, s1 = ggms.mixture_params(seed)
        m2, mu2, s2 = oracle.mixture_params(seed)
        assert m1 == m2 and np.array_equal(mu1, mu2) and np.array_equal(s1, s2)


def test_invalid_arguments_rejected_without_device_work():
    import ggms
    lib = ggms._lib.load()
    out = np.zeros(4, dtype=np.uint32)
    assert lib.ggms_generate_ks(99, 4, 100, 1, out.ctypes.data_as(C.POINTER(C.c_uint32))) == ggms.EINVAL
    assert lib.ggms_generate_ks(0, 4, 0, 1, out.ctypes.data_as(C.POINTER(C.c_uint32))) == ggms.EINVAL
    assert lib.ggms_select(None, 1, 0, None, 10, None, 1, None, None) == ggms.EINVAL


def _tile_ends_u32(n, tile, groups):
    """Tile bounds exactly as the streaming kernels compute them, in uint32:
    t0 = w * tile, t1 = n - t0 > tile ? t0 + tile : n."""
    w = np.arange(groups, dtype=np.uint64)
    t0 = (w * np.uint64(tile)).astype(np.uint32)
    n32 = np.uint32(n)
    rem = (n32 - t0).astype(np.uint32)
    t1 = np.where(rem > np.uint32(tile), (t0 + np.uint32(tile)).astype(np.uint32), n32).astype(np.uint32)
    return t0, t1


def test_first_pass_tiling_covers_every_element_once():
    """The host tiling (select.hip first_pass_tiling) and the kernels' uint32 tile
    bounds put every element of [0, n) in exactly one tile, up to the length cap."""
    import ggms
    lib = ggms._lib.load()
    tile, groups = C.c_uint32(), C.c_uint32()
    cases = [1, 63, 64, 16383, 16384, 16385, (1 << 20) + 3, 1 << 28, (1 << 31) - 1, (1 << 31) - (1 << 23) + 1,
             (1 << 31) - 2, (1 << 30) + 5]
    for n in cases:
        for wgs in (1, 256, 512, 4096):
            assert lib.ggms_first_pass_tiling(n, wgs, C.byref(tile), C.byref(groups)) == ggms.OK
            t, g = tile.value, groups.value
            assert 1 <= g <= wgs and t % 64 == 0
            assert (g - 1) * t < n <= g * t, (n, wgs, t, g)
            assert g <= max(1, -(-n // 16384))  # no more workgroups than 16384-element tiles
            t0, t1 = _tile_ends_u32(n, t, g)
            assert t0[0] == 0 and t1[-1] == n
            assert np.all(t1[:-1] == t0[1:]) and np.all(t1 > t0), (n, wgs)
            # the last tile's end computed the unguarded way would wrap past 2^32 only above the cap
            assert int(t0[-1]) + t < 1 << 32
    # above the reference's int length the call is refused
    for n in (0, 1 << 31, (1 << 32) - 1, 1 << 40):
        assert lib.ggms_first_pass_tiling(n, 256, C.byref(tile), C.byref(groups)) == ggms.EINVAL


def test_uint32_tile_end_guard_against_wrap():
    """The kernels' tile_end form stays exact where t0 + tile passes 2^32 (the
    form it replaced, t0 + tile < n ? t0 + tile : n, returned 0 there)."""
    n, tile = (1 << 32) - 1, 1 << 23
    g = -(-n // tile)
    t0, t1 = _tile_ends_u32(n, tile, g)
    assert t1[-1] == np.uint32(n) and np.all(t1[:-1] == t0[1:])
    naive = (t0[-1] + np.uint32(tile)).astype(np.uint32)
    assert naive == 0  # the wrap the guard avoids


def test_reference_shim_header_compiles():
    """include/ggms.hpp (the reference's C++ template call shape) compiles for
    all four instantiated types and returns the reference's (T)1 on success."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        import pytest
        pytest.skip("g++ not available")
    src = os.path.join(ROOT, "tests", "shim", "shim_main.cpp")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
                        "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
