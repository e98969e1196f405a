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
