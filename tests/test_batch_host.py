"""Host-side checks of the batched-selection API (no GPU): struct layout,
defaults, size classes, argument validation before any device work, and the
autotuner table format."""
import ctypes as C
import json

import numpy as np


def test_batch_opts_layout_and_defaults():
    import ggms
    from ggms import _lib
    from ggms.batch import default_opts
    assert C.sizeof(_lib.BatchClass) == 16
    assert C.sizeof(_lib.BatchOpts) == 4 * 4 + 8 + 24 * 16
    o = default_opts()
    assert (o.workgroups, o.block_threads, o.use_graph, o.seed) == (0, 0, 0, 0x5EED)
    for c in range(_lib.GGMS_BATCH_CLASSES):
        k = o.classes[c]
        assert 16 <= k.num_buckets <= 8192 and 2 <= k.pieces <= 64 and k.sample_size > k.pieces
        assert k.rebucket == 8192
    o2 = default_opts({12: dict(num_buckets=999, sample_size=300, pieces=10)})
    assert (o2.classes[12].num_buckets, o2.classes[12].sample_size, o2.classes[12].pieces) == (999, 300, 10)
    assert o2.classes[11].num_buckets == o.classes[11].num_buckets


def test_size_class():
    from ggms.batch import size_class
    assert [size_class(n) for n in (1, 2, 3, 4, 5, 1024, 1025, 1 << 22, (1 << 22) + 1)] == [0, 1, 2, 2, 3, 10, 11, 22,
                                                                                               23]
    assert size_class(1 << 40) == 23


def _create(lib, off, koff, nprob=None, dks=1, dout=1, vals=1, opts=None):
    h = C.c_void_p()
    off = np.ascontiguousarray(off, dtype=np.uint64)
    koff = np.ascontiguousarray(koff, dtype=np.uint32)
    return lib.ggms_batch_plan_create(0, 0, C.c_void_p(vals), len(off) - 1 if nprob is None else nprob,
                                      off.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_void_p(dks),
                                      koff.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_void_p(dout),
                                      opts, C.byref(h))


def test_batch_plan_rejects_bad_shapes_without_device_work():
    import ggms
    lib = ggms._lib.load()
    E = ggms.EINVAL
    assert _create(lib, [0, 10], [0, 1], nprob=0) == E                 # no problems
    assert _create(lib, [0, 10, 5], [0, 1, 2]) == E                    # offsets decrease
    assert _create(lib, [0, 10], [0, 2049]) == E                       # too many ks
    assert _create(lib, [0, 10, 20], [0, 3, 1]) == E                   # koffsets decrease
    assert _create(lib, [0, 1 << 32], [0, 1]) == E                     # problem of 2^32 values
    assert _create(lib, [0, 10], [0, 1], dks=0) == E                   # ks pointer missing
    assert _create(lib, [0, 10], [0, 1], vals=0) == E                  # values pointer missing


def test_autotune_table_roundtrip(tmp_path):
    from ggms import autotune
    from ggms.batch import default_opts
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"block_threads": 512, "classes": {"14": dict(num_buckets=2048, sample_size=1024,
                                                                        pieces=32, rebucket=4096)}}))
    classes, bt = autotune.load(str(p))
    assert bt == 512 and classes == {14: dict(num_buckets=2048, sample_size=1024, pieces=32, rebucket=4096)}
    o = default_opts(classes)
    assert (o.classes[14].num_buckets, o.classes[14].rebucket) == (2048, 4096)
    assert autotune.load(str(tmp_path / "missing.json")) == (None, 0)
