"""The reference-format sweep on the GPU: rows for every algorithm beside the
GPU sort-and-choose, zero error flags (every value equals sort-and-choose's, as
the reference's harness checks, DistributedSMOS/runTestsDistributedSMOS.cpp:427-461),
and the summary files."""
import os

import pytest

pytestmark = pytest.mark.gpu


def test_reference_format_sweep(gpu_ctx, tmp_path):
    from ggms import refbench as R
    path = R.sweep("float32", "normal", "uniform_random", 16, 17, 101, 400, 501, 2,
                   ("bucket", "iterative", "distributed", "distributed_bucket"), str(tmp_path), ranks=1)
    lines = [l for l in open(path).read().split("\n") if l]
    assert len(lines) == 2 * 2 * 2
    for l in lines:
        size, nk, dist, kd, pairs, err = R.parse_row(l)
        assert dist == "NORMAL FLOATS" and kd == "Uniform Random Ks" and err == 0
        assert [n for n, _ in pairs] == ["Sort and Choose Multiselect", "Bucket Multiselect", "IterativeSMOS",
                                         "DistributedSMOS", "DistributedBucketMultiSelect"]
        assert l.split(",")[1] == "1"
    base = os.path.basename(path)
    for p in R.summarize(base, str(tmp_path)):
        assert os.path.getsize(p) > 0
