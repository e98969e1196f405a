"""CPU emulation of device kernels (no GPU): the kernel sources are compiled
with g++ against a small HIP shim (tests/emu/hip/hip_runtime.h) and run with
one thread per work-item, checked barriers and 64-lane waves (ballots,
readlanes, shuffles and scans are rendezvous of a wave's active lanes).  This covers
kernel logic that the GPU parity tests cannot isolate: barrier agreement on
every path (including the error paths), the batched kernel's planning and
work stack, the LDS sort network, the finish's multi-rank radix select and the
first-pass splitter's range table.

Each driver compares its results with std::sort / a binary-search map."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ggms-distributed_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "emu")


_SHARED = re.compile(r"__shared__\s+(?:__align__\(\d+\)\s+)?(?P<type>[A-Za-z_][\w:<>]*)\s+(?P<decls>[^;]+);")


def _per_workgroup_shared(src):
    """Every `__shared__ T a[n], b;` becomes per-workgroup slots:
    `auto& a = emu_shared<T[n], ID>(); auto& b = emu_shared<T, ID+1>();`."""
    counter = [0]

    def repl(m):
        out = []
        for d in m.group("decls").split(","):
            d = d.strip()
            name = re.match(r"[A-Za-z_]\w*", d).group(0)
            dims = d[len(name):].strip()
            counter[0] += 1
            out.append(f"auto& {name} = emu_shared<{m.group('type')}{dims}, {counter[0]}>();")
        return " ".join(out)

    return _SHARED.sub(repl, src)


def _prepare(dst):
    """Copy the kernel headers with the substitutions the emulation needs: the
    dynamic-LDS declaration becomes a pointer the driver sets, every other
    __shared__ variable a per-workgroup slot, the wave-wide inclusive scan
    calls the emulated 64-thread scan, and the nontemporal streaming load (a
    clang vector builtin) becomes a plain load."""
    base = {"select_kernels.hpp": 1000, "splitter.hpp": 2000, "batch": 3000}
    for f in ["ggms_device.hpp", "philox.hpp", "ggms_internal.hpp", "select_kernels.hpp", "splitter.hpp"]:
        s = open(os.path.join(CSRC, f)).read()
        s = s.replace("extern __shared__ __align__(16) unsigned char smem[];", "unsigned char* smem = g_dyn_smem;")
        s = _per_workgroup_shared(s) if f in base else s
        if f in base:
            s = re.sub(r"emu_shared<(.*?), (\d+)>\(\)", lambda m: f"emu_shared<{m.group(1)}, {base[f] + int(m.group(2))}>()", s)
        if f == "select_kernels.hpp":
            s, n = re.subn(r"__device__ __forceinline__ uint32_t wave_incl_scan_u32\(uint32_t v\) \{.*?\n\}\n",
                           "__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) "
                           "{ return emu_wave_incl_scan(v); }\n", s, flags=re.S)
            assert n == 1
            s, n = re.subn(r"__device__ __forceinline__ uint4 ld_stream\(const uint4\* p\) \{.*?\n\}\n",
                           "__device__ __forceinline__ uint4 ld_stream(const uint4* p) { return *p; }\n", s, flags=re.S)
            assert n == 1
        open(os.path.join(dst, f), "w").write(s)
    s = open(os.path.join(CSRC, "batch.hip")).read()
    s = s[: s.index("using namespace ggms;")]
    s = s.replace("extern __shared__ __align__(16) unsigned char smem[];", "unsigned char* smem = g_dyn_smem;")
    s = _per_workgroup_shared(s)
    s = re.sub(r"emu_shared<(.*?), (\d+)>\(\)", lambda m: f"emu_shared<{m.group(1)}, {base['batch'] + int(m.group(2))}>()", s)
    open(os.path.join(dst, "batch_kernel.hpp"), "w").write(s)


@pytest.fixture(scope="module")
def gen_dir(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("emu")
    _prepare(str(d))
    return str(d)


def _build_and_run(gen_dir, driver, args=(), timeout=900, asan=False):
    exe = os.path.join(gen_dir, driver.replace(".cpp", "") + ("_asan" if asan else ""))
    opt = ["-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"] if asan else ["-O2"]
    cmd = ["g++", "-std=c++20", *opt, "-fno-strict-aliasing", "-pthread", "-I", EMU, "-I", os.path.join(ROOT, "include"), "-I", gen_dir,
           os.path.join(EMU, driver), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0") if asan else None
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    return r.stdout


def test_lds_bitonic_sort_network(gen_dir):
    out = _build_and_run(gen_dir, "sort_driver.cpp")
    assert "sort ok" in out


def test_finish_multi_rank_radix_select(gen_dir):
    """Up to eight ranks of one segment in a single radix sweep per digit; the one-wave radix
    select (every wave of the workgroup on its own segment and rank)."""
    out = _build_and_run(gen_dir, "select_driver.cpp", [60])
    assert "select ok" in out


def test_first_pass_splitter_ranges_and_cell_table(gen_dir):
    out = _build_and_run(gen_dir, "plan0_driver.cpp")
    assert out.count(" ok") == 28


def test_batched_kernel_paths(gen_dir):
    """Error statuses (bad k, NaN in a large and in a small problem), random
    mixed batches with duplicates, and tiny first-pass tables that force the
    re-bucketing work stack."""
    out = _build_and_run(gen_dir, "batch_driver.cpp", [256])
    assert out.count(", ok") == 5
    # the wave operations ran with many lanes at once, and with some lanes
    # masked off (divergent branches, loops left early), as on the GPU
    m = re.search(r"waves: (\d+) operations, (\d+) with several active lanes, (\d+) with some lanes inactive", out)
    assert m and int(m.group(2)) > 1000 and int(m.group(3)) > 100, out[-500:]


def _write_batch(path, parts):
    import numpy as np
    with open(path, "wb") as f:
        f.write(np.uint32(len(parts)).tobytes())
        for v, ks in parts:
            v = np.ascontiguousarray(v, dtype=np.float32)
            ks = np.asarray(ks, dtype=np.uint32)
            f.write(np.array([v.size, ks.size], dtype=np.uint32).tobytes())
            f.write(v.tobytes())
            f.write(ks.tobytes())


def test_batched_kernel_replays_the_gpu_nan_batches(gen_dir):
    """The GPU NaN batch value for value (numpy's generator, the library's defaults:
    1024 threads, one workgroup per problem, seed and class table), and its earlier form
    with invalid ks passed to the kernel itself: every status and answer right, with
    AddressSanitizer watching every access."""
    from test_gpu_batch import nan_batch_parts
    files = []
    for bad in (False, True):
        path = os.path.join(gen_dir, f"nan_batch_{int(bad)}.bin")
        _write_batch(path, nan_batch_parts(bad_ks=bad))
        files.append(path)
    out = _build_and_run(gen_dir, "batch_driver.cpp", [256, *files], asan=True)
    assert out.count(", ok") == 4, out


def test_batched_kernel_paths_under_address_sanitizer(gen_dir):
    """The same batches with the driver built with AddressSanitizer (host code): every
    access of the kernel to its work stacks, scratch regions, outputs and dynamic LDS
    (allocated at exactly the launch's size) stays inside its buffer."""
    out = _build_and_run(gen_dir, "batch_driver.cpp", [256], asan=True)
    assert out.count(", ok") == 5
