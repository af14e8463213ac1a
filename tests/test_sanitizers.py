"""The host library's CPU parts and the oracle under sanitizers
(tests/native/host_selftest.cpp, built by tests/native/Makefile): scenes,
the median / SAH / binned-SAH builders and their OpenMP tasks
(kd_tree.cpp:71-75, kd_tree_sah.cpp:127-131 in the reference), the flat
layout and files, the scene loader and the oracle's traversal twin, brute
force and GPU-build twins, with the self-test's own result checks."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
BUILD = os.path.join(NATIVE, "_build")
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")
LLVM = "/opt/rocm/llvm"
ARCHER = os.path.join(LLVM, "lib", "libarcher.so")


def build(target):
    r = subprocess.run(["make", "-C", NATIVE, target], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


def test_asan_ubsan_selftest(tmp_path):
    """AddressSanitizer (with leak checks) and UndefinedBehaviorSanitizer:
    no report, and every result check of the self-test passes."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    build("asan")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="8")
    r = subprocess.run([os.path.join(BUILD, "host_selftest_asan"), SCENES, str(tmp_path)], capture_output=True,
                       text=True, env=env, timeout=900)
    assert r.returncode == 0 and "selftest ok" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-6000:]


def tsan_reports(stderr):
    """The ThreadSanitizer reports of a run, each as its text."""
    parts = re.split(r"^==================\n", stderr, flags=re.MULTILINE)
    return [p for p in parts if "WARNING: ThreadSanitizer" in p]


def on_openmp_task_memory(report):
    """A race on memory that the LLVM OpenMP runtime's task allocator handed
    out (its task descriptors, recycled across threads through the runtime's
    own lock-free lists, which ThreadSanitizer does not see): no data of the
    code under test lives there."""
    m = re.search(r"Location is heap block .*?\n((?:\s+#\d+ .*\n)+)", report)
    return bool(m) and re.search(r"#1 bget\(.*kmp_alloc\.cpp \(libomp\.so", m.group(1)) is not None


def test_tsan_selftest_openmp_builders(tmp_path):
    """ThreadSanitizer with the LLVM OpenMP runtime and its Archer tool (so
    that task and barrier synchronisation is seen) over 1- and 8-thread tree
    builds: no race on the builders' or the oracle's data (the libstdc++
    parallel sort's shared thread count is suppressed in tsan.supp), and the
    builds agree with each other and with brute force."""
    clang = os.path.join(LLVM, "bin", "clang++")
    if not (os.path.exists(clang) and os.path.exists(ARCHER)):
        pytest.skip("clang++ with the LLVM OpenMP runtime and Archer not available")
    build("tsan")
    env = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER, OMP_NUM_THREADS="8",
               TSAN_OPTIONS="halt_on_error=0 ignore_noninstrumented_modules=1 "
                            f"suppressions={os.path.join(NATIVE, 'tsan.supp')}")
    r = subprocess.run([os.path.join(BUILD, "host_selftest_tsan"), SCENES, str(tmp_path)], capture_output=True,
                       text=True, env=env, timeout=900)
    assert "selftest ok" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
    reports = tsan_reports(r.stderr)
    real = [rep for rep in reports if not on_openmp_task_memory(rep)]
    assert not real, real[0][:6000]
    assert r.returncode in (0, 66), r.returncode  # 66: ThreadSanitizer reported (only filtered) warnings


def test_task_memory_filter_recognises_only_runtime_memory():
    """The filter above accepts a race on the runtime's task memory and
    rejects one on the program's own heap."""
    runtime = ("WARNING: ThreadSanitizer: data race (pid=1)\n  Write of size 8 at 0x1 by thread T4:\n"
               "    #0 f kd_tree.cpp:295:1 (x+0x1)\n\n  Location is heap block of size 1048576 at 0x1 allocated by "
               "main thread:\n    #0 malloc tsan_interceptors_posix.cpp:667:5 (x+0x2)\n"
               "    #1 bget(kmp_info*, long) kmp_alloc.cpp (libomp.so+0x46b3d)\n    #2 g kd_tree.cpp:79:13 (x+0x3)\n")
    own = runtime.replace("bget(kmp_info*, long) kmp_alloc.cpp (libomp.so+0x46b3d)",
                          "std::vector<int>::reserve vector.tcc:72 (x+0x4)")
    assert on_openmp_task_memory(runtime)
    assert not on_openmp_task_memory(own)
