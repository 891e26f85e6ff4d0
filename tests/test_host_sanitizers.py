"""Sanitizer builds of the host engine and server (CPU only).

The host C++ (JSON, parameters, TIFF and zstd IO, the schedules with their IO
threads and pinned double buffers, the chunk planner, the server's
dispatcher / scheduler / worker threads) is compiled with g++ under
AddressSanitizer + UndefinedBehaviorSanitizer and, separately, under
ThreadSanitizer, and linked against tests/sanitize/host_device.cpp: a host
stand-in for the device entry points that keeps the HIP runtime's ordering
(one worker thread per stream, events, asynchronous copies from pinned
memory) and touches exactly the buffers each kernel reads and writes.
tests/sanitize/host_drive.cpp then runs every schedule, the in-memory entry
from two threads, the chunk planner, error paths and both server executions.
Each build must run to the end with no sanitizer report.  (The reference has
none: SURVEY's race list - a non-blocking upload from a vector freed at scope
end, reconstruction_lpln.hpp:64-65 - is the class of bug the stand-in's
asynchronous copies expose.)
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.tiff_util import write_f32

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "xpulse_recon_amd", "csrc", "host")
SRCS = [os.path.join(HOST, f) for f in ("json.cpp", "params.cpp", "io.cpp", "recon.cpp",
                                        "capi.cpp", "serve.cpp")]
SRCS += [os.path.join(ROOT, "xpulse_recon_amd", "csrc", "geometry.cpp"),
         os.path.join(ROOT, "tests", "sanitize", "host_device.cpp"),
         os.path.join(ROOT, "tests", "sanitize", "host_drive.cpp")]
FLAGS = {"asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                  "-fno-omit-frame-pointer"],
         "tsan": ["-fsanitize=thread"]}
ENV = {"asan": {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:halt_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"},
       "tsan": {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"}}


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("sanitize")
    procs = {}
    for kind, flags in FLAGS.items():
        exe = str(out / f"host_drive_{kind}")
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", *flags, "-o", exe, *SRCS,
               "-ldl", "-lpthread"]
        procs[kind] = (exe, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                             text=True))
    exes = {}
    for kind, (exe, p) in procs.items():
        log, _ = p.communicate(timeout=600)
        assert p.returncode == 0, f"{kind} build failed:\n{log[-4000:]}"
        exes[kind] = exe
    # the TIFF projections of a small scan: 24 views of 24 x 16
    data = out / "data"
    (data / "in").mkdir(parents=True)
    rng = np.random.default_rng(7)
    for a in range(24):
        write_f32(data / "in" / f"proj_{a:04d}.tif",
                  rng.uniform(0.2, 0.9, (16, 24)).astype(np.float32))
    return exes, data


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_engine_and_server_under_sanitizer(builds, kind):
    exes, data = builds
    env = dict(os.environ, **ENV[kind])
    r = subprocess.run([exes[kind], str(data)], capture_output=True, text=True, timeout=900,
                       env=env)
    out = r.stdout + r.stderr
    reports = [ln for ln in out.splitlines()
               if "ERROR: AddressSanitizer" in ln or "WARNING: ThreadSanitizer" in ln
               or "runtime error:" in ln or "ERROR: LeakSanitizer" in ln]
    assert not reports, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert "host_drive: 0 failure(s)" in out


@pytest.mark.parametrize("kind,arg,report", [
    ("tsan", "--race", "WARNING: ThreadSanitizer: data race"),
    ("asan", "--overflow", "ERROR: AddressSanitizer: heap-buffer-overflow")])
def test_stand_in_exposes_the_bugs_it_is_for(builds, kind, arg, report):
    """The stand-in is only useful if it makes these visible: a host write to a
    pinned buffer whose asynchronous upload has not been waited for (a race
    with the stream's copy), and a kernel extent past its allocation."""
    exes, _ = builds
    env = dict(os.environ, **ENV[kind])
    r = subprocess.run([exes[kind], arg], capture_output=True, text=True, timeout=300, env=env)
    assert report in r.stdout + r.stderr, (r.stdout + r.stderr)[-4000:]
