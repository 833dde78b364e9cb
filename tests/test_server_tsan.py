"""The serving loop's threading under ThreadSanitizer (CPU only).

server.cpp (admission, reservations, completion) and driver.cpp (the
heat-bath loop it runs per request) are compiled with g++
-fsanitize=thread against a CPU stand-in for the device engine
(tests/tsan/engine_stub.cpp: no HIP calls; each phase sleeps and moves the
thread's device-byte counters).  Two client threads submit 40 requests to a
four-slot server while a third polls its statistics; the program checks
every result against the same request run alone, and TSan stops it with
status 66 at the first data race.  A second scenario checks admission: a
request too large to fit beside the running one waits, and a small request
queued after it starts first.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "vhci_amd", "csrc")
TSAN = os.path.join(ROOT, "tests", "tsan")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_server_threads_race_free(tmp_path):
    exe = str(tmp_path / "server_tsan")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(TSAN, "server_main.cpp"), os.path.join(TSAN, "engine_stub.cpp"),
           os.path.join(CSRC, "server.cpp"), os.path.join(CSRC, "driver.cpp"), os.path.join(CSRC, "problem.cpp"),
           "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-6000:])
    assert "0 differing from alone" in r.stdout
    # a large request waiting for memory does not hold back a smaller one queued after it
    assert "while the first ran: yes" in r.stdout and "passed over 1 time(s); all done: yes" in r.stdout
