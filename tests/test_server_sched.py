"""The serving form's scheduler (resmcm_ari_amd/csrc/server_sched.h: queues, lanes, admission by
predicted bytes, shutdown), the same code the server runs, stress-tested on the CPU with a stub
executor under ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer
(tests/sched/sched_stress.cpp): an 8-slot server then a 4-slot server in one process, immediate
shutdown after hand-overs to the large lane, a budget that admits one job at a time, no lanes.
Any sanitizer report, failed check or hang (watchdog) fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "sched", "sched_stress.cpp")
INC = os.path.join(ROOT, "resmcm_ari_amd", "csrc")

SANITIZERS = {
    "thread": ["-fsanitize=thread"],
    "address_undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("san", list(SANITIZERS))
def test_scheduler_under_sanitizer(tmp_path, san):
    exe = str(tmp_path / f"sched_{san}")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", *SANITIZERS[san], "-I" + INC, SRC, "-o", exe],
                   check=True, capture_output=True, text=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               ASAN_OPTIONS="halt_on_error=1 detect_leaks=1")
    p = subprocess.run([exe, "10", "180"], capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "sched stress ok" in p.stdout
    assert "WARNING: ThreadSanitizer" not in p.stderr and "ERROR: AddressSanitizer" not in p.stderr
