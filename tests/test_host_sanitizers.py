"""The host-only code under AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU): tests/cxx/host_check, built with g++ by tests/cxx/Makefile from the
library's writer, importer and simulation set-up sources and the CPU oracle,
drives them through their edge cases (long and malformed simulation-file
lines, synthetic mass tables, empty clouds, both file layouts, every oracle
field term).  Any sanitizer report or leak fails the run."""
import glob
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cxx", "host_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="build it with make -C tests/cxx")
def test_host_code_under_sanitizers(tmp_path):
    inis = sorted(glob.glob(os.path.join(HERE, "golden", "ini", "*.ini")))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([EXE, str(tmp_path)] + inis, capture_output=True, text=True, timeout=300,
                       env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "host_check ok" in p.stdout
    assert "runtime error" not in p.stderr and "Sanitizer" not in p.stderr
