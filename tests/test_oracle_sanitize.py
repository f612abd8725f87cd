"""The CPU restatement (oracle/) built with AddressSanitizer and
UndefinedBehaviorSanitizer runs every restated plugin on small 2D and 3D
problems without a report (oracle/sanitize_main.c).  CPU only."""
import os
import subprocess

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_under_asan_ubsan(tmp_path):
    r = subprocess.run(["make", "-s", "-C", ORACLE_DIR, "sanitize-run", "OUT=%s" % tmp_path],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "oracle sanitizer run: clean" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
