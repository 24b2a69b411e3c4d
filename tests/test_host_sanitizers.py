"""The library's host code and the oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only).  tests/sanitize/Makefile builds a check program from the oracle and the host sources of
librkdg that need no GPU (mesh generator, ADCIRC / .meta readers and writers, partitioner,
renumbering, model builder, ensemble merge, L2 residual, modal / VTU writers) with
-fsanitize=address,undefined; it runs them on the reference's fixtures, generated meshes
and malformed files.  Any sanitizer report makes it exit with a nonzero status."""
import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"


def test_host_code_under_sanitizers(tmp_path):
    jobs = str(min(8, os.cpu_count() or 2))
    b = subprocess.run(["make", "-C", str(SAN), "-j", jobs], capture_output=True, text=True)
    assert b.returncode == 0, b.stdout[-2000:] + b.stderr[-2000:]
    # verify_asan_link_order=0: the process may start with other libraries preloaded
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    g = ROOT / "tests" / "golden"
    r = subprocess.run([str(SAN / "_build" / "host_check"), str(g / "weir" / "weir.14"), str(g / "weir" / "weir.bcis"),
                        str(g / "sample_fort.14"), str(tmp_path)], capture_output=True, text=True, env=env,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "host_check: ok" in r.stdout, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
