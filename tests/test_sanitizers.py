"""CPU: the host layer of libcalaman_expm and the CPU oracle under AddressSanitizer
and UndefinedBehaviorSanitizer.

Two builds, no GPU needed:
* ``make -C calaman_g_amd/csrc host-sanitize``: every library source with the
  sanitizers on its host half, linked with tests/cpp/host_sanitize.cpp, which
  drives the plan rules, workspace sizes and argument checks over ordinary,
  extreme and malformed inputs;
* ``make -C oracle sanitize``: oracle.c with tests/cpp/oracle_sanitize.c, every
  entry point on small, padded, singular and non-finite inputs.
A sanitizer report aborts the program (``-fno-sanitize-recover``), so the test
is the exit status plus the absence of a report in the output.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(build_dir, target, exe, env_extra):
    jobs = str(min(8, os.cpu_count() or 1))
    b = subprocess.run(["make", "-j", jobs, "-C", os.path.join(ROOT, build_dir), target],
                       capture_output=True, text=True, timeout=900)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    env = dict(os.environ, **env_extra)
    r = subprocess.run([os.path.join(ROOT, exe)], capture_output=True, text=True, timeout=600,
                       env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "ok: 0 failure(s)" in out
    return out


@pytest.mark.timeout(1200)
def test_host_layer_clean_under_asan_ubsan():
    # the HIP runtime keeps allocations for the process lifetime: leaks are not checked
    _run(os.path.join("calaman_g_amd", "csrc"), "host-sanitize",
         os.path.join("tests", "cpp", "host_sanitize"), {"ASAN_OPTIONS": "detect_leaks=0"})


@pytest.mark.timeout(600)
def test_oracle_clean_under_asan_ubsan():
    _run("oracle", "sanitize", os.path.join("oracle", "oracle_sanitize"),
         {"ASAN_OPTIONS": "detect_leaks=1", "OMP_NUM_THREADS": "2"})
