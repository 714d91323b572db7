"""Host code under the sanitizers: the FP64 oracle and the driver's INI parser
built with -fsanitize=address,undefined (g++, host only) and run on small
cases; any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined"]


def _build_and_run(tmp_path, sources, extra, args=()):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "prog")
    subprocess.run(["g++", *FLAGS, *extra, "-o", exe, *sources], check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)
    report = p.stdout + p.stderr
    assert p.returncode == 0, report
    assert "AddressSanitizer" not in report and "runtime error" not in report, report
    return p.stdout


def test_oracle_under_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, [os.path.join(REPO, "oracle", "shrt_oracle.cpp"),
                                    os.path.join(REPO, "tests", "host", "oracle_sanitize.cpp")],
                         ["-fopenmp", "-ffp-contract=off"])
    assert out.count("oracle ") == 3
    assert "remaps 1" in out


def test_ini_parser_under_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, [os.path.join(REPO, "shearing-rt_amd", "cli", "ini.cpp"),
                                    os.path.join(REPO, "tests", "host", "ini_sanitize.cpp")], [],
                         [os.path.join(REPO, "tests", "golden", "config_sample.ini")])
    assert "missing 42" in out
