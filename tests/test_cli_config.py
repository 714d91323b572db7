"""The shrt driver reads config.ini like the reference's input_data
(file_access.cu:123-188): boost::property_tree::read_ini's dialect and
readEntry's typed lookups with a warning and the default for a missing or
unconvertible key.  The fixture config_sample.ini is the reference's own
input_file/config_sample.txt.  Runs the built driver with --print-config
(no GPU is touched)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "shearing-rt_amd", "shrt")
SAMPLE = os.path.join(REPO, "tests", "golden", "config_sample.ini")

WARN = 'WARNING: readEntry: Key "{}" in section [{}] not found. Using default.'


def run(text=None, path=None, args=(), tmp_path=None):
    if path is None:
        path = os.path.join(tmp_path, "config.ini")
        with open(path, "w") as f:
            f.write(text)
    if not os.path.exists(BIN):
        raise RuntimeError("the driver is not built: make -C shearing-rt_amd")
    p = subprocess.run([BIN, "--config", path, "--print-config", *args], capture_output=True, text=True,
                       timeout=60)
    lines = p.stdout.splitlines()
    warns = [l for l in lines if l.startswith("WARNING")]
    vals = dict(l.split(" = ", 1) for l in lines if " = " in l and not l.startswith("WARNING"))
    return p.returncode, warns, vals, p.stdout


def test_reference_sample_config(tmp_path):
    rc, warns, v, _ = run(path=SAMPLE, args=("--dble",))
    assert rc == 0
    # the sample lacks these two keys; "eps" and "rho_eps2" are present but unread
    assert warns == [WARN.format("cfl number", "simulation"), WARN.format("linear eq", "problem")]
    assert v["nx"] == "256" and v["ny"] == "512" and v["cuda thread num"] == "1024"
    assert float(v["Lx"]) == 3.1415926535897931 and float(v["Ly"]) == 3.1415926535897931
    assert float(v["time step"]) == 1e-3 and float(v["time max"]) == 30
    assert float(v["cfl number"]) == 0.1
    assert float(v["kappa"]) == 8e-5 and float(v["nu"]) == 1e-3
    assert float(v["rho_eps2"]) == 0.1 * 1e-2   # derived, not read (file_access.cu:164)
    assert v["initial noise"] == "1" and v["linear eq"] == "0" and v["write output"] == "1"
    assert v["output loop count"] == "100" and float(v["output time step"]) == 1.0


def test_fp32_reads_floats(tmp_path):
    """cureal = float in the default build: values are parsed as float."""
    rc, _, v, _ = run(path=SAMPLE)
    import numpy as np
    assert float(v["time step"]) == float(np.float32(1e-3))
    assert float(v["Lx"]) == float(np.float32(np.pi))
    assert v["precision"] == "fp32"


def test_dialect(tmp_path):
    text = """
; a comment
# another
top = 1
   [simulation]
   nx    =    64
ny=32
time step = 2.5e-4
Lx = 6.28 # trailing text is part of the value
cuda thread num = 1e3
[problem]
initial noise = false
linear eq = True
nu =
sigma = -0.5
[output]
write output = 0
output loop count = 7
"""
    rc, warns, v, _ = run(text, tmp_path=str(tmp_path), args=("--dble",))
    assert rc == 0
    assert v["nx"] == "64" and v["ny"] == "32" and float(v["time step"]) == 2.5e-4
    assert WARN.format("Lx", "simulation") in warns and float(v["Lx"]) == 3.141592653589793
    assert WARN.format("cuda thread num", "simulation") in warns and v["cuda thread num"] == "1024"
    assert v["initial noise"] == "0"
    assert WARN.format("linear eq", "problem") in warns and v["linear eq"] == "0"
    assert WARN.format("nu", "problem") in warns and float(v["nu"]) == 1e-3
    assert float(v["sigma"]) == -0.5
    assert v["write output"] == "0" and v["output loop count"] == "7"


@pytest.mark.parametrize("text,reason", [
    ("[simulation]\nnx = 1\nnx = 2\n", "(3): duplicate key name"),
    ("[simulation]\nnx 64\n", "(2): '=' character not found in line"),
    ("[simulation\nnx = 1\n", "(1): unmatched '['"),
    ("[a]\nx = 1\n[a]\ny = 2\n", "(3): duplicate section name"),
    ("[a]\n= 3\n", "(2): key expected"),
])
def test_parse_errors(tmp_path, text, reason):
    rc, _, _, out = run(text, tmp_path=str(tmp_path))
    assert rc == 1
    assert out.startswith("ERROR: unable to read config file: ") and out.endswith(reason)


def test_empty_section_may_repeat(tmp_path):
    rc, _, v, _ = run("[simulation]\n[simulation]\nnx = 128\n", tmp_path=str(tmp_path))
    assert rc == 0 and v["nx"] == "128"


def test_missing_file(tmp_path):
    rc, _, _, out = run(path=os.path.join(str(tmp_path), "nope.ini"))
    assert rc == 1 and out.startswith("ERROR: unable to read config file: ")
