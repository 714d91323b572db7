"""End-to-end: the shrt driver (main.cu's loop on the MI355X library) on a small
config, against the reference's output schedule and file formats and against
the oracle's fields at the output steps."""
import os
import subprocess

import numpy as np
import pytest

from conftest import oracle, rel_l2

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "shearing-rt_amd", "shrt")

CONFIG = """[simulation]
nx = 32
ny = 32
cuda thread num = 1024
time step = 1e-3
time max = 0.0185
cfl number = 0.5
Lx = 3.1415926535897932384626433832795
Ly = 3.1415926535897932384626433832795

[problem]
initial noise = true
linear eq = false
nu = 1e-3
kappa = 8e-5
sigma = 1.0
g = 1.0
rho0_prime = 1.0
rho0 = 1.0
rho_eps1 = 1e-2

[output]
write output = true
output time step = 0.0075
output loop count = 1
"""


def schedule():
    """The output steps of main.cu:46-66 for this config, in FP64 as the run
    computes them (time += delt, next_output_time += output time step)."""
    t, nxt, steps = 0.0, 0.0075, []
    while t <= 0.0185:
        t += 1e-3
        if t > nxt:
            nxt += 0.0075
            steps.append((round(t / 1e-3), "%09.6f" % t))
    return steps


OUT = schedule()   # [(8, '00.008000'), (15, '00.015000')]


def read_table(path, ncol):
    rows = [l.split() for l in open(path) if l.strip()]
    a = np.array(rows, dtype=np.float64)
    assert a.shape[1] == ncol
    return a


@pytest.fixture(scope="module")
def run_dir(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("shrt_run")
    with open(d / "config.ini", "w") as f:
        f.write(CONFIG)
    print("driver run in", d, flush=True)
    p = subprocess.run([BIN, "--dble"], cwd=d, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return d, p.stdout


def test_output_schedule(run_dir):
    d, out = run_dir
    names = set(os.listdir(d))
    assert OUT == [(8, "00.008000"), (15, "00.015000")]
    for _, t in [(0, "00.000000")] + OUT:
        assert f"vel_t{t}.dat" in names and f"ksre_t{t}.dat" in names and f"ksim_t{t}.dat" in names
    assert {"n00000_t00.000000.dat", "n00008_t00.008000.dat", "n00015_t00.015000.dat"} <= names
    assert len([n for n in names if n.startswith("n000")]) == 3
    assert {"kx_ensp.txt", "ky_ensp.txt", "after_kdata.txt", "pbk.txt", "max_amp.dat", "energy.dat"} <= names
    # output_pbk at the output steps only, labelled with the time after the step
    assert sorted(n for n in names if n.startswith("pbk_t")) == [f"pbk_t{t}.dat" for _, t in OUT]
    assert {"kx_ensp_t00.008000.dat", "ky_ensp_t00.015000.dat"} <= names
    # k_data_bef at istep 0 and istep 1 only (sigma != 0, nonlinear)
    assert sorted(n for n in names if n.startswith("befk")) == ["befk_n00000_t00.000000.dat",
                                                                 "befk_n00001_t00.001000.dat"]
    assert not any(n.startswith("aftk") for n in names)
    assert out.startswith("yy[ny] = 3.141593\n")
    assert out.count("time = ") == 3
    assert len(open(d / "max_amp.dat").read().splitlines()) == 1
    assert len(open(d / "energy.dat").read().splitlines()) == 3


def test_field_files_match_oracle(run_dir):
    d, _ = run_dir
    o = oracle.Oracle(nx=32, ny=32, Lx=np.pi, Ly=np.pi, delt=1e-3, cfl_num=0.5, nu=1e-3, kappa=8e-5,
                      sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2, noise=1, linear=0)
    o.initialize_reference()
    done = 0
    for istep, name in ((0, "n00000_t00.000000.dat"), (8, "n00008_t00.008000.dat"), (15, "n00015_t00.015000.dat")):
        o.step(istep - done)
        done = istep
        a = read_table(d / name, 5)
        assert a.shape[0] == 33 * 33
        grid = a.reshape(33, 33, 5)
        # periodic closure of file_access.cu:266-286
        assert np.array_equal(grid[32, :32, 2:], grid[0, :32, 2:])
        assert np.array_equal(grid[:32, 32, 2:], grid[:32, 0, 2:])
        assert np.allclose(grid[:32, 0, 0], np.arange(32) * np.pi / 32)
        for col, ring in ((2, "aomg"), (3, "aphi"), (4, "arho")):
            ref = o.ktox_shear(o.buffer(ring))
            assert rel_l2(grid[:32, :32, col], ref) < 1e-9, (name, ring)


def test_spectral_dumps(run_dir):
    d, _ = run_dir
    o = oracle.Oracle(nx=32, ny=32, Lx=np.pi, Ly=np.pi, delt=1e-3, cfl_num=0.5, nu=1e-3, kappa=8e-5,
                      sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2, noise=1, linear=0)
    o.initialize_reference()
    o.step(1)
    text = open(d / "befk_n00001_t00.001000.dat").read()
    omg = o.buffer("aomg")
    lines = [l for l in text.splitlines() if l and not l.startswith("/")]
    re = np.array([[float(x) for x in l.split()] for l in lines[:omg.shape[0]]])
    assert re.shape == omg.shape
    assert rel_l2(re, omg.real) < 1e-9
    o.step(7)
    ks = read_table(d / "ksre_t00.008000.dat", 5)
    nkx, nky = o.shape_k
    assert ks.shape[0] == nkx * nky
    assert np.all(np.diff(ks[::nky, 0]) > 0)   # rows in ascending kx (kx_index order)
    assert rel_l2(np.sort(ks[:, 2]), np.sort(np.abs(o.buffer("aomg").real).ravel())) < 1e-9


def test_pbk_files_match_oracle(run_dir):
    """pbk_t*.dat holds the bracket advect_omg computed in the step that ended
    past the output time: that of the state before the step."""
    d, _ = run_dir
    o = oracle.Oracle(nx=32, ny=32, Lx=np.pi, Ly=np.pi, delt=1e-3, cfl_num=0.5, nu=1e-3, kappa=8e-5,
                      sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2, noise=1, linear=0)
    o.initialize_reference()
    assert os.path.getsize(d / "pbk.txt") == 0
    done = 0
    for istep, t in OUT:
        o.step(istep - 1 - done)
        done = istep - 1
        o.check_cfl()
        ref = o.poisson_bracket(o.buffer("aomg"))
        a = read_table(d / f"pbk_t{t}.dat", 3)
        grid = a.reshape(33, 33, 3)
        assert np.array_equal(grid[32, :32, 2], grid[0, :32, 2])
        assert np.array_equal(grid[:32, 32, 2], grid[:32, 0, 2])
        assert grid[32, 32, 2] == grid[0, 0, 2]
        assert rel_l2(grid[:32, :32, 2], ref) < 1e-9, t


# --- FP32 driver: the reference's default build keeps time in float ----------

CONFIG_F32 = CONFIG.replace("time step = 1e-3", "time step = 0.01").replace(
    "time max = 0.0185", "time max = 0.35").replace("output time step = 0.0075", "output time step = 0.1")


def schedule_f32(dt=0.01, out=0.1, tmax=0.35):
    """main.cu:46-66 with cureal = float: time += delt and next_output_time +=
    output time step in float, as the default (FP32) build runs them."""
    f = np.float32
    t, nxt, steps, n = f(0), f(out), [], 0
    while t <= f(tmax):
        t = f(t + f(dt))
        n += 1
        if t > nxt:
            nxt = f(nxt + f(out))
            steps.append((n, "%09.6f" % t))
    return steps


def f32(x):
    return float(np.float32(x))


@pytest.fixture(scope="module")
def run_f32(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("shrt_f32")
    with open(d / "config.ini", "w") as fh:
        fh.write(CONFIG_F32)
    print("FP32 driver run in", d, flush=True)
    p = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return d


def test_fp32_output_schedule_follows_float_time(run_f32):
    """Float accumulation moves the third output to step 31 (t = 0.31 in float),
    where FP64 time would give step 30: the FP32 driver follows the float loop."""
    d = run_f32
    out = schedule_f32()
    assert out == [(11, "00.110000"), (20, "00.200000"), (31, "00.310000")]
    names = set(os.listdir(d))
    fields = sorted(n for n in names if n.startswith("n000"))
    assert fields == ["n00000_t00.000000.dat"] + [f"n{i:05d}_t{t}.dat" for i, t in out]
    assert sorted(n for n in names if n.startswith("pbk_t")) == [f"pbk_t{t}.dat" for _, t in out]


def test_fp32_field_files_match_oracle(run_f32):
    d = run_f32
    o = oracle.Oracle(nx=32, ny=32, Lx=f32(np.pi), Ly=f32(np.pi), delt=f32(0.01), cfl_num=f32(0.5),
                      nu=f32(1e-3), kappa=f32(8e-5), sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
                      rho_eps1=f32(1e-2), noise=1, linear=0)
    o.initialize_reference()
    done = 0
    for istep, t in schedule_f32():
        o.step(istep - done)
        done = istep
        grid = read_table(d / f"n{istep:05d}_t{t}.dat", 5).reshape(33, 33, 5)
        for col, ring in ((2, "aomg"), (3, "aphi"), (4, "arho")):
            ref = o.ktox_shear(o.buffer(ring))
            assert rel_l2(grid[:32, :32, col], ref) < 1e-5, (istep, ring)


# --- a remap inside the run: update_shear's befk / aftk dumps -----------------

CONFIG_REMAP = CONFIG.replace("time step = 1e-3", "time step = 0.0078125").replace(
    "time max = 0.0185", "time max = 1.01").replace("output time step = 0.0075", "output time step = 0.5").replace(
    "Lx = 3.1415926535897932384626433832795", "Lx = 6.283185307179586").replace(
    "Ly = 3.1415926535897932384626433832795", "Ly = 6.283185307179586")


def read_kdump(path, nkx, nky):
    """k_data_bef / k_data_aft (file_access.cu:490-625): omega, phi, rho, each
    real then imaginary rows of nky values."""
    rows = [l.split() for l in open(path).read().splitlines() if l.strip() and not l.startswith("/")]
    a = np.array(rows, dtype=np.float64)
    assert a.shape == (6 * nkx, nky)
    a = a.reshape(3, 2, nkx, nky)
    return {f: a[i, 0] + 1j * a[i, 1] for i, f in enumerate(("omega", "phi", "rho"))}


@pytest.fixture(scope="module", params=[8, 4])
def run_remap(request, gpu, tmp_path_factory):
    prec = request.param
    d = tmp_path_factory.mktemp(f"shrt_remap{prec}")
    with open(d / "config.ini", "w") as fh:
        fh.write(CONFIG_REMAP)
    print(f"remap driver run prec={prec} in", d, flush=True)
    p = subprocess.run([BIN] + (["--dble"] if prec == 8 else []), cwd=d, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "not the last of its batch" not in p.stderr
    return prec, d


def test_remap_dumps_match_oracle(run_remap):
    """sigma = 1, Lx = 2 pi, delt = 2^-7: the remap fires on the 128th step (t = 1).
    The driver ends a batch there and writes befk (the fields before
    shearing_field) and aftk (after), named with the step and time after the step
    (shear.cu:236, 270), besides the befk of istep 0 and 1."""
    prec, d = run_remap
    names = set(os.listdir(d))
    assert sorted(n for n in names if n.startswith("befk")) == [
        "befk_n00000_t00.000000.dat", "befk_n00001_t00.007812.dat", "befk_n00128_t01.000000.dat"]
    assert sorted(n for n in names if n.startswith("aftk")) == ["aftk_n00128_t01.000000.dat"]
    rd = (lambda x: x) if prec == 8 else f32
    o = oracle.Oracle(nx=32, ny=32, Lx=rd(2 * np.pi), Ly=rd(2 * np.pi), delt=0.0078125, cfl_num=0.5,
                      nu=rd(1e-3), kappa=rd(8e-5), sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
                      rho_eps1=rd(1e-2), noise=1, linear=0)
    o.initialize_reference()
    o.step(128)
    st = o.state()
    assert st["remap_count"] == 1 and st["last_remap_step"] == 127
    nkx, nky = o.shape_k
    tol = 1e-9 if prec == 8 else 1e-5
    bef = read_kdump(d / "befk_n00128_t01.000000.dat", nkx, nky)
    aft = read_kdump(d / "aftk_n00128_t01.000000.dat", nkx, nky)
    for f, ring in (("omega", "aomg"), ("phi", "aphi"), ("rho", "arho")):
        assert rel_l2(bef[f], o.kdata_bef(f)) < tol, ("befk", f)
        assert rel_l2(aft[f], o.buffer(ring)) < tol, ("aftk", f)
    # the output after the remap (step 129, t = 1.0078125) in the new frame
    o.step(1)
    grid = read_table(d / "n00129_t01.007812.dat", 5).reshape(33, 33, 5)
    for col, ring in ((2, "aomg"), (4, "arho")):
        assert rel_l2(grid[:32, :32, col], o.ktox_shear(o.buffer(ring))) < tol, ring
