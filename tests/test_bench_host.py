"""Host logic of bench.py, on the CPU: the algorithmic bytes per launch, the
rocprofv3 counter summary (calibration by the plain copy, kernel-name
matching, per-dispatch averaging) and the one-GPU guard."""
import importlib.util
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
bench = importlib.util.module_from_spec(spec)
spec.loader.exec_module(bench)


def test_kernel_bytes_at_2048():
    # the 2048^2 sizes (M = nkx*nky, I = nkxh*ny), FP32 complex = 8 bytes
    s = {"M": 1365 * 683, "I": 683 * 2048}
    kb = bench.kernel_bytes(s, 8)
    assert kb["p1"] == 8 * (2 * s["M"] + 6 * s["I"]) == 82_058_352
    assert kb["p2"] == 8 * 8 * s["I"] == 89_522_176
    assert kb["p3"] == 8 * (2 * s["I"] + 12 * s["M"])
    assert sum(kb.values()) == 283_461_392   # the bench line's bytes_per_step


def _csv(path, rows):
    with open(path, "w") as f:
        f.write("Kernel_Name,Counter_Value\n")
        for name, v in rows:
            f.write(f"\"{name}\",{v}\n")


def test_pmc_summary(tmp_path):
    cal = 1 << 20
    # the copy reads FETCH_SIZE = half its bytes and WRITE_SIZE = all of them (KiB)
    _csv(tmp_path / "cf.csv", [("void shrt::k_copy16<float>(x)", cal / 2 / 1024)])
    _csv(tmp_path / "cw.csv", [("void shrt::k_copy16<float>(x)", cal / 1024)])
    p1 = "void shrt::k_p1_pair<float, 2048, false>(shrt::StepArgs<float>, shrt::StepArgs<float> const*)"
    p2 = "void shrt::k_p2<float, 2048, false, true>(shrt::StepArgs<float>, shrt::StepArgs<float> const*)"
    p3 = "void shrt::k_p3<double, 1024, false, 1, false>(shrt::StepArgs<double>, shrt::StepArgs<double> const*)"
    _csv(tmp_path / "f.csv", [(p1, 100.0), (p1, 300.0), (p2, 50.0), (p3, 10.0), ("other", 1.0)])
    _csv(tmp_path / "w.csv", [(p1, 400.0), (p1, 400.0), (p2, 25.0), (p3, 5.0)])
    out = bench.pmc_summary(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                            str(tmp_path / "cw.csv"), cal)
    assert out["calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert out["calibration"]["write_factor"] == pytest.approx(1.0)
    # averages over dispatches, then 2 x fetch + 1 x write, KiB -> bytes
    assert out["p1_2048_fp32"]["bytes"] == pytest.approx((2 * 200.0 + 400.0) * 1024)
    assert out["p2_2048_fp32"]["read_bytes"] == pytest.approx(2 * 50.0 * 1024)
    assert out["p3_1024_fp64"]["write_bytes"] == pytest.approx(5.0 * 1024)
    assert set(out) == {"calibration", "p1_2048_fp32", "p2_2048_fp32", "p3_1024_fp64"}


def test_fused_kernel_bytes_and_counter_key(tmp_path):
    # P3 of step n with P1 of step n+1: P3's bytes plus P1's writes, no state read
    s = {"M": 1365 * 683, "I": 683 * 2048}
    kb = bench.kernel_bytes(s, 8)
    assert bench.fused_bytes(s, 8) == kb["p3"] + 8 * 6 * s["I"] == kb["p3"] + kb["p1"] - 8 * 2 * s["M"]
    cal = 1 << 20
    _csv(tmp_path / "cf.csv", [("void shrt::k_copy16<float>(x)", cal / 2 / 1024)])
    _csv(tmp_path / "cw.csv", [("void shrt::k_copy16<float>(x)", cal / 1024)])
    p3 = "void shrt::k_p3<float, 2048, false, 1, false, false, false>(shrt::StepArgs<float>, a)"
    p31 = "void shrt::k_p3<float, 2048, true, 1, true, false, false>(shrt::StepArgs<float>, a)"
    _csv(tmp_path / "f.csv", [(p3, 10.0), (p31, 30.0)])
    _csv(tmp_path / "w.csv", [(p3, 5.0), (p31, 20.0)])
    out = bench.pmc_summary(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                            str(tmp_path / "cw.csv"), cal)
    assert out["p3_2048_fp32"]["kernel"] == p3 and out["p31_2048_fp32"]["kernel"] == p31
    assert out["p31_2048_fp32"]["bytes"] == pytest.approx((2 * 30.0 + 20.0) * 1024)


def test_step_bytes_of_a_batch():
    # a fused batch of G steps launches P1 and P3 once, P2 G times, the fused kernel G - 1 times
    s = {"M": 1365 * 683, "I": 683 * 2048}
    kb = bench.kernel_bytes(s, 8)
    assert bench.step_bytes(s, 8, False, 100) == sum(kb.values()) == 283_461_392
    assert bench.step_bytes(s, 8, True, 1) == sum(kb.values())
    G = 100
    total = kb["p1"] + kb["p3"] + G * kb["p2"] + (G - 1) * bench.fused_bytes(s, 8)
    assert bench.step_bytes(s, 8, True, G) == pytest.approx(total / G)
    assert bench.step_bytes(s, 8, True, G) == pytest.approx(268_693_839.2)


def test_pmc_summary_needs_the_copy(tmp_path):
    _csv(tmp_path / "x.csv", [("void shrt::k_p2<float, 2048, false, true>(a)", 1.0)])
    with pytest.raises(RuntimeError):
        bench.pmc_summary(*(str(tmp_path / "x.csv") for _ in range(4)), 1024)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
