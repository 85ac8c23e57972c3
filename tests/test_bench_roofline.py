"""bench.py's roofline arithmetic on the CPU: the per-launch work of a kernel family does not depend on
how many evaluations the launch counts were summed over, and the fraction is algorithmic work over
time over the published peak."""
import importlib.util
import os
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


T = types.SimpleNamespace(nts=100, nw_b=99)
TUNING = {"p0_split_tile": 2, "sigma_split_tile": 2}


def _stats(steps):
    # launches and milliseconds of one evaluation, summed over `steps` identical evaluations
    one = {"P0 GEMM": (64, 450.0), "Sigma GEMM": (64, 780.0), "P0 prep": (64, 550.0),
           "Sigma epilogue": (64, 855.0), "Dyson inverse": (1, 66.0)}
    return {k: [n * steps, ms * steps] for k, (n, ms) in one.items()}


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_roofline_independent_of_step_count(precision):
    b = _bench()
    r1 = b.roofline(_stats(1), precision, b.WORKLOAD, T, 4, TUNING, 1)
    r3 = b.roofline(_stats(3), precision, b.WORKLOAD, T, 4, TUNING, 3)
    assert r1["stage"] == r3["stage"] == "Sigma epilogue"
    for name in r1["all_stages"]:
        a, c = r1["all_stages"][name], r3["all_stages"][name]
        for key in ("achieved", "frac", "launches", "algorithmic_flops_per_launch", "algorithmic_bytes"):
            assert a[key] == pytest.approx(c[key], rel=1e-12), (name, key)


def test_roofline_fraction_is_algorithmic_work_over_peak():
    b = _bench()
    r = b.roofline(_stats(2), "split", b.WORKLOAD, T, 4, TUNING, 2)
    w = b.WORKLOAD
    # Sigma GEMM: 8 NQ^2 nk nao^2 (nts/2) flop per launch (one q, the half tau grid)
    s = r["all_stages"]["Sigma GEMM"]
    per_launch = 8.0 * w["NQ"] ** 2 * w["nk"] * w["nao"] ** 2 * (T.nts // 2)
    assert s["algorithmic_flops_per_launch"] == pytest.approx(per_launch)
    assert s["achieved"] == pytest.approx(per_launch / (780.0 / 64 * 1e-3) / 1e12)
    assert s["frac"] == pytest.approx(s["achieved"] / b.PEAK_INT8_TOPS)
    assert s["kernel"].startswith("zgemm_ozaki_pipe_kernel<4>")
    # Sigma epilogue: 16 NQ nao^3 nk nts flop per launch, FP64 peak
    e = r["all_stages"]["Sigma epilogue"]
    assert e["algorithmic_flops_per_launch"] == pytest.approx(16.0 * w["NQ"] * w["nao"] ** 3 * w["nk"] * T.nts)
    assert e["frac"] == pytest.approx(e["achieved"] / b.PEAK_FP64_TFLOPS)
    assert r["traffic"] == pytest.approx(e["traffic"]) and r["traffic"] > 0


def test_epilogue_label_and_traffic_follow_the_tuned_form():
    b = _bench()
    r1 = b.roofline(_stats(1), "split", b.WORKLOAD, T, 4, dict(TUNING, epi_form=0), 1)["all_stages"]["Sigma epilogue"]
    r3 = b.roofline(_stats(1), "split", b.WORKLOAD, T, 4, dict(TUNING, epi_form=3), 1)["all_stages"]["Sigma epilogue"]
    assert r1["kernel"].startswith("sigma_epi4_kernel<7>") and r1["traffic"] > 0
    assert r3["kernel"].startswith("sigma_epi_mfma_kernel<2,double>") and r3["traffic"] is None
    assert r1["achieved"] == pytest.approx(r3["achieved"])
