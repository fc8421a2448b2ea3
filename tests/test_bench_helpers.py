"""Host-side helpers of bench.py that turn rocprofv3 CSV output into the
summaries kept with the profiles: dispatch overlap and per-kernel tables."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_concurrency_known_answer():
    # [0,10) and [5,15) overlap for 5 ns; [20,30) runs alone after a 5 ns gap
    c = bench.dispatch_concurrency([(0, 10), (5, 15), (20, 30)])
    assert c["dispatches"] == 3
    assert c["window_ms"] == pytest.approx(30e-6)
    assert c["busy_ms"] == pytest.approx(25e-6)
    assert c["mean_running"] == pytest.approx(30 / 30)
    assert c["mean_running_busy"] == pytest.approx(30 / 25)
    assert c["max_running"] == 2


def test_concurrency_touching_spans_do_not_overlap():
    # one dispatch ending where the next starts is back to back, not concurrent
    c = bench.dispatch_concurrency([(0, 10), (10, 20), (20, 30)])
    assert c["max_running"] == 1
    assert c["mean_running"] == pytest.approx(1.0)


def test_concurrency_nested_and_identical():
    c = bench.dispatch_concurrency([(0, 100), (10, 20), (10, 20), (50, 60)])
    assert c["max_running"] == 3
    # area = 100 + 10 + 10 + 10 over a window of 100
    assert c["mean_running"] == pytest.approx(1.3)
    assert c["busy_ms"] == pytest.approx(100e-6)


def test_summarize_trace_and_counters(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    (d / "x_kernel_trace.csv").write_text(
        "Kernel_Name,Grid_Size_X,Start_Timestamp,End_Timestamp\n"
        "wave_worm_kernel(a),32768,0,2000000\n"
        "wave_worm_kernel(a),32768,1000000,4000000\n"
        "other_kernel(b),64,0,5\n")
    (d / "x_counter_collection.csv").write_text(
        "Kernel_Name,Grid_Size,Dispatch_Id,Counter_Name,Counter_Value\n"
        "wave_worm_kernel(a),32768,1,SQ_WAVES,256\n"
        "wave_worm_kernel(a),32768,1,SQ_WAVES,256\n"
        "wave_worm_kernel(a),32768,2,SQ_WAVES,1024\n")
    out = bench.summarize_rocprof(str(d), last=1, regex="wave_")
    assert "other_kernel" not in out
    # durations 2 ms and 3 ms; the last one alone is 3 ms
    assert "| wave_worm_kernel | 32768 | 2 | 2.5000 | 2.0000 | 3.0000 | 3.0000 (1) |" in out
    # last 1 dispatch: one 3 ms span
    assert "concurrency over the last 1 dispatches: window 3.000 ms" in out
    # counter rows of one dispatch are summed (two XCD rows of 256), last 1 = dispatch 2
    assert "| wave_worm_kernel | 32768 | SQ_WAVES | 2 | 1024 (1) |" in out
    out_all = bench.summarize_rocprof(str(d), last=0, regex="wave_")
    assert "| wave_worm_kernel | 32768 | SQ_WAVES | 2 | 768 (2) |" in out_all
    assert "window 4.000 ms, busy 4.000 ms, mean dispatches running 1.25" in out_all


def test_sweep_budgets_known_answer():
    # costliest walker 4 sweeps; a walker 8x cheaper 32; the cap holds at 400
    b = bench.sweep_budgets([100.0, 12.5, 0.01, 50.0])
    assert b.tolist() == [4, 32, 400, 8]
    assert b.dtype.name == "int64"


class _FakeBatch:
    def __init__(self, ns):
        self.ns = ns

    def get_walker(self, k, with_ops=True, with_spins=True):
        assert not with_ops and not with_spins
        return {"num_operators": self.ns[k]}, None, None


def test_operator_count_costs_layout_weight():
    # 32 KiB of worm LDS holds 8-byte records up to n = 4096; beyond, x3
    ns = [0, 1000, 4096, 4097, 20000]
    cost, n, in_lds = bench.operator_count_costs(_FakeBatch(ns), len(ns), 32768)
    assert in_lds.tolist() == [True, True, True, False, False]
    assert cost.tolist() == [1.0, 1000.0, 4096.0, 3 * 4097.0, 60000.0]
    # a smaller worm LDS moves the limit to 8n <= bytes
    _, _, in_lds = bench.operator_count_costs(_FakeBatch(ns), len(ns), 16384)
    assert in_lds.tolist() == [True, True, False, False, False]
