"""The counter report's host logic, on the CPU: which dispatches count, and
that the lines have the reference's format (unsteady.c:537, 541-545, 558),
so the reference's own log parsing (results/log.sh: grep AE, strip "AE: ")
reads them."""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "memory-aware-lbm_amd"))

from lbm_amd import counters  # noqa: E402


def test_sum_counters_keeps_sweep_kernels_only(tmp_path):
    p = tmp_path / "run_counter_collection.csv"
    rows = [
        ("k_init(lbm::Problem, double*)", "TCC_MISS_sum", 1000),
        ("void lbm::k_wpanel_in<double, 2, 4>(lbm::Problem)", "TCC_MISS_sum", 7),
        ("void lbm::k_wpanel_edge<double>(lbm::Problem)", "TCC_MISS_sum", 5),
        ("void lbm::k_wpanel_edge<double>(lbm::Problem)", "TCP_TCC_READ_REQ_sum", 11),
        ("lbm::k_bump(unsigned int*, unsigned int)", "TCP_TCC_READ_REQ_sum", 1),
    ]
    with open(p, "w") as fh:
        fh.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n")
        for i, (k, c, v) in enumerate(rows):
            fh.write(f'{i},"{k}",{c},{v}\n')
    assert counters.sum_counters(str(p)) == {"TCC_MISS_sum": 12.0, "TCP_TCC_READ_REQ_sum": 12.0}


def test_report_lines_have_the_reference_format():
    totals = {"TCP_TCC_READ_REQ_sum": 700.0, "TCP_TCC_WRITE_REQ_sum": 300.0,
              "TCC_MISS_sum": 512.0, "TCC_EA0_RDREQ_DRAM_sum": 100.0,
              "TCC_EA0_WRREQ_DRAM_sum": 28.0, "TCP_UTCL1_TRANSLATION_MISS_sum": 3.0}
    lines = counters.report(totals, units=256, steps=20, seconds=1.25)
    assert lines[:4] == ["Global_E[00]: 1000", "Global_E[01]: 512",
                         "Global_E[02]: 128", "Global_E[03]: 3"]
    ae = [l for l in lines if "AE" in l]
    assert len(ae) == 1
    # integer division per event, as the reference's global_CM[i]/thread_count
    assert ae[0].replace("AE: ", "").split() == ["3", "2", "0", "0"]
    assert re.fullmatch(r"time eclapsed in 20 steps, t_total=1\.250, t_collid=\d+\.\d{3}, "
                        r"t_stream=\d+\.\d{3}, t_zgb=\d+\.\d{3}", lines[-1])


def test_passes_stay_within_one_run_limits():
    # at most 4 TCP and 4 TCC hardware counters in one rocprofv3 pass
    # (FETCH-style derived counters excepted, none are used here), and every
    # event's counters are collected by some pass
    for p in counters.PASSES:
        assert sum(c.startswith("TCP_") for c in p) <= 4
        assert sum(c.startswith("TCC_") for c in p) <= 4
    collected = {c for p in counters.PASSES for c in p}
    assert {c for ev in counters.EVENTS for c in ev} <= collected


def test_batch_launches_are_time_overlap_groups(tmp_path):
    """The profile report's batch launches: dispatches that overlap in time
    form one launch (whatever their submission order: the deep launch
    enqueues its interior before its frame), a launch has the steps of its
    deepest interior dispatch, and the next launch starts after the last one
    ended."""
    from lbm_amd import profile_report as pr
    ms = 1_000_000
    rows = [  # (dispatch id, kernel, start, end) in ns
        (1, "void lbm::k_batch_win<double, 2, 6>(lbm::ChannelWaves const*)", 0, 6 * ms),
        (2, "void lbm::k_batch_frame<double, 16>(lbm::ChannelWaves const*)", 10, 1 * ms),
        (3, "void lbm::k_batch_frame<double, 8>(lbm::ChannelWaves const*)", 1 * ms, 2 * ms),
        (4, "void lbm::k_batch_wedge<double, 4>(lbm::ChannelWaves const*)", 2 * ms, 3 * ms),
        (5, "void lbm::k_batch_win<double, 2, 2>(lbm::ChannelWaves const*)", 3 * ms, 4 * ms),
        (6, "lbm::k_bump(unsigned int*, unsigned int)", 6 * ms, 6 * ms + 10),
        (7, "void lbm::k_batch_win<double, 2, 6>(lbm::ChannelWaves const*)", 7 * ms, 12 * ms),
        (8, "void lbm::k_batch_frame<double, 16>(lbm::ChannelWaves const*)", 7 * ms, 13 * ms),
        (9, "void lbm::k_batch_wedge<double, 4>(lbm::ChannelWaves const*)", 14 * ms, 15 * ms),
        (10, "void lbm::k_batch_win<double, 2, 2>(lbm::ChannelWaves const*)", 14 * ms, 17 * ms),
    ]
    with open(tmp_path / "run_kernel_trace.csv", "w") as fh:
        fh.write("Dispatch_Id,Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for d, k, t0, t1 in rows:
            fh.write(f'{d},"{k}",{t0},{t1}\n')
    got = pr.batch_marches(str(tmp_path))
    assert sorted(got) == [("double", 2), ("double", 6)]
    assert [round(v, 6) for v in got[("double", 6)]] == [6.0, 6.0]
    assert [round(v, 6) for v in got[("double", 2)]] == [3.0]


def test_panel_launch_traffic_sums_frame_passes():
    """Whole-launch traffic of the panel sweeps: every perimeter dispatch
    (k_wpanel_edge for two steps, k_frame for the deep sweeps' frame passes)
    belongs to the next interior dispatch; other kernels start afresh."""
    from lbm_amd import profile_report as pr
    disp = [(1, "k_frame<double, 16>", 1.0), (2, "k_frame<double, 16>", 2.0),
            (3, "k_frame<double, 8>", 3.0), (4, "k_wpanel_in<double, 2, 6>", 100.0),
            (5, "k_bump", 0.5), (6, "k_wpanel_edge<float, 4>", 4.0),
            (7, "k_wpanel_in<float, 2, 2>", 50.0), (8, "k_frame<double, 16>", 5.0),
            (9, "k_collide<double, true, 2>", 7.0), (10, "k_wpanel_in<double, 1, 8>", 90.0)]
    got = pr.groups(disp)
    assert got == {"panel6_launch<double>": [106.0], "panel_launch<float>": [54.0],
                   "panel8_launch<double>": [90.0]}
