"""The bench's counter arithmetic on synthetic rocprofv3 outputs (CPU only).

``bench.py --counters-from DIR`` turns the CSVs of the PMC passes into the
per-kernel figures ``roofline_for`` divides by live HIP-event times.  These
tests write CSVs with chosen counter values for one kernel and check every
derived figure against its definition: the measured clock (GRBM_GUI_ACTIVE
summed over 8 XCDs), TA / TD / LDS busy shares per CU-cycle, the bytes the
vector-memory path moves (TA busy cycles x 64 B), HBM bytes from FETCH_SIZE
and WRITE_SIZE calibrated by a 512 MiB device copy, the LDS bank-conflict
share, rocprofv3's VALUBusy and VALUUtilization, and the roofline fraction
against 256 CUs x 64 B/clk x 2.4 GHz.
"""
import csv
import json
import os

import pytest

import bench

KERNEL = "void dd::(anonymous namespace)::k_probe<1>(dd::Geo2, float const*)"
COPY = "__amd_rocclr_copyBuffer"
MS = 1.0                  # dispatch time in the PMC runs
CYC = 2.0e6               # cycles per XCD in that time: a 2.0 GHz clock
CUS = 256


def _write(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value",
                    "Start_Timestamp", "End_Timestamp"])
        for disp, name, counter, value, ms in rows:
            w.writerow([disp, name, counter, value, 0, int(ms * 1e6)])


def _pass(base, name, counters, copy=None):
    rows = []
    for disp in (1, 2):  # two dispatches with the same values: means are the values
        rows += [(disp, KERNEL, c, v, MS) for c, v in counters.items()]
    if copy:
        rows += [(3, COPY, c, v, 0.07) for c, v in copy.items()]
    _write(os.path.join(base, name, "run_counter_collection.csv"), rows)


@pytest.fixture()
def counters(tmp_path, monkeypatch):
    base = str(tmp_path / "pmc")
    _pass(base, "fanarc_A", {"GRBM_GUI_ACTIVE": 8 * CYC, "TA_TA_BUSY_sum": CUS * 0.5 * CYC,
                             "SQ_INSTS_VMEM_RD": 1.0e6, "SQ_INSTS_VALU": 4.0e6,
                             "SQ_WAVES": 1.0e4})
    _pass(base, "fanarc_B", {"TD_TD_BUSY_sum": CUS * 0.75 * CYC,
                             "TCP_TOTAL_CACHE_ACCESSES_sum": 3.0e6})
    # the copy of 512 MiB: FETCH_SIZE reads half the bytes it fetches (scale 2),
    # WRITE_SIZE all it writes (scale 1)
    _pass(base, "fanarc_FETCH", {"FETCH_SIZE": 1000.0}, copy={"FETCH_SIZE": 262144.0})
    _pass(base, "fanarc_WRITE", {"WRITE_SIZE": 500.0}, copy={"WRITE_SIZE": 524288.0})
    _pass(base, "fanarc_LDS", {"GRBM_GUI_ACTIVE": 8 * CYC, "SQ_LDS_IDX_ACTIVE": CUS * 0.4 * CYC,
                               "SQ_LDS_BANK_CONFLICT": 0.25 * CUS * 0.4 * CYC,
                               "SQ_INSTS_LDS": 2.0e6})
    act = CUS * 0.6 * CYC
    _pass(base, "fanarc_VALU", {"GRBM_GUI_ACTIVE": 8 * CYC, "SQ_ACTIVE_INST_VALU": act,
                                "SQ_THREAD_CYCLES_VALU": 0.9 * 64 * act, "SQ_INSTS_VALU": 1000.0,
                                "SQ_INSTS_VALU_ADD_F64": 100.0, "SQ_INSTS_VALU_FMA_F64": 50.0})
    monkeypatch.setattr(bench, "COUNTER_KEYS", {"fanarc": {"probe:k_probe<1>": ["k_probe<1>"]}})
    out = str(tmp_path / "kernel_counters.json")
    bench.summarize_counters(base, out)
    with open(out) as f:
        data = json.load(f)
    monkeypatch.setattr(bench, "COUNTERS", out)
    return data


def test_derived_figures(counters):
    cal = counters["calibration"]
    assert cal["fetch_scale"] == pytest.approx(2.0) and cal["write_scale"] == pytest.approx(1.0)
    e = counters["kernels"]["probe:k_probe<1>"]
    assert e["kernels"] == ["k_probe<1>"]
    assert e["pmc_ms"] == pytest.approx(MS)
    assert e["measured_clock_hz"] == pytest.approx(CYC / (MS * 1e-3))
    assert e["ta_busy_frac"] == pytest.approx(0.5)
    assert e["td_busy_frac"] == pytest.approx(0.75)
    assert e["vmem_bytes"] == pytest.approx(CUS * 0.5 * CYC * 64)
    assert e["l1_sector_bytes"] == pytest.approx(3.0e6 * 32)
    assert e["hbm_bytes"] == pytest.approx((1000.0 * 2 + 500.0) * 1024)
    assert e["lds_busy_frac"] == pytest.approx(0.4)
    assert e["lds_bank_conflict_frac"] == pytest.approx(0.25)
    assert e["lds_instructions"] == pytest.approx(2.0e6)
    assert e["valu_busy"] == pytest.approx(0.6)
    assert e["valu_lanes_active"] == pytest.approx(0.9)
    assert e["valu_f64_share"] == pytest.approx(0.15)


def test_roofline_divides_by_the_live_time(counters):
    live_ms = 0.8
    r = bench.roofline_for("probe:k_probe<1>", "k_probe", live_ms, b_alg=4.0e9, step_ms=2.0)
    vmem = CUS * 0.5 * CYC * 64
    peak = CUS * 64 * 2.4e9
    assert r["bound"] == "vmem" and r["unit"] == "GB/s"
    assert r["achieved"] == pytest.approx(vmem / (live_ms * 1e-3) / 1e9)
    assert r["peak"] == pytest.approx(peak / 1e9)
    assert r["frac"] == pytest.approx(vmem / (live_ms * 1e-3) / peak)
    assert r["traffic"] == pytest.approx((1000.0 * 2 + 500.0) * 1024)
    assert r["hbm"]["frac"] == pytest.approx(r["traffic"] / (live_ms * 1e-3) / 8e12)
    assert r["effective_GBps"] == pytest.approx(4.0e9 / (live_ms * 1e-3) / 1e9)
    assert r["step_effective_frac_of_hbm_peak"] == pytest.approx(2 * 4.0e9 / 2.0e-3 / 8e12)
    busy = r["busy_in_counter_run"]
    assert busy["ta"] == pytest.approx(0.5) and busy["td"] == pytest.approx(0.75)
    assert busy["lds"] == pytest.approx(0.4) and busy["lds_bank_conflict"] == pytest.approx(0.25)
    assert busy["valu"] == pytest.approx(0.6) and busy["valu_lanes_active"] == pytest.approx(0.9)


def test_missing_optional_passes_and_kernels(tmp_path, monkeypatch):
    """Without LDS / VALU passes the figures are simply absent; a key whose
    kernel a run did not launch is skipped."""
    base = str(tmp_path / "pmc")
    _pass(base, "fanarc_A", {"GRBM_GUI_ACTIVE": 8 * CYC, "TA_TA_BUSY_sum": CUS * CYC,
                             "SQ_INSTS_VMEM_RD": 1.0, "SQ_INSTS_VALU": 1.0, "SQ_WAVES": 1.0})
    _pass(base, "fanarc_B", {"TD_TD_BUSY_sum": CUS * CYC, "TCP_TOTAL_CACHE_ACCESSES_sum": 1.0})
    _pass(base, "fanarc_FETCH", {"FETCH_SIZE": 1.0}, copy={"FETCH_SIZE": 262144.0})
    _pass(base, "fanarc_WRITE", {"WRITE_SIZE": 1.0}, copy={"WRITE_SIZE": 524288.0})
    monkeypatch.setattr(bench, "COUNTER_KEYS", {"fanarc": {"probe:k_probe<1>": ["k_probe<1>"],
                                                           "absent:k_none": ["k_none<1>"]}})
    out = str(tmp_path / "kernel_counters.json")
    bench.summarize_counters(base, out)
    with open(out) as f:
        ks = json.load(f)["kernels"]
    assert set(ks) == {"probe:k_probe<1>"}
    e = ks["probe:k_probe<1>"]
    assert e["ta_busy_frac"] == pytest.approx(1.0)
    for k in ("lds_busy_frac", "valu_busy", "valu_lanes_active"):
        assert k not in e
