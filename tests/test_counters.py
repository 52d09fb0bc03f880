"""The GPU counterparts of the reference's PAPI event sets
(simplemoc_amd/counters.py, src/papi.c:19-134): every set's passes fit
one rocprofv3 counter run, the counter CSV is read per kernel, and the
summaries compute the reference's figures (src/papi.c:457-476) from
known totals."""
import pytest

from simplemoc_amd import counters as ct


def test_every_set_fits_one_run_per_pass():
    ct.check_sets()
    for sid, (_, _, passes) in ct.EVENT_SETS.items():
        for p in passes:
            cost = ct.pass_cost([c for c, _ in p])
            for blk, n in cost.items():
                assert n <= ct.BLOCK_LIMITS[blk], (sid, blk, n)


def test_pass_cost_counts_derived_and_instances():
    assert ct.pass_cost(["FETCH_SIZE", "WRITE_SIZE"]) == {"TCC": 5}
    assert ct.pass_cost(["TCC_HIT_sum", "TCC_HIT", "TCC_MISS_sum"]) == {"TCC": 2}
    assert ct.pass_cost(["SQ_WAVES", "GRBM_GUI_ACTIVE"]) == {"SQ": 1, "GRBM": 1}


def test_a_set_over_the_limits_is_refused(monkeypatch):
    sets = dict(ct.EVENT_SETS)
    sets[99] = ("bad", "-", [[("FETCH_SIZE", ""), ("WRITE_SIZE", "")]])
    monkeypatch.setattr(ct, "EVENT_SETS", sets)
    with pytest.raises(ValueError, match="TCC"):
        ct.check_sets()


def write_csv(path, rows):
    cols = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
            "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
            "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
            "Counter_Value", "Start_Timestamp", "End_Timestamp"]
    with open(path, "w") as f:
        f.write(",".join(f'"{c}"' for c in cols) + "\n")
        for d, k, c, v, t0, t1 in rows:
            f.write(f'{d},{d},"Agent 1",1,1,1,256,1,"{k}",256,0,0,8,0,32,"{c}",{v},{t0},{t1}\n')


def test_counter_csv_is_totalled_per_kernel(tmp_path):
    p = tmp_path / "run_counter_collection.csv"
    att = "k_attenuate<64, 2, true, false, 0, 4>(moc::AttArgs)"
    write_csv(p, [(1, att, "FETCH_SIZE", 100.0, 0, 1000), (1, att, "GRBM_GUI_ACTIVE", 5.0, 0, 1000),
                  (2, att, "FETCH_SIZE", 300.0, 2000, 5000), (2, att, "GRBM_GUI_ACTIVE", 7.0, 2000, 5000),
                  (3, "k_prep(float const*)", "FETCH_SIZE", 1.0, 0, 10)])
    t = ct.read_counters(p)
    assert t["k_attenuate"]["dispatches"] == 2
    assert t["k_attenuate"]["ns"] == 4000
    assert t["k_attenuate"]["FETCH_SIZE"] == 400.0
    assert t["k_prep"]["dispatches"] == 1


def test_summaries_from_known_totals():
    # 2 GB fetched (as FETCH_SIZE counts it: half) + 1 GB written in 0.5 s
    v = {"ns": 5e8, "FETCH_SIZE": 1e9 / 1024, "WRITE_SIZE": 1e9 / 1024}
    assert ct.summary(1, v) == [f"Bandwidth: {(2e9 * 1 + 1e9) / 0.5 / 1e9:.3f} (GB/s)"]
    v = {"ns": 1e9, "SQ_INSTS_VALU_FLOPS_FP32": 3e12, "SQ_INSTS_VALU_FLOPS_FP32_TRANS": 1e12}
    assert ct.summary(0, v) == [f"GFLOPs: {4000.0 * 64:.3f} (all 64 lanes active)"]
    v = {"SQ_WAIT_ANY": 25.0, "SQ_WAVE_CYCLES": 100.0}
    assert ct.summary(3, v) == ["GPU Stalled Wave Cycles: 25.00%"]
    v = {"TCC_MISS_sum": 1.0, "TCC_REQ_sum": 4.0}
    assert ct.summary(5, v) == ["L2 miss rate: 25.00%"]
    v = {"TCP_TOTAL_CACHE_ACCESSES_sum": 10.0, "TCP_TCC_READ_REQ_sum": 4.0, "TCC_HIT_sum": 3.0,
         "TCC_MISS_sum": 1.0}
    assert ct.summary(4, v)[0].endswith("60.00%") and ct.summary(4, v)[1].endswith("75.00%")


def test_report_uses_the_reference_layout():
    t = {"k_attenuate": {"dispatches": 2, "ns": 4e6, "SQ_WAIT_ANY": 5.0, "SQ_WAVE_CYCLES": 10.0},
         "k_prep": {"dispatches": 1, "ns": 1.0}}
    r = ct.report(3, t)
    assert "Count          \tSymbol      \tDescription" in r
    assert "k_attenuate: 2 dispatches, 4.000 ms" in r
    assert "GPU Stalled Wave Cycles: 50.00%" in r
    assert "k_prep" not in r


def test_main_needs_a_program():
    with pytest.raises(SystemExit):
        ct.main(["--set", "0"])
