"""rocprofv3 CSV readers of dragon_amd.profiling on small synthetic files."""
import csv
import json
import os

from dragon_amd import profiling as P


def write_csv(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_kernel_class_names():
    assert P.kernel_class("void dg::k_sweep_march<3, true, double, 1, false>(dg::SweepArgs)") == "pred_sweep_y"
    assert P.kernel_class("void dg::k_sweep_wave<3, true, float, 2, true>(dg::SweepArgs)") == "flux_sweep_z"
    assert P.kernel_class("void dg::k_emf3<true>(dg::Soa, ...)") == "emf_pred"
    assert P.kernel_class("dg::p64::k_emf3_tiled(dg::SoaT<double>, ...)") == "emf (predictor + final)"
    assert P.kernel_class("void dg::k_apply<3, true>(...)") == "apply"
    assert P.kernel_class("dg::k_bc_cells(dg::BCArgs)") == "boundary"


def test_traffic_scaled_by_copy_probe(tmp_path):
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    mul = "void at::native::vectorized_elementwise_kernel<4, at::native::neg_kernel_cuda(at::TensorIteratorBase&)::{lambda()#1}::operator()() const::{lambda(double)#1}>"
    for cnt, d in (("FETCH_SIZE", "f"), ("WRITE_SIZE", "w")):
        rows = [(1, mul, cnt, 2 * (1 << 20)), (2, mul, cnt, 2 * (1 << 20)),  # counts 2x the known 1 GiB
                (3, "void dg::k_init_B<3>(...)", cnt, 10.0),
                (4, "void dg::k_apply<3, true>(...)", cnt, 100.0), (5, "void dg::k_apply<3, true>(...)", cnt, 100.0),
                (6, "void dg::k_init_B<3>(...)", cnt, 10.0),
                (7, "void dg::k_apply<3, true>(...)", cnt, 60.0), (8, "void dg::k_divb<3>(...)", cnt, 5.0)]
        write_csv(str(tmp_path / d / "x" / "p_counter_collection.csv"), hdr, rows)
    out = P.traffic(str(tmp_path / "f"), str(tmp_path / "w"), steps=2, forms=["fp64", "mixed"])
    assert abs(out["copy_probe"]["fetch_factor"] - 0.5) < 1e-12
    # fp64: 200 KiB counted over 2 steps, x0.5 -> 50 KiB per step each way
    assert abs(out["fp64"]["fetch_bytes_per_step"] - 50 * 1024) < 1e-6
    assert abs(out["fp64"]["bytes_per_step"] - 100 * 1024) < 1e-6
    assert abs(out["mixed"]["write_bytes_per_step"] - 15 * 1024) < 1e-6
    json.dumps(out)


def test_stats_table(tmp_path):
    write_csv(str(tmp_path / "kt" / "a" / "kt_kernel_stats.csv"),
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"],
              [("void dg::k_apply<3, true>(...)", 4, 4e6, 1e6, 60.0, 9e5, 1.1e6, 1.0)])
    t = P.stats_table(str(tmp_path / "kt"), "rocprofv3 ... -- python3 bench.py")
    assert "| apply | 4 | 1000.0 | 4.00 | 60.00 |" in t


def test_valu_table(tmp_path):
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]
    k = "void dg::p32::k_sweep_wave<3, true, float, 2, false>(dg::p32::SweepArgs)"
    rows = []
    for did in (1, 2):  # two launches of 1 ms each, 256000 waves of 1200 VALU instructions
        for c, v in (("SQ_WAVES", 256000), ("SQ_INSTS_VALU", 256000 * 1200), ("SQ_WAVE_CYCLES", 1000), ("SQ_WAIT_ANY", 250)):
            rows.append((did, k, c, v, 0, 1_000_000))
    rows.append((3, "void dg::p64::k_init_B<3>(...)", "SQ_WAVES", 5, 0, 10))
    write_csv(str(tmp_path / "p" / "x" / "p_counter_collection.csv"), hdr, rows)
    text = P.valu_table(str(tmp_path / "p"), "cmd")
    line = [l for l in text.splitlines() if "pred_sweep_z" in l]
    assert len(line) == 1 and "init" not in text
    cells = [c.strip() for c in line[0].strip("|").split("|")]
    assert cells[0] == "mixed" and cells[2] == "2" and cells[3] == "256000" and cells[4] == "1200"
    # 256000 * 1200 * 4 cycles / (1 ms * 2.4 GHz * 1024 SIMDs) = 0.5
    assert cells[8] == "0.50" and cells[9] == "0.25"
    # per step: the counted run advanced the form 2 steps, one launch each
    per = P.valu_per_step(str(tmp_path / "p"), steps=2, command="cmd")
    assert set(per) == {"mixed"}
    assert per["mixed"]["valu_insts_per_step"] == 256000 * 1200
    assert per["mixed"]["kernel_ns_per_step"] == 1_000_000
    assert per["mixed"]["valu_issue_share_counter_run"] == 0.5
