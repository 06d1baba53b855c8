#!/usr/bin/env python3
"""Throughput of the unsplit Godunov step on one MI355X.

Workload (the headline metric's configuration): 3D driven-free MHD
turbulence, 512^3 periodic, gamma=5/3, unsplit CTU + HLLD + CT, van Leer,
rho=1, p=1, B=(0.2,0,0), solenoidal random velocity at rms Mach 1.
A step is one pass of {boundary fill + CFL reduction + unsplit step} over the
whole grid.  Inputs are built directly in HBM before the timed region.

Prints one JSON line.  `value` is zone-updates/s of the faster validated
form; both forms are reported.  The roofline uses the algorithmic bytes of
the step (128 B per zone-update: 8 FP64 fields read + written once) over the
step kernels' device time measured with HIP events on the library stream.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_ALG_MHD = 128  # bytes per zone-update: rho, m(3), E, face Bx, By, Bz; read + write, FP64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_state_gpu(n, seed, torch, dev):
    """C4 initial state in the library's SoA layout, built on the GPU."""
    from dragon_amd.gpu_states import turbulence_3d_soa
    return turbulence_3d_soa(n, torch, dev, seed)


def run_form(n, precision, W, Bf, steps, warmup, torch, rho_out=None):
    import dragon_amd as D
    d = 1.0 / n
    cfg = D.Config(mhd=1, limiter=D.LIMITER_VANLEER, ctu=1, ct_energy=D.CT_CONSV_BETA_GATED, precision=precision)
    G = D.Grid3D(n, n, n, d, d, d, 1, cfg)
    G.upload_device(W.data_ptr(), Bf.data_ptr())
    G.boundary = D.Periodic()
    G.initialize_B_fields()
    if warmup > 0:
        G.advance(1e9, True, warmup)
    torch.cuda.synchronize()
    G.timing(1)
    t0 = time.perf_counter()
    done, fails = G.advance(1e9, True, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms, ksteps = G.timing(-1)
    kern = G.kernel_times()
    G.timing(0)
    divb = G.divb()
    per = {k: {"ms_per_launch": ms / max(nl, 1), "launches_per_step": nl / max(ksteps, 1)} for k, (ms, nl) in kern.items()}
    out = {
        "steps": done, "failures": fails, "wall_s": t1 - t0,
        "zone_updates_per_s": n ** 3 * done / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / max(done, 1),
        "kernel_ms_per_step": kms / max(ksteps, 1),
        "kernels": per,
        "max_divB_normalised": divb,
    }
    out["t_advanced"] = G.t_advanced
    if rho_out is not None:
        Wf = torch.empty_like(W)
        Bo = torch.empty_like(Bf)
        G.download_device(Wf.data_ptr(), Bo.data_ptr())
        torch.cuda.synchronize()
        g = 3
        sl = slice(g, g + n)
        r = Wf[0, sl, sl, sl]
        # evolution check: how far rho moved from the initial state
        out["rel_L1_rho_change"] = float((r - W[0, sl, sl, sl]).abs().sum() / W[0, sl, sl, sl].abs().sum())
        rho_out.append(r.clone())
        del Wf, Bo
    del G
    return out


def copy_probe(torch, dev, nbytes=1 << 30, reps=3):
    """HBM copy ceiling: an elementwise read+write of nbytes each way (torch's
    negation kernel, used nowhere else in this program), timed with HIP
    events.  Also the known-size kernel that scales the FETCH_SIZE /
    WRITE_SIZE counters."""
    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        torch.neg(a, out=b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GBs": 2 * nbytes / (ms * 1e-3) / 1e9}


def host_cpu():
    """Host CPU model (lscpu's "Model name", read from /proc/cpuinfo) and the
    logical CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(n, repeats=3, target_s=3.0, max_steps=40):
    """Oracle (our reference-equivalent C++ port, the CPU path's arithmetic
    and multi-pass structure) on the host cores, on a bounded sample of the
    same workload (n^3 instead of 512^3): the median of `repeats` timed runs
    on all usable cores (at most 16, the host share of one GPU), and one
    single-thread run."""
    import oracle as O
    from dragon_amd import problems as P
    model, ncpu, usable = host_cpu()
    threads = max(1, min(16, usable))
    pb = P.turbulence_3d(n)
    shape, dd, cfg, bcs, w, B, meta = pb

    def grid(t):
        c = O.Cfg(threads=t)
        for kk, vv in cfg.items():
            setattr(c, kk, vv)
        G = O.Grid(shape, dd, c, ghosts=1)
        G.set_w(w)
        G.set_B(B)
        G.add_bc(O.PERIODIC, O.FALL)
        G.init_B()
        G.advance(1e9, True, 1)  # warm
        return G

    def timed(G, budget):
        t0 = time.perf_counter()
        steps = 0
        while steps < max_steps:
            G.advance(1e9, True, 1)
            steps += 1
            if time.perf_counter() - t0 > budget:
                break
        return n ** 3 * steps / (time.perf_counter() - t0), steps

    runs = [timed(grid(threads), target_s) for _ in range(repeats)]
    rates = sorted(r[0] for r in runs)
    one, one_steps = timed(grid(1), target_s)
    return {"value": rates[len(rates) // 2], "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"C4 switch set at {n}^3 (BC+CFL+step per step), oracle C++ port; median of {repeats} runs of "
                      f"{[r[1] for r in runs]} steps on {threads} threads",
            "runs": rates, "single_thread": one, "single_thread_steps": one_steps,
            "host_cpu_model": model, "host_logical_cpus": ncpu, "usable_cpus": usable}


def bench_batch(torch, npatch=4096, steps=10, seed=20260705, baseline_patches=256):
    """The batched-rotor workload: npatch independent 2D MHD patches (HLLD +
    CT + CTU, gamma=1.4, minmod, outflow), each step a fixed sequence of
    launches over all patches with per-patch dt, mixed FP32; also the FP64
    batch, the launches replayed from a HIP graph, and per-patch launches of
    the single-grid path on the first `baseline_patches` patches."""
    import dragon_amd as D
    from dragon_amd import problems as P
    sizes, d, cfg, make = P.rotor_batch(npatch, seed)
    t0 = time.perf_counter()
    ics = [make(p) for p in range(npatch)]
    log(f"batch: {npatch} patch initial states in {time.perf_counter() - t0:.1f} s")
    out = {"workload": f"{npatch} 2D MHD rotor patches, n log-uniform in [32,256), HLLD+CT+CTU, gamma=1.4, minmod, outflow",
           "zones_per_step": int((sizes[:, 0] * sizes[:, 1]).sum()), "steps": steps}
    rho = {}
    for form, prec in (("mixed", D.MIXED), ("fp64", D.FP64)):
        conf = D.Config(precision=prec)
        for k, v in cfg.items():
            setattr(conf, k, v)
        b = D.Batch(sizes, d, conf)
        for p, (w, B) in enumerate(ics):
            b.set_state(p, w, B)
        b.initialize_B_fields()
        tiles, cls_ms = b.autotune(reps=2)
        b.run(2)
        b.sync()
        ms_eager = b.run(steps, graph=False, timed=True)
        b.sync()
        b.run(steps, graph=True)  # capture + first replay
        b.sync()
        ms_graph = b.run(steps, graph=True, timed=True)
        b.sync()
        z = b.zones
        out[form] = {"tiles": tiles, "ysweep": b.ysweep(), "autotune_class_ms": cls_ms,
                     "eager_zone_updates_per_s": z * steps / (ms_eager * 1e-3),
                     "graph_zone_updates_per_s": z * steps / (ms_graph * 1e-3),
                     "ms_per_step_eager": ms_eager / steps, "ms_per_step_graph": ms_graph / steps}
        rho[form] = [b.get_state(p)[0][3:-3, 3:-3, 0] for p in range(0, npatch, 97)]
        log(f"  batch {form}: {out[form]}")
        del b
    num = sum(float(np.abs(a - c).sum()) for a, c in zip(rho["mixed"], rho["fp64"]))
    den = sum(float(np.abs(c).sum()) for c in rho["fp64"])
    out["mixed_rel_L1_rho_vs_fp64_sampled_patches"] = num / den
    # per-patch launches of the single-grid path (host-driven, one grid per patch)
    conf = D.Config(precision=D.MIXED)
    for k, v in cfg.items():
        setattr(conf, k, v)
    grids = []
    zsub = 0
    for p in range(baseline_patches):
        n = int(sizes[p, 0])
        G = D.Grid2D(n, n, float(d[p, 0]), float(d[p, 1]), 1, conf)
        G.set_state(*ics[p])
        G.initialize_B_fields()
        grids.append(G)
        zsub += n * n
    for G in grids:
        G.advance(1e9, True, 1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for G in grids:
        G.advance(1e9, True, 2)
    torch.cuda.synchronize()
    dtw = time.perf_counter() - t1
    out["per_patch_launches"] = {"patches": baseline_patches, "zone_updates_per_s": zsub * 2 / dtw,
                                 "note": "single-grid path, one grid per patch, BC+CFL+step launched per patch with host syncs"}
    best = max(out["mixed"]["eager_zone_updates_per_s"], out["mixed"]["graph_zone_updates_per_s"])
    out["speedup_vs_per_patch_launches"] = best / out["per_patch_launches"]["zone_updates_per_s"]
    out["roofline"] = {"bound": "hbm", "achieved": round(best * B_ALG_MHD / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(best * B_ALG_MHD / 1e9 / HBM_PEAK_GBS, 5)}
    del grids
    return out


def read_valu():
    p = os.path.join(ROOT, "profiles", "valu_c4_512.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def read_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_c4_512.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20260704)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--cpu-n", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-rotor measurement")
    ap.add_argument("--batch-patches", type=int, default=4096)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (requested --gpus {a.gpus}); the multi-GPU "
              "decomposition of this path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    import torch
    import dragon_amd as D
    dev = torch.device("cuda:0")
    log(f"building {a.n}^3 initial state on the GPU")
    W, Bf = make_state_gpu(a.n, a.seed, torch, dev)
    log(f"initial state: max|v| {W[2:5].abs().max().item():.4f}, rho mean {W[0, 3:-3, 3:-3, 3:-3].mean().item():.4f}")
    copy = copy_probe(torch, dev)
    log(f"copy probe: {copy}")
    forms = {}
    rhos = {}
    for f in a.forms.split(","):
        prec = D.MIXED if f == "mixed" else D.FP64
        log(f"running form {f}: warmup {a.warmup}, timed {a.steps}")
        keep = []
        try:
            forms[f] = run_form(a.n, prec, W, Bf, a.steps, a.warmup, torch, keep)
            rhos[f] = keep[0]
        except D.DragonError as e:
            forms[f] = {"error": str(e), "zone_updates_per_s": 0.0}
        log(f"  {f}: {forms[f]}")
    del W, Bf
    if "fp64" in rhos and "mixed" in rhos:
        # same initial state and step count; FP64 GPU is bit-identical to the oracle
        r64 = rhos["fp64"]
        forms["mixed"]["rel_L1_rho_vs_fp64"] = float((rhos["mixed"] - r64).abs().sum() / r64.abs().sum())
    rhos.clear()
    torch.cuda.empty_cache()
    best = max(forms, key=lambda f: forms[f]["zone_updates_per_s"])
    r = forms[best]
    zones = a.n ** 3
    achieved = B_ALG_MHD * zones / (r["kernel_ms_per_step"] * 1e-3) / 1e9
    traffic = read_traffic()
    valu = read_valu()
    dom = max(r["kernels"].items(), key=lambda kv: kv[1]["ms_per_launch"] * kv[1]["launches_per_step"]) if r.get("kernels") else None
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic.get(best, {}).get("bytes_per_step") if traffic else None,
            "traffic_note": "FETCH_SIZE + WRITE_SIZE over the step kernels, per step, scaled by the copy probe's "
                            "counters (profiles/traffic_c4_512.json, an earlier counter run of this command); "
                            "FETCH_SIZE counts L2 misses, including those served by the MALL, so it can exceed HBM bytes",
            "kernel": "unsplit step: its kernel sequence on one stream, HIP events around each step",
            "alg_bytes_per_launch": B_ALG_MHD * zones,
            "ms_per_launch": r["kernel_ms_per_step"],
            "dominant_kernel": {"name": dom[0], **dom[1]} if dom else None,
            "copy_ceiling_GBs": round(copy["GBs"], 1)}
    if valu and best in valu:
        # VALU issue: wave-level VALU instructions per step x 4 cycles (a wave64 instruction on a 16-lane SIMD)
        # over the cycles of 1024 SIMDs (256 CUs x 4) at the 2.4 GHz peak engine clock for the step's device time
        inst = valu[best]["valu_insts_per_step"]
        roof["valu_issue"] = {"frac": round(inst * 4 / (1024 * 2.4e9 * r["kernel_ms_per_step"] * 1e-3), 4),
                              "valu_insts_per_step": inst, "cycles_per_inst": 4, "simds": 1024, "clock_GHz": 2.4,
                              "source": "SQ_INSTS_VALU summed over the step's kernels per step, from a counter run of this command "
                                        "(profiles/valu_c4_512.json); the fraction is a lower bound when the clock runs below 2.4 GHz"}
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_n)
    line = {
        "metric": "zone-updates/s, 3D MHD HLLD+CT 512^3, mixed-FP32 vs FP64 GPU; % HBM roofline",
        "value": r["zone_updates_per_s"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": r["steps"], "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if best == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"3D MHD turbulence {a.n}^3 periodic, CTU+HLLD+CT, van Leer, gamma=5/3, rms Mach 1, B0=0.2",
                   "grid": [a.n] * 3, "form": best},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": forms,
        "copy_probe": copy,
    }
    if "mixed" in forms and "fp64" in forms and forms["fp64"]["zone_updates_per_s"] > 0:
        line["mixed_speedup_vs_fp64"] = forms["mixed"]["zone_updates_per_s"] / forms["fp64"]["zone_updates_per_s"]
    if not a.no_batch:
        try:
            line["batch"] = bench_batch(torch, npatch=a.batch_patches)
        except D.DragonError as e:  # the headline line must still print
            line["batch"] = {"error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
