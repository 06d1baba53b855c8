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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_ALG_MHD = 128  # bytes per zone-update: rho, m(3), E, face Bx, By, Bz; read + write, FP64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_state_gpu(n, seed, torch, dev, mach=1.0, gamma=5.0 / 3.0, B0=0.2):
    """C4 initial state in the library's SoA layout, built on the GPU."""
    from dragon_amd import problems as P
    g = 3
    k, re, im = P.turbulence_modes(seed)
    x = (torch.arange(n, dtype=torch.float64, device=dev) + 0.5) / n
    v = torch.zeros(3, n, n, n, dtype=torch.float64, device=dev)
    two_pi_i = 2j * math.pi
    for q in range(len(k)):
        px = torch.exp(two_pi_i * float(k[q, 0]) * x)
        py = torch.exp(two_pi_i * float(k[q, 1]) * x)
        pz = torch.exp(two_pi_i * float(k[q, 2]) * x)
        pyz = py[:, None] * pz[None, :]
        for c in range(3):
            a = complex(re[q, c], im[q, c])
            # Re(a * px * pyz) without materialising the complex cube per component
            ax = a * px
            v[c] += 2.0 * (ax.real[:, None, None] * pyz.real[None] - ax.imag[:, None, None] * pyz.imag[None])
    rms = torch.sqrt((v * v).sum(0).mean()).item()
    v *= mach * math.sqrt(gamma) / rms
    N = n + 2 * g
    W = torch.zeros(8, N, N, N, dtype=torch.float64, device=dev)
    sl = slice(g, g + n)
    W[0, sl, sl, sl] = 1.0
    W[1, sl, sl, sl] = 1.0
    W[2:5, sl, sl, sl] = v
    del v
    Nf = n + 1 + 2 * g
    Bf = torch.zeros(3, Nf, Nf, Nf, dtype=torch.float64, device=dev)
    sf = slice(g, g + n + 1)
    Bf[0, sf, sf, sf] = B0
    torch.cuda.synchronize()
    return W, Bf


def run_form(n, precision, W, Bf, steps, warmup, torch):
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
    kms, ksteps = G.timing(0)
    divb = G.divb()
    out = {
        "steps": done, "failures": fails, "wall_s": t1 - t0,
        "zone_updates_per_s": n ** 3 * done / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / max(done, 1),
        "kernel_ms_per_step": kms / max(ksteps, 1),
        "max_divB_normalised": divb,
    }
    del G
    return out


def cpu_baseline(n, target_s=12.0, max_steps=40):
    """Oracle (our reference-equivalent C++ port) on the host cores, on a
    bounded sample of the same workload (n^3 instead of 512^3)."""
    import oracle as O
    from dragon_amd import problems as P
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    threads = max(1, min(16, os.cpu_count() or 1))
    pb = P.turbulence_3d(n)
    shape, dd, cfg, bcs, w, B, meta = pb
    c = O.Cfg(threads=threads)
    for kk, vv in cfg.items():
        setattr(c, kk, vv)
    G = O.Grid(shape, dd, c, ghosts=1)
    G.set_w(w)
    G.set_B(B)
    G.add_bc(O.PERIODIC, O.FALL)
    G.init_B()
    G.advance(1e9, True, 1)  # warm
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps:
        G.advance(1e9, True, 1)
        steps += 1
        if time.perf_counter() - t0 > target_s:
            break
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"C4 switch set at {n}^3, {steps} steps (BC+CFL+step), oracle C++ port, {threads} threads"}


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
    forms = {}
    for f in a.forms.split(","):
        prec = D.MIXED if f == "mixed" else D.FP64
        log(f"running form {f}: warmup {a.warmup}, timed {a.steps}")
        try:
            forms[f] = run_form(a.n, prec, W, Bf, a.steps, a.warmup, torch)
        except D.DragonError as e:
            forms[f] = {"error": str(e), "zone_updates_per_s": 0.0}
        log(f"  {f}: {forms[f]}")
    del W, Bf
    torch.cuda.empty_cache()
    best = max(forms, key=lambda f: forms[f]["zone_updates_per_s"])
    r = forms[best]
    zones = a.n ** 3
    achieved = B_ALG_MHD * zones / (r["kernel_ms_per_step"] * 1e-3) / 1e9
    traffic = read_traffic()
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic.get(best, {}).get("bytes_per_step") if traffic else None,
            "kernel": "unsplit step (all step kernels, HIP events)",
            "alg_bytes_per_launch": B_ALG_MHD * zones}
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_n)
    line = {
        "metric": "zone-updates/s, 3D MHD HLLD+CT 512^3, mixed-FP32 vs FP64 GPU; % HBM roofline",
        "value": r["zone_updates_per_s"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": r["steps"], "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 reconstruction/Riemann + f64 state" if best == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"3D MHD turbulence {a.n}^3 periodic, CTU+HLLD+CT, van Leer, gamma=5/3, rms Mach 1, B0=0.2",
                   "grid": [a.n] * 3, "form": best},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    if "mixed" in forms and "fp64" in forms and forms["fp64"]["zone_updates_per_s"] > 0:
        line["mixed_speedup_vs_fp64"] = forms["mixed"]["zone_updates_per_s"] / forms["fp64"]["zone_updates_per_s"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
