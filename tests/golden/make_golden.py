"""Writes the golden fixtures in tests/golden/ from the reference's OWN code.

Run in the development container, where oracle/build_ref.sh has compiled the
reference's per-function sources into oracle/_ref/libref_<variant>.so:

    python tests/golden/make_golden.py

Each fixture holds seeded random inputs and the reference's outputs for one
function family; tests/test_oracle_golden.py checks the oracle against them
without needing the reference.  Variants: "mhd" / "hydro" (TESTMODE, runtime
solver/limiter switches), "mhdprod" (production switches: beta-gated CT
energy), "mhdg14" (gamma = 1.4).

golden_ctu.npz: Hydro/Godunov/CTU.cpp's correctState (2D, 3D) and
computeCTUFlux_X/Y/Z (the 3D 12-solve corner fluxes), hydro build, HLLC and
the exact solver.  golden_zerob.npz: Riemann fluxes, MUSCL-Hancock half
states and CFL speeds on MHD states whose field is exactly zero on one or both
sides, or below the HLLD floor (|B|^2 < 1e-100), in a near-uniform flow.
Both use their own seeds, so the older fixtures are unchanged.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

SEED = 20261202


def rand_states(rng, n, mhd, zero_bx_every=0, low_beta=False):
    rho = rng.uniform(0.1, 3.0, n)
    p = rng.uniform(0.01, 3.0, n) if not low_beta else rng.uniform(1e-3, 0.05, n)
    v = rng.normal(0, 1.0, (n, 3))
    B = rng.normal(0, 2.0, (n, 3)) if mhd else np.zeros((n, 3))
    if mhd and zero_bx_every:
        B[::zero_bx_every, 0] = 0.0
    return np.column_stack([rho, p, v, B])


def riemann_cases(ref, rng, n, mhd, solvers):
    out = {}
    L = rand_states(rng, n, mhd, zero_bx_every=7)
    R = rand_states(rng, n, mhd, zero_bx_every=7)
    # pair some states to force exact-equality / contact / supersonic branches
    R[: n // 20] = L[: n // 20]
    R[n // 20: n // 10, 1] = L[n // 20: n // 10, 1]
    R[n // 20: n // 10, 2:5] = 0.0
    L[n // 20: n // 10, 2:5] = 0.0
    if mhd:
        R[n // 20: n // 10, 5:8] = L[n // 20: n // 10, 5:8]
    L[n // 10: n // 8, 2] += 12.0
    R[n // 10: n // 8, 2] += 12.0
    out["L"], out["R"] = L, R
    out["dir"] = rng.integers(0, 3, n)
    out["dt_dx"] = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0.0, 3.0, n))
    for s in solvers:
        ref.set(riemann=s)
        F = np.zeros((n, 8))
        err = np.zeros(n, dtype=np.int8)
        for q in range(n):
            try:
                F[q] = ref.riemann_flux(L[q], R[q], out["dt_dx"][q], int(out["dir"][q]))
            except O.OracleError as e:
                err[q] = 1 if "inputs were not physical" in str(e) else 2
        out[f"F_{s}"] = F
        out[f"err_{s}"] = err
    return out


def muscl_cases(ref, rng, n, mhd):
    out = {}
    w = [rand_states(rng, n, mhd) for _ in range(3)]
    dB = rng.normal(0, 0.1, (n, 3)) if mhd else np.zeros((n, 3))
    dt = rng.uniform(0.0, 0.8, n)
    out["wL"], out["wC"], out["wR"], out["dB"], out["dt"] = w[0], w[1], w[2], dB, dt
    for lim in range(5):
        ref.set(limiter=lim)
        Lo, Ro = np.zeros((n, 8)), np.zeros((n, 8))
        for q in range(n):
            Lo[q], Ro[q] = ref.muscl(w[0][q], w[1][q], w[2][q], dt[q], dB[q])
        out[f"L_{lim}"], out[f"R_{lim}"] = Lo, Ro
    ab = rng.normal(0, 1.0, (n, 2))
    out["ab"] = ab
    for lim in range(5):
        ref.set(limiter=lim)
        out[f"lim_{lim}"] = np.array([ref.limit(a, b) for a, b in ab])
    return out


def state_cases(ref, rng, n, mhd):
    W = rand_states(rng, n, mhd, low_beta=False)
    U = np.array([ref.to_cons(w) for w in W])
    return {"W": W, "U": U, "P": np.array([ref.to_prim(u) for u in U]),
            "F": np.array([ref.flux_prim(w) for w in W]), "S": np.array([ref.speeds(w) for w in W]),
            "phys": np.array([ref.is_physical(w) for w in W]),
            "cfl": np.array([[ref.cfl_speed(w, 0.1, 0.2, 0.3), ref.cfl_speed(w, 0.1, 0.2, 0.0), ref.cfl_speed(w, 0.1)] for w in W])}


def ct_cases(ref, rng):
    """CT pieces on small arrays: corner EMF (+upwinding), body EMF, Faraday,
    body fields (2D 5x4, 3D 4x3x3, 2 ghosts)."""
    out = {}
    import ctypes as C
    dp = C.POINTER(C.c_double)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    for dim, shp in ((2, (5, 4)), (3, (4, 3, 3))):
        g = 2
        cells = tuple(s + 2 * g for s in shp)
        faces = tuple(s + 1 + 2 * g for s in shp)
        FX, FY, FZ = (rng.normal(0, 1, cells + (8,)) for _ in range(3))
        ref_e = rng.normal(0, 1, faces + (3,))
        W = np.column_stack([rng.uniform(0.2, 2, np.prod(cells)), rng.uniform(0.2, 2, np.prod(cells)),
                             rng.normal(0, 1, (np.prod(cells), 3)), rng.normal(0, 1, (np.prod(cells), 3))]).reshape(cells + (8,))
        Bf = rng.normal(0, 1, faces + (3,))
        E0 = rng.normal(0, 1, faces + (3,))
        for gg in (0, 1):
            E = E0.copy()
            if dim == 2:
                ref.lib.rf_electric2D(*shp, g, P(FX), P(FY), P(ref_e), gg, P(E))
            else:
                ref.lib.rf_electric3D(*shp, g, P(FX), P(FY), P(FZ), P(ref_e), gg, P(E))
            out[f"E{dim}_{gg}"] = E
            Eb = E0.copy()
            (ref.lib.rf_body_electric2D if dim == 2 else ref.lib.rf_body_electric3D)(*shp, g, P(W), gg, P(Eb))
            out[f"Ebody{dim}_{gg}"] = Eb
            B2 = Bf.copy()
            if dim == 2:
                ref.lib.rf_faraday2D(*shp, g, P(E0), 0.3, 0.7, gg, P(B2))
            else:
                ref.lib.rf_faraday3D(*shp, g, P(E0), 0.3, 0.7, 0.5, gg, P(B2))
            out[f"Bfar{dim}_{gg}"] = B2
        for ct in range(3):
            ref.set(ct=ct)
            W2 = W.copy()
            (ref.lib.rf_body_fields2D if dim == 2 else ref.lib.rf_body_fields3D)(*shp, g, P(Bf), P(W2))
            out[f"Wbody{dim}_{ct}"] = W2
        out[f"in{dim}_FX"], out[f"in{dim}_FY"], out[f"in{dim}_FZ"] = FX, FY, FZ
        out[f"in{dim}_ref"], out[f"in{dim}_W"], out[f"in{dim}_B"], out[f"in{dim}_E0"] = ref_e, W, Bf, E0
    return out


def main():
    rng = np.random.default_rng(SEED)
    n = 600
    for variant, mhd, solvers in (("hydro", False, [O.HLLX, O.HLL, O.HLLE, O.EXACT]),
                                  ("mhd", True, [O.HLLX, O.HLL, O.HLLE]),
                                  ("mhdg14", True, [O.HLLX])):
        ref = O.RefLib(variant)
        ref.set()
        data = {"gamma": np.array(ref.gamma)}
        for k, v in riemann_cases(ref, rng, n, mhd, solvers).items():
            data["riemann_" + k] = v
        ref.set()
        for k, v in muscl_cases(ref, rng, 300, mhd).items():
            data["muscl_" + k] = v
        ref.set()
        for k, v in state_cases(ref, rng, 200, mhd).items():
            data["state_" + k] = v
        if mhd and variant == "mhd":
            for k, v in ct_cases(ref, rng).items():
                data["ct_" + k] = v
        path = os.path.join(HERE, f"golden_{variant}.npz")
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path), "bytes")
    # production build: beta-gated CT energy choice on low-beta / high-beta cells
    ref = O.RefLib("mhdprod")
    import ctypes as C
    dp = C.POINTER(C.c_double)
    shp, g = (6, 5), 2
    cells = tuple(s + 2 * g for s in shp)
    faces = tuple(s + 1 + 2 * g for s in shp)
    W = np.column_stack([rng.uniform(0.2, 2, np.prod(cells)), 10 ** rng.uniform(-3, 0.5, np.prod(cells)),
                         rng.normal(0, 1, (np.prod(cells), 3)), rng.normal(0, 2, (np.prod(cells), 3))]).reshape(cells + (8,))
    Bf = rng.normal(0, 2, faces + (3,))
    W2 = W.copy()
    ref.lib.rf_body_fields2D(*shp, g, Bf.ctypes.data_as(dp), W2.ctypes.data_as(dp))
    path = os.path.join(HERE, "golden_mhdprod.npz")
    np.savez_compressed(path, W=W, B=Bf, Wout=W2)
    print("wrote", path, os.path.getsize(path), "bytes")


def ctu_cases(ref, rng):
    """correctState 2D/3D and computeCTUFlux_X/Y/Z on small arrays (2 ghosts)."""
    import ctypes as C
    dp = C.POINTER(C.c_double)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    g = 2

    def states(shape):
        n = int(np.prod(shape))
        w = np.zeros((n, 8))
        w[:, 0] = rng.uniform(0.5, 2.0, n)
        w[:, 1] = rng.uniform(0.5, 2.0, n)
        w[:, 2:5] = rng.normal(0, 0.5, (n, 3))
        return w.reshape(shape + (8,))

    def fluxes(shape):
        f = rng.normal(0, 0.2, shape + (8,))
        f[..., 5:] = 0.0
        return f

    out = {}
    shp = (5, 4)
    cells = tuple(s + 2 * g for s in shp)
    for fdim in (0, 1):
        L, R, F = states(cells), states(cells), fluxes(cells)
        out[f"cs2_{fdim}_in"] = np.stack([L, R, F])
        ref.lib.rf_correct_state2D(*shp, g, P(L), P(R), P(F), 0.13, fdim)
        out[f"cs2_{fdim}_out"] = np.stack([L, R])
    shp = (3, 2, 4)
    cells = tuple(s + 2 * g for s in shp)
    for sdim in range(3):
        for fdim in range(3):
            if sdim == fdim:
                continue
            L, R, F = states(cells), states(cells), fluxes(cells)
            out[f"cs3_{sdim}{fdim}_in"] = np.stack([L, R, F])
            ref.lib.rf_correct_state3D(*shp, g, P(L), P(R), P(F), 0.11, sdim, fdim)
            out[f"cs3_{sdim}{fdim}_out"] = np.stack([L, R])
    for solver in (O.HLLX, O.EXACT):
        ref.set(riemann=solver, limiter=O.MINMOD)
        for axis in range(3):
            for dimt in range(3):
                if dimt == axis:
                    continue
                L, R, T = states(cells), states(cells), fluxes(cells)
                F = fluxes(cells) * 5.0
                out[f"ctu_{solver}_{axis}{dimt}_in"] = np.stack([L, R, T, F])
                err = C.create_string_buffer(256)
                rc = ref.lib.rf_ctu_flux3D(axis, *shp, g, P(L), P(R), P(T), P(F), 0.2, 0.2 / 3.0, dimt, err, 256)
                assert rc == 0, err.value
                out[f"ctu_{solver}_{axis}{dimt}_out"] = F
    return out


def zerob_cases(ref, rng, n=150):
    """MHD states in a near-uniform flow whose field is exactly zero on one or
    both sides, tiny (~1e-6, the field-loop amplitude) or below the HLLD
    zero-field floor (|B|^2 < 1e-100)."""
    def states(kind):
        rho = 1.0 + 0.01 * rng.normal(size=n)
        p = 5.0 + 0.01 * rng.normal(size=n)
        v = np.tile([2.0, 0.5, 0.0], (n, 1)) + 1e-3 * rng.normal(size=(n, 3))
        if kind == "zero":
            B = np.zeros((n, 3))
        elif kind == "tiny":
            B = 1e-6 * rng.normal(size=(n, 3))
            B[::3, 0] = 0.0
        else:  # below the floor
            B = 1e-52 * rng.normal(size=(n, 3))
        return np.column_stack([rho, p, v, B])

    out = {}
    kinds = [("zero", "zero"), ("tiny", "zero"), ("zero", "tiny"), ("tiny", "tiny"), ("sub", "sub"), ("sub", "zero")]
    for a, b in kinds:
        L, R, Rn = states(a), states(b), states(a)
        key = f"{a}_{b}"
        out[f"{key}_L"], out[f"{key}_R"], out[f"{key}_Rn"] = L, R, Rn
        d = np.arange(n) % 3
        for s in (O.HLLX, O.HLL, O.HLLE):
            ref.set(riemann=s)
            out[f"{key}_F_{s}"] = np.array([ref.riemann_flux(L[q], R[q], 0.1, int(d[q])) for q in range(n)])
        for lim in range(5):
            ref.set(limiter=lim)
            LR = [ref.muscl(L[q], R[q], Rn[q], 0.05, R[q, 5:8] - L[q, 5:8]) for q in range(n)]
            out[f"{key}_M_{lim}"] = np.array([np.concatenate(x) for x in LR])
        ref.set()
        out[f"{key}_cfl"] = np.array([[ref.cfl_speed(w, 0.1, 0.2, 0.0), ref.cfl_speed(w, 0.1, 0.2, 0.3)] for w in L])
    return out


def main_new():
    ref = O.RefLib("hydro")
    data = ctu_cases(ref, np.random.default_rng(SEED + 11))
    path = os.path.join(HERE, "golden_ctu.npz")
    np.savez_compressed(path, **data)
    print("wrote", path, os.path.getsize(path), "bytes")
    ref = O.RefLib("mhd")
    data = zerob_cases(ref, np.random.default_rng(SEED + 12))
    path = os.path.join(HERE, "golden_zerob.npz")
    np.savez_compressed(path, **data)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if "--new-only" not in sys.argv:
        main()
    main_new()
