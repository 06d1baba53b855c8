"""Initial conditions of the benchmark workloads and of the reference's own
accuracy examples, in the reference's ghosted array layout.

Each builder returns ``(shape, d, config_kwargs, bcs, w, B, meta)``:
``w`` is (n+2g)^D x 8 (PrimitiveState rho, p, vx, vy, vz, Bx, By, Bz) with
only interior cells set, ``B`` the (n+1+2g)^D x 3 face field (interior faces
set) or None -- exactly what the reference's main.cpp:26-69 leaves in a fresh
grid before the first boundary fill.  ``bcs`` is a list of
(kind, faces[, state]) tuples; kinds follow dragon_gpu.h (0 outflow,
1 periodic, 2 reflective, 3 fixed, 4 ignore).  The same arrays feed the HIP
path and the CPU oracle, so parity runs start from identical bits.
"""
from __future__ import annotations

import math

import numpy as np

G_MHD_CTU = 3   # validGhosts for MHD + CTU (Splitting.cpp:19-27)
G_MUSCL = 2


def ghosts_for(mhd: bool, ctu: bool = True, muscl: bool = True, requested: int = 1) -> int:
    if mhd and ctu:
        return max(requested, 3)
    if muscl:
        return max(requested, 2)
    return max(requested, 1)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n doubles in [0,1) from splitmix64, 53-bit mantissa (x >> 11) * 2^-53."""
    M = (1 << 64) - 1
    out = np.empty(n, dtype=np.float64)
    s = seed & M
    # vectorised: state_k = seed + (k+1)*gamma
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(s) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return out


def empty_state(shape, g, mhd):
    D = len(shape)
    w = np.zeros(tuple(s + 2 * g for s in shape) + (8,))
    B = np.zeros(tuple(s + 1 + 2 * g for s in shape) + (3,)) if (mhd and D >= 2) else None
    return w, B


def interior(a, g, shape, extra=0):
    sl = tuple(slice(g, g + s + extra) for s in shape)
    return a[sl]


def centers(n, d, origin=0.0):
    return (np.arange(n) + 0.5) * d + origin


# ------------------------------------------------------------------- C1
def sod_1d(n=1024, gamma=5.0 / 3.0):
    """1D Sod shock tube on [0,1] (Examples/1D/Sod/Problem.cpp)."""
    dx = 1.0 / n
    g = ghosts_for(False)
    w, _ = empty_state((n,), g, False)
    x = (np.arange(n) + 0.5) * dx
    left = x < 0.5
    wi = interior(w, g, (n,))
    wi[:, 0] = np.where(left, 1.0, 0.125)
    wi[:, 1] = np.where(left, 1.0, 0.1)
    cfg = dict(mhd=0, gamma=gamma, solver=5, limiter=0, cfl_coeff=0.3)
    return (n,), (dx,), cfg, [], w, None, {"t_end": 0.2}


# ------------------------------------------------------------------- C2
def kelvin_helmholtz_2d(n=1024, seed=20260701, amp=0.01, gamma=5.0 / 3.0):
    """2D KH on [-0.5,0.5]^2, periodic: rho=2, vx=+0.5 for |y|<0.25, else
    rho=1, vx=-0.5; p=2.5; i.i.d. U[-amp, amp) perturbations of vx and vy
    (splitmix64, vx then vy, row-major (i, j))."""
    d = 1.0 / n
    g = ghosts_for(False)
    w, _ = empty_state((n, n), g, False)
    x = centers(n, d, -0.5)
    y = centers(n, d, -0.5)
    X, Y = np.meshgrid(x, y, indexing="ij")
    band = np.abs(Y) < 0.25
    u = splitmix64(seed, 2 * n * n)
    dvx = (2.0 * u[: n * n] - 1.0).reshape(n, n) * amp
    dvy = (2.0 * u[n * n:] - 1.0).reshape(n, n) * amp
    wi = interior(w, g, (n, n))
    wi[..., 0] = np.where(band, 2.0, 1.0)
    wi[..., 1] = 2.5
    wi[..., 2] = np.where(band, 0.5, -0.5) + dvx
    wi[..., 3] = dvy
    cfg = dict(mhd=0, gamma=gamma, solver=5, limiter=2, ctu=1)
    return (n, n), (d, d), cfg, [(1, 63)], w, None, {"seed": seed}


# ------------------------------------------------------------------- C3
def orszag_tang_2d(n=2048, gamma=5.0 / 3.0, s=1.0 / math.sqrt(4.0 * math.pi)):
    """2D Orszag-Tang on [0,1]^2, periodic, face-centred B given directly:
    Bx(x_{i-1/2}, y_j) = -sin(2 pi y_j) s, By(x_i, y_{j-1/2}) = sin(4 pi x_i) s.
    Call initialize_B_fields() after loading."""
    d = 1.0 / n
    g = ghosts_for(True)
    w, B = empty_state((n, n), g, True)
    x = centers(n, d)
    y = centers(n, d)
    X, Y = np.meshgrid(x, y, indexing="ij")
    wi = interior(w, g, (n, n))
    wi[..., 0] = 25.0 / (36.0 * math.pi)
    wi[..., 1] = 5.0 / (12.0 * math.pi)
    wi[..., 2] = -np.sin(2 * math.pi * Y)
    wi[..., 3] = np.sin(2 * math.pi * X)
    Bi = interior(B, g, (n, n), extra=1)          # face indices 0..n
    c = (np.arange(n + 1) + 0.5) * d               # cell-centre coordinate of each index
    Bi[..., 0] = (-np.sin(2 * math.pi * c) * s)[None, :]   # Bx on x-face i-1/2, row j
    Bi[..., 1] = (np.sin(4 * math.pi * c) * s)[:, None]    # By on y-face j-1/2, column i
    cfg = dict(mhd=1, gamma=gamma, solver=5, limiter=1, ctu=1, ct_energy=2)
    return (n, n), (d, d), cfg, [(1, 63)], w, B, {"init_B": True}


# ------------------------------------------------------------------- C4
def turbulence_modes(seed=20260704, kmin=1.0, kmax=4.0):
    """Solenoidal Gaussian modes k (integer, 2pi units) with kmin<=|k|<=kmax,
    half-space, amplitude ~ |k|^-1 (power ~ |k|^-2), Box-Muller from
    splitmix64.  Returns (k[m,3], a_re[m,3], a_im[m,3])."""
    r = int(math.ceil(kmax))
    ks = []
    for kx in range(-r, r + 1):
        for ky in range(-r, r + 1):
            for kz in range(-r, r + 1):
                kk = math.sqrt(kx * kx + ky * ky + kz * kz)
                if kk < kmin or kk > kmax:
                    continue
                if (kx, ky, kz) <= (0, 0, 0) and (-kx, -ky, -kz) > (0, 0, 0):
                    continue  # keep one of each +/- pair
                ks.append((kx, ky, kz))
    k = np.array(ks, dtype=np.float64)
    m = len(k)
    u = splitmix64(seed, 12 * m).reshape(m, 12)
    # Box-Muller: 6 pairs -> 6 normals per mode (3 re, 3 im)
    u1 = np.maximum(u[:, 0::2], 2.0 ** -53)
    u2 = u[:, 1::2]
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * math.pi * u2)
    kn = np.sqrt((k * k).sum(1))
    amp = 1.0 / kn
    re = z[:, :3] * amp[:, None]
    im = z[:, 3:] * amp[:, None]
    # projection a -= k (k.a)/|k|^2
    for a in (re, im):
        a -= k * ((k * a).sum(1) / (kn * kn))[:, None]
    return k, re, im


def turbulence_velocity(n, seed=20260704, mach=1.0, gamma=5.0 / 3.0, p=1.0, rho=1.0):
    """v(x) = sum_m 2 Re(a_m e^{2 pi i k.x}) at cell centres, scaled to rms |v|
    = mach * sqrt(gamma p / rho).  Direct mode sum with separable phases."""
    k, re, im = turbulence_modes(seed)
    x = (np.arange(n) + 0.5) / n
    v = np.zeros((n, n, n, 3))
    for q in range(len(k)):
        px = np.exp(2j * math.pi * k[q, 0] * x)
        py = np.exp(2j * math.pi * k[q, 1] * x)
        pz = np.exp(2j * math.pi * k[q, 2] * x)
        ph = px[:, None, None] * py[None, :, None] * pz[None, None, :]
        for c in range(3):
            a = complex(re[q, c], im[q, c])
            v[..., c] += 2.0 * (a * ph).real
    rms = math.sqrt((v * v).sum(-1).mean())
    target = mach * math.sqrt(gamma * p / rho)
    return v * (target / rms)


def turbulence_3d(n=512, seed=20260704, mach=1.0, gamma=5.0 / 3.0, B0=0.2, velocity=None):
    """3D driven-free MHD turbulence box, periodic [0,1]^3: rho=1, p=1,
    face B = (B0, 0, 0), solenoidal random velocity (rms Mach `mach`).
    Call initialize_B_fields() after loading."""
    d = 1.0 / n
    g = ghosts_for(True)
    w, B = empty_state((n, n, n), g, True)
    wi = interior(w, g, (n, n, n))
    wi[..., 0] = 1.0
    wi[..., 1] = 1.0
    v = turbulence_velocity(n, seed, mach, gamma) if velocity is None else velocity
    wi[..., 2:5] = v
    Bi = interior(B, g, (n, n, n), extra=1)
    Bi[..., 0] = B0
    cfg = dict(mhd=1, gamma=gamma, solver=5, limiter=2, ctu=1, ct_energy=2)
    return (n, n, n), (d, d, d), cfg, [(1, 63)], w, B, {"init_B": True}


# ------------------------------------------------------------------- C5
def rotor_params(npatch=4096, seed=20260705):
    """Per-patch (n, B0, r0, u0) of the batched rotor workload: n log-uniform
    in [32, 256], B0 ~ U[2.5, 5], r0 ~ U[0.08, 0.12], u0 ~ U[1, 2]."""
    u = splitmix64(seed, 4 * npatch).reshape(npatch, 4)
    n = np.floor(np.exp(math.log(32.0) + u[:, 0] * (math.log(256.0) - math.log(32.0)))).astype(np.int64)
    n = np.clip(n, 32, 255)
    B0 = 2.5 + 2.5 * u[:, 1]
    r0 = 0.08 + 0.04 * u[:, 2]
    u0 = 1.0 + u[:, 3]
    return n, B0, r0, u0


def rotor_2d(n, B0=5.0, r0=0.1, u0=2.0, gamma=1.4, limiter=0, ny=None):
    """MHD rotor on [0,1]^2 with outflow boundaries (Examples/MHDRotor):
    ambient rho=1, p=1, Bx = B0/sqrt(4 pi); disk r<r0 rho=10 spinning at
    edge speed u0, linear taper to ambient by r1 = 1.15 r0.  ny: cells in y
    (default n), dy = 1/ny."""
    ny = n if ny is None else ny
    d = 1.0 / n
    dy = 1.0 / ny
    g = ghosts_for(True)
    w, B = empty_state((n, ny), g, True)
    x = centers(n, d) - 0.5
    y = centers(ny, dy) - 0.5
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    r1 = 1.15 * r0
    om = u0 / r0
    f = np.clip((r1 - r) / (r1 - r0), 0.0, 1.0)
    inside = r < r0
    taper = (r >= r0) & (r < r1)
    rho = np.where(inside, 10.0, np.where(taper, 1.0 + 9.0 * f, 1.0))
    scale = np.where(inside, om, np.where(taper, om * f * r0 / np.maximum(r, 1e-300), 0.0))
    wi = interior(w, g, (n, ny))
    wi[..., 0] = rho
    wi[..., 1] = 1.0
    wi[..., 2] = -scale * Y
    wi[..., 3] = scale * X
    Bi = interior(B, g, (n, ny), extra=1)
    Bi[..., 0] = B0 / math.sqrt(4.0 * math.pi)
    cfg = dict(mhd=1, gamma=gamma, solver=5, limiter=limiter, ctu=1, ct_energy=2)
    return (n, ny), (d, dy), cfg, [], w, B, {"init_B": True}


def rotor_batch(npatch=4096, seed=20260705):
    """The batched rotor workload: per-patch sizes and parameters from
    rotor_params; returns (sizes, spacings, cfg, make) where make(p) builds
    patch p's (w, B) as rotor_2d does."""
    n, B0, r0, u0 = rotor_params(npatch, seed)
    sizes = np.stack([n, n], axis=1)
    d = 1.0 / sizes
    cfg = rotor_2d(8)[2]

    def make(p):
        pb = rotor_2d(int(n[p]), float(B0[p]), float(r0[p]), float(u0[p]))
        return pb[4], pb[5]

    return sizes, d, cfg, make


# ----------------------------------------- reference accuracy examples
def diagonal_2d(n, limiter=1):
    """Examples/Diagonal/Problem2D.cpp: rho = 1 + 0.1 sin 2pi(x+y), p=1, v=(1,1), periodic, HLLC."""
    d = 1.0 / n
    g = ghosts_for(False)
    w, _ = empty_state((n, n), g, False)
    x = (np.arange(n) + 0.5) * d
    X, Y = np.meshgrid(x, x, indexing="ij")
    wi = interior(w, g, (n, n))
    wi[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * (X + Y))
    wi[..., 1] = 1.0
    wi[..., 2] = 1.0
    wi[..., 3] = 1.0
    cfg = dict(mhd=0, solver=2, limiter=limiter, ctu=1)
    return (n, n), (d, d), cfg, [(1, 63)], w, None, {"t_end": 1.0, "frame": 0.05}


def diagonal_3d(n, limiter=1):
    """Examples/Diagonal/Problem3D.cpp: rho = 1 + 0.1 sin 2pi(x+y+z), v=(1,1,1)."""
    d = 1.0 / n
    g = ghosts_for(False)
    w, _ = empty_state((n, n, n), g, False)
    x = (np.arange(n) + 0.5) * d
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    wi = interior(w, g, (n, n, n))
    wi[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * (X + Y + Z))
    wi[..., 1] = 1.0
    wi[..., 2:5] = 1.0
    cfg = dict(mhd=0, solver=2, limiter=limiter, ctu=1)
    return (n, n, n), (d, d, d), cfg, [(1, 63)], w, None, {"t_end": 1.0, "frame": 0.05}


def diagonal_error(w, g, shape, t):
    """L1/L2/Linf of rho against the advected sine (Diagonal/Problem2D.cpp:76-90)."""
    n = shape[0]
    x = (np.arange(n) + 0.5) / n
    grids = np.meshgrid(*([x] * len(shape)), indexing="ij")
    s = sum(grids)
    exact = 1.0 + 0.1 * np.sin(2.0 * np.pi * (s - len(shape) * t))
    err = np.abs(interior(w, g, shape)[..., 0] - exact)
    N = err.size
    return err.sum() / N, math.sqrt((err * err).sum() / N), err.max()


def faraday_from_potential(A, d, g, shape):
    """B = curl A by CT::Faraday(A, B, -1/dx, -1/dy[, -1/dz], 0) on zeroed
    faces (main.cpp:35-43, 57-67; B_fields.cpp:21-66), A at corners."""
    D = len(shape)
    B = np.zeros(tuple(s + 1 + 2 * g for s in shape) + (3,))
    Bi = interior(B, g, shape, extra=1)
    if D == 2:
        nx, ny = shape
        cx, cy = -1 / d[0], -1 / d[1]
        Bi[:, :ny, 0] = 0.0 - (A[:, 1:, 2] - A[:, :-1, 2]) * cy
        Bi[:nx, :, 1] = 0.0 - (-(A[1:, :, 2] - A[:-1, :, 2]) * cx)
        Bi[:nx, :ny, 2] = 0.0 - ((A[1:, :-1, 1] - A[:-1, :-1, 1]) * cx - (A[:-1, 1:, 0] - A[:-1, :-1, 0]) * cy)
    else:
        nx, ny, nz = shape
        cx, cy, cz = -1 / d[0], -1 / d[1], -1 / d[2]
        Bi[:, :ny, :nz, 0] = 0.0 - ((A[:, 1:, :-1, 2] - A[:, :-1, :-1, 2]) * cy - (A[:, :-1, 1:, 1] - A[:, :-1, :-1, 1]) * cz)
        Bi[:nx, :, :nz, 1] = 0.0 - ((A[:-1, :, 1:, 0] - A[:-1, :, :-1, 0]) * cz - (A[1:, :, :-1, 2] - A[:-1, :, :-1, 2]) * cx)
        Bi[:nx, :ny, :, 2] = 0.0 - ((A[1:, :-1, :, 1] - A[:-1, :-1, :, 1]) * cx - (A[:-1, 1:, :, 0] - A[:-1, :-1, :, 0]) * cy)
    return B


def loop_advection(n, dim=2, gaussian=False):
    """Examples/LoopAdvection/Problem{2,3}D.cpp: field loop Az = B0 max(r0-r, 0)
    (or Gaussian) advected by v=(2,0.5[,..]) on [0,2]x[0,1][x[0,1]], p=5."""
    B0 = 1e-6
    r0 = 0.1 if gaussian else 0.2
    dx = 2.0 / n
    shape = (n, n // 2) if dim == 2 else (n, n // 2, n // 2)
    g = ghosts_for(True)
    w, _ = empty_state(shape, g, True)
    wi = interior(w, g, shape)
    wi[..., 0] = 1.0
    wi[..., 1] = 5.0
    vel = (2.0, 0.5, 0.0) if dim == 2 else (2.0, 1.0, 1.0)
    wi[..., 2], wi[..., 3], wi[..., 4] = vel
    xs = [np.arange(s + 1) * dx for s in shape]
    P = np.meshgrid(*xs, indexing="ij")
    x = P[0] - 1.0
    y = P[1] - 0.5
    r = np.sqrt(x * x + y * y)
    Az = B0 * np.exp(-0.5 * (r / r0) ** 2) if gaussian else B0 * np.maximum(r0 - r, 0.0)
    A = np.zeros(Az.shape + (3,))
    A[..., 2] = Az
    B = faraday_from_potential(A, (dx,) * dim, g, shape)
    cfg = dict(mhd=1, solver=5, limiter=1, ctu=1, ct_energy=0)
    return shape, (dx,) * dim, cfg, [(1, 63)], w, B, {"init_B": True, "t_end": 4.0, "frame": 0.1, "B0": B0}


def loop_error(w_final, w_init, g, shape, B0=1e-6):
    """|B|^2 error / B0^2 (LoopAdvection/Problem2D.cpp:80-107; the reference
    divides by n^D, not by the cell count)."""
    a = interior(w_final, g, shape)[..., 5:8]
    b = interior(w_init, g, shape)[..., 5:8]
    err = np.abs((a * a).sum(-1) - (b * b).sum(-1)) / (B0 * B0)
    N = shape[0] ** len(shape)
    return err.sum() / N, math.sqrt((err * err).sum() / N), err.max()


def frames(t_end, frame):
    """Frame times of main.cpp:133-150: advance(frame) while time < t_end."""
    t, out = 0.0, []
    while t < t_end:
        out.append(frame)
        t += frame
    return out, t
