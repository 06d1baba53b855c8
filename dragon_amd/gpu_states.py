"""Initial states built on the GPU with torch, in the library's SoA device
layout (field-major, ghosts included), for the grids too large to build on the
host quickly: the 512^3 MHD turbulence box.  The arrays feed
Grid.upload_device directly."""
from __future__ import annotations

import math

from . import problems as P


def turbulence_3d_soa(n, torch, dev, seed=20260704, mach=1.0, gamma=5.0 / 3.0, B0=0.2):
    """3D MHD turbulence (problems.turbulence_3d's state): rho=1, p=1, face
    B=(B0,0,0), solenoidal velocity from the same seeded modes, rms Mach
    `mach`.  Returns (W[8, N, N, N], B[3, N+1, N+1, N+1]) with N = n + 6."""
    g = P.ghosts_for(True)
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
