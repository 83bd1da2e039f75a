"""Writes tests/golden/kat.json: closed-form known answers for the SPH passes.

Every value here is derived analytically (or by integer enumeration of lattice
offsets), independently of the oracle's pair loops, from the reference's
constants:
  h = fl32(4.0 * 0.005) (Main.cpp:113, rho_pres_comp.glsl:5,40), m = 0.02,
  visc = 3000, rho0 = 1000 (Main.cpp:112-115), k = 2000 (rho_pres_comp.glsl:33),
  PI = fl32(3.141592741) (rho_pres_comp.glsl:8), G = (0,-9806.65,0)
  (force_comp.glsl:33), dt = fl32(1/10000) (integrate_comp.glsl:33), damping
  fl32(0.3) (integrate_comp.glsl:8).
Run: python tests/golden/make_golden.py
"""
import json
import math
import os
from fractions import Fraction

import numpy as np

F32 = np.float32
H = float(F32(4.0) * F32(0.005))
M = float(F32(0.02))
VISC = float(F32(3000.0))
RHO0 = 1000.0
K = 2000.0
PI = float(F32(3.141592741))
G = float(F32(-9806.65))
DT = float(F32(1.0) / F32(10000.0))
DAMP = float(F32(0.3))


def w_poly6(r):
    """m * 315 (h^2 - r^2)^3 / (64 pi h^9) for r < h, else 0."""
    return M * 315.0 * (H * H - r * r) ** 3 / (64.0 * PI * H ** 9) if r < H else 0.0


def lattice_density(spacing_over_h: Fraction):
    """Interior density of an infinite cubic lattice: sum over integer offsets n
    with |n| s < h of W(|n| s), by exact integer enumeration of the offsets."""
    s = spacing_over_h
    lim = int(math.floor(1 / s)) + 1
    total = 0.0
    count = 0
    for i in range(-lim, lim + 1):
        for j in range(-lim, lim + 1):
            for k in range(-lim, lim + 1):
                n2 = i * i + j * j + k * k
                if Fraction(n2) * s * s < 1:  # |n| s < h, decided exactly
                    count += 1
                    q = 1.0 - float(Fraction(n2) * s * s)  # (h^2 - r^2)/h^2
                    total += q ** 3
    rho = M * 315.0 / (64.0 * PI * H ** 3) * total
    return rho, count


def main():
    out = {}
    rho1 = M * 315.0 / (64.0 * PI * H ** 3)
    out["isolated"] = {
        "rho": rho1,
        "p": max(K * (rho1 - RHO0), 0.0),
        "force": [0.0, rho1 * G, 0.0],
        "vel1": [0.0, DT * G, 0.0],
    }
    # Pair along x at r = 0.6 h, equal velocities (viscosity vanishes).
    r = float(F32(0.6) * F32(H))
    rho_pair = rho1 + w_poly6(r)
    p_pair = K * (rho_pair - RHO0)
    mag = M * (2 * p_pair) / (2 * rho_pair) * 45.0 / (PI * H ** 6) * (H - r) ** 2
    out["pair"] = {"r": r, "rho": rho_pair, "p": p_pair,
                   "force0": [-mag, rho_pair * G, 0.0],   # particle 0 at x=0 is pushed to -x
                   "force1": [mag, rho_pair * G, 0.0]}
    # Same pair with particle 1 moving at +u along z: viscosity pulls 0 along +z, 1 along -z.
    u = 0.25
    fv = VISC * M * u / rho_pair * 45.0 / (PI * H ** 6) * (H - r)
    out["pair_visc"] = {"u": u, "fv0_z": fv, "fv1_z": -fv}
    # Free fall of an isolated particle for n steps: v_n = n dt g, x_n = x_0 + dt^2 g n(n+1)/2.
    n = 50
    out["free_fall"] = {"n": n, "v": n * DT * G, "dy": DT * DT * G * n * (n + 1) / 2}
    # Lattice sums (exact offset enumeration).
    rho_half, cnt_half = lattice_density(Fraction(1, 2))
    rho_quarter, cnt_quarter = lattice_density(Fraction(1, 4))
    out["lattice"] = {"half_h": {"rho": rho_half, "neighbours": cnt_half},
                      "quarter_h": {"rho": rho_quarter, "neighbours": cnt_quarter}}
    out["constants"] = {"h": H, "m": M, "visc": VISC, "rho0": RHO0, "k": K, "pi": PI, "g": G, "dt": DT, "damping": DAMP}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
