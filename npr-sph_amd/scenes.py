"""Synthetic scene definitions shared by the solver's callers (bench, tests).

Every scene is a seeded jittered lattice (sph_lattice_t) inside a wall box.
Conventions, all derived from the reference's own constants:

* h = smoothing_coeff * PARTICLE_RADIUS = 4 * 0.005 (float32), Main.cpp:113 and
  rho_pres_comp.glsl:5,40.
* Lattice spacing 0.5 h (number density 1e6 per unit volume), per-axis jitter
  uniform in [-0.05 h, +0.05 h), zero initial velocity.
* The box is the reference BoundaryUniform (Main.cpp:120-121,
  lower (-0.1,-0.35,-0.1), upper (0.5,1.0,0.5)) scaled about the origin so the
  fluid fills 40% of it at that number density.

Workloads:
  dam10k    10,000 particles, 25x16x25 block in the corner of the reference box
  dam1m     1,000,000 particles, 100^3 block, box scaled by (2.5/0.486)^(1/3)
  dam8m     8,000,000 particles, 200^3 block, box twice that of dam1m
  batch256  256 scenes, counts log-uniform in [2000, 200000], per-scene box,
            dam-break block or random column chosen per scene from the seed
  reference the literal make_grid scene of Main.cpp:488-505 (10x100x10 at
            spacing PARTICLE_RADIUS from the origin, no jitter)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

REF_LOWER = (-0.1, -0.35, -0.1)
REF_UPPER = (0.5, 1.0, 0.5)
REF_VOLUME = 0.6 * 1.35 * 0.6  # 0.486
NUMBER_DENSITY = 1.0e6          # particles per unit volume at spacing 0.5h = 0.01
FILL = 0.4
DEFAULT_SEED = 0x5EED5EED


def h32() -> np.float32:
    return np.float32(4.0) * np.float32(0.005)


def spacing32() -> np.float32:
    return np.float32(0.5) * h32()


def jitter32() -> np.float32:
    return np.float32(0.05) * h32()


def splitmix64(seed: int, counter: int) -> int:
    """SplitMix64 output function as a counter-based stream (same definition as
    the device generator)."""
    m = 0xFFFFFFFFFFFFFFFF
    z = (seed + (counter + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def uniform01(seed: int, counter: int) -> float:
    return (splitmix64(seed, counter) >> 11) * (1.0 / 9007199254740992.0)


@dataclass
class Scene:
    count: int
    lower: tuple
    upper: tuple
    nx: int
    ny: int
    nz: int
    origin: tuple
    spacing: float
    jitter: float
    seed: int
    layout: str = "dam"

    def lattice_args(self):
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, spacing=self.spacing, origin=self.origin,
                    jitter=self.jitter, seed=self.seed, count=self.count)


def _box(scale: float):
    lo = tuple(float(np.float32(v * scale)) for v in REF_LOWER)
    hi = tuple(float(np.float32(v * scale)) for v in REF_UPPER)
    return lo, hi


def _corner_origin(lo, sp):
    return tuple(float(np.float32(v) + np.float32(0.5) * sp) for v in lo)


def box_scale(count: int) -> float:
    """Scale of the reference box that puts `count` particles at 40% fill."""
    return (count / (FILL * NUMBER_DENSITY * REF_VOLUME)) ** (1.0 / 3.0)


def dam10k(seed: int = DEFAULT_SEED) -> Scene:
    sp, jt = spacing32(), jitter32()
    lo, hi = REF_LOWER, REF_UPPER
    return Scene(10000, lo, hi, 25, 16, 25, _corner_origin(lo, sp), float(sp), float(jt), seed)


def dam_block(n_side: int, scale: float, seed: int = DEFAULT_SEED) -> Scene:
    sp, jt = spacing32(), jitter32()
    lo, hi = _box(scale)
    n = n_side ** 3
    return Scene(n, lo, hi, n_side, n_side, n_side, _corner_origin(lo, sp), float(sp), float(jt), seed)


def dam1m(seed: int = DEFAULT_SEED) -> Scene:
    return dam_block(100, box_scale(1_000_000), seed)


def dam8m(seed: int = DEFAULT_SEED) -> Scene:
    return dam_block(200, 2.0 * box_scale(1_000_000), seed)


def scaled_dam(count: int, seed: int = DEFAULT_SEED, layout: str = "dam") -> Scene:
    """A scene of `count` particles in a box scaled to 40% fill.  "dam": block
    of side floor(count^(1/3)) in the lower corner, topped up layer by layer;
    "column": narrower footprint floor(0.7 count^(1/3)) at a seeded x/z offset."""
    sp, jt = spacing32(), jitter32()
    scale = box_scale(count)
    lo, hi = _box(scale)
    side = count ** (1.0 / 3.0)
    if layout == "dam":
        nx = nz = max(1, int(math.floor(side + 1e-9)))
    else:
        nx = nz = max(1, int(math.floor(0.7 * side)))
    ny = -(-count // (nx * nz))
    org = list(_corner_origin(lo, sp))
    if layout != "dam":
        for a, key in ((0, 1), (2, 2)):
            room = (hi[a] - lo[a]) - (nx * float(sp)) - float(sp)
            org[a] = float(np.float32(org[a] + max(room, 0.0) * uniform01(seed, key)))
    return Scene(count, lo, hi, nx, ny, nz, tuple(org), float(sp), float(jt), seed, layout)


def batch256(seed: int = DEFAULT_SEED, nscenes: int = 256, nmin: int = 2000, nmax: int = 200000):
    scenes = []
    lmin, lmax = math.log(nmin), math.log(nmax)
    for s in range(nscenes):
        u = uniform01(seed, 2 * s)
        count = int(math.floor(math.exp(lmin + (lmax - lmin) * u)))
        count = min(max(count, nmin), nmax)
        layout = "dam" if uniform01(seed, 2 * s + 1) < 0.5 else "column"
        scenes.append(scaled_dam(count, splitmix64(seed, 100000 + s), layout))
    return scenes


def reference(seed: int = 0) -> Scene:
    """make_grid of Main.cpp:488-505 (index = 1000 i + 10 j + k, no jitter)."""
    pr = float(np.float32(0.005))
    return Scene(10000, REF_LOWER, REF_UPPER, 10, 100, 10, (0.0, 0.0, 0.0), pr, 0.0, seed, "reference")


WORKLOADS = {"dam10k": dam10k, "dam1m": dam1m, "dam8m": dam8m}
