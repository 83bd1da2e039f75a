"""Python face of the FP64 CPU oracle (oracle/sph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP solver, never
as the thing measured or shipped.

Every function restates one reference pass (citations relative to the
reference repository root):
  density()    NPR-SPH/rho_pres_comp.glsl:35-58
  forces()     NPR-SPH/force_comp.glsl:35-66
  integrate()  NPR-SPH/integrate_comp.glsl:35-81
  lattice()    NPR-SPH/Main.cpp:488-539 plus seeded jitter
  grid_build() float32 cell hash + stable counting sort (no reference counterpart)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "sph_oracle.c")

# -ffp-contract=off keeps the float32 cell hash free of fused multiply-adds so
# it equals the GPU hash bit for bit; x86-64-v3 runs on any AVX2 host.
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-fPIC", "-shared", "-std=c11"]

PARTICLE_DTYPE = np.dtype([("pos", "<f4", 4), ("vel", "<f4", 4), ("force", "<f4", 4), ("extras", "<f4", 4)])
assert PARTICLE_DTYPE.itemsize == 64


def build(verbose: bool = False) -> str:
    cmd = ["gcc", *CFLAGS, "-o", LIB_PATH, SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return LIB_PATH


class Constants(C.Structure):
    _fields_ = [("mass", C.c_float), ("smoothing_coeff", C.c_float), ("visc", C.c_float), ("resting_rho", C.c_float)]


class Boundary(C.Structure):
    _fields_ = [("upper", C.c_float * 4), ("lower", C.c_float * 4)]


class Fixed(C.Structure):
    _fields_ = [("particle_radius", C.c_float), ("gas_const", C.c_float), ("gravity", C.c_float * 3),
                ("damping", C.c_float), ("dt", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("inv_cs", C.c_float), ("dims", C.c_int32 * 3), ("ncells", C.c_uint32),
                ("sub", C.c_int32), ("cs", C.c_double)]


def default_constants() -> Constants:
    """ConstantsUniform defaults, Main.cpp:110-116."""
    return Constants(0.02, 4.0, 3000.0, 1000.0)


def default_boundary() -> Boundary:
    """BoundaryUniform defaults, Main.cpp:118-122."""
    b = Boundary()
    b.upper[:] = (0.5, 1.0, 0.5, 1.0)
    b.lower[:] = (-0.1, -0.35, -0.1, 1.0)
    return b


def default_fixed() -> Fixed:
    """Shader-baked constants: PARTICLE_RADIUS (rho_pres_comp.glsl:5), GAS_CONST
    (:33), G (force_comp.glsl:33), DAMPING and dt (integrate_comp.glsl:8,33)."""
    f = Fixed()
    f.particle_radius = 0.005
    f.gas_const = 2000.0
    f.gravity[:] = (0.0, -9806.65, 0.0)
    f.damping = 0.3
    f.dt = np.float32(1.0) / np.float32(10000.0)
    return f


def make_boundary(lower, upper) -> Boundary:
    b = Boundary()
    b.lower[:] = (*[float(v) for v in lower], 1.0)
    b.upper[:] = (*[float(v) for v in upper], 1.0)
    return b


_lib = None
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_P = np.ctypeslib.ndpointer(dtype=PARTICLE_DTYPE, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC):
            build()
        L = C.CDLL(LIB_PATH)
        cp, bp, fp = C.POINTER(Constants), C.POINTER(Boundary), C.POINTER(Fixed)
        L.orc_smoothing_length.argtypes = [cp, fp]
        L.orc_smoothing_length.restype = C.c_float
        L.orc_grid_geometry.argtypes = [C.c_float, bp, C.c_uint64, C.c_int, C.POINTER(Grid)]
        L.orc_grid_geometry.restype = C.c_int
        L.orc_cell_key.argtypes = [C.POINTER(Grid), C.c_float, C.c_float, C.c_float]
        L.orc_cell_key.restype = C.c_uint32
        L.orc_grid_build.argtypes = [C.c_uint64, _F, C.c_uint64, C.POINTER(Grid), _U32, _U32, _U32]
        L.orc_grid_build.restype = None
        L.orc_density_bf.argtypes = [C.c_uint64, _D, cp, fp, _D, _D]
        L.orc_density_bf.restype = None
        L.orc_density_grid.argtypes = [C.c_uint64, _D, cp, fp, bp, _D, _D]
        L.orc_density_grid.restype = C.c_int
        L.orc_forces_bf.argtypes = [C.c_uint64, _D, _D, _D, _D, cp, fp, _D, _D]
        L.orc_forces_bf.restype = None
        L.orc_forces_grid.argtypes = [C.c_uint64, _D, _D, _D, _D, cp, fp, bp, _D, _D]
        L.orc_forces_grid.restype = C.c_int
        L.orc_count_pairs_grid.argtypes = [C.c_uint64, _D, cp, fp, bp, C.c_double, C.POINTER(C.c_uint64)]
        L.orc_count_pairs_grid.restype = C.c_uint64
        L.orc_integrate.argtypes = [C.c_uint64, _D, _D, _D, _D, bp, fp, C.c_double, _U8]
        L.orc_integrate.restype = None
        L.orc_step.argtypes = [C.c_uint64, _D, _D, _D, _D, _D, cp, bp, fp, C.c_int]
        L.orc_step.restype = C.c_int
        L.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_lattice.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _F, C.c_float, C.c_float,
                                  C.c_uint64, _P]
        L.orc_lattice.restype = None
        L.orc_reference_grid.argtypes = [_P]
        L.orc_reference_grid.restype = None
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


def _c(c):
    return default_constants() if c is None else c


def _b(b):
    return default_boundary() if b is None else b


def _f(f):
    return default_fixed() if f is None else f


def smoothing_length(c=None, f=None) -> np.float32:
    return np.float32(lib().orc_smoothing_length(C.byref(_c(c)), C.byref(_f(f))))


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ------------------------------------------------------------------ state ---

def particles(n: int) -> np.ndarray:
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["pos"][:, 3] = 1.0
    return p


def pos64(p: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(p["pos"][:, :3], dtype=np.float64).reshape(-1)


def vel64(p: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(p["vel"][:, :3], dtype=np.float64).reshape(-1)


# ----------------------------------------------------------------- passes ---

def density(x: np.ndarray, c=None, f=None, b=None, grid: bool = True):
    """rho_pres_comp.glsl:35-58 in FP64.  x: (n,3) or flat 3n positions.
    Returns (rho[n], p[n])."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    n = x.size // 3
    rho = np.empty(n)
    p = np.empty(n)
    if grid:
        rc = lib().orc_density_grid(n, x, C.byref(_c(c)), C.byref(_f(f)), C.byref(_b(b)), rho, p)
        if rc != 0:
            raise ValueError("invalid grid geometry")
    else:
        lib().orc_density_bf(n, x, C.byref(_c(c)), C.byref(_f(f)), rho, p)
    return rho, p


def forces(x, v, rho, p, c=None, f=None, b=None, grid: bool = True):
    """force_comp.glsl:35-66 in FP64.  Returns (force (n,3), scale[n]) where
    scale bounds the magnitude of the summed terms."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    n = x.size // 3
    out = np.empty(3 * n)
    scale = np.empty(n)
    if grid:
        rc = lib().orc_forces_grid(n, x, v, rho, p, C.byref(_c(c)), C.byref(_f(f)), C.byref(_b(b)), out, scale)
        if rc != 0:
            raise ValueError("invalid grid geometry")
    else:
        lib().orc_forces_bf(n, x, v, rho, p, C.byref(_c(c)), C.byref(_f(f)), out, scale)
    return out.reshape(n, 3), scale


def count_pairs(x, c=None, f=None, b=None, margin: float = 1e-6):
    """Ordered neighbour pairs (j != i, r < h) in FP64 over the cell stencil
    (force_comp.glsl:50-57), and how many of them lie within `margin` of
    r^2 = h^2, where float32 rounding may decide the side."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    amb = C.c_uint64()
    total = lib().orc_count_pairs_grid(x.size // 3, x, C.byref(_c(c)), C.byref(_f(f)), C.byref(_b(b)), float(margin),
                                       C.byref(amb))
    if total == 2**64 - 1:
        raise ValueError("invalid grid geometry")
    return int(total), int(amb.value)


def integrate(x, v, force, rho, b=None, f=None, wall_eps: float = 0.0):
    """integrate_comp.glsl:35-81 in FP64.  Returns (x', v', near_wall flags)."""
    x = np.array(x, dtype=np.float64).reshape(-1)
    v = np.array(v, dtype=np.float64).reshape(-1)
    force = np.ascontiguousarray(force, dtype=np.float64).reshape(-1)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = x.size // 3
    flags = np.zeros(n, dtype=np.uint8)
    lib().orc_integrate(n, x, v, force, rho, C.byref(_b(b)), C.byref(_f(f)), float(wall_eps), flags)
    return x.reshape(n, 3), v.reshape(n, 3), flags


class State64:
    """FP64 particle state for multi-step oracle runs (Main.cpp:295-303)."""

    def __init__(self, p: np.ndarray):
        n = len(p)
        self.n = n
        self.x = pos64(p)
        self.v = vel64(p)
        self.f = np.zeros(3 * n)
        self.rho = np.zeros(n)
        self.p = np.zeros(n)

    def step(self, c=None, b=None, f=None, grid: bool = True, nsteps: int = 1):
        for _ in range(nsteps):
            rc = lib().orc_step(self.n, self.x, self.v, self.f, self.rho, self.p, C.byref(_c(c)), C.byref(_b(b)),
                                C.byref(_f(f)), 1 if grid else 0)
            if rc != 0:
                raise ValueError("invalid grid geometry")

    def to_particles(self, template: np.ndarray | None = None) -> np.ndarray:
        out = particles(self.n) if template is None else template.copy()
        out["pos"][:, :3] = self.x.reshape(-1, 3)
        out["vel"][:, :3] = self.v.reshape(-1, 3)
        out["force"][:, :3] = self.f.reshape(-1, 3)
        out["extras"][:, 0] = self.rho
        out["extras"][:, 1] = self.p
        return out


# ------------------------------------------------------------------- grid ---

def grid_geometry(h, b=None, n: int = 0, sub: int = 2) -> Grid:
    """Grid geometry rule shared with the solver (sub = cells per smoothing length)."""
    g = Grid()
    if lib().orc_grid_geometry(float(h), C.byref(_b(b)), int(n), int(sub), C.byref(g)) != 0:
        raise ValueError("invalid grid geometry")
    return g


def grid_build(pos: np.ndarray, g: Grid):
    """Stable counting sort by float32 cell key.  pos: (n,>=3) float32.
    Returns keys[n], cell_start[ncells+2], perm[n]."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n = pos.shape[0]
    stride = pos.shape[1] if pos.ndim == 2 else 3
    flat = pos.reshape(-1) if n else np.zeros(1, np.float32)
    keys = np.empty(max(n, 1), dtype=np.uint32)
    start = np.empty(g.ncells + 2, dtype=np.uint32)
    perm = np.empty(max(n, 1), dtype=np.uint32)
    lib().orc_grid_build(n, flat, stride, C.byref(g), keys, start, perm)
    return keys[:n], start, perm[:n]


def cell_key(g: Grid, x, y, z) -> int:
    return int(lib().orc_cell_key(C.byref(g), float(x), float(y), float(z)))


# ------------------------------------------------------------- generators ---

def splitmix64(seed: int, counter: int) -> int:
    return int(lib().orc_splitmix64(seed & 0xFFFFFFFFFFFFFFFF, counter & 0xFFFFFFFFFFFFFFFF))


def lattice(nx, ny, nz, spacing, origin, jitter, seed, count=None) -> np.ndarray:
    """Seeded jittered lattice (sph_lattice_t in include/sph.h)."""
    count = nx * ny * nz if count is None else count
    out = np.zeros(max(count, 1), dtype=PARTICLE_DTYPE)
    org = np.ascontiguousarray(origin, dtype=np.float32)
    lib().orc_lattice(nx, ny, nz, count, org, np.float32(spacing), np.float32(jitter), seed, out)
    return out[:count]


def reference_grid() -> np.ndarray:
    """make_grid/init_particles of Main.cpp:488-539 (10x100x10, no jitter)."""
    out = np.zeros(10000, dtype=PARTICLE_DTYPE)
    lib().orc_reference_grid(out)
    return out
