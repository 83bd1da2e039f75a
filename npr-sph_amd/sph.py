"""Host-side mirror of NPR-SPH's simulation surface over the C ABI (include/sph.h).

The reference drives its solver from NPR-SPH/Main.cpp: init_particles
(:510-539) uploads the particle SSBO, sendUniforms (:254-282) uploads the
constant blocks, and display (:291-304) dispatches rho_pres_comp, force_comp and
integrate_comp with a barrier after each.  `Simulation` offers the same
operations (upload, set_constants/set_boundary, rho_pres, force, integrate, step)
on top of libnprsph.so.  Errors surface as `SphError` carrying the library's
message, the counterpart of InitShader's logged failure (InitShader.cpp:47-112).

The library is loaded from this directory; there is no fallback path: if the
HIP library is missing or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnprsph.so")

PARTICLE_DTYPE = np.dtype([("pos", "<f4", 4), ("vel", "<f4", 4), ("force", "<f4", 4), ("extras", "<f4", 4)])

SPH_OK, SPH_ERR_INVALID, SPH_ERR_HIP, SPH_ERR_NOMEM, SPH_ERR_STATE = 0, -1, -2, -3, -4

# Every symbol include/sph.h declares.
EXPORTS = [
    "sph_defaults", "sph_create", "sph_create_batch", "sph_create_error", "sph_destroy", "sph_last_error",
    "sph_size", "sph_num_scenes", "sph_set_constants", "sph_set_boundary", "sph_set_fixed", "sph_get_params",
    "sph_upload", "sph_download", "sph_init_lattice", "sph_rho_pres", "sph_force", "sph_integrate", "sph_step",
    "sph_device_buffer", "sph_stream", "sph_synchronize", "sph_grid_info", "sph_total_cells", "sph_download_grid",
    "sph_metrics", "sph_set_timing", "sph_kernel_times", "sph_get_tuning", "sph_set_tuning", "sph_autotune",
    "sph_mlp_set", "sph_mlp_get", "sph_mlp_train", "sph_mlp_predict", "sph_coarse_step", "sph_copy_state",
    "sph_save", "sph_load", "sph_metrics_noalloc",
]

# Snapshot files (sph_save / sph_load): 64-byte header, per-scene counts, records.
SNAPSHOT_MAGIC = b"NPRSPHv1"
SNAPSHOT_HEADER = np.dtype([("magic", "S8"), ("header_bytes", "<u4"), ("record_bytes", "<u4"), ("count", "<u8"),
                            ("nscenes", "<u4"), ("reserved0", "<u4"), ("frame", "<u8"), ("reserved", "V24")])
assert SNAPSHOT_HEADER.itemsize == 64


def write_snapshot(path: str, particles: np.ndarray, counts=None, frame: int = 0) -> None:
    """Write particles (PARTICLE_DTYPE, original index order) in the sph_save format."""
    particles = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
    counts = np.asarray([len(particles)] if counts is None else counts, dtype="<u8")
    if int(counts.sum()) != len(particles):
        raise ValueError("scene counts must add up to the particle count")
    h = np.zeros(1, dtype=SNAPSHOT_HEADER)
    h["magic"], h["header_bytes"], h["record_bytes"] = SNAPSHOT_MAGIC, 64, PARTICLE_DTYPE.itemsize
    h["count"], h["nscenes"], h["frame"] = len(particles), len(counts), frame
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(counts.tobytes())
        f.write(particles.tobytes())


def read_snapshot(path: str):
    """(particles, per-scene counts, frame) of a file written by sph_save,
    without the library: what a renderer or an offline tool reads."""
    with open(path, "rb") as f:
        h = np.frombuffer(f.read(64), dtype=SNAPSHOT_HEADER)
        if len(h) != 1 or h["magic"][0] != SNAPSHOT_MAGIC:
            raise ValueError(f"{path}: not a snapshot file")
        if h["header_bytes"][0] != 64 or h["record_bytes"][0] != PARTICLE_DTYPE.itemsize:
            raise ValueError(f"{path}: unsupported snapshot layout")
        counts = np.frombuffer(f.read(8 * int(h["nscenes"][0])), dtype="<u8")
        n = int(h["count"][0])
        particles = np.frombuffer(f.read(n * PARTICLE_DTYPE.itemsize), dtype=PARTICLE_DTYPE)
        if len(particles) != n or len(counts) != int(h["nscenes"][0]):
            raise ValueError(f"{path}: truncated snapshot")
    return particles.copy(), counts.copy(), int(h["frame"][0])

MLP_PARAMS = 4995


class SphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sph error {code}: {msg}")
        self.code = code


class Constants(C.Structure):
    """ConstantsUniform, Main.cpp:110-116."""
    _fields_ = [("mass", C.c_float), ("smoothing_coeff", C.c_float), ("visc", C.c_float), ("resting_rho", C.c_float)]


class Boundary(C.Structure):
    """BoundaryUniform, Main.cpp:118-122."""
    _fields_ = [("upper", C.c_float * 4), ("lower", C.c_float * 4)]


class Fixed(C.Structure):
    """Shader-baked constants (PARTICLE_RADIUS, GAS_CONST, G, DAMPING, dt)."""
    _fields_ = [("particle_radius", C.c_float), ("gas_const", C.c_float), ("gravity", C.c_float * 3),
                ("damping", C.c_float), ("dt", C.c_float)]


class Lattice(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("count", C.c_uint32),
                ("origin", C.c_float * 3), ("spacing", C.c_float), ("jitter", C.c_float), ("seed", C.c_uint64)]


# Cells are this many times narrower along x than along y and z
# (SPH_GRID_X_SPLIT in include/sph.h): the x coordinate of a cell uses
# GRID_X_SPLIT * inv_cell.
GRID_X_SPLIT = 2


class GridInfo(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("inv_cell", C.c_float), ("cell", C.c_double), ("dims", C.c_int32 * 3),
                ("ncells", C.c_uint32), ("cell_base", C.c_uint32), ("first", C.c_uint32), ("count", C.c_uint32),
                ("cells_per_h", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("cells_per_h", C.c_int32), ("prune", C.c_int32), ("block", C.c_int32), ("blocks_per_xcd", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MlpNorm(C.Structure):
    _fields_ = [("v_scale", C.c_float), ("rho_scale", C.c_float), ("pi_scale", C.c_float), ("acc_scale", C.c_float)]


class TrainOpts(C.Structure):
    _fields_ = [("first", C.c_uint32), ("stride", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("update", C.c_int32)]


class TrainResult(C.Structure):
    _fields_ = [("loss", C.c_double), ("pairs", C.c_uint64), ("targets", C.c_uint32), ("weighted", C.c_uint32)]


class CoarseOpts(C.Structure):
    _fields_ = [("use_network", C.c_int32), ("judge_every", C.c_int32), ("drift_limit", C.c_float),
                ("cooldown", C.c_int32)]


class CoarseStats(C.Structure):
    _fields_ = [("coarse_steps", C.c_int64), ("plain_steps", C.c_int64), ("judged", C.c_int64),
                ("fallbacks", C.c_int64), ("last_drift", C.c_double), ("max_drift", C.c_double),
                ("pairs", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Metrics(C.Structure):
    _fields_ = [("mean_rho", C.c_double), ("kinetic", C.c_double), ("com", C.c_double * 3),
                ("max_speed", C.c_double), ("nonfinite", C.c_uint64)]


_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library(path: str | None = None):
    """Load libnprsph.so and declare its signatures (no device work)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not built; run __graft_entry__.build()")
    L = C.CDLL(p)
    vp = C.c_void_p
    cp, bp, fp = C.POINTER(Constants), C.POINTER(Boundary), C.POINTER(Fixed)
    sig = {
        "sph_defaults": (None, [cp, bp, fp]),
        "sph_create": (C.c_int, [C.POINTER(vp), C.c_size_t, cp, bp, fp, vp]),
        "sph_create_batch": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(C.c_size_t), cp, bp, fp, vp]),
        "sph_create_error": (C.c_char_p, []),
        "sph_destroy": (None, [vp]),
        "sph_last_error": (C.c_char_p, [vp]),
        "sph_size": (C.c_size_t, [vp]),
        "sph_num_scenes": (C.c_int, [vp]),
        "sph_set_constants": (C.c_int, [vp, C.c_int, cp]),
        "sph_set_boundary": (C.c_int, [vp, C.c_int, bp]),
        "sph_set_fixed": (C.c_int, [vp, C.c_int, fp]),
        "sph_get_params": (C.c_int, [vp, C.c_int, cp, bp, fp]),
        "sph_upload": (C.c_int, [vp, vp, C.c_size_t, C.c_size_t]),
        "sph_download": (C.c_int, [vp, vp, C.c_size_t, C.c_size_t]),
        "sph_init_lattice": (C.c_int, [vp, C.c_int, C.POINTER(Lattice)]),
        "sph_rho_pres": (C.c_int, [vp]),
        "sph_force": (C.c_int, [vp]),
        "sph_integrate": (C.c_int, [vp]),
        "sph_step": (C.c_int, [vp, C.c_int]),
        "sph_device_buffer": (C.c_int, [vp, C.POINTER(vp), C.POINTER(C.c_size_t)]),
        "sph_stream": (vp, [vp]),
        "sph_synchronize": (C.c_int, [vp]),
        "sph_grid_info": (C.c_int, [vp, C.c_int, C.POINTER(GridInfo)]),
        "sph_total_cells": (C.c_size_t, [vp]),
        "sph_download_grid": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t]),
        "sph_save": (C.c_int, [vp, C.c_char_p, C.c_uint64]),
        "sph_load": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_uint64)]),
        "sph_metrics": (C.c_int, [vp, C.c_int, C.POINTER(Metrics)]),
        "sph_metrics_noalloc": (C.c_int, [vp, C.c_int, C.POINTER(Metrics)]),
        "sph_set_timing": (C.c_int, [vp, C.c_int]),
        "sph_kernel_times": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "sph_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "sph_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "sph_autotune": (C.c_int, [vp, C.c_int, C.POINTER(Tuning), C.POINTER(C.c_double)]),
        "sph_mlp_set": (C.c_int, [vp, vp, C.c_size_t, C.POINTER(MlpNorm)]),
        "sph_mlp_get": (C.c_int, [vp, vp, C.c_size_t, C.POINTER(MlpNorm)]),
        "sph_mlp_train": (C.c_int, [vp, vp, C.POINTER(TrainOpts), vp, C.POINTER(TrainResult)]),
        "sph_mlp_predict": (C.c_int, [vp, C.c_int, vp]),
        "sph_coarse_step": (C.c_int, [vp, C.c_int, C.POINTER(CoarseOpts), C.POINTER(CoarseStats)]),
        "sph_copy_state": (C.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def defaults():
    c, b, f = Constants(), Boundary(), Fixed()
    load_library().sph_defaults(C.byref(c), C.byref(b), C.byref(f))
    return c, b, f


def make_boundary(lower, upper) -> Boundary:
    b = Boundary()
    b.lower[:] = (*[float(v) for v in lower], 1.0)
    b.upper[:] = (*[float(v) for v in upper], 1.0)
    return b


def make_lattice(nx, ny, nz, spacing, origin, jitter, seed, count=None) -> Lattice:
    lat = Lattice()
    lat.nx, lat.ny, lat.nz = int(nx), int(ny), int(nz)
    lat.count = int(nx * ny * nz if count is None else count)
    lat.origin[:] = [float(v) for v in origin]
    lat.spacing = float(np.float32(spacing))
    lat.jitter = float(np.float32(jitter))
    lat.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return lat


class Simulation:
    """One solver context (one or more scenes sharing every launch)."""

    def __init__(self, counts, constants=None, boundaries=None, fixed=None, stream=None):
        L = load_library()
        self._L = L
        counts = [int(counts)] if np.isscalar(counts) else [int(x) for x in counts]
        ns = len(counts)
        arr = (C.c_size_t * ns)(*counts)

        def _arr(kind, vals):
            if vals is None:
                return None
            vals = list(vals) if isinstance(vals, (list, tuple)) else [vals] * ns
            a = (kind * ns)(*vals)
            return a

        self._c = _arr(Constants, constants)
        self._b = _arr(Boundary, boundaries)
        self._f = _arr(Fixed, fixed)
        h = C.c_void_p()
        rc = L.sph_create_batch(C.byref(h), ns, arr, self._c, self._b, self._f, stream)
        if rc != SPH_OK:
            raise SphError(rc, L.sph_create_error().decode())
        self._h = h
        self.counts = counts
        self.n = sum(counts)
        self.first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != SPH_OK:
            raise SphError(rc, self._L.sph_last_error(self._h).decode())
        return rc

    @property
    def handle(self):
        return self._h

    # -- parameters (sendUniforms, Main.cpp:254-282) -----------------------
    def set_constants(self, c: Constants, scene: int = -1):
        self._check(self._L.sph_set_constants(self._h, scene, C.byref(c)))

    def set_boundary(self, b: Boundary, scene: int = -1):
        self._check(self._L.sph_set_boundary(self._h, scene, C.byref(b)))

    def set_fixed(self, f: Fixed, scene: int = -1):
        self._check(self._L.sph_set_fixed(self._h, scene, C.byref(f)))

    def params(self, scene: int = 0):
        c, b, f = Constants(), Boundary(), Fixed()
        self._check(self._L.sph_get_params(self._h, scene, C.byref(c), C.byref(b), C.byref(f)))
        return c, b, f

    # -- state --------------------------------------------------------------
    def upload(self, particles: np.ndarray, first: int = 0):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        self._check(self._L.sph_upload(self._h, p.ctypes.data, first, len(p)))

    def download(self, first: int = 0, n: int | None = None) -> np.ndarray:
        n = self.n - first if n is None else n
        out = np.empty(n, dtype=PARTICLE_DTYPE)
        self._check(self._L.sph_download(self._h, out.ctypes.data, first, n))
        return out

    def download_into(self, out: np.ndarray, first: int = 0) -> np.ndarray:
        """sph_download into a caller's contiguous PARTICLE_DTYPE array."""
        if out.dtype != PARTICLE_DTYPE or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous PARTICLE_DTYPE array")
        self._check(self._L.sph_download(self._h, out.ctypes.data, first, len(out)))
        return out

    def init_lattice(self, lat: Lattice, scene: int = 0):
        self._check(self._L.sph_init_lattice(self._h, scene, C.byref(lat)))

    # -- the three passes (Main.cpp:295-303) -------------------------------
    def rho_pres(self):
        self._check(self._L.sph_rho_pres(self._h))

    def force(self):
        self._check(self._L.sph_force(self._h))

    def integrate(self):
        self._check(self._L.sph_integrate(self._h))

    def step(self, nsteps: int = 1):
        self._check(self._L.sph_step(self._h, int(nsteps)))

    def prepare(self):
        """sph_step(ctx, 0): build the sorted state and capture the step graphs
        without advancing, so that the next step() call only launches."""
        self._check(self._L.sph_step(self._h, 0))

    def synchronize(self):
        self._check(self._L.sph_synchronize(self._h))

    def stream(self):
        return self._L.sph_stream(self._h)

    def device_buffer(self):
        ptr, n = C.c_void_p(), C.c_size_t()
        self._check(self._L.sph_device_buffer(self._h, C.byref(ptr), C.byref(n)))
        return ptr.value, n.value

    # -- diagnostics --------------------------------------------------------
    def grid_info(self, scene: int = 0) -> GridInfo:
        g = GridInfo()
        self._check(self._L.sph_grid_info(self._h, scene, C.byref(g)))
        return g

    def total_cells(self) -> int:
        return int(self._L.sph_total_cells(self._h))

    def download_grid(self):
        nc = self.total_cells() + 1
        start = np.empty(nc, dtype=np.uint32)
        perm = np.empty(self.n, dtype=np.uint32)
        self._check(self._L.sph_download_grid(self._h, start.ctypes.data, nc, perm.ctypes.data, self.n))
        return start, perm

    def save(self, path: str, frame: int = 0):
        """Write the particle buffer to `path` (sph_save)."""
        self._check(self._L.sph_save(self._h, os.fsencode(path), int(frame)))

    def load(self, path: str) -> int:
        """Replace the state with a file written by save(); returns its frame (sph_load)."""
        fr = C.c_uint64()
        self._check(self._L.sph_load(self._h, os.fsencode(path), C.byref(fr)))
        return fr.value

    def metrics(self, scene: int = 0) -> Metrics:
        m = Metrics()
        self._check(self._L.sph_metrics(self._h, scene, C.byref(m)))
        return m

    def metrics_noalloc(self, scene: int = 0) -> Metrics:
        """sph_metrics_noalloc: the same reductions with no allocation (static
        device scratch; scenes of at most 4,194,304 particles)."""
        m = Metrics()
        self._check(self._L.sph_metrics_noalloc(self._h, scene, C.byref(m)))
        return m

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(self._L.sph_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, cells_per_h=2, prune=1, block=128, blocks_per_xcd=0):
        t = Tuning(cells_per_h, prune, block, blocks_per_xcd)
        self._check(self._L.sph_set_tuning(self._h, C.byref(t)))

    def autotune(self, trial_steps: int = 10):
        t, ms = Tuning(), C.c_double()
        self._check(self._L.sph_autotune(self._h, int(trial_steps), C.byref(t), C.byref(ms)))
        return t, ms.value

    # -- learned coarse step -------------------------------------------------
    def mlp_set(self, params, norm: MlpNorm):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._L.sph_mlp_set(self._h, p.ctypes.data, p.size, C.byref(norm)))

    def mlp_get(self):
        p = np.empty(MLP_PARAMS, dtype=np.float32)
        nm = MlpNorm()
        self._check(self._L.sph_mlp_get(self._h, p.ctypes.data, p.size, C.byref(nm)))
        return p, nm

    def mlp_train(self, later: "Simulation", first=0, stride=1, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
                  update=True, want_grad=False):
        o = TrainOpts(first, stride, lr, beta1, beta2, eps, 1 if update else 0)
        r = TrainResult()
        g = np.empty(MLP_PARAMS, dtype=np.float32) if want_grad else None
        self._check(self._L.sph_mlp_train(self._h, later.handle, C.byref(o), None if g is None else g.ctypes.data,
                                          C.byref(r)))
        return (r, g) if want_grad else r

    def mlp_predict(self, path: int = 0) -> np.ndarray:
        out = np.empty((self.n, 3), dtype=np.float32)
        self._check(self._L.sph_mlp_predict(self._h, int(path), out.ctypes.data))
        return out

    def coarse_step(self, n=1, use_network=True, judge_every=0, drift_limit=0.01, cooldown=0) -> CoarseStats:
        o = CoarseOpts(1 if use_network else 0, judge_every, drift_limit, cooldown)
        st = CoarseStats()
        self._check(self._L.sph_coarse_step(self._h, int(n), C.byref(o), C.byref(st)))
        return st

    def copy_state_from(self, other: "Simulation"):
        self._check(self._L.sph_copy_state(self._h, other.handle))

    def set_timing(self, on: bool):
        self._check(self._L.sph_set_timing(self._h, 1 if on else 0))

    def kernel_times(self) -> dict:
        k = 64
        names = (C.c_char * 32 * k)()
        ms = (C.c_double * k)()
        cnt = (C.c_int64 * k)()
        got = self._L.sph_kernel_times(self._h, k, C.cast(names, C.c_void_p), C.cast(ms, C.c_void_p),
                                       C.cast(cnt, C.c_void_p))
        if got < 0:
            self._check(got)
        return {names[i].value.decode(): (ms[i], cnt[i]) for i in range(got)}
