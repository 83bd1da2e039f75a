"""The C ABI library loads and exports what include/sph.h declares (no GPU calls).

Also pins the ABI structs to the reference's byte layout: the 64-byte
std430 Particle (Main.cpp:93-99), the 16-byte ConstantsUniform
(Main.cpp:110-116) and the 32-byte BoundaryUniform (Main.cpp:118-122).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg, gpu_lib):
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(gpu_lib, n), f"libnprsph.so lacks {n}"
    assert sorted(pkg.sph.EXPORTS) == names


def test_no_oracle_symbols_in_product(pkg):
    """The product library never links the checker."""
    out = os.popen(f"nm -D --defined-only {pkg.sph.library_path()}").read()
    assert "orc_" not in out
    assert "sph_rho_pres" in out


def test_struct_layouts(pkg):
    s = pkg.sph
    assert s.PARTICLE_DTYPE.itemsize == 64
    assert C.sizeof(s.Constants) == 16
    assert C.sizeof(s.Boundary) == 32
    assert C.sizeof(s.Fixed) == 28
    assert C.sizeof(s.Lattice) == 48
    assert C.sizeof(s.GridInfo) == 56
    assert C.sizeof(s.Tuning) == 16
    assert C.sizeof(s.MlpNorm) == 16 and C.sizeof(s.TrainOpts) == 28 and C.sizeof(s.TrainResult) == 24
    assert C.sizeof(s.CoarseOpts) == 16 and C.sizeof(s.CoarseStats) == 56
    assert C.sizeof(s.Metrics) == 56
    assert s.PARTICLE_DTYPE.fields["pos"][1] == 0 and s.PARTICLE_DTYPE.fields["extras"][1] == 48


def test_grid_layout_constants_agree(pkg):
    """The x split and the z-band height of the cell grid are declared once in
    the header; the Python binding, the bench's candidate counter and the
    oracle's CPU grid build use the same values."""
    hdr = open(os.path.join(ROOT, "include", "sph.h")).read()
    split = re.search(r"#define SPH_GRID_X_SPLIT (\d+)", hdr).group(1)
    band = re.search(r"#define SPH_GRID_BAND (\d+)", hdr).group(1)
    assert int(split) == pkg.sph.GRID_X_SPLIT
    bench = open(os.path.join(ROOT, "bench.py")).read()
    assert re.search(r"^X_SPLIT = (\d+)$", bench, re.M).group(1) == split
    assert re.search(r"^CELL_BAND = (\d+)$", bench, re.M).group(1) == band
    oracle = open(os.path.join(ROOT, "oracle", "sph_oracle.c")).read()
    assert re.search(r"#define ORC_XSPLIT (\d+)", oracle).group(1) == split
    assert re.search(r"#define ORC_BAND (\d+)", oracle).group(1) == band


def test_defaults_are_the_references(pkg, gpu_lib):
    c, b, f = pkg.sph.defaults()
    assert (c.mass, c.smoothing_coeff, c.visc, c.resting_rho) == (np.float32(0.02), 4.0, 3000.0, 1000.0)
    assert list(b.upper) == [0.5, 1.0, 0.5, 1.0]
    assert list(b.lower) == [np.float32(-0.1), np.float32(-0.35), np.float32(-0.1), 1.0]
    assert f.particle_radius == np.float32(0.005) and f.gas_const == 2000.0
    assert list(f.gravity) == [0.0, np.float32(-9806.65), 0.0]
    assert f.damping == np.float32(0.3) and f.dt == np.float32(1.0) / np.float32(10000.0)


def test_null_context_calls_fail_cleanly(gpu_lib):
    """Every entry point rejects a NULL context without touching the GPU."""
    L = gpu_lib
    assert L.sph_rho_pres(None) == -1
    assert L.sph_force(None) == -1
    assert L.sph_integrate(None) == -1
    assert L.sph_step(None, 1) == -1
    assert L.sph_size(None) == 0
    assert L.sph_last_error(None) == b"null context"
    vp = C.c_void_p()
    assert L.sph_create_batch(C.byref(vp), 0, None, None, None, None, None) == -1
    assert b"scene" in L.sph_create_error()


def test_snapshot_format_round_trip(pkg, tmp_path):
    """The snapshot reader and writer agree with each other and with the header
    layout sph.h documents (no device work)."""
    sph = pkg.sph
    rng = np.random.default_rng(1)
    p = np.zeros(37, dtype=sph.PARTICLE_DTYPE)
    p["pos"] = rng.normal(size=(37, 4)).astype(np.float32)
    p["vel"] = rng.normal(size=(37, 4)).astype(np.float32)
    path = str(tmp_path / "s.snap")
    sph.write_snapshot(path, p, counts=[30, 7], frame=123)
    raw = open(path, "rb").read()
    assert raw[:8] == b"NPRSPHv1" and len(raw) == 64 + 2 * 8 + 37 * 64
    assert int.from_bytes(raw[16:24], "little") == 37 and int.from_bytes(raw[24:28], "little") == 2
    assert int.from_bytes(raw[32:40], "little") == 123
    assert raw[64 + 16:64 + 16 + 16] == p["pos"][0].tobytes()  # pos at offset 0 of the first record
    q, counts, frame = sph.read_snapshot(path)
    assert q.tobytes() == p.tobytes() and list(counts) == [30, 7] and frame == 123
    with pytest.raises(ValueError):
        sph.write_snapshot(path, p, counts=[30, 8])
