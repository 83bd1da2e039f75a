"""Pins the FP64 oracle (oracle/sph_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible transcription mistake fails it:
  - closed forms (tests/golden/kat.json, derived analytically): isolated
    particle, a pair inside h (pressure sign and magnitude, viscosity), free
    fall, lattice sums counted by exact integer enumeration;
  - the strict `r < h` test, the self term, the lower-before-upper wall order;
  - invariants: permutation equivariance, zero viscous force under uniform
    velocity, NaN exactly at coincident pairs;
  - brute force vs grid (the neighbour grid changes no result);
  - the float32 cell hash and stable counting sort against an independent
    numpy implementation, the lattice generator against pure Python.
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def _one(orc, pos, vel=None):
    p = orc.particles(len(pos))
    p["pos"][:, :3] = np.asarray(pos, dtype=np.float32)
    if vel is not None:
        p["vel"][:, :3] = np.asarray(vel, dtype=np.float32)
    return p


def test_constants_match_reference(orc):
    c, f = orc.default_constants(), orc.default_fixed()
    assert (c.mass, c.smoothing_coeff, c.visc, c.resting_rho) == (np.float32(0.02), 4.0, 3000.0, 1000.0)
    assert float(orc.smoothing_length()) == GOLD["constants"]["h"]
    assert f.dt == GOLD["constants"]["dt"] and f.damping == GOLD["constants"]["damping"]
    assert f.gravity[1] == GOLD["constants"]["g"]


@pytest.mark.parametrize("grid", [False, True])
def test_isolated_particle(orc, grid):
    p = _one(orc, [[0.1, 0.2, 0.1]])
    rho, pr = orc.density(orc.pos64(p), grid=grid)
    assert rho[0] == pytest.approx(GOLD["isolated"]["rho"], rel=1e-13)
    assert pr[0] == pytest.approx(GOLD["isolated"]["p"], rel=1e-12)
    f, _ = orc.forces(orc.pos64(p), orc.vel64(p), rho, pr, grid=grid)
    assert np.allclose(f[0], GOLD["isolated"]["force"], rtol=1e-13, atol=0)
    x, v, _ = orc.integrate(orc.pos64(p), orc.vel64(p), f, rho)
    assert np.allclose(v[0], GOLD["isolated"]["vel1"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("grid", [False, True])
def test_pair_inside_h(orc, grid):
    r = GOLD["pair"]["r"]
    p = _one(orc, [[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    rho, pr = orc.density(orc.pos64(p), grid=grid)
    assert np.allclose(rho, GOLD["pair"]["rho"], rtol=1e-12)
    assert np.allclose(pr, GOLD["pair"]["p"], rtol=1e-12)
    f, _ = orc.forces(orc.pos64(p), orc.vel64(p), rho, pr, grid=grid)
    assert np.allclose(f[0], GOLD["pair"]["force0"], rtol=1e-11)
    assert np.allclose(f[1], GOLD["pair"]["force1"], rtol=1e-11)
    # Particle 1 moving along +z: the viscous term drags 0 along +z and 1 along -z.
    u = GOLD["pair_visc"]["u"]
    v = np.array([[0, 0, 0], [0, 0, u]], dtype=np.float64)
    pv = _one(orc, [[0.0, 0.0, 0.0], [r, 0.0, 0.0]], vel=v)
    f2, _ = orc.forces(orc.pos64(pv), orc.vel64(pv), rho, pr, grid=grid)
    assert f2[0, 2] == pytest.approx(GOLD["pair_visc"]["fv0_z"], rel=1e-6)
    assert f2[1, 2] == pytest.approx(GOLD["pair_visc"]["fv1_z"], rel=1e-6)


def test_pair_at_exactly_h_does_not_interact(orc):
    h = float(orc.smoothing_length())
    p = _one(orc, [[0.0, 0.0, 0.0], [h, 0.0, 0.0]])
    assert float(p["pos"][1, 0]) == h  # exactly representable
    rho, pr = orc.density(orc.pos64(p), grid=False)
    assert np.all(rho == GOLD["isolated"]["rho"]) or np.allclose(rho, GOLD["isolated"]["rho"], rtol=1e-14)
    f, _ = orc.forces(orc.pos64(p), orc.vel64(p), rho, pr, grid=False)
    assert np.allclose(f[:, 0], 0.0, atol=0)


def test_free_fall_closed_form(orc):
    p = _one(orc, [[0.2, 0.5, 0.2]])
    st = orc.State64(p)
    n = GOLD["free_fall"]["n"]
    st.step(nsteps=n, grid=True)
    assert st.v[1] == pytest.approx(GOLD["free_fall"]["v"], rel=1e-12)
    assert st.x[1] - 0.5 == pytest.approx(GOLD["free_fall"]["dy"], rel=1e-9)


def test_wall_clamp_and_order(orc):
    b = orc.default_boundary()
    f = orc.default_fixed()
    x = np.array([[-0.0999, 0.0, 0.0]])
    v = np.array([[-10.0, 0.0, 0.0]])
    force = np.zeros((1, 3))
    rho = np.array([1000.0])
    nx, nv, _ = orc.integrate(x, v, force, rho, b=b, f=f)
    assert nx[0, 0] == np.float64(np.float32(-0.1))
    assert nv[0, 0] == pytest.approx(10.0 * float(np.float32(0.3)), rel=1e-15)
    # Degenerate box (lower > upper): a position below lower is clamped to lower
    # and the upper test is skipped (else-if), so lower wins.
    bd = orc.make_boundary((0.5, 0.5, 0.5), (0.0, 0.0, 0.0))
    nx, nv, _ = orc.integrate(np.array([[0.2, 0.2, 0.2]]), np.zeros((1, 3)), force, rho, b=bd, f=f)
    assert np.allclose(nx[0], 0.5)


def test_lattice_interior_density_matches_enumeration(orc):
    """A 9^3 jitter-free lattice at spacing 0.5h: the centre particle sees the
    infinite-lattice sum (27 neighbours incl. self)."""
    h = orc.smoothing_length()
    s = np.float32(0.5) * h
    p = orc.lattice(9, 9, 9, s, (0.0, 0.0, 0.0), 0.0, 1)
    rho, _ = orc.density(orc.pos64(p), grid=False)
    centre = (4 * 9 + 4) * 9 + 4
    assert rho[centre] == pytest.approx(GOLD["lattice"]["half_h"]["rho"], rel=1e-6)


def test_reference_scene_density(orc):
    """make_grid (Main.cpp:488-505) at spacing 0.25h: an interior particle of the
    10x100x10 column sees the 251-neighbour lattice sum."""
    p = orc.reference_grid()
    idx = np.arange(10000)
    i, j, k = idx // 1000, (idx // 10) % 100, idx % 10
    expect = np.stack([i, j, k], 1).astype(np.float32) * np.float32(0.005)
    assert np.array_equal(p["pos"][:, :3], expect)
    rho, _ = orc.density(orc.pos64(p), grid=True)
    centre = 5 * 1000 + 50 * 10 + 5
    assert rho[centre] == pytest.approx(GOLD["lattice"]["quarter_h"]["rho"], rel=1e-5)


def _random_cloud(orc, n, seed, spread=0.06):
    rng = np.random.default_rng(seed)
    p = orc.particles(n)
    p["pos"][:, :3] = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    p["vel"][:, :3] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    return p


@pytest.mark.parametrize("seed", [1, 2])
def test_grid_equals_bruteforce(orc, seed):
    p = _random_cloud(orc, 1500, seed)
    # some particles beyond the box: clamped into border cells, must still interact
    p["pos"][:20, 0] = np.float32(0.6)
    p["pos"][20:40, 1] = np.float32(-0.4)
    x, v = orc.pos64(p), orc.vel64(p)
    r1, p1 = orc.density(x, grid=False)
    r2, p2 = orc.density(x, grid=True)
    assert np.allclose(r1, r2, rtol=1e-13, atol=0)
    f1, s1 = orc.forces(x, v, r1, p1, grid=False)
    f2, s2 = orc.forces(x, v, r1, p1, grid=True)
    assert np.all(np.abs(f1 - f2).max(axis=1) <= 1e-12 * s1)


def test_count_pairs_equals_bruteforce(orc):
    """Ordered pairs (j != i, r < h) over the cell stencil equal an O(N^2)
    numpy count, including particles beyond the walls and a coincident pair;
    a jitter-free 0.5h lattice interior has 26 neighbours per particle."""
    p = _random_cloud(orc, 1500, 5)
    p["pos"][:20, 0] = np.float32(0.6)
    p["pos"][21, :3] = p["pos"][22, :3]
    x = orc.pos64(p).reshape(-1, 3)
    h = float(orc.smoothing_length())
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    want = int(np.count_nonzero(np.sqrt(d2) < h))
    got, amb = orc.count_pairs(x, margin=1e-6)
    assert got == want and want > 0
    assert amb == int(np.count_nonzero(np.abs(d2 / (h * h) - 1.0) < 1e-6))
    s = np.float32(0.5) * orc.smoothing_length()
    lat = orc.lattice(3, 3, 3, s, (0.0, 0.0, 0.0), 0.0, 1)
    got, _ = orc.count_pairs(orc.pos64(lat))
    # 27 sites: pairs within h on the lattice offsets {-1,0,1}^3 minus the origin
    # and the (+-2,0,0)-type offsets at exactly r = h (strict <) are absent.
    idx = np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)], dtype=float)
    dd = ((idx[:, None, :] - idx[None, :, :]) ** 2).sum(-1)
    assert got == int(np.count_nonzero((dd > 0) & (np.sqrt(dd) * 0.5 < 1.0)))


def test_permutation_equivariance(orc):
    p = _random_cloud(orc, 800, 3)
    perm = np.random.default_rng(4).permutation(800)
    x, v = orc.pos64(p), orc.vel64(p)
    r, pr = orc.density(x, grid=False)
    f, s = orc.forces(x, v, r, pr, grid=False)
    q = p[perm]
    rq, prq = orc.density(orc.pos64(q), grid=False)
    fq, _ = orc.forces(orc.pos64(q), orc.vel64(q), rq, prq, grid=False)
    assert np.allclose(rq, r[perm], rtol=1e-13)
    assert np.all(np.abs(fq - f[perm]).max(axis=1) <= 1e-12 * s[perm])


def test_uniform_velocity_has_no_viscous_force(orc):
    p = _random_cloud(orc, 600, 5)
    x = orc.pos64(p)
    r, pr = orc.density(x, grid=False)
    v0 = np.zeros(3 * 600)
    v1 = np.tile([1.5, -2.0, 0.25], 600)
    f0, s = orc.forces(x, v0, r, pr, grid=False)
    f1, _ = orc.forces(x, v1, r, pr, grid=False)
    assert np.all(np.abs(f0 - f1).max(axis=1) <= 1e-13 * s)


def test_lattice_interior_net_pressure_force_vanishes(orc):
    """Spacing 2^-7 makes every position and difference exact, so the lattice is
    exactly symmetric about the centre and only FP64 summation rounding remains."""
    p = orc.lattice(11, 11, 11, np.float32(2.0 ** -7), (0.0, 0.0, 0.0), 0.0, 1)
    x, v = orc.pos64(p), orc.vel64(p)
    r, pr = orc.density(x, grid=False)
    f, s = orc.forces(x, v, r, pr, grid=False)
    centre = (5 * 11 + 5) * 11 + 5
    fp = f[centre] - r[centre] * np.array([0, GOLD["constants"]["g"], 0])
    assert np.abs(fp).max() <= 1e-12 * s[centre]
    # an off-centre particle near the lattice edge is pushed outwards (repulsive pressure)
    edge = (5 * 11 + 5) * 11 + 0
    assert f[edge, 2] - r[edge] * 0 < 0


def test_coincident_pair_gives_nan_force(orc):
    p = _one(orc, [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.05, 0.05, 0.05]])
    x, v = orc.pos64(p), orc.vel64(p)
    r, pr = orc.density(x, grid=False)
    assert np.all(np.isfinite(r))
    f, _ = orc.forces(x, v, r, pr, grid=False)
    assert np.isnan(f[0]).all() and np.isnan(f[1]).all() and np.isfinite(f[2]).all()
    f2, _ = orc.forces(x, v, r, pr, grid=True)
    assert np.array_equal(np.isnan(f), np.isnan(f2))


def test_nan_particle_is_inert(orc):
    p = _one(orc, [[0.0, 0.0, 0.0], [np.nan, 0.0, 0.0], [0.005, 0.0, 0.0]])
    x, v = orc.pos64(p), orc.vel64(p)
    for grid in (False, True):
        r, pr = orc.density(x, grid=grid)
        assert r[1] == 0.0 and pr[1] == 0.0
        f, _ = orc.forces(x, v, r, pr, grid=grid)
        assert np.all(np.isfinite(f[[0, 2]]))


# --------------------------------------------------------------- grid ---

def _numpy_keys(pos, g):
    """Independent float32 cell hash (numpy scalar semantics, no FMA)."""
    lo = np.array(g.lo[:3], dtype=np.float32)
    # x cells are half as wide as y/z cells: 2 * inv_cs along x (exact)
    inv = np.array([np.float32(2) * np.float32(g.inv_cs), g.inv_cs, g.inv_cs], dtype=np.float32)
    dims = np.array(g.dims[:3])
    f = (pos.astype(np.float32) - lo) * inv
    nan = np.isnan(f).any(axis=1)
    c = np.where(f < 0, 0, np.where(f >= dims.astype(np.float32), dims - 1,
                                     np.trunc(np.nan_to_num(f, nan=0.0, posinf=0.0, neginf=0.0)))).astype(np.int64)
    band, zb = c[:, 2] // 8, c[:, 2] % 8  # bands of 8 z-planes
    key = ((band * dims[1] + c[:, 1]) * 8 + zb) * dims[0] + c[:, 0]
    return np.where(nan, g.ncells, key).astype(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 257, 5000])
def test_grid_build_matches_numpy(orc, n):
    rng = np.random.default_rng(n + 11)
    b = orc.default_boundary()
    pos = rng.uniform(-0.2, 1.1, (n, 3)).astype(np.float32)
    if n > 10:
        pos[:3] = np.nan
        pos[3:6] = np.inf
        pos[6:9] = -np.inf
        pos[9, 0] = np.float32(0.5)  # exactly on the upper wall
    g = orc.grid_geometry(orc.smoothing_length(), b, n, sub=1 + n % 3)
    keys, start, perm = orc.grid_build(pos, g)
    ref = _numpy_keys(pos, g)
    assert np.array_equal(keys, ref)
    order = np.lexsort((np.arange(n), ref))  # stable: by key, then index
    assert np.array_equal(perm, order.astype(np.uint32))
    counts = np.bincount(ref.astype(np.int64), minlength=g.ncells + 1)
    assert np.array_equal(start, np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32))


def test_grid_geometry_rule(orc):
    """Cell edge h(1+2^-10)/sub along y and z, half that along x; dims cover
    the box; the grid holds at most 2 max(2^19, 16n) cells."""
    h = orc.smoothing_length()
    g = orc.grid_geometry(h, orc.default_boundary(), 10000)
    assert g.cs == pytest.approx(float(h) * (1 + 2 ** -10) / 2, rel=1e-15)
    assert tuple(g.dims) == (120, 135, 60)  # box 0.6 x 1.35 x 0.6, x cells of cs/2
    g1 = orc.grid_geometry(h, orc.default_boundary(), 10000, sub=1)
    assert tuple(g1.dims) == (60, 68, 30) and g1.sub == 1
    g3 = orc.grid_geometry(h, orc.default_boundary(), 10 ** 6, sub=3)
    assert g3.cs == pytest.approx(float(h) * (1 + 2 ** -10) / 3, rel=1e-15) and g3.sub == 3
    with pytest.raises(ValueError):
        orc.grid_geometry(h, orc.default_boundary(), 10, sub=4)
    big = orc.make_boundary((-100, -100, -100), (100, 100, 100))
    g2 = orc.grid_geometry(h, big, 1000)
    assert max(g2.dims) <= 2048 and g2.ncells <= 2 ** 20
    # cells are counted in whole bands of 8 z-planes (the last band padded)
    for gg in (g, g1, g2, g3):
        assert gg.ncells == gg.dims[0] * gg.dims[1] * 8 * -(-gg.dims[2] // 8)
    # neighbour safety: cell edge * 2 >= h always
    assert 2 * g2.cs >= float(h)
    # a huge, sparse scene: the cell budget stops at 2^30 (cell indices fit 31 bits)
    g4 = orc.grid_geometry(h, orc.make_boundary((-40, -40, -40), (40, 40, 40)), 200_000_000)
    assert g4.ncells <= 2 ** 30 and max(g4.dims) <= 2048
    # every axis covers the box: x in cells of cs/2, y and z in cells of cs
    for gg, bb in ((g, orc.default_boundary()), (g2, big)):
        ext = [float(bb.upper[a]) - float(bb.lower[a]) for a in range(3)]
        assert gg.dims[0] == min(2048, max(1, math.ceil(ext[0] / (gg.cs / 2))))
        assert gg.dims[1] == min(2048, max(1, math.ceil(ext[1] / gg.cs)))


# --------------------------------------------------------- generators ---

def _py_splitmix(seed, k):
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_splitmix_known_answers(orc):
    # SplitMix64 from state 0: the first outputs of the published generator.
    assert _py_splitmix(0, 0) == 0xE220A8397B1DCDAF
    assert _py_splitmix(0, 1) == 0x6E789E6AA1B965F4
    for s, k in [(0, 0), (1, 2), (0xDEADBEEF, 12345), ((1 << 64) - 1, 7)]:
        assert orc.splitmix64(s, k) == _py_splitmix(s, k)


def test_lattice_matches_python(orc):
    nx, ny, nz, sp, jit, seed = 3, 4, 5, np.float32(0.01), np.float32(0.001), 99
    org = np.array([-0.1, -0.35, -0.1], dtype=np.float32)
    p = orc.lattice(nx, ny, nz, sp, org, jit, seed, count=55)
    assert len(p) == 55
    scale = np.float32(jit * np.float32(1.0 / 8388608.0))
    for k in (0, 1, 17, 54):
        idx = (k // (nz * ny), (k // nz) % ny, k % nz)
        for a in range(3):
            u = np.int32((_py_splitmix(seed, 3 * k + a) >> 40) - (1 << 23))
            jv = np.float32(np.float32(u) * scale)
            base = np.float32(org[a] + np.float32(np.float32(idx[a]) * sp))
            assert p["pos"][k, a] == np.float32(base + jv)
        assert p["pos"][k, 3] == 1.0
    assert np.all(np.abs(p["pos"][:, :3] - np.round((p["pos"][:, :3] - org) / sp) * sp - org) <= jit * 1.0001)
