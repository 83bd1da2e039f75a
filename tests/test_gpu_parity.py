"""GPU parity of the three passes against the FP64 oracle, through the C ABI.

Each pass is compared from the identical float32 input state:
  rho_pres  vs rho_pres_comp.glsl:35-58   |drho| <= 2e-5 rho
  force     vs force_comp.glsl:35-66      |dF|   <= 2e-5 * sum_j |term_ij| (oracle `scale`)
  integrate vs integrate_comp.glsl:35-81  |dv|, |dx| <= 4 ulp-scale of their operands; axes
            whose FP64 result lies within 1e-6 of a wall may take either clamp branch.
The neighbour grid (cell starts and slot permutation) must equal the CPU
build bit for bit, and a graph-captured step must equal the three passes bit
for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO_RTOL = 2e-5
FORCE_TOL = 2e-5
F32_EPS = np.float64(np.finfo(np.float32).eps)


def _lattice(orc, sc, count=None):
    return orc.lattice(sc.nx, sc.ny, sc.nz, sc.spacing, sc.origin, sc.jitter, sc.seed,
                       count=sc.count if count is None else count)


def check_density(orc, before, after, ob, grid=False, c=None, f=None):
    rho, pr = orc.density(orc.pos64(before), c=c, f=f, b=ob, grid=grid)
    g_rho = after["extras"][:, 0].astype(np.float64)
    g_p = after["extras"][:, 1].astype(np.float64)
    err = np.abs(g_rho - rho)
    assert np.all(err <= RHO_RTOL * rho + 1e-30), f"rho rel err {np.max(err / np.maximum(rho, 1e-30))}"
    k = 2000.0 if f is None else float(f.gas_const)
    perr = np.abs(g_p - pr)
    assert np.all(perr <= k * RHO_RTOL * rho + 4 * F32_EPS * pr + 1e-30), "pressure mismatch"


def check_force(orc, state, after, ob, grid=False, c=None, f=None):
    frc, scale = orc.forces(orc.pos64(state), orc.vel64(state), state["extras"][:, 0].astype(np.float64),
                            state["extras"][:, 1].astype(np.float64), c=c, f=f, b=ob, grid=grid)
    g = after["force"][:, :3].astype(np.float64)
    # Coincident pairs give NaN through normalize(0) in the reference; the GPU
    # must produce non-finite forces for exactly those particles.
    bad = ~np.isfinite(frc).all(axis=1)
    assert np.array_equal(bad, ~np.isfinite(g).all(axis=1)), "non-finite force pattern differs"
    err = np.sqrt(((g[~bad] - frc[~bad]) ** 2).sum(axis=1))
    sc = scale[~bad]
    assert np.all(err <= FORCE_TOL * sc + 1e-30), f"force rel err {np.max(err / np.maximum(sc, 1e-300))}"


def check_integrate(orc, state, after, ob, f=None):
    x, v, near = orc.integrate(orc.pos64(state), orc.vel64(state), state["force"][:, :3],
                               state["extras"][:, 0], b=ob, f=f, wall_eps=1e-6)
    gx = after["pos"][:, :3].astype(np.float64)
    gv = after["vel"][:, :3].astype(np.float64)
    dt = float(np.float32(1e-4)) if f is None else float(f.dt)
    acc = state["force"][:, :3].astype(np.float64) / state["extras"][:, [0]].astype(np.float64)
    vscale = np.abs(state["vel"][:, :3]) + dt * np.abs(acc) + np.abs(v)
    xscale = np.abs(state["pos"][:, :3]) + dt * np.abs(v) + np.abs(x)
    strict = np.ones_like(gx, dtype=bool)
    for a in range(3):
        strict[:, a] = (near & (1 << a)) == 0
    bad = ~np.isfinite(x) | ~np.isfinite(v)
    assert np.array_equal(bad, ~np.isfinite(gx) | ~np.isfinite(gv)), "non-finite state pattern differs"
    strict &= ~bad
    assert np.all(np.abs(gv - v)[strict] <= 4 * F32_EPS * vscale[strict] + 1e-30), "velocity mismatch"
    assert np.all(np.abs(gx - x)[strict] <= 4 * F32_EPS * xscale[strict] + 1e-30), "position mismatch"
    # Near a wall either branch is acceptable, but the result must be one of them.
    if (~strict).any():
        lo = np.array(ob.lower[:3], dtype=np.float64)
        hi = np.array(ob.upper[:3], dtype=np.float64)
        on_wall = np.isclose(gx, lo[None, :], atol=0, rtol=0) | np.isclose(gx, hi[None, :], atol=0, rtol=0)
        ok = np.abs(gx - x) <= 1e-5 + 4 * F32_EPS * xscale
        assert np.all((ok | on_wall | bad)[~strict])


def assert_same_grid(start, perm, cstart, cperm, void_cells):
    """Cell starts equal; slot order equal except inside NaN cells, whose
    members interact with nothing and may sit in any order."""
    assert np.array_equal(start, cstart)
    mask = np.ones(len(perm), dtype=bool)
    for c in void_cells:
        b, e = int(start[c]), int(start[c + 1])
        assert np.array_equal(np.sort(perm[b:e]), np.sort(cperm[b:e]))
        mask[b:e] = False
    assert np.array_equal(perm[mask], cperm[mask])


def test_passes_dam10k_vs_bruteforce(pkg, orc, gpu_lib):
    """dam10k: each pass equals the O(N^2) FP64 transcription (the reference's own algorithm)."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.upload(p0)
        sim.rho_pres()
        s1 = sim.download()
        check_density(orc, p0, s1, ob, grid=False)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=False)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)
        # Untouched lanes are preserved bit for bit.
        for name, lane in (("pos", 3), ("vel", 3), ("force", 3), ("extras", 2), ("extras", 3)):
            assert np.array_equal(s3[name][:, lane].view(np.uint32), p0[name][:, lane].view(np.uint32))


def test_passes_after_steps_dam10k(pkg, orc, gpu_lib):
    """Parity holds from an evolved (disordered, wall-touching) state too."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.upload(p0)
        sim.step(200)
        st = sim.download()
        sim.rho_pres()
        s1 = sim.download()
        check_density(orc, st, s1, ob, grid=False)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=False)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)


def test_grid_bit_exact(pkg, orc, gpu_lib):
    """Cell starts and the (cell, AoS index) slot order equal the CPU stable counting sort."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        for nsteps in (0, 150):
            sim.upload(p0)
            if nsteps:
                sim.step(nsteps)
            st = sim.download()
            sim.rho_pres()
            start, perm = sim.download_grid()
            gi = sim.grid_info(0)
            g = orc.grid_geometry(orc.smoothing_length(), ob, sc.count, gi.cells_per_h)
            assert tuple(gi.dims) == tuple(g.dims) and gi.ncells == g.ncells
            assert np.float32(gi.inv_cell) == np.float32(g.inv_cs)
            keys, cstart, cperm = orc.grid_build(st["pos"][:, :3], g)
            assert_same_grid(start, perm, cstart, cperm, [g.ncells])


def test_wide_box_grown_cells(pkg, orc, gpu_lib):
    """A box 12 units wide and a few thousand particles: the default cell edge
    would exceed the grid's cell budget (and the 2048-cell x axis, whose cells
    are half as wide), so the edge grows. The grid equals the CPU build bit for
    bit and each pass the O(N^2) FP64 transcription."""
    lower, upper = (-6.0, -0.35, -0.1), (6.0, 1.0, 0.5)
    rng = np.random.default_rng(3)
    n = 3000
    p = orc.particles(n)
    centres = rng.uniform(-5.9, 5.9, 60)
    p["pos"][:, 0] = (centres[rng.integers(0, 60, n)] + rng.normal(0.0, 0.01, n)).astype(np.float32)
    p["pos"][:, 1] = rng.uniform(-0.30, -0.25, n).astype(np.float32)
    p["pos"][:, 2] = rng.uniform(0.0, 0.05, n).astype(np.float32)
    p["vel"][:, :3] = rng.normal(0.0, 0.3, (n, 3)).astype(np.float32)
    b = pkg.sph.make_boundary(lower, upper)
    ob = orc.make_boundary(lower, upper)
    g = orc.grid_geometry(orc.smoothing_length(), ob, n, 2)
    assert g.cs > float(orc.smoothing_length()) * (1 + 2 ** -10) / 2 * 1.1  # the edge grew
    assert g.dims[0] <= 2048 and g.ncells <= 2 * max(2 ** 19, 16 * n)
    with pkg.Simulation(n, boundaries=b) as sim:
        sim.upload(p)
        sim.rho_pres()
        s1 = sim.download()
        gi = sim.grid_info(0)
        assert tuple(gi.dims) == tuple(g.dims) and gi.ncells == g.ncells
        start, perm = sim.download_grid()
        keys, cstart, cperm = orc.grid_build(p["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])
        check_density(orc, p, s1, ob, grid=False)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=False)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)


@pytest.mark.parametrize("case", range(8))
def test_random_boxes_vs_oracle(pkg, orc, gpu_lib, case):
    """Seeded random scenes: boxes from thin to wide, particles in clusters
    and some beyond the walls, every cells_per_h. The grid equals the CPU
    build bit for bit and each pass the O(N^2) FP64 transcription."""
    rng = np.random.default_rng(1000 + case)
    h = float(orc.smoothing_length())
    ext = rng.uniform([0.05, 0.05, 0.05], [3.0, 1.5, 3.0])
    lo = rng.uniform(-1.0, 1.0, 3)
    lower, upper = tuple(lo), tuple(lo + ext)
    n = int(rng.integers(500, 2500))
    p = orc.particles(n)
    nc = int(rng.integers(1, 12))
    centres = lo + rng.uniform(0.0, 1.0, (nc, 3)) * ext
    pos = centres[rng.integers(0, nc, n)] + rng.normal(0.0, 0.6 * h, (n, 3))
    out = rng.random(n) < 0.02  # a few beyond the walls
    pos[out] += rng.normal(0.0, 0.2, (int(out.sum()), 3))
    p["pos"][:, :3] = pos.astype(np.float32)
    p["vel"][:, :3] = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
    sub = 1 + case % 3
    b = pkg.sph.make_boundary(lower, upper)
    ob = orc.make_boundary(lower, upper)
    with pkg.Simulation(n, boundaries=b) as sim:
        sim.set_tuning(cells_per_h=sub, prune=case % 2, block=128 if case < 4 else 256)
        sim.upload(p)
        sim.rho_pres()
        s1 = sim.download()
        g = orc.grid_geometry(orc.smoothing_length(), ob, n, sub)
        gi = sim.grid_info(0)
        assert tuple(gi.dims) == tuple(g.dims) and gi.ncells == g.ncells, (case, tuple(gi.dims), tuple(g.dims))
        start, perm = sim.download_grid()
        keys, cstart, cperm = orc.grid_build(p["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])
        check_density(orc, p, s1, ob, grid=False)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=False)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)


TUNINGS = [(sub, prune, block, bpx) for sub in (1, 2, 3) for prune in (0, 1) for block in (128, 256)
           for bpx in (0, 3)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: "%d-%d-%d-%d" % t)
def test_fused_step_equals_passes(pkg, orc, gpu_lib, tuning):
    """sph_step (the fused force+integrate kernel k_force<true,...> in captured
    graphs) equals 37 x (rho_pres, force, integrate) bit for bit at every tuning."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as a, pkg.Simulation(sc.count, boundaries=b) as c:
        a.set_tuning(*tuning)
        c.set_tuning(*tuning)
        a.upload(p0)
        c.upload(p0)
        a.step(37)  # 2 x 16-step graph + 5 single-step graphs
        for _ in range(37):
            c.rho_pres()
            c.force()
            c.integrate()
        assert a.download().tobytes() == c.download().tobytes()


def test_fused_step_every_tuning_dam1m_settled(pkg, orc, gpu_lib):
    """The timed path on the timed state: dam1m after 400 steps (collapsed, about
    150 neighbours per particle).  From that state, 17 fused steps (one 16-step
    graph and one 1-step graph) at every tuning give the same bits as the
    cells_per_h setting's pruned 128-wide one-tile-per-workgroup form; 17 steps
    of the three-pass surface equal them at the autotuner's usual picks
    (1,0,128,0) and (2,1,128,0); and one fused step at each cells_per_h, and at
    (1,0,128,0), matches the FP64 grid transcription pass by pass."""
    sc = pkg.scenes.dam1m()
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(400)
        st = sim.download()
        ref = {}
        for t in [(s, 1, 128, 0) for s in (1, 2, 3)] + [x for x in TUNINGS if x[1:] != (1, 128, 0)]:
            sim.set_tuning(*t)
            sim.upload(st)
            sim.step(17)
            got = sim.download().tobytes()
            if t[0] not in ref:
                ref[t[0]] = got
            else:
                assert got == ref[t[0]], t
        for t in [(1, 0, 128, 0), (2, 1, 128, 0)]:
            sim.set_tuning(*t)
            sim.upload(st)
            for _ in range(17):
                sim.rho_pres()
                sim.force()
                sim.integrate()
            assert sim.download().tobytes() == ref[t[0]], t
        for t in [(1, 0, 128, 0), (2, 1, 128, 0), (3, 1, 256, 3)]:
            sim.set_tuning(*t)
            sim.upload(st)
            sim.step(1)
            after = sim.download()
            check_density(orc, st, after, ob, grid=True)
            s1 = st.copy()
            s1["extras"][:, :2] = after["extras"][:, :2]
            check_force(orc, s1, after, ob, grid=True)
            s2 = s1.copy()
            s2["force"][:, :3] = after["force"][:, :3]
            check_integrate(orc, s2, after, ob)


def test_device_lattice_matches_cpu(pkg, orc, gpu_lib):
    sc = pkg.scenes.dam10k()
    lat = pkg.sph.make_lattice(**sc.lattice_args())
    with pkg.Simulation(sc.count) as sim:
        sim.init_lattice(lat)
        got = sim.download()
    assert got.tobytes() == _lattice(orc, sc).tobytes()


def test_dam1m_one_step_vs_grid_oracle(pkg, orc, gpu_lib):
    """1M particles: per-pass parity against the FP64 grid transcription."""
    sc = pkg.scenes.dam1m()
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(300)  # collapsed, wall-loaded state with NaN (coincident-corner) particles
        st = sim.download()
        sim.rho_pres()
        s1 = sim.download()
        check_density(orc, st, s1, ob, grid=True)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=True)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)
        start, perm = sim.download_grid()
        g = orc.grid_geometry(orc.smoothing_length(), ob, sc.count, sim.grid_info(0).cells_per_h)
        _, cstart, cperm = orc.grid_build(s2["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])


def test_dam8m_one_step_vs_grid_oracle(pkg, orc, gpu_lib):
    """8M particles (the large mixed run's scene) during its collapse: per-pass
    parity against the FP64 grid transcription and the bit-exact grid.  Its
    5 GB working set exceeds the Infinity Cache, unlike the 1M state."""
    sc = pkg.scenes.dam8m()
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(150)
        st = sim.download()
        sim.rho_pres()
        s1 = sim.download()
        check_density(orc, st, s1, ob, grid=True)
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=True)
        sim.integrate()
        s3 = sim.download()
        check_integrate(orc, s2, s3, ob)
        start, perm = sim.download_grid()
        g = orc.grid_geometry(orc.smoothing_length(), ob, sc.count, sim.grid_info(0).cells_per_h)
        _, cstart, cperm = orc.grid_build(s2["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])


def test_dam64m_cells_beyond_2e28_vs_grid_oracle(pkg, orc, gpu_lib):
    """64M particles in dam1m's box scaled 4x per axis: 321M cells, so cell keys
    and the cell-start array run past 2^28 and the scan folds 150k block sums.
    Three steps from the lattice, then the density and force passes against the
    FP64 grid transcription and the grid bit for bit against the CPU build."""
    sc = pkg.scenes.dam_block(400, 4.0 * pkg.scenes.box_scale(1_000_000))
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    print(f"dam64m: {sc.count} particles", flush=True)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(3)
        st = sim.download()
        sim.rho_pres()
        s1 = sim.download()
        check_density(orc, st, s1, ob, grid=True)
        del st
        sim.force()
        s2 = sim.download()
        check_force(orc, s1, s2, ob, grid=True)
        del s1
        gi = sim.grid_info(0)
        g = orc.grid_geometry(orc.smoothing_length(), ob, sc.count, gi.cells_per_h)
        assert gi.ncells == g.ncells and g.ncells > 2 ** 28
        start, perm = sim.download_grid()
        _, cstart, cperm = orc.grid_build(s2["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])


def test_reference_make_grid_scene(pkg, orc, gpu_lib):
    """The reference's literal initial state (make_grid, Main.cpp:488-505: 10x100x10
    at spacing PARTICLE_RADIUS = h/4, ~251 neighbours) through all three passes,
    then 50 steps later, against the O(N^2) FP64 transcription."""
    p0 = orc.reference_grid()
    ob = orc.default_boundary()
    with pkg.Simulation(10000) as sim:
        sim.upload(p0)
        for k in range(2):
            st = sim.download()
            sim.rho_pres()
            s1 = sim.download()
            check_density(orc, st, s1, ob, grid=False)
            sim.force()
            s2 = sim.download()
            check_force(orc, s1, s2, ob, grid=False)
            sim.integrate()
            s3 = sim.download()
            check_integrate(orc, s2, s3, ob)
            if k == 0:
                sim.step(50)


def test_short_trajectory_statistics_vs_fp64(pkg, orc, gpu_lib):
    """dam10k for 40 steps in FP32 on the GPU and in FP64 on the CPU from the same
    state: the chaotic trajectories stay close in aggregate (mean density,
    kinetic energy, centre of mass)."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    ob = orc.make_boundary(sc.lower, sc.upper)
    st = orc.State64(p0)
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.upload(p0)
        for _ in range(4):
            sim.step(10)
            st.step(b=ob, grid=True, nsteps=10)
            g = sim.download()
            gv = g["vel"][:, :3].astype(np.float64)
            cv = st.v.reshape(-1, 3)
            ke_g, ke_c = (gv * gv).sum(), (cv * cv).sum()
            assert abs(ke_g - ke_c) <= 1e-3 * ke_c
            assert abs(g["extras"][:, 0].astype(np.float64).mean() - st.rho.mean()) <= 1e-4 * st.rho.mean()
            assert np.allclose(g["pos"][:, :3].astype(np.float64).mean(0), st.x.reshape(-1, 3).mean(0), atol=1e-5)
