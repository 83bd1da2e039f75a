"""GPU tests of the solver's surface beyond the three passes: tuning variants,
batched scenes, autotuning, parameter updates, edge cases and error paths.
All calls go through the C ABI (libnprsph.so)."""
import json
import os

import numpy as np
import pytest

from test_gpu_parity import assert_same_grid, check_density, check_force, check_integrate

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def _lattice(orc, sc):
    return orc.lattice(sc.nx, sc.ny, sc.nz, sc.spacing, sc.origin, sc.jitter, sc.seed, count=sc.count)


def _passes(sim):
    sim.rho_pres()
    a = sim.download()
    sim.force()
    b = sim.download()
    sim.integrate()
    return a, b, sim.download()


@pytest.mark.parametrize("sub", [1, 2, 3])
def test_tuning_variants_match_oracle_and_each_other(pkg, orc, gpu_lib, sub):
    """Every tuning gives oracle parity; prune/block/persistence never change bits."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.upload(p0)
        sim.step(120)
        st = sim.download()
        results = []
        for prune in (0, 1):
            for block in (128, 256):
                for bpx in (0, 3):
                    sim.set_tuning(sub, prune, block, bpx)
                    sim.upload(st)
                    results.append(_passes(sim))
        ref = results[0]
        for r in results[1:]:
            for x, y in zip(ref, r):
                assert x.tobytes() == y.tobytes()
        s1, s2, s3 = ref
        check_density(orc, st, s1, ob, grid=False)
        check_force(orc, s1, s2, ob, grid=False)
        check_integrate(orc, s2, s3, ob)
        start, perm = sim.download_grid()
        g = orc.grid_geometry(orc.smoothing_length(), ob, sc.count, sub)
        assert sim.grid_info(0).cells_per_h == sub and tuple(sim.grid_info(0).dims) == tuple(g.dims)
        _, cstart, cperm = orc.grid_build(s2["pos"][:, :3], g)
        assert_same_grid(start, perm, cstart, cperm, [g.ncells])


def test_batch_equals_unbatched(pkg, orc, gpu_lib):
    """Scenes advanced together in one launch per pass equal their own runs bit for bit."""
    scenes = [s for s in pkg.scenes.batch256() if s.count < 20000][:6]
    assert {s.layout for s in scenes} == {"dam", "column"}
    sph = pkg.sph
    bnds = [sph.make_boundary(s.lower, s.upper) for s in scenes]
    with pkg.Simulation([s.count for s in scenes], boundaries=bnds) as batch:
        for i, s in enumerate(scenes):
            batch.init_lattice(sph.make_lattice(**s.lattice_args()), scene=i)
        batch.step(40)
        got = batch.download()
        metrics = [batch.metrics(i) for i in range(len(scenes))]
    off = 0
    for i, s in enumerate(scenes):
        with pkg.Simulation(s.count, boundaries=bnds[i]) as one:
            one.init_lattice(sph.make_lattice(**s.lattice_args()))
            one.step(40)
            alone = one.download()
            m = one.metrics(0)
        assert got[off:off + s.count].tobytes() == alone.tobytes(), f"scene {i} differs"
        assert metrics[i].mean_rho == m.mean_rho
        off += s.count


def test_batch256_full_draw_equals_unbatched(pkg, gpu_lib):
    """The whole 256-scene draw (about 9.5M particles) in one batch: the largest,
    the smallest and two more of its scenes equal their own runs bit for bit."""
    scenes = pkg.scenes.batch256()
    assert len(scenes) == 256
    sph = pkg.sph
    bnds = [sph.make_boundary(s.lower, s.upper) for s in scenes]
    counts = [s.count for s in scenes]
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    picks = sorted({int(np.argmax(counts)), int(np.argmin(counts)), 17, 200})
    with pkg.Simulation(counts, boundaries=bnds) as batch:
        for i, s in enumerate(scenes):
            batch.init_lattice(sph.make_lattice(**s.lattice_args()), scene=i)
        batch.step(40)
        got = batch.download()
        metrics = {i: batch.metrics(i) for i in picks}
    for i in picks:
        s = scenes[i]
        with pkg.Simulation(s.count, boundaries=bnds[i]) as one:
            one.init_lattice(sph.make_lattice(**s.lattice_args()))
            one.step(40)
            alone = one.download()
            m = one.metrics(0)
        assert got[first[i]:first[i] + s.count].tobytes() == alone.tobytes(), f"scene {i} differs"
        assert metrics[i].mean_rho == m.mean_rho


def test_batch_scene_parity_vs_oracle(pkg, orc, gpu_lib):
    scenes = [s for s in pkg.scenes.batch256() if s.count < 8000][:3]
    sph = pkg.sph
    with pkg.Simulation([s.count for s in scenes], boundaries=[sph.make_boundary(s.lower, s.upper) for s in scenes]) as sim:
        for i, s in enumerate(scenes):
            sim.init_lattice(sph.make_lattice(**s.lattice_args()), scene=i)
        sim.step(30)
        st = sim.download()
        sim.rho_pres()
        s1 = sim.download()
        sim.force()
        s2 = sim.download()
    off = 0
    for s in scenes:
        ob = orc.make_boundary(s.lower, s.upper)
        sl = slice(off, off + s.count)
        check_density(orc, st[sl], s1[sl], ob, grid=False)
        check_force(orc, s1[sl], s2[sl], ob, grid=False)
        off += s.count


def test_autotune_keeps_state_and_picks_supported(pkg, orc, gpu_lib):
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.upload(p0)
        sim.step(10)
        before = sim.download()
        best, ms = sim.autotune(trial_steps=3)
        assert sim.download().tobytes() == before.tobytes()
        assert best.cells_per_h in (1, 2, 3) and best.block in (128, 256) and ms > 0
        assert sim.tuning().as_dict() == best.as_dict()
        sim.step(5)  # still usable


def test_isolated_particle_closed_form(pkg, gpu_lib):
    with pkg.Simulation(1) as sim:
        p = np.zeros(1, dtype=pkg.sph.PARTICLE_DTYPE)
        p["pos"][0] = (0.1, 0.2, 0.1, 1.0)
        sim.upload(p)
        sim.rho_pres()
        sim.force()
        got = sim.download()
        assert got["extras"][0, 0] == pytest.approx(GOLD["isolated"]["rho"], rel=2e-6)
        assert got["extras"][0, 1] == pytest.approx(GOLD["isolated"]["p"], rel=4e-6)
        assert np.allclose(got["force"][0, :3], GOLD["isolated"]["force"], rtol=4e-6)
        sim.integrate()
        got = sim.download()
        assert got["vel"][0, 1] == pytest.approx(GOLD["isolated"]["vel1"][1], rel=1e-6)


def test_edge_cases_vs_oracle(pkg, orc, gpu_lib):
    """Ragged count, particles outside the box, a NaN particle, a coincident pair
    (NaN force, as normalize(0)), and a crowded cell."""
    rng = np.random.default_rng(7)
    n = 2053
    p = orc.particles(n)
    p["pos"][:, :3] = rng.uniform(-0.05, 0.05, (n, 3)).astype(np.float32)
    p["vel"][:, :3] = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    p["pos"][0, :3] = (0.7, 0.0, 0.0)          # beyond the upper x wall
    p["pos"][1, :3] = (0.69, 0.0, 0.0)         # its neighbour, also outside
    p["pos"][2, :3] = (np.nan, 0.0, 0.0)
    p["pos"][3, :3] = p["pos"][4, :3]          # coincident pair
    p["pos"][10:200, :3] = (rng.uniform(0, 1e-3, (190, 3)) + 0.02).astype(np.float32)  # crowded cell
    ob = orc.default_boundary()
    with pkg.Simulation(n) as sim:
        sim.upload(p)
        s1, s2, s3 = _passes(sim)
    check_density(orc, p, s1, ob, grid=False)
    check_force(orc, s1, s2, ob, grid=False)
    check_integrate(orc, s2, s3, ob)
    assert np.isnan(s2["force"][3, :3]).all() and np.isnan(s2["force"][4, :3]).all()
    assert s1["extras"][2, 0] == 0.0 and s1["extras"][0, 0] > 0 and s1["extras"][1, 0] > GOLD["isolated"]["rho"]


def test_force_on_uploaded_extras(pkg, orc, gpu_lib):
    """sph_force with no sph_rho_pres before it takes rho and p from the uploaded
    extras (force_comp.glsl reads them from the buffer); here they are not the
    state's own density, so a stale or missing copy would show."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    ob = orc.make_boundary(sc.lower, sc.upper)
    rng = np.random.default_rng(5)
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.upload(p0)
        sim.step(60)
        st = sim.download()
        st["extras"][:, 0] = rng.uniform(2e4, 6e4, sc.count).astype(np.float32)
        st["extras"][:, 1] = rng.uniform(0.0, 1e8, sc.count).astype(np.float32)
        sim.upload(st)
        sim.force()
        got = sim.download()
        # and after a density pass the packed copies follow the new density
        sim.upload(st)
        sim.rho_pres()
        s1 = sim.download()
        sim.force()
        s2 = sim.download()
    assert np.array_equal(got["extras"], st["extras"])
    check_force(orc, st, got, ob, grid=False)
    check_force(orc, s1, s2, ob, grid=False)


def test_regrid_right_after_enqueued_steps(pkg, orc, gpu_lib):
    """A parameter change that moves the grid replaces the captured step graphs;
    issued right after sph_step returns (graphs possibly still running), it
    gives the same bits as issuing it after a synchronisation."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    c7 = pkg.sph.Constants(0.02, 7.0, 3000.0, 1000.0)
    with pkg.Simulation(sc.count, boundaries=b) as eager, pkg.Simulation(sc.count, boundaries=b) as synced:
        for sim in (eager, synced):
            sim.upload(p0)
        eager.step(48)
        eager.set_constants(c7)
        eager.step(20)
        synced.step(48)
        synced.synchronize()
        synced.set_constants(c7)
        synced.step(20)
        assert eager.download().tobytes() == synced.download().tobytes()


def test_event_timed_steps_equal_graph_steps(pkg, orc, gpu_lib):
    """sph_set_timing runs the steps as separate launches bracketed by HIP events.
    They give the same bits as the captured graphs, including the AoS record,
    which only a call's last step writes: a three-pass frame issued between
    two multi-step calls reads it."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as timed, pkg.Simulation(sc.count, boundaries=b) as graphed:
        for sim in (timed, graphed):
            sim.upload(p0)
        timed.set_timing(True)
        for sim in (timed, graphed):
            sim.step(21)
            _passes(sim)
            sim.step(19)
        assert timed.download().tobytes() == graphed.download().tobytes()
        times = timed.kernel_times()
        for k in ("scan_down", "scatter", "rank_reorder", "density"):
            assert times[k][1] == 41, (k, times[k])  # 40 steps and the frame's rho_pres
        # The fused step also counts the scan's block sums; only the builds that
        # follow the separate count sum them from the cell counts.
        assert times["scan_reduce"][1] == 2, times["scan_reduce"]
        # The fused step counts the next grid's cells itself: only the builds
        # after an upload or a three-pass integrate run the separate count.
        assert times["hash_count"][1] == 2, times["hash_count"]
        assert "scan_top" not in times  # a grid this small folds the block sums into scan_down
        assert times["force_integrate"][1] == 40
        assert times["force"][1] == 1 and times["integrate"][1] == 1
        assert all(ms > 0.0 for ms, _ in times.values())
        timed.set_timing(False)
        timed.step(5)
        assert timed.kernel_times() == {}


def test_set_constants_regrids(pkg, orc, gpu_lib):
    """The reference's smoothing slider (Main.cpp:244) changes h; the grid follows."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.upload(p0)
        dims0 = tuple(sim.grid_info(0).dims)
        c = pkg.sph.Constants(0.02, 7.0, 3000.0, 1000.0)
        sim.set_constants(c)
        assert tuple(sim.grid_info(0).dims) != dims0
        s1, s2, _ = _passes(sim)
    oc = orc.Constants(0.02, 7.0, 3000.0, 1000.0)
    check_density(orc, p0, s1, ob, grid=False, c=oc)
    check_force(orc, s1, s2, ob, grid=False, c=oc)


def test_error_paths_keep_context_usable(pkg, gpu_lib):
    with pkg.Simulation(100) as sim:
        with pytest.raises(pkg.SphError) as e:
            sim.set_constants(pkg.sph.Constants(0.02, -1.0, 3000.0, 1000.0))
        assert e.value.code == -1
        with pytest.raises(pkg.SphError):
            sim.upload(np.zeros(5, dtype=pkg.sph.PARTICLE_DTYPE), first=99)
        with pytest.raises(pkg.SphError):
            sim.set_tuning(cells_per_h=4)
        with pytest.raises(pkg.SphError):
            sim.init_lattice(pkg.sph.make_lattice(2, 2, 2, 0.01, (0, 0, 0), 0.0, 1))  # count 8 != 100
        c, _, _ = sim.params(0)
        assert c.smoothing_coeff == 4.0
        sim.step(3)
        assert sim.download().shape == (100,)
    with pytest.raises(pkg.SphError):
        pkg.Simulation(0)


def test_metrics_match_numpy(pkg, orc, gpu_lib):
    sc = pkg.scenes.dam10k()
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(50)
        st = sim.download()
        m = sim.metrics(0)
    v = st["vel"][:, :3].astype(np.float64)
    assert m.mean_rho == pytest.approx(st["extras"][:, 0].astype(np.float64).mean(), rel=1e-12)
    assert m.kinetic == pytest.approx(0.5 * float(np.float32(0.02)) * (v * v).sum(), rel=1e-10)
    assert m.max_speed == pytest.approx(np.sqrt((v * v).sum(1)).max(), rel=1e-12)
    assert np.allclose(m.com, st["pos"][:, :3].astype(np.float64).mean(0), rtol=1e-10)
    assert m.nonfinite == 0
    # a NaN particle is counted and left out of the kinetic energy and centre of mass
    st2 = st.copy()
    st2["vel"][5, 0] = np.nan
    with pkg.Simulation(sc.count, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as sim:
        sim.upload(st2)
        m2 = sim.metrics(0)
    keep = np.ones(len(st), bool)
    keep[5] = False
    assert m2.nonfinite == 1 and np.isfinite(m2.kinetic)
    assert m2.kinetic == pytest.approx(0.5 * float(np.float32(0.02)) * (v[keep] ** 2).sum(), rel=1e-10)
    assert np.allclose(m2.com, st["pos"][keep, :3].astype(np.float64).mean(0), rtol=1e-10)


def test_snapshot_save_load_resume(pkg, gpu_lib, tmp_path):
    """sph_save writes the documented format (byte for byte what the numpy
    writer produces from the downloaded state); sph_load into a fresh context
    resumes the run bit for bit; mismatched or damaged files are refused."""
    sph = pkg.sph
    scenes = [s for s in pkg.scenes.batch256() if s.count < 8000][:2]
    bnds = [sph.make_boundary(s.lower, s.upper) for s in scenes]
    counts = [s.count for s in scenes]
    path, ref = str(tmp_path / "a.snap"), str(tmp_path / "b.snap")
    with pkg.Simulation(counts, boundaries=bnds) as a:
        for i, s in enumerate(scenes):
            a.init_lattice(sph.make_lattice(**s.lattice_args()), scene=i)
        a.step(25)
        a.save(path, frame=25)
        sph.write_snapshot(ref, a.download(), counts=counts, frame=25)
        assert open(path, "rb").read() == open(ref, "rb").read()
        a.step(20)
        want = a.download()
    with pkg.Simulation(counts, boundaries=bnds) as b:
        assert b.load(path) == 25
        b.step(20)
        assert b.download().tobytes() == want.tobytes()
        with pytest.raises(pkg.SphError):
            b.load(str(tmp_path / "missing.snap"))
        open(str(tmp_path / "short.snap"), "wb").write(open(path, "rb").read()[:-64])
        with pytest.raises(pkg.SphError):
            b.load(str(tmp_path / "short.snap"))
        b.step(1)  # still usable
    with pkg.Simulation(sum(counts)) as c:  # one scene of the same total size
        with pytest.raises(pkg.SphError):
            c.load(path)


def test_dense_cluster_long_rows(pkg, orc, gpu_lib):
    """1400 particles within 0.2h next to a sparse cloud: stencil rows far longer
    than one LDS stage round (128 candidates), waves that mix cluster and sparse
    lanes.  Every pass matches the oracle and the fused step equals the passes."""
    rng = np.random.default_rng(11)
    n = 3200
    p = orc.particles(n)
    p["pos"][:, :3] = rng.uniform(-0.08, 0.08, (n, 3)).astype(np.float32)
    p["pos"][:1400, :3] = (rng.uniform(0, 4e-3, (1400, 3)) + 0.01).astype(np.float32)
    p["vel"][:, :3] = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    ob = orc.default_boundary()
    with pkg.Simulation(n) as sim, pkg.Simulation(n) as fused:
        sim.upload(p)
        s1, s2, s3 = _passes(sim)
        fused.upload(p)
        fused.step(1)
        assert fused.download().tobytes() == s3.tobytes()
    check_density(orc, p, s1, ob, grid=False)
    check_force(orc, s1, s2, ob, grid=False)
    check_integrate(orc, s2, s3, ob)


@pytest.mark.parametrize("sub", [1, 2, 3])
def test_staged_rows_every_tuning(pkg, orc, gpu_lib, sub):
    """A 128 x 8 x 16 slab along x in a widened box: almost every wave's
    particles share one cell row, so the gathers take the LDS-staged path
    (dam10k's short rows mostly take the global-memory one).  Every tuning
    matches the oracle and gives the same bits."""
    h = orc.smoothing_length()
    sp = 0.5 * h
    lower, upper = (-0.1, -0.35, -0.1), (1.4, 0.3, 0.3)
    origin = (lower[0] + 0.5 * sp, lower[1] + 0.5 * sp, lower[2] + 0.5 * sp)
    n = 128 * 8 * 16
    p0 = orc.lattice(128, 8, 16, sp, origin, 0.05 * h, 7, count=n)
    ob = orc.make_boundary(lower, upper)
    with pkg.Simulation(n, boundaries=pkg.sph.make_boundary(lower, upper)) as sim:
        sim.upload(p0)
        sim.step(40)
        st = sim.download()
        results = []
        for prune in (0, 1):
            for block in (128, 256):
                for bpx in (0, 2):
                    sim.set_tuning(sub, prune, block, bpx)
                    sim.upload(st)
                    results.append(_passes(sim))
    for r in results[1:]:
        for x, y in zip(results[0], r):
            assert x.tobytes() == y.tobytes()
    s1, s2, s3 = results[0]
    check_density(orc, st, s1, ob, grid=False)
    check_force(orc, s1, s2, ob, grid=False)
    check_integrate(orc, s2, s3, ob)


def test_counts_left_by_the_fused_step(pkg, orc, gpu_lib):
    """The fused step counts its new state into the next grid's cells.  Those
    counts serve the next build whichever call makes it (another step, the
    three-pass surface, a coarse step), and are discarded when the state or
    the grid is replaced (upload, regrid).  Every path equals a fresh context
    fed the same AoS state."""
    sc = pkg.scenes.dam10k()
    p0 = _lattice(orc, sc)
    b = pkg.sph.make_boundary(sc.lower, sc.upper)

    def fresh_passes(state):
        with pkg.Simulation(sc.count, boundaries=b) as f:
            f.upload(state)
            return _passes(f)

    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.upload(p0)
        sim.step(5)
        st = sim.download()
        got = _passes(sim)  # the build reuses the fused step's counts
        for x, y in zip(got, fresh_passes(st)):
            assert x.tobytes() == y.tobytes()
        sim.step(3)
        sim.upload(p0)  # counted state dropped: stale counts must not leak into the next build
        sim.step(4)
        with pkg.Simulation(sc.count, boundaries=b) as ref:
            ref.upload(p0)
            ref.step(4)
            assert sim.download().tobytes() == ref.download().tobytes()
        sim.step(2)
        sim.set_tuning(3, 1, 128, 0)  # regrid after a counting step
        sim.step(2)
        st = sim.download()
        sim.set_tuning(2, 1, 128, 0)
        sim.upload(st)
        sim.step(3)
        with pkg.Simulation(sc.count, boundaries=b) as ref:
            ref.upload(st)
            ref.step(3)
            assert sim.download().tobytes() == ref.download().tobytes()
