"""GPU tests of the learned coarse step through the C ABI, against the FP64
reference in oracle/surrogate_ref.py.

* sph_mlp_train gradients (FP32 forward/backward kernels, deterministic
  reduction) equal the FP64 analytic gradient of the same batch to 1e-4
  relative norm; the loss to 1e-5.
* sph_mlp_predict (FP32 path) equals the FP64 forward to 1e-4.
* Adam steps on the GPU lower the loss of a fixed batch.
* One coarse step (network, and the SPH force held for 4 fine steps) equals
  the FP64 coarse integrator from the same state.
* The judge with a zero drift limit keeps the 4-plain-step result.
"""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ref():
    spec = importlib.util.spec_from_file_location("surrogate_ref", os.path.join(ROOT, "oracle", "surrogate_ref.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


NORM = dict(v_scale=0.1, rho_scale=1e-5, pi_scale=5e-4, acc_scale=2e4)


def _norm(pkg):
    return pkg.sph.MlpNorm(NORM["v_scale"], NORM["rho_scale"], NORM["pi_scale"], NORM["acc_scale"])


def _scene(pkg, orc, n=1500, steps=30):
    sc = pkg.scenes.dam10k()
    p0 = orc.lattice(sc.nx, sc.ny, sc.nz, sc.spacing, sc.origin, sc.jitter, sc.seed, count=sc.count)[:n].copy()
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    sim = pkg.Simulation(n, boundaries=b)
    sim.upload(p0)
    sim.step(steps)
    return sim, sc


def _cpu_pairs(state, perm, h, slots):
    """Neighbour pairs (r < h, j != i) of the given target slots, by brute force
    in float32 arithmetic on the same positions."""
    pos = state["pos"][perm, :3]
    rows_t, rows_j = [], []
    h2 = np.float32(h) * np.float32(h)
    margin = []
    for t, s in enumerate(slots):
        d = (pos[s][None, :] - pos).astype(np.float32)
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).astype(np.float32)
        near = np.nonzero((r2 < h2) & (np.arange(len(pos)) != s))[0]
        margin.append(np.min(np.abs(r2.astype(np.float64) / float(h2) - 1.0)))
        rows_t += [t] * len(near)
        rows_j += list(near)
    return np.array(rows_t), np.array(rows_j), float(np.min(margin))


def _features(ref, state, perm, slots, rows_t, rows_j, h):
    st = state[perm]
    x = st["pos"][:, :3].astype(np.float64)
    v = st["vel"][:, :3].astype(np.float64)
    rho = st["extras"][:, 0].astype(np.float64)
    p = st["extras"][:, 1].astype(np.float64)
    inv_rho = (np.float32(1.0) / st["extras"][:, 0]).astype(np.float64)
    i = np.asarray(slots)[rows_t]
    j = rows_j
    return ref.pair_features(x[i], x[j], v[i], v[j], rho[i], rho[j], p[i], p[j], inv_rho[j], float(h), NORM)


def _targets(state_now, state_later, perm, slots, sc):
    a = state_now[perm][slots]
    b = state_later[perm][slots]
    dt = np.float32(1.0) / np.float32(10000.0)
    g = np.array([0.0, np.float32(-9806.65), 0.0])
    acc = (b["vel"][:, :3].astype(np.float64) - a["vel"][:, :3]) / (4 * float(dt)) - g
    lo, hi = np.array(sc.lower, dtype=np.float32), np.array(sc.upper, dtype=np.float32)
    ok = np.ones(len(slots), dtype=bool)
    for st in (a, b):
        ps = st["pos"][:, :3]
        ok &= np.all(np.isfinite(ps), axis=1) & np.all((np.abs(ps - lo) > 1e-6) & (np.abs(ps - hi) > 1e-6), axis=1)
    tgt = np.where(ok[:, None], acc / NORM["acc_scale"], 0.0)
    return tgt, ok.astype(np.float64)


def test_training_gradient_matches_fp64(pkg, orc, ref, gpu_lib):
    sim, sc = _scene(pkg, orc)
    later = pkg.Simulation(sim.n, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper))
    later.copy_state_from(sim)
    later.step(4)
    params = ref.init_params(7)
    params[ref.NPARAMS - 195:] += np.random.default_rng(8).normal(0, 0.05, 195).astype(np.float32)
    sim.mlp_set(params, _norm(pkg))
    first, stride = 1, 2
    res, grad = sim.mlp_train(later, first=first, stride=stride, update=False, want_grad=True)
    now = sim.download()
    lat = later.download()
    _, perm = sim.download_grid()
    slots = np.arange(first, sim.n, stride)
    h = orc.smoothing_length()
    rows_t, rows_j, margin = _cpu_pairs(now, perm, h, slots)
    assert margin > 1e-6, "fixture has a pair on the r = h boundary"
    assert res.pairs == len(rows_t) and res.targets == len(slots)
    f = _features(ref, now, perm, slots, rows_t, rows_j, h)
    tgt, w = _targets(now, lat, perm, slots, sc)
    assert res.weighted == int(w.sum())
    L, g, _ = ref.loss_and_grad(params.astype(np.float64), f, rows_t, tgt, w)
    assert res.loss == pytest.approx(L, rel=1e-5)
    assert np.linalg.norm(grad - g) <= 1e-4 * np.linalg.norm(g)
    sim.close()
    later.close()


def test_predict_fp32_matches_fp64(pkg, orc, ref, gpu_lib):
    sim, sc = _scene(pkg, orc, n=1200, steps=20)
    params = ref.init_params(11)
    sim.mlp_set(params, _norm(pkg))
    pred = sim.mlp_predict(path=0)
    now = sim.download()
    _, perm = sim.download_grid()
    slots = np.arange(sim.n)
    h = orc.smoothing_length()
    rows_t, rows_j, margin = _cpu_pairs(now, perm, h, slots)
    assert margin > 1e-6
    f = _features(ref, now, perm, slots, rows_t, rows_j, h)
    m, _ = ref.forward(params.astype(np.float64), f)
    cpu = np.zeros((sim.n, 3))
    np.add.at(cpu, rows_t, m)
    cpu_aos = np.zeros_like(cpu)
    cpu_aos[perm] = cpu
    scale = np.abs(m).sum() / len(m) * 200
    assert np.max(np.abs(pred - cpu_aos)) <= 1e-4 * scale
    sim.close()


def test_adam_lowers_loss(pkg, orc, ref, gpu_lib):
    sim, sc = _scene(pkg, orc, n=1500, steps=40)
    later = pkg.Simulation(sim.n, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper))
    later.copy_state_from(sim)
    later.step(4)
    sim.mlp_set(ref.init_params(5), _norm(pkg))
    losses = [sim.mlp_train(later, lr=3e-3).loss for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0]
    p, nm = sim.mlp_get()
    assert np.all(np.isfinite(p)) and nm.acc_scale == NORM["acc_scale"]
    sim.close()
    later.close()


@pytest.mark.parametrize("use_network", [False, True])
def test_coarse_step_matches_fp64_integrator(pkg, orc, ref, gpu_lib, use_network):
    sim, sc = _scene(pkg, orc, n=2000, steps=25)
    params = ref.init_params(13)
    sim.mlp_set(params, _norm(pkg))
    s0 = sim.download()
    if use_network:
        pred = sim.mlp_predict(path=1).astype(np.float64)  # the coarse step runs the MFMA kernel
        acc = NORM["acc_scale"] * pred + np.array([0.0, np.float32(-9806.65), 0.0])
    else:
        sim.rho_pres()
        sim.force()
        sf = sim.download()
        acc = sf["force"][:, :3].astype(np.float64) / sf["extras"][:, [0]].astype(np.float64)
    sim.upload(s0)
    st = sim.coarse_step(1, use_network=use_network)
    assert st.coarse_steps == 1 and st.plain_steps == 0
    got = sim.download()
    dt = float(np.float32(1.0) / np.float32(10000.0))
    x, v = ref.coarse_integrate(s0["pos"][:, :3].astype(np.float64), s0["vel"][:, :3].astype(np.float64), acc,
                                sc.lower, sc.upper, dt, float(np.float32(0.3)))
    vscale = np.abs(s0["vel"][:, :3]) + 4 * dt * np.abs(acc) + 1e-3
    # Axes whose unclamped FP64 position lies within 1e-5 of a wall may take
    # either clamp branch in float32.
    xu = s0["pos"][:, :3].astype(np.float64) + 4 * dt * s0["vel"][:, :3] + 10 * dt * dt * acc
    lo, hi = np.array(sc.lower)[None, :], np.array(sc.upper)[None, :]
    strict = (np.abs(xu - lo) > 1e-5) & (np.abs(xu - hi) > 1e-5)
    assert strict.mean() > 0.9
    assert np.all((np.abs(got["vel"][:, :3] - v) <= 2e-5 * vscale)[strict])
    assert np.all((np.abs(got["pos"][:, :3] - x) <= 1e-6)[strict])
    sim.close()


def test_judge_zero_limit_keeps_plain_result(pkg, orc, ref, gpu_lib):
    sim, sc = _scene(pkg, orc, n=2000, steps=10)
    sim.mlp_set(ref.init_params(17), _norm(pkg))
    s0 = sim.download()
    st = sim.coarse_step(3, use_network=True, judge_every=1, drift_limit=0.0, cooldown=0)
    assert st.judged == 3 and st.fallbacks == 3 and st.coarse_steps == 0 and st.plain_steps == 12
    got = sim.download()
    with pkg.Simulation(sim.n, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as plain:
        plain.upload(s0)
        plain.step(12)
        want = plain.download()
    assert got["pos"].tobytes() == want["pos"].tobytes() and got["vel"].tobytes() == want["vel"].tobytes()
    # cooldown: after a fallback the next coarse steps run plain.
    sim.upload(s0)
    st = sim.coarse_step(4, use_network=True, judge_every=4, drift_limit=0.0, cooldown=3)
    assert st.fallbacks == 1 and st.plain_steps == 16 and st.coarse_steps == 0
    sim.close()


def test_predict_mfma_matches_fp32(pkg, orc, ref, gpu_lib):
    """BF16 MFMA messages equal the FP32 path to BF16 accuracy, on a scene with
    more than 512 neighbour hits for some particles (hit-list wrap-around)."""
    sim, sc = _scene(pkg, orc, n=4000, steps=5)
    params = ref.init_params(19)
    params[ref.NPARAMS - 195:] += np.random.default_rng(20).normal(0, 0.05, 195).astype(np.float32)
    sim.mlp_set(params, _norm(pkg))
    p32 = sim.mlp_predict(path=0)
    pbf = sim.mlp_predict(path=1)
    assert np.all(np.isfinite(pbf))
    rel = np.linalg.norm(pbf - p32) / np.linalg.norm(p32)
    assert rel < 2e-2, rel
    # crowded: every particle within h of every other (hit lists longer than one LDS list)
    n = 700
    with pkg.Simulation(n) as dense:
        p = orc.particles(n)
        p["pos"][:, :3] = np.random.default_rng(1).uniform(0.0, 0.008, (n, 3)).astype(np.float32)
        dense.upload(p)
        dense.mlp_set(params, _norm(pkg))
        a = dense.mlp_predict(path=0)
        b = dense.mlp_predict(path=1)
        assert np.linalg.norm(b - a) / np.linalg.norm(a) < 2e-2


@pytest.mark.parametrize("sub", [1, 3])
def test_predict_mfma_every_stencil_size(pkg, orc, ref, gpu_lib, sub):
    """The message kernel at 1 and 3 cells per h (9 and 49 stencil rows, one
    lane each) matches the FP32 path to BF16 accuracy and the 2-cells-per-h
    messages up to the summation order, also where hit lists overflow."""
    sim, sc = _scene(pkg, orc, n=4000, steps=5)
    params = ref.init_params(19)
    params[ref.NPARAMS - 195:] += np.random.default_rng(20).normal(0, 0.05, 195).astype(np.float32)
    sim.mlp_set(params, _norm(pkg))
    p2 = sim.mlp_predict(path=1)
    sim.set_tuning(cells_per_h=sub)
    p32 = sim.mlp_predict(path=0)
    pbf = sim.mlp_predict(path=1)
    assert np.all(np.isfinite(pbf))
    assert np.linalg.norm(pbf - p32) / np.linalg.norm(p32) < 2e-2
    assert np.linalg.norm(pbf - p2) / np.linalg.norm(p2) < 1e-4
    sim.close()
    n = 700
    with pkg.Simulation(n) as dense:
        p = orc.particles(n)
        p["pos"][:, :3] = np.random.default_rng(1).uniform(0.0, 0.008, (n, 3)).astype(np.float32)
        dense.upload(p)
        dense.set_tuning(cells_per_h=sub)
        dense.mlp_set(params, _norm(pkg))
        a = dense.mlp_predict(path=0)
        b = dense.mlp_predict(path=1)
        assert np.linalg.norm(b - a) / np.linalg.norm(a) < 2e-2


def test_trained_network_beats_no_network(pkg, orc, gpu_lib):
    """The committed weights (trained on the GPU on 200 dam1m rollouts) predict
    the 4-step mean acceleration of held-out states better than the method
    without the network (the SPH force held over the 4 steps) and than zero."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "npr-sph_amd"))
    import train_surrogate
    params, norm, meta = train_surrogate.load_weights()
    assert meta["rollouts"] >= 200 and meta["workload"] == "dam1m"
    sph = pkg.sph
    sc = pkg.scenes.dam1m(seed=4242)  # seed not used in training
    b = sph.make_boundary(sc.lower, sc.upper)
    G = np.array([0.0, np.float32(-9806.65), 0.0])
    dt = float(np.float32(1.0) / np.float32(10000.0))
    with pkg.Simulation(sc.count, boundaries=b) as sim, pkg.Simulation(sc.count, boundaries=b) as later:
        sim.init_lattice(sph.make_lattice(**sc.lattice_args()))
        sim.step(600)
        later.copy_state_from(sim)
        later.step(4)
        sim.mlp_set(params, norm)
        pred = sim.mlp_predict(path=1).astype(np.float64)
        a, l = sim.download(), later.download()
        sim.rho_pres()
        sim.force()
        f = sim.download()
    with np.errstate(invalid="ignore", divide="ignore"):
        sph_acc = f["force"][:, :3].astype(np.float64) / f["extras"][:, [0]] - G
    tgt = (l["vel"][:, :3].astype(np.float64) - a["vel"][:, :3]) / (4 * dt) - G
    lo, hi = np.array(sc.lower, np.float32), np.array(sc.upper, np.float32)
    ok = np.all(np.isfinite(tgt), 1) & np.all(np.isfinite(sph_acc), 1)
    for st in (a, l):
        ps = st["pos"][:, :3]
        ok &= np.all((np.abs(ps - lo) > 1e-6) & (np.abs(ps - hi) > 1e-6), 1)
    A = norm.acc_scale

    def loss(p):
        return float(np.mean(np.sum((p[ok] - tgt[ok] / A) ** 2, 1)))

    l_net, l_sph, l_zero = loss(pred), loss(sph_acc / A), loss(np.zeros_like(tgt))
    assert l_net < 0.5 * l_zero and l_net < 0.1 * l_sph, (l_net, l_sph, l_zero)


def test_coarse_step_counts_every_pair(pkg, orc, gpu_lib):
    """sph_coarse_stats_t.pairs equals the number of ordered neighbour pairs
    (r < h, j != i) of the state the message pass ran on, counted by brute force
    in float32 on the CPU."""
    params = np.random.default_rng(3).normal(0, 0.05, pkg.sph.MLP_PARAMS).astype(np.float32)
    sim, sc = _scene(pkg, orc)
    with sim:
        sim.mlp_set(params, _norm(pkg))
        st = sim.download()
        stats = sim.coarse_step(1, use_network=True)
    pos = st["pos"][:, :3].astype(np.float32)
    h2 = np.float32(orc.smoothing_length()) * np.float32(orc.smoothing_length())
    total, margin = 0, np.inf
    for i in range(len(pos)):
        d = (pos[i][None, :] - pos).astype(np.float32)
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).astype(np.float32)
        r2[i] = np.inf
        total += int(np.count_nonzero(r2 < h2))
        margin = min(margin, float(np.min(np.abs(r2.astype(np.float64) / float(h2) - 1.0))))
    assert margin > 1e-6  # no pair so close to r = h that float32 rounding could decide it
    assert stats.pairs == total and total > 0


def test_repeated_rejections_keep_the_plain_trajectory(pkg, orc, ref, gpu_lib):
    """With a zero drift limit every judge rejects: after each rejection the
    cooldown runs plain steps and the next step is judged again, so no coarse
    step is ever accepted and the state equals the plain run bit for bit."""
    sim, sc = _scene(pkg, orc, n=2000, steps=10)
    sim.mlp_set(ref.init_params(23), _norm(pkg))
    s0 = sim.download()
    st = sim.coarse_step(10, use_network=True, judge_every=50, drift_limit=0.0, cooldown=2)
    # judged at 0, 3, 6, 9 (each after a 2-step cooldown)
    assert (st.judged, st.fallbacks, st.coarse_steps, st.plain_steps) == (4, 4, 0, 40)
    got = sim.download()
    with pkg.Simulation(sim.n, boundaries=pkg.sph.make_boundary(sc.lower, sc.upper)) as plain:
        plain.upload(s0)
        plain.step(40)
        want = plain.download()
    assert got["pos"].tobytes() == want["pos"].tobytes() and got["vel"].tobytes() == want["vel"].tobytes()
    # A limit no drift reaches accepts every judged step.
    sim.upload(s0)
    st = sim.coarse_step(3, use_network=True, judge_every=1, drift_limit=1e30, cooldown=2)
    assert (st.judged, st.fallbacks, st.coarse_steps, st.plain_steps) == (3, 0, 3, 0)
    sim.close()


def test_dam8m_learned_leg(pkg, orc, gpu_lib):
    """The large mixed run's learned leg at its real size: dam8m (8,000,000
    particles) after 150 steps, with the committed weights.
    * The BF16 MFMA messages (the kernel fused into the cell-sorted gather)
      equal the FP32 message path to BF16 accuracy over all particles.
    * The pairs one message pass evaluates equal the FP64 grid count of
      neighbour pairs (r < h, j != i) of that state, up to pairs within 1e-6
      of r = h, whose side float32 rounding decides.
    * Three coarse steps judged every step by 4 plain steps: each judged step
      is accepted or replaced, the drift is a finite relative figure and the
      state stays finite except the reference's own coincident-pair NaNs."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "npr-sph_amd"))
    import train_surrogate
    params, norm, _ = train_surrogate.load_weights()
    sc = pkg.scenes.dam8m()
    b = pkg.sph.make_boundary(sc.lower, sc.upper)
    ob = orc.make_boundary(sc.lower, sc.upper)
    with pkg.Simulation(sc.count, boundaries=b) as sim:
        sim.init_lattice(pkg.sph.make_lattice(**sc.lattice_args()))
        sim.step(150)
        sim.mlp_set(params, norm)
        s0 = sim.download()
        p32 = sim.mlp_predict(path=0)
        pbf = sim.mlp_predict(path=1)
        ok = np.all(np.isfinite(p32), axis=1)
        assert ok.mean() > 0.999
        assert np.array_equal(ok, np.all(np.isfinite(pbf), axis=1))
        rel = np.linalg.norm(pbf[ok] - p32[ok]) / np.linalg.norm(p32[ok])
        assert rel < 2e-2, rel
        del p32, pbf
        sim.upload(s0)
        st = sim.coarse_step(1, use_network=True, judge_every=0, drift_limit=0.0, cooldown=0)
        want, amb = orc.count_pairs(orc.pos64(s0), b=ob, margin=1e-6)
        assert want > 100 * sc.count
        assert abs(st.pairs - want) <= amb, (st.pairs, want, amb)
        sim.upload(s0)
        st = sim.coarse_step(3, use_network=True, judge_every=1, drift_limit=0.01, cooldown=0)
        assert st.judged == 3 and st.coarse_steps + st.fallbacks == 3
        assert st.plain_steps == 4 * st.fallbacks
        assert np.isfinite(st.max_drift) and 0.0 <= st.max_drift < 1.0
        m = sim.metrics(0)
        assert np.isfinite(m.mean_rho) and m.mean_rho > 1000.0
        # Non-finite particles: the reference's coincident-pair NaNs (normalize(0)
        # where corner walls clamp particles onto one point), as in 12 plain steps.
        sim.upload(s0)
        sim.step(12)
        mp = sim.metrics(0)
        assert mp.nonfinite < 1e-3 * sc.count
        assert m.nonfinite <= 2 * mp.nonfinite + 100, (m.nonfinite, mp.nonfinite)
