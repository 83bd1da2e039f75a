"""Pins the FP64 reference of the learned coarse step (oracle/surrogate_ref.py).

The GPU training kernels are checked against it, so it is checked first:
GELU against the erf form, analytic gradients against central finite
differences, Adam against its closed-form first step, and the coarse
integrator against 4 fine semi-implicit Euler steps with a constant
acceleration.
"""
import importlib.util
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ref():
    spec = importlib.util.spec_from_file_location("surrogate_ref", os.path.join(ROOT, "oracle", "surrogate_ref.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_parameter_count(ref):
    assert ref.NPARAMS == 4995
    p = ref.init_params(3)
    assert np.allclose(ref.pack(ref.unpack(p)), p)


def test_gelu_matches_erf_form(ref):
    x = np.linspace(-12, 12, 200001)
    exact = 0.5 * x * (1 + np.vectorize(math.erf)(x / math.sqrt(2)))
    assert np.max(np.abs(ref.gelu(x) - exact)) < 3e-3
    assert ref.gelu(np.array([0.0]))[0] == 0.0


def test_gelu_grad_finite_difference(ref):
    x = np.concatenate([np.linspace(-5, 5, 1001), [-3.0001, 2.9999, 3.0001]])
    x = x[np.abs(np.abs(x) - 3.0) > 1e-3]  # the derivative jumps at |x| = 3 by construction
    e = 1e-6
    fd = (ref.gelu(x + e) - ref.gelu(x - e)) / (2 * e)
    assert np.max(np.abs(fd - ref.gelu_grad(x))) < 1e-6


def _batch(ref, seed=0, npairs=300, ntargets=40):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 1, (npairs, ref.NIN))
    seg = np.sort(rng.integers(0, ntargets, npairs))
    target = rng.normal(0, 1, (ntargets, 3))
    weight = (rng.uniform(size=ntargets) > 0.2).astype(np.float64)
    return f, seg, target, weight


def test_loss_gradient_finite_difference(ref):
    params = ref.init_params(1).astype(np.float64)
    params += np.random.default_rng(2).normal(0, 0.05, params.shape)  # make the output layer matter
    f, seg, target, weight = _batch(ref)
    L, g, _ = ref.loss_and_grad(params, f, seg, target, weight)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(ref.NPARAMS, 60, replace=False), [0, 575, 576, 639, 640, 4735, 4736, 4799, 4800,
                                                                      4991, 4992, 4994]])
    for i in idx:
        e = 1e-6
        pp, pm = params.copy(), params.copy()
        pp[i] += e
        pm[i] -= e
        fd = (ref.loss_and_grad(pp, f, seg, target, weight)[0] - ref.loss_and_grad(pm, f, seg, target, weight)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_adam_first_step(ref):
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 0.0])
    p1, m, v = ref.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # First bias-corrected step moves each parameter by lr * sign(g) (g != 0).
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-7)


def test_coarse_integrator_is_four_fine_steps(ref):
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 0.1, (50, 3))
    v = rng.normal(0, 1, (50, 3))
    a = rng.normal(0, 1e3, (50, 3))
    dt = 1e-4
    nx, nv = ref.coarse_integrate(x, v, a, (-10, -10, -10), (10, 10, 10), dt, 0.3)
    fx, fv = x.copy(), v.copy()
    for _ in range(4):
        fv = fv + dt * a
        fx = fx + dt * fv
    assert np.allclose(nx, fx, rtol=0, atol=1e-14) and np.allclose(nv, fv, rtol=0, atol=1e-12)
    # Wall: lower tested first, velocity reflected with damping.
    cx, cv = ref.coarse_integrate(np.array([[0.0, 0, 0]]), np.array([[-100.0, 0, 0]]), np.zeros((1, 3)),
                                  (-0.001, -1, -1), (1, 1, 1), dt, 0.3)
    assert cx[0, 0] == -0.001 and cv[0, 0] == pytest.approx(30.0)
