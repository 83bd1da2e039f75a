"""FP64 CPU reference of the learned coarse step's pair-message MLP.

TEST INFRASTRUCTURE ONLY: the checker for the solver's MLP kernels (forward,
backward, loss, Adam, coarse integrator).  No reference counterpart: NPR-SPH
has no learned component; the network is defined by this build.

Network (4995 parameters, flat float32 layout shared with the solver):
  W1[9][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]
  z1 = f W1 + b1, h1 = gelu(z1), z2 = h1 W2 + b2, h2 = gelu(z2), m = h2 W3 + b3
Pair features f_ij (i = target, j = neighbour with r < h, j != i):
  f0..2 = (x_i - x_j) / h
  f3..5 = (v_i - v_j) * v_scale
  f6    = rho_i * rho_scale,  f7 = rho_j * rho_scale
  f8    = (p_i + p_j) * 0.5 / rho_j * pi_scale     (the force_comp.glsl:59 pressure term)
Acceleration a_i = acc_scale * sum_j m_ij + G.
Coarse step (4 fine steps of semi-implicit Euler with a held fixed):
  v' = v + 4 dt a,  x' = x + 4 dt v + 10 dt^2 a, then the wall clamp of
  integrate_comp.glsl:46-77.
gelu is a polynomial form of x*Phi(x): 0.5 x + q(min(x^2, 9)) + 0.5 max(|x|-3, 0),
q(s) = s (c0 + s (c1 + s (c2 + s c3))), max |error| 2.8e-3 against the erf form.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 9, 64, 3
SIZES = [("w1", (NIN, NH)), ("b1", (NH,)), ("w2", (NH, NH)), ("b2", (NH,)), ("w3", (NH, NOUT)), ("b3", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)
GELU_C = np.array([3.88549402e-01, -5.37665106e-02, 4.76835308e-03, -1.70698904e-04], dtype=np.float32).astype(np.float64)
GELU_CLAMP = 9.0
GELU_KNEE = 3.0


def unpack(flat):
    flat = np.asarray(flat, dtype=np.float64)
    assert flat.size == NPARAMS
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def gelu(x):
    s = np.minimum(x * x, GELU_CLAMP)
    c = GELU_C
    q = s * (c[0] + s * (c[1] + s * (c[2] + s * c[3])))
    return 0.5 * x + q + 0.5 * np.maximum(np.abs(x) - GELU_KNEE, 0.0)


def gelu_grad(x):
    s = x * x
    c = GELU_C
    dq = c[0] + s * (2 * c[1] + s * (3 * c[2] + s * 4 * c[3]))
    inner = (s < GELU_CLAMP).astype(np.float64)
    tail = (np.abs(x) > GELU_KNEE).astype(np.float64) * 0.5 * np.sign(x)
    return 0.5 + inner * dq * 2.0 * x + tail


def init_params(seed: int = 0):
    """Scaled-uniform (Glorot) initialisation; small output layer."""
    rng = np.random.default_rng(seed)
    d = {}
    d["w1"] = rng.uniform(-1, 1, (NIN, NH)) * np.sqrt(6.0 / (NIN + NH))
    d["b1"] = np.zeros(NH)
    d["w2"] = rng.uniform(-1, 1, (NH, NH)) * np.sqrt(6.0 / (2 * NH))
    d["b2"] = np.zeros(NH)
    d["w3"] = rng.uniform(-1, 1, (NH, NOUT)) * np.sqrt(6.0 / (NH + NOUT)) * 0.1
    d["b3"] = np.zeros(NOUT)
    return pack(d).astype(np.float32)


def pair_features(xi, xj, vi, vj, rhoi, rhoj, pi, pj, inv_rhoj, h, norm):
    """Features of pairs (arrays over pairs), from float32 state values."""
    f = np.empty((len(rhoi), NIN))
    f[:, 0:3] = (xi - xj) / h
    f[:, 3:6] = (vi - vj) * norm["v_scale"]
    f[:, 6] = rhoi * norm["rho_scale"]
    f[:, 7] = rhoj * norm["rho_scale"]
    f[:, 8] = (pi + pj) * (0.5 * inv_rhoj) * norm["pi_scale"]
    return f


def forward(params, f):
    P = unpack(params)
    z1 = f @ P["w1"] + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["w2"] + P["b2"]
    h2 = gelu(z2)
    m = h2 @ P["w3"] + P["b3"]
    return m, (f, z1, h1, z2, h2)


def loss_and_grad(params, f, seg, target, weight):
    """Loss L = sum_t w_t |sum_{pairs of t} m - target_t|^2 / sum_t w_t and its
    gradient.  seg[p] = target index of pair p."""
    P = unpack(params)
    m, (f, z1, h1, z2, h2) = forward(params, f)
    nt = len(target)
    pred = np.zeros((nt, NOUT))
    np.add.at(pred, seg, m)
    wsum = max(weight.sum(), 1e-300)
    diff = pred - target
    L = float((weight[:, None] * diff * diff).sum() / wsum)
    dpred = 2.0 * weight[:, None] * diff / wsum
    dm = dpred[seg]
    g = {}
    g["w3"] = h2.T @ dm
    g["b3"] = dm.sum(0)
    dz2 = (dm @ P["w3"].T) * gelu_grad(z2)
    g["w2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["w2"].T) * gelu_grad(z1)
    g["w1"] = f.T @ dz1
    g["b1"] = dz1.sum(0)
    return L, pack(g), pred


def adam(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, 2015), bias-corrected; FP64."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def coarse_integrate(x, v, a, lower, upper, dt, damping):
    """4 fine semi-implicit Euler steps with a held fixed, then the wall clamp
    of integrate_comp.glsl:46-77 (lower wall tested first)."""
    nv = v + 4.0 * dt * a
    nx = x + 4.0 * dt * v + 10.0 * dt * dt * a
    lo = np.asarray(lower)[None, :]
    hi = np.asarray(upper)[None, :]
    below = nx < lo
    above = (~below) & (nx > hi)
    nx = np.where(below, lo, np.where(above, hi, nx))
    nv = np.where(below | above, nv * -damping, nv)
    return nx, nv
