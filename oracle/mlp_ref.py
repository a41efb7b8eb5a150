"""CPU checker for the contact-seed network (test infrastructure only: the
tests, smoke() and bench.py's baseline leg may use it; the library never
does).  No reference counterpart: GAZPART seeds its contact solve from the
centre line (ellipsoid_contact_detection.f90:53-62).

Network 12 -> 64 -> 64 -> 4 (ReLU): outputs 0..2 predict the converged
common normal of the contact solve in the body frame of the first particle,
output 3 is the logit of the solve reporting a contact.  Loss per pair:
flag * |y[0:3] - t|^2 + softplus(y3) - flag * y3 (the normal is learned on
contacts only; the solve's converged normal elsewhere depends on its path,
not on the pose).  The parameters travel as one float32 vector:
    W1 (64 x 12), b1 (64), W2 (64 x 64), b2 (64), W3 (4 x 64), b3 (4).

Two modes:
  exact  - float64 throughout (pinned by finite differences in the tests);
  bf16   - what the GPU kernels compute: every matrix-product operand
           (inputs, weights, activations, output and hidden gradients)
           rounded to bfloat16 (round to nearest even), products exact,
           sums in float64 here (float32 accumulate on the GPU), biases and
           bias gradients in float32 from the unrounded values.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 12, 64, 4
SHAPES = (("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NOUT, NH)), ("b3", (NOUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)   # 5252


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1)
                           for name, _ in SHAPES])


def init_params(seed=0):
    """He-uniform weights, zero biases (float32 values)."""
    g = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shp[1])
            p[name] = g.uniform(-lim, lim, shp).astype(np.float32)
        else:
            p[name] = np.zeros(shp, dtype=np.float32)
    return pack(p).astype(np.float32)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# features
def rotm(q):
    """Rotation matrix of unit quaternions q (..., 4), scalar first; the
    convention of the reference's quaternion routines."""
    q0, q1, q2, q3 = (q[..., k] for k in range(4))
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - 2 * (q2 * q2 + q3 * q3)
    R[..., 0, 1] = 2 * (q1 * q2 - q0 * q3)
    R[..., 0, 2] = 2 * (q1 * q3 + q0 * q2)
    R[..., 1, 0] = 2 * (q1 * q2 + q0 * q3)
    R[..., 1, 1] = 1 - 2 * (q1 * q1 + q3 * q3)
    R[..., 1, 2] = 2 * (q2 * q3 - q0 * q1)
    R[..., 2, 0] = 2 * (q1 * q3 - q0 * q2)
    R[..., 2, 1] = 2 * (q2 * q3 + q0 * q1)
    R[..., 2, 2] = 1 - 2 * (q1 * q1 + q2 * q2)
    return R


def qmul(a, b):
    w1, x1, y1, z1 = (a[..., k] for k in range(4))
    w2, x2, y2, z2 = (b[..., k] for k in range(4))
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=-1)


def rotvec(q):
    """Rotation vector (axis * angle, angle in [0, pi]) of unit quaternions."""
    q = np.where(q[..., :1] < 0, -q, q)
    v = q[..., 1:]
    s = np.linalg.norm(v, axis=-1)
    ang = 2.0 * np.arctan2(s, q[..., 0])
    fac = np.where(s > 1e-12, ang / np.where(s > 1e-12, s, 1.0), 2.0)
    return v * fac[..., None]


def features(b1, q1, a1, b2, q2, a2):
    """(n, 12) inputs: R1^T (b2 - b1) / s, a1 / s, a2 / s, rotvec(q1* q2),
    s = max(a1) + max(a2); b the (minimum-image) centres, q unit
    quaternions, a the three semi-axes."""
    s = a1.max(axis=-1) + a2.max(axis=-1)
    R1 = rotm(q1)
    d = np.einsum("nji,nj->ni", R1, b2 - b1) / s[:, None]
    qc = q1 * np.array([1.0, -1.0, -1.0, -1.0])
    rv = rotvec(qmul(qc, q2))
    return np.concatenate([d, a1 / s[:, None], a2 / s[:, None], rv], axis=1)


def target(q1, cv):
    """Body-frame (particle 1) unit normal of a lab-frame normal cv."""
    return np.einsum("nji,nj->ni", rotm(q1), cv)


def seed_ok(y, p_min=0.9):
    """The seed is used when the contact probability exceeds p_min."""
    return 1.0 / (1.0 + np.exp(-y[:, 3])) > p_min


def seed_angles(q1, y):
    """Solve angles (azimuth, elevation) of the normal R1 y / |y|."""
    y = y[:, :3]
    n = y / np.linalg.norm(y, axis=-1, keepdims=True)
    cv = np.einsum("nij,nj->ni", rotm(q1), n)
    return np.stack([np.arctan2(cv[:, 1], cv[:, 0]), np.arcsin(np.clip(cv[:, 2], -1, 1))], axis=1)


# ---------------------------------------------------------------------------
# network
def forward(theta, X, mode="bf16"):
    """Returns y (n, 4) and the cache for backward."""
    p = unpack(theta)
    rnd = bf16 if mode == "bf16" else (lambda v: np.asarray(v, dtype=np.float64))
    fl = f32 if mode == "bf16" else (lambda v: np.asarray(v, dtype=np.float64))
    Xb = rnd(X)
    W1, W2, W3 = rnd(p["W1"]), rnd(p["W2"]), rnd(p["W3"])
    z1 = fl(fl(Xb @ W1.T) + fl(p["b1"]))
    h1 = np.maximum(z1, 0.0)
    H1 = rnd(h1)
    z2 = fl(fl(H1 @ W2.T) + fl(p["b2"]))
    h2 = np.maximum(z2, 0.0)
    H2 = rnd(h2)
    y = fl(fl(H2 @ W3.T) + fl(p["b3"]))
    return y, dict(Xb=Xb, W1=W1, W2=W2, W3=W3, z1=z1, H1=H1, z2=z2, H2=H2)


def softplus(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def pair_loss(y, T):
    """T (n, 4): unit normal (3) and contact flag (0 / 1)."""
    f = T[:, 3]
    return f * ((y[:, :3] - T[:, :3]) ** 2).sum(axis=1) + softplus(y[:, 3]) - f * y[:, 3]


def loss(theta, X, T, mode="exact"):
    y, _ = forward(theta, X, mode)
    return float(pair_loss(y, T).sum() / X.shape[0])


def backward(theta, X, T, mode="bf16"):
    """Gradient of the mean pair loss (see the module docstring)."""
    y, c = forward(theta, X, mode)
    rnd = bf16 if mode == "bf16" else (lambda v: np.asarray(v, dtype=np.float64))
    fl = f32 if mode == "bf16" else (lambda v: np.asarray(v, dtype=np.float64))
    n = X.shape[0]
    f = T[:, 3:4]
    dy = np.empty_like(y)
    dy[:, :3] = 2.0 * f * (y[:, :3] - T[:, :3]) / n
    dy[:, 3] = (1.0 / (1.0 + np.exp(-y[:, 3])) - T[:, 3]) / n
    dy = fl(dy)
    dYb = rnd(dy)
    g = {}
    g["W3"] = dYb.T @ c["H2"]
    g["b3"] = dy.sum(axis=0)
    dh2 = fl(dYb @ c["W3"]) * (c["z2"] > 0)
    dH2 = rnd(dh2)
    g["W2"] = dH2.T @ c["H1"]
    g["b2"] = dh2.sum(axis=0)
    dh1 = fl(dH2 @ c["W2"]) * (c["z1"] > 0)
    dH1 = rnd(dh1)
    g["W1"] = dH1.T @ c["Xb"]
    g["b1"] = dh1.sum(axis=0)
    return pack(g), y


def adam(theta, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on float32 masters (float32 arithmetic as on the GPU
    up to the order of operations)."""
    theta = theta.astype(np.float32)
    m = (b1 * m + (1 - b1) * grad).astype(np.float32)
    v = (b2 * v + (1 - b2) * grad * grad).astype(np.float32)
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    theta = (theta - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return theta, m, v
