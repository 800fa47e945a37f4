"""CPU oracle of the learned Newton seed: features, network, loss and gradients.

TEST INFRASTRUCTURE ONLY.  Only tests, the smoke check and the benchmark's CPU
leg use this module, and only as the checker: the product computes all of
this on the GPU.

The reference has no learned seed.  Its Newton seed is the centroid of the
closest stage-1 triangle (include/ScatObjects/SurfacePatch.hpp:272-303 and
triangle_uv_centroid, :499-510), and Newton runs from it (:593-660).  The
learned seed is a network that predicts the same Newton's converged
(u, v, t) for a (ray, patch) pair, from the ray in the patch's local frame and
15 points of the patch.  It predicts (u, v) as a correction to the FP64
tangent-plane estimate (u_a, v_a) and t relative to t_c (ray_features), so the
bf16 inputs' rounding only scales the correction.  Everything below is this build's own definition; the
HIP kernels implement exactly this definition.

Pinning:
  * `forward`/`backward` in exact mode (float64, no bf16 rounding) are checked
    against torch.autograd in float64 and against central finite differences;
  * `bf16` is checked against torch's float32 -> bfloat16 conversion (round to
    nearest even);
  * `gelu` is checked against torch's tanh-form GELU, and its derivative by
    finite differences;
  * `patch_frame`/`ray_features` are checked on a planar patch with a
    closed-form frame and on rigid motions (features are invariant under a
    rotation + translation of patch and ray together).

Network (parameter layout, float32, padded so every weight block is a whole
number of 16x16 tiles; pads are zero and stay zero):
    W1 [64][64] (inputs 51..63 zero)  b1 [64]
    W2 [64][64]                       b2 [64]
    W3 [64][64]                       b3 [64]
    W4 [16][64] (rows 3..15 zero)     b4 [16] (3..15 zero)
    z1 = W1 x + b1, h1 = gelu(z1); z2 = W2 h1 + b2, h2 = gelu(z2);
    z3 = W3 h2 + b3, h3 = gelu(z3); y = W4 h3 + b4 (y[0:3] used).
GPU rounding points (emulated by ``bf16=True``): x, every W, h1..h3 and the
back-propagated dY and dZ1..dZ3 are rounded to bfloat16 (round to nearest
even) as matrix-core operands; accumulation, biases, GELU and its derivative
are float32.

Loss over a batch of n samples: L = (1/n) sum_i sum_k (y_ik - target_ik)^2 for
k < 3; dY = (2/n)(y - target).
"""
from __future__ import annotations

import math

import numpy as np

N_IN = 51
N_RAY = 6
N_PATCH = 45
H = 64
N_OUT = 3
OUT_PAD = 16

OFF_W1 = 0
OFF_B1 = OFF_W1 + H * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + OUT_PAD * H
N_PARAMS = OFF_B4 + OUT_PAD          # 13520 (11843 of them live)

GELU_K0 = math.sqrt(2.0 / math.pi)
GELU_K1 = 0.044715

# the 15 feature points: u on the 5 Gauss-Lobatto points, v on {-1, 0, 1}
FEATURE_U = np.array([-1.0, -math.sqrt(3.0 / 7.0), 0.0, math.sqrt(3.0 / 7.0), 1.0])
FEATURE_V = np.array([-1.0, 0.0, 1.0])


def unpack(p: np.ndarray):
    """Views (W1, b1, W2, b2, W3, b3, W4, b4) of a flat parameter vector."""
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(H, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(H, H), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4].reshape(OUT_PAD, H), p[OFF_B4:])


def live_mask() -> np.ndarray:
    """1 where a parameter is live, 0 on the zero pads."""
    m = np.ones(N_PARAMS, np.float32)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(m)
    W1[:, N_IN:] = 0
    W4[N_OUT:, :] = 0
    b4[N_OUT:] = 0
    return m


def bf16(x) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest even), returned as float32 values."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def gelu(z):
    """tanh-form GELU and its derivative (float64 arithmetic)."""
    z = np.asarray(z, dtype=np.float64)
    a = GELU_K0 * (z + GELU_K1 * z ** 3)
    T = np.tanh(a)
    g = 0.5 * z * (1.0 + T)
    gp = 0.5 * (1.0 + T) + 0.5 * z * (1.0 - T * T) * GELU_K0 * (1.0 + 3.0 * GELU_K1 * z * z)
    return g, gp


def _f32(a, on):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64) if on else a


def forward(params, x, bf16_ops: bool = True):
    """Network outputs y [n][3] and the cache the backward pass needs.

    x is [n][51] (or [n][64] with zero pads).  With bf16_ops the GPU's operand
    roundings are applied (float64 accumulation stands in for float32).
    """
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    X = np.zeros((n, H))
    X[:, :x.shape[1]] = x
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, dtype=np.float64) for a in unpack(params)]
    rb = (lambda a: bf16(a).astype(np.float64)) if bf16_ops else (lambda a: a)
    Ws = [rb(W1), rb(W2), rb(W3), rb(W4)]
    bs = [b1, b2, b3, b4]
    acts = [rb(X)]
    gps = []
    zs = []
    for l in range(3):
        z = _f32(acts[-1] @ Ws[l].T + bs[l], bf16_ops)
        g, gp = gelu(z)
        zs.append(z)
        gps.append(_f32(gp, bf16_ops))
        acts.append(rb(_f32(g, bf16_ops)))
    y = _f32(acts[-1] @ Ws[3].T + bs[3], bf16_ops)
    cache = dict(acts=acts, gps=gps, zs=zs, Ws=Ws, bf16=bf16_ops)
    return y[:, :N_OUT], cache


def loss_grad(params, x, target, bf16_ops: bool = True):
    """(loss, flat gradient [N_PARAMS]) of the mean squared error over the batch."""
    y, c = forward(params, x, bf16_ops)
    t = np.asarray(target, dtype=np.float64)
    n = y.shape[0]
    r = y - t
    loss = float(np.sum(r * r) / n)
    rb = (lambda a: bf16(a).astype(np.float64)) if bf16_ops else (lambda a: a)
    dY = np.zeros((n, OUT_PAD))
    dY[:, :N_OUT] = _f32(2.0 * r / n, bf16_ops)
    g = np.zeros(N_PARAMS)
    gW = list(unpack(g))
    acts, gps, Ws = c["acts"], c["gps"], c["Ws"]
    dz = rb(dY)
    # layer 4
    gW[6][:] = dz.T @ acts[3]
    gW[7][:] = dz.sum(axis=0)
    for l in (2, 1, 0):
        dh = _f32(dz @ Ws[l + 1], bf16_ops)
        dz = rb(_f32(dh * gps[l], bf16_ops))
        gW[2 * l][:] = dz.T @ acts[l]
        gW[2 * l + 1][:] = dz.sum(axis=0)
    g *= live_mask()
    return loss, g


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights from SplitMix64(seed), zero biases (the library's bbrt_mlp_create)."""
    from billardballraytracing3d_amd.scenes import SplitMix64
    rng = SplitMix64(seed)
    p = np.zeros(N_PARAMS, np.float32)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    for W, fi, fo in ((W1[:, :N_IN], N_IN, H), (W2, H, H), (W3, H, H), (W4[:N_OUT], H, N_OUT)):
        a = math.sqrt(6.0 / (fi + fo))
        W[...] = rng.uniform(W.size, -a, a).reshape(W.shape).astype(np.float32)
    return p


# ---------------------------------------------------------------------------
# Features
# ---------------------------------------------------------------------------
def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def patch_frame(X, Xu, nu, P, Xv=None):
    """Local frame of a patch from X(0,0), X_u(0,0), nu(0,0) (the stored normal field)
    and the 15 feature points P[a*3+b] = X(FEATURE_U[a], FEATURE_V[b]); X_v(0,0) gives
    the tangent-plane coordinate rows (a_u, a_v) (zero when Xv is None).

    Returns (c, e1, e2, n, s, f[45], a_u, a_v); all float64 except f (float32 values).
    Order of every operation is the one the GPU kernel uses (no FMA contraction).
    """
    c = np.array(X, dtype=np.float64)
    ln = math.sqrt(_dot(nu, nu))
    n = np.array([nu[0] / ln, nu[1] / ln, nu[2] / ln])
    k = _dot(Xu, n)
    w = np.array([Xu[0] - k * n[0], Xu[1] - k * n[1], Xu[2] - k * n[2]])
    lw = math.sqrt(_dot(w, w))
    e1 = np.array([w[0] / lw, w[1] / lw, w[2] / lw])
    e2 = np.array([n[1] * e1[2] - n[2] * e1[1], n[2] * e1[0] - n[0] * e1[2], n[0] * e1[1] - n[1] * e1[0]])
    s = 0.0
    for q in range(15):
        dq = np.array(P[q], dtype=np.float64) - c
        s = max(s, math.sqrt(_dot(dq, dq)))
    f = np.zeros(N_PATCH, np.float32)
    for q in range(15):
        dq = np.array(P[q], dtype=np.float64) - c
        f[3 * q] = _dot(dq, e1) / s
        f[3 * q + 1] = _dot(dq, e2) / s
        f[3 * q + 2] = _dot(dq, n) / s
    au = np.zeros(3); av = np.zeros(3)
    if Xv is not None:
        # rows of the pseudo-inverse of [X_u X_v]: u = a_u . (P - c), v = a_v . (P - c) on the tangent plane
        guu = _dot(Xu, Xu); guv = _dot(Xu, Xv); gvv = _dot(Xv, Xv)
        det = guu * gvv - guv * guv
        i00 = gvv / det; i01 = -guv / det; i11 = guu / det
        au = np.array([i00 * Xu[0] + i01 * Xv[0], i00 * Xu[1] + i01 * Xv[1], i00 * Xu[2] + i01 * Xv[2]])
        av = np.array([i01 * Xu[0] + i11 * Xv[0], i01 * Xu[1] + i11 * Xv[1], i01 * Xu[2] + i11 * Xv[2]])
    return c, e1, e2, n, s, f, au, av


PLANE_CLAMP = 1.5


def ray_features(frame, o, d):
    """(r[6] float32, t_c, u_a, v_a): the ray in the patch frame, origin moved to the
    point of the ray closest to the frame centre (at parameter t_c); (u_a, v_a) are
    the tangent-plane coordinates of the ray's hit on the plane through the centre,
    clamped to [-1.5, 1.5] (t_c's point when the ray is parallel to the plane)."""
    c, e1, e2, n, s, _, au, av = frame
    o = np.asarray(o, dtype=np.float64); d = np.asarray(d, dtype=np.float64)
    cmo = c - o
    tc = _dot(cmo, d)
    q = np.array([o[0] + tc * d[0] - c[0], o[1] + tc * d[1] - c[1], o[2] + tc * d[2] - c[2]])
    r = np.array([_dot(q, e1) / s, _dot(q, e2) / s, _dot(q, n) / s, _dot(d, e1), _dot(d, e2), _dot(d, n)],
                 dtype=np.float32)
    dn = _dot(d, n)
    tp = _dot(cmo, n) / dn if abs(dn) > 1e-12 else tc
    pc = np.array([o[0] + tp * d[0] - c[0], o[1] + tp * d[1] - c[1], o[2] + tp * d[2] - c[2]])
    ua = min(max(_dot(au, pc), -PLANE_CLAMP), PLANE_CLAMP)
    va = min(max(_dot(av, pc), -PLANE_CLAMP), PLANE_CLAMP)
    return r, tc, ua, va


def seed_from_output(frame, feats, y):
    """Newton seed (u0, v0, t0) from network outputs y = (u - u_a, v - v_a, (t - t_c)/s)."""
    _, tc, ua, va = feats
    s = frame[4]
    return ua + float(y[0]), va + float(y[1]), tc + s * float(y[2])


def target_from_root(frame, feats, u, v, t):
    _, tc, ua, va = feats
    return np.array([u - ua, v - va, (t - tc) / frame[4]], dtype=np.float32)


def scene_patch_frame(sc, k: int):
    """patch_frame of patch k of a Scene, with X, X_u, nu evaluated by the oracle's
    lagrange_basis_1d + eval_scalar_field restatement (SurfacePatch.hpp:518-555)."""
    from oracle import oracle as O
    un = sc.unodes[sc.u_off[k]:sc.u_off[k + 1]]
    vn = sc.vnodes[sc.v_off[k]:sc.v_off[k + 1]]
    F = sc.patch_fields(k)
    im, jm = int(sc.imax[k]), int(sc.jmax[k])

    def ev(u, v, fields):
        lu = O.basis1d(un, u); lv = O.basis1d(vn, v)
        return np.array([O.eval_field(lu, lv, F[f], im, jm) for f in fields])

    X = ev(0.0, 0.0, (0, 1, 2)); Xu = ev(0.0, 0.0, (3, 4, 5)); Xv = ev(0.0, 0.0, (6, 7, 8))
    nu = ev(0.0, 0.0, (9, 10, 11))
    P = [ev(FEATURE_U[a], FEATURE_V[b], (0, 1, 2)) for a in range(5) for b in range(3)]
    return patch_frame(X, Xu, nu, P, Xv)
