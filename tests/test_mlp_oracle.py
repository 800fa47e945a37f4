"""Pins of the learned-seed oracle (oracle/mlp_oracle.py) against torch and closed forms."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def _rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(M.N_PARAMS) * scale).astype(np.float32)
    return p * M.live_mask()


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000)).astype(np.float32)
    # exact ties: low 16 bits = 0x8000 with both parities of bit 16
    ties = (np.arange(1000, dtype=np.uint32) << 17 | 0x3F800000 | 0x8000).view(np.float32)
    ties2 = (np.arange(1000, dtype=np.uint32) << 17 | 0x3F810000 | 0x8000).view(np.float32)
    for x in (a, ties, ties2, -a):
        want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
        np.testing.assert_array_equal(M.bf16(x), want)


def test_gelu_matches_torch_tanh_form_and_derivative():
    z = np.linspace(-9, 9, 4001)
    g, gp = M.gelu(z)
    zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    gt = torch.nn.functional.gelu(zt, approximate="tanh")
    gt.sum().backward()
    np.testing.assert_allclose(g, gt.detach().numpy(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(gp, zt.grad.numpy(), rtol=0, atol=1e-13)
    h = 1e-6
    fd = (M.gelu(z + h)[0] - M.gelu(z - h)[0]) / (2 * h)
    np.testing.assert_allclose(gp, fd, rtol=0, atol=1e-8)


def _torch_loss(p, x, t):
    o = [M.OFF_W1, M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3, M.OFF_W4, M.OFF_B4, M.N_PARAMS]
    W1, b1, W2, b2, W3, b3, W4, b4 = [p[o[i]:o[i + 1]] for i in range(8)]
    W1, W2, W3, W4 = W1.reshape(64, 64), W2.reshape(64, 64), W3.reshape(64, 64), W4.reshape(16, 64)
    X = torch.zeros((x.shape[0], 64), dtype=torch.float64)
    X[:, :x.shape[1]] = torch.tensor(x, dtype=torch.float64)
    g = lambda z: torch.nn.functional.gelu(z, approximate="tanh")
    h = g(X @ W1.T + b1)
    h = g(h @ W2.T + b2)
    h = g(h @ W3.T + b3)
    y = (h @ W4.T + b4)[:, :3]
    return ((y - torch.tensor(t)) ** 2).sum() / x.shape[0], y


def test_exact_forward_backward_match_torch_autograd():
    p = _rand_params(1).astype(np.float64)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((37, M.N_IN))
    t = rng.standard_normal((37, 3))
    y, _ = M.forward(p, x, bf16_ops=False)
    loss, g = M.loss_grad(p, x, t, bf16_ops=False)
    pt = torch.tensor(p, requires_grad=True)
    lt, yt = _torch_loss(pt, x, t)
    lt.backward()
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-13, atol=1e-13)
    assert abs(loss - lt.item()) <= 1e-12 * max(1.0, abs(loss))
    gt = pt.grad.numpy() * M.live_mask()
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-12)


def test_exact_gradient_matches_finite_differences():
    p = _rand_params(3).astype(np.float64)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((9, M.N_IN)); t = rng.standard_normal((9, 3))
    _, g = M.loss_grad(p, x, t, bf16_ops=False)
    live = np.flatnonzero(M.live_mask())
    idx = rng.choice(live, 60, replace=False)
    idx = np.concatenate([idx, [M.OFF_B1, M.OFF_B2 + 5, M.OFF_B3 + 63, M.OFF_B4, M.OFF_B4 + 2]])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (M.loss_grad(pp, x, t, False)[0] - M.loss_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_mode_is_a_small_perturbation_of_exact():
    p = _rand_params(5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((256, M.N_IN)); t = rng.standard_normal((256, 3))
    l0, g0 = M.loss_grad(p, x, t, False)
    l1, g1 = M.loss_grad(p, x, t, True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)
    # the pads never receive gradient
    assert np.all(g1[M.live_mask() == 0] == 0)


def test_init_params_glorot_bounds_and_pads():
    p = M.init_params(7)
    W1, b1, W2, b2, W3, b3, W4, b4 = M.unpack(p)
    assert np.all(W1[:, 51:] == 0) and np.all(W4[3:] == 0) and np.all(b1 == 0) and np.all(b4 == 0)
    a1 = math.sqrt(6.0 / (51 + 64)); a4 = math.sqrt(6.0 / (64 + 3))
    assert np.abs(W1[:, :51]).max() < a1 and np.abs(W1[:, :51]).max() > 0.9 * a1
    assert np.abs(W4[:3]).max() < a4


def _rot(seed):
    q = np.random.default_rng(seed).standard_normal(4); q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def test_features_planar_patch_closed_form():
    R = _rot(8); E1, E2, N = R[:, 0], R[:, 1], R[:, 2]
    c0 = np.array([0.3, -1.2, 2.0]); a, b = 0.2, 0.05
    X = c0; Xu = a * E1; Xv = b * E2; nu = 3.0 * N
    P = [c0 + a * M.FEATURE_U[i] * E1 + b * M.FEATURE_V[j] * E2 for i in range(5) for j in range(3)]
    fr = M.patch_frame(X, Xu, nu, P, Xv)
    c, e1, e2, n, s, f, au, av = fr
    np.testing.assert_allclose(au, E1 / a, atol=1e-12); np.testing.assert_allclose(av, E2 / b, atol=1e-12)
    np.testing.assert_allclose(e1, E1, atol=1e-15); np.testing.assert_allclose(e2, E2, atol=1e-15)
    np.testing.assert_allclose(n, N, atol=1e-15)
    assert abs(s - math.hypot(a, b)) < 1e-15
    want = np.array([[a * M.FEATURE_U[i] / s, b * M.FEATURE_V[j] / s, 0.0] for i in range(5) for j in range(3)])
    np.testing.assert_allclose(f.reshape(15, 3), want, atol=1e-7)
    # a ray through c0 + 0.1 E1 + 0.02 E2 along -N (from 5 units above): closest point to c0 is that point
    d = -N; o = c0 + 0.1 * E1 + 0.02 * E2 + 5.0 * N
    feats = M.ray_features(fr, o, d)
    r, tc, ua, va = feats
    assert abs(tc - 5.0) < 1e-14
    np.testing.assert_allclose(r, [0.1 / s, 0.02 / s, 0.0, 0.0, 0.0, -1.0], atol=1e-7)
    # on a plane the tangent-plane estimate is the exact root: (0.1/a, 0.02/b) = (0.5, 0.4)
    assert abs(ua - 0.5) < 1e-14 and abs(va - 0.4) < 1e-14
    u0, v0, t0 = M.seed_from_output(fr, feats, [0.0, 0.0, 0.0])
    assert abs(u0 - 0.5) < 1e-14 and abs(v0 - 0.4) < 1e-14 and abs(t0 - 5.0) < 1e-14
    np.testing.assert_allclose(M.target_from_root(fr, feats, 0.5, 0.4, 5.0 + 2 * s), [0.0, 0.0, 2.0], atol=1e-7)
    # an oblique ray through the same point: the plane hit is still exact; far off the patch it is clamped
    d2 = -N + 0.3 * E1; d2 /= np.linalg.norm(d2)
    o2 = c0 + 0.1 * E1 + 0.02 * E2 - 4.0 * d2
    _, _, ua2, va2 = M.ray_features(fr, o2, d2)
    assert abs(ua2 - 0.5) < 1e-12 and abs(va2 - 0.4) < 1e-12
    o3 = c0 + 3.0 * E1 + 5.0 * N
    _, _, ua3, va3 = M.ray_features(fr, o3, d)
    assert ua3 == M.PLANE_CLAMP and abs(va3) < 1e-14


def test_features_invariant_under_rigid_motion():
    rng = np.random.default_rng(9)
    X = rng.standard_normal(3); Xu = rng.standard_normal(3); nu = rng.standard_normal(3)
    P = [X + 0.3 * rng.standard_normal(3) for _ in range(15)]
    o = rng.standard_normal(3) * 3; d = rng.standard_normal(3); d /= np.linalg.norm(d)
    Xv = rng.standard_normal(3)
    f0 = M.patch_frame(X, Xu, nu, P, Xv); r0, tc0, ua0, va0 = M.ray_features(f0, o, d)
    R = _rot(10); T = np.array([4.0, -2.0, 7.0])
    f1 = M.patch_frame(R @ X + T, R @ Xu, R @ nu, [R @ p + T for p in P], R @ Xv)
    r1, tc1, ua1, va1 = M.ray_features(f1, R @ o + T, R @ d)
    np.testing.assert_allclose(f1[5], f0[5], atol=2e-6)
    np.testing.assert_allclose(r1, r0, atol=2e-6)
    assert abs(tc1 - tc0) < 1e-12 and abs(ua1 - ua0) < 1e-9 and abs(va1 - va0) < 1e-9
