"""The oracle's patch arithmetic pinned by what the mathematics fixes.

SurfacePatch.hpp is not buildable in this image (it needs <mkl.h>), so the
coarse grid, cell padding and Newton are pinned by closed forms:
  * barycentric weights of explicit node sets;
  * polynomial reproduction: data from a bidegree <= (n-1) polynomial is
    interpolated exactly by both the first form (the coarse grid,
    PatchInterpolation.hpp:54-159) and the second form (the Newton basis,
    SurfacePatch.hpp:518-555), and one-hot rows at nodes;
  * a paraboloid patch X = (u, v, a u^2 + b v^2): cell sagitta = |a+b| / 99^2
    (SurfacePatch.hpp:446-492), Newton on a vertical ray lands on
    (u, v, t) = (x0, y0, a x0^2 + b y0^2 - z0) (SurfacePatch.hpp:593-660);
  * each Newton failure forced on purpose: singular Jacobian, no convergence,
    out of domain.
Each check fails for a dropped term, a swapped index or a sign error.
"""
import numpy as np
import pytest

from oracle import oracle as O
from billardballraytracing3d_amd import scenes as S


def lob(n):
    return S.lobatto_nodes(n)


def test_barycentric_weights_closed_form():
    # nodes -1, 0, 1: w_j = 1 / prod_{i != j} (x_j - x_i) = (1/2, -1, 1/2)
    np.testing.assert_array_equal(O.bary_weights([-1.0, 0.0, 1.0]), [0.5, -1.0, 0.5])
    x = np.array([-1.0, -0.5, 0.25, 1.0])
    w = O.bary_weights(x)
    for j in range(4):
        assert w[j] == pytest.approx(1.0 / np.prod([x[j] - x[i] for i in range(4) if i != j]), rel=1e-15)


@pytest.mark.parametrize("n", [3, 5, 7])
@pytest.mark.parametrize("kind", ["lobatto", "cheb1"])
def test_first_form_reproduces_polynomials(n, kind):
    nodes = S.patch_nodes(kind, n)
    rng = np.random.default_rng(n)
    C = rng.normal(size=(n, n))            # bidegree (n-1, n-1) coefficients
    P = lambda u, v: np.einsum("ij,...i,...j->...", C, np.power.outer(u, np.arange(n)),
                               np.power.outer(v, np.arange(n)))
    F = P(nodes[:, None] * np.ones((1, n)), np.ones((n, 1)) * nodes[None, :])
    x = np.linspace(-1, 1, 100)           # the coarse grid, including the nodes +-1
    got = O.interp2d(nodes, nodes, F, x, x)
    want = P(x[:, None] * np.ones((1, 100)), np.ones((100, 1)) * x[None, :])
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    if kind == "lobatto":   # x = -1 and +1 are nodes: those rows are the node data exactly
        np.testing.assert_array_equal(got[0, 0], F[0, 0])
        np.testing.assert_array_equal(got[-1, -1], F[-1, -1])


@pytest.mark.parametrize("n", [3, 5])
def test_second_form_basis(n):
    nodes = lob(n)
    for x in [-0.93, -0.2, 0.31, 0.999]:
        b = O.basis1d(nodes, x)
        assert abs(b.sum() - 1.0) < 1e-15
        # reproduces x^k exactly
        for k in range(n):
            assert abs(b @ nodes ** k - x ** k) < 1e-14
    # one-hot on a node and within 1e-12 of it; not one-hot 1e-11 away
    b = O.basis1d(nodes, nodes[1])
    np.testing.assert_array_equal(b, np.eye(n)[1])
    b = O.basis1d(nodes, nodes[1] + 5e-13)
    np.testing.assert_array_equal(b, np.eye(n)[1])
    b = O.basis1d(nodes, nodes[1] + 1e-11)   # second form, not the one-hot row
    assert b[0] != 0.0 and abs(b[1] - 1.0) < 1e-9


def test_eval_field_skips_zero_rows_and_sums_rows():
    lu = np.array([0.0, 2.0, 0.0]); lv = np.array([1.0, 10.0, 100.0])
    f = np.arange(9, dtype=np.float64)
    assert O.eval_field(lu, lv, f, 3, 3) == 2.0 * (3.0 + 40.0 + 500.0)
    f2 = f.copy(); f2[0] = np.inf   # row 0 skipped: no 0 * inf = NaN
    assert O.eval_field(lu, lv, f2, 3, 3) == 2.0 * (3.0 + 40.0 + 500.0)


def test_solve3x3_against_lu():
    rng = np.random.default_rng(3)
    for _ in range(200):
        J = rng.normal(size=(3, 3)); b = rng.normal(size=3)
        ok, x = O.solve3x3(J, b)
        assert ok
        np.testing.assert_allclose(x, np.linalg.solve(J, b), rtol=1e-9, atol=1e-12)
    ok, _ = O.solve3x3(np.array([[1.0, 2, 3], [2, 4, 6], [0, 1, 1]]), np.ones(3))
    assert not ok
    ok, _ = O.solve3x3(np.diag([1e-5, 1e-5, 1e-5]), np.ones(3))   # |det| = 1e-15 < 1e-14
    assert not ok


def paraboloid_scene(a, b, n=3, scale=1.0, z0=0.0):
    """One tensor patch X(u,v) = (s u, s v, z0 + a u^2 + b v^2) on Lobatto nodes (exact for n >= 3)."""
    nodes = lob(n)
    U, V = np.meshgrid(nodes, nodes, indexing="ij")
    x, y = scale * U, scale * V
    z = z0 + a * U ** 2 + b * V ** 2
    xu = scale * np.ones_like(U); yu = 0 * U; zu = 2 * a * U
    xv = 0 * U; yv = scale * np.ones_like(U); zv = 2 * b * V
    nrm = np.stack([-zu * scale, -zv * scale, scale * scale * np.ones_like(U)])
    nrm /= np.linalg.norm(nrm, axis=0)
    fields = np.stack([x, y, z, xu, yu, zu, xv, yv, zv, nrm[0], nrm[1], nrm[2]]).reshape(12, -1)
    imax = np.array([n], np.int32)
    off = np.array([0, n], np.int64)
    return S.Scene(imax=imax, jmax=imax.copy(), u_off=off, v_off=off.copy(), f_off=np.array([0, n * n], np.int64),
                   unodes=nodes.copy(), vnodes=nodes.copy(), fields=np.ascontiguousarray(fields.ravel()),
                   obstacle=np.zeros(1, np.int32), order=np.full(1, 2, np.int32))


@pytest.mark.parametrize("a,b", [(0.3, 0.2), (-0.5, 0.1), (0.0, 0.0), (1.0, -1.0)])
def test_paraboloid_sagitta_closed_form(a, b):
    osc = O.OracleScene(paraboloid_scene(a, b))
    xyz, sag, cn = osc.coarse(0)
    assert np.max(np.abs(sag - abs(a + b) / 99.0 ** 2)) < 1e-15
    box, cent, ms = osc.patch_info()
    assert abs(ms[0] - abs(a + b) / 99.0 ** 2) < 1e-15
    # coarse vertex (i, j) is X(u_i, v_j), u_i = -1 + 2 i / 99
    u = -1.0 + 2.0 * np.arange(100) / 99.0
    U, V = np.meshgrid(u, u, indexing="ij")
    np.testing.assert_allclose(xyz[:, 2], (a * U ** 2 + b * V ** 2).ravel(), atol=1e-14)
    np.testing.assert_allclose(xyz[:, 0], U.ravel(), atol=1e-15)
    # box = node box inflated by the max sagitta; centroid = node mean
    assert box[0, 0] == -1.0 - ms[0] and box[0, 3] == 1.0 + ms[0]


@pytest.mark.parametrize("mode", [0, 1])
def test_paraboloid_newton_closed_form(mode):
    a, b = 0.3, 0.2
    osc = O.OracleScene(paraboloid_scene(a, b))
    rng = np.random.default_rng(5)
    for _ in range(40):
        x0, y0 = rng.uniform(-0.95, 0.95, 2)
        o = np.array([x0, y0, -5.0]); d = np.array([0.0, 0.0, 1.0])
        st, t, p, n, it = osc.patch_intersect(0, o, d, mode)
        assert st == 1
        want_t = a * x0 ** 2 + b * y0 ** 2 + 5.0
        tol = 1e-6 if mode == 0 else 1e-13
        assert abs(t - want_t) < tol
        assert np.max(np.abs(p - [x0, y0, want_t - 5.0])) < tol
        assert 1 <= it <= 20
        if mode == 1:   # normal = the nodal unit normals interpolated at (x0, y0), then normalised
            nodes = lob(3)
            bu, bv = O.basis1d(nodes, x0), O.basis1d(nodes, y0)
            U, V = np.meshgrid(nodes, nodes, indexing="ij")
            nod = np.stack([-2 * a * U, -2 * b * V, np.ones_like(U)])
            nod /= np.linalg.norm(nod, axis=0)
            nn = np.einsum("i,j,cij->c", bu, bv, nod)
            nn /= np.linalg.norm(nn)
            assert np.max(np.abs(n - nn)) < 1e-12


def test_newton_quadratic_convergence():
    """Newton from a rough seed on the paraboloid: step sizes shrink quadratically."""
    a, b = 0.6, -0.4
    osc = O.OracleScene(paraboloid_scene(a, b))
    x0, y0 = 0.4, -0.3
    o = np.array([x0, y0, -2.0]); d = np.array([0.0, 0.0, 1.0])
    want = a * x0 ** 2 + b * y0 ** 2 + 2.0
    errs = []
    for it_cap_seed in [(0.0, 0.0, 1.0)]:
        st, t, p, n, it = osc.newton(0, o, d, *it_cap_seed, mode=1)
        assert st == 1 and abs(t - want) < 1e-13
        errs.append(it)
    assert errs[0] <= 8


def test_newton_forced_failures():
    osc = O.OracleScene(paraboloid_scene(0.0, 0.0))   # flat z = 0
    # singular: ray in the plane of the patch (d in span(Xu, Xv)) -> det = 0 at the first iteration
    st, *_ = osc.newton(0, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], 0.1, 0.1, 1.0, mode=0)
    assert st == -1
    # out of domain: converges to u = 1.5 on the extended plane
    st, *_ = osc.newton(0, [1.5, 0.0, -1.0], [0.0, 0.0, 1.0], 0.9, 0.0, 0.5, mode=0)
    assert st == -3
    # no convergence: a horizontal ray below a bowl z = u^2 + v^2 + 1 never meets it
    bowl = O.OracleScene(paraboloid_scene(1.0, 1.0, z0=1.0))
    st, *_ = bowl.newton(0, [-3.0, 0.0, 0.0], [1.0, 0.0, 0.0], 0.5, 0.0, 3.0, mode=0)
    assert st in (-2, -1)
    st, *_ = bowl.newton(0, [-3.0, 0.0, 0.0], [1.0, 0.0, 0.0], 0.5, 0.3, 3.0, mode=0)
    assert st == -2


def test_stage1_picks_closest_padded_triangle():
    """Vertical ray on a paraboloid: stage 1 hits; brute-force MT tests = 6 * 99^2 per candidate."""
    osc = O.OracleScene(paraboloid_scene(0.3, 0.2))
    # from below the ray meets the back face (normals point +z): RayTracingDriver.hpp:220-226
    res = osc.trace([[0.1, 0.2, -5.0]], [[0.0, 0.0, 1.0]], max_bounces=1, mode=1)
    assert res.status[0] == 2 and res.err_kind[0] == 4 and res.bounces[0] == 0 and res.err_patch[0] == 0
    res = osc.trace([[0.1, 0.2, 5.0]], [[0.0, 0.0, -1.0]], max_bounces=1, mode=1)
    assert res.status[0] == 1 and res.bounces[0] == 1 and res.hit_patch[0, 0] == 0
    assert res.counters[2] == 1 and res.counters[3] == 6 * 99 * 99
    np.testing.assert_allclose(res.hit_xyz[0, 0], [0.1, 0.2, 0.3 * 0.01 + 0.2 * 0.04], atol=1e-13)
