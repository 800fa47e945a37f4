"""Synthetic billiard scenes and ray sources, seeded and bit-reproducible.

The reference reads curved surface patches from IFGF-RP ``.txt`` files
(``include/ScatObjects/SurfacePatch.hpp:313-371``) and has no scene generator
of its own.  This module makes the scenes the benchmark workloads are quoted
on: spheres and ellipsoids meshed as curved Lagrange triangles of order 2
(6 nodes) or 4 (15 nodes) whose nodes sit on the exact surface.

The reference only understands tensor-product ``imax x jmax`` patches, so each
order-p triangle is written as an exact ``(p+1) x (p+1)`` tensor patch: the
triangle map ``P(lambda)`` (total degree p in barycentric coordinates) is
composed with the bilinear map of ``[-1,1]^2`` onto the triangle that sends
the corners ``(-1,-1), (1,-1), (1,1), (-1,1)`` to ``A, B, mid(BC), C``.  The
composition has bidegree ``(p, p)`` so the tensor Lagrange interpolant through
``(p+1)^2`` Chebyshev-Lobatto nodes reproduces the triangle (and its first
derivatives) exactly.  Lobatto nodes include the patch edges: the reference
bounds a patch by the box of its nodes (SurfacePatch.hpp:380-407), and
interior-only nodes leave the patch rim outside that box, so rays tunnel
through the rim (2% of first hits on the two-sphere scene).  The Jacobian of that map vanishes only at the corner sent to
``mid(BC)``.  Normals are the normalised gradient of the implicit
sphere/ellipsoid function at each node (outward).

Random numbers come from SplitMix64 with an explicit integer-to-double map
(``(z >> 11) * 2**-53``), so CPU and GPU consume bit-identical inputs.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# Field order of the 12 nodal fields the bounce loop reads
# (SurfacePatch::Data, SurfacePatch.hpp:74-80; dS and mask are not used).
FIELD_NAMES = ("x", "y", "z", "dxdu", "dydu", "dzdu", "dxdv", "dydv", "dzdv", "nuX", "nuY", "nuZ")
N_FIELDS = 12

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


# --------------------------------------------------------------------------
# Random numbers
# --------------------------------------------------------------------------
class SplitMix64:
    """SplitMix64 stream (Steele, Lea & Flood 2014) with a vectorised draw."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + _GOLDEN * k
            self.state = self.state + _GOLDEN * np.uint64(n)
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        """n doubles in [lo, hi): lo + (hi - lo) * ((z >> 11) * 2**-53)."""
        z = self.next_u64(n)
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """n integers uniform in [lo, hi] (inclusive), by multiply-shift."""
        z = self.next_u64(n) >> np.uint64(11)
        span = hi - lo + 1
        return lo + ((z.astype(np.float64) * (1.0 / 9007199254740992.0)) * span).astype(np.int64)


def random_rotation(rng: SplitMix64) -> np.ndarray:
    """Uniform random rotation from a unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.uniform(3)
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    q = np.array([a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                  b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


# --------------------------------------------------------------------------
# Polyhedra and subdivision
# --------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def _octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4],
                  [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]], dtype=np.int64)
    return v, f


def _orient_outward(v, f):
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, a + b + c) < 0
    f = f.copy()
    f[flip, 1], f[flip, 2] = f[flip, 2].copy(), f[flip, 1].copy()
    return f


def sphere_triangles(kind: str, freq: int, rot: Optional[np.ndarray] = None) -> np.ndarray:
    """Flat-subdivided, radially projected triangles of a unit sphere.

    kind "ico" gives 20*freq^2 triangles, "oct" gives 8*freq^2.  Returns an
    array (T, 3, 3) of corner positions on the unit sphere, outward winding.
    """
    v, f = _icosahedron() if kind == "ico" else _octahedron()
    f = _orient_outward(v, f)
    if rot is not None:
        v = v @ rot.T
    # corner (i, j) of the subdivision, per face; triangles in the order "up (i,j), then down (i,j)"
    idx = []
    for i in range(freq):
        for j in range(freq - i):
            idx.append(((i, j), (i + 1, j), (i, j + 1)))
            if i + j < freq - 1:
                idx.append(((i + 1, j), (i + 1, j + 1), (i, j + 1)))
    ij = np.asarray(idx, dtype=np.int64)                             # (t, 3, 2)
    wi = ij[..., 0].astype(np.float64)[None, :, :, None]
    wj = ij[..., 1].astype(np.float64)[None, :, :, None]
    w0 = (freq - ij[..., 0] - ij[..., 1]).astype(np.float64)[None, :, :, None]
    F = v[f]                                                         # (faces, 3, 3)
    A, B, C = (F[:, k][:, None, None, :] for k in range(3))
    p = A * w0 / freq + B * wi / freq + C * wj / freq                # (faces, t, 3, 3)
    p = p / np.sqrt((p * p).sum(axis=-1, keepdims=True))
    return np.ascontiguousarray(p.reshape(-1, 3, 3), dtype=np.float64)


# --------------------------------------------------------------------------
# Order-p Lagrange triangle -> exact (p+1)x(p+1) tensor patch
# --------------------------------------------------------------------------
def chebyshev_nodes(n: int) -> np.ndarray:
    """Chebyshev points of the first kind on (-1, 1), ascending."""
    k = np.arange(n, dtype=np.float64)
    return -np.cos((2.0 * k + 1.0) * math.pi / (2.0 * n))


def lobatto_nodes(n: int) -> np.ndarray:
    """Chebyshev-Lobatto points on [-1, 1] (endpoints included), ascending, exact -1/0/1."""
    k = np.arange(n, dtype=np.float64)
    x = -np.cos(k * math.pi / (n - 1))
    x[0] = -1.0; x[-1] = 1.0
    if n % 2 == 1:
        x[n // 2] = 0.0
    return x


def patch_nodes(kind: str, n: int) -> np.ndarray:
    return lobatto_nodes(n) if kind == "lobatto" else chebyshev_nodes(n)


def _tri_node_indices(p: int):
    return [(p - i - j, i, j) for i in range(p + 1) for j in range(p + 1 - i)]


def _univariate(p, k, x):
    """L_k(x) = prod_{m<k} (p x - m)/(m+1) and its derivative, vectorised."""
    val = np.ones_like(x)
    der = np.zeros_like(x)
    for m in range(k):
        fac = (p * x - m) / (m + 1.0)
        der = der * fac + val * (p / (m + 1.0))
        val = val * fac
    return val, der


def _bilinear_lambda(u, v):
    """Barycentric coordinates (and d/du, d/dv) of the quad->triangle map.

    Corners (-1,-1)->A=(1,0,0), (1,-1)->B=(0,1,0), (1,1)->mid(BC)=(0,.5,.5),
    (-1,1)->C=(0,0,1).
    """
    eA = np.array([1.0, 0.0, 0.0]); eB = np.array([0.0, 1.0, 0.0])
    eM = np.array([0.0, 0.5, 0.5]); eC = np.array([0.0, 0.0, 1.0])
    u = u[..., None]; v = v[..., None]
    lam = 0.25 * ((1 - u) * (1 - v) * eA + (1 + u) * (1 - v) * eB + (1 + u) * (1 + v) * eM + (1 - u) * (1 + v) * eC)
    dlu = 0.25 * (-(1 - v) * eA + (1 - v) * eB + (1 + v) * eM - (1 + v) * eC)
    dlv = 0.25 * (-(1 - u) * eA - (1 + u) * eB + (1 + u) * eM + (1 - u) * eC)
    return lam, dlu, dlv


def triangle_basis(p: int, lam: np.ndarray):
    """Lagrange basis of the equispaced order-p triangle and its gradient.

    lam: (..., 3).  Returns phi (..., K) and dphi (..., K, 3) with K=(p+1)(p+2)/2.
    """
    idx = _tri_node_indices(p)
    phis, grads = [], []
    for (a, b, c) in idx:
        va, da = _univariate(p, a, lam[..., 0])
        vb, db = _univariate(p, b, lam[..., 1])
        vc, dc = _univariate(p, c, lam[..., 2])
        phis.append(va * vb * vc)
        grads.append(np.stack([da * vb * vc, va * db * vc, va * vb * dc], axis=-1))
    return np.stack(phis, axis=-1), np.stack(grads, axis=-2)


def triangle_nodes(p: int, corners: np.ndarray) -> np.ndarray:
    """Order-p triangle nodes radially projected onto the unit sphere.

    corners: (T, 3, 3) -> (T, K, 3) in _tri_node_indices order.
    """
    idx = np.array(_tri_node_indices(p), dtype=np.float64) / p  # (K,3)
    pts = np.einsum("kc,tcd->tkd", idx, corners)
    return pts / np.linalg.norm(pts, axis=-1, keepdims=True)


@dataclass
class Obstacle:
    kind: str            # "sphere" or "ellipsoid"
    center: np.ndarray   # (3,)
    axes: np.ndarray     # (3,) semi-axes; equal for a sphere
    mesh: str            # "ico" or "oct"
    freq: int
    order: int           # 2 or 4
    rot_seed: int
    node_kind: str = "lobatto"   # tensor node set: "lobatto" or "cheb1"
    n_nodes: int = 0             # nodes per patch side; 0 means order + 1
    rotate: bool = True          # False: the polyhedron in its standard orientation (octahedron vertices on the axes)


def obstacle_patches(ob: Obstacle) -> dict:
    """Tensor patches of one obstacle: dict of nodes, fields (T, 12, n*n) and triangle nodes."""
    rot = random_rotation(SplitMix64(ob.rot_seed)) if ob.rotate else None
    corners = sphere_triangles(ob.mesh, ob.freq, rot)
    p = ob.order
    tn = triangle_nodes(p, corners)                       # (T,K,3) on unit sphere
    nodes = patch_nodes(ob.node_kind, ob.n_nodes if ob.n_nodes > 0 else p + 1)
    uu, vv = np.meshgrid(nodes, nodes, indexing="ij")     # flat index i*jmax + j
    lam, dlu, dlv = _bilinear_lambda(uu.ravel(), vv.ravel())
    phi, dphi = triangle_basis(p, lam)                    # (n2,K), (n2,K,3)
    dphi_du = np.einsum("nkc,nc->nk", dphi, dlu)
    dphi_dv = np.einsum("nkc,nc->nk", dphi, dlv)
    X = np.einsum("nk,tkd->tnd", phi, tn)
    Xu = np.einsum("nk,tkd->tnd", dphi_du, tn)
    Xv = np.einsum("nk,tkd->tnd", dphi_dv, tn)
    ax = np.asarray(ob.axes, dtype=np.float64)
    c = np.asarray(ob.center, dtype=np.float64)
    Xw = X * ax + c
    Xuw = Xu * ax
    Xvw = Xv * ax
    g = X / ax  # gradient of sum((x-c)_i^2 / a_i^2) up to a factor 2, at Xw
    nrm = g / np.linalg.norm(g, axis=-1, keepdims=True)
    fields = np.concatenate([Xw, Xuw, Xvw, nrm], axis=-1)  # (T, n2, 12)
    fields = np.ascontiguousarray(fields.transpose(0, 2, 1))  # (T, 12, n2)
    tri_nodes_w = tn * ax + c
    return {"nodes": nodes, "fields": fields, "tri_nodes": tri_nodes_w, "order": p}


@dataclass
class Scene:
    """A set of tensor patches (all obstacles, concatenated in patch-index order)."""
    imax: np.ndarray          # (N,) int32
    jmax: np.ndarray          # (N,) int32
    u_off: np.ndarray         # (N+1,) int64 offsets into unodes/uweights
    v_off: np.ndarray         # (N+1,) int64
    f_off: np.ndarray         # (N+1,) int64 node offsets; patch k's fields at 12*f_off[k]
    unodes: np.ndarray        # float64
    vnodes: np.ndarray        # float64
    fields: np.ndarray        # float64, length 12*f_off[N]
    obstacle: np.ndarray      # (N,) int32 patch -> obstacle id
    order: np.ndarray         # (N,) int32 triangle order of each patch
    tri_nodes: List[np.ndarray] = field(default_factory=list)  # per obstacle (T,K,3)
    obstacles: List[Obstacle] = field(default_factory=list)

    @property
    def n_patches(self) -> int:
        return int(self.imax.shape[0])

    def patch_fields(self, k: int) -> np.ndarray:
        n2 = int(self.imax[k] * self.jmax[k])
        o = 12 * int(self.f_off[k])
        return self.fields[o:o + 12 * n2].reshape(12, n2)


def build_scene(obstacles: Sequence[Obstacle]) -> Scene:
    imax, jmax, unodes, fields, obst, order, tri_nodes = [], [], [], [], [], [], []
    for oi, ob in enumerate(obstacles):
        pt = obstacle_patches(ob)
        T = pt["fields"].shape[0]
        n = pt["nodes"].shape[0]
        imax.append(np.full(T, n, np.int32)); jmax.append(np.full(T, n, np.int32))
        unodes.append(np.tile(pt["nodes"], T))
        fields.append(pt["fields"].reshape(-1))
        obst.append(np.full(T, oi, np.int32)); order.append(np.full(T, ob.order, np.int32))
        tri_nodes.append(pt["tri_nodes"])
    imax = np.concatenate(imax); jmax = np.concatenate(jmax)
    u_off = np.concatenate([[0], np.cumsum(imax.astype(np.int64))]).astype(np.int64)
    v_off = np.concatenate([[0], np.cumsum(jmax.astype(np.int64))]).astype(np.int64)
    f_off = np.concatenate([[0], np.cumsum(imax.astype(np.int64) * jmax)]).astype(np.int64)
    un = np.ascontiguousarray(np.concatenate(unodes))
    return Scene(imax=imax, jmax=jmax, u_off=u_off, v_off=v_off, f_off=f_off,
                 unodes=un, vnodes=un.copy(), fields=np.ascontiguousarray(np.concatenate(fields)),
                 obstacle=np.concatenate(obst), order=np.concatenate(order),
                 tri_nodes=tri_nodes, obstacles=list(obstacles))


def subset_scene(sc: Scene, patches: Sequence[int]) -> Scene:
    """A scene made of the listed patches (in the listed order)."""
    ks = [int(k) for k in patches]
    imax = sc.imax[ks].copy(); jmax = sc.jmax[ks].copy()
    un = np.concatenate([sc.unodes[sc.u_off[k]:sc.u_off[k + 1]] for k in ks])
    vn = np.concatenate([sc.vnodes[sc.v_off[k]:sc.v_off[k + 1]] for k in ks])
    fl = np.concatenate([sc.fields[12 * sc.f_off[k]:12 * sc.f_off[k + 1]] for k in ks])
    u_off = np.concatenate([[0], np.cumsum(imax.astype(np.int64))]).astype(np.int64)
    v_off = np.concatenate([[0], np.cumsum(jmax.astype(np.int64))]).astype(np.int64)
    f_off = np.concatenate([[0], np.cumsum(imax.astype(np.int64) * jmax)]).astype(np.int64)
    return Scene(imax=imax, jmax=jmax, u_off=u_off, v_off=v_off, f_off=f_off, unodes=un, vnodes=vn,
                 fields=fl, obstacle=sc.obstacle[ks].copy(), order=sc.order[ks].copy())


# --------------------------------------------------------------------------
# The benchmark scenes
# --------------------------------------------------------------------------
def two_spheres_scene(seed: int = 0, freq: int = 4, order: int = 2) -> Scene:
    """Two unit spheres at (+-1.5,0,0), icosphere frequency 4 (320 P2 patches each)."""
    obs = [Obstacle("sphere", np.array([-1.5, 0.0, 0.0]), np.array([1.0, 1.0, 1.0]), "ico", freq, order,
                    rot_seed=(seed << 8) + 0),
           Obstacle("sphere", np.array([1.5, 0.0, 0.0]), np.array([1.0, 1.0, 1.0]), "ico", freq, order,
                    rot_seed=(seed << 8) + 1)]
    return build_scene(obs)


def sinai_obstacles(seed: int = 0, freq: int = 16, order: int = 4) -> List[Obstacle]:
    """4 spheres r~U(0.5,1.2) + ellipsoid (1.5,1,0.6), centres U([-4,4]^3), non-overlapping."""
    rng = SplitMix64(0x5137A1 + seed)
    radii = list(rng.uniform(4, 0.5, 1.2)) + [1.5]
    axes = [np.array([r, r, r]) for r in radii[:4]] + [np.array([1.5, 1.0, 0.6])]
    centres: List[np.ndarray] = []
    for r in radii:
        while True:
            c = rng.uniform(3, -4.0, 4.0)
            if all(np.linalg.norm(c - c2) > r + r2 + 0.05 for c2, r2 in zip(centres, radii)):
                centres.append(c)
                break
    obs = []
    for i in range(5):
        kind = "sphere" if i < 4 else "ellipsoid"
        obs.append(Obstacle(kind, centres[i], axes[i], "oct", freq, order, rot_seed=(seed << 8) + 16 + i))
    return obs


def sinai_scene(seed: int = 0, freq: int = 16, order: int = 4) -> Scene:
    return build_scene(sinai_obstacles(seed, freq, order))


def curved_ellipsoid_scene(seed: int = 0, freq: int = 16, order: int = 4) -> Scene:
    """One high-curvature ellipsoid, semi-axes (5, 1, 0.5), 2,048 order-4 patches (octahedral frequency 16)."""
    ob = Obstacle("ellipsoid", np.zeros(3), np.array([5.0, 1.0, 0.5]), "oct", freq, order, rot_seed=(seed << 8) + 40)
    return build_scene([ob])


def plane_grid(corner0, corner1, nx: int, ny: int) -> np.ndarray:
    """Inclusive nx x ny grid between two corners on an axis-aligned plane.

    Restates PlaneSourceDriver::build_plane_grid (PlaneSourceDriver.hpp:190-243):
    exactly one coordinate must agree within 1e-9*scale; nx runs along the
    first free axis and ny along the second (x/y/z order); point k = i*ny + j.
    """
    c0 = [float(v) for v in corner0]
    c1 = [float(v) for v in corner1]
    scale = max([1.0] + [abs(v) for v in c0 + c1])
    tol = 1e-9 * scale
    const = [a for a in range(3) if abs(c0[a] - c1[a]) < tol]
    if len(const) != 1:
        raise ValueError("corner0 and corner1 must agree in exactly one coordinate "
                         f"to define an axis-aligned plane (got {len(const)} matching coordinates)")
    ca = const[0]
    fa, fb = [a for a in range(3) if a != ca]
    ga = np.array([c0[fa] if nx == 1 else c0[fa] + (c1[fa] - c0[fa]) * float(i) / (nx - 1) for i in range(nx)])
    gb = np.array([c0[fb] if ny == 1 else c0[fb] + (c1[fb] - c0[fb]) * float(j) / (ny - 1) for j in range(ny)])
    pts = np.empty((nx * ny, 3))
    pts[:, ca] = c0[ca]
    pts[:, fa] = np.repeat(ga, ny)
    pts[:, fb] = np.tile(gb, nx)
    return pts


def cone_rays(n: int, seed: int, half_extent: float, z0: float, cone_full_deg: float,
              centre_xy=(0.0, 0.0)):
    """n rays: origins U([-h,h]^2) on z=z0, directions solid-angle uniform in a cone about +z.

    cone_full_deg is the full apex angle (the convention of BounceMapDriver.hpp:88-94).
    """
    rng = SplitMix64(0xC0FFEE + seed)
    u = rng.uniform(4 * n).reshape(n, 4)
    o = np.empty((n, 3))
    o[:, 0] = centre_xy[0] - half_extent + 2.0 * half_extent * u[:, 0]
    o[:, 1] = centre_xy[1] - half_extent + 2.0 * half_extent * u[:, 1]
    o[:, 2] = z0
    cmax = math.cos(math.radians(cone_full_deg) / 2.0)
    mu = cmax + (1.0 - cmax) * u[:, 2]
    st = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    ph = 2.0 * math.pi * u[:, 3]
    d = np.stack([st * np.cos(ph), st * np.sin(ph), mu], axis=1)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


def two_spheres_plane_source(seed: int = 0, nx: int = 64, ny: int = 64):
    """Two P2 unit spheres + 64x64 plane source at z=-5 along +z; 20 bounces."""
    sc = two_spheres_scene(seed)
    o = plane_grid((-3.0, -3.0, -5.0), (3.0, 3.0, -5.0), nx, ny)
    d = np.tile(np.array([0.0, 0.0, 1.0]), (o.shape[0], 1))
    return sc, o, np.ascontiguousarray(d), 20


def sinai_cone(seed: int = 0, n_rays: int = 1 << 20):
    """Sinai-like P4 scene (10,240 patches) + n_rays in a 10 degree cone from z=-8; 100 bounces."""
    sc = sinai_scene(seed)
    o, d = cone_rays(n_rays, seed, 6.0, -8.0, 10.0)
    return sc, o, d, 100


def batch_obstacles(rng: "SplitMix64", patch_lo: float, patch_hi: float) -> List[Obstacle]:
    """One scene of the batched bounce map: 1-8 obstacles (sphere or ellipsoid, 50/50), patch counts
    log-uniform in [patch_lo, patch_hi] snapped to an icosphere (20 f^2) or octasphere (8 f^2), order 2 or 4,
    non-overlapping centres in [-4, 4]^3."""
    n_obs = int(rng.integers(1, 1, 8)[0])
    specs = []
    for _ in range(n_obs):
        kind = "sphere" if rng.uniform(1)[0] < 0.5 else "ellipsoid"
        target = math.exp(rng.uniform(1, math.log(patch_lo), math.log(patch_hi))[0])
        f_ico = max(2, int(round(math.sqrt(target / 20.0))))
        f_oct = max(3, int(round(math.sqrt(target / 8.0))))
        mesh, f = ("ico", f_ico) if abs(math.log(20 * f_ico ** 2 / target)) <= abs(math.log(8 * f_oct ** 2 / target)) \
            else ("oct", f_oct)
        order = 2 if rng.uniform(1)[0] < 0.5 else 4
        if kind == "sphere":
            r = rng.uniform(1, 0.5, 1.2)[0]
            axes = np.array([r, r, r])
        else:
            axes = rng.uniform(3, 0.5, 1.5)
        specs.append((kind, mesh, f, order, axes))
    centres: List[np.ndarray] = []
    obs = []
    for i, (kind, mesh, f, order, axes) in enumerate(specs):
        rad = float(np.max(axes))
        for _ in range(10000):
            c = rng.uniform(3, -4.0, 4.0)
            if all(np.linalg.norm(c - c2) > rad + r2 + 0.05 for c2, r2 in centres):
                break
        centres.append((c, rad))
        obs.append(Obstacle(kind, c, axes, mesh, f, order, rot_seed=int(rng.integers(1, 0, 1 << 30)[0])))
    return obs


def batch_config(seed: int = 0, n_scenes: int = 256, patch_lo: float = 80, patch_hi: float = 5120,
                 rays_lo: float = 1000, rays_hi: float = 64000, cone_deg: float = 30.0,
                 total_rays: Optional[int] = None):
    """The batched bounce map: n_scenes independent scenes (batch_obstacles) with ray counts log-uniform in
    [rays_lo, rays_hi], origins uniform on z = -8 over [-6, 6]^2, directions uniform in a cone about +z.
    With total_rays the drawn counts are scaled (keeping their proportions) to sum to it.

    Returns (scenes, origins, dirs, ray_begin): the rays of scene i are ray_begin[i] .. ray_begin[i+1]-1.
    """
    rng = SplitMix64(0xBA7C4 + seed)
    scenes, counts = [], []
    for k in range(n_scenes):
        scenes.append(build_scene(batch_obstacles(rng, patch_lo, patch_hi)))
        counts.append(math.exp(rng.uniform(1, math.log(rays_lo), math.log(rays_hi))[0]))
    if total_rays is not None:
        scale = total_rays / sum(counts)
        counts = [c * scale for c in counts]
    counts = [max(1, int(round(c))) for c in counts]
    if total_rays is not None:
        counts[-1] = max(1, counts[-1] + total_rays - sum(counts))
    os_, ds_ = [], []
    begin = [0]
    for k, nr in enumerate(counts):
        o, d = cone_rays(nr, (seed << 16) + k, 6.0, -8.0, cone_deg)
        os_.append(o); ds_.append(d)
        begin.append(begin[-1] + nr)
    return scenes, np.ascontiguousarray(np.concatenate(os_)), np.ascontiguousarray(np.concatenate(ds_)), \
        np.array(begin, np.int64)


# --------------------------------------------------------------------------
# IFGF-RP patch files
# --------------------------------------------------------------------------
def write_patch_dir(sc: Scene, directory: str, prefix: str = "patch") -> List[str]:
    """Writes one reference-format .txt per patch, named prefix<i>.txt, i = 1..N.

    Field order follows SurfacePatch::load_patch_file (SurfacePatch.hpp:320-354):
    header, uNodes, vNodes, x, y, z, mask, dxdu..dzdv, dS, nuX..nuZ.  Values are
    written with repr(), which round-trips binary64 exactly.
    """
    os.makedirs(directory, exist_ok=True)
    paths = []
    for k in range(sc.n_patches):
        im, jm = int(sc.imax[k]), int(sc.jmax[k])
        F = sc.patch_fields(k)
        xu, yu, zu = F[3], F[4], F[5]
        xv, yv, zv = F[6], F[7], F[8]
        dS = np.sqrt((yu * zv - zu * yv) ** 2 + (zu * xv - xu * zv) ** 2 + (xu * yv - yu * xv) ** 2)
        rows = [f"{im} {jm} 1 {k + 1}",
                " ".join(repr(float(v)) for v in sc.unodes[sc.u_off[k]:sc.u_off[k + 1]]),
                " ".join(repr(float(v)) for v in sc.vnodes[sc.v_off[k]:sc.v_off[k + 1]])]
        for f in range(3):
            rows.append(" ".join(repr(float(v)) for v in F[f]))
        rows.append(" ".join("1" for _ in range(im * jm)))
        for f in range(3, 9):
            rows.append(" ".join(repr(float(v)) for v in F[f]))
        rows.append(" ".join(repr(float(v)) for v in dS))
        for f in range(9, 12):
            rows.append(" ".join(repr(float(v)) for v in F[f]))
        path = os.path.join(directory, f"{prefix}{k + 1}.txt")
        with open(path, "w") as fh:
            fh.write("\n".join(rows) + "\n")
        paths.append(path)
    return paths


def read_patch_file(path: str):
    """Parses one IFGF-RP patch file (SurfacePatch.hpp:313-371) -> (imax, jmax, unodes, vnodes, fields[12,n])."""
    with open(path) as fh:
        tok = fh.read().split()
    if len(tok) < 4:
        raise RuntimeError("SurfacePatch: bad header in patch file: " + path)
    im, jm, km = int(tok[0]), int(tok[1]), int(tok[2])
    n = im * jm * km
    pos = 4
    need = 4 + im + jm + 14 * n
    if len(tok) < need:
        raise RuntimeError("SurfacePatch: truncated patch file: " + path)

    def take(c):
        nonlocal pos
        out = np.array([float(t) for t in tok[pos:pos + c]])
        pos += c
        return out
    un = take(im); vn = take(jm)
    x, y, z = take(n), take(n), take(n)
    pos += n  # mask
    d = [take(n) for _ in range(6)]
    pos += n  # dS
    nu = [take(n) for _ in range(3)]
    rest = len(tok) - pos
    if rest not in (0, 4):
        raise RuntimeError("SurfacePatch: malformed sharp-edge data in patch file: " + path)
    return im, jm, un, vn, np.stack([x, y, z] + d + nu)
