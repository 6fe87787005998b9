"""CPU pins of the oracle's scene volumes and meshes (oracle/rk_oracle.c vol_sdf, mesh_sdf,
scene_sdf_all; geom/sdf_kernels.py:151-222, 346-380).

Volumes: trilinear interpolation reproduces any affine function exactly, with its gradient, and
the grid values at the grid points; outside the box the distance is the box distance plus the
padding.  Meshes: a closed box mesh's signed distance equals the analytic box SDF of the
primitive path everywhere, inside and out (the pseudonormal sign and the region tests), and the
distance to an irregular mesh equals a brute-force minimum over densely sampled triangle points
(to the sampling resolution).  The scene keeps the first strict minimum in the order spheres,
boxes, geometries.
"""
import numpy as np
import pytest

from oracle import oracle as O
from robokit_amd.geom import SdfVolume, TriMesh, box_mesh


def _affine_volume(a, b, lo, hi, h, pad=0.05):
    return SdfVolume.from_function(lambda p: p @ np.asarray(a, np.float64) + b, lo, hi, h, pad)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_volume_affine_exact_and_outside(dtype):
    rng = np.random.default_rng(0)
    a, b = np.array([0.3, -0.8, 0.52]), 0.1
    vol = _affine_volume(a, b, [-0.4, -0.3, 0.0], [0.5, 0.35, 0.6], 0.07)
    p = rng.uniform(vol.aabb_min, vol.aabb_max, (200, 3))
    d, n = O.scene_query(p, geoms=[vol], dtype=dtype)
    tol = 1e-7 if dtype == "f64" else 2e-6  # the grid values are float32
    np.testing.assert_allclose(d, p @ a + b, atol=tol)
    # differences of float32 grid values over a 7 cm voxel: ~1e-6 in the direction
    np.testing.assert_allclose(n, np.broadcast_to(a / np.linalg.norm(a), n.shape), atol=2e-5 if dtype == "f32" else 2e-6)
    gp = vol.grid_points().reshape(-1, 3)
    inside = np.all((gp >= vol.aabb_min) & (gp <= vol.aabb_max), 1)
    dg, _ = O.scene_query(gp[inside], geoms=[vol], dtype=dtype)
    np.testing.assert_allclose(dg, vol.values.reshape(-1)[inside], atol=tol)
    q = np.array([[1.5, 0.0, 0.3], [0.0, -1.3, -0.4]])
    dq, nq = O.scene_query(q, geoms=[vol], dtype=dtype)
    cl = np.clip(q, vol.aabb_min, vol.aabb_max)
    np.testing.assert_allclose(dq, np.linalg.norm(q - cl, axis=1) + vol.padding, rtol=1e-6)
    np.testing.assert_allclose(nq, (q - cl) / np.linalg.norm(q - cl, axis=1, keepdims=True), atol=1e-6)


def _box_sdf(p, c, h):
    q = np.abs(p - c) - h
    return np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_box_mesh_equals_the_box_sdf(dtype):
    rng = np.random.default_rng(1)
    c, h = np.array([0.2, -0.1, 0.4]), np.array([0.15, 0.3, 0.1])
    mesh = box_mesh(c, h)
    p = np.concatenate([rng.uniform(c - 2 * h, c + 2 * h, (300, 3)), rng.uniform(c - 0.9 * h, c + 0.9 * h, (100, 3))])
    d, n = O.scene_query(p, geoms=[mesh], dtype=dtype)
    ref = _box_sdf(p, c, h)
    np.testing.assert_allclose(d, ref, atol=5e-8 if dtype == "f64" else 2e-6)  # float32 vertices
    assert (d < 0).sum() >= 100  # the sign: every interior point is negative
    # the primitive path gives the same distances (box primitive = [centre, half extents])
    dp, _ = O.scene_query(p, boxes=np.concatenate([c, h])[None], dtype=dtype)
    np.testing.assert_allclose(d, dp, atol=5e-8 if dtype == "f64" else 2e-6)
    # off the edges and corners the mesh normal is the box face normal (the primitive's gradient);
    # it is the direction from the nearest point, so its rounding grows as 1 / distance: points
    # at least 1 cm from the surface
    q = np.abs(p - c) - h
    srt = np.sort(q, 1)
    clear = (srt[:, 2] - srt[:, 1] > 0.02) & (srt[:, 1] < 0) & (np.abs(d) > 0.01)
    _, npr = O.scene_query(p[clear], boxes=np.concatenate([c, h])[None], dtype=dtype)
    np.testing.assert_allclose(n[clear], npr, atol=2e-5 if dtype == "f32" else 1e-10)
    # everywhere the normal is the distance gradient: a central difference of the distance
    e = 1e-5
    far = np.abs(d) > 0.01
    if dtype == "f64":
        g = np.stack([(O.scene_query(p[far] + e * u, geoms=[mesh], dtype=dtype)[0]
                       - O.scene_query(p[far] - e * u, geoms=[mesh], dtype=dtype)[0]) / (2 * e) for u in np.eye(3)], 1)
        edge = np.abs(np.sort(q[far], 1)[:, 1] - np.sort(q[far], 1)[:, 2]) > 1e-4  # off the medial axis
        np.testing.assert_allclose(n[far][edge], g[edge], atol=1e-6)


def _icosphere(sub=1, r=0.2, c=(0.1, 0.0, 0.3), jitter=0.0, seed=0):
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7],
         [9, 8, 1]]
    v = [np.array(x, float) / np.linalg.norm(x) for x in v]
    for _ in range(sub):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]

        f = [g for a, b, cc in f for g in ([a, m(a, b), m(cc, a)], [b, m(b, cc), m(a, b)], [cc, m(cc, a), m(b, cc)],
                                            [m(a, b), m(b, cc), m(cc, a)])]
    v = np.asarray(v) * r * (1 + jitter * np.random.default_rng(seed).uniform(-1, 1, (len(v), 1))) + np.asarray(c)
    return TriMesh(v, np.asarray(f))


def _sampled_distance(mesh, p, k=120):
    """min over the triangles of the distance to a dense barycentric grid of each (k per side)"""
    V, F = mesh.vertices.astype(np.float64), mesh.faces
    u, w = np.meshgrid(np.linspace(0, 1, k), np.linspace(0, 1, k))
    keep = u + w <= 1
    u, w = u[keep], w[keep]
    pts = (V[F[:, 0]][:, None] * (1 - u - w)[None, :, None] + V[F[:, 1]][:, None] * u[None, :, None]
           + V[F[:, 2]][:, None] * w[None, :, None]).reshape(-1, 3)
    return np.array([np.sqrt(((pts - x) ** 2).sum(1).min()) for x in p])


def test_irregular_mesh_distance_and_sign():
    mesh = _icosphere(sub=1, jitter=0.15, seed=3)
    rng = np.random.default_rng(4)
    c = np.array([0.1, 0.0, 0.3])
    dirs = rng.normal(size=(60, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p = c + dirs * rng.uniform(0.05, 0.45, (60, 1))
    d, _ = O.scene_query(p, geoms=[mesh], dtype="f64")
    ref = _sampled_distance(mesh, p)
    edge = 0.42 * 0.2 * 1.15 / 119  # sampling resolution of the longest edge
    np.testing.assert_allclose(np.abs(d), ref, atol=edge)
    # the sign: the jittered sphere is star-shaped around c, so a point is inside iff it is
    # nearer c than the surface along its ray; the points near the shell are left out
    r = np.linalg.norm(p - c, axis=1)
    assert np.all(d[r < 0.13] < 0) and np.all(d[r > 0.28] > 0)


def test_scene_order_first_strict_minimum():
    """a sphere primitive and a mesh at the same distance: the primitive (earlier) decides the normal"""
    mesh = box_mesh([0.0, 0.0, 0.0], [0.125, 0.125, 0.125])  # binary fractions: exact distances
    p = np.array([[0.375, 0.0, 0.0]])
    d_m, n_m = O.scene_query(p, geoms=[mesh], dtype="f64")
    sph = np.array([[0.375, 0.375, 0.0, 0.125]])  # surface 0.25 from p, as the box's face
    d, n = O.scene_query(p, spheres=sph, geoms=[mesh], dtype="f64")
    assert d[0] == 0.25 and d_m[0] == 0.25
    np.testing.assert_allclose(n[0], [0, -1, 0])  # the sphere's gradient, not the mesh's face normal
    np.testing.assert_allclose(n_m[0], [1, 0, 0])
