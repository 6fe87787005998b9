"""Scene geometry beyond primitives: signed-distance volumes and triangle meshes.

  SdfVolume        geom/sdf_volume.py:25-155 (a dense grid of signed distances, its box and
                   padding; from_mesh bakes a mesh on the GPU)
  TriMesh          the triangle meshes a MeshGeom holds (geom/geoms.py:502-600)
  box_mesh         a closed box as 12 outward-facing triangles

The device evaluates both (include/rkhip.h rk_geom_upload_volume / rk_geom_upload_mesh): a volume
by trilinear interpolation inside its box and box distance + padding outside; a mesh exactly, by
the nearest triangle, signed by the angle-weighted pseudonormal of the nearest feature (negative
inside a closed, outward-oriented mesh).  Scenes (lm.PrimitiveScene) list them after their
obstacle spheres and boxes; the collision terms and the trajectory kernel query them there.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(eq=False)
class TriMesh:
    """vertices [nv, 3] float32, faces [nf, 3] int32 (counter-clockwise seen from outside)"""

    vertices: np.ndarray
    faces: np.ndarray

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, np.float32).reshape(-1, 3)
        self.faces = np.ascontiguousarray(self.faces, np.int32).reshape(-1, 3)
        if len(self.vertices) < 3 or len(self.faces) < 1:
            raise ValueError("a mesh needs at least 3 vertices and 1 face")
        if self.faces.min() < 0 or self.faces.max() >= len(self.vertices):
            raise ValueError("face vertex index out of range")

    @property
    def bounds(self):
        v = self.vertices[np.unique(self.faces)]
        return v.min(0), v.max(0)


def box_mesh(center, half_extents) -> TriMesh:
    """the closed box centre +- half_extents as 12 triangles, outward normals"""
    c = np.asarray(center, np.float64).reshape(3)
    h = np.asarray(half_extents, np.float64).reshape(3)
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    v = c + corners * h
    # corner index = 4 (x > 0) + 2 (y > 0) + (z > 0); two triangles per face, counter-clockwise
    # seen from outside
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    faces = []
    for a, b, cc, d in quads:
        faces += [(a, b, cc), (a, cc, d)]
    return TriMesh(v.astype(np.float32), np.asarray(faces, np.int32))


@dataclass(eq=False)
class SdfVolume:
    """Signed distances ``values`` [nz, ny, nx] at origin + (i, j, k) * voxel_size; sampled inside
    [aabb_min, aabb_max], box distance + padding outside (geom/sdf_volume.py:25-40)."""

    values: np.ndarray
    origin: np.ndarray
    voxel_size: float
    aabb_min: np.ndarray
    aabb_max: np.ndarray
    padding: float = 0.0

    def __post_init__(self):
        self.values = np.ascontiguousarray(self.values, np.float32)
        if self.values.ndim != 3 or min(self.values.shape) < 2:
            raise ValueError("values must be [nz, ny, nx] with at least 2 points per axis")
        self.origin = np.asarray(self.origin, np.float32).reshape(3)
        self.aabb_min = np.asarray(self.aabb_min, np.float32).reshape(3)
        self.aabb_max = np.asarray(self.aabb_max, np.float32).reshape(3)
        if not self.voxel_size > 0:
            raise ValueError("voxel_size must be positive")

    @property
    def shape_xyz(self):
        nz, ny, nx = self.values.shape
        return nx, ny, nz

    def grid_points(self) -> np.ndarray:
        """world positions of the grid points [nz, ny, nx, 3]"""
        nx, ny, nz = self.shape_xyz
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        idx = np.stack([i, j, k], -1).astype(np.float64)
        return (self.origin.astype(np.float64) + idx * float(self.voxel_size)).astype(np.float32)

    @classmethod
    def from_function(cls, fn, aabb_min, aabb_max, voxel_size: float, padding: float = 0.0) -> "SdfVolume":
        """a volume over the box [aabb_min, aabb_max] (grid points from aabb_min, enough to cover
        it), values = fn(points [..., 3]) evaluated on the host"""
        lo = np.asarray(aabb_min, np.float32).reshape(3)
        hi = np.asarray(aabb_max, np.float32).reshape(3)
        n = (np.ceil((hi - lo) / voxel_size).astype(int) + 1).clip(2)
        vol = cls(np.zeros((n[2], n[1], n[0]), np.float32), lo, voxel_size, lo, hi, padding)
        vol.values = np.ascontiguousarray(np.asarray(fn(vol.grid_points()), np.float32))
        return vol

    @classmethod
    def from_mesh(cls, mesh: TriMesh, voxel_size: float, padding: float = 0.1, ctx=None) -> "SdfVolume":
        """Bake ``mesh`` (SdfVolume.from_mesh, geom/sdf_volume.py:80-155): the grid spans the mesh's
        bounds widened by ``padding`` (ceil(extent / voxel_size) + 1 points per axis) and holds the
        mesh's exact signed distance at each grid point, evaluated on the GPU (rk_scene_query)."""
        from . import native as N

        lo, hi = mesh.bounds
        lo = (lo - np.float32(padding)).astype(np.float32)
        hi = (hi + np.float32(padding)).astype(np.float32)
        n = np.ceil((hi - lo) / np.float32(voxel_size)).astype(int) + 1
        vol = cls(np.zeros((n[2], n[1], n[0]), np.float32), lo, voxel_size, lo, hi, padding)
        ctx = ctx or N.Context(0)
        d, _ = ctx.scene_query(vol.grid_points().reshape(-1, 3), geoms=[mesh])
        vol.values = np.ascontiguousarray(d.reshape(n[2], n[1], n[0]))
        return vol


def geom_kind(g) -> int:
    """0 for a volume, 1 for a mesh"""
    if isinstance(g, SdfVolume):
        return 0
    if isinstance(g, TriMesh):
        return 1
    raise TypeError(f"scene geometries are SdfVolume or TriMesh, got {type(g).__name__}")


def as_geom(g):
    """a scene geometry from a TriMesh / SdfVolume, or a (vertices, faces) pair"""
    if isinstance(g, (TriMesh, SdfVolume)):
        return g
    if isinstance(g, tuple) and len(g) == 2:
        return TriMesh(*g)
    raise TypeError(f"not a scene geometry: {type(g).__name__}")
