"""Benchmark problem builders (host setup).

* ``cantilever``: x=0 face clamped, unit downward (-z) load split equally
  over the nodes of the bottom edge (x = nx, z = 0); hex8 grid or jittered
  tet4 mesh.
* ``multiload``: x=0 face clamped, ``n_loads`` unit point loads at distinct
  random nodes of the x = nx face with directions uniform on the sphere;
  ``default_rng(seed)`` draws the nodes first, then an (n_loads, 3) normal
  block that is normalised row-wise.
* ``heat_patch``: n^3 cube, uniform volumetric source, T = 0 on a
  ``patch x patch`` node patch centred on the z = 0 face.
"""
from __future__ import annotations

import numpy as np

from .meshes import MeshHex, jittered_tet, structured_hex
from .task import LinearElasticity, LinearHeatConduction


def cantilever(nx, ny, nz, kind="hex8", jitter=0.0, seed=0, E=1.0, nu=0.3):
    mesh = structured_hex(nx, ny, nz) if kind == "hex8" else jittered_tet(nx, ny, nz, jitter, seed)
    i, j, k = MeshHex.node_ijk(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    load = np.nonzero((i == nx) & (k == 0))[0]
    f = np.zeros(3 * mesh.nvertices)
    f[3 * load + 2] = -1.0 / load.size
    return LinearElasticity.from_nodes(mesh, clamp, "all", [f], load, None, E, nu)


def multiload(nx, ny, nz, n_loads=8, seed=0, E=1.0, nu=0.3):
    mesh = structured_hex(nx, ny, nz)
    i, j, k = MeshHex.node_ijk(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    face = np.nonzero(i == nx)[0]
    rng = np.random.default_rng(seed)
    nodes = rng.choice(face, size=n_loads, replace=False)
    g = rng.standard_normal((n_loads, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    forces = []
    for a in range(n_loads):
        f = np.zeros(3 * mesh.nvertices)
        f[3 * nodes[a]:3 * nodes[a] + 3] = g[a]
        forces.append(f)
    return LinearElasticity.from_nodes(mesh, clamp, "all", forces, np.sort(nodes), None, E, nu)


def heat_patch(n, patch=16, source=1e-2, k=1.0):
    mesh = structured_hex(n, n, n)
    i, j, kk = MeshHex.node_ijk(n, n, n)
    lo = (n + 1 - patch) // 2
    sel = (kk == 0) & (i >= lo) & (i < lo + patch) & (j >= lo) & (j < lo + patch)
    nodes = np.nonzero(sel)[0]
    emit = np.zeros(mesh.nvertices)
    w = source * mesh.h[0] * mesh.h[1] * mesh.h[2] / 8.0
    for a in range(8):
        np.add.at(emit, mesh.t[a], w)
    return LinearHeatConduction(mesh, nodes, [0.0], emit, None, k)
