"""Structured and mixed mesh generators for the benchmark workloads and tests.

Node numbering follows a refined integer lattice (x fastest, then y, then z), which keeps the
DOFs of neighbouring elements close together in memory.  Jitter, when requested, moves the
corner lattice points by seeded uniform noise (numpy PCG64), never normal to the domain
boundary; higher-order nodes are placed by the parent cell's multilinear map so shared nodes
are computed once and are identical for all elements that use them.
"""
from __future__ import annotations

import numpy as np

from .mesh import HEX8, HEX20, NNODE, QUA4, QUA8, QUA9, TET4, TET10, TRI3, TRI6, TRI10, TRI15, Mesh

# reference node coordinates of each kind (gemlab GeoKind conventions)
REF_NODES = {
    TRI3: [(0, 0), (1, 0), (0, 1)],
    TRI6: [(0, 0), (1, 0), (0, 1), (.5, 0), (.5, .5), (0, .5)],
    TRI10: [(0, 0), (1, 0), (0, 1), (1 / 3, 0), (2 / 3, 1 / 3), (0, 2 / 3), (2 / 3, 0), (1 / 3, 2 / 3), (0, 1 / 3),
            (1 / 3, 1 / 3)],
    TRI15: [(0, 0), (1, 0), (0, 1), (.5, 0), (.5, .5), (0, .5), (.25, 0), (.75, 0), (.75, .25), (.25, .75),
            (0, .75), (0, .25), (.25, .25), (.5, .25), (.25, .5)],
    QUA4: [(-1, -1), (1, -1), (1, 1), (-1, 1)],
    QUA8: [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)],
    QUA9: [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)],
    TET4: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    TET10: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, 0, 0), (.5, .5, 0), (0, .5, 0), (0, 0, .5),
            (.5, 0, .5), (0, .5, .5)],
    HEX8: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],
    HEX20: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1),
            (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1), (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1),
            (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)],
}


def _jitter_corners(shape_pts, h, jitter, rng, ndim):
    """Uniform ±jitter*h per axis on the corner lattice, zero normal to the boundary."""
    d = rng.uniform(-jitter, jitter, size=tuple(shape_pts) + (ndim,)) * np.asarray(h)
    for ax in range(ndim):
        sl0 = [slice(None)] * ndim
        sl1 = [slice(None)] * ndim
        sl0[ax] = 0
        sl1[ax] = -1
        d[tuple(sl0) + (ax,)] = 0.0
        d[tuple(sl1) + (ax,)] = 0.0
    return d


def hex20_block(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, seed=None, jitter=0.0, attribute=1):
    """nx*ny*nz Hex20 over [0,lx]x[0,ly]x[0,lz]; corner jitter ±jitter*h (seeded), mid-edge nodes at the
    midpoints of the jittered corners."""
    h = (lx / nx, ly / ny, lz / nz)
    # doubled lattice (2nx+1, 2ny+1, 2nz+1); a Hex20 node has at most one odd index
    I, J, K = np.meshgrid(np.arange(2 * nx + 1), np.arange(2 * ny + 1), np.arange(2 * nz + 1), indexing="ij")
    odd = (I % 2) + (J % 2) + (K % 2)
    is_node = odd <= 1
    # numbering: x fastest, then y, then z
    order = np.transpose(is_node, (2, 1, 0)).ravel()
    ids_flat = np.full(order.shape, -1, dtype=np.int64)
    ids_flat[order] = np.arange(int(order.sum()))
    ids = np.transpose(ids_flat.reshape(2 * nz + 1, 2 * ny + 1, 2 * nx + 1), (2, 1, 0))
    npoint = int(order.sum())
    # corner positions
    cx, cy, cz = np.meshgrid(np.arange(nx + 1) * h[0], np.arange(ny + 1) * h[1], np.arange(nz + 1) * h[2], indexing="ij")
    corners = np.stack([cx, cy, cz], axis=-1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        corners = corners + _jitter_corners((nx + 1, ny + 1, nz + 1), h, jitter, rng, 3)
    X = np.zeros((npoint, 3))
    # corners
    X[ids[0::2, 0::2, 0::2].ravel()] = corners.reshape(-1, 3)
    # x-edges (odd I)
    X[ids[1::2, 0::2, 0::2].ravel()] = (0.5 * (corners[:-1, :, :] + corners[1:, :, :])).reshape(-1, 3)
    X[ids[0::2, 1::2, 0::2].ravel()] = (0.5 * (corners[:, :-1, :] + corners[:, 1:, :])).reshape(-1, 3)
    X[ids[0::2, 0::2, 1::2].ravel()] = (0.5 * (corners[:, :, :-1] + corners[:, :, 1:])).reshape(-1, 3)
    offs = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2),
            (1, 0, 0), (2, 1, 0), (1, 2, 0), (0, 1, 0), (1, 0, 2), (2, 1, 2), (1, 2, 2), (0, 1, 2),
            (0, 0, 1), (2, 0, 1), (2, 2, 1), (0, 2, 1)]
    # elements ordered x fastest
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    conn = np.stack([ids[2 * ei + a, 2 * ej + b, 2 * ek + c] for (a, b, c) in offs], axis=1).astype(np.int32)
    ncell = conn.shape[0]
    return Mesh.from_arrays(3, X, np.full(ncell, HEX20, np.int32), np.full(ncell, attribute, np.int32),
                            np.arange(ncell + 1, dtype=np.int64) * 20, conn.ravel())


def qua8_block(nx, ny, lx, ly, seed=None, jitter=0.0, attribute=1, x0=0.0, y0=0.0):
    """nx*ny Qua8 over [x0,x0+lx]x[y0,y0+ly] (x fastest numbering of the doubled lattice)."""
    h = (lx / nx, ly / ny)
    I, J = np.meshgrid(np.arange(2 * nx + 1), np.arange(2 * ny + 1), indexing="ij")
    is_node = ((I % 2) + (J % 2)) <= 1
    order = is_node.T.ravel()
    ids_flat = np.full(order.shape, -1, dtype=np.int64)
    ids_flat[order] = np.arange(int(order.sum()))
    ids = ids_flat.reshape(2 * ny + 1, 2 * nx + 1).T
    npoint = int(order.sum())
    cx, cy = np.meshgrid(x0 + np.arange(nx + 1) * h[0], y0 + np.arange(ny + 1) * h[1], indexing="ij")
    corners = np.stack([cx, cy], axis=-1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        corners = corners + _jitter_corners((nx + 1, ny + 1), h, jitter, rng, 2)
    X = np.zeros((npoint, 2))
    X[ids[0::2, 0::2].ravel()] = corners.reshape(-1, 2)
    X[ids[1::2, 0::2].ravel()] = (0.5 * (corners[:-1, :] + corners[1:, :])).reshape(-1, 2)
    X[ids[0::2, 1::2].ravel()] = (0.5 * (corners[:, :-1] + corners[:, 1:])).reshape(-1, 2)
    offs = [(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1)]
    ej, ei = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej = ei.ravel(), ej.ravel()
    conn = np.stack([ids[2 * ei + a, 2 * ej + b] for (a, b) in offs], axis=1).astype(np.int32)
    ncell = conn.shape[0]
    return Mesh.from_arrays(2, X, np.full(ncell, QUA8, np.int32), np.full(ncell, attribute, np.int32),
                            np.arange(ncell + 1, dtype=np.int64) * 8, conn.ravel())


# ----------------------------------------------------------------------------------------------
# generic lattice generator (any kind; triangles split quads, tetrahedra split hexahedra)
# ----------------------------------------------------------------------------------------------
_LAT = 12  # lattice subdivisions per parent cell (multiple of 1, 2, 3 and 4)

_TRI_SPLIT = [((0, 0), (1, 0), (1, 1)), ((0, 0), (1, 1), (0, 1))]
_TET_SPLIT = [((0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)), ((0, 0, 0), (1, 1, 0), (0, 1, 0), (1, 1, 1)),
              ((0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1)), ((0, 0, 0), (0, 1, 1), (0, 0, 1), (1, 1, 1)),
              ((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)), ((0, 0, 0), (1, 0, 1), (1, 0, 0), (1, 1, 1))]


def _sub_elements(kind):
    """Local lattice coordinates (fractions of the parent cell) of each sub-element's nodes."""
    ref = np.asarray(REF_NODES[kind], dtype=np.float64)
    if kind in (QUA4, QUA8, QUA9, HEX8, HEX20):
        return [(ref + 1.0) / 2.0]
    out = []
    split = _TRI_SPLIT if kind in (TRI3, TRI6, TRI10, TRI15) else _TET_SPLIT
    for verts in split:
        v = np.asarray(verts, dtype=np.float64)
        bary = np.concatenate([1.0 - ref.sum(axis=1, keepdims=True), ref], axis=1)
        out.append(bary @ v)
    return out


def block(kind, n, lengths, seed=None, jitter=0.0, attribute=1, kinds_per_cell=None):
    """Structured block of `kind` (or per-parent kinds from `kinds_per_cell`, shape n) over a box.

    Every node sits on a lattice of _LAT subdivisions per parent cell; nodes of different elements
    that fall on the same lattice point are merged.  Returns a Mesh."""
    n = tuple(int(v) for v in n)
    ndim = len(n)
    h = np.asarray(lengths, dtype=np.float64) / np.asarray(n)
    if kinds_per_cell is None:
        kinds_per_cell = np.full(n, kind, dtype=np.int32)
    rng = np.random.default_rng(seed)
    corners = np.stack(np.meshgrid(*[np.arange(k + 1) * hh for k, hh in zip(n, h)], indexing="ij"), axis=-1)
    if jitter > 0:
        corners = corners + _jitter_corners(tuple(k + 1 for k in n), h, jitter, rng, ndim)
    keys_all, conn_all, kinds_all = [], [], []
    parent_idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), axis=-1).reshape(-1, ndim)
    if ndim == 2:
        parent_idx = parent_idx[np.lexsort((parent_idx[:, 0], parent_idx[:, 1]))]
    else:
        parent_idx = parent_idx[np.lexsort((parent_idx[:, 0], parent_idx[:, 1], parent_idx[:, 2]))]
    subs_cache = {}
    for p in parent_idx:
        k = int(kinds_per_cell[tuple(p)])
        if k not in subs_cache:
            subs_cache[k] = _sub_elements(k)
        for frac in subs_cache[k]:
            lat = np.rint((p[None, :] + frac) * _LAT).astype(np.int64)
            keys_all.append(lat)
            kinds_all.append(k)
    lat_all = np.concatenate(keys_all, axis=0)
    mult = np.asarray([1] + list(np.cumprod([_LAT * k + 1 for k in n])[:-1]), dtype=np.int64)
    key = (lat_all * mult).sum(axis=1)
    uniq, inv = np.unique(key, return_inverse=True)
    # positions: multilinear map of the parent cell at the lattice point
    lat_u = np.zeros((uniq.shape[0], ndim), dtype=np.int64)
    lat_u[inv] = lat_all
    par = np.minimum(lat_u // _LAT, np.asarray(n) - 1)
    t = (lat_u - par * _LAT) / _LAT
    X = np.zeros((uniq.shape[0], ndim))
    for corner in np.ndindex(*([2] * ndim)):
        c = np.asarray(corner)
        w = np.prod(np.where(c[None, :] == 1, t, 1.0 - t), axis=1)
        X += w[:, None] * corners[tuple((par + c).T)]
    ptr = [0]
    for lat in keys_all:
        ptr.append(ptr[-1] + lat.shape[0])
    conn = inv.astype(np.int32)
    kinds_arr = np.asarray(kinds_all, dtype=np.int32)
    ncell = kinds_arr.shape[0]
    m = Mesh.from_arrays(ndim, X, kinds_arr, np.full(ncell, attribute, np.int32), np.asarray(ptr, np.int64), conn)
    _fix_orientation(m)
    return m


def _fix_orientation(m: Mesh):
    """Flips simplices with negative orientation (swap local nodes 1 and 2 and matching higher nodes)."""
    for e in range(m.ncell):
        k = int(m.kinds[e])
        if k not in (TRI3, TRI6, TRI10, TRI15, TET4, TET10):
            continue
        pts = m.cell_pts[m.cell_ptr[e]:m.cell_ptr[e + 1]]
        X = m.coords[pts[:m.ndim + 1]]
        d = X[1:] - X[0]
        if np.linalg.det(d) < 0:
            ref = np.asarray(REF_NODES[k], dtype=np.float64)
            # mirror r <-> s (swap local axes 0 and 1): node i maps to the node at the swapped position
            sw = ref.copy()
            sw[:, [0, 1]] = ref[:, [1, 0]]
            perm = [int(np.argmin(np.abs(ref - s).sum(axis=1))) for s in sw]
            pts[:] = pts[perm]


def mixed_2d(n_elements=200_000, lx=100.0, ly=100.0, seed=20261203, jitter=0.1):
    """Unstructured-type 2D mix: 30% Qua8, 25% Qua4, 20% Tri6, 15% Qua9, 10% Tri15 (by element count).

    Each parent cell of a structured grid carries one quadrilateral or two triangles of a seeded
    random kind.  Nodes shared by neighbours are merged on the common lattice; a mid-side node
    next to an element that has no node there is used by one element only (non-conforming but
    well-posed)."""
    frac = {QUA8: 0.30, QUA4: 0.25, TRI6: 0.20, QUA9: 0.15, TRI15: 0.10}
    parents_per_elem = {QUA8: 1.0, QUA4: 1.0, QUA9: 1.0, TRI6: 0.5, TRI15: 0.5}
    p_parent = {k: frac[k] * parents_per_elem[k] for k in frac}
    tot = sum(p_parent.values())
    n_par = n_elements * tot
    nx = int(round(np.sqrt(n_par * lx / ly)))
    ny = int(round(n_par / nx))
    rng = np.random.default_rng(seed)
    ks = np.asarray(list(p_parent.keys()), dtype=np.int32)
    pr = np.asarray([p_parent[k] / tot for k in ks])
    kinds = rng.choice(ks, size=(nx, ny), p=pr)
    return block(None, (nx, ny), (lx, ly), seed=seed + 1, jitter=jitter, kinds_per_cell=kinds)
