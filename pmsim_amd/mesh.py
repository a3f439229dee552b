"""Mesh, cell kinds and boundary-feature search for the solid element loop.

Mirrors the parts of gemlab's ``Mesh``/``Features`` that pmsim's solid path uses
(call sites: ``tests/test_spo_754_footing.rs:19-35``, ``tests/test_von_mises_2x2_elements_2d.rs:77-83``):
reading the ``.msh`` text format (``data/spo/spo_754_footing.msh:1-10,454-460``), the
local node numbering of each cell kind, and the search for boundary edges/faces
lying on a coordinate plane with an extra point filter.

Kind codes are shared with the C ABI (``include/pmsim_amd.h``).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence

import numpy as np

# --- cell kinds (codes of the C ABI) ---------------------------------------------------------
TRI3, TRI6, TRI10, TRI15, QUA4, QUA8, QUA9, TET4, TET10, HEX8, HEX20 = range(11)
LIN2, LIN3, LIN4, LIN5 = 100, 101, 102, 103

KIND_NAMES = {
    "tri3": TRI3, "tri6": TRI6, "tri10": TRI10, "tri15": TRI15,
    "qua4": QUA4, "qua8": QUA8, "qua9": QUA9,
    "tet4": TET4, "tet10": TET10, "hex8": HEX8, "hex20": HEX20,
}
NNODE = {TRI3: 3, TRI6: 6, TRI10: 10, TRI15: 15, QUA4: 4, QUA8: 8, QUA9: 9, TET4: 4, TET10: 10, HEX8: 8, HEX20: 20}
NDIM_OF = {k: (2 if k <= QUA9 else 3) for k in NNODE}
# default number of integration points per kind (the rules used when ParamSolid.ngauss is None); Tri3 3,
# Qua8 9 and Hex8 8 are the counts pmsim writes in its own state files (data/results/artificial/*.json)
DEFAULT_NGAUSS = {TRI3: 3, TRI6: 3, TRI10: 6, TRI15: 12, QUA4: 4, QUA8: 9, QUA9: 9, TET4: 1, TET10: 4, HEX8: 8, HEX20: 27}

# local edges of 2D cells: [end0, end1, interior nodes ordered from end0 to end1]
EDGES_2D = {
    TRI3: [(0, 1), (1, 2), (2, 0)],
    TRI6: [(0, 1, 3), (1, 2, 4), (2, 0, 5)],
    TRI10: [(0, 1, 3, 6), (1, 2, 4, 7), (2, 0, 5, 8)],
    TRI15: [(0, 1, 6, 3, 7), (1, 2, 8, 4, 9), (2, 0, 10, 5, 11)],
    QUA4: [(0, 1), (1, 2), (2, 3), (3, 0)],
    QUA8: [(0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7)],
    QUA9: [(0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7)],
}
EDGE_KIND = {2: LIN2, 3: LIN3, 4: LIN4, 5: LIN5}

# local faces of 3D cells: corners (cyclic) followed by the mid-edge nodes between them
_HEX_FACES = [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 3, 2, 1), (4, 5, 6, 7)]
_HEX20_MID = {(0, 1): 8, (1, 2): 9, (2, 3): 10, (3, 0): 11, (4, 5): 12, (5, 6): 13, (6, 7): 14, (7, 4): 15,
              (0, 4): 16, (1, 5): 17, (2, 6): 18, (3, 7): 19}
_TET_FACES = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)]
_TET10_MID = {(0, 1): 4, (1, 2): 5, (0, 2): 6, (0, 3): 7, (1, 3): 8, (2, 3): 9}


def _mid(table, a, b):
    return table[(a, b)] if (a, b) in table else table[(b, a)]


def _faces(kind):
    if kind == HEX8:
        return [(f, QUA4) for f in _HEX_FACES]
    if kind == HEX20:
        return [(tuple(f) + tuple(_mid(_HEX20_MID, f[i], f[(i + 1) % 4]) for i in range(4)), QUA8) for f in _HEX_FACES]
    if kind == TET4:
        return [(f, TRI3) for f in _TET_FACES]
    if kind == TET10:
        return [(tuple(f) + tuple(_mid(_TET10_MID, f[i], f[(i + 1) % 3]) for i in range(3)), TRI6) for f in _TET_FACES]
    raise ValueError("not a 3D kind")


@dataclasses.dataclass
class Facet:
    """A boundary edge (2D) or face (3D): its points, facet kind and owning cell."""
    points: tuple
    kind: int
    cell: int


class Mesh:
    """Points and cells (gemlab ``Mesh`` subset): ``coords`` (npoint×ndim), per-cell kind/attribute/points."""

    def __init__(self, ndim: int, coords, cells: Sequence[tuple]):
        self.ndim = int(ndim)
        self.coords = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, ndim))
        kinds, attrs, ptr, pts = [], [], [0], []
        for (attr, kind, points) in cells:
            k = KIND_NAMES[kind] if isinstance(kind, str) else int(kind)
            if len(points) != NNODE[k]:
                raise ValueError("wrong number of points for cell kind")
            if NDIM_OF[k] != self.ndim:
                raise ValueError("cell kind does not match ndim")
            kinds.append(k)
            attrs.append(int(attr))
            pts.extend(int(p) for p in points)
            ptr.append(len(pts))
        self.kinds = np.asarray(kinds, dtype=np.int32)
        self.attrs = np.asarray(attrs, dtype=np.int32)
        self.cell_ptr = np.asarray(ptr, dtype=np.int64)
        self.cell_pts = np.asarray(pts, dtype=np.int32)

    @classmethod
    def from_arrays(cls, ndim, coords, kinds, attrs, cell_ptr, cell_pts):
        m = cls.__new__(cls)
        m.ndim = int(ndim)
        m.coords = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(-1, ndim))
        m.kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        m.attrs = np.ascontiguousarray(attrs, dtype=np.int32)
        m.cell_ptr = np.ascontiguousarray(cell_ptr, dtype=np.int64)
        m.cell_pts = np.ascontiguousarray(cell_pts, dtype=np.int32)
        return m

    @property
    def npoint(self):
        return self.coords.shape[0]

    @property
    def ncell(self):
        return self.kinds.shape[0]

    def cell_points(self, e):
        return self.cell_pts[self.cell_ptr[e]:self.cell_ptr[e + 1]]

    @classmethod
    def read(cls, path: str) -> "Mesh":
        """Reads the ``.msh`` text format: header ``ndim npoint ncell``, then points ``id marker x y {z}``,
        then cells ``id attribute kind points...``; ``#`` starts a comment line."""
        rows = []
        with open(path) as f:
            for line in f:
                s = line.strip()
                if not s or s.startswith("#"):
                    continue
                rows.append(s.split())
        ndim, npoint, ncell = (int(v) for v in rows[0])
        coords = np.zeros((npoint, ndim))
        for r in rows[1:1 + npoint]:
            coords[int(r[0])] = [float(v) for v in r[2:2 + ndim]]
        cells = [None] * ncell
        for r in rows[1 + npoint:1 + npoint + ncell]:
            cells[int(r[0])] = (int(r[1]), r[2], [int(v) for v in r[3:]])
        return cls(ndim, coords, cells)


class Features:
    """Boundary edges/faces search (gemlab ``Features::search_edges/search_faces`` semantics as used by pmsim:
    a boundary facet is kept when ALL its points lie on the plane and pass the filter)."""

    def __init__(self, mesh: Mesh, extract_all: bool = False):
        self.mesh = mesh
        count = {}
        first = {}
        for e in range(mesh.ncell):
            k = int(mesh.kinds[e])
            pts = mesh.cell_points(e)
            if mesh.ndim == 2:
                locs = [(loc, EDGE_KIND[len(loc)]) for loc in EDGES_2D[k]]
            else:
                locs = _faces(k)
            for loc, fk in locs:
                gp = tuple(int(pts[i]) for i in loc)
                n_corner = 2 if mesh.ndim == 2 else (3 if fk in (TRI3, TRI6) else 4)
                key = tuple(sorted(gp[:n_corner]))
                count[key] = count.get(key, 0) + 1
                if key not in first:
                    first[key] = Facet(gp, fk, e)
        self.boundary = [first[k] for k in first if count[k] == 1]

    def _search(self, at: Callable, flt: Callable):
        X = self.mesh.coords
        out = [f for f in self.boundary if all(at(X[p]) and flt(X[p]) for p in f.points)]
        if not out:
            raise ValueError("cannot find any feature with the given constraints")
        return out

    def search_edges(self, at, flt: Optional[Callable] = None) -> List[Facet]:
        if self.mesh.ndim != 2:
            raise ValueError("search_edges is implemented for 2D meshes")
        return self._search(at, flt or any_x)

    def search_faces(self, at, flt: Optional[Callable] = None) -> List[Facet]:
        if self.mesh.ndim != 3:
            raise ValueError("search_faces requires a 3D mesh")
        return self._search(at, flt or any_x)

    def get_edge(self, a: int, b: int) -> Facet:
        """The boundary edge whose end points are a and b (gemlab's Features::get_edge)."""
        for f in self.boundary:
            if sorted((int(f.points[0]), int(f.points[1]))) == sorted((int(a), int(b))):
                return f
        raise ValueError("cannot find edge with given key")

    def cells_via_edges(self, edges: Sequence[Facet]) -> List[int]:
        """Ids (ascending) of the cells that own the given boundary edges (gemlab's get_cells_via_2d_edges)."""
        return sorted({int(f.cell) for f in edges})

    def search_point_ids(self, at, flt: Optional[Callable] = None) -> List[int]:
        """Ids (ascending) of the boundary points that satisfy ``at`` and the filter."""
        X = self.mesh.coords
        flt = flt or any_x
        ids = sorted({p for f in self.boundary for p in f.points if at(X[p]) and flt(X[p])})
        if not ids:
            raise ValueError("cannot find any point with the given constraints")
        return ids


def any_x(_x) -> bool:
    return True


TOL_AT = 1e-8  # distance below which a point lies on a searched plane, circle or point


class At:
    """Geometric selectors for boundary searches (gemlab's ``At``): each returns a predicate on a
    point's coordinates.  ``At.X/Y/Z(v)``: coordinate plane; ``At.XY(x, y)`` / ``At.XYZ(x, y, z)``: one
    point; ``At.Circle(cx, cy, r)``: points at distance r from the centre (2D)."""

    @staticmethod
    def X(v):
        v = float(v)
        return lambda x: abs(x[0] - v) <= TOL_AT

    @staticmethod
    def Y(v):
        v = float(v)
        return lambda x: abs(x[1] - v) <= TOL_AT

    @staticmethod
    def Z(v):
        v = float(v)
        return lambda x: abs(x[2] - v) <= TOL_AT

    @staticmethod
    def XY(px, py):
        px, py = float(px), float(py)
        return lambda x: abs(x[0] - px) <= TOL_AT and abs(x[1] - py) <= TOL_AT

    @staticmethod
    def XYZ(px, py, pz):
        px, py, pz = float(px), float(py), float(pz)
        return lambda x: abs(x[0] - px) <= TOL_AT and abs(x[1] - py) <= TOL_AT and abs(x[2] - pz) <= TOL_AT

    @staticmethod
    def Circle(cx, cy, r):
        cx, cy, r = float(cx), float(cy), float(r)
        return lambda x: abs(math.hypot(x[0] - cx, x[1] - cy) - r) <= TOL_AT * max(1.0, r)


def edges_path(edges: Sequence[Facet]) -> List[int]:
    """Point ids along the path formed by 2D edges joined end to end (gemlab's Edges::any_path), each
    edge's interior points between its two ends; starts at an end of the path (a point used once)."""
    if not edges:
        return []
    ends = {}
    for i, f in enumerate(edges):
        for p in (f.points[0], f.points[1]):
            ends.setdefault(int(p), []).append(i)
    start = next((p for p, es in ends.items() if len(es) == 1), int(edges[0].points[0]))
    path, used, cur = [start], set(), start
    while True:
        nxt = [i for i in ends[cur] if i not in used]
        if not nxt:
            break
        i = nxt[0]
        used.add(i)
        f = edges[i]
        a, b = int(f.points[0]), int(f.points[1])
        inner = [int(p) for p in f.points[2:]]
        if cur == b:
            a, b, inner = b, a, inner[::-1]
        path.extend(inner + [b])
        cur = b
    return path


def points_of(facets: Sequence[Facet]) -> List[int]:
    s = set()
    for f in facets:
        s.update(f.points)
    return sorted(s)
