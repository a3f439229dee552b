"""Post-processing of simulation results (pmsim's ``PostProc``, src/fem/post_processing.rs).

Reads the files a run wrote (summary, mesh, base and state files; ``file_io``) and gives stresses at
the integration points and at the nodes:

* ``gauss_coords`` / ``gauss_stress``: IP coordinates x = sum_m N_m(xi_p) X_m and stress tensor
  components [sxx, syy, szz, sxy(, syz, szx)] (Mandel shear / sqrt 2) per IP
  (post_processing.rs:147-230);
* ``nodal_stress``: the IP values extrapolated to the cell's nodes with the extrapolation matrix E
  (``pm_extrapolation_matrix``; post_processing.rs:455-520);
* ``gauss_stresses`` / ``nodal_stresses``: a patch of cells, filtered by coordinates, nodal values
  averaged over the cells that share a node, sorted by y then x (2D) or z, y, x (3D) within a
  tolerance (post_processing.rs:286-453, 546-653; util/tensor_components_map.rs, util/spatial_tensor.rs);
* ``values_along_x`` / ``values_along_edges``: a DOF along a line of boundary points or a path of
  edges (post_processing.rs:729-811).

Strains are not recorded by this library's solver, so the strain forms are absent.
``footing_load_settlement`` restates examples/spo_754_footing_analysis.rs: the normalised settlement
of the footing's centre and the normalised footing pressure, -2 int(sigma_yy dx) / c over the footing
nodes' extrapolated stresses (trapezoidal rule).
"""
import ctypes as C
import dataclasses
import functools
import json
import math
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import file_io as fio
from . import mesh as M
from .solver import PmError, lib

SQ2 = math.sqrt(2.0)
TOL_COMPARE_POINTS = 1e-5  # relative to the mesh's extent: coordinates closer than this sort as equal


def gauss_rule(kind: int, ngauss: Optional[int] = None):
    """Natural coordinates (nip x 3) and weights of a kind's integration rule."""
    L = lib()
    n = C.c_int32(0)
    if L.pm_gauss_rule(int(kind), int(ngauss or 0), C.byref(n), None, None):
        raise PmError(L.pm_global_last_error().decode())
    xi = np.zeros(3 * n.value)
    w = np.zeros(n.value)
    L.pm_gauss_rule(int(kind), int(ngauss or 0), C.byref(n), xi.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    return xi.reshape(-1, 3), w


def shape_values(kind: int, xi: np.ndarray) -> np.ndarray:
    """Shape functions (npoint x nnode) at natural coordinates xi (npoint x 3)."""
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
    N = np.zeros(len(xi) * M.NNODE[kind])
    if lib().pm_shape_values(int(kind), len(xi), xi.ravel(), N):
        raise PmError(lib().pm_global_last_error().decode())
    return N.reshape(len(xi), -1)


@functools.lru_cache(maxsize=None)
def _extrap(kind: int, ngauss: int):
    L = lib()
    nip = len(gauss_rule(kind, ngauss)[1])
    E = np.zeros(M.NNODE[kind] * nip)
    if L.pm_extrapolation_matrix(int(kind), int(ngauss), E):
        raise PmError(L.pm_global_last_error().decode())
    E = E.reshape(M.NNODE[kind], nip)
    E.setflags(write=False)
    return E


def extrapolation_matrix(kind: int, ngauss: Optional[int] = None) -> np.ndarray:
    """E (nnode x nip): nodal values = E @ IP values."""
    return _extrap(int(kind), int(ngauss or 0))


@dataclasses.dataclass
class SpatialTensor:
    """Tensor components at a sorted set of locations (src/util/spatial_tensor.rs)."""
    k2id: List[int]
    xx: np.ndarray
    yy: np.ndarray
    zz: np.ndarray
    txx: np.ndarray
    tyy: np.ndarray
    tzz: np.ndarray
    txy: np.ndarray
    tyz: np.ndarray
    tzx: np.ndarray

    @property
    def id2k(self) -> Dict[int, int]:
        return {i: k for k, i in enumerate(self.k2id)}


def _sorted_order(X: np.ndarray, tol: np.ndarray) -> List[int]:
    """Indices of the rows of X sorted by the last coordinate, then the previous ones, each compared
    within its tolerance (russell_lab's argsort2/argsort3 as called at post_processing.rs:390-408)."""
    nd = X.shape[1]

    def cmp(a, b):
        for d in range(nd - 1, -1, -1):
            if abs(X[a, d] - X[b, d]) > tol[d]:
                return -1 if X[a, d] < X[b, d] else 1
        return 0

    return sorted(range(X.shape[0]), key=functools.cmp_to_key(cmp))


class PostProc:
    """Results of one run, read from ``{dir}/{fn_stem}-*`` files (PostProc::new, post_processing.rs:57-83)."""

    def __init__(self, dir: str, fn_stem: str):
        io = fio.FileIo(dir, fn_stem)
        s = fio.FileIo.read_summary(io.path_summary())
        io.active, io.counter, io.indices, io.times = s["active"], s["counter"], s["indices"], s["times"]
        self.file_io = io
        self.mesh = M.Mesh.read(io.path_mesh())
        with open(io.path_base()) as f:
            base = json.load(f)
        self.ngauss = {int(a): (p["Solid"]["ngauss"] or 0) for a, p in base["amap"]["all"].items()}
        lo, hi = self.mesh.coords.min(axis=0), self.mesh.coords.max(axis=0)
        self.tol = TOL_COMPARE_POINTS * (hi - lo)

    # ---- states -------------------------------------------------------------------------------
    def n_state(self) -> int:
        return len(self.file_io.indices)

    def read_state(self, index: int) -> dict:
        st = fio.read_state(self.file_io.path_state(index))
        # IP row offsets per cell
        st["ip_ptr"] = np.concatenate([[0], np.cumsum(st["ngauss"])]).astype(np.int64)
        return st

    def eq(self, point: int, dof: int) -> int:
        return self.mesh.ndim * int(point) + int(dof)

    # ---- integration points -------------------------------------------------------------------
    def _ng(self, cell: int) -> int:
        """Integration points of a cell: the attribute's ngauss, else the kind's default rule."""
        return self.ngauss.get(int(self.mesh.attrs[cell]), 0) or M.DEFAULT_NGAUSS[int(self.mesh.kinds[cell])]

    def gauss_coords(self, cell: int) -> np.ndarray:
        kind = int(self.mesh.kinds[cell])
        xi, _ = gauss_rule(kind, self._ng(cell))
        X = self.mesh.coords[self.mesh.cell_points(cell)]
        return shape_values(kind, xi) @ X

    def gauss_stress(self, state: dict, cell: int) -> np.ndarray:
        nd = self.mesh.ndim
        ns = 2 * nd
        a, b = state["ip_ptr"][cell], state["ip_ptr"][cell + 1]
        t = np.array(state["ip"][a:b, :ns], dtype=float)
        t[:, 3:] /= SQ2
        return t

    def _tensors(self, X: List[np.ndarray], T: List[np.ndarray], ids: Optional[List[int]]) -> SpatialTensor:
        """Sorted results; ids are the point ids (nodes) or, when None, the positions after sorting
        (integration points, post_processing.rs:412-418)."""
        nd = self.mesh.ndim
        X = np.asarray(X).reshape(-1, nd)
        T = np.asarray(T).reshape(-1, 2 * nd)
        order = _sorted_order(X, self.tol)
        X, T = X[order], T[order]
        z = X[:, 2] if nd == 3 else np.zeros(0)
        e = np.zeros(0)
        k2id = list(range(len(order))) if ids is None else [ids[k] for k in order]
        return SpatialTensor(k2id, X[:, 0], X[:, 1], z, T[:, 0], T[:, 1], T[:, 2], T[:, 3],
                             T[:, 4] if nd == 3 else e, T[:, 5] if nd == 3 else e)

    def gauss_stresses(self, state: dict, cells: Sequence[int], filter: Callable = lambda x, y, z: True) -> SpatialTensor:
        nd = self.mesh.ndim
        X, T = [], []
        for c in cells:
            xs, ts = self.gauss_coords(c), self.gauss_stress(state, c)
            for p in range(len(xs)):
                x = xs[p]
                if filter(x[0], x[1], x[2] if nd == 3 else 0.0):
                    X.append(x); T.append(ts[p])
        return self._tensors(X, T, None)

    # ---- nodes --------------------------------------------------------------------------------
    def nodal_stress(self, state: dict, cell: int) -> np.ndarray:
        E = extrapolation_matrix(int(self.mesh.kinds[cell]), self._ng(cell))
        return E @ self.gauss_stress(state, cell)

    def nodal_stresses(self, state: dict, cells: Sequence[int], filter: Callable = lambda x, y, z: True) -> SpatialTensor:
        nd = self.mesh.ndim
        acc: Dict[int, np.ndarray] = {}
        cnt: Dict[int, int] = {}
        for c in cells:
            tn = self.nodal_stress(state, c)
            for m, p in enumerate(self.mesh.cell_points(c)):
                p = int(p)
                acc[p] = acc[p] + tn[m] if p in acc else tn[m].copy()
                cnt[p] = cnt.get(p, 0) + 1
        X, T, ids = [], [], []
        for p in acc:
            x = self.mesh.coords[p]
            if filter(x[0], x[1], x[2] if nd == 3 else 0.0):
                X.append(x); T.append(acc[p] / cnt[p]); ids.append(p)
        return self._tensors(X, T, ids)

    # ---- primary values along lines -----------------------------------------------------------
    def values_along_x(self, features: M.Features, state: dict, dof: int, y: float, filter: Callable = M.any_x):
        """(ids, x, values) of a DOF at the boundary points with coordinate y, sorted by x."""
        ids = features.search_point_ids(M.At.Y(y), filter)
        ids = sorted(ids, key=lambda i: self.mesh.coords[i][0])
        xs = [float(self.mesh.coords[i][0]) for i in ids]
        return ids, xs, [float(state["u"][self.eq(i, dof)]) for i in ids]

    def values_along_edges(self, state: dict, edges: Sequence[M.Facet], dof: int):
        """(ids, coords, values) of a DOF along the path formed by the edges, from the end with the
        smaller y (then smaller x)."""
        ids = M.edges_path(edges)
        if len(ids) < 2:
            raise ValueError("not enough points along the path of edges")
        a, b = self.mesh.coords[ids[0]], self.mesh.coords[ids[-1]]
        if b[1] < a[1] or (abs(b[1] - a[1]) < TOL_COMPARE_POINTS and b[0] < a[0]):
            ids = ids[::-1]
        return ids, [self.mesh.coords[i].tolist() for i in ids], [float(state["u"][self.eq(i, dof)]) for i in ids]


def footing_load_settlement(post: PostProc, footing_edges: Sequence[M.Facet], corner: int, width: float,
                            cohesion: float):
    """examples/spo_754_footing_analysis.rs: for every state, the settlement -u_y(corner) / width and the
    pressure -2 int(sigma_yy dx) / cohesion over the footing's top nodes (nodal stresses extrapolated
    in the cells below the footing, trapezoidal rule in x)."""
    cells = sorted({f.cell for f in footing_edges})
    ytop = float(post.mesh.coords[:, 1].max())
    settlement, pressure = [], []
    for index in range(post.n_state()):
        st = post.read_state(index)
        settlement.append(-st["u"][post.eq(corner, 1)] / width)
        res = post.nodal_stresses(st, cells, lambda x, y, z: y == ytop)
        area = float(np.sum(np.diff(res.xx) * (res.tyy[1:] + res.tyy[:-1]) / 2.0))
        pressure.append(-2.0 * area / cohesion)
    return np.asarray(settlement), np.asarray(pressure)
