"""pmsim's on-disk formats for the solid path, so results can be checked by pmsim's own comparator and
pmsim's state files can be read back.

* State files ``{dir}/{stem}-{index:020}.json``: the serde layout of ``FemState``
  (src/fem/fem_state.rs:14-96, written by ``write_json`` at :231-245): scalar fields, the primary
  vectors as ``{"data": [...]}``, and per cell the ``SecondaryValues`` (src/fem/secondary_values.rs:8-37)
  with one ``LocalState`` per IP (src/material/local_state.rs:9-22): elastic flag, internal variables
  (von Mises: ``[z, lambda]``), stress as ``{"vec": {"data": [...]}, "mandel": ...}`` and strain.
* ``{stem}-summary.json`` (FileIo: active, dir, fn_stem, counter, indices, times; src/fem/file_io.rs:14-52),
  ``{stem}-base.json`` (FemBase: attribute map, element-DOF map, DOF numbering; src/fem/fem_base.rs),
  ``{stem}-mesh.msh`` (the text mesh format).
* ``compare_results``: pmsim's comparison against HYPLAS/SPO reference data
  (src/util/compare_results.rs:42-155): step k of the reference against state file k+1, displacement
  and stress (Mandel: shear components x sqrt 2) maximum absolute differences against two tolerances.
"""
import gzip
import json
import math
import os
from typing import List, Optional

import numpy as np

from . import mesh as M

KIND_NAMES = {M.TRI3: "Tri3", M.TRI6: "Tri6", M.TRI10: "Tri10", M.TRI15: "Tri15", M.QUA4: "Qua4", M.QUA8: "Qua8",
              M.QUA9: "Qua9", M.TET4: "Tet4", M.TET10: "Tet10", M.HEX8: "Hex8", M.HEX20: "Hex20"}
DOF_NAMES = ["Ux", "Uy", "Uz"]
SQ2 = math.sqrt(2.0)


def _vec(a):
    return {"data": [float(v) for v in np.asarray(a, dtype=float).ravel()]}


def _mandel(ndim):
    return "Symmetric2D" if ndim == 2 else "Symmetric"


def state_dict(ndim, t, ddt, u, ddu, ip_state, nip_per_cell, model_per_cell, ell=0.0):
    """FemState as pmsim serializes it.  ``ip_state`` rows are this library's IP records
    [sigma (Mandel), z, lambda, elastic flag, epbar] in cell order; the internal variables written are
    pmsim's ([z, lambda] for von Mises, none for linear elasticity); epbar has no place in pmsim's state."""
    ns = 2 * ndim
    ip_state = np.asarray(ip_state, dtype=float)
    gauss, row = [], 0
    for nip, model in zip(nip_per_cell, model_per_cell):
        solid = []
        for _ in range(int(nip)):
            r = ip_state[row]
            row += 1
            solid.append({"elastic": bool(r[ns + 2] != 0.0), "int_vars": _vec(r[ns:ns + 2] if model == 1 else []),
                          "stress": {"vec": _vec(r[:ns]), "mandel": _mandel(ndim)}, "strain": None})
        gauss.append({"ngauss": int(nip), "solid": solid, "porous_liq": [], "porous_liq_gas": [], "porous_sld_liq": [],
                      "porous_sld_liq_gas": []})
    out = {"ell": float(ell), "t": float(t), "ddt": float(ddt)}
    for k in range(1, 9):
        out[f"alpha{k}"] = 0.0
    out["beta1"] = 0.0
    out["beta2"] = 0.0
    out["ddu"] = _vec(ddu)
    out["u"] = _vec(u)
    for k in ("v", "a", "u_star", "v_star", "a_star"):
        out[k] = {"data": []}
    out["gauss"] = gauss
    return out


def read_state(path):
    """Reads a pmsim state file: t, u, ddu and, per IP in cell order, [sigma..., z, lambda, elastic]
    (internal variables padded with zeros for models that have fewer)."""
    with open(path) as f:
        d = json.load(f)
    rows = []
    for sv in d["gauss"]:
        for ls in sv["solid"]:
            iv = list(ls["int_vars"]["data"]) + [0.0, 0.0]
            rows.append(list(ls["stress"]["vec"]["data"]) + iv[:2] + [1.0 if ls["elastic"] else 0.0])
    return {"t": d["t"], "ddt": d["ddt"], "u": np.asarray(d["u"]["data"]), "ddu": np.asarray(d["ddu"]["data"]),
            "ip": np.asarray(rows), "ngauss": [sv["ngauss"] for sv in d["gauss"]]}


def base_dict(d):
    """FemBase as pmsim serializes it, from a flat problem description (``problem.describe``)."""
    ndim = int(d["ndim"])
    amap = {}
    for a, (model, E, nu, H, z0, ng) in d["params"].items():
        ss = ({"LinearElastic": {"young": E, "poisson": nu}} if model == 0 else
              {"VonMises": {"young": E, "poisson": nu, "hh": H, "z_ini": z0}})
        amap[str(a)] = {"Solid": {"density": float(d.get("densities", {}).get(a, 1.0)), "stress_strain": ss,
                                  "ngauss": int(ng) if ng else None}}
    emap = {}
    kinds, attrs = np.asarray(d["kinds"]), np.asarray(d["attrs"])
    for k, a in sorted(set(zip(kinds.tolist(), attrs.tolist()))):
        nn = M.NNODE[k]
        dofs = [[[DOF_NAMES[i], m * ndim + i] for i in range(ndim)] for m in range(nn)]
        emap[f"{a},{KIND_NAMES[k]}"] = {"dofs": dofs, "n_equation": nn * ndim, "eq_first_pl": None, "eq_first_pg": None,
                                        "eq_first_tt": None}
    npoint = len(d["coords"])
    dof_list = [{DOF_NAMES[i]: p * ndim + i for i in range(ndim)} for p in range(npoint)]
    return {"amap": {"all": amap}, "emap": {"all": emap}, "dofs": {"list": dof_list, "ndof": npoint * ndim}}


def write_msh(mesh, path):
    """The text mesh format read by ``Mesh.read``."""
    names = {M.TRI3: "tri3", M.TRI6: "tri6", M.TRI10: "tri10", M.TRI15: "tri15", M.QUA4: "qua4", M.QUA8: "qua8",
             M.QUA9: "qua9", M.TET4: "tet4", M.TET10: "tet10", M.HEX8: "hex8", M.HEX20: "hex20"}
    with open(path, "w") as f:
        f.write("# header\n# ndim npoint ncell\n")
        f.write(f"{mesh.ndim} {mesh.npoint} {mesh.ncell}\n\n# points\n# id marker x y {{z}}\n")
        for i, x in enumerate(mesh.coords):
            f.write(f"{i} 0 " + " ".join(repr(float(v)) for v in x) + "\n")
        f.write("\n# cells\n# id attribute kind points\n")
        for e in range(mesh.ncell):
            pts = mesh.cell_pts[mesh.cell_ptr[e]:mesh.cell_ptr[e + 1]]
            f.write(f"{e} {int(mesh.attrs[e])} {names[int(mesh.kinds[e])]} " + " ".join(str(int(p)) for p in pts) + "\n")


class FileIo:
    """Output files of a simulation (src/fem/file_io.rs): ``{dir}/{stem}-...``."""

    def __init__(self, dir: str, fn_stem: str):
        self.active = True
        self.dir = dir
        self.fn_stem = fn_stem
        self.counter = 0
        self.indices: List[int] = []
        self.times: List[float] = []

    def path_mesh(self): return f"{self.dir}/{self.fn_stem}-mesh.msh"
    def path_base(self): return f"{self.dir}/{self.fn_stem}-base.json"
    def path_summary(self): return f"{self.dir}/{self.fn_stem}-summary.json"
    def path_state(self, index: int): return f"{self.dir}/{self.fn_stem}-{index:020d}.json"

    def write_setup(self, mesh, d):
        os.makedirs(self.dir, exist_ok=True)
        write_msh(mesh, self.path_mesh())
        with open(self.path_base(), "w") as f:
            json.dump(base_dict(d), f)

    def write_state(self, state: dict):
        os.makedirs(self.dir, exist_ok=True)
        index = self.counter
        with open(self.path_state(index), "w") as f:
            json.dump(state, f)
        self.indices.append(index)
        self.times.append(float(state["t"]))
        self.counter += 1

    def write_summary(self):
        with open(self.path_summary(), "w") as f:
            json.dump({"active": self.active, "dir": self.dir, "fn_stem": self.fn_stem, "counter": self.counter,
                       "indices": self.indices, "times": self.times}, f)

    @staticmethod
    def read_summary(path):
        with open(path) as f:
            return json.load(f)


def _load_reference(ref_path):
    op = gzip.open if ref_path.endswith(".gz") else open
    with op(ref_path, "rt") as f:
        return json.load(f)["all"]


def compare_results(mesh, file_io: FileIo, ref_path: str, tol_displacement: float, tol_stress: float,
                    verbose: int = 0) -> bool:
    """pmsim's ``compare_results`` for SPO reference data: the state file of index k+1 against reference
    step k (index 0, the initial state, is skipped); every displacement component and every stress
    component at every IP must agree within its tolerance (shear stresses converted to Mandel)."""
    ndim = mesh.ndim
    if mesh.npoint < 1:
        raise ValueError("there must be at least one point in the mesh")
    if mesh.ncell < 1:
        raise ValueError("there must be at least one cell")
    ref = _load_reference(ref_path)
    if mesh.npoint != len(ref[0]["displacement"]):
        raise ValueError("the number of points in the mesh must equal the corresponding number in the reference data")
    if mesh.ncell != len(ref[0]["stresses"]):
        raise ValueError("the number of elements in the mesh must be equal to the reference number of elements (stresses)")
    summary = FileIo.read_summary(file_io.path_summary())
    if len(summary["indices"]) != len(ref) + 1:
        raise ValueError("the number of steps must equal the reference's number of steps + 1")
    all_good = True
    du_max = ds_max = -math.inf
    for index in range(1, len(summary["indices"])):
        step = ref[index - 1]
        st = read_state(file_io.path_state(index))
        u = st["u"].reshape(-1, ndim)
        du = np.abs(u - np.asarray(step["displacement"], dtype=float))
        du_max = max(du_max, float(du.max()))
        if (du > tol_displacement).any():
            all_good = False
        row = 0
        for e in range(mesh.ncell):
            ngauss = len(step["stresses"][e])
            if ngauss < 1:
                raise ValueError("there must be at least on integration point in reference data (stress)")
            for ip in range(ngauss):
                b = np.asarray(step["stresses"][e][ip], dtype=float)[: 2 * ndim].copy()
                b[3:] *= SQ2
                a = st["ip"][row, : 2 * ndim]
                row += 1
                ds = np.abs(a - b)
                ds_max = max(ds_max, float(ds.max()))
                if (ds > tol_stress).any():
                    all_good = False
            row += st["ngauss"][e] - ngauss
        if verbose > 0:
            print(f"step {index}: max |du| = {du.max():.2e}")
    if verbose > 0:
        print(f"diff_displacement_max = {du_max:9.2e}\ndiff_stress_max       = {ds_max:9.2e}")
    return all_good
