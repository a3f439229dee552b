"""Writes the golden fixtures under tests/golden/ from the reference's own data and test vectors.

Run in the build container only (it reads /root/reference, which is absent on GPU hosts):

    python tests/golden/make_fixtures.py

Outputs (data only: inputs and expected outputs):
* spo_*.json.gz   -- HYPLAS per-step displacements, IP stresses and plastic flags/epbar
                     (data/spo/*_ref.json, loaded by src/util/reference_data_spo.rs:38-151)
* *.msh           -- meshes shipped with the reference (data/spo/*.msh)
* *.tsv           -- SPO curves digitised from the book, shipped with the reference (data/spo/*.tsv):
                     plate deflection-load and deflection profiles (753), footing load-displacement (754)
* sgm_5d17_ref.json -- Smith-Griffiths-Margetts Example 5.17 reference data (data/sgm/)
* smith_meshes.json -- point coordinates and cell connectivity of the Smith sample meshes and Bhatti's
                     Example 1.6 bracket (src/base/sample_meshes.rs:195-735)
* smith_u.json    -- expected displacement vectors of tests/test_solid_smith_*.rs and
                     tests/test_solid_bhatti_1d6_plane_stress.rs
* meshes/*.msh    -- meshes of tests/test_solid_felippa_thick_cylinder_axisym.rs and
                     tests/test_durand_farias_example4.rs (data/meshes/)
* artificial/     -- state, summary, base and mesh files written by pmsim itself for three small linear
                     elastic runs (data/results/artificial/): the on-disk formats' known answers
"""
import gzip
import json
import os
import re
import shutil

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def spo(name):
    with open(os.path.join(REF, "data/spo", name + ".json")) as f:
        d = json.load(f)
    steps = []
    for s in d["all"]:
        steps.append(dict(load_factor=s["load_factor"], n_iterations=len(s["iterations"]),
                          displacement=s["displacement"], stresses=s["stresses"],
                          plast_apex_epbar=s["plast_apex_epbar"]))
    # mtime=0: the same data gives the same bytes
    with gzip.GzipFile(os.path.join(OUT, name + ".json.gz"), "wb", mtime=0) as f:
        f.write(json.dumps(dict(all=steps)).encode())


def meshes():
    src = open(os.path.join(REF, "src/base/sample_meshes.rs")).read()
    out = {}
    for m in re.finditer(r"pub fn (smith_example_\w+|bhatti_example_1d6_bracket)\(\) -> Mesh \{(.*?)\n    \}\n", src, re.S):
        name, body = m.group(1), m.group(2)
        ndim = int(re.search(r"ndim:\s*(\d+)", body).group(1))
        pts = {}
        for p in re.finditer(r"Point \{ id:\s*(\d+),\s*marker:\s*-?\d+,\s*coords: vec!\[([^\]]*)\]", body):
            pts[int(p.group(1))] = [float(v) for v in p.group(2).split(",") if v.strip()]
        cells = []
        for c in re.finditer(r"Cell \{ id:\s*(\d+),\s*attribute:\s*(\d+),\s*kind: GeoKind::(\w+),\s*points: vec!\[([^\]]*)\]", body):
            cells.append([int(c.group(2)), c.group(3).lower(), [int(v) for v in c.group(4).split(",") if v.strip()]])
        coords = [pts[i] for i in range(len(pts))]
        out[name] = dict(ndim=ndim, coords=coords, cells=cells)
    with open(os.path.join(OUT, "smith_meshes.json"), "w") as f:
        json.dump(out, f, indent=0)


def smith_vectors():
    out = {}
    for fn in sorted(os.listdir(os.path.join(REF, "tests"))):
        if not (fn.startswith("test_solid_smith_") or fn.startswith("test_solid_bhatti_")):
            continue
        src = open(os.path.join(REF, "tests", fn)).read()
        m = re.search(r"let uu_correct = &\[(.*?)\];", src, re.S)
        if not m:
            continue
        vals = [float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", m.group(1))]
        out[fn[:-3]] = vals
    with open(os.path.join(OUT, "smith_u.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    for n in ["test_von_mises_single_element_2d_ref", "test_von_mises_2x2_elements_2d_ref",
              "spo_751_pres_cylin_ref", "spo_753_circ_plate_ref", "spo_754_footing_ref", "spo_755_tensile_perf_plast_ref"]:
        spo(n)
    for n in ["spo_751_pres_cylin_qua4.msh", "spo_753_circ_plate.msh", "spo_754_footing.msh", "spo_755_tensile.msh",
              "spo_753_plate_deflection_load.tsv", "spo_753_profiles.tsv", "spo_754_footing_load_displacement.tsv"]:
        shutil.copy(os.path.join(REF, "data/spo", n), os.path.join(OUT, n))
    shutil.copy(os.path.join(REF, "data/sgm/sgm_5d17_ref.json"), os.path.join(OUT, "sgm_5d17_ref.json"))
    meshes()
    smith_vectors()
    os.makedirs(os.path.join(OUT, "meshes"), exist_ok=True)
    for n in ["test_durand_farias_example4_qua4.msh", "test_solid_felippa_thick_cylinder_axisym.msh"]:
        shutil.copy(os.path.join(REF, "data/meshes", n), os.path.join(OUT, "meshes", n))
        os.chmod(os.path.join(OUT, "meshes", n), 0o644)
    art = os.path.join(OUT, "artificial")
    os.makedirs(art, exist_ok=True)
    for n in sorted(os.listdir(os.path.join(REF, "data/results/artificial"))):
        if n.startswith("artificial-elastic-2d-") or n.startswith("artificial-elastic-3d-"):
            shutil.copy(os.path.join(REF, "data/results/artificial", n), os.path.join(art, n))
            os.chmod(os.path.join(art, n), 0o644)
    print("fixtures written to", OUT)
