"""pmsim's on-disk formats: state/summary/base/mesh files against the files pmsim itself wrote
(tests/golden/artificial, copied from the reference's data/results/artificial), and compare_results
against the HYPLAS fixtures with states produced by the oracle."""
import json
import os

import numpy as np
import pytest

import cases
from oracle import oracle as orc
from pmsim_amd import Config, Elem, Essential, FemBase, Mesh, Natural, ParamSolid, StressStrain, describe
from pmsim_amd import file_io as fio
from pmsim_amd import mesh as M

ART = os.path.join(os.path.dirname(__file__), "golden", "artificial")
STEMS = ["artificial-elastic-2d", "artificial-elastic-2d-qua8", "artificial-elastic-3d"]


def _strip_strain(d):
    for sv in d["gauss"]:
        for ls in sv["solid"]:
            ls["strain"] = None
    return d


@pytest.mark.parametrize("stem", STEMS)
def test_state_files_round_trip(stem):
    """Reading pmsim's state files and writing them back gives the same JSON (strain aside: this
    library does not record strains)."""
    for k in range(3):
        path = os.path.join(ART, f"{stem}-{k:020d}.json")
        ref = json.load(open(path))
        st = fio.read_state(path)
        ndim = len(ref["gauss"][0]["solid"][0]["stress"]["vec"]["data"]) // 2
        ip = np.hstack([st["ip"], np.zeros((len(st["ip"]), 1))])  # + epbar column
        models = [0] * len(st["ngauss"])  # linear elastic: no internal variables
        out = fio.state_dict(ndim, st["t"], st["ddt"], st["u"], st["ddu"], ip, st["ngauss"], models, ell=ref["ell"])
        assert json.loads(json.dumps(out)) == _strip_strain(ref)


@pytest.mark.parametrize("stem", STEMS)
def test_base_mesh_and_summary_match_pmsim(stem, tmp_path):
    mesh = Mesh.read(os.path.join(ART, f"{stem}-mesh.msh"))
    attrs = sorted(set(int(a) for a in mesh.attrs))
    base = FemBase(mesh, [(a, Elem.Solid(ParamSolid(1.0, StressStrain.LinearElastic(1500.0, 0.25), None))) for a in attrs])
    d = describe(mesh, base, Config(mesh), Essential(), Natural())
    assert fio.base_dict(d) == json.load(open(os.path.join(ART, f"{stem}-base.json")))
    io = fio.FileIo(str(tmp_path), stem)
    io.write_setup(mesh, d)
    assert open(io.path_mesh()).read().split() == open(os.path.join(ART, f"{stem}-mesh.msh")).read().split()
    ref_sum = json.load(open(os.path.join(ART, f"{stem}-summary.json")))
    for k in range(3):
        io.write_state(json.load(open(os.path.join(ART, f"{stem}-{k:020d}.json"))))
    io.dir = ref_sum["dir"]  # the reference ran in another directory
    got = {"active": io.active, "dir": io.dir, "fn_stem": io.fn_stem, "counter": io.counter, "indices": io.indices,
           "times": io.times}
    assert got == ref_sum


def test_state_path_naming():
    io = fio.FileIo("/tmp/pmsim", "footing")
    assert io.path_state(3) == "/tmp/pmsim/footing-00000000000000000003.json"
    assert io.path_summary() == "/tmp/pmsim/footing-summary.json"
    assert io.path_base() == "/tmp/pmsim/footing-base.json"
    assert io.path_mesh() == "/tmp/pmsim/footing-mesh.msh"


def _oracle_files(case, tmp_path, stem):
    mesh, base, cfg, ess, nat = case
    d = describe(mesh, base, cfg, ess, nat)
    u, ip, _ = orc.Oracle(d).solve()
    nip = [int(d["params"][int(a)][5]) or M.DEFAULT_NGAUSS[int(k)] for k, a in zip(d["kinds"], d["attrs"])]
    model = [int(d["params"][int(a)][0]) for a in d["attrs"]]
    io = fio.FileIo(str(tmp_path), stem)
    io.write_setup(mesh, d)
    prev = None
    for k in range(len(u)):
        uk = u[k][: mesh.npoint * mesh.ndim]
        ddu = uk - prev if prev is not None else np.zeros_like(uk)
        io.write_state(fio.state_dict(mesh.ndim, float(k), 1.0, uk, ddu, ip[k], nip, model))
        prev = uk
    io.write_summary()
    return mesh, io


def test_compare_results_on_hyplas_fixture(tmp_path):
    """pmsim's comparator semantics on the single-element von Mises test (tolerances of
    tests/test_von_mises_single_element_2d.rs): the oracle's states pass, a perturbed stress fails."""
    ref = os.path.join(os.path.dirname(__file__), "golden", "test_von_mises_single_element_2d_ref.json.gz")
    mesh, io = _oracle_files(cases.single_element(), tmp_path, "single")
    assert fio.compare_results(mesh, io, ref, 1e-13, 1e-10)
    p = io.path_state(2)
    st = json.load(open(p))
    st["gauss"][0]["solid"][0]["stress"]["vec"]["data"][1] += 1e-8
    json.dump(st, open(p, "w"))
    assert not fio.compare_results(mesh, io, ref, 1e-13, 1e-10)
    with pytest.raises(ValueError, match="number of steps"):
        io.indices.append(99)
        io.write_summary()
        fio.compare_results(mesh, io, ref, 1e-13, 1e-10)


def test_compare_results_on_footing_fixture(tmp_path):
    ref = os.path.join(os.path.dirname(__file__), "golden", "spo_754_footing_ref.json.gz")
    mesh, io = _oracle_files(cases.spo_754_footing(), tmp_path, "footing")
    assert fio.compare_results(mesh, io, ref, 1e-10, 5e-5)


def test_msh_writer_round_trip(tmp_path):
    mesh = cases.spo_754_footing()[0]
    p = str(tmp_path / "m.msh")
    fio.write_msh(mesh, p)
    back = Mesh.read(p)
    assert np.array_equal(back.coords, mesh.coords)
    assert np.array_equal(back.kinds, mesh.kinds) and np.array_equal(back.cell_pts, mesh.cell_pts)
