"""The HIP solver's output files in pmsim's formats, checked with pmsim's comparator semantics against
the HYPLAS fixtures (tolerances of the reference's own tests)."""
import math
import os

import numpy as np
import pytest

import cases
from pmsim_amd import file_io as fio
from pmsim_amd import post
from pmsim_amd.solver import FemState, SolverImplicit
from test_post import footing_curve, write_oracle_run

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,case,tol_u,tol_s", [
    ("test_von_mises_single_element_2d_ref", cases.single_element, 1e-13, 1e-10),
    ("test_von_mises_2x2_elements_2d_ref", cases.four_qua8, 1e-12, 1e-9),
    ("spo_754_footing_ref", cases.spo_754_footing, 1e-10, 5e-5),
    ("spo_751_pres_cylin_ref", cases.spo_751_pres_cylin, 1e-11, 1e-13),
    ("spo_753_circ_plate_ref", cases.spo_753_circ_plate, 2e-9, 1e-6),
])
def test_solver_files_pass_compare_results(name, case, tol_u, tol_s, tmp_path):
    mesh, base, cfg, ess, nat = case()
    solver = SolverImplicit(mesh, base, cfg, ess, nat)
    io = fio.FileIo(str(tmp_path), name)
    solver.solve(FemState(mesh, base, ess, cfg), file_io=io)
    assert os.path.exists(io.path_summary()) and os.path.exists(io.path_base()) and os.path.exists(io.path_mesh())
    assert fio.compare_results(mesh, io, os.path.join(GOLDEN, name + ".json.gz"), tol_u, tol_s)


def test_footing_load_settlement_from_solver_files(tmp_path):
    """examples/spo_754_footing_analysis.rs on the files the HIP solver wrote: the same curve as from the
    oracle's solution (1e-9 relative), within 2% of de Souza Neto et al.'s curve, plateau at 2 + pi."""
    mesh, base, cfg, ess, nat = cases.spo_754_footing()
    io = fio.FileIo(str(tmp_path / "gpu"), "footing")
    SolverImplicit(mesh, base, cfg, ess, nat).solve(FemState(mesh, base, ess, cfg), file_io=io)
    s, p = footing_curve(post.PostProc(str(tmp_path / "gpu"), "footing"), mesh)
    write_oracle_run(cases.spo_754_footing(), tmp_path / "orc", "footing")
    so, po = footing_curve(post.PostProc(str(tmp_path / "orc"), "footing"), mesh)
    assert np.abs(s - so).max() <= 1e-9 * np.abs(so).max()
    assert np.abs(p - po).max() <= 1e-9 * np.abs(po).max()
    book = np.loadtxt(os.path.join(GOLDEN, "spo_754_footing_load_displacement.tsv"), skiprows=1)
    assert (np.abs(p[1:] - book[1:, 1]) / book[1:, 1]).max() < 0.02
    assert abs(p[-4:].mean() - (2.0 + math.pi)) / (2.0 + math.pi) < 0.005


def test_durand_farias_example4_from_solver_files(tmp_path):
    """tests/test_durand_farias_example4.rs: the HIP solver's files post-processed into nodal stresses
    along the axis under the flexible footing, against the half-space solution within pmsim's
    tolerances and against the same post-processing of the oracle's solution (1e-10 relative)."""
    from pmsim_amd import At, Features
    mesh, base, cfg, ess, nat = cases.durand_farias_example4()
    io = fio.FileIo(str(tmp_path / "gpu"), "df4")
    SolverImplicit(mesh, base, cfg, ess, nat).solve(FemState(mesh, base, ess, cfg), file_io=io)
    write_oracle_run(cases.durand_farias_example4(), tmp_path / "orc", "df4")
    feat = Features(mesh)
    cells = feat.cells_via_edges(feat.search_edges(At.X(0.0)))
    res = {}
    for k in ("gpu", "orc"):
        pp = post.PostProc(str(tmp_path / k), "df4")
        res[k] = pp.nodal_stresses(pp.read_state(pp.n_state() - 1), cells, lambda x, y, z: x < 0.3)
    g, o = res["gpu"], res["orc"]
    assert g.k2id == o.k2id
    for name in ("txx", "tyy", "txy"):
        assert np.abs(getattr(g, name) - getattr(o, name)).max() <= 1e-10 * cases.DF_QN
    for x, y, sxx, syy, sxy in zip(g.xx, g.yy, g.txx, g.tyy, g.txy):
        a = cases.strip_load_stress(x, y)
        assert abs(sxx - a[0]) / cases.DF_QN <= 0.25 and abs(syy - a[1]) / cases.DF_QN <= 0.23
        assert abs(sxy - a[2]) / cases.DF_QN <= 0.09
