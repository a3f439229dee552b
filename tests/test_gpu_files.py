"""The HIP solver's output files in pmsim's formats, checked with pmsim's comparator semantics against
the HYPLAS fixtures (tolerances of the reference's own tests)."""
import os

import pytest

import cases
from pmsim_amd import file_io as fio
from pmsim_amd.solver import FemState, SolverImplicit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,case,tol_u,tol_s", [
    ("test_von_mises_single_element_2d_ref", cases.single_element, 1e-13, 1e-10),
    ("test_von_mises_2x2_elements_2d_ref", cases.four_qua8, 1e-12, 1e-9),
    ("spo_754_footing_ref", cases.spo_754_footing, 1e-10, 5e-5),
])
def test_solver_files_pass_compare_results(name, case, tol_u, tol_s, tmp_path):
    mesh, base, cfg, ess, nat = case()
    solver = SolverImplicit(mesh, base, cfg, ess, nat)
    io = fio.FileIo(str(tmp_path), name)
    solver.solve(FemState(mesh, base, ess, cfg), file_io=io)
    assert os.path.exists(io.path_summary()) and os.path.exists(io.path_base()) and os.path.exists(io.path_mesh())
    assert fio.compare_results(mesh, io, os.path.join(GOLDEN, name + ".json.gz"), tol_u, tol_s)
