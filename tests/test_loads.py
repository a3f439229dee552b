"""Distributed loads Qn / Qx / Qy / Qz in the oracle (BcDistributed::calc_f_ext, src/fem/bc_distributed.rs:156-236),
pinned by pmsim's own known answers (bc_distributed.rs:446-512) and by the total force of a uniform traction."""
import numpy as np
import pytest

import loads_cases as LC
from oracle import oracle as orc
from pmsim_amd import Nbc, describe


@pytest.mark.parametrize("nbc", [Nbc.Qn, Nbc.Qx, Nbc.Qy])
@pytest.mark.parametrize("edge", ["top", "left", "right", "bottom"])
def test_oracle_qua8_edge_loads_known_answers(edge, nbc):
    o = orc.Oracle(describe(*LC.qua8_edge_problem(edge, nbc)))
    f = o.f_ext(0.0).reshape(-1, 2)
    want = np.zeros_like(f)
    for p, v in LC.expected_qua8(edge, nbc).items():
        want[p] = v
    assert np.abs(f - want).max() < 1e-14


def test_oracle_qz_on_3d_edge_known_answer():
    o = orc.Oracle(describe(*LC.hex8_edge_qz_problem()))
    f = o.f_ext(0.0).reshape(-1, 3)
    want = np.zeros_like(f)
    want[4, 2] = want[5, 2] = LC.Q / 2.0
    assert np.abs(f - want).max() < 1e-14


@pytest.mark.parametrize("kind", ["qua4", "qua8", "qua9", "tri6", "tri15"])
def test_oracle_traction_resultant_2d(kind):
    """The nodal forces of Qx on the right edge and Qy on the top add up to the traction times the edge
    length (exact for any shape functions that form a partition of unity)."""
    mesh, base, cfg, ess, nat = LC.traction_block_2d(kind, q=3.0)
    o = orc.Oracle(describe(mesh, base, cfg, ess, nat))
    f = o.f_ext(0.0).reshape(-1, 2)
    # prescribed rows are kept with LMM off only for the loaded faces' corners on the supports (RSM skips them)
    free_x = [p for p in range(mesh.npoint) if mesh.coords[p][0] > 0.0]
    free_y = [p for p in range(mesh.npoint) if mesh.coords[p][1] > 0.0]
    right = [p for p in free_x if mesh.coords[p][0] == 2.0]
    top = [p for p in free_y if mesh.coords[p][1] == 1.0]
    # the corner (2, 0) carries a Qx share but its uy is supported; (0, 1) carries a Qy share, ux supported
    assert abs(f[right, 0].sum() - 3.0 * 1.0) < 1e-13
    assert abs(f[top, 1].sum() + 1.5 * 2.0) < 1e-13
    assert np.abs(f[:, 0][[p for p in range(mesh.npoint) if mesh.coords[p][0] < 2.0]]).max() < 1e-14
