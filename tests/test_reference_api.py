"""The reference's own unit tests (tests/test_problem.py, test_mesh.py,
test_material.py), restated against toptimization_amd's setup modules."""

import tempfile
from pathlib import Path

import numpy as np
import pytest

from toptimization_amd.material import compute_Ke_2d, compute_Ke_3d
from toptimization_amd.mesh import build_edof_2d, element_centers_2d, grid_from_edof
from toptimization_amd.problem import Problem, load_problem

CANTILEVER_YAML = """\
problem: {name: "Test Cantilever", dimension: 2}
domain: {nelx: 10, nely: 5}
material: {E: 1.0, nu: 0.3, E_min: 1.0e-9}
boundary_conditions:
  fixed:
    - nodes: {type: "edge", edge: "left"}
      dofs: ["x", "y"]
loads:
  - nodes: {type: "point", location: [10, 2]}
    force: [0.0, -1.0]
optimization: {volume_fraction: 0.5, penalty: 3.0, filter_radius: 1.5, filter_type: "density",
               max_iterations: 100, tolerance: 0.01, move_limit: 0.2}
solver: {max_cg_iterations: 500, cg_tolerance: 1.0e-8, preconditioner: "jacobi"}
output: {directory: "results/test", save_interval: 5, format: "png"}
"""


def _write(text):
    f = tempfile.NamedTemporaryFile(suffix=".yaml", delete=False, mode="w")
    f.write(text)
    f.close()
    return Path(f.name)


@pytest.fixture(scope="module")
def prob():
    return load_problem(_write(CANTILEVER_YAML))


def test_problem_counts(prob):
    assert prob.name == "Test Cantilever" and prob.dim == 2
    assert (prob.nelx, prob.nely) == (10, 5)
    assert (prob.n_nodes, prob.n_dofs, prob.n_elem) == (66, 132, 50)
    assert len(prob.fixed_dofs) == 12
    assert prob.fixed_dofs.min() >= 0 and prob.fixed_dofs.max() < prob.n_dofs
    assert len(prob.force_dofs) > 0 and np.any(prob.force_values != 0)
    assert prob.volfrac == pytest.approx(0.5)
    np.testing.assert_array_equal(np.union1d(prob.free_dofs, prob.fixed_dofs), np.arange(prob.n_dofs))
    assert len(np.intersect1d(prob.free_dofs, prob.fixed_dofs)) == 0
    # load at node (10, 2): dof 2*(10*6+2)+1
    assert prob.force_dofs.tolist() == [2 * (10 * 6 + 2) + 1]


def test_problem_validation_errors():
    bad = CANTILEVER_YAML.replace('preconditioner: "jacobi"', 'preconditioner: "ilu"')
    with pytest.raises(ValueError):
        load_problem(_write(bad))
    with pytest.raises(ValueError):
        load_problem(_write(CANTILEVER_YAML.replace("force: [0.0, -1.0]", "force: [0.0, 0.0]")))
    with pytest.raises(ValueError):
        load_problem(_write(CANTILEVER_YAML.replace('filter_type: "density"', 'filter_type: "x"')))
    ok = load_problem(_write(CANTILEVER_YAML.replace('preconditioner: "jacobi"',
                                                     'preconditioner: "block_jacobi_2x2", precision: "mixed"')))
    assert ok.preconditioner == "block_jacobi_2x2" and ok.precision == "mixed"


def test_reference_example_files_parse():
    ex = Path("/root/reference/examples")
    if not ex.exists():
        pytest.skip("reference examples not mounted")
    for y in sorted(ex.glob("*.yaml")):
        p = load_problem(y)
        assert p.n_dofs == 2 * (p.nelx + 1) * (p.nely + 1)
    mbb = load_problem(ex / "mbb_beam_2d.yaml")
    assert len(mbb.fixed_dofs) == 61 + 1   # left edge x only + bottom-right y
    assert mbb.force_dofs.tolist() == [2 * 60 + 1]


def make_problem_2d(nelx=4, nely=3):
    return Problem(name="test", dim=2, nelx=nelx, nely=nely, nelz=1, E=1.0, nu=0.3, E_min=1e-9,
                   fixed_dofs=np.array([0, 1]), force_dofs=np.array([2 * (nelx + 1) * (nely + 1) - 1]),
                   force_values=np.array([-1.0]), volfrac=0.5, penalty=3.0, rmin=1.5,
                   filter_type="density", max_iter=100, tol=0.01, move_limit=0.2, max_cg_iter=500,
                   cg_tol=1e-8, preconditioner="jacobi", output_dir=Path("results"),
                   save_interval=10, output_format="png")


def test_edof_2d():
    p = make_problem_2d()
    edof = build_edof_2d(p)
    assert edof.shape == (12, 8) and edof.dtype == np.int32
    assert edof.min() >= 0 and edof.max() < 2 * 5 * 4
    np.testing.assert_array_equal(edof[0], [0, 1, 8, 9, 10, 11, 2, 3])
    assert all(len(set(r)) == 8 for r in edof)
    assert len(set(edof[0]) & set(edof[3])) == 4
    assert grid_from_edof(edof, 12) == (4, 3)


def test_element_centers_2d():
    c = element_centers_2d(make_problem_2d())
    assert c.shape == (12, 2) and c.dtype == np.float32
    np.testing.assert_allclose(c[0], [0.5, 0.5])
    np.testing.assert_allclose(c[11], [3.5, 2.5])


def test_ke_2d_properties():
    Ke = compute_Ke_2d(1.0, 0.3)
    assert Ke.shape == (8, 8)
    np.testing.assert_allclose(Ke, Ke.T, atol=1e-12)
    assert np.linalg.eigvalsh(Ke).min() >= -1e-10
    assert np.linalg.matrix_rank(Ke, tol=1e-10) == 5
    np.testing.assert_allclose(compute_Ke_2d(2.0, 0.3), 2 * Ke, rtol=1e-10)
    assert np.all(np.diag(Ke) > 0)


def test_ke_3d_properties():
    Ke = compute_Ke_3d(1.0, 0.3)
    assert Ke.shape == (24, 24)
    np.testing.assert_allclose(Ke, Ke.T, atol=1e-12)
    assert np.linalg.eigvalsh(Ke).min() >= -1e-8
    assert np.linalg.matrix_rank(Ke, tol=1e-8) == 18
