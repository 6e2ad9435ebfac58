"""tests/08_steady_flow_sources of the reference from its input files
(flow_vtk.con): the square (-1, 1)^2 of 4016 triangles, homogeneous Dirichlet
data and the water source density 2 (1 - x^2) + 2 (1 - y^2) taken at the element
centres (darcy_flow_mh.cc:612-614), exact solution (1 - x^2)(1 - y^2).

Checked against the element pressures of ref_output/flow_vtk.con/flow-000000.vtu
(extracted by tests/golden/extract_test08.py): the oracle agrees to 4.3e-11 on the
scale 1, so the source term's size, sign and evaluation point are pinned (the
discretization error against the exact solution is 3.5e-4).  GPU: the same, and
within 1e-9 of the oracle."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O


def case08(om):
    E, NB = om.n_elements, om.n_bc_elements
    c = om.centre.reshape(-1, 3)[:E]
    f = 2 * (1 - c[:, 0] ** 2) + 2 * (1 - c[:, 1] ** 2)
    return (np.ones(E), np.ones(E), np.ones(E), f, np.ones(NB, np.int32), np.zeros(NB), np.zeros(NB), np.zeros(NB)), c


def gold08():
    return np.array(open(cases.ref_path("test08_pressure.txt")).read().split(), float)


def test_case08_oracle_against_reference_output():
    om = O.Mesh(cases.ref_path("test08.msh"))
    args, c = case08(om)
    D = O.Darcy(om, *args)
    D.solve(0, rtol=1e-15)
    p = D.solution[D.S:D.S + D.E]
    assert np.max(np.abs(p - gold08())) < 1e-9
    assert np.max(np.abs(p - (1 - c[:, 0] ** 2) * (1 - c[:, 1] ** 2))) < 5e-4


@pytest.mark.gpu
def test_case08_gpu_against_reference_output():
    om = O.Mesh(cases.ref_path("test08.msh"))
    gm = G.Mesh(cases.ref_path("test08.msh"))
    args, c = case08(om)
    D = O.Darcy(om, *args)
    D.solve(0, rtol=1e-15)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(*args)
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    x = g.get_solution_vector()
    assert np.linalg.norm(x - D.solution) <= 1e-9 * np.linalg.norm(D.solution)
    assert np.max(np.abs(x[D.S:D.S + D.E] - gold08())) < 1e-9
