"""Element output fields of the flow (DarcyFlowMHOutput::make_element_scalar /
make_element_vector, darcy_flow_mh_output.cc:261-325): element pressure,
piezometric head and the Darcy velocity at the element centre from the RT0 side
fluxes (MHFEValues::RT0_value, mh_fe_values.cc:76-118), on lines, triangles and
tetrahedra.  tests/01 of the reference from its inputs: the three $ElementData
sections of ref_output/flow_gmsh.con/test1.msh, written through the MSH writer,
against the reference's file as its ndiff compares them (same tokens, numbers
within 2e-5 of the 6 printed digits, and within 2e-10 absolute:
the reference solved to accuracy 1e-7, and a channel element's centre velocity is
half the difference of its two side fluxes), on the oracle (CPU) and on the GPU;
GPU against the oracle to 1e-9."""
import re

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff


def gold01():
    txt = open(cases.ref_path("test01_output.msh")).read()
    return re.sub(r"\$NodeData.*?\$EndNodeData\n", "", txt, flags=re.S)


def write01(gm, path, p, h, v):
    out = G.OutputMSH(gm, path)
    out.write_head()
    out.write_elem_data("pressure_elements", "L", p, 0.0, 0)
    out.write_elem_data("piezo_head_elements", "L", h, 0.0, 0)
    out.write_elem_data("velocity_elements", "L/T", v, 0.0, 0)


def solve01_oracle():
    om = O.Mesh(cases.ref_path("test01_mesh.msh"))
    d = cases.darcy_case("test01", om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                aniso=d["aniso"])
    D.solve(0, rtol=1e-15)
    return om, d, D


def test_element_output_oracle_test01(tmp_path):
    om, d, D = solve01_oracle()
    gm = G.Mesh(cases.ref_path("test01_mesh.msh"))
    p = str(tmp_path / "t.msh")
    write01(gm, p, *D.element_output())
    n = ndiff(open(p).read(), gold01(), 2e-5, 2e-10, columns=False)
    assert n > 831 * 5


def test_rt0_velocity_of_uniform_flow():
    """A linear pressure field p = 1 - x on a mixed mesh: every element velocity
    is K (1, 0, 0) projected on the element (exact in RT0), here checked on the
    tetrahedra of the generated cube, whose velocity is the full gradient."""
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "c.msh")
        cases.generated_mesh("cube", 3, path, n_fractures=0, n_channels=0, seed=1)
        om = O.Mesh(path)
        E, NB = om.n_elements, om.n_bc_elements
        cen = om.bc_centre.reshape(-1, 3)[:NB]
        bct = np.ones(NB, np.int32)
        bcp = 1.0 - cen[:, 0] - cen[:, 2]  # pressure 1 - x - z: piezometric head 1 - x
        D = O.Darcy(om, np.full(E, 2.0), np.ones(E), np.ones(E), np.zeros(E), bct, bcp, np.zeros(NB), np.zeros(NB))
        D.solve(0, rtol=1e-15)
        p, h, v = D.element_output()
        assert np.allclose(v, [[2.0, 0.0, 0.0]] * E, atol=1e-10)
        c = om.centre.reshape(-1, 3)[:E]
        assert np.allclose(h, 1.0 - c[:, 0], atol=1e-10)


@pytest.mark.gpu
def test_element_output_gpu_test01(tmp_path):
    om, d, D = solve01_oracle()
    gm = G.Mesh(cases.ref_path("test01_mesh.msh"))
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    pg, hg, vg = g.element_output()
    po, ho, vo = D.element_output()
    assert np.linalg.norm(pg - po) <= 1e-9 * np.linalg.norm(po)
    assert np.linalg.norm(hg - ho) <= 1e-9 * np.linalg.norm(ho)
    assert np.linalg.norm(vg - vo) <= 1e-9 * np.linalg.norm(vo)
    p = str(tmp_path / "t.msh")
    write01(gm, p, pg, hg, vg)
    ndiff(open(p).read(), gold01(), 2e-5, 2e-10, columns=False)
