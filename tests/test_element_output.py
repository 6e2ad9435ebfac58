"""Output fields of the flow (DarcyFlowMHOutput::make_element_scalar /
make_element_vector / make_node_scalar_param, darcy_flow_mh_output.cc:261-325,
424-541): element pressure, piezometric head, the Darcy velocity at the element
centre from the RT0 side fluxes (MHFEValues::RT0_value, mh_fe_values.cc:76-118),
on lines, triangles and tetrahedra, and the pressure interpolated to the nodes.
tests/01 of the reference from its inputs: the whole of
ref_output/flow_gmsh.con/test1.msh (mesh, $NodeData and three $ElementData
sections), written through the MSH writer, against the reference's file as its
ndiff compares them (same tokens, numbers within 2e-5 of the 6 printed digits,
and within 2e-10 absolute: the reference solved to accuracy 1e-7, and a channel
element's centre velocity is half the difference of its two side fluxes), on the
oracle (CPU) and on the GPU; GPU against the oracle to 1e-9.

The node interpolation is pinned beyond the file's 6 digits by two properties of
its weights: they sum to 1 at every node (so a constant solution gives that
constant at every node), and a node's value is a convex combination of the
element pressures and side traces around it."""
import re

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff


def gold01():
    return open(cases.ref_path("test01_output.msh")).read()


def write01(gm, path, pn, p, h, v):
    out = G.OutputMSH(gm, path)
    out.write_head()
    out.write_node_data("pressure_nodes", "L", pn, 0.0, 0)
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
    write01(gm, p, D.node_pressure(), *D.element_output())
    n = ndiff(open(p).read(), gold01(), 2e-5, 2e-10, columns=False)
    assert n > 831 * 5 + 192 * 2


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


def test_node_pressure_weights(tmp_path):
    """Constant solution -> the constant at every node (weights sum to 1); a node
    value lies between the smallest and largest value around the node.  The
    constant solution: Dirichlet 2.5 on every boundary of the flat (z = 0) mixed
    1D/2D mesh of tests/06, where pressure and piezometric head coincide."""
    om = O.Mesh(cases.ref_path("test6_21d.msh"))
    E, NB = om.n_elements, om.n_bc_elements
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), np.ones(NB, np.int32), np.full(NB, 2.5),
                np.zeros(NB), np.zeros(NB))
    D.solve(0, rtol=1e-15)
    assert np.allclose(D.solution[D.S:], 2.5, rtol=1e-13)
    pn = D.node_pressure()
    used = np.zeros(om.n_nodes, bool)
    en = om.elem_nodes.reshape(-1, 4)[:E]
    used[en[en >= 0]] = True
    assert np.allclose(pn[used], 2.5, rtol=1e-12)
    assert np.all(pn[~used] == 0.0)
    om1, _, D1 = solve01_oracle()
    pn = D1.node_pressure()
    lo = np.full(om1.n_nodes, np.inf)
    hi = np.full(om1.n_nodes, -np.inf)
    en = om1.elem_nodes.reshape(-1, 4)[:om1.n_elements]
    pe = D1.solution[D1.S:D1.S + D1.E]
    tr = D1.solution[D1.S + D1.E:]
    for e in range(om1.n_elements):
        for v in en[e][en[e] >= 0]:
            lo[v], hi[v] = min(lo[v], pe[e]), max(hi[v], pe[e])
    for s_, v in zip(*side_node_pairs(om1)):
        lo[v], hi[v] = min(lo[v], tr[s_]), max(hi[v], tr[s_])
    assert np.all(pn >= lo - 1e-12) and np.all(pn <= hi + 1e-12)


def side_node_pairs(om):
    """(edge index, node) for every node of every side (ref_element.cc:51-74)."""
    tab = {1: [[0], [1]], 2: [[0, 1], [1, 2], [0, 2]], 3: [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]]}
    en, ee = om.elem_nodes.reshape(-1, 4), om.elem_edge.reshape(-1, 4)
    edges, nodes = [], []
    for e in range(om.n_elements):
        dim = int(om.elem_dim[e])
        for s in range(dim + 1):
            for k in tab[dim][s]:
                edges.append(int(ee[e][s]))
                nodes.append(int(en[e][k]))
    return edges, nodes


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
    ng, no = g.node_pressure(), D.node_pressure()
    assert np.linalg.norm(ng - no) <= 1e-9 * np.linalg.norm(no)
    p = str(tmp_path / "t.msh")
    write01(gm, p, ng, pg, hg, vg)
    ndiff(open(p).read(), gold01(), 2e-5, 2e-10, columns=False)
