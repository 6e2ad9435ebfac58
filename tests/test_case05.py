"""tests/05_flow_boundary_conditions of the reference from its input files: the
unit square of 942 triangles with exact solution u = xy and, in three variants,
Dirichlet (all four sides), Neumann (south x, north -x) and Robin (south -2x,
north 3x, sigma 0.5) data on two of the sides, formula values taken at the
boundary element centre (darcy_flow_mh.cc:572-591).

Checked against ref_output/flow_<case>.con: the VTU fields (node pressure,
element pressure, element velocity; extracted by tests/golden/extract_test05.py)
and water_balance.out.  The reference solved to accuracy 1e-7, which bounds the
agreement: 1e-6 on pressures of scale 1 and 3e-6 on velocities; the
discretization error against u = xy is 1.4e-4, so a boundary term of the wrong
size or sign, or data taken at the wrong point, fails the comparison.  GPU: the
same, and the GPU fields within 1e-9 of the oracle's."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff

VARIANTS = ["dirichlet", "neumann", "robin"]


def golden(variant):
    return {l.split()[0]: np.array(l.split()[1:], float) for l in open(cases.ref_path("test05_%s_fields.txt" % variant))}


def check_fields(variant, pn, p, v):
    g = golden(variant)
    assert np.max(np.abs(pn - g["pressure_nodes"])) < 1e-6, variant
    assert np.max(np.abs(p - g["pressure_elements"])) < 1e-6, variant
    assert np.max(np.abs(v.reshape(-1) - g["velocity_elements"])) < 3e-6, variant


def check_balance(gm, path, bal, variant):
    G.write_water_balance(gm, path, float("inf"), *bal, np.zeros(1))
    ref = open(cases.ref_path("test05_%s_water_balance.out" % variant)).read()
    # the totals of about 1e-6 are the reference's solver residual
    return ndiff(open(path).read(), ref, 2e-5, 2e-6)


@pytest.mark.parametrize("variant", VARIANTS)
def test_case05_oracle_against_reference_output(tmp_path, variant):
    om = O.Mesh(cases.ref_path("test05.msh"))
    gm = G.Mesh(cases.ref_path("test05.msh"))
    d = cases.darcy_case("test05_" + variant, om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D.solve(0, rtol=1e-15)
    p, h, v = D.element_output()
    check_fields(variant, D.node_pressure(), p, v)
    assert np.array_equal(h, p)  # the square lies in z = 0
    c = om.centre.reshape(-1, 3)[:om.n_elements]
    assert np.max(np.abs(p - c[:, 0] * c[:, 1])) < 2e-4  # u = xy at the element centres
    n = check_balance(gm, str(tmp_path / "water_balance.out"), D.water_balance(), variant)
    assert n > 20


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS)
def test_case05_gpu_against_reference_output(tmp_path, variant):
    om = O.Mesh(cases.ref_path("test05.msh"))
    gm = G.Mesh(cases.ref_path("test05.msh"))
    d = cases.darcy_case("test05_" + variant, om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D.solve(0, rtol=1e-15)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    x = g.get_solution_vector()
    assert np.linalg.norm(x - D.solution) <= 1e-9 * np.linalg.norm(D.solution)
    pg, hg, vg = g.element_output()
    ng = g.node_pressure()
    po, ho, vo = D.element_output()
    no = D.node_pressure()
    for a, b in [(pg, po), (vg, vo), (ng, no)]:
        assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)
    check_fields(variant, ng, pg, vg)
    p = str(tmp_path / "wb.out")
    g.write_water_balance(p)
    ndiff(open(p).read(), open(cases.ref_path("test05_%s_water_balance.out" % variant)).read(), 2e-5, 2e-6)
