"""tests/03_transport_small_12d of the reference from its input files
(flow_old_gmsh.con): the model of tests/02 (1D-2D flow with anisotropy 10 on the
lines, dual porosity, linear and Freundlich sorption) on a 97-element mesh with
output every 0.01 to t = 1, so the step sequence runs through 100 output marks.

Checked against ref_output/flow_old_gmsh.con: the flow output test3.msh, the
water balance, and the transport output and mass balance at t = 0.1, 0.5 and 1
(sections extracted unchanged by tests/golden/extract_test03.py), within the
printed 6 digits.  GPU: the same from the GPU flow and transport."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff
from test_case02 import PHASES, write_flow

NAMES = ["age", "U235"]
SEL = (10, 50, 100)  # output indices of t = 0.1, 0.5, 1


def load03():
    ref = cases.ref_path
    gm, om = G.Mesh(ref("test03.msh")), O.Mesh(ref("test03.msh"))
    b = G.OldBcdInput(gm)
    bc = b.read_flow(ref("test03.fbc"))
    bcc = b.read_transport(ref("test03.tbc"), 2)
    E = om.n_elements
    aniso = np.zeros((E, 9))
    a = np.where(om.elem_region == 17, 10.0, 1.0)
    for k in (0, 4, 8):
        aniso[:, k] = a
    return gm, om, bc, bcc, E, aniso


def run(T, update, conc, gm, tr_path, mb_path, write_mb):
    out = G.OutputMSH(gm, tr_path)
    for k in range(101):
        t = 0.01 * k
        update(t)
        if k in SEL:
            c = conc()
            out.current_step = k  # the reference's stream counts every output time
            out.write_data(t, [("%s_mobile" % n, "M/L^3", c[i]) for i, n in enumerate(NAMES)])
            write_mb(mb_path)


def sections(path):
    txt = open(path).read()
    return "".join("$ElementData\n" + p.split("$EndElementData\n")[0] + "$EndElementData\n"
                   for p in txt.split("$ElementData\n")[1:])


def compare(flow_path, tr_path, mb_path):
    ref = cases.ref_path
    ndiff(open(flow_path).read(), open(ref("test03_flow.msh")).read(), 2e-5, 1e-6, columns=False)
    n = ndiff(sections(tr_path), open(ref("test03_transport_sel.msh")).read(), 2e-5, 1e-8, columns=False)
    assert n > 3 * 2 * 97
    ndiff(open(mb_path).read(), open(ref("test03_mass_balance_sel.txt")).read(), 2e-5, 1e-9)


def test_case03_oracle_against_reference_output(tmp_path):
    gm, om, bc, bcc, E, aniso = load03()
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *bc, aniso=aniso)
    D.solve(0, rtol=1e-15)
    p, h, v = D.element_output()
    fp = str(tmp_path / "test3.msh")
    write_flow(gm, fp, D.node_pressure(), p, v)
    pw = str(tmp_path / "water_balance.txt")
    G.write_water_balance(gm, pw, float("inf"), *D.water_balance(), np.zeros(2))
    ndiff(open(pw).read(), open(cases.ref_path("test03_water_balance.txt")).read(), 2e-5, 1e-6)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, 0.25), bcc, np.zeros((2, E)), 1.0, 0.01)
    T.set_phases(1, 1, **PHASES)
    tp, mp = str(tmp_path / "t.msh"), str(tmp_path / "mb.txt")
    run(T, T.update, T.conc, gm, tp, mp,
        lambda path: G.write_mass_balance(gm, path, T.time, 0.0, NAMES, T.mass_balance()))
    compare(fp, tp, mp)


@pytest.mark.gpu
def test_case03_gpu_against_reference_output(tmp_path):
    gm, om, bc, bcc, E, aniso = load03()
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *bc, aniso=aniso)
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    p, h, v = g.element_output()
    fp = str(tmp_path / "test3.msh")
    write_flow(gm, fp, g.node_pressure(), p, v)
    T = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), bcc, np.zeros((2, E)), 1.0, 0.01)
    T.set_velocity_field(darcy=g)
    T.set_dual_porosity_sorption(True, True, **PHASES)
    tp, mp = str(tmp_path / "t.msh"), str(tmp_path / "mb.txt")
    run(T, T.update_solution, T.conc, gm, tp, mp, lambda path: T.write_mass_balance(path, NAMES))
    compare(fp, tp, mp)
