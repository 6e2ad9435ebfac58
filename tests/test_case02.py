"""tests/02_transport_12d of the reference from its input files (flow_gmsh.con):
steady flow on 542 triangles and 135 lines coupled through vb neighbours
(anisotropy 10 on the lines, legacy .fbc), then explicit transport of two
substances to t = 5 with dual porosity (por_m = por_imm = 0.25, alpha 0.01) and
sorption (A linear, B Freundlich with exponent 0.5; phi 0.5), boundary
concentrations from the legacy .tbc (ConvectionTransport::transport_dual_porosity,
transport_sorption, compute_sorption, transport.cc:1026-1160).

Checked against ref_output/flow_gmsh.con: the flow output test2.msh (node
pressure, element pressure, element velocity), transport.msh (both substances at
the 11 output times) and mass_balance.txt, whose region mass sums every phase
times por_m (transport.cc:1354-1362).  The printed 6 digits bound the agreement
(rtol 2e-5); the reference's flow solve leaves absolute differences up to 5.5e-7
on velocities of scale 14.6 and 1e-8 on concentrations of scale 1 in the
steepest front tails.  Without dual porosity, or with B linear instead of
Freundlich, the concentrations and masses miss the reference's by 2e-3 and more.

Langmuir (no reference output uses it) is pinned by its isotherm: after each
step the dissolved concentration solves c + N c1 c0 c / (1 + c0 c) = pore mass
and the sorbed part is the rest of the pore mass; the mass of the mobile and
immobile phases is conserved by the dual porosity exchange.

GPU: the same against the reference's files; and against the oracle (the phase
update uses exp and pow of the device library, so agreement is to 1e-12 rather
than bitwise), with decays acting on both phases, on a generated 1D/2D/3D mesh."""
import os

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff

NAMES = ["A", "B"]


def load02():
    ref = cases.ref_path
    gm, om = G.Mesh(ref("test02.msh")), O.Mesh(ref("test02.msh"))
    b = G.OldBcdInput(gm)
    bc = b.read_flow(ref("test02.fbc"))
    bcc = b.read_transport(ref("test02.tbc"), 2)
    E = om.n_elements
    aniso = np.zeros((E, 9))
    a = np.where(om.elem_region == 48, 10.0, 1.0)  # rid 48: anisotropy 10, rid 184: 1
    for k in (0, 4, 8):
        aniso[:, k] = a
    return dict(gm=gm, om=om, bc=bc, bcc=bcc, E=E, aniso=aniso)


PHASES = dict(por_imm=0.25, alpha=0.01, phi=0.5, sorp_type=np.array([[1], [2]]), sorp_coef0=0.02,
              sorp_coef1=np.array([[0.0], [0.5]]))


def write_flow(gm, path, pn, p, v):
    out = G.OutputMSH(gm, path)
    out.write_head()
    out.write_node_data("pressure_nodes", "L", pn, 0.0, 0)
    out.write_elem_data("pressure_elements", "L", p, 0.0, 0)
    out.write_elem_data("velocity_elements", "L/T", v, 0.0, 0)


def run_transport(T, update, conc, gm, tr_path, mb_path, write_mb):
    out = G.OutputMSH(gm, tr_path)
    hist = []
    for k in range(11):
        update(0.5 * k)
        c = conc()
        hist.append(c.copy())
        out.write_data(0.5 * k, [("%s_mobile" % n, "M/L^3", c[i]) for i, n in enumerate(NAMES)])
        write_mb(mb_path)
    return hist


def compare(tmp_path, flow_path, tr_path, mb_path):
    ref = cases.ref_path
    ndiff(open(flow_path).read(), open(ref("test02_flow.msh")).read(), 2e-5, 1e-6, columns=False)
    n = ndiff(open(tr_path).read(), open(ref("test02_transport.msh")).read(), 2e-5, 1e-8, columns=False)
    assert n > 11 * 2 * 677
    n = ndiff(open(mb_path).read(), open(ref("test02_mass_balance.txt")).read(), 2e-5, 1e-9)
    assert n > 11 * 40


def test_case02_oracle_against_reference_output(tmp_path):
    c = load02()
    gm, om, E = c["gm"], c["om"], c["E"]
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    D.solve(0, rtol=1e-15)
    p, h, v = D.element_output()
    fp = str(tmp_path / "test2.msh")
    write_flow(gm, fp, D.node_pressure(), p, v)
    pw = str(tmp_path / "water_balance.txt")
    G.write_water_balance(gm, pw, float("inf"), *D.water_balance(), np.zeros(2))
    ndiff(open(pw).read(), open(cases.ref_path("test02_water_balance.txt")).read(), 2e-5, 1e-6)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, 0.25), c["bcc"], np.zeros((2, E)), 5.0, 0.5)
    T.set_phases(1, 1, **PHASES)
    tp, mp = str(tmp_path / "transport.msh"), str(tmp_path / "mass_balance.txt")
    run_transport(T, T.update, T.conc, gm, tp, mp,
                  lambda path: G.write_mass_balance(gm, path, T.time, 0.0, NAMES, T.mass_balance()))
    compare(tmp_path, fp, tp, mp)


def small_case(seed=3):
    """Two substances on the mixed 1D/2D mesh of tests/06 with its flow."""
    om = O.Mesh(cases.ref_path("test6_21d.msh"))
    d = cases.darcy_case("test06_21", om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D.solve(0, rtol=1e-15)
    E, NB = om.n_elements, om.n_bc_elements
    rng = np.random.default_rng(seed)
    bcc = rng.uniform(0.2, 1.0, (2, NB))
    ic = rng.uniform(0.0, 0.5, (2, E))
    return om, D, bcc, ic


def test_langmuir_isotherm_and_dual_exchange():
    om, D, bcc, ic = small_case()
    E = om.n_elements
    por_m, por_imm, phi = 0.3, 0.2, 0.4
    c0, c1 = np.array([[0.7], [1.5]]), np.array([[2.0], [0.3]])
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, por_m), bcc, ic, 1.0, 0.25)
    T.set_phases(1, 1, por_imm, alpha=0.5, phi=phi, sorp_type=3, sorp_coef0=c0, sorp_coef1=c1)
    n = 1 - (por_m + por_imm)
    for t in (0.25, 0.5, 1.0):
        T.update(t)
        cm, ci, sm, si = (T.conc_phase(k) for k in range(4))
        for conc, sorb, Nv, N in ((cm, sm, por_m / n, n * phi / por_m), (ci, si, por_imm / n, n * (1 - phi) / por_imm)):
            avg = conc + sorb / Nv
            resid = conc + N * c1 * c0 * conc / (1 + c0 * conc) - avg
            assert np.max(np.abs(resid)) < 1e-7 * np.max(np.abs(avg)), t
            assert np.all(sorb >= -1e-15)
    # dual porosity alone conserves por_m c_m + por_imm c_i element by element
    T = O.Transport(om, np.zeros(D.S), np.ones(E), np.full(E, por_m), bcc, ic, 1.0, 0.25)
    T.set_phases(1, 0, por_imm, alpha=0.5)
    T.update(0.25)
    dt = T.dt
    steps = round(0.25 / dt)
    cm, ci = T.conc_phase(0), T.conc_phase(1)
    assert np.allclose(por_m * cm + por_imm * ci, por_m * ic, rtol=1e-13, atol=1e-16)
    # closed form with no flow: c_m - c_i decays by exp(-alpha (pm + pi) / (pm pi) dt) per step
    f = np.exp(-0.5 * (por_m + por_imm) / (por_m * por_imm) * dt) ** steps
    assert np.allclose(cm - ci, ic * f, rtol=1e-12)


@pytest.mark.gpu
def test_case02_gpu_against_reference_output(tmp_path):
    c = load02()
    gm, om, E = c["gm"], c["om"], c["E"]
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    p, h, v = g.element_output()
    fp = str(tmp_path / "test2.msh")
    write_flow(gm, fp, g.node_pressure(), p, v)
    T = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), c["bcc"], np.zeros((2, E)), 5.0, 0.5)
    T.set_velocity_field(darcy=g)
    T.set_dual_porosity_sorption(True, True, **PHASES)
    tp, mp = str(tmp_path / "transport.msh"), str(tmp_path / "mass_balance.txt")
    run_transport(T, T.update_solution, T.conc, gm, tp, mp, lambda path: T.write_mass_balance(path, NAMES))
    compare(tmp_path, fp, tp, mp)


@pytest.mark.gpu
@pytest.mark.parametrize("sorp_type", [1, 2, 3])
def test_phases_gpu_against_oracle(tmp_path, sorp_type):
    """Generated 1D/2D/3D mesh with real flow; three substances, dual porosity,
    sorption of one isotherm type and a decay chain 0 -> 1 -> 2 that acts on the
    mobile and immobile phases; every phase against the oracle's."""
    path = str(tmp_path / "cube4.msh")
    cases.generated_mesh("cube", 4, path, n_fractures=2, n_channels=4, channel_len=4, seed=8)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=2)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-14)
    flux = D0.solution[:D0.S]
    E, NB, ns = om.n_elements, om.n_bc_elements, 3
    rng = np.random.default_rng(sorp_type)
    bcc = rng.uniform(0.5, 1.0, (ns, NB))
    ic = rng.uniform(0.0, 0.3, (ns, E))
    por = rng.uniform(0.2, 0.3, E)
    kw = dict(por_imm=rng.uniform(0.1, 0.2, E), alpha=rng.uniform(0.1, 1.0, (ns, E)), phi=rng.uniform(0.3, 0.7, E),
              sorp_type=sorp_type, sorp_coef0=rng.uniform(0.05, 0.5, (ns, E)),
              sorp_coef1=(rng.uniform(0.4, 0.8, (ns, E)) if sorp_type == 2 else rng.uniform(0.5, 2.0, (ns, E))))
    decays = [dict(parent=0, half_life=0.7, products=[1]), dict(parent=1, half_life=1.3, products=[2])]
    T0 = O.Transport(om, flux, d["cs"], por, bcc, ic, 1.0, 0.25)
    T0.set_phases(1, 1, **kw)
    T0.set_reactions([0, 1], [[1], [2]], [0.7, 1.3], [[1.0], [1.0]])
    T1 = G.ConvectionTransport(gm, por, d["cs"], bcc, ic, 1.0, 0.25)
    T1.set_velocity_field(side_flux=flux)
    T1.set_dual_porosity_sorption(True, True, **kw)
    T1.set_linear_reactions(decays)
    for t in (0.25, 0.5, 0.75, 1.0):
        assert T0.update(t) == T1.update_solution(t)
        for k in range(4):
            a, b = T1.conc_phase(k), T0.conc_phase(k)
            assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-300), (t, k)
    m0, m1 = T0.mass_balance(), T1.mass_balance()
    assert np.allclose(m1["totals"], m0["totals"], rtol=1e-11, atol=1e-15)
