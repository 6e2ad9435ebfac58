"""tests/12_radioactive_decay of the reference from its input files: six substances
transported on a three-tetrahedron mesh with the legacy .fbc/.tbc boundary files
and first-order reactions applied after every convection step
(Linear_reaction, linear_reaction.cc:144-300, driven as
transport_operator_splitting.cc:244-287):
  * flow_decay_gmsh.con: decay chain D -> F -> B, half-lives 2.5;
  * flow_lin_react_gmsh.con: D -> E with kinetic constant 0.277258872.
Checked against ref_output/*/transport_*.msh (every substance at the 21 output
times, 6 significant digits) and, written through the MSH writer, byte for
byte against those files.  GPU tests: the same from the GPU flow and transport,
and the GPU step with reactions bitwise equal to the oracle, also on a
generated mixed-dimensional mesh with real flow and branched decays."""
import math

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

NAMES = list("ABCDEF")
INIT = [0.01, 0.02, 0.03, 0.04, 0.05, 0.06]
CASES = {
    "decay": ("test12_transport_decay.msh",
              [dict(parent="D", half_life=2.5, products=["F"]), dict(parent="F", half_life=2.5, products=["B"])]),
    "lin_react": ("test12_transport_lin_react.msh", [dict(parent="D", kinetic=0.277258872, products=["E"])]),
}


def load_case12():
    ref = cases.ref_path
    gm = G.Mesh(ref("test12.msh"))
    b = G.OldBcdInput(gm)
    bc = b.read_flow(ref("test12.fbc"))
    bcc = b.read_transport(ref("test12.tbc"), 6)
    E = gm.n_elements
    aniso = np.zeros((E, 9))
    aniso[:, [0, 4, 8]] = 0.01  # flow_*.con: anisotropy 0.01 on region 1
    init = np.repeat(np.array(INIT)[:, None], E, axis=1)
    return dict(gm=gm, bc=bc, bcc=bcc, aniso=aniso, init=init, E=E)


def oracle_reactions(decays):
    parent, prods, hl, br = [], [], [], []
    for d in decays:
        parent.append(NAMES.index(d["parent"]))
        prods.append([NAMES.index(p) for p in d["products"]])
        hl.append(d["half_life"] if "half_life" in d else math.log(2) / d["kinetic"])
        br.append(d.get("branch_ratios", [1.0]))
    return parent, prods, hl, br


def read_golden(name):
    out = {}
    for s in open(cases.ref_path(name)).read().split("$ElementData\n")[1:]:
        L = s.split("\n")
        out[(L[1].strip('"').split("_")[0], round(float(L[3]), 6))] = np.array(
            [float(l.split()[1]) for l in L[8:8 + int(L[7])]])
    return out


def run_and_write(T, conc, update, gm, path):
    """Steps to t = 0, 0.5, .., 10, writing every substance after each output time
    as the reference's transport output stream does."""
    out = G.OutputMSH(gm, path)
    hist = []
    for k in range(21):
        update(0.5 * k)
        c = conc()
        hist.append(c.copy())
        out.write_data(0.5 * k, [("%s_mobile" % n, "M/L^3", c[i]) for i, n in enumerate(NAMES)])
    return hist


@pytest.mark.parametrize("case", sorted(CASES))
def test_case12_oracle_against_reference_output(tmp_path, case):
    golden, decays = CASES[case]
    c = load_case12()
    om = O.Mesh(cases.ref_path("test12.msh"))
    E = c["E"]
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    D.solve(0, rtol=1e-15)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.ones(E), c["bcc"], c["init"], 10.0, 0.5)
    T.set_reactions(*oracle_reactions(decays))
    path = str(tmp_path / "transport.msh")
    hist = run_and_write(T, T.conc, T.update, c["gm"], path)
    gold = read_golden(golden)
    for k, cc in enumerate(hist):
        for i, n in enumerate(NAMES):
            g = gold[(n, round(0.5 * k, 6))]
            assert np.all(np.abs(cc[i] - g) <= 5e-6 * np.abs(g)), (case, k, n, cc[i], g)
    # the chain conserves the total of the substances it links (no flow here)
    assert abs(hist[-1][[1, 3, 5]].sum() - hist[0][[1, 3, 5]].sum()) < 1e-15 * 3 * E or case != "decay"
    # no flow: D is multiplied by 0.5^(dt / half_life) in each of the 20 steps of 0.5
    hl = oracle_reactions(decays)[2][0]
    assert abs(hist[-1][3, 0] - 0.04 * (0.5 ** (0.5 / hl)) ** 20) < 1e-15
    assert open(path).read() == open(cases.ref_path(golden)).read()


def test_reaction_input_errors():
    c = load_case12()
    # creating a transport handle needs the GPU; the argument checks of the binding do not
    with pytest.raises((G.FlowError, ValueError)):
        T = G.ConvectionTransport.__new__(G.ConvectionTransport)
        T.h = None
        T.set_linear_reactions([dict(parent="D", half_life=1.0, products=["E", "F"], branch_ratios=[1.0])], NAMES)
    with pytest.raises(ValueError):
        T = G.ConvectionTransport.__new__(G.ConvectionTransport)
        T.h = None
        T.set_linear_reactions([dict(parent="X", half_life=1.0, products=["E"])], NAMES)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_case12_gpu_against_reference_output(tmp_path, case):
    golden, decays = CASES[case]
    c = load_case12()
    gm, E = c["gm"], c["E"]
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    T = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["bcc"], c["init"], 10.0, 0.5)
    T.set_velocity_field(darcy=g)
    T.set_linear_reactions(decays, NAMES)
    hist = run_and_write(T, T.conc, T.update_solution, gm, str(tmp_path / "t.msh"))
    gold = read_golden(golden)
    for k, cc in enumerate(hist):
        for i, n in enumerate(NAMES):
            gv = gold[(n, round(0.5 * k, 6))]
            assert np.all(np.abs(cc[i] - gv) <= 5e-6 * np.abs(gv)), (case, k, n)
    # the same steps on the oracle's fluxes: bitwise equal
    om = O.Mesh(cases.ref_path("test12.msh"))
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    D.solve(0, rtol=1e-15)
    flux = D.solution[:D.S]
    T0 = O.Transport(om, flux, np.ones(E), np.ones(E), c["bcc"], c["init"], 10.0, 0.5)
    T0.set_reactions(*oracle_reactions(decays))
    T1 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["bcc"], c["init"], 10.0, 0.5)
    T1.set_velocity_field(side_flux=flux)
    T1.set_linear_reactions(decays, NAMES)
    for k in range(1, 21):
        assert T0.update(0.5 * k) == T1.update_solution(0.5 * k)
        assert np.array_equal(T0.conc(), T1.conc()), k


@pytest.mark.gpu
@pytest.mark.parametrize("nsub", [2, 6, 8])
def test_reactions_with_flow_generated_mesh(tmp_path, nsub):
    """Real flow on a mixed 1D/2D/3D mesh, nsub substances, a branched decay
    (0 -> 1 (0.3) + 2 (0.7)) and a chain (2 -> 3): bitwise equal to the oracle,
    with mass balance totals equal."""
    path = str(tmp_path / "cube4.msh")
    cases.generated_mesh("cube", 4, path, n_fractures=2, n_channels=4, channel_len=4, seed=8)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=2)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-14)
    flux = D0.solution[:D0.S]
    E, NB = om.n_elements, om.n_bc_elements
    bcc = np.ones((nsub, NB)) * (1.0 + np.arange(nsub))[:, None] / nsub
    ic = np.zeros((nsub, E))
    ic[0] = 0.5
    por = np.full(E, 0.3)
    if nsub == 2:
        decays = [dict(parent=0, half_life=0.3, products=[1])]
    else:
        decays = [dict(parent=0, half_life=0.4, products=[1, 2], branch_ratios=[0.3, 0.7]),
                  dict(parent=2, half_life=1.1, products=[3])]
    parent = [dd["parent"] for dd in decays]
    prods = [dd["products"] for dd in decays]
    hl = [dd["half_life"] for dd in decays]
    br = [dd.get("branch_ratios", [1.0]) for dd in decays]
    T0 = O.Transport(om, flux, d["cs"], por, bcc, ic, 1.0, 0.25)
    T0.set_reactions(parent, prods, hl, br)
    T1 = G.ConvectionTransport(gm, por, d["cs"], bcc, ic, 1.0, 0.25)
    T1.set_velocity_field(side_flux=flux)
    T1.set_linear_reactions(decays)
    for t in (0.25, 0.5, 0.75, 1.0):
        assert T0.update(t) == T1.update_solution(t)
        assert np.array_equal(T0.conc(), T1.conc()), t
    m0, m1 = T0.mass_balance(), T1.mass_balance()
    assert np.allclose(m1["totals"], m0["totals"], rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
def test_cpp_adapter_decay_chain(tmp_path):
    """examples/decay_chain.cpp: the same case through the C++ adapter (legacy files,
    steady flow, reactions, MSH output stream); its output file equals the
    reference's transport_decay.msh byte for byte."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "examples", "decay_chain")
    assert os.path.exists(exe), "build() compiles the example"
    out, mb = str(tmp_path / "t.msh"), str(tmp_path / "mb.txt")
    r = subprocess.run([exe, cases.ref_path("test12.msh"), cases.ref_path("test12.fbc"), cases.ref_path("test12.tbc"),
                        out, mb], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "decay chain:" in r.stdout
    assert open(out).read() == open(cases.ref_path("test12_transport_decay.msh")).read()
    assert open(mb).read().count("MASS BALANCE") == 21
