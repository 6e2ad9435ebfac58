"""tests/16_concentration_sources of the reference, end to end from its own input
files with no hand conversion: the mesh (test16.msh), the legacy flow and
transport boundary files (test16.fbc, test16.tbc), the three concentration-source
fields from `$ElementData` sections (test16_sources.msh), the steady flow, then
explicit upwind transport with sources to t = 5.7 with output every 0.8
(flow.con).  Checked against ref_output/flow.con/mass_balance.txt at the eight
output times after t = 0 (6 significant digits) and water_balance.txt.  The
t = 0 block is not compared: its source column is evaluated before the first
step, which this interface does not report."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

TIMES = [0.8, 1.6, 2.4, 3.2, 4.0, 4.8, 5.6, 5.7]


def load_case16():
    ref = cases.ref_path
    gm = G.Mesh(ref("test16.msh"))
    b = G.OldBcdInput(gm)
    bt, bp, bf, bs = b.read_flow(ref("test16.fbc"))
    bcc = b.read_transport(ref("test16.tbc"), 1)
    src = []
    for name in ("density", "sigma", "concentration"):
        f = G.FieldElementwise(gm, ref("test16_sources.msh"), name)
        assert f.set_time(0.0)
        src.append(f.value())
    E = gm.n_elements
    # flow.con: cross_section 1, anisotropy 1 on region 7; por_m 1, init_conc 0
    return dict(gm=gm, K=np.ones(E), cs=np.ones(E), sigma=np.ones(E), f=np.zeros(E), bc=(bt, bp, bf, bs), bcc=bcc,
                src=src, por=np.ones(E), init=np.zeros(E))


def read_mass_balance(path):
    """Per output time: boundary (total, outward, inward), region (mass, source) and
    cumulative (A init, B sources, C flux, A+B-C, D current) rows of mass_balance.txt."""
    out, t = {}, None
    for line in open(path):
        if line.startswith("# Time:"):
            t = round(float(line.split()[2]), 6)
            out[t] = {}
        elif line.startswith("  -3 "):
            out[t]["flux"] = [float(v) for v in line.split()[4:7]]
        elif line.startswith("  7 "):
            out[t]["region"] = [float(v) for v in line.split()[3:5]]
        elif line.startswith("  conc "):
            out[t]["cumul"] = [float(v) for v in line.split()[1:6]]
    return out


def _close(a, b):
    return abs(a - b) <= 2e-5 * abs(b) + 1e-10


def check_against_golden(mb, t, gold):
    g = gold[round(t, 6)]
    tot = mb["totals"][0]  # initial mass, integrated flux, integrated sources, current mass
    got = {"flux": [mb["bal"][0, 0], mb["plus"][0, 0], mb["minus"][0, 0]],
           "region": [mb["mass"][0, 0], mb["src"][0, 0]],
           "cumul": [tot[0], tot[2], tot[1], tot[0] + tot[2] - tot[1], tot[3]]}
    for k in ("flux", "region", "cumul"):
        for a, b in zip(got[k], g[k]):
            assert _close(a, b), (t, k, got[k], g[k])


def test_case16_boundary_files_match_the_case_description():
    """All 136 .fbc rows are Dirichlet conditions on distinct boundary elements; the
    .tbc concentrations are zero (flow.con: bc_conc = 0); each source field has one
    non-zero element."""
    c = load_case16()
    bt, bp, _, _ = c["bc"]
    assert (bt == 1).sum() == 136 == c["gm"].n_bc_elements
    assert np.count_nonzero(c["bcc"]) == 0
    assert [int(np.count_nonzero(s)) for s in c["src"]] == [1, 1, 1]


def test_case16_oracle_against_reference_output():
    c = load_case16()
    om = O.Mesh(cases.ref_path("test16.msh"))
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    D.solve(0, rtol=1e-15)
    bal, out, inn = D.water_balance()
    # water_balance.txt: outflow 2, inflow -2, balance -1.35259e-06 (an iterative solve to r_tol 1e-12)
    assert abs(out[0] - 2.0) < 1e-5 and abs(inn[0] + 2.0) < 1e-5 and abs(bal[0]) < 2e-6
    T = O.Transport(om, D.solution[:D.S], c["cs"], c["por"], c["bcc"], c["init"], 5.7, 0.8)
    T.set_sources(*c["src"])
    gold = read_mass_balance(cases.ref_path("test16_mass_balance.txt"))
    for t in TIMES:
        T.update(t)
        check_against_golden(T.mass_balance(), t, gold)


@pytest.mark.gpu
def test_case16_gpu_against_reference_output():
    """GPU flow -> GPU transport with sources; and the GPU transport on the oracle's
    fluxes bitwise equal to the oracle's concentrations."""
    c = load_case16()
    gm = c["gm"]
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    bal, out, inn = g.water_balance()
    assert abs(out[0] - 2.0) < 1e-5 and abs(inn[0] + 2.0) < 1e-5 and abs(bal[0]) < 2e-6
    T = G.ConvectionTransport(gm, c["por"], c["cs"], c["bcc"], c["init"], 5.7, 0.8)
    T.set_velocity_field(darcy=g)
    T.set_sources(*c["src"])
    gold = read_mass_balance(cases.ref_path("test16_mass_balance.txt"))
    for t in TIMES:
        T.update_solution(t)
        check_against_golden(T.mass_balance(), t, gold)

    om = O.Mesh(cases.ref_path("test16.msh"))
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"])
    D.solve(0, rtol=1e-15)
    flux = D.solution[:D.S]
    T0 = O.Transport(om, flux, c["cs"], c["por"], c["bcc"], c["init"], 5.7, 0.8)
    T0.set_sources(*c["src"])
    T1 = G.ConvectionTransport(gm, c["por"], c["cs"], c["bcc"], c["init"], 5.7, 0.8)
    T1.set_velocity_field(side_flux=flux)
    T1.set_sources(*c["src"])
    for t in TIMES:
        assert T0.update(t) == T1.update_solution(t)
        assert np.array_equal(T0.conc(), T1.conc()), t
