"""GPU parity of the explicit upwind transport step against the CPU oracle
and the reference's test 13 goldens (through the C ABI)."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

pytestmark = pytest.mark.gpu


def _flow13(om):
    d = cases.darcy_case("test13", om)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                 aniso=d["aniso"])
    D0.solve(0, rtol=1e-15)
    return d, D0


def read_test13_golden():
    txt = open(cases.ref_path("test13_transport.msh")).read().split("$ElementData")[1:]
    gold = {}
    for blk in txt:
        L = blk.strip().split("\n")
        t = float(L[3])
        n = int(L[7])
        gold[round(t, 6)] = np.array([float(l.split()[1]) for l in L[8:8 + n]])
    return gold


def test_transport_bitexact_vs_oracle():
    """Same fluxes in -> identical transport matrix and concentrations (bitwise)."""
    path = cases.ref_path("test13.msh")
    om, gm = O.Mesh(path), G.Mesh(path)
    d, D0 = _flow13(om)
    flux = D0.solution[:D0.S]
    E = om.n_elements
    bcc = cases.test13_bc_conc(om)
    T0 = O.Transport(om, flux, np.ones(E), np.ones(E), bcc, np.zeros(E), 100.0, 0.5)
    T1 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), bcc, np.zeros(E), 100.0, 0.5)
    T1.set_velocity_field(side_flux=flux)
    for k in range(1, 201):
        t = 0.5 * k
        s0, s1 = T0.update(t), T1.update_solution(t)
        assert s0 == s1
        c0, c1 = T0.conc(), T1.conc()
        assert np.array_equal(c0, c1), (t, np.abs(c0 - c1).max())
        if k == 1:
            p0, col0, v0 = T0.matrix()
            p1, col1, v1 = T1.matrix()
            assert np.array_equal(p0, p1) and np.array_equal(col0, col1) and np.array_equal(v0, v1)
    mb0, mb1 = T0.mass_balance(), T1.mass_balance()
    assert np.allclose(mb1["mass"], mb0["mass"], rtol=1e-13, atol=0)
    assert np.allclose(mb1["bal"], mb0["bal"], rtol=1e-12, atol=1e-15)
    assert abs(mb1["totals"][0, 1] - mb0["totals"][0, 1]) <= 1e-12 * abs(mb0["totals"][0, 1])


def test_transport_test13_golden():
    """GPU flow -> GPU transport against tests/13 transport.msh (%g) and mass_balance.txt."""
    path = cases.ref_path("test13.msh")
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("test13", om)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    E = gm.n_elements
    bcc = cases.test13_bc_conc(om)
    T = G.ConvectionTransport(gm, np.ones(E), np.ones(E), bcc, np.zeros(E), 100.0, 0.5)
    T.set_velocity_field(darcy=g)
    gold = read_test13_golden()
    worst = 0.0
    for k in range(1, 201):
        t = 0.5 * k
        T.update_solution(t)
        c = T.conc()[0]
        gv = gold[round(t, 6)]
        m = np.abs(gv) > 1e-4
        worst = max(worst, (np.abs(c - gv)[m] / np.abs(gv[m])).max() if m.any() else 0.0)
        assert np.abs(c - gv)[~m].max(initial=0.0) < 1e-9
    assert worst < 2e-5
    mb = T.mass_balance()
    assert abs(mb["bal"][0, 0] - (-0.0281415)) < 1e-6
    assert np.allclose(mb["mass"][0], [5.0, 6.72162, 7.07107], rtol=1e-5)
    tot = mb["totals"][0]
    # cumulative balance: initial + sources - integrated flux - current mass ~ 0 (reference: rel. err 1.45e-13)
    err = tot[0] + tot[2] - tot[1] - tot[3]
    assert abs(err) < 1e-11 * abs(tot[3])


def test_transport_generated_multisubstance(tmp_path):
    """4 substances on a mixed 1D/2D/3D mesh: bitwise equal to the oracle for the same fluxes."""
    path = str(tmp_path / "cube5.msh")
    cases.generated_mesh("cube", 5, path, n_fractures=2, n_channels=6, channel_len=4, seed=5)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=4)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-14)
    flux = D0.solution[:D0.S]
    E, NB = om.n_elements, om.n_bc_elements
    bcc = np.ones((4, NB)) * np.array([1.0, 0.5, 0.25, 2.0])[:, None]
    ic = np.zeros((4, E))
    por = np.full(E, 0.3)
    T0 = O.Transport(om, flux, d["cs"], por, bcc, ic, 1.0, 0.25)
    T1 = G.ConvectionTransport(gm, por, d["cs"], bcc, ic, 1.0, 0.25)
    T1.set_velocity_field(side_flux=flux)
    for t in (0.25, 0.5, 0.75, 1.0):
        assert T0.update(t) == T1.update_solution(t)
        assert np.array_equal(T0.conc(), T1.conc())


def test_transport_sources_bitexact(tmp_path):
    """Concentration sources (transport.cc:501-544): GPU equals the oracle bitwise;
    with por = cs = 1 the cumulative balance initial + sources - flux = mass closes."""
    path = str(tmp_path / "sq12.msh")
    cases.generated_mesh("square", 12, path)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("square", om, seed=8)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-14)
    flux = D0.solution[:D0.S]
    E, NB = om.n_elements, om.n_bc_elements
    rng = np.random.default_rng(1)
    ns = 3
    dens, sig, sc = rng.uniform(0, 0.2, (ns, E)), rng.uniform(0, 2.0, (ns, E)), rng.uniform(0, 1.5, (ns, E))
    bcc = np.zeros((ns, NB))
    ic = rng.uniform(0, 1, (ns, E))
    one = np.ones(E)
    T0 = O.Transport(om, flux, one, one, bcc, ic, 0.6, 0.2)
    T0.set_sources(dens, sig, sc)
    T1 = G.ConvectionTransport(gm, one, one, bcc, ic, 0.6, 0.2)
    T1.set_velocity_field(side_flux=flux)
    T1.set_sources(dens, sig, sc)
    for t in (0.2, 0.4, 0.6):
        assert T0.update(t) == T1.update_solution(t)
        assert np.array_equal(T0.conc(), T1.conc())
    mb0, mb1 = T0.mass_balance(), T1.mass_balance()
    for k in range(ns):
        a, flx, src, cur = mb1["totals"][k]
        assert abs(a + src - flx - cur) < 1e-12 * max(abs(cur), 1.0)
        assert np.allclose(mb1["totals"][k], mb0["totals"][k], rtol=1e-12, atol=1e-14)
    assert np.allclose(mb1["src"], mb0["src"], rtol=1e-12, atol=1e-14)
