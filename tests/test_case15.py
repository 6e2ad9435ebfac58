"""tests/15_unsteady_and_transport of the reference (flow_trans_explicit.con):
unsteady lumped MH flow (source -x, dt 0.01 to t = 2) coupled to explicit
transport as HC_ExplicitSequential::run_simulation does it
(hc_explicit_sequential.cc:151-240): the water runs one step ahead and each
transport interval [t, t + dt] uses the flux field of t + dt (velocity
interpolation time between the transport's solved and planned times, theta
0.5); the transport rebuilds its matrix from every new field; boundary data
from test15_tbc_000, then test15_tbc_001 from t = 0.1.

Checked against the reference's VTK output (element pressures and concentrations
at t = 0, 0.1, .., 2.0, extracted by tests/golden/extract_test15.py).  The
reference's flow solver stopped at accuracy 1e-7, so agreement is to 1e-5
(pressure, relative) and 2e-4 (concentration, absolute, boundary value 10);
coupling with the start-of-interval flux instead misses by 3e-2 at t = 0.5.
GPU: flow against the oracle, transport bitwise against the oracle on the
oracle's fluxes, and the whole GPU coupling against the reference's output."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

DT, STEPS = 0.01, 200


def load_case15():
    ref = cases.ref_path
    gm = G.Mesh(ref("test15.msh"))
    om = O.Mesh(ref("test15.msh"))
    b = G.OldBcdInput(gm)
    bc = b.read_flow(ref("test15.fbc"))
    tbc = [b.read_transport(ref("test15_tbc_%03d" % k), 1) for k in range(2)]
    E = gm.n_elements
    f = -om.centre.reshape(-1, 3)[:E, 0]  # water_source_density = -x at the element centre
    gp = [np.array(l.split(), float) for l in open(ref("test15_flow_pressure.txt"))]
    gc = [np.array(l.split(), float) for l in open(ref("test15_transport_conc.txt"))]
    return dict(gm=gm, om=om, bc=bc, tbc=tbc, f=f, E=E, gp=gp, gc=gc)


def check_output(c, j, p, conc):
    gp, gc = c["gp"][j][1:], c["gc"][j][1:]
    assert np.max(np.abs(p - gp)) <= 1e-5 * max(np.max(np.abs(gp)), 1e-300), j
    assert np.max(np.abs(conc - gc)) <= 2e-4, j


def test_case15_oracle_against_reference_output():
    c = load_case15()
    E, om = c["E"], c["om"]
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), c["f"], *c["bc"])
    D.set_unsteady(1, np.ones(E), np.zeros(E), DT)
    check_output(c, 0, D.solution[D.S:D.S + E], np.zeros(E))
    T = None
    for k in range(1, STEPS + 1):
        D.unsteady_step(DT, rtol=1e-15)
        flux = D.solution[:D.S].copy()
        if T is None:
            T = O.Transport(om, flux, np.ones(E), np.ones(E), c["tbc"][0], np.zeros(E), 2.0, 0.1)
        else:
            T.set_velocity(flux)
        T.update(DT * k)
        if k % 10 == 0:
            check_output(c, k // 10, D.solution[D.S:D.S + E], T.conc()[0])
        if k == 10:
            T.set_bc_conc(c["tbc"][1])


@pytest.mark.gpu
def test_case15_gpu():
    c = load_case15()
    E, om, gm = c["E"], c["om"], c["gm"]
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), c["f"], *c["bc"])
    D.set_unsteady(1, np.ones(E), np.zeros(E), DT)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), c["f"], *c["bc"])
    g.set_unsteady(True, np.ones(E), np.zeros(E))
    T0 = T1 = T2 = None
    S = gm.n_sides
    for k in range(1, STEPS + 1):
        D.unsteady_step(DT, rtol=1e-15)
        g.unsteady_step(DT, r_tol=1e-15, a_tol=0.0, max_it=100000)
        x = g.get_solution_vector()
        assert np.linalg.norm(x - D.solution) <= 1e-9 * np.linalg.norm(D.solution), k
        flux = D.solution[:D.S].copy()
        if T0 is None:
            T0 = O.Transport(om, flux, np.ones(E), np.ones(E), c["tbc"][0], np.zeros(E), 2.0, 0.1)
            T1 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["tbc"][0], np.zeros(E), 2.0, 0.1)
            T2 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["tbc"][0], np.zeros(E), 2.0, 0.1)
        else:
            T0.set_velocity(flux)
        T1.set_velocity_field(side_flux=flux)
        T2.set_velocity_field(darcy=g)
        assert T0.update(DT * k) == T1.update_solution(DT * k)
        T2.update_solution(DT * k)
        assert np.array_equal(T0.conc(), T1.conc()), k
        if k % 10 == 0:
            check_output(c, k // 10, x[S:S + E], T2.conc()[0])
        if k == 10:
            T0.set_bc_conc(c["tbc"][1])
            T1.set_bc_conc(c["tbc"][1])
            T2.set_bc_conc(c["tbc"][1])
