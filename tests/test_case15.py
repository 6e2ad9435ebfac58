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


@pytest.mark.gpu
def test_case15_gpu_coupling_loop():
    """test15 driven through the equations' time interface exactly as
    HC_ExplicitSequential::run_simulation (hc_explicit_sequential.cc:186-231)
    does it: the transport's upper step constraint is the water's dt, the
    water solves while its solved time is below the velocity interpolation
    time (theta 0.5 between the transport's solved and planned times), the
    transport takes its governor's steps on the last velocity field.  Outputs
    at every save_step mark against the reference's."""
    c = load_case15()
    E, gm = c["E"], c["gm"]
    S = gm.n_sides
    water = G.DarcyFlowMH_Steady(gm)
    water.set_data(np.ones(E), np.ones(E), np.ones(E), c["f"], *c["bc"])
    water.set_unsteady(True, np.ones(E), np.zeros(E))
    water.set_time(2.0, DT, 0.1)
    assert water.time() == (0.0, DT, False)
    trans = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["tbc"][0], np.zeros(E), 2.0, 0.1)
    theta = 0.5
    velocity_changed, n_water, n_trans, checked, bc_switched = False, 0, 0, 0, False
    while not (water.time()[2] and trans.time()[2]):
        trans.set_time_upper_constraint(water.time()[1])
        vit = theta * trans.planned_time() + (1 - theta) * trans.solved_time()
        if water.solved_time() < vit:
            water.update_solution(r_tol=1e-15, a_tol=0.0, max_it=100000)
            water.choose_next_time()
            velocity_changed = True
            n_water += 1
        else:
            if velocity_changed:
                trans.set_velocity_field(darcy=water)
                velocity_changed = False
            trans.update_solution()
            n_trans += 1
            t = trans.solved_time()
            k = int(round(t / 0.1))
            if abs(t - 0.1 * k) <= 1e-9:
                x = water.get_solution_vector()
                check_output(c, k, x[S:S + E], trans.conc()[0])
                checked += 1
                if k == 1 and not bc_switched:
                    trans.set_bc_conc(c["tbc"][1])
                    bc_switched = True
        assert n_water + n_trans <= 2 * STEPS + 20
    print("water steps %d, transport steps %d, outputs checked %d" % (n_water, n_trans, checked), flush=True)
    # the transport governor's step: the water's dt as its upper constraint,
    # rounded to whole steps until its next mark (time_governor.cc:261-300); the
    # water's fixed step drifts by ulps near t = 2, so the last interval before
    # the end mark is cut into one more step
    assert n_water == STEPS and STEPS <= n_trans <= STEPS + 8 and checked == 20
    assert abs(water.solved_time() - 2.0) < 1e-9 and abs(trans.solved_time() - 2.0) < 1e-9


@pytest.mark.gpu
def test_steady_water_time_governor():
    """A steady Darcy handle carries the steady TimeGovernor
    (time_governor.cc:151-168, 261-330): t = 0, dt = inf, planned inf before
    the solve; t = inf, dt = 0 and at its end after it."""
    c = load_case15()
    E, gm = c["E"], c["gm"]
    w = G.DarcyFlowMH_Steady(gm)
    w.set_data(np.ones(E), np.ones(E), np.ones(E), c["f"], *c["bc"])
    assert w.time() == (0.0, np.inf, False)
    assert w.planned_time() == np.inf and w.solved_time() == 0.0
    w.set_time_upper_constraint(0.5)  # no effect on a steady governor's estimate
    w.choose_next_time()
    assert w.planned_time() == np.inf
    w.update_solution(r_tol=1e-12, a_tol=0.0, max_it=10000)
    assert w.time() == (np.inf, 0.0, True) and w.solved_time() == np.inf
    with pytest.raises(G.FlowError):
        w.set_time(2.0, 0.01, 0.1)  # an unsteady governor needs the unsteady form
