"""tests/20_limited_sorption of the reference (test_20.con), the part without the
reaction module's sorption: steady flow on a 5 x 1 rectangle of 1306 triangles
(conductivity 0.125, Dirichlet p = x at the boundary element centres) and
transport with boundary concentration 0.05 t on every boundary, a formula of time
that the reference re-evaluates at the start of every step
(FieldFormula::set_time, field_formula_impl.hh:70-122; compute_one_step,
transport.cc:598-627), to t = 15 with output every 0.5.  Substance A (sorption
type none; B and C equal it in the reference's output) against
ref_output/test_20.con/transport_test_20 at t = 0, 2.5, .., 15 (extracted by
tests/golden/extract_test20.py).

The reference's flow solve stopped at a_tol 1e-7; a relative flux perturbation
of 1e-6 moves the concentrations by 6e-6, and the agreement is 8.1e-6 on the
scale 0.75 (test bound 2e-5).  Boundary data taken at the end of each step
instead of its start would be 2e-3 off.  GPU: the same, and bitwise equal to the
oracle on the oracle's fluxes."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O


def load20():
    om = O.Mesh(cases.ref_path("test20.msh"))
    E, NB = om.n_elements, om.n_bc_elements
    cen = om.bc_centre.reshape(-1, 3)[:NB]
    bc = (np.ones(NB, np.int32), cen[:, 0].copy(), np.zeros(NB), np.zeros(NB))
    gold = {round(float(l.split()[0]), 6): np.array(l.split()[1:], float)
            for l in open(cases.ref_path("test20_A_mobile.txt"))}
    return om, E, NB, bc, gold


def bc_fn(NB):
    return lambda t: np.full((1, NB), 0.05 * t)


def run(update, conc, gold):
    out = {}
    for k in range(31):  # the reference's output times
        update(0.5 * k)
        t = round(0.5 * k, 6)
        if t in gold:
            out[t] = conc()[0].copy()
    return out


def check(out, gold):
    assert sorted(out) == sorted(gold)
    for t in gold:
        assert np.max(np.abs(out[t] - gold[t])) < 2e-5, t


def test_case20_oracle_against_reference_output():
    om, E, NB, bc, gold = load20()
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, 0.25), np.zeros((1, NB)), np.zeros((1, E)),
                    15.0, 0.5)
    T.set_bc_function(bc_fn(NB))
    check(run(T.update, T.conc, gold), gold)


@pytest.mark.gpu
def test_case20_gpu_against_reference_output():
    om, E, NB, bc, gold = load20()
    gm = G.Mesh(cases.ref_path("test20.msh"))
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    T = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), np.zeros((1, NB)), np.zeros((1, E)), 15.0, 0.5)
    T.set_velocity_field(darcy=g)
    T.set_bc_conc_function(bc_fn(NB))
    check(run(T.update_solution, T.conc, gold), gold)
    # bitwise against the oracle on the oracle's fluxes
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    T0 = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, 0.25), np.zeros((1, NB)), np.zeros((1, E)),
                     15.0, 0.5)
    T0.set_bc_function(bc_fn(NB))
    T1 = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), np.zeros((1, NB)), np.zeros((1, E)), 15.0, 0.5)
    T1.set_velocity_field(side_flux=D.solution[:D.S])
    T1.set_bc_conc_function(bc_fn(NB))
    for k in range(1, 11):
        assert T0.update(0.5 * k) == T1.update_solution(0.5 * k)
        assert np.array_equal(T0.conc(), T1.conc()), k
