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
oracle on the oracle's fluxes.

The sorbing substances (the reaction module's Sorption with tabulated isotherms,
sorption.cc:71-285, isotherm.cc:33-258: As-lin linear 0.6, Bs-lang Langmuir
(0.4, 0.6), Cs-freund Freundlich (0.6, 0.4), solubility 1, 1000 table steps)
against the reference at t = 7.5 and 15: within 1.1e-5, 3.6e-5 and 1.1e-4; a flow
solved to the reference's 1e-7 instead of to round-off moves them by 1.4e-4,
1.4e-4 and 2.8e-4, which sets the bounds 5e-5, 1e-4, 3e-4.  The isotherm roots
come from bisection to the last bit where the reference brackets to 30 bits
(boost's TOMS 748), a 1e-9 difference.  GPU: every species within 1e-12 of the
oracle (the device pow)."""
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


SORB = dict(subst=[0, 1, 3, 5], types=[[0, 1, 3, 2]], mult=[[0, 0.6, 0.4, 0.6]], second=[[0, 0, 0.6, 0.4]],
            rock_density=1.0, molar_mass=[1.0] * 4, solubility=[1.0] * 4)
SORB_TOL = {"As-lin": 5e-5, "Bs-lang": 1e-4, "Cs-freund": 3e-4}
NAMES = ["A", "As-lin", "B", "Bs-lang", "C", "Cs-freund"]


def sorbing_gold():
    out = {}
    for l in open(cases.ref_path("test20_sorbing.txt")):
        w = l.split()
        out[(w[0], round(float(w[1]), 6))] = np.array(w[2:], float)
    return out


def run6(update, conc):
    res = {}
    for k in range(31):
        update(0.5 * k)
        if k in (15, 30):
            res[round(0.5 * k, 6)] = conc().copy()
    return res


def check6(res, gold, A):
    for (name, t), g in gold.items():
        assert np.max(np.abs(res[t][NAMES.index(name)] - g)) < SORB_TOL[name], (name, t)
    for t in res:
        for i in (0, 2, 4):
            assert np.max(np.abs(res[t][i] - A[t])) < 2e-5, t


def test_case20_sorption_oracle_against_reference_output():
    om, E, NB, bc, goldA = load20()
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.full(E, 0.25), np.zeros((6, NB)), np.zeros((6, E)),
                    15.0, 0.5)
    T.set_bc_function(lambda t: np.full((6, NB), 0.05 * t))
    T.set_reaction_sorption(**SORB)
    check6(run6(T.update, T.conc), sorbing_gold(), goldA)
    # the sorbed amount sits on the isotherm: linear s = 0.6 c for As-lin
    s = T.reaction_sorbed()
    c = T.conc()
    assert np.allclose(s[1], 0.6 * c[1], rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_case20_sorption_gpu():
    om, E, NB, bc, goldA = load20()
    gm = G.Mesh(cases.ref_path("test20.msh"))
    D = O.Darcy(om, np.full(E, 0.125), np.ones(E), np.ones(E), np.zeros(E), *bc)
    D.solve(0, rtol=1e-15)
    flux = D.solution[:D.S]
    T0 = O.Transport(om, flux, np.ones(E), np.full(E, 0.25), np.zeros((6, NB)), np.zeros((6, E)), 15.0, 0.5)
    T0.set_bc_function(lambda t: np.full((6, NB), 0.05 * t))
    T0.set_reaction_sorption(**SORB)
    T1 = G.ConvectionTransport(gm, np.full(E, 0.25), np.ones(E), np.zeros((6, NB)), np.zeros((6, E)), 15.0, 0.5)
    T1.set_velocity_field(side_flux=flux)
    T1.set_bc_conc_function(lambda t: np.full((6, NB), 0.05 * t))
    T1.set_reaction_sorption(**SORB)
    for k in range(31):
        assert T0.update(0.5 * k) == T1.update_solution(0.5 * k)
        a, b = T1.conc(), T0.conc()
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-300), k
    assert np.max(np.abs(T1.reaction_sorbed() - T0.reaction_sorbed())) <= 1e-12
    res = {round(t, 6): T1.conc() for t in (15.0,)}
    gold = {key: v for key, v in sorbing_gold().items() if key[1] == 15.0}
    check6(res, gold, goldA)
