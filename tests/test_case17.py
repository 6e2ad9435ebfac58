"""tests/17_decay_pade_approximation of the reference (flow_decay_vtk.con): the
test-12 mesh and boundary files with five substances and the decay chain
A -> B -> C -> D -> E, reactions by the (3, 3) Pade approximant of the matrix
exponential (Pade_approximant, pade_approximant.cc:97-290).  Checked against the
reference's VTK output (16 printed digits, extracted by
tests/golden/extract_test17.py) at all 21 output times; the step matrix against
an independent numpy evaluation of the same approximant; GPU against the oracle."""
import math

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_case12 import load_case12

NAMES = list("ABCDE")
DECAYS = [dict(parent="A", half_life=1.3863, products=["B"]), dict(parent="B", half_life=2.3105, products=["C"]),
          dict(parent="C", half_life=1.5403, products=["D"]), dict(parent="D", half_life=1.1552, products=["E"])]
INIT = [0.01, 0.02, 0.03, 0.04, 0.05]


def gold17():
    out = []
    for line in open(cases.ref_path("test17_transport_conc.txt")):
        v = np.array(line.split(), float)
        out.append((v[0], v[1:].reshape(5, 3)))
    return out


def pade_step_matrix(dt, p=3, q=3):
    """exp(A dt) by the (p, q) Pade approximant, numpy; c_new = P c."""
    n = 5
    A = np.zeros((n, n))
    for d in DECAYS:
        i = NAMES.index(d["parent"])
        r = dt / d["half_life"]
        A[i, i] = -math.log(2) * r
        for pr in d["products"]:
            A[i, NAMES.index(pr)] = math.log(2) * r
    f = math.factorial
    N = sum(f(p + q - j) * f(p) / (f(p + q) * f(j) * f(p - j)) * np.linalg.matrix_power(A, j) for j in range(p + 1))
    D = sum((-1) ** j * f(p + q - j) * f(q) / (f(p + q) * f(j) * f(q - j)) * np.linalg.matrix_power(A, j)
            for j in range(q + 1))
    return np.linalg.solve(D, N).T


def oracle_run(c, init, times):
    om = O.Mesh(cases.ref_path("test12.msh"))
    E = c["E"]
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    D.solve(0, rtol=1e-15)
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.ones(E), c["bcc"][:5], init, 10.0, 0.5)
    T.set_reactions([NAMES.index(d["parent"]) for d in DECAYS], [[NAMES.index(p) for p in d["products"]] for d in DECAYS],
                    [d["half_life"] for d in DECAYS], [[1.0]] * len(DECAYS))
    T.set_pade(3, 3)
    hist = []
    for t in times:
        T.update(t)
        hist.append(T.conc().copy())
    return hist, D.solution[:D.S]


def test_case17_oracle_against_reference_output():
    c = load_case12()
    E = c["E"]
    gold = gold17()
    init = np.repeat(np.array(INIT)[:, None], E, axis=1)
    hist, _ = oracle_run(c, init, [t for t, _ in gold])
    for (t, g), h in zip(gold, hist):
        assert np.allclose(h, g, rtol=1e-12, atol=1e-18), t


def test_pade_step_matrix_independent():
    """No flow (fluxes ~1e-17): one step of 0.5 from each unit vector is a column of P."""
    c = load_case12()
    E = c["E"]
    P = pade_step_matrix(0.5)
    for k in range(5):
        init = np.zeros((5, E))
        init[k] = 1.0
        hist, _ = oracle_run(c, init, [0.5])
        assert np.allclose(hist[0][:, 0], P[:, k], rtol=1e-13, atol=1e-15), k
    # the approximant is close to, not equal to, the exponential
    assert np.allclose(np.diag(P), [math.exp(-math.log(2) * 0.5 / d["half_life"]) for d in DECAYS] + [1.0],
                       rtol=1e-6)


@pytest.mark.gpu
def test_case17_gpu():
    c = load_case12()
    gm, E = c["gm"], c["E"]
    gold = gold17()
    init = np.repeat(np.array(INIT)[:, None], E, axis=1)
    hist0, flux = oracle_run(c, init, [t for t, _ in gold])
    T1 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["bcc"][:5], init, 10.0, 0.5)
    T1.set_velocity_field(side_flux=flux)
    T1.set_pade_reactions(DECAYS, NAMES, 3, 3)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), *c["bc"], aniso=c["aniso"])
    g.update_solution(r_tol=1e-14, a_tol=1e-14)
    T2 = G.ConvectionTransport(gm, np.ones(E), np.ones(E), c["bcc"][:5], init, 10.0, 0.5)
    T2.set_velocity_field(darcy=g)
    T2.set_pade_reactions(DECAYS, NAMES, 3, 3)
    for (t, gv), h0 in zip(gold, hist0):
        T1.update_solution(t)
        T2.update_solution(t)
        assert np.allclose(T1.conc(), h0, rtol=1e-13, atol=1e-18), t
        assert np.allclose(T2.conc(), gv, rtol=1e-12, atol=1e-18), t
