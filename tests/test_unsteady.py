"""Unsteady Darcy flow, mixed-hybrid (Unsteady_MH) and lumped mixed-hybrid
(Unsteady_LMH), darcy_flow_mh.cc:1601-1827: implicit Euler steps on the same
condensation as the steady solve.

Pins of the oracle's restatement (CPU):
  * tests/10_unsteady_2d (flow_LMH.con) of the reference: element pressures at
    t = 0.1 .. 0.5 (ref_output test10-00000k.vtu, extracted by
    tests/golden/extract_test10.py) and water_balance.txt;
  * MH: every step satisfies the element storage balance
    sum of side fluxes + S|e|/dt (p - p_prev) = |e| cs f;
  * zero storativity gives the steady solution in one step (MH and LMH);
  * the MH element pressures approach the series solution of the heat equation
    that tests/10 documents (exact_solution.py).
GPU (through the C ABI): MH and LMH steps against the oracle's, including a
time-step change, vb coupling and sources; LMH against the test-10 files."""
import math

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_balance_files import ndiff


def case10(om):
    E, NB = om.n_elements, om.n_bc_elements
    lab = {r["id"]: r["label"] for r in om.regions}
    bct, bcp = np.zeros(NB, np.int32), np.zeros(NB)
    for b, r in enumerate(om.bc_region):
        if lab[int(r)] == ".bc_inflow":
            bct[b], bcp[b] = 1, 100.0
        elif lab[int(r)] == ".bc_outflow":
            bct[b], bcp[b] = 1, 0.0
    aniso = np.zeros((E, 9))
    aniso[:, [0, 4, 8]] = 0.02  # flow_*.con: anisotropy 0.02, storativity 1, init_pressure 0
    return dict(K=np.ones(E), cs=np.ones(E), sigma=np.ones(E), f=np.zeros(E), bc=(bct, bcp, np.zeros(NB), np.zeros(NB)),
                aniso=aniso, stor=np.ones(E), init=np.zeros(E))


def gold10():
    return [np.array(l.split(), float) for l in open(cases.ref_path("test10_lmh_pressure.txt"))]


def test_lmh_oracle_against_test10(tmp_path):
    om = O.Mesh(cases.ref_path("test10.msh"))
    gm = G.Mesh(cases.ref_path("test10.msh"))
    c = case10(om)
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"], aniso=c["aniso"])
    D.set_unsteady(1, c["stor"], c["init"], 0.01)
    gold = gold10()
    p = str(tmp_path / "water_balance.txt")
    G.write_water_balance(gm, p, 0.0, *D.water_balance(), np.zeros(1))
    for k in range(1, 51):
        D.unsteady_step(0.01, rtol=1e-15)
        if k % 10 == 0:
            g = gold[k // 10]
            pe = D.solution[D.S:D.S + D.E]
            # the reference stopped CG at a_tol 1e-7: agreement to 1e-7 of the pressure scale 100
            assert np.max(np.abs(pe - g[1:])) < 1e-5, k
            G.write_water_balance(gm, p, 0.01 * k, *D.water_balance(), np.zeros(1))
    ndiff(open(p).read(), open(cases.ref_path("test10_lmh_water_balance.txt")).read(), 1e-5, 1e-6)


@pytest.mark.parametrize("lumped", [0, 1])
def test_zero_storativity_is_steady(lumped):
    path = cases.ref_path("test6_32d.msh")
    om = O.Mesh(path)
    d = cases.darcy_case("test06_32", om)
    args = (om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0 = O.Darcy(*args)
    D0.solve(0, rtol=1e-15)
    D1 = O.Darcy(*args)
    D1.set_unsteady(lumped, np.zeros(om.n_elements), np.full(om.n_elements, 3.0), 0.1)
    D1.unsteady_step(0.1, rtol=1e-15)
    if lumped:
        # LMH moves the element source into the side fluxes after the solve (darcy_flow_mh.cc:1807-1826)
        e_src = np.array([om.measure[e] * d["cs"][e] * d["f"][e] / (om.elem_dim[e] + 1) for e in range(om.n_elements)])
        q = D1.solution[:D1.S].copy()
        off = np.concatenate([[0], np.cumsum(om.elem_dim + 1)])
        for e in range(om.n_elements):
            q[off[e]:off[e + 1]] += e_src[e]
        assert np.allclose(q, D0.solution[:D0.S], rtol=1e-10, atol=1e-12)
        assert np.allclose(D1.solution[D1.S:], D0.solution[D0.S:], rtol=1e-10, atol=1e-12)
    else:
        assert np.allclose(D1.solution, D0.solution, rtol=1e-10, atol=1e-12)


def test_mh_element_storage_balance():
    """Each MH step: sum_i q_i + (vb terms) - |e| cs f + S|e|/dt (p - p_prev) = 0."""
    path = cases.ref_path("test6_21d.msh")
    om = O.Mesh(path)
    d = cases.darcy_case("test06_21", om)
    E = om.n_elements
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    stor = 0.5 + np.arange(E) % 3
    D.set_unsteady(0, stor, np.linspace(0, 1, E), 0.05)
    pprev = D.solution[D.S:D.S + E].copy()
    for k, dt in enumerate([0.05, 0.05, 0.2, 0.2]):
        D.unsteady_step(dt, rtol=1e-15)
        p = D.solution[D.S:D.S + E]
        r = D.element_balance() + stor * om.measure[:E] / dt * (p - pprev)
        assert np.max(np.abs(r)) < 1e-9 * max(1.0, np.max(np.abs(D.solution[:D.S]))), k
        pprev = p.copy()


def test_mh_approaches_heat_equation_series():
    """tests/10 exact_solution.py: u = 100 x + 200/pi sum (-1)^n/n exp(-0.02 n^2 pi^2 t) sin(n pi x);
    MH with dt = 0.01 on the 942-triangle mesh at t = 0.5 within 1.5 of the scale 100."""
    om = O.Mesh(cases.ref_path("test10.msh"))
    c = case10(om)
    D = O.Darcy(om, c["K"], c["cs"], c["sigma"], c["f"], *c["bc"], aniso=c["aniso"])
    D.set_unsteady(0, c["stor"], c["init"], 0.01)
    for _ in range(50):
        D.unsteady_step(0.01, rtol=1e-15)
    cen = om.centre.reshape(-1, 3)[:om.n_elements]
    # which end is x = 1: the .bc_inflow side (pressure 100)
    x = cen[:, 0] if np.mean(D.solution[D.S:D.S + D.E][cen[:, 0] > 0.9]) > 50 else 1.0 - cen[:, 0]
    u = 100.0 * x + 200.0 / math.pi * sum((-1) ** n / n * math.exp(-0.02 * n * n * math.pi ** 2 * 0.5) *
                                         np.sin(n * math.pi * x) for n in range(1, 400))
    assert np.max(np.abs(D.solution[D.S:D.S + D.E] - u)) < 1.5


@pytest.mark.gpu
@pytest.mark.parametrize("lumped", [0, 1])
def test_unsteady_gpu_against_oracle(tmp_path, lumped):
    path = str(tmp_path / "cube4.msh")
    cases.generated_mesh("cube", 4, path, n_fractures=2, n_channels=4, channel_len=4, seed=3)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=5)
    E = om.n_elements
    f = np.where(np.arange(E) % 7 == 0, 0.3, 0.0)
    stor = 0.2 + (np.arange(E) % 5) * 0.1
    init = np.linspace(-0.5, 1.0, E)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], f, d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D.set_unsteady(lumped, stor, init, 0.02)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], f, d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    g.set_unsteady(bool(lumped), stor, init)
    assert np.array_equal(g.get_solution_vector(), D.solution)
    for k, dt in enumerate([0.02, 0.02, 0.02, 0.1, 0.1]):
        D.unsteady_step(dt, rtol=1e-15)
        g.unsteady_step(dt, r_tol=1e-15, a_tol=0.0, max_it=100000)
        x, y = g.get_solution_vector(), D.solution
        S = D.S
        assert np.linalg.norm(x[:S] - y[:S]) <= 1e-9 * np.linalg.norm(y[:S]), k
        assert np.linalg.norm(x[S:] - y[S:]) <= 1e-9 * np.linalg.norm(y[S:]), k
    # the steady solve afterwards sees the steady system again
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], f, d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-15)
    g.update_solution(r_tol=1e-15, a_tol=0.0, max_it=100000)
    x = g.get_solution_vector()
    assert np.linalg.norm(x - D0.solution) <= 1e-9 * np.linalg.norm(D0.solution)


@pytest.mark.gpu
def test_lmh_gpu_against_test10(tmp_path):
    gm = G.Mesh(cases.ref_path("test10.msh"))
    om = O.Mesh(cases.ref_path("test10.msh"))
    c = case10(om)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(c["K"], c["cs"], c["sigma"], c["f"], *c["bc"], aniso=c["aniso"])
    g.set_unsteady(True, c["stor"], c["init"])
    gold = gold10()
    p = str(tmp_path / "wb.txt")
    g.write_water_balance(p, 0.0)
    S, E = gm.n_sides, gm.n_elements
    for k in range(1, 51):
        g.unsteady_step(0.01, r_tol=1e-14, a_tol=0.0, max_it=100000)
        if k % 10 == 0:
            x = g.get_solution_vector()
            assert np.max(np.abs(x[S:S + E] - gold[k // 10][1:])) < 1e-5, k
            g.write_water_balance(p, 0.01 * k)
    ndiff(open(p).read(), open(cases.ref_path("test10_lmh_water_balance.txt")).read(), 1e-5, 1e-6)
