"""Pins the CPU oracle against the reference's own golden vectors and against
what the mathematics fixes (closed forms, brute force).  CPU only."""
import numpy as np
import pytest

import cases
import meshgen
import oracle as O


def test_simplest_cube_topology():
    """test_units/mesh/mesh_test.cpp:53-81 known answers."""
    m = O.Mesh(cases.ref_path("simplest_cube.msh"))
    assert m.n_elements == 9
    assert m.n_bc_elements == 18
    assert list(m.bc_region[:4]) == [101, 101, 102, 102]
    assert m.bc_region[4] == -3 and m.bc_region[17] == -3
    assert m.n_edges == 28
    assert m.n_vb == 6


def test_square_counts():
    """Structured 100x100 '/'-split square: F = 2n(n+1) + n^2 = 30,200 edges."""
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.msh")
        cases.generated_mesh("square", 100, p)
        m = O.Mesh(p)
        assert (m.n_elements, m.n_edges, m.n_sides) == (20000, 30200, 60000)
        # every interior edge has two sides, boundary edges one
        deg = np.diff(m.edge_ptr)
        assert deg.max() == 2 and (deg == 1).sum() == 400


def _load_case(name, mesh):
    om = O.Mesh(cases.ref_path(mesh))
    d = cases.darcy_case(name, om)
    D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                aniso=d["aniso"])
    return om, D


def test_test01_raw_output_golden():
    """tests/01_steady_flow_123d/ref_output/flow_gmsh.con/raw_output.txt, 8 digits."""
    om, D = _load_case("test01", "test01_mesh.msh")
    D.solve(0, rtol=1e-15)
    sol, S, E = D.solution, D.S, D.E
    lines = open(cases.ref_path("test01_raw_output.txt")).read().split("\n")
    i0 = lines.index("$FlowField") + 3
    off = np.concatenate([[0], np.cumsum(om.elem_dim + 1)])
    err = 0.0
    for e in range(E):
        t = lines[i0 + e].split()
        assert int(t[0]) == om.elem_id[e]
        ns = int(t[5])
        assert ns == om.elem_dim[e] + 1
        err = max(err, abs(float(t[1]) - sol[S + e]))
        for i in range(ns):
            err = max(err, abs(float(t[6 + i]) - sol[S + E + om.elem_edge[4 * e + i]]))
            err = max(err, abs(float(t[6 + ns + i]) - sol[off[e] + i]))
    assert err < 2e-8


def test_test01_water_balance_golden():
    om, D = _load_case("test01", "test01_mesh.msh")
    D.solve(0, rtol=1e-15)
    bal, plus, minus = D.water_balance()
    gold = {".1d_channel": (2.25456e-05, 2.25456e-05, 0.0), ".2d_fraction_1": (-2.4826e-05, 1.05424, -1.05426),
            ".2d_fraction_2": (-3.31144e-06, 1.05426, -1.05427), ".3d_cube": (5.59198e-06, 0.102171, -0.102165)}
    for r in om.regions:
        if r["label"] in gold:
            g = gold[r["label"]]
            i = r["sub_idx"]
            assert abs(bal[i] - g[0]) < 2e-10
            assert abs(plus[i] - g[1]) <= 1e-5 * abs(g[1]) + 1e-12
            assert abs(minus[i] - g[2]) <= 1e-5 * abs(g[2]) + 1e-12


def test_two_stage_schur_equals_dense_mh():
    """Brute force: dense LU of the full saddle-point MH system equals the
    two-stage Schur + CG path (darcy_flow_mh.cc:1009-1131)."""
    om, D = _load_case("test01", "test01_mesh.msh")
    D.solve(0, rtol=1e-15)
    x = D.solution.copy()
    om2, D2 = _load_case("test01", "test01_mesh.msh")
    D2.solve(2)
    assert np.abs(D2.solution - x).max() < 1e-11 * np.abs(x).max()
    assert D.mh_residual() < 1e-12


def test_test06_analytic():
    """tests/06_flow_coupling description: p2 = -2 (head -3) on the 3D-2D fracture,
    p1 = -2.5 on the 2D-1D channel."""
    om, D = _load_case("test06_32", "test6_32d.msh")
    D.solve(0, rtol=1e-15)
    p = D.solution[D.S:D.S + D.E]
    assert np.abs(p[om.elem_dim == 2] + 2.0).max() < 1e-6
    om, D = _load_case("test06_21", "test6_21d.msh")
    D.solve(0, rtol=1e-15)
    p = D.solution[D.S:D.S + D.E]
    assert np.abs(p[om.elem_dim == 1] + 2.5).max() < 1e-6


def test_local_matrix_closed_forms():
    """RT0 mass matrices: line [[1,-1/2],[-1/2,1]] L/(3 K cs); reference triangle
    and tetrahedron against an independent quadrature of the RT0 basis; SPD; 1/(K cs) scaling."""
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.msh")
        with open(p, "w") as f:
            f.write('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n1 0 0 0\n2 2 0 0\n3 0 1 0\n4 0 0 1\n5 3 0 0\n'
                    '$EndNodes\n$Elements\n3\n1 1 2 1 1 2 5\n2 2 2 2 2 1 2 3\n3 4 2 3 3 1 2 3 4\n$EndElements\n')
        m = O.Mesh(p)
        # line of length 1 (nodes 2,5)
        loc, inv, det = m.local_matrix(0, K=2.0, cs=0.5)
        assert np.allclose(loc, np.array([[1, -0.5], [-0.5, 1]]) / 3.0, rtol=1e-15)
        assert np.allclose(loc @ inv, np.eye(2), atol=1e-14)
        for e, d in ((1, 2), (2, 3)):
            loc, inv, det = m.local_matrix(e)
            # independent RT0 basis + quadrature: phi_i(x) = (x - P_opp(i)) / (d |e|) scaled
            en = m.elem_nodes[4 * e:4 * e + d + 1]
            P = m.nodes.reshape(-1, 3)[en]
            sides = {2: [(0, 1), (1, 2), (0, 2)], 3: [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]}[d]
            opp = [[k for k in range(d + 1) if k not in s][0] for s in sides]
            vol = m.measure[e]
            # exact quadrature for quadratics on a simplex: vertices + edge midpoints (2D) / Keast-like (3D)
            if d == 2:
                pts = [(P[0] + P[1]) / 2, (P[1] + P[2]) / 2, (P[0] + P[2]) / 2]
                w = [vol / 3] * 3
            else:
                a, b = 0.5854101966249685, 0.1381966011250105
                bary = [[a, b, b, b], [b, a, b, b], [b, b, a, b], [b, b, b, a]]
                pts = [np.dot(np.array(bb), P) for bb in bary]
                w = [vol / 4] * 4
            def phi(i, x):
                # flux-normalised RT0: int over side i of phi.n = 1
                smeas = {2: lambda s: np.linalg.norm(P[s[1]] - P[s[0]]),
                         3: lambda s: 0.5 * np.linalg.norm(np.cross(P[s[1]] - P[s[0]], P[s[2]] - P[s[0]]))}[d](sides[i])
                return (x - P[opp[i]]) / (d * vol)
            M = np.zeros((d + 1, d + 1))
            for i in range(d + 1):
                for j in range(d + 1):
                    M[i, j] = sum(wq * phi(i, x) @ phi(j, x) for wq, x in zip(w, pts))
            assert np.allclose(loc, M, rtol=1e-12, atol=1e-14)
            assert np.all(np.linalg.eigvalsh(loc) > 0)
            loc2, _, _ = m.local_matrix(e, K=4.0, cs=0.25)
            assert np.allclose(loc2, loc, rtol=1e-14)
            loc3, _, _ = m.local_matrix(e, K=10.0, cs=1.0)
            assert np.allclose(loc3, loc / 10.0, rtol=1e-14)


def test_test13_transport_golden():
    """tests/13_edge_solute_mixing: transport.msh concentrations (%g) at all 200
    output times and the final mass balance of mass_balance.txt."""
    om, D = _load_case("test13", "test13.msh")
    D.solve(0, rtol=1e-15)
    E = om.n_elements
    T = O.Transport(om, D.solution[:D.S], np.ones(E), np.ones(E), cases.test13_bc_conc(om), np.zeros(E), 100.0, 0.5)
    txt = open(cases.ref_path("test13_transport.msh")).read().split("$ElementData")[1:]
    gold = {}
    for blk in txt:
        L = blk.strip().split("\n")
        gold[round(float(L[3]), 6)] = np.array([float(l.split()[1]) for l in L[8:8 + int(L[7])]])
    worst = 0.0
    total_steps = 0
    for k in range(1, 201):
        t = 0.5 * k
        total_steps += T.update(t)
        c = T.conc()[0]
        g = gold[round(t, 6)]
        m = np.abs(g) > 1e-4
        if m.any():
            worst = max(worst, (np.abs(c - g)[m] / np.abs(g[m])).max())
        assert np.abs(c - g)[~m].max(initial=0.0) < 1e-9
    assert worst < 2e-5
    assert total_steps == 600
    mb = T.mass_balance()
    assert abs(mb["bal"][0, 0] + 0.0281415) < 1e-6
    assert abs(mb["plus"][0, 0] - 0.846884) < 1e-6
    assert abs(mb["minus"][0, 0] + 0.875025) < 1e-6
    assert np.allclose(mb["mass"][0], [5.0, 6.72162, 7.07107], rtol=1e-5)
    tot = mb["totals"][0]
    assert abs(tot[1] + 18.7927) < 1e-4
    assert abs(tot[0] + tot[2] - tot[1] - tot[3]) < 1e-11 * tot[3]


def test_transport_invariants():
    """Monotone upwind with dt <= CFL: 0 <= c <= 1 under unit inflow; rows of
    tm = I + dt A sum to 1 - dt*(outflow) and are non-negative."""
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "c.msh")
        cases.generated_mesh("cube", 4, p, n_fractures=2, n_channels=4, channel_len=4, seed=1)
        om = O.Mesh(p)
        d = cases.darcy_case("cube", om, seed=3)
        D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
        D.solve(0, rtol=1e-14)
        E = om.n_elements
        T = O.Transport(om, D.solution[:D.S], d["cs"], np.ones(E), np.ones(om.n_bc_elements), np.zeros(E), 1.0, 1.0)
        T.update(1.0)
        c = T.conc()[0]
        assert c.min() >= -1e-14 and c.max() <= 1 + 1e-12
        ptr, col, val = T.matrix()
        assert val.min() >= -1e-15
        assert T.dt <= T.cfl * (1 + 1e-12)


def test_transport_sources_balance():
    """Oracle with sources: the cumulative balance closes (por = cs = 1) and with
    zero source fields the result equals the run without sources bitwise."""
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.msh")
        cases.generated_mesh("square", 8, p)
        om = O.Mesh(p)
        d = cases.darcy_case("square", om, seed=2)
        D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
        D.solve(0, rtol=1e-14)
        E, NB = om.n_elements, om.n_bc_elements
        one = np.ones(E)
        rng = np.random.default_rng(0)
        T = O.Transport(om, D.solution[:D.S], one, one, np.zeros(NB), np.zeros(E), 0.5, 0.25)
        T.set_sources(rng.uniform(0, 0.3, E), rng.uniform(0, 1, E), rng.uniform(0, 1, E))
        T.update(0.25); T.update(0.5)
        a, flx, src, cur = T.mass_balance()["totals"][0]
        assert src > 0 and abs(a + src - flx - cur) < 1e-12
        T1 = O.Transport(om, D.solution[:D.S], one, one, np.ones(NB), np.zeros(E), 0.5, 0.25)
        T2 = O.Transport(om, D.solution[:D.S], one, one, np.ones(NB), np.zeros(E), 0.5, 0.25)
        T2.set_sources(np.zeros(E), np.zeros(E), np.zeros(E))
        T1.update(0.5); T2.update(0.5)
        assert np.array_equal(T1.conc(), T2.conc())


def test_cpu_baseline_solvers_agree(tmp_path):
    """The CPU-baseline variants of the edge solve (IC(0)-CG, threaded
    Jacobi-CG) reach the dense LU solution of the same edge system; IC(0) needs
    fewer iterations than Jacobi (a wrong factor would not)."""
    import meshgen
    m = meshgen.square_mesh(24)
    om = O.Mesh(arrays=m)
    d = cases.darcy_case("square", om, seed=3)
    args = (om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D = O.Darcy(*args)
    D.solve(1)
    x_lu = D.solution.copy()
    S = om.n_sides
    it = {}
    for meth, th in ((0, 1), (0, 4), (3, 1)):
        O.set_threads(th)
        D = O.Darcy(*args)
        it[(meth, th)] = D.solve(meth, rtol=1e-13)
        e = np.linalg.norm(D.solution - x_lu) / np.linalg.norm(x_lu)
        assert e < 1e-10, (meth, th, e)
    O.set_threads(1)
    assert it[(3, 1)] < 0.6 * it[(0, 1)]
    assert abs(it[(0, 4)] - it[(0, 1)]) <= 3
