"""GPU parity of the steady MH Darcy solve against the CPU oracle and the
reference's golden outputs.  Every call goes through the C ABI
(libflow123d_gpu.so)."""
import os
import sys

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

pytestmark = pytest.mark.gpu

REF_CASES = [("test01", "test01_mesh.msh"), ("test06_32", "test6_32d.msh"), ("test06_21", "test6_21d.msh"),
             ("test13", "test13.msh")]


def solve_both(path, case, seed=0, rtol=1e-15):
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case(case, om, seed=seed)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                 aniso=d["aniso"])
    D0.solve(0, rtol=rtol, max_it=200000)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    info = g.update_solution(r_tol=rtol, a_tol=0.0, max_it=200000)
    print("case", case, os.path.basename(path), "iters gpu", info["iterations"], "oracle", D0.stats["iters"],
          flush=True)
    return om, gm, D0, g, info


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def check_solution(D0, g, tol=1e-10):
    S, E = D0.S, D0.E
    x0 = D0.solution
    x1 = g.get_solution_vector()
    assert x1.shape == x0.shape
    eq, ep, el = rel(x1[:S], x0[:S]), rel(x1[S:S + E], x0[S:S + E]), rel(x1[S + E:], x0[S + E:])
    print("rel err flux %.2e pressure %.2e trace %.2e" % (eq, ep, el), flush=True)
    assert eq <= tol and ep <= tol and el <= tol
    return x1


@pytest.mark.parametrize("case,mesh", REF_CASES)
def test_darcy_reference_meshes(case, mesh):
    om, gm, D0, g, info = solve_both(cases.ref_path(mesh), case)
    assert info["converged"] == 1
    x = check_solution(D0, g)
    # condensed edge system: same CSR pattern, values to round-off
    p0, c0, v0, r0 = D0.edge_system()
    p1, c1, v1, r1 = g.edge_system()
    assert np.array_equal(p0, p1) and np.array_equal(c0, c1)
    scale = np.abs(v0).max()
    assert np.abs(v1 - v0).max() <= 1e-13 * scale
    assert np.abs(r1 - r0).max() <= 1e-12 * max(np.abs(r0).max(), 1e-300)
    # element balance (mass conservation of the element rows)
    eb = g.element_balance()
    assert np.abs(eb).max() <= 1e-9 * max(np.abs(x[:D0.S]).max(), 1.0)


def test_darcy_test01_golden():
    """tests/01_steady_flow_123d raw_output.txt (8 significant digits, PETSc rtol 1e-10)."""
    om, gm, D0, g, info = solve_both(cases.ref_path("test01_mesh.msh"), "test01")
    sol = g.get_solution_vector()
    S, E = gm.n_sides, gm.n_elements
    lines = open(cases.ref_path("test01_raw_output.txt")).read().split("\n")
    i0 = lines.index("$FlowField") + 3
    err = 0.0
    for e in range(E):
        t = lines[i0 + e].split()
        assert int(t[0]) == gm.elem_id[e]
        ns = int(t[5])
        err = max(err, abs(float(t[1]) - sol[S + e]))
        for i in range(ns):
            err = max(err, abs(float(t[6 + i]) - sol[S + E + gm.elem_edge[4 * e + i]]))
            err = max(err, abs(float(t[6 + ns + i]) - sol[gm.side_off[e] + i]))
    assert err < 2e-8
    # water balance per boundary region vs water_balance.out
    bal, plus, minus = g.water_balance()
    ref = {".1d_channel": 2.25456e-05, ".2d_fraction_1": -2.4826e-05, ".2d_fraction_2": -3.31144e-06,
           ".3d_cube": 5.59198e-06}
    for r in gm.regions:
        if r["boundary"] and r["label"] in ref:
            assert abs(bal[r["sub_idx"]] - ref[r["label"]]) < 2e-10


def test_darcy_test06_analytic():
    """tests/06_flow_coupling: fracture pressure -2 (head -3), channel pressure -2.5."""
    om, gm, D0, g, _ = solve_both(cases.ref_path("test6_32d.msh"), "test06_32")
    sol = g.get_solution_vector()
    p = sol[gm.n_sides:gm.n_sides + gm.n_elements]
    fr = gm.elem_dim == 2
    assert np.abs(p[fr] + 2.0).max() < 1e-6
    om, gm, D0, g, _ = solve_both(cases.ref_path("test6_21d.msh"), "test06_21")
    sol = g.get_solution_vector()
    p = sol[gm.n_sides:gm.n_sides + gm.n_elements]
    ch = gm.elem_dim == 1
    assert np.abs(p[ch] + 2.5).max() < 1e-6


@pytest.mark.parametrize("kind,n,kw", [("square", 20, {}), ("square", 100, {}),
                                       ("cube", 6, dict(n_fractures=3, n_channels=8, channel_len=5, seed=3))])
def test_darcy_generated(tmp_path, kind, n, kw):
    path = str(tmp_path / ("%s%d.msh" % (kind, n)))
    cases.generated_mesh(kind, n, path, **kw)
    om, gm, D0, g, info = solve_both(path, kind, seed=11, rtol=1e-14)
    check_solution(D0, g, tol=1e-9)


def test_darcy_linear_patch(tmp_path):
    """MH-RT0 reproduces a linear head exactly: K const, p = 1 - x on the unit square."""
    path = str(tmp_path / "sq8.msh")
    cases.generated_mesh("square", 8, path)
    gm, om = G.Mesh(path), O.Mesh(path)
    d = cases.darcy_case("square", om)
    d["K"][:] = 2.5
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    g.update_solution(r_tol=1e-15, a_tol=0.0, max_it=10000)
    sol = g.get_solution_vector()
    cen = om.centre.reshape(-1, 3)
    p = sol[gm.n_sides:gm.n_sides + gm.n_elements]
    assert np.abs(p - (1.0 - cen[:, 0])).max() < 1e-12
    # boundary inflow at x=0: total flux = -K * 1 (unit height)
    bal, plus, minus = g.water_balance()
    for r in gm.regions:
        if r["label"] == ".left":
            assert abs(bal[r["sub_idx"]] + 2.5) < 1e-11
        if r["label"] == ".right":
            assert abs(bal[r["sub_idx"]] - 2.5) < 1e-11


def test_darcy_repeat_and_new_realization(tmp_path):
    """Two solves in a row are bitwise identical; a new K realization is re-solved correctly."""
    path = str(tmp_path / "sq16.msh")
    cases.generated_mesh("square", 16, path)
    gm, om = G.Mesh(path), O.Mesh(path)
    d = cases.darcy_case("square", om, seed=1)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    g.update_solution(r_tol=1e-14, a_tol=0.0)
    a = g.get_solution_vector()
    g.update_solution(r_tol=1e-14, a_tol=0.0)
    b = g.get_solution_vector()
    assert np.array_equal(a, b)
    d2 = cases.darcy_case("square", om, seed=2)
    g.set_conductivity(d2["K"])
    g.update_solution(r_tol=1e-14, a_tol=0.0)
    c = g.get_solution_vector()
    D0 = O.Darcy(om, d2["K"], d2["cs"], d2["sigma"], d2["f"], d2["bc_type"], d2["bc_p"], d2["bc_flux"],
                 d2["bc_sigma"])
    D0.solve(0, rtol=1e-15)
    assert rel(c, D0.solution) < 1e-9


def test_cpp_adapter_example(tmp_path):
    """The C++ adapter (DarcyFlowMH_Steady / ConvectionTransport method set) over the
    C ABI; the flow output file it writes (node pressure, element pressure, head,
    velocity) against the oracle's fields on the same data."""
    import subprocess
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "examples", "steady_flow")
    assert os.path.exists(exe), "build() compiles the example"
    msh = str(tmp_path / "flow.msh")
    out = subprocess.run([exe, cases.ref_path("test01_mesh.msh"), msh], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "darcy:" in out.stdout and "transport:" in out.stdout
    om = O.Mesh(cases.ref_path("test01_mesh.msh"))
    E, NB = om.n_elements, om.n_bc_elements
    bnd = [r["id"] for r in om.regions if r["boundary"] and r["id"] != -3][:2]
    bct = np.isin(om.bc_region, bnd).astype(np.int32)
    bcp = (om.bc_region == bnd[0]).astype(float)
    D = O.Darcy(om, np.ones(E), np.ones(E), np.ones(E), np.zeros(E), bct, bcp, np.zeros(NB), np.zeros(NB))
    D.solve(0, rtol=1e-15)
    sec = {}
    for part in open(msh).read().split("Data\n1\n\"")[1:]:
        L = part.split("\n")
        n = int(L[6])
        sec[L[0].split("_[")[0]] = np.array([[float(x) for x in l.split()[1:]] for l in L[7:7 + n]])
    assert sorted(sec) == ["piezo_head_elements", "pressure_elements", "pressure_nodes", "velocity_elements"]
    p, h, v = D.element_output()
    for name, ref in [("pressure_nodes", D.node_pressure()), ("pressure_elements", p), ("piezo_head_elements", h),
                      ("velocity_elements", v)]:
        got = sec[name].reshape(ref.shape)
        assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-9 * np.max(np.abs(ref))), name



class _Solved:
    """The oracle solution of a solve made elsewhere, as check_solution reads it."""

    def __init__(self, solution, iters, S, E):
        self.solution, self.S, self.E, self.stats = solution, S, E, {"iters": iters}


def _oracle_cube54(out):
    import meshgen
    mesh = meshgen.cube_mesh(54, n_fractures=2, n_channels=0, seed=1)
    om = O.Mesh(arrays=mesh)
    d = cases.darcy_case("cube", om, seed=7)
    O.set_threads(min(16, len(os.sched_getaffinity(0))))  # the oracle's CG, at most 16 host threads
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-14, max_it=200000)
    np.savez(out, solution=D0.solution, iters=D0.stats["iters"])


def test_darcy_split_cg_beyond_cache():
    """Edge systems whose matrix and vectors exceed the last-level cache run the
    CG iteration in its split form (a direction pass, then the SpMV with
    non-temporal matrix loads over XCD-contiguous slices), smaller ones the
    fused kernel that every other test here runs: a 54^3 x 6 cube with two
    fractures (1.9 M edge unknowns) against the oracle in FP64 and in FP32 +
    FP64 reliable updates, and the FP64 solve repeated bit for bit."""
    import meshgen
    mesh = meshgen.cube_mesh(54, n_fractures=2, n_channels=0, seed=1)
    om, gm = O.Mesh(arrays=mesh), G.Mesh(arrays=mesh)
    d = cases.darcy_case("cube", om, seed=7)
    # the oracle's 1.9 M-unknown solve (multi-threaded) runs in a process of its
    # own, so nothing of it stays in this one for the tests that follow
    import multiprocessing as mp
    import tempfile
    out = os.path.join(tempfile.mkdtemp(), "oracle.npz")
    p = mp.get_context("spawn").Process(target=_oracle_cube54, args=(out,))
    p.start()
    p.join(600)
    assert p.exitcode == 0, "oracle process failed"
    ref = np.load(out)
    D0 = _Solved(ref["solution"], int(ref["iters"]), om.n_sides, om.n_elements)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    nnz = G.lib().fl_darcy_edge_nnz(g.h)
    # the split form's condition (stored entries >= nnz): 12 B per entry + 56 B per row > 224 MiB
    assert 12.0 * nnz + 56.0 * gm.n_edges > 224.0 * (1 << 20)
    i64 = g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=200000)
    print("F", gm.n_edges, "nnz", nnz, "iterations gpu", i64["iterations"], "oracle", D0.stats["iters"], flush=True)
    assert i64["converged"] == 1
    x1 = check_solution(D0, g, tol=1e-9)
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=200000)
    assert np.array_equal(g.get_solution_vector(), x1)
    i32 = g.update_solution(r_tol=1e-12, a_tol=0.0, max_it=200000, precision=1)
    print("mixed iterations", i32["iterations"], "reliable updates", i32["reliable_updates"], flush=True)
    assert i32["converged"] == 1
    check_solution(D0, g, tol=1e-8)


@pytest.mark.parametrize("case,mesh", REF_CASES)
def test_darcy_mixed_precision_reference_meshes(case, mesh):
    """FP32 CG with FP64 reliable updates: the stop test is on the true FP64
    residual, so the result matches the FP64 oracle like the FP64 path does."""
    path = cases.ref_path(mesh)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case(case, om)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                 aniso=d["aniso"])
    D0.solve(0, rtol=1e-15)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    info = g.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=1)
    print(case, "mixed iters", info["iterations"], "reliable updates", info["reliable_updates"])
    assert info["converged"] == 1 and info["residual"] <= 1e-13 * info["rhs_norm"] * (1 + 1e-12)
    check_solution(D0, g, tol=1e-9)


@pytest.mark.parametrize("kind,n,kw", [("square", 100, {}),
                                       ("cube", 8, dict(n_fractures=3, n_channels=10, channel_len=6, seed=4))])
def test_darcy_mixed_precision_generated(tmp_path, kind, n, kw):
    path = str(tmp_path / "m.msh")
    cases.generated_mesh(kind, n, path, **kw)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case(kind, om, seed=5)
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-15, max_it=200000)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    i64 = g.update_solution(r_tol=1e-12, a_tol=0.0, max_it=200000, precision=0)
    i32 = g.update_solution(r_tol=1e-12, a_tol=0.0, max_it=200000, precision=1)
    print(kind, "iterations fp64", i64["iterations"], "mixed", i32["iterations"], "ru", i32["reliable_updates"])
    assert i32["converged"] == 1
    check_solution(D0, g, tol=1e-8)
