"""Batched Monte-Carlo Darcy solves (many conductivity realizations, one
launch sequence) against the CPU oracle and the single-solve GPU path.  Every
call goes through the C ABI."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def parts(x, S, E):
    return x[:S], x[S:S + E], x[S + E:]


def setup(path, case, seed=0):
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case(case, om, seed=seed)
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
               aniso=d["aniso"])
    return om, gm, d, g


def oracle_solve(om, d, K, rtol=1e-15):
    D0 = O.Darcy(om, K, d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"],
                 aniso=d["aniso"])
    D0.solve(0, rtol=rtol, max_it=200000)
    return D0.solution


@pytest.mark.parametrize("case,mesh", [("test01", "test01_mesh.msh"), ("test06_32", "test6_32d.msh"),
                                       ("test06_21", "test6_21d.msh"), ("test13", "test13.msh")])
def test_batch_reference_meshes(case, mesh):
    """Dirichlet / Neumann / anisotropy / vb coupling of the reference's own
    inputs, three K realizations (the case's K times a random factor field)."""
    om, gm, d, g = setup(cases.ref_path(mesh), case)
    R = 3
    rng = np.random.default_rng(7)
    K = d["K"][None, :] * np.exp(0.7 * rng.standard_normal((R, om.n_elements)))
    b = G.DarcyBatch(g, R)
    b.set_conductivity(K)
    info = b.update_solution(r_tol=1e-15, a_tol=0.0, max_it=100000, precision=0)
    print(case, "iterations", info["iterations"], "residual", info["residual"], flush=True)
    S, E = om.n_sides, om.n_elements
    for r in range(R):
        x0 = oracle_solve(om, d, K[r])
        x1 = b.get_solution_vector(r)
        eq, ep, el = [rel(a, c) for a, c in zip(parts(x1, S, E), parts(x0, S, E))]
        print("  r=%d rel err flux %.2e pressure %.2e trace %.2e" % (r, eq, ep, el), flush=True)
        assert eq <= 1e-10 and ep <= 1e-10 and el <= 1e-10


@pytest.fixture(scope="module")
def cube6(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("mc") / "cube6.msh")
    cases.generated_mesh("cube", 6, path, n_fractures=3, n_channels=8, channel_len=5, seed=3)
    om, gm, d, g = setup(path, "cube", seed=1)
    return om, gm, d, g


def test_batch_matches_single_solves(cube6):
    """70 realizations (padded to 128 lanes): FP64 batch vs the single FP64 GPU
    solve of each sampled realization; uneven iteration counts."""
    om, gm, d, g = cube6
    R = 70
    K = cases.mc_conductivity(om, R, seed=3)
    b = G.DarcyBatch(g, R)
    b.set_conductivity(K)
    info = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=0)
    it = info["iterations"]
    print("iterations min %d max %d; grid %d xcd %d chunk %d" % (it.min(), it.max(), info["grid"],
                                                                  info["xcd_map"], info["chunk"]), flush=True)
    assert info["converged"].all() and info["n_converged"] == R
    assert it.max() > it.min()
    assert info["compactions"] >= 1  # 70 realizations in two 64-slot chunks: the tail runs in one
    assert (info["residual"] <= 2e-13).all()
    S, E = om.n_sides, om.n_elements
    for r in (0, 31, 63, 64, 69):
        g.set_conductivity(K[r])
        g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
        x0 = g.get_solution_vector()
        x1 = b.get_solution_vector(r)
        errs = [rel(a, c) for a, c in zip(parts(x1, S, E), parts(x0, S, E))]
        print("  r=%d rel err flux %.2e pressure %.2e trace %.2e" % (r, *errs), flush=True)
        assert max(errs) <= 1e-9
    g.set_conductivity(d["K"])


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_slot_refill(cube6, precision):
    """100 realizations through 64 slots: converged slots are refilled from
    the queue (hardest first); every realization matches its single solve."""
    om, gm, d, g = cube6
    R = 100
    K = cases.mc_conductivity(om, R, seed=8)
    b = G.DarcyBatch(g, R, slots=64)
    b.set_conductivity(K)
    info = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=precision)
    it = info["iterations"]
    print("refill precision %d: iterations min %d max %d total %d" % (precision, it.min(), it.max(),
                                                                      info["total_iterations"]), flush=True)
    assert info["converged"].all() and (it > 0).all()
    S, E = om.n_sides, om.n_elements
    for r in (0, 63, 64, 99):
        g.set_conductivity(K[r])
        g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
        x0 = g.get_solution_vector()
        x1 = b.get_solution_vector(r)
        errs = [rel(a, c) for a, c in zip(parts(x1, S, E), parts(x0, S, E))]
        print("  r=%d rel err flux %.2e pressure %.2e trace %.2e" % (r, *errs), flush=True)
        assert max(errs) <= 1e-8
    g.set_conductivity(d["K"])


def test_batch_mixed_precision(cube6):
    """FP32 CG with FP64 reliable updates: flux within 1e-8 of the FP64 oracle.
    With sigma up to 3 the flux error is up to ~2e4 times the relative
    residual, so the stop test is 1e-13."""
    om, gm, d, g = cube6
    R = 64
    K = cases.mc_conductivity(om, R, seed=4)
    b = G.DarcyBatch(g, R)
    b.set_conductivity(K)
    i64 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=0)
    i32 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=1)
    st, ru = b.diagnostics()
    print("fp64 iterations total %d, mixed %d, reliable updates %d, fp64 finish %d" % (
        i64["total_iterations"], i32["total_iterations"], ru.sum(), i32["fp64_finish"]), flush=True)
    assert i32["converged"].all() and (st == 1).all()
    assert (i32["residual"] <= 1e-13 * 1.01).all()
    assert i32["total_iterations"] < 2 * i64["total_iterations"]
    S = om.n_sides
    for r in (0, 17, 63):
        x0 = oracle_solve(om, d, K[r])
        x1 = b.get_solution_vector(r)
        e = rel(x1[:S], x0[:S])
        print("  r=%d flux rel err vs oracle %.2e" % (r, e), flush=True)
        assert e <= 1e-8


def test_batch_launch_shapes_and_repeat(cube6):
    """Repeated solves are bitwise identical; another launch shape (grid,
    XCD mapping, chunk) changes only the reduction order; K from a device
    pointer gives the same result as from the host."""
    import torch
    om, gm, d, g = cube6
    R = 64
    K = cases.mc_conductivity(om, R, seed=5)
    b = G.DarcyBatch(g, R)
    b.set_conductivity(K)
    b.set_launch(512, 0, 16)
    b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000)
    a = np.stack([b.get_solution_vector(r) for r in (0, 40)])
    b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000)
    a2 = np.stack([b.get_solution_vector(r) for r in (0, 40)])
    assert np.array_equal(a, a2)
    b.set_launch(256, 1, 7)
    info = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000)
    assert info["converged"].all()
    assert info["grid"] == 256 and info["xcd_map"] == 1 and info["chunk"] == 7
    b.set_launch(1024, 2, 16)
    c2 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000)
    assert c2["xcd_map"] == 2 and c2["converged"].all()
    c = np.stack([b.get_solution_vector(r) for r in (0, 40)])
    assert rel(c, a) < 1e-11
    Kd = torch.tensor(K, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    b.set_conductivity(Kd.data_ptr())
    b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000)
    e = np.stack([b.get_solution_vector(r) for r in (0, 40)])
    assert np.array_equal(e, c)
    with pytest.raises(G.FlowError):
        b.get_solution_vector(R)


def test_batch_autotune(cube6):
    om, gm, d, g = cube6
    R = 128
    K = cases.mc_conductivity(om, R, seed=6)
    b = G.DarcyBatch(g, R)
    b.set_conductivity(K)
    b.autotune()
    info = b.update_solution(r_tol=1e-12, a_tol=0.0, max_it=100000, precision=1)
    print("autotuned grid %d xcd %d" % (info["grid"], info["xcd_map"]), flush=True)
    assert info["grid"] in (512, 1024, 2048) and info["xcd_map"] in (0, 1, 2) and info["converged"].all()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_handover(cube6, precision):
    """The last live realizations are finished by the single-solve FP64 path
    (warm start from the batch iterate): same answers as the batch alone."""
    om, gm, d, g = cube6
    R = 64
    K = cases.mc_conductivity(om, R, seed=9)
    b = G.DarcyBatch(g, R, slots=64)
    b.set_conductivity(K)
    b.set_handover(0)
    i0 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=precision)
    x0 = [b.get_solution_vector(r) for r in (0, 30, 63)]
    b.set_handover(48)
    i1 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=precision)
    x1 = [b.get_solution_vector(r) for r in (0, 30, 63)]
    print("handovers %d (%d iterations) vs none; totals %d / %d" % (
        i1["handovers"], i1["handover_iterations"], i1["total_iterations"], i0["total_iterations"]), flush=True)
    # with hand-over off only realizations whose FP32 recursion stagnated go to the FP64 path
    assert i0["handovers"] == i0["fp64_finish"] and i1["handovers"] >= 1
    assert i1["converged"].all() and i0["converged"].all()
    S = om.n_sides
    for a, c in zip(x1, x0):
        assert rel(a[:S], c[:S]) <= 1e-8
    # the base solver's own conductivity is restored
    g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
    xs = g.get_solution_vector()
    D0 = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D0.solve(0, rtol=1e-15, max_it=200000)
    assert rel(xs, D0.solution) <= 1e-9


def test_batch_two_level(cube6):
    """Two-level preconditioner (aggregation coarse space in flexible CG): same
    solutions as the single FP64 solves, fewer iterations than Jacobi; a
    solution-shaped second coarse column lowers them further."""
    import torch
    om, gm, d, g = cube6
    R = 64
    K = cases.mc_conductivity(om, R, seed=10)
    b = G.DarcyBatch(g, R, slots=64)
    b.set_conductivity(K)
    b.set_handover(0)
    nag, eagg = b.aggregates()
    assert nag >= 2 and eagg.min() == 0 and eagg.max() == nag - 1
    jac = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=0)
    S, E = om.n_sides, om.n_elements
    lam = np.stack([b.get_solution_vector(r)[S + E:] for r in range(R)])
    b.set_coarse(1)
    c1 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=0)
    x1 = [b.get_solution_vector(r) for r in (0, 33)]
    phi = torch.tensor(lam, dtype=torch.float32, device="cuda")
    b.set_coarse_basis(phi.data_ptr())
    b.set_coarse(2)
    c2 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=0)
    c2m = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=1)
    print("aggregates %d; total iterations jacobi %d, constant %d, constant+solution %d (mixed %d)" % (
        nag, jac["total_iterations"], c1["total_iterations"], c2["total_iterations"], c2m["total_iterations"]),
        flush=True)
    assert c1["converged"].all() and c2["converged"].all() and c2m["converged"].all()
    assert c1["total_iterations"] < jac["total_iterations"]
    assert c2["total_iterations"] < c1["total_iterations"]
    for r, x in zip((0, 33), x1):
        g.set_conductivity(K[r])
        g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
        x0 = g.get_solution_vector()
        errs = [rel(a, c) for a, c in zip(parts(x, S, E), parts(x0, S, E))]
        e2 = rel(b.get_solution_vector(r)[:S], x0[:S])
        print("  r=%d rel err flux %.2e pressure %.2e trace %.2e; two-column flux %.2e" % (r, *errs, e2), flush=True)
        assert max(errs) <= 1e-8 and e2 <= 1e-8
    g.set_conductivity(d["K"])


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_two_level_narrow_tail(cube6, precision):
    """The two-level tail continues in the batch with 16, then 4 lanes per edge
    row (several rows per wavefront): every realization's solution equals the
    one of the full-width run and of the single FP64 solve."""
    om, gm, d, g = cube6
    R = 80
    K = cases.mc_conductivity(om, R, seed=12)
    b = G.DarcyBatch(g, R, slots=64)
    b.set_conductivity(K)
    b.set_handover(0)
    b.set_coarse(1)
    b.set_narrow(False)
    wide = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=precision)
    S, E = om.n_sides, om.n_elements
    xw = [b.get_solution_vector(r) for r in range(R)]
    b.set_narrow(True)
    nar = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=precision)
    xn = [b.get_solution_vector(r) for r in range(R)]
    print("precision %d: iterations wide %d, narrow %d; narrowings+compactions %d / %d" % (
        precision, wide["total_iterations"], nar["total_iterations"], nar["compactions"], wide["compactions"]),
        flush=True)
    assert wide["converged"].all() and nar["converged"].all()
    assert nar["compactions"] > wide["compactions"]  # the tail was narrowed (16, then 4 lanes per row)
    worst = max(rel(a[:S], c[:S]) for a, c in zip(xn, xw))
    print("  worst flux difference narrow vs wide %.2e" % worst, flush=True)
    assert worst <= 1e-8
    hard = int(np.argmax(nar["iterations"]))  # the longest solve ran its end in the narrowest mode
    for r in (0, hard):
        g.set_conductivity(K[r])
        g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=100000)
        x0 = g.get_solution_vector()
        errs = [rel(a, c) for a, c in zip(parts(xn[r], S, E), parts(x0, S, E))]
        print("  r=%d (%d iterations) rel err flux %.2e pressure %.2e trace %.2e" % (
            r, nar["iterations"][r], *errs), flush=True)
        assert max(errs) <= 1e-8
    g.set_conductivity(d["K"])


@pytest.mark.parametrize("coarse", [0, 1])
def test_batch_handover8_fp32(cube6, coarse):
    """Hand-over of the live realizations once at most 8 are left in the FP32
    batch (one-level, and two-level with its narrowed tail; the count is read
    at chunk ends, so fewer than 8 may be left by then): the batch terminates
    with every realization converged, equal to the batch run without
    hand-over."""
    om, gm, d, g = cube6
    R = 96
    K = cases.mc_conductivity(om, R, seed=21)
    b = G.DarcyBatch(g, R, slots=64)
    b.set_conductivity(K)
    b.set_coarse(coarse)
    b.set_handover(0)
    i0 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=1)
    x0 = [b.get_solution_vector(r) for r in range(R)]
    b.set_handover(8)
    i1 = b.update_solution(r_tol=1e-13, a_tol=0.0, max_it=100000, precision=1)
    x1 = [b.get_solution_vector(r) for r in range(R)]
    print("coarse %d: handovers %d (%d iterations), chunks %d / %d" % (
        coarse, i1["handovers"], i1["handover_iterations"], i1["chunks"], i0["chunks"]), flush=True)
    assert i0["converged"].all() and i1["converged"].all()
    assert 1 <= i1["handovers"] and i0["handovers"] == 0
    S = om.n_sides
    worst = max(rel(a[:S], c[:S]) for a, c in zip(x1, x0))
    print("  worst flux difference %.2e" % worst, flush=True)
    assert worst <= 1e-8
    g.set_conductivity(d["K"])
