"""fl_darcy_update_solution_noalloc (CG scratch in a file-scope device array,
no allocation during the call) against fl_darcy_update_solution through the
C ABI: two calls in a row, the second on a smaller problem, each bitwise
equal to the allocating entry point's result."""
import ctypes as C

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

pytestmark = pytest.mark.gpu


def problem(tmp_path, n, seed):
    path = str(tmp_path / ("cube%d.msh" % n))
    cases.generated_mesh("cube", n, path, n_fractures=2, n_channels=4, channel_len=4, seed=seed)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=seed)
    return om, gm, d


def solve_both(gm, d, norm):
    L = G.lib()
    out = []
    for entry in (L.fl_darcy_update_solution, L.fl_darcy_update_solution_noalloc):
        g = G.DarcyFlowMH_Steady(gm)
        g.set_data(d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
        o = G.SolverOpts(1e-13, 0.0, 100000, 0, norm)
        info = G.SolveInfo()
        rc = entry(g.h, C.byref(o), C.byref(info))
        assert rc == 0, L.fl_last_error().decode()
        out.append((g.get_solution_vector(), info.iterations, info.residual))
    return out


def test_noalloc_matches_allocating_solve(tmp_path):
    for n, seed, norm in ((6, 3, 1), (4, 5, 0)):  # the second call is smaller
        om, gm, d = problem(tmp_path, n, seed)
        (x0, it0, r0), (x1, it1, r1) = solve_both(gm, d, norm)
        print("n=%d: iterations %d / %d, residual %.3e / %.3e, max |diff| %.1e" % (
            n, it0, it1, r0, r1, np.abs(x0 - x1).max()), flush=True)
        assert it0 == it1 and r0 == r1
        assert np.array_equal(x0, x1)
        D = O.Darcy(om, d["K"], d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
        D.solve(0, rtol=1e-15, max_it=200000)
        S = om.n_sides
        assert np.linalg.norm(x1[:S] - D.solution[:S]) <= 1e-9 * np.linalg.norm(D.solution[:S])


def test_noalloc_rejects_fp32():
    L = G.lib()
    o = G.SolverOpts(1e-12, 0.0, 100, 1, 0)
    info = G.SolveInfo()
    # no handle is touched: the precision check comes first
    assert L.fl_darcy_update_solution_noalloc(None, C.byref(o), C.byref(info)) == -8
