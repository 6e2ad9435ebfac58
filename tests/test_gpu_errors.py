"""Failure statuses of the GPU path through the C ABI: singular local
matrices (src/flow/mh_fe_values.cc:54-59), bad field values, CG stopped at
its iteration limit (src/la/solve.cc:454-458) or diverging.  Every case is a
computed check on finite launches; none faults."""
import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O
from test_error_paths import NUM_ZERO, flat_square

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube4(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("err") / "cube4.msh")
    cases.generated_mesh("cube", 4, path, n_fractures=2, n_channels=4, channel_len=4, seed=2)
    om, gm = O.Mesh(path), G.Mesh(path)
    d = cases.darcy_case("cube", om, seed=0)
    return om, gm, d


def darcy(gm, d, **over):
    g = G.DarcyFlowMH_Steady(gm)
    a = dict(d)
    a.update(over)
    g.set_data(a["K"], a["cs"], a["sigma"], a["f"], a["bc_type"], a["bc_p"], a["bc_flux"], a["bc_sigma"])
    return g


def expect(status, fn, match=None):
    with pytest.raises(G.FlowError, match=match) as ei:
        fn()
    assert ei.value.status == status, str(ei.value)
    print(str(ei.value), flush=True)


def test_zero_area_element_is_reported():
    m, tri = flat_square()
    om, gm = O.Mesh(arrays=m), G.Mesh(arrays=m)
    E, NB = om.n_elements, om.n_bc_elements
    bad = [e for e in range(E) if not abs(om.local_matrix(e)[2]) >= NUM_ZERO]
    assert len(bad) >= 1
    bct = np.zeros(NB, np.int32)
    bct[:2] = 1
    g = G.DarcyFlowMH_Steady(gm)
    g.set_data(np.ones(E), np.ones(E), np.ones(E), np.zeros(E), bct, np.zeros(NB), np.zeros(NB), np.zeros(NB))
    expect(G.E_SINGULAR_ELEMENT, lambda: g.update_solution(r_tol=1e-12),
           "Singular local matrix of the element %d " % gm.elem_id[bad[0]])


def test_underflowing_determinant_matches_oracle(cube4):
    om, gm, d = cube4
    e0 = int(np.flatnonzero(om.elem_dim == 3)[7])
    K = d["K"].copy()
    K[e0] = 1e80
    g = darcy(gm, d, K=K)
    expect(G.E_SINGULAR_ELEMENT, lambda: g.update_solution(r_tol=1e-12),
           "Singular local matrix of the element %d " % om.elem_id[e0])
    D = O.Darcy(om, K, d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    with pytest.raises(RuntimeError, match="element %d$" % om.elem_id[e0]):
        D.solve(0)
    # the flag is cleared: the same handle solves again with good data
    g.set_conductivity(d["K"])
    info = g.update_solution(r_tol=1e-12, a_tol=0.0)
    assert info["status"] == 0 and info["converged"] == 1


@pytest.mark.parametrize("field,value", [("K", np.nan), ("K", 0.0), ("K", -2.0), ("cs", np.inf), ("cs", 0.0)])
def test_bad_field_values_are_rejected(cube4, field, value):
    om, gm, d = cube4
    e0 = 11
    a = d[field].copy()
    a[e0] = value
    name = "conductivity" if field == "K" else "cross_section"
    expect(G.E_BAD_FIELD, lambda: darcy(gm, d, **{field: a}), "%s of the element %d " % (name, om.elem_id[e0]))


def test_bad_conductivity_in_a_batch(cube4):
    """host and device K arrays of a batch: the message names realization and element"""
    import torch
    om, gm, d = cube4
    g = darcy(gm, d)
    R = 5
    KR = cases.mc_conductivity(om, R, seed=2)
    KR[3, 17] = -1.0
    b = G.DarcyBatch(g, R)
    expect(G.E_BAD_FIELD, lambda: b.set_conductivity(KR),
           "conductivity of realization 3, element %d " % om.elem_id[17])
    KR[3, 17] = np.nan
    Kd = torch.tensor(KR, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    expect(G.E_BAD_FIELD, lambda: b.set_conductivity(Kd.data_ptr()),
           "conductivity of realization 3, element %d " % om.elem_id[17])
    KR[3, 17] = 1.0
    b.set_conductivity(KR)
    info = b.update_solution(r_tol=1e-12, a_tol=0.0, max_it=100000)
    assert info["status"] == 0 and info["converged"].all()


def test_iteration_limit_is_a_warning(cube4):
    om, gm, d = cube4
    g = darcy(gm, d)
    info = g.update_solution(r_tol=1e-14, a_tol=0.0, max_it=3)
    assert info["status"] == G.NOT_CONVERGED and info["converged"] == 0
    assert "iteration limit" in G.lib().fl_last_error().decode()
    x = g.get_solution_vector()
    assert np.isfinite(x).all()
    b = G.DarcyBatch(g, 3)
    b.set_conductivity(cases.mc_conductivity(om, 3, seed=5))
    bi = b.update_solution(r_tol=1e-14, a_tol=0.0, max_it=3)
    assert bi["status"] == G.NOT_CONVERGED and not bi["converged"].any()
    info = g.update_solution(r_tol=1e-12, a_tol=0.0, max_it=100000)
    assert info["status"] == 0 and info["converged"] == 1


def test_nan_source_diverges(cube4):
    om, gm, d = cube4
    f = d["f"].copy()
    f[4] = np.nan
    g = darcy(gm, d, f=f)
    expect(G.E_DIVERGED, lambda: g.update_solution(r_tol=1e-12, a_tol=0.0, max_it=1000), "CG diverged")
