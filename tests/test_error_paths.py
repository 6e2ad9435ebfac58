"""Failure cases of the reference on the CPU: a singular local matrix stops
the solve (src/flow/mh_fe_values.cc:54-59, NUM_ZERO of src/system/math_fce.h:51
in the inverses of math_fce.cc:147,168,211), restated by the oracle; the
status codes the library reports for them (include/flow123d_gpu.h)."""
import os
import re

import numpy as np
import pytest

import cases
import flowgpu as G
import meshgen
import oracle as O

NUM_ZERO = np.finfo(np.float64).tiny / np.finfo(np.float64).eps
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def flat_square(n=4):
    """square mesh whose interior node 6 is moved onto the midpoint of two of
    its neighbours, so that the triangle of the three is flat"""
    m = meshgen.square_mesh(n)
    nodes = m["nodes"].copy()
    en = m["elem_nodes"]
    tri = [e for e in range(len(en)) if m["elem_type"][e] == 2 and 6 in en[e][:3]][0]
    a, b = [v for v in en[tri][:3] if v != 6]
    nodes[6] = 0.5 * (nodes[a] + nodes[b])
    m["nodes"] = nodes
    return m, tri


def test_flat_triangle_gives_nan_determinant():
    """A zero-area element: its RT0 local matrix divides by the measure, so
    the determinant is not finite and the reference's |det| < NUM_ZERO test
    (false for NaN) would let it pass; the library checks !(|det| >= NUM_ZERO)."""
    m, tri = flat_square()
    om = O.Mesh(arrays=m)
    bulk = np.flatnonzero(m["elem_type"] == 2)
    e = int(np.searchsorted(bulk, tri))
    _, _, det = om.local_matrix(e)
    assert not np.isfinite(det)
    dets = [om.local_matrix(k)[2] for k in range(om.n_elements) if k != e]
    assert all(np.isfinite(d) and abs(d) >= NUM_ZERO for d in dets)


def test_oracle_stops_on_underflowing_determinant(tmp_path):
    """K = 1e80 on one tetrahedron scales its local matrix by 1e-80, the 4x4
    determinant underflows below NUM_ZERO and the solve stops naming the
    element (the reference's xprintf(Err) there); K = 1e60 still passes."""
    path = str(tmp_path / "cube2.msh")
    cases.generated_mesh("cube", 2, path, n_fractures=0, n_channels=0)
    om = O.Mesh(path)
    d = cases.darcy_case("cube", om, seed=0)
    e0 = int(np.flatnonzero(om.elem_dim == 3)[5])
    K = d["K"].copy()
    K[e0] = 1e80
    assert abs(om.local_matrix(e0, K=1e80)[2]) < NUM_ZERO
    D = O.Darcy(om, K, d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    with pytest.raises(RuntimeError, match="Singular local matrix of the element %d$" % om.elem_id[e0]):
        D.solve(0)
    K[e0] = 1e60
    assert abs(om.local_matrix(e0, K=1e60)[2]) >= NUM_ZERO
    D = O.Darcy(om, K, d["cs"], d["sigma"], d["f"], d["bc_type"], d["bc_p"], d["bc_flux"], d["bc_sigma"])
    D.solve(0)


def test_status_codes_match_header():
    txt = open(os.path.join(ROOT, "include", "flow123d_gpu.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define FL_(\w+) \(?(-?\d+)\)?", txt)}
    assert defs == {"NOT_CONVERGED": G.NOT_CONVERGED, "E_SINGULAR_ELEMENT": G.E_SINGULAR_ELEMENT,
                    "E_BAD_FIELD": G.E_BAD_FIELD, "E_DIVERGED": G.E_DIVERGED, "E_CALLBACK": G.E_CALLBACK,
                    "E_SORPTION": G.E_SORPTION}
    # and the library's own enumeration
    src = open(os.path.join(ROOT, "svn-export_amd", "csrc", "common.hh")).read()
    assert {int(v) for v in re.findall(r"kErr\w+ = (-\d+)", src)} == {v for v in defs.values() if v < 0}
