"""Host-side checks that need no GPU: the product's topology and numbering
(built by libflow123d_gpu.so's own host module) equal the oracle's
restatement of Mesh::make_neighbours_and_edges bit-for-bit, and the C ABI
library exports every symbol declared in include/flow123d_gpu.h."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases
import flowgpu as G
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TOPO = ["elem_dim", "elem_id", "elem_region", "elem_nodes", "elem_edge", "elem_bnd", "edge_ptr", "edge_side",
        "vb_elem", "vb_edge", "bnd_edge", "bnd_bcele", "bc_region"]
COUNTS = ["n_nodes", "n_elements", "n_bc_elements", "n_sides", "n_edges", "n_boundaries", "n_vb", "n_regions",
          "n_file_bc"]


def compare_topology(path):
    a, b = G.Mesh(path), O.Mesh(path)
    for k in COUNTS:
        assert getattr(a, k) == getattr(b, k), k
    for k in TOPO:
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape and np.array_equal(x, y), k
    assert a.regions == b.regions


@pytest.mark.parametrize("mesh", ["simplest_cube.msh", "test01_mesh.msh", "test13.msh", "test6_32d.msh",
                                  "test6_21d.msh"])
def test_topology_reference_meshes(mesh):
    compare_topology(cases.ref_path(mesh))


@pytest.mark.parametrize("kind,n,kw", [("square", 30, {}),
                                       ("cube", 7, dict(n_fractures=4, n_channels=12, channel_len=6, seed=9)),
                                       ("cube", 6, dict(n_fractures=3, n_channels=5, channel_len=8, seed=2,
                                                        fracture_extent=(0.3, 0.8)))])
def test_topology_generated(tmp_path, kind, n, kw):
    path = str(tmp_path / "m.msh")
    cases.generated_mesh(kind, n, path, **kw)
    compare_topology(path)


def test_mesh_from_arrays_equals_gmsh(tmp_path):
    import meshgen
    m = meshgen.cube_mesh(4, n_fractures=2, n_channels=3, channel_len=3, seed=4)
    path = str(tmp_path / "m.msh")
    meshgen.write_gmsh(m, path)
    a, b = G.Mesh(path), G.Mesh(arrays=m)
    for k in TOPO:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "flow123d_gpu.h")).read()
    names = sorted(set(re.findall(r"\b(fl_[a-z_]+)\s*\(", hdr)))
    assert len(names) >= 25
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares the same set
    assert set(G.SIGNATURES) == set(names)


def test_no_device_fails_loudly():
    """Without a GPU the solver handles refuse to be created (no CPU fallback)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    m = G.Mesh(cases.ref_path("simplest_cube.msh"))
    with pytest.raises(G.FlowError):
        G.DarcyFlowMH_Steady(m)
