"""Problem definitions of the parity cases: the reference's own test inputs
(tests/01, 06, 13 of Flow123d, fixtures under tests/golden/ref) and the
generated benchmark configurations.  Boundary data are given per boundary
element (file ones first, implicit ones after), as the reference's fields
are evaluated on boundary elements (darcy_flow_mh.cc:569-592)."""
import os

import numpy as np

import meshgen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref")


def ref_path(name):
    return os.path.join(GOLD, name)


def _region_label(mesh, rid):
    for r in mesh.regions:
        if r["id"] == rid:
            return r["label"]
    return None


def darcy_case(name, om, seed=0):
    """Field and BC arrays for case `name` on the oracle mesh `om`.
    Returns dict(K, cs, sigma, f, aniso, bc_type, bc_p, bc_flux, bc_sigma)."""
    E, NB = om.n_elements, om.n_bc_elements
    K, cs, sigma, f = np.ones(E), np.ones(E), np.ones(E), np.zeros(E)
    aniso = None
    bct, bcp, bcf, bcs = np.zeros(NB, np.int32), np.zeros(NB), np.zeros(NB), np.zeros(NB)
    cen = om.bc_centre.reshape(-1, 3)[:NB]
    lab = [_region_label(om, r) for r in om.elem_region]
    blab = [_region_label(om, r) for r in om.bc_region]
    if name == "test01":
        # tests/01_steady_flow_123d/flow_gmsh.con: anisotropy per region,
        # Dirichlet p = x - z on the four '.'-regions
        an = {"1d_channel": 10.0, "2d_fraction_1": 1.0, "2d_fraction_2": 1.0, "3d_cube": 0.1}
        aniso = np.zeros((E, 9))
        for e in range(E):
            a = an[lab[e]]
            aniso[e, [0, 4, 8]] = a
        for b in range(NB):
            if blab[b] in (".1d_channel", ".2d_fraction_1", ".2d_fraction_2", ".3d_cube"):
                bct[b] = 1
                bcp[b] = cen[b, 0] - cen[b, 2]
    elif name == "test06_32":
        # tests/06_flow_coupling/flow_32d.con; bc_piezo_head h -> pressure h - z
        for e in range(E):
            if lab[e] == "2d_fraction":
                K[e], sigma[e], cs[e], f[e] = 16.0, 1.0, 10.0, -2.0 / 10.0
            else:
                K[e] = 2.0
        for b in range(NB):
            if blab[b] == ".2d_fraction":
                bct[b] = 2
            elif blab[b] == ".3d_cube":
                bct[b] = 1
                bcp[b] = cen[b, 2] - cen[b, 2]
    elif name == "test06_21":
        # tests/06_flow_coupling/flow_21d.con
        for e in range(E):
            if lab[e] == "plane":
                K[e], cs[e] = 5.0, 10.0
            else:
                K[e], cs[e], sigma[e], f[e] = 50.0, 20.0, 2.0, -1.0 / 20.0 * (10.0 * 5.0)
        for b in range(NB):
            if blab[b] == ".plane":
                bct[b] = 1
                bcp[b] = cen[b, 0] - cen[b, 2]
            elif blab[b] == ".channel":
                bct[b] = 2
    elif name.startswith("test05_"):
        # tests/05_flow_boundary_conditions/flow_{dirichlet,neumann,robin}.con: exact
        # solution u = xy; formula data evaluated at the boundary element centre
        # (darcy_flow_mh.cc:577-589)
        variant = name[len("test05_"):]
        for b in range(NB):
            x, y = cen[b, 0], cen[b, 1]
            if blab[b] == ".bc_east":
                bct[b], bcp[b] = 1, y
            elif blab[b] == ".bc_west":
                bct[b], bcp[b] = 1, 0.0
            elif variant == "dirichlet":
                bct[b], bcp[b] = 1, (0.0 if blab[b] == ".bc_south" else x)
            elif variant == "neumann":
                bct[b], bcf[b] = 2, (x if blab[b] == ".bc_south" else -x)
            elif variant == "robin":
                bct[b], bcs[b] = 3, 0.5
                bcp[b] = -2.0 * x if blab[b] == ".bc_south" else 3.0 * x
            else:
                raise KeyError(name)
    elif name == "test13":
        # tests/13_edge_solute_mixing: anisotropy per region, old .fbc Dirichlet data
        an = {17: 1.0, 18: 0.1, 19: 1.0}
        aniso = np.zeros((E, 9))
        for e in range(E):
            aniso[e, [0, 4, 8]] = an[int(om.elem_region[e])]
        id2e = {int(i): e for e, i in enumerate(om.elem_id)}
        for (eid, sid, val) in TEST13_FBC:
            e = id2e[eid]
            bnd = om.elem_bnd[4 * e + sid]
            bce = om.bnd_bcele[bnd]
            bct[bce] = 1
            bcp[bce] = val
    elif name.startswith("square") or name.startswith("cube"):
        # generated configurations: lognormal K, Dirichlet head 1 at x=0 and 0 at x=1
        cen_e = om.centre.reshape(-1, 3)
        Kf = meshgen.lognormal_field(cen_e, sigma=1.0, corr_len=0.1, seed=seed)
        K[:] = Kf
        for e in range(E):
            if lab[e] == "fracture":
                K[e] *= 1e3
                cs[e] = 1e-3
            elif lab[e] == "channel":
                K[e] *= 1e3
                cs[e] = 1e-5
        for b in range(NB):
            if blab[b] == ".left":
                bct[b], bcp[b] = 1, 1.0 - cen[b, 2]
            elif blab[b] == ".right":
                bct[b], bcp[b] = 1, 0.0 - cen[b, 2]
    else:
        raise KeyError(name)
    return dict(K=K, cs=cs, sigma=sigma, f=f, aniso=aniso, bc_type=bct, bc_p=bcp, bc_flux=bcf, bc_sigma=bcs)


# tests/13_edge_solute_mixing/input/test13.fbc: (element id, side, Dirichlet value)
TEST13_FBC = [(20, 2, 10.0), (83, 1, -5.0), (135, 1, 5.0)]


def test13_bc_conc(om):
    """tests/13 .tbc: boundary condition 0 (element 20 side 2) has concentration 1."""
    bc = np.zeros(om.n_bc_elements)
    id2e = {int(i): e for e, i in enumerate(om.elem_id)}
    e = id2e[20]
    bc[om.bnd_bcele[om.elem_bnd[4 * e + 2]]] = 1.0
    return bc


def generated_mesh(kind, n, path, **kw):
    if kind == "square":
        m = meshgen.square_mesh(n)
    else:
        m = meshgen.cube_mesh(n, **kw)
    meshgen.write_gmsh(m, path)
    return m


def mc_conductivity(om, n_real, seed=100, sigma=(0.5, 3.0), corr=(0.05, 0.3), base=None):
    """Monte-Carlo conductivity realizations [n_real, E]: lognormal fields with
    sigma and correlation length drawn per realization, per-realization
    seeds; fracture / channel elements scaled by 1e3 as in darcy_case."""
    rng = np.random.default_rng(seed)
    cen_e = om.centre.reshape(-1, 3)
    lab = np.array([_region_label(om, r) for r in om.elem_region])
    scale = np.where((lab == "fracture") | (lab == "channel"), 1e3, 1.0)
    K = np.empty((n_real, om.n_elements))
    for r in range(n_real):
        s = rng.uniform(*sigma)
        ell = rng.uniform(*corr)
        K[r] = meshgen.lognormal_field(cen_e, sigma=s, corr_len=ell, seed=seed * 1000 + r, n_features=256) * scale
        if base is not None:
            K[r] *= base
    return K
