"""Seeded generators for the benchmark and test inputs.

* ``square_mesh(n)``  -- unit square, n x n cells, each split into two P1
  triangles along the '/' diagonal; boundary lines ``.left`` (x=0) and
  ``.right`` (x=1) come first in the element list, the rest of the boundary
  falls to IMPLICIT BOUNDARY (no flow).  n=100 gives 20,000 triangles and
  30,200 edges.
* ``cube_mesh(n, n_fractures, n_channels)`` -- unit cube, n^3 cells, each split
  into 6 tetrahedra (Kuhn / Freudenthal split along the (0,0,0)-(1,1,1)
  diagonal, conforming across cells); planar fractures are whole grid planes
  meshed conformingly by the tetrahedron faces they contain (2D elements);
  channels are random walks along grid lines inside fracture planes (1D
  elements lying on fracture-triangle sides, so they couple to the fractures).
  Boundary triangles ``.left`` (x=0) and ``.right`` (x=1) come first.
* ``lognormal_field(points, sigma, corr_len, seed)`` -- ln K Gaussian with
  mean 0 and exponential covariance sigma^2 exp(-r/corr_len), sampled at
  arbitrary points by random Fourier features (frequencies drawn from the
  exponential covariance's spectral density, a multivariate Student-t with one
  degree of freedom scaled by 1/corr_len).
* ``write_gmsh(mesh, path, element_data=None)`` -- gmsh 2.2 ASCII writer in the
  layout read by the reference's GmshMeshReader (msh_gmshreader.cc:76-246).

A mesh is a dict of numpy arrays: ``nodes`` [n,3] float64, ``elem_type``
(gmsh type: 1 line, 2 triangle, 4 tetrahedron), ``elem_region`` (physical
id), ``elem_nodes`` [n,4] (0-based node index, -1 padded), and
``physical`` = list of (dim, id, name).  Node ids and element ids are 1-based
positions.
"""
import numpy as np

REGION_LEFT, REGION_RIGHT = 101, 102
REGION_ROCK, REGION_FRACTURE, REGION_CHANNEL = 1, 2, 3


def square_mesh(n):
    h = 1.0 / n
    xs = np.arange(n + 1) * h
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)

    def nid(i, j):  # i: x index, j: y index
        return j * (n + 1) + i

    j = np.arange(n)
    left = np.stack([nid(0, j), nid(0, j + 1)], 1)
    right = np.stack([nid(n, j), nid(n, j + 1)], 1)
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    I, J = I.ravel(), J.ravel()
    a, b, c, d = nid(I, J), nid(I + 1, J), nid(I + 1, J + 1), nid(I, J + 1)
    t1 = np.stack([a, b, c], 1)
    t2 = np.stack([a, c, d], 1)
    tris = np.stack([t1, t2], 1).reshape(-1, 3)
    ne_l = 2 * n
    en = np.full((ne_l + len(tris), 4), -1, np.int64)
    en[:n, :2] = left
    en[n:2 * n, :2] = right
    en[ne_l:, :3] = tris
    etype = np.concatenate([np.ones(ne_l, np.int32), np.full(len(tris), 2, np.int32)])
    ereg = np.concatenate([np.full(n, REGION_LEFT), np.full(n, REGION_RIGHT), np.full(len(tris), REGION_ROCK)]).astype(np.int32)
    phys = [(1, REGION_LEFT, ".left"), (1, REGION_RIGHT, ".right"), (2, REGION_ROCK, "plane")]
    return dict(nodes=nodes, elem_type=etype, elem_region=ereg, elem_nodes=en, physical=phys)


# Kuhn split of the unit cube: each permutation p of the axes gives the tetrahedron
# v0=(0,0,0), v1=v0+e_p0, v2=v1+e_p1, v3=(1,1,1).
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def cube_mesh(n, n_fractures=8, n_channels=200, channel_len=25, seed=0, fracture_extent=None):
    """Unit cube tetrahedral mesh with conforming fractures and channels.

    fracture_extent: None for whole planes, else (lo, hi) fraction range of the
    random rectangular extent of each fracture in both in-plane directions.
    """
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    N1 = n + 1
    g = np.arange(N1) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def nid(i, j, k):
        return (i * N1 + j) * N1 + k

    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    tets = []
    for p in _PERMS:
        v = [np.zeros(3, int)]
        for ax in p:
            w = v[-1].copy()
            w[ax] += 1
            v.append(w)
        tets.append(np.stack([nid(I + q[0], J + q[1], K + q[2]) for q in v], 1))
    # element order: cell-major, 6 tets per cell
    tets = np.stack(tets, 1).reshape(-1, 4)

    def face_tris(axis, pos, a0, a1, b0, b1):
        """Triangles of the grid plane x_axis = pos over cells [a0,a1) x [b0,b1)
        of the two in-plane axes, split along the low-high diagonal (matches the
        Kuhn tetrahedron faces)."""
        u, w = [ax for ax in range(3) if ax != axis]
        A, B = np.meshgrid(np.arange(a0, a1), np.arange(b0, b1), indexing="ij")
        A, B = A.ravel(), B.ravel()

        def node(da, db):
            idx = [None] * 3
            idx[axis] = np.full(A.shape, pos)
            idx[u] = A + da
            idx[w] = B + db
            return nid(idx[0], idx[1], idx[2])

        t1 = np.stack([node(0, 0), node(1, 0), node(1, 1)], 1)
        t2 = np.stack([node(0, 0), node(0, 1), node(1, 1)], 1)
        return np.stack([t1, t2], 1).reshape(-1, 3)

    left = face_tris(0, 0, 0, n, 0, n)
    right = face_tris(0, n, 0, n, 0, n)

    fr_tris, fr_planes = [], []
    used = set()
    while len(fr_planes) < n_fractures:
        axis = int(rng.integers(0, 3))
        pos = int(rng.integers(2, n - 1))
        if (axis, pos) in used:
            continue
        used.add((axis, pos))
        if fracture_extent is None:
            a0, a1, b0, b1 = 0, n, 0, n
        else:
            la = int(round(n * rng.uniform(*fracture_extent)))
            lb = int(round(n * rng.uniform(*fracture_extent)))
            a0 = int(rng.integers(0, n - la + 1)); b0 = int(rng.integers(0, n - lb + 1))
            a1, b1 = a0 + la, b0 + lb
        fr_tris.append(face_tris(axis, pos, a0, a1, b0, b1))
        fr_planes.append((axis, pos, a0, a1, b0, b1))
    fr_tris = np.concatenate(fr_tris) if fr_tris else np.zeros((0, 3), int)
    # remove duplicate fracture triangles (none for distinct planes)
    segs = set()
    seg_list = []
    for _ in range(n_channels):
        axis, pos, a0, a1, b0, b1 = fr_planes[int(rng.integers(0, len(fr_planes)))]
        u, w = [ax for ax in range(3) if ax != axis]
        cur = [0, 0, 0]
        cur[axis] = pos
        cur[u] = int(rng.integers(a0, a1 + 1))
        cur[w] = int(rng.integers(b0, b1 + 1))
        lo = {u: a0, w: b0}
        hi = {u: a1, w: b1}
        d_ax = u if rng.random() < 0.5 else w
        d_sg = 1 if rng.random() < 0.5 else -1
        for _s in range(channel_len):
            if rng.random() < 0.2:  # turn
                d_ax = w if d_ax == u else u
                d_sg = 1 if rng.random() < 0.5 else -1
            nxt = list(cur)
            nxt[d_ax] += d_sg
            if nxt[d_ax] < lo[d_ax] or nxt[d_ax] > hi[d_ax]:
                d_sg = -d_sg
                nxt = list(cur)
                nxt[d_ax] += d_sg
                if nxt[d_ax] < lo[d_ax] or nxt[d_ax] > hi[d_ax]:
                    break
            a_, b_ = nid(*cur), nid(*nxt)
            key = (min(a_, b_), max(a_, b_))
            if key not in segs:
                segs.add(key)
                seg_list.append((a_, b_))
            cur = nxt
    segs_arr = np.array(seg_list, np.int64).reshape(-1, 2)

    nb = len(left) + len(right)
    parts_nodes = [left, right, tets, fr_tris, segs_arr]
    counts = [len(p) for p in parts_nodes]
    en = np.full((sum(counts), 4), -1, np.int64)
    o = 0
    types, regs = [], []
    for p, t, r in zip(parts_nodes, [2, 2, 4, 2, 1], [REGION_LEFT, REGION_RIGHT, REGION_ROCK, REGION_FRACTURE, REGION_CHANNEL]):
        en[o:o + len(p), :p.shape[1]] = p
        o += len(p)
        types.append(np.full(len(p), t, np.int32))
        regs.append(np.full(len(p), r, np.int32))
    phys = [(2, REGION_LEFT, ".left"), (2, REGION_RIGHT, ".right"), (3, REGION_ROCK, "rock"),
            (2, REGION_FRACTURE, "fracture"), (1, REGION_CHANNEL, "channel")]
    return dict(nodes=nodes, elem_type=np.concatenate(types), elem_region=np.concatenate(regs), elem_nodes=en,
                physical=phys, n_boundary=nb, fractures=fr_planes)


def bulk_centroids(mesh):
    """Centroids of the bulk (non-boundary) elements in file order."""
    bnd_ids = {pid for (_, pid, name) in mesh["physical"] if name.startswith(".")}
    bulk = ~np.isin(mesh["elem_region"], list(bnd_ids))
    en = mesh["elem_nodes"][bulk]
    cnt = (en >= 0).sum(1)
    P = mesh["nodes"]
    s = np.zeros((len(en), 3))
    for k in range(4):
        m = en[:, k] >= 0
        s[m] += P[en[m, k]]
    return s / cnt[:, None], bulk


def lognormal_field(points, sigma=1.0, corr_len=0.1, seed=0, n_features=1024, dim=None):
    """exp(G) with G a zero-mean Gaussian field of covariance sigma^2 exp(-r/corr_len)."""
    rng = np.random.default_rng(seed)
    pts = np.asarray(points, np.float64)
    d = pts.shape[1] if dim is None else dim
    pts = pts[:, :d]
    z = rng.standard_normal((n_features, d))
    w = rng.chisquare(1, n_features)
    omega = z / (corr_len * np.sqrt(w))[:, None]
    phi = rng.uniform(0, 2 * np.pi, n_features)
    out = np.empty(len(pts))
    amp = sigma * np.sqrt(2.0 / n_features)
    chunk = max(1, (1 << 24) // n_features)
    for s in range(0, len(pts), chunk):
        arg = pts[s:s + chunk] @ omega.T + phi
        out[s:s + chunk] = amp * np.cos(arg).sum(1)
    return np.exp(out)


def write_gmsh(mesh, path, element_data=None):
    """gmsh 2.2 ASCII.  element_data: optional dict name -> per-bulk-element values
    written as $ElementData (msh_gmshreader.cc:306-397 layout)."""
    P = mesh["nodes"]
    en = mesh["elem_nodes"]
    et = mesh["elem_type"]
    er = mesh["elem_region"]
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write("$PhysicalNames\n%d\n" % len(mesh["physical"]))
        for d, i, name in mesh["physical"]:
            f.write('%d %d "%s"\n' % (d, i, name))
        f.write("$EndPhysicalNames\n$Nodes\n%d\n" % len(P))
        ids = np.arange(1, len(P) + 1)
        np.savetxt(f, np.column_stack([ids, P]), fmt="%d %.17g %.17g %.17g")
        f.write("$EndNodes\n$Elements\n%d\n" % len(et))
        nn = {1: 2, 2: 3, 4: 4}
        eid = np.arange(1, len(et) + 1)
        # contiguous runs of one element type, written block-wise
        cut = np.flatnonzero(np.diff(et)) + 1
        for a, b in zip(np.r_[0, cut], np.r_[cut, len(et)]):
            k = nn[int(et[a])]
            blk = np.column_stack([eid[a:b], et[a:b], np.full(b - a, 2), er[a:b], er[a:b], en[a:b, :k] + 1])
            np.savetxt(f, blk, fmt="%d")
        f.write("$EndElements\n")
        if element_data:
            _, bulk = bulk_centroids(mesh)
            bulk_ids = eid[bulk]
            for name, vals in element_data.items():
                f.write('$ElementData\n1\n"%s"\n1\n0.0\n3\n0\n1\n%d\n' % (name, len(bulk_ids)))
                np.savetxt(f, np.column_stack([bulk_ids, vals]), fmt="%d %.17g")
                f.write("$EndElementData\n")
