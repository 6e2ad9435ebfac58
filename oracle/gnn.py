"""CPU oracle of the learned coarse basis (test infrastructure only).

FP64 numpy restatement of the graph network the GPU library trains and runs
(svn-export_amd/csrc/gnn.hip), used by the tests to check the GPU forward pass,
loss and gradients, and itself checked against finite differences.  There is
no counterpart in the reference: the network produces the second column phi of
the two-level preconditioner's coarse space (constant + phi per aggregate).

Graph: the bipartite element-edge graph of the mesh (element e -- edge f when
f is one of e's condensed unknowns: its sides' edges and its vb edges).

    He0 = relu(Xe We0 + be0)            Hf0 = relu(Xf Wf0 + bf0)
    layer l = 1, 2, 3:
        Mf  = mean over the elements of f of He(l-1)
        Hf(l) = relu(Hf(l-1) A1_l + Mf A2_l + a_l)
        l < 3:  Me = mean over the edges of e of Hf(l)
                He(l) = relu(He(l-1) B1_l + Me B2_l + b_l)
    phi = Hf3 w_out + c_out

Loss (per realization, solution lambda of the edge system): the part of the
solution that the coarse columns {1, phi} of each aggregate cannot represent,
    L = sum_a ( |l_a|^2 - (l_a . p_a)^2 / |p_a|^2 ) / sum_a |l_a|^2
with l_a, p_a lambda and phi restricted to aggregate a, minus their means.
"""
import numpy as np

H = 64
NFE = 16  # element input features (padded)
NFF = 16  # edge input features (padded)


def param_shapes():
    s = [("We0", (NFE, H)), ("be0", (H,)), ("Wf0", (NFF, H)), ("bf0", (H,))]
    for l in range(1, 4):
        s += [("A1_%d" % l, (H, H)), ("A2_%d" % l, (H, H)), ("a_%d" % l, (H,))]
        if l < 3:
            s += [("B1_%d" % l, (H, H)), ("B2_%d" % l, (H, H)), ("b_%d" % l, (H,))]
    s += [("w_out", (H,)), ("c_out", (1,))]
    return s


def init_params(seed=0):
    """He-style initialisation; the same order and shapes as the GPU's flat parameter vector."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, shp in param_shapes():
        if len(shp) == 2:
            P[name] = rng.standard_normal(shp) * np.sqrt(2.0 / shp[0])
        else:
            P[name] = np.zeros(shp)
    P["w_out"] = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return P


def flatten(P):
    return np.concatenate([np.asarray(P[n], np.float64).ravel() for n, _ in param_shapes()])


def unflatten(v):
    P, o = {}, 0
    for n, shp in param_shapes():
        k = int(np.prod(shp))
        P[n] = np.asarray(v[o:o + k], np.float64).reshape(shp)
        o += k
    return P


class Graph:
    """Incidence of the element-edge graph: el_ptr/el_edge (edges of each
    element), edge_ptr/edge_el (elements of each edge)."""

    def __init__(self, el_ptr, el_edge, n_edges):
        self.el_ptr = np.asarray(el_ptr, np.int64)
        self.el_edge = np.asarray(el_edge, np.int64)
        self.E = len(self.el_ptr) - 1
        self.F = n_edges
        deg_e = np.diff(self.el_ptr)
        self.el_of = np.repeat(np.arange(self.E), deg_e)
        self.deg_e = deg_e.astype(np.float64)
        self.deg_f = np.bincount(self.el_edge, minlength=self.F).astype(np.float64)

    def mean_e2f(self, He):
        M = np.zeros((self.F, He.shape[1]))
        np.add.at(M, self.el_edge, He[self.el_of])
        return M / np.maximum(self.deg_f, 1)[:, None]

    def mean_f2e(self, Hf):
        M = np.zeros((self.E, Hf.shape[1]))
        np.add.at(M, self.el_of, Hf[self.el_edge])
        return M / self.deg_e[:, None]

    def mean_e2f_T(self, dM):  # adjoint of mean_e2f
        dHe = np.zeros((self.E, dM.shape[1]))
        np.add.at(dHe, self.el_of, (dM / np.maximum(self.deg_f, 1)[:, None])[self.el_edge])
        return dHe

    def mean_f2e_T(self, dM):  # adjoint of mean_f2e
        dHf = np.zeros((self.F, dM.shape[1]))
        np.add.at(dHf, self.el_edge, (dM / self.deg_e[:, None])[self.el_of])
        return dHf


def relu(x):
    return np.maximum(x, 0.0)


def forward(P, g, Xe, Xf):
    c = {}
    He = relu(Xe @ P["We0"] + P["be0"])
    Hf = relu(Xf @ P["Wf0"] + P["bf0"])
    c["He0"], c["Hf0"] = He, Hf
    for l in range(1, 4):
        Mf = g.mean_e2f(He)
        Hf_new = relu(Hf @ P["A1_%d" % l] + Mf @ P["A2_%d" % l] + P["a_%d" % l])
        c["Mf%d" % l], c["Hf%d" % l] = Mf, Hf_new
        if l < 3:
            Me = g.mean_f2e(Hf_new)
            He_new = relu(He @ P["B1_%d" % l] + Me @ P["B2_%d" % l] + P["b_%d" % l])
            c["Me%d" % l], c["He%d" % l] = Me, He_new
            He = He_new
        Hf = Hf_new
    phi = Hf @ P["w_out"] + P["c_out"][0]
    return phi, c


def backward(P, g, Xe, Xf, c, dphi):
    G = {n: np.zeros(s) for n, s in param_shapes()}
    Hf3 = c["Hf3"]
    G["w_out"] = Hf3.T @ dphi
    G["c_out"][0] = dphi.sum()
    dHf = {3: np.outer(dphi, P["w_out"])}
    dHe = {l: np.zeros((g.E, H)) for l in range(0, 3)}
    for l in (1, 2):
        dHf[l] = np.zeros((g.F, H))
    dHf[0] = np.zeros((g.F, H))
    for l in range(3, 0, -1):
        if l < 3:  # element update of layer l
            dZ = dHe[l] * (c["He%d" % l] > 0)
            G["B1_%d" % l] += c["He%d" % (l - 1)].T @ dZ
            G["B2_%d" % l] += c["Me%d" % l].T @ dZ
            G["b_%d" % l] += dZ.sum(0)
            dHe[l - 1] += dZ @ P["B1_%d" % l].T
            dHf[l] += g.mean_f2e_T(dZ @ P["B2_%d" % l].T)
        dZ = dHf[l] * (c["Hf%d" % l] > 0)
        G["A1_%d" % l] += c["Hf%d" % (l - 1)].T @ dZ
        G["A2_%d" % l] += c["Mf%d" % l].T @ dZ
        G["a_%d" % l] += dZ.sum(0)
        dHf[l - 1] += dZ @ P["A1_%d" % l].T
        dHe[l - 1] += g.mean_e2f_T(dZ @ P["A2_%d" % l].T)
    dZe = dHe[0] * (c["He0"] > 0)
    dZf = dHf[0] * (c["Hf0"] > 0)
    G["We0"] = Xe.T @ dZe
    G["be0"] = dZe.sum(0)
    G["Wf0"] = Xf.T @ dZf
    G["bf0"] = dZf.sum(0)
    return G


def coarse_fit_loss(phi, lam, agg, n_agg):
    """Loss and d loss / d phi (see the module docstring)."""
    cnt = np.bincount(agg, minlength=n_agg).astype(np.float64)
    lm = np.bincount(agg, weights=lam, minlength=n_agg) / np.maximum(cnt, 1)
    pm = np.bincount(agg, weights=phi, minlength=n_agg) / np.maximum(cnt, 1)
    lt, pt = lam - lm[agg], phi - pm[agg]
    ll = np.bincount(agg, weights=lt * lt, minlength=n_agg)
    cc = np.bincount(agg, weights=lt * pt, minlength=n_agg)
    nn = np.bincount(agg, weights=pt * pt, minlength=n_agg)
    tot = ll.sum()
    ok = nn > 1e-30
    res = ll - np.where(ok, cc * cc / np.where(ok, nn, 1.0), 0.0)
    L = res.sum() / tot
    a = np.where(ok, cc / np.where(ok, nn, 1.0), 0.0)
    dphi = (-2.0 * a[agg] * lt + 2.0 * (a * a)[agg] * pt) / tot
    return L, dphi


def element_features(logK, dim, log_cs):
    """[E, 16]: standardised ln K and its square, dimension one-hot, ln cs / 10, bias."""
    mu, sd = logK.mean(), max(logK.std(), 1e-12)
    k = (logK - mu) / sd
    X = np.zeros((len(logK), NFE))
    X[:, 0] = k
    X[:, 1] = k * k
    X[:, 2] = dim == 1
    X[:, 3] = dim == 2
    X[:, 4] = dim == 3
    X[:, 5] = log_cs / 10.0
    X[:, 6] = 1.0
    return X


def edge_features(pos, agg, n_agg, dirichlet):
    """[F, 16]: position relative to the aggregate centroid (scaled by the
    aggregate's RMS radius), absolute position, Dirichlet flag, bias."""
    cnt = np.bincount(agg, minlength=n_agg).astype(np.float64)
    cen = np.stack([np.bincount(agg, weights=pos[:, d], minlength=n_agg) for d in range(3)], 1) / \
        np.maximum(cnt, 1)[:, None]
    rel = pos - cen[agg]
    rad = np.sqrt(np.bincount(agg, weights=(rel * rel).sum(1), minlength=n_agg) / np.maximum(cnt, 1))
    rel = rel / np.maximum(rad[agg], 1e-30)[:, None]
    X = np.zeros((len(pos), NFF))
    X[:, 0:3] = rel
    X[:, 3:6] = pos
    X[:, 6] = dirichlet
    X[:, 7] = 1.0
    return X
