"""The CPU oracle of the learned coarse basis, pinned by finite differences
and by closed forms of the coarse-fit loss."""
import numpy as np

import gnn as N


def small_graph(seed=0, E=30, F=45):
    rng = np.random.default_rng(seed)
    ptr, edges = [0], []
    for e in range(E):
        k = rng.integers(2, 5)
        edges += list(rng.choice(F, size=k, replace=False))
        ptr.append(len(edges))
    return N.Graph(ptr, edges, F)


def test_loss_closed_forms():
    rng = np.random.default_rng(1)
    F, na = 60, 4
    agg = rng.integers(0, na, F)
    lam = rng.standard_normal(F)
    # phi equal to lambda (up to an affine map per aggregate) represents it exactly
    L, _ = N.coarse_fit_loss(3.0 * lam + 7.0, lam, agg, na)
    assert abs(L) < 1e-12
    # phi constant on every aggregate represents nothing beyond the constant column
    L, d = N.coarse_fit_loss(np.ones(F), lam, agg, na)
    assert abs(L - 1.0) < 1e-12
    # brute force: least squares of lambda on [1, phi] per aggregate
    phi = rng.standard_normal(F)
    L, _ = N.coarse_fit_loss(phi, lam, agg, na)
    res = tot = 0.0
    for a in range(na):
        m = agg == a
        A = np.stack([np.ones(m.sum()), phi[m]], 1)
        c = np.linalg.lstsq(A, lam[m], rcond=None)[0]
        res += ((lam[m] - A @ c) ** 2).sum()
        tot += ((lam[m] - lam[m].mean()) ** 2).sum()
    assert abs(L - res / tot) < 1e-12


def test_loss_gradient_fd():
    rng = np.random.default_rng(2)
    F, na = 40, 3
    agg = rng.integers(0, na, F)
    lam, phi = rng.standard_normal(F), rng.standard_normal(F)
    L, d = N.coarse_fit_loss(phi, lam, agg, na)
    h = 1e-6
    for f in rng.choice(F, 8, replace=False):
        p = phi.copy()
        p[f] += h
        Lp, _ = N.coarse_fit_loss(p, lam, agg, na)
        p[f] -= 2 * h
        Lm, _ = N.coarse_fit_loss(p, lam, agg, na)
        assert abs((Lp - Lm) / (2 * h) - d[f]) < 1e-6 * max(1.0, abs(d[f]))


def test_mean_aggregation_adjoints():
    g = small_graph(3)
    rng = np.random.default_rng(3)
    He, Hf = rng.standard_normal((g.E, 5)), rng.standard_normal((g.F, 5))
    # <mean_e2f(He), Hf> == <He, mean_e2f_T(Hf)>
    assert abs((g.mean_e2f(He) * Hf).sum() - (He * g.mean_e2f_T(Hf)).sum()) < 1e-12
    assert abs((g.mean_f2e(Hf) * He).sum() - (Hf * g.mean_f2e_T(He)).sum()) < 1e-12


def test_network_gradient_fd():
    g = small_graph(4)
    rng = np.random.default_rng(4)
    Xe = rng.standard_normal((g.E, N.NFE))
    Xf = rng.standard_normal((g.F, N.NFF))
    agg = rng.integers(0, 3, g.F)
    lam = rng.standard_normal(g.F)
    P = N.init_params(5)
    phi, c = N.forward(P, g, Xe, Xf)
    L, dphi = N.coarse_fit_loss(phi, lam, agg, 3)
    G = N.backward(P, g, Xe, Xf, c, dphi)
    v, gv = N.flatten(P), N.flatten(G)
    h = 1e-6
    idx = rng.choice(len(v), 40, replace=False)
    for i in idx:
        w = v.copy()
        w[i] += h
        Lp = N.coarse_fit_loss(N.forward(N.unflatten(w), g, Xe, Xf)[0], lam, agg, 3)[0]
        w[i] -= 2 * h
        Lm = N.coarse_fit_loss(N.forward(N.unflatten(w), g, Xe, Xf)[0], lam, agg, 3)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - gv[i]) < 1e-5 * max(1.0, abs(fd)), (i, fd, gv[i])
