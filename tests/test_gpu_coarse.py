"""Hand-written batched Cholesky + inverse of the coarse space (coarse.hh)
against host inverses: orders that fill one tile, several tiles and need
padding; conditioning up to 1e6; a matrix that is not positive definite is
flagged and gets a zero inverse."""
import numpy as np
import pytest

import flowgpu as G

pytestmark = pytest.mark.gpu


def spd(rng, n, cond):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    a = (q * ev) @ q.T
    return 0.5 * (a + a.T)


@pytest.mark.parametrize("n", [1, 37, 64, 130, 300])
def test_coarse_inverse_matches_host(n):
    rng = np.random.default_rng(n)
    conds = [1e2, 1e4, 1e6]
    a = np.stack([spd(rng, n, c) for c in conds])
    print("n %d" % n, flush=True)
    inv, info = G.coarse_invert(a)
    assert (info == 0).all()
    for k, c in enumerate(conds):
        ref = np.linalg.inv(a[k])
        err = np.linalg.norm(inv[k] - ref) / np.linalg.norm(ref)
        res = np.linalg.norm(a[k] @ inv[k].astype(np.float64) - np.eye(n), 2)
        print("  cond %.0e: rel err %.2e, ||A X - I|| %.2e" % (c, err, res), flush=True)
        # FP32 rounding of an FP64-accurate inverse: a few ulp plus cond * eps64
        assert err <= 2e-7 + c * 1e-15
        assert np.allclose(inv[k], inv[k].T)


def test_coarse_inverse_aggregate_pattern():
    """A coarse-operator-like matrix: a 3D grid of 8^3 aggregates with
    26-neighbour coupling and two columns per aggregate (order 1024)."""
    rng = np.random.default_rng(5)
    k = 8
    idx = np.arange(k ** 3).reshape(k, k, k)
    n = 2 * k ** 3
    lap = np.zeros((k ** 3, k ** 3))
    for (i, j, l), v in np.ndenumerate(idx):
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dl in (-1, 0, 1):
                    a, b, c = i + di, j + dj, l + dl
                    if (di, dj, dl) != (0, 0, 0) and 0 <= a < k and 0 <= b < k and 0 <= c < k:
                        w = rng.uniform(0.5, 2.0)
                        lap[v, idx[a, b, c]] -= w
                        lap[idx[a, b, c], v] -= w
    lap -= np.diag(lap.sum(1))  # graph Laplacian (singular)
    lap += np.eye(k ** 3) * 1e-3
    blk = np.array([[1.0, 0.3], [0.3, 0.5]])
    a = np.kron(lap, blk)
    inv, info = G.coarse_invert(a)
    ref = np.linalg.inv(a)
    err = np.linalg.norm(inv[0] - ref) / np.linalg.norm(ref)
    print("order %d: rel err %.2e (cond %.1e)" % (n, err, np.linalg.cond(a)), flush=True)
    assert info[0] == 0 and err <= 1e-6


def test_coarse_inverse_flags_indefinite():
    rng = np.random.default_rng(7)
    good = spd(rng, 70, 1e3)
    bad = good.copy()
    bad[40, 40] = -5.0
    inv, info = G.coarse_invert(np.stack([good, bad, good]))
    assert list(info) == [0, 1, 0]
    assert not inv[1].any()
    ref = np.linalg.inv(good)
    for k in (0, 2):
        assert np.linalg.norm(inv[k] - ref) / np.linalg.norm(ref) <= 1e-6
