"""Pins the FP64 gradient oracle (oracle/oracle_grad.py): its forward maps agree with the C oracle's
ray-triangle / closest-point routines, and its analytic gradients agree with central finite
differences of those forward maps."""
import numpy as np

import oracle_grad as OG
import oracle_py as O


def _random_hits(rng, n):
    a = rng.standard_normal((n, 3))
    b = a + rng.standard_normal((n, 3))
    c = a + rng.standard_normal((n, 3))
    bc = rng.dirichlet([2.0, 2.0, 2.0], n)  # interior points
    p = bc[:, :1] * a + bc[:, 1:2] * b + bc[:, 2:3] * c
    d = rng.standard_normal((n, 3))
    nrm = np.cross(b - a, c - a)
    cos = np.abs(np.einsum("ij,ij->i", d, nrm)) / np.linalg.norm(d, axis=1) / np.linalg.norm(nrm, axis=1)
    keep = cos > 0.3  # well-conditioned hits only: finite differences are exact to O(h^2) there
    t = rng.uniform(0.5, 3.0, n)
    o = p - t[:, None] * d
    return o[keep], d[keep], a[keep], b[keep], c[keep]


def test_ray_forward_matches_c_oracle():
    rng = np.random.default_rng(1)
    o, d, a, b, c = _random_hits(rng, 300)
    dist, point, bary = OG.ray_forward(o, d, a, b, c)
    for i in range(len(o)):
        f = lambda x: np.asarray(x, np.float32)
        # the C oracle takes float32 inputs: compare on float32-representable data
        oi, di, ai, bi, ci = (f(x[i]).astype(np.float64) for x in (o, d, a, b, c))
        r = O.ray_triangle(oi, di, ai, bi, ci, backface_culling=False)
        assert r is not None
        _, pt = r
        dd, pp, bb = OG.ray_forward(oi[None], di[None], ai[None], bi[None], ci[None])
        assert np.allclose(pp[0], pt, rtol=1e-9, atol=1e-9)
        assert abs(dd[0] - np.linalg.norm(pt - oi)) <= 1e-9 * max(1.0, dd[0])
        bc = O.barycoord(pt, ai, bi, ci)
        assert bc is not None and np.allclose(bb[0], bc, rtol=1e-7, atol=1e-7)


def _fd(fun, args, k, h):
    """Central difference of scalar fun w.r.t. every component of args[k] (n, 3)."""
    out = np.zeros_like(args[k])
    for j in range(3):
        plus = [x.copy() for x in args]
        minus = [x.copy() for x in args]
        plus[k][:, j] += h
        minus[k][:, j] -= h
        out[:, j] = (fun(*plus) - fun(*minus)) / (2 * h)
    return out


def test_ray_grads_match_finite_differences():
    rng = np.random.default_rng(2)
    o, d, a, b, c = _random_hits(rng, 400)
    n = len(o)
    g_dist, g_point, g_bary = rng.standard_normal(n), rng.standard_normal((n, 3)), rng.standard_normal((n, 3))

    def loss(o, d, a, b, c):
        dist, point, bary = OG.ray_forward(o, d, a, b, c)
        return g_dist * dist + np.einsum("ij,ij->i", g_point, point) + np.einsum("ij,ij->i", g_bary, bary)

    g = OG.ray_grads(o, d, a, b, c, g_dist, g_point, g_bary)
    args = [o, d, a, b, c]
    for k, name in enumerate(["origins", "directions", "a", "b", "c"]):
        fd = _fd(loss, args, k, 1e-6)
        err = np.linalg.norm(fd - g[name], axis=1) / np.maximum(np.linalg.norm(g[name], axis=1), 1e-3)
        assert err.max() < 1e-5, (name, err.max())
    # a planted mistake in the adjoint (transposed system) is caught by the same check
    M = OG._ray_system(d, a, b, c)
    dlen = np.linalg.norm(d, axis=1)
    gx = np.stack([g_dist * dlen + np.einsum("ij,ij->i", g_point, d), g_bary[:, 1] - g_bary[:, 0],
                   g_bary[:, 2] - g_bary[:, 0]], axis=1)
    wrong = np.linalg.solve(M, gx[:, :, None])[:, :, 0] + g_point
    fd = _fd(loss, args, 0, 1e-6)
    assert (np.linalg.norm(fd - wrong, axis=1) / np.linalg.norm(fd, axis=1)).max() > 1e-2


def test_closest_on_triangle_matches_c_oracle():
    rng = np.random.default_rng(3)
    n = 2000
    a = rng.standard_normal((n, 3))
    b = a + rng.standard_normal((n, 3))
    c = a + rng.standard_normal((n, 3))
    q = a + 2.0 * rng.standard_normal((n, 3))
    p, bary = OG.closest_on_triangle(q, a, b, c)
    regions = set()
    for i in range(n):
        ref = O.closest_point_triangle(q[i], a[i], b[i], c[i])
        assert np.allclose(p[i], ref, rtol=1e-12, atol=1e-12), i
        regions.add(tuple(np.round(bary[i], 12) == 0) + tuple(np.round(bary[i], 12) == 1))
    assert len(regions) == 7  # every Voronoi region was exercised
    assert np.allclose(bary.sum(axis=1), 1.0)


def test_closest_grads_match_finite_differences():
    rng = np.random.default_rng(4)
    n = 1500
    a = rng.standard_normal((n, 3))
    b = a + rng.standard_normal((n, 3))
    c = a + rng.standard_normal((n, 3))
    q = a + 2.0 * rng.standard_normal((n, 3))
    D = OG.closest_distance(q, a, b, c)
    keep = D > 0.05
    q, a, b, c = q[keep], a[keep], b[keep], c[keep]
    g_dist = rng.standard_normal(len(q))
    loss = lambda q, a, b, c: g_dist * OG.closest_distance(q, a, b, c)
    g = OG.closest_grads(q, a, b, c, g_dist)
    for k, name in enumerate(["points", "a", "b", "c"]):
        fd = _fd(loss, [q, a, b, c], k, 1e-6)
        err = np.linalg.norm(fd - g[name], axis=1) / np.maximum(np.linalg.norm(g[name], axis=1), 1e-3)
        assert np.quantile(err, 0.999) < 1e-5 and err.max() < 1e-3, (name, err.max())
    # translation invariance: moving q and the triangle together leaves D unchanged, so the
    # gradients sum to zero
    tot = g["points"] + g["a"] + g["b"] + g["c"]
    assert np.abs(tot).max() < 1e-12
