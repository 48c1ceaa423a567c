"""The oracle's intersectsBox (OrientedBox's separating-axis test restated in double) pinned by closed
forms (containment, distance, a touching face, a rotated slab across and beside a triangle) and by
agreement with an independent formulation (the triangle clipped to the box in the box frame) on
random boxes and triangles."""
import numpy as np
import pytest

import oracle_py as O

I4 = np.eye(4).T.reshape(-1)


def _mat(rot=None, t=(0, 0, 0), s=(1, 1, 1)):
    m = np.eye(4)
    r = np.eye(3) if rot is None else rot
    m[:3, :3] = r @ np.diag(s)
    m[:3, 3] = t
    return m.T.reshape(-1)  # column-major, as Matrix4.elements


def _rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


TRI = (np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0]))


@pytest.mark.parametrize("clip", [False, True])
def test_closed_forms(clip):
    f = lambda mn, mx, m: O.box_triangle(mn, mx, m, *TRI, clip=clip)
    assert f([0.1, 0.1, -0.1], [0.2, 0.2, 0.1], I4) == 1            # a small box through the face
    assert f([-5, -5, -5], [5, 5, 5], I4) == 1                      # the triangle inside the box
    assert f([2, 2, -1], [3, 3, 1], I4) == 0                        # beside it
    assert f([0.6, 0.6, -1], [0.9, 0.9, 1], I4) == 0                # past the hypotenuse (x + y > 1)
    assert f([0.2, 0.2, 0.0], [0.3, 0.3, 1.0], I4) == 1             # a face touching the triangle's plane
    assert f([0.2, 0.2, 1e-9], [0.3, 0.3, 1.0], I4) == 0            # just above it
    slab = _mat(_rot([1, 1, 0], 0.7), t=(0.3, 0.3, 0.0))
    assert f([-0.05, -2, -2], [0.05, 2, 2], slab) == 1              # a thin rotated slab across the face
    above = _mat(_rot([0, 0, 1], 0.3), t=(0.3, 0.3, 0.5))
    assert f([-1, -1, -0.1], [1, 1, 0.1], above) == 0               # a rotated box 0.4 above the plane


def test_sat_matches_clipping_on_random_cases():
    rng = np.random.default_rng(4)
    agree = hits = 0
    n = 20000
    for _ in range(n):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        m = _mat(R, t=rng.uniform(-1, 1, 3), s=rng.uniform(0.3, 2.0, 3))
        mn = rng.uniform(-1, 0, 3)
        mx = mn + rng.uniform(0.05, 1.0, 3)
        a, b, c = (rng.uniform(-1.5, 1.5, 3) for _ in range(3))
        s1 = O.box_triangle(mn, mx, m, a, b, c)
        s2 = O.box_triangle(mn, mx, m, a, b, c, clip=True)
        agree += s1 == s2
        hits += s1
    assert agree == n and 0.1 * n < hits < 0.9 * n
