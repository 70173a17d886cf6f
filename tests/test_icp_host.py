"""Host-side pieces of the point-to-plane adjoint binding (no GPU needed):
the exponential map's left Jacobian and the chain rule that turns the
adjoint kernel's left-perturbation gradient into d cost / d xi.

The gradient the kernel returns is J^T r with J_i = [(x_i - c) x n_i ; n_i]
(icp.cpp:318-320's rows about a centre c).  Here that quantity is formed in
float64 from its definition and carried through `twist_gradient`; the result
must equal central finite differences of the cost itself at nonzero twists,
for the origin and for an off-origin centre.  Dropping the left Jacobian,
the (v - c) cross term or its sign fails these checks.
"""
import math

import numpy as np
import pytest

from conftest import rng
from surface_match_amd.icp import left_jacobian_so3, twist_gradient


def _rodrigues(w):
    t = np.linalg.norm(w)
    if t == 0:
        return np.eye(3)
    k = w / t
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _pose(xi, T0):
    U = np.eye(4)
    U[:3, :3] = _rodrigues(xi[:3])
    U[:3, 3] = xi[3:]
    return U @ T0


def _cost(xi, T0, x, q, n):
    T = _pose(xi, T0)
    y = x @ T[:3, :3].T + T[:3, 3]
    r = np.einsum("ij,ij->i", n, y - q)
    return 0.5 * float(r @ r)


def _left_grad(xi, T0, x, q, n, c):
    T = _pose(xi, T0)
    y = x @ T[:3, :3].T + T[:3, 3]
    r = np.einsum("ij,ij->i", n, y - q)
    J = np.concatenate([np.cross(y - c, n), n], 1)
    return J.T @ r


@pytest.fixture(scope="module")
def problem():
    r = rng(7)
    x = r.normal(size=(200, 3)) * 20.0
    n = r.normal(size=(200, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    T0 = np.eye(4)
    T0[:3, :3] = _rodrigues(np.array([0.3, -0.2, 0.5]))
    T0[:3, 3] = [5.0, -3.0, 2.0]
    q = x @ T0[:3, :3].T + T0[:3, 3] + r.normal(size=(200, 3)) * 0.5
    return x, q, n, T0


def test_left_jacobian_matches_finite_differences():
    """R(w + h e_k) R(w)^T = exp([Jl(w) h e_k]x) + O(h^2): the k-th column of
    Jl is the axis of that small rotation divided by h."""
    for w in (np.array([0.4, -1.1, 0.7]), np.array([1e-8, 0.0, 2e-8]), np.array([2.5, 0.1, -0.3])):
        Jl = left_jacobian_so3(w)
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            D = _rodrigues(w + e) @ _rodrigues(w - e).T
            axis = np.array([D[2, 1] - D[1, 2], D[0, 2] - D[2, 0], D[1, 0] - D[0, 1]]) / 2.0
            assert np.allclose(axis / (2 * h), Jl[:, k], atol=1e-7), (w, k)


@pytest.mark.parametrize("centre", [np.zeros(3), np.array([12.0, -7.0, 30.0])])
def test_twist_gradient_matches_finite_differences(problem, centre):
    x, q, n, T0 = problem
    for xi in (np.zeros(6), np.array([0.05, -0.12, 0.2, 1.5, -2.0, 0.7]),
               np.array([0.9, 0.3, -0.6, -4.0, 3.0, 10.0])):
        g = _left_grad(xi, T0, x, q, n, centre)
        got = twist_gradient(xi[None], g[None], centre[None])[0]
        fd = np.zeros(6)
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1e-6
            fd[k] = (_cost(xi + e, T0, x, q, n) - _cost(xi - e, T0, x, q, n)) / 2e-6
        assert np.allclose(got, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max()), (xi, got, fd)


def test_twist_gradient_without_chain_rule_is_wrong(problem):
    """The raw left-perturbation gradient is not d cost / d xi away from
    xi = 0: the check above is sensitive to the chain rule."""
    x, q, n, T0 = problem
    xi = np.array([0.9, 0.3, -0.6, -4.0, 3.0, 10.0])
    g = _left_grad(xi, T0, x, q, n, np.zeros(3))
    fd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1e-6
        fd[k] = (_cost(xi + e, T0, x, q, n) - _cost(xi - e, T0, x, q, n)) / 2e-6
    assert not np.allclose(g, fd, rtol=1e-3)
