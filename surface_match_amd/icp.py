"""Batched ICP refinement and the point-to-plane adjoint on the GPU.

``IcpTarget`` mirrors ``ICP::regist(src, dst, kdtree, initPoses)``
(icp.h:86-90, icp.cpp:214-278) with the target gridded once.  ``refine``
runs either the batched persistent launch (all hypotheses in one kernel,
replayed from a captured hipGraph) or the conventional per-hypothesis
refiner.  ``adjoint`` returns the point-to-plane cost, its se(3) gradient
and the Gauss-Newton J^T J for many poses at once.

``PointToPlaneCost`` (needs torch) wraps ``adjoint`` as a
torch.autograd.Function of the pose twist, evaluated on host tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _pose16(T):
    T = np.asarray(T, np.float32)
    if T.shape[-2:] == (3, 4):
        T4 = np.zeros(T.shape[:-2] + (4, 4), np.float32)
        T4[..., :3, :] = T
        T4[..., 3, 3] = 1.0
        T = T4
    # row-major 4x4 -> column-major 16 floats
    return np.ascontiguousarray(np.swapaxes(T, -1, -2).reshape(-1, 16), np.float32)


class IcpTarget:
    def __init__(self, pts, nrm, cell: float = 0.0):
        self._pts = N.f32(pts)
        self._nrm = N.f32(nrm)
        h = C.c_void_p()
        N.check(N.lib().ppf_icp_target_create(N.ptr(self._pts), N.ptr(self._nrm), len(self._pts),
                                              cell, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().ppf_icp_target_destroy(h)
            self._h = None

    def refine(self, src, src_nrm, init_poses, iterations=10, reject_dist=10.0, mse_min=0.0,
               mse_max=1e30, tolerance=0.0, mode=1, batched=True, block=0, scoring_dist=1.0):
        """Returns (poses [H,4,4], results dict of arrays, device ms).

        batched: True / "spread" = every pose spread over workgroups of
        `block` source points; "persistent" = one workgroup per pose;
        False = the conventional one-pose-at-a-time refiner."""
        src, src_nrm = N.f32(src), N.f32(src_nrm)
        P = _pose16(init_poses)
        H = len(P)
        crit = N.IcpCriteria(iterations, reject_dist, mse_min, mse_max, tolerance)
        out = (N.IcpResult * H)()
        ms = C.c_float(0)
        N.check(N.lib().ppf_icp_refine(self._h, N.ptr(src), N.ptr(src_nrm), len(src), N.ptr(P), H,
                                       C.byref(crit), mode, _batched_code(batched), block, scoring_dist,
                                       out, C.byref(ms)))
        poses = np.array([np.array(o.pose16, np.float32).reshape(4, 4).T for o in out])
        res = {k: np.array([getattr(o, k) for o in out]) for k in
               ("type", "mse", "convergeRate", "iterations", "converged", "inliers")}
        return poses, res, ms.value

    def adjoint(self, src, poses, xi=None, reject_dist=10.0, center=1, with_corr=False):
        """Per pose: dict(cost, count, center, grad[6], JtJ[6x6]) at T = U(xi) T0."""
        src = N.f32(src)
        P = _pose16(poses)
        H = len(P)
        xi = np.zeros((H, 6)) if xi is None else np.ascontiguousarray(xi, np.float64).reshape(H, 6)
        out = np.zeros((H, 36))
        corr = np.zeros((H, len(src)), np.int32) if with_corr else None
        N.check(N.lib().ppf_icp_adjoint(self._h, N.ptr(src), len(src), N.ptr(P), N.ptr(xi), H,
                                        reject_dist, center, N.ptr(out), N.ptr(corr)))
        JtJ = np.zeros((H, 6, 6))
        iu = np.triu_indices(6)
        JtJ[:, iu[0], iu[1]] = out[:, 11:32]
        JtJ[:, iu[1], iu[0]] = out[:, 11:32]
        r = dict(cost=out[:, 0], count=out[:, 1], center=out[:, 2:5], grad=out[:, 5:11], JtJ=JtJ)
        if with_corr:
            r["corr"] = corr
        return r


def refine_many(targets, pose_target, src, src_nrm, init_poses, iterations=10, reject_dist=10.0,
                mse_min=0.0, mse_max=1e30, tolerance=0.0, mode=1, batched=True, block=0,
                scoring_dist=1.0):
    """Pose h refined against targets[pose_target[h]] (IcpTargets of several
    scenes), all in one batched refinement (batched True / "spread" /
    "persistent") or one at a time (False); each result equals
    targets[pose_target[h]].refine of that pose alone.  Returns what
    IcpTarget.refine returns."""
    src, src_nrm = N.f32(src), N.f32(src_nrm)
    P = _pose16(init_poses)
    H = len(P)
    pt = np.ascontiguousarray(pose_target, np.int32).reshape(-1)
    if len(pt) != H:
        raise ValueError("one target index per pose")
    hs = (C.c_void_p * len(targets))(*[t._h.value for t in targets])
    crit = N.IcpCriteria(iterations, reject_dist, mse_min, mse_max, tolerance)
    out = (N.IcpResult * max(H, 1))()
    ms = C.c_float(0)
    N.check(N.lib().ppf_icp_refine_multi(hs, len(targets), N.ptr(pt), N.ptr(src), N.ptr(src_nrm),
                                         len(src), N.ptr(P), H, C.byref(crit), mode,
                                         _batched_code(batched), block, scoring_dist, out,
                                         C.byref(ms)))
    poses = np.array([np.array(out[h].pose16, np.float32).reshape(4, 4).T for h in range(H)])
    res = {k: np.array([getattr(out[h], k) for h in range(H)]) for k in
           ("type", "mse", "convergeRate", "iterations", "converged", "inliers")}
    return poses, res, ms.value


def _batched_code(b):
    if b is False or b == 0:
        return 0
    if b == "persistent":
        return 2
    if b is True or b == "spread" or b == 1:
        return 1
    raise ValueError(f"batched must be True/'spread', 'persistent' or False, not {b!r}")


def _skew(a):
    return np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])


def left_jacobian_so3(w):
    """Jl(w) of SO(3): R(w + dw) R(w)^T = exp([Jl(w) dw]x) to first order,
    Jl = I + (1 - cos t) / t^2 [w]x + (t - sin t) / t^3 [w]x^2, t = |w|."""
    w = np.asarray(w, np.float64)
    t = float(np.linalg.norm(w))
    K = _skew(w)
    if t < 1e-6:
        return np.eye(3) + 0.5 * K + K @ K / 6.0
    return np.eye(3) + (1.0 - np.cos(t)) / t ** 2 * K + (t - np.sin(t)) / t ** 3 * (K @ K)


def twist_gradient(xi, g, center):
    """d cost / d xi for T = [R(w), v] T0 from the adjoint kernel's g = J^T r,
    the gradient in a left perturbation [R(om), c - R(om) c + u] of T.  With
    R(w + dw) R(w)^T = I + [phi]x, phi = Jl(w) dw, the change of xi is that
    perturbation with om = phi and u = dv + (v - c) x phi, so
    d/dw = Jl(w)^T (g_om + g_u x (v - c)) and d/dv = g_u."""
    xi = np.asarray(xi, np.float64).reshape(-1, 6)
    g = np.asarray(g, np.float64).reshape(-1, 6)
    c = np.asarray(center, np.float64).reshape(-1, 3)
    out = np.empty_like(g)
    for h in range(len(xi)):
        w, v = xi[h, :3], xi[h, 3:]
        out[h, :3] = left_jacobian_so3(w).T @ (g[h, :3] + np.cross(g[h, 3:], v - c[h]))
        out[h, 3:] = g[h, 3:]
    return out


def point_to_plane_cost():
    """torch.autograd.Function: cost(xi) for H poses T = [R(w), v] T0 and its
    gradient d cost / d xi (the adjoint kernel's J^T r carried through the
    exponential map's Jacobian; exact for the correspondences at T)."""
    import torch

    class PointToPlaneCost(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xi, target: IcpTarget, src, poses, reject_dist=10.0, center=1):
            x = xi.detach().cpu().numpy().astype(np.float64)
            r = target.adjoint(src, poses, x, reject_dist, center)
            grad = twist_gradient(x, r["grad"], r["center"])
            ctx.save_for_backward(torch.from_numpy(grad).to(device=xi.device, dtype=xi.dtype))
            return torch.from_numpy(r["cost"]).to(device=xi.device, dtype=xi.dtype)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g[:, None] * grad, None, None, None, None, None

    return PointToPlaneCost
