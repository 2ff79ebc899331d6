"""Host tessellation of RenderMan patches into meshes (render.rs
src/geometry/patches.rs): bilinear and bicubic PatchMesh / Patch under
RiSpec basis matrices, and NuPatch NURBS surfaces by Cox-de Boor basis
functions (trim curves ignored, as render.rs does). Each returns
(positions (n, 3) f32, triangles (m, 3) u32, smooth normals (n, 3) f32, st
(n, 2) f32); grid vertices are shared across patch boundaries, so sheets
are welded. Patches are diced once at scene build; the GPU traces the
resulting meshes like any other.
"""
import numpy as np


def smooth_normals(positions, tris):
    """subdiv.rs smooth_normals: per vertex the normalised sum of the
    unnormalised face normals (f64) of the triangles using it."""
    p = np.asarray(positions, np.float64)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    n = np.cross(b - a, c - a)
    acc = np.zeros_like(p)
    for k in range(3):
        np.add.at(acc, t[:, k], n)
    ln = np.maximum(np.sqrt((acc * acc).sum(1)), 1e-12)
    return (acc / ln[:, None]).astype(np.float32)


def _grid_tris(gu, gv, wrap_u, wrap_v):
    """Two triangles (a, b, c) and (a, c, d) per grid cell, wrapping seams."""
    cu = gu if wrap_u else gu - 1
    cv = gv if wrap_v else gv - 1
    out = []
    for y in range(cv):
        for x in range(cu):
            a = y * gu + x
            b = y * gu + (x + 1) % gu
            c = ((y + 1) % gv) * gu + (x + 1) % gu
            d = ((y + 1) % gv) * gu + x
            out += [a, b, c, a, c, d]
    return np.asarray(out, np.uint32).reshape(-1, 3)


def _basis_weights(m, t):
    """[t^3 t^2 t 1] M for each t (patches.rs basis_weights)."""
    t = np.asarray(t, np.float64)
    tv = np.stack([t * t * t, t * t, t, np.ones_like(t)], -1)
    return tv @ np.asarray(m, np.float64)


def _sample_params(npatch, segs, wrap):
    """Grid samples of one direction: the patch index and the local
    parameter of each, over npatch patches of segs spans each."""
    g = npatch * segs + (0 if wrap else 1)
    s = np.arange(g)
    p = np.minimum(s // segs, npatch - 1)
    return p, s / segs - p, s / (npatch * segs), g


def tessellate_bicubic(points, nu, nv, u_wrap, v_wrap, basis_u, ustep, basis_v, vstep, segs):
    """patches.rs tessellate_bicubic; None for an invalid control grid."""
    pts = np.asarray(points, np.float64).ravel()
    if pts.size < nu * nv * 3 or (nu < 4 and not u_wrap) or (nv < 4 and not v_wrap):
        return None
    npu = nu // ustep if u_wrap else (nu - 4) // ustep + 1
    npv = nv // vstep if v_wrap else (nv - 4) // vstep + 1
    if npu <= 0 or npv <= 0:
        return None
    ctrl = pts[:nu * nv * 3].reshape(nv, nu, 3)
    pu, lu, su, gu = _sample_params(npu, segs, u_wrap)
    pv, lv, sv, gv = _sample_params(npv, segs, v_wrap)
    wu = _basis_weights(basis_u, lu)  # (gu, 4)
    wv = _basis_weights(basis_v, lv)  # (gv, 4)
    pos = np.zeros((gv, gu, 3))
    for jj in range(4):
        rows = (pv * vstep + jj) % nv
        for ii in range(4):
            cols = (pu * ustep + ii) % nu
            pos += ctrl[rows][:, cols] * (wv[:, jj][:, None, None] * wu[:, ii][None, :, None])
    positions = pos.reshape(-1, 3).astype(np.float32)
    st = np.stack(np.meshgrid(su, sv), -1).reshape(-1, 2).astype(np.float32)
    tris = _grid_tris(gu, gv, u_wrap, v_wrap)
    return positions, tris, smooth_normals(positions, tris), st


def tessellate_bilinear(points, nu, nv, u_wrap, v_wrap, segs):
    """patches.rs tessellate_bilinear."""
    pts = np.asarray(points, np.float64).ravel()
    if pts.size < nu * nv * 3 or nu < 2 or nv < 2:
        return None
    npu = nu if u_wrap else nu - 1
    npv = nv if v_wrap else nv - 1
    ctrl = pts[:nu * nv * 3].reshape(nv, nu, 3)
    pu, u, su, gu = _sample_params(npu, segs, u_wrap)
    pv, v, sv, gv = _sample_params(npv, segs, v_wrap)
    c = lambda i, j: ctrl[(j % nv)][:, (i % nu)]
    u_ = u[None, :, None]
    v_ = v[:, None, None]
    a = c(pu, pv) * (1 - u_) + c(pu + 1, pv) * u_
    b = c(pu, pv + 1) * (1 - u_) + c(pu + 1, pv + 1) * u_
    positions = (a * (1 - v_) + b * v_).reshape(-1, 3).astype(np.float32)
    st = np.stack(np.meshgrid(su, sv), -1).reshape(-1, 2).astype(np.float32)
    tris = _grid_tris(gu, gv, u_wrap, v_wrap)
    return positions, tris, smooth_normals(positions, tris), st


def bspline_basis(knots, n_ctrl, order, t):
    """patches.rs bspline_basis, step for step: from the knot span (clamped
    to the valid range) the recursion updates the values in place from the
    span downwards, so at each level a value reads its upper neighbour as
    already raised to that level. That is render.rs's tessellation, kept
    for identical surfaces; it is not the textbook Cox-de Boor basis (for
    order 2 on knots 0 0 1 1 it gives [t (1 - t), t], not [1 - t, t])."""
    degree = order - 1
    basis = [0.0] * (n_ctrl + degree + 1)
    hi = min(n_ctrl, max(len(knots) - order, 0))
    span = degree
    for i in range(degree, hi):
        if t < knots[i + 1] or i == hi - 1:
            span = i
            break
        span = i + 1
    basis[span] = 1.0
    k = lambda i: knots[i] if i < len(knots) else 0.0
    for p in range(1, degree + 1):
        for i in range(span, max(span - p, 0) - 1, -1):
            val = 0.0
            d1 = k(i + p) - knots[i]
            if d1 > 1e-12:
                val += basis[i] * (t - knots[i]) / d1
            d2 = k(i + p + 1) - k(i + 1)
            if d2 > 1e-12:
                val += basis[i + 1] * (knots[i + p + 1] - t) / d2
            basis[i] = val
    return basis[:n_ctrl]


def tessellate_nurbs(nu, uorder, uknot, umin, umax, nv, vorder, vknot, vmin, vmax, points, rational, segs_u,
                     segs_v):
    """patches.rs tessellate_nurbs: (segs_u + 1) x (segs_v + 1) samples of
    the surface over [umin, umax] x [vmin, vmax]; homogeneous "Pw" points
    divide by the summed weight."""
    stride = 4 if rational else 3
    pts = np.asarray(points, np.float64).ravel()
    if pts.size < nu * nv * stride or uorder < 1 or vorder < 1:
        return None
    ctrl = pts[:nu * nv * stride].reshape(nv, nu, stride)
    us = [umin + (umax - umin) * i / segs_u for i in range(segs_u + 1)]
    vs = [vmin + (vmax - vmin) * j / segs_v for j in range(segs_v + 1)]
    bu = np.asarray([bspline_basis(list(uknot), nu, uorder, u) for u in us])  # (gu, nu)
    bv = np.asarray([bspline_basis(list(vknot), nv, vorder, v) for v in vs])  # (gv, nv)
    xyz = np.einsum("gj,hi,jic->ghc", bv, bu, ctrl[..., :3])
    if rational:
        w = np.einsum("gj,hi,ji->gh", bv, bu, ctrl[..., 3])
        ok = np.abs(w) > 1e-12
        xyz = np.where(ok[..., None], xyz / np.where(ok, w, 1.0)[..., None], xyz)
    positions = xyz.reshape(-1, 3).astype(np.float32)
    su = np.arange(segs_u + 1) / segs_u
    sv = np.arange(segs_v + 1) / segs_v
    st = np.stack(np.meshgrid(su, sv), -1).reshape(-1, 2).astype(np.float32)
    tris = _grid_tris(segs_u + 1, segs_v + 1, False, False)
    return positions, tris, smooth_normals(positions, tris), st
