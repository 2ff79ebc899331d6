"""Host tessellation of RenderMan surfaces into meshes (render.rs
src/geometry/patches.rs, subdiv.rs, earclip.rs): bilinear and bicubic
PatchMesh / Patch under RiSpec basis matrices, NuPatch NURBS surfaces by
render.rs's basis recursion (trim curves ignored, as render.rs does),
Catmull-Clark subdivision surfaces, and planar polygons with holes
(`triangulate_loops`). Each returns
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


def _subdivide_once(pos, faces, creases, corners, holes, interp):
    """One Catmull-Clark round as subdiv.rs subdivide_once: face points,
    edge points (boundary or sharpness >= 1: the midpoint; semi-sharp: a lerp
    of the smooth point towards it), vertex points (corner >= 1: pinned;
    interpolateboundary with two boundary edges, or two creases: the (6P +
    A + B) / 8 crease rule, semi-sharp creases lerped from the smooth rule;
    three creases: pinned; else (F + 2M + (n - 3) P) / n, corners lerped);
    one quad per face corner, holes dropped; sharpnesses decay by one."""
    key = lambda a, b: (a, b) if a < b else (b, a)
    edge_faces, vedges, vfaces = {}, {}, {}
    for fi, f in enumerate(faces):
        for i in range(len(f)):
            a, b = f[i], f[(i + 1) % len(f)]
            k = key(a, b)
            edge_faces.setdefault(k, []).append(fi)
            for v in (a, b):
                lst = vedges.setdefault(v, [])
                if k not in lst:
                    lst.append(k)
            lst = vfaces.setdefault(a, [])
            if fi not in lst:
                lst.append(fi)
    crease = {key(a, b): max(s, 0.0) for a, b, s in creases}
    corner = {v: max(s, 0.0) for v, s in corners}
    hole = set(h for h in holes if h < len(faces))
    new = []
    fpi = []
    for f in faces:
        new.append(pos[f].mean(0) if len(f) else np.zeros(3))
        fpi.append(len(new) - 1)
    epi = {}
    for k, fl in edge_faces.items():
        a, b = pos[k[0]], pos[k[1]]
        mid = (a + b) * 0.5
        sharp = crease.get(k, 0.0)
        if len(fl) < 2 or sharp >= 1.0:
            pt = mid
        else:
            fp = sum(new[fpi[f]] for f in fl)
            smooth = (a + b + fp) * (1.0 / (2.0 + len(fl)))
            pt = smooth + (mid - smooth) * min(sharp, 1.0) if sharp > 0.0 else smooth
        new.append(pt)
        epi[k] = len(new) - 1
    vpi = []
    for v in range(len(pos)):
        p = pos[v]
        edges = vedges.get(v, [])
        vf = vfaces.get(v, [])
        bnd = [e for e in edges if len(edge_faces.get(e, ())) < 2]
        sharp_e = [(e, crease[e]) for e in edges if crease.get(e, 0.0) > 0.0]
        cs = corner.get(v, 0.0)
        other = lambda e: e[1] if e[0] == v else e[0]

        def smooth_point():
            n = max(len(edges), 1)
            favg = sum((new[fpi[f]] for f in vf), np.zeros(3)) / max(len(vf), 1)
            mavg = sum(((pos[e[0]] + pos[e[1]]) * 0.5 for e in edges), np.zeros(3)) / n
            return (favg + mavg * 2.0 + p * (n - 3.0)) / n

        def crease_point(pair):
            return (p * 6.0 + pos[other(pair[0])] + pos[other(pair[1])]) / 8.0

        if cs >= 1.0:
            pt = p
        elif interp and len(bnd) >= 2:
            pt = crease_point(bnd[:2])
        elif len(sharp_e) >= 3:
            pt = p
        elif len(sharp_e) == 2:
            avg = (sharp_e[0][1] + sharp_e[1][1]) * 0.5
            cp = crease_point([sharp_e[0][0], sharp_e[1][0]])
            pt = cp if avg >= 1.0 else smooth_point() + (cp - smooth_point()) * avg
        else:
            sp = smooth_point()
            pt = sp + (p - sp) * min(cs, 1.0) if cs > 0.0 else sp
        new.append(pt)
        vpi.append(len(new) - 1)
    nf = []
    for fi, f in enumerate(faces):
        if fi in hole:
            continue
        n = len(f)
        for i in range(n):
            prev, cur, nxt = f[(i + n - 1) % n], f[i], f[(i + 1) % n]
            nf.append([vpi[cur], epi[key(cur, nxt)], fpi[fi], epi[key(prev, cur)]])
    nc = []
    for a, b, s in creases:
        ns = s - 1.0
        if ns > 0.0 and key(a, b) in epi:
            ep = epi[key(a, b)]
            nc += [(vpi[a], ep, ns), (vpi[b], ep, ns)]
    ncor = [(vpi[v], s if s >= 1e9 else s - 1.0) for v, s in corners if s - 1.0 > 0.0 or s >= 1e9]
    return np.asarray(new, np.float64), nf, nc, ncor


def tessellate_subdiv(positions, faces, creases=(), corners=(), holes=(), interpolate_boundary=False, levels=3):
    """SubdivCage::tessellate (subdiv.rs:275-326): sharpnesses >= 10 become
    infinite, `levels` uniform Catmull-Clark rounds, faces fanned into
    triangles with smooth normals; no st."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    fl = [list(f) for f in faces]
    cr = [(a, b, 1e9 if s >= 10.0 else s) for a, b, s in creases]
    co = [(v, 1e9 if s >= 10.0 else s) for v, s in corners]
    ho = list(holes)
    for _ in range(levels):
        pos, fl, cr, co = _subdivide_once(pos, fl, cr, co, ho, interpolate_boundary)
        ho = []
    p32 = pos.astype(np.float32)
    tris = [[f[0], f[i], f[i + 1]] for f in fl for i in range(1, len(f) - 1)]
    tris = np.asarray(tris, np.uint32).reshape(-1, 3)
    return p32, tris, smooth_normals(p32, tris), None


def _cross2(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def _proper_cross(a, b, c, d):
    """Segments ab and cd cross at a point interior to both."""
    d1, d2 = _cross2(c, d, a), _cross2(c, d, b)
    d3, d4 = _cross2(a, b, c), _cross2(a, b, d)
    return d1 * d2 < 0.0 and d3 * d4 < 0.0


def triangulate_loops(points, loops):
    """Triangles covering a planar polygon given as loops of vertex indices
    into `points` (n, 3): the first loop is the outline, the others holes
    (GeneralPolygon, scene_builder.rs:1413-1446, earclip.rs). The loops are
    projected on the plane that drops the largest component of the
    outline's Newell normal, cyclically ((y, z), (z, x) or (x, y)), the
    outline turned counter-clockwise and the holes clockwise there; each
    hole, in order of its rightmost x, is joined to the ring by a two-way
    bridge from that rightmost vertex to the nearest ring vertex whose
    segment properly crosses no edge of the ring or of any hole; the ring
    is then cut into ears (convex corners whose triangle holds no other
    ring vertex, copies of its own corners aside), resuming the scan next
    to the last ear, and what is left when no ear is found is fanned.
    Triangles are (m, 3) indices into `points` (bridges reuse vertices),
    counter-clockwise in the projection as render.rs emits them."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    outline = [int(i) for i in loops[0]]
    nrm = np.zeros(3)
    for i in range(len(outline)):  # Newell's normal
        a, b = p[outline[i]], p[outline[(i + 1) % len(outline)]]
        nrm += ((a[1] - b[1]) * (a[2] + b[2]), (a[2] - b[2]) * (a[0] + b[0]), (a[0] - b[0]) * (a[1] + b[1]))
    ax = int(np.argmax(np.abs(nrm)))  # ties go to the lower axis
    q = p[:, [(ax + 1) % 3, (ax + 2) % 3]]

    def area(loop):
        xy = q[loop]
        return 0.5 * float(np.sum(xy[:, 0] * np.roll(xy[:, 1], -1) - np.roll(xy[:, 0], -1) * xy[:, 1]))

    ring = outline if area(outline) >= 0.0 else outline[::-1]
    holes = []
    for h in loops[1:]:
        h = [int(i) for i in h]
        if len(h) >= 3:
            holes.append(h if area(h) <= 0.0 else h[::-1])
    holes.sort(key=lambda h: -q[h, 0].max())
    for k, h in enumerate(holes):
        m = int(np.argmax(q[h, 0]))
        mp = q[h[m]]
        edges = [(ring[i], ring[(i + 1) % len(ring)]) for i in range(len(ring))]
        for o in holes[k:]:
            edges += [(o[i], o[(i + 1) % len(o)]) for i in range(len(o))]
        order = np.argsort(((q[ring] - mp) ** 2).sum(1), kind="stable")
        bridge = None
        for ri in order:
            rp = q[ring[ri]]
            if not any(_proper_cross(rp, mp, q[e0], q[e1]) for e0, e1 in edges):
                bridge = int(ri)
                break
        if bridge is None:
            return None
        ring = ring[:bridge + 1] + h[m:] + h[:m + 1] + ring[bridge:]
    n = len(ring)
    if n < 3:
        return None
    xy = q[ring]
    alive = list(range(n))
    tris = []
    start = 0
    while len(alive) > 3:
        m = len(alive)
        found = -1
        for step in range(m):
            i = (start + step) % m
            ia, ib, ic = alive[i - 1], alive[i], alive[(i + 1) % m]
            a, b, c = xy[ia], xy[ib], xy[ic]
            if _cross2(a, b, c) <= 1e-12:
                continue  # reflex or flat corner
            rest = xy[[j for j in alive if j not in (ia, ib, ic)]]
            if len(rest):
                same = np.zeros(len(rest), bool)
                for v in (a, b, c):
                    same |= (np.abs(rest - v) < 1e-12).all(1)
                d1 = (b[0] - a[0]) * (rest[:, 1] - a[1]) - (b[1] - a[1]) * (rest[:, 0] - a[0])
                d2 = (c[0] - b[0]) * (rest[:, 1] - b[1]) - (c[1] - b[1]) * (rest[:, 0] - b[0])
                d3 = (a[0] - c[0]) * (rest[:, 1] - c[1]) - (a[1] - c[1]) * (rest[:, 0] - c[0])
                inside = (d1 >= 0.0) & (d2 >= 0.0) & (d3 >= 0.0)
                if (inside & ~same).any():
                    continue
            found = i
            break
        if found < 0:
            break
        tris.append((ring[alive[found - 1]], ring[alive[found]], ring[alive[(found + 1) % m]]))
        del alive[found]
        start = max(found - 1, 0)
    if len(alive) >= 3:
        tris += [(ring[alive[0]], ring[alive[i]], ring[alive[i + 1]]) for i in range(1, len(alive) - 1)]
    return np.asarray(tris, np.uint32).reshape(-1, 3)
