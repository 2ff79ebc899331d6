"""Direct illumination of a Lambertian plane in closed form: the scenes and
expected values that pin next-event estimation and its MIS pairing with
emitter hits (render.rs pt/mod.rs:1032-1209 sample_light /
area_light_contribution, 1000-1023 light_pdf_solid_angle, 709-723 the
emitter-hit weight).

Every scene is a matte plane y = 0 of albedo RHO (one value per channel, so
a swapped channel shows) under one or more lights, seen by an orthographic
camera that looks straight down from below the lights with its rays
through pixel centres (filter width 1e-9), so every sample of a pixel
lands on the same plane point x. Emitter geometry (rect quad, disk fan,
sphere) carries emission with `area_light` set and no lobes, so a path
ends where it meets a light. With the firefly clamp off:

* max_depth 2 runs both strategies: the expectation is the plane's
  outgoing radiance rho/pi * E(x), E the irradiance from the lights;
* max_depth 1 keeps only the light-sampling (NEE) arm, weighted by the
  power heuristic: rho/pi * integral of L cos(theta) w_light over the
  light's solid angle, w_light = p_l^2 / (p_l^2 + p_bsdf^2), p_bsdf =
  cos(theta)/pi. Its value is computed here by Gauss-Legendre quadrature.

Irradiance closed forms (textbook view factors, e.g. Howell's catalogue):
parallel rectangle by superposing the corner form factor; parallel disk
at horizontal offset a (on axis: r^2 / (r^2 + h^2)); sphere wholly above
the plane: pi L (R/d)^2 cos(theta); point light I cos(theta) / d^2;
distant light I cos(theta). Each is cross-checked against brute-force
quadrature in the CPU tests, so a wrong formula cannot pin a wrong
estimator.
"""
import ctypes

import numpy as np

import rrhip
import scenes

RHO = np.array([0.8, 0.5, 0.3])
CAM_Y = 0.5          # camera plane height (below every light)
HALF = (1.6, 1.2)    # orthographic half extents along x and z


def _plane_scene(w, h):
    s = scenes.Scene(w, h, 40.0, eye=(0, CAM_Y, 0), look_at=(0, CAM_Y - 1, 0), up=(0, 0, 1))
    s.camera.projection = 1
    s.camera.ortho_half = (ctypes.c_double * 2)(*HALF)
    s.camera.filter_width = 1e-9
    m = s.add_material(scenes.matte(tuple(RHO)))
    big = 40.0
    s.add_polygon([(-big, 0, -big), (big, 0, -big), (big, 0, big), (-big, 0, big)], m)
    return s


def _emitter(s, radiance, li):
    return s.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0, emission=radiance,
                                          area_light=li))


def pixel_points(w, h):
    """Plane points seen by the pixel centres: x = u * HALF[0] along right
    (+x), z = v * HALF[1] along up (+z), u, v the NDC of the centre."""
    i, j = np.meshgrid(np.arange(w), np.arange(h))
    u = (i.ravel() + 0.5) / w * 2 - 1
    v = 1 - (j.ravel() + 0.5) / h * 2
    return np.stack([u * HALF[0], np.zeros(u.size), v * HALF[1]], 1)


# ---- lights ------------------------------------------------------------------

class Rect:
    """Rect light parallel to the plane: x in [x0, x1], z in [z0, z1] at
    height hy, radiance L (rgb)."""

    def __init__(self, x0, x1, z0, z1, hy, L):
        self.x0, self.x1, self.z0, self.z1, self.hy, self.L = x0, x1, z0, z1, hy, np.asarray(L, float)

    def add(self, s):
        c = (self.x0, self.hy, self.z0)
        li = s.add_light(scenes.light(rrhip.LIGHT_RECT, intensity=1.0, color=tuple(self.L), p=c,
                                      e1=(self.x1 - self.x0, 0, 0), e2=(0, 0, self.z1 - self.z0)))
        m = _emitter(s, tuple(self.L), li)
        s.add_polygon([(self.x0, self.hy, self.z0), (self.x1, self.hy, self.z0),
                       (self.x1, self.hy, self.z1), (self.x0, self.hy, self.z1)], m)

    def form_factor(self, pts):
        """Differential area at pts (normal +y) to the rectangle, by
        superposing the corner form factor f(a, b) (odd in a and b)."""
        h = self.hy

        def f(a, b):
            ra = np.sqrt(h * h + a * a)
            rb = np.sqrt(h * h + b * b)
            return (a / ra * np.arctan(b / ra) + b / rb * np.arctan(a / rb)) / (2 * np.pi)
        ax0, ax1 = self.x0 - pts[:, 0], self.x1 - pts[:, 0]
        bz0, bz1 = self.z0 - pts[:, 2], self.z1 - pts[:, 2]
        return f(ax1, bz1) - f(ax0, bz1) - f(ax1, bz0) + f(ax0, bz0)

    def irradiance(self, pts):
        return np.pi * self.form_factor(pts)[:, None] * self.L

    def quad_points(self, n):
        """Gauss-Legendre nodes over the rectangle: (points, area weights)."""
        g, wg = np.polynomial.legendre.leggauss(n)
        xs = 0.5 * (self.x1 - self.x0) * g + 0.5 * (self.x1 + self.x0)
        zs = 0.5 * (self.z1 - self.z0) * g + 0.5 * (self.z1 + self.z0)
        X, Z = np.meshgrid(xs, zs, indexing="ij")
        W = np.outer(wg, wg) * 0.25 * (self.x1 - self.x0) * (self.z1 - self.z0)
        q = np.stack([X.ravel(), np.full(X.size, self.hy), Z.ravel()], 1)
        return q, W.ravel(), np.array([0.0, -1.0, 0.0]), (self.x1 - self.x0) * (self.z1 - self.z0)


class Disk:
    """Disk light parallel to the plane, centre (cx, hy, cz), radius r."""

    def __init__(self, cx, cz, hy, r, L, fan=256):
        self.cx, self.cz, self.hy, self.r, self.L, self.fan = cx, cz, hy, r, np.asarray(L, float), fan

    def add(self, s):
        li = s.add_light(scenes.light(rrhip.LIGHT_DISK, intensity=1.0, color=tuple(self.L),
                                      p=(self.cx, self.hy, self.cz), e1=(self.r, 0, 0), e2=(0, 0, self.r)))
        m = _emitter(s, tuple(self.L), li)
        # Emitter geometry: a fan circumscribing the disk would catch rays
        # outside it, so the fan is inscribed; the sliver it misses
        # (1 - sin(2pi/n)/(2pi/n) of the area, 1e-4 at n = 256) is the only
        # mismatch between geometry and light, and it is below the tests'
        # resolution.
        a = 2 * np.pi * np.arange(self.fan) / self.fan
        ring = [(self.cx + self.r * np.cos(t), self.hy, self.cz + self.r * np.sin(t)) for t in a]
        for k in range(self.fan):
            s.add_triangle((self.cx, self.hy, self.cz), ring[k], ring[(k + 1) % self.fan], m)

    def form_factor(self, pts):
        h, r = self.hy, self.r
        a = np.hypot(pts[:, 0] - self.cx, pts[:, 2] - self.cz)
        on_axis = r * r / (r * r + h * h)
        with np.errstate(divide="ignore", invalid="ignore"):
            H, R = h / a, r / a
            Z = 1 + H * H + R * R
            off = 0.5 * (1 - (Z - 2 * R * R) / np.sqrt(Z * Z - 4 * R * R))
        return np.where(a > 1e-9 * h, off, on_axis)

    def irradiance(self, pts):
        return np.pi * self.form_factor(pts)[:, None] * self.L

    def quad_points(self, n):
        """Gauss-Legendre in r^2 (uniform area) and midpoint in phi."""
        g, wg = np.polynomial.legendre.leggauss(n)
        s = 0.5 * g + 0.5                      # s = (rho / r)^2
        ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
        S, P = np.meshgrid(s, ph, indexing="ij")
        rho = self.r * np.sqrt(S)
        W = np.outer(wg * 0.5, np.full(2 * n, 1.0 / (2 * n))) * np.pi * self.r * self.r
        q = np.stack([self.cx + rho.ravel() * np.cos(P.ravel()), np.full(S.size, self.hy),
                      self.cz + rho.ravel() * np.sin(P.ravel())], 1)
        return q, W.ravel(), np.array([0.0, -1.0, 0.0]), np.pi * self.r * self.r


class SphereLight:
    """Sphere light of radius R centred at c (wholly above the plane)."""

    def __init__(self, c, R, L, geometry=True):
        self.c, self.R, self.L = np.asarray(c, float), R, np.asarray(L, float)
        self.geometry = geometry
        assert self.c[1] > R

    def add(self, s):
        li = s.add_light(scenes.light(rrhip.LIGHT_SPHERE, intensity=1.0, color=tuple(self.L),
                                      p=tuple(self.c), radius=self.R))
        if self.geometry:
            s.add_sphere(tuple(self.c), self.R, _emitter(s, tuple(self.L), li))

    def irradiance(self, pts):
        d = self.c - pts
        dist = np.linalg.norm(d, axis=1)
        cos = d[:, 1] / dist
        return (np.pi * (self.R / dist) ** 2 * cos)[:, None] * self.L


class Point:
    def __init__(self, p, I):
        self.p, self.I = np.asarray(p, float), np.asarray(I, float)

    def add(self, s):
        s.add_light(scenes.light(rrhip.LIGHT_POINT, intensity=1.0, color=tuple(self.I), p=tuple(self.p)))

    def irradiance(self, pts):
        d = self.p - pts
        d2 = (d * d).sum(1)
        return (d[:, 1] / np.sqrt(d2) / d2)[:, None] * self.I


class Distant:
    def __init__(self, travel, I):
        t = np.asarray(travel, float)
        self.t, self.I = t / np.linalg.norm(t), np.asarray(I, float)

    def add(self, s):
        s.add_light(scenes.light(rrhip.LIGHT_DISTANT, intensity=1.0, color=tuple(self.I), p=tuple(self.t)))

    def irradiance(self, pts):
        return np.full((pts.shape[0], 1), max(-self.t[1], 0.0)) * self.I


def scene(lights, w=24, h=18):
    s = _plane_scene(w, h)
    for l in lights:
        l.add(s)
    return s


def expected_total(lights, pts):
    """rho/pi * E(x): both strategies (max_depth 2)."""
    e = sum(l.irradiance(pts) for l in lights)
    return RHO / np.pi * e


def _weight(pl, pb, heuristic):
    if heuristic == "power":
        return pl * pl / (pl * pl + pb * pb)
    if heuristic == "balance":
        return pl / (pl + pb)
    return np.ones_like(pl * pb)  # "none": the whole integral (both arms)


def nee_arm(light, pts, n=48, heuristic="power", pick=1.0):
    """Expected NEE-arm radiance of an area light (rect or disk) alone:
    rho/pi * sum over the light area of L cos cos' / r^2 * w, with the
    light strategy's solid-angle pdf p_l = pick r^2 / (cos' A) and the
    Lambert pdf cos / pi; `heuristic` "power" (render.rs), "balance" (a
    plausible mistake the tests must reject) or "none" (weight 1: the
    irradiance integral itself, to check the closed forms)."""
    q, wq, nl, area = light.quad_points(n)
    out = np.zeros((pts.shape[0], 3))
    for k in range(pts.shape[0]):
        d = q - pts[k]
        r2 = (d * d).sum(1)
        r = np.sqrt(r2)
        cos = d[:, 1] / r
        cosl = np.abs(d @ nl) / r
        pl = pick * r2 / (cosl * area)
        pb = cos / np.pi
        wgt = _weight(pl, pb, heuristic)
        out[k] = (wq * cos * cosl / r2 * wgt).sum() * light.L
    return RHO / np.pi * out


def _frame(axis):
    a = np.array([1.0, 0, 0]) if abs(axis[0]) < 0.9 else np.array([0, 0, 1.0])
    t1 = np.cross(axis, a)
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(axis, t1)


def _cone_integral(c, R, o, integrand, n):
    """Gauss-Legendre in cos and midpoint in phi over the cone of
    directions from o that meet the sphere (c, R): sum of integrand(wi) d
    omega."""
    g, wg = np.polynomial.legendre.leggauss(n)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    d = c - o
    dist = np.linalg.norm(d)
    axis = d / dist
    cmax = np.sqrt(max(1 - (R / dist) ** 2, 0.0))
    t1, t2 = _frame(axis)
    CT, PH = np.meshgrid(0.5 * (1 - cmax) * g + 0.5 * (1 + cmax), ph, indexing="ij")
    ST = np.sqrt(1 - CT * CT)
    wi = (ST * np.cos(PH))[..., None] * t1 + (ST * np.sin(PH))[..., None] * t2 + CT[..., None] * axis
    dw = np.outer(wg * 0.5 * (1 - cmax), np.full(2 * n, 2 * np.pi / (2 * n)))
    return (dw * integrand(wi)).sum()


RAY_OFFSET = 1e-4    # render.rs pt/mod.rs:30
SHADOW_MARGIN = 1e-3  # visibility_to's tmax = distance - 1e-3 (pt/mod.rs:1212-1224)


def _clip_loss(c, R, p, integrand, nq=128, nphi=256, offset=RAY_OFFSET):
    """The share of the cone from p that render.rs's shadow rays lose to
    the light's own sphere: the ray to the cone sample at distance s_t
    leaves from p + offset n (n = +y; render.rs: RAY_OFFSET) and runs
    s_t - SHADOW_MARGIN; it is blocked when it enters the sphere before
    that. A shift of the origin moves the entry by offset (n.N) / |wi.N|
    (N the sphere normal there), so only |wi.N| < 0.1 can block: q = 1 -
    (b/R)^2 < 0.012 for the impact parameter b. Midpoint rule over q and
    phi. Origins on the surface itself (offset 0, as error-bounded origins
    are up to a push of ~1e-7 here) lose nothing."""
    d = c - p
    dist = np.linalg.norm(d)
    axis = d / dist
    t1, t2 = _frame(axis)
    q = (np.arange(nq) + 0.5) / nq * 0.012
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    Q, PH = np.meshgrid(q, ph, indexing="ij")
    b = R * np.sqrt(1 - Q)
    st = b / dist
    ct = np.sqrt(1 - st * st)
    wi = (st * np.cos(PH))[..., None] * t1 + (st * np.sin(PH))[..., None] * t2 + ct[..., None] * axis
    s_t = dist * ct - np.sqrt(np.maximum(R * R - dist * dist * st * st, 0.0))
    o2 = p + np.array([0.0, offset, 0.0])
    oc = c - o2
    B = wi @ oc
    C = oc @ oc - R * R
    disc = B * B - C
    s_e = B - np.sqrt(np.maximum(disc, 0.0))
    blocked = (disc >= 0) & (s_e > 1e-6) & (s_e < s_t - SHADOW_MARGIN)
    dcos_dq = R * R / (2 * dist * dist * ct)
    dw = dcos_dq * (0.012 / nq) * (2 * np.pi / nphi)
    return (dw * blocked * integrand(wi)).sum()


def sphere_nee_arm(light, pts, n=64, heuristic="power", clip=False, offset=RAY_OFFSET):
    """The NEE arm of a sphere light: uniform cone sampling, p_l = 1 / (2 pi
    (1 - cos_max)), integrated over the cone; with `clip`, less what shadow
    rays leaving from p + offset n lose to the light's own sphere
    (_clip_loss; render.rs: offset RAY_OFFSET)."""
    out = np.zeros((pts.shape[0], 3))
    for k in range(pts.shape[0]):
        d = light.c - pts[k]
        dist = np.linalg.norm(d)
        cmax = np.sqrt(max(1 - (light.R / dist) ** 2, 0.0))
        pl = 1.0 / (2 * np.pi * (1 - cmax))

        def integrand(wi):
            cos = np.maximum(wi[..., 1], 0.0)
            return cos * _weight(pl, cos / np.pi, heuristic)
        v = _cone_integral(light.c, light.R, pts[k], integrand, n)
        if clip and offset > 0:
            v -= _clip_loss(light.c, light.R, pts[k], integrand, offset=offset)
        out[k] = v * light.L
    return RHO / np.pi * out


def sphere_bsdf_arm(light, pts, n=64, offset=RAY_OFFSET):
    """The emitter-hit arm of a sphere light: BSDF rays leave from p +
    offset n (render.rs: RAY_OFFSET), so they meet the sphere over the cone
    seen from there, weighted by the power heuristic against the light pdf
    from p."""
    out = np.zeros((pts.shape[0], 3))
    for k in range(pts.shape[0]):
        d = light.c - pts[k]
        dist = np.linalg.norm(d)
        cmax = np.sqrt(max(1 - (light.R / dist) ** 2, 0.0))
        pl = 1.0 / (2 * np.pi * (1 - cmax))

        def integrand(wi):
            cos = np.maximum(wi[..., 1], 0.0)
            pb = cos / np.pi
            return cos * pb * pb / (pb * pb + pl * pl)
        o2 = pts[k] + np.array([0.0, offset, 0.0])
        out[k] = _cone_integral(light.c, light.R, o2, integrand, n) * light.L
    return RHO / np.pi * out


def sphere_total(light, pts, offset=RAY_OFFSET):
    """Both strategies of a sphere light with its emitter sphere, as
    render.rs's estimator computes them with origins pushed by `offset`
    (render.rs: RAY_OFFSET; error-bounded origins: 0)."""
    return sphere_nee_arm(light, pts, clip=True, offset=offset) + sphere_bsdf_arm(light, pts, offset=offset)


# ---- statistics ----------------------------------------------------------------

def z_scores(acc, expected):
    """Per pixel-channel z of the sample mean against the expectation, the
    aggregate z of the summed differences (pixels are independent PCG32
    streams), and the number of pixel-channels whose samples are all equal
    (variance 0) yet differ from the expectation by more than 1e-9
    relative."""
    n = acc[:, 6:7]
    mean = acc[:, :3] / n
    var = np.maximum(acc[:, 3:6] / n - mean * mean, 0.0) * n / np.maximum(n - 1, 1)
    se = np.sqrt(var / n)
    d = mean - expected
    flat = se <= 1e-12 * np.maximum(np.abs(expected), 1e-30)
    flat_bad = int((np.abs(d[flat]) > 1e-9 * np.maximum(np.abs(expected[flat]), 1e-12)).sum())
    z = np.where(flat, 0.0, d / np.where(flat, 1.0, se))
    zagg = d.sum(0) / np.sqrt((se * se).sum(0))
    return z, zagg, mean, flat_bad


def outlier_bound(npix, p=0.0027, alpha=1e-4):
    """Largest count of pixels beyond 3 sigma that a correct render reaches
    with probability >= alpha (binomial upper tail)."""
    from scipy.stats import binom
    return int(binom.isf(alpha, npix, p)) + 1


def check(acc, expected, label, say=print, zmax=4.0):
    """Pixels beyond 3 sigma (in any channel) within the binomial bound and
    every channel's aggregate z below zmax."""
    z, zagg, mean, flat_bad = z_scores(acc, expected)
    out = int((np.abs(z) > 3).any(1).sum())
    bound = outlier_bound(z.shape[0])
    rel = (mean.sum(0) - expected.sum(0)) / expected.sum(0)
    say(f"{label}: pixels outside 3 sigma {out}/{z.shape[0]} (bound {bound}), aggregate z "
        f"{np.round(zagg, 2)}, image-sum rel err {np.round(rel, 5)}")
    assert flat_bad == 0, f"{label}: {flat_bad} zero-variance pixel-channels off the expectation"
    assert out <= bound, f"{label}: {out} pixels outside 3 sigma"
    assert np.all(np.abs(zagg) < zmax), f"{label}: aggregate z {zagg}"
    return zagg


def rejects(acc, wrong, zmin=6.0):
    """True when the render is inconsistent with a wrong expectation."""
    _, zagg, _, flat_bad = z_scores(acc, wrong)
    return flat_bad > 0 or bool(np.any(np.abs(zagg) > zmin))


# ---- the scenes the tests render ---------------------------------------------------

def rect_light():
    return [Rect(-0.6, 0.9, -0.5, 0.7, 1.0, (6.0, 5.0, 4.0))]


def disk_light():
    return [Disk(0.2, -0.1, 0.9, 0.8, (5.0, 6.0, 7.0))]


def sphere_light(geometry=True):
    """A sphere light of 20-25 degrees angular radius, with or without its
    emitter sphere. render.rs's shadow ray to a cone sample starts at p + n
    1e-4 but is aimed from p (pt/mod.rs:1118-1124, 1212-1224), so near the
    silhouette it can enter the light's own sphere before its end: with
    the sphere present the light arm loses ~0.1-0.2 % (sphere_total)."""
    return [SphereLight((0.3, 1.1, 0.2), 0.45, (4.0, 4.5, 5.0), geometry=geometry)]


def point_light():
    return [Point((-0.4, 1.3, 0.3), (3.0, 2.0, 1.0))]


def distant_light():
    return [Distant((0.3, -1.0, 0.2), (0.9, 1.1, 1.3))]


def mixed_lights():
    """Three lights of different kinds: the light-BVH pick pmf (< 1 and
    position dependent) then enters both the NEE divisor and the
    emitter-hit MIS weight."""
    return [Rect(-1.2, -0.2, -0.8, 0.4, 1.2, (4.0, 4.0, 4.0)),
            SphereLight((0.9, 1.0, 0.5), 0.3, (6.0, 5.0, 4.0)),
            Point((0.5, 1.5, -0.7), (2.0, 2.5, 3.0))]


CASES = {
    "rect": rect_light,
    "disk": disk_light,
    "sphere": sphere_light,
    "point": point_light,
    "distant": distant_light,
    "mixed": mixed_lights,
}
