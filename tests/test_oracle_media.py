"""Oracle participating media (medium.rs, pt/volume.rs restated) pinned by
render.rs's own volume KATs (volume.rs:296-364, medium.rs:107-150), the
Henyey-Greenstein closed forms, and the fBm cloud density formula."""
import ctypes

import numpy as np

import orc
import rrhip


def medium(sigma_a, sigma_s, g=0.0, density=None, lo=(-1, -1, -1), hi=(1, 1, 1)):
    m = rrhip.MediumDesc()
    m.sigma_a = (ctypes.c_double * 3)(*[float(x) for x in np.broadcast_to(sigma_a, 3)])
    m.sigma_s = (ctypes.c_double * 3)(*[float(x) for x in np.broadcast_to(sigma_s, 3)])
    m.g = g
    if density is not None:
        dens = np.ascontiguousarray(density, np.float32)
        m._keep = dens
        m.nz, m.ny, m.nx = dens.shape
        m.density = dens.ctypes.data
        m.grid_lo = (ctypes.c_double * 3)(*lo)
        m.grid_hi = (ctypes.c_double * 3)(*hi)
    return m


def test_hg_normalises_and_has_mean_cosine_g():
    """medium.rs:73-101 evaluates hg_phase(g, -wo.wi), i.e. the textbook
    Henyey-Greenstein function of cos(wo, wi) with wo pointing back along
    the incoming ray: its integral is 1, hg_sample returns the phase value
    as pdf, and the sampled wi has mean cosine g with wo (so positive g
    scatters back towards the previous vertex, render.rs's convention,
    which the kernels keep)."""
    mu = np.linspace(-1, 1, 200001)
    for g in (-0.6, 0.0, 0.3, 0.8):
        vals = np.array([orc.hg_phase(g, x) for x in mu[::50]])
        integral = np.trapezoid(vals, mu[::50]) * 2 * np.pi
        assert abs(integral - 1.0) < 2e-4, (g, integral)
        wo = np.array([0.0, 0.0, 1.0])
        wi, pdf = orc.hg_sample(g, wo, 5, 7, 200000)
        cos_wo = wi @ wo
        want = [(1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * c) ** 1.5) for c in cos_wo[:100]]
        assert np.allclose(pdf[:100], want, rtol=1e-9)
        assert abs(cos_wo.mean() - g) < 0.01, (g, cos_wo.mean())
        assert np.allclose(np.linalg.norm(wi, axis=1), 1.0)


def test_homogeneous_distance_sampling_is_unbiased():
    """volume.rs:296-317: sigma_t = 1, pass weight -> e^-1.3, scatter weight
    -> 0.6 (1 - e^-1.3)."""
    m = medium(0.4, 0.6)
    sc, t, w = orc.medium_sample(m, (0, 0, 0), (0, 0, 1), 1.3, (1, 1, 1), 3, 7, 200000)
    n = sc.size
    pass_w = w[~sc].sum(0) / n
    scat_w = w[sc].sum(0) / n
    assert np.all(np.abs(pass_w - np.exp(-1.3)) < 0.01)
    assert np.all(np.abs(scat_w - 0.6 * (1 - np.exp(-1.3))) < 0.01)
    assert np.all(t[sc] < 1.3)


def test_chromatic_homogeneous_sampling_is_unbiased_per_channel():
    """Spectral one-sample MIS (volume.rs:60-102): each channel's expected
    pass weight is its own transmittance."""
    st = np.array([0.5, 1.0, 2.0])
    m = medium(0.3 * st, 0.7 * st)
    sc, t, w = orc.medium_sample(m, (0, 0, 0), (0, 0, 1), 0.9, (1.0, 0.5, 0.25), 11, 3, 300000)
    pass_w = w[~sc].sum(0) / sc.size
    assert np.all(np.abs(pass_w - np.exp(-st * 0.9)) < 0.01), pass_w


def test_delta_tracking_matches_closed_form_in_constant_density():
    """A density grid of ones over [-1,1]^3 is a homogeneous medium inside
    the box: delta tracking from inside to the far wall must pass with
    probability e^{-sigma_t L} and scatter with weight sigma_s/sigma_t."""
    ones = np.ones((8, 8, 8), np.float32)
    m = medium(0.5, 1.5, density=ones)
    L = 1.0 + 0.3  # from z = -0.3 to the box face at z = 1 (t_max beyond it)
    sc, t, w = orc.medium_sample(m, (0.1, 0.2, -0.3), (0, 0, 1), 5.0, (1, 1, 1), 9, 1, 200000)
    assert abs((~sc).mean() - np.exp(-2.0 * L)) < 0.01
    assert np.allclose(w[sc], 0.75)
    assert np.all(w[~sc] == 1.0) and np.all(t[sc] < L)


def test_ratio_tracking_matches_quadrature():
    """volume.rs:324-364 on the voxel fBm cloud: ratio-tracked
    transmittance vs exp(-Riemann optical depth)."""
    dens = rrhip.gen_density_grid(32, 5, frequency=0.8 * 2, octaves=3, coverage=0.6, sharpness=3.0,
                                  lo=(-2, -2, -2), hi=(2, 2, 2))
    m = medium(0.3, 0.7, density=dens, lo=(-2, -2, -2), hi=(2, 2, 2))
    o = np.array([0.3, 0.2, -1.0])
    d = np.array([0.2, 0.1, 1.0])
    d /= np.linalg.norm(d)
    dist = 2.5
    steps = 20000
    ts = (np.arange(steps) + 0.5) / steps * dist
    rho = orc.medium_density(m, o[None, :] + ts[:, None] * d[None, :])
    expected = np.exp(-(rho.sum() * dist / steps))
    tr = orc.medium_transmittance(m, o, d, dist, 11, 4, 30000)
    assert 0.05 < expected < 0.95
    assert abs(tr[:, 0].mean() - expected) < 0.02, (tr[:, 0].mean(), expected)


def test_density_is_trilinear_and_zero_outside():
    rng = np.random.default_rng(1)
    grid = rng.uniform(0, 1, (5, 4, 3)).astype(np.float32)  # nz=5, ny=4, nx=3
    m = medium(0.1, 0.1, density=grid, lo=(0, 0, 0), hi=(2, 3, 4))
    # vertices reproduce the samples
    for (k, j, i) in [(0, 0, 0), (4, 3, 2), (2, 1, 1), (3, 2, 0)]:
        p = (2 * i / 2, 3 * j / 3, 4 * k / 4)
        assert abs(orc.medium_density(m, p)[0] - grid[k, j, i]) < 1e-7
    # a linear field is reproduced exactly between vertices
    zz, yy, xx = np.meshgrid(np.arange(5), np.arange(4), np.arange(3), indexing="ij")
    lin = (0.1 * xx + 0.05 * yy + 0.02 * zz).astype(np.float32)
    ml = medium(0.1, 0.1, density=lin, lo=(0, 0, 0), hi=(2, 3, 4))
    pts = rng.uniform((0, 0, 0), (2, 3, 4), (200, 3))
    want = 0.1 * pts[:, 0] / 1.0 + 0.05 * pts[:, 1] / 1.0 + 0.02 * pts[:, 2] / 1.0
    assert np.abs(orc.medium_density(ml, pts) - want).max() < 1e-6
    assert orc.medium_density(ml, (2.01, 1, 1))[0] == 0.0
    assert orc.medium_density(ml, (-0.01, 1, 1))[0] == 0.0


def test_cloud_density_grid_follows_the_fbm_formula():
    """rrhip_gen_density_grid = medium.rs:55-63 on the grid vertices, with the
    oracle's fBm (displace.rs) and the seeded offset."""
    n, seed = 9, 3
    grid = rrhip.gen_density_grid(n, seed, frequency=2.0, octaves=4, coverage=0.5, sharpness=3.0)
    u = orc.pcg32_seq(seed, 0x5eed, 3).astype(np.float64) / 2 ** 32
    off = 1000.0 * u
    idx = np.array([(0, 0, 0), (8, 8, 8), (3, 5, 1), (7, 2, 4)])
    pts = -1.0 + 2.0 * idx / (n - 1)
    f = orc.fbm(pts, frequency=2.0, octaves=4, gain=0.5, lacunarity=2.0, offset=off)
    want = np.clip((f * 0.5 + 0.5 - 0.5) * 3.0, 0, 1)
    got = grid[idx[:, 2], idx[:, 1], idx[:, 0]]
    assert np.abs(got - want).max() < 1e-6
    assert grid.min() >= 0 and grid.max() <= 1 and 0.05 < (grid > 0).mean() < 0.95


def test_equiangular_sampling_inverts_its_cdf_and_normalises():
    """volume.rs:195-232: t = delta + h tan(theta_a + u span) inverts the
    CDF (atan((t - delta) / h) - theta_a) / span; the returned pdf equals
    equiangular_pdf(t) and integrates to 1 over [0, t_max]; the samples
    follow it (binned counts)."""
    rng = np.random.default_rng(5)
    for _ in range(4):
        o = rng.uniform(-2, 2, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        light = rng.uniform(-3, 3, 3)
        t_max = rng.uniform(0.5, 8.0)
        u = rng.uniform(0, 1, 200000)
        t, pdf, pdf_at = orc.equiangular(o, d, t_max, light, u)
        delta = (light - o) @ d
        h = max(np.linalg.norm(light - (o + d * delta)), 1e-6)
        th_a, th_b = np.arctan(-delta / h), np.arctan((t_max - delta) / h)
        cdf = (np.arctan((t - delta) / h) - th_a) / (th_b - th_a)
        assert np.allclose(cdf, u, atol=1e-9)
        assert np.allclose(pdf, pdf_at, rtol=1e-12)
        grid = np.linspace(0, t_max, 200001)
        dens = h / ((th_b - th_a) * (h * h + (grid - delta) ** 2))
        assert abs(np.trapezoid(dens, grid) - 1.0) < 1e-6
        hist, edges = np.histogram(t, bins=20, range=(0, t_max))
        want = np.array([np.trapezoid(dens[(grid >= a) & (grid <= b)], grid[(grid >= a) & (grid <= b)])
                         for a, b in zip(edges[:-1], edges[1:])]) * t.size
        assert np.all(np.abs(hist - want) < 5 * np.sqrt(want) + 5)


def test_distance_pdf_approx_is_a_density():
    """volume.rs:234-240: mean extinction s, s e^{-st} before t_max plus
    the tail mass e^{-s t_max} sums to 1."""
    m = medium((0.2, 0.4, 0.6), (0.3, 0.1, 0.5))
    s = (0.5 + 0.5 + 1.1) / 3
    t_max = 3.0
    ts = np.linspace(0, t_max, 100001)[:-1]
    vals = np.array([orc.distance_pdf_approx(m, t, t_max) for t in ts[::1000]])
    assert np.allclose(vals, s * np.exp(-s * ts[::1000]), rtol=1e-12)
    tail = orc.distance_pdf_approx(m, t_max, t_max)
    assert abs(tail - np.exp(-s * t_max)) < 1e-12
    assert abs((1 - np.exp(-s * t_max)) + tail - 1.0) < 1e-12


def single_scatter_scene(w=1, h=1):
    """A camera looking down +z through a homogeneous atmosphere at nothing,
    a point light off the axis: with one bounce the pixel holds the single-
    scattered radiance (equiangular + collision NEE)."""
    import scenes
    s = scenes.Scene(w, h, 0.01, eye=(0, 0, 0), look_at=(0, 0, 1))
    s.add_light(scenes.light(rrhip.LIGHT_POINT, intensity=20.0, color=(1.0, 0.8, 0.6), p=(0.8, 0.3, 4.0)))
    s.atmosphere = s.add_medium((0.05, 0.1, 0.15), (0.3, 0.25, 0.2), 0.3)
    return s


def single_scatter_quadrature():
    """L = int_0^inf sigma_s e^{-sigma_t t} hg(g, d.wi) I / r^2
    e^{-sigma_t (r - 1e-3)} dt along the axis (hg_phase(g, -wo.wi) with wo =
    -d, as pt/mod.rs:524-530 evaluates it)."""
    sa, ss, g = np.array([0.05, 0.1, 0.15]), np.array([0.3, 0.25, 0.2]), 0.3
    st = sa + ss
    light, I = np.array([0.8, 0.3, 4.0]), 20.0 * np.array([1.0, 0.8, 0.6])
    t = np.linspace(0, 150, 3_000_001)
    to_l = light[None, :] - np.stack([np.zeros_like(t), np.zeros_like(t), t], 1)
    r = np.linalg.norm(to_l, axis=1)
    cos = to_l[:, 2] / r
    ph = (1 - g * g) / (4 * np.pi * (1 + g * g + 2 * g * cos) ** 1.5)
    f = ss[None, :] * np.exp(-np.outer(t, st)) * (ph / r ** 2)[:, None] * I[None, :] * np.exp(-np.outer(r - 1e-3, st))
    return np.trapezoid(f, t, axis=0)


def test_single_scatter_towards_a_point_light_matches_quadrature():
    s = single_scatter_scene()
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=1, firefly_clamp=0.0), 0, 200000)
    n = acc[0, 6]
    mean = acc[0, :3] / n
    se = np.sqrt(np.maximum(acc[0, 3:6] / n - mean ** 2, 0) / n)
    want = single_scatter_quadrature()
    assert np.all(np.abs(mean - want) < 4 * se + 1e-4 * want), (mean, want, se)


def hull_sphere_scene(w, h, sigma_a, sigma_s, g=0.0):
    """A unit sphere hull (no lobes, medium 0 inside, no atmosphere) seen
    from z = -4 against a unit dome."""
    import scenes
    s = scenes.Scene(w, h, 30.0, eye=(0, 0, -4), look_at=(0, 0, 0))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0))
    med = s.add_medium(sigma_a, sigma_s, g)
    s.add_sphere((0, 0, 0), 1.0, s.add_material(scenes.hull(med)))
    return s


def center_chords(w, h, fov=30.0, eye=(0, 0, -4)):
    """Chord through the unit sphere of each pixel-centre ray, and the
    pixels whose whole footprint lies inside or outside the silhouette."""
    th = np.tan(np.radians(fov) / 2)
    ys, xs = np.mgrid[0:h, 0:w]
    u = ((xs + 0.5) / w * 2 - 1) * th * w / h
    v = (1 - (ys + 0.5) / h * 2) * th
    d = np.stack([u, v, np.ones_like(u)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.array(eye, float)
    b = d @ o
    c = o @ o - 1.0
    disc = b * b - c
    chord = np.where(disc > 0, 2 * np.sqrt(np.maximum(disc, 0)), 0.0)
    # distance of the ray from the centre against a margin of a few pixel
    # footprints at the sphere
    dist_axis = np.sqrt(np.maximum(o @ o - b * b, 0))
    margin = 4.0 * (2 * th / h) * 2.0
    clean = np.abs(dist_axis - 1.0) > margin
    return chord.reshape(-1), clean.reshape(-1)


def test_hull_sphere_absorbs_by_its_chord():
    """Entering the hull pushes its medium and leaving pops it
    (pt/mod.rs:662-683): a purely absorbing interior transmits
    exp(-sigma_a chord) of the dome."""
    W = H = 24
    s = hull_sphere_scene(W, H, (0.8, 0.5, 0.2), (0, 0, 0))
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=8, firefly_clamp=0.0), 0, 1024)
    n = acc[:, 6:7]
    m = acc[:, :3] / n
    se = np.sqrt(np.maximum(acc[:, 3:6] / n - m * m, 0) / n)
    chord, clean = center_chords(W, H)
    want = np.exp(-np.outer(chord, [0.8, 0.5, 0.2]))
    inside = clean & (chord > 0)
    assert inside.sum() > 150 and (clean & (chord == 0)).sum() >= 30
    # outside the silhouette: the dome exactly; inside: exp(-sigma_a chord)
    assert np.array_equal(m[clean & (chord == 0)], np.ones_like(m[clean & (chord == 0)]))
    z = np.abs(m - want)[inside] / np.maximum(se[inside], 1e-12)
    assert z.max() < 5.5 and z.mean() < 1.2, (z.max(), z.mean())


def test_hull_white_furnace():
    """A conservative cloud inside a hull, lit by a unit dome: every pixel
    is 1 (scattering inside, NEE walking out through the hull)."""
    s = hull_sphere_scene(16, 16, (0, 0, 0), (2.0, 2.0, 2.0), 0.3)
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=64, firefly_clamp=0.0), 0, 256)
    n = acc[:, 6]
    mean = acc[:, :3].sum() / (3 * n.sum())
    var = np.maximum(acc[:, 3:6] / acc[:, 6:7] - (acc[:, :3] / acc[:, 6:7]) ** 2, 0)
    se = np.sqrt(var.mean() / n.mean() / acc.shape[0])
    assert abs(mean - 1.0) < 4 * se + 1e-3, (mean, se)


def cutout_scene(w, h, presence):
    """A camera at y = 0.5 looking straight down at a matte 0.5 ground; a
    black quad of the given presence at y = 1 over x < 0 (above the camera,
    never seen) between the ground and a distant light shining straight
    down."""
    import scenes
    s = scenes.Scene(w, h, 60.0, eye=(0, 0.5, 0), look_at=(0, 0, 0), up=(0, 0, 1))
    s.add_light(scenes.light(rrhip.LIGHT_DISTANT, intensity=1.0, p=(0, -1, 0)))
    ground = s.add_material(scenes.matte((0.5, 0.5, 0.5)))
    s.add_polygon([(-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5)], ground)
    cut = s.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0, presence=presence))
    s.add_polygon([(-5, 1, -5), (0, 1, -5), (0, 1, 5), (-5, 1, 5)], cut)
    return s


def test_cutout_attenuates_shadow_rays_and_passes_camera_rays():
    """pt/mod.rs:1292-1303: a blocker of presence p lets 1 - p of the light
    through; pt/mod.rs:702-710: a camera ray passes a cutout with
    probability 1 - p."""
    W = H = 16
    s = cutout_scene(W, H, 0.6)
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=4, firefly_clamp=0.0), 0, 16)
    m = (acc[:, :3] / acc[:, 6:7])[:, 0].reshape(H, W)
    # image x runs along world x: left half (x < 0) is under the cutout
    lit, shaded = m[:, W // 2 + 2:], m[:, :W // 2 - 2]
    assert np.allclose(lit, 0.5 / np.pi, rtol=1e-9)
    assert np.allclose(shaded, 0.5 / np.pi * 0.4, rtol=1e-9)
    # camera rays through a cutout: a black quad of presence 0.3 in front of a dome
    import scenes
    s2 = scenes.Scene(8, 8, 20.0, eye=(0, 0, -2), look_at=(0, 0, 0))
    s2.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0))
    q = s2.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0, presence=0.3))
    s2.add_polygon([(-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0)], q)
    acc2, _ = orc.render(s2.desc(), rrhip.RenderParams.make(max_depth=4, firefly_clamp=0.0), 0, 4096)
    mean = acc2[:, 0].sum() / acc2[:, 6].sum()
    assert abs(mean - 0.7) < 4 * np.sqrt(0.21 / acc2[:, 6].sum()), mean


def fbm_medium(sigma_a, sigma_s, frequency=0.3, octaves=5, gain=0.55, lacunarity=2.0, coverage=0.5,
               sharpness=4.0):
    m = medium(sigma_a, sigma_s)
    m.procedural = rrhip.MEDIUM_FBM
    m.octaves, m.frequency, m.gain, m.lacunarity = octaves, frequency, gain, lacunarity
    m.coverage, m.sharpness = coverage, sharpness
    return m


def test_procedural_cloud_density_follows_medium_rs():
    """DensityField::Fbm (medium.rs:55-63): clamp((fbm(p) * 0.5 + 0.5 - (1 -
    coverage)) * sharpness, 0, 1) at the world point, with the oracle's
    fBm (pinned by displace.rs's tests) and the parameters as f32."""
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    m = fbm_medium(0.2, 0.6, frequency=0.7, octaves=4, gain=0.6, lacunarity=2.1, coverage=0.55, sharpness=3.5)
    pts = np.random.default_rng(2).uniform(-4, 4, (400, 3))
    n = orc.fbm(pts, f32(0.7), 4, f32(0.6), f32(2.1), (0.0, 0.0, 0.0)) * 0.5 + 0.5
    want = np.clip((n - (1.0 - f32(0.55))) * f32(3.5), 0.0, 1.0)
    np.testing.assert_array_equal(orc.medium_density(m, pts), want)
    assert 0.0 < want.mean() < 1.0 and (want == 0.0).any() and (want > 0.0).any()


def test_procedural_cloud_limits_match_closed_forms():
    """Constant limits of the cloud field through delta and ratio tracking:
    sharpness 0 makes the medium vanish (transmittance 1, no scattering);
    coverage 2 with sharpness 100 fills it (|fbm| < 1.3 keeps the ramp
    above 1), so transmittance is exp(-sigma_t d) and the scattered
    fraction sigma_s / sigma_t (1 - T), as for a homogeneous medium."""
    o, d = np.array([0.1, -0.3, 0.2]), np.array([0.0, 0.0, 1.0])
    empty = fbm_medium(0.3, 0.5, sharpness=0.0)
    assert np.all(orc.medium_transmittance(empty, o, d, 3.0, 5, 1, 2000) == 1.0)
    sc, _, _ = orc.medium_sample(empty, o, d, 3.0, (1, 1, 1), 5, 2, 2000)
    assert not sc.any()
    full = fbm_medium(0.3, 0.5, coverage=2.0, sharpness=100.0)
    assert np.all(orc.medium_density(full, np.random.default_rng(3).uniform(-9, 9, (300, 3))) == 1.0)
    tr = orc.medium_transmittance(full, o, d, 2.0, 7, 3, 40000)
    assert abs(tr[:, 0].mean() - np.exp(-0.8 * 2.0)) < 0.01
    sc, t, w = orc.medium_sample(full, o, d, 2.0, (1, 1, 1), 9, 4, 40000)
    assert abs(sc.mean() - (1.0 - np.exp(-1.6))) < 0.01  # channel-equal sigma_t: majorant = sigma_t
    assert np.allclose(w[sc], 0.5 / 0.8)


def test_procedural_cloud_ratio_tracking_matches_quadrature():
    """volume.rs:153-189 ratio tracking through the procedural cloud vs
    exp(-Riemann optical depth) of the field along the segment."""
    m = fbm_medium(0.4, 0.8, frequency=0.9, octaves=4, coverage=0.6, sharpness=3.0)
    o = np.array([0.3, 0.2, -1.0])
    d = np.array([0.2, 0.1, 1.0])
    d /= np.linalg.norm(d)
    dist, steps = 3.0, 20000
    ts = (np.arange(steps) + 0.5) / steps * dist
    rho = orc.medium_density(m, o[None, :] + ts[:, None] * d[None, :])
    expected = np.exp(-1.2 * rho.sum() * dist / steps)
    tr = orc.medium_transmittance(m, o, d, dist, 13, 5, 30000)
    assert 0.05 < expected < 0.95
    assert abs(tr[:, 0].mean() - expected) < 0.02, (tr[:, 0].mean(), expected)
