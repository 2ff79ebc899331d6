"""Scene descriptions of the path-tracing workloads, built as the C-ABI
structs of include/rrhip.h (rrhip_scene_desc and friends).

`Scene` keeps every array it hands to C alive; `desc()` returns the
rrhip_scene_desc. Materials follow render.rs's PbrParams mappings
(src/scene/pbr.rs:103-130): `matte` = from_matte (diffuse only, specular_ior
1), `metal` = from_metal, `glass` = a pure dielectric (glassRefractionGain 1,
diffuseGain 0, specularIor 1 so no extra F0 = 0.04 coat).
"""
import ctypes

import numpy as np

import rrhip


def _d3(v):
    return rrhip.D3(*[float(x) for x in v])


def material(diffuse_gain=1.0, diffuse_color=(0.5, 0.5, 0.5), diffuse_roughness=0.0,
             specular_face_color=(0, 0, 0), specular_edge_color=(1, 1, 1), specular_roughness=0.2,
             specular_ior=1.5, clearcoat_gain=0.0, clearcoat_roughness=0.1, fuzz_gain=0.0,
             fuzz_color=(1, 1, 1), glass_gain=0.0, glass_ior=1.5, glass_roughness=0.05,
             refraction_color=(1, 1, 1), glow=(0, 0, 0), emission=(0, 0, 0), area_light=-1, presence=1.0,
             interior=None, object_id=0):
    """PbrParams defaults of pbr.rs:50-75; presence (pbr.rs:39), the
    interior medium index (material.rs:37, None: no medium) and the object
    id of the id AOV (material.rs:39, 0: none)."""
    m = rrhip.MaterialDesc()
    m.diffuse_gain = diffuse_gain
    m.diffuse_color = _d3(diffuse_color)
    m.diffuse_roughness = diffuse_roughness
    m.specular_face_color = _d3(specular_face_color)
    m.specular_edge_color = _d3(specular_edge_color)
    m.specular_roughness = specular_roughness
    m.specular_ior = specular_ior
    m.clearcoat_gain = clearcoat_gain
    m.clearcoat_roughness = clearcoat_roughness
    m.fuzz_gain = fuzz_gain
    m.fuzz_color = _d3(fuzz_color)
    m.glass_gain = glass_gain
    m.glass_ior = glass_ior
    m.glass_roughness = glass_roughness
    m.refraction_color = _d3(refraction_color)
    m.glow = _d3(glow)
    m.emission = _d3(emission)
    m.area_light = area_light
    m.cutout = 1.0 - presence
    m.interior = 0 if interior is None else interior + 1
    m.id = object_id
    return m


def hull(interior):
    """An invisible volume boundary (no lobes, no emission) around the
    medium `interior`."""
    return material(diffuse_gain=0.0, specular_ior=1.0, interior=interior)


def matte(color, **kw):
    """pbr.rs:103-110 (PbrParams::from_matte)."""
    return material(diffuse_gain=1.0, diffuse_color=color, specular_ior=1.0, **kw)


def metal(color, roughness):
    """pbr.rs:122-130 (PbrParams::from_metal)."""
    return material(diffuse_gain=0.0, specular_face_color=color, specular_edge_color=(1, 1, 1),
                    specular_roughness=max(roughness, 0.02))


def glass(ior=1.5, roughness=0.05):
    return material(diffuse_gain=0.0, specular_ior=1.0, glass_gain=1.0, glass_ior=ior,
                    glass_roughness=roughness)


def sigma_a_from_color(color, beta_n):
    """Fiber absorption that gives a hair volume the colour `color` under
    azimuthal roughness beta_n (hair.rs:64-77, after Chiang et al. 2016,
    "A practical and controllable hair and fur model", eq. 9)."""
    b = beta_n
    denom = 5.969 - 0.215 * b + 2.532 * b ** 2 - 10.73 * b ** 3 + 5.574 * b ** 4 + 0.245 * b ** 5
    return tuple((np.log(np.clip(c, 1e-4, 0.999)) / denom) ** 2 for c in color)


def hair(color=(0.35, 0.2, 0.06), beta_m=0.3, beta_n=0.3, eta=1.55, absorption=None, **kw):
    """A hair fiber (Bxdf "PxrMarschnerHair", HairParams::from_bxdf_params,
    hair.rs:79-98): absorption given, or from `color`; curve hits scatter
    with the fiber BSDF, other hits see the brown matte render.rs keeps as
    the material's surface (scene_builder.rs:1851-1857)."""
    m = matte((0.35, 0.22, 0.12), **kw)
    m.hair = 1
    bn = min(max(beta_n, 0.01), 1.0)
    if absorption is None:
        sa = sigma_a_from_color(color, bn)
    else:
        a = list(absorption)
        sa = tuple(a[:3]) if len(a) >= 3 else (a[0],) * 3
    m.hair_sigma_a = _d3(sa)
    m.hair_beta_m, m.hair_beta_n, m.hair_eta = beta_m, beta_n, eta
    return m


def dice_curve(ctrl, cubic, width_root, width_tip, segs_per_span):
    """Dices one curve into rounded-cone segments (curves.rs:191-258):
    the control points of a linear curve (cubic None), or for a cubic one
    (cubic = (basis 4x4, step)) segs_per_span points per span plus the last
    span's end; widths lerp root -> tip by sample index and become radii
    (width / 2); v = index / (samples - 1). Returns (p0, p1) as n x 4 f32
    (xyz, radius) and (v0, v1) as f32."""
    ctrl = np.asarray(ctrl, np.float64).reshape(-1, 3)
    if cubic is None:
        pts = ctrl
    else:
        basis, step = np.asarray(cubic[0], np.float64).reshape(4, 4), int(cubic[1])
        n = ctrl.shape[0]
        if n < 4:
            return (np.zeros((0, 4), np.float32),) * 2 + (np.zeros(0, np.float32),) * 2
        spans = (n - 4) // step + 1
        rows = []
        for span in range(spans):
            g = ctrl[span * step:span * step + 4]
            count = segs_per_span + 1 if span == spans - 1 else segs_per_span
            u = np.arange(count, dtype=np.float64) / segs_per_span
            uv = np.stack([u * u * u, u * u, u, np.ones_like(u)], 1)
            w = uv[:, 0:1] * basis[0] + uv[:, 1:2] * basis[1] + uv[:, 2:3] * basis[2] + uv[:, 3:4] * basis[3]
            rows.append(w[:, 0:1] * g[0] + w[:, 1:2] * g[1] + w[:, 2:3] * g[2] + w[:, 3:4] * g[3])
        pts = np.concatenate(rows)
    if pts.shape[0] < 2:
        return (np.zeros((0, 4), np.float32),) * 2 + (np.zeros(0, np.float32),) * 2
    k = np.arange(pts.shape[0], dtype=np.float64) / (pts.shape[0] - 1)
    r = (width_root + (width_tip - width_root) * k) * 0.5
    ends = np.hstack([pts, r[:, None]]).astype(np.float32)
    v = k.astype(np.float32)
    return ends[:-1], ends[1:], v[:-1], v[1:]


def light(kind, intensity=1.0, color=(1, 1, 1), p=(0, 0, 0), e1=(0, 0, 0), e2=(0, 0, 0), radius=0.0):
    l = rrhip.LightDesc()
    l.kind = kind
    l.intensity = intensity
    l.color = _d3(color)
    l.p = _d3(p)
    l.e1 = _d3(e1)
    l.e2 = _d3(e2)
    l.radius = radius
    return l


class Scene:
    def __init__(self, width, height, fov, eye, look_at, up=(0, 1, 0), filter=0, filter_width=1.0,
                 background=(0, 0, 0)):
        self.camera = rrhip.CameraDesc()
        self.camera.width = width
        self.camera.height = height
        self.camera.filter = filter
        self.camera.fov_deg = fov
        self.camera.filter_width = filter_width
        self.camera.eye = _d3(eye)
        self.camera.look_at = _d3(look_at)
        self.camera.up = _d3(up)
        self.background = background
        self.materials, self.lights, self.triangles, self.spheres = [], [], [], []
        self.meshes, self.instances = [], []
        self.curve_sets = []
        self.textures = []
        self.patterns = []
        self.bindings = []
        self.media = []
        self.atmosphere = -1
        self._keep = []

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_light(self, l):
        self.lights.append(l)
        return len(self.lights) - 1

    def add_triangle(self, v0, v1, v2, mat):
        t = rrhip.TriangleDesc()
        t.v0, t.v1, t.v2 = _d3(v0), _d3(v1), _d3(v2)
        t.material = mat
        self.triangles.append(t)

    def add_polygon(self, verts, mat):
        """Fan triangulation as render.rs's Polygon (scene_builder.rs:939-947)."""
        for i in range(1, len(verts) - 1):
            self.add_triangle(verts[0], verts[i], verts[i + 1], mat)

    def add_sphere(self, center, radius, mat):
        s = rrhip.SphereDesc()
        s.center = _d3(center)
        s.radius = radius
        s.material = mat
        self.spheres.append(s)

    def add_mesh(self, positions, indices, normals=None, st=None, positions1=None):
        """A mesh for instances; st: optional per-vertex texture
        coordinates (n x 2); positions1: optional vertices at shutter close
        (deformation motion, same count as positions)."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        stv = None if st is None else np.ascontiguousarray(st, np.float32).reshape(-1, 2)
        p1 = None if positions1 is None else np.ascontiguousarray(positions1, np.float32).reshape(-1, 3)
        assert p1 is None or p1.shape == pos.shape
        self._keep += [pos, idx, nrm, stv, p1]
        self.meshes.append(rrhip.MeshDesc(pos.ctypes.data, pos.shape[0], idx.ctypes.data, idx.shape[0],
                                          None if nrm is None else nrm.ctypes.data,
                                          None if stv is None else stv.ctypes.data,
                                          None if p1 is None else p1.ctypes.data))
        return len(self.meshes) - 1

    def add_texture(self, rgb):
        """A texture from level 0 (height x width x 3 linear rgb, rows from
        the top); returns its index."""
        t = np.ascontiguousarray(rgb, np.float32)
        assert t.ndim == 3 and t.shape[2] == 3
        self._keep.append(t)
        self.textures.append(rrhip.TextureDesc(t.shape[1], t.shape[0], t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return len(self.textures) - 1

    def add_pattern(self, kind, udim=None, stops=None, **kw):
        """A pattern node (rrhip.PatternDesc.make's keywords); udim: a dict
        tile -> texture index; stops: ramp [(position, (r, g, b)), ...].
        Returns its index."""
        p = rrhip.PatternDesc.make(kind, **kw)
        if udim is not None:
            u = np.full(100, -1, np.int32)
            for tile, tex in udim.items():
                u[tile - 1001] = tex
            self._keep.append(u)
            p.udim = u.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        if stops:
            pos = np.ascontiguousarray([a for a, _ in stops], np.float64)
            col = np.ascontiguousarray([c for _, c in stops], np.float64).reshape(-1, 3)
            self._keep += [pos, col]
            p.num_stops = len(stops)
            p.positions = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            p.colors = col.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self.patterns.append(p)
        return len(self.patterns) - 1

    def bind(self, mat, field, pattern):
        """Drive material `mat`'s field (a name of rrhip.BIND_FIELDS or its
        index) by pattern node `pattern`."""
        f = rrhip.BIND_FIELDS.index(field) if isinstance(field, str) else int(field)
        self.bindings.append(rrhip.BindingDesc(mat, f, pattern, 0))

    def add_curves(self, p0, p1, v0=None, v1=None):
        """A curve set (rrhip_curves_desc): per segment the start and end
        as xyz + radius (n x 4 each) and the curve parameter at both ends
        (default 0 and 1). Returns its index for add_instance(kind=curves)."""
        a = np.ascontiguousarray(p0, np.float32).reshape(-1, 4)
        b = np.ascontiguousarray(p1, np.float32).reshape(-1, 4)
        n = a.shape[0]
        va = np.zeros(n, np.float32) if v0 is None else np.ascontiguousarray(v0, np.float32).reshape(n)
        vb = np.ones(n, np.float32) if v1 is None else np.ascontiguousarray(v1, np.float32).reshape(n)
        self._keep += [a, b, va, vb]
        self.curve_sets.append(rrhip.CurvesDesc(a.ctypes.data, b.ctypes.data, va.ctypes.data, vb.ctypes.data, n))
        return len(self.curve_sets) - 1

    def add_instance(self, mesh, mat, object_to_world_3x4, object_to_world1_3x4=None, curves=False):
        """An instance of mesh `mesh` (of curve set `mesh` with curves=True);
        with object_to_world1_3x4 it moves from the first transform at
        shutter open to the second at close."""
        i = rrhip.InstanceDesc()
        i.mesh = mesh
        i.material = mat
        i.kind = rrhip.INSTANCE_CURVES if curves else rrhip.INSTANCE_MESH
        i.object_to_world = (ctypes.c_double * 12)(*[float(x) for x in np.asarray(object_to_world_3x4).ravel()])
        if object_to_world1_3x4 is not None:
            i.object_to_world1 = (ctypes.c_double * 12)(*[float(x) for x in
                                                          np.asarray(object_to_world1_3x4).ravel()])
            i.moving = 1
        self.instances.append(i)

    def add_medium(self, sigma_a, sigma_s, g=0.0, emission=(0, 0, 0), max_distance=0.0, density=None,
                   grid_lo=(-1, -1, -1), grid_hi=(1, 1, 1), fbm=None):
        """A medium (medium.rs Medium); density: None (homogeneous) or a
        (nz, ny, nx) array of [0, 1] values on the vertices of the grid box;
        fbm: None or a dict of render.rs's cloud field parameters
        (frequency, octaves, gain, lacunarity, coverage, sharpness; the RIB
        defaults where missing), evaluated per lookup."""
        m = rrhip.MediumDesc()
        m.sigma_a, m.sigma_s, m.emission = _d3(sigma_a), _d3(sigma_s), _d3(emission)
        m.g = g
        m.max_distance = max_distance
        if fbm is not None:
            f = dict(frequency=0.3, octaves=5, gain=0.55, lacunarity=2.0, coverage=0.5, sharpness=4.0)
            f.update(fbm)
            m.procedural = rrhip.MEDIUM_FBM
            m.octaves = int(f["octaves"])
            m.frequency, m.gain, m.lacunarity = f["frequency"], f["gain"], f["lacunarity"]
            m.coverage, m.sharpness = f["coverage"], f["sharpness"]
        if density is not None:
            dens = np.ascontiguousarray(density, np.float32)
            self._keep.append(dens)
            m.nz, m.ny, m.nx = dens.shape
            m.density = dens.ctypes.data
            m.grid_lo, m.grid_hi = _d3(grid_lo), _d3(grid_hi)
        self.media.append(m)
        return len(self.media) - 1

    def _arr(self, ctype, items):
        if not items:
            return None
        a = (ctype * len(items))(*items)
        self._keep.append(a)
        return ctypes.cast(a, ctypes.POINTER(ctype))

    def desc(self, bins=0, max_leaf=0):
        d = rrhip.SceneDesc()
        d.struct_size = ctypes.sizeof(rrhip.SceneDesc)
        d.num_meshes = len(self.meshes)
        d.meshes = self._arr(rrhip.MeshDesc, self.meshes)
        d.build = rrhip.BuildParams(bins, max_leaf, 0, 0)
        d.num_instances = len(self.instances)
        d.instances = self._arr(rrhip.InstanceDesc, self.instances)
        d.num_triangles = len(self.triangles)
        d.triangles = self._arr(rrhip.TriangleDesc, self.triangles)
        d.num_spheres = len(self.spheres)
        d.spheres = self._arr(rrhip.SphereDesc, self.spheres)
        d.num_materials = len(self.materials)
        d.materials = self._arr(rrhip.MaterialDesc, self.materials)
        d.num_lights = len(self.lights)
        d.lights = self._arr(rrhip.LightDesc, self.lights)
        d.camera = self.camera
        d.background = _d3(self.background)
        d.num_media = len(self.media)
        d.atmosphere = self.atmosphere
        d.media = self._arr(rrhip.MediumDesc, self.media)
        qs = getattr(self, "quadrics", [])
        d.num_quadrics = len(qs)
        d.quadrics = self._arr(rrhip.QuadricDesc, qs)
        d.num_curve_sets = len(self.curve_sets)
        d.curve_sets = self._arr(rrhip.CurvesDesc, self.curve_sets)
        d.num_textures = len(self.textures)
        d.textures = self._arr(rrhip.TextureDesc, self.textures)
        d.num_patterns = len(self.patterns)
        d.patterns = self._arr(rrhip.PatternDesc, self.patterns)
        d.num_bindings = len(self.bindings)
        d.bindings = self._arr(rrhip.BindingDesc, self.bindings)
        env = getattr(self, "env", None)
        if env is not None:
            env = np.ascontiguousarray(env, np.float32)
            self._keep.append(env)
            d.env_height, d.env_width = env.shape[0], env.shape[1]
            d.env_rgb = env.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self._keep.append(d)
        return d

    def add_quadric(self, kind, mat, object_to_world, radius=1.0, zmin=-1.0, zmax=1.0, height=1.0,
                    thetamax=360.0, p1=(0, 0, 0), p2=(0, 0, 0)):
        """A RenderMan quadric (rrhip_quadric_desc): object_to_world a 3x4
        or 4x4 row-major matrix."""
        q = rrhip.QuadricDesc()
        q.kind, q.material = kind, mat
        q.radius, q.zmin, q.zmax, q.height, q.thetamax = radius, zmin, zmax, height, thetamax
        q.p1 = (ctypes.c_double * 3)(*p1)
        q.p2 = (ctypes.c_double * 3)(*p2)
        m = np.asarray(object_to_world, np.float64)[:3].reshape(-1)
        for k in range(12):
            q.object_to_world[k] = m[k]
        if not hasattr(self, "quadrics"):
            self.quadrics = []
        self.quadrics.append(q)

    def set_env(self, rgb):
        """Lat-long environment map of the dome light: (height, width, 3)
        linear radiance, rows from the +y pole (envmap.rs)."""
        self.env = np.ascontiguousarray(rgb, np.float32)

    @property
    def npix(self):
        return self.camera.width * self.camera.height


def box_room(width=256, height=256, emission=15.0, offset=(0.0, 0.0, 0.0), instanced=False, light_y=0.99):
    """The box-room workload: a closed room spanning [-1,1]^3 of 10
    axis-aligned quads (floor, ceiling, back and front walls each split in
    two halves at x = 0, left and right walls whole: 20 triangles), walls
    albedo 0.75 except left (0.75, 0.1, 0.1) and right (0.1, 0.75, 0.1),
    spheres r = 0.3 at (-0.4, -0.7, 0.2) and r = 0.35 at (0.4, -0.65, -0.3),
    all Lambert (Oren-Nayar sigma = 0); a 0.5 x 0.5 rect light at y = 0.99
    with emission 15, whose quad is also a white matte emitter surface (as
    render.rs's AreaLightSource + Polygon, scene_builder.rs:924-936). Camera
    inside the room at z = -0.9 looking at +z, 60 degree vertical fov, box
    filter of width 1.

    offset moves the whole scene (geometry, light, camera); instanced puts
    the walls and the light quad in one mesh placed by an instance whose
    transform carries the offset (the spheres stay loose); light_y moves
    the light (and its quad) below the ceiling."""
    o = np.asarray(offset, np.float64)

    def at(p):
        return tuple(np.asarray(p, np.float64) + o)

    s = Scene(width, height, 60.0, eye=at((0, 0, -0.9)), look_at=at((0, 0, 1)), up=(0, 1, 0))
    white = s.add_material(matte((0.75, 0.75, 0.75)))
    red = s.add_material(matte((0.75, 0.1, 0.1)))
    green = s.add_material(matte((0.1, 0.75, 0.1)))
    li = s.add_light(light(rrhip.LIGHT_RECT, intensity=emission, color=(1, 1, 1),
                           p=at((-0.25, light_y, -0.25)), e1=(0.5, 0, 0), e2=(0, 0, 0.5)))
    emitter = s.add_material(matte((1, 1, 1), emission=(emission,) * 3, area_light=li))
    quads = []
    for x0, x1 in ((-1, 0), (0, 1)):
        quads.append(([(x0, -1, -1), (x1, -1, -1), (x1, -1, 1), (x0, -1, 1)], white))   # floor
        quads.append(([(x0, 1, -1), (x0, 1, 1), (x1, 1, 1), (x1, 1, -1)], white))       # ceiling
        quads.append(([(x0, -1, 1), (x1, -1, 1), (x1, 1, 1), (x0, 1, 1)], white))       # back
        quads.append(([(x0, -1, -1), (x0, 1, -1), (x1, 1, -1), (x1, -1, -1)], white))   # front
    quads.append(([(-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)], red))    # left
    quads.append(([(1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)], green))      # right
    # The light quad (corner, corner+e1, corner+e1+e2, corner+e2).
    y = light_y
    quads.append(([(-0.25, y, -0.25), (0.25, y, -0.25), (0.25, y, 0.25), (-0.25, y, 0.25)], emitter))
    if instanced:
        # one mesh per material (an instance carries one material)
        xf = np.hstack([np.eye(3), o[:, None]])
        for mat in (white, red, green, emitter):
            pos, idx = [], []
            for verts, m in quads:
                if m != mat:
                    continue
                b = len(pos)
                pos += verts
                idx += [(b, b + i, b + i + 1) for i in range(1, len(verts) - 1)]
            s.add_instance(s.add_mesh(np.asarray(pos, np.float32), np.asarray(idx, np.uint32)), mat, xf)
    else:
        for verts, m in quads:
            s.add_polygon([at(v) for v in verts], m)
    s.add_sphere(at((-0.4, -0.7, 0.2)), 0.3, white)
    s.add_sphere(at((0.4, -0.65, -0.3)), 0.35, white)
    return s


def plastic(color, roughness=0.1):
    """RenderMan "plastic" as PbrParams::from_plastic (pbr.rs:112-120)."""
    return material(diffuse_gain=0.9, diffuse_color=color, specular_ior=1.5,
                    specular_roughness=max(roughness, 0.02))


def sphere_zoo(width=96, height=96):
    """A background- and direct-light-dominated scene for adaptive
    sampling, laid out as render.rs's quadric_zoo fixture
    (tests/fixtures/quadric_zoo.rib) with every quadric replaced by a
    sphere of similar size at its position and with its colour and surface:
    two rows of four over a grey ground quad at y = -2.2, lit by a distant
    key light (from (4, 8, -6) towards (0, 0, 7), intensity 0.85) and a
    point light at (-5, 4, 2) (intensity 0.35); black background, camera
    at the origin looking down +z with a 40 degree fov."""
    s = Scene(width, height, 40.0, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0))
    rows = [((-3.4, 1.2, 7), 0.9, plastic((0.9, 0.25, 0.2), 0.15)),
            ((-1.1, 1.2, 7), 0.75, metal((0.95, 0.7, 0.15), 0.08)),
            ((1.1, 1.2, 7), 0.7, plastic((0.2, 0.55, 0.9))),
            ((3.4, 0.8, 7), 0.8, plastic((0.25, 0.8, 0.35))),
            ((-3.4, -1.0, 7), 0.7, plastic((0.7, 0.3, 0.85))),
            ((-1.1, -1.1, 7), 0.6, metal((0.9, 0.45, 0.6), 0.15)),
            ((1.1, -1.1, 7), 0.75, metal((0.85, 0.85, 0.9), 0.05)),
            ((3.4, -1.1, 7), 0.8, plastic((0.3, 0.85, 0.8)))]
    for c, r, m in rows:
        s.add_sphere(c, r, s.add_material(m))
    ground = s.add_material(matte((0.75, 0.75, 0.75)))
    s.add_polygon([(-12, -2.2, 0.5), (12, -2.2, 0.5), (12, -2.2, 18), (-12, -2.2, 18)], ground)
    d = np.array([0.0, 0.0, 7.0]) - np.array([4.0, 8.0, -6.0])
    s.add_light(light(rrhip.LIGHT_DISTANT, intensity=0.85, p=tuple(d / np.linalg.norm(d))))
    s.add_light(light(rrhip.LIGHT_POINT, intensity=0.35, p=(-5, 4, 2)))
    return s


def cloud_room(width=128, height=128, n=128, seed=1, sigma_s=(2.0, 2.0, 2.0), sigma_a=(0.2, 0.2, 0.2), g=0.3,
               emission=15.0, spheres=None):
    """The box room filled with a heterogeneous atmosphere: an n^3 fBm cloud
    density grid over the room's [-1, 1]^3 (medium.rs's DensityField::Fbm:
    frequency 2, 4 octaves, coverage 0.5, sharpness 3, offset from `seed`),
    sigma_s 2, sigma_a 0.2, Henyey-Greenstein g = 0.3. `spheres` replaces
    the two room spheres with (centre, radius) pairs when given."""
    s = box_room(width, height, emission)
    if spheres is not None:
        s.spheres = []
        white = 0
        for c, r in spheres:
            s.add_sphere(c, r, white)
    dens = rrhip.gen_density_grid(n, seed, frequency=2.0, octaves=4, coverage=0.5, sharpness=3.0)
    s.atmosphere = s.add_medium(sigma_a, sigma_s, g, density=dens)
    return s


def cornell_fixture(width=96, height=96):
    """render.rs's tests/fixtures/cornell.rib as a scene description:
    AreaLightSource intensity 17, lightcolor (1, 0.9, 0.75) on the ceiling
    quad (white matte emitter surface), white (0.73) floor / ceiling / back,
    red (0.63, 0.065, 0.05) left and green (0.14, 0.45, 0.091) right walls,
    a white matte sphere and a "metal" roughness-0.08 sphere (0.95, 0.93,
    0.88) of radius 0.35; RIB default camera at the origin looking +z with
    fov 36 and a box filter of width 1."""
    s = Scene(width, height, 36.0, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0))
    lc = (1.0, 0.9, 0.75)
    quad = [(-0.25, 0.995, 3.95), (0.25, 0.995, 3.95), (0.25, 0.995, 4.45), (-0.25, 0.995, 4.45)]
    c = np.array(quad[0])
    li = s.add_light(light(rrhip.LIGHT_RECT, intensity=17.0, color=lc, p=c,
                           e1=np.array(quad[1]) - c, e2=np.array(quad[3]) - c))
    emitter = s.add_material(matte((1, 1, 1), emission=tuple(17.0 * x for x in lc), area_light=li))
    s.add_polygon(quad, emitter)
    white = s.add_material(matte((0.73, 0.73, 0.73)))
    s.add_polygon([(-1, -1, 3), (1, -1, 3), (1, -1, 5), (-1, -1, 5)], white)
    s.add_polygon([(-1, 1, 3), (-1, 1, 5), (1, 1, 5), (1, 1, 3)], white)
    s.add_polygon([(-1, -1, 5), (1, -1, 5), (1, 1, 5), (-1, 1, 5)], white)
    red = s.add_material(matte((0.63, 0.065, 0.05)))
    s.add_polygon([(-1, -1, 3), (-1, -1, 5), (-1, 1, 5), (-1, 1, 3)], red)
    green = s.add_material(matte((0.14, 0.45, 0.091)))
    s.add_polygon([(1, -1, 3), (1, 1, 3), (1, 1, 5), (1, -1, 5)], green)
    s.add_sphere((-0.42, -0.65, 4.4), 0.35, s.add_material(matte((0.73, 0.73, 0.73))))
    s.add_sphere((0.4, -0.65, 3.9), 0.35, s.add_material(metal((0.95, 0.93, 0.88), 0.08)))
    return s


def furnace(width=64, height=64, albedo=1.0, mat=None, radius=1.5):
    """render.rs's white furnace (tests/pbr_tests.rs:31-54): a sphere of
    radius 1.5 at distance 4 inside a constant dome of radiance 1, 30 degree
    fov. The default material is an albedo-1 diffuse (PxrSurface
    diffuseGain 1, diffuseColor 1, default specularIor 1.5 -> F0 0.04 spec)
    unless `mat` is given."""
    s = Scene(width, height, 30.0, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0))
    if mat is None:
        mat = material(diffuse_gain=1.0, diffuse_color=(albedo,) * 3, specular_ior=1.0)
    m = s.add_material(mat)
    s.add_light(light(rrhip.LIGHT_DOME, intensity=1.0))
    s.add_sphere((0, 0, 4), radius, m)
    return s


def random_rotation(u1, u2, u3):
    """Uniform random rotation from three uniforms (Shoemake's quaternion)."""
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    t1, t2 = 2 * np.pi * u2, 2 * np.pi * u3
    x, y, z, w = a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_field(width=1920, height=1080, n_instances=10000, levels=5, n_point_lights=64, seed=1,
                    box=(200.0, 20.0, 200.0)):
    """The instanced path-tracing workload: `n_instances` TLAS instances of
    one displaced icosphere (levels = 5: 20,480 triangles; amplitude 0.15,
    frequency 3, 4 octaves), each with a uniformly random rotation, uniform
    scale in [0.5, 2] and position uniform in a 200 x 20 x 200 box (y up);
    materials chosen 50/50 between GGX metal (pbr.rs from_metal, colour
    uniform in [0.5, 1]^3) and a rough dielectric of IOR 1.5, roughness
    U(0.05, 1); `n_point_lights` point lights with intensity log-uniform in
    [1, 100] placed 25-40 above the field, plus one 60 x 60 rect area light
    at y = 60 facing down. A matte 0.5 ground quad at y = -2 (a recorded
    addition: the configuration names no floor). Camera at (0, 45, -140)
    looking at the origin, 45 degree vertical fov."""
    rng = np.random.default_rng(seed)
    s = Scene(width, height, 45.0, eye=(0, 45, -140), look_at=(0, 0, 0))
    pos, idx, nrm = rrhip.gen_icosphere(levels, seed)
    mesh = s.add_mesh(pos, idx, nrm)
    ground = s.add_material(matte((0.5, 0.5, 0.5)))
    half = 0.75 * max(box[0], box[2])
    s.add_polygon([(-half, -2, -half), (half, -2, -half), (half, -2, half), (-half, -2, half)], ground)
    for _ in range(n_point_lights):
        p = (rng.uniform(-box[0] / 2, box[0] / 2), rng.uniform(25, 40), rng.uniform(-box[2] / 2, box[2] / 2))
        s.add_light(light(rrhip.LIGHT_POINT, intensity=float(np.exp(rng.uniform(0, np.log(100)))),
                          color=(1, 1, 1), p=p))
    corner = np.array([-30.0, 60.0, -30.0])
    li = s.add_light(light(rrhip.LIGHT_RECT, intensity=3.0, color=(1, 0.95, 0.9), p=corner,
                           e1=(0, 0, 60), e2=(60, 0, 0)))
    emitter = s.add_material(matte((1, 1, 1), emission=(3.0, 2.85, 2.7), area_light=li))
    s.add_polygon([corner, corner + (0, 0, 60), corner + (60, 0, 60), corner + (60, 0, 0)], emitter)
    for _ in range(n_instances):
        R = random_rotation(*rng.uniform(0, 1, 3)) * rng.uniform(0.5, 2.0)
        t = (rng.uniform(-box[0] / 2, box[0] / 2), rng.uniform(0, box[1]), rng.uniform(-box[2] / 2, box[2] / 2))
        rough = rng.uniform(0.05, 1.0)
        if rng.uniform() < 0.5:
            m = s.add_material(metal(tuple(rng.uniform(0.5, 1.0, 3)), rough))
        else:
            m = s.add_material(glass(1.5, rough))
        xf = np.concatenate([R, np.array(t).reshape(3, 1)], axis=1)
        s.add_instance(mesh, m, xf)
    return s


def batch_rooms(n_jobs=32, seed=1, grid=128):
    """The batched-jobs workload: `n_jobs` box rooms, each with 1-40 white
    matte spheres of radius U(0.05, 0.3) centred uniformly in [-0.7, 0.7]^3,
    a square resolution drawn from {64, 128, 256, 512}, spp uniform in
    4..64, and every other job filled with the fBm cloud atmosphere
    (grid^3 density, g 0.3). Returns [(Scene, spp, max_depth)]: depth 8 for
    surface-only rooms, 64 for rooms with media (render.rs's budget,
    pt/mod.rs:484)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        res = int(rng.choice([64, 128, 256, 512]))
        k = int(rng.integers(1, 41))
        spheres = [(tuple(rng.uniform(-0.7, 0.7, 3)), float(rng.uniform(0.05, 0.3))) for _ in range(k)]
        spp = int(rng.integers(4, 65))
        media = j % 2 == 1
        if media:
            s = cloud_room(res, res, n=grid, seed=seed * 1000 + j, spheres=spheres)
        else:
            s = box_room(res, res)
            s.spheres = []
            for c, r in spheres:
                s.add_sphere(c, r, 0)
        jobs.append((s, spp, 64 if media else 8))
    return jobs


def _place(t, rx_deg=0.0, scale=1.0):
    """Row-major 3x4 object-to-world: scale, then a rotation about x, then a
    translation (RIB's Translate; Rotate; order)."""
    a = np.radians(rx_deg)
    r = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]]) * scale
    return np.hstack([r, np.asarray(t, float).reshape(3, 1)])


def quadric_zoo(width=96, height=54):
    """render.rs's quadric_zoo fixture (tests/fixtures/quadric_zoo.rib):
    sphere, torus, cylinder and cone on the top row, paraboloid,
    hyperboloid, disk and a 270-degree partial sphere below, with the
    fixture's sizes, placements (Translate then Rotate about x), colours,
    "plastic" / "metal" surfaces, distant key light, point fill light and
    ground quad."""
    s = Scene(width, height, 40.0, eye=(0, 0, 0), look_at=(0, 0, 1))
    Q = rrhip
    d = np.array([0.0, 0.0, 7.0]) - np.array([4.0, 8.0, -6.0])
    s.add_light(light(rrhip.LIGHT_DISTANT, intensity=0.85, p=tuple(d / np.linalg.norm(d))))
    s.add_light(light(rrhip.LIGHT_POINT, intensity=0.35, p=(-5, 4, 2)))
    add = s.add_quadric
    add(Q.QUADRIC_SPHERE, s.add_material(plastic((0.9, 0.25, 0.2), 0.15)), _place((-3.4, 1.2, 7)),
        radius=0.9, zmin=-0.9, zmax=0.9)
    add(Q.QUADRIC_TORUS, s.add_material(metal((0.95, 0.7, 0.15), 0.08)), _place((-1.1, 1.2, 7), -55),
        radius=0.65, height=0.28, zmin=0.0, zmax=360.0)
    add(Q.QUADRIC_CYLINDER, s.add_material(plastic((0.2, 0.55, 0.9))), _place((1.1, 1.2, 7), -90),
        radius=0.55, zmin=-0.85, zmax=0.85)
    add(Q.QUADRIC_CONE, s.add_material(plastic((0.25, 0.8, 0.35))), _place((3.4, 0.35, 7), -90),
        radius=0.75, height=1.7)
    add(Q.QUADRIC_PARABOLOID, s.add_material(plastic((0.7, 0.3, 0.85))), _place((-3.4, -1.6, 7), -90),
        radius=0.8, zmin=0.0, zmax=1.5)
    add(Q.QUADRIC_HYPERBOLOID, s.add_material(metal((0.9, 0.45, 0.6), 0.15)), _place((-1.1, -1.1, 7), -90),
        p1=(0.7, 0.0, -0.6), p2=(0.35, 0.35, 0.6))
    add(Q.QUADRIC_DISK, s.add_material(metal((0.85, 0.85, 0.9), 0.05)), _place((1.1, -1.1, 7), -65),
        radius=0.8, height=0.0)
    add(Q.QUADRIC_SPHERE, s.add_material(plastic((0.3, 0.85, 0.8))), _place((3.4, -1.1, 7)),
        radius=0.8, zmin=-0.8, zmax=0.8, thetamax=270.0)
    ground = s.add_material(matte((0.75, 0.75, 0.75)))
    s.add_polygon([(-12, -2.2, 0.5), (12, -2.2, 0.5), (12, -2.2, 18), (-12, -2.2, 18)], ground)
    return s


def hair_strands(n=400, seed=12345):
    """Control points (n x 4 x 3) of the hair-ball strands of render.rs's
    hair parity scene (tests/metal_pt_tests.rs:246-300): roots on a unit
    sphere by golden-angle spiral in longitude and a stepped latitude
    (z = -0.2 + 1.15 frac(0.377 i), clamped to [-0.2, 0.98]), each strand
    growing outward by 45 % with a 0.25 droop and a jitter of 0.08 t^2 from
    a 64-bit LCG (draws (state >> 33) / 2^31 - 1), coordinates rounded to 4
    decimals as the RIB text carries them."""
    state = seed
    mask = (1 << 64) - 1

    def draw():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        return (state >> 33) / float(1 << 31) - 1.0

    out = np.zeros((n, 4, 3))
    for i in range(n):
        phi = i * 0.61803 * 2.0 * np.pi
        z = min(max(-0.2 + 1.15 * ((i * 0.377) % 1.0), -0.2), 0.98)
        r = np.sqrt(max(1.0 - z * z, 0.0))
        nx, ny, nz = r * np.cos(phi), z, r * np.sin(phi)
        for k in range(4):
            t = k / 3.0
            g = 1.0 + 0.45 * t
            x = nx * g + 0.08 * t * t * draw()
            y = ny * g - 0.25 * t * t
            zz = nz * g + 0.08 * t * t * draw()
            out[i, k] = [round(x, 4), round(y, 4), round(zz, 4)]
    return out


def hair_ball(width=96, height=96, n=400, seed=12345):
    """render.rs's hair parity scene (tests/metal_pt_tests.rs:246-300): 400
    linear 4-point strands (1,200 rounded-cone segments, width 0.03 root to
    0.008 tip) of PxrMarschnerHair fiber (colour 0.5 0.3 0.12, roughness
    0.3, azimuthal roughness 0.35) around the origin, placed at (0, 0.4, 7);
    a grey matte ground at y = -1.3; a distant light from (4, 8, -6) towards
    (0, 0, 6) of intensity 1.4; background (0.4, 0.45, 0.55); 35-degree
    RenderMan camera at the origin."""
    s = Scene(width, height, 35.0, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0), filter=0, filter_width=1.0,
              background=(0.4, 0.45, 0.55))
    d = np.array([0.0, 0.0, 6.0]) - np.array([4.0, 8.0, -6.0])
    s.add_light(light(rrhip.LIGHT_DISTANT, intensity=1.4, p=tuple(d / np.linalg.norm(d))))
    ground = s.add_material(matte((0.6, 0.6, 0.6)))
    s.add_polygon([(-30, -1.3, -5), (30, -1.3, -5), (30, -1.3, 40), (-30, -1.3, 40)], ground)
    cs = s.add_curves(*dice_linear_strands(hair_strands(n, seed), 0.03, 0.008))
    fur = s.add_material(hair(color=(0.5, 0.3, 0.12), beta_m=0.3, beta_n=0.35))
    s.add_instance(cs, fur, _place((0.0, 0.4, 7.0)), curves=True)
    return s


def dice_linear_strands(strands, width_root, width_tip):
    """dice_curve over many linear curves of equal point count at once
    (strands: n x m x 3): the same segments, radii and v, bit for bit."""
    strands = np.asarray(strands, np.float64)
    n, m, _ = strands.shape
    k = np.arange(m, dtype=np.float64) / (m - 1)
    r = (width_root + (width_tip - width_root) * k) * 0.5
    ends = np.concatenate([strands, np.broadcast_to(r[None, :, None], (n, m, 1))], 2).astype(np.float32)
    v = np.broadcast_to(k.astype(np.float32), (n, m))
    return (ends[:, :-1].reshape(-1, 4), ends[:, 1:].reshape(-1, 4), v[:, :-1].reshape(-1).copy(),
            v[:, 1:].reshape(-1).copy())


def fur_ball(width=1920, height=1080, n=100000, seed=7):
    """A dense groom for throughput: the hair ball's layout (roots over
    the upper part of a unit sphere, strands growing 45 % outward with a
    droop, same fiber, light, ground and camera) with n strands of 4 points
    whose roots and jitter come from numpy's PCG64 (seed), 3 n segments."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.2, 0.98, n)
    r = np.sqrt(np.maximum(1 - z * z, 0))
    root = np.stack([r * np.cos(phi), z, r * np.sin(phi)], 1)
    t = np.arange(4) / 3.0
    g = 1.0 + 0.45 * t
    strands = root[:, None, :] * g[None, :, None]
    strands[:, :, 1] -= 0.25 * t * t
    jit = rng.uniform(-1, 1, (n, 4, 2)) * (0.08 * t * t)[None, :, None]
    strands[:, :, 0] += jit[:, :, 0]
    strands[:, :, 2] += jit[:, :, 1]
    s = Scene(width, height, 35.0, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0), filter=0, filter_width=1.0,
              background=(0.4, 0.45, 0.55))
    d = np.array([0.0, 0.0, 6.0]) - np.array([4.0, 8.0, -6.0])
    s.add_light(light(rrhip.LIGHT_DISTANT, intensity=1.4, p=tuple(d / np.linalg.norm(d))))
    ground = s.add_material(matte((0.6, 0.6, 0.6)))
    s.add_polygon([(-30, -1.3, -5), (30, -1.3, -5), (30, -1.3, 40), (-30, -1.3, 40)], ground)
    cs = s.add_curves(*dice_linear_strands(strands, 0.012, 0.003))
    fur = s.add_material(hair(color=(0.5, 0.3, 0.12), beta_m=0.3, beta_n=0.35))
    s.add_instance(cs, fur, _place((0.0, 0.4, 7.0)), curves=True)
    return s


BEZIER_BASIS = [[-1, 3, -3, 1], [3, -6, 3, 0], [-3, 3, 0, 0], [1, 0, 0, 0]]


def curve_zoo(width=48, height=48, hair_fiber=True, dome=0.5):
    """Every kind of curve geometry: a capsule, a segment tapering 0.35 ->
    0.1, a Bezier strand diced into 8 segments (width 0.16 -> 0.04), three
    point spheres, and a second copy of the capsule set rotated 30 degrees
    about x and scaled by 0.6; the capsule and strand of hair fiber (or
    matte with hair_fiber False), the rest matte; a rect light above, a
    constant dome, a ground quad. Object ids 1 (fiber), 2 (red), 3 (blue),
    4 (ground)."""
    s = Scene(width, height, 40.0, eye=(0, 0.5, -6), look_at=(0, 0, 0))
    s.add_light(light(rrhip.LIGHT_RECT, intensity=6.0, p=(-1, 3, -1), e1=(2, 0, 0), e2=(0, 0, 2)))
    if dome > 0:
        s.add_light(light(rrhip.LIGHT_DOME, intensity=dome))
    fur = s.add_material(hair(color=(0.6, 0.35, 0.15), beta_m=0.35, beta_n=0.4, object_id=1) if hair_fiber
                         else matte((0.6, 0.35, 0.15), object_id=1))
    red = s.add_material(matte((0.8, 0.2, 0.2), object_id=2))
    blue = s.add_material(matte((0.2, 0.3, 0.8), object_id=3))
    ground = s.add_material(matte((0.7, 0.7, 0.7), object_id=4))
    s.add_polygon([(-10, -1.2, -10), (10, -1.2, -10), (10, -1.2, 10), (-10, -1.2, 10)], ground)
    capsule = s.add_curves([[-1.5, 0.8, 0, 0.3]], [[0.2, 0.8, 0, 0.3]], [0.0], [1.0])
    taper = s.add_curves([[0.6, 0.6, 0, 0.35]], [[2.0, 1.0, 0.3, 0.1]], [0.0], [1.0])
    p0, p1, v0, v1 = dice_curve([[-2, -1, 0.5], [-1, 1, 0], [0.5, -1.5, -0.5], [1.8, 0.0, 0.5]],
                                (BEZIER_BASIS, 3), 0.16, 0.04, 8)
    strand = s.add_curves(p0, p1, v0, v1)
    pts = np.array([[0.8, -0.6, -0.5, 0.2], [1.4, -0.8, 0.2, 0.15], [2.0, -0.5, -0.3, 0.1]], np.float32)
    points = s.add_curves(pts, pts, np.full(3, 0.5), np.full(3, 0.5))
    s.add_instance(capsule, fur, _place((0, 0, 0)), curves=True)
    s.add_instance(taper, red, _place((0, 0, 0)), curves=True)
    s.add_instance(strand, fur, _place((0, 0, 0)), curves=True)
    s.add_instance(points, blue, _place((0, 0, 0)), curves=True)
    s.add_instance(capsule, red, _place((0.3, -1.6, 1.0), -30, 0.6), curves=True)
    return s


def smooth_texture(w, h, seed=0, cells=6):
    """A smooth random rgb texture (bilinear upsampling of a coarse random
    grid, periodic), values in [0.05, 0.95]."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.05, 0.95, (cells, cells, 3))
    ys = (np.arange(h) + 0.5) / h * cells - 0.5
    xs = (np.arange(w) + 0.5) / w * cells - 0.5
    y0, x0 = np.floor(ys).astype(int), np.floor(xs).astype(int)
    ay, ax = (ys - y0)[:, None, None], (xs - x0)[None, :, None]
    a = lambda yy, xx: g[yy[:, None] % cells, xx[None, :] % cells]
    img = (a(y0, x0) * (1 - ax) * (1 - ay) + a(y0, x0 + 1) * ax * (1 - ay) + a(y0 + 1, x0) * (1 - ax) * ay +
           a(y0 + 1, x0 + 1) * ax * ay)
    return img.astype(np.float32)


def stripes_texture(w, h, period=8):
    """Black and white columns `period` texels wide (cache.rs's EWA test
    texture)."""
    v = ((np.arange(w) // period) % 2 == 0).astype(np.float32)
    return np.repeat(np.broadcast_to(v[None, :], (h, w))[..., None], 3, axis=2).copy()


def pattern_zoo(width=64, height=64, lights=True):
    """Every pattern kind on every kind of texture coordinates: a ground
    quad (mesh st, tiled 4 x 4) with a smooth texture at grazing angles
    (EWA), a RenderMan sphere with a checker mixed over a ramp, a loose
    sphere with a triplanar stripe texture, a mesh quad spanning UDIM tiles
    1001 and 1002 (colour-corrected), a metal quad whose specular roughness
    follows a ramp over t, a box whose diffuse roughness and colour follow
    a fractal, a curve set coloured by a ramp over v; a rect light and a
    distant light. Object ids 1..7 in that order."""
    s = Scene(width, height, 45.0, eye=(0, 1.2, -7), look_at=(0, 0.3, 0))
    if lights:
        s.add_light(light(rrhip.LIGHT_RECT, intensity=5.0, p=(-1.5, 4, -2), e1=(3, 0, 0), e2=(0, 0, 3)))
        s.add_light(light(rrhip.LIGHT_DISTANT, intensity=1.5, p=(0.3, -1, 0.4)))
    tex = s.add_texture(smooth_texture(512, 512, seed=1, cells=24))
    stripes = s.add_texture(stripes_texture(64, 64, 4))
    tile_a = s.add_texture(smooth_texture(32, 32, seed=2, cells=4))
    tile_b = s.add_texture(smooth_texture(16, 32, seed=3, cells=3))
    # ground: mesh with st tiled 4x4, texture with scale 1
    ground = s.add_material(matte((0.5, 0.5, 0.5), object_id=1))
    gm = s.add_mesh([[-8, -1, -8], [8, -1, -8], [8, -1, 8], [-8, -1, 8]], [0, 1, 2, 0, 2, 3],
                    st=[[0, 0], [4, 0], [4, 4], [0, 4]])
    s.add_instance(gm, ground, np.eye(4)[:3])
    s.bind(ground, "diffuseColor", s.add_pattern(rrhip.PATTERN_TEXTURE, texture=tex))
    # RenderMan sphere: checker mixed with a ramp over s by a constant 0.3
    ball = s.add_material(material(diffuse_color=(0.5, 0.5, 0.5), specular_roughness=0.3, object_id=2))
    s.add_quadric(rrhip.QUADRIC_SPHERE, ball, _place((-1.8, 0.0, 0.5)), radius=1.0, zmin=-1.0, zmax=1.0)
    chk = s.add_pattern(rrhip.PATTERN_CHECKER, color_a=(0.9, 0.8, 0.2), color_b=(0.1, 0.2, 0.7), scale=(12, 6))
    ramp = s.add_pattern(rrhip.PATTERN_RAMP, stops=[(0.0, (0.9, 0.1, 0.1)), (0.5, (0.1, 0.9, 0.1)),
                                                    (1.0, (0.1, 0.1, 0.9))])
    s.bind(ball, "diffuseColor", s.add_pattern(rrhip.PATTERN_MIX, input=(chk, ramp, -1),
                                               constant=[(0, 0, 0), (0, 0, 0), (0.3, 0, 0)]))
    # loose sphere: triplanar stripes
    tri = s.add_material(matte((0.5, 0.5, 0.5), object_id=3))
    s.add_sphere((0.2, -0.3, 1.5), 0.7, tri)
    s.bind(tri, "diffuseColor", s.add_pattern(rrhip.PATTERN_TRIPLANAR, texture=stripes, frequency=0.8))
    # UDIM quad over tiles 1001 and 1002, colour-corrected
    ud = s.add_material(matte((0.5, 0.5, 0.5), object_id=4))
    um = s.add_mesh([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], [0, 1, 2, 0, 2, 3],
                    st=[[0, 0], [2, 0], [2, 1], [0, 1]])
    s.add_instance(um, ud, _place((1.0, -0.2, 2.5), 0, 1.2))
    un = s.add_pattern(rrhip.PATTERN_TEXTURE, udim={1001: tile_a, 1002: tile_b}, wrap=rrhip.WRAP_CLAMP)
    s.bind(ud, "diffuseColor", s.add_pattern(rrhip.PATTERN_COLOR_CORRECT, input=(un, -1, -1), cc_gain=(1.2, 1.0, 0.8),
                                             cc_offset=(0.02, 0.0, 0.05), gamma=1.4, saturation=1.3))
    # metal quad: specular roughness from a ramp over t
    met = s.add_material(material(diffuse_gain=0.0, specular_face_color=(0.9, 0.7, 0.5), specular_roughness=0.3,
                                  object_id=5))
    mm = s.add_mesh([[0, 0, 0], [1.6, 0, 0], [1.6, 1.6, 0], [0, 1.6, 0]], [0, 1, 2, 0, 2, 3],
                    st=[[0, 0], [1, 0], [1, 1], [0, 1]])
    s.add_instance(mm, met, _place((2.2, -0.9, 1.0), -25, 1.0))
    s.bind(met, "specularRoughness", s.add_pattern(rrhip.PATTERN_RAMP, use_t=1, stops=[(0.0, (0.05, 0, 0)),
                                                                                         (1.0, (0.6, 0, 0))]))
    # box (loose-triangle-free mesh, no st): fractal roughness and colour
    bx = s.add_material(matte((0.5, 0.5, 0.5), object_id=6))
    frac = s.add_pattern(rrhip.PATTERN_FRACTAL, frequency=3.0, octaves=4)
    s.bind(bx, "diffuseRoughness", frac)
    s.bind(bx, "diffuseColor", s.add_pattern(rrhip.PATTERN_COLOR_CORRECT, input=(frac, -1, -1), cc_gain=(0.9, 0.6, 0.3)))
    bpos = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32) - 0.5
    bidx = [0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1, 2, 3, 7, 2, 7, 6, 0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3]
    bm = s.add_mesh(bpos, bidx)
    s.add_instance(bm, bx, _place((-0.3, -0.55, -0.6), 30, 0.9))
    # curves: ramp over v
    cv = s.add_material(matte((0.5, 0.5, 0.5), object_id=7))
    p0, p1, v0, v1 = dice_curve([[-3.2, -1, -1], [-2.5, 1.5, -1.2], [-1.0, 0.5, -1.5], [-0.4, 1.8, -1.0]],
                                (BEZIER_BASIS, 3), 0.12, 0.06, 8)
    s.add_instance(s.add_curves(p0, p1, v0, v1), cv, _place((0, 0, 0)), curves=True)
    s.bind(cv, "diffuseColor", s.add_pattern(rrhip.PATTERN_RAMP, use_t=1, stops=[(0.0, (0.9, 0.9, 0.1)),
                                                                                  (1.0, (0.1, 0.6, 0.9))]))
    return s
