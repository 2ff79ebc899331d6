"""RIB front end: RenderMan scene files to a scene description for the GPU
path tracer.

Two stages, as in render.rs:

* `parse` (src/parser/parser.rs): the ASCII RIB token stream -> requests,
  each a name with a list of values (a number, a string, a numeric array
  or a string array; src/parser/ast.rs:1-45). `#` comments run to the end
  of the line.
* `Builder` (src/parser/scene_builder.rs): requests -> `scenes.Scene`
  with the builder's semantics: attribute and transform stacks, the
  legacy "matte" / "plastic" / "metal" surfaces and `Bxdf "PxrSurface"`
  parameters (src/scene/pbr.rs:100-185), polygons fan-triangulated into
  loose triangles (an area-light quad also becomes a rect light), the
  RenderMan quadrics as transformed quadric objects, PointsPolygons meshes
  as instances, ObjectBegin / ObjectInstance sharing meshes, classic and
  Pxr lights, Interior / Atmosphere media, the camera options and the id
  AOV's object ids.

Curves (linear, or cubic under the current v Basis, diced by ShadingRate)
and Points become curve sets placed by instances; `Bxdf
"PxrMarschnerHair"` makes hair fibers of them (scene_builder.rs:685-697,
1050-1170).
Patterns (Pattern requests and PxrSurface "reference" connections) build
the pattern graph and its bindings; .tex files are read by `rtex`.
PatchMesh / Patch / NuPatch, SubdivisionMesh, GeneralPolygon /
PointsGeneralPolygons and Displace "noise" are diced or triangulated on
the host (`tessellate`) into meshes. Requests the GPU path has no
counterpart (Procedural "DynamicLoad") raise `RibUnsupported` by default; with `strict=False` they are skipped
with one warning each, as render.rs skips requests it does not implement
(a motion block over other requests then keeps its first sample, as
render.rs does). Transform motion blocks move the mesh instances created
under them, and before WorldBegin the camera; PointsPolygons /
PointsGeneralPolygons motion blocks give the mesh a shutter-close
endpoint (scene_builder.rs:272-287, 1637-1692). Inline archives
(ArchiveBegin / ReadArchive) and file archives are replayed.
Requests that carry no render effect (Display, Declare, ShadingRate,
Sides, ...) are accepted silently.
"""
import math
import os
import warnings

import numpy as np

import rgbe
import rrhip
import rtex
import scenes
import tessellate


class RibError(ValueError):
    """Malformed RIB text."""


class RibUnsupported(RibError):
    """A request whose semantics this backend does not implement."""


# ---- tokenizer ---------------------------------------------------------------------

_NUM_START = set("0123456789+-.")


def _number(text, i):
    j = i
    n = len(text)
    if j < n and text[j] in "+-":
        j += 1
    while j < n and (text[j].isdigit() or text[j] == "."):
        j += 1
    if j < n and text[j] in "eE":
        k = j + 1
        if k < n and text[k] in "+-":
            k += 1
        if k < n and text[k].isdigit():
            j = k
            while j < n and text[j].isdigit():
                j += 1
    try:
        return float(text[i:j]), j
    except ValueError:
        raise RibError(f"bad number {text[i:j]!r} at offset {i}") from None


def _string(text, i):
    """A double-quoted string starting at text[i] == '"' (backslash escapes)."""
    out = []
    j = i + 1
    n = len(text)
    while j < n and text[j] != '"':
        if text[j] == "\\" and j + 1 < n:
            j += 1
            out.append({"n": "\n", "t": "\t"}.get(text[j], text[j]))
        else:
            out.append(text[j])
        j += 1
    if j >= n:
        raise RibError(f"unterminated string at offset {i}")
    return "".join(out), j + 1


def parse(text):
    """RIB text -> [(name, [values])]: a value is a float, a str, a tuple of
    floats (numeric array) or a list of str (string array)."""
    reqs = []
    i, n = 0, len(text)
    cur = None
    while i < n:
        c = text[i]
        if c.isspace():
            i += 1
        elif c == "#":
            e = text.find("\n", i)
            i = n if e < 0 else e + 1
        elif c == '"':
            s, i = _string(text, i)
            if cur is None:
                raise RibError(f"value before any request at offset {i}")
            cur[1].append(s)
        elif c == "[":
            i += 1
            nums, strs = [], []
            while True:
                while i < n and (text[i].isspace() or text[i] == "#"):
                    if text[i] == "#":
                        e = text.find("\n", i)
                        i = n if e < 0 else e
                    i += 1
                if i >= n:
                    raise RibError("unterminated array")
                if text[i] == "]":
                    i += 1
                    break
                if text[i] == '"':
                    s, i = _string(text, i)
                    strs.append(s)
                elif text[i] in _NUM_START:
                    v, i = _number(text, i)
                    nums.append(v)
                else:
                    raise RibError(f"unexpected {text[i]!r} in array at offset {i}")
            if nums and strs:
                raise RibError("array mixes numbers and strings")
            if cur is None:
                raise RibError("array before any request")
            cur[1].append(list(strs) if strs else tuple(nums))
        elif c in _NUM_START:
            v, i = _number(text, i)
            if cur is None:
                raise RibError("number before any request")
            cur[1].append(v)
        elif c.isalpha() or c == "_":
            j = i
            while j < n and (text[j].isalnum() or text[j] == "_"):
                j += 1
            cur = (text[i:j], [])
            reqs.append(cur)
            i = j
        else:
            raise RibError(f"unexpected character {c!r} at offset {i}")
    return reqs


# ---- value access (ast.rs RibValue / RibRequest / ParamList) ----------------------------

def _as_number(v):
    if isinstance(v, float):
        return v
    if isinstance(v, tuple) and len(v) == 1:
        return v[0]
    return None


def _as_str(v):
    if isinstance(v, str):
        return v
    if isinstance(v, list) and len(v) == 1:
        return v[0]
    return None


def _as_numbers(v):
    if isinstance(v, tuple):
        return v
    if isinstance(v, float):
        return (v,)
    return None


def _num(vals, i):
    return _as_number(vals[i]) if i < len(vals) else None


def _str(vals, i):
    return _as_str(vals[i]) if i < len(vals) else None


class _Params:
    """Token / value pairs from `start`; a token matches a name by its last
    word, so inline declarations ("float intensity") resolve."""

    def __init__(self, vals, start):
        self.items = []
        i = start
        while i + 1 < len(vals):
            tok = _as_str(vals[i])
            if tok is None:
                break
            self.items.append((tok, vals[i + 1]))
            i += 2

    def get(self, name):
        for tok, v in self.items:
            w = tok.split()
            if w and w[-1] == name:
                return v
        return None

    def number(self, name, default=None):
        v = self.get(name)
        r = None if v is None else _as_number(v)
        return default if r is None else r

    def numbers(self, name):
        v = self.get(name)
        return None if v is None else _as_numbers(v)

    def string(self, name):
        v = self.get(name)
        return None if v is None else _as_str(v)

    def color(self, name, default):
        v = self.numbers(name)
        return tuple(v[:3]) if v is not None and len(v) >= 3 else default


def _pstart(vals):
    """Light / Bxdf / AreaLightSource parameters start after the type and
    an optional handle: the handle is there when the value count is even."""
    return 2 if len(vals) % 2 == 0 else 1


# ---- transforms (math/matrix.rs: column-vector 4x4) ---------------------------------

def _translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def _scale(x, y, z):
    return np.diag([x, y, z, 1.0])


def _rotate(deg, x, y, z):
    """Rotation by `deg` degrees about axis (x, y, z) (matrix.rs:81-113)."""
    a = np.array([x, y, z], float)
    a /= np.linalg.norm(a)
    x, y, z = a
    r = math.radians(deg)
    c, s = math.cos(r), math.sin(r)
    t = 1.0 - c
    m = np.eye(4)
    m[:3, :3] = [[c + x * x * t, x * y * t - z * s, x * z * t + y * s],
                 [y * x * t + z * s, c + y * y * t, y * z * t - x * s],
                 [z * x * t - y * s, z * y * t + x * s, c + z * z * t]]
    return m


def _from_rib(m):
    """RIB's 16 row-major row-vector numbers -> the column-vector matrix
    (scene_builder.rs:1925-1938: a transpose)."""
    if len(m) != 16:
        return None
    return np.array(m, float).reshape(4, 4).T


def _pt(m, p):
    return tuple(float(x) for x in m[:3, :3] @ np.asarray(p, float) + m[:3, 3])


def _vec(m, v):
    return m[:3, :3] @ np.asarray(v, float)


# ---- scene builder -------------------------------------------------------------------

_FILTERS = {"box": 0, "triangle": 1, "gaussian": 2}

# Requests with no effect on the rendered image (recorded or ignored by
# render.rs too: scene_builder.rs:271, 310, 541-575, 662-678).
_NO_EFFECT = {"version", "WorldEnd", "FrameBegin", "FrameEnd", "Display", "Declare", "Option",
              "Clipping", "ClippingPlane", "CropWindow", "FrameAspectRatio", "PixelVariance",
              "Exposure", "Quantize", "Hider", "Integrator", "TextureCoordinates",
              "ShadingInterpolation", "RelativeDetail", "Opacity", "Orientation",
              "ReverseOrientation", "Sides", "Shutter", "Exterior",
              "CoordinateSystem", "ScopedCoordinateSystem", "IfBegin", "ElseIf", "Else", "IfEnd",
              "Attribute"}

# Requests render.rs implements that this backend does not.
_UNSUPPORTED = set()

# RiSpec cubic bases and their default steps (patches.rs:10-53); the
# matrix M of p(u) = [u^3 u^2 u 1] M G, row-major.
_BASES = {
    "bezier": ([[-1, 3, -3, 1], [3, -6, 3, 0], [-3, 3, 0, 0], [1, 0, 0, 0]], 3),
    "b-spline": ([[-1 / 6, 3 / 6, -3 / 6, 1 / 6], [3 / 6, -6 / 6, 3 / 6, 0], [-3 / 6, 0, 3 / 6, 0],
                  [1 / 6, 4 / 6, 1 / 6, 0]], 1),
    "catmull-rom": ([[-0.5, 1.5, -1.5, 0.5], [1, -2.5, 2, -0.5], [-0.5, 0, 0.5, 0], [0, 1, 0, 0]], 1),
    "hermite": ([[2, 1, -2, 1], [-3, -2, 3, -1], [0, 1, 0, 0], [1, 0, 0, 0]], 2),
    "power": (np.eye(4).tolist(), 4),
}
_BASES["bspline"] = _BASES["b-spline"]
_BASES["catmullrom"] = _BASES["catmull-rom"]


dice_curve = scenes.dice_curve

_QUADRIC_ARGS = {"Sphere": 4, "Cylinder": 4, "Cone": 3, "Torus": 5, "Disk": 3, "Paraboloid": 4,
                 "Hyperboloid": 7}


class _State:
    """GraphicsState defaults (scene_builder.rs:55-80)."""

    def __init__(self):
        self.color = (1.0, 1.0, 1.0)
        self.surface = "matte"
        self.roughness = 0.1
        self.area_light = None  # (intensity, color)
        self.bxdf = None        # PxrSurface parameters as keyword arguments of scenes.material
        self.interior = None
        self.identifier = None
        self.motion = None      # (t0, t1) of the last transform motion block
        self.hair = None        # PxrMarschnerHair parameters as keyword arguments of scenes.hair
        self.bxdf_bindings = []  # PxrSurface "reference" connections: (field name, pattern node)
        self.displace = None    # Displace "noise" parameters (displace.rs DisplaceParams) for diced meshes
        self.basis_u = _BASES["bezier"]
        self.basis_v = _BASES["bezier"]
        self.shading_rate = 1.0

    def copy(self):
        s = _State()
        s.__dict__.update(self.__dict__)
        return s


def _pxr_surface(p):
    """PbrParams::from_bxdf_params (pbr.rs:133-185) as scenes.material
    keyword arguments."""
    kw = dict(diffuse_gain=p.number("diffuseGain", 1.0),
              diffuse_color=p.color("diffuseColor", (0.18, 0.18, 0.18)),
              diffuse_roughness=p.number("diffuseRoughness", 0.0),
              specular_face_color=p.color("specularFaceColor", (0.0, 0.0, 0.0)),
              specular_edge_color=p.color("specularEdgeColor", (1.0, 1.0, 1.0)),
              specular_roughness=p.number("specularRoughness", 0.2),
              specular_ior=p.number("specularIor", 1.5),
              clearcoat_gain=p.number("clearcoatGain", 0.0),
              clearcoat_roughness=p.number("clearcoatRoughness", 0.1),
              fuzz_gain=p.number("fuzzGain", 0.0),
              fuzz_color=p.color("fuzzColor", (1.0, 1.0, 1.0)),
              glass_gain=p.number("glassRefractionGain", 0.0),
              glass_ior=p.number("glassIor", 1.5),
              glass_roughness=p.number("glassRoughness", 0.05),
              refraction_color=p.color("refractionColor", (1.0, 1.0, 1.0)),
              presence=p.number("presence", 1.0))
    gg = p.number("glowGain", 0.0)
    gc = p.color("glowColor", (1.0, 1.0, 1.0))
    kw["glow"] = tuple(gg * c for c in gc)
    for k in ("specular_roughness", "glass_roughness", "clearcoat_roughness"):
        kw[k] = min(max(kw[k], 0.005), 1.0)
    return kw


class Builder:
    """SceneBuilder (scene_builder.rs) over scenes.Scene. After `run`,
    `scene()` returns the scene and `spp` the PixelSamples product."""

    def __init__(self, base_dir=None, strict=True):
        self.base_dir = base_dir
        self.strict = strict
        self.width, self.height, self.fov = 640, 480, 45.0
        self.projection = 0
        self.dof = None
        self.filter = (0, 1.0)
        self.screen_window = None
        self.pixel_samples = (1, 1)
        self.background = (0.0, 0.0, 0.0)
        self.state = _State()
        self.attr_stack = []
        self.xf_stack = [np.eye(4)]
        self.object_defs = {}
        self.defining = None  # (handle, [(mesh, local transform, material)])
        self.object_ids = {}
        self.next_object_id = 1
        self.warned = set()
        # scene data, collected until the camera is known
        self.materials, self.lights, self.tris, self.quadrics = [], [], [], []
        self.meshes, self.instances, self.media = [], [], []
        self.curve_sets, self.curve_instances = [], []
        self.atmosphere = -1
        self.env = None
        self.archives = {}
        self.recording = None  # (name, [requests]) between ArchiveBegin and ArchiveEnd
        self.motion = None     # requests of an open MotionBegin block
        self.archive_depth = 0
        self.camera_motion_inv = None
        self.pending_deform_p = None  # "P" of a deformation block's last sample
        # pattern graph (scene_builder.rs:1703-1848): textures (level 0 of each
        # .tex file, by path), nodes as scenes.Scene.add_pattern arguments,
        # handles, and the materials' bindings
        self.textures, self.tex_index = [], {}
        self.patterns, self.pattern_handles, self.bindings = [], {}, []

    # -- helpers
    def _skip(self, what):
        if self.strict:
            raise RibUnsupported(what)
        if what not in self.warned:
            self.warned.add(what)
            warnings.warn(f"RIB: {what}; skipping")

    def _warn(self, what):
        """warn_once: the request is still honoured without the named part."""
        if what not in self.warned:
            self.warned.add(what)
            warnings.warn(f"RIB: {what}")

    @property
    def xf(self):
        return self.xf_stack[-1]

    def _apply(self, m):
        self.xf_stack[-1] = self.xf_stack[-1] @ m

    def _object_id(self):
        """assign_object_id (scene_builder.rs:1888-1903)."""
        name = self.state.identifier or f"object_{self.next_object_id}"
        if name in self.object_ids:
            return self.object_ids[name]
        oid = self.next_object_id
        self.next_object_id += 1
        self.object_ids[name] = oid
        return oid

    def _material(self):
        """make_material (scene_builder.rs:1850-1879) -> a new material index."""
        st = self.state
        if st.hair is not None:
            m = scenes.hair(**st.hair)
        elif st.bxdf is not None:
            m = scenes.material(**st.bxdf)
        elif st.surface == "plastic":
            m = scenes.plastic(st.color, st.roughness)
        elif st.surface == "metal":
            m = scenes.metal(st.color, st.roughness)
        else:
            m = scenes.matte(st.color)
        m.interior = 0 if st.interior is None else st.interior + 1
        m.id = self._object_id()
        self.materials.append(m)
        if st.hair is None and st.bxdf is not None:
            for field, node in st.bxdf_bindings:
                self.bindings.append((len(self.materials) - 1, field, node))
        return len(self.materials) - 1

    def _emitter(self, color, scale, light_index):
        """The black matte stand-in of a Pxr area light (scene_builder.rs:1501-1507)."""
        m = scenes.matte((0.0, 0.0, 0.0), emission=tuple(c * scale for c in color), area_light=light_index)
        self.materials.append(m)
        return len(self.materials) - 1

    def _quadric(self, kind, mat, m, **kw):
        self.quadrics.append((kind, mat, m[:3, :4].copy(), kw))

    # -- requests
    def run(self, reqs):
        """scene_builder.rs:246-266: archive recording and motion blocks
        capture requests verbatim before they are processed."""
        for name, vals in reqs:
            if self.recording is not None:
                if name == "ArchiveEnd":
                    self.archives[self.recording[0]] = self.recording[1]
                    self.recording = None
                else:
                    self.recording[1].append((name, vals))
                continue
            if self.motion is not None and name != "MotionEnd":
                self.motion.append((name, vals))
                continue
            self.process(name, vals)
        return self

    def _read_archive(self, name):
        if self.archive_depth >= 16:
            raise RibError(f'ReadArchive "{name}": nesting deeper than 16')
        reqs = self.archives.get(name)
        if reqs is None:
            path = name if self.base_dir is None else os.path.join(self.base_dir, name)
            if not os.path.exists(path):
                return self._skip(f'ReadArchive "{name}" not found')
            with open(path, "r", encoding="utf-8", errors="replace") as f:
                reqs = parse(f.read())
        self.archive_depth += 1
        try:
            self.run(reqs)
        finally:
            self.archive_depth -= 1

    def _procedural(self, vals):
        """Procedural (scene_builder.rs:381-438): "DelayedReadArchive" reads
        its archive at once (bounds ignored); "RunProgram" runs the program
        (relative to the RIB's directory, extra arguments split on
        whitespace, in that directory) and replays the RIB it prints;
        "DynamicLoad" has no counterpart."""
        kind = _str(vals, 0) or ""
        raw = vals[1] if len(vals) > 1 else None
        if isinstance(raw, str):
            args = [raw]
        elif isinstance(raw, list):
            args = [a for a in raw if isinstance(a, str)]
        else:
            args = []
        if kind == "DelayedReadArchive":
            self._warn('Procedural "DelayedReadArchive" loads eagerly (no lazy loading)')
            if args:
                self._read_archive(args[0])
        elif kind == "RunProgram":
            if not args:
                return self._warn('Procedural "RunProgram" without a program; skipping')
            import subprocess
            prog = args[0] if self.base_dir is None or os.path.isabs(args[0]) else \
                os.path.join(self.base_dir, args[0])
            cmd = [prog] + (args[1].split() if len(args) > 1 else [])
            try:
                out = subprocess.run(cmd, cwd=self.base_dir, capture_output=True, timeout=600)
            except (OSError, subprocess.SubprocessError) as e:
                return self._warn(f"RunProgram {args[0]}: {e}")
            if out.returncode != 0:
                return self._warn(f"RunProgram {args[0]} exited with {out.returncode}")
            try:
                reqs = parse(out.stdout.decode("utf-8", errors="replace"))
            except RibError as e:
                return self._warn(f"RunProgram {args[0]}: output failed to parse: {e}")
            if self.archive_depth >= 16:
                raise RibError(f'RunProgram "{args[0]}": nesting deeper than 16')
            self.archive_depth += 1
            try:
                self.run(reqs)
            finally:
                self.archive_depth -= 1
        else:
            self._skip(f'Procedural "{kind}" is not implemented (DynamicLoad)')

    def process(self, name, vals):
        st = self.state
        if name == "WorldBegin":
            if st.motion is not None:  # a motion block before the world moves the camera
                t0, t1 = st.motion
                self.camera_motion_inv = np.linalg.inv(t1 @ np.linalg.inv(t0))
                st.motion = None
            return
        if name in _NO_EFFECT and name not in ("Option", "Attribute"):
            return
        if name in _UNSUPPORTED:
            return self._skip(f"request {name} is not supported by this backend")
        if name == "ArchiveBegin":
            self.recording = (_str(vals, 0) or "", [])
        elif name == "ArchiveEnd":
            pass  # a stray end
        elif name == "ReadArchive":
            if _str(vals, 0) is not None:
                self._read_archive(_str(vals, 0))
        elif name == "Procedural":
            self._procedural(vals)
        elif name == "MotionBegin":
            self.motion = []
        elif name == "MotionEnd":
            block, self.motion = self.motion or [], None
            self._motion_block(block)
        elif name == "AttributeBegin":
            self.attr_stack.append(st.copy())
            self.xf_stack.append(self.xf.copy())
        elif name == "AttributeEnd":
            if self.attr_stack:
                self.state = self.attr_stack.pop()
            if len(self.xf_stack) > 1:
                self.xf_stack.pop()
        elif name == "TransformBegin":
            self.xf_stack.append(self.xf.copy())
        elif name == "TransformEnd":
            if len(self.xf_stack) > 1:
                self.xf_stack.pop()
        elif name == "ObjectBegin":
            handle = _str(vals, 0) or (str(_num(vals, 0)) if _num(vals, 0) is not None else "0")
            self.attr_stack.append(st.copy())
            self.xf_stack.append(np.eye(4))
            self.defining = (handle, [])
        elif name == "ObjectEnd":
            if self.defining is not None:
                self.object_defs[self.defining[0]] = self.defining[1]
                self.defining = None
            if self.attr_stack:
                self.state = self.attr_stack.pop()
            if len(self.xf_stack) > 1:
                self.xf_stack.pop()
        elif name == "ObjectInstance":
            handle = _str(vals, 0) or (str(_num(vals, 0)) if _num(vals, 0) is not None else "0")
            entries = self.object_defs.get(handle)
            if entries is None:
                self._skip(f'ObjectInstance "{handle}" has no definition')
                return
            m1 = self._endpoint(self.xf)
            for mesh, local, mat in entries:
                self.instances.append((mesh, mat, (self.xf @ local)[:3, :4].copy(),
                                       None if m1 is None else (m1 @ local)[:3, :4].copy()))
        elif name == "Format":
            if _num(vals, 0) is not None and _num(vals, 1) is not None:
                self.width, self.height = int(_num(vals, 0)), int(_num(vals, 1))
        elif name == "Projection":
            kind = _str(vals, 0)
            if kind == "perspective":
                self.projection = 0
                self.fov = _Params(vals, 1).number("fov", self.fov)
            elif kind == "orthographic":
                self.projection = 1
            else:
                self._skip(f'Projection "{kind}" is not supported')
        elif name == "DepthOfField":
            f = [_num(vals, i) for i in range(3)]
            if all(x is not None for x in f):
                self.dof = tuple(f)
        elif name == "PixelFilter":
            fname = _str(vals, 0)
            if fname is not None:
                xw = _num(vals, 1) if _num(vals, 1) is not None else 2.0
                yw = _num(vals, 2) if _num(vals, 2) is not None else xw
                w = (xw + yw) * 0.5
                self.filter = (_FILTERS[fname], max(w, 1e-3)) if fname in _FILTERS else (0, 1.0)
        elif name == "ScreenWindow":
            f = [_num(vals, i) for i in range(4)]
            if all(x is not None for x in f):
                self.screen_window = tuple(f)
        elif name == "PixelSamples":
            if _num(vals, 0) is not None and _num(vals, 1) is not None:
                self.pixel_samples = (max(int(_num(vals, 0)), 1), max(int(_num(vals, 1)), 1))
        elif name == "Option":
            if _str(vals, 0) == "background":
                c = _Params(vals, 1).numbers("color")
                if c is not None and len(c) >= 3:
                    self.background = tuple(c[:3])
        elif name == "Attribute":
            if _str(vals, 0) == "identifier":
                n = _Params(vals, 1).string("name")
                if n is not None:
                    st.identifier = n
        elif name == "Identity":
            self.xf_stack[-1] = np.eye(4)
        elif name in ("Transform", "ConcatTransform"):
            m = _as_numbers(vals[0]) if vals else None
            m = None if m is None else _from_rib(m)
            if m is not None:
                if name == "Transform":
                    self.xf_stack[-1] = m
                else:
                    self._apply(m)
        elif name == "Translate":
            f = [_num(vals, i) for i in range(3)]
            if all(x is not None for x in f):
                self._apply(_translate(*f))
        elif name == "Rotate":
            f = [_num(vals, i) for i in range(4)]
            if all(x is not None for x in f):
                self._apply(_rotate(*f))
        elif name == "Scale":
            f = [_num(vals, i) for i in range(3)]
            if all(x is not None for x in f):
                self._apply(_scale(*f))
        elif name == "Color":
            f = [_num(vals, i) for i in range(3)]
            if all(x is not None for x in f):
                st.color = tuple(f)
            elif vals and _as_numbers(vals[0]) is not None and len(_as_numbers(vals[0])) >= 3:
                st.color = tuple(_as_numbers(vals[0])[:3])
        elif name == "Surface":
            s = _str(vals, 0)
            if s is not None:
                st.surface = s
                st.roughness = _Params(vals, 1).number("roughness", 0.1)
        elif name == "Bxdf":
            kind = _str(vals, 0) or ""
            p = _Params(vals, _pstart(vals))
            if kind == "PxrMarschnerHair":  # HairParams::from_bxdf_params (hair.rs:79-98)
                hp = dict(beta_m=p.number("roughness", 0.3), beta_n=p.number("azimuthalRoughness", 0.3),
                          eta=p.number("ior", 1.55), color=p.color("color", (0.35, 0.2, 0.06)))
                a = p.numbers("absorption")
                if a:
                    hp["absorption"] = a
                st.hair, st.bxdf, st.bxdf_bindings = hp, None, []
                return
            st.hair = None
            if kind != "PxrSurface":
                self._skip(f'Bxdf "{kind}" treated as PxrSurface')
            if p.number("subsurfaceGain", 0.0) > 0.0:
                self._skip("PxrSurface subsurface is not supported by this backend")
            st.bxdf = _pxr_surface(p)
            st.bxdf_bindings = self._bxdf_bindings(p)
        elif name == "Pattern":
            self._pattern(vals)
        elif name == "Displace":
            self._displace(vals)
        elif name == "AreaLightSource":
            p = _Params(vals, _pstart(vals))
            st.area_light = (p.number("intensity", 1.0), p.color("lightcolor", (1.0, 1.0, 1.0)))
        elif name == "LightSource":
            self._light_source(vals)
        elif name == "Light":
            self._pxr_light(vals)
        elif name in ("Interior", "Atmosphere"):
            self._medium(name, vals)
        elif name == "Basis":
            self._basis(vals)
        elif name == "ShadingRate":
            if _num(vals, 0) is not None:
                st.shading_rate = _num(vals, 0)
        elif name == "Curves":
            self._curves(vals)
        elif name == "Points":
            self._points(vals)
        elif name in _QUADRIC_ARGS:
            self._quadric_request(name, vals)
        elif name == "Polygon":
            self._polygon(vals)
        elif name == "PointsPolygons":
            self._points_polygons(vals)
        elif name == "GeneralPolygon":
            self._general_polygon(vals)
        elif name == "PointsGeneralPolygons":
            self._points_general_polygons(vals)
        elif name in ("PatchMesh", "Patch"):
            self._patch_mesh(name, vals)
        elif name == "NuPatch":
            self._nu_patch(vals)
        elif name in ("SubdivisionMesh", "HierarchicalSubdivisionMesh"):
            if name == "HierarchicalSubdivisionMesh":
                self._warn("HierarchicalSubdivisionMesh treated as SubdivisionMesh (string tags ignored)")
            self._subdivision_mesh(vals)
        else:
            self._skip(f"request {name} is not implemented")

    _XF_REQS = ("Translate", "Rotate", "Scale", "ConcatTransform", "Transform", "Identity")

    def _motion_block(self, block):
        """finish_motion_block (scene_builder.rs:1637-1684): a block of
        transforms records the shutter-open (first sample) and -close (last
        sample) transforms; anything else keeps its first sample."""
        if not block:
            return
        if all(n in self._XF_REQS for n, _ in block):
            base = self.xf.copy()
            self.process(*block[0])
            t0 = self.xf.copy()
            if len(block) > 1:
                self.xf_stack[-1] = base
                self.process(*block[-1])
                t1 = self.xf.copy()
                self.xf_stack[-1] = t0
                self.state.motion = (t0, t1)
            return
        if all(n == block[0][0] for n, _ in block) and block[0][0] in ("PointsPolygons", "PointsGeneralPolygons"):
            # deformation motion: the last sample's "P" is the mesh's
            # shutter-close endpoint (pending_deform_p)
            if len(block) > 1:
                lead = 2 if block[0][0] == "PointsPolygons" else 3
                self.pending_deform_p = _Params(block[-1][1], lead).numbers("P")
            self.process(*block[0])
            self.pending_deform_p = None
            return
        self._skip(f'MotionBegin over "{block[0][0]}" is not supported by this backend')
        self.process(*block[0])

    def _endpoint(self, current):
        """motion_endpoint (scene_builder.rs:1686-1692): the shutter-close
        transform of geometry created now, t1 * t0^-1 * current."""
        if self.state.motion is None:
            return None
        t0, t1 = self.state.motion
        return t1 @ np.linalg.inv(t0) @ current

    def _light_source(self, vals):
        """parse_light (scene_builder.rs:1582-1635): positions are taken as
        given (not through the current transform), colour white."""
        kind = _str(vals, 0)
        p = _Params(vals, _pstart(vals))
        inten = p.number("intensity", 1.0)
        if kind == "pointlight":
            if p.string("iesProfile") is not None:
                return self._skip("IES profiles are not supported by this backend")
            f = p.numbers("from")
            pos = tuple(f[:3]) if f is not None and len(f) >= 3 else (0.0, 0.0, 0.0)
            self.lights.append(scenes.light(rrhip.LIGHT_POINT, intensity=inten, p=pos))
        elif kind == "distantlight":
            f = p.numbers("from")
            t = p.numbers("to")
            a = np.array(f[:3] if f is not None and len(f) >= 3 else (0.0, 0.0, -10.0))
            b = np.array(t[:3] if t is not None and len(t) >= 3 else (0.0, 0.0, 0.0))
            d = b - a
            self.lights.append(scenes.light(rrhip.LIGHT_DISTANT, intensity=inten,
                                            p=tuple(d / np.linalg.norm(d))))
        # other classic shaders: no light (render.rs returns None)

    def _pxr_light(self, vals):
        """build_light (scene_builder.rs:1488-1580)."""
        kind = _str(vals, 0)
        if kind is None:
            return
        p = _Params(vals, _pstart(vals))
        scale = p.number("intensity", 1.0) * 2.0 ** p.number("exposure", 0.0)
        color = p.color("lightColor", (1.0, 1.0, 1.0))
        xf = self.xf
        origin = _pt(xf, (0, 0, 0))
        li = len(self.lights)
        if kind == "PxrRectLight":
            corner = _pt(xf, (-0.5, -0.5, 0.0))
            self.lights.append(scenes.light(rrhip.LIGHT_RECT, intensity=scale, color=color, p=corner,
                                            e1=tuple(_vec(xf, (1, 0, 0))), e2=tuple(_vec(xf, (0, 1, 0)))))
            mat = self._emitter(color, scale, li)
            v = [_pt(xf, (x, y, 0.0)) for x, y in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5))]
            self.tris.append((v[0], v[1], v[2], mat))
            self.tris.append((v[0], v[2], v[3], mat))
        elif kind == "PxrSphereLight":
            radius = float(np.linalg.norm(_vec(xf, (0.5, 0, 0))))
            self.lights.append(scenes.light(rrhip.LIGHT_SPHERE, intensity=scale, color=color, p=origin,
                                            radius=radius))
            self._quadric(rrhip.QUADRIC_SPHERE, self._emitter(color, scale, li), xf, radius=0.5, zmin=-0.5,
                          zmax=0.5, thetamax=360.0)
        elif kind == "PxrDiskLight":
            self.lights.append(scenes.light(rrhip.LIGHT_DISK, intensity=scale, color=color, p=origin,
                                            e1=tuple(_vec(xf, (0.5, 0, 0))), e2=tuple(_vec(xf, (0, 0.5, 0)))))
            self._quadric(rrhip.QUADRIC_DISK, self._emitter(color, scale, li), xf, radius=0.5, height=0.0,
                          thetamax=360.0)
        elif kind == "PxrDistantLight":
            d = _vec(xf, (0, 0, -1))
            ang = p.number("angleExtent", 0.53)
            self.lights.append(scenes.light(rrhip.LIGHT_DISTANT, intensity=scale, color=color,
                                            p=tuple(d / np.linalg.norm(d)), radius=math.radians(ang * 0.5)))
        elif kind == "PxrDomeLight":
            mp = p.string("lightColorMap")
            if mp is not None:
                if self.env is not None:
                    return self._skip("a second environment map")
                path = mp if self.base_dir is None else os.path.join(self.base_dir, mp)
                self.env = rgbe.read_hdr(path)
            self.lights.append(scenes.light(rrhip.LIGHT_DOME, intensity=scale, color=color))
        else:
            self._skip(f'Light "{kind}" is not implemented')

    def _medium(self, name, vals):
        """scene_builder.rs:328-373: homogeneous media, or with density
        "fbm" / "noise" render.rs's procedural cloud field (frequency 0.3,
        5 octaves, gain 0.55, lacunarity 2, coverage 0.5, sharpness 4 by
        default); another density name warns and stays homogeneous."""
        p = _Params(vals, 1)
        dens = p.string("density")
        fbm = None
        if dens in ("fbm", "noise"):
            fbm = dict(frequency=p.number("frequency", 0.3), octaves=int(p.number("octaves", 5.0)),
                       gain=p.number("gain", 0.55), lacunarity=p.number("lacunarity", 2.0),
                       coverage=p.number("coverage", 0.5), sharpness=p.number("sharpness", 4.0))
        elif dens is not None:
            self._warn(f'{name} density "{dens}" not implemented; homogeneous')
        far = 1200.0 if name == "Atmosphere" else 0.0  # 0: unbounded
        idx = len(self.media)
        self.media.append(dict(sigma_a=p.color("sigma_a", (0.1, 0.1, 0.1)),
                               sigma_s=p.color("sigma_s", (0.5, 0.5, 0.5)), g=p.number("g", 0.0),
                               emission=p.color("emission", (0.0, 0.0, 0.0)),
                               max_distance=p.number("maxdistance", far), fbm=fbm))
        if name == "Atmosphere":
            self.atmosphere = idx
        else:
            self.state.interior = idx

    def _quadric_request(self, name, vals):
        v = [_num(vals, i) for i in range(_QUADRIC_ARGS[name])]
        if any(x is None for x in v):
            return
        if self.defining is not None:
            return self._skip(f"{name} inside ObjectBegin is not instanced")
        mat = self._material()
        Q = rrhip
        if name == "Sphere":
            self._quadric(Q.QUADRIC_SPHERE, mat, self.xf, radius=v[0], zmin=v[1], zmax=v[2], thetamax=v[3])
        elif name == "Cylinder":
            self._quadric(Q.QUADRIC_CYLINDER, mat, self.xf, radius=v[0], zmin=v[1], zmax=v[2], thetamax=v[3])
        elif name == "Cone":
            self._quadric(Q.QUADRIC_CONE, mat, self.xf, height=v[0], radius=v[1], thetamax=v[2])
        elif name == "Torus":
            self._quadric(Q.QUADRIC_TORUS, mat, self.xf, radius=v[0], height=v[1], zmin=v[2], zmax=v[3],
                          thetamax=v[4])
        elif name == "Disk":
            self._quadric(Q.QUADRIC_DISK, mat, self.xf, height=v[0], radius=v[1], thetamax=v[2])
        elif name == "Paraboloid":
            self._quadric(Q.QUADRIC_PARABOLOID, mat, self.xf, radius=v[0], zmin=v[1], zmax=v[2], thetamax=v[3])
        else:
            self._quadric(Q.QUADRIC_HYPERBOLOID, mat, self.xf, p1=tuple(v[0:3]), p2=tuple(v[3:6]), thetamax=v[6])

    def _basis(self, vals):
        """Basis <u> <ustep> <v> <vstep> (scene_builder.rs:576-601): values
        0 and 2 a name or a 16-number matrix, 1 and 3 the steps (default:
        the named basis's, 1 for a matrix); the v basis dices curves, both
        dice bicubic patches."""
        def one(v, step):
            if v is None:
                return None
            name = _as_str(v)
            if name is not None:
                if name not in _BASES:
                    return None
                m, dflt = _BASES[name]
                return m, int(step) if step is not None else dflt
            nums = _as_numbers(v)
            if nums is None or len(nums) != 16:
                return None
            return np.asarray(nums, float).reshape(4, 4).tolist(), int(step) if step is not None else 1
        a = one(vals[0] if len(vals) > 0 else None, _num(vals, 1))
        if a is not None:
            self.state.basis_u = a
        b = one(vals[2] if len(vals) > 2 else None, _num(vals, 3))
        if b is not None:
            self.state.basis_v = b

    def _curve_set(self, p0, p1, v0, v1):
        """add_curve_set (scene_builder.rs:1156-1170): a curve set placed by an
        instance at the current transform (moving under a motion block)."""
        mat = self._material()
        m1 = self._endpoint(self.xf)
        self.curve_sets.append((p0, p1, v0, v1))
        self.curve_instances.append((len(self.curve_sets) - 1, mat, self.xf[:3, :4].copy(),
                                     None if m1 is None else m1[:3, :4].copy()))

    def _curves(self, vals):
        """curves_request (scene_builder.rs:1050-1119): "linear" or cubic
        (the v basis) curves, "width" root -> tip (or "constantwidth", 0.01),
        8 / ShadingRate segments per cubic span (2-24); "periodic" is taken
        as nonperiodic."""
        if self.defining is not None:
            return self._skip("Curves inside ObjectBegin is not instanced")
        kind = _str(vals, 0) or "cubic"
        nverts = _as_numbers(vals[1]) if len(vals) > 1 else None
        if nverts is None:
            return self._skip("Curves without [nvertices]")
        if (_str(vals, 2) or "nonperiodic") == "periodic" and "periodic" not in self.warned:
            self.warned.add("periodic")  # render.rs warns and dices the open curve
            warnings.warn('RIB: Curves "periodic" taken as nonperiodic')
        p = _Params(vals, 3)
        P = p.numbers("P")
        if P is None:
            return self._skip('Curves without "P"')
        w = p.numbers("width")
        if w:
            wr, wt = w[0], w[-1]
        else:
            wr = wt = p.number("constantwidth", 0.01)
        cubic = None if kind == "linear" else self.state.basis_v
        segs = int(min(max(8.0 / max(self.state.shading_rate, 0.1), 2), 24))
        parts = []
        cur = 0
        for nv in nverts:
            nv = int(nv)
            if (cur + nv) * 3 > len(P):
                break
            parts.append(dice_curve(np.asarray(P[3 * cur:3 * (cur + nv)], float), cubic, wr, wt, segs))
            cur += nv
        parts = [q for q in parts if len(q[0])]
        if not parts:
            return
        self._curve_set(*(np.concatenate([q[k] for q in parts]) for k in range(4)))

    def _points(self, vals):
        """points_request (scene_builder.rs:1121-1154): each point a
        zero-length segment of radius width / 2 ("width" per point, else
        "constantwidth", 0.05), v = 0.5."""
        if self.defining is not None:
            return self._skip("Points inside ObjectBegin is not instanced")
        p = _Params(vals, 0)
        P = p.numbers("P")
        if P is None:
            return self._skip('Points without "P"')
        n = len(P) // 3
        if n == 0:
            return
        widths = p.numbers("width")
        cw = p.number("constantwidth", 0.05)
        w = np.array([widths[i] if widths is not None and i < len(widths) else cw for i in range(n)], float)
        ends = np.hstack([np.asarray(P[:3 * n], float).reshape(n, 3), (w * 0.5)[:, None]]).astype(np.float32)
        half = np.full(n, 0.5, np.float32)
        self._curve_set(ends, ends.copy(), half, half.copy())

    def _polygon(self, vals):
        """scene_builder.rs:911-949: fan triangles through the current
        transform; under an AreaLightSource the material emits and a quad
        (4 vertices) also becomes a rect light."""
        pts = _Params(vals, 0).numbers("P")
        if pts is None or len(pts) < 9:
            return
        if self.defining is not None:
            return self._skip("Polygon inside ObjectBegin is not instanced")
        verts = [_pt(self.xf, pts[3 * i:3 * i + 3]) for i in range(len(pts) // 3)]
        mat = self._material()
        al = self.state.area_light
        if al is not None:
            inten, col = al
            m = self.materials[mat]
            m.emission = scenes._d3(tuple(c * inten for c in col))
            if len(pts) == 12:
                c = np.array(verts[0])
                m.area_light = len(self.lights)
                self.lights.append(scenes.light(rrhip.LIGHT_RECT, intensity=inten, color=col, p=verts[0],
                                                e1=tuple(np.array(verts[1]) - c), e2=tuple(np.array(verts[3]) - c)))
        for i in range(1, len(verts) - 1):
            self.tris.append((verts[0], verts[i], verts[i + 1], mat))

    def _points_polygons(self, vals):
        """points_polygons + add_tessellated_mesh (scene_builder.rs:982-1051,
        1195-1233): fan-triangulated faces, per-vertex normals when "N" has
        one per point; an instance of the mesh at the current transform, or
        an entry of the object being defined."""
        nverts = _as_numbers(vals[0]) if len(vals) > 0 else None
        idx = _as_numbers(vals[1]) if len(vals) > 1 else None
        if nverts is None or idx is None:
            return
        p = _Params(vals, 2)
        P = p.numbers("P")
        if P is None:
            return self._skip('PointsPolygons without "P"')
        N = p.numbers("N")
        if N is not None and len(N) != len(P):
            N = None
        ST = p.numbers("st")
        if ST is not None and len(ST) * 3 != len(P) * 2:  # one (s, t) per point, else none
            ST = None
        P1 = self.pending_deform_p
        if P1 is not None and len(P1) != len(P):  # an endpoint of another vertex count is dropped
            P1 = None
        tri = []
        cur = 0
        for nv in nverts:
            nv = int(nv)
            if cur + nv > len(idx):
                break
            poly = idx[cur:cur + nv]
            for i in range(1, nv - 1):
                tri += [poly[0], poly[i], poly[i + 1]]
            cur += nv
        if not tri:
            return
        self._add_mesh(np.array(P, np.float32).reshape(-1, 3), np.array(tri, np.uint32).reshape(-1, 3),
                       None if N is None else np.array(N, np.float32).reshape(-1, 3),
                       None if ST is None else np.array(ST, np.float32).reshape(-1, 2),
                       None if P1 is None else np.array(P1, np.float32).reshape(-1, 3))

    def _general_polygon(self, vals):
        """general_polygon (scene_builder.rs:1413-1446): GeneralPolygon
        [nverts per loop] "P", the first loop the outline and the rest
        holes, triangulated with its holes, smooth normals, no st."""
        nloops = _as_numbers(vals[0]) if len(vals) > 0 else None
        if nloops is None:
            return
        P = _Params(vals, 1).numbers("P")
        if P is None:
            return self._warn('GeneralPolygon without "P"; skipping')
        pos = np.asarray(P, np.float64)[:len(P) // 3 * 3].reshape(-1, 3)
        loops, cur = [], 0
        for nv in nloops:
            nv = int(nv)
            if cur + nv > len(pos):
                break
            loops.append(list(range(cur, cur + nv)))
            cur += nv
        if not loops or len(loops[0]) < 3:
            return
        tri = tessellate.triangulate_loops(pos, loops)
        if tri is None or len(tri) == 0:
            return self._warn("GeneralPolygon: triangulation failed; skipping")
        p32 = pos.astype(np.float32)
        self._add_mesh(p32, tri, tessellate.smooth_normals(p32, tri), None)

    def _points_general_polygons(self, vals):
        """PointsGeneralPolygons [nloops] [nverts] [vertices] params: a
        polygon of one loop is fanned as PointsPolygons fans it (the
        reference's handling, scene_builder.rs:796-810); a polygon with holes
        is triangulated with them, where render.rs instead fills each loop
        as a polygon of its own (with a warning). "P", "N", "st" and a
        deformation endpoint are per point, as for PointsPolygons."""
        nloops = _as_numbers(vals[0]) if len(vals) > 0 else None
        nverts = _as_numbers(vals[1]) if len(vals) > 1 else None
        idx = _as_numbers(vals[2]) if len(vals) > 2 else None
        if nloops is None or nverts is None or idx is None:
            return
        P = _Params(vals, 3).numbers("P")
        if P is None:
            return self._skip('PointsGeneralPolygons without "P"')
        pos = np.asarray(P, np.float64)[:len(P) // 3 * 3].reshape(-1, 3)
        npts = len(pos)
        polys, li, cur = [], 0, 0
        for nl in nloops:
            loops = []
            for _ in range(int(nl)):
                if li >= len(nverts):
                    break
                nv = int(nverts[li])
                li += 1
                if cur + nv > len(idx):
                    break
                loops.append([int(i) for i in idx[cur:cur + nv]])
                cur += nv
            if loops:
                polys.append(loops)
        if any(v < 0 or v >= npts for loops in polys for lp in loops for v in lp):
            return self._warn("PointsGeneralPolygons: vertex index out of range; skipping")
        tri = []
        for loops in polys:
            if len(loops[0]) < 3:
                continue
            if len(loops) == 1:
                lp = loops[0]
                tri += [(lp[0], lp[i], lp[i + 1]) for i in range(1, len(lp) - 1)]
            else:
                t = tessellate.triangulate_loops(pos, loops)
                if t is None:
                    self._warn("PointsGeneralPolygons: triangulation failed; polygon skipped")
                    continue
                tri += [tuple(int(x) for x in r) for r in t]
        if not tri:
            return
        p = _Params(vals, 3)
        N, ST = p.numbers("N"), p.numbers("st")
        if N is not None and len(N) != len(P):
            N = None
        if ST is not None and len(ST) * 3 != len(P) * 2:
            ST = None
        P1 = self.pending_deform_p
        if P1 is not None and len(P1) != len(P):
            P1 = None
        self._add_mesh(pos.astype(np.float32), np.asarray(tri, np.uint32).reshape(-1, 3),
                       None if N is None else np.array(N, np.float32).reshape(-1, 3),
                       None if ST is None else np.array(ST, np.float32).reshape(-1, 2),
                       None if P1 is None else np.array(P1, np.float32).reshape(-1, 3))

    def _add_mesh(self, pos, tri, nrm, st, pos1=None):
        """add_tessellated_mesh (scene_builder.rs:1195-1233): the mesh, and an
        instance of it at the current transform with the current material
        (emitting under an AreaLightSource), or an entry of the object
        being defined."""
        if self.state.displace is not None:  # displace_mesh (displace.rs:101-123)
            dp = self.state.displace
            base = nrm if nrm is not None else tessellate.smooth_normals(pos, tri)
            d = rrhip.gen_fbm(pos.astype(np.float64), dp["frequency"], dp["octaves"], dp["gain"], dp["lacunarity"],
                              dp["offset"]) * dp["amplitude"]
            pos = (pos.astype(np.float64) + base.astype(np.float64) * d[:, None]).astype(np.float32)
            nrm = tessellate.smooth_normals(pos, tri)
            pos1 = None  # the displaced mesh is rebuilt without a deformation endpoint
        mesh = len(self.meshes)
        self.meshes.append((pos, tri, nrm, st, pos1))
        mat = self._material()
        if self.state.area_light is not None:
            inten, col = self.state.area_light
            self.materials[mat].emission = scenes._d3(tuple(c * inten for c in col))
        if self.defining is not None:
            self.defining[1].append((mesh, self.xf.copy(), mat))
        else:
            m1 = self._endpoint(self.xf)
            self.instances.append((mesh, mat, self.xf[:3, :4].copy(), None if m1 is None else m1[:3, :4].copy()))

    # -- patches
    def _dice_segments(self):
        """dice_segments (scene_builder.rs:1173-1175): 16 / ShadingRate
        spans per patch, 2..64."""
        return int(min(max(16.0 / max(self.state.shading_rate, 0.05), 2.0), 64.0))

    def _patch_mesh(self, name, vals):
        """patch_mesh (scene_builder.rs:1324-1361): PatchMesh "type" nu uwrap
        nv vwrap params, or Patch "type" params (one 2x2 / 4x4 grid);
        bilinear or bicubic under the current Basis, diced by ShadingRate."""
        ptype = _str(vals, 0)
        if ptype is None:
            return
        if name == "Patch":
            nu = nv = 2 if ptype == "bilinear" else 4
            uw = vw = False
            start = 1
        else:
            nu, nv = _num(vals, 1), _num(vals, 3)
            if nu is None or nv is None:
                return
            nu, nv = int(nu), int(nv)
            uw, vw = _str(vals, 2) == "periodic", _str(vals, 4) == "periodic"
            start = 5
        P = _Params(vals, start).numbers("P")
        if P is None:
            return self._warn(f'{name} without "P"; skipping')
        segs = self._dice_segments()
        if ptype == "bilinear":
            m = tessellate.tessellate_bilinear(P, nu, nv, uw, vw, segs)
        elif ptype == "bicubic":
            (bu, us), (bv, vs) = self.state.basis_u, self.state.basis_v
            m = tessellate.tessellate_bicubic(P, nu, nv, uw, vw, bu, us, bv, vs, segs)
        else:
            return self._warn(f'{name} type "{ptype}" not supported')
        if m is None:
            return self._warn(f"{name}: invalid control grid; skipping")
        self._add_mesh(*m)

    def _nu_patch(self, vals):
        """nu_patch (scene_builder.rs:1363-1410): NuPatch nu uorder [uknot]
        umin umax nv vorder [vknot] vmin vmax with "P" or homogeneous "Pw";
        segments 16 / ShadingRate per knot span, 4..128; trim curves
        ignored."""
        nu, uo, umin, umax = _num(vals, 0), _num(vals, 1), _num(vals, 3), _num(vals, 4)
        nv, vo, vmin, vmax = _num(vals, 5), _num(vals, 6), _num(vals, 8), _num(vals, 9)
        uk = _as_numbers(vals[2]) if len(vals) > 2 else None
        vk = _as_numbers(vals[7]) if len(vals) > 7 else None
        if None in (nu, uo, umin, umax, nv, vo, vmin, vmax) or uk is None or vk is None:
            return
        p = _Params(vals, 10)
        pts, rational = (p.numbers("Pw"), True) if p.numbers("Pw") is not None else (p.numbers("P"), False)
        if pts is None:
            return self._warn('NuPatch without "P"/"Pw"; skipping')
        if p.get("trimcurve") is not None:
            self._warn("NuPatch trim curves not supported yet; rendering untrimmed")
        base = self._dice_segments()
        nu, uo, nv, vo = int(nu), int(uo), int(nv), int(vo)
        su = min(max(base * (max(nu - uo, 0) + 1), 4), 128)
        sv = min(max(base * (max(nv - vo, 0) + 1), 4), 128)
        m = tessellate.tessellate_nurbs(nu, uo, uk, umin, umax, nv, vo, vk, vmin, vmax, pts, rational, su, sv)
        if m is None:
            return self._warn("NuPatch: invalid control data; skipping")
        self._add_mesh(*m)

    def _displace(self, vals):
        """Displace (scene_builder.rs:776-795): the built-in "noise"
        displacement with DisplaceParams' defaults (amplitude 0.1, frequency
        1, 4 octaves, gain 0.5, lacunarity 2, offset 0) for the meshes diced
        under it."""
        name = _str(vals, 0) or ""
        if name != "noise":
            self._warn(f'Displace "{name}": only the built-in "noise" displacement exists; using it')
        p = _Params(vals, _pstart(vals))
        off = p.numbers("offset")
        self.state.displace = dict(amplitude=p.number("amplitude", 0.1), frequency=p.number("frequency", 1.0),
                                   octaves=int(p.number("octaves", 4)), gain=p.number("gain", 0.5),
                                   lacunarity=p.number("lacunarity", 2.0),
                                   offset=tuple(off[:3]) if off is not None and len(off) >= 3 else (0.0, 0.0, 0.0))

    def _subdiv_levels(self):
        """subdiv_levels (scene_builder.rs:1178-1192): uniform depth from
        ShadingRate."""
        sr = self.state.shading_rate
        return 5 if sr <= 0.15 else 4 if sr <= 0.35 else 3 if sr <= 1.5 else 2 if sr <= 6.0 else 1

    def _subdivision_mesh(self, vals):
        """subdivision_mesh (scene_builder.rs:1236-1322): SubdivisionMesh
        "scheme" [nverts] [verts] (["tags"] [nargs] [ints] [floats])? params,
        uniform Catmull-Clark with crease / corner / hole /
        interpolateboundary tags (sharpness 10 when a tag has none)."""
        scheme = _str(vals, 0)
        if scheme is None:
            return
        if scheme != "catmull-clark":
            self._warn(f'SubdivisionMesh scheme "{scheme}" treated as catmull-clark')
        nverts = _as_numbers(vals[1]) if len(vals) > 1 else None
        verts = _as_numbers(vals[2]) if len(vals) > 2 else None
        if nverts is None or verts is None:
            return
        creases, corners, holes, interp = [], [], [], False
        start = 3
        if len(vals) > 3 and isinstance(vals[3], list):  # the tag block
            tags = vals[3]
            nargs = (_as_numbers(vals[4]) if len(vals) > 4 else None) or ()
            ints = (_as_numbers(vals[5]) if len(vals) > 5 else None) or ()
            floats = (_as_numbers(vals[6]) if len(vals) > 6 else None) or ()
            ic = fc = 0
            for ti, tag in enumerate(tags):
                ni = int(nargs[2 * ti]) if 2 * ti < len(nargs) else 0
                nf = int(nargs[2 * ti + 1]) if 2 * ti + 1 < len(nargs) else 0
                ia = [int(x) for x in ints[min(ic, len(ints)):min(ic + ni, len(ints))]]
                fa = list(floats[min(fc, len(floats)):min(fc + nf, len(floats))])
                if tag == "crease":
                    sharp = fa[0] if fa else 10.0
                    creases += [(a, b, sharp) for a, b in zip(ia, ia[1:])]
                elif tag == "corner":
                    corners += [(v, fa[i] if i < len(fa) else (fa[0] if fa else 10.0)) for i, v in enumerate(ia)]
                elif tag == "hole":
                    holes += ia
                elif tag == "interpolateboundary":
                    interp = True
                else:
                    self._warn(f'SubdivisionMesh tag "{tag}" not supported; ignoring')
                ic += ni
                fc += nf
            start = 7
        P = _Params(vals, start).numbers("P")
        if P is None:
            return self._warn('SubdivisionMesh without "P"; skipping')
        faces, cur = [], 0
        for nv in nverts:
            nv = int(nv)
            if cur + nv > len(verts):
                break
            faces.append([int(x) for x in verts[cur:cur + nv]])
            cur += nv
        m = tessellate.tessellate_subdiv(np.asarray(P, np.float64).reshape(-1, 3), faces, creases, corners, holes,
                                         interp, self._subdiv_levels())
        self._add_mesh(*m)

    # -- patterns
    def _bxdf_bindings(self, p):
        """PxrSurface "reference <type> <param>" ["handle:output"]
        connections (scene_builder.rs:705-728) as (field, node) pairs."""
        out = []
        for tok, v in p.items:
            words = tok.split()
            if "reference" not in words:
                continue
            param, target = words[-1], _as_str(v)
            if target is None:
                continue
            handle = target.split(":")[0]
            if param not in rrhip.BIND_FIELDS:
                self._warn(f'reference to unsupported Bxdf param "{param}"; ignoring')
                continue
            if handle not in self.pattern_handles:
                self._warn(f'Bxdf references unknown Pattern "{handle}"; ignoring')
                continue
            out.append((param, self.pattern_handles[handle]))
        return out

    def _pattern_input(self, p, name, default):
        """pattern_input (scene_builder.rs:1740-1758): a referenced node, else
        the constant (a colour, or one number as grey), else the default."""
        for tok, v in p.items:
            words = tok.split()
            if "reference" in words and words[-1] == name:
                target = _as_str(v)
                if target is not None and target.split(":")[0] in self.pattern_handles:
                    return self.pattern_handles[target.split(":")[0]], (0.0, 0.0, 0.0)
        v = p.numbers(name)
        if v is not None and len(v) >= 3:
            return -1, tuple(v[:3])
        if v:
            return -1, (v[0], v[0], v[0])
        return -1, default

    def _open_texture(self, filename):
        """open_texture (scene_builder.rs:1703-1736): a .tex file, or the
        UDIM tiles 1001..1100 that exist for a "<UDIM>" name; (texture,
        udim dict), texture -1 when none opens (renders magenta)."""
        if not filename:
            self._warn('Pattern texture without "filename"; rendering magenta')
            return -1, None

        def open_one(name):
            path = name if self.base_dir is None else os.path.join(self.base_dir, name)
            if path in self.tex_index:
                return self.tex_index[path]
            if not os.path.exists(path):
                return None
            try:
                level0 = rtex.read_tex(path)
            except (OSError, rtex.TexError) as e:
                self._warn(f"texture {path}: {e}")
                return None
            self.tex_index[path] = len(self.textures)
            self.textures.append(level0)
            return self.tex_index[path]

        if "<UDIM>" in filename:
            tiles = {}
            for tile in range(1001, 1101):
                t = open_one(filename.replace("<UDIM>", str(tile)))
                if t is not None:
                    tiles[tile] = t
            if not tiles:
                self._warn(f'no UDIM tiles found for "{filename}"')
                return -1, None
            return -1, tiles
        t = open_one(filename)
        if t is None:
            self._warn(f"texture {filename}: cannot open; rendering magenta")
            return -1, None
        return t, None

    def _pattern(self, vals):
        """pattern_request (scene_builder.rs:1760-1848): one graph node under
        its handle."""
        ptype = _str(vals, 0) or ""
        start = 2 if len(vals) % 2 == 0 else 1
        handle = _str(vals, 1) if start == 2 else None
        if handle is None:
            return self._warn("Pattern without a handle; skipping")
        p = _Params(vals, start)
        wrap = {"clamp": rrhip.WRAP_CLAMP, "black": rrhip.WRAP_BLACK}.get(p.string("wrapMode") or "periodic",
                                                                          rrhip.WRAP_PERIODIC)
        if ptype in ("PxrTexture", "texture"):
            tex, udim = self._open_texture(p.string("filename") or "")
            kind, kw = rrhip.PATTERN_TEXTURE, dict(texture=tex, udim=udim, wrap=wrap,
                                                    scale=(p.number("scaleS", 1.0), p.number("scaleT", 1.0)))
        elif ptype in ("PxrChecker", "checker"):
            kind, kw = rrhip.PATTERN_CHECKER, dict(color_a=p.color("colorA", (1.0, 1.0, 1.0)),
                                                    color_b=p.color("colorB", (0.0, 0.0, 0.0)),
                                                    scale=(p.number("sScale", 8.0), p.number("tScale", 8.0)))
        elif ptype in ("PxrFractal", "fractal", "noise"):
            kind, kw = rrhip.PATTERN_FRACTAL, dict(frequency=p.number("frequency", 1.0),
                                                    octaves=int(p.number("layers", 4.0)), gain=p.number("gain", 0.5),
                                                    lacunarity=p.number("lacunarity", 2.0))
        elif ptype in ("PxrMix", "mix"):
            ins = [self._pattern_input(p, n, d) for n, d in (("color1", (0.0, 0.0, 0.0)), ("color2", (1.0, 1.0, 1.0)),
                                                             ("mix", (0.0, 0.0, 0.0)))]
            kind, kw = rrhip.PATTERN_MIX, dict(input=[i for i, _ in ins], constant=[c for _, c in ins])
        elif ptype in ("PxrColorCorrect", "colorCorrect"):
            i, c = self._pattern_input(p, "inputRGB", (0.5, 0.5, 0.5))
            kind, kw = rrhip.PATTERN_COLOR_CORRECT, dict(input=(i, -1, -1), constant=[c], cc_gain=p.color(
                "gain", (1.0, 1.0, 1.0)), cc_offset=p.color("offset", (0.0, 0.0, 0.0)), gamma=p.number("gamma", 1.0),
                saturation=p.number("saturation", 1.0))
        elif ptype in ("PxrRamp", "ramp"):
            pos = p.numbers("positions") or [0.0, 1.0]
            col = p.numbers("colors") or [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
            cols = [tuple(col[3 * k:3 * k + 3]) for k in range(len(col) // 3)]
            n = min(len(pos), len(cols))
            kind, kw = rrhip.PATTERN_RAMP, dict(stops=list(zip(pos[:n], cols[:n])) or None,
                                                 use_t=1 if p.string("axis") == "t" else 0)
        elif ptype == "triplanar":
            tex, udim = self._open_texture(p.string("filename") or "")
            kind, kw = rrhip.PATTERN_TRIPLANAR, dict(texture=tex, udim=udim, wrap=wrap,
                                                      frequency=p.number("frequency", 1.0))
        else:
            return self._warn(f'Pattern "{ptype}" not implemented; skipping')
        self.pattern_handles[handle] = len(self.patterns)
        self.patterns.append((kind, kw))

    # -- result
    @property
    def spp(self):
        return self.pixel_samples[0] * self.pixel_samples[1]

    def scene(self):
        """The scene (SceneBuilder::build, scene_builder.rs:208-244): the
        camera at the origin of the baked frame looking down +z."""
        s = scenes.Scene(self.width, self.height, self.fov, eye=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0),
                         filter=self.filter[0], filter_width=self.filter[1], background=self.background)
        s.camera.projection = self.projection
        if self.dof is not None:
            fstop, fl, fd = self.dof
            if math.isfinite(fstop) and fstop > 0.0:
                s.camera.lens_radius = fl / (2.0 * fstop)
                s.camera.focal_distance = fd
        if self.screen_window is not None:
            l, r, b, t = self.screen_window
            s.camera.ortho_half[0] = abs(r - l) * 0.5
            s.camera.ortho_half[1] = abs(t - b) * 0.5
        if self.camera_motion_inv is not None:
            for k, v in enumerate(self.camera_motion_inv[:3, :4].reshape(-1)):
                s.camera.motion_inv[k] = float(v)
            s.camera.has_motion_inv = 1
        s.materials = list(self.materials)
        s.lights = list(self.lights)
        for v0, v1, v2, mat in self.tris:
            s.add_triangle(v0, v1, v2, mat)
        for kind, mat, m, kw in self.quadrics:
            s.add_quadric(kind, mat, m, **kw)
        for pos, idx, nrm, st, pos1 in self.meshes:
            s.add_mesh(pos, idx, nrm, st, pos1)
        for mesh, mat, m, m1 in self.instances:
            s.add_instance(mesh, mat, m, m1)
        for p0, p1, v0, v1 in self.curve_sets:
            s.add_curves(p0, p1, v0, v1)
        for cs, mat, m, m1 in self.curve_instances:
            s.add_instance(cs, mat, m, m1, curves=True)
        for md in self.media:
            s.add_medium(**md)
        s.atmosphere = self.atmosphere
        for tex in self.textures:
            s.add_texture(tex)
        for kind, kw in self.patterns:
            s.add_pattern(kind, **kw)
        for mat, field, node in self.bindings:
            s.bind(mat, field, node)
        if self.env is not None:
            s.set_env(self.env)
        return s


def load(path, strict=True):
    """Parses and builds a RIB file; returns (scene, spp) with spp the
    PixelSamples product. Relative resource paths (environment maps)
    resolve against the file's directory."""
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        text = f.read()
    b = Builder(base_dir=os.path.dirname(os.path.abspath(path)), strict=strict).run(parse(text))
    return b.scene(), b.spp


def loads(text, base_dir=None, strict=True):
    """`load` for RIB text."""
    b = Builder(base_dir=base_dir, strict=strict).run(parse(text))
    return b.scene(), b.spp
