"""pbrt-v3 scene files to the flat scenes of the GPU path tracer.

Two layers, as in the reference's core/parser:

* ``parse_to_ast(text)`` / ``parse_file(path)`` -- tokens (numbers, quoted
  strings, bare identifiers, brackets, ``#`` comments) to a list of
  :class:`Directive` with typed :class:`Param` lists. Blocks become one
  directive holding its children (``WorldBlock``, ``AttributeBlock``,
  ``TransformBlock``, ``ObjectBlock``), and ``Include`` is replaced by the
  directives of the included file. Follows parse_to_ast
  (core/parser/mod.rs:19-37), the directive parsers
  (parser/parsers/directives.rs), the parameter parsers
  (parsers/parameters.rs:20-349: integer, float, point2, vector2,
  point3|point, vector3|vector, normal3|normal, rgb|color, xyz, spectrum,
  blackbody, spectrum file, string, bool, texture), the transforms
  (parsers/transforms.rs) and the AST (core/parser/ast.rs:397-560).
* ``load(ast)`` -- walks the AST with pbrt-v3's graphics state (transform
  stack, material, area light, reverse orientation, named materials and
  textures, object instances) and returns the flat scene, the camera and the
  render settings. The reference stops at the AST (its ``parse`` that would
  set the state is a TODO, parser/mod.rs:39-46), so the loader follows
  pbrt-v3's scene API (api.cpp) for the subset the GPU path tracer renders.

Loaded: LookAt + perspective camera (fov), image film (resolution,
filename), box/triangle/gaussian/mitchell/sinc pixel filters, the halton,
sobol and random samplers (other samplers render with Halton) and
pixelsamples, the path integrator's maxdepth,
Identity/Translate/Scale/Rotate/Transform/ConcatTransform,
CoordinateSystem/CoordSysTransform, ReverseOrientation, matte/metal/glass
materials (constant textures allowed), named materials, diffuse area lights on
triangle meshes, trianglemesh, plymesh and sphere shapes (spheres under
translation and uniform scale), object instancing, and one extension for the
renderer's analytic sky: ``LightSource "sky" "rgb zenith" [...] "rgb horizon"
[...]``. Everything else raises :class:`PbrtError` naming the file, line and
directive.

Differences from the reference parser, all toward pbrt-v3: ``Include`` paths
are relative to the including file (the reference changes the process working
directory, mod.rs:22); ``Transform``/``ConcatTransform`` also take a bracketed
matrix; ``MakeNamedMaterial``/``NamedMaterial``, ``Identity`` and bracketed
string/bool values are accepted.
"""
import math
import os
import re

import numpy as np

from . import scenes as S


class PbrtError(ValueError):
    pass


class Param:
    """One typed parameter: ``type`` is the canonical type name (``point3``
    for "point", ``rgb`` for "color", ...); ``values`` is a flat list."""
    __slots__ = ("type", "name", "values")

    def __init__(self, type_, name, values):
        self.type, self.name, self.values = type_, name, values

    def __eq__(self, other):
        return isinstance(other, Param) and (self.type, self.name, self.values) == (other.type, other.name,
                                                                                   other.values)

    def __repr__(self):
        return "Param(%r, %r, %r)" % (self.type, self.name, self.values)


class Directive:
    """``name`` is the directive keyword, or WorldBlock / AttributeBlock /
    TransformBlock / ObjectBlock for a block (``children`` then holds its
    directives). ``args`` are the positional operands (numbers, strings, a
    16-number list for Transform/ConcatTransform)."""
    __slots__ = ("name", "args", "params", "children", "line", "file")

    def __init__(self, name, args=(), params=(), children=None, line=0, file="<string>"):
        self.name, self.args, self.params, self.children = name, list(args), list(params), children
        self.line, self.file = line, file

    def param(self, name):
        for p in self.params:
            if p.name == name:
                return p
        return None

    def where(self):
        return "%s:%d: %s" % (self.file, self.line, self.name)

    def __repr__(self):
        kids = "" if self.children is None else ", %d children" % len(self.children)
        return "Directive(%r, %r, %r%s)" % (self.name, self.args, self.params, kids)


PARAM_TYPES = {"integer": "integer", "float": "float", "point2": "point2", "vector2": "vector2",
               "point3": "point3", "point": "point3", "vector3": "vector3", "vector": "vector3",
               "normal3": "normal3", "normal": "normal3", "rgb": "rgb", "color": "rgb", "xyz": "xyz",
               "spectrum": "spectrum", "blackbody": "blackbody", "string": "string", "bool": "bool",
               "texture": "texture"}
# numbers per value group; a bracketed list must hold whole groups
_GROUP = {"integer": 1, "float": 1, "point2": 2, "vector2": 2, "point3": 3, "vector3": 3, "normal3": 3,
          "rgb": 3, "xyz": 3, "blackbody": 2, "spectrum": 2}
_NEEDS_BRACKETS = {"point2", "vector2", "point3", "vector3", "normal3", "rgb", "xyz", "blackbody"}

_TOKEN = re.compile(r'(?P<ws>[ \t\r\n]+)|(?P<comment>\#[^\n]*)|(?P<str>"[^"]*")|(?P<lb>\[)|(?P<rb>\])|'
                    r'(?P<num>[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)(?![A-Za-z_\d.])|'
                    r'(?P<id>[A-Za-z_][A-Za-z0-9_]*)')


def tokenize(text, filename="<string>"):
    """-> list of (kind, value, line, raw); kind is num, str, id, [ or ]."""
    out, pos, line = [], 0, 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise PbrtError("%s:%d: unexpected character %r" % (filename, line, text[pos]))
        k = m.lastgroup
        if k == "str":
            out.append(("str", m.group(k)[1:-1], line, m.group(k)))
        elif k == "lb":
            out.append(("[", "[", line, "["))
        elif k == "rb":
            out.append(("]", "]", line, "]"))
        elif k == "num":
            out.append(("num", float(m.group(k)), line, m.group(k)))
        elif k == "id":
            out.append(("id", m.group(k), line, m.group(k)))
        line += text.count("\n", pos, m.end())
        pos = m.end()
    return out


# operand shapes: n = number, s = quoted string, i = bare identifier, m = 16 numbers, p = parameter list
_OPERANDS = {"Identity": "", "Translate": "nnn", "Scale": "nnn", "Rotate": "nnnn", "LookAt": "nnnnnnnnn",
             "CoordinateSystem": "s", "CoordSysTransform": "s", "Transform": "m", "ConcatTransform": "m",
             "ActiveTransform": "i", "TransformTimes": "nn", "ReverseOrientation": "",
             "Camera": "sp", "Sampler": "sp", "Film": "sp", "PixelFilter": "sp", "Integrator": "sp",
             "Accelerator": "sp", "Shape": "sp", "LightSource": "sp", "AreaLightSource": "sp", "Material": "sp",
             "MakeNamedMaterial": "sp", "NamedMaterial": "s", "Texture": "sssp", "MakeNamedMedium": "sp",
             "MediumInterface": "ss", "ObjectInstance": "s", "Include": "s"}
_BLOCKS = {"WorldBegin": ("WorldEnd", "WorldBlock", False), "AttributeBegin": ("AttributeEnd", "AttributeBlock", False),
           "TransformBegin": ("TransformEnd", "TransformBlock", False),
           "ObjectBegin": ("ObjectEnd", "ObjectBlock", True)}
_ACTIVE = ("StartTime", "EndTime", "All")


class _Parser:
    def __init__(self, tokens, filename, base_dir, depth):
        self.t, self.i, self.file, self.base, self.depth = tokens, 0, filename, base_dir, depth

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def line(self):
        tok = self.peek()
        return tok[2] if tok else (self.t[-1][2] if self.t else 1)

    def fail(self, what):
        raise PbrtError("%s:%d: %s" % (self.file, self.line(), what))

    def take(self, kind, what):
        tok = self.peek()
        if tok is None or tok[0] != kind:
            self.fail("expected %s, found %s" % (what, "end of input" if tok is None else repr(tok[3])))
        self.i += 1
        return tok[1]

    def params(self):
        out = []
        while True:
            tok = self.peek()
            if tok is None or tok[0] != "str":
                return out
            decl = tok[1].split()
            if len(decl) != 2 or decl[0] not in PARAM_TYPES:
                return out  # a quoted string that declares no parameter ends the list
            self.i += 1
            ptype, name = PARAM_TYPES[decl[0]], decl[1]
            out.append(Param(ptype, name, self.param_values(ptype, name)))

    def param_values(self, ptype, name):
        tok = self.peek()
        if tok is None:
            self.fail("parameter %r has no value" % name)
        if tok[0] == "[":
            self.i += 1
            items = []
            while self.peek() is not None and self.peek()[0] not in ("]", "["):
                items.append(self.peek())
                self.i += 1
            self.take("]", "']'")
        else:
            if ptype in _NEEDS_BRACKETS:
                self.fail("parameter %r (%s) needs a bracketed list" % (name, ptype))
            if tok[0] not in ("num", "str"):
                self.fail("parameter %r has no value" % name)
            items = [tok]
            self.i += 1
        kinds = {k for k, _, _, _ in items}
        if ptype == "spectrum" and kinds == {"str"} and len(items) == 1:
            return [items[0][1]]  # spectrum file name
        if ptype in _GROUP:
            if kinds - {"num"}:
                self.fail("parameter %r (%s) needs numbers" % (name, ptype))
            vals = [v for _, v, _, _ in items]
            if ptype == "integer":
                if any(not re.fullmatch(r"[+-]?\d+", raw) for _, _, _, raw in items):
                    self.fail("parameter %r (integer) needs integers" % name)
                vals = [int(v) for v in vals]
            if len(vals) % _GROUP[ptype] or not vals:
                self.fail("parameter %r (%s) needs a multiple of %d values" % (name, ptype, _GROUP[ptype]))
            return vals
        if ptype == "bool":
            if len(items) != 1 or items[0][0] != "str" or items[0][1] not in ("true", "false"):
                self.fail("parameter %r (bool) needs \"true\" or \"false\"" % name)
            return [items[0][1] == "true"]
        if kinds != {"str"}:
            self.fail("parameter %r (%s) needs quoted strings" % (name, ptype))
        return [v for _, v, _, _ in items]

    def directive(self):
        kind, name, line, raw = self.peek()
        if kind != "id":
            self.fail("expected a directive, found %r" % raw)
        self.i += 1
        if name in _BLOCKS:
            end, block, named = _BLOCKS[name]
            args = [self.take("str", "a quoted object name")] if named else []
            children = []
            while True:
                tok = self.peek()
                if tok is None:
                    raise PbrtError("%s:%d: %s without %s" % (self.file, line, name, end))
                if tok[0] == "id" and tok[1] == end:
                    self.i += 1
                    break
                children.extend(self.directive())
            return [Directive(block, args, (), children, line, self.file)]
        if name not in _OPERANDS:
            raise PbrtError("%s:%d: unknown directive %r" % (self.file, line, name))
        args, params = [], []
        for c in _OPERANDS[name]:
            if c == "n":
                args.append(self.take("num", "a number"))
            elif c == "s":
                args.append(self.take("str", "a quoted string"))
            elif c == "i":
                v = self.take("id", "StartTime, EndTime or All")
                if v not in _ACTIVE:
                    raise PbrtError("%s:%d: ActiveTransform needs StartTime, EndTime or All" % (self.file, line))
                args.append(v)
            elif c == "m":
                br = self.peek() is not None and self.peek()[0] == "["
                if br:
                    self.i += 1
                args.append([self.take("num", "a number") for _ in range(16)])
                if br:
                    self.take("]", "']'")
            elif c == "p":
                params = self.params()
        if name == "Include":
            if self.depth > 32:
                raise PbrtError("%s:%d: Include nested too deeply" % (self.file, line))
            path = args[0] if os.path.isabs(args[0]) else os.path.join(self.base, args[0])
            return parse_file(path, self.depth + 1)
        return [Directive(name, args, params, None, line, self.file)]

    def all(self):
        out = []
        while self.peek() is not None:
            out.extend(self.directive())
        return out


def parse_to_ast(text, filename="<string>", base_dir=".", _depth=0):
    """Directives of a pbrt-v3 scene description; ``Include`` paths are
    resolved against ``base_dir``."""
    return _Parser(tokenize(text, filename), filename, base_dir, _depth).all()


def parse_file(path, _depth=0):
    """parse_to_ast (core/parser/mod.rs:19-37) of a file; includes are
    relative to the file's directory."""
    with open(path) as f:
        text = f.read()
    return parse_to_ast(text, path, os.path.dirname(os.path.abspath(path)), _depth)


# ---------------------------------------------------------------------------------- transforms
def translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scale(x, y, z):
    return np.diag([float(x), float(y), float(z), 1.0])


def rotate(theta, x, y, z):
    """Rotation by theta degrees about axis (x, y, z) (Rodrigues' formula, the
    matrix of pbrt-v3 Rotate, PBR book 3e 2.7.6)."""
    a = np.array([x, y, z], np.float64)
    n = np.linalg.norm(a)
    if n == 0:
        raise PbrtError("Rotate about a zero axis")
    a /= n
    s, c = math.sin(math.radians(theta)), math.cos(math.radians(theta))
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    m = np.eye(4)
    m[:3, :3] = c * np.eye(3) + (1 - c) * np.outer(a, a) + s * k
    return m


def look_at(pos, look, up):
    """World-to-camera matrix of pbrt-v3 LookAt (PBR book 3e 2.7.7): the
    camera looks down +z with +y up."""
    pos, look, up = (np.asarray(v, np.float64) for v in (pos, look, up))
    d = look - pos
    if np.linalg.norm(d) == 0:
        raise PbrtError("LookAt with the eye at the look-at point")
    d = d / np.linalg.norm(d)
    r = np.cross(up / np.linalg.norm(up), d)
    if np.linalg.norm(r) == 0:
        raise PbrtError("LookAt with the up vector along the viewing direction")
    r = r / np.linalg.norm(r)
    u = np.cross(d, r)
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = r, u, d, pos
    return np.linalg.inv(c2w)


def roughness_to_alpha(roughness):
    """TrowbridgeReitzDistribution::roughness_to_alpha
    (core/microfacet/trowbridge_reitz_distribution.rs:25-30), in float32."""
    f = np.float32
    x = f(np.log(max(f(roughness), f(1e-3))))
    return float(f(1.62142) + f(0.819955) * x + f(0.1734) * x * x + f(0.0171201) * x * x * x
                 + f(0.000640711) * x * x * x * x)


# --------------------------------------------------------------------------------------- PLY
_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
              "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
              "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def read_ply(path):
    """Vertex positions (float64, n x 3) and triangles (int64, m x 3) of a PLY
    file (ascii, binary_little_endian or binary_big_endian); quads are split
    into (0,1,2) and (0,2,3), as pbrt-v3's plymesh shape splits them."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise PbrtError("%s: not a PLY file" % path)
    nl = data.find(b"\n", end)
    header, body = data[:end].decode("ascii", "replace").split("\n"), data[nl + 1:]
    fmt, elements = None, []
    for ln in header:
        w = ln.split()
        if not w or w[0] in ("ply", "comment", "obj_info"):
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append((w[1], int(w[2]), []))
        elif w[0] == "property":
            if not elements:
                raise PbrtError("%s: property before element" % path)
            try:
                if w[1] == "list":
                    elements[-1][2].append((w[4], _PLY_TYPES[w[2]], _PLY_TYPES[w[3]]))
                else:
                    elements[-1][2].append((w[2], _PLY_TYPES[w[1]], None))
            except (KeyError, IndexError):
                raise PbrtError("%s: bad property line %r" % (path, ln))
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise PbrtError("%s: PLY format %r not supported" % (path, fmt))
    endian = ">" if fmt == "binary_big_endian" else "<"
    verts, faces = None, []
    words = body.split() if fmt == "ascii" else None
    pos = 0
    for name, count, props in elements:
        cols = {p[0]: [] for p in props}
        if fmt != "ascii" and all(p[2] is None for p in props):
            # fixed-size records: one structured read
            dt = np.dtype([(p[0], endian + p[1]) for p in props])
            rec = np.frombuffer(body, dt, count, pos)
            pos += count * dt.itemsize
            cols = {p[0]: rec[p[0]].astype(np.float64) for p in props}
        else:
            for _ in range(count):
                for pname, t, t2 in props:
                    if fmt == "ascii":
                        if t2 is not None:
                            k = int(words[pos])
                            cols[pname].append([float(v) for v in words[pos + 1:pos + 1 + k]])
                            pos += 1 + k
                        else:
                            cols[pname].append(float(words[pos]))
                            pos += 1
                    else:
                        dt = np.dtype(endian + t)
                        if t2 is not None:
                            k = int(np.frombuffer(body, dt, 1, pos)[0])
                            pos += dt.itemsize
                            dt2 = np.dtype(endian + t2)
                            cols[pname].append(np.frombuffer(body, dt2, k, pos).tolist())
                            pos += k * dt2.itemsize
                        else:
                            cols[pname].append(float(np.frombuffer(body, dt, 1, pos)[0]))
                            pos += dt.itemsize
        if name == "vertex":
            verts = np.stack([np.asarray(cols[c], np.float64) for c in ("x", "y", "z")], 1).reshape(-1, 3)
        elif name == "face":
            key = "vertex_indices" if "vertex_indices" in cols else "vertex_index"
            for vi in cols[key]:
                vi = [int(v) for v in vi]
                if len(vi) == 3:
                    faces.append(vi)
                elif len(vi) == 4:
                    faces.append([vi[0], vi[1], vi[2]])
                    faces.append([vi[0], vi[2], vi[3]])
                else:
                    raise PbrtError("%s: faces of %d vertices are not supported" % (path, len(vi)))
    if verts is None:
        raise PbrtError("%s: no vertex element" % path)
    return verts, np.asarray(faces, np.int64).reshape(-1, 3)


def write_ply(path, verts, tris):
    """Binary little-endian PLY of a triangle mesh."""
    verts = np.ascontiguousarray(verts, "<f4").reshape(-1, 3)
    tris = np.ascontiguousarray(tris, "<i4").reshape(-1, 3)
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
            "property float z\nelement face %d\nproperty list uchar int vertex_indices\nend_header\n"
            % (len(verts), len(tris)))
    face = np.zeros(len(tris), np.dtype([("n", "u1"), ("i", "<i4", 3)]))
    face["n"], face["i"] = 3, tris
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(verts.tobytes())
        f.write(face.tobytes())


# ------------------------------------------------------------------------------------ loader
class Loaded:
    """Result of :func:`load`: ``desc`` (scenes.SceneDesc), ``camera``
    (scenes.camera record), ``spp``, ``max_depth``, the film ``filename``, the
    pixel ``filter`` (film.filter_desc record, None for pbrt-v3's default box),
    the ``sampler`` ("halton", pbrt-v3's default, "sobol" or "random") and the
    warnings for directives and parameters that were skipped."""

    def __init__(self):
        self.desc = S.SceneDesc()
        self.camera = None
        self.spp = 16
        self.max_depth = 5
        self.filename = None
        self.filter = None  # film.filter_desc record; None: box of radius 0.5
        self.sampler = "halton"  # pbrt-v3's default sampler
        self.warnings = []


class _GState:
    """pbrt-v3 GraphicsState: what AttributeBegin saves and AttributeEnd
    restores besides the transform."""

    def __init__(self):
        self.material = None
        self.area_L = None
        self.reverse = False
        self.named = {}
        self.textures = {}

    def copy(self):
        g = _GState()
        g.material, g.area_L, g.reverse = self.material, self.area_L, self.reverse
        g.named, g.textures = dict(self.named), dict(self.textures)
        return g


def _uniform_scale(m):
    """Scale factor of a 3x3 matrix that is a positive multiple of the
    identity, else None."""
    if np.array_equal(m, np.diag(np.diag(m))) and m[0, 0] == m[1, 1] == m[2, 2] and m[0, 0] > 0:
        return float(m[0, 0])
    return None


def _apply(ctm, p):
    if np.array_equal(ctm, np.eye(4)):
        return np.asarray(p, np.float32)
    return (np.c_[np.asarray(p, np.float64), np.ones(len(p))] @ ctm.T)[:, :3].astype(np.float32)


class _Loader:
    def __init__(self, base_dir):
        self.base_dir = base_dir
        self.out = Loaded()
        self.ctm = np.eye(4)
        self.gs = _GState()
        self.coords = {}
        self.objects = {}
        self.obj = None  # name of the object being defined
        self.lookat = None
        self.fov = 90.0
        self.res = (1280, 720)
        self.has_camera = False
        self.in_world = False
        self.prims = []  # ("mesh", verts, idx, material, emission) | ("sphere", (x, y, z, r), material)
        self.mat_index = {}

    # -- parameters
    def warn(self, d, what):
        self.out.warnings.append("%s: %s" % (d.where(), what))

    def get(self, d, name, types, used):
        used.add(name)
        p = d.param(name)
        if p is not None and p.type not in types:
            raise PbrtError("%s: parameter %r: type %s not supported here (%s)" % (d.where(), name, p.type,
                                                                                   "/".join(types)))
        return p

    def check_unused(self, d, used):
        for p in d.params:
            if p.name not in used:
                self.warn(d, "parameter %r ignored" % p.name)

    def one(self, d, name, default, used, kind=float):
        p = self.get(d, name, ("integer",) if kind is int else ("float",), used)
        if p is None:
            return default
        if len(p.values) != 1:
            raise PbrtError("%s: parameter %r needs one value" % (d.where(), name))
        return kind(p.values[0])

    def spectrum(self, d, name, default, used):
        p = self.get(d, name, ("rgb", "texture", "spectrum", "xyz", "blackbody"), used)
        if p is None:
            return tuple(float(v) for v in default)
        if p.type == "texture":
            t = self.gs.textures.get(("spectrum", p.values[0]))
            if t is None:
                raise PbrtError("%s: %r names no constant spectrum texture" % (d.where(), p.values[0]))
            return t
        if p.type != "rgb" or len(p.values) != 3:
            raise PbrtError("%s: parameter %r: only one rgb value is supported" % (d.where(), name))
        return tuple(float(v) for v in p.values)

    def scalar_tex(self, d, name, default, used):
        p = self.get(d, name, ("float", "texture"), used)
        if p is None:
            return default
        if p.type == "texture":
            t = self.gs.textures.get(("float", p.values[0]))
            if t is None:
                raise PbrtError("%s: %r names no constant float texture" % (d.where(), p.values[0]))
            return t
        if len(p.values) != 1:
            raise PbrtError("%s: parameter %r needs one value" % (d.where(), name))
        return float(p.values[0])

    def material(self, d, kind):
        """Material record of pbrt-v3 matte/metal/glass; identical records
        share one table entry."""
        used = {"type"}
        if self.get(d, "bumpmap", ("texture",), used) is not None:
            raise PbrtError("%s: bump maps are not supported" % d.where())
        if kind == "matte":
            m = S.matte(self.spectrum(d, "Kd", (0.5, 0.5, 0.5), used))
            if self.scalar_tex(d, "sigma", 0.0, used) != 0.0:
                raise PbrtError("%s: matte with sigma != 0 (Oren-Nayar) is not supported" % d.where())
        elif kind == "metal":
            r = self.scalar_tex(d, "roughness", 0.01, used)
            ur = self.scalar_tex(d, "uroughness", r, used)
            vr = self.scalar_tex(d, "vroughness", r, used)
            if ur != vr:
                raise PbrtError("%s: anisotropic metal roughness is not supported" % d.where())
            p = self.get(d, "remaproughness", ("bool",), used)
            remap = True if p is None else p.values[0]
            m = S.metal(roughness_to_alpha(ur) if remap else ur, eta=self.spectrum(d, "eta", S.COPPER_ETA, used),
                        k=self.spectrum(d, "k", S.COPPER_K, used))
        elif kind == "glass":
            ior = self.scalar_tex(d, "index", 1.5, used)
            if self.spectrum(d, "Kr", (1, 1, 1), used) != (1.0, 1.0, 1.0) or \
                    self.spectrum(d, "Kt", (1, 1, 1), used) != (1.0, 1.0, 1.0):
                raise PbrtError("%s: glass with Kr or Kt != 1 is not supported" % d.where())
            r = self.scalar_tex(d, "roughness", 0.0, used)
            if self.scalar_tex(d, "uroughness", r, used) != 0 or self.scalar_tex(d, "vroughness", r, used) != 0:
                raise PbrtError("%s: rough glass is not supported" % d.where())
            used.add("remaproughness")
            m = S.glass(ior)
        else:
            raise PbrtError("%s: material %r is not supported (matte, metal, glass)" % (d.where(), kind))
        self.check_unused(d, used)
        key = m.tobytes()
        if key not in self.mat_index:
            self.mat_index[key] = self.out.desc.add_material(m)
        return self.mat_index[key]

    def current_material(self):
        if self.gs.material is None:
            self.gs.material = self.material(Directive("Material", ["matte"]), "matte")
        return self.gs.material

    def sink(self):
        return self.objects[self.obj] if self.obj is not None else self.prims

    # -- shapes
    def add_mesh(self, d, p_obj, idx, normals):
        ctm = self.ctm
        w = _apply(ctm, p_obj)
        idx = np.asarray(idx, np.int64).reshape(-1, 3)
        if idx.size and (idx.min() < 0 or idx.max() >= len(w)):
            raise PbrtError("%s: vertex index out of range" % d.where())
        if normals is not None:
            # pbrt-v3 Triangle::Intersect: with per-vertex normals the geometric
            # normal is turned toward the shading normal
            nw = normals @ np.linalg.inv(ctm[:3, :3])
            a, b, c = (w[idx[:, k]].astype(np.float64) for k in range(3))
            ng = np.cross(a - c, b - c)
            ns = nw[idx[:, 0]] + nw[idx[:, 1]] + nw[idx[:, 2]]
            flip = np.einsum("ij,ij->i", ng, ns) < 0
        else:
            swaps = bool(np.linalg.det(ctm[:3, :3]) < 0)
            flip = np.full(len(idx), self.gs.reverse != swaps)
        idx = idx.copy()
        idx[flip] = idx[flip][:, [0, 2, 1]]
        if self.gs.area_L is not None and self.obj is not None:
            raise PbrtError("%s: area lights inside object instances are not supported" % d.where())
        self.sink().append(("mesh", w, idx, self.current_material(), self.gs.area_L))

    def shape(self, d):
        kind = d.args[0]
        used = set()
        if kind in ("trianglemesh", "plymesh"):
            normals = None
            if kind == "trianglemesh":
                P = self.get(d, "P", ("point3",), used)
                if P is None:
                    raise PbrtError("%s: trianglemesh needs \"point P\"" % d.where())
                p = np.asarray(P.values, np.float64).reshape(-1, 3)
                I = self.get(d, "indices", ("integer",), used)
                if I is None:
                    if len(p) != 3:
                        raise PbrtError("%s: trianglemesh needs \"integer indices\"" % d.where())
                    idx = np.array([[0, 1, 2]])
                else:
                    if len(I.values) % 3:
                        raise PbrtError("%s: index count is not a multiple of 3" % d.where())
                    idx = np.asarray(I.values, np.int64).reshape(-1, 3)
                N = self.get(d, "N", ("normal3",), used)
                if N is not None:
                    normals = np.asarray(N.values, np.float64).reshape(-1, 3)
                    if len(normals) != len(p):
                        raise PbrtError("%s: N needs one normal per vertex" % d.where())
                for nm in ("uv", "st", "S", "faceIndices"):
                    used.add(nm)
                    if d.param(nm) is not None:
                        self.warn(d, "parameter %r ignored (no textures or shading frames)" % nm)
            else:
                fn = self.get(d, "filename", ("string",), used)
                if fn is None:
                    raise PbrtError("%s: plymesh needs \"string filename\"" % d.where())
                path = fn.values[0]
                if not os.path.isabs(path):
                    base = self.base_dir if d.file == "<string>" else os.path.dirname(os.path.abspath(d.file))
                    path = os.path.join(base, path)
                p, idx = read_ply(path)
            for nm in ("alpha", "shadowalpha"):
                used.add(nm)
                if d.param(nm) is not None:
                    raise PbrtError("%s: alpha masks are not supported" % d.where())
            self.check_unused(d, used)
            self.add_mesh(d, p, idx, normals)
        elif kind == "sphere":
            r = self.one(d, "radius", 1.0, used)
            for nm in ("zmin", "zmax", "phimax"):
                used.add(nm)
                if d.param(nm) is not None:
                    raise PbrtError("%s: partial spheres are not supported" % d.where())
            self.check_unused(d, used)
            if self.gs.area_L is not None:
                raise PbrtError("%s: only triangle meshes may be area lights" % d.where())
            if self.gs.reverse:
                raise PbrtError("%s: inside-out spheres are not supported" % d.where())
            s = _uniform_scale(self.ctm[:3, :3])
            if s is None:
                raise PbrtError("%s: spheres only under translation and uniform scale" % d.where())
            self.sink().append(("sphere", (self.ctm[0, 3], self.ctm[1, 3], self.ctm[2, 3], r * s),
                                self.current_material()))
        else:
            raise PbrtError("%s: shape %r is not supported (trianglemesh, plymesh, sphere)" % (d.where(), kind))

    def instance(self, d):
        """pbrt-v3 ObjectInstance: the object's shapes (already under the
        transform current at their definition) under the current transform."""
        name = d.args[0]
        if name not in self.objects:
            raise PbrtError("%s: unknown object %r" % (d.where(), name))
        if self.obj is not None:
            raise PbrtError("%s: ObjectInstance inside ObjectBegin" % d.where())
        ctm = self.ctm
        for item in self.objects[name]:
            if item[0] == "mesh":
                _, w, idx, mat, em = item
                w = _apply(ctm, w)
                if np.linalg.det(ctm[:3, :3]) < 0:
                    idx = idx[:, [0, 2, 1]]
                self.prims.append(("mesh", w, idx, mat, em))
            else:
                _, (x, y, z, r), mat = item
                s = _uniform_scale(ctm[:3, :3])
                if s is None:
                    raise PbrtError("%s: sphere instances only under translation and uniform scale" % d.where())
                c = ctm @ np.array([x, y, z, 1.0])
                self.prims.append(("sphere", (c[0], c[1], c[2], r * s), mat))

    # -- directives
    def run(self, dirs):
        for d in dirs:
            self.directive(d)

    def directive(self, d):
        n = d.name
        if n == "Identity":
            self.ctm = np.eye(4)
        elif n == "Translate":
            self.ctm = self.ctm @ translate(*d.args)
        elif n == "Scale":
            self.ctm = self.ctm @ scale(*d.args)
        elif n == "Rotate":
            self.ctm = self.ctm @ rotate(*d.args)
        elif n == "LookAt":
            if self.in_world:
                raise PbrtError("%s: LookAt inside the world block is not supported" % d.where())
            if not np.array_equal(self.ctm, np.eye(4)):
                raise PbrtError("%s: the camera transform must be a single LookAt" % d.where())
            self.lookat = list(d.args)
            self.ctm = look_at(d.args[0:3], d.args[3:6], d.args[6:9])
        elif n in ("Transform", "ConcatTransform"):
            m = np.array(d.args[0], np.float64).reshape(4, 4).T  # files list the matrix column by column
            self.ctm = m if n == "Transform" else self.ctm @ m
        elif n == "CoordinateSystem":
            self.coords[d.args[0]] = self.ctm.copy()
        elif n == "CoordSysTransform":
            if d.args[0] not in self.coords:
                raise PbrtError("%s: unknown coordinate system %r" % (d.where(), d.args[0]))
            self.ctm = self.coords[d.args[0]].copy()
        elif n == "ReverseOrientation":
            self.gs.reverse = not self.gs.reverse
        elif n in ("ActiveTransform", "TransformTimes", "Accelerator", "MakeNamedMedium", "MediumInterface"):
            self.warn(d, "ignored")
        elif n == "PixelFilter":
            self.pixel_filter(d)
        elif n == "Camera":
            self.camera(d)
        elif n == "Film":
            used = {"filename"}
            self.res = (self.one(d, "xresolution", 1280, used, int), self.one(d, "yresolution", 720, used, int))
            if self.res[0] <= 0 or self.res[1] <= 0:
                raise PbrtError("%s: resolution must be positive" % d.where())
            p = self.get(d, "filename", ("string",), used)
            self.out.filename = p.values[0] if p is not None else None
            self.check_unused(d, used)
        elif n == "Sampler":
            used = set()
            kind = d.args[0]
            if kind in ("halton", "random", "sobol"):
                self.out.sampler = kind
            else:
                self.out.sampler = "halton"
                self.warn(d, "sampler %r rendered with the Halton sampler" % kind)
            self.out.spp = self.one(d, "pixelsamples", 16, used, int)
            if self.out.spp <= 0:
                raise PbrtError("%s: pixelsamples must be positive" % d.where())
            self.check_unused(d, used)
        elif n == "Integrator":
            if d.args[0] != "path":
                raise PbrtError("%s: integrator %r is not supported (path)" % (d.where(), d.args[0]))
            used = set()
            self.out.max_depth = self.one(d, "maxdepth", 5, used, int)
            self.check_unused(d, used)
        elif n == "WorldBlock":
            if self.in_world:
                raise PbrtError("%s: nested WorldBegin" % d.where())
            if not self.has_camera:
                raise PbrtError("%s: WorldBegin before a Camera directive" % d.where())
            self.in_world = True
            self.ctm = np.eye(4)
            self.coords["world"] = np.eye(4)
            self.run(d.children)
        elif n == "AttributeBlock":
            saved = (self.ctm.copy(), self.gs.copy())
            self.run(d.children)
            self.ctm, self.gs = saved
        elif n == "TransformBlock":
            saved = self.ctm.copy()
            self.run(d.children)
            self.ctm = saved
        elif n == "ObjectBlock":
            if self.obj is not None:
                raise PbrtError("%s: ObjectBegin inside ObjectBegin" % d.where())
            saved = (self.ctm.copy(), self.gs.copy())
            self.objects[d.args[0]] = []
            self.obj = d.args[0]
            self.run(d.children)
            self.obj = None
            self.ctm, self.gs = saved
        elif n == "ObjectInstance":
            self.instance(d)
        elif n == "Material":
            self.gs.material = self.material(d, d.args[0])
        elif n == "MakeNamedMaterial":
            p = self.get(d, "type", ("string",), set())
            if p is None:
                raise PbrtError("%s: MakeNamedMaterial needs \"string type\"" % d.where())
            self.gs.named[d.args[0]] = self.material(d, p.values[0])
        elif n == "NamedMaterial":
            if d.args[0] not in self.gs.named:
                raise PbrtError("%s: unknown material %r" % (d.where(), d.args[0]))
            self.gs.material = self.gs.named[d.args[0]]
        elif n == "Texture":
            name, ttype, cls = d.args
            if cls != "constant":
                raise PbrtError("%s: texture class %r is not supported (constant)" % (d.where(), cls))
            used = set()
            if ttype in ("spectrum", "color"):
                self.gs.textures[("spectrum", name)] = self.spectrum(d, "value", (1, 1, 1), used)
            elif ttype == "float":
                self.gs.textures[("float", name)] = self.scalar_tex(d, "value", 1.0, used)
            else:
                raise PbrtError("%s: texture type %r is not supported" % (d.where(), ttype))
            self.check_unused(d, used)
        elif n == "AreaLightSource":
            if d.args[0] != "diffuse":
                raise PbrtError("%s: area light %r is not supported (diffuse)" % (d.where(), d.args[0]))
            used = {"samples", "nsamples"}
            L = self.spectrum(d, "L", (1, 1, 1), used)
            sc = self.spectrum(d, "scale", (1, 1, 1), used)
            p = self.get(d, "twosided", ("bool",), used)
            if p is not None and p.values[0]:
                raise PbrtError("%s: two-sided area lights are not supported" % d.where())
            self.check_unused(d, used)
            self.gs.area_L = tuple(float(np.float32(a) * np.float32(b)) for a, b in zip(L, sc))
        elif n == "LightSource":
            if d.args[0] != "sky":
                raise PbrtError("%s: light %r is not supported (diffuse area lights and \"sky\")"
                                % (d.where(), d.args[0]))
            if self.out.desc.sky is not None:
                raise PbrtError("%s: only one sky" % d.where())
            used = set()
            self.out.desc.sky = (self.spectrum(d, "zenith", (1, 1, 1), used),
                                 self.spectrum(d, "horizon", (1, 1, 1), used))
            self.check_unused(d, used)
        elif n == "Shape":
            self.shape(d)
        else:
            raise PbrtError("%s: not supported" % d.where())

    def pixel_filter(self, d):
        """pbrt-v3 filter parameters: xwidth/ywidth (radius), Gaussian alpha,
        Mitchell B and C, Lanczos tau; defaults per filter as pbrt-v3's."""
        from . import film
        kind = d.args[0]
        if kind not in film.NAMES:
            raise PbrtError("%s: filter %r is not supported (box, triangle, gaussian, mitchell, sinc)"
                            % (d.where(), kind))
        t = film.NAMES[kind]
        used = set()
        r = film.DEFAULTS[t]["radius"]
        rx, ry = self.one(d, "xwidth", r, used), self.one(d, "ywidth", r, used)
        kw = {}
        if t == film.GAUSSIAN:
            kw["alpha"] = self.one(d, "alpha", 2.0, used)
        elif t == film.MITCHELL:
            kw["b"], kw["c"] = self.one(d, "B", 1.0 / 3.0, used), self.one(d, "C", 1.0 / 3.0, used)
        elif t == film.LANCZOS:
            kw["tau"] = self.one(d, "tau", 3.0, used)
        if not (0 < rx <= 32 and 0 < ry <= 32):
            raise PbrtError("%s: filter radius must be in (0, 32]" % d.where())
        self.check_unused(d, used)
        self.out.filter = film.filter_desc(t, radius=(rx, ry), **kw)

    def camera(self, d):
        if self.in_world:
            raise PbrtError("%s: Camera inside the world block" % d.where())
        if d.args[0] != "perspective":
            raise PbrtError("%s: camera %r is not supported (perspective)" % (d.where(), d.args[0]))
        if self.lookat is None or not np.allclose(self.ctm, look_at(self.lookat[0:3], self.lookat[3:6],
                                                                    self.lookat[6:9]), rtol=0, atol=0):
            raise PbrtError("%s: the camera transform must be a single LookAt" % d.where())
        used = {"shutteropen", "shutterclose"}
        self.fov = self.one(d, "fov", 90.0, used)
        if self.one(d, "lensradius", 0.0, used) != 0.0:
            raise PbrtError("%s: thin-lens cameras are not supported" % d.where())
        used.add("focaldistance")
        for nm in ("screenwindow", "frameaspectratio"):
            used.add(nm)
            if d.param(nm) is not None:
                raise PbrtError("%s: %s is not supported" % (d.where(), nm))
        self.check_unused(d, used)
        self.coords["camera"] = np.linalg.inv(self.ctm)
        self.has_camera = True

    def finish(self):
        out, desc, a = self.out, self.out.desc, self.lookat
        out.camera = S.camera(tuple(a[0:3]), tuple(a[3:6]), tuple(a[6:9]), self.fov, self.res[0], self.res[1])
        verts, idx, tri_mat, tri_em, spheres, sph_mat = [], [], [], [], [], []
        base = 0
        for item in self.prims:
            if item[0] == "mesh":
                _, w, ii, mat, em = item
                verts.append(w)
                idx.append(ii + base)
                base += len(w)
                tri_mat.extend([mat] * len(ii))
                tri_em.extend([tuple(em) if em is not None else (0.0, 0.0, 0.0)] * len(ii))
            else:
                spheres.append(item[1])
                sph_mat.append(item[2])
        desc.verts = np.concatenate(verts).astype(np.float32) if verts else np.zeros((0, 3), np.float32)
        desc.idx = np.concatenate(idx).astype(np.uint32) if idx else np.zeros((0, 3), np.uint32)
        desc.tri_material, desc.tri_emission = tri_mat, tri_em
        desc.spheres = np.asarray(spheres, np.float32).reshape(-1, 4)
        desc.sphere_material = sph_mat
        if not desc.materials:
            desc.add_material(S.matte((0.5, 0.5, 0.5)))
        return out


def load(ast, base_dir="."):
    """Flat scene, camera and settings of a parsed pbrt-v3 description.
    Triangles keep file order, spheres follow them in file order."""
    ld = _Loader(base_dir)
    ld.run(ast)
    if not ld.has_camera:
        raise PbrtError("the scene has no Camera directive")
    if not ld.in_world:
        raise PbrtError("the scene has no WorldBegin block")
    return ld.finish()


def load_file(path):
    return load(parse_file(path), os.path.dirname(os.path.abspath(path)))


def load_text(text, base_dir="."):
    return load(parse_to_ast(text, "<string>", base_dir), base_dir)


# ------------------------------------------------------------------------------------ writer
def _f(x):
    return repr(float(np.float32(x)))


def _vals(a):
    return "[ " + " ".join(_f(v) for v in np.asarray(a).reshape(-1)) + " ]"


def _material_text(m):
    m = np.asarray(m).reshape(-1)[0]
    t = int(m["type"])
    if t == S.MATTE:
        return 'Material "matte" "rgb Kd" %s' % _vals(m["kd"])
    if t == S.METAL:
        return ('Material "metal" "float roughness" [ %s ] "bool remaproughness" "false" "rgb eta" %s "rgb k" %s'
                % (_f(m["alpha"]), _vals(m["eta"]), _vals(m["k"])))
    return 'Material "glass" "float index" [ %s ]' % _f(m["ior"])


def _filter_text(f):
    from . import film
    f = np.asarray(f).reshape(-1)[0]
    t = int(f["type"])
    name = {film.BOX: "box", film.TRIANGLE: "triangle", film.GAUSSIAN: "gaussian", film.MITCHELL: "mitchell",
            film.LANCZOS: "sinc"}[t]
    out = 'PixelFilter "%s" "float xwidth" [ %s ] "float ywidth" [ %s ]' % (name, _f(f["radius"][0]),
                                                                          _f(f["radius"][1]))
    if t == film.GAUSSIAN:
        out += ' "float alpha" [ %s ]' % _f(f["alpha"])
    elif t == film.MITCHELL:
        out += ' "float B" [ %s ] "float C" [ %s ]' % (_f(f["b"]), _f(f["c"]))
    elif t == film.LANCZOS:
        out += ' "float tau" [ %s ]' % _f(f["tau"])
    return out


def write_text(desc, camera, spp=16, max_depth=5, filename=None, pixel_filter=None, sampler="random"):
    """pbrt-v3 text of a flat scene. Triangles keep their order: each run of
    triangles with one material and emission becomes one trianglemesh (its
    own vertices, in first-use order); spheres become translated spheres; the
    analytic sky becomes LightSource "sky". Loading the text gives back the
    same geometry, per-primitive materials and emission, bit for bit."""
    c = np.asarray(camera).reshape(-1)[0]
    film = 'Film "image" "integer xresolution" [ %d ] "integer yresolution" [ %d ]' % (int(c["width"]),
                                                                                     int(c["height"]))
    if filename is not None:
        film += ' "string filename" "%s"' % filename
    lines = ["LookAt %s  %s  %s" % (" ".join(_f(v) for v in c["pos"]), " ".join(_f(v) for v in c["look"]),
                                    " ".join(_f(v) for v in c["up"])),
             'Camera "perspective" "float fov" [ %s ]' % _f(c["fov"]), film,
             'Sampler "%s" "integer pixelsamples" [ %d ]' % (sampler, spp),
             'Integrator "path" "integer maxdepth" [ %d ]' % max_depth]
    if pixel_filter is not None:
        lines.append(_filter_text(pixel_filter))
    lines.append("WorldBegin")
    a = desc.arrays()
    mats = a["materials"]
    if a["sky"] is not None:
        lines.append('LightSource "sky" "rgb zenith" %s "rgb horizon" %s' % (_vals(a["sky"][:3]),
                                                                           _vals(a["sky"][3:])))
    n_t = len(a["idx"])
    pm, pe = a["prim_material"], a["prim_emission"]
    t = 0
    while t < n_t:
        u = t + 1
        while u < n_t and pm[u] == pm[t] and np.array_equal(pe[u], pe[t]):
            u += 1
        run = a["idx"][t:u].reshape(-1)
        used, first = np.unique(run, return_index=True)
        order = used[np.argsort(first)]
        local = np.empty(int(order.max()) + 1, np.int64)
        local[order] = np.arange(len(order))
        lines.append("AttributeBegin")
        lines.append("  " + _material_text(mats[pm[t]]))
        if np.any(pe[t] != 0):
            lines.append('  AreaLightSource "diffuse" "rgb L" %s' % _vals(pe[t]))
        lines.append('  Shape "trianglemesh" "integer indices" [ %s ]' % " ".join(str(v) for v in local[run]))
        lines.append('    "point P" %s' % _vals(a["verts"][order]))
        lines.append("AttributeEnd")
        t = u
    for s, m in zip(a["spheres"], pm[n_t:]):
        lines.append("AttributeBegin")
        lines.append("  Translate %s %s %s" % (_f(s[0]), _f(s[1]), _f(s[2])))
        lines.append("  " + _material_text(mats[m]))
        lines.append('  Shape "sphere" "float radius" [ %s ]' % _f(s[3]))
        lines.append("AttributeEnd")
    lines.append("WorldEnd")
    return "\n".join(lines) + "\n"


def render(loaded, spp=None, wavefront=False, seed=0, builder="SAH"):
    """Render a loaded scene on the GPU: (H, W, 3) float32 tensor of the
    pixel values (mean radiance through the scene's pixel filter, with the
    scene's sampler), by the megakernel or the wavefront renderer."""
    from .render import PathIntegrator, Scene, WavefrontPathIntegrator
    spp = loaded.spp if spp is None else spp
    sc = Scene(loaded.desc, builder=builder)
    if wavefront:
        pi = WavefrontPathIntegrator(sc, loaded.camera, spp=spp, max_depth=loaded.max_depth, seed=seed,
                                     max_paths=1 << 20, sampler=loaded.sampler)
    else:
        pi = PathIntegrator(loaded.camera, spp=spp, max_depth=loaded.max_depth, seed=seed, sampler=loaded.sampler)
    if loaded.filter is not None:
        from .film import Filter, render_filtered, resolve
        f = loaded.filter.reshape(-1)[0]
        return resolve(render_filtered(pi, sc, Filter(int(f["type"]), radius=tuple(f["radius"]), alpha=f["alpha"],
                                                      b=f["b"], c=f["c"], tau=f["tau"])))
    return pi.render(sc)


def main(argv=None):
    import argparse
    from .imageio import write_image
    ap = argparse.ArgumentParser(description="render a pbrt-v3 scene file on the GPU and write the image "
                                             "(.pfm, .png or .bmp by the output's suffix)")
    ap.add_argument("scene")
    ap.add_argument("-o", "--output", help="output .pfm, .png or .bmp (default: the Film filename; a suffix "
                                           "other than those becomes .pfm)")
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--wavefront", action="store_true")
    ap.add_argument("--builder", default="SAH", choices=("SAH", "SAH_GPU", "LBVH"))
    a = ap.parse_args(argv)
    ld = load_file(a.scene)
    for w in ld.warnings:
        print("warning:", w)
    img = render(ld, a.spp, a.wavefront, builder=a.builder)
    out = a.output
    if out is None:
        out = ld.filename or "pbrt.pfm"
        if os.path.splitext(out)[1].lower() not in (".pfm", ".png", ".bmp"):
            print("warning: no writer for %r; writing PFM" % out)
            out = os.path.splitext(out)[0] + ".pfm"
    write_image(out, img.cpu().numpy())
    print("wrote", out)


if __name__ == "__main__":
    main()
