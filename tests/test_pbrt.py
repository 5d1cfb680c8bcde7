"""pbrt-v3 scene files: the parser against the reference parser's own test
cases (core/parser/parsers/tests.rs), the loader's graphics-state semantics
against closed forms, the writer by exact round trips, PLY and PFM files."""
import math
import os

import numpy as np
import pytest

from rs_illume_amd import pbrt as P
from rs_illume_amd import scenes as S
from rs_illume_amd.imageio import parse_pfm, pfm_bytes, read_pfm, write_pfm


def one_param(text):
    """Parameters of a Shape directive holding just `text`."""
    (d,) = P.parse_to_ast('Shape "x" ' + text)
    return d.params


# ------------------------------------------------------------------ parser (tests.rs cases)
@pytest.mark.parametrize("text,value", [("3.14", 3.14), ("-3.14", -3.14), ("3e10", 3e10), ("3E10", 3e10),
                                        ("3e-10", 3e-10), ("0.5", 0.5), (".5", 0.5), ("4", 4.0),
                                        ("4350", 4350.0), ("-1", -1.0), ("0", 0.0)])
def test_numbers(text, value):
    """tests.rs:65-81 (test_parse_number, test_parse_int)."""
    (d,) = P.parse_to_ast("Translate %s 0 0" % text)
    assert d.args[0] == value


def test_comments_and_whitespace():
    """tests.rs:49-54, 648-652: comments anywhere between tokens."""
    ast = P.parse_to_ast(" \t#test\n  #test\n#test\nScale 1 2 3#test\nTranslate 1 2 3")
    assert [d.name for d in ast] == ["Scale", "Translate"]  # tests.rs:687-702
    assert ast[0].args == [1.0, 2.0, 3.0] and ast[1].args == [1.0, 2.0, 3.0]
    assert P.parse_to_ast(" \t\n  ") == []


@pytest.mark.parametrize("text,ptype,values", [
    ('"integer test" 20', "integer", [20]), ('"integer test" [20]', "integer", [20]),
    ('"integer test" [20 42 36]', "integer", [20, 42, 36]),
    ('"float test" 20.5', "float", [20.5]), ('"float test" [20.5]', "float", [20.5]),
    ('"float test" [20.5 42 36.0]', "float", [20.5, 42.0, 36.0]),
    ('"point2 test" [20.5 10]', "point2", [20.5, 10.0]), ('"vector2 test" [20.5 10]', "vector2", [20.5, 10.0]),
    ('"point3 test" [20.5 10 0]', "point3", [20.5, 10.0, 0.0]),
    ('"point test" [20.5 10 0]', "point3", [20.5, 10.0, 0.0]),
    ('"vector3 test" [20.5 10 0]', "vector3", [20.5, 10.0, 0.0]),
    ('"vector test" [20.5 10 0]', "vector3", [20.5, 10.0, 0.0]),
    ('"normal3 test" [20.5 10 0]', "normal3", [20.5, 10.0, 0.0]),
    ('"normal test" [20.5 10 0]', "normal3", [20.5, 10.0, 0.0]),
    ('"rgb test" [.2 .5 .3]', "rgb", [0.2, 0.5, 0.3]), ('"color test" [.2 .5 .3]', "rgb", [0.2, 0.5, 0.3]),
    ('"xyz test" [.4 .5 .6]', "xyz", [0.4, 0.5, 0.6]),
    ('"spectrum test" [300 .3  400 .6   410 .65  415 .8  500 .2  600 .1]', "spectrum",
     [300, .3, 400, .6, 410, .65, 415, .8, 500, .2, 600, .1]),
    ('"blackbody test" [.4 .5]', "blackbody", [0.4, 0.5]),
    ('"spectrum test" "filename.test"', "spectrum", ["filename.test"]),
    ('"string test" "test value"', "string", ["test value"]),
    ('"bool test" "true"', "bool", [True]), ('"bool test" "false"', "bool", [False]),
    ('"texture Kd" "checks"', "texture", ["checks"]),
])
def test_parameters(text, ptype, values):
    """tests.rs:281-646: every parameter type, scalar and bracketed forms."""
    (p,) = one_param(text)
    assert (p.type, p.name) == (ptype, "Kd" if ptype == "texture" else "test")
    assert p.values == values


@pytest.mark.parametrize("text", ['"integer test" 3.5', '"integer test" [1 2.0]', '"point3 test" [1 2]',
                                  '"rgb test" 0.5', '"bool test" "yes"', '"float test" ["a"]',
                                  '"string test" 3', '"float test" [1 2'])
def test_bad_parameters(text):
    with pytest.raises(P.PbrtError):
        P.parse_to_ast('Shape "x" ' + text)


def test_parameter_list():
    """tests.rs:654-685: a list with comments between its parameters."""
    ps = one_param('"bool test" "true" #test\n#test\n"string test2" "test string"')
    assert ps == [P.Param("bool", "test", [True]), P.Param("string", "test2", ["test string"])]


def test_transform_directives():
    """tests.rs:122-279: LookAt, Translate, Scale, Rotate, CoordinateSystem,
    CoordSysTransform, Transform, ConcatTransform, ActiveTransform,
    TransformTimes."""
    ast = P.parse_to_ast("""LookAt 0 0 0 10 10 10 0 1 0
        Translate 4 5.5 1.0
        Scale 4 5.5 1.0
        Rotate 45 4 5.5 1.0
        CoordinateSystem "test"
        CoordSysTransform "test"
        Transform 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16
        ConcatTransform [1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16]
        ActiveTransform StartTime ActiveTransform EndTime ActiveTransform All
        TransformTimes 100 250.6
        ReverseOrientation Identity""")
    assert [d.name for d in ast] == ["LookAt", "Translate", "Scale", "Rotate", "CoordinateSystem",
                                     "CoordSysTransform", "Transform", "ConcatTransform", "ActiveTransform",
                                     "ActiveTransform", "ActiveTransform", "TransformTimes", "ReverseOrientation",
                                     "Identity"]
    assert ast[0].args == [0, 0, 0, 10, 10, 10, 0, 1, 0]
    assert ast[3].args == [45, 4, 5.5, 1.0]
    assert ast[4].args == ["test"]
    assert ast[6].args == [[float(v) for v in range(1, 17)]] == ast[7].args
    assert [d.args[0] for d in ast[8:11]] == ["StartTime", "EndTime", "All"]
    assert ast[11].args == [100.0, 250.6]
    with pytest.raises(P.PbrtError):
        P.parse_to_ast("ActiveTransform Sometimes")


def test_camera_directive():
    """tests.rs:704-719."""
    (c,) = P.parse_to_ast('Camera "realistic"')
    assert c.args == ["realistic"] and c.params == []
    (c,) = P.parse_to_ast('Camera "realistic" "string test" "test"')
    assert c.params == [P.Param("string", "test", ["test"])]


MESH = '''Shape "trianglemesh"
            "integer indices" [0 1 2 0 2 3]
            "point P" [ -20 -20 0   20 -20 0   20 20 0   -20 20 0 ]
            "float st" [ 0 0   1 0    1 1   0 1 ]'''
CHECKS = '''Texture "checks" "spectrum" "checkerboard"
            "float uscale" [8] "float vscale" [8]
            "rgb tex1" [.1 .1 .1] "rgb tex2" [.8 .8 .8]'''
WORLD = '''WorldBegin

    # uniform blue-ish illumination from all directions
    LightSource "infinite" "rgb L" [.4 .45 .5]

    # approximate the sun
    LightSource "distant"  "point from" [ -30 40  100 ]
    "blackbody L" [3000 1.5]

    AttributeBegin
    Material "glass"
    Shape "sphere" "float radius" 1
    AttributeEnd

    AttributeBegin
    %s
    Material "matte" "texture Kd" "checks"
    Translate 0 0 -1
    %s
    AttributeEnd

WorldEnd''' % (CHECKS, MESH)
ROOT = '''LookAt 3 4 1.5  # eye
    .5 .5 0  # look at point
    0 0 1    # up vector
Camera "perspective" "float fov" 45

Sampler "halton" "integer pixelsamples" 128
Integrator "path"
Film "image" "string filename" "simple.png"
    "integer xresolution" [400] "integer yresolution" [400]

''' + WORLD


def test_blocks():
    """tests.rs:721-789: object, attribute, transform and world blocks hold
    their directives as children."""
    (o,) = P.parse_to_ast('ObjectBegin "shape1"\nShape "sphere" "float radius" 1\n%s\nObjectEnd' % MESH)
    assert (o.name, o.args, len(o.children)) == ("ObjectBlock", ["shape1"], 2)
    (a,) = P.parse_to_ast("AttributeBegin\n%s\nMaterial \"matte\" \"texture Kd\" \"checks\"\nTranslate 0 0 -1\n"
                          "%s\nAttributeEnd" % (CHECKS, MESH))
    assert (a.name, len(a.children)) == ("AttributeBlock", 4)
    assert a.children[0].args == ["checks", "spectrum", "checkerboard"] and len(a.children[0].params) == 4
    (t,) = P.parse_to_ast('TransformBegin\nTranslate 1 0 1\nShape "sphere"\nTransformEnd')
    assert (t.name, len(t.children)) == ("TransformBlock", 2)
    (w,) = P.parse_to_ast(WORLD)
    assert (w.name, len(w.children)) == ("WorldBlock", 4)
    for bad in ("AttributeBegin Shape \"sphere\"", "WorldBegin", "ObjectBegin Shape \"sphere\" ObjectEnd"):
        with pytest.raises(P.PbrtError):
            P.parse_to_ast(bad)


def test_root_example():
    """tests.rs:816-855: the example scene of the pbrt-v3 file-format page
    parses to six top-level directives; the loader names the first directive
    it cannot render (the infinite light) with its line."""
    ast = P.parse_to_ast(ROOT)
    assert [d.name for d in ast] == ["LookAt", "Camera", "Sampler", "Integrator", "Film", "WorldBlock"]
    assert ast[0].args == [3, 4, 1.5, .5, .5, 0, 0, 0, 1]
    assert ast[4].param("xresolution").values == [400]
    with pytest.raises(P.PbrtError, match=r"<string>:14: LightSource: light 'infinite'"):
        P.load(ast)


def test_include(tmp_path):
    """tests.rs:791-814: Include splices the included file's directives."""
    (tmp_path / "shapes.pbrt").write_text('Shape "sphere" "float radius" 1\n' + MESH)
    ast = P.parse_file_text = P.parse_to_ast('Include "shapes.pbrt"\nShape "sphere" "float radius" 1',
                                             base_dir=str(tmp_path))
    assert [d.name for d in ast] == ["Shape"] * 3
    assert ast[1].args == ["trianglemesh"] and ast[1].file.endswith("shapes.pbrt")
    sub = tmp_path / "sub"
    sub.mkdir()
    (sub / "inner.pbrt").write_text('Include "../shapes.pbrt"')
    (tmp_path / "main.pbrt").write_text('Include "sub/inner.pbrt"')
    assert len(P.parse_file(str(tmp_path / "main.pbrt"))) == 2
    (tmp_path / "loop.pbrt").write_text('Include "loop.pbrt"')
    with pytest.raises(P.PbrtError, match="nested too deeply"):
        P.parse_file(str(tmp_path / "loop.pbrt"))


def test_unknown_directive_names_line():
    with pytest.raises(P.PbrtError, match=r"<string>:3: unknown directive 'Frobnicate'"):
        P.parse_to_ast("Scale 1 1 1\n\nFrobnicate 2")


# ------------------------------------------------------------------------- transforms
def test_rotate_closed_forms():
    r = P.rotate(90, 0, 0, 1)
    assert np.allclose(r[:3, :3] @ [1, 0, 0], [0, 1, 0], atol=1e-15)
    assert np.allclose(r[:3, :3] @ [0, 1, 0], [-1, 0, 0], atol=1e-15)
    for ang, ax in ((37.0, (1, 2, 3)), (-110.0, (0.3, -1, 0.2))):
        m = P.rotate(ang, *ax)[:3, :3]
        a = np.asarray(ax, float) / np.linalg.norm(ax)
        assert np.allclose(m @ m.T, np.eye(3), atol=1e-14) and np.isclose(np.linalg.det(m), 1.0)
        assert np.allclose(m @ a, a, atol=1e-14)                      # the axis is fixed
        assert np.isclose((np.trace(m) - 1) / 2, math.cos(math.radians(ang)))  # rotation angle


def test_look_at_closed_forms():
    pos, look, up = np.array([1.0, 2, 3]), np.array([4.0, -1, 7]), np.array([0.0, 1, 0])
    m = P.look_at(pos, look, up)
    h = lambda p: (m @ np.r_[p, 1])[:3]
    assert np.allclose(h(pos), 0, atol=1e-14)
    assert np.allclose(h(look), [0, 0, np.linalg.norm(look - pos)], atol=1e-12)
    assert h(pos + up)[1] > 0 and abs(h(pos + np.cross(up, look - pos))[0]) > 0


def test_roughness_to_alpha():
    """trowbridge_reitz_distribution.rs:25-30: alpha(1) is the constant term;
    roughness is clamped below at 1e-3."""
    assert P.roughness_to_alpha(1.0) == float(np.float32(1.62142))
    assert P.roughness_to_alpha(0.0) == P.roughness_to_alpha(1e-3)
    x = math.log(0.3)
    assert math.isclose(P.roughness_to_alpha(0.3), 1.62142 + 0.819955 * x + 0.1734 * x * x
                        + 0.0171201 * x ** 3 + 0.000640711 * x ** 4, rel_tol=1e-6)


# ---------------------------------------------------------------------------- loader
HEAD = 'LookAt 0 0 5 0 0 0 0 1 0\nCamera "perspective" "float fov" [ 30 ]\n' \
       'Film "image" "integer xresolution" [ 32 ] "integer yresolution" [ 24 ]\n'
TRI = 'Shape "trianglemesh" "integer indices" [0 1 2] "point P" [0 0 0  1 0 0  0 1 0]'


def world(body, head=HEAD):
    return P.load_text(head + "WorldBegin\n" + body + "\nWorldEnd\n")


def tri_normals(desc):
    a = desc.arrays()
    v = a["verts"][a["idx"]].astype(np.float64)
    n = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def test_settings_and_camera():
    ld = P.load_text(HEAD + 'Sampler "halton" "integer pixelsamples" 64\nIntegrator "path" "integer maxdepth" [ 7 ]\n'
                     'WorldBegin\n' + TRI + '\nWorldEnd')
    assert (ld.spp, ld.max_depth) == (64, 7)
    c = ld.camera[0]
    assert list(c["pos"]) == [0, 0, 5] and list(c["look"]) == [0, 0, 0] and list(c["up"]) == [0, 1, 0]
    assert (c["fov"], c["width"], c["height"]) == (30, 32, 24)
    d = world(TRI, 'LookAt 0 0 5 0 0 0 0 1 0\nCamera "perspective"\n')
    assert (d.camera[0]["fov"], d.camera[0]["width"], d.camera[0]["height"]) == (90, 1280, 720)
    assert (d.spp, d.max_depth) == (16, 5)  # pbrt-v3 defaults


def test_ctm_composition_and_blocks():
    """pbrt-v3 post-multiplies the current transform: Translate then Scale
    maps p to T(S(p)). AttributeEnd restores transform and material,
    TransformEnd the transform only."""
    ld = world('''AttributeBegin
          Translate 1 0 0
          Scale 2 2 2
          Material "matte" "rgb Kd" [0.1 0.2 0.3]
          TransformBegin
            Translate 0 0 1
            Material "matte" "rgb Kd" [0.4 0.5 0.6]
          TransformEnd
          %s
        AttributeEnd
        %s
        Transform [1 0 0 0  0 1 0 0  0 0 1 0  5 6 7 1]
        %s
        Identity
        Rotate 90 0 0 1
        %s''' % (TRI, TRI, TRI, TRI))
    a = ld.desc.arrays()
    v = a["verts"][a["idx"]]
    assert np.array_equal(v[0], [[1, 0, 0], [3, 0, 0], [1, 2, 0]])      # T(1,0,0) S(2): transform block undone
    assert np.array_equal(v[1], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])      # attribute block undone
    assert np.array_equal(v[2], [[5, 6, 7], [6, 6, 7], [5, 7, 7]])      # column-major matrix: a translation
    assert np.allclose(v[3], [[0, 0, 0], [0, 1, 0], [-1, 0, 0]], atol=1e-7)
    kd = a["materials"][a["prim_material"]]["kd"]
    assert np.allclose(kd[0], [0.4, 0.5, 0.6])                           # material survives TransformEnd
    assert np.allclose(kd[1], [0.5, 0.5, 0.5])                           # default matte after AttributeEnd


def test_orientation():
    """ReverseOrientation and handedness-swapping transforms turn the
    geometric normal (pbrt-v3 Triangle); both together cancel; per-vertex
    normals decide alone."""
    ld = world('\n'.join([TRI,
                          'AttributeBegin ReverseOrientation %s AttributeEnd' % TRI,
                          'AttributeBegin Scale 1 1 -1 %s AttributeEnd' % TRI,
                          'AttributeBegin ReverseOrientation Scale -1 1 1 %s AttributeEnd' % TRI,
                          'AttributeBegin ReverseOrientation Shape "trianglemesh" "integer indices" [0 1 2] '
                          '"point P" [0 0 0  1 0 0  0 1 0] "normal N" [0 0 -1 0 0 -1 0 0 -1] AttributeEnd']))
    n = tri_normals(ld.desc)
    # n = normalize((p0-p2) x (p1-p2)) of the world-space points, negated when
    # ReverseOrientation xor the transform swaps handedness: Scale 1 1 -1 leaves
    # the z=0 triangle in place and negates n; Scale -1 1 1 mirrors the points
    # (n = -z) and ReverseOrientation cancels the swap
    assert np.allclose(n[:, 2], [1, -1, -1, -1, -1])


def test_materials_textures_named():
    ld = world('''Texture "c" "color" "constant" "rgb value" [0.25 0.5 0.75]
        Texture "r" "float" "constant" "float value" 1.0
        MakeNamedMaterial "copper" "string type" "metal" "texture roughness" "r"
        MakeNamedMaterial "clear" "string type" "glass" "float index" 1.33
        Material "matte" "texture Kd" "c"
        %s
        NamedMaterial "copper"
        %s
        NamedMaterial "clear"
        %s
        Material "metal" "float roughness" 0.05 "bool remaproughness" "false" "rgb eta" [1 1 1] "rgb k" [2 2 2]
        %s''' % (TRI, TRI, TRI, TRI))
    a = ld.desc.arrays()
    m = a["materials"][a["prim_material"]]
    assert m["type"].tolist() == [S.MATTE, S.METAL, S.GLASS, S.METAL]
    assert np.allclose(m["kd"][0], [0.25, 0.5, 0.75])
    assert m["alpha"][1] == np.float32(1.62142)        # roughness 1 remapped
    assert tuple(m["eta"][1]) == tuple(np.float32(S.COPPER_ETA))
    assert m["ior"][2] == np.float32(1.33)
    assert m["alpha"][3] == np.float32(0.05) and list(m["k"][3]) == [2, 2, 2]
    # identical records share one table entry
    ld2 = world('Material "matte" %s Material "matte" %s' % (TRI, TRI))
    assert len(ld2.desc.materials) == 1


def test_area_lights_and_sky():
    ld = world('''LightSource "sky" "rgb zenith" [0.1 0.2 0.3] "rgb horizon" [1 1 1]
        AttributeBegin
          AreaLightSource "diffuse" "rgb L" [4 5 6] "rgb scale" [0.5 0.5 0.5]
          %s
        AttributeEnd
        %s''' % (TRI, TRI))
    a = ld.desc.arrays()
    assert np.array_equal(a["prim_emission"], [[2, 2.5, 3], [0, 0, 0]])
    assert np.allclose(a["sky"], [0.1, 0.2, 0.3, 1, 1, 1])


def test_spheres_and_instances():
    ld = world('''AttributeBegin Translate 1 2 3 Scale 2 2 2 Material "glass" Shape "sphere" "float radius" 0.5
        AttributeEnd
        ObjectBegin "pair"
          Translate 0 0 1
          %s
          Shape "sphere" "float radius" 1
        ObjectEnd
        AttributeBegin Translate 10 0 0 ObjectInstance "pair" AttributeEnd
        AttributeBegin Translate 20 0 0 Scale 3 3 3 ObjectInstance "pair" AttributeEnd''' % TRI)
    a = ld.desc.arrays()
    assert np.array_equal(a["spheres"], [[1, 2, 3, 1], [10, 0, 1, 1], [20, 0, 3, 3]])
    v = a["verts"][a["idx"]]
    assert np.array_equal(v[0], [[10, 0, 1], [11, 0, 1], [10, 1, 1]])
    assert np.array_equal(v[1], [[20, 0, 3], [23, 0, 3], [20, 3, 3]])
    assert a["materials"][a["prim_material"][2]]["type"] == S.GLASS


@pytest.mark.parametrize("body,msg", [
    ('LightSource "infinite" "rgb L" [1 1 1]', "light 'infinite'"),
    ('Shape "disk"', "shape 'disk'"),
    ('Scale 1 2 1 Shape "sphere"', "uniform scale"),
    ('AreaLightSource "diffuse" Shape "sphere"', "only triangle meshes"),
    ('AreaLightSource "diffuse" "bool twosided" "true"', "two-sided"),
    ('Material "plastic"', "material 'plastic'"),
    ('Material "matte" "float sigma" 20', "Oren-Nayar"),
    ('Material "metal" "float uroughness" 0.1 "float vroughness" 0.2', "anisotropic"),
    ('Material "matte" "texture Kd" "nope"', "no constant spectrum texture"),
    ('Texture "t" "spectrum" "checkerboard"', "texture class"),
    ('NamedMaterial "nope"', "unknown material"),
    ('ObjectInstance "nope"', "unknown object"),
    ('Shape "trianglemesh" "integer indices" [0 1 5] "point P" [0 0 0 1 0 0 0 1 0]', "out of range"),
    ('Shape "sphere" "float zmax" 0.5', "partial spheres"),
])
def test_unsupported_is_an_error(body, msg):
    with pytest.raises(P.PbrtError, match=msg):
        world(body)


def test_camera_errors():
    with pytest.raises(P.PbrtError, match="single LookAt"):
        P.load_text('Scale -1 1 1 LookAt 0 0 5 0 0 0 0 1 0 Camera "perspective" WorldBegin WorldEnd')
    with pytest.raises(P.PbrtError, match="single LookAt"):
        P.load_text('LookAt 0 0 5 0 0 0 0 1 0 Translate 1 0 0 Camera "perspective" WorldBegin WorldEnd')
    with pytest.raises(P.PbrtError, match="camera 'orthographic'"):
        P.load_text('LookAt 0 0 5 0 0 0 0 1 0 Camera "orthographic" WorldBegin WorldEnd')
    with pytest.raises(P.PbrtError, match="before a Camera"):
        P.load_text('WorldBegin WorldEnd')


def test_warnings_for_ignored_parameters():
    ld = world(TRI.replace('"point P"', '"float st" [0 0 1 0 0 1] "point P"'),
               HEAD + 'PixelFilter "gaussian"\nSampler "sobol" "integer pixelsamples" 8 "integer dimension" 5\n')
    assert not any("PixelFilter" in w for w in ld.warnings)
    assert any("'dimension' ignored" in w for w in ld.warnings)
    assert ld.sampler == "sobol"
    assert P.load_text(HEAD + 'Sampler "02sequence" WorldBegin %s WorldEnd' % TRI).sampler == "halton"
    assert P.load_text(HEAD + 'Sampler "random" WorldBegin %s WorldEnd' % TRI).sampler == "random"
    assert P.load_text(HEAD + 'WorldBegin %s WorldEnd' % TRI).sampler == "halton"  # pbrt-v3's default
    assert any("'st' ignored" in w for w in ld.warnings)


@pytest.mark.parametrize("text,expect", [
    ('PixelFilter "box"', dict(type=0, radius=(0.5, 0.5))),
    ('PixelFilter "gaussian"', dict(type=2, radius=(2, 2), alpha=2)),
    ('PixelFilter "gaussian" "float xwidth" 1.5 "float ywidth" 1 "float alpha" 3', dict(type=2, radius=(1.5, 1), alpha=3)),
    ('PixelFilter "mitchell"', dict(type=3, radius=(2, 2), b=np.float32(1 / 3), c=np.float32(1 / 3))),
    ('PixelFilter "mitchell" "float B" 0.5 "float C" 0.25', dict(type=3, b=0.5, c=0.25)),
    ('PixelFilter "sinc"', dict(type=4, radius=(4, 4), tau=3)),
    ('PixelFilter "triangle" "float xwidth" 1', dict(type=1, radius=(1, 2))),
])
def test_pixel_filters(text, expect):
    """pbrt-v3 filter parameters and defaults (xwidth/ywidth, alpha, B, C, tau)."""
    ld = world(TRI, HEAD + text + "\n")
    f = ld.filter[0]
    for k, v in expect.items():
        assert np.allclose(f[k], v), (k, f[k], v)


def test_pixel_filter_errors():
    with pytest.raises(P.PbrtError, match="filter 'bspline'"):
        world(TRI, HEAD + 'PixelFilter "bspline"\n')
    with pytest.raises(P.PbrtError, match="radius"):
        world(TRI, HEAD + 'PixelFilter "box" "float xwidth" 0\n')


# ------------------------------------------------------------------------ round trips
def assert_same_scene(d0, d1):
    a, b = d0.arrays(), d1.arrays()
    assert np.array_equal(a["verts"][a["idx"]], b["verts"][b["idx"]])
    assert np.array_equal(a["spheres"], b["spheres"])
    assert np.array_equal(a["prim_emission"], b["prim_emission"])
    assert a["materials"][a["prim_material"]].tobytes() == b["materials"][b["prim_material"]].tobytes()
    assert (a["sky"] is None) == (b["sky"] is None)
    if a["sky"] is not None:
        assert np.array_equal(a["sky"], b["sky"])


ROUND_TRIP = {
    "cornell": lambda: (S.cornell_box(0), S.cornell_camera(48, 40)),
    "cornell_glossy": lambda: (S.cornell_box(0, glossy=True, light_size=0.2, light_L=26.5625), S.cornell_camera(32, 32)),
    "procedural": lambda: S.procedural_scene(22)[:2],
    "terrain_sky": lambda: (S.terrain_scene(1, n=16, n_spheres=20), S.terrain_views(1, 1)[0]),
}


@pytest.mark.parametrize("name", sorted(ROUND_TRIP))
def test_write_load_round_trip(name, tmp_path):
    desc, cam = ROUND_TRIP[name]()
    text = P.write_text(desc, cam, spp=8, max_depth=6, filename="out.pfm")
    path = tmp_path / "scene.pbrt"
    path.write_text(text)
    ld = P.load_file(str(path))
    assert_same_scene(desc, ld.desc)
    assert ld.camera.tobytes() == np.asarray(cam).reshape(1).tobytes()
    assert (ld.spp, ld.max_depth, ld.filename) == (8, 6, "out.pfm")
    assert ld.warnings == []
    # and the written text is a fixed point
    assert P.write_text(ld.desc, ld.camera, 8, 6, "out.pfm") == text


def test_round_trip_keeps_the_filter():
    from rs_illume_amd.film import filter_desc
    desc, cam = S.cornell_box(0), S.cornell_camera(8, 8)
    f = filter_desc("mitchell", radius=(1.5, 2.5), b=0.25, c=0.375)
    ld = P.load_text(P.write_text(desc, cam, pixel_filter=f))
    assert ld.filter.tobytes() == f.tobytes()


def test_round_trip_renders_identically_in_the_oracle(oracle):
    """The scene read back renders the same image as the one written (the
    oracle path tracer, a small Cornell box), bit for bit."""
    desc, cam = S.cornell_box(0), S.cornell_camera(12, 10)
    ld = P.load_text(P.write_text(desc, cam))
    i0, _ = oracle.OracleRenderScene(desc.arrays()).render_sum(cam, 4, 5, seed=2)
    i1, _ = oracle.OracleRenderScene(ld.desc.arrays()).render_sum(ld.camera, 4, 5, seed=2)
    assert np.array_equal(i0, i1)


# ------------------------------------------------------------------------------- PLY
def test_ply_round_trip_and_plymesh(tmp_path):
    rng = np.random.default_rng(3)
    v = rng.normal(size=(40, 3)).astype(np.float32)
    t = rng.integers(0, 40, size=(60, 3))
    P.write_ply(str(tmp_path / "m.ply"), v, t)
    v2, t2 = P.read_ply(str(tmp_path / "m.ply"))
    assert np.array_equal(v2, v.astype(np.float64)) and np.array_equal(t2, t)
    ld = world('Translate 1 0 0 Shape "plymesh" "string filename" "m.ply"'.replace("m.ply", str(tmp_path / "m.ply")))
    a = ld.desc.arrays()
    assert np.array_equal(a["verts"][a["idx"]], (v.astype(np.float64) + [1, 0, 0]).astype(np.float32)[t])
    # relative to the scene file's directory
    (tmp_path / "s.pbrt").write_text(HEAD + 'WorldBegin Shape "plymesh" "string filename" "m.ply" WorldEnd')
    assert len(P.load_file(str(tmp_path / "s.pbrt")).desc.idx) == 60


def test_ply_ascii_quads_and_big_endian(tmp_path):
    (tmp_path / "q.ply").write_text("ply\nformat ascii 1.0\ncomment quad\nelement vertex 4\nproperty float x\n"
                                    "property float y\nproperty float z\nproperty float nx\nelement face 1\n"
                                    "property list uchar int vertex_indices\nend_header\n"
                                    "0 0 0 9\n1 0 0 9\n1 1 0 9\n0 1 0 9\n4 0 1 2 3\n")
    v, t = P.read_ply(str(tmp_path / "q.ply"))
    assert np.array_equal(v, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]) and t.tolist() == [[0, 1, 2], [0, 2, 3]]
    head = ("ply\nformat binary_big_endian 1.0\nelement vertex 3\nproperty double x\nproperty double y\n"
            "property double z\nelement face 1\nproperty list uchar uint vertex_indices\nend_header\n").encode()
    body = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], ">f8").tobytes() + b"\x03" + np.array([2, 1, 0], ">u4").tobytes()
    (tmp_path / "b.ply").write_bytes(head + body)
    v, t = P.read_ply(str(tmp_path / "b.ply"))
    assert np.array_equal(v, [[0, 0, 0], [2, 0, 0], [0, 2, 0]]) and t.tolist() == [[2, 1, 0]]


# ------------------------------------------------------------------------------- PFM
def test_pfm_layout():
    """imageio.rs:264-310: 'PF', width height, -1 (little endian), then rows
    from the bottom of the image up."""
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)  # 2 rows, 3 columns
    data = pfm_bytes(img)
    assert data.startswith(b"PF\n3 2\n-1\n")
    px = np.frombuffer(data[len(b"PF\n3 2\n-1\n"):], "<f4")
    assert px.size == 18
    assert px[:3].tolist() == [9, 10, 11]       # bottom-left pixel first
    assert px[-3:].tolist() == [6, 7, 8]        # top-right pixel last
    assert np.array_equal(parse_pfm(data), img)


def test_pfm_read_variants(tmp_path):
    g = np.array([[1, 2], [3, 4]], np.float32)
    data = b"Pf\n2 2\n2.0\n" + np.ascontiguousarray(g[::-1], ">f4").tobytes()   # big endian, scale 2
    assert np.array_equal(parse_pfm(data)[..., 0], 2 * g)
    rgb = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    write_pfm(str(tmp_path / "x.pfm"), rgb)
    assert np.array_equal(read_pfm(str(tmp_path / "x.pfm")), rgb)
    for bad in (b"P6\n1 1\n-1\n", b"PF\n2 2\n-1\n" + b"\0" * 8, b"PF\n0 1\n-1\n"):
        with pytest.raises(ValueError):
            parse_pfm(bad)
    with pytest.raises(ValueError):
        pfm_bytes(np.zeros((2, 2)))


# ------------------------------------------------------------------------- 8-bit images
def test_gamma_and_bytes_closed_forms():
    """imageio.rs:42-55 with gamma_correct (pbrt.rs:176-183): the linear
    segment below 0.0031308, the power segment above it, rounding by +0.5 and
    truncation, clamping, and NaN -> 0 (Rust's saturating float cast)."""
    from rs_illume_amd.imageio import gamma_correct, to_bytes
    v = np.array([0.0, 0.001, 0.0031308, 0.18, 0.5, 1.0], np.float32)
    v64 = v.astype(np.float64)
    exact = np.where(v64 < 0.0031308, 12.92 * v64, 1.055 * v64 ** (1 / 2.4) - 0.055)
    assert np.allclose(gamma_correct(v), exact, rtol=3e-7, atol=0)
    # float32 as in the reference: 1.055f * 1 - 0.055f rounds to 1 - 2^-24
    assert gamma_correct(np.float32(1.0)) == np.float32(1.055) - np.float32(0.055) == np.float32(1 - 2 ** -24)
    # 0.5 -> 0.7354 -> 188.03; 0.18 -> 0.4614 -> 118.16; 0.001 -> 0.01292 -> 3.79
    assert to_bytes(np.array([0, 0.001, 0.18, 0.5, 1, 1e6, -3, np.nan, np.inf, -np.inf],
                             np.float32)).tolist() == [0, 3, 118, 188, 255, 255, 0, 0, 255, 0]
    # every byte value is reached, and the map is monotone
    ramp = np.linspace(0, 1, 4097, dtype=np.float32)
    b = to_bytes(ramp)
    assert set(b.tolist()) == set(range(256)) and np.all(np.diff(b.astype(int)) >= 0)


def test_png_and_bmp_decode_to_the_bytes(tmp_path):
    """An independent decoder (Pillow) reads back exactly the 8-bit values,
    top row first, for both formats; write_image picks the format by suffix."""
    import io
    from PIL import Image
    from rs_illume_amd.imageio import bmp_bytes, png_bytes, to_bytes, write_image
    img = np.random.default_rng(3).random((9, 13, 3)).astype(np.float32) * 1.3
    img[0, 0] = [np.nan, -1, 7]
    for fn, fmt in ((png_bytes, "PNG"), (bmp_bytes, "BMP")):
        im = Image.open(io.BytesIO(fn(img)))
        assert im.format == fmt and im.size == (13, 9)
        assert np.array_equal(np.asarray(im.convert("RGB")), to_bytes(img))
    for ext in ("png", "bmp", "pfm"):
        write_image(str(tmp_path / ("a." + ext)), img)
    assert np.array_equal(np.asarray(Image.open(str(tmp_path / "a.png"))), to_bytes(img))
    assert np.array_equal(read_pfm(str(tmp_path / "a.pfm"))[1:], img[1:])
    for bad in ("a.tga", "a.exr", "a"):
        with pytest.raises(ValueError):
            write_image(str(tmp_path / bad), img)
    with pytest.raises(ValueError):
        png_bytes(np.zeros((0, 4, 3), np.float32))
