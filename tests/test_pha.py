"""FORBILD .pha parsing (CPU): values the format fixes, checked by hand on the
published head and thorax phantoms (the package's data files), and the
behaviour of the preprocessor and the arithmetic evaluator."""

from __future__ import annotations

import math
import pathlib

import numpy as np
import pytest

from aprict_amd.volume import forbild, pha

DATA = pathlib.Path(pha.__file__).resolve().parents[1] / "data" / "forbild"
HEAD = DATA / "head" / "HeadPhantom.pha"
THORAX = DATA / "thorax" / "ThoraxPhantom.pha"
D = math.pi / 180


def vec(s, k):
    return np.array(tuple(s[k].tolist()), dtype=np.float64)


def test_head_phantom_shapes():
    sh = pha.convert_ph_file_to_shapes(HEAD)
    assert len(sh) == 479
    # { "5" [Ellipsoid: dx=9.6 dy=12 dz=12.5 ] DTMAT(1.800) }
    s = sh[0]
    assert s["type"] == forbild.TYPE_IDS["Ellipsoid"]
    assert (float(s["dx"]), float(s["dy"]), float(s["dz"])) == pytest.approx((9.6, 12.0, 12.5))
    assert float(s["rho"]) == pytest.approx(1.8)
    assert (float(s["x"]), float(s["y"]), float(s["z"])) == (0.0, 0.0, 0.0)
    # { "12"[Ellipsoid: y=-3.6 dx=1.8 dy= 3.6 dz= 3.6 ] DTMAT(1.045) }  (spaces after '=')
    s = sh[2]
    assert float(s["y"]) == pytest.approx(-3.6) and float(s["dy"]) == pytest.approx(3.6)
    assert float(s["rho"]) == pytest.approx(1.045)
    # { "7" ... DTMAT(0)} : density zero
    assert float(sh[3]["rho"]) == 0.0
    # { "14"[Ellipt_Cyl: y=3.6 dx=1.2 dy=4.0 l=0.5*cos(15)
    #        achse(0,-cos(30),sin(30)) a_y(0,sin(30),cos(30))] }: degrees, German axis name,
    # and a_x completed as a_y x axis
    s = next(x for x in sh if x["type"] == forbild.TYPE_IDS["Ellipt_Cyl"])
    assert float(s["l"]) == pytest.approx(0.5 * math.cos(15 * D), rel=1e-6)
    np.testing.assert_allclose(vec(s, "axis"), [0, -math.cos(30 * D), math.sin(30 * D)], atol=1e-6)
    np.testing.assert_allclose(vec(s, "a_y"), [0, math.sin(30 * D), math.cos(30 * D)], atol=1e-6)
    np.testing.assert_allclose(vec(s, "a_x"), [1, 0, 0], atol=1e-6)
    # { "8" [Ellipsoid_free: ... a_z(sin(15),0,cos(15)) a_x(cos(15)*cos(120),sin(120),-sin(15)*cos(120))] }
    s = next(x for x in sh if x["type"] == forbild.TYPE_IDS["Ellipsoid_free"])
    az = np.array([math.sin(15 * D), 0, math.cos(15 * D)])
    ax = np.array([math.cos(15 * D) * math.cos(120 * D), math.sin(120 * D), -math.sin(15 * D) * math.cos(120 * D)])
    np.testing.assert_allclose(vec(s, "a_z"), az, atol=1e-6)
    np.testing.assert_allclose(vec(s, "a_x"), ax, atol=1e-6)
    np.testing.assert_allclose(vec(s, "a_y"), np.cross(az, ax), atol=1e-6)
    assert float(s["dx"]) == pytest.approx(1.206483 * math.cos(15 * D), rel=1e-6)
    # { "16"[Cone_y: y=-11.15 z=-0.2  r1=0.5 r2=0.2 l=1.5 ] DTMAT(1.800) }
    cones = [x for x in sh if x["type"] == forbild.TYPE_IDS["Cone_y"]]
    assert len(cones) == 2
    c = cones[0]
    assert (float(c["y"]), float(c["z"]), float(c["r1"]), float(c["r2"]), float(c["l"])) == pytest.approx(
        (-11.15, -0.2, 0.5, 0.2, 1.5))
    # { [Ellipsoid: x=9.1 ... x<9.1] }: clip plane x <= 9.1 (mode 1)
    clipped = [x for x in sh if x["cas1"]["mode"] != 0]
    assert len(clipped) >= 1
    c1 = clipped[0]["cas1"]
    assert int(c1["mode"]) == 1 and float(c1["dotprod"]) == pytest.approx(9.1)
    np.testing.assert_array_equal(vec(clipped[0]["cas1"], "axis"), [1, 0, 0])
    assert np.all(np.isfinite(sh["rho"]))


def test_thorax_phantom_includes_macros_and_clip_order():
    sh = pha.convert_ph_file_to_shapes(THORAX)
    assert len(sh) == 271
    # { [ Cylinder_z: x=-20  y=0  z=17.5  l=15  r=5 ] formula=H2O  rho=WATER }
    s = sh[0]
    assert s["type"] == forbild.TYPE_IDS["Cylinder_z"]
    assert (float(s["x"]), float(s["z"]), float(s["l"]), float(s["r"]), float(s["rho"])) == pytest.approx(
        (-20, 17.5, 15, 5, 1.0))
    assert sh[2]["type"] == forbild.TYPE_IDS["Ellipt_Cyl_z"]
    # { [ Cylinder: axis(3,7,0) ... r(3,7,sqrt(58))>2.982109 r(1,2,-sqrt(5))<-2.2393277 ] }:
    # the '<' plane comes first
    aorta = [x for x in sh if x["type"] == forbild.TYPE_IDS["Cylinder"] and x["cas2"]["mode"] != 0][0]
    assert int(aorta["cas1"]["mode"]) == 1 and int(aorta["cas2"]["mode"]) == 2
    np.testing.assert_allclose(vec(aorta["cas1"], "axis"), np.array([1, 2, -math.sqrt(5)]) / math.sqrt(10), atol=1e-6)
    assert float(aorta["cas1"]["dotprod"]) == pytest.approx(-2.2393277)
    np.testing.assert_allclose(vec(aorta["cas2"], "axis"), np.array([3, 7, math.sqrt(58)]) / math.sqrt(116),
                               atol=1e-6)
    np.testing.assert_allclose(vec(aorta, "axis"), [3, 7, 0], atol=1e-6)
    # Wirbelkoerper(0,-5,15,2) (a function-like macro from an included file): eight shapes, the
    # last a Cylinder_z at y=-5-4.05 clipped by y<Y-4.05
    cyl = [x for x in sh if x["type"] == forbild.TYPE_IDS["Cylinder_z"] and float(x["r"]) == pytest.approx(0.4)
           and float(x["z"]) == pytest.approx(15.0)]
    assert len(cyl) == 1
    c = cyl[0]
    assert float(c["y"]) == pytest.approx(-9.05) and float(c["l"]) == pytest.approx(2.0)
    assert int(c["cas1"]["mode"]) == 1 and float(c["cas1"]["dotprod"]) == pytest.approx(-9.05)
    np.testing.assert_array_equal(vec(c["cas1"], "axis"), [0, 1, 0])


def test_preprocessor_and_format_rules():
    text = """
    #define A 2
    #define TWICE(v) (2*(v))
    /* block
       comment */ #define UNUSED 1
    #ifdef A
    { [Sphere: x=TWICE(A) r=1 ] rho=0.5 }   // line comment
    #else
    { [Sphere: x=99 r=1 ] rho=0.5 }
    #endif
    #ifndef A
    { [Box: dx=1 dy=1 dz=1 ] rho=1 }
    #endif
    { [Sphere: y=cos(60) z=cos(pi/3) r=2.5mm] dichte=1 }
    Detector{[Plane_xy: x=0mm] x1=1 x2=2 nx=3}
    { [Box: dx=1 dy=2 dz=3 z>1 x<2 y > 5] rho=1 }
    """
    sh = pha.convert_ph_str_to_shapes(text)
    assert len(sh) == 3
    assert float(sh[0]["x"]) == 4.0 and float(sh[0]["rho"]) == 0.5
    # cos of a plain number takes degrees, of any other expression radians
    assert float(sh[1]["y"]) == pytest.approx(0.5) and float(sh[1]["z"]) == pytest.approx(0.5)
    assert float(sh[1]["r"]) == 2.5 and float(sh[1]["rho"]) == 1.0
    b = sh[2]
    assert int(b["cas1"]["mode"]) == 1 and int(b["cas2"]["mode"]) == 2
    np.testing.assert_array_equal(vec(b["cas1"], "axis"), [1, 0, 0])   # '<' planes first
    np.testing.assert_array_equal(vec(b["cas2"], "axis"), [0, 0, 1])
    assert int(b["cas3"]["mode"]) == 0                                   # 'y > 5' has spaces: not a plane


def test_evaluator_refuses_anything_but_arithmetic():
    assert pha.evaluate("2*sqrt(4)+3**2/3") == pytest.approx(7.0)
    for bad in ("__import__('os')", "x.y", "open('f')", "[1,2]", "1 if 1 else 2", "lambda: 1"):
        with pytest.raises(ValueError):
            pha.evaluate(bad)
    with pytest.raises(ValueError):
        pha.convert_ph_str_to_shapes("{ [Sphere: r=1 foo=2] rho=1 }")
    with pytest.raises(ValueError):
        pha.convert_ph_str_to_shapes("{ [Banana: r=1] rho=1 }")
