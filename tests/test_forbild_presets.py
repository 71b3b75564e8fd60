"""FORBILD presets (CPU): the lung phantom grown to 3 generations equals the
published 3-generation file (the check of
tests/test_FORBILD_lungs_construction.py:27-64: every published shape, rounded
to 2 decimals, is among the generated ones, and the counts agree); the file
presets parse, come out in millimetres, and the phantom transforms act as
their definitions say."""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import phantoms
from aprict_amd.volume import FBVolume, forbild

FORB = phantoms.FORBILD
SCALARS = ("rho", "x", "y", "z", "dx", "dy", "dz", "r", "r1", "r2", "l")


def _rounded(s, digits=2):
    """Hashable rounded view of one shape."""
    vals = [int(s["type"])] + [round(float(s[k]), digits) + 0.0 for k in SCALARS]
    for key in ("a_x", "a_y", "a_z", "axis"):
        vals += [round(float(c), digits) + 0.0 for c in s[key].tolist()]
    for cas in ("cas1", "cas2", "cas3"):
        vals += [int(s[cas]["mode"]), round(float(s[cas]["dotprod"]), digits) + 0.0]
        vals += [round(float(c), digits) + 0.0 for c in s[cas]["axis"].tolist()]
    return tuple(vals)


def test_lungs_generated_to_3_generations_equal_the_published_file():
    gen = FORB.LungPhantom(generation=3)
    ref = FORB.LungPhantom3Generations()
    assert len(gen) == len(ref) == 60             # 15 branches x (tube + cap) x (outer + inner)
    got = {_rounded(s) for s in gen.shapes}
    missing = [i for i, s in enumerate(ref.shapes) if _rounded(s) not in got]
    assert not missing, f"published shapes not generated: {missing[:5]}"
    # outer tubes (rho 1) come before inner ones (rho 0.1), as in the file
    rho = gen.shapes["rho"]
    assert np.all(rho[:30] == np.float32(1.0)) and np.all(rho[30:] == np.float32(0.1))


def test_lung_options():
    g4 = FORB.LungPhantom(generation=4)
    assert len(g4) == 4 * (2 ** 5 - 1)             # 31 branches x (tube + cap) x (outer + inner)
    t = FORB.LungPhantom(generation=3, taper=True, fix=True)
    cones = t.shapes[t.shapes["type"] == forbild.TYPE_IDS["Cone"]]
    assert len(cones) == 2 * 14                     # outer and inner tube of every branch below the trachea
    # a tapered tube starts at its parent's radius: the main bronchi widen by
    # 1 / 0.86 towards the trachea (first two outer cones, left then right)
    np.testing.assert_allclose(cones["r1"][:2], cones["r2"][:2] / 0.86, rtol=1e-6)
    np.testing.assert_allclose(cones["r2"][:2], [5.59, 6.45], rtol=1e-6)
    with pytest.raises(ValueError):
        FORB.LungPhantom(generation=1)


def test_file_presets_in_millimetres():
    head = FORB.HeadPhantom()
    assert len(head) == 479
    s = head.shapes[0]
    assert (float(s["dx"]), float(s["dy"]), float(s["dz"])) == pytest.approx((96.0, 120.0, 125.0))
    for name, n in (("JawPhantom", None), ("HipPhantom", None), ("AbdomenPhantom", None), ("ThoraxPhantom", 271)):
        v = getattr(FORB, name)()
        assert len(v) > 0 and np.all(np.isfinite(v.shapes["rho"]))
        if n is not None:
            assert len(v) == n
    # a fresh object per call
    a, b = FORB.HeadPhantom(), FORB.HeadPhantom()
    a.scale_len(2.0)
    assert float(b.shapes[0]["dx"]) == pytest.approx(96.0)


def test_composite_phantom_builds():
    v = FORB.CustomComposite()
    assert len(v) > 4000
    assert float(v.shapes[0]["l"]) > 0 and np.all(np.isfinite(v.shapes["z"]))


def test_phantom_transforms():
    box = forbild.box((1.0, 2.0, 3.0), 2.0, 4.0, 6.0, 0.5, clips=[((0.0, 0.0, 1.0), 3.5, 1)])   # z <= 3.5
    cyl = forbild.cylinder((1.0, -1.0, 2.0), 0.5, 3.0, (0.6, 0.0, 0.8), 1.0)
    v = FBVolume(forbild.stack([box, cyl]))
    v.translate(axis="z", amount=10.0)
    assert float(v.shapes[0]["z"]) == 13.0 and float(v.shapes[0]["cas1"]["dotprod"]) == 13.5
    v.flip(axis=2)
    assert float(v.shapes[0]["z"]) == -13.0 and float(v.shapes[0]["cas1"]["axis"]["z"]) == -1.0
    assert float(v.shapes[0]["cas1"]["dotprod"]) == 13.5           # the plane -z <= 13.5 mirrors z <= 13.5
    assert float(v.shapes[1]["axis"]["z"]) == pytest.approx(-0.8) and float(v.shapes[1]["axis"]["x"]) == pytest.approx(0.6)
    v.scale_len(2.0)
    assert float(v.shapes[0]["dz"]) == 12.0 and float(v.shapes[0]["cas1"]["dotprod"]) == 27.0
    v.multiply_attenuation(4.0)
    assert float(v.shapes[0]["rho"]) == 2.0
    w = FBVolume(forbild.stack([forbild.sphere((0, 0, 0), 1.0, 3.0)])).superimpose_onto(v)
    assert len(w) == 3 and float(w.shapes[2]["rho"]) == 3.0       # the superimposed shape comes last (wins)
    with pytest.raises(ValueError):
        v.flip(axis="w")
