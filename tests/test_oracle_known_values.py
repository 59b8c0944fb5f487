"""The CPU oracle against the known answers of the reference's own tests
(tests/golden/reference_known_values.json) and closed forms."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
KV = json.load(open(os.path.join(HERE, "golden", "reference_known_values.json")))


@pytest.mark.parametrize("case", KV["cp2"], ids=[c["src"] for c in KV["cp2"]])
def test_cp2_known_values(case):
    r = O.cp2_eval(case["box"], case["cuts"], ball=case["ball"])
    assert r["status"] == 0
    if "area" in case:
        assert abs(r["area"] - case["area"]) < case["tol"]
    if case.get("nb_points") is not None:
        assert r["nb_points"] == case["nb_points"]
    if "perimeter" in case:
        assert abs(r["measures"].sum() - case["perimeter"]) < case["tol_perimeter"]


@pytest.mark.parametrize("case", KV["cp3"], ids=[c["src"] for c in KV["cp3"]])
def test_cp3_known_values(case):
    r = O.cp3_eval(case["p0"], case["p1"], case["cuts"])
    assert abs(r["volume"] - case["volume"]) < case["tol"]
    assert abs(r["areas"].sum() - case["area"]) < case["tol"]


def test_rng_golden():
    ref = json.load(open(os.path.join(HERE, "golden", "rng_first8.json")))
    for seed, vals in ref.items():
        assert np.array_equal(O.rng_uniform(int(seed), 8), np.array(vals))


def test_cp2_ball_closed_forms():
    # disc segment cut by the line x = c: area r^2 acos(c/r) - c sqrt(r^2-c^2) remains on the far side
    for c in (-0.9, -0.3, 0.0, 0.25, 0.8):
        r = O.cp2_eval([-10, -10, 10, 10], [[c, 0, 1, 0]], ball=[0, 0, 1])
        kept = math.pi - (math.acos(c) - c * math.sqrt(1 - c * c))
        assert abs(r["area"] - kept) < 1e-12
        # boundary: chord 2 sqrt(1-c^2) with the neighbour id 0, arc with the ball id (-2)
        line = r["measures"][r["ids"] == 0].sum()
        arc = r["measures"][r["ids"] == -2].sum()
        assert abs(line - 2 * math.sqrt(1 - c * c)) < 1e-12
        assert abs(arc - (2 * math.pi - 2 * math.acos(c))) < 1e-12


def _halfspace_measure(dim, box, cuts):
    """Volume of box & {p : (o - p).n >= 0} by scipy's half-space intersection + hull."""
    from scipy.spatial import ConvexHull, HalfspaceIntersection
    hs = []
    for d in range(dim):
        e = np.zeros(dim)
        e[d] = 1
        hs.append(list(e) + [-box[1][d]])
        hs.append(list(-e) + [box[0][d]])
    for c in cuts:
        o, nv = np.array(c[:dim]), np.array(c[dim:])
        hs.append(list(nv) + [-(o @ nv)])
    hs = np.array(hs)
    return ConvexHull(HalfspaceIntersection(hs, np.zeros(dim)).intersections).volume


def test_cp2_random_convex_area_matches_halfspace_hull():
    rng = np.random.default_rng(3)
    for _ in range(50):
        cuts = []
        for _ in range(rng.integers(3, 12)):
            a = rng.uniform(0, 2 * math.pi)
            d = rng.uniform(0.2, 1.0)
            cuts.append([d * math.cos(a), d * math.sin(a), math.cos(a), math.sin(a)])
        r = O.cp2_eval([-10, -10, 10, 10], cuts)
        ref = _halfspace_measure(2, ([-10, -10], [10, 10]), cuts)
        assert abs(ref - r["area"]) < 1e-10 * max(1, ref)


def test_cp3_unit_cube_and_corner_cut():
    r = O.cp3_eval([0, 0, 0], [1, 1, 1], [])
    assert abs(r["volume"] - 1) < 1e-15 and r["nb_faces"] == 6
    # cut off the corner x+y+z > 2.5: tetra of volume (0.5)^3/6
    n = 1 / math.sqrt(3)
    r = O.cp3_eval([0, 0, 0], [1, 1, 1], [[2.5 / 3, 2.5 / 3, 2.5 / 3, n, n, n]])
    assert abs(r["volume"] - (1 - 0.5 ** 3 / 6)) < 1e-14
    assert r["nb_faces"] == 7
    new = r["areas"][r["ids"] == 0][0]
    assert abs(new - math.sqrt(3) / 8) < 1e-14     # equilateral triangle of side sqrt(2)/2


def test_cp3_random_polytopes_match_halfspace_hull():
    rng = np.random.default_rng(5)
    for _ in range(30):
        cuts = []
        for _ in range(rng.integers(4, 25)):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            d = rng.uniform(0.3, 1.0)
            cuts.append(list(d * v) + list(v))
        r = O.cp3_eval([-2, -2, -2], [2, 2, 2], cuts)
        ref = _halfspace_measure(3, ([-2, -2, -2], [2, 2, 2]), cuts)
        assert abs(ref - r["volume"]) < 1e-10 * max(1, ref)
        # faces: the area vector sum of a closed surface vanishes; areas positive
        assert (r["areas"] > -1e-15).all()
