"""Writes tests/golden/*.json.

reference_known_values.json holds the known answers that the reference's own
tests assert for the 2D/3D clipper (data only: boxes, cuts, balls and the
expected measures / vertex counts) and for the pixel density (coefficients,
diracs and expected integrals), with the file:line each comes from.  The
values are transcribed from those assertions; nothing of the reference is run.

rng_first8.json pins the synthetic-input generator (splitmix64 -> xoshiro256**,
u = (x >> 11) * 2^-53): the first 8 values per seed, computed here by a
pure-Python implementation of the published algorithm (independent of the C
code it checks).
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
s2 = math.sqrt


def n(x, y):
    l = math.hypot(x, y)
    return [x / l, y / l]


def known_values():
    cases = []
    # tests/test_ConvexPolyhedron2.cpp:16-31 "diam": octagon
    cuts = []
    for y in (-1, 0, 1):
        for x in (-1, 0, 1):
            if x or y:
                cuts.append(n(x, y) + n(x, y))
    cases.append(dict(src="tests/test_ConvexPolyhedron2.cpp:16-31", box=[-10, -10, 10, 10], cuts=cuts, ball=None,
                      area=3.313708499, tol=1e-4))
    # :81-95 only_sphere
    cases.append(dict(src="tests/test_ConvexPolyhedron2.cpp:81-95", box=[-10, -10, 10, 10], cuts=[], ball=[1, 0, 1],
                      area=math.pi, tol=1e-6))
    kv = "tests/test_ConvexPolyhedron2.cpp"
    B = [-10, -10, 10, 10]
    ball = [0, 0, 1]
    table = [
        (":161-164 full", [], math.pi, None),
        (":166-171 no modification", [[1, 0, 1, 0]], math.pi, 0),
        (":173-178 full removal", [[1, 0, -1, 0]], 0.0, 0),
        (":191-196 1 intersection at the right", [[s2(.5), 0, 1, 0]], 3 * math.pi / 4 + .5, 2),
        (":198-203 1 intersection at the left", [[-s2(.5), 0, 1, 0]], math.pi / 4 - .5, 2),
        (":205-211 2 intersections", [[-s2(.5), 0, -1, 0], [s2(.5), 0, 1, 0]], math.pi / 2 + 1, 4),
        (":213-219 second cancels first", [[-s2(.5), 0, -1, 0], [s2(.5), 0, -1, 0]], math.pi / 4 - .5, 2),
        (":221-227 one worthless", [[-s2(.5), 0, -1, 0], [-.9, 0, -1, 0]], 3 * math.pi / 4 + .5, 2),
        (":229-235 line and circle", [[0, 0, -1, 0], [0, 0, 0, 1]], math.pi / 4, 3),
        (":237-243 .", [[0, 0, -1, 0], [0, 0, 0, -1]], math.pi / 4, 3),
        (":245-253 only lines at the end", [[.2, 0, 1, 0], [0, .2, 0, 1], [-.2, 0, -1, 0], [0, -.2, 0, -1]], .16, 4),
        (":255-261 degeneracies", [[0, -1, 1, 0], [0, -1] + n(1, 1)], math.pi / 4 - .5, None),
    ]
    for src, cuts, area, nb in table:
        cases.append(dict(src=kv + src, box=B, cuts=cuts, ball=ball, area=area, nb_points=nb, tol=1e-7))
    for i in range(12):   # :180-189 1 intersection at the middle, 12 directions
        a = 2 * math.pi * i / 12
        cases.append(dict(src=kv + ":180-189", box=B, cuts=[[0, 0, math.cos(a), math.sin(a)]], ball=ball,
                          area=math.pi / 2, nb_points=2, tol=1e-7))
    # :114-154 cases_1: vertex counts (perimeters are the commented expectations)
    c1 = [
        ([[s2(.5), 0, 1, 0], [s2(.6), .1, 1, 0]], 2, 3 * math.pi / 2 + s2(2)),
        ([[s2(.5), 0, 1, 0], [0, 0, 0, 1]], 3, 4.77041),
        ([[s2(.5), 0, 1, 0], [0, 0, 0, -1]], 3, 4.77041),
        ([[s2(.5), 0, 1, 0], [s2(.2), .1, 1, 0]], 2, 5.85774),
        ([[s2(.5), 0, 1, 0], [s2(.6), .1, -1, 0]], 0, 0.0),
        ([[.7, 0, 1, 0], [0, .7, 0, 1], [-.7, 0, -1, 0], [0, -.7, 0, -1]], 4, 5.6),
    ]
    for cuts, nb, per in c1:
        cases.append(dict(src=kv + ":114-154 cases_1", box=B, cuts=cuts, ball=ball, nb_points=nb, perimeter=per,
                          tol_perimeter=1e-4))
    # :313-336 integration_only_lines (Unit)
    cases.append(dict(src=kv + ":321,333", box=[0, 0, 2, 1], cuts=[], ball=None, area=2.0, tol=1e-5))
    cases.append(dict(src=kv + ":324-336", box=[0, 0, 3, 3], cuts=[], ball=[0, 0, 2], area=math.pi, tol=1e-5))
    cases.append(dict(src=kv + ":328-337", box=[0, 0, 8, 8], cuts=[], ball=[4, 4, 2], area=4 * math.pi, tol=1e-5))
    cp3 = []
    # tests/test_ConvexPolyhedron3.cpp:62-76 (commented "diam"): volume 4.77730, boundary 14.3319
    cuts3 = []
    for z in (-1, 0, 1):
        for y in (-1, 0, 1):
            for x in (-1, 0, 1):
                if x or y or z:
                    l = math.sqrt(x * x + y * y + z * z)
                    v = [x / l, y / l, z / l]
                    cuts3.append(v + v)
    cp3.append(dict(src="tests/test_ConvexPolyhedron3.cpp:62-76", p0=[-5, -5, -5], p1=[5, 5, 5], cuts=cuts3,
                    volume=4.77730, area=14.3319, tol=1e-4))
    # tests/test_scaled_image.cpp:56-92: one pixel over [0,1]^2 carrying the quadratic
    # 2 + 3x + 4y + 5x^2 + 6xy + 7y^2 (6 coefficients), unit weights
    img = []
    img.append(dict(src="tests/test_scaled_image.cpp:56-69", lo=[0, 0], hi=[1, 1], sizes=[1, 1], nb_coeffs=6,
                    data=[2, 3, 4, 5, 6, 7], positions=[[0.5, 0.5]], weights=[1.0], integrals=[11.0], tol=1e-6))
    img.append(dict(src="tests/test_scaled_image.cpp:71-92", lo=[0, 0], hi=[1, 1], sizes=[1, 1], nb_coeffs=6,
                    data=[2, 3, 4, 5, 6, 7], positions=[[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]],
                    weights=[1.0] * 4, integrals=[1.28125, 2.46875, 2.84375, 4.40625], tol=1e-6))
    # tests/test_ConvexPolyhedron2.cpp:265-310 "centroid" and :322-336 "integration_only_lines":
    # polygon / disc pieces and their centroids
    B = [-10, -10, 10, 10]
    cen = [
        dict(src=kv + ":273-277 full ball", box=B, cuts=[], ball=[0, 0, 1], centroid=[0.0, 0.0], tol=1e-6),
        dict(src=kv + ":279-284 half disc", box=B, cuts=[[0, 0, 1, 0]], ball=[0, 0, 1], centroid=[-0.424413, 0.0],
             tol=1e-4),
        dict(src=kv + ":286-292 quarter disc", box=B, cuts=[[0, 0, 1, 0], [0, 0, 0, 1]], ball=[0, 0, 1],
             centroid=[-0.424413, -0.424413], tol=1e-4),
        dict(src=kv + ":294-300 quarter disc r 0.5", box=B, cuts=[[0, 0, 1, 0], [0, 0, 0, 1]], ball=[0, 0, 0.5],
             centroid=[-0.212207, -0.212207], tol=1e-4),
        dict(src=kv + ":302-310 only lines", box=B, cuts=[[1, 0, 1, 0], [0, 1, 0, 1], [-1, 0, -1, 0], [0, -1, 0, -1]],
             ball=[0, 0, 10], centroid=[0.0, 0.0], tol=1e-6),
        dict(src=kv + ":322-336 box", box=[0, 0, 2, 1], cuts=[], ball=None, centroid=[1.0, 0.5], tol=1e-5),
    ]
    # tests/test_der_pos.cpp:118-135: the 36 diracs (6 x 6 lattice of pitch 0.08 from
    # 0.05) and their weights on which the reference checks
    # get_der_centroids_and_integrals_wrt_weight_and_positions against finite
    # differences, with InBallW05 on the unit box
    g = [0.05, 0.13, 0.21, 0.29, 0.37, 0.45]
    der_pos = dict(src="tests/test_der_pos.cpp:118-135", box=[0, 0, 1, 1], ball=True,
                   positions=[[x, y] for y in g for x in g],
                   weights=[0.00314773, 0.00393608, 0.00428933, 0.0042071, 0.00371394, 0.00284918,
                            0.00393608, 0.0048636, 0.00525743, 0.00514763, 0.00452475, 0.00332225,
                            0.00428933, 0.00525743, 0.00566234, 0.0055339, 0.00484392, 0.0034862,
                            0.0042071, 0.00514763, 0.0055339, 0.00541476, 0.00476184, 0.00345417,
                            0.00371394, 0.00452475, 0.00484392, 0.00476184, 0.00426395, 0.00322731,
                            0.00284918, 0.00332225, 0.0034862, 0.00345417, 0.00322731, 0.00270565])
    return dict(cp2=cases, cp3=cp3, scaled_image=img, cp2_centroid=cen, der_pos=der_pos)


def rng_first8():
    M = (1 << 64) - 1

    def splitmix(state):
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return state, z ^ (z >> 31)

    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M

    out = {}
    for seed in (0, 1, 7, 12345):
        st = seed
        s = []
        for _ in range(4):
            st, v = splitmix(st)
            s.append(v)
        vals = []
        for _ in range(8):
            r = (rotl((s[1] * 5) & M, 7) * 9) & M
            t = (s[1] << 17) & M
            s[2] ^= s[0]
            s[3] ^= s[1]
            s[1] ^= s[2]
            s[0] ^= s[3]
            s[2] ^= t
            s[3] = rotl(s[3], 45)
            vals.append((r >> 11) * 2.0 ** -53)
        out[str(seed)] = vals
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_known_values.json"), "w") as f:
        json.dump(known_values(), f, indent=1)
    with open(os.path.join(HERE, "rng_first8.json"), "w") as f:
        json.dump(rng_first8(), f, indent=1)
