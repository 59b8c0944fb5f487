"""ConvexPolyhedronAssembly::measure of the C++ host layer (ConvexPolyhedronAssembly.tcc:115-120)
against closed forms, on the CPU: a hexagon and a triangle (the compiled example), and
in 3D a regular octahedron (4 sqrt(3) a^3 for apothem a), a tetrahedron (|det| / 6)
and a polyhedron cut by the box of its face points widened by their extent."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include "sdot_mi355x/PowerDiagram.h"
struct Pc { enum { dim = 3, allow_ball_cut = 0, allow_translations = 0 }; using TI = std::size_t; using TF = double; using CI = std::size_t; };
int main() {
    using namespace sdot_mi355x;
    const double a = 0.4, s3 = std::sqrt(3.0);
    ConvexPolyhedronAssembly<Pc> oct;
    std::vector<Point3> p, n;
    for (int sx = -1; sx <= 1; sx += 2) for (int sy = -1; sy <= 1; sy += 2) for (int sz = -1; sz <= 1; sz += 2) {
        n.push_back(Point3{{sx / s3, sy / s3, sz / s3}});
        p.push_back(Point3{{0.5 + a * sx / s3, 0.5 + a * sy / s3, 0.5 + a * sz / s3}});
    }
    oct.add_convex_polyhedron(p, n, 2.0);
    ConvexPolyhedronAssembly<Pc> tet;
    const double q[12] = {0.05, 0.05, 0.05, 0.95, 0.1, 0.1, 0.1, 0.95, 0.15, 0.2, 0.2, 0.95};
    tet.add_simplex(q, 1.5);
    // a cube of side 0.5 from its six faces, and a box beside it
    ConvexPolyhedronAssembly<Pc> cube;
    std::vector<Point3> cp, cn;
    for (int d = 0; d < 3; ++d) for (int sgn = -1; sgn <= 1; sgn += 2) {
        Point3 pp{{0.5, 0.5, 0.5}}, nn{{0, 0, 0}};
        pp[d] = 0.5 + 0.25 * sgn;
        nn[d] = sgn;
        cp.push_back(pp);
        cn.push_back(nn);
    }
    cube.add_convex_polyhedron(cp, cn, 3.0);
    cube.add_box(Point3{{0, 0, 0}}, Point3{{1, 2, 0.5}}, 0.5);
    std::printf("{\"oct\": %.15f, \"oct_exact\": %.15f, \"tet\": %.15f, \"cube\": %.15f}\n",
                oct.measure(), 2.0 * 4 * s3 * a * a * a, tet.measure(), cube.measure());
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_assembly_measure_closed_forms(tmp_path):
    exe = os.path.join(ROOT, "sdot_amd", "host", "host_example")
    assert os.path.exists(exe), "build first (make -C sdot_amd)"
    r = subprocess.run([exe, "measure"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    src = tmp_path / "m3.cpp"
    src.write_text(PROGRAM)
    out = tmp_path / "m3"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "sdot_amd", "host"), str(src), "-o",
                           str(out), "-L", os.path.join(ROOT, "sdot_amd"), "-lsdot_mi355x",
                           "-Wl,-rpath," + os.path.join(ROOT, "sdot_amd")])
    m = json.loads(subprocess.run([str(out)], capture_output=True, text=True, timeout=60, check=True).stdout)
    assert abs(m["oct"] - m["oct_exact"]) < 1e-13
    q = np.array([[0.05, 0.05, 0.05], [0.95, 0.1, 0.1], [0.1, 0.95, 0.15], [0.2, 0.2, 0.95]])
    assert abs(m["tet"] - 1.5 * abs(np.linalg.det(q[1:] - q[0])) / 6) < 1e-15
    assert abs(m["cube"] - (3.0 * 0.125 + 0.5 * 1.0)) < 1e-14
