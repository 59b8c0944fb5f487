"""The `.xyw` dirac files of the reference's benchmarks (benchmarks/set_up_diracs_2D.h:171-181):
round trip through the Python helpers and the C++ host layer, bit-exact."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_xyw_round_trip(tmp_path):
    import sdot_amd
    rng = np.random.default_rng(3)
    pos, w = rng.uniform(size=(50, 2)), rng.uniform(size=50) * 1e-3
    f = tmp_path / "d.xyw"
    sdot_amd.write_xyw(f, pos, w)
    p2, w2 = sdot_amd.read_xyw(f)
    assert np.array_equal(p2, pos) and np.array_equal(w2, w)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_xyw_host_layer(tmp_path):
    import sdot_amd
    src = tmp_path / "x.cpp"
    src.write_text(r"""
#include "sdot_mi355x/PowerDiagram.h"
int main(int argc, char **argv) {
    std::vector<sdot_mi355x::Point<2>> p;
    std::vector<double> w;
    if (!sdot_mi355x::read_xyw(argv[1], p, w)) return 1;
    return sdot_mi355x::write_xyw(argv[2], p, w) ? 0 : 2;
}
""")
    exe = tmp_path / "x"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "sdot_amd", "host"), str(src), "-o",
                           str(exe), "-L", os.path.join(ROOT, "sdot_amd"), "-lsdot_mi355x",
                           "-Wl,-rpath," + os.path.join(ROOT, "sdot_amd")])
    rng = np.random.default_rng(4)
    pos, w = rng.uniform(size=(30, 2)), rng.uniform(size=30)
    a, b = tmp_path / "a.xyw", tmp_path / "b.xyw"
    sdot_amd.write_xyw(a, pos, w)
    subprocess.check_call([str(exe), str(a), str(b)])
    p2, w2 = sdot_amd.read_xyw(b)
    assert np.array_equal(p2, pos) and np.array_equal(w2, w)
