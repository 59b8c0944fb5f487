"""The C++ host layer (reference-style SpZGrid / ConvexPolyhedronAssembly /
OptimalTransportSolver / get_der_integrals_wrt_weights over the C-ABI) driven
by a compiled program, as a reference user would."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def test_host_example_solves_and_exports_csr(tmp_path):
    exe = os.path.join(ROOT, "sdot_amd", "host", "host_example")
    assert os.path.exists(exe), "build first (make -C sdot_amd)"
    out = subprocess.run([exe, "10000", str(tmp_path / "cells.vtk")], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert out.returncode == 0, r
    assert r["status"] == 0 and r["der_status"] == 0
    assert r["max_rel_mass_err"] < 1e-6
    assert abs(r["mass_sum"] - 1) < 1e-12
    # get_centroids: the mass-weighted mean centroid is the centre of the unit square
    assert r["centroid_mean_err"] < 1e-12
    # the weight columns of the mass rows of the position Jacobian equal the Hessian
    assert r["der_pos_status"] == 0 and r["der_pos_vs_hessian"] < 1e-9
    # display_vtk through VtkOutput::save: one polygon per cell
    assert r["vtk_polygons"] == 10000


def test_host_example_assembly_of_two_shapes():
    """add_convex_polyhedron + add_simplex through the host layer: the masses sum to
    the assembly's closed-form measure and the solve reaches the targets."""
    exe = os.path.join(ROOT, "sdot_amd", "host", "host_example")
    out = subprocess.run([exe, "shapes"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert out.returncode == 0, r
    assert abs(r["mass_sum"] - r["exact"]) < 1e-12 and r["status"] == 0 and r["max_rel_mass_err"] < 1e-6
