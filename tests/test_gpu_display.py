"""display_vtk (src/sdot/PowerDiagram/display_vtk.h:13-25, VtkOutput::save,
src/sdot/Display/VtkOutput.tcc:17-79): the cells drawn on the device, written as a
legacy ASCII VTK file, read back and checked against what fixes them -- every 2D
polygon's area (shoelace) equals its cell's mass (density 1), its centroid the
device's centroid; with the ball the arcs' 20 chords lose at most the closed-form
segment area; per piece (pixels, two overlapping shapes) the pieces' areas add up to
the cell; in 3D the faces close a polyhedron whose divergence-theorem volume is the
mass and whose face areas add up to the boundary measure."""
import numpy as np
import pytest

from tests.gpu_parity import inputs

pytestmark = pytest.mark.gpu


def read_vtk(path):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[2] == "ASCII"
    assert lines[3] == "DATASET UNSTRUCTURED_GRID"
    k = 4
    np_ = int(lines[k].split()[1])
    pts = np.array([[float(v) for v in lines[k + 1 + i].split()] for i in range(np_)]).reshape(-1, 3)
    k += 1 + np_
    head = lines[k].split()
    assert head[0] == "CELLS"
    nc, items = int(head[1]), int(head[2])
    polys = []
    for i in range(nc):
        v = [int(x) for x in lines[k + 1 + i].split()]
        assert v[0] == len(v) - 1
        polys.append(np.array(v[1:]))
    assert sum(len(p) + 1 for p in polys) == items
    k += 1 + nc
    assert lines[k] == f"CELL_TYPES {nc}"
    assert all(lines[k + 1 + i] == "7" for i in range(nc))
    k += 1 + nc
    assert lines[k] == f"CELL_DATA {nc}" and lines[k + 1] == "FIELD FieldData 1"
    name, one, cnt, typ = lines[k + 2].split()
    assert one == "1" and int(cnt) == nc and typ == "float"
    owner = np.array([int(float(x)) for x in lines[k + 3].split()])
    assert len(owner) == nc
    return pts, polys, owner, name


def area2(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


def centroid2(p):
    x, y = p[:, 0], p[:, 1]
    cr = x * np.roll(y, -1) - np.roll(x, -1) * y
    a = 0.5 * cr.sum()
    return np.array([((x + np.roll(x, -1)) * cr).sum(), ((y + np.roll(y, -1)) * cr).sum()]) / (6 * a)


def open_polygon(pts, idx):
    p = pts[idx]
    assert np.array_equal(p[0], p[-1])   # closed by its first point, as for_each_approx_seg ends
    return p[:-1, :2]


def test_display_2d_box_polygons_are_the_cells(tmp_path):
    import sdot_amd
    n = 3000
    pos, w = inputs(2, n, 91, wscale=0.2 / n)
    pd = sdot_amd.PowerDiagram(2, "fp64")
    pd.add_box_shape([0, 0], [1, 1])
    pd.set_diracs(pos, w)
    path = tmp_path / "cells.vtk"
    pd.display_vtk(path, "num_dirac")
    m = pd.integrals()
    cen, _ = pd.centroids()
    pts, polys, owner, name = read_vtk(path)
    # one polygon per non-empty cell (close pairs of unequal weights leave some cells
    # empty), in dirac order
    assert name == "num_dirac" and np.array_equal(owner, np.nonzero(m > 0)[0])
    assert (m > 0).sum() > 0.95 * n
    assert np.abs(pts[:, 2]).max() == 0
    for k, idx in enumerate(polys):
        i = owner[k]
        p = open_polygon(pts, idx)
        assert abs(area2(p) - m[i]) <= 1e-12, i
        # the centroid divides moments by the area: tiny cells amplify rounding
        assert np.abs(centroid2(p) - cen[i]).max() <= 1e-14 / m[i] + 1e-12, i


def test_display_2d_ball_arcs(tmp_path):
    import sdot_amd
    n = 2000
    pos, w = inputs(2, n, 92, wscale=0.6 / n, wshift=0.3 / n)
    pd = sdot_amd.PowerDiagram(2, "fp64", ball=True)
    pd.add_box_shape([0, 0], [1, 1])
    pd.set_diracs(pos, w)
    path = tmp_path / "ball.vtk"
    pd.display_vtk(path)
    m = pd.integrals()
    arc, _ = pd.boundary_measures()
    pts, polys, owner, _ = read_vtk(path)
    drawn = set(owner.tolist())
    assert drawn == set(np.nonzero(m > 0)[0].tolist())
    full = 0
    for k, idx in enumerate(polys):
        i = owner[k]
        r = np.sqrt(w[i])
        p = open_polygon(pts, idx)
        d = np.hypot(p[:, 0] - pos[i, 0], p[:, 1] - pos[i, 1])
        # every point in the closed disc (the pass's weights carry tie-breaking offsets of
        # 1e-12 of the squared mean spacing, hence the relative 1e-10)
        assert d.max() <= r * (1 + 1e-10)
        # the arcs, total angle T = arc length / r, drawn as chords of 1/20 of each arc:
        # a loss between 0 and r^2/2 (T - 20 sin(T / 20)) (superadditive in the arcs)
        T = arc[i] / r
        loss = m[i] - area2(p)
        assert -1e-13 <= loss <= 0.5 * r * r * (T - 20 * np.sin(T / 20)) + 1e-13, (i, loss, T)
        if len(idx) == 21 and abs(arc[i] - 2 * np.pi * r) < 1e-9 * r:
            full += 1   # a whole disc: 21 points, the regular 20-gon
            assert abs(area2(p) - 10 * r * r * np.sin(2 * np.pi / 20)) < 1e-9 * r * r
    print(f"[display ball] {len(polys)} polygons, {full} whole discs")


def test_display_2d_pieces_pixels_and_two_shapes(tmp_path):
    import sdot_amd
    n = 600
    pos, w = inputs(2, n, 93, wscale=0.1 / n)
    # pixels: 7 x 5 of density 1 over the unit square; the pieces of a cell add up to it
    pd = sdot_amd.PowerDiagram(2, "fp64")
    pd.set_image_density([0, 0], [1, 1], np.ones((5, 7)))
    pd.set_diracs(pos, w)
    path = tmp_path / "pix.vtk"
    pd.display_vtk(path)
    m = pd.integrals()
    pts, polys, owner, _ = read_vtk(path)
    acc = np.zeros(n)
    for k, idx in enumerate(polys):
        p = open_polygon(pts, idx)
        acc[owner[k]] += area2(p)
        # each piece inside one pixel
        ix = np.floor(p[:, 0].mean() * 7)
        iy = np.floor(p[:, 1].mean() * 5)
        assert p[:, 0].min() >= ix / 7 - 1e-12 and p[:, 0].max() <= (ix + 1) / 7 + 1e-12
        assert p[:, 1].min() >= iy / 5 - 1e-12 and p[:, 1].max() <= (iy + 1) / 5 + 1e-12
    assert np.abs(acc - m).max() <= 1e-12
    assert len(polys) > n
    # two overlapping boxes of density 1: a cell's mass is the sum of its two pieces
    pd = sdot_amd.PowerDiagram(2, "fp64")
    pd.add_box_shape([0, 0], [0.7, 1])
    pd.add_box_shape([0.4, 0], [1, 1])
    pd.set_diracs(pos, w)
    path = tmp_path / "two.vtk"
    pd.display_vtk(path)
    m = pd.integrals()
    pts, polys, owner, _ = read_vtk(path)
    acc = np.zeros(n)
    for k, idx in enumerate(polys):
        acc[owner[k]] += area2(open_polygon(pts, idx))
    assert np.abs(acc - m).max() <= 1e-12


def test_display_3d_faces_close_the_cells(tmp_path):
    import sdot_amd
    n = 1500
    pos, w = inputs(3, n, 94, wscale=0.05 * n ** (-2 / 3))
    pd = sdot_amd.PowerDiagram(3, "fp64")
    pd.add_box_shape([0, 0, 0], [1, 1, 1])
    pd.set_diracs(pos, w)
    path = tmp_path / "cells3.vtk"
    pd.display_vtk(path)
    m = pd.integrals()
    _, per = pd.boundary_measures()
    pts, polys, owner, _ = read_vtk(path)
    vol = np.zeros(n)
    area = np.zeros(n)
    for k, idx in enumerate(polys):
        p = pts[idx]
        va = 0.5 * np.cross(p, np.roll(p, -1, axis=0)).sum(axis=0)   # vector area, outward
        vol[owner[k]] += np.dot(p[0], va) / 3
        area[owner[k]] += np.linalg.norm(va)
    assert np.abs(vol - m).max() <= 1e-12
    assert np.abs(area - per).max() <= 1e-10 * per.max()


def test_display_rejects_periodic_images(tmp_path):
    import sdot_amd
    pos, w = inputs(2, 50, 95)
    pd = sdot_amd.PowerDiagram(2, "fp64")
    pd.add_box_shape([0, 0], [1, 1])
    pd.set_translations([(1, 0), (-1, 0)])
    pd.set_diracs(pos, np.zeros(50))
    with pytest.raises(sdot_amd.SdotError) as e:
        pd.display_vtk(tmp_path / "x.vtk")
    assert e.value.status == sdot_amd.SDOT_NOT_SUPPORTED
