"""Checkpoint and VTK output of the host forest (amr_io.c): the reference
uses p4est_save and p4est's VTK writer, absent here; these pin the project
formats by round trips and by reading the files back independently."""
import re
import struct

import numpy as np
import pytest

import amrhip


def adaptive_forest(seed=1, n=4):
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(2, 1, (1, 0), -1.0, 0.5, 0.75, n, 1, 5)
    for _ in range(3):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    s = f.soln()
    s[:] = rng.uniform(0.5, 2.0, s.shape)
    return f


def test_checkpoint_round_trip(tmp_path):
    f = adaptive_forest()
    path = tmp_path / "ck.bin"
    f.save(path, istep=1234, t=0.125)
    g, istep, t = amrhip.Forest.load(path)
    assert istep == 1234 and t == 0.125
    assert g.quads() == f.quads() and g.n == f.n
    assert np.array_equal(g.soln(), f.soln())
    # the restored forest describes the same mesh to the physics module
    a, b = f.external(), g.external()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert g.is_balanced()


def test_checkpoint_rejects_garbage(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(b"not a checkpoint at all" * 10)
    with pytest.raises(ValueError):
        amrhip.Forest.load(p)
    f = adaptive_forest()
    good = tmp_path / "good.bin"
    f.save(good)
    data = good.read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[:len(data) // 2])      # truncated patches
    with pytest.raises(ValueError):
        amrhip.Forest.load(tmp_path / "cut.bin")


def test_from_quads_validates():
    f = adaptive_forest(3)
    q = f.quads()
    g = amrhip.Forest.from_quads(2, 1, (1, 0), -1.0, 0.5, 0.75, 4, 1, 5, q)
    assert g.quads() == q
    with pytest.raises(ValueError):           # order broken
        amrhip.Forest.from_quads(2, 1, (1, 0), -1.0, 0.5, 0.75, 4, 1, 5, [q[1], q[0]] + q[2:])
    with pytest.raises(ValueError):           # a leaf missing: the tree is not tiled
        amrhip.Forest.from_quads(2, 1, (1, 0), -1.0, 0.5, 0.75, 4, 1, 5, q[:-1])
    t, lv, x, y = q[0]
    with pytest.raises(ValueError):           # misaligned corner
        amrhip.Forest.from_quads(2, 1, (1, 0), -1.0, 0.5, 0.75, 4, 1, 5, [(t, lv, x + 1, y)] + q[1:])


def read_vtu(path):
    """minimal reader of the raw appended .vtu this library writes"""
    data = path.read_bytes()
    head, _, rest = data.partition(b"<AppendedData encoding=\"raw\">\n_")
    text = head.decode()
    npts = int(re.search(r'NumberOfPoints="(\d+)"', text).group(1))
    ncell = int(re.search(r'NumberOfCells="(\d+)"', text).group(1))
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: NumberOfComponents="(\d+)")?(?: Name="(\w+)")? '
                         r'format="appended" offset="(\d+)"/>', text):
        typ, comp, name, off = m.group(1), m.group(2), m.group(3) or "points", int(m.group(4))
        nb = struct.unpack_from("<I", rest, off)[0]
        dt = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[typ]
        arrays[name] = np.frombuffer(rest, dtype=dt, count=nb // np.dtype(dt).itemsize, offset=off + 4)
    return npts, ncell, arrays


def test_vtk_per_quadrant_matches_reference_fields(tmp_path):
    f = adaptive_forest(5)
    path = tmp_path / "soln.vtu"
    f.write_vtk(path)
    npts, ncell, a = read_vtu(path)
    nq = f.nquad
    assert ncell == nq and npts == 4 * nq
    assert np.array_equal(a["level"], f.levels())
    assert np.array_equal(a["treeid"], [q[0] for q in f.quads()])
    assert np.array_equal(a["types"], np.full(nq, 8))
    assert np.array_equal(a["offsets"], 4 * np.arange(1, nq + 1))
    q = f.soln().reshape(nq, -1, 4)
    avg = q.mean(axis=1)
    for v, name in enumerate(["rho", "rhou", "rhov", "rhoe"]):
        got = a[name].reshape(nq, 4)
        assert np.allclose(got, avg[:, v:v + 1], rtol=1e-14, atol=0)
    # corners: the quadrant's geometry in z order
    _, _, geom, _ = f.external()
    pts = a["points"].reshape(nq, 4, 3)
    assert np.allclose(pts[:, :, :2], geom.reshape(nq, 4, 3)[:, :, :2], rtol=0, atol=1e-14)


def test_vtk_subcells(tmp_path):
    f = adaptive_forest(6)
    path = tmp_path / "sub.vtu"
    f.write_vtk(path, subcells=True)
    npts, ncell, a = read_vtu(path)
    n = f.n
    assert ncell == f.nquad * n * n and npts == 4 * ncell
    q = f.soln().reshape(f.nquad * n * n, 4)
    for v, name in enumerate(["rho", "rhou", "rhov", "rhoe"]):
        assert np.array_equal(a[name], q[:, v])
    # pixels tile each quadrant: total area = domain area 2 x 1 trees of 0.75
    p = a["points"].reshape(ncell, 4, 3)
    area = np.sum((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]))
    assert area == pytest.approx(2 * 0.75 * 0.75, rel=1e-12)
