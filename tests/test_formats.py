"""On-disk and wire formats of the host side (eq_amd/host/field_io.hpp):
the dolfin::File "compressed" series (.pvd + .vtu, src/fHSL.cpp:606-636),
diffusionPETSc::RecordData's .vtr (diffuclass.cpp:560-583) and the
controller's JSON frame log (src/main.cpp:81-187).  The writers run in a
small C++ driver on fields given to it; every value read back must be the
double that went in, bit for bit, and the mesh must be the "right"
triangulation of the trap."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import vtk_read  # noqa: E402

EXE = os.path.join(ROOT, "tests", "host", "bin", "io_check")


@pytest.fixture(scope="module")
def io_exe():
    if not os.path.exists(EXE):
        sys.path.insert(0, ROOT)
        import __graft_entry__
        __graft_entry__.build_io_check()
    return EXE


def _run(io_exe, tmp_path, nx, ny, W, H, dt, fields, cells):
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<4i3d", nx, ny, len(fields), len(cells), W, H, dt))
        for u in fields:
            f.write(np.ascontiguousarray(u, np.float64).tobytes())
        f.write(np.ascontiguousarray(cells, np.float64).tobytes())
    out = tmp_path / "out"
    out.mkdir()
    subprocess.run([io_exe, str(inp), str(out)], check=True, timeout=120)
    return out


def _fields(nx, ny, nsteps, seed):
    rng = np.random.default_rng(seed)
    fs = []
    for k in range(nsteps):
        u = rng.standard_normal((nx + 1) * (ny + 1)) * 10.0 ** rng.integers(-300, 300, (nx + 1) * (ny + 1))
        u[:4] = [0.0, -0.0, 1.0 / 3.0, 5e-324]
        fs.append(u)
    return fs


@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (64, 33)])
def test_pvd_series_round_trip(io_exe, tmp_path, shape):
    nx, ny = shape
    W, H, dt = nx * 0.25, ny * 0.5, 0.1
    fields = _fields(nx, ny, 3, nx * 100 + ny)
    cells = np.random.default_rng(1).random((4, 5))
    out = _run(io_exe, tmp_path, nx, ny, W, H, dt, fields, cells)
    N = (nx + 1) * (ny + 1)
    j, i = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    xs = (W * j.ravel()) / nx
    ys = (H * i.ravel()) / ny
    for stem, compressed in (("z", True), ("r", False)):
        series = vtk_read.read_pvd(out / f"{stem}.pvd")
        assert [t for t, _ in series] == [(k + 1) * dt for k in range(3)]
        assert [f for _, f in series] == [f"{stem}{k:06d}.vtu" for k in range(3)]
        for k, (_, name) in enumerate(series):
            v = vtk_read.read(out / name)
            assert v["type"] == "UnstructuredGrid"
            assert (v["attrs"].get("compressor") == "vtkZLibDataCompressor") == compressed
            assert int(v["piece"]["NumberOfPoints"]) == N and int(v["piece"]["NumberOfCells"]) == 2 * nx * ny
            a = v["arrays"]
            assert a["u"].tobytes() == np.asarray(fields[k]).tobytes()  # bit-exact, natural order
            assert np.array_equal(a["Points"][:, 0], xs) and np.array_equal(a["Points"][:, 1], ys)
            assert not a["Points"][:, 2].any()
            tri = a["connectivity"].reshape(-1, 3)
            assert np.array_equal(a["offsets"], 3 * np.arange(1, 2 * nx * ny + 1))
            assert np.all(a["types"] == 5)
            P = a["Points"][:, :2]
            e1 = P[tri[:, 1]] - P[tri[:, 0]]
            e2 = P[tri[:, 2]] - P[tri[:, 0]]
            area2 = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
            assert np.all(area2 > 0)  # counter-clockwise
            assert abs(area2.sum() / 2 - W * H) <= 1e-12 * W * H
            # "right" diagonal: both triangles of cell (i, j) hold v00 and v11
            c = tri.reshape(ny, nx, 2, 3)
            v00 = (np.arange(ny)[:, None] * (nx + 1) + np.arange(nx)[None, :])
            for t in range(2):
                assert np.all((c[:, :, t, :] == v00[..., None]).any(-1))
                assert np.all((c[:, :, t, :] == (v00 + nx + 2)[..., None]).any(-1))
            # every interior edge is shared by exactly two triangles
            edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
            _, cnt = np.unique(edges, axis=0, return_counts=True)
            assert cnt.max() == 2 and (cnt == 1).sum() == 2 * (nx + ny)


def test_vtr_round_trip(io_exe, tmp_path):
    nx, ny, h = 9, 4, 0.5
    fields = _fields(nx, ny, 2, 5)
    out = _run(io_exe, tmp_path, nx, ny, nx * h, ny * h, 0.1, fields, np.zeros((0, 5)))
    for k in range(2):
        v = vtk_read.read(out / f"g{k + 1:04d}.vtr")
        assert v["type"] == "RectilinearGrid" and v["extent"] == f"0 {nx} 0 {ny} 0 0"
        a = v["arrays"]
        assert np.array_equal(a["x"], np.arange(nx + 1) * h) and np.array_equal(a["y"], np.arange(ny + 1) * h)
        assert a["g"].tobytes() == np.asarray(fields[k]).tobytes()


def test_json_frames(io_exe, tmp_path):
    nx, ny, dt = 5, 3, 0.1
    fields = _fields(nx, ny, 3, 9)
    cells = np.random.default_rng(2).random((6, 5)) * [100, 50, 3.1, 4, 1]
    out = _run(io_exe, tmp_path, nx, ny, 1.0, 1.0, dt, fields, cells)
    path = out / "io-3_7-jsonData.json"
    text = path.read_text()
    doc = json.loads(text)
    # nlohmann::json writes object keys sorted
    assert list(doc.keys()) == sorted(doc.keys())
    assert doc["jsonSimVersion"] == "v00_02" and doc["nodeID"] == 3 and doc["simNumber"] == 7
    assert doc["timeSinceEpoch"] == 1234567890123
    assert doc["parameters"] == {"dt": dt, "name": 'io "check"'}
    assert doc["cellDataFormat"] == [["id", "strainType"], ["x", "y", "angle", "length", "spring", "vargs"]]
    assert len(doc["frames"]) == 3
    for k, fr in enumerate(doc["frames"]):
        assert list(fr.keys()) == ["cells", "divisions", "externalHSL", "simTime"]
        assert fr["simTime"] == (k + 1) * dt
        hsl = np.array(fr["externalHSL"], np.float64)
        assert hsl.shape == (2, (nx + 1) * (ny + 1))
        assert hsl[0].tobytes() == np.asarray(fields[k]).tobytes()  # shortest text reads back exactly
        assert np.array_equal(hsl[1], 2.0 * fields[k])
        assert [c[0] for c in fr["cells"]] == [[c * 1000003, c % 3] for c in range(6)]
        assert np.array(fr["cells"][2][1]).tobytes() == cells[2].tobytes()
        if k == 0:
            assert len(fr["divisions"]) == 6 and fr["divisions"][4][1] == 4
        else:
            assert fr["divisions"] == []
