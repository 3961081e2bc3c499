"""Output writers (host only): the VTU file is parsed back independently and
every array checked; result.csv matches the fields; checkpoint headers are
validated without a device."""
import re
import struct

import numpy as np
import pytest

import nflow


def parse_vtu(path):
    raw = open(path, "rb").read()
    start = raw.index(b'<AppendedData encoding="raw">\n_') + len(b'<AppendedData encoding="raw">\n_')
    head = raw[:start].decode()
    assert 'header_type="UInt64"' in head
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="(\d+)" format="appended" '
                         r'offset="(\d+)"/>', head):
        typ, name, comps, off = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
        nbytes = struct.unpack_from("<Q", raw, start + off)[0]
        dt = {"Float32": "<f4", "Int32": "<i4", "UInt8": "u1"}[typ]
        arrays[name] = np.frombuffer(raw, dtype=dt, count=nbytes // np.dtype(dt).itemsize, offset=start + off + 8)
        if comps > 1:
            arrays[name] = arrays[name].reshape(-1, comps)
    npts = int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))
    ncell = int(re.search(r'NumberOfCells="(\d+)"', head).group(1))
    assert raw.endswith(b"\n</AppendedData>\n</VTKFile>\n")
    return npts, ncell, arrays


def fields(shape, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]


def test_vtu_round_trip(tmp_path):
    ny, nz, nx = 3, 4, 5
    r, ux, uy, uz = fields((ny, nz, nx))
    t = np.full((ny, nz, nx), nflow.NODE_FLUID, np.int32)
    t[0] = nflow.NODE_FIXEDWALL
    t[-1] = nflow.NODE_MOVINGWALL
    t[:, :, 0] = nflow.NODE_INLET
    p = tmp_path / "out.vtu"
    nflow.write_vtu(str(p), r, ux, uy, uz, t)
    npts, ncell, a = parse_vtu(p)
    assert npts == nx * ny * nz and ncell == (nx - 1) * (ny - 1) * (nz - 1)
    pts = a["Points"]
    y, z, x = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
    assert np.array_equal(pts[:, 0], x.ravel()) and np.array_equal(pts[:, 1], y.ravel())
    assert np.array_equal(pts[:, 2], z.ravel())
    assert np.array_equal(a["Velocity"], np.stack([ux.ravel(), uy.ravel(), uz.ravel()], 1))
    assert np.array_equal(a["Density"], r.ravel())
    assert np.array_equal(a["offsets"], 8 * np.arange(1, ncell + 1)) and (a["types"] == 12).all()
    conn = a["connectivity"].reshape(ncell, 8)
    # every cell is a unit cube: its 8 vertices span exactly 1 in x, y and z
    for c in conn[:: max(1, ncell // 7)]:
        v = pts[c]
        assert np.array_equal(v.max(0) - v.min(0), [1, 1, 1]) and len({tuple(q) for q in v}) == 8
    assert np.array_equal(a["Inlet"], (t.ravel() & nflow.NODE_INLET) != 0)
    assert np.array_equal(a["Moving Wall"], (t.ravel() & nflow.NODE_MOVINGWALL) != 0)
    assert not a["Struct"].any()


def test_vtu_without_flags_and_degenerate_box(tmp_path):
    r, ux, uy, uz = fields((1, 2, 3))
    p = tmp_path / "flat.vtu"
    nflow.write_vtu(str(p), r, ux, uy, uz)
    npts, ncell, a = parse_vtu(p)
    assert npts == 6 and ncell == 0 and "Fluid" not in a


def test_csv(tmp_path):
    ny, nz, nx = 2, 3, 4
    r, ux, uy, uz = fields((ny, nz, nx), seed=1)
    p = tmp_path / "result.csv"
    nflow.write_csv(str(p), r, ux, uy, uz)
    lines = open(p).read().splitlines()
    assert lines[0] == "x,y,z,rho,u,v,w" and len(lines) == 1 + nx * ny * nz
    rows = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    y, z, x = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
    assert np.array_equal(rows[:, 0], x.ravel()) and np.array_equal(rows[:, 1], y.ravel())
    assert np.allclose(rows[:, 3], r.ravel(), atol=1e-6) and np.allclose(rows[:, 6], uz.ravel(), atol=1e-6)


def test_checkpoint_header_validation(tmp_path):
    good = tmp_path / "good.ckp"
    hdr = b"NFLWCKP1" + struct.pack("<iiiiiiqi", 64, 1, 7, 5, 3, 1, 1234, 0) + bytes(20)
    assert len(hdr) == 64
    good.write_bytes(hdr + bytes(4 * 9 * 7 * 5 * 19))
    assert nflow.checkpoint_info(str(good)) == (7, 5, 3, 1234)
    bad = tmp_path / "bad.ckp"
    bad.write_bytes(b"NOTACKPT" + hdr[8:])
    with pytest.raises(nflow.NflowError):
        nflow.checkpoint_info(str(bad))
    with pytest.raises(nflow.NflowError):
        nflow.checkpoint_info(str(tmp_path / "missing.ckp"))
    with pytest.raises(nflow.NflowError):
        nflow.write_vtu(str(tmp_path / "no_such_dir" / "x.vtu"), *fields((2, 2, 2)))
