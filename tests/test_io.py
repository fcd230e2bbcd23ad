"""Checkpoint files in the reference's HDF5 layout (``sopht/utils/io.py:13-427,646-694``).

* Round trips mirror ``tests/test_utils/test_io.py`` of the reference (scalar, vector and
  several Eulerian fields; Lagrangian scalar and vector fields; 2D and 3D; single and double).
* The layout itself is checked with the HDF5 distribution's own tools, independently of this
  package's reader: ``h5ls -r`` for the object tree and shapes, ``h5dump`` for stored types,
  attribute dataspaces and values (so a missing ``(1, ...)`` reshape, an untransposed
  Lagrangian array, a mis-split vector component or a wrong type fails here).
* The reader is checked on files rewritten by ``h5repack`` (latest-format groups, chunked
  and gzip-compressed datasets): the values must not change.
* XDMF sidecars: every HDF reference resolves to a dataset with as many elements as the
  DataItem's Dimensions say.
"""

import re
import subprocess
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

import sopht_amd.utils as spu
from sopht_amd.utils import io as sio

H5_BIN = Path("/opt/conda/bin")
HAVE_TOOLS = all((H5_BIN / t).exists() for t in ("h5ls", "h5dump", "h5repack"))
needs_tools = pytest.mark.skipif(not HAVE_TOOLS, reason="HDF5 command-line tools not installed")
pytestmark = pytest.mark.skipif(not sio._LIB_PATH.exists(), reason="libsopht_amd_h5.so not built")


def h5ls(path):
    """{object path: shape tuple or 'Group'} from ``h5ls -r``."""
    out = subprocess.run([str(H5_BIN / "h5ls"), "-r", str(path)], capture_output=True, text=True, check=True).stdout
    tree = {}
    for line in out.splitlines():
        m = re.match(r"^(\S+)\s+(Group|Dataset \{([^}]*)\})", line)
        if m:
            tree[m.group(1)] = "Group" if m.group(2) == "Group" else tuple(
                int(s) for s in m.group(3).split(",") if s.strip())
    return tree


def h5dump_dataset(path, obj):
    """(HDF5 type name, shape, values) of one dataset, printed by ``h5dump``."""
    out = subprocess.run([str(H5_BIN / "h5dump"), "-d", obj, "-y", "-w", "0", "-m", "%.17g", str(path)],
                         capture_output=True, text=True, check=True).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    shape = tuple(int(s) for s in re.search(r"SIMPLE \{ \( ([^)]*) \)", out).group(1).split(","))
    body = out[out.index("DATA {") + 6: out.rindex("}", 0, out.rindex("}"))]
    values = np.array([float(v) for v in re.findall(r"[-+0-9.eEnaif]+", body)])
    return dtype, shape, values.reshape(shape)


def h5dump_attr(path, obj, name):
    """(HDF5 type name, dataspace text, values) of one attribute."""
    out = subprocess.run([str(H5_BIN / "h5dump"), "-a", f"{obj.rstrip('/')}/{name}", "-m", "%.17g", str(path)],
                         capture_output=True, text=True, check=True).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    space = re.search(r"DATASPACE\s+(SCALAR|SIMPLE \{[^}]*\})", out).group(1)
    body = out[out.index("DATA {") + 6:]
    values = [float(v) for line in re.findall(r"\(\d+\):\s*([^\n]+)", body) for v in line.split(",") if v.strip()]
    return dtype, space, np.array(values)


def grid_of(dim):
    return np.array([8, 10, 12][:dim][::-1])  # z, y, x (x fastest), distinct extents


@pytest.fixture
def tmp_cwd(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    return tmp_path


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("dim", [2, 3])
def test_eulerian_scalar_vector_multiple_round_trip(tmp_cwd, dim, precision, rng):
    """``test_io.py:18-171`` of the reference, all three Eulerian cases in one file."""
    real_t = spu.get_real_t(precision)
    grid = grid_of(dim)
    s = rng.standard_normal(tuple(grid)).astype(real_t)
    v = rng.standard_normal((dim, *grid)).astype(real_t)
    origin, dx = 0.25 * np.ones(dim), 0.01 * np.ones(dim)
    io = spu.IO(dim=dim, real_dtype=real_t)
    io.define_eulerian_grid(origin=origin, dx=dx, grid_size=grid)
    io.add_as_eulerian_fields_for_io(scalar_field=s)
    io.add_as_eulerian_fields_for_io(vector_field=v)
    io.save(h5_file_name="eul.h5", time=0.37)

    s2, v2 = np.zeros_like(s), np.zeros_like(v)
    io = spu.IO(dim=dim, real_dtype=real_t)
    io.define_eulerian_grid(origin=origin, dx=dx, grid_size=grid)
    io.add_as_eulerian_fields_for_io(scalar_field=s2, vector_field=v2)
    t = io.load(h5_file_name="eul.h5")
    np.testing.assert_array_equal(s, s2)
    np.testing.assert_array_equal(v, v2)
    assert t == 0.37 and isinstance(t, np.float64)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("connect", [False, True])
def test_lagrangian_scalar_vector_round_trip(tmp_cwd, dim, precision, connect, rng):
    """``test_io.py:174-330`` of the reference (positions, scalar and vector fields)."""
    real_t = spu.get_real_t(precision)
    n = 37
    pos = rng.random((dim, n)).astype(real_t)
    sf = rng.random(n).astype(real_t)
    vf = rng.random((dim, n)).astype(real_t)
    io = spu.IO(dim=dim, real_dtype=real_t)
    io.add_as_lagrangian_fields_for_io(lagrangian_grid=pos, lagrangian_grid_connect=connect, scalar_field=sf,
                                       vector_field=vf)
    io.save(h5_file_name="lag.h5", time=2.0)
    pos2, sf2, vf2 = np.zeros_like(pos), np.zeros_like(sf), np.zeros_like(vf)
    io = spu.IO(dim=dim, real_dtype=real_t)
    io.add_as_lagrangian_fields_for_io(lagrangian_grid=pos2, scalar_field=sf2, vector_field=vf2)
    assert io.load(h5_file_name="lag.h5") == 2.0
    np.testing.assert_array_equal(pos, pos2)
    np.testing.assert_array_equal(sf, sf2)
    np.testing.assert_array_equal(vf, vf2)
    if connect:
        np.testing.assert_array_equal(io.lagrangian_grid_connection["Lagrangian_grid_0"], np.arange(n))
    else:
        assert io.lagrangian_grid_connection == {}


def save_mixed(dim, real_t, rng, name="mixed.h5"):
    grid = grid_of(dim)
    s = rng.standard_normal(tuple(grid)).astype(real_t)
    v = rng.standard_normal((dim, *grid)).astype(real_t)
    pos = rng.random((dim, 7))
    vf = rng.random((dim, 7)).astype(real_t)
    sf = rng.random(7).astype(real_t)
    io = spu.IO(dim=dim, real_dtype=real_t)
    io.define_eulerian_grid(origin=np.array([0.5, 0.25, 0.125][:dim]), dx=np.array([0.1, 0.2, 0.3][:dim]),
                            grid_size=grid)
    io.add_as_eulerian_fields_for_io(s=s, v=v)
    io.add_as_lagrangian_fields_for_io(pos, "body", True, f=vf, m=sf)
    io.save(h5_file_name=name, time=0.5)
    return io, dict(s=s, v=v, pos=pos, vf=vf, sf=sf, grid=grid)


@needs_tools
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_layout_matches_reference_via_hdf5_tools(tmp_cwd, dim, precision, rng):
    """The object tree, shapes, types and values of ``io.py:218-296``, read by h5ls/h5dump."""
    real_t = spu.get_real_t(precision)
    _, d = save_mixed(dim, real_t, rng)
    grid = tuple(int(g) for g in d["grid"])
    ftype = "H5T_IEEE_F32LE" if real_t is np.float32 else "H5T_IEEE_F64LE"
    tree = h5ls("mixed.h5")
    expected = {
        "/": "Group", "/Eulerian": "Group", "/Eulerian/Scalar": "Group", "/Eulerian/Vector": "Group",
        "/Eulerian/Parameters": "Group", "/Eulerian/Scalar/s": (1, *grid),
        **{f"/Eulerian/Vector/v_{k}": (1, *grid) for k in range(dim)},
        "/Lagrangian": "Group", "/Lagrangian/body": "Group", "/Lagrangian/body/Grid": (7, dim),
        "/Lagrangian/body/Connection": (7,), "/Lagrangian/body/Scalar": "Group",
        "/Lagrangian/body/Vector": "Group", "/Lagrangian/body/Scalar/m": (7,),
        "/Lagrangian/body/Vector/f": (7, dim),
    }
    assert tree == expected

    t, shape, vals = h5dump_dataset("mixed.h5", "/Eulerian/Vector/v_1")
    assert t == ftype and shape == (1, *grid)
    np.testing.assert_array_equal(vals[0], d["v"][1].astype(np.float64))
    t, shape, vals = h5dump_dataset("mixed.h5", "/Eulerian/Scalar/s")
    np.testing.assert_array_equal(vals[0], d["s"].astype(np.float64))
    t, shape, vals = h5dump_dataset("mixed.h5", "/Lagrangian/body/Grid")
    assert t == "H5T_IEEE_F64LE"
    np.testing.assert_array_equal(vals, d["pos"].T)
    t, shape, vals = h5dump_dataset("mixed.h5", "/Lagrangian/body/Vector/f")
    assert t == ftype
    np.testing.assert_array_equal(vals, d["vf"].T.astype(np.float64))
    t, shape, vals = h5dump_dataset("mixed.h5", "/Lagrangian/body/Connection")
    assert t == "H5T_STD_I64LE"
    np.testing.assert_array_equal(vals, np.arange(7))

    t, space, vals = h5dump_attr("mixed.h5", "/", "time")
    assert (t, space) == ("H5T_IEEE_F64LE", "SCALAR") and vals.tolist() == [0.5]
    t, space, vals = h5dump_attr("mixed.h5", "/Eulerian/Parameters", "origin")
    assert t == "H5T_IEEE_F64LE" and space.startswith(f"SIMPLE {{ ( {dim} )")
    np.testing.assert_array_equal(vals, [0.5, 0.25, 0.125][:dim])
    t, _, vals = h5dump_attr("mixed.h5", "/Eulerian/Parameters", "dx")
    np.testing.assert_array_equal(vals, [0.1, 0.2, 0.3][:dim])
    t, _, vals = h5dump_attr("mixed.h5", "/Eulerian/Parameters", "grid_size")
    assert t == "H5T_STD_I64LE"
    np.testing.assert_array_equal(vals, grid)


REPACK = {
    "latest_format": ["-L"],
    "chunked_gzip": ["-l", "/Eulerian/Scalar/s,/Eulerian/Vector/v_0:CHUNK=1x4x5x6", "-f", "GZIP=6"],
}


@needs_tools
@pytest.mark.parametrize("mode", sorted(REPACK))
def test_reader_on_repacked_files(tmp_cwd, mode, rng):
    """Latest-format groups (link messages instead of symbol tables) and chunked, compressed
    datasets read back to the same values."""
    real_t = np.float32
    io, d = save_mixed(3, real_t, rng, "orig.h5")
    subprocess.run([str(H5_BIN / "h5repack"), *REPACK[mode], "orig.h5", "repacked.h5"], check=True,
                   capture_output=True)
    s2, v2 = np.zeros_like(d["s"]), np.zeros_like(d["v"])
    pos2, vf2, sf2 = np.zeros_like(d["pos"]), np.zeros_like(d["vf"]), np.zeros_like(d["sf"])
    io2 = spu.IO(dim=3, real_dtype=real_t)
    io2.define_eulerian_grid(origin=io.eulerian_origin, dx=io.eulerian_dx, grid_size=io.eulerian_grid_size)
    io2.add_as_eulerian_fields_for_io(s=s2, v=v2)
    io2.add_as_lagrangian_fields_for_io(pos2, "body", f=vf2, m=sf2)
    assert io2.load("repacked.h5") == 0.5
    for a, b in ((d["s"], s2), (d["v"], v2), (d["pos"], pos2), (d["vf"], vf2), (d["sf"], sf2)):
        np.testing.assert_array_equal(a, b)


@needs_tools
@pytest.mark.parametrize("dim", [2, 3])
def test_xdmf_references_resolve(tmp_cwd, dim, rng):
    """Every HDF DataItem of the sidecars names a dataset of the stated element count
    (``io.py:429-604``); the Eulerian topology is the z-y-x grid (2D as one z slice)."""
    save_mixed(dim, np.float64, rng)
    tree = h5ls("mixed.h5")
    for xmf in ("mixed_eulerian.xmf", "mixed_body.xmf"):
        root = ET.parse(xmf).getroot()
        items = [e for e in root.iter("DataItem") if e.get("Format") == "HDF"]
        assert items
        for e in items:
            fname, obj = e.text.strip().split(":")
            assert fname == "mixed.h5"
            dims = [int(x) for x in e.get("Dimensions").split()]
            assert int(np.prod(dims)) == int(np.prod(tree[obj])), (xmf, obj)
    topo = ET.parse("mixed_eulerian.xmf").getroot().find(".//Topology")
    want = [int(g) for g in grid_of(dim)]
    assert [int(x) for x in topo.get("Dimensions").split()] == (want if dim == 3 else [1, *want])
    assert ET.parse("mixed_body.xmf").getroot().find(".//Topology").get("TopologyType") == "Polyline"


def test_reference_errors(tmp_cwd, rng):
    """Validation errors of ``io.py:42-45,87-91,104-106,121-126,153-158,181-186,352-385``."""
    with pytest.raises(ValueError, match="Invalid dimension"):
        spu.IO(dim=4)
    io = spu.IO(dim=3)
    with pytest.raises(ValueError, match="Eulerian mesh is not defined"):
        io.add_as_eulerian_fields_for_io(a=np.zeros((4, 4, 4)))
    with pytest.raises(TypeError):
        io.define_eulerian_grid(origin=[0, 0, 0], dx=np.ones(3), grid_size=np.ones(3))
    io.define_eulerian_grid(origin=np.zeros(3), dx=np.ones(3), grid_size=np.array([4, 4, 4]))
    with pytest.raises(ValueError, match="Unable to identify eulerian field type"):
        io.add_as_eulerian_fields_for_io(a=np.zeros((4, 4)))
    with pytest.raises(ValueError, match="2D"):
        io.add_as_lagrangian_fields_for_io(np.zeros(3))
    with pytest.raises(ValueError, match="Invalid lagrangian grid dimension"):
        io.add_as_lagrangian_fields_for_io(np.zeros((2, 5)))
    with pytest.raises(ValueError, match="Unable to identify lagrangian field type"):
        io.add_as_lagrangian_fields_for_io(np.zeros((3, 5)), f=np.zeros((2, 4)))

    a = rng.random((4, 4, 4))
    io = spu.IO(dim=3)
    io.define_eulerian_grid(origin=np.zeros(3), dx=np.ones(3), grid_size=np.array([4, 4, 4]))
    io.add_as_eulerian_fields_for_io(a=a)
    io.save("e.h5")
    for kw, msg in ((dict(origin=np.ones(3)), "origin"), (dict(dx=2 * np.ones(3)), "dx")):
        other = spu.IO(dim=3)
        grid = dict(origin=np.zeros(3), dx=np.ones(3), grid_size=np.array([4, 4, 4]))
        grid.update(kw)
        other.define_eulerian_grid(**grid)
        other.add_as_eulerian_fields_for_io(a=np.zeros_like(a))
        with pytest.raises(ValueError, match=f"Eulerian {msg} does not match"):
            other.load("e.h5")
    other = spu.IO(dim=3)
    other.define_eulerian_grid(origin=np.zeros(3), dx=np.ones(3), grid_size=np.array([4, 4, 4]))
    other.add_as_eulerian_fields_for_io(b=np.zeros_like(a))
    with pytest.raises(ValueError, match="Unable to find scalar field b"):
        other.load("e.h5")
    with pytest.raises(FileNotFoundError):
        other.load("missing.h5")
    with pytest.raises(TypeError):
        bad = spu.IO(dim=3)
        bad.add_as_lagrangian_fields_for_io(np.zeros((3, 2)), f=np.zeros(2, dtype=np.complex64))
        bad.save("c.h5")


def test_eulerian_field_io_from_position_field(tmp_cwd, rng):
    """``EulerianFieldIO`` (``io.py:646-694``): origin per axis (z, y, x), dx from x."""
    nz, ny, nx, dx = 4, 6, 8, 0.125
    z, y, x = (np.arange(n) * dx + dx / 2 for n in (nz, ny, nx))
    pos = np.flipud(np.array(np.meshgrid(z, y, x, indexing="ij")))
    w = rng.random((3, nz, ny, nx))
    io = spu.EulerianFieldIO(position_field=pos, eulerian_fields_dict={"vorticity": w})
    np.testing.assert_array_equal(io.eulerian_origin, [dx / 2] * 3)
    np.testing.assert_array_equal(io.eulerian_dx, [dx] * 3)
    np.testing.assert_array_equal(io.eulerian_grid_size, [nz, ny, nx])
    io.save("f.h5", time=1.0)
    w2 = np.zeros_like(w)
    assert spu.EulerianFieldIO(pos, {"vorticity": w2}).load("f.h5") == 1.0
    np.testing.assert_array_equal(w, w2)


def test_restart_simulation_picks_latest(tmp_cwd, rng):
    """``restart_sim.py:11-37``: the highest-numbered ``sopht_*``/``forcing_grid_*`` pair is
    loaded from the working directory; none raises FileNotFoundError."""
    pos = rng.random((3, 10))
    dX = np.zeros((3, 10))
    w = np.zeros((3, 4, 4, 4))
    flow = spu.IO(dim=3)
    flow.define_eulerian_grid(origin=np.zeros(3), dx=np.ones(3), grid_size=np.array([4, 4, 4]))
    flow.add_as_eulerian_fields_for_io(vorticity=w)
    forcing = spu.IO(dim=3)
    forcing.add_as_lagrangian_fields_for_io(pos, "body", position_mismatch=dX)
    with pytest.raises(FileNotFoundError):
        spu.restart_simulation(None, flow, None, forcing, "restart_data")
    for k, t in ((3, 0.03), (12, 0.12), (7, 0.07)):
        w[...] = k
        dX[...] = -k
        flow.save(f"sopht_{k:04d}.h5", time=t)
        forcing.save(f"forcing_grid_{k:04d}.h5", time=t)
    w[...] = 0
    dX[...] = 0
    assert spu.restart_simulation(None, flow, None, forcing, "restart_data") == 0.12
    assert (w == 12).all() and (dX == -12).all()


def test_h5_tools_and_library_agree_on_version():
    """The tools that pin the layout come from the same HDF5 as the library."""
    if not HAVE_TOOLS:
        pytest.skip("HDF5 command-line tools not installed")
    out = subprocess.run([str(H5_BIN / "h5dump"), "-V"], capture_output=True, text=True).stdout
    major, minor, rel = (int(x) for x in re.search(r"(\d+)\.(\d+)\.(\d+)", out).groups())
    assert sio._h5().sopht_h5_library_version() == major * 10000 + minor * 100 + rel


def test_long_object_paths_round_trip(tmp_cwd):
    """Objects whose paths exceed any fixed buffer (2 kB here) are written, found and read
    back, and a missing one of the same length is reported as missing."""
    name = "g" * 700
    path = f"/{name}/{name}/field"
    data = np.arange(12, dtype=np.float64).reshape(3, 4)
    with sio._File("long.h5", "w") as f:
        f.create_group(f"/{name}")
        f.create_group(f"/{name}/{name}")
        f.write(path, data)
        f.set_attr(path, "time", 0.5)
    with sio._File("long.h5", "r") as f:
        np.testing.assert_array_equal(f.read(path), data)
        assert float(f.attr(path, "time")) == 0.5
        with pytest.raises(KeyError):
            f.read(path + "x")
