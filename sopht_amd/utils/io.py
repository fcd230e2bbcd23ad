"""Snapshot save/load in SophT's HDF5 layout, plus its XDMF sidecars for Paraview.

Mirrors ``IO`` and ``EulerianFieldIO`` of ``sopht/utils/io.py:13-427,646-694`` (same methods,
keyword names, field-type rules and error messages), with the file operations done by the C
layer ``libsopht_amd_h5.so`` over the HDF5 C library instead of h5py. Fields may be numpy
arrays or torch tensors (device tensors are staged through host memory on save and written
back in place on load), so the simulator's device fields are registered directly.

File layout (one file per save), identical to the reference's:

* root attribute ``time``;
* ``/Eulerian/Scalar/<name>`` of shape ``(1, *grid_size)`` per scalar field and
  ``/Eulerian/Vector/<name>_<d>`` of the same shape per vector component;
  ``/Eulerian/Parameters`` with attributes ``origin``, ``dx``, ``grid_size`` (z, y, x order);
* ``/Lagrangian/<grid>/Grid`` of shape ``(N, dim)``, optional ``Connection`` (``arange(N)``),
  ``Scalar/<name>`` of shape ``(N,)`` and ``Vector/<name>`` of shape ``(N, dim)``.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .field import VectorField

_LIB_PATH = Path(__file__).resolve().parent.parent / "libsopht_amd_h5.so"
_MAX_RANK = 8
_ERR_INVALID, _ERR_NOT_FOUND = -1, -2
_CODE_OF = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.int32): 2, np.dtype(np.int64): 3}
_DTYPE_OF = {v: k for k, v in _CODE_OF.items()}

_lib = None


def _h5():
    """The checkpoint C layer, loaded on first use with argument types set."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(f"{_LIB_PATH.name} is not built; run `python -m sopht_amd.build`")
    lib = ctypes.CDLL(str(_LIB_PATH))
    i64, i32, p, cs = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
    pi64, pi32 = ctypes.POINTER(i64), ctypes.POINTER(i32)
    sig = {
        "sopht_h5_last_error": (cs, []),
        "sopht_h5_library_version": (i32, []),
        "sopht_h5_file_create": (i32, [cs, pi64]),
        "sopht_h5_file_open": (i32, [cs, pi64]),
        "sopht_h5_file_close": (i32, [i64]),
        "sopht_h5_group_create": (i32, [i64, cs]),
        "sopht_h5_dataset_write": (i32, [i64, cs, i32, i32, pi64, p]),
        "sopht_h5_dataset_shape": (i32, [i64, cs, pi32, pi64, pi32]),
        "sopht_h5_dataset_read": (i32, [i64, cs, i32, i64, p]),
        "sopht_h5_attr_write": (i32, [i64, cs, cs, i32, i32, pi64, p]),
        "sopht_h5_attr_shape": (i32, [i64, cs, cs, pi32, pi64, pi32]),
        "sopht_h5_attr_read": (i32, [i64, cs, cs, i32, i64, p]),
        "sopht_h5_list": (i32, [i64, p, i64, pi64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _lib = lib
    return lib


SIGNATURES = (
    "sopht_h5_last_error", "sopht_h5_library_version", "sopht_h5_file_create", "sopht_h5_file_open",
    "sopht_h5_file_close", "sopht_h5_group_create", "sopht_h5_dataset_write", "sopht_h5_dataset_shape",
    "sopht_h5_dataset_read", "sopht_h5_attr_write", "sopht_h5_attr_shape", "sopht_h5_attr_read",
    "sopht_h5_list",
)


def _check(rc: int, what: str = "") -> None:
    if rc == 0:
        return
    msg = _h5().sopht_h5_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == _ERR_INVALID:
        raise ValueError(msg)
    if rc == _ERR_NOT_FOUND:
        raise KeyError(msg)
    raise OSError(msg)


def _code(dtype) -> int:
    try:
        return _CODE_OF[np.dtype(dtype)]
    except KeyError:
        raise TypeError(f"unsupported dtype {np.dtype(dtype)} for HDF5 IO "
                        f"(float32, float64, int32, int64)") from None


def _dims(shape):
    d = (ctypes.c_int64 * _MAX_RANK)(*[int(s) for s in shape])
    return d


class _File:
    """An open HDF5 file (the role of ``h5py.File`` inside a ``with`` block)."""

    def __init__(self, name: str, mode: str):
        h = ctypes.c_int64()
        if mode == "w":
            _check(_h5().sopht_h5_file_create(str(name).encode(), ctypes.byref(h)))
        else:
            rc = _h5().sopht_h5_file_open(str(name).encode(), ctypes.byref(h))
            if rc == _ERR_NOT_FOUND:
                raise FileNotFoundError(_h5().sopht_h5_last_error().decode(errors="replace"))
            _check(rc)
        self.id = h.value

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        rc = _h5().sopht_h5_file_close(self.id)
        if exc_type is None:  # a close failure must not hide the error that is propagating
            _check(rc)
        return False

    def create_group(self, path: str) -> None:
        _check(_h5().sopht_h5_group_create(self.id, path.encode()))

    def write(self, path: str, data) -> None:
        a = np.array(data, order="C", copy=None)  # keeps 0-d arrays 0-d (ascontiguousarray does not)
        _check(_h5().sopht_h5_dataset_write(self.id, path.encode(), _code(a.dtype), a.ndim, _dims(a.shape),
                                            a.ctypes.data), path)

    def read(self, path: str) -> np.ndarray:
        rank, dt = ctypes.c_int(), ctypes.c_int()
        dims = (ctypes.c_int64 * _MAX_RANK)()
        _check(_h5().sopht_h5_dataset_shape(self.id, path.encode(), ctypes.byref(rank), dims, ctypes.byref(dt)), path)
        if dt.value not in _DTYPE_OF:
            raise TypeError(f"dataset {path} has a type this reader does not map")
        out = np.empty(tuple(dims[: rank.value]), dtype=_DTYPE_OF[dt.value])
        _check(_h5().sopht_h5_dataset_read(self.id, path.encode(), dt.value, out.size, out.ctypes.data), path)
        return out

    def set_attr(self, obj: str, name: str, value) -> None:
        a = np.array(value, order="C", copy=None)  # a Python float stays a scalar dataspace, as in h5py
        if a.dtype == np.dtype(bool) or a.dtype.kind not in "fi":
            raise TypeError(f"unsupported attribute type {a.dtype}")
        if a.dtype.kind == "i" and a.dtype.itemsize not in (4, 8):
            a = a.astype(np.int64)
        _check(_h5().sopht_h5_attr_write(self.id, obj.encode(), name.encode(), _code(a.dtype), a.ndim,
                                         _dims(a.shape), a.ctypes.data), f"{obj}.attrs[{name}]")

    def attr(self, obj: str, name: str):
        rank, dt = ctypes.c_int(), ctypes.c_int()
        dims = (ctypes.c_int64 * _MAX_RANK)()
        _check(_h5().sopht_h5_attr_shape(self.id, obj.encode(), name.encode(), ctypes.byref(rank), dims,
                                         ctypes.byref(dt)), f"{obj}.attrs[{name}]")
        if dt.value not in _DTYPE_OF:
            raise TypeError(f"attribute {name} of {obj} has a type this reader does not map")
        out = np.empty(tuple(dims[: rank.value]), dtype=_DTYPE_OF[dt.value])
        _check(_h5().sopht_h5_attr_read(self.id, obj.encode(), name.encode(), dt.value, max(out.size, 1),
                                        out.ctypes.data), f"{obj}.attrs[{name}]")
        return out[()] if rank.value == 0 else out

    def keys(self) -> list[str]:
        need = ctypes.c_int64()
        _check(_h5().sopht_h5_list(self.id, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(_h5().sopht_h5_list(self.id, buf, need.value, ctypes.byref(need)))
        text = buf.value.decode()
        return [k for k in text.split("\n") if k]


def _is_tensor(x) -> bool:
    return type(x).__module__.startswith("torch")


def _host(x) -> np.ndarray:
    """A host numpy view/copy of a numpy array or torch tensor (device tensors are copied)."""
    if _is_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _np_dtype(x) -> np.dtype:
    """The numpy dtype of a numpy array or torch tensor (no data is copied)."""
    if _is_tensor(x):
        import torch

        return torch.empty((), dtype=x.dtype).numpy().dtype
    return np.asarray(x).dtype


def _assign(dst, src: np.ndarray) -> None:
    """dst[...] = src for a numpy array or a (device) torch tensor, casting like numpy does."""
    if _is_tensor(dst):
        import torch

        dst.copy_(torch.from_numpy(np.ascontiguousarray(src)))
    else:
        dst[...] = src


def _shape(x) -> tuple:
    return tuple(int(s) for s in x.shape)


class IO:
    r"""Field save and load in the reference's HDF5 hierarchy, with XDMF files for Paraview.

    Mirrors ``sopht/utils/io.py:13-604``::

        /  (attr time)
        |- Eulerian/{Scalar, Vector, Parameters (attrs origin, dx, grid_size)}
        |- Lagrangian/<grid name>/{Grid, [Connection], Scalar, Vector}
    """

    def __init__(self, dim: int, real_dtype: type = np.float64) -> None:
        self.dim = dim
        if self.dim not in (2, 3):
            raise ValueError("Invalid dimension (only 2D and 3D)")
        self.real_dtype = real_dtype
        # io.py:50; the XDMF "Precision" written below is the stored type's byte width.
        self.precision = 8 if real_dtype is np.float32 else 16

        self.eulerian_grid_defined = False
        self.eulerian_fields: dict = {}
        self.eulerian_fields_type: dict = {}

        self.lagrangian_fields: dict = {}
        self.lagrangian_fields_type: dict = {}
        self.lagrangian_grids: dict = {}
        self.lagrangian_fields_with_grid_name: dict = {}
        self.lagrangian_grid_count = 0
        self.lagrangian_grid_connection: dict = {}

    def define_eulerian_grid(self, origin: np.ndarray, dx: np.ndarray, grid_size: np.ndarray) -> None:
        """Eulerian grid mesh: origin, spacing and grid size, each (dim,) in z-y-x order
        (``io.py:64-95``)."""
        if not (isinstance(origin, np.ndarray) and isinstance(dx, np.ndarray) and isinstance(grid_size, np.ndarray)):
            raise TypeError("Invalid input types for Eulerian grid definition")
        self.eulerian_origin = origin
        self.eulerian_dx = dx
        self.eulerian_grid_size = grid_size
        self.eulerian_grid_defined = True

    def add_as_eulerian_fields_for_io(self, **fields_for_io) -> None:
        """Register Eulerian fields by keyword name; scalar ``grid_size`` or vector
        ``(dim, *grid_size)`` shapes (``io.py:97-126``)."""
        if not self.eulerian_grid_defined:
            raise ValueError("Eulerian mesh is not defined!")
        grid = tuple(int(g) for g in self.eulerian_grid_size)
        for field_name, field in fields_for_io.items():
            self.eulerian_fields[field_name] = field
            shape = _shape(field)
            if shape == grid:
                self.eulerian_fields_type[field_name] = "Scalar"
            elif shape == (self.dim, *grid):
                self.eulerian_fields_type[field_name] = "Vector"
            else:
                raise ValueError(
                    f"Unable to identify eulerian field type (scalar / vector) based on field dimension {shape}"
                )

    def add_as_lagrangian_fields_for_io(
        self,
        lagrangian_grid,
        lagrangian_grid_name: str | None = None,
        lagrangian_grid_connect: bool = False,
        **fields_for_io,
    ) -> None:
        """Register a (dim, N) Lagrangian grid and the fields that live on it
        (``io.py:128-186``)."""
        gshape = _shape(lagrangian_grid)
        if len(gshape) != 2:
            raise ValueError("lagrangian grid has to be a 2D (dim, N) array.")
        if gshape[0] != self.dim:
            raise ValueError("Invalid lagrangian grid dimension (only 2D and 3D)")
        if lagrangian_grid_name is None:
            lagrangian_grid_name = f"Lagrangian_grid_{self.lagrangian_grid_count}"
            self.lagrangian_grid_count += 1
        if lagrangian_grid_connect:
            self.lagrangian_grid_connection[lagrangian_grid_name] = np.arange(gshape[1])
        self.lagrangian_grids[lagrangian_grid_name] = lagrangian_grid
        self.lagrangian_fields_with_grid_name[lagrangian_grid_name] = []
        for field_name, field in fields_for_io.items():
            self.lagrangian_fields[field_name] = field
            self.lagrangian_fields_with_grid_name[lagrangian_grid_name].append(field_name)
            shape = _shape(field)
            if shape[0] == gshape[1]:
                self.lagrangian_fields_type[field_name] = "Scalar"
            elif shape == gshape:
                self.lagrangian_fields_type[field_name] = "Vector"
            else:
                raise ValueError(
                    f"Unable to identify lagrangian field type (scalar / vector) based on field dimension {shape}"
                )

    def save(self, h5_file_name: str, time: float = 0.0) -> None:
        """Write every registered field to ``h5_file_name`` (``io.py:188-199``)."""
        self._save(h5_file_name, time)

    def _save(self, h5_file_name: str, time: float = 0.0) -> None:
        """``io.py:201-320``: groups, datasets and attributes as laid out in the module
        docstring, then the XDMF sidecars."""
        grid = tuple(int(g) for g in self.eulerian_grid_size) if self.eulerian_grid_defined else ()
        with _File(h5_file_name, "w") as f:
            f.set_attr("/", "time", time)
            if self.eulerian_grid_defined:
                f.create_group("/Eulerian")
                f.create_group("/Eulerian/Scalar")
                f.create_group("/Eulerian/Vector")
                for field_name, field in self.eulerian_fields.items():
                    field_type = self.eulerian_fields_type[field_name]
                    host = _host(field)
                    if field_type == "Scalar":
                        f.write(f"/Eulerian/Scalar/{field_name}", host.reshape(1, *grid))
                    elif field_type == "Vector":
                        for d in range(self.dim):
                            f.write(f"/Eulerian/Vector/{field_name}_{d}", host[d].reshape(1, *grid))
                    else:
                        raise ValueError("Unsupported eulerian_field_type ('Scalar' and 'Vector' only)")
                f.create_group("/Eulerian/Parameters")
                f.set_attr("/Eulerian/Parameters", "origin", self.eulerian_origin)
                f.set_attr("/Eulerian/Parameters", "dx", self.eulerian_dx)
                f.set_attr("/Eulerian/Parameters", "grid_size", self.eulerian_grid_size)

            f.create_group("/Lagrangian")
            for grid_name, lag_grid in self.lagrangian_grids.items():
                base = f"/Lagrangian/{grid_name}"
                f.create_group(base)
                f.write(f"{base}/Grid", _host(lag_grid).T)
                if grid_name in self.lagrangian_grid_connection:
                    f.write(f"{base}/Connection", self.lagrangian_grid_connection[grid_name])
                f.create_group(f"{base}/Scalar")
                f.create_group(f"{base}/Vector")
                for field_name in self.lagrangian_fields_with_grid_name[grid_name]:
                    field_type = self.lagrangian_fields_type[field_name]
                    host = _host(self.lagrangian_fields[field_name])
                    if field_type == "Scalar":
                        f.write(f"{base}/Scalar/{field_name}", host)
                    elif field_type == "Vector":
                        f.write(f"{base}/Vector/{field_name}", np.moveaxis(host, 0, -1))
                    else:
                        raise ValueError("Unsupported lagrangian_field_type ('Scalar' and 'Vector' only)")

        if self.eulerian_fields:
            self.generate_xdmf_eulerian(h5_file_name=h5_file_name, time=time)
        if self.lagrangian_fields:
            self.generate_xdmf_lagrangian(h5_file_name=h5_file_name, time=time)

    def load(self, h5_file_name: str):
        """Read the registered fields back in place and return the saved time; the Eulerian
        grid parameters must match the file's (``io.py:322-427``)."""
        with _File(h5_file_name, "r") as f:
            keys = f.keys()
            time = f.attr("/", "time")

            if self.eulerian_fields:
                if not self.eulerian_grid_defined:
                    raise ValueError("Eulerian grid is not defined!")
                for field_name, field in self.eulerian_fields.items():
                    field_type = self.eulerian_fields_type[field_name]
                    if field_type == "Scalar":
                        if f"Eulerian/{field_type}/{field_name}" not in keys:
                            raise ValueError(f"Unable to find scalar field {field_name} in loaded file!")
                        _assign(field, f.read(f"/Eulerian/{field_type}/{field_name}")[0])
                    elif field_type == "Vector":
                        for d in range(self.dim):
                            if f"Eulerian/{field_type}/{field_name}_{d}" not in keys:
                                raise ValueError(f"Unable to find vector field {field_name}_{d} in loaded file!")
                            _assign(field[d], f.read(f"/Eulerian/{field_type}/{field_name}_{d}")[0])
                    else:
                        raise ValueError("Unsupported lagrangian_field_type ('Scalar' and 'Vector' only)")
                params = "/Eulerian/Parameters"
                if not np.allclose(self.eulerian_origin, f.attr(params, "origin")):
                    raise ValueError("Eulerian origin does not match loaded file!")
                if not np.allclose(self.eulerian_dx, f.attr(params, "dx")):
                    raise ValueError("Eulerian dx does not match loaded file!")
                if not np.allclose(self.eulerian_grid_size, f.attr(params, "grid_size")):
                    raise ValueError("Eulerian grid size does not match loaded file!")

            if self.lagrangian_fields:
                for grid_name, lag_grid in self.lagrangian_grids.items():
                    base = f"Lagrangian/{grid_name}"
                    if f"{base}/Grid" not in keys:
                        raise ValueError(f"Unable to find grid '{grid_name}' in loaded file!")
                    _assign(lag_grid, f.read(f"/{base}/Grid").T)
                    if f"{base}/Connection" in keys:
                        self.lagrangian_grid_connection[grid_name] = f.read(f"/{base}/Connection")
                    for field_name in self.lagrangian_fields_with_grid_name[grid_name]:
                        field_type = self.lagrangian_fields_type[field_name]
                        path = f"{base}/{field_type}/{field_name}"
                        if field_type not in ("Scalar", "Vector"):
                            raise ValueError("Unsupported lagrangian_field_type ('Scalar' and 'Vector' only)")
                        if path not in keys:
                            raise ValueError(
                                f"Unable to find {field_type.lower()} field {field_name} on "
                                f"grid {grid_name} in loaded file!"
                            )
                        data = f.read(f"/{path}")
                        _assign(self.lagrangian_fields[field_name],
                                data if field_type == "Scalar" else np.moveaxis(data, -1, 0))
        return time

    # XDMF sidecars (``io.py:429-604``): XML descriptions that point Paraview at the datasets.

    def _xdmf_precision(self, dtype) -> int:
        return np.dtype(dtype).itemsize

    @staticmethod
    def _numbers(values) -> str:
        return "    ".join(repr(float(v)) if isinstance(v, (float, np.floating)) else str(int(v)) for v in values)

    @staticmethod
    def _xdmf_document(time: float, body: str) -> str:
        return (
            '<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n'
            '<Xdmf xmlns:xi="http://www.w3.org/2003/XInclude" Version="2.2">\n'
            f'  <Domain>\n    <Grid GridType="Uniform">\n      <Time Value="{time}"/>\n{body}'
            "    </Grid>\n  </Domain>\n</Xdmf>\n"
        )

    @staticmethod
    def _hdf_item(dims: str, number_type: str, precision: int, ref: str) -> str:
        return (f'        <DataItem Dimensions="{dims}" NumberType="{number_type}" Precision="{precision}" '
                f'Format="HDF">{ref}</DataItem>\n')

    def generate_xdmf_eulerian(self, h5_file_name: str, time: float = 0.0) -> None:
        """``<file>_eulerian.xmf``: a 3DCORECTMesh (2D grids as one z slice) with one node
        attribute per scalar field or vector component (``io.py:429-509``)."""
        grid = [int(g) for g in self.eulerian_grid_size]
        origin = [float(v) for v in self.eulerian_origin]
        dx = [float(v) for v in self.eulerian_dx]
        if self.dim == 2:
            grid, origin, dx = [1, *grid], [0.0, *origin], [0.0, *dx]
        dims = self._numbers(grid)
        body = (f'      <Topology TopologyType="3DCORECTMesh" Dimensions="{dims}"/>\n'
                '      <Geometry GeometryType="ORIGIN_DXDYDZ">\n'
                f'        <DataItem Name="Origin" Dimensions="3" NumberType="Float" Precision="8" '
                f'Format="XML">{self._numbers(origin)}</DataItem>\n'
                f'        <DataItem Name="Spacing" Dimensions="3" NumberType="Float" Precision="8" '
                f'Format="XML">{self._numbers(dx)}</DataItem>\n'
                "      </Geometry>\n")
        for field_name, field in self.eulerian_fields.items():
            field_type = self.eulerian_fields_type[field_name]
            prec = self._xdmf_precision(_np_dtype(field))
            names = [field_name] if field_type == "Scalar" else [f"{field_name}_{d}" for d in range(self.dim)]
            for name in names:
                body += (f'      <Attribute Name="{name}" Active="1" AttributeType="Scalar" Center="Node">\n'
                         + self._hdf_item(dims, "Float", prec, f"{h5_file_name}:/Eulerian/{field_type}/{name}")
                         + "      </Attribute>\n")
        Path(h5_file_name.replace(".h5", "_eulerian.xmf")).write_text(self._xdmf_document(time, body))

    def generate_xdmf_lagrangian(self, h5_file_name: str, time: float) -> None:
        """``<file>_<grid>.xmf`` per Lagrangian grid: Polyvertex (or Polyline with a
        Connection dataset) geometry and one node attribute per field (``io.py:511-604``)."""
        for grid_name, lag_grid in self.lagrangian_grids.items():
            n = _shape(lag_grid)[1]
            grid_dims = self._numbers([n, self.dim])
            geo_prec = self._xdmf_precision(_np_dtype(lag_grid))
            if grid_name in self.lagrangian_grid_connection:
                topology = ('      <Topology TopologyType="Polyline">\n'
                            + self._hdf_item(f"1    {n}", "Int", 8, f"{h5_file_name}:/Lagrangian/{grid_name}/Connection")
                            + "      </Topology>\n")
            else:
                topology = f'      <Topology TopologyType="Polyvertex" NumberOfElements="{n}"/>\n'
            body = (topology
                    + f'      <Geometry GeometryType="{"XYZ" if self.dim == 3 else "XY"}">\n'
                    + self._hdf_item(grid_dims, "Float", geo_prec, f"{h5_file_name}:/Lagrangian/{grid_name}/Grid")
                    + "      </Geometry>\n")
            for field_name in self.lagrangian_fields_with_grid_name[grid_name]:
                field_type = self.lagrangian_fields_type[field_name]
                field = self.lagrangian_fields[field_name]
                dims = str(n) if field_type == "Scalar" else grid_dims
                prec = self._xdmf_precision(_np_dtype(field))
                body += (f'      <Attribute Name="{field_name}" Active="1" AttributeType="{field_type}" '
                         'Center="Node">\n'
                         + self._hdf_item(dims, "Float", prec,
                                          f"{h5_file_name}:/Lagrangian/{grid_name}/{field_type}/{field_name}")
                         + "      </Attribute>\n")
            Path(h5_file_name.replace(".h5", f"_{grid_name}.xmf")).write_text(self._xdmf_document(time, body))


class EulerianFieldIO(IO):
    """IO for the fields of one Eulerian grid, defined from its position field
    (``io.py:646-694``): origin = the per-axis minimum (z, y, x), dx from the first two x
    coordinates, grid size from the field's shape."""

    def __init__(self, position_field, eulerian_fields_dict: dict) -> None:
        pos = _host(position_field)
        grid_dim = pos.shape[0]
        grid_size = pos.shape[1:]
        super().__init__(dim=grid_dim, real_dtype=pos.dtype.type)
        if grid_dim == 2:
            origin = np.array([pos[VectorField.y_axis_idx()].min(), pos[VectorField.x_axis_idx()].min()])
        elif grid_dim == 3:
            origin = np.array([pos[VectorField.z_axis_idx()].min(), pos[VectorField.y_axis_idx()].min(),
                               pos[VectorField.x_axis_idx()].min()])
        else:
            raise ValueError("Position field of invalid shape.")
        x = pos[VectorField.x_axis_idx()].reshape(-1)
        dx = x[1] - x[0]
        self.define_eulerian_grid(origin=origin, dx=dx * np.ones(grid_dim), grid_size=np.array(grid_size))
        self.add_as_eulerian_fields_for_io(**eulerian_fields_dict)
