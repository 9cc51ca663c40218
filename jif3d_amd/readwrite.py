"""jif3D's netCDF data and model files, read and written through jif3d_amd.netcdf.

Data files (jif3D/Gravity/ReadWriteGravityData.cpp:20-235,
jif3D/Magnetics/ReadWriteMagneticData.cpp): dimension StationNumber, variables MeasPosX,
MeasPosY, MeasPosZ (m) and
  scalar gravity    Scalar_gravity, dGz (m/s2)
  FTG               Uxx, Uxy, Uxz, Uyx, Uyy, Uyz, Uzx, Uzy, Uzz and d<name> (1/s2)
  total field       T, dT (nT)
  three components  Bx, By, Bz, dBx, dBy, dBz (nT)
Model files (jif3D/ModelBase/NetCDFModelTools.cpp:96-279, ThreeDGravityModel.cpp:75-165,
ThreeDSusceptibilityModel.h:112-175): cell boundary coordinates Northing / Easting / Depth
(n values: the far boundary of each cell) with <axis>_Origin and <axis>_bnds, nbound, the
cell values as (Depth, Easting, Northing) with a units attribute, and the background
layers bg_thicknesses, bg_densities / bg_susceptibilities on dimension bg_layers.

Reading accepts classic and netCDF-4 files; writing produces classic (CDF-2) files, which
netCDF-C and therefore jif3D read.  Missing error variables read as zeros, as in the
reference.
"""
import numpy as np

from .model import (ThreeDGravityModel, ThreeDSusceptibilityModel, ThreeDMagnetizationModel,
                    ScalarGravityData, TensorGravityData, TotalFieldMagneticData,
                    ThreeComponentMagneticData)
from .netcdf import Dataset, write_classic

TENSOR_NAMES = ("Uxx", "Uxy", "Uxz", "Uyx", "Uyy", "Uyz", "Uzx", "Uzy", "Uzz")
MAG_NAMES = ("Bx", "By", "Bz")
DENSITY_UNIT, OLD_DENSITY_UNIT = "kg/m3", "g/cm3"
SUSCEPTIBILITY_UNIT = " "


# ---------------------------------------------------------------------------
# measurements
# ---------------------------------------------------------------------------
def _positions(ds):
    x, y, z = (np.asarray(ds[n].read(), dtype=np.float64).ravel()
               for n in ("MeasPosX", "MeasPosY", "MeasPosZ"))
    if not (x.size == y.size == z.size):
        raise ValueError("measurement coordinates differ in length")
    return x, y, z


def _component_data(ds, names, nsites):
    k = len(names)
    data = np.zeros(nsites * k)
    err = np.zeros(nsites * k)
    for i, nm in enumerate(names):
        data[i::k] = np.asarray(ds[nm].read(), dtype=np.float64).ravel()
        if "d" + nm in ds:
            err[i::k] = np.asarray(ds["d" + nm].read(), dtype=np.float64).ravel()
    return data, err


def _read_data(path, names, data_cls, error_name=None):
    ds = Dataset(path)
    x, y, z = _positions(ds)
    if error_name is None:
        data, err = _component_data(ds, names, x.size)
    else:
        data = np.asarray(ds[names[0]].read(), dtype=np.float64).ravel()
        err = (np.asarray(ds[error_name].read(), dtype=np.float64).ravel()
               if error_name in ds else np.zeros(data.size))
    d = data_cls()
    d.SetMeasurementPoints(x, y, z)
    d.SetDataAndErrors(data, err)
    return d


def _write_data(path, data, columns, unit):
    x, y, z = data.positions
    ns = x.size
    variables = {"MeasPosX": (["StationNumber"], x, {"units": "m"}),
                 "MeasPosY": (["StationNumber"], y, {"units": "m"}),
                 "MeasPosZ": (["StationNumber"], z, {"units": "m"})}
    for name, values in columns:
        variables[name] = (["StationNumber"], np.asarray(values, dtype=np.float64),
                           {"units": unit})
    write_classic(path, {"StationNumber": ns}, variables)


def ReadScalarGravityMeasurements(path):
    """ReadWriteGravityData.cpp:127-151: positions, Scalar_gravity, dGz (zeros if absent)."""
    return _read_data(path, ("Scalar_gravity",), ScalarGravityData, error_name="dGz")


def SaveScalarGravityMeasurements(path, data):
    d, e = data.GetData(), data.GetErrors()
    _write_data(path, data, [("Scalar_gravity", d), ("dGz", e if e.size else np.zeros(d.size))],
                "m/s2")


def ReadTensorGravityMeasurements(path):
    """ReadWriteGravityData.cpp:162-199: nine components per site, row-major."""
    return _read_data(path, TENSOR_NAMES, TensorGravityData)


def SaveTensorGravityMeasurements(path, data):
    d, e = data.GetData(), data.GetErrors()
    e = e if e.size else np.zeros(d.size)
    cols = [(nm, d[i::9]) for i, nm in enumerate(TENSOR_NAMES)]
    cols += [("d" + nm, e[i::9]) for i, nm in enumerate(TENSOR_NAMES)]
    _write_data(path, data, cols, "1/s2")


def ReadTotalFieldMagneticMeasurements(path):
    return _read_data(path, ("T",), TotalFieldMagneticData, error_name="dT")


def SaveTotalFieldMagneticMeasurements(path, data):
    d, e = data.GetData(), data.GetErrors()
    _write_data(path, data, [("T", d), ("dT", e if e.size else np.zeros(d.size))], "nT")


def ReadMagneticComponentMeasurements(path):
    return _read_data(path, MAG_NAMES, ThreeComponentMagneticData)


def SaveMagneticComponentMeasurements(path, data):
    d, e = data.GetData(), data.GetErrors()
    e = e if e.size else np.zeros(d.size)
    cols = [(nm, d[i::3]) for i, nm in enumerate(MAG_NAMES)]
    cols += [("d" + nm, e[i::3]) for i, nm in enumerate(MAG_NAMES)]
    _write_data(path, data, cols, "nT")


def IdentifyGravityDatafileType(path):
    """'scalar', 'ftg' or 'unknown' (ReadWriteGravityData.cpp:68-84)."""
    ds = Dataset(path)
    if "Scalar_gravity" in ds:
        return "scalar"
    if "Uxx" in ds:
        return "ftg"
    return "unknown"


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------
def _read_coordinates(ds, name):
    if name not in ds.dimensions or ds.dimensions[name] == 0:
        raise ValueError(f"Cell coordinate is empty: {name}")
    n = ds.dimensions[name]
    c = np.empty(n + 1)
    c[1:] = np.asarray(ds[name].read(), dtype=np.float64).ravel()[:n]
    origin = name + "_Origin"
    c[0] = float(np.asarray(ds[origin].read()).ravel()[0]) if origin in ds else 0.0
    if np.any(np.diff(c) < 0):
        raise ValueError(f"Cell coordinates in netcdf file are not in increasing order: {name}")
    return c


def _read_cells(ds, name, unit):
    if name not in ds:
        raise ValueError(f"Cannot read variable: {name}")
    var = ds[name]
    if unit is not None and var.attrs.get("units") != unit:
        raise ValueError("Units in file do not match expected units !")
    xc, yc, zc = (_read_coordinates(ds, a) for a in ("Northing", "Easting", "Depth"))
    nx, ny, nz = xc.size - 1, yc.size - 1, zc.size - 1
    vals = np.asarray(var.read(), dtype=np.float64).reshape(nz, ny, nx)
    return xc, yc, zc, np.transpose(vals, (2, 1, 0)).copy()  # [x][y][z]


def _set_mesh(model, xc, yc, zc):
    model.SetMeshSize(xc.size - 1, yc.size - 1, zc.size - 1)
    model.SetOrigin(xc[0], yc[0], zc[0])
    model.SetXCellSizes(np.diff(xc))
    model.SetYCellSizes(np.diff(yc))
    model.SetZCellSizes(np.diff(zc))


def _model_file_vars(model, cells):
    """Dimensions and variables of Write3DModelToNetCDF; cells: [(name, unit, long name,
    values[x][y][z])]."""
    dims = {"nbound": 2}
    variables = {"nbound": (["nbound"], np.array([1, 2], dtype=np.int32),
                            {"long_name": "Boundary index variable"})}
    for axis, c in zip(("Northing", "Easting", "Depth"), model.coords):
        c = np.asarray(c, dtype=np.float64)
        dims[axis] = c.size - 1
        variables[axis] = ([axis], c[1:], {"units": "m", "long_name": axis + " coordinate",
                                           "bounds": axis + "_bnds"})
        bnds = np.stack([c[:-1], c[1:]], axis=1)
        variables[axis + "_bnds"] = ([axis, "nbound"], bnds, {"units": "m"})
        variables[axis + "_Origin"] = ([], np.array([c[0]]), {})
    for name, unit, long_name, values in cells:
        v = np.transpose(np.asarray(values, dtype=np.float64).reshape(model.nx, model.ny, model.nz),
                         (2, 1, 0))
        variables[name] = (["Depth", "Easting", "Northing"], v,
                           {"units": unit, "long_name": long_name})
    return dims, variables


def _add_background(dims, variables, model, name, long_name, unit, values):
    if len(values):
        dims["bg_layers"] = len(values)
        variables[name] = (["bg_layers"], np.asarray(values, dtype=np.float64),
                           {"long_name": long_name, "units": unit})
        variables["bg_thicknesses"] = (["bg_layers"],
                                       np.asarray(model.bg_thicknesses, dtype=np.float64),
                                       {"long_name": "Background Thicknesses", "units": "m"})


def WriteGravityModel(path, model):
    """ThreeDGravityModel::WriteNetCDF (ThreeDGravityModel.cpp:75-103)."""
    dims, variables = _model_file_vars(model, [("Density", DENSITY_UNIT, "Density",
                                                 model.GetDensities())])
    _add_background(dims, variables, model, "bg_densities", "Background Densities",
                    DENSITY_UNIT, model.GetBackgroundDensities())
    write_classic(path, dims, variables, {"Conventions": "CF-1.3"})


def ReadGravityModel(path):
    """ThreeDGravityModel::ReadNetCDF (ThreeDGravityModel.cpp:105-165): densities in
    g/cm3 are converted to kg/m3."""
    ds = Dataset(path)
    unit = ds["Density"].attrs.get("units") if "Density" in ds else None
    factor = 1000.0 if unit == OLD_DENSITY_UNIT else 1.0
    xc, yc, zc, vals = _read_cells(ds, "Density", OLD_DENSITY_UNIT if factor != 1.0
                                   else DENSITY_UNIT)
    m = ThreeDGravityModel()
    _set_mesh(m, xc, yc, zc)
    m.SetDensities(factor * vals.ravel())
    if "bg_layers" in ds.dimensions:
        m.SetBackgroundThicknesses(np.asarray(ds["bg_thicknesses"].read(), dtype=np.float64))
        m.SetBackgroundDensities(factor * np.asarray(ds["bg_densities"].read(), dtype=np.float64))
    return m


def WriteSusceptibilityModel(path, model):
    dims, variables = _model_file_vars(model, [("Susceptibility", SUSCEPTIBILITY_UNIT,
                                                 "Susceptibility", model.GetSusceptibilities())])
    _add_background(dims, variables, model, "bg_susceptibilities",
                    "Background Susceptibilities", SUSCEPTIBILITY_UNIT,
                    model.GetBackgroundSusceptibilities())
    write_classic(path, dims, variables, {"Conventions": "CF-1.3"})


def ReadSusceptibilityModel(path):
    ds = Dataset(path)
    xc, yc, zc, vals = _read_cells(ds, "Susceptibility", SUSCEPTIBILITY_UNIT)
    m = ThreeDSusceptibilityModel()
    _set_mesh(m, xc, yc, zc)
    m.SetSusceptibilities(vals.ravel())
    if "bg_layers" in ds.dimensions:
        m.SetBackgroundThicknesses(np.asarray(ds["bg_thicknesses"].read(), dtype=np.float64))
        m.SetBackgroundSusceptibilities(np.asarray(ds["bg_susceptibilities"].read(),
                                                   dtype=np.float64))
    return m


def WriteMagnetizationModel(path, model):
    """ThreeDMagnetizationModel::WriteNetCDF (ThreeDMagnetizationModel.cpp:28-98)."""
    n = model.ncells
    p = model.GetModelParameters()
    cells = [(nm, "A/m", f"{axis}-component magnetization", p[i * n:(i + 1) * n])
             for i, (nm, axis) in enumerate((("Mx", "X"), ("My", "Y"), ("Mz", "Z")))]
    dims, variables = _model_file_vars(model, cells)
    write_classic(path, dims, variables, {"Conventions": "CF-1.3"})


def ReadMagnetizationModel(path):
    ds = Dataset(path)
    xc, yc, zc, mx = _read_cells(ds, "Mx", "A/m")
    shape = (zc.size - 1, yc.size - 1, xc.size - 1)
    my, mz = (np.transpose(np.asarray(ds[nm].read(), dtype=np.float64).reshape(shape),
                           (2, 1, 0)) for nm in ("My", "Mz"))
    m = ThreeDMagnetizationModel()
    _set_mesh(m, xc, yc, zc)
    m.SetMagnetization(mx.ravel(), my.ravel(), mz.ravel())
    return m
