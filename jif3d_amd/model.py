"""Model and survey containers for the prism path, mirroring jif3D's own classes.

* ThreeDModelBase / ThreeDGravityModel / ThreeDSusceptibilityModel /
  ThreeDMagnetizationModel follow jif3D/ModelBase/ThreeDModelBase.h:44-286,
  jif3D/Gravity/ThreeDGravityModel.h:69-202,
  jif3D/Magnetics/ThreeDSusceptibilityModel.h:29-110 and
  jif3D/Magnetics/ThreeDMagnetizationModel.h: a rectilinear mesh given by cell
  sizes, n+1 corner coordinates per axis computed as a running sum from the
  origin (jif3D/ModelBase/ThreeDModelBase.cpp:96-113), cell values stored with z
  fastest (offset = (xi*ny + yi)*nz + zi, ThreeDModelBase.h:157-168) and a 1-D
  layered background whose values are appended to the model vector
  (ThreeDGravityModel.h:91-104).
* GeneralData / ScalarGravityData / TensorGravityData / TotalFieldMagneticData
  hold measurement positions (jif3D/DataBase/GeneralData.h:95-108).  z is
  positive down; sites above the surface have z < 0.  SI units: densities in
  kg/m^3, susceptibilities in SI, fields in nT.
"""
import numpy as np


class ThreeDModelBase:
    def __init__(self):
        self._sizes = [np.zeros(0), np.zeros(0), np.zeros(0)]
        self._coords = [np.zeros(1), np.zeros(1), np.zeros(1)]
        self._data = np.zeros((0, 0, 0))

    # mesh -----------------------------------------------------------------
    def SetMeshSize(self, nx, ny, nz):
        self._sizes = [np.zeros(nx), np.zeros(ny), np.zeros(nz)]
        origin = [c[0] if len(c) else 0.0 for c in self._coords]
        self._coords = [np.full(n + 1, o) for n, o in zip((nx, ny, nz), origin)]
        self._data = np.zeros((nx, ny, nz))

    def _set_sizes(self, axis, sizes):
        sizes = np.asarray(sizes, dtype=np.float64).copy()
        origin = self._coords[axis][0] if len(self._coords[axis]) else 0.0
        coords = np.empty(len(sizes) + 1)
        coords[0] = origin
        s = origin
        for i, v in enumerate(sizes):  # running sum, ThreeDModelBase.cpp:104-110
            s += v
            coords[i + 1] = s
        self._sizes[axis] = sizes
        self._coords[axis] = coords

    def SetXCellSizes(self, s):
        self._set_sizes(0, s)

    def SetYCellSizes(self, s):
        self._set_sizes(1, s)

    def SetZCellSizes(self, s):
        self._set_sizes(2, s)

    def SetOrigin(self, x, y, z):
        """Shift every corner coordinate as (val + new) - old (ThreeDModelBase.cpp:158-176)."""
        for axis, o in enumerate((x, y, z)):
            c = self._coords[axis]
            if len(c) == 0:
                c = np.zeros(1)
            old = c[0]
            self._coords[axis] = (c + o) - old

    def GetXCellSizes(self):
        return self._sizes[0]

    def GetYCellSizes(self):
        return self._sizes[1]

    def GetZCellSizes(self):
        return self._sizes[2]

    def GetXCoordinates(self):
        return self._coords[0]

    def GetYCoordinates(self):
        return self._coords[1]

    def GetZCoordinates(self):
        return self._coords[2]

    def GetXOrigin(self):
        return float(self._coords[0][0])

    def GetYOrigin(self):
        return float(self._coords[1][0])

    def GetZOrigin(self):
        return float(self._coords[2][0])

    @property
    def nx(self):
        return len(self._sizes[0])

    @property
    def ny(self):
        return len(self._sizes[1])

    @property
    def nz(self):
        return len(self._sizes[2])

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    @property
    def coords(self):
        return tuple(self._coords)

    @property
    def sizes(self):
        return tuple(self._sizes)

    def GetNModelElements(self):
        return self.ncells

    def IndexToOffset(self, xi, yi, zi):
        return self.nz * (xi * self.ny + yi) + zi

    def OffsetToIndex(self, offset):
        zi = offset % self.nz
        xi = (offset - zi) // self.nz
        yi = xi % self.ny
        xi = (xi - yi) // self.ny
        return xi, yi, zi

    def CheckConsistency(self):
        """Mirror of ThreeDGravMagCalculator::CheckModelConsistency (ThreeDGravMagCalculator.h:122-146)."""
        shape = self._data.shape
        if shape != (self.nx, self.ny, self.nz):
            raise ValueError("Model dimensions do not match the cell-size specification")


class _LayeredModel(ThreeDModelBase):
    def __init__(self):
        super().__init__()
        self.bg_values = np.zeros(0)
        self.bg_thicknesses = np.zeros(0)

    def SetBackgroundThicknesses(self, t):
        self.bg_thicknesses = np.asarray(t, dtype=np.float64).copy()

    def GetBackgroundThicknesses(self):
        return self.bg_thicknesses

    def GetModelParameters(self):
        return np.concatenate([self._data.ravel(), self.bg_values])

    def GetNModelParm(self):
        return self.ncells + len(self.bg_values)

    def SetModelParameters(self, m):
        m = np.asarray(m, dtype=np.float64)
        self._data = m[: self.ncells].reshape(self.nx, self.ny, self.nz).copy()
        self.bg_values = m[self.ncells:].copy()


class ThreeDGravityModel(_LayeredModel):
    """Densities in kg/m^3 on the mesh plus background layer densities."""

    def SetDensities(self, d):
        self._data = np.asarray(d, dtype=np.float64).reshape(self.nx, self.ny, self.nz).copy()

    def GetDensities(self):
        return self._data

    def SetBackgroundDensities(self, d):
        self.bg_values = np.asarray(d, dtype=np.float64).copy()

    def GetBackgroundDensities(self):
        return self.bg_values


class ThreeDSusceptibilityModel(_LayeredModel):
    """Susceptibilities (SI) on the mesh plus background layer susceptibilities."""

    def SetSusceptibilities(self, d):
        self._data = np.asarray(d, dtype=np.float64).reshape(self.nx, self.ny, self.nz).copy()

    def GetSusceptibilities(self):
        return self._data

    def SetBackgroundSusceptibilities(self, d):
        self.bg_values = np.asarray(d, dtype=np.float64).copy()

    def GetBackgroundSusceptibilities(self):
        return self.bg_values


class ThreeDMagnetizationModel(ThreeDModelBase):
    """Three magnetisation components per cell (A/m); model vector [Mx; My; Mz]."""

    def __init__(self):
        super().__init__()
        self.bg_thicknesses = np.zeros(0)
        self._my = np.zeros((0, 0, 0))
        self._mz = np.zeros((0, 0, 0))

    def SetMeshSize(self, nx, ny, nz):
        super().SetMeshSize(nx, ny, nz)
        self._my = np.zeros((nx, ny, nz))
        self._mz = np.zeros((nx, ny, nz))

    def SetMagnetization(self, mx, my, mz):
        shp = (self.nx, self.ny, self.nz)
        self._data = np.asarray(mx, dtype=np.float64).reshape(shp).copy()
        self._my = np.asarray(my, dtype=np.float64).reshape(shp).copy()
        self._mz = np.asarray(mz, dtype=np.float64).reshape(shp).copy()

    def GetModelParameters(self):
        return np.concatenate([self._data.ravel(), self._my.ravel(), self._mz.ravel()])

    def GetNModelParm(self):
        return 3 * self.ncells

    def SetModelParameters(self, m):
        m = np.asarray(m, dtype=np.float64)
        n = self.ncells
        self.SetMagnetization(m[:n], m[n:2 * n], m[2 * n:3 * n])


class GeneralData:
    """Measurement positions and optionally data with errors
    (jif3D/DataBase/GeneralData.h:19-142)."""

    def __init__(self):
        self._x, self._y, self._z = [], [], []
        self._data = np.zeros(0)
        self._errors = np.zeros(0)

    def GetData(self):
        return self._data

    def GetErrors(self):
        return self._errors

    def SetDataAndErrors(self, d, e):
        d, e = np.asarray(d, dtype=np.float64), np.asarray(e, dtype=np.float64)
        if d.shape != e.shape:
            raise ValueError("Data and Errors do not have the same size !")
        self._data, self._errors = d.copy(), e.copy()

    def SetErrors(self, e):
        e = np.asarray(e, dtype=np.float64)
        if e.shape != self._data.shape:
            raise ValueError("Data and Errors do not have the same size !")
        self._errors = e.copy()

    def AddMeasurementPoint(self, x, y, z):
        self._x.append(float(x))
        self._y.append(float(y))
        self._z.append(float(z))

    def SetMeasurementPoints(self, x, y, z):
        self._x = list(np.asarray(x, dtype=np.float64))
        self._y = list(np.asarray(y, dtype=np.float64))
        self._z = list(np.asarray(z, dtype=np.float64))

    def CopyMeasurementConfiguration(self, other):
        self._x, self._y, self._z = list(other._x), list(other._y), list(other._z)

    def GetMeasPosX(self):
        return np.asarray(self._x, dtype=np.float64)

    def GetMeasPosY(self):
        return np.asarray(self._y, dtype=np.float64)

    def GetMeasPosZ(self):
        return np.asarray(self._z, dtype=np.float64)

    @property
    def nsites(self):
        return len(self._x)

    @property
    def positions(self):
        return self.GetMeasPosX(), self.GetMeasPosY(), self.GetMeasPosZ()


class ScalarGravityData(GeneralData):
    pass


class TensorGravityData(GeneralData):
    pass


class TotalFieldMagneticData(GeneralData):
    pass


class ThreeComponentMagneticData(GeneralData):
    pass
