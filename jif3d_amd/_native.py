"""ctypes binding of libjif3d_amd.so (C ABI in include/jif3d_gravmag.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libjif3d_amd.so")

API_VERSION = 1  # JGM_API_VERSION of include/jif3d_gravmag.h this binding follows
JGM_OK = 0
(JGM_ERR_INVALID_ARGUMENT, JGM_ERR_NOT_READY, JGM_ERR_DEVICE, JGM_ERR_OUT_OF_MEMORY,
 JGM_ERR_UNSUPPORTED) = 1, 2, 3, 4, 5
KIND_SCALAR, KIND_TENSOR, KIND_MAGNETIC, KIND_SCALAR_TENSOR = 1, 2, 3, 4
KIND_MAGNETIC_GRADIENT, KIND_MAGNETIZATION = 5, 6
FP64, MIXED = 0, 1
TF_NONE, TF_FTG_INVARIANT, TF_TOTAL_FIELD, TF_TOTAL_FIELD_ANOMALY = 0, 1, 2, 3
GEMM_FP32, GEMM_BF16, GEMM_BF16_REFINED, GEMM_FP64 = 0, 1, 2, 3
RAW_PER_SITE = {KIND_SCALAR: 1, KIND_TENSOR: 9, KIND_MAGNETIC: 3, KIND_SCALAR_TENSOR: 10,
                KIND_MAGNETIC_GRADIENT: 1, KIND_MAGNETIZATION: 3}

EXPORTS = [
    "jgm_last_error", "jgm_api_version", "jgm_raw_per_site", "jgm_create", "jgm_destroy",
    "jgm_set_stream", "jgm_synchronize", "jgm_set_mesh", "jgm_set_background", "jgm_set_sites",
    "jgm_set_field", "jgm_set_transform", "jgm_set_mixed_threshold", "jgm_set_mixed_far_radius",
    "jgm_data_per_site",
    "jgm_n_params", "jgm_forward", "jgm_forward_host", "jgm_forward_noalloc", "jgm_adjoint",
    "jgm_adjoint_host",
    "jgm_forward_batch", "jgm_adjoint_batch", "jgm_forward_batch_host", "jgm_adjoint_batch_host",
    "jgm_storedj_forward_batch", "jgm_storedj_gradient_batch", "jgm_storedj_forward_batch_host",
    "jgm_storedj_gradient_batch_host", "jgm_storedj_build_host", "jgm_storedj_host_resident",
    "jgm_sensitivity_rows", "jgm_model_size", "jgm_storedj_build", "jgm_storedj_rows",
    "jgm_storedj_apply",
    "jgm_storedj_apply_t", "jgm_storedj_forward", "jgm_storedj_gradient",
    "jgm_storedj_forward_host", "jgm_storedj_gradient_host", "jgm_storedj_release", "jgm_device_alloc", "jgm_device_free",
    "jgm_copy_to_device", "jgm_copy_to_host", "jgm_copy_device", "jgm_profile_enable", "jgm_profile_reset",
    "jgm_profile_count", "jgm_profile_get", "jgm_storedj_build_fp64", "jgm_storedj_apply_f64",
    "jgm_storedj_apply_t_f64", "jgm_set_mixed_check", "jgm_mixed_check_report",
    "jgm_storedj_build_file", "jgm_storedj_file_backed",
]


class JgmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"jif3d_amd error {status}: {msg}")
        self.status = status


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(
            f"{LIBPATH} is missing: build it with `python -m jif3d_amd.build` "
            "(the prism kernels have no CPU fallback)")
    lib = ctypes.CDLL(LIBPATH)
    vp, sz, d, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
    pd = ctypes.POINTER(ctypes.c_double)
    lib.jgm_last_error.restype = ctypes.c_char_p
    lib.jgm_last_error.argtypes = []
    lib.jgm_api_version.restype = i
    lib.jgm_raw_per_site.restype = sz
    lib.jgm_raw_per_site.argtypes = [i]
    sigs = {
        "jgm_create": [i, vp, ctypes.POINTER(vp)],
        "jgm_destroy": [vp],
        "jgm_set_stream": [vp, vp],
        "jgm_synchronize": [vp],
        "jgm_set_mesh": [vp, sz, sz, sz, pd, pd, pd, pd, pd, pd],
        "jgm_set_background": [vp, sz, pd],
        "jgm_set_sites": [vp, sz, pd, pd, pd],
        "jgm_set_field": [vp, d, d, d],
        "jgm_set_transform": [vp, i, d, d],
        "jgm_set_mixed_threshold": [vp, d],
        "jgm_set_mixed_far_radius": [vp, d],
        "jgm_forward": [vp, i, i, vp, vp],
        "jgm_forward_host": [vp, i, i, pd, pd],
        "jgm_forward_noalloc": [vp, i, i, vp, vp],
        "jgm_adjoint": [vp, i, i, vp, vp, vp],
        "jgm_adjoint_host": [vp, i, i, pd, pd, pd],
        "jgm_forward_batch": [vp, i, i, sz, vp, sz, vp, sz],
        "jgm_adjoint_batch": [vp, i, i, sz, vp, sz, vp, sz, vp, sz],
        "jgm_forward_batch_host": [vp, i, i, sz, pd, pd],
        "jgm_adjoint_batch_host": [vp, i, i, sz, pd, pd, pd],
        "jgm_sensitivity_rows": [vp, i, sz, sz, vp],
        "jgm_storedj_build": [vp, i, i],
        "jgm_storedj_build_host": [vp, i, i, sz],
        "jgm_storedj_apply": [vp, i, vp, sz, vp],
        "jgm_storedj_apply_t": [vp, i, vp, sz, vp],
        "jgm_storedj_forward": [vp, i, vp, vp],
        "jgm_storedj_gradient": [vp, i, vp, vp, vp],
        "jgm_storedj_forward_host": [vp, i, pd, pd],
        "jgm_storedj_gradient_host": [vp, i, pd, pd, pd],
        "jgm_storedj_forward_batch": [vp, i, sz, vp, sz, vp, sz],
        "jgm_storedj_gradient_batch": [vp, i, sz, vp, sz, vp, sz, vp, sz],
        "jgm_storedj_forward_batch_host": [vp, i, sz, pd, pd],
        "jgm_storedj_gradient_batch_host": [vp, i, sz, pd, pd, pd],
        "jgm_storedj_release": [vp],
        "jgm_storedj_build_fp64": [vp, i],
        "jgm_storedj_build_file": [vp, i, i, sz, ctypes.c_char_p],
        "jgm_set_mixed_check": [vp, sz, sz, d],
        "jgm_mixed_check_report": [vp, pd, pd, ctypes.POINTER(sz)],
        "jgm_storedj_apply_f64": [vp, vp, sz, vp],
        "jgm_storedj_apply_t_f64": [vp, vp, sz, vp],
        "jgm_device_alloc": [vp, sz, ctypes.POINTER(vp)],
        "jgm_device_free": [vp, vp],
        "jgm_copy_to_device": [vp, vp, vp, sz],
        "jgm_copy_to_host": [vp, vp, vp, sz],
        "jgm_copy_device": [vp, vp, vp, sz],
        "jgm_profile_enable": [vp, i],
        "jgm_profile_reset": [vp],
        "jgm_profile_get": [vp, sz, ctypes.POINTER(ctypes.c_char_p), pd,
                            ctypes.POINTER(sz)],
    }
    for name, args in sigs.items():
        f = getattr(lib, name)
        f.restype = i
        f.argtypes = args
    lib.jgm_data_per_site.restype = sz
    lib.jgm_data_per_site.argtypes = [vp, i]
    lib.jgm_n_params.restype = sz
    lib.jgm_n_params.argtypes = [vp]
    lib.jgm_model_size.restype = sz
    lib.jgm_model_size.argtypes = [vp, i]
    lib.jgm_storedj_host_resident.restype = i
    lib.jgm_storedj_host_resident.argtypes = [vp]
    lib.jgm_storedj_file_backed.restype = i
    lib.jgm_storedj_file_backed.argtypes = [vp]
    lib.jgm_storedj_rows.restype = sz
    lib.jgm_storedj_rows.argtypes = [vp]
    lib.jgm_profile_count.restype = sz
    lib.jgm_profile_count.argtypes = [vp]
    if lib.jgm_api_version() != API_VERSION:
        raise RuntimeError(f"{LIBPATH} has C ABI version {lib.jgm_api_version()}, this binding "
                           f"expects {API_VERSION}: rebuild with `python -m jif3d_amd.build`")
    _lib = lib
    return lib


def check(status):
    if status != JGM_OK:
        raise JgmError(status, load().jgm_last_error().decode())


def _pd(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check_rhs(x, y, nin, nout, dtype=np.float32):
    """Shapes of a batched stored-J product: x (nrhs, nin) and y (nrhs, nout), of dtype."""
    if len(x.shape) != 2 or tuple(y.shape) != (x.shape[0], nout) or x.shape[1] != nin:
        raise ValueError(f"stored-J operands must be (nrhs, {nin}) and (nrhs, {nout}), "
                         f"got {tuple(x.shape)} and {tuple(y.shape)}")
    if np.dtype(x.dtype) != np.dtype(dtype) or np.dtype(y.dtype) != np.dtype(dtype):
        raise ValueError(f"stored-J operands are {np.dtype(dtype).name}")


def _devptr(t):
    """Device address of a DeviceArray (or of any object exposing data_ptr())."""
    if isinstance(t, DeviceArray):
        return ctypes.c_void_p(t.ptr)
    return ctypes.c_void_p(t.data_ptr())


def _is_device(a):
    return isinstance(a, DeviceArray) or (hasattr(a, "data_ptr") and getattr(a, "is_cuda", False))


class DeviceArray:
    """A device allocation made through the library (jgm_device_alloc)."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(ctx.lib.jgm_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, arr, dtype=None):
        arr = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        d = cls(ctx, arr.shape, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(self.ctx.lib.jgm_copy_to_device(self.ctx.handle, ctypes.c_void_p(self.ptr),
                                              arr.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.lib.jgm_copy_to_host(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx.lib.jgm_device_free(self.ctx.handle, ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Owns one jgm_context: mesh, background, sites, field, transform."""

    def __init__(self, device=0, stream=None):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.jgm_create(int(device), stream, ctypes.byref(h)))
        self.handle = h
        self.device = device
        self.nsites = 0
        self.storedj_kind = 0

    def close(self):
        if getattr(self, "handle", None):
            self.lib.jgm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        check(self.lib.jgm_set_stream(self.handle, stream))

    def synchronize(self):
        check(self.lib.jgm_synchronize(self.handle))

    def set_mesh(self, mesh):
        xc, yc, zc = (_f64(v) for v in mesh.coords)
        xs, ys, zs = (_f64(v) for v in mesh.sizes)
        check(self.lib.jgm_set_mesh(self.handle, mesh.nx, mesh.ny, mesh.nz, _pd(xc), _pd(yc),
                                    _pd(zc), _pd(xs), _pd(ys), _pd(zs)))
        self.ncells = int(mesh.nx) * int(mesh.ny) * int(mesh.nz)
        t = _f64(getattr(mesh, "bg_thicknesses", np.zeros(0)))
        check(self.lib.jgm_set_background(self.handle, len(t), _pd(t) if len(t) else None))

    def set_sites(self, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        check(self.lib.jgm_set_sites(self.handle, len(x), _pd(x), _pd(y), _pd(z)))
        self.nsites = len(x)

    def set_field(self, inclination, declination, fieldstrength):
        check(self.lib.jgm_set_field(self.handle, float(inclination), float(declination),
                                     float(fieldstrength)))

    def set_transform(self, tf, inclination=0.0, declination=0.0):
        check(self.lib.jgm_set_transform(self.handle, int(tf), float(inclination),
                                         float(declination)))

    def set_mixed_threshold(self, q0):
        check(self.lib.jgm_set_mixed_threshold(self.handle, float(q0)))

    def set_mixed_check(self, nsites, ncells, tolerance=5e-7):
        """A-posteriori check of the mixed path (nsites = 0 turns it off)."""
        check(self.lib.jgm_set_mixed_check(self.handle, int(nsites), int(ncells),
                                           float(tolerance)))

    def mixed_check_report(self):
        """(sampled forward error, sampled adjoint error, calls recomputed in FP64)."""
        fe, ae, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_size_t()
        check(self.lib.jgm_mixed_check_report(self.handle, ctypes.byref(fe), ctypes.byref(ae),
                                              ctypes.byref(n)))
        return fe.value, ae.value, n.value

    def set_mixed_far_radius(self, q1):
        check(self.lib.jgm_set_mixed_far_radius(self.handle, float(q1)))

    def data_per_site(self, kind):
        return self.lib.jgm_data_per_site(self.handle, kind)

    def n_params(self):
        return self.lib.jgm_n_params(self.handle)

    def model_size(self, kind):
        return self.lib.jgm_model_size(self.handle, kind)

    def forward(self, kind, prec, model, out=None):
        """Host arrays in -> host array out; torch device tensors -> device result."""
        if _is_device(model):
            if out is None:
                raise ValueError("device forward needs an output tensor")
            check(self.lib.jgm_forward(self.handle, kind, prec, _devptr(model), _devptr(out)))
            return out
        m = _f64(model)
        res = np.zeros(self.nsites * self.data_per_site(kind))
        check(self.lib.jgm_forward_host(self.handle, kind, prec, _pd(m), _pd(res)))
        return res

    def adjoint(self, kind, prec, model, misfit, out=None):
        if _is_device(misfit):
            if out is None:
                raise ValueError("device adjoint needs an output tensor")
            mp = _devptr(model) if model is not None else None
            check(self.lib.jgm_adjoint(self.handle, kind, prec, mp, _devptr(misfit),
                                       _devptr(out)))
            return out
        m = _f64(model) if model is not None else None
        r = _f64(misfit)
        g = np.zeros(self.model_size(kind))
        check(self.lib.jgm_adjoint_host(self.handle, kind, prec, _pd(m), _pd(r), _pd(g)))
        return g

    def forward_batch(self, kind, prec, models, out=None, nmodels=None):
        """Forward of several models.  Host: models (B, model_size) -> (B, nout).
        Device arrays: packed (B * model_size) in, (B * nout) out, nmodels = B."""
        nout = self.nsites * self.data_per_site(kind)
        if _is_device(models):
            if out is None or nmodels is None:
                raise ValueError("device forward_batch needs an output array and nmodels")
            check(self.lib.jgm_forward_batch(self.handle, kind, prec, int(nmodels),
                                             _devptr(models), self.model_size(kind),
                                             _devptr(out), nout))
            return out
        m = np.ascontiguousarray(np.asarray(models, dtype=np.float64).reshape(
            -1, self.model_size(kind)))
        res = np.zeros((m.shape[0], nout))
        check(self.lib.jgm_forward_batch_host(self.handle, kind, prec, m.shape[0], _pd(m),
                                              _pd(res)))
        return res

    def adjoint_batch(self, kind, prec, models, misfits, out=None, nmodels=None):
        """2 J^T D^T r for several (model, misfit) pairs.  Host: models (B, model_size) or
        None (no transform), misfits (B, nout) -> (B, model_size).  Device: packed."""
        np_ = self.model_size(kind)
        nout = self.nsites * self.data_per_site(kind)
        if _is_device(misfits):
            if out is None or nmodels is None:
                raise ValueError("device adjoint_batch needs an output array and nmodels")
            mp = _devptr(models) if models is not None else None
            check(self.lib.jgm_adjoint_batch(self.handle, kind, prec, int(nmodels), mp, np_,
                                             _devptr(misfits), nout, _devptr(out), np_))
            return out
        r = np.ascontiguousarray(np.asarray(misfits, dtype=np.float64).reshape(-1, nout))
        m = None
        if models is not None:
            m = np.ascontiguousarray(np.asarray(models, dtype=np.float64).reshape(-1, np_))
            if m.shape[0] != r.shape[0]:
                raise ValueError("models and misfits differ in batch size")
        g = np.zeros((r.shape[0], np_))
        check(self.lib.jgm_adjoint_batch_host(self.handle, kind, prec, r.shape[0],
                                              _pd(m) if m is not None else None, _pd(r),
                                              _pd(g)))
        return g

    def sensitivity_rows(self, kind, site0, nsites, out):
        check(self.lib.jgm_sensitivity_rows(self.handle, kind, site0, nsites, _devptr(out)))
        return out

    def storedj_build(self, kind, prec, host_block_cells=None, fp64=False, directory=None):
        """Build the stored J in HBM, or with host_block_cells in page-locked host
        memory, streamed to the device in blocks of that many cells (with directory: in a
        file there, read back block by block); fp64 keeps J in FP64 (HBM only) for the
        GEMM_FP64 products."""
        if directory is not None:
            check(self.lib.jgm_storedj_build_file(self.handle, kind, prec,
                                                  int(host_block_cells or 16384),
                                                  os.fsencode(directory)))
        elif fp64:
            if host_block_cells:
                raise ValueError("the FP64 stored J is held in HBM only")
            check(self.lib.jgm_storedj_build_fp64(self.handle, kind))
        elif host_block_cells:
            check(self.lib.jgm_storedj_build_host(self.handle, kind, prec, int(host_block_cells)))
        else:
            check(self.lib.jgm_storedj_build(self.handle, kind, prec))
        self.storedj_kind = kind

    def storedj_host_resident(self):
        return bool(self.lib.jgm_storedj_host_resident(self.handle))

    def storedj_file_backed(self):
        return bool(self.lib.jgm_storedj_file_backed(self.handle))

    def storedj_rows(self):
        return self.lib.jgm_storedj_rows(self.handle)

    def storedj_cols(self):
        """Parameter columns of the stored J: three per cell for the magnetization kind."""
        return 3 * self.ncells if self.storedj_kind == KIND_MAGNETIZATION else self.ncells

    def alloc(self, shape, dtype=np.float64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, arr, dtype=None):
        return DeviceArray.from_host(self, arr, dtype)

    def copy_device(self, dst, src):
        n = min(dst.nbytes, src.nbytes)
        check(self.lib.jgm_copy_device(self.handle, ctypes.c_void_p(dst.ptr),
                                       ctypes.c_void_p(src.ptr), n))

    def profile_enable(self, on=True):
        check(self.lib.jgm_profile_enable(self.handle, 1 if on else 0))

    def profile_reset(self):
        check(self.lib.jgm_profile_reset(self.handle))

    def profile(self):
        """{kernel name: (total ms, launches)} of the timed launches so far."""
        res = {}
        for i in range(self.lib.jgm_profile_count(self.handle)):
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            n = ctypes.c_size_t()
            check(self.lib.jgm_profile_get(self.handle, i, ctypes.byref(name), ctypes.byref(ms),
                                           ctypes.byref(n)))
            res[name.value.decode()] = (ms.value, n.value)
        return res

    def storedj_apply(self, prec, M, Y):
        """Y = J M for device arrays of FP32 right-hand sides, one per row: M of shape
        (nrhs, storedj_cols()), Y of shape (nrhs, storedj_rows())."""
        _check_rhs(M, Y, self.storedj_cols(), self.storedj_rows())
        check(self.lib.jgm_storedj_apply(self.handle, prec, _devptr(M), M.shape[0],
                                         _devptr(Y)))
        return Y

    def storedj_apply_t(self, prec, R, G):
        """G = J^T R for device arrays: R of shape (nrhs, storedj_rows()), G of shape
        (nrhs, storedj_cols())."""
        _check_rhs(R, G, self.storedj_rows(), self.storedj_cols())
        check(self.lib.jgm_storedj_apply_t(self.handle, prec, _devptr(R), R.shape[0],
                                           _devptr(G)))
        return G

    def storedj_apply_f64(self, M, Y):
        """Y = J M on the FP64 J: device FP64 arrays, M (nrhs, storedj_cols()), Y (nrhs,
        storedj_rows())."""
        _check_rhs(M, Y, self.storedj_cols(), self.storedj_rows(), np.float64)
        check(self.lib.jgm_storedj_apply_f64(self.handle, _devptr(M), M.shape[0], _devptr(Y)))
        return Y

    def storedj_apply_t_f64(self, R, G):
        """G = J^T R on the FP64 J: R (nrhs, storedj_rows()), G (nrhs, storedj_cols())."""
        _check_rhs(R, G, self.storedj_rows(), self.storedj_cols(), np.float64)
        check(self.lib.jgm_storedj_apply_t_f64(self.handle, _devptr(R), R.shape[0],
                                               _devptr(G)))
        return G

    def storedj_forward(self, prec, model, out=None):
        """Cached forward over the stored J (host arrays, or device arrays with out)."""
        if _is_device(model):
            if out is None:
                raise ValueError("device forward needs an output array")
            check(self.lib.jgm_storedj_forward(self.handle, prec, _devptr(model), _devptr(out)))
            return out
        m = _f64(model)
        res = np.zeros(self.nsites * self.data_per_site(self.storedj_kind))
        check(self.lib.jgm_storedj_forward_host(self.handle, prec, _pd(m), _pd(res)))
        return res

    def storedj_gradient(self, prec, model, misfit, out=None):
        """Cached gradient 2 [J B]^T D^T misfit over the stored J."""
        if _is_device(misfit):
            if out is None:
                raise ValueError("device gradient needs an output array")
            mp = _devptr(model) if model is not None else None
            check(self.lib.jgm_storedj_gradient(self.handle, prec, mp, _devptr(misfit),
                                                _devptr(out)))
            return out
        m = _f64(model) if model is not None else None
        r = _f64(misfit)
        g = np.zeros(self.model_size(self.storedj_kind))
        check(self.lib.jgm_storedj_gradient_host(self.handle, prec, _pd(m), _pd(r), _pd(g)))
        return g

    def storedj_forward_batch(self, prec, models):
        """Cached forward of several models (host (B, n_params) -> (B, nout)); up to 64
        models share one pass over J."""
        n = self.model_size(self.storedj_kind)
        m = np.ascontiguousarray(np.asarray(models, dtype=np.float64).reshape(-1, n))
        res = np.zeros((m.shape[0], self.nsites * self.data_per_site(self.storedj_kind)))
        check(self.lib.jgm_storedj_forward_batch_host(self.handle, prec, m.shape[0], _pd(m),
                                                      _pd(res)))
        return res

    def storedj_gradient_batch(self, prec, models, misfits):
        """Cached gradients of several (model, misfit) pairs (models may be None without
        a transform)."""
        nout = self.nsites * self.data_per_site(self.storedj_kind)
        r = np.ascontiguousarray(np.asarray(misfits, dtype=np.float64).reshape(-1, nout))
        m = None
        if models is not None:
            m = np.ascontiguousarray(np.asarray(models, dtype=np.float64).reshape(
                -1, self.model_size(self.storedj_kind)))
        g = np.zeros((r.shape[0], self.model_size(self.storedj_kind)))
        check(self.lib.jgm_storedj_gradient_batch_host(self.handle, prec, r.shape[0],
                                                       _pd(m) if m is not None else None,
                                                       _pd(r), _pd(g)))
        return g

    def storedj_release(self):
        check(self.lib.jgm_storedj_release(self.handle))

    def configure(self, mesh, data, field=None):
        self.set_mesh(mesh)
        x, y, z = data.positions
        self.set_sites(x, y, z)
        if field is not None:
            self.set_field(*field)
