"""fwi3d_amd — MI355X-native time-domain wave propagation and adjoint gradients.

Python bindings (ctypes) of the C ABI declared in include/fwi3d_hip.h and
implemented by fwi3d_amd/libfwi3d_amd.so (HIP, gfx950). The library is the product;
this module only marshals arguments. Importing works without a GPU; creating an
Engine needs one, and every call fails loudly (RuntimeError) if the native library
is missing or returns an error — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfwi3d_amd.so")

ACOUSTIC, ELASTIC, VTI = 0, 1, 2
PHYSICS = {"acoustic": ACOUSTIC, "elastic": ELASTIC, "vti": VTI}
SNAP = {"f32": 0, "bf16": 1, "f16": 2}

# Exported symbols of include/fwi3d_hip.h (checked by the CPU test suite).
EXPORTS = (
    "fwi3d_create", "fwi3d_destroy", "fwi3d_last_error", "fwi3d_injector_tables",
    "fwi3d_propagate", "fwi3d_propagate_device", "fwi3d_profile", "fwi3d_profile_read",
    "fwi3d_op_param_defaults", "fwi3d_analyze_model", "fwi3d_apply_forward", "fwi3d_apply_jacobianT",
    "fwi3d_apply_forward_device", "fwi3d_apply_jacobianT_device", "fwi3d_fwi_shot_device",
    "fwi3d_lbfgs_param_defaults", "fwi3d_lbfgs_run", "fwi3d_tune", "fwi3d_autotune",
    "fwi3d_batch_create", "fwi3d_batch_destroy", "fwi3d_batch_footprint", "fwi3d_batch_add", "fwi3d_batch_step",
    "fwi3d_batch_query", "fwi3d_batch_wait", "fwi3d_batch_status", "fwi3d_batch_autotune", "fwi3d_batch_tiles",
    "fwi3d_batch_snap_footprint", "fwi3d_batch_add_adjoint", "fwi3d_batch_release", "fwi3d_batch_retired",
    "fwi3d_server_create", "fwi3d_server_submit", "fwi3d_server_poll", "fwi3d_server_wait", "fwi3d_server_error",
    "fwi3d_server_result", "fwi3d_server_stats", "fwi3d_server_destroy", "fwi3d_dipole_tables",
    "fwi3d_snapshots_read", "fwi3d_propagate_device_noalloc",
)


class Param(C.Structure):
    """fwi3d_param (include/fwi3d_hip.h)."""
    _fields_ = [("nt", C.c_int), ("sub", C.c_int), ("dt", C.c_float),
                ("bc", C.c_int * 6), ("taper", C.c_int * 6),
                ("taper_strength", C.c_float), ("free_surface_stiffness", C.c_float),
                ("snap_precision", C.c_int)]


class InjectorStruct(C.Structure):
    """fwi3d_injector (include/fwi3d_hip.h)."""
    _fields_ = [("npts", C.c_int), ("kind", C.c_int),
                ("xind", C.c_void_p), ("yind", C.c_void_p), ("zind", C.c_void_p),
                ("xw", C.c_void_p), ("yw", C.c_void_p), ("zw", C.c_void_p)]


class OpParam(C.Structure):
    """fwi3d_op_param (include/fwi3d_hip.h): the reference's param fields read by the
    shot operator; dt, courant, sub, nt, vmin, vmax are updated by each call."""
    _fields_ = [("nmodels", C.c_int), ("dt", C.c_float), ("courant", C.c_float), ("sub", C.c_int),
                ("nt", C.c_int), ("vmin", C.c_float), ("vmax", C.c_float), ("resampling", C.c_int),
                ("sinc_half_length", C.c_int), ("bc", C.c_int * 6), ("taper", C.c_int * 6),
                ("taper_strength", C.c_float), ("free_surface_stiffness", C.c_float), ("mt", C.c_int),
                ("seismotype", C.c_int), ("soft_clip", C.c_int)] + \
               [(k, C.c_float) for k in ("vpmin", "vpmax", "vsmin", "vsmax", "rhomin", "rhomax", "deltamin",
                                         "deltamax", "epsilonmin", "epsilonmax", "gammamin", "gammamax")] + \
               [("snap_precision", C.c_int), ("gl", C.c_float)]


class ShotStruct(C.Structure):
    """fwi3d_shot (include/fwi3d_hip.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
               [(k, C.c_float) for k in ("dx", "dy", "dz", "ox", "oy", "oz")] + \
               [("nt_data", C.c_int), ("dt_data", C.c_float), ("sxyz", C.c_void_p), ("nrcv", C.c_int),
                ("rxyz", C.c_void_p), ("rdipaz", C.c_void_p)]


class LbfgsParam(C.Structure):
    """fwi3d_lbfgs_param (include/fwi3d_hip.h)."""
    _fields_ = [("niter", C.c_int), ("m", C.c_int), ("max_trial", C.c_int), ("threshold", C.c_double),
                ("c1", C.c_double), ("c2", C.c_double), ("a0", C.c_double), ("a1", C.c_double),
                ("step_max", C.c_double), ("version", C.c_int)]


class RequestStruct(C.Structure):
    """fwi3d_request (include/fwi3d_hip.h)."""
    _fields_ = [("kind", C.c_int), ("par", OpParam), ("shot", ShotStruct), ("d_model", C.c_void_p),
                ("d_src", C.c_void_p), ("d_obs", C.c_void_p), ("d_data", C.c_void_p), ("d_grad64", C.c_void_p)]


class ServerCounters(C.Structure):
    """fwi3d_server_counters (include/fwi3d_hip.h)."""
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
                ("queued", C.c_int), ("active", C.c_int), ("hbm_in_use", C.c_longlong), ("hbm_peak", C.c_longlong),
                ("peak_active", C.c_int), ("steps", C.c_longlong)]


# int (*)(void *user, const float *d_model, float *d_grad, double *f)
OBJECTIVE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double))

_lib = None


def lib():
    """Load the native library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"fwi3d_amd native library not built: {LIB_PATH} (run python -m fwi3d_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.fwi3d_last_error.restype = C.c_char_p
        L.fwi3d_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        L.fwi3d_destroy.argtypes = [C.c_void_p]
        L.fwi3d_last_error.argtypes = [C.c_void_p]
        L.fwi3d_profile.argtypes = [C.c_void_p, C.c_int]
        L.fwi3d_profile_read.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.fwi3d_op_param_defaults.restype = None
        L.fwi3d_op_param_defaults.argtypes = [C.POINTER(OpParam)]
        L.fwi3d_lbfgs_param_defaults.restype = None
        L.fwi3d_lbfgs_param_defaults.argtypes = [C.POINTER(LbfgsParam)]
        L.fwi3d_lbfgs_run.argtypes = [C.c_void_p, C.POINTER(LbfgsParam), C.c_longlong, C.c_void_p, C.c_void_p,
                                      C.c_void_p, OBJECTIVE, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.fwi3d_batch_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.fwi3d_batch_destroy.argtypes = [C.c_void_p]
        L.fwi3d_batch_footprint.restype = C.c_longlong
        L.fwi3d_batch_footprint.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.fwi3d_batch_add.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.c_float, C.POINTER(C.c_int)]
        L.fwi3d_batch_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.fwi3d_batch_query.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.fwi3d_batch_wait.argtypes = [C.c_void_p]
        L.fwi3d_batch_status.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.fwi3d_batch_autotune.argtypes = [C.c_void_p, C.c_int]
        L.fwi3d_batch_tiles.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.fwi3d_batch_snap_footprint.restype = C.c_longlong
        L.fwi3d_batch_snap_footprint.argtypes = [C.c_int] * 7
        L.fwi3d_batch_add_adjoint.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float,
                                              C.c_float, C.c_float, C.POINTER(C.c_int)]
        L.fwi3d_batch_release.argtypes = [C.c_void_p, C.c_int]
        L.fwi3d_batch_retired.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.fwi3d_server_create.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.POINTER(C.c_void_p)]
        L.fwi3d_server_submit.argtypes = [C.c_void_p, C.POINTER(RequestStruct), C.POINTER(C.c_int)]
        L.fwi3d_server_poll.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.fwi3d_server_wait.argtypes = [C.c_void_p, C.c_int]
        L.fwi3d_server_error.restype = C.c_char_p
        L.fwi3d_server_error.argtypes = [C.c_void_p, C.c_int]
        L.fwi3d_server_result.argtypes = [C.c_void_p, C.c_int, C.POINTER(OpParam)]
        L.fwi3d_server_stats.argtypes = [C.c_void_p, C.POINTER(ServerCounters)]
        L.fwi3d_server_destroy.argtypes = [C.c_void_p]
        L.fwi3d_snapshots_read.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        _lib = L
    return _lib


def make_param(nt, dt, sub=0, bc=(1, 1, 1, 1, 1, 1), taper=(0,) * 6, taper_strength=0.05,
               free_surface_stiffness=1.05, snap_precision="f32"):
    sp = SNAP[snap_precision] if isinstance(snap_precision, str) else int(snap_precision)
    return Param(int(nt), int(sub), float(dt), (C.c_int * 6)(*bc), (C.c_int * 6)(*taper),
                 float(taper_strength), float(free_surface_stiffness), sp)


class Injector:
    """Injector tables for point sources/receivers (delta_m3: kind 0, ddelta_m3: kind 1),
    computed by the library (fwi3d_injector_tables)."""

    def __init__(self, kind, loc, shape, d, o=(0.0, 0.0, 0.0)):
        ny, nx, nz = shape
        loc = np.ascontiguousarray(np.asarray(loc, dtype=np.float32).reshape(-1, 3))
        self.npts = int(loc.shape[0])
        self.kind = int(kind)
        nw = 6 if kind else 3
        self.xind = np.zeros((self.npts, 3), np.int32)
        self.yind = np.zeros_like(self.xind)
        self.zind = np.zeros_like(self.xind)
        self.xw = np.zeros((self.npts, 2, nw), np.float32)
        self.yw = np.zeros_like(self.xw)
        self.zw = np.zeros_like(self.xw)
        rc = lib().fwi3d_injector_tables(
            self.kind, C.c_void_p(loc.ctypes.data), self.npts, nx, ny, nz,
            C.c_float(d[0]), C.c_float(d[1]), C.c_float(d[2]),
            C.c_float(o[0]), C.c_float(o[1]), C.c_float(o[2]),
            C.c_void_p(self.xind.ctypes.data), C.c_void_p(self.yind.ctypes.data), C.c_void_p(self.zind.ctypes.data),
            C.c_void_p(self.xw.ctypes.data), C.c_void_p(self.yw.ctypes.data), C.c_void_p(self.zw.ctypes.data))
        if rc != 0:
            raise ValueError("point outside the grid or invalid injector arguments")

    @classmethod
    def dipole(cls, loc5, gl, shape, d, o=(0.0, 0.0, 0.0)):
        """DAS channels (dipole_m3, kind 2; fwi3d_dipole_tables): loc5 [npts][5] = x, y, z,
        dip, azimuth; gauge length gl. Indices [npts][6], weights [npts][2][6]."""
        ny, nx, nz = shape
        loc5 = np.ascontiguousarray(np.asarray(loc5, dtype=np.float32).reshape(-1, 5))
        self = cls.__new__(cls)
        self.npts, self.kind = int(loc5.shape[0]), 2
        self.xind = np.zeros((self.npts, 6), np.int32)
        self.yind = np.zeros_like(self.xind)
        self.zind = np.zeros_like(self.xind)
        self.xw = np.zeros((self.npts, 2, 6), np.float32)
        self.yw = np.zeros_like(self.xw)
        self.zw = np.zeros_like(self.xw)
        rc = lib().fwi3d_dipole_tables(
            C.c_void_p(loc5.ctypes.data), self.npts, C.c_float(gl), nx, ny, nz,
            C.c_float(d[0]), C.c_float(d[1]), C.c_float(d[2]), C.c_float(o[0]), C.c_float(o[1]), C.c_float(o[2]),
            C.c_void_p(self.xind.ctypes.data), C.c_void_p(self.yind.ctypes.data), C.c_void_p(self.zind.ctypes.data),
            C.c_void_p(self.xw.ctypes.data), C.c_void_p(self.yw.ctypes.data), C.c_void_p(self.zw.ctypes.data))
        if rc != 0:
            raise ValueError("DAS channel end outside the grid or invalid arguments")
        return self

    def cstruct(self):
        return InjectorStruct(self.npts, self.kind, self.xind.ctypes.data, self.yind.ctypes.data,
                              self.zind.ctypes.data, self.xw.ctypes.data, self.yw.ctypes.data, self.zw.ctypes.data)


def _vp(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def n_model(physics):
    return {ACOUSTIC: 2, ELASTIC: 3, VTI: 6}[physics]


def n_comp(physics):
    return 1 if physics == ACOUSTIC else 3


def n_grad(physics):
    return 2 if physics == ACOUSTIC else 3


class Engine:
    """One device context (fwi3d_ctx): a HIP stream plus the device-resident state of
    an operator instance (wavefields, snapshots kept from forward to adjoint)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        self._batches = []
        self._servers = []
        rc = lib().fwi3d_create(C.byref(self._h), int(device))
        if rc != 0:
            raise RuntimeError(f"fwi3d_create failed ({rc})")

    def close(self):
        if self._h:
            for s in list(self._servers):  # stop the server threads before the context goes
                try:
                    s.close()
                except RuntimeError:  # a server that had failed; its requests already say so
                    pass
            self._servers = []
            for b in self._batches:  # the context destroys its batches with it
                b._h = C.c_void_p()
            self._batches = []
            lib().fwi3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().fwi3d_last_error(self._h)
            raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def propagate(self, physics, model, src, inj, ext, par, d, adj=False, grad=None):
        """Host-array propagate (fwi3d_propagate). model [np][ny][nx][nz] converted
        parameters; src [ns][nsrc][nt]; returns rcv [nc][nrcv][nt] for a forward call.
        grad (float32 [ng][ny][nx][nz]) is accumulated in place for an adjoint call."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        model = np.ascontiguousarray(model, dtype=np.float32)
        src = np.ascontiguousarray(src, dtype=np.float32)
        ny, nx, nz = model.shape[-3:]
        rcv = None if adj else np.zeros((n_comp(ph), ext.npts, par.nt), np.float32)
        if grad is not None:
            assert grad.dtype == np.float32 and grad.flags.c_contiguous
        I = inj.cstruct()
        E = ext.cstruct() if ext is not None else InjectorStruct(0, 0, None, None, None, None, None, None)
        rc = lib().fwi3d_propagate(self._h, ph, int(bool(adj)), _vp(model), _vp(src), _vp(rcv), C.byref(I),
                                   C.byref(E), _vp(grad), C.byref(par), nx, ny, nz,
                                   C.c_float(d[0]), C.c_float(d[1]), C.c_float(d[2]))
        self._check(rc, "fwi3d_propagate")
        return rcv

    def propagate_device(self, physics, model_ptr, src_ptr, rcv_ptr, inj, ext, par, shape, d, adj=False,
                         grad64_ptr=None):
        """Device-pointer propagate (fwi3d_propagate_device): pointers are integers
        (e.g. torch.Tensor.data_ptr()) to HBM-resident arrays in reference layouts."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        I = inj.cstruct()
        E = ext.cstruct() if ext is not None else InjectorStruct(0, 0, None, None, None, None, None, None)
        rc = lib().fwi3d_propagate_device(self._h, ph, int(bool(adj)), C.c_void_p(model_ptr), C.c_void_p(src_ptr),
                                          C.c_void_p(rcv_ptr) if rcv_ptr else None, C.byref(I), C.byref(E),
                                          C.c_void_p(grad64_ptr) if grad64_ptr else None, C.byref(par), nx, ny, nz,
                                          C.c_float(d[0]), C.c_float(d[1]), C.c_float(d[2]))
        self._check(rc, "fwi3d_propagate_device")

    def propagate_device_noalloc(self, physics, model_ptr, src_ptr, rcv_ptr, inj, ext, par, shape, d, adj=False):
        """Allocation-free device-pointer propagate (fwi3d_propagate_device_noalloc): the
        shot's buffers live in the library's static device workspace (128 MiB per GPU);
        no snapshot store (par.sub must be 0) and no gradient. Calls on one GPU must not
        overlap."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        I = inj.cstruct()
        E = ext.cstruct() if ext is not None else InjectorStruct(0, 0, None, None, None, None, None, None)
        rc = lib().fwi3d_propagate_device_noalloc(self._h, ph, int(bool(adj)), C.c_void_p(model_ptr),
                                                  C.c_void_p(src_ptr), C.c_void_p(rcv_ptr) if rcv_ptr else None,
                                                  C.byref(I), C.byref(E), C.byref(par), nx, ny, nz, C.c_float(d[0]),
                                                  C.c_float(d[1]), C.c_float(d[2]))
        self._check(rc, "fwi3d_propagate_device_noalloc")

    def tune(self, ychunk=None, tx=None, graphs=None, fused=None, fchunk=None, overlap=None, echunk=None):
        """Tiling of the time-step kernels, graph replay of time loops, and the elastic /
        VTI step form (fused, the default: the fused kernel over the planes 6 cells inside
        the y faces; fchunk: its y planes per workgroup, 0 by grid size; overlap: 0 one
        stream, 1 / 2 its y-face launches on a second stream from the step's start / from
        the edge tiles on, 3 its edge tiles on a second stream beside the middle tiles;
        echunk: y planes per workgroup of its side tiles, 0 by grid size) (fwi3d_tune). Results do not depend on the tiling, the overlap or
        graph replay; the two elastic forms agree to rounding."""
        if ychunk is not None:
            self._check(lib().fwi3d_tune(self._h, 0, int(ychunk)), "fwi3d_tune")
        if tx is not None:
            self._check(lib().fwi3d_tune(self._h, 1, int(tx)), "fwi3d_tune")
        if fused is not None:
            self._check(lib().fwi3d_tune(self._h, 3, 1 if fused else 0), "fwi3d_tune")
        if fchunk is not None:
            self._check(lib().fwi3d_tune(self._h, 4, int(fchunk)), "fwi3d_tune")
        if overlap is not None:
            self._check(lib().fwi3d_tune(self._h, 5, int(overlap)), "fwi3d_tune")
        if echunk is not None:
            self._check(lib().fwi3d_tune(self._h, 6, int(echunk)), "fwi3d_tune")
        if graphs is not None:
            self._check(lib().fwi3d_tune(self._h, 2, int(bool(graphs))), "fwi3d_tune")

    def autotune(self, physics, shape, steps=10):
        """Measure every tile shape on this grid (fwi3d_autotune); the fastest is used for
        this physics and grid from now on. Returns (tx, ychunk)."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        tx, yc = C.c_int(), C.c_int()
        self._check(lib().fwi3d_autotune(self._h, ph, nx, ny, nz, int(steps), C.byref(tx), C.byref(yc)),
                    "fwi3d_autotune")
        return tx.value, yc.value

    def snapshots(self, shape):
        """The snapshot store of the last forward call with sub > 0 on grid shape
        (ny, nx, nz), decoded to FP32: [slots][ncomp][ny][nx][nz] (fwi3d_snapshots_read)."""
        slots, ncomp = C.c_int(), C.c_int()
        self._check(lib().fwi3d_snapshots_read(self._h, None, 0, C.byref(slots), C.byref(ncomp)),
                    "fwi3d_snapshots_read")
        out = np.zeros((slots.value, ncomp.value) + tuple(shape), np.float32)
        self._check(lib().fwi3d_snapshots_read(self._h, C.c_void_p(out.ctypes.data), out.size, C.byref(slots),
                                               C.byref(ncomp)), "fwi3d_snapshots_read")
        return out

    def profile(self, enable=True, steps_only=False):
        """Start (reset) or stop HIP-event timing; steps_only records whole time steps only."""
        level = (2 if steps_only else 1) if enable else 0
        self._check(lib().fwi3d_profile(self._h, level), "fwi3d_profile")

    def profile_read(self, kernel=0):
        """(milliseconds, launches) of kernel class 0 (interior time step), 1 (flux
        pre-pass, elastic/VTI), 2 (boundary-shell time step) or 3 (z-slab time step,
        elastic/VTI)."""
        ms = C.c_double()
        n = C.c_longlong()
        self._check(lib().fwi3d_profile_read(self._h, int(kernel), C.byref(ms), C.byref(n)), "fwi3d_profile_read")
        return ms.value, n.value


class Batch:
    """Batched forward modelling (fwi3d_batch_*): shots of any physics and grid join
    between time steps and advance together, one launch per kernel class and step.
    Pointers are device integers (torch.Tensor.data_ptr()); the caller keeps the
    model, source and receiver buffers alive until the shot's query() reports done."""

    def __init__(self, engine):
        self.engine = engine
        self._h = C.c_void_p()
        engine._check(lib().fwi3d_batch_create(engine._h, C.byref(self._h)), "fwi3d_batch_create")
        engine._batches.append(self)

    def close(self):
        if self._h:
            lib().fwi3d_batch_destroy(self._h)
            self._h = C.c_void_p()
            if self in self.engine._batches:
                self.engine._batches.remove(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def footprint(physics, shape, inj, ext):
        """Device bytes one shot holds while it is in the batch."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        I, E = inj.cstruct(), ext.cstruct()
        return int(lib().fwi3d_batch_footprint(ph, nx, ny, nz, C.byref(I), C.byref(E)))

    def add(self, physics, model_ptr, src_ptr, rcv_ptr, inj, ext, par, shape, d):
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        I, E = inj.cstruct(), ext.cstruct()
        sid = C.c_int(-1)
        self.engine._check(lib().fwi3d_batch_add(self._h, ph, C.c_void_p(model_ptr), C.c_void_p(src_ptr),
                                                 C.c_void_p(rcv_ptr) if rcv_ptr else None, C.byref(I), C.byref(E),
                                                 C.byref(par), nx, ny, nz, C.c_float(d[0]), C.c_float(d[1]),
                                                 C.c_float(d[2]), C.byref(sid)), "fwi3d_batch_add")
        return sid.value

    @staticmethod
    def snap_footprint(physics, shape, nt, sub, snap_precision="f32"):
        """Device bytes of the snapshot store of a forward shot with imaging interval sub."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        sp = SNAP[snap_precision] if isinstance(snap_precision, str) else int(snap_precision)
        ny, nx, nz = shape
        return int(lib().fwi3d_batch_snap_footprint(ph, nx, ny, nz, int(nt), int(sub), sp))

    def add_adjoint(self, fwd_id, physics, model_ptr, src_ptr, inj, grad64_ptr, par, shape, d):
        """Adjoint shot of forward shot fwd_id (which must have left the batch): inject
        src [ncomp][inj.npts][nt] at the receivers and accumulate the FP64 imaging
        condition into grad64 from fwd_id's snapshots."""
        ph = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        ny, nx, nz = shape
        I = inj.cstruct()
        sid = C.c_int(-1)
        self.engine._check(lib().fwi3d_batch_add_adjoint(self._h, int(fwd_id), ph, C.c_void_p(model_ptr),
                                                         C.c_void_p(src_ptr), C.byref(I), C.c_void_p(grad64_ptr),
                                                         C.byref(par), nx, ny, nz, C.c_float(d[0]), C.c_float(d[1]),
                                                         C.c_float(d[2]), C.byref(sid)), "fwi3d_batch_add_adjoint")
        return sid.value

    def release(self, fwd_id):
        """Drop the snapshot store of forward shot fwd_id."""
        self.engine._check(lib().fwi3d_batch_release(self._h, int(fwd_id)), "fwi3d_batch_release")

    def retired(self):
        """Ids of the shots that left the batch during the last step() call."""
        n = C.c_int(0)
        self.engine._check(lib().fwi3d_batch_retired(self._h, None, 0, C.byref(n)), "fwi3d_batch_retired")
        ids = (C.c_int * max(1, n.value))()
        self.engine._check(lib().fwi3d_batch_retired(self._h, ids, n.value, C.byref(n)), "fwi3d_batch_retired")
        return list(ids)[: n.value]

    def step(self, nsteps=1):
        """Advance every shot in the batch by up to nsteps time steps; returns the
        number of shots still in the batch."""
        n = C.c_int(0)
        self.engine._check(lib().fwi3d_batch_step(self._h, int(nsteps), C.byref(n)), "fwi3d_batch_step")
        return n.value

    def run(self):
        while self.step(1 << 20) > 0:
            pass
        self.wait()

    def query(self, sid):
        """1 when the shot's traces are complete, 0 while it runs, -1 if unknown."""
        st = C.c_int(-2)
        self.engine._check(lib().fwi3d_batch_query(self._h, int(sid), C.byref(st)), "fwi3d_batch_query")
        return st.value

    def wait(self):
        self.engine._check(lib().fwi3d_batch_wait(self._h), "fwi3d_batch_wait")

    def autotune(self, window=10):
        """Time y chunks 4, 8, 16, 32 over the next 4 x window steps, then keep the best."""
        self.engine._check(lib().fwi3d_batch_autotune(self._h, int(window)), "fwi3d_batch_autotune")

    def tiles(self):
        """(y chunk in force or 0, [ms per step of each candidate])."""
        yc, ms = C.c_int(), (C.c_float * 4)()
        self.engine._check(lib().fwi3d_batch_tiles(self._h, C.byref(yc), ms), "fwi3d_batch_tiles")
        return yc.value, list(ms)

    def status(self):
        """(shots in the batch, device bytes they hold, global step)."""
        n, b, it = C.c_int(), C.c_longlong(), C.c_int()
        self.engine._check(lib().fwi3d_batch_status(self._h, C.byref(n), C.byref(b), C.byref(it)),
                           "fwi3d_batch_status")
        return n.value, b.value, it.value


def op_param(**kw):
    """fwi3d_op_param with the reference's defaults (fwi3d_op_param_defaults) and the
    given fields; `bc`/`taper` take 6-tuples, `resampling` "sinc"/"linear" or 0/1."""
    p = OpParam()
    lib().fwi3d_op_param_defaults(C.byref(p))
    for k, v in kw.items():
        if k in ("bc", "taper"):
            setattr(p, k, (C.c_int * 6)(*v))
        elif k == "resampling" and isinstance(v, str):
            p.resampling = 1 if v == "linear" else 0
        elif k == "snap_precision" and isinstance(v, str):
            p.snap_precision = SNAP[v]
        else:
            setattr(p, k, v)
    return p


class Shot:
    """fwi3d_shot with its position arrays kept alive."""

    def __init__(self, shape, d, o, nt_data, dt_data, sxyz, rxyz, rdipaz=None):
        """rdipaz [nrcv][2]: fiber dip and azimuth of DAS channels (used when gl > 0)."""
        ny, nx, nz = shape
        self.shape = tuple(shape)
        self.sxyz = np.ascontiguousarray(np.asarray(sxyz, np.float32).reshape(3))
        self.rxyz = np.ascontiguousarray(np.asarray(rxyz, np.float32).reshape(-1, 3))
        self.rdipaz = None if rdipaz is None else np.ascontiguousarray(np.asarray(rdipaz, np.float32).reshape(-1, 2))
        self.nrcv = int(self.rxyz.shape[0])
        self.nt_data = int(nt_data)
        self.c = ShotStruct(nx, ny, nz, d[0], d[1], d[2], o[0], o[1], o[2], int(nt_data), float(dt_data),
                            self.sxyz.ctypes.data, self.nrcv, self.rxyz.ctypes.data,
                            None if self.rdipaz is None else self.rdipaz.ctypes.data)


def analyze_model(par, model, shot):
    """analyzeModel on a copy of the raw model: returns the clipped copy; par gets the
    time axis (fwi3d_analyze_model, host code)."""
    m = np.ascontiguousarray(model, dtype=np.float32).copy()
    rc = lib().fwi3d_analyze_model(C.byref(par), C.c_void_p(m.ctypes.data), C.byref(shot.c))
    if rc != 0:
        raise ValueError(f"fwi3d_analyze_model failed ({rc})")
    return m


class WeOp:
    """Nonlinear wave-equation operator of one shot, the interface of nl_we_op_e /
    nl_we_op_a / nl_we_op_vti (src/include/we_op.hpp:13-170): the physics follows
    par.nmodels (2 acoustic, 3 elastic, 6 VTI). Snapshots of apply_forward stay on the
    engine for the next apply_jacobianT, as on the reference's operator instance."""

    def __init__(self, engine, shot, src, par):
        self.eng = engine
        self.shot = shot
        self.src = np.ascontiguousarray(src, dtype=np.float32)
        self.par = par

    @property
    def nrcomp(self):
        return 1 if self.par.nmodels == 2 else 3

    def apply_forward(self, add, model, data=None):
        """model [nmodels][ny][nx][nz] raw parameters -> data [nrcomp][nrcv][nt_data]."""
        model = np.ascontiguousarray(model, dtype=np.float32)
        if data is None:
            data = np.zeros((self.nrcomp, self.shot.nrcv, self.shot.nt_data), np.float32)
        assert data.dtype == np.float32 and data.flags.c_contiguous
        rc = lib().fwi3d_apply_forward(self.eng._h, C.byref(self.par), C.byref(self.shot.c), int(bool(add)),
                                       _vp(model), _vp(self.src), _vp(data))
        self.eng._check(rc, "fwi3d_apply_forward")
        return data

    def apply_jacobianT(self, add, grad, model0, data):
        """data residual -> grad [min(3,nmodels)][ny][nx][nz] (float32, in place)."""
        model0 = np.ascontiguousarray(model0, dtype=np.float32)
        data = np.ascontiguousarray(data, dtype=np.float32)
        assert grad.dtype == np.float32 and grad.flags.c_contiguous
        rc = lib().fwi3d_apply_jacobianT(self.eng._h, C.byref(self.par), C.byref(self.shot.c), int(bool(add)),
                                         _vp(grad), _vp(model0), _vp(data))
        self.eng._check(rc, "fwi3d_apply_jacobianT")
        return grad


def apply_forward_device(engine, par, shot, add, model_ptr, src_ptr, data_ptr):
    rc = lib().fwi3d_apply_forward_device(engine._h, C.byref(par), C.byref(shot.c), int(bool(add)),
                                          C.c_void_p(model_ptr), C.c_void_p(src_ptr), C.c_void_p(data_ptr))
    engine._check(rc, "fwi3d_apply_forward_device")


def apply_jacobianT_device(engine, par, shot, add, grad64_ptr, model_ptr, data_ptr):
    rc = lib().fwi3d_apply_jacobianT_device(engine._h, C.byref(par), C.byref(shot.c), int(bool(add)),
                                            C.c_void_p(grad64_ptr), C.c_void_p(model_ptr), C.c_void_p(data_ptr))
    engine._check(rc, "fwi3d_apply_jacobianT_device")


def fwi_shot_device(engine, par, shot, model_ptr, src_ptr, obs_ptr, grad64_ptr, resid_ptr=None):
    """One FWI shot on device buffers: returns the misfit contribution (FP64); the
    gradient accumulates into grad64 [min(3,nmodels)][ny][nx][nz]."""
    mis = C.c_double()
    rc = lib().fwi3d_fwi_shot_device(engine._h, C.byref(par), C.byref(shot.c), C.c_void_p(model_ptr),
                                     C.c_void_p(src_ptr), C.c_void_p(obs_ptr),
                                     C.c_void_p(resid_ptr) if resid_ptr else None, C.c_void_p(grad64_ptr),
                                     C.byref(mis))
    engine._check(rc, "fwi3d_fwi_shot_device")
    return mis.value


def lbfgs_run(engine, n, model_ptr, grad_ptr, objective, niter, h0_ptr=None, **kw):
    """l-BFGS on device vectors (fwi3d_lbfgs_run). `objective(model_ptr, grad_ptr) -> f`
    evaluates the misfit and writes the gradient into the buffer at grad_ptr (the same
    buffers as given here). Returns (f history, (iterations, feval, geval))."""
    p = LbfgsParam()
    lib().fwi3d_lbfgs_param_defaults(C.byref(p))
    p.niter = int(niter)
    for k, v in kw.items():
        setattr(p, k, v)
    err = []

    def cb(user, m, g, f):
        try:
            f[0] = float(objective(m, g))
            return 0
        except Exception as e:  # reported after the run
            err.append(e)
            return 1

    fn = OBJECTIVE(cb)
    hist = (C.c_double * (p.niter + 1))()
    cnt = (C.c_int * 3)()
    rc = lib().fwi3d_lbfgs_run(engine._h, C.byref(p), int(n), C.c_void_p(model_ptr), C.c_void_p(grad_ptr),
                               C.c_void_p(h0_ptr) if h0_ptr else None, fn, None, hist, cnt)
    if err:
        raise err[0]
    engine._check(rc, "fwi3d_lbfgs_run")
    return list(hist)[: cnt[0] + 1], (cnt[0], cnt[1], cnt[2])


REQ_MODEL, REQ_GRADIENT = 0, 1


class Server:
    """Serving loop (fwi3d_server_*): a native server thread admits shot-modelling and
    shot-gradient requests by HBM footprint into one continuously refilled batch on the
    engine's device. While it exists the engine belongs to it. Device pointers are
    integers (torch.Tensor.data_ptr()); the caller keeps the buffers alive until the
    request is done."""

    def __init__(self, engine, hbm_budget, max_active=64):
        self.engine = engine
        self._h = C.c_void_p()
        self._shots = {}
        engine._check(lib().fwi3d_server_create(engine._h, int(hbm_budget), int(max_active), C.byref(self._h)),
                      "fwi3d_server_create")
        engine._servers.append(self)

    def submit(self, kind, par, shot, model_ptr, src_ptr, data_ptr=None, obs_ptr=None, grad64_ptr=None):
        """kind REQ_MODEL: data <- apply_forward(model); REQ_GRADIENT: one FWI shot
        (misfit, gradient += into grad64, residual into data if given). Returns an id."""
        r = RequestStruct(int(kind), par, shot.c, C.c_void_p(model_ptr), C.c_void_p(src_ptr),
                          C.c_void_p(obs_ptr) if obs_ptr else None, C.c_void_p(data_ptr) if data_ptr else None,
                          C.c_void_p(grad64_ptr) if grad64_ptr else None)
        rid = C.c_int(-1)
        self.engine._check(lib().fwi3d_server_submit(self._h, C.byref(r), C.byref(rid)), "fwi3d_server_submit")
        self._shots[rid.value] = shot  # position arrays are copied; keep the Shot for symmetry
        return rid.value

    def poll(self, rid):
        """(state, misfit, latency_s, queue_s); state 0 queued, 1 running, 2 done, -1 failed."""
        st, mis, lat, qs = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        self.engine._check(lib().fwi3d_server_poll(self._h, int(rid), C.byref(st), C.byref(mis), C.byref(lat),
                                                   C.byref(qs)), "fwi3d_server_poll")
        return st.value, mis.value, lat.value, qs.value

    def wait(self, rid=-1):
        rc = lib().fwi3d_server_wait(self._h, int(rid))
        if rc != 0:
            msg = lib().fwi3d_server_error(self._h, int(rid)) if rid >= 0 else b""
            raise RuntimeError(f"request {rid} failed ({rc}): {msg.decode() if msg else ''}")

    def error(self, rid):
        msg = lib().fwi3d_server_error(self._h, int(rid))
        return msg.decode() if msg else ""

    def result_param(self, rid):
        p = OpParam()
        self.engine._check(lib().fwi3d_server_result(self._h, int(rid), C.byref(p)), "fwi3d_server_result")
        return p

    def stats(self):
        c = ServerCounters()
        self.engine._check(lib().fwi3d_server_stats(self._h, C.byref(c)), "fwi3d_server_stats")
        return {k: getattr(c, k) for k, _ in ServerCounters._fields_}

    def close(self):
        if self._h:
            rc = lib().fwi3d_server_destroy(self._h)
            self._h = C.c_void_p()
            if self in self.engine._servers:
                self.engine._servers.remove(self)
            self.engine._check(rc, "fwi3d_server_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
