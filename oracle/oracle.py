"""ORACLE — test infrastructure only.

ctypes bindings of the CPU restatement (oracle/_build/libcpu_ref.so) and, where it
was built, of the reference's own operators (oracle/_ref/libref_ops_{f,d}.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this,
as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_REF = os.path.join(HERE, "_build", "libcpu_ref.so")
REF_OPS = {"f": os.path.join(HERE, "_ref", "libref_ops_f.so"), "d": os.path.join(HERE, "_ref", "libref_ops_d.so")}

PHYS = {"acoustic": 0, "elastic": 1, "vti": 2}
FACES = ("top", "bottom", "left", "right", "front", "back")


class Inj(C.Structure):
    _fields_ = [("npts", C.c_int), ("kind", C.c_int),
                ("xi", C.c_void_p), ("yi", C.c_void_p), ("zi", C.c_void_p),
                ("xw", C.c_void_p), ("yw", C.c_void_p), ("zw", C.c_void_p)]


class Par(C.Structure):
    _fields_ = [("nt", C.c_int), ("sub", C.c_int), ("dt", C.c_double),
                ("bc", C.c_int * 6), ("taper", C.c_int * 6),
                ("taper_strength", C.c_double), ("free_surface_stiffness", C.c_double), ("lag", C.c_int)]


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _dt(s):
    return np.float32 if s == "f" else np.float64


_lib = None
_refs = {}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(CPU_REF):
            raise RuntimeError(f"oracle library missing: {CPU_REF} (run `make -C oracle`)")
        _lib = C.CDLL(CPU_REF)
    return _lib


def ref_available(s="f"):
    return os.path.exists(REF_OPS[s])


def ref(s="f"):
    if s not in _refs:
        if not ref_available(s):
            raise RuntimeError(f"reference operator library missing: {REF_OPS[s]} (run `make -C oracle ref`)")
        _refs[s] = C.CDLL(REF_OPS[s])
    return _refs[s]


class Ops:
    """Operator-level calls with the same argument meaning for the oracle
    (`prefix='orc'`) and the compiled reference (`prefix='ref'`)."""

    def __init__(self, prefix="orc", s="f"):
        self.s = s
        self.T = _dt(s)
        self.cT = C.c_float if s == "f" else C.c_double
        self.L = lib() if prefix == "orc" else ref(s)
        self.p = prefix

    def _f(self, name):
        return getattr(self.L, f"{self.p}_{name}_{self.s}")

    def _chk(self, *arrs):
        for a in arrs:
            if a is not None:
                assert a.dtype == self.T and a.flags.c_contiguous

    def d1(self, axis, inp, out, h, par=None, add=False):
        ny, nx, nz = inp.shape[-3:]
        self._chk(inp, out, par)
        self._f("d1")(axis, int(add), _ptr(inp), _ptr(out), nx, ny, nz, self.cT(h), _ptr(par))
        return out

    def d2(self, axis, inp, out, b, h, a=1.0, add=False):
        ny, nx, nz = inp.shape[-3:]
        self._chk(inp, out, b)
        self._f("d2")(axis, int(add), _ptr(inp), _ptr(out), _ptr(b), nx, ny, nz, self.cT(h), self.cT(a))
        return out

    def apply_h(self, axis, inp, out, h, inv=False, add=False):
        ny, nx, nz = inp.shape[-3:]
        self._chk(inp, out)
        self._f("apply_h")(axis, int(inv), int(add), _ptr(inp), _ptr(out), nx, ny, nz, self.cT(h))
        return out

    def taper(self, u, tp, strength):
        ny, nx, nz = u.shape[-3:]
        ncomp = u.size // (nx * ny * nz)
        self._chk(u)
        tpa = (C.c_int * 6)(*tp)
        self._f("taper")(_ptr(u), ncomp, nx, ny, nz, tpa, self.cT(strength))
        return u

    def asat(self, face, bc, curr, prev, nxt, model, d, dt, alpha):
        ny, nx, nz = curr.shape[-3:]
        self._chk(curr, prev, nxt, model)
        self._f("asat")(face, bc, _ptr(curr), _ptr(prev), _ptr(nxt), _ptr(model), nx, ny, nz,
                        self.cT(d[0]), self.cT(d[1]), self.cT(d[2]), self.cT(dt), self.cT(alpha))
        return nxt

    def ascale(self, nxt, model, d, dt, bc):
        ny, nx, nz = nxt.shape[-3:]
        self._chk(nxt, model)
        self._f("ascale")(_ptr(nxt), _ptr(model), nx, ny, nz, self.cT(d[0]), self.cT(d[1]), self.cT(d[2]),
                          self.cT(dt), (C.c_int * 6)(*bc))
        return nxt

    def esat(self, face, bc, vti, curr, prev, nxt, model, d, dt):
        ny, nx, nz = curr.shape[-3:]
        self._chk(curr, prev, nxt, model)
        self._f("esat")(face, bc, int(vti), _ptr(curr), _ptr(prev), _ptr(nxt), _ptr(model), nx, ny, nz,
                        self.cT(d[0]), self.cT(d[1]), self.cT(d[2]), self.cT(dt))
        return nxt

    def escale(self, vti, nxt, model, d, dt, bc):
        ny, nx, nz = nxt.shape[-3:]
        self._chk(nxt, model)
        self._f("escale")(int(vti), _ptr(nxt), _ptr(model), nx, ny, nz, self.cT(d[0]), self.cT(d[1]),
                          self.cT(d[2]), self.cT(dt), (C.c_int * 6)(*bc))
        return nxt


class Injector:
    """Index/weight tables of a point injector (kind 0 delta_m3, 1 ddelta_m3)."""

    def __init__(self, kind, loc, shape, d, o, s="f"):
        ny, nx, nz = shape
        T = _dt(s)
        cT = C.c_float if s == "f" else C.c_double
        loc = np.ascontiguousarray(np.asarray(loc, dtype=T).reshape(-1, 3))
        self.npts = loc.shape[0]
        self.kind = kind
        nw = 6 if kind else 3
        self.xi = np.zeros((self.npts, 3), np.int32)
        self.yi = np.zeros_like(self.xi)
        self.zi = np.zeros_like(self.xi)
        self.xw = np.zeros((self.npts, 2, nw), T)
        self.yw = np.zeros_like(self.xw)
        self.zw = np.zeros_like(self.xw)
        getattr(lib(), f"orc_injector_{s}")(kind, _ptr(loc), self.npts, nx, ny, nz,
                                            cT(d[0]), cT(d[1]), cT(d[2]), cT(o[0]), cT(o[1]), cT(o[2]),
                                            _ptr(self.xi), _ptr(self.yi), _ptr(self.zi),
                                            _ptr(self.xw), _ptr(self.yw), _ptr(self.zw))

    @classmethod
    def dipole(cls, loc5, gl, shape, d, o, s="f"):
        """dipole_m3 tables of DAS channels loc5 [npts][5] = (x, y, z, dip, azimuth) with
        gauge length gl: indices [npts][6], weights [npts][2][6] (entries 0-2 the end at
        -gl/2 along the fiber, 3-5 the end at +gl/2)."""
        ny, nx, nz = shape
        T = _dt(s)
        cT = C.c_float if s == "f" else C.c_double
        loc5 = np.ascontiguousarray(np.asarray(loc5, dtype=T).reshape(-1, 5))
        self = cls.__new__(cls)
        self.npts, self.kind = loc5.shape[0], 2
        self.xi = np.zeros((self.npts, 6), np.int32)
        self.yi = np.zeros_like(self.xi)
        self.zi = np.zeros_like(self.xi)
        self.xw = np.zeros((self.npts, 2, 6), T)
        self.yw = np.zeros_like(self.xw)
        self.zw = np.zeros_like(self.xw)
        getattr(lib(), f"orc_dipole_injector_{s}")(_ptr(loc5), self.npts, cT(gl), nx, ny, nz, cT(d[0]), cT(d[1]),
                                                   cT(d[2]), cT(o[0]), cT(o[1]), cT(o[2]), _ptr(self.xi),
                                                   _ptr(self.yi), _ptr(self.zi), _ptr(self.xw), _ptr(self.yw),
                                                   _ptr(self.zw))
        return self

    def as_double(self):
        """Same stencil (indices and weights of this table) with FP64 weights, for FP64
        oracle runs on the identical discrete geometry."""
        import copy
        o = copy.copy(self)
        o.xw, o.yw, o.zw = (w.astype(np.float64) for w in (self.xw, self.yw, self.zw))
        return o

    def cstruct(self):
        return Inj(self.npts, self.kind, self.xi.ctypes.data, self.yi.ctypes.data, self.zi.ctypes.data,
                   self.xw.ctypes.data, self.yw.ctypes.data, self.zw.ctypes.data)


def inject(inj, src, nt, comp, it, out, s="f"):
    """Add source samples at step it (component comp) into out [ny][nx][nz]."""
    ny, nx, nz = out.shape[-3:]
    I = inj.cstruct()
    getattr(lib(), f"orc_inject_{s}")(C.byref(I), _ptr(np.ascontiguousarray(src, dtype=_dt(s))), nt, comp, it,
                                      _ptr(out), nx, ny, nz)
    return out


def extract(ext, u, nt, comp, it, rcv, s="f"):
    """rcv[comp][r][it] += receiver values of field u [ny][nx][nz]."""
    ny, nx, nz = u.shape[-3:]
    E = ext.cstruct()
    getattr(lib(), f"orc_extract_{s}")(C.byref(E), _ptr(np.ascontiguousarray(u, dtype=_dt(s))), nt, comp, it,
                                       _ptr(rcv), nx, ny, nz)
    return rcv


def dipole_to_strain(adj, p, dipaz, gl, s="f"):
    """misc.cpp dipole_to_strain on traces p [3][nrcv][nt] in place."""
    T = _dt(s)
    cT = C.c_float if s == "f" else C.c_double
    assert p.dtype == T and p.flags.c_contiguous
    dipaz = np.ascontiguousarray(np.asarray(dipaz, T).reshape(-1, 2))
    getattr(lib(), f"orc_dipole_to_strain_{s}")(int(adj), _ptr(p), _ptr(dipaz), p.shape[1], p.shape[2], cT(gl))
    return p


def make_par(nt, sub, dt, bc=(1, 1, 1, 1, 1, 1), taper=(0,) * 6, taper_strength=0.05, free_surface_stiffness=1.05,
             lag=1):
    """lag (test-only): 1 feeds the elastic/VTI imaging condition the previous step's
    adjoint derivatives, as the reference does; 0 the current step's."""
    return Par(nt, sub, dt, (C.c_int * 6)(*bc), (C.c_int * 6)(*taper), taper_strength, free_surface_stiffness, lag)


def n_snap_slots(nt, sub):
    return 1 + nt // sub if sub > 0 else 0


def propagate(physics, model, src, inj, ext, par, d, adj=False, grad=None, snaps=None, s="f"):
    """Run the oracle's propagate. model: converted parameters [np][ny][nx][nz].
    src: [ns][nsrc][nt]. Returns rcv [nrc][nr][nt] (forward) or None (adjoint; grad updated)."""
    T = _dt(s)
    cT = C.c_float if s == "f" else C.c_double
    ph = PHYS[physics] if isinstance(physics, str) else physics
    ny, nx, nz = model.shape[-3:]
    ncomp = 1 if ph == 0 else 3
    model = np.ascontiguousarray(model, dtype=T)
    src = np.ascontiguousarray(src, dtype=T)
    rcv = None
    if not adj:
        rcv = np.zeros((ncomp, ext.npts, par.nt), T)
    if snaps is None and par.sub > 0:
        snaps = np.zeros((n_snap_slots(par.nt, par.sub), ncomp, ny, nx, nz), T)
    I, E = inj.cstruct(), ext.cstruct()
    rc = getattr(lib(), f"orc_propagate_{s}")(ph, int(adj), _ptr(model), _ptr(src), _ptr(rcv), C.byref(I), C.byref(E),
                                              _ptr(grad), _ptr(snaps), C.byref(par), nx, ny, nz,
                                              cT(d[0]), cT(d[1]), cT(d[2]))
    if rc != 0:
        raise RuntimeError("oracle propagate failed")
    return rcv, snaps


def spatial(physics, curr, prev, model, bc, d, dt, free_surface_stiffness=1.05, s="d"):
    """The spatial operator of one time step as propagate applies it (volume terms and
    the SATs of the faces, before the source, the leapfrog update and the absorbing
    scaling): returns next."""
    T = _dt(s)
    cT = C.c_float if s == "f" else C.c_double
    ph = PHYS[physics] if isinstance(physics, str) else physics
    ny, nx, nz = model.shape[-3:]
    curr, prev, model = (np.ascontiguousarray(a, dtype=T) for a in (curr, prev, model))
    nxt = np.zeros_like(curr)
    getattr(lib(), f"orc_spatial_{s}")(ph, _ptr(curr), _ptr(prev), _ptr(nxt), _ptr(model), (C.c_int * 6)(*bc),
                                       nx, ny, nz, cT(d[0]), cT(d[1]), cT(d[2]), cT(dt), cT(free_surface_stiffness))
    return nxt


def grad_event(physics, model, snaps, itl, nt, sub, dt, adjf, derivf, grad, d, s="d"):
    """One imaging event of propagate's adjoint (grad accumulates in place): forward
    snapshots itl-1..itl+1 of the store `snaps`, the adjoint field adjf for the rho / K
    term and the derivatives of derivf for the lambda / mu / 1/rho terms (elastic/VTI;
    the acoustic event differentiates adjf)."""
    T = _dt(s)
    cT = C.c_float if s == "f" else C.c_double
    ph = PHYS[physics] if isinstance(physics, str) else physics
    ny, nx, nz = model.shape[-3:]
    for a in (model, snaps, adjf, derivf, grad):
        assert a.dtype == T and a.flags.c_contiguous
    getattr(lib(), f"orc_grad_event_{s}")(ph, _ptr(model), _ptr(snaps), snaps.shape[0], int(itl), int(nt), int(sub),
                                          cT(dt), _ptr(adjf), _ptr(derivf), _ptr(grad), nx, ny, nz,
                                          cT(d[0]), cT(d[1]), cT(d[2]))
    return grad


# ---------------------------------------------------------------------------------
# Shot operator (apply_forward / apply_jacobianT / FWI shot) of the restatement.

class OpPar(C.Structure):
    """orc_op_par: the reference's param fields read by the shot operator (double)."""
    _fields_ = [("nmodels", C.c_int), ("dt", C.c_double), ("courant", C.c_double), ("sub", C.c_int),
                ("nt", C.c_int), ("vmin", C.c_double), ("vmax", C.c_double), ("resampling", C.c_int),
                ("sinc_half_length", C.c_int), ("bc", C.c_int * 6), ("taper", C.c_int * 6),
                ("taper_strength", C.c_double), ("free_surface_stiffness", C.c_double), ("mt", C.c_int),
                ("seismotype", C.c_int), ("soft_clip", C.c_int)] + \
               [(k, C.c_double) for k in ("vpmin", "vpmax", "vsmin", "vsmax", "rhomin", "rhomax", "deltamin",
                                          "deltamax", "epsilonmin", "epsilonmax", "gammamin", "gammamax")] + \
               [("lag", C.c_int), ("gl", C.c_double)]


class ShotC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
               [(k, C.c_double) for k in ("dx", "dy", "dz", "ox", "oy", "oz")] + \
               [("nt_data", C.c_int), ("dt_data", C.c_double), ("sxyz", C.c_void_p), ("nrcv", C.c_int),
                ("rxyz", C.c_void_p), ("rdipaz", C.c_void_p)]


# defaults of the reference's param (src/include/param.hpp:8-71)
OP_DEFAULTS = dict(nmodels=3, dt=-1.0, courant=0.5, sub=0, nt=0, vmin=0.1, vmax=8.0, resampling=0,
                   sinc_half_length=11, bc=(1,) * 6, taper=(0,) * 6, taper_strength=0.05,
                   free_surface_stiffness=1.05, mt=0, seismotype=0, soft_clip=0, vpmin=0.2, vpmax=8.0, vsmin=0.1,
                   vsmax=5.0, rhomin=0.1, rhomax=8.0, deltamin=-0.5, deltamax=1.0, epsilonmin=0.0, epsilonmax=1.0,
                   gammamin=0.0, gammamax=1.0, lag=1, gl=0.0)


def op_par(**kw):
    v = dict(OP_DEFAULTS)
    v.update(kw)
    p = OpPar()
    for k, x in v.items():
        if k in ("bc", "taper"):
            setattr(p, k, (C.c_int * 6)(*x))
        else:
            setattr(p, k, x)
    return p


class ShotGeom:
    """Geometry of one shot with its position arrays kept alive in the variant's type."""

    def __init__(self, shape, d, o, nt_data, dt_data, sxyz, rxyz, s="f", rdipaz=None):
        T = _dt(s)
        self.s = s
        self.shape = tuple(shape)
        self.sxyz = np.ascontiguousarray(np.asarray(sxyz, T).reshape(3))
        self.rxyz = np.ascontiguousarray(np.asarray(rxyz, T).reshape(-1, 3))
        self.rdipaz = None if rdipaz is None else np.ascontiguousarray(np.asarray(rdipaz, T).reshape(-1, 2))
        ny, nx, nz = shape
        self.c = ShotC(nx, ny, nz, d[0], d[1], d[2], o[0], o[1], o[2], int(nt_data), float(dt_data),
                       self.sxyz.ctypes.data, self.rxyz.shape[0], self.rxyz.ctypes.data,
                       None if self.rdipaz is None else self.rdipaz.ctypes.data)
        self.nrcv = self.rxyz.shape[0]
        self.nt_data = int(nt_data)


def _ncomp(nmodels):
    """Wavefield (snapshot) components."""
    return 1 if nmodels == 2 else 3


def _nrcomp(par):
    """Data components: 1 for pressure and for DAS strain."""
    return 1 if (par.nmodels == 2 or par.gl > 0) else 3


def analyze(par, model, geom):
    """analyzeModel on a copy of model; returns (clipped model, par updated in place)."""
    T = _dt(geom.s)
    m = np.ascontiguousarray(model, dtype=T).copy()
    getattr(lib(), f"orc_analyze_{geom.s}")(C.byref(par), _ptr(m), C.byref(geom.c))
    return m


def convert(model, nmodels, s="f"):
    m = np.ascontiguousarray(model, dtype=_dt(s)).copy()
    n = m.size // nmodels
    getattr(lib(), f"orc_convert_{s}")(_ptr(m), C.c_longlong(n), nmodels)
    return m


def resample(method, adj, add, mod, dat, dm, dd, hl=11, s="f"):
    """Resampler between model traces mod [ntr][nm] (step dm) and data traces dat
    [ntr][nd] (step dd); fwd writes dat, adj writes mod (in place)."""
    cT = C.c_float if s == "f" else C.c_double
    ntr, nm = mod.shape
    nd = dat.shape[1]
    getattr(lib(), f"orc_resample_{s}")(method, int(adj), int(add), _ptr(mod), _ptr(dat), ntr, nm, cT(dm), nd, cT(dd), hl)


def ht(inv, x, dt, s="f"):
    cT = C.c_float if s == "f" else C.c_double
    ntr, nt = x.shape
    getattr(lib(), f"orc_ht_{s}")(int(inv), _ptr(x), ntr, nt, cT(dt))
    return x


def time_derivative(adj, x, dt, s="f"):
    cT = C.c_float if s == "f" else C.c_double
    out = np.zeros_like(x)
    ntr, nt = x.shape
    getattr(lib(), f"orc_dt_{s}")(int(adj), _ptr(x), _ptr(out), ntr, nt, cT(dt))
    return out


def _snap_buffer(par, model, geom):
    """Snapshot store sized from an analysis on a copy of par."""
    p2 = OpPar.from_buffer_copy(par)
    analyze(p2, model, geom)
    ny, nx, nz = geom.shape
    sub = p2.sub
    if sub <= 0:
        return None
    return np.zeros((n_snap_slots(p2.nt, sub), _ncomp(par.nmodels), ny, nx, nz), _dt(geom.s))


def apply_forward(par, geom, model, src, add=False, data=None):
    """apply_forward: raw model, wavelets src [nscomp][nt_data] -> data
    [nrcomp][nrcv][nt_data]. Returns (data, snaps); par is updated in place."""
    T = _dt(geom.s)
    model = np.ascontiguousarray(model, dtype=T)
    src = np.ascontiguousarray(src, dtype=T)
    if data is None:
        data = np.zeros((_nrcomp(par), geom.nrcv, geom.nt_data), T)
    snaps = _snap_buffer(par, model, geom)
    rc = getattr(lib(), f"orc_apply_forward_{geom.s}")(C.byref(par), C.byref(geom.c), int(add), _ptr(model), _ptr(src),
                                                      _ptr(data), _ptr(snaps))
    if rc != 0:
        raise RuntimeError("oracle apply_forward failed")
    return data, snaps


def apply_jacobianT(par, geom, model0, data, snaps, add=False, grad=None):
    T = _dt(geom.s)
    ny, nx, nz = geom.shape
    model0 = np.ascontiguousarray(model0, dtype=T)
    data = np.ascontiguousarray(data, dtype=T)
    if grad is None:
        grad = np.zeros((min(3, par.nmodels), ny, nx, nz), T)
    rc = getattr(lib(), f"orc_apply_jacobianT_{geom.s}")(C.byref(par), C.byref(geom.c), int(add), _ptr(grad),
                                                        _ptr(model0), _ptr(data), _ptr(snaps))
    if rc != 0:
        raise RuntimeError("oracle apply_jacobianT failed")
    return grad


def fwi_shot(par, geom, model, src, obs, grad=None):
    """Returns (misfit, residual, grad); grad accumulates."""
    T = _dt(geom.s)
    ny, nx, nz = geom.shape
    model = np.ascontiguousarray(model, dtype=T)
    src = np.ascontiguousarray(src, dtype=T)
    obs = np.ascontiguousarray(obs, dtype=T)
    resid = np.zeros_like(obs)
    if grad is None:
        grad = np.zeros((min(3, par.nmodels), ny, nx, nz), T)
    snaps = _snap_buffer(par, model, geom)
    mis = C.c_double()
    rc = getattr(lib(), f"orc_fwi_shot_{geom.s}")(C.byref(par), C.byref(geom.c), _ptr(model), _ptr(src), _ptr(obs),
                                                 _ptr(resid), _ptr(grad), _ptr(snaps), C.byref(mis))
    if rc != 0:
        raise RuntimeError("oracle fwi_shot failed")
    return mis.value, resid, grad
