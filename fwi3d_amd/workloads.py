"""Synthetic workloads: models, wavelets and acquisition geometry of the benchmark
configurations, in the reference's units (km, s, km/s, g/cc; src/include/param.hpp:28-31)
and layouts ([param][y][x][z], traces [comp][point][t]).

Model conversions follow the reference's single-precision formulas
(nl_we_op_a / _e / _vti ::convert_model, src/cpp/we_op.cpp:1535-1552, 448-466, 1083-1106).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


def ricker(freq, nt, dt, delay=None):
    """w(t) = (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - 1.2/f."""
    t = np.arange(nt, dtype=np.float64) * dt
    tau = t - (1.2 / freq if delay is None else delay)
    a = (math.pi * freq * tau) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(F32)


def gardner(vp):
    """Gardner density in g/cc from vp in km/s: 0.31 * (1000 vp)^0.25."""
    return (0.31 * (1000.0 * np.asarray(vp, np.float64)) ** 0.25).astype(F32)


def acoustic_convert(vp, rho):
    """(vp, rho) -> (K, 1/rho) with the reference's float arithmetic."""
    vp = vp.astype(F32)
    rho = rho.astype(F32)
    K = (rho * vp) * vp
    b = (1.0 / rho.astype(np.float64)).astype(F32)
    return np.ascontiguousarray(np.stack([K, b]))


def elastic_convert(vp, vs, rho):
    """(vp, vs, rho) -> (lambda, mu, rho)."""
    vp, vs, rho = vp.astype(F32), vs.astype(F32), rho.astype(F32)
    lam = rho * (vp * vp - F32(2) * vs * vs)
    mu = (rho * vs) * vs
    return np.ascontiguousarray(np.stack([lam, mu, rho]))


def vti_convert(vp, vs, rho, delta, eps, gamma):
    """(vp, vs, rho, delta, eps, gamma) -> (lambda, mu, rho, c13, eps, c66)."""
    vp, vs, rho = vp.astype(F32), vs.astype(F32), rho.astype(F32)
    delta, eps, gamma = delta.astype(F32), eps.astype(F32), gamma.astype(F32)
    lam = rho * (vp * vp - F32(2) * vs * vs)
    mu = (rho * vs) * vs
    c13 = (np.sqrt((F32(2) * (lam + F32(2) * mu) * (lam + mu) * delta + (lam + mu) * (lam + mu)).astype(np.float64))
           .astype(F32) - mu)
    c66 = (F32(1) + F32(2) * gamma) * mu
    return np.ascontiguousarray(np.stack([lam, mu, rho, c13, eps, c66]))


@dataclass
class Shot:
    """One shot: model, wavelet(s), geometry and propagation settings."""
    physics: str
    model: np.ndarray            # converted parameters [np][ny][nx][nz]
    raw: np.ndarray              # (vp, rho) / (vp, vs, rho[, delta, eps, gamma])
    src: np.ndarray              # [ns][1][nt]
    src_loc: np.ndarray          # [1][3] (x, y, z)
    src_kind: int                # 0 delta_m3, 1 ddelta_m3
    rcv_loc: np.ndarray          # [nr][3]
    d: tuple                     # (dx, dy, dz)
    nt: int
    dt: float
    sub: int = 0
    bc: tuple = (1, 2, 2, 2, 2, 2)
    taper: tuple = (0, 20, 20, 20, 20, 20)
    taper_strength: float = 0.05
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        # C order: device entry points take bare pointers ([param][y][x][z], z fastest)
        self.model = np.ascontiguousarray(self.model)
        self.raw = np.ascontiguousarray(self.raw)

    @property
    def shape(self):
        return self.model.shape[-3:]


def _surface_grid(cx, cy, z, n, spacing):
    off = (np.arange(n) - (n - 1) / 2.0) * spacing
    xs, ys = np.meshgrid(cx + off, cy + off, indexing="xy")
    return np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, z)], axis=1).astype(F32)


def config_a(n=100, nt=1000, dt=0.001, h=0.01):
    """Acoustic variable-density forward, 100^3, 15 Hz Ricker at z=50 m, 10x10
    receivers at z=20 m with 50 m spacing; vp 1.5->2.5 km/s with depth, Gardner rho."""
    nz = nx = ny = n
    vz = 1.5 + 1.0 * np.arange(nz) / (nz - 1)
    vp = np.broadcast_to(vz, (ny, nx, nz)).astype(F32)
    rho = gardner(vp)
    cx = (nx - 1) * h / 2
    cy = (ny - 1) * h / 2
    src = ricker(15.0, nt, dt)[None, None, :]
    rcv = _surface_grid(cx, cy, 0.02, 10, 0.05)
    return Shot("acoustic", acoustic_convert(vp, rho), np.stack([vp, rho]), src,
                np.array([[cx, cy, 0.05]], F32), 0, rcv, (h, h, h), nt, dt,
                meta={"name": "acoustic_fwd_100^3", "freq": 15.0})


def layered_elastic(n, h=0.01, anomaly=True, nlayers=5):
    """5 flat layers (boundaries at iz = round(k nz/5)), vp 2.0 -> 3.5 km/s, optional
    +5% Gaussian vp anomaly of radius 300 m at the centre; vs = vp/1.8, Gardner rho."""
    nz = nx = ny = n
    vals = np.linspace(2.0, 3.5, nlayers)
    iz = np.arange(nz)
    bounds = [round(k * nz / nlayers) for k in range(nlayers + 1)]
    vz = np.empty(nz)
    for k in range(nlayers):
        vz[bounds[k]:bounds[k + 1]] = vals[k]
    vp = np.broadcast_to(vz, (ny, nx, nz)).astype(np.float64).copy()
    if anomaly:
        c = (n - 1) * h / 2
        z = iz * h
        x = np.arange(nx) * h
        y = np.arange(ny) * h
        r2 = (y[:, None, None] - c) ** 2 + (x[None, :, None] - c) ** 2 + (z[None, None, :] - c) ** 2
        vp = vp * (1 + 0.05 * np.exp(-r2 / 0.3 ** 2))
    vp = vp.astype(F32)
    vs = (vp / F32(1.8)).astype(F32)
    rho = gardner(vp)
    return vp, vs, rho


def config_b(n=200, nt=2000, dt=0.0008, h=0.01, anomaly=True, sub=4, nrec=40):
    """Elastic isotropic forward + adjoint gradient: explosive (moment tensor) 12 Hz
    Ricker at z=30 m, 40x40 three-component receivers at z=20 m (50 m spacing)."""
    vp, vs, rho = layered_elastic(n, h, anomaly)
    c = (n - 1) * h / 2
    nrec = min(nrec, int(((n - 1) * h - 0.04) / 0.05) + 1)  # the receiver spread stays inside small grids
    w = ricker(12.0, nt, dt)
    src = np.zeros((6, 1, nt), F32)
    src[0, 0] = w  # mxx
    src[1, 0] = w  # myy
    src[2, 0] = w  # mzz
    rcv = _surface_grid(c, c, 0.02, nrec, 0.05)
    return Shot("elastic", elastic_convert(vp, vs, rho), np.stack([vp, vs, rho]), src,
                np.array([[c, c, 0.03]], F32), 1, rcv, (h, h, h), nt, dt, sub=sub,
                meta={"name": f"elastic_grad_{n}^3", "freq": 12.0})


def smooth_random(shape, sigma, seed):
    """Seeded white noise, Gaussian-filtered (sigma in cells), scaled to [0, 1]."""
    from scipy.ndimage import gaussian_filter
    r = np.random.default_rng(seed).standard_normal(shape)
    s = gaussian_filter(r, sigma, mode="nearest")
    s = (s - s.min()) / max(s.max() - s.min(), 1e-30)
    return s


def vti_model(shape, h=0.01, seed=0):
    """VTI model: vp linear in depth 1.8 -> 4.0 km/s, eps in [0, 0.2] and delta in
    [0, 0.1] smooth random fields, gamma 0, vs = vp/2, Gardner rho."""
    ny, nx, nz = shape
    vz = 1.8 + 2.2 * np.arange(nz) / (nz - 1)
    vp = np.broadcast_to(vz, shape).astype(F32)
    vs = (vp / F32(2)).astype(F32)
    rho = gardner(vp)
    eps = (0.2 * smooth_random(shape, 5, 1000 + seed)).astype(F32)
    dlt = (0.1 * smooth_random(shape, 5, 2000 + seed)).astype(F32)
    gam = np.zeros(shape, F32)
    return vp, vs, rho, dlt, eps, gam


def config_c_shot(grid_index, shot_index, nt=1500, dt=0.0008, h=0.01):
    """One VTI shot of the batched configuration: grids (nz,nx,ny) = (128,128,128),
    (200,160,160), (96,192,192); 15 Hz explosive source at z=20 m on a 4x3 surface
    grid over the untapered interior; 16x16 receivers at z=20 m."""
    grids = [(128, 128, 128), (200, 160, 160), (96, 192, 192)]
    nz, nx, ny = grids[grid_index]
    raw = vti_model((ny, nx, nz), h, seed=grid_index)
    tap = 20
    xs = np.linspace((tap + 5) * h, (nx - 1 - tap - 5) * h, 4)
    ys = np.linspace((tap + 5) * h, (ny - 1 - tap - 5) * h, 3)
    pos = [(x, y) for y in ys for x in xs][shot_index]
    w = ricker(15.0, nt, dt)
    src = np.zeros((6, 1, nt), F32)
    src[0, 0] = src[1, 0] = src[2, 0] = w
    cx, cy = (nx - 1) * h / 2, (ny - 1) * h / 2
    sp = min((nx - 2 * tap) * h, (ny - 2 * tap) * h) / 16
    rcv = _surface_grid(cx, cy, 0.02, 16, sp)
    return Shot("vti", vti_convert(*raw), np.stack(raw), src, np.array([[pos[0], pos[1], 0.02]], F32), 1, rcv,
                (h, h, h), nt, dt, meta={"name": f"vti_{nz}x{nx}x{ny}", "grid": grid_index, "freq": 15.0})


# ---------------------------------------------------------------------------------
# Serving stream: a seeded mix of modelling and gradient requests.

@dataclass
class Request:
    """One request of the serving stream; its model is generated from `seed`."""
    index: int
    kind: str        # "model" or "gradient"
    physics: str     # "acoustic", "elastic", "vti"
    n: int           # cubic grid n^3
    nt: int
    dt: float
    seed: int


def serve_mix(count, seed=42):
    """60% acoustic modelling on n^3 grids, n ~ U{100..200}; 30% elastic shot gradients,
    n ~ U{128..256}; 10% VTI modelling on 160^3; nt ~ U{1000..3000}. dt 1 ms (acoustic)
    or 0.8 ms (elastic, VTI), as in the acoustic / elastic / VTI configurations."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        u = rng.random()
        nt = int(rng.integers(1000, 3001))
        s = int(rng.integers(0, 2**31 - 1))
        if u < 0.6:
            out.append(Request(i, "model", "acoustic", int(rng.integers(100, 201)), nt, 0.001, s))
        elif u < 0.9:
            out.append(Request(i, "gradient", "elastic", int(rng.integers(128, 257)), nt, 0.0008, s))
        else:
            out.append(Request(i, "model", "vti", 160, nt, 0.0008, s))
    return out


def poisson_arrivals(count, rate, seed=42):
    """Arrival times (s) of a Poisson stream of `rate` requests per second."""
    rng = np.random.default_rng(seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def cell_updates(req):
    """Cell-updates of a request: n^3 (nt-1), twice for a gradient (forward + adjoint)."""
    return float(req.n) ** 3 * (req.nt - 1) * (2 if req.kind == "gradient" else 1)
