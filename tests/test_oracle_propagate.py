"""Pin the oracle's propagate/compute_gradients composition (step order, initial
condition, injection timing, taper-twice, SAT/scale order, lagged adjoint
derivatives, snapshot-slot quirks) against a composition of the reference's own
compiled operators (oracle/_ref) in the step order of src/cpp/we_op.cpp.

The composition below follows we_op.cpp:
  acoustic  propagate 1607-1800, compute_gradients 1555-1605
  elastic   propagate  567-870,  compute_gradients  488-565
  vti       propagate 1218-1531, compute_gradients 1108-1216
with injection/extraction written directly from injector.cpp:84-140, 241-311."""
import numpy as np
import pytest

from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available("f") or not O.ref_available("d"),
                               reason="reference operator library not built (no reference checkout)")


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / np.linalg.norm(np.asarray(b, np.float64).ravel()))


def np_inject(inj, src, it, comp, out):
    """Add src samples at time it into out (one displacement/pressure component)."""
    nw = 6 if inj.kind else 3
    for r in range(inj.npts):
        xw, yw, zw = inj.xw[r, 0], inj.yw[r, 0], inj.zw[r, 0]
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    iy, ix, iz = inj.yi[r, a], inj.xi[r, b], inj.zi[r, c]
                    if inj.kind == 0:
                        v = yw[a] * xw[b] * zw[c] * src[comp, r, it]
                    else:
                        icx, icy, icz = {0: (0, 3, 4), 1: (3, 1, 5), 2: (4, 5, 2)}[comp]
                        v = (-xw[b + 3] * yw[a] * zw[c] * src[icx, r, it] - xw[b] * yw[a + 3] * zw[c] * src[icy, r, it]
                             - xw[b] * yw[a] * zw[c + 3] * src[icz, r, it])
                    out[iy, ix, iz] += v
    assert nw


def np_extract(ext, u, it, comp, rcv):
    for r in range(ext.npts):
        xw, yw, zw = ext.xw[r, 1], ext.yw[r, 1], ext.zw[r, 1]
        acc = u.dtype.type(0)
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    acc += yw[a] * xw[b] * zw[c] * u[ext.yi[r, a], ext.xi[r, b], ext.zi[r, c]]
        rcv[comp, r, it] = acc


def ref_propagate(physics, model, src, inj, ext, par, d, s, adj=False, grad=None, snaps=None):
    R = O.Ops("ref", s)
    T = np.float32 if s == "f" else np.float64
    dx, dy, dz = (T(v) for v in d)
    nt, sub, dt = par.nt, par.sub, T(par.dt)
    bc, tp, ts = list(par.bc), list(par.taper), T(par.taper_strength)
    shape = model.shape[-3:]
    ncomp = 1 if physics == "acoustic" else 3
    vti = physics == "vti"
    u = np.zeros((3, ncomp) + shape, T)
    prev, curr, nxt = u[0], u[1], u[2]
    rcv = None if adj else np.zeros((ncomp, ext.npts, nt), T)
    if snaps is None and sub > 0:
        snaps = np.zeros((O.n_snap_slots(nt, sub), ncomp) + shape, T)
    nslots = O.n_snap_slots(nt, sub) if sub > 0 else 0
    D1 = lambda a, f, h: R.d1(a, np.ascontiguousarray(f), np.zeros_like(f), h)
    hs = (dx, dy, dz)

    if physics == "acoustic":
        K, b = model[0], model[1]
        alpha = T(0.2508560249) / T(par.free_surface_stiffness)
        np_inject(inj, src, 0, 0, prev[0])
        prev[0] *= T(0.5) * K * dt * dt
    else:
        lam, mu, rho = model[0], model[1], model[2]
        for c in range(3):
            np_inject(inj, src, 0, c, prev[c])
        prev *= T(0.5) * dt * dt / rho
        dux = np.zeros((3, 3) + shape, T)  # dux[a][c] = D_a u_c (persists: lagged in the gradient)
    tmpgrad = None

    def grad_event(itl, adjf):
        sdt = T(sub) * dt
        fn = snaps[itl]
        fm = snaps[itl - 1] if itl > 0 else None
        fp = snaps[itl + 1] if itl + 1 < nslots else np.zeros_like(fn)
        last = nt // sub
        if physics == "acoustic":
            f = [D1(a, fn[0], hs[a]) for a in range(3)]
            g_ = [D1(a, adjf[0], hs[a]) for a in range(3)]
            dot = g_[0] * f[0] + g_[1] * f[1] + g_[2] * f[2]
            k2 = T(1.0) / (sdt * model[0] * model[0])
            if 0 < itl < last:
                grad[0] -= k2 * adjf[0] * (fp[0] - 2 * fn[0] + fm[0]); grad[1] += sdt * dot
            elif itl == 0:
                grad[0] -= k2 * adjf[0] * (fp[0] - fn[0]); grad[1] += T(0.5) * sdt * dot
            else:
                grad[0] -= k2 * adjf[0] * (-2 * fn[0] + fm[0]); grad[1] += sdt * dot
            return
        t = [D1(0, fn[0], dx), D1(1, fn[1], dy), D1(2, fn[2], dz), D1(0, fn[2], dx), D1(2, fn[0], dz),
             D1(0, fn[1], dx), D1(1, fn[0], dy), D1(1, fn[2], dy), D1(2, fn[1], dz)]
        A = dux  # lagged adjoint derivatives A[a][c]
        w = T(0.5) if itl == 0 else T(1)
        if not vti:
            grad[0] += w * sdt * (A[0][0] + A[1][1] + A[2][2]) * (t[0] + t[1] + t[2])
            grad[1] += w * sdt * ((A[2][0] + A[0][2]) * (t[3] + t[4]) + (A[1][0] + A[0][1]) * (t[5] + t[6])
                                  + (A[2][1] + A[1][2]) * (t[7] + t[8]) + 2 * A[0][0] * t[0] + 2 * A[1][1] * t[1]
                                  + 2 * A[2][2] * t[2])
        else:
            L, U, E, C13, C66 = model[0], model[1], model[4], model[3], model[5]
            dl = ((C13 + U) ** 2 - (L + U) ** 2) / (2 * (L + U) * (L + 2 * U))
            gm = (C66 - U) / (2 * U)
            v1 = np.sqrt(2 * (L + 2 * U) * (L + U) * dl + (L + U) ** 2)
            v2 = ((1 + 2 * dl) * L + (1 + 3 * dl) * U) / v1
            v3 = ((1 + 3 * dl) * L + (1 + 4 * dl) * U) / v1 - 1
            v4 = (A[0][0] + A[1][1]) * t[2] + A[2][2] * (t[0] + t[1])
            grad[0] += w * sdt * ((1 + 2 * E) * (A[0][0] + A[1][1]) * (t[0] + t[1]) + A[2][2] * t[2] + v2 * v4)
            grad[1] += w * sdt * (2 * (1 + 2 * E) * (A[0][0] * t[0] + A[1][1] * t[1]) + 2 * A[2][2] * t[2]
                                  + 4 * (E - gm) * (A[0][0] * t[1] + A[0][0] * t[0]) + (A[1][2] + A[2][1]) * (t[7] + t[8])
                                  + (A[0][2] + A[2][0]) * (t[3] + t[4]) + (1 + 2 * gm) * (A[0][1] + A[1][0]) * (t[5] + t[6])
                                  + v3 * v4)
        if 0 < itl < last:
            d2t = fp - 2 * fn + fm
        elif itl == 0:
            d2t = fp - fn
        else:
            d2t = -fn + fm
        grad[2] += T(1.0) / sdt * (adjf * d2t).sum(axis=0)

    for it in range(nt - 1):
        if sub > 0 and it % sub == 0 and grad is None:
            snaps[it // sub] = curr
        if grad is None:
            for c in range(ncomp):
                np_extract(ext, curr[c], it, c, rcv)
        if grad is not None and sub > 0 and (nt - 1 - it) % sub == 0 and it != 0:
            grad_event((nt - 1 - it) // sub, curr)
        if physics == "acoustic":
            R.d2(0, curr[0], nxt[0], b, dx, 1.0)
            R.d2(1, curr[0], nxt[0], b, dy, 1.0, add=True)
            R.d2(2, curr[0], nxt[0], b, dz, 1.0, add=True)
            np_inject(inj, src, it, 0, nxt[0])
            for f in range(6):
                R.asat(f, bc[f], curr[0], prev[0], nxt[0], model, d, dt, alpha)
            nxt[0] = dt * dt * nxt[0] * K + 2 * curr[0] - prev[0]
            R.ascale(nxt[0], model, d, dt, bc)
            R.taper(curr, tp, ts)
            R.taper(nxt, tp, ts)
        else:
            if not vti:
                bfield = {(c, a): (mu, 2.0 if a == c else 1.0) for c in range(3) for a in range(3)}
            else:
                c11b = (2 * (1 + 2 * model[4]) * mu).astype(T)
                bfield = {(0, 0): (c11b, 1.0), (0, 1): (model[5], 1.0), (0, 2): (mu, 1.0),
                          (1, 0): (model[5], 1.0), (1, 1): (c11b, 1.0), (1, 2): (mu, 1.0),
                          (2, 0): (mu, 1.0), (2, 1): (mu, 1.0), (2, 2): (mu, 2.0)}
            for c in range(3):
                for a in range(3):
                    bb, fa = bfield[(c, a)]
                    R.d2(a, curr[c], nxt[c], np.ascontiguousarray(bb), hs[a], fa, add=(a != 0))
            dux[0][0] = D1(0, curr[0], dx); dux[1][1] = D1(1, curr[1], dy); dux[2][2] = D1(2, curr[2], dz)
            if not vti:
                tmp = lam * (dux[0][0] + dux[1][1] + dux[2][2])
                R.d1(0, tmp, nxt[0], dx, add=True); R.d1(1, tmp, nxt[1], dy, add=True); R.d1(2, tmp, nxt[2], dz, add=True)
            else:
                E, C13, C66 = model[4], model[3], model[5]
                tmp = ((1 + 2 * E) * lam * dux[0][0] + ((1 + 2 * E) * (lam + 2 * mu) - 2 * C66) * dux[1][1] + C13 * dux[2][2]).astype(T)
                R.d1(0, tmp, nxt[0], dx, add=True)
                tmp = ((1 + 2 * E) * lam * dux[1][1] + ((1 + 2 * E) * (lam + 2 * mu) - 2 * C66) * dux[0][0] + C13 * dux[2][2]).astype(T)
                R.d1(1, tmp, nxt[1], dy, add=True)
                tmp = (C13 * (dux[0][0] + dux[1][1]) + lam * dux[2][2]).astype(T)
                R.d1(2, tmp, nxt[2], dz, add=True)
            gxy = np.ascontiguousarray(model[5]) if vti else mu
            dux[0][1] = D1(0, curr[1], dx); dux[0][2] = D1(0, curr[2], dx)
            R.d1(1, dux[0][1], nxt[0], dy, par=gxy, add=True)
            R.d1(2, dux[0][2], nxt[0], dy, par=mu, add=True)  # spacing dy as in we_op.cpp:678
            dux[1][0] = D1(1, curr[0], dy); dux[1][2] = D1(1, curr[2], dy)
            R.d1(0, dux[1][0], nxt[1], dx, par=gxy, add=True)
            R.d1(2, dux[1][2], nxt[1], dz, par=mu, add=True)
            dux[2][0] = D1(2, curr[0], dz); dux[2][1] = D1(2, curr[1], dz)
            R.d1(0, dux[2][0], nxt[2], dx, par=mu, add=True)
            R.d1(1, dux[2][1], nxt[2], dy, par=mu, add=True)
            for c in range(3):
                np_inject(inj, src, it, c, nxt[c])
            for f in range(6):
                R.esat(f, bc[f], vti, curr, prev, nxt, model, d, dt)
            nxt[:] = dt * dt * nxt / rho + 2 * curr - prev
            R.escale(vti, nxt, model, d, dt, bc)
            R.taper(curr, tp, ts)
            R.taper(nxt, tp, ts)
        prev, curr, nxt = curr, nxt, prev
    if sub > 0 and (nt - 1) % sub == 0 and grad is None:
        snaps[(nt - 1) // sub] = curr
    if grad is None:
        for c in range(ncomp):
            np_extract(ext, curr[c], nt - 1, c, rcv)
    if grad is not None and sub > 0:
        grad_event(0, curr)
    return rcv, snaps


def tiny_case(physics, s, kind, shape=(14, 15, 16), nt=40, seed=11):
    T = np.float32 if s == "f" else np.float64
    rng = np.random.default_rng(seed)
    h = (0.01, 0.011, 0.009)
    vp = rng.uniform(2.0, 3.0, shape)
    if physics == "acoustic":
        rho = rng.uniform(1.8, 2.4, shape)
        model = np.stack([rho * vp * vp, 1 / rho])
    else:
        vs = vp / rng.uniform(1.7, 2.0, shape)
        rho = rng.uniform(1.8, 2.4, shape)
        lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
        if physics == "elastic":
            model = np.stack([lam, mu, rho])
        else:
            eps, dl, gm = rng.uniform(0, .2, shape), rng.uniform(0, .1, shape), rng.uniform(0, .1, shape)
            c13 = np.sqrt(2 * (lam + 2 * mu) * (lam + mu) * dl + (lam + mu) ** 2) - mu
            model = np.stack([lam, mu, rho, c13, eps, (1 + 2 * gm) * mu])
    model = np.ascontiguousarray(model.astype(T))
    ns = 1 if physics == "acoustic" else (6 if kind else 3)
    t = np.arange(nt) * 0.0008
    w = (1 - 2 * (np.pi * 60 * (t - 0.02)) ** 2) * np.exp(-(np.pi * 60 * (t - 0.02)) ** 2)
    src = np.ascontiguousarray((rng.uniform(0.5, 1.5, (ns, 1, 1)) * w).astype(T))
    ny, nx, nz = shape
    sloc = np.array([[6.3 * h[0], 7.6 * h[1], 5.2 * h[2]]])
    rloc = np.array([[1.0 * h[0], 2.0 * h[1], 0.0], [12.4 * h[0], 10.3 * h[1], 14.8 * h[2]], [7 * h[0], 0.0, 9.1 * h[2]],
                     [0.0, 12.2 * h[1], 3.3 * h[2]], [14 * h[0], 13 * h[1], 15 * h[2]]])
    inj = O.Injector(kind, sloc, shape, h, (0, 0, 0), s)
    ext = O.Injector(0, rloc, shape, h, (0, 0, 0), s)
    return model, src, inj, ext, h


BCS = [((1, 2, 2, 2, 2, 2), (0, 3, 3, 3, 3, 3)), ((2, 1, 1, 2, 2, 1), (2, 0, 0, 3, 2, 0)), ((1, 1, 1, 1, 1, 1), (0,) * 6)]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("physics,kind", [("acoustic", 0), ("elastic", 1), ("elastic", 0), ("vti", 1)])
@pytest.mark.parametrize("bci", range(len(BCS)))
def test_forward_composition_matches_reference_operators(s, physics, kind, bci):
    bc, tp = BCS[bci]
    if physics == "acoustic" and bci == 1:
        bc = (3, 1, 2, 3, 1, 2)
    model, src, inj, ext, h = tiny_case(physics, s, kind)
    par = O.make_par(40, 0, 0.0008, bc=bc, taper=tp, taper_strength=0.3)
    got, _ = O.propagate(physics, model, src, inj, ext, par, h, s=s)
    want, _ = ref_propagate(physics, model, src, inj, ext, par, h, s)
    tol = 2e-5 if s == "f" else 1e-11
    assert np.abs(want).max() > 0
    assert rel(got, want) < tol, rel(got, want)


@needs_ref
@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
@pytest.mark.parametrize("nt,sub", [(41, 4), (40, 3), (43, 3)])
def test_gradient_composition_matches_reference_operators(physics, nt, sub):
    """Imaging condition incl. the lagged adjoint derivatives, the first/middle/last
    sample branches and the zero snapshot slot, in FP64."""
    s = "d"
    model, src, inj, ext, h = tiny_case(physics, s, 1 if physics != "acoustic" else 0, nt=nt)
    par = O.make_par(nt, sub, 0.0008, bc=(1, 2, 2, 2, 2, 2), taper=(0, 3, 3, 3, 3, 3), taper_strength=0.3)
    rcv, snaps = O.propagate(physics, model, src, inj, ext, par, h, s=s)
    rcv2, snaps2 = ref_propagate(physics, model, src, inj, ext, par, h, s)
    assert rel(rcv, rcv2) < 1e-11
    res = np.ascontiguousarray(rcv[:, :, ::-1])
    ng = 2 if physics == "acoustic" else 3
    g1 = np.zeros((ng,) + model.shape[-3:])
    g2 = np.zeros_like(g1)
    O.propagate(physics, model, res, ext, inj, par, h, adj=True, grad=g1, snaps=snaps, s=s)
    ref_propagate(physics, model, res, ext, inj, par, h, s, adj=True, grad=g2, snaps=snaps2)
    for k in range(ng):
        assert np.abs(g2[k]).max() > 0
        assert rel(g1[k], g2[k]) < 1e-10, (k, rel(g1[k], g2[k]))
