// ORACLE — test infrastructure only (see sbp.hpp). CPU restatement of the trace-
// domain and model-domain steps around propagate: the operator's apply_forward /
// apply_jacobianT and the FWI objective of one shot.
//
// Follows:
//   analyzeModel ............ src/cpp/we_op.cpp:234-446
//   convert_model ........... src/cpp/we_op.cpp:448-466 (e), 1083-1106 (vti), 1535-1552 (a)
//   sinc / linear resampler . src/cpp/operator.cpp:419-520
//   applyHt, Dt ............. src/cpp/misc.cpp:44-94
//   apply_forward ........... src/cpp/we_op.cpp:872-946
//   apply_jacobianT ......... src/cpp/we_op.cpp:948-1078
//   residual + misfit ....... src/cpp/optimization.cpp:34-197, src/include/optimization.hpp:403-424
#pragma once
#include "wave.hpp"

namespace orc {

template <class T>
struct OpParams {
    int nmodels = 3;
    T dt = -1, courant = (T)0.5;
    int sub = 0;
    int resampling = 0;  // 0 sinc, 1 linear
    int sinc_half_length = 11;
    int bc[6] = {1, 1, 1, 1, 1, 1};
    int taper[6] = {0, 0, 0, 0, 0, 0};
    T taper_strength = (T)0.05, free_surface_stiffness = (T)1.05;
    int mt = 0, seismotype = 0, soft_clip = 0;
    T vpmin = (T)0.2, vpmax = 8, vsmin = (T)0.1, vsmax = 5, rhomin = (T)0.1, rhomax = 8;
    T deltamin = (T)-0.5, deltamax = 1, epsilonmin = 0, epsilonmax = 1, gammamin = 0, gammamax = 1;
    int lag = 1;  // test-only: lagged (1, the reference) or lag-free (0) adjoint derivatives
    T gl = 0;     // DAS gauge length (> 0: dipole receivers and strain data, elastic/VTI)
    // outputs of analyze_model
    int nt = 0;
    T vmax = 0, vmin = 0;
};

// Clip and check the raw model, and derive dt, nt, sub (we_op.cpp:234-446).
// model: [nmodels][N] raw parameters (modified in place); data axis (nt_data, dt_data).
template <class T>
void analyze_model(T* model, const Grid& g, T dx, T dy, T dz, int nt_data, T dt_data, OpParams<T>& par) {
    const size_t N = g.n();
    auto clip = [&](int k, T lo, T hi) {
        T* m = model + (size_t)k * N;
        for (size_t i = 0; i < N; i++) m[i] = std::max(lo, std::min(hi, m[i]));
    };
    auto mn = [&](int k) { T v = model[(size_t)k * N]; for (size_t i = 1; i < N; i++) v = std::min(v, model[(size_t)k * N + i]); return v; };
    auto mx = [&](int k) { T v = model[(size_t)k * N]; for (size_t i = 1; i < N; i++) v = std::max(v, model[(size_t)k * N + i]); return v; };
    if (par.nmodels >= 3) {
        if (par.nmodels == 6) {
            // Thomsen delta admissibility, over the first nx*nz points only (we_op.cpp:273)
            const T* vp = model;
            const T* vs = model + N;
            T* del = model + 3 * N;
            for (int i = 0; i < g.nx * g.nz; i++) {
                T v1 = (T)(-0.5 * (1 - vs[i] * vs[i] / (vp[i] * vp[i])));
                T v2 = (T)(2.0 / (vp[i] * vp[i] / (vs[i] * vs[i]) - 1));
                if (del[i] < v1) del[i] = (T)(0.9999 * v1);
                else if (del[i] > v2) del[i] = (T)(0.9999 * v2);
            }
        }
        if (!par.soft_clip) {
            clip(0, par.vpmin, par.vpmax);
            clip(1, par.vsmin, par.vsmax);
            clip(2, par.rhomin, par.rhomax);
            if (par.nmodels == 6) {
                clip(3, par.deltamin, par.deltamax);
                clip(4, par.epsilonmin, par.epsilonmax);
                clip(5, par.gammamin, par.gammamax);
            }
            for (size_t i = 0; i < N; i++)
                if ((double)model[i] < std::sqrt(2.0) * (double)model[N + i]) model[N + i] = (T)((double)model[i] / std::sqrt(2.0001));
        }
        par.vmax = mx(0);
        par.vmin = mn(1);
        if (par.nmodels == 6) {
            T m = 0;
            for (size_t i = 0; i < N; i++) m = std::max(m, (T)((double)model[i] * std::sqrt((double)(1 + 2 * model[4 * N + i]))));
            par.vmax = std::max(m, par.vmax);
        }
    } else {
        clip(0, par.vpmin, par.vpmax);
        clip(1, par.rhomin, par.rhomax);
        par.vmax = mx(0);
        par.vmin = mn(0);
    }
    const T tmax = (nt_data - 1) * dt_data;
    const T hmin = std::min(std::min(dx, dy), dz);
    if (par.dt == 0) {
        par.dt = par.courant * hmin / par.vmax;
    } else if (par.dt > 0) {
        par.dt = std::min(par.dt, dt_data);
        par.courant = par.dt * par.vmax / hmin;
    } else {
        par.dt = par.courant * hmin / par.vmax;
        if (dt_data > par.dt) par.dt = (T)((double)dt_data / std::ceil((double)(dt_data / par.dt)));
        else par.dt = dt_data;
        par.courant = par.dt * par.vmax / hmin;
    }
    par.nt = (int)(std::round((double)(tmax / par.dt)) + 1);
    if (par.sub < 0) par.sub = (int)std::round((double)(dt_data / par.dt));
}

// Raw parameters -> propagation parameters (forward conversion).
template <class T>
void convert_model(T* m, size_t n, int nmodels) {
    if (nmodels == 2) {
        for (size_t i = 0; i < n; i++) {
            m[i] = m[n + i] * m[i] * m[i];
            m[n + i] = (T)(1.0 / (double)m[n + i]);
        }
        return;
    }
    for (size_t i = 0; i < n; i++) {
        m[i] = m[2 * n + i] * (m[i] * m[i] - 2 * m[n + i] * m[n + i]);
        m[n + i] = m[2 * n + i] * m[n + i] * m[n + i];
        if (nmodels == 6) {
            m[3 * n + i] = (T)(std::sqrt((double)(2 * (m[i] + 2 * m[n + i]) * (m[i] + m[n + i]) * m[3 * n + i] +
                                                  (m[i] + m[n + i]) * (m[i] + m[n + i]))) - (double)m[n + i]);
            m[5 * n + i] = (1 + 2 * m[5 * n + i]) * m[n + i];
        }
    }
}

// Windowed-sinc resampler between the time axes (nm, dm) and (nd, dd), ntr traces:
// forward maps model (nm) to data (nd); adjoint data to model. The tap factors follow
// the reference's promotions: the sinc argument is rounded to T and sin() is the
// double one (no std overloads in scope there); the cosine window is double; each
// product is added through a double add (operator.cpp:419-422, 474-520).
template <class T>
inline T sinc_fn(T x) { return x == (T)0 ? (T)1 : (T)(std::sin((double)x) / (double)x); }

template <class T>
inline double sinc_win(T t, int itm, T dm, int hl, T a) {
    return (double)a + (double)(1 - a) * std::cos(M_PI * (double)((T)itm - t / dm) / hl);
}

template <class T>
void sinc_resample(bool adj, bool add, T* mod, T* dat, int ntr, int nm, T dm, int nd, T dd, int hl, T a = (T)0.5) {
    if (adj && !add) std::fill(mod, mod + (size_t)ntr * nm, (T)0);
    for (int ix = 0; ix < ntr; ix++)
        for (int itd = 0; itd < nd; itd++) {
            const T t = itd * dd;
            const int itm0 = (int)std::floor((double)(t / dm));
            T& out = dat[(size_t)ix * nd + itd];
            if (!adj) out = add ? out : (T)0;
            for (int itm = std::max(0, itm0 - hl + 1); itm <= std::min(nm - 1, itm0 + hl); itm++) {
                const T s = sinc_fn<T>((T)(M_PI * (double)(t - itm * dm) / (double)dm));
                const double w = sinc_win<T>(t, itm, dm, hl, a);
                if (!adj) out = (T)((double)out + (double)(mod[(size_t)ix * nm + itm] * s) * w);
                else {
                    T& m = mod[(size_t)ix * nm + itm];
                    m = (T)((double)m + (double)(out * s) * w);
                }
            }
        }
}

template <class T>
void linear_resample(bool adj, bool add, T* mod, T* dat, int ntr, int nm, T dm, int nd, T dd) {
    if (adj && !add) std::fill(mod, mod + (size_t)ntr * nm, (T)0);
    for (int ix = 0; ix < ntr; ix++)
        for (int it = 0; it < nd; it++) {
            const T q = it * dd / dm;
            int i = std::min(nm - 1, (int)std::floor((double)q));
            int j = std::min(nm - 1, (int)std::ceil((double)q));
            T wj = (it * dd - i * dm) / dm;
            T wi = 1 - wj;
            if (!adj) {
                T& o = dat[(size_t)ix * nd + it];
                o = (add ? o : (T)0) + wi * mod[(size_t)ix * nm + i] + wj * mod[(size_t)ix * nm + j];
            } else {
                mod[(size_t)ix * nm + i] += wi * dat[(size_t)ix * nd + it];
                mod[(size_t)ix * nm + j] += wj * dat[(size_t)ix * nd + it];
            }
        }
}

// Trapezoid time quadrature (or its inverse), in place allowed.
template <class T>
void apply_ht(bool inv, T* x, int ntr, int nt, T dt) {
    for (int ix = 0; ix < ntr; ix++) {
        T* r = x + (size_t)ix * nt;
        for (int it = 0; it < nt; it++) {
            const bool end = (it == 0 || it == nt - 1);
            if (!inv) r[it] = end ? (T)((double)r[it] * 0.5 * (double)dt) : r[it] * dt;
            else r[it] = end ? (T)((double)r[it] / (0.5 * (double)dt)) : r[it] / dt;
        }
    }
}

// Centred time derivative with one-sided ends, or its adjoint (misc.cpp:70-94).
template <class T>
void dt_op(bool adj, const T* in, T* out, int ntr, int nt, T dt) {
    for (int ix = 0; ix < ntr; ix++) {
        const T* a = in + (size_t)ix * nt;
        T* o = out + (size_t)ix * nt;
        if (!adj) {
            o[0] = (a[1] - a[0]) / dt;
            for (int it = 1; it < nt - 1; it++) o[it] = (a[it + 1] - a[it - 1]) / (2 * dt);
            o[nt - 1] = (a[nt - 1] - a[nt - 2]) / dt;
        } else {
            o[0] = (a[1] + a[0]) / dt;
            for (int it = 1; it < nt - 1; it++) o[it] = (a[it + 1] - a[it - 1]) / (2 * dt);
            o[nt - 1] = (-a[nt - 1] - a[nt - 2]) / dt;
        }
    }
}

template <class T>
struct ShotGeom {
    Grid g;
    T dx, dy, dz, ox, oy, oz;
    int nt_data;
    T dt_data;
    const T* sxyz;  // [3]
    int nrcv;
    const T* rxyz;  // [nrcv][3]
    const T* rdipaz = nullptr;  // [nrcv][2] fiber dip and azimuth of DAS channels (gl > 0)
};

// Per-physics overrides of analyzeGeometry (we_op.cpp:11-18).
template <class T>
void normalize_params(OpParams<T>& p) {
    if (p.nmodels == 2) {
        p.mt = 0;
        p.gl = 0;
        p.soft_clip = 0;
        p.free_surface_stiffness = std::max(p.free_surface_stiffness, (T)1.0);
    }
    for (int f = 0; f < 6; f++) p.bc[f] = std::max(0, p.bc[f]);
}

template <class T>
PropParams<T> prop_params(const OpParams<T>& p) {
    PropParams<T> q;
    q.nt = p.nt;
    q.sub = p.sub;
    q.dt = p.dt;
    for (int i = 0; i < 6; i++) { q.bc[i] = p.bc[i]; q.taper[i] = p.taper[i]; }
    q.taper_strength = p.taper_strength;
    q.free_surface_stiffness = p.free_surface_stiffness;
    q.lag = p.lag;
    return q;
}

// Components of the data (nrcomp: 1 for pressure and for DAS strain) and of the
// wavefield (and of the extracted traces before the strain conversion).
template <class T>
inline int n_comp(const OpParams<T>& p) { return (p.nmodels == 2 || p.gl > 0) ? 1 : 3; }
template <class T>
inline int n_field(const OpParams<T>& p) { return p.nmodels == 2 ? 1 : 3; }

// Receiver injector of a shot: delta_m3 points or, with gl > 0, dipole_m3 channels.
template <class T>
Injector<T> receiver_injector(const OpParams<T>& par, const ShotGeom<T>& G) {
    if (par.gl <= 0) return make_injector<T>(0, G.rxyz, G.nrcv, G.g, G.dx, G.dy, G.dz, G.ox, G.oy, G.oz);
    if (!G.rdipaz) throw std::runtime_error("DAS channels need their dip and azimuth");
    std::vector<T> loc5((size_t)5 * G.nrcv);
    for (int r = 0; r < G.nrcv; r++) {
        for (int k = 0; k < 3; k++) loc5[5 * r + k] = G.rxyz[3 * r + k];
        loc5[5 * r + 3] = G.rdipaz[2 * r];
        loc5[5 * r + 4] = G.rdipaz[2 * r + 1];
    }
    return make_dipole<T>(loc5.data(), G.nrcv, par.gl, G.g, G.dx, G.dy, G.dz, G.ox, G.oy, G.oz);
}

// dipole_to_strain (src/cpp/misc.cpp:11-42) on traces p [3][nrcv][nt]: forward projects
// the end-to-end difference on the fiber and divides by gl into component 0 (zeroing
// 1 and 2); adjoint spreads component 0 back onto the three components.
template <class T>
void dipole_to_strain(bool adj, T* p, const T* dipaz, int nrcv, int nt, T gl) {
    for (int r = 0; r < nrcv; r++) {
        const T cd = (T)std::cos((double)dipaz[2 * r]), sd = (T)std::sin((double)dipaz[2 * r]);
        const T ca = (T)std::cos((double)dipaz[2 * r + 1]), sa = (T)std::sin((double)dipaz[2 * r + 1]);
        T* p0 = p + (size_t)r * nt;
        T* p1 = p + ((size_t)nrcv + r) * nt;
        T* p2 = p + ((size_t)2 * nrcv + r) * nt;
        for (int it = 0; it < nt; it++) {
            if (!adj) {
                p0[it] = (cd * ca * p0[it] + cd * sa * p1[it] + sd * p2[it]) / gl;
                p1[it] = 0;
                p2[it] = 0;
            } else {
                p2[it] = sd * p0[it] / gl;
                p1[it] = cd * sa * p0[it] / gl;
                p0[it] *= cd * ca / gl;
            }
        }
    }
}
template <class T>
inline int n_scomp(const OpParams<T>& p) { return p.nmodels == 2 ? 1 : (p.mt ? 6 : 3); }

// apply_forward of one shot (we_op.cpp:872-946). src: [nscomp][nt_data] wavelets; data:
// [nrcomp][nrcv][nt_data] (overwritten or added). par is updated (analysis). When
// par.sub > 0 after analysis, the snapshots go to `snaps`
// ([n_snap_slots(nt,sub)][ncomp][N], caller-sized from a prior analysis).
template <class T>
void apply_forward(OpParams<T>& par, const ShotGeom<T>& G, const T* raw_model, const T* src, bool add, T* data,
                   T* snaps) {
    normalize_params(par);
    const size_t N = G.g.n();
    std::vector<T> model(raw_model, raw_model + par.nmodels * N);
    analyze_model(model.data(), G.g, G.dx, G.dy, G.dz, G.nt_data, G.dt_data, par);
    convert_model(model.data(), N, par.nmodels);
    const int nt = par.nt;
    const T alpha = par.dt / G.dt_data;
    const bool acoustic = par.nmodels == 2;
    const int nscomp = n_scomp(par);
    const int nrcomp = n_comp(par);
    const int nfc = n_field(par);
    std::vector<T> s(nscomp * (size_t)nt), r((size_t)nfc * G.nrcv * nt, 0);
    std::vector<T> srcc(src, src + (size_t)nscomp * G.nt_data);
    if (par.resampling == 1) linear_resample<T>(false, false, srcc.data(), s.data(), nscomp, G.nt_data, G.dt_data, nt, par.dt);
    else sinc_resample<T>(false, false, srcc.data(), s.data(), nscomp, G.nt_data, G.dt_data, nt, par.dt, par.sinc_half_length);
    auto inj = make_injector<T>(par.mt ? 1 : 0, G.sxyz, 1, G.g, G.dx, G.dy, G.dz, G.ox, G.oy, G.oz);
    auto ext = receiver_injector<T>(par, G);
    std::vector<T> own;
    T* sp = nullptr;
    if (par.sub > 0) {
        const size_t ns = (size_t)n_snap_slots(nt, par.sub) * nfc * N;
        if (snaps) { sp = snaps; std::fill(sp, sp + ns, (T)0); }
        else { own.assign(ns, (T)0); sp = own.data(); }
    }
    auto P = prop_params(par);
    if (acoustic) propagate_acoustic<T>(false, model.data(), s.data(), r.data(), inj, ext, nullptr, sp, P, G.g, G.dx, G.dy, G.dz);
    else propagate_elastic<T>(par.nmodels == 6, false, model.data(), s.data(), r.data(), inj, ext, nullptr, sp, P, G.g, G.dx, G.dy, G.dz);
    if (par.gl > 0) dipole_to_strain<T>(false, r.data(), G.rdipaz, G.nrcv, nt, par.gl);
    r.resize((size_t)nrcomp * G.nrcv * nt);
    if (par.seismotype == 1) {
        std::vector<T> rd(r.size());
        dt_op<T>(false, r.data(), rd.data(), nrcomp * G.nrcv, nt, par.dt);
        r.swap(rd);
    }
    for (auto& v : r) v *= alpha;
    if (par.resampling == 1) linear_resample<T>(true, add, data, r.data(), nrcomp * G.nrcv, G.nt_data, G.dt_data, nt, par.dt);
    else sinc_resample<T>(true, add, data, r.data(), nrcomp * G.nrcv, G.nt_data, G.dt_data, nt, par.dt, par.sinc_half_length);
}

// apply_jacobianT of one shot (we_op.cpp:948-1078): residual data [nrcomp][nrcv][nt_data]
// -> gradient [min(3,nmodels)][N] w.r.t. the raw parameters (vp, vs, rho / vp, rho),
// overwritten or added, using the snapshots of the matching apply_forward.
template <class T>
void apply_jacobianT(OpParams<T>& par, const ShotGeom<T>& G, const T* raw_model0, const T* dat, bool add, T* grad,
                     const T* snaps) {
    normalize_params(par);
    const size_t N = G.g.n();
    std::vector<T> model(raw_model0, raw_model0 + par.nmodels * N);
    analyze_model(model.data(), G.g, G.dx, G.dy, G.dz, G.nt_data, G.dt_data, par);
    convert_model(model.data(), N, par.nmodels);
    const int nt = par.nt;
    const T alpha = par.dt / G.dt_data;
    const bool acoustic = par.nmodels == 2;
    const int nrcomp = n_comp(par);
    const int nfc = n_field(par);
    const int nm = std::min(3, par.nmodels);
    std::vector<T> temp;
    if (add) temp.assign(grad, grad + nm * N);
    std::fill(grad, grad + nm * N, (T)0);
    std::vector<T> r((size_t)nfc * G.nrcv * nt, 0), dd(dat, dat + (size_t)nrcomp * G.nrcv * G.nt_data);
    if (par.resampling == 1) linear_resample<T>(false, false, dd.data(), r.data(), nrcomp * G.nrcv, G.nt_data, G.dt_data, nt, par.dt);
    else sinc_resample<T>(false, false, dd.data(), r.data(), nrcomp * G.nrcv, G.nt_data, G.dt_data, nt, par.dt, par.sinc_half_length);
    apply_ht<T>(true, r.data(), nfc * G.nrcv, nt, par.dt);
    for (int i = 0; i < nfc * G.nrcv; i++) std::reverse(r.begin() + (size_t)i * nt, r.begin() + (size_t)(i + 1) * nt);
    const T nalpha = -alpha;
    for (auto& v : r) v *= nalpha;
    if (par.seismotype == 1) {
        std::vector<T> rd(r.size());
        dt_op<T>(true, r.data(), rd.data(), nfc * G.nrcv, nt, par.dt);
        r.swap(rd);
    }
    if (par.gl > 0) dipole_to_strain<T>(true, r.data(), G.rdipaz, G.nrcv, nt, par.gl);
    auto srcinj = make_injector<T>(par.mt ? 1 : 0, G.sxyz, 1, G.g, G.dx, G.dy, G.dz, G.ox, G.oy, G.oz);
    auto rinj = receiver_injector<T>(par, G);
    std::vector<T> snapc;
    if (par.sub > 0) snapc.assign(snaps, snaps + (size_t)n_snap_slots(nt, par.sub) * nfc * N);
    auto P = prop_params(par);
    if (acoustic) propagate_acoustic<T>(true, model.data(), r.data(), nullptr, rinj, srcinj, grad, snapc.data(), P, G.g, G.dx, G.dy, G.dz);
    else propagate_elastic<T>(par.nmodels == 6, true, model.data(), r.data(), nullptr, rinj, srcinj, grad, snapc.data(), P, G.g, G.dx, G.dy, G.dz);
    for (int k = 0; k < nm; k++) {
        T* gk = grad + (size_t)k * N;
        apply_h<T>(AX_X, false, false, gk, gk, G.g, G.dx);
        apply_h<T>(AX_Y, false, false, gk, gk, G.g, G.dy);
        apply_h<T>(AX_Z, false, false, gk, gk, G.g, G.dz);
    }
    const T* pm = model.data();
    if (nm >= 3) {
        for (size_t i = 0; i < N; i++) {
            T g0 = grad[i], g1 = grad[N + i], g2 = grad[2 * N + i];
            grad[2 * N + i] = (pm[i] / pm[2 * N + i]) * g0 + (pm[N + i] / pm[2 * N + i]) * g1 + g2;
            grad[N + i] = (T)(2 * std::sqrt((double)(pm[N + i] * pm[2 * N + i])) * (double)(g1 - 2 * g0));
            grad[i] = (T)(2 * std::sqrt((double)(pm[2 * N + i] * (pm[i] + 2 * pm[N + i]))) * (double)g0);
        }
    } else {
        for (size_t i = 0; i < N; i++) {
            T g0 = grad[i], g1 = grad[N + i];
            grad[N + i] = pm[N + i] * pm[i] * g0 - pm[N + i] * pm[N + i] * g1;
            grad[i] = (T)(2 * std::sqrt((double)(pm[i] / pm[N + i])) * (double)g0);
        }
    }
    if (add)
        for (size_t i = 0; i < nm * N; i++) grad[i] += temp[i];
}

// One shot of nlls_fwi::compute_res_and_grad without weights, filters or
// normalisation (optimization.cpp:35-188): forward, r = syn - obs, Ht r, and
// apply_jacobianT(add) into grad. Returns 0.5*dt*(sum r^2 - (r_0^2+r_{nt-1}^2)/2)
// summed over traces in double (optimization.hpp:410-424 before / dnorm).
template <class T>
double fwi_shot(OpParams<T>& par, const ShotGeom<T>& G, const T* model, const T* src, const T* obs, T* resid, T* grad,
                T* snaps) {
    const int ntr = n_comp(par) * G.nrcv;
    const int nt = G.nt_data;
    std::vector<T> syn((size_t)ntr * nt, (T)0);
    apply_forward<T>(par, G, model, src, false, syn.data(), snaps);
    double f = 0;
    for (int ix = 0; ix < ntr; ix++)
        for (int it = 0; it < nt; it++) {
            const size_t i = (size_t)ix * nt + it;
            const T r = syn[i] - obs[i];
            if (resid) resid[i] = r;
            f += (double)r * (double)r * ((it == 0 || it == nt - 1) ? 0.5 : 1.0);
            syn[i] = r;
        }
    apply_ht<T>(false, syn.data(), ntr, nt, G.dt_data);
    apply_jacobianT<T>(par, G, model, syn.data(), true, grad, snaps);
    return 0.5 * (double)G.dt_data * f;
}

}  // namespace orc
