// ORACLE — test infrastructure only (see sbp.hpp). C-ABI of the CPU restatement,
// built by oracle/Makefile into oracle/_build/libcpu_ref.so. Every function has a
// float (_f) and a double (_d) variant; the double build is the FP64 oracle.
#include "pipeline.hpp"

extern "C" {

// Injector tables (the arrays of the reference's injector: src/include/injector.hpp:8-16).
typedef struct {
    int npts, kind;  // kind 0 delta_m3, 1 ddelta_m3
    const int *xi, *yi, *zi;
    const void *xw, *yw, *zw;
} orc_inj;

// Propagation parameters (subset of the reference's param struct: src/include/param.hpp:8-71).
typedef struct {
    int nt, sub;
    double dt;
    int bc[6], taper[6];
    double taper_strength, free_surface_stiffness;
    int lag;  // 1: lagged adjoint derivatives in the elastic/VTI imaging condition (the reference); 0: lag-free
} orc_par;

// Shot-operator parameters (src/include/param.hpp:8-71): in/out like the reference's _par.
typedef struct {
    int nmodels;
    double dt, courant;
    int sub, nt;
    double vmin, vmax;
    int resampling, sinc_half_length;
    int bc[6], taper[6];
    double taper_strength, free_surface_stiffness;
    int mt, seismotype, soft_clip;
    double vpmin, vpmax, vsmin, vsmax, rhomin, rhomax, deltamin, deltamax, epsilonmin, epsilonmax, gammamin, gammamax;
    int lag;  // as in orc_par
    double gl;  // DAS gauge length (> 0: dipole receivers, strain data)
} orc_op_par;

// Geometry of one shot; positions are arrays of the variant's type.
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, ox, oy, oz;
    int nt_data;
    double dt_data;
    const void* sxyz;
    int nrcv;
    const void* rxyz;
    const void* rdipaz;  // [nrcv][2] dip, azimuth of DAS channels (or NULL)
} orc_shot;
}

namespace {

template <class T>
orc::Injector<T> to_inj(const orc_inj* s) {
    orc::Injector<T> I;
    I.npts = s->npts;
    I.kind = s->kind;
    const int nw = I.nw();
    const int npi = I.npi();
    I.xi.assign(s->xi, s->xi + npi * s->npts);
    I.yi.assign(s->yi, s->yi + npi * s->npts);
    I.zi.assign(s->zi, s->zi + npi * s->npts);
    const T* xw = (const T*)s->xw;
    const T* yw = (const T*)s->yw;
    const T* zw = (const T*)s->zw;
    I.xw.assign(xw, xw + 2 * nw * s->npts);
    I.yw.assign(yw, yw + 2 * nw * s->npts);
    I.zw.assign(zw, zw + 2 * nw * s->npts);
    return I;
}

template <class T>
orc::PropParams<T> to_par(const orc_par* p) {
    orc::PropParams<T> q;
    q.nt = p->nt;
    q.sub = p->sub;
    q.dt = (T)p->dt;
    for (int i = 0; i < 6; i++) { q.bc[i] = p->bc[i]; q.taper[i] = p->taper[i]; }
    q.taper_strength = (T)p->taper_strength;
    q.free_surface_stiffness = (T)p->free_surface_stiffness;
    q.lag = p->lag;
    return q;
}

template <class T>
int propagate_t(int physics, int adj, const T* model, const T* src, T* rcv, const orc_inj* inj, const orc_inj* ext,
                T* grad, T* snaps, const orc_par* par, int nx, int ny, int nz, T dx, T dy, T dz) {
    try {
        orc::Grid g{nx, ny, nz};
        auto I = to_inj<T>(inj);
        auto E = to_inj<T>(ext);
        auto P = to_par<T>(par);
        if (physics == orc::PHYS_ACOUSTIC)
            orc::propagate_acoustic<T>(adj, model, src, rcv, I, E, grad, snaps, P, g, dx, dy, dz);
        else
            orc::propagate_elastic<T>(physics == orc::PHYS_VTI, adj, model, src, rcv, I, E, grad, snaps, P, g, dx, dy, dz);
        return 0;
    } catch (...) {
        return -1;
    }
}

template <class T>
orc::OpParams<T> to_op(const orc_op_par* p) {
    orc::OpParams<T> q;
    q.nmodels = p->nmodels; q.dt = (T)p->dt; q.courant = (T)p->courant; q.sub = p->sub; q.nt = p->nt;
    q.vmin = (T)p->vmin; q.vmax = (T)p->vmax; q.resampling = p->resampling; q.sinc_half_length = p->sinc_half_length;
    for (int i = 0; i < 6; i++) { q.bc[i] = p->bc[i]; q.taper[i] = p->taper[i]; }
    q.taper_strength = (T)p->taper_strength; q.free_surface_stiffness = (T)p->free_surface_stiffness;
    q.mt = p->mt; q.seismotype = p->seismotype; q.soft_clip = p->soft_clip;
    q.vpmin = (T)p->vpmin; q.vpmax = (T)p->vpmax; q.vsmin = (T)p->vsmin; q.vsmax = (T)p->vsmax;
    q.rhomin = (T)p->rhomin; q.rhomax = (T)p->rhomax; q.deltamin = (T)p->deltamin; q.deltamax = (T)p->deltamax;
    q.epsilonmin = (T)p->epsilonmin; q.epsilonmax = (T)p->epsilonmax; q.gammamin = (T)p->gammamin; q.gammamax = (T)p->gammamax;
    q.lag = p->lag;
    q.gl = (T)p->gl;
    return q;
}

template <class T>
void from_op(const orc::OpParams<T>& q, orc_op_par* p) {
    p->nmodels = q.nmodels; p->dt = q.dt; p->courant = q.courant; p->sub = q.sub; p->nt = q.nt;
    p->vmin = q.vmin; p->vmax = q.vmax; p->mt = q.mt; p->soft_clip = q.soft_clip;
    p->free_surface_stiffness = q.free_surface_stiffness;
    p->gl = q.gl;
    for (int i = 0; i < 6; i++) p->bc[i] = q.bc[i];
}

template <class T>
orc::ShotGeom<T> to_shot(const orc_shot* s) {
    orc::ShotGeom<T> G;
    G.g = orc::Grid{s->nx, s->ny, s->nz};
    G.dx = (T)s->dx; G.dy = (T)s->dy; G.dz = (T)s->dz; G.ox = (T)s->ox; G.oy = (T)s->oy; G.oz = (T)s->oz;
    G.nt_data = s->nt_data; G.dt_data = (T)s->dt_data;
    G.sxyz = (const T*)s->sxyz; G.nrcv = s->nrcv; G.rxyz = (const T*)s->rxyz; G.rdipaz = (const T*)s->rdipaz;
    return G;
}

}  // namespace

#define ORC_OP_EXPORTS(T, S)                                                                                           \
    int orc_analyze_##S(orc_op_par* par, T* model, const orc_shot* shot) {                                             \
        auto P = to_op<T>(par);                                                                                        \
        auto G = to_shot<T>(shot);                                                                                     \
        orc::normalize_params(P);                                                                                      \
        orc::analyze_model<T>(model, G.g, G.dx, G.dy, G.dz, G.nt_data, G.dt_data, P);                                  \
        from_op<T>(P, par);                                                                                            \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_convert_##S(T* model, long long n, int nmodels) {                                                          \
        orc::convert_model<T>(model, (size_t)n, nmodels);                                                              \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_resample_##S(int method, int adj, int add, T* mod, T* dat, int ntr, int nm, T dm, int nd, T dd, int hl) {  \
        if (method == 1) orc::linear_resample<T>(adj, add, mod, dat, ntr, nm, dm, nd, dd);                            \
        else orc::sinc_resample<T>(adj, add, mod, dat, ntr, nm, dm, nd, dd, hl);                                       \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_ht_##S(int inv, T* x, int ntr, int nt, T dt) {                                                             \
        orc::apply_ht<T>(inv, x, ntr, nt, dt);                                                                         \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_dt_##S(int adj, const T* in, T* out, int ntr, int nt, T dt) {                                              \
        orc::dt_op<T>(adj, in, out, ntr, nt, dt);                                                                      \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_apply_forward_##S(orc_op_par* par, const orc_shot* shot, int add, const T* model, const T* src, T* data,   \
                              T* snaps) {                                                                              \
        try {                                                                                                          \
            auto P = to_op<T>(par);                                                                                    \
            orc::apply_forward<T>(P, to_shot<T>(shot), model, src, add, data, snaps);                                  \
            from_op<T>(P, par);                                                                                        \
            return 0;                                                                                                  \
        } catch (...) {                                                                                                \
            return -1;                                                                                                 \
        }                                                                                                              \
    }                                                                                                                  \
    int orc_apply_jacobianT_##S(orc_op_par* par, const orc_shot* shot, int add, T* grad, const T* model0,             \
                                const T* data, const T* snaps) {                                                       \
        try {                                                                                                          \
            auto P = to_op<T>(par);                                                                                    \
            orc::apply_jacobianT<T>(P, to_shot<T>(shot), model0, data, add, grad, snaps);                              \
            from_op<T>(P, par);                                                                                        \
            return 0;                                                                                                  \
        } catch (...) {                                                                                                \
            return -1;                                                                                                 \
        }                                                                                                              \
    }                                                                                                                  \
    int orc_fwi_shot_##S(orc_op_par* par, const orc_shot* shot, const T* model, const T* src, const T* obs, T* resid,  \
                         T* grad, T* snaps, double* misfit) {                                                          \
        try {                                                                                                          \
            auto P = to_op<T>(par);                                                                                    \
            *misfit = orc::fwi_shot<T>(P, to_shot<T>(shot), model, src, obs, resid, grad, snaps);                      \
            from_op<T>(P, par);                                                                                        \
            return 0;                                                                                                  \
        } catch (...) {                                                                                                \
            return -1;                                                                                                 \
        }                                                                                                              \
    }

#define ORC_EXPORTS(T, S)                                                                                              \
    int orc_d1_##S(int axis, int add, const T* in, T* out, int nx, int ny, int nz, T h, const T* par) {               \
        orc::d1<T>(axis, add, in, out, orc::Grid{nx, ny, nz}, h, par);                                                 \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_d2_##S(int axis, int add, const T* in, T* out, const T* b, int nx, int ny, int nz, T h, T a) {            \
        try { orc::d2<T>(axis, add, in, out, b, orc::Grid{nx, ny, nz}, h, a); } catch (...) { return -1; }            \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_apply_h_##S(int axis, int inv, int add, const T* in, T* out, int nx, int ny, int nz, T h) {               \
        orc::apply_h<T>(axis, inv, add, in, out, orc::Grid{nx, ny, nz}, h);                                            \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_taper_##S(T* u, int ncomp, int nx, int ny, int nz, const int* tp, T a) {                                  \
        orc::taper<T>(u, ncomp, orc::Grid{nx, ny, nz}, tp, a);                                                         \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_asat_##S(int face, int bc, const T* curr, const T* prev, T* next, const T* model, int nx, int ny, int nz,  \
                     T dx, T dy, T dz, T dt, T alpha) {                                                                \
        orc::Grid g{nx, ny, nz};                                                                                       \
        orc::asat_face<T>(face, bc, curr, prev, next, model, model + g.n(), g, dx, dy, dz, dt, alpha);                \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_ascale_##S(T* next, const T* model, int nx, int ny, int nz, T dx, T dy, T dz, T dt, const int* bc) {      \
        orc::Grid g{nx, ny, nz};                                                                                       \
        orc::ascale<T>(next, model, model + g.n(), g, dx, dy, dz, dt, bc);                                            \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_esat_##S(int face, int bc, int vti, const T* curr, const T* prev, T* next, const T* model, int nx,         \
                     int ny, int nz, T dx, T dy, T dz, T dt) {                                                         \
        orc::Grid g{nx, ny, nz};                                                                                       \
        const size_t n = g.n();                                                                                        \
        orc::EMod<T> M;                                                                                                \
        M.vti = vti;                                                                                                   \
        for (int k = 0; k < 6; k++) M.m[k] = (vti || k < 3) ? model + k * n : nullptr;                                \
        const T* c[3] = {curr, curr + n, curr + 2 * n};                                                                \
        const T* p[3] = {prev, prev + n, prev + 2 * n};                                                                \
        T* x[3] = {next, next + n, next + 2 * n};                                                                      \
        if (bc == 1 || bc == 2) orc::esat_face<T>(face, bc, c, p, x, M, g, dx, dy, dz, dt);                          \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_escale_##S(int vti, T* next, const T* model, int nx, int ny, int nz, T dx, T dy, T dz, T dt,              \
                       const int* bc) {                                                                                \
        orc::Grid g{nx, ny, nz};                                                                                       \
        const size_t n = g.n();                                                                                        \
        orc::EMod<T> M;                                                                                                \
        M.vti = vti;                                                                                                   \
        for (int k = 0; k < 6; k++) M.m[k] = (vti || k < 3) ? model + k * n : nullptr;                                \
        T* x[3] = {next, next + n, next + 2 * n};                                                                      \
        orc::escale<T>(x, M, g, dx, dy, dz, dt, bc);                                                                   \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_injector_##S(int kind, const T* loc, int npts, int nx, int ny, int nz, T dx, T dy, T dz, T ox, T oy,      \
                         T oz, int* xi, int* yi, int* zi, T* xw, T* yw, T* zw) {                                       \
        auto I = orc::make_injector<T>(kind, loc, npts, orc::Grid{nx, ny, nz}, dx, dy, dz, ox, oy, oz);                \
        std::copy(I.xi.begin(), I.xi.end(), xi);                                                                       \
        std::copy(I.yi.begin(), I.yi.end(), yi);                                                                       \
        std::copy(I.zi.begin(), I.zi.end(), zi);                                                                       \
        std::copy(I.xw.begin(), I.xw.end(), xw);                                                                       \
        std::copy(I.yw.begin(), I.yw.end(), yw);                                                                       \
        std::copy(I.zw.begin(), I.zw.end(), zw);                                                                       \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_dipole_injector_##S(const T* loc5, int npts, T gl, int nx, int ny, int nz, T dx, T dy, T dz, T ox, T oy,  \
                                T oz, int* xi, int* yi, int* zi, T* xw, T* yw, T* zw) {                                \
        auto I = orc::make_dipole<T>(loc5, npts, gl, orc::Grid{nx, ny, nz}, dx, dy, dz, ox, oy, oz);                  \
        std::copy(I.xi.begin(), I.xi.end(), xi);                                                                       \
        std::copy(I.yi.begin(), I.yi.end(), yi);                                                                       \
        std::copy(I.zi.begin(), I.zi.end(), zi);                                                                       \
        std::copy(I.xw.begin(), I.xw.end(), xw);                                                                       \
        std::copy(I.yw.begin(), I.yw.end(), yw);                                                                       \
        std::copy(I.zw.begin(), I.zw.end(), zw);                                                                       \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_dipole_to_strain_##S(int adj, T* p, const T* dipaz, int nrcv, int nt, T gl) {                              \
        orc::dipole_to_strain<T>(adj != 0, p, dipaz, nrcv, nt, gl);                                                    \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_inject_##S(const orc_inj* inj, const T* src, int nt, int comp, int it, T* out, int nx, int ny, int nz) {  \
        orc::inject<T>(to_inj<T>(inj), src, nt, comp, it, out, orc::Grid{nx, ny, nz});                                 \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_extract_##S(const orc_inj* inj, const T* u, int nt, int comp, int it, T* rcv, int nx, int ny, int nz) {   \
        orc::extract<T>(to_inj<T>(inj), u, nt, comp, it, rcv, orc::Grid{nx, ny, nz});                                  \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_spatial_##S(int physics, const T* curr, const T* prev, T* next, const T* model, const int* bc, int nx,    \
                        int ny, int nz, T dx, T dy, T dz, T dt, T free_surface_stiffness) {                            \
        orc::Grid g{nx, ny, nz};                                                                                       \
        const size_t n = g.n();                                                                                        \
        if (physics == orc::PHYS_ACOUSTIC) {                                                                           \
            const T alpha = (T)0.2508560249 / free_surface_stiffness;                                                  \
            orc::acoustic_volume<T>(curr, next, model + n, g, dx, dy, dz);                                             \
            orc::acoustic_sats<T>(bc, curr, prev, next, model, model + n, g, dx, dy, dz, dt, alpha);                  \
            return 0;                                                                                                  \
        }                                                                                                              \
        const bool vti = physics == orc::PHYS_VTI;                                                                     \
        const orc::EMod<T> M = orc::emod<T>(vti, model, n);                                                            \
        std::vector<T> c11b(vti ? n : 0), du(9 * n), tmp(n);                                                           \
        for (size_t i = 0; i < c11b.size(); i++) c11b[i] = M.c11b(i);                                                  \
        orc::elastic_volume<T>(vti, M, c11b.data(), curr, next, du.data(), tmp.data(), g, dx, dy, dz);                 \
        orc::elastic_sats<T>(M, bc, curr, prev, next, g, dx, dy, dz, dt);                                              \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_grad_event_##S(int physics, const T* model, const T* snaps, int nslots, int itl, int nt, int sub, T dt,   \
                           const T* adjf, const T* derivf, T* grad, int nx, int ny, int nz, T dx, T dy, T dz) {        \
        orc::Grid g{nx, ny, nz};                                                                                       \
        const size_t n = g.n();                                                                                        \
        if (physics == orc::PHYS_ACOUSTIC) {                                                                           \
            std::vector<T> tmp(6 * n);                                                                                 \
            orc::acoustic_imaging<T>(model, snaps, nslots, itl, nt, sub, dt, adjf, tmp.data(), grad, g, dx, dy, dz);  \
            return 0;                                                                                                  \
        }                                                                                                              \
        const bool vti = physics == orc::PHYS_VTI;                                                                     \
        std::vector<T> du(9 * n), tmp(9 * n);                                                                          \
        orc::all_derivatives<T>(derivf, du.data(), g, dx, dy, dz);                                                     \
        orc::elastic_imaging<T>(vti, orc::emod<T>(vti, model, n), snaps, nslots, itl, nt, sub, dt, adjf, du.data(),   \
                                tmp.data(), grad, g, dx, dy, dz);                                                      \
        return 0;                                                                                                      \
    }                                                                                                                  \
    int orc_propagate_##S(int physics, int adj, const T* model, const T* src, T* rcv, const orc_inj* inj,             \
                          const orc_inj* ext, T* grad, T* snaps, const orc_par* par, int nx, int ny, int nz, T dx,     \
                          T dy, T dz) {                                                                                \
        return propagate_t<T>(physics, adj, model, src, rcv, inj, ext, grad, snaps, par, nx, ny, nz, dx, dy, dz);      \
    }

extern "C" {
ORC_EXPORTS(float, f)
ORC_EXPORTS(double, d)
ORC_OP_EXPORTS(float, f)
ORC_OP_EXPORTS(double, d)
}
