// Reference-side binding of the engine, as a maintainer would add it to fwi3d
// (src/include/we_op.hpp, src/cpp/we_op.cpp). Compiled with g++ -fsyntax-only against the
// reference's own headers and include/fwi3d_hip.h by tests/test_integration_shim.py, so
// the types, member names and argument orders below are the reference's.
//
// The only change to the reference's class is one member on nl_we_op (we_op.hpp:18-24):
//     mutable fwi3d_ctx * _gpu = nullptr;     // one engine context per operator instance
// and its release in the destructor; the method bodies then call the functions below
// with _gpu.
#include <string>
#include <vector>

#include "fwi3d_hip.h"
#include "we_op.hpp"

// ---- 1. propagate (we_op.hpp:93, 127, 163): the body of nl_we_op_{e,vti,a}::propagate ----
static fwi3d_injector gpu_tables(const injector* I, int kind) {
    fwi3d_injector t = {I->_npts, kind, I->_xind.data(), I->_yind.data(), I->_zind.data(),
                        I->_xw.data(), I->_yw.data(), I->_zw.data()};
    return t;
}

// physics: FWI3D_ELASTIC (nl_we_op_e), FWI3D_VTI (nl_we_op_vti), FWI3D_ACOUSTIC (nl_we_op_a).
// The injector kind follows the classes apply_forward / apply_jacobianT build: ddelta_m3
// for a moment-tensor source, dipole_m3 for DAS channels (par.gl > 0), delta_m3 otherwise.
void gpu_propagate(fwi3d_ctx*& ctx, int physics, bool adj, const data_t* model, const data_t* src, data_t* rcv,
                   const injector* inj, const injector* ext, data_t* grad, const param& par, int nx, int ny, int nz,
                   data_t dx, data_t dy, data_t dz) {
    if (ctx == nullptr) successCheck(fwi3d_create(&ctx, 0) == 0, "fwi3d_create failed\n");
    fwi3d_param p = {};
    p.nt = par.nt;
    p.sub = par.sub;
    p.dt = par.dt;
    const int bc[6] = {par.bc_top, par.bc_bottom, par.bc_left, par.bc_right, par.bc_front, par.bc_back};
    const int tp[6] = {par.taper_top, par.taper_bottom, par.taper_left, par.taper_right, par.taper_front,
                       par.taper_back};
    for (int i = 0; i < 6; i++) {
        p.bc[i] = bc[i];
        p.taper[i] = tp[i];
    }
    p.taper_strength = par.taper_strength;
    p.free_surface_stiffness = par.free_surface_stiffness;
    p.snap_precision = FWI3D_SNAP_F32;
    const int src_kind = par.mt ? 1 : 0;
    const int rcv_kind = par.gl > 0 ? 2 : 0;
    fwi3d_injector ti = gpu_tables(inj, adj ? rcv_kind : src_kind);
    fwi3d_injector te = gpu_tables(ext, adj ? src_kind : rcv_kind);
    const int rc = fwi3d_propagate(ctx, physics, adj, model, src, rcv, &ti, &te, grad, &p, nx, ny, nz, dx, dy, dz);
    successCheck(rc == 0, std::string("fwi3d_propagate: ") + fwi3d_last_error(ctx) + "\n");
}

// ---- 2. the shot operator (we_op.hpp:50-51, 98-99): apply_forward / apply_jacobianT ----
fwi3d_op_param gpu_op_param(const param& par) {
    fwi3d_op_param p;
    fwi3d_op_param_defaults(&p);
    p.nmodels = par.nmodels;
    p.dt = par.dt;
    p.courant = par.courant;
    p.sub = par.sub;
    p.resampling = par.resampling == "linear";
    p.sinc_half_length = par.sinc_half_length;
    const int bc[6] = {par.bc_top, par.bc_bottom, par.bc_left, par.bc_right, par.bc_front, par.bc_back};
    const int tp[6] = {par.taper_top, par.taper_bottom, par.taper_left, par.taper_right, par.taper_front,
                       par.taper_back};
    for (int i = 0; i < 6; i++) {
        p.bc[i] = bc[i];
        p.taper[i] = tp[i];
    }
    p.taper_strength = par.taper_strength;
    p.free_surface_stiffness = par.free_surface_stiffness;
    p.mt = par.mt;
    p.seismotype = par.seismotype;
    p.soft_clip = par.soft_clip;
    p.vpmin = par.vpmin; p.vpmax = par.vpmax; p.vsmin = par.vsmin; p.vsmax = par.vsmax;
    p.rhomin = par.rhomin; p.rhomax = par.rhomax; p.deltamin = par.deltamin; p.deltamax = par.deltamax;
    p.epsilonmin = par.epsilonmin; p.epsilonmax = par.epsilonmax; p.gammamin = par.gammamin; p.gammamax = par.gammamax;
    p.gl = par.gl;
    return p;
}

// the fields analyzeModel updates in the operator's _par (we_op.cpp:378-444)
void gpu_param_back(const fwi3d_op_param& p, param& par) {
    par.dt = p.dt;
    par.courant = p.courant;
    par.sub = p.sub;
    par.nt = p.nt;
    par.vmin = p.vmin;
    par.vmax = p.vmax;
    par.tmax = (par.nt - 1) * par.dt;
}

// Geometry of the operator's shot (its _par holds one shot, we_op.hpp:59-79): receiver
// positions flattened into (x, y, z) triplets and the fibers' (dip, azimuth) pairs.
struct GpuShot {
    std::vector<data_t> rxyz, rdipaz;
    fwi3d_shot s;
};
void gpu_shot(const hypercube<data_t>& domain, const vecReg<data_t>& src, const param& par, GpuShot& g) {
    const axis<data_t> X = domain.getAxis(2), Y = domain.getAxis(3), Z = domain.getAxis(1);
    const axis<data_t> T = src.getHyper()->getAxis(1);
    g.rxyz.clear();
    g.rdipaz.clear();
    for (size_t r = 0; r < par.rxyz[0].size(); r++) {
        const std::vector<data_t>& v = par.rxyz[0][r];
        g.rxyz.insert(g.rxyz.end(), v.begin(), v.begin() + 3);
        g.rdipaz.insert(g.rdipaz.end(), v.begin() + 3, v.begin() + 5);
    }
    const fwi3d_shot s = {X.n, Y.n, Z.n, X.d, Y.d, Z.d, X.o, Y.o, Z.o, T.n, T.d, par.sxyz[0].data(),
                          (int)par.rxyz[0].size(), g.rxyz.data(), g.rdipaz.data()};
    g.s = s;
}

// nl_we_op_e::apply_forward(add, pmod, pdat) becomes
//     gpu_apply_forward(_gpu, _domain, *_src, _par, add, pmod, pdat);
void gpu_apply_forward(fwi3d_ctx*& ctx, const hypercube<data_t>& domain, const vecReg<data_t>& src, param& par,
                       bool add, const data_t* pmod, data_t* pdat) {
    if (ctx == nullptr) successCheck(fwi3d_create(&ctx, 0) == 0, "fwi3d_create failed\n");
    GpuShot g;
    gpu_shot(domain, src, par, g);
    fwi3d_op_param p = gpu_op_param(par);
    const int rc = fwi3d_apply_forward(ctx, &p, &g.s, add, pmod, src.getCVals(), pdat);
    successCheck(rc == 0, std::string("fwi3d_apply_forward: ") + fwi3d_last_error(ctx) + "\n");
    gpu_param_back(p, par);
}

// nl_we_op_e::apply_jacobianT(add, pmod, pmod0, pdat) becomes
//     gpu_apply_jacobianT(_gpu, _domain, *_src, _par, add, pmod, pmod0, pdat);
// (the same context as the apply_forward whose snapshots it consumes)
void gpu_apply_jacobianT(fwi3d_ctx*& ctx, const hypercube<data_t>& domain, const vecReg<data_t>& src, param& par,
                         bool add, data_t* pmod, const data_t* pmod0, const data_t* pdat) {
    successCheck(ctx != nullptr, "apply_jacobianT before apply_forward on this operator\n");
    GpuShot g;
    gpu_shot(domain, src, par, g);
    fwi3d_op_param p = gpu_op_param(par);
    const int rc = fwi3d_apply_jacobianT(ctx, &p, &g.s, add, pmod, pmod0, pdat);
    successCheck(rc == 0, std::string("fwi3d_apply_jacobianT: ") + fwi3d_last_error(ctx) + "\n");
    gpu_param_back(p, par);
}
