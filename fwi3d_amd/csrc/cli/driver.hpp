// Shared host code of the drop-in executables (WE_MODELING_3D.x, FWI3D.x): the checks
// and set-up steps of the reference's drivers, restated on this engine's C ABI.
//   analyze_geometry  src/cpp/we_op.cpp:5-136 (model axes, component counts, bounds checks)
//   analyze_wavelet   src/cpp/we_op.cpp:138-230 (one wavelet duplicated per shot/component,
//                     or one per shot and component given)
//   per-shot operator set-up: nl_we_op_e(domain, allsrc, s, par), src/include/we_op.hpp:59-79
// Errors print a message and exit with status 123, as successCheck does
// (src/include/misc.hpp:4-12).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../../include/fwi3d_hip.h"
#include "cube_io.hpp"
#include "params.hpp"

namespace fwi3d_cli {

[[noreturn]] inline void die(const std::string& msg) {
    std::fprintf(stderr, "\n========ERROR========\n%s\n", msg.c_str());
    std::exit(123);
}
inline void check(bool ok, const std::string& msg) {
    if (!ok) die(msg);
}
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}
inline void abi_check(int rc, fwi3d_ctx* c, const char* what) {
    if (rc != 0) die(std::string(what) + " failed (" + std::to_string(rc) + "): " + (c ? fwi3d_last_error(c) : ""));
}

// Model axes: z, x, y, parameter.
struct Grid3 {
    int nz, nx, ny, nm;
    float dz, dx, dy, oz, ox, oy;
    size_t n() const { return (size_t)nz * nx * ny; }
};

inline Grid3 analyze_geometry(const Cube& model, RunParam& par) {
    check(model.axes.size() == 4, "The model must have four axes");
    const Axis &Z = model.axes[0], &X = model.axes[1], &Y = model.axes[2];
    par.nmodels = model.axes[3].n;
    check(par.nmodels == 2 || par.nmodels == 3 || par.nmodels == 6, "The model must contain 2, 3, or 6 parameters");
    if (par.nmodels == 2) {
        par.mt = false;
        par.gl = 0;
        par.free_surface_stiffness = std::max(par.free_surface_stiffness, 1.0f);
        par.soft_clip = false;
        par.nscomp = 1;
        par.nrcomp = 1;
        par.model_parameterization = 1;
    } else {
        par.nscomp = par.mt ? 6 : 3;
        par.nrcomp = par.gl > 0 ? 1 : 3;
    }
    par.bc_top = std::max(0, par.bc_top); par.bc_bottom = std::max(0, par.bc_bottom);
    par.bc_left = std::max(0, par.bc_left); par.bc_right = std::max(0, par.bc_right);
    par.bc_front = std::max(0, par.bc_front); par.bc_back = std::max(0, par.bc_back);
    if (par.verbose > 0) {
        std::fprintf(stderr, "\n==========================\n Subsurface model geometry\n==========================\n");
        std::fprintf(stderr, "xmin=%.5f km, xmax=%.5f km, dx=%0.5f km, nx=%d\n", X.o, (X.n - 1) * X.d + X.o, X.d, X.n);
        std::fprintf(stderr, "ymin=%.5f km, ymax=%.5f km, dy=%0.5f km, ny=%d\n", Y.o, (Y.n - 1) * Y.d + Y.o, Y.d, Y.n);
        std::fprintf(stderr, "zmin=%.5f km, zmax=%.5f km, dz=%0.5f km, nz=%d\n", Z.o, (Z.n - 1) * Z.d + Z.o, Z.d, Z.n);
        std::fprintf(stderr, "Number of sources = %d\n", par.ns);
    }
    const float prec = std::min(X.d, Z.d) * 1e-3f;
    auto inside = [&](float v, const Axis& a) { return v - a.o >= -prec && -prec <= -v + a.o + (a.n - 1) * a.d; };
    par.nr = 0;
    for (int s = 0; s < par.ns; s++) {
        const auto& sx = par.sxyz[s];
        check(inside(sx[0], X) && inside(sx[1], Y) && inside(sx[2], Z), "Shot falls outside subsurface model boundaries");
        const int nr = (int)par.rxyz[s].size();
        par.nr += nr;
        for (int r = 0; r < nr; r++) {
            const auto& rr = par.rxyz[s][r];
            if (par.gl <= 0) {
                check(inside(rr[0], X) && inside(rr[1], Y) && inside(rr[2], Z),
                      "Receiver " + std::to_string(r) + " at x=" + std::to_string(rr[0]) + " km, y=" + std::to_string(rr[1]) +
                          " km, z=" + std::to_string(rr[2]) + " km, falls outside subsurface model boundaries");
            } else {
                const float hx = par.gl / 2 * std::cos(rr[3]) * std::cos(rr[4]);
                const float hy = par.gl / 2 * std::cos(rr[3]) * std::sin(rr[4]);
                const float hz = par.gl / 2 * std::sin(rr[3]);
                auto in_open = [](float v, const Axis& a) { return !(v < a.o || v > a.o + (a.n - 1) * a.d); };
                check(in_open(rr[0] - hx, X) && in_open(rr[0] + hx, X) && in_open(rr[1] - hy, Y) && in_open(rr[1] + hy, Y) &&
                          in_open(rr[2] - hz, Z) && in_open(rr[2] + hz, Z),
                      "DAS channel " + std::to_string(r) + " must be away from the subsurface model boundaries by half of "
                      "the gauge length along the fiber");
            }
        }
    }
    return Grid3{Z.n, X.n, Y.n, par.nmodels, Z.d, X.d, Y.d, Z.o, X.o, Y.o};
}

// Wavelets of every shot: [ns][nscomp][nt].
inline Cube analyze_wavelet(const Cube& src, const RunParam& par) {
    const Axis T = src.axes[0];
    Cube all;
    if (src.axes.size() == 1) {
        all.axes = {T, Axis{par.nscomp, 0.f, 1.f, ""}, Axis{par.ns, 0.f, 1.f, ""}};
        if (par.nscomp == 1) all.axes = {T, Axis{par.ns, 0.f, 1.f, ""}};
        all.alloc();
        const int nt = T.n;
        for (int s = 0; s < par.ns; s++) {
            float* o = all.v.data() + (size_t)s * par.nscomp * nt;
            for (int it = 0; it < nt; it++) {
                const float w = src.v[it];
                if (par.nscomp == 1) {
                    o[it] = w;
                } else if (par.nscomp == 3) {
                    o[it] = w * std::cos(par.fangle) * std::cos(par.fazimuth);
                    o[nt + it] = w * std::cos(par.fangle) * std::sin(par.fazimuth);
                    o[2 * nt + it] = w * std::sin(par.fangle);
                } else {
                    const float m[6] = {par.mxx, par.myy, par.mzz, par.mxy, par.mxz, par.myz};
                    for (int c = 0; c < 6; c++) o[c * nt + it] = w * m[c];
                }
            }
        }
        return all;
    }
    const long ntr = (long)(src.count() / T.n);
    const long want = (long)par.ns * par.nscomp;
    check(ntr == want, std::to_string(want) + " wavelets are expected, " + std::to_string(ntr) + " are provided");
    return src;
}

inline fwi3d_op_param op_param(const RunParam& p) {
    fwi3d_op_param q;
    fwi3d_op_param_defaults(&q);
    q.nmodels = p.nmodels;
    q.dt = p.dt;
    q.courant = p.courant;
    q.sub = p.sub;
    q.resampling = p.resampling == "linear" ? 1 : 0;
    q.sinc_half_length = p.sinc_half_length;
    const int bc[6] = {p.bc_top, p.bc_bottom, p.bc_left, p.bc_right, p.bc_front, p.bc_back};
    const int tp[6] = {p.taper_top, p.taper_bottom, p.taper_left, p.taper_right, p.taper_front, p.taper_back};
    for (int i = 0; i < 6; i++) { q.bc[i] = bc[i]; q.taper[i] = tp[i]; }
    q.taper_strength = p.taper_strength;
    q.free_surface_stiffness = p.free_surface_stiffness;
    q.mt = p.mt ? 1 : 0;
    q.seismotype = p.seismotype;
    q.soft_clip = p.soft_clip ? 1 : 0;
    q.vpmin = p.vpmin; q.vpmax = p.vpmax; q.vsmin = p.vsmin; q.vsmax = p.vsmax; q.rhomin = p.rhomin; q.rhomax = p.rhomax;
    q.deltamin = p.deltamin; q.deltamax = p.deltamax; q.epsilonmin = p.epsilonmin; q.epsilonmax = p.epsilonmax;
    q.gammamin = p.gammamin; q.gammamax = p.gammamax;
    q.gl = p.gl;
    return q;
}

// One shot's geometry (positions kept alive by the caller's vectors).
struct ShotGeo {
    std::vector<float> sxyz, rxyz, rdipaz;
    fwi3d_shot shot{};
};

inline ShotGeo shot_geometry(const RunParam& p, const Grid3& g, int s, int nt_data, float dt_data) {
    ShotGeo o;
    o.sxyz = p.sxyz[s];
    for (const auto& r : p.rxyz[s]) {
        o.rxyz.insert(o.rxyz.end(), r.begin(), r.begin() + 3);
        o.rdipaz.insert(o.rdipaz.end(), r.begin() + 3, r.begin() + 5);
    }
    o.shot.nx = g.nx; o.shot.ny = g.ny; o.shot.nz = g.nz;
    o.shot.dx = g.dx; o.shot.dy = g.dy; o.shot.dz = g.dz;
    o.shot.ox = g.ox; o.shot.oy = g.oy; o.shot.oz = g.oz;
    o.shot.nt_data = nt_data;
    o.shot.dt_data = dt_data;
    o.shot.sxyz = o.sxyz.data();
    o.shot.nrcv = (int)p.rxyz[s].size();
    o.shot.rxyz = o.rxyz.data();
    o.shot.rdipaz = o.rdipaz.data();
    return o;
}

// Device buffer with an upload helper.
struct DevArray {
    void* p = nullptr;
    size_t bytes = 0;
    explicit DevArray(size_t b) : bytes(b) { hip_check(hipMalloc(&p, b ? b : 4), "hipMalloc"); }
    ~DevArray() { if (p) (void)hipFree(p); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    template <class T>
    T* as() const { return (T*)p; }
    void upload(const void* h, size_t b) { hip_check(hipMemcpy(p, h, b, hipMemcpyHostToDevice), "hipMemcpy"); }
    void download(void* h, size_t b) const { hip_check(hipMemcpy(h, p, b, hipMemcpyDeviceToHost), "hipMemcpy"); }
};

inline std::vector<std::string> args_of(int argc, char** argv) { return std::vector<std::string>(argv, argv + argc); }

}  // namespace fwi3d_cli
