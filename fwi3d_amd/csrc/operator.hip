// Shot operator around propagate(): model analysis and conversion, time resampling
// of sources and traces, the adjoint-source chain, the gradient quadrature and chain
// rule, and the per-shot FWI residual and misfit. Device-resident; the host only
// derives the time axis from two reduced scalars.
//
//   analyzeGeometry / analyzeModel ... src/cpp/we_op.cpp:5-25, 234-446
//   convert_model ..................... src/cpp/we_op.cpp:448-466, 1080-1106, 1535-1552
//   apply_forward ..................... src/cpp/we_op.cpp:872-946
//   apply_jacobianT ................... src/cpp/we_op.cpp:948-1078
//   sinc / linear resampler ........... src/cpp/operator.cpp:419-520
//   applyHt, Dt ....................... src/cpp/misc.cpp:44-94
//   compute_res_and_grad, getFunc ..... src/cpp/optimization.cpp:5-197, src/include/optimization.hpp:410-424
//
// Rounding follows the reference's single-precision build where it is cheap to do
// so (float products, double window/sinc factors accumulated through a double add),
// so that traces agree with it to a few ulps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/fwi3d_hip.h"
#include "context.hpp"
#include "sbp_tables.hpp"

using namespace fwi3d;

// Products and sums in this file round as written (no fused multiply-add), as the
// reference's host build does.
#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------------
// Model preparation: clip, admissibility fixes, conversion to propagation
// parameters and the velocity extremes, in one pass.

struct ClipBounds {
    float vpmin, vpmax, vsmin, vsmax, rhomin, rhomax, delmin, delmax, epsmin, epsmax, gammin, gammax;
    int hard;  // 0 when soft_clip (no clipping and no vs fix)
};

// Order-preserving float -> uint key (for atomicMin/atomicMax on floats of any sign).
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float unkey(unsigned k) {
    const unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fmaxf(lo, fminf(hi, v)); }

__global__ void key_init(unsigned* red) {
    if (threadIdx.x == 0) {
        red[0] = 0u;           // max key
        red[1] = 0xffffffffu;  // min key
    }
}

// raw [NM][N] (vp, vs, rho[, delta, eps, gamma]) or (vp, rho) -> conv [NM][N]
// (lambda, mu, rho[, c13, eps, c66]) or (K, 1/rho); red[0] = key(max velocity),
// red[1] = key(min velocity) as analyzeModel defines them (we_op.cpp:378-390, 416-417).
template <int NM>
__global__ __launch_bounds__(kThreads) void prep_model(const float* __restrict__ raw, float* __restrict__ conv,
                                                       long long N, long long nxz, ClipBounds B,
                                                       unsigned* __restrict__ red) {
    unsigned kmax = 0u, kmin = 0xffffffffu;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        if (NM == 2) {
            const float vp = B.hard ? clampf(raw[i], B.vpmin, B.vpmax) : raw[i];
            const float rho = B.hard ? clampf(raw[N + i], B.rhomin, B.rhomax) : raw[N + i];
            kmax = max(kmax, fkey(vp));
            kmin = min(kmin, fkey(vp));
            conv[i] = rho * vp * vp;
            conv[N + i] = (float)(1.0 / (double)rho);
        } else {
            float vp = raw[i], vs = raw[N + i], rho = raw[2 * N + i];
            float del = 0.f, eps = 0.f, gam = 0.f;
            if (NM == 6) {
                del = raw[3 * N + i];
                eps = raw[4 * N + i];
                gam = raw[5 * N + i];
                if (i < nxz) {  // the admissibility scan covers the first nx*nz points only (we_op.cpp:273)
                    const float v1 = (float)(-0.5 * (double)(1.f - vs * vs / (vp * vp)));
                    const float v2 = (float)(2.0 / (double)(vp * vp / (vs * vs) - 1.f));
                    if (del < v1) del = (float)(0.9999 * (double)v1);
                    else if (del > v2) del = (float)(0.9999 * (double)v2);
                }
            }
            if (B.hard) {
                vp = clampf(vp, B.vpmin, B.vpmax);
                vs = clampf(vs, B.vsmin, B.vsmax);
                rho = clampf(rho, B.rhomin, B.rhomax);
                if (NM == 6) {
                    del = clampf(del, B.delmin, B.delmax);
                    eps = clampf(eps, B.epsmin, B.epsmax);
                    gam = clampf(gam, B.gammin, B.gammax);
                }
                if ((double)vp < sqrt(2.0) * (double)vs) vs = (float)((double)vp / sqrt(2.0001));
            }
            kmax = max(kmax, fkey(vp));
            if (NM == 6) kmax = max(kmax, fkey((float)((double)vp * sqrt((double)(1.f + 2.f * eps)))));
            kmin = min(kmin, fkey(vs));
            const float lam = rho * (vp * vp - 2.f * vs * vs);
            const float mu = rho * vs * vs;
            conv[i] = lam;
            conv[N + i] = mu;
            conv[2 * N + i] = rho;
            if (NM == 6) {
                const float a = lam + mu;
                conv[3 * N + i] = (float)(sqrt((double)(2.f * (lam + 2.f * mu) * a * del + a * a)) - (double)mu);
                conv[4 * N + i] = eps;
                conv[5 * N + i] = (1.f + 2.f * gam) * mu;
            }
        }
    }
    // wave64 reduction, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
        kmax = max(kmax, (unsigned)__shfl_xor((int)kmax, o));
        kmin = min(kmin, (unsigned)__shfl_xor((int)kmin, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&red[0], kmax);
        atomicMin(&red[1], kmin);
    }
}

// ---------------------------------------------------------------------------------
// Time-axis resampling of ntr traces between a "model" axis (nm, dm) and a "data"
// axis (nd, dd) (operator.cpp:419-520). fwd: model -> data; adj: data -> model. Each
// output sample gathers its taps in the order in which the reference scatters them,
// so float accumulation matches.

struct Axis {
    int n;
    float d;
};

__device__ __forceinline__ float sinc_tap(float t, int itm, float dm) {
    const float x = (float)(kPi * (double)(t - (float)itm * dm) / (double)dm);
    return x == 0.f ? 1.f : (float)(sin((double)x) / (double)x);
}
__device__ __forceinline__ double sinc_window(float t, int itm, float dm, int hl, float a) {
    return (double)a + (double)(1.f - a) * cos(kPi * (double)((float)itm - t / dm) / (double)hl);
}

// out[tr][itd] (= add*out) + sum_itm (scale*in[tr][itm]) * w(itd, itm)
__global__ __launch_bounds__(kThreads) void resample_fwd(const float* __restrict__ in, Axis M, float* __restrict__ out,
                                                         Axis D, int ntr, int method, int hl, float scale, int add) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)ntr * D.n) return;
    const int tr = (int)(gid / D.n), itd = (int)(gid % D.n);
    const float* m = in + (size_t)tr * M.n;
    float* o = out + (size_t)tr * D.n + itd;
    if (method == 1) {
        const float q = (float)itd * D.d / M.d;
        const int i = min(M.n - 1, (int)floorf(q));
        const int j = min(M.n - 1, (int)ceilf(q));
        const float wj = ((float)itd * D.d - (float)i * M.d) / M.d;
        const float wi = 1.f - wj;
        const float base = add ? *o : 0.f;
        *o = base + wi * (scale * m[i]) + wj * (scale * m[j]);
        return;
    }
    const float t = (float)itd * D.d;
    const int itm0 = (int)floorf(t / M.d);
    float acc = add ? *o : 0.f;
    const int lo = max(0, itm0 - hl + 1), hi = min(M.n - 1, itm0 + hl);
    for (int itm = lo; itm <= hi; itm++) {
        const float v = scale * m[itm];
        acc = (float)((double)acc + (double)(v * sinc_tap(t, itm, M.d)) * sinc_window(t, itm, M.d, hl, 0.5f));
    }
    *o = acc;
}

// mod[tr][itm] (= add*mod) + sum_itd (scale*dat[tr][itd]) * w(itd, itm)
__global__ __launch_bounds__(kThreads) void resample_adj(float* __restrict__ mod, Axis M, const float* __restrict__ dat,
                                                         Axis D, int ntr, int method, int hl, float scale, int add) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)ntr * M.n) return;
    const int tr = (int)(gid / M.n), itm = (int)(gid % M.n);
    const float* d = dat + (size_t)tr * D.n;
    float* o = mod + (size_t)tr * M.n + itm;
    float acc = add ? *o : 0.f;
    const double r = (double)M.d / (double)D.d;
    if (method == 1) {
        // taps: floor(itd*dd/dm) or ceil(itd*dd/dm) equal to itm (clamped to nm-1)
        const int lo = max(0, (int)floor((itm - 1) * r) - 2);
        const int hi = itm == M.n - 1 ? D.n - 1 : min(D.n - 1, (int)ceil((itm + 1) * r) + 2);
        for (int itd = lo; itd <= hi; itd++) {
            const float q = (float)itd * D.d / M.d;
            const int i = min(M.n - 1, (int)floorf(q));
            const int j = min(M.n - 1, (int)ceilf(q));
            if (i != itm && j != itm) continue;
            const float wj = ((float)itd * D.d - (float)i * M.d) / M.d;
            const float wi = 1.f - wj;
            const float v = scale * d[itd];
            if (i == itm) acc += wi * v;
            if (j == itm) acc += wj * v;
        }
        *o = acc;
        return;
    }
    // taps: itm in [itm0 - hl + 1, itm0 + hl] with itm0 = floor(itd*dd/dm)
    const int lo = max(0, (int)floor((itm - hl) * r) - 2);
    const int hi = min(D.n - 1, (int)ceil((itm + hl) * r) + 2);
    for (int itd = lo; itd <= hi; itd++) {
        const float t = (float)itd * D.d;
        const int itm0 = (int)floorf(t / M.d);
        if (itm < itm0 - hl + 1 || itm > itm0 + hl) continue;
        const float v = scale * d[itd];
        acc = (float)((double)acc + (double)(v * sinc_tap(t, itm, M.d)) * sinc_window(t, itm, M.d, hl, 0.5f));
    }
    *o = acc;
}

// Time derivative (misc.cpp:70-94) of ntr traces, or its adjoint; out = in' (no add).
__global__ __launch_bounds__(kThreads) void time_derivative(const float* __restrict__ in, float* __restrict__ out,
                                                            int ntr, int nt, float dt, int adj) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)ntr * nt) return;
    const int it = (int)(gid % nt);
    const float* a = in + (gid - it);
    float v;
    if (it == 0) v = adj ? (a[1] + a[0]) / dt : (a[1] - a[0]) / dt;
    else if (it == nt - 1) v = adj ? (-a[nt - 1] - a[nt - 2]) / dt : (a[nt - 1] - a[nt - 2]) / dt;
    else v = (a[it + 1] - a[it - 1]) / (2.f * dt);
    out[gid] = v;
}

// Adjoint source from the residual at the propagation rate: inverse time
// quadrature, time reversal and the factor -alpha (we_op.cpp:1001-1003).
__global__ __launch_bounds__(kThreads) void adjoint_source(const float* __restrict__ r, float* __restrict__ out,
                                                           int ntr, int nt, float dt, float nalpha) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)ntr * nt) return;
    const int it = (int)(gid % nt);
    const int src = nt - 1 - it;
    const float v = r[gid - it + src];
    const float w = (src == 0 || src == nt - 1) ? (float)((double)v / (0.5 * (double)dt)) : v / dt;
    out[gid] = w * nalpha;
}

// Per-shot residual r = syn - obs (written to `resid`), the FP64 trapezoid misfit
// sum_tr [ sum_t r^2 - (r_0^2 + r_{nt-1}^2)/2 ] and the time quadrature Ht r in
// place of syn (optimization.cpp:111-116, optimization.hpp:410-424, misc.cpp:44-55).
__global__ __launch_bounds__(kThreads) void residual_misfit(float* __restrict__ syn, const float* __restrict__ obs,
                                                            float* __restrict__ resid, int ntr, int nt, float dt,
                                                            double* __restrict__ misfit) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (gid < (long long)ntr * nt) {
        const int it = (int)(gid % nt);
        const float r = syn[gid] - obs[gid];
        if (resid) resid[gid] = r;
        const bool end = (it == 0 || it == nt - 1);
        e = (double)r * (double)r * (end ? 0.5 : 1.0);
        syn[gid] = end ? (float)((double)r * 0.5 * (double)dt) : r * dt;
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
    __shared__ double part[kThreads / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; w++) s += part[w];
        atomicAdd(misfit, s);
    }
}

// Gradient post-processing (we_op.cpp:1033-1075): H quadrature on all three axes,
// then the chain rule from (lambda, mu, rho) to (vp, vs, rho) or from (K, 1/rho) to
// (vp, rho), then out (+)= result. g64: raw FP64 gradient [nm][N]; pm: converted model.
template <class OutT>
__global__ __launch_bounds__(kThreads) void gradient_finish(const double* __restrict__ g64,
                                                            const float* __restrict__ pm, OutT* __restrict__ out,
                                                            int nx, int ny, int nz, float dx, float dy, float dz, int nm,
                                                            int add, double scale) {
    const long long N = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int iz = (int)(i % nz);
    const int ix = (int)((i / nz) % nx);
    const int iy = (int)(i / ((long long)nz * nx));
    auto hw = [](int k, int n) -> double {
        const double h[4] = {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0};
        if (k < 4) return h[k];
        if (k >= n - 4) return h[n - 1 - k];
        return 1.0;
    };
    const double w = hw(ix, nx) * (double)dx * hw(iy, ny) * (double)dy * hw(iz, nz) * (double)dz;
    if (nm >= 3) {
        const double g0 = g64[i] * w, g1 = g64[N + i] * w, g2 = g64[2 * N + i] * w;
        const double lam = pm[i], mu = pm[N + i], rho = pm[2 * N + i];
        const double grho = (lam / rho) * g0 + (mu / rho) * g1 + g2;
        const double gvs = 2.0 * sqrt(mu * rho) * (g1 - 2.0 * g0);
        const double gvp = 2.0 * sqrt(rho * (lam + 2.0 * mu)) * g0;
        out[i] = (OutT)((add ? (double)out[i] : 0.0) + scale * gvp);
        out[N + i] = (OutT)((add ? (double)out[N + i] : 0.0) + scale * gvs);
        out[2 * N + i] = (OutT)((add ? (double)out[2 * N + i] : 0.0) + scale * grho);
    } else {
        const double g0 = g64[i] * w, g1 = g64[N + i] * w;
        const double K = pm[i], b = pm[N + i];
        const double grho = b * K * g0 - b * b * g1;
        const double gvp = 2.0 * sqrt(K / b) * g0;
        out[i] = (OutT)((add ? (double)out[i] : 0.0) + scale * gvp);
        out[N + i] = (OutT)((add ? (double)out[N + i] : 0.0) + scale * grho);
    }
}

// dipole_to_strain (src/cpp/misc.cpp:11-42) on traces p [3][nrcv][nt]; trig [nrcv] =
// (cos dip, sin dip, cos az, sin az). Forward: component 0 <- the fiber projection / gl,
// components 1 and 2 <- 0; adjoint: the three components from component 0.
__global__ void dipole_strain(float* __restrict__ p, const float4* __restrict__ trig, int nrcv, int nt, float gl,
                              int adj) {
    const long long n = (long long)nrcv * nt;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / nt);
        const float4 t = trig[r];  // cd, sd, ca, sa
        float* p0 = p + i;
        float* p1 = p + n + i;
        float* p2 = p + 2 * n + i;
        if (!adj) {
            *p0 = (t.x * t.z * *p0 + t.x * t.w * *p1 + t.y * *p2) / gl;
            *p1 = 0.f;
            *p2 = 0.f;
        } else {
            const float v = *p0;
            *p2 = t.y * v / gl;
            *p1 = t.x * t.w * v / gl;
            *p0 = v * (t.x * t.z / gl);
        }
    }
}

__global__ void zero_f64(double* p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

inline unsigned nblocks(long long n) { return (unsigned)std::max(1LL, (n + kThreads - 1) / kThreads); }

// ---------------------------------------------------------------------------------
// Host side

struct ShotDims {
    int nx, ny, nz;
    long long N;
    int nm;      // nmodels
    int ncomp;   // propagated components (1 acoustic, 3 elastic)
    int nscomp;  // source components
    int nrcomp;  // receiver components
    int ng;      // gradient parameters (min(3, nmodels))
    int physics;
};

ShotDims shot_dims(const fwi3d_op_param* par, const fwi3d_shot* s) {
    ShotDims d;
    d.nx = s->nx; d.ny = s->ny; d.nz = s->nz;
    d.N = (long long)s->nx * s->ny * s->nz;
    d.nm = par->nmodels;
    const bool ac = par->nmodels == 2;
    d.ncomp = ac ? 1 : 3;
    d.nscomp = ac ? 1 : (par->mt ? 6 : 3);
    d.nrcomp = (ac || par->gl > 0) ? 1 : 3;  // DAS channels record one strain component
    d.ng = std::min(3, par->nmodels);
    d.physics = ac ? FWI3D_ACOUSTIC : (par->nmodels == 3 ? FWI3D_ELASTIC : FWI3D_VTI);
    return d;
}

void check_op(const fwi3d_op_param* par, const fwi3d_shot* s) {
    REQUIRE(par && s, "null parameters or shot geometry");
    REQUIRE(par->nmodels == 2 || par->nmodels == 3 || par->nmodels == 6,
            "The model must contain two (Vp, rho), three (Vp, Vs, rho) or six (Vp, Vs, rho, delta, eps, gamma) parameters");
    REQUIRE(s->nx >= 12 && s->ny >= 12 && s->nz >= 12, "every grid axis needs at least 12 points (SBP closures)");
    REQUIRE(s->dx > 0 && s->dy > 0 && s->dz > 0, "grid spacing must be positive");
    REQUIRE(s->nt_data >= 2 && s->dt_data > 0, "the data time axis needs nt >= 2 and dt > 0");
    REQUIRE(s->sxyz != nullptr, "null source position");
    REQUIRE(s->nrcv >= 0 && (s->nrcv == 0 || s->rxyz != nullptr), "null receiver positions");
    REQUIRE(par->resampling == 0 || par->resampling == 1, "resampling must be 0 (sinc) or 1 (linear)");
    REQUIRE(par->sinc_half_length >= 1, "sinc_half_length must be >= 1");
    REQUIRE(par->seismotype == 0 || par->seismotype == 1, "seismotype must be 0 or 1");
    REQUIRE(par->courant > 0 || par->dt > 0, "courant must be positive");
    REQUIRE(!(par->gl > 0) || par->nmodels == 2 || s->nrcv == 0 || s->rdipaz != nullptr,
            "DAS channels (gl > 0) need their dip and azimuth (shot rdipaz)");
}

// The per-physics overrides of analyzeGeometry (we_op.cpp:11-18) on a copy.
fwi3d_op_param normalized(const fwi3d_op_param* in) {
    fwi3d_op_param p = *in;
    if (p.nmodels == 2) {
        p.mt = 0;
        p.gl = 0;
        p.free_surface_stiffness = std::max(p.free_surface_stiffness, 1.0f);
        p.soft_clip = 0;
    }
    if (!(p.gl > 0)) p.gl = 0;
    for (int f = 0; f < 6; f++) p.bc[f] = std::max(0, p.bc[f]);
    return p;
}

// Time axis of the propagation from the velocity extremes (we_op.cpp:378-444), in
// the reference's single-precision arithmetic.
void time_axis(fwi3d_op_param* par, float vmax, float vmin, const fwi3d_shot* s) {
    par->vmax = vmax;
    par->vmin = vmin;
    REQUIRE(vmax > 0, "maximum velocity must be positive");
    const float tmax = (float)(s->nt_data - 1) * s->dt_data;
    const float hmin = std::min(std::min(s->dx, s->dy), s->dz);
    if (par->dt == 0) {
        par->dt = par->courant * hmin / par->vmax;
    } else if (par->dt > 0) {
        par->dt = std::min(par->dt, s->dt_data);
        par->courant = par->dt * par->vmax / hmin;
    } else {
        par->dt = par->courant * hmin / par->vmax;
        if (s->dt_data > par->dt) par->dt = (float)((double)s->dt_data / ::ceil((double)(s->dt_data / par->dt)));
        else par->dt = s->dt_data;
        par->courant = par->dt * par->vmax / hmin;
    }
    par->nt = (int)(::round((double)(tmax / par->dt)) + 1);
    if (par->sub < 0) par->sub = (int)::round((double)(s->dt_data / par->dt));
    REQUIRE(par->nt >= 2, "the propagation needs at least 2 time samples");
}

ClipBounds clip_bounds(const fwi3d_op_param* p) {
    ClipBounds b;
    b.vpmin = p->vpmin; b.vpmax = p->vpmax; b.vsmin = p->vsmin; b.vsmax = p->vsmax;
    b.rhomin = p->rhomin; b.rhomax = p->rhomax; b.delmin = p->deltamin; b.delmax = p->deltamax;
    b.epsmin = p->epsilonmin; b.epsmax = p->epsilonmax; b.gammin = p->gammamin; b.gammax = p->gammamax;
    b.hard = p->nmodels == 2 ? 1 : (p->soft_clip ? 0 : 1);
    return b;
}

// Device buffers of the shot operator (each role has its own growable allocation on
// the context, so a later call never moves an earlier result).
struct OpBufs {
    float* conv;    // [nm][N] converted model
    float* srcp;    // [nscomp][nt] source at the propagation rate
    float* rcvp;    // [ncomp][nrcv][nt] traces at the propagation rate
    float* tmp;     // [ncomp][nrcv][nt]
    double* g64;    // [ng][N] raw gradient (jacobianT only)
    unsigned* red;  // [2] velocity extreme keys
};

OpBufs op_buffers(fwi3d_ctx* c, const ShotDims& d, const fwi3d_shot* s, bool need_grad) {
    OpBufs B{};
    c->opmodel.ensure((size_t)d.nm * d.N * 4);
    c->opscal.ensure(256);
    B.conv = c->opmodel.as<float>();
    B.red = c->opscal.as<unsigned>();
    if (need_grad) {
        c->opgrad.ensure((size_t)d.ng * d.N * 8);
        B.g64 = c->opgrad.as<double>();
    }
    return B;
}

void trace_buffers(fwi3d_ctx* c, OpBufs& B, const ShotDims& d, int nt, const fwi3d_shot* s) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t ntr = (size_t)d.ncomp * s->nrcv;
    const size_t b_src = al((size_t)d.nscomp * nt * 4), b_rcv = al(std::max<size_t>(1, ntr * nt) * 4);
    c->optrace.ensure(b_src + 2 * b_rcv);
    char* p = c->optrace.as<char>();
    B.srcp = (float*)p;
    B.rcvp = (float*)(p + b_src);
    B.tmp = (float*)(p + b_src + b_rcv);
}

// analyzeModel + convert_model on the device; sets par's time axis.
void analyze_convert(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* s, const ShotDims& d, const float* d_raw,
                     float* conv, unsigned* red) {
    const ClipBounds B = clip_bounds(par);
    hipLaunchKernelGGL(key_init, dim3(1), dim3(64), 0, c->stream, red);
    const unsigned grid = (unsigned)std::min<long long>(4096, (d.N + kThreads - 1) / kThreads);
    const long long nxz = (long long)d.nx * d.nz;
    if (d.nm == 2)
        hipLaunchKernelGGL(prep_model<2>, dim3(grid), dim3(kThreads), 0, c->stream, d_raw, conv, d.N, nxz, B, red);
    else if (d.nm == 3)
        hipLaunchKernelGGL(prep_model<3>, dim3(grid), dim3(kThreads), 0, c->stream, d_raw, conv, d.N, nxz, B, red);
    else
        hipLaunchKernelGGL(prep_model<6>, dim3(grid), dim3(kThreads), 0, c->stream, d_raw, conv, d.N, nxz, B, red);
    HIP_OK(hipGetLastError());
    unsigned keys[2];
    HIP_OK(hipMemcpyAsync(keys, red, sizeof(keys), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    time_axis(par, unkey(keys[0]), unkey(keys[1]), s);
}

fwi3d_param prop_param(const fwi3d_op_param* p) {
    fwi3d_param q{};
    q.nt = p->nt;
    q.sub = p->sub;
    q.dt = p->dt;
    for (int f = 0; f < 6; f++) { q.bc[f] = p->bc[f]; q.taper[f] = p->taper[f]; }
    q.taper_strength = p->taper_strength;
    q.free_surface_stiffness = p->free_surface_stiffness;
    q.snap_precision = p->snap_precision;
    return q;
}

// Injector tables of one shot (host arrays; the engine uploads them per call).
struct HostInj {
    std::vector<int> xi, yi, zi;
    std::vector<float> xw, yw, zw;
    fwi3d_injector t{};
    void build(int kind, const float* loc, int npts, const fwi3d_shot* s) {
        const int nw = kind ? 6 : 3;
        xi.assign(3 * (size_t)npts, 0); yi = xi; zi = xi;
        xw.assign(2 * (size_t)nw * npts, 0.f); yw = xw; zw = xw;
        if (npts > 0) {
            const int rc = fwi3d_injector_tables(kind, loc, npts, s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->ox, s->oy,
                                                 s->oz, xi.data(), yi.data(), zi.data(), xw.data(), yw.data(), zw.data());
            REQUIRE(rc == FWI3D_OK, "a source or receiver position lies outside the grid");
        }
        t = fwi3d_injector{npts, kind, xi.data(), yi.data(), zi.data(), xw.data(), yw.data(), zw.data()};
    }
    // DAS channels (dipole_m3) of gauge length gl at rxyz with fiber dip/azimuth rdipaz
    void build_dipole(const float* rxyz, const float* rdipaz, int npts, float gl, const fwi3d_shot* s) {
        xi.assign(6 * (size_t)npts, 0); yi = xi; zi = xi;
        xw.assign(12 * (size_t)npts, 0.f); yw = xw; zw = xw;
        if (npts > 0) {
            std::vector<float> loc5(5 * (size_t)npts);
            for (int r = 0; r < npts; r++) {
                for (int k = 0; k < 3; k++) loc5[5 * r + k] = rxyz[3 * r + k];
                loc5[5 * r + 3] = rdipaz[2 * r];
                loc5[5 * r + 4] = rdipaz[2 * r + 1];
            }
            const int rc = fwi3d_dipole_tables(loc5.data(), npts, gl, s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->ox,
                                               s->oy, s->oz, xi.data(), yi.data(), zi.data(), xw.data(), yw.data(),
                                               zw.data());
            REQUIRE(rc == FWI3D_OK, "a DAS channel end lies outside the grid");
        }
        t = fwi3d_injector{npts, 2, xi.data(), yi.data(), zi.data(), xw.data(), yw.data(), zw.data()};
    }
    // the receivers of a shot: points, or DAS channels when gl > 0
    void build_receivers(const fwi3d_op_param* par, const fwi3d_shot* s) {
        if (par->gl > 0) build_dipole(s->rxyz, s->rdipaz, s->nrcv, par->gl, s);
        else build(0, s->rxyz, s->nrcv, s);
    }
};

// Per-channel (cos dip, sin dip, cos az, sin az) of DAS channels on the device, each the
// double cosine / sine rounded to float as the reference's dipole_to_strain
// (misc.cpp:11-42) computes them.
const float4* das_trig(fwi3d_ctx* c, const fwi3d_shot* s) {
    std::vector<float4> h((size_t)std::max(1, s->nrcv));
    for (int r = 0; r < s->nrcv; r++) {
        const double dip = s->rdipaz[2 * r], az = s->rdipaz[2 * r + 1];
        h[r] = make_float4((float)std::cos(dip), (float)std::sin(dip), (float)std::cos(az), (float)std::sin(az));
    }
    c->opdas.ensure(h.size() * sizeof(float4));
    HIP_OK(hipMemcpyAsync(c->opdas.p, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));  // h is a stack vector
    return c->opdas.as<const float4>();
}

void propagate_or_throw(fwi3d_ctx* c, int physics, int adj, const float* model, const float* src, float* rcv,
                        const fwi3d_injector* inj, const fwi3d_injector* ext, double* g64, const fwi3d_param* p,
                        const fwi3d_shot* s) {
    const int rc = fwi3d_propagate_device(c, physics, adj, model, src, rcv, inj, ext, g64, p, s->nx, s->ny, s->nz,
                                          s->dx, s->dy, s->dz);
    if (rc != FWI3D_OK) throw Err{rc, fwi3d::thread_error()};
}

// apply_forward (we_op.cpp:872-946): raw model and source wavelets at the data rate
// -> data [nrcomp][nrcv][nt_data] (add or overwrite). par is updated (dt, nt, sub).
void forward_device(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* s, int add, const float* d_model,
                    const float* d_src, float* d_data) {
    *par = normalized(par);
    const ShotDims d = shot_dims(par, s);
    OpBufs B = op_buffers(c, d, s, false);
    analyze_convert(c, par, s, d, d_model, B.conv, B.red);
    const int nt = par->nt;
    trace_buffers(c, B, d, nt, s);
    const Axis Td{s->nt_data, s->dt_data}, Tp{nt, par->dt};
    const float alpha = par->dt / s->dt_data;
    hipLaunchKernelGGL(resample_fwd, dim3(nblocks((long long)d.nscomp * nt)), dim3(kThreads), 0, c->stream, d_src, Td,
                       B.srcp, Tp, d.nscomp, par->resampling, par->sinc_half_length, 1.0f, 0);
    HostInj si, ri;
    si.build(par->mt ? 1 : 0, s->sxyz, 1, s);
    ri.build_receivers(par, s);
    const fwi3d_param pp = prop_param(par);
    propagate_or_throw(c, d.physics, 0, B.conv, B.srcp, B.rcvp, &si.t, &ri.t, nullptr, &pp, s);
    if (par->gl > 0 && s->nrcv > 0)  // dipole differences -> strain along the fiber (component 0)
        hipLaunchKernelGGL(dipole_strain, dim3(nblocks((long long)s->nrcv * nt)), dim3(kThreads), 0, c->stream, B.rcvp,
                           das_trig(c, s), s->nrcv, nt, par->gl, 0);
    const int ntr = d.nrcomp * s->nrcv;
    const float* r = B.rcvp;
    if (par->seismotype == 1) {
        hipLaunchKernelGGL(time_derivative, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, c->stream, B.rcvp,
                           B.tmp, ntr, nt, par->dt, 0);
        r = B.tmp;
    }
    if (ntr > 0)
        hipLaunchKernelGGL(resample_adj, dim3(nblocks((long long)ntr * s->nt_data)), dim3(kThreads), 0, c->stream,
                           d_data, Td, r, Tp, ntr, par->resampling, par->sinc_half_length, alpha, add);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
}

// apply_jacobianT (we_op.cpp:948-1078): data residual [nrcomp][nrcv][nt_data] ->
// gradient w.r.t. the raw parameters, out (+)= scale * grad. Needs the snapshots of
// the matching forward call on this context.
template <class OutT>
void jacobianT_device(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* s, int add, OutT* d_grad,
                      const float* d_model0, const float* d_data, double scale) {
    *par = normalized(par);
    const ShotDims d = shot_dims(par, s);
    OpBufs B = op_buffers(c, d, s, true);
    analyze_convert(c, par, s, d, d_model0, B.conv, B.red);
    const int nt = par->nt;
    trace_buffers(c, B, d, nt, s);
    const Axis Td{s->nt_data, s->dt_data}, Tp{nt, par->dt};
    const float alpha = par->dt / s->dt_data;
    const int ntr = d.nrcomp * s->nrcv;
    if (ntr > 0) {
        hipLaunchKernelGGL(resample_fwd, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, c->stream, d_data, Td,
                           B.tmp, Tp, ntr, par->resampling, par->sinc_half_length, 1.0f, 0);
        hipLaunchKernelGGL(adjoint_source, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, c->stream, B.tmp,
                           B.rcvp, ntr, nt, par->dt, -alpha);
    }
    float* asrc = B.rcvp;
    if (par->seismotype == 1 && ntr > 0) {
        hipLaunchKernelGGL(time_derivative, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, c->stream, B.rcvp,
                           B.tmp, ntr, nt, par->dt, 1);
        asrc = B.tmp;
    }
    if (par->gl > 0 && s->nrcv > 0)  // strain -> the three dipole components (writes all three)
        hipLaunchKernelGGL(dipole_strain, dim3(nblocks((long long)s->nrcv * nt)), dim3(kThreads), 0, c->stream, asrc,
                           das_trig(c, s), s->nrcv, nt, par->gl, 1);
    hipLaunchKernelGGL(zero_f64, dim3(1024), dim3(kThreads), 0, c->stream, B.g64, (long long)d.ng * d.N);
    HostInj si, ri;
    si.build(par->mt ? 1 : 0, s->sxyz, 1, s);
    ri.build_receivers(par, s);
    const fwi3d_param pp = prop_param(par);
    propagate_or_throw(c, d.physics, 1, B.conv, asrc, nullptr, &ri.t, &si.t, B.g64, &pp, s);
    hipLaunchKernelGGL(gradient_finish<OutT>, dim3(nblocks(d.N)), dim3(kThreads), 0, c->stream, B.g64, B.conv, d_grad,
                       d.nx, d.ny, d.nz, s->dx, s->dy, s->dz, d.ng, add, scale);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
}

int fail(fwi3d_ctx* c, const Err& e) {
    fwi3d::set_error(c, e.msg);
    return e.code;
}

}  // namespace

extern "C" {

void fwi3d_op_param_defaults(fwi3d_op_param* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nmodels = 3;
    p->dt = -1.0f;
    p->courant = 0.5f;
    p->sub = 0;
    p->resampling = 0;
    p->sinc_half_length = 11;
    for (int f = 0; f < 6; f++) { p->bc[f] = 1; p->taper[f] = 0; }
    p->taper_strength = 0.05f;
    p->free_surface_stiffness = 1.05f;
    p->vpmin = 0.2f; p->vpmax = 8.f; p->vsmin = 0.1f; p->vsmax = 5.f; p->rhomin = 0.1f; p->rhomax = 8.f;
    p->deltamin = -0.5f; p->deltamax = 1.f; p->epsilonmin = 0.f; p->epsilonmax = 1.f; p->gammamin = 0.f; p->gammamax = 1.f;
    p->vmax = 8.f; p->vmin = 0.1f;
    p->snap_precision = FWI3D_SNAP_F32;
}

int fwi3d_apply_forward_device(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* shot, int add,
                               const float* d_model, const float* d_src, float* d_data) {
    if (!c) return FWI3D_EINVAL;
    try {
        check_op(par, shot);
        REQUIRE(d_model && d_src && (d_data || shot->nrcv == 0), "null model, source or data");
        HIP_OK(hipSetDevice(c->device));
        forward_device(c, par, shot, add, d_model, d_src, d_data);
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_apply_jacobianT_device(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* shot, int add, double* d_grad,
                                 const float* d_model0, const float* d_data) {
    if (!c) return FWI3D_EINVAL;
    try {
        check_op(par, shot);
        REQUIRE(d_grad && d_model0 && (d_data || shot->nrcv == 0), "null gradient, model or data");
        HIP_OK(hipSetDevice(c->device));
        jacobianT_device<double>(c, par, shot, add, d_grad, d_model0, d_data, 1.0);
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_fwi_shot_device(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* shot, const float* d_model,
                          const float* d_src, const float* d_obs, float* d_resid, double* d_grad, double* misfit) {
    if (!c) return FWI3D_EINVAL;
    try {
        check_op(par, shot);
        REQUIRE(d_model && d_src && d_grad && (d_obs || shot->nrcv == 0), "null model, source, data or gradient");
        HIP_OK(hipSetDevice(c->device));
        const ShotDims d = shot_dims(par, shot);
        const int ntr = d.nrcomp * shot->nrcv;
        const size_t ntd = (size_t)ntr * shot->nt_data;
        // synthetic data of the shot into a buffer of its own (the operator's buffers are
        // re-carved per call)
        const size_t syn_bytes = (std::max<size_t>(1, ntd) * 4 + 255) / 256 * 256;
        c->opsyn.ensure(syn_bytes + 256);  // synthetic traces, then the FP64 misfit slot
        float* syn = c->opsyn.as<float>();
        double* dmis = (double*)(c->opsyn.as<char>() + syn_bytes);
        forward_device(c, par, shot, 0, d_model, d_src, syn);
        hipLaunchKernelGGL(zero_f64, dim3(1), dim3(64), 0, c->stream, dmis, 1LL);
        if (ntd > 0)
            hipLaunchKernelGGL(residual_misfit, dim3(nblocks((long long)ntd)), dim3(kThreads), 0, c->stream, syn, d_obs,
                               d_resid, ntr, shot->nt_data, shot->dt_data, dmis);
        jacobianT_device<double>(c, par, shot, 1, d_grad, d_model, syn, 1.0);
        double m = 0;
        HIP_OK(hipMemcpy(&m, dmis, sizeof(double), hipMemcpyDeviceToHost));
        if (misfit) *misfit = 0.5 * (double)shot->dt_data * m;
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// Host-array variants: upload, run the device path, download.
int fwi3d_apply_forward(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* shot, int add, const float* model,
                        const float* src, float* data) {
    if (!c) return FWI3D_EINVAL;
    try {
        check_op(par, shot);
        REQUIRE(model && src && (data || shot->nrcv == 0), "null model, source or data");
        HIP_OK(hipSetDevice(c->device));
        const ShotDims d = shot_dims(par, shot);
        const size_t bm = (size_t)d.nm * d.N * 4, bs = (size_t)d.nscomp * shot->nt_data * 4,
                     bd = (size_t)d.nrcomp * shot->nrcv * shot->nt_data * 4;
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        c->tmpio.ensure(al(bm) + al(bs) + al(std::max<size_t>(bd, 4)));
        char* p = c->tmpio.as<char>();
        float* dm = (float*)p;
        float* ds = (float*)(p + al(bm));
        float* dd = (float*)(p + al(bm) + al(bs));
        HIP_OK(hipMemcpyAsync(dm, model, bm, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(ds, src, bs, hipMemcpyHostToDevice, c->stream));
        if (add && bd) HIP_OK(hipMemcpyAsync(dd, data, bd, hipMemcpyHostToDevice, c->stream));
        forward_device(c, par, shot, add, dm, ds, dd);
        if (bd) HIP_OK(hipMemcpy(data, dd, bd, hipMemcpyDeviceToHost));
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_apply_jacobianT(fwi3d_ctx* c, fwi3d_op_param* par, const fwi3d_shot* shot, int add, float* grad,
                          const float* model0, const float* data) {
    if (!c) return FWI3D_EINVAL;
    try {
        check_op(par, shot);
        REQUIRE(grad && model0 && (data || shot->nrcv == 0), "null gradient, model or data");
        HIP_OK(hipSetDevice(c->device));
        const ShotDims d = shot_dims(par, shot);
        const size_t bm = (size_t)d.nm * d.N * 4, bg = (size_t)d.ng * d.N * 4,
                     bd = (size_t)d.nrcomp * shot->nrcv * shot->nt_data * 4;
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        c->tmpio.ensure(al(bm) + al(bg) + al(std::max<size_t>(bd, 4)));
        char* p = c->tmpio.as<char>();
        float* dm = (float*)p;
        float* dg = (float*)(p + al(bm));
        float* dd = (float*)(p + al(bm) + al(bg));
        HIP_OK(hipMemcpyAsync(dm, model0, bm, hipMemcpyHostToDevice, c->stream));
        if (bd) HIP_OK(hipMemcpyAsync(dd, data, bd, hipMemcpyHostToDevice, c->stream));
        if (add) HIP_OK(hipMemcpyAsync(dg, grad, bg, hipMemcpyHostToDevice, c->stream));
        jacobianT_device<float>(c, par, shot, add, dg, dm, dd, 1.0);
        HIP_OK(hipMemcpy(grad, dg, bg, hipMemcpyDeviceToHost));
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

// Host-side analyzeModel (we_op.cpp:234-446) on a caller's model copy, in place:
// clipping, admissibility fixes and the time axis. Returns the time axis in par.
int fwi3d_analyze_model(fwi3d_op_param* par, float* model, const fwi3d_shot* s) {
    try {
        check_op(par, s);
        REQUIRE(model != nullptr, "null model");
        *par = normalized(par);
        const long long N = (long long)s->nx * s->ny * s->nz;
        const ClipBounds B = clip_bounds(par);
        float vmax = -INFINITY, vmin = INFINITY;
        auto clip = [](float v, float lo, float hi) { return std::max(lo, std::min(hi, v)); };
        if (par->nmodels == 2) {
            for (long long i = 0; i < N; i++) {
                model[i] = clip(model[i], B.vpmin, B.vpmax);
                model[N + i] = clip(model[N + i], B.rhomin, B.rhomax);
                vmax = std::max(vmax, model[i]);
                vmin = std::min(vmin, model[i]);
            }
        } else {
            if (par->nmodels == 6) {
                for (long long i = 0; i < (long long)s->nx * s->nz; i++) {
                    const float vp = model[i], vs = model[N + i];
                    float& del = model[3 * N + i];
                    const float v1 = (float)(-0.5 * (double)(1.f - vs * vs / (vp * vp)));
                    const float v2 = (float)(2.0 / (double)(vp * vp / (vs * vs) - 1.f));
                    if (del < v1) del = (float)(0.9999 * (double)v1);
                    else if (del > v2) del = (float)(0.9999 * (double)v2);
                }
            }
            for (long long i = 0; i < N; i++) {
                if (B.hard) {
                    model[i] = clip(model[i], B.vpmin, B.vpmax);
                    model[N + i] = clip(model[N + i], B.vsmin, B.vsmax);
                    model[2 * N + i] = clip(model[2 * N + i], B.rhomin, B.rhomax);
                    if (par->nmodels == 6) {
                        model[3 * N + i] = clip(model[3 * N + i], B.delmin, B.delmax);
                        model[4 * N + i] = clip(model[4 * N + i], B.epsmin, B.epsmax);
                        model[5 * N + i] = clip(model[5 * N + i], B.gammin, B.gammax);
                    }
                    if ((double)model[i] < std::sqrt(2.0) * (double)model[N + i])
                        model[N + i] = (float)((double)model[i] / std::sqrt(2.0001));
                }
                vmax = std::max(vmax, model[i]);
                if (par->nmodels == 6)
                    vmax = std::max(vmax, (float)((double)model[i] * std::sqrt((double)(1.f + 2.f * model[4 * N + i]))));
                vmin = std::min(vmin, model[N + i]);
            }
        }
        time_axis(par, vmax, vmin, s);
        return FWI3D_OK;
    } catch (const Err& e) {
        return e.code;
    }
}

}  // extern "C"

#include "serve.hpp"
