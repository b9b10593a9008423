// Elastic isotropic and VTI kernels (gfx950): one time step of
// nl_we_op_e::propagate (src/cpp/we_op.cpp:646-836) / nl_we_op_vti::propagate
// (we_op.cpp:1290-1497) in two y-marching launches.
//
//   elastic_flux  first derivatives of the displacement -> normal-stress fields
//                 T (lambda*div u; VTI: three c11c/c12/c13 combinations) and the six
//                 shear fluxes F_ac = G * D_c u_a (a != c; G = mu, or c66 for the
//                 x-y coupling in VTI), written to HBM.
//   elastic_step  D2(b) u_c along x,y,z + D_c T_c + D_a F_ac + traction SATs, then
//                 the leapfrog update, absorbing-boundary scaling and cos^2 taper.
//
// The displacement components are stored one after the other ([3][y][x][z]).
// Flux layout in the aux buffer: [T (1 or 3)][F_yx F_zx F_xy F_zy F_xz F_yz].
// The stored-field taper convention is the acoustic one (acoustic.hpp).
#pragma once
#include "acoustic.hpp"
#include "common.hpp"

namespace fwi3d {

// Model fields (padded): elastic {lam, mu, rho}; VTI {lam, mu, rho, c13, eps, c66, c11b}.
struct EModel {
    const float* __restrict__ lam;
    const float* __restrict__ mu;
    const float* __restrict__ rho;
    const float* __restrict__ c13;
    const float* __restrict__ eps;
    const float* __restrict__ c66;
    const float* __restrict__ c11b;
};

enum { FX_YX = 0, FX_ZX = 1, FX_XY = 2, FX_ZY = 3, FX_XZ = 4, FX_YZ = 5 };

struct EArgs {
    const float* __restrict__ curr;  // [3][n]
    const float* __restrict__ prev;
    float* __restrict__ next;
    float* __restrict__ aux;         // [nT + 6][n]
    EModel m;
    Geom g;
    Bound B;
    float dt, dt2;
    int prev_tapered;
    int ychunk, ntz, ntx, nyc;
};

template <bool VTI>
struct Coef {
    // normal-traction coefficients of D_a u_a on a face with normal axis n (axes: 0 x, 1 y, 2 z)
    __device__ static float cnorm(const EModel& M, int n, int a, long long p) {
        if (!VTI) return M.lam[p];
        const float e2 = 1.f + 2.f * M.eps[p];
        if (n == 2) return a == 2 ? M.lam[p] : M.c13[p];
        if (a == n) return e2 * M.lam[p];
        return a == 2 ? M.c13[p] : e2 * (M.lam[p] + 2.f * M.mu[p]) - 2.f * M.c66[p];
    }
    __device__ static float knn(const EModel& M, int n, long long p) {
        if (!VTI || n == 2) return 2.f * M.mu[p];
        return 2.f * (1.f + 2.f * M.eps[p]) * M.mu[p];
    }
    __device__ static float pimp(const EModel& M, int n, long long p) {
        if (!VTI || n == 2) return 0.5f * sqrtf((M.lam[p] + 2.f * M.mu[p]) * M.rho[p]);
        return 0.5f * sqrtf((1.f + 2.f * M.eps[p]) * (M.lam[p] + 2.f * M.mu[p]) * M.rho[p]);
    }
    __device__ static float gshear(const EModel& M, int n, int c, long long p) {
        if (VTI && ((n == 0 && c == 1) || (n == 1 && c == 0))) return M.c66[p];
        return M.mu[p];
    }
};

// Axis helpers: 0 x, 1 y, 2 z.
__device__ __forceinline__ long long ax_stride(const Geom& g, int a) { return a == 0 ? g.sx : (a == 1 ? g.sy : 1); }
__device__ __forceinline__ int ax_len(const Geom& g, int a) { return a == 0 ? g.nx : (a == 1 ? g.ny : g.nz); }
__device__ __forceinline__ float ax_h(const Geom& g, int a) { return a == 0 ? g.dx : (a == 1 ? g.dy : g.dz); }

// Divisor of the absorbing SATs' implicit part for component c at boundary point p.
template <bool VTI>
__device__ __forceinline__ float e_scale(const Geom& g, const Bound& B, const EModel& M, long long p, int ix, int iy,
                                         int iz, int c, float dt) {
    const bool bx = (ix == 0) | (ix == g.nx - 1), by = (iy == 0) | (iy == g.ny - 1), bz = (iz == 0) | (iz == g.nz - 1);
    if (!(bx | by | bz)) return 1.f;
    const bool corner = bx & by & bz;
    const float vs = sqrtf(M.mu[p] / M.rho[p]);
    const float l2m = M.lam[p] + 2.f * M.mu[p];
    float sc = 0.f;
    bool any = false;
    // faces: top(z0) bottom(z1) left(x0) right(x1) front(y0) back(y1)
#pragma unroll
    for (int f = 0; f < 6; f++) {
        if (B.bc[f] != 2) continue;
        const int ax = f < 2 ? 2 : (f < 4 ? 0 : 1);
        const int i = ax == 0 ? ix : (ax == 1 ? iy : iz);
        const int n = ax_len(g, ax);
        if (!((f & 1) ? i == n - 1 : i == 0)) continue;
        any = true;
        float v;
        if (c == ax) {
            const float F = (VTI && ax != 2 && !corner) ? 1.f + 2.f * M.eps[p] : 1.f;
            v = sqrtf(F * l2m / M.rho[p]);
        } else {
            v = vs;
        }
        sc += v * dt / (2.f * ax_h(g, ax) * SAT_H0);
    }
    return any ? 1.f + sc : 1.f;
}

// Traction SAT terms of component c at face point p (only called within 1 cell of a face).
template <bool VTI>
__device__ float e_sat(const EArgs& A, long long p, int ix, int iy, int iz, int c, float pv) {
    const Geom& g = A.g;
    const EModel& M = A.m;
    const long long n3 = g.n;
    float acc = 0.f;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int bc = A.B.bc[f];
        if (bc != 1 && bc != 2) continue;
        const int nax = f < 2 ? 2 : (f < 4 ? 0 : 1);
        const bool hi = f & 1;
        const int i = nax == 0 ? ix : (nax == 1 ? iy : iz);
        const int len = ax_len(g, nax);
        if (!(hi ? i == len - 1 : i == 0)) continue;
        // tangential axes in the order of the reference's traction terms
        const int t1 = nax == 2 ? 0 : (nax == 0 ? 1 : 0);
        const int t2 = nax == 2 ? 1 : 2;
        const float sg = hi ? 1.f : -1.f;
        const float hn = ax_h(g, nax);
        const long long sn = ax_stride(g, nax);
        const long long inward = hi ? -sn : sn;
        const float adh = 1.0f / (hn * SAT_H0);
        const int cidx[3] = {ix, iy, iz};
        auto D = [&](int a, int comp) -> float {
            return d1_point(A.curr + comp * n3, p, cidx[a], ax_len(g, a), ax_stride(g, a)) / ax_h(g, a);
        };
        float term;
        if (c == nax) {
            float tang = Coef<VTI>::cnorm(M, nax, t1, p) * D(t1, t1) + Coef<VTI>::cnorm(M, nax, t2, p) * D(t2, t2) +
                         Coef<VTI>::cnorm(M, nax, nax, p) * D(nax, nax);
            term = sg * tang + Coef<VTI>::knn(M, nax, p) * sat_s(A.curr + nax * n3, p, inward) / hn;
            if (bc == 2) term -= Coef<VTI>::pimp(M, nax, p) * pv / A.dt;
        } else {
            const float G = Coef<VTI>::gshear(M, nax, c, p);
            term = sg * G * D(c, nax) + G * sat_s(A.curr + c * n3, p, inward) / hn;
            if (bc == 2) term -= 0.5f * sqrtf(M.mu[p] * M.rho[p]) * pv / A.dt;
        }
        acc -= adh * term;
    }
    return acc;
}

// ------------------------------------------------------------------------------------
// Pass 1: stresses and fluxes.
template <int TX, bool VTI>
__global__ __launch_bounds__(64 * TX) void elastic_flux(EArgs A) {
    constexpr int TZ = 64;
    __shared__ float su[3][TX + 4][TZ + 4];
    const Geom g = A.g;
    const EModel& M = A.m;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = L % A.ntz, tx = (L / A.ntz) % A.ntx, cy = L / (A.ntz * A.ntx);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = tz * TZ, x0 = tx * TX, iz = z0 + lz, ix = x0 + lx;
    const bool valid = (iz < g.nz) && (ix < g.nx);
    const int y0 = cy * A.ychunk, y1 = min(g.ny, y0 + A.ychunk);
    const long long col = (long long)ix * g.sx + iz;
    const long long n3 = g.n;
    auto ldc = [&](const float* __restrict__ f, int y) -> float {
        return (valid && y >= 0 && y < g.ny) ? f[(long long)y * g.sy + col] : 0.f;
    };
    auto ldp = [&](const float* __restrict__ f, int x, int y, int z) -> float {
        return (x >= 0 && x < g.nx && z >= 0 && z < g.nz) ? f[g.at(x, y, z)] : 0.f;
    };
    float q[3][5];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        q[c][0] = ldc(A.curr + c * n3, y0 - 2);
        q[c][1] = ldc(A.curr + c * n3, y0 - 1);
        q[c][2] = ldc(A.curr + c * n3, y0);
        q[c][3] = ldc(A.curr + c * n3, y0 + 1);
    }
    const float idx = 1.0f / g.dx, idy = 1.0f / g.dy, idz = 1.0f / g.dz;
    (void)idx; (void)idy; (void)idz;
    for (int iy = y0; iy < y1; iy++) {
#pragma unroll
        for (int c = 0; c < 3; c++) q[c][4] = ldc(A.curr + c * n3, iy + 2);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float* u = A.curr + c * n3;
            su[c][lx + 2][lz + 2] = q[c][2];
            if (lx < 2) {
                su[c][lx][lz + 2] = ldp(u, x0 - 2 + lx, iy, iz);
                su[c][TX + 2 + lx][lz + 2] = ldp(u, x0 + TX + lx, iy, iz);
            }
            if (lz < 2) {
                su[c][lx + 2][lz] = ldp(u, ix, iy, z0 - 2 + lz);
                su[c][lx + 2][TZ + 2 + lz] = ldp(u, ix, iy, z0 + TZ + lz);
            }
        }
        __syncthreads();
        if (valid) {
            const long long p = (long long)iy * g.sy + col;
            const bool inx = ix >= 4 && ix < g.nx - 4, iny = iy >= 4 && iy < g.ny - 4, inz = iz >= 4 && iz < g.nz - 4;
            float dxu[3], dyu[3], dzu[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                dxu[c] = inx ? D1C1 * (su[c][lx + 3][lz + 2] - su[c][lx + 1][lz + 2]) + D1C2 * (su[c][lx + 4][lz + 2] - su[c][lx][lz + 2])
                             : d1_point(A.curr + c * n3, p, ix, g.nx, g.sx);
                dyu[c] = iny ? D1C1 * (q[c][3] - q[c][1]) + D1C2 * (q[c][4] - q[c][0])
                             : d1_point(A.curr + c * n3, p, iy, g.ny, g.sy);
                dzu[c] = inz ? D1C1 * (su[c][lx + 2][lz + 3] - su[c][lx + 2][lz + 1]) + D1C2 * (su[c][lx + 2][lz + 4] - su[c][lx + 2][lz])
                             : d1_point(A.curr + c * n3, p, iz, g.nz, 1);
                dxu[c] = dxu[c] / g.dx;
                dyu[c] = dyu[c] / g.dy;
                dzu[c] = dzu[c] / g.dz;
            }
            const float mu = M.mu[p];
            float* out = A.aux;
            int nT;
            if (!VTI) {
                out[p] = M.lam[p] * (dxu[0] + dyu[1] + dzu[2]);
                nT = 1;
            } else {
                const float lam = M.lam[p], e2 = 1.f + 2.f * M.eps[p], c13 = M.c13[p], c66 = M.c66[p];
                const float c11c = e2 * lam, c12 = e2 * (lam + 2.f * mu) - 2.f * c66;
                out[p] = c11c * dxu[0] + c12 * dyu[1] + c13 * dzu[2];
                out[n3 + p] = c11c * dyu[1] + c12 * dxu[0] + c13 * dzu[2];
                out[2 * n3 + p] = c13 * (dxu[0] + dyu[1]) + lam * dzu[2];
                nT = 3;
            }
            const float gxy = VTI ? M.c66[p] : mu;
            float* F = out + nT * n3;
            F[FX_YX * n3 + p] = dxu[1] * gxy;  // G * D_x u_y  (outer D_y, component x)
            F[FX_ZX * n3 + p] = dxu[2] * mu;   // mu * D_x u_z (outer D_z, component x)
            F[FX_XY * n3 + p] = dyu[0] * gxy;  // G * D_y u_x  (outer D_x, component y)
            F[FX_ZY * n3 + p] = dyu[2] * mu;   // mu * D_y u_z (outer D_z, component y)
            F[FX_XZ * n3 + p] = dzu[0] * mu;   // mu * D_z u_x (outer D_x, component z)
            F[FX_YZ * n3 + p] = dzu[1] * mu;   // mu * D_z u_y (outer D_y, component z)
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            q[c][0] = q[c][1]; q[c][1] = q[c][2]; q[c][2] = q[c][3]; q[c][3] = q[c][4];
        }
    }
}

// ------------------------------------------------------------------------------------
// Pass 2: the time step.
//
// LDS planes (x and z neighbours): u_x u_y u_z | b-fields | T_x T_z | F_zx F_xy F_zy F_xz.
// Register queues (y neighbours): u (3), b-fields, T_y, F_yx, F_yz.
// b-fields: elastic {mu}; VTI {mu, c66, c11b}.
template <int TX, bool VTI>
__global__ __launch_bounds__(64 * TX) void elastic_step(EArgs A) {
    constexpr int TZ = 64;
    constexpr int NB = VTI ? 3 : 1;   // D2 coefficient fields
    constexpr int NT = VTI ? 3 : 1;   // normal-stress fields
    constexpr int PU = 0, PB = 3, PT = 3 + NB, PF = PT + (VTI ? 2 : 1);
    constexpr int NPL = PF + 4;
    __shared__ float s[NPL][TX + 4][TZ + 4];
    const Geom g = A.g;
    const EModel& M = A.m;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = L % A.ntz, tx = (L / A.ntz) % A.ntx, cy = L / (A.ntz * A.ntx);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = tz * TZ, x0 = tx * TX, iz = z0 + lz, ix = x0 + lx;
    const bool valid = (iz < g.nz) && (ix < g.nx);
    const int y0 = cy * A.ychunk, y1 = min(g.ny, y0 + A.ychunk);
    const long long col = (long long)ix * g.sx + iz;
    const long long n3 = g.n;
    auto ldc = [&](const float* __restrict__ f, int y) -> float {
        return (valid && y >= 0 && y < g.ny) ? f[(long long)y * g.sy + col] : 0.f;
    };
    auto ldp = [&](const float* __restrict__ f, int x, int y, int z) -> float {
        return (x >= 0 && x < g.nx && z >= 0 && z < g.nz) ? f[g.at(x, y, z)] : 0.f;
    };
    const float* bf[3] = {M.mu, VTI ? M.c66 : M.mu, VTI ? M.c11b : M.mu};
    const float* Tf[3] = {A.aux, A.aux + (VTI ? n3 : 0), A.aux + (VTI ? 2 * n3 : 0)};
    const float* Ff = A.aux + NT * n3;
    // y queues: u(3), b(NB), T_y, F_yx, F_yz
    constexpr int NQ = 3 + NB + 3;
    const float* qsrc[NQ];
#pragma unroll
    for (int c = 0; c < 3; c++) qsrc[c] = A.curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) qsrc[3 + k] = bf[k];
    qsrc[3 + NB] = Tf[1];
    qsrc[4 + NB] = Ff + FX_YX * n3;
    qsrc[5 + NB] = Ff + FX_YZ * n3;
    float q[NQ][5];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        q[k][0] = ldc(qsrc[k], y0 - 2);
        q[k][1] = ldc(qsrc[k], y0 - 1);
        q[k][2] = ldc(qsrc[k], y0);
        q[k][3] = ldc(qsrc[k], y0 + 1);
    }
    // LDS plane sources
    const float* psrc[NPL];
#pragma unroll
    for (int c = 0; c < 3; c++) psrc[PU + c] = A.curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) psrc[PB + k] = bf[k];
    psrc[PT] = Tf[0];
    if (VTI) psrc[PT + 1] = Tf[2];
    psrc[PF + 0] = Ff + FX_ZX * n3;
    psrc[PF + 1] = Ff + FX_XY * n3;
    psrc[PF + 2] = Ff + FX_ZY * n3;
    psrc[PF + 3] = Ff + FX_XZ * n3;

    const float ih2[3] = {1.0f / (g.dx * g.dx), 1.0f / (g.dy * g.dy), 1.0f / (g.dz * g.dz)};

    for (int iy = y0; iy < y1; iy++) {
#pragma unroll
        for (int k = 0; k < NQ; k++) q[k][4] = ldc(qsrc[k], iy + 2);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            float center;
            if (k < 3 + NB) center = q[k][2];  // u and b planes come from the queues
            else center = ldc(psrc[k], iy);
            s[k][lx + 2][lz + 2] = center;
            if (lx < 2) {
                s[k][lx][lz + 2] = ldp(psrc[k], x0 - 2 + lx, iy, iz);
                s[k][TX + 2 + lx][lz + 2] = ldp(psrc[k], x0 + TX + lx, iy, iz);
            }
            if (lz < 2) {
                s[k][lx + 2][lz] = ldp(psrc[k], ix, iy, z0 - 2 + lz);
                s[k][lx + 2][TZ + 2 + lz] = ldp(psrc[k], ix, iy, z0 + TZ + lz);
            }
        }
        __syncthreads();
        if (valid) {
            const long long p = (long long)iy * g.sy + col;
            const bool in2x = ix >= 6 && ix < g.nx - 6, in2y = iy >= 6 && iy < g.ny - 6, in2z = iz >= 6 && iz < g.nz - 6;
            const bool in1x = ix >= 4 && ix < g.nx - 4, in1y = iy >= 4 && iy < g.ny - 4, in1z = iz >= 4 && iz < g.nz - 4;
            // LDS accessors along x and z for plane k at offset o
            auto SX = [&](int k, int o) { return s[k][lx + 2 + o][lz + 2]; };
            auto SZ = [&](int k, int o) { return s[k][lx + 2][lz + 2 + o]; };
            // D1 (times h) of an LDS plane / queue / global field
            auto d1x = [&](int k) {
                return in1x ? D1C1 * (SX(k, 1) - SX(k, -1)) + D1C2 * (SX(k, 2) - SX(k, -2)) : d1_point(psrc[k], p, ix, g.nx, g.sx);
            };
            auto d1z = [&](int k) {
                return in1z ? D1C1 * (SZ(k, 1) - SZ(k, -1)) + D1C2 * (SZ(k, 2) - SZ(k, -2)) : d1_point(psrc[k], p, iz, g.nz, 1);
            };
            auto d1y = [&](int k) {
                return in1y ? D1C1 * (q[k][3] - q[k][1]) + D1C2 * (q[k][4] - q[k][0]) : d1_point(qsrc[k], p, iy, g.ny, g.sy);
            };
            // D2(b) (times h^2) of component c with b-field index kb along axis a
            auto d2 = [&](int c, int kb, int a) -> float {
                if (a == 0)
                    return in2x ? d2_interior(SX(PU + c, -2), SX(PU + c, -1), SX(PU + c, 0), SX(PU + c, 1), SX(PU + c, 2),
                                              SX(PB + kb, -2), SX(PB + kb, -1), SX(PB + kb, 0), SX(PB + kb, 1), SX(PB + kb, 2))
                                : d2_point(A.curr + c * n3, bf[kb], p, ix, g.nx, g.sx);
                if (a == 1)
                    return in2y ? d2_interior(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4], q[3 + kb][0], q[3 + kb][1],
                                              q[3 + kb][2], q[3 + kb][3], q[3 + kb][4])
                                : d2_point(A.curr + c * n3, bf[kb], p, iy, g.ny, g.sy);
                return in2z ? d2_interior(SZ(PU + c, -2), SZ(PU + c, -1), SZ(PU + c, 0), SZ(PU + c, 1), SZ(PU + c, 2),
                                          SZ(PB + kb, -2), SZ(PB + kb, -1), SZ(PB + kb, 0), SZ(PB + kb, 1), SZ(PB + kb, 2))
                            : d2_point(A.curr + c * n3, bf[kb], p, iz, g.nz, 1);
            };
            const bool near = (iz == 0) | (iz == g.nz - 1) | (ix == 0) | (ix == g.nx - 1) | (iy == 0) | (iy == g.ny - 1);
            const float rho = M.rho[p];
            float res[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                // b-field and factor of each D2 term (we_op.cpp:647-657, 1290-1300)
                int kb[3];
                float fa[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    fa[a] = (!VTI && a == c) ? 2.0f : ((VTI && c == 2 && a == 2) ? 2.0f : 1.0f);
                    kb[a] = 0;
                    if (VTI && c < 2 && a < 2) kb[a] = (a == c) ? 2 : 1;
                }
                float acc = (fa[0] * ih2[0]) * d2(c, kb[0], 0);
                acc = acc + (fa[1] * ih2[1]) * d2(c, kb[1], 1);
                acc = acc + (fa[2] * ih2[2]) * d2(c, kb[2], 2);
                // normal stress and shear fluxes (we_op.cpp:660-688)
                if (c == 0) {
                    acc = acc + d1x(PT) / g.dx;
                    acc = acc + d1y(4 + NB) / g.dy;     // D_y F_yx
                    acc = acc + d1z(PF + 0) / g.dy;     // D_z F_zx (spacing dy as in the reference)
                } else if (c == 1) {
                    acc = acc + d1y(3 + NB) / g.dy;     // D_y T_y
                    acc = acc + d1x(PF + 1) / g.dx;     // D_x F_xy
                    acc = acc + d1z(PF + 2) / g.dz;     // D_z F_zy
                } else {
                    acc = acc + d1z(VTI ? PT + 1 : PT) / g.dz;
                    acc = acc + d1x(PF + 3) / g.dx;     // D_x F_xz
                    acc = acc + d1y(5 + NB) / g.dy;     // D_y F_yz
                }
                float pv = A.prev[c * n3 + p];
                if (A.prev_tapered) pv = taper_apply(pv, A.B, ix, iy, iz);
                if (near) acc += e_sat<VTI>(A, p, ix, iy, iz, c, pv);
                float nv = A.dt2 * acc / rho + 2.f * q[c][2] - pv;
                if (near) nv = nv / e_scale<VTI>(g, A.B, M, p, ix, iy, iz, c, A.dt);
                res[c] = taper_apply(nv, A.B, ix, iy, iz);
            }
#pragma unroll
            for (int c = 0; c < 3; c++) A.next[c * n3 + p] = res[c];
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            q[k][0] = q[k][1]; q[k][1] = q[k][2]; q[k][2] = q[k][3]; q[k][3] = q[k][4];
        }
    }
}

// Source injection of one displacement component: next += taper(dt^2 inj / rho / scale);
// mode 1: initial prev = 0.5 dt^2 inj(src[0]) / rho.
template <bool VTI>
__global__ void elastic_inject(float* __restrict__ out, EModel M, Geom g, Bound B, InjTab I, const float* __restrict__ src,
                               int nt, int it, float dt, int comp, int mode) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= I.npts * 27) return;
    const int r = t / 27, qq = t % 27;
    const int a = qq / 9, bb = (qq / 3) % 3, c = qq % 3;
    const int ix = I.xi[3 * r + bb], iy = I.yi[3 * r + a], iz = I.zi[3 * r + c];
    const long long p = g.at(ix, iy, iz);
    const float v = inj_value(I, src, nt, it, comp, r, a, bb, c);
    float add;
    if (mode == 1) {
        add = v * (0.5f * dt * dt / M.rho[p]);
    } else {
        add = dt * dt * v / M.rho[p];
        add = add / e_scale<VTI>(g, B, M, p, ix, iy, iz, comp, dt);
        add = taper_apply(add, B, ix, iy, iz);
    }
    atomicAdd(out + p, add);
}

// Imaging condition of nl_we_op_e / nl_we_op_vti ::compute_gradients
// (we_op.cpp:488-565, 1108-1216) with FP64 accumulation. `lag` is the adjoint field
// of the previous step (the reference uses the derivatives computed one step earlier);
// `adj` the current adjoint field (rho term). Gradients in the unpadded layout.
template <bool VTI>
__global__ void elastic_grad(const float* __restrict__ fm, const float* __restrict__ fn, const float* __restrict__ fp,
                             const float* __restrict__ adj, const float* __restrict__ lag, EModel M, Geom g, float sdt,
                             int branch, double* __restrict__ gl, double* __restrict__ gm, double* __restrict__ gr) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y, iy = blockIdx.z;
    if (iz >= g.nz) return;
    const long long p = g.at(ix, iy, iz);
    const long long n3 = g.n;
    float F[3][3], Aj[3][3];  // [deriv axis][component]
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F[0][c] = d1_point(fn + c * n3, p, ix, g.nx, g.sx) / g.dx;
        F[1][c] = d1_point(fn + c * n3, p, iy, g.ny, g.sy) / g.dy;
        F[2][c] = d1_point(fn + c * n3, p, iz, g.nz, 1) / g.dz;
        Aj[0][c] = d1_point(lag + c * n3, p, ix, g.nx, g.sx) / g.dx;
        Aj[1][c] = d1_point(lag + c * n3, p, iy, g.ny, g.sy) / g.dy;
        Aj[2][c] = d1_point(lag + c * n3, p, iz, g.nz, 1) / g.dz;
    }
    // forward-field terms named as in the reference: temp[0..8]
    const float t0 = F[0][0], t1 = F[1][1], t2 = F[2][2], t3 = F[0][2], t4 = F[2][0], t5 = F[0][1], t6 = F[1][0],
                t7 = F[1][2], t8 = F[2][1];
    const float w = branch == 0 ? 0.5f : 1.0f;
    float gL, gM;
    if (!VTI) {
        gL = w * sdt * (Aj[0][0] + Aj[1][1] + Aj[2][2]) * (t0 + t1 + t2);
        gM = w * sdt * ((Aj[2][0] + Aj[0][2]) * (t3 + t4) + (Aj[1][0] + Aj[0][1]) * (t5 + t6) +
                        (Aj[2][1] + Aj[1][2]) * (t7 + t8) + 2.f * Aj[0][0] * t0 + 2.f * Aj[1][1] * t1 +
                        2.f * Aj[2][2] * t2);
    } else {
        const float L = M.lam[p], U = M.mu[p], E = M.eps[p], C13 = M.c13[p], C66 = M.c66[p];
        const float del = ((C13 + U) * (C13 + U) - (L + U) * (L + U)) / (2.f * (L + U) * (L + 2.f * U));
        const float gam = (C66 - U) / (2.f * U);
        const float v1 = sqrtf(2.f * (L + 2.f * U) * (L + U) * del + (L + U) * (L + U));
        const float v2 = ((1.f + 2.f * del) * L + (1.f + 3.f * del) * U) / v1;
        const float v3 = ((1.f + 3.f * del) * L + (1.f + 4.f * del) * U) / v1 - 1.f;
        const float v4 = (Aj[0][0] + Aj[1][1]) * t2 + Aj[2][2] * (t0 + t1);
        gL = w * sdt * ((1.f + 2.f * E) * (Aj[0][0] + Aj[1][1]) * (t0 + t1) + Aj[2][2] * t2 + v2 * v4);
        gM = w * sdt * (2.f * (1.f + 2.f * E) * (Aj[0][0] * t0 + Aj[1][1] * t1) + 2.f * Aj[2][2] * t2 +
                        4.f * (E - gam) * (Aj[0][0] * t1 + Aj[0][0] * t0) + (Aj[1][2] + Aj[2][1]) * (t7 + t8) +
                        (Aj[0][2] + Aj[2][0]) * (t3 + t4) + (1.f + 2.f * gam) * (Aj[0][1] + Aj[1][0]) * (t5 + t6) +
                        v3 * v4);
    }
    float rt = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const long long qc = c * n3 + p;
        const float fpv = fp ? fp[qc] : 0.f;
        const float d2t = branch == 1 ? fpv - 2.f * fn[qc] + fm[qc] : (branch == 0 ? fpv - fn[qc] : -fn[qc] + fm[qc]);
        rt += adj[qc] * d2t;
    }
    const long long o = ((long long)iy * g.nx + ix) * g.nz + iz;
    gl[o] += (double)gL;
    gm[o] += (double)gM;
    gr[o] += (double)(1.0f / sdt * rt);
}

}  // namespace fwi3d
