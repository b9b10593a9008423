// Elastic isotropic and VTI kernels (gfx950): one time step of
// nl_we_op_e::propagate (src/cpp/we_op.cpp:646-836) / nl_we_op_vti::propagate
// (we_op.cpp:1290-1497) in four launches (batch_kernels.hpp):
//
//   flux      (whole grid) first derivatives of the displacement -> normal stress T
//             (lambda*div u; VTI: three c11c/c12/c13 combinations) and the six shear
//             fluxes F_ac = G * D_c u_a (a != c; G = mu, or c66 for the x-y coupling in
//             VTI), written to HBM (F_zx and F_xz only near the faces).
//   interior  points 6+ cells inside every face: 9x D2(b) u_c + D_c T_c + D_a F_ac from
//             LDS x-z planes and y register queues, leapfrog update, cos^2 taper.
//   boundary  the edges and corners of the 6-cell shell, point-wise: the same terms
//             with SBP closure rows, the traction SATs, the update, the absorbing
//             scaling and the taper.
//   slabs     the rest of the shell, marching: the z slabs (elastic_zslab_tile) and
//             the x and y slabs (elastic_side_tile).
//
// The displacement components are stored one after the other ([3][y][x][z]).
// Flux layout in the aux buffer: [T (1 or 3)][F_yx F_zx F_xy F_zy F_xz F_yz].
// The stored-field taper convention is the acoustic one (acoustic.hpp).
#pragma once
#include "acoustic.hpp"
#include "common.hpp"

namespace fwi3d {

enum { FX_YX = 0, FX_ZX = 1, FX_XY = 2, FX_ZY = 3, FX_XZ = 4, FX_YZ = 5 };

// Kernel arguments. Model fields (padded): elastic {lam, mu, rho}; VTI adds
// {c13, eps, c66, c11b}.
struct EArgs {
    const float* curr;  // [3][n]
    const float* prev;
    float* next;
    float* aux;         // [nT + 6][n]
    const float* lam;
    const float* mu;
    const float* rho;
    const float* c13;
    const float* eps;
    const float* c66;
    const float* c11b;
    const float* tz2;
    const float* tx2;
    const float* ty2;
    Geom g;
    int bc[6];
    float dt, dt2;
    int prev_tapered;
    // y-march tiling: tiles of 64 (z) x TX (x) x ychunk (y) from the corner (org, org, org);
    // the y march of the last chunk stops at yend. ysplit (flux pass only): two chunks,
    // the 8 planes at each y face (the planes the slab and edge kernels read fluxes from
    // when the fused kernel updates the planes between them)
    int ychunk, ntz, ntx, nyc, org, yend, ysplit;
};

// ------------------------------------------------------------------------------------
// Pass 1: stresses and fluxes over the whole grid.
// Launched by elastic_flux_batch (batch_kernels.hpp) at 5 waves per SIMD.
template <int TX, bool VTI, class LFn>
__device__ __forceinline__ void elastic_flux_tile(const EArgs A, const LFn& tile_id) {
    constexpr int TZ = 64;
    __shared__ float su[3][TX + 4][TZ + 4];
    const float* __restrict__ curr = A.curr;
    const GAS float* __restrict__ gcur = gas(A.curr);
    GAS float* __restrict__ aux = gas(A.aux);
    const GAS float* __restrict__ lamf = gas(A.lam);
    const GAS float* __restrict__ muf = gas(A.mu);
    const GAS float* __restrict__ c13f = gas(A.c13);
    const GAS float* __restrict__ epsf = gas(A.eps);
    const GAS float* __restrict__ c66f = gas(A.c66);
    const int nx = A.g.nx, ny = A.g.ny, nz = A.g.nz, sx = A.g.sx, sy = A.g.sy, n3 = A.g.n;
    const float dx = A.g.dx, dy = A.g.dy, dz = A.g.dz;
    // y chunks fastest in the workgroup order: the chunks above and below, which re-read
    // this chunk's edge planes, run on the same XCD at the same time (measured faster
    // for this kernel; the interior kernel prefers z tiles fastest)
    const int L = tile_id();
    const int cy = L % A.nyc, tz = (L / A.nyc) % A.ntz, tx = L / (A.nyc * A.ntz);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = A.org + tz * TZ, x0 = A.org + tx * TX, iz = z0 + lz, ix = x0 + lx;
    const bool valid = (iz < nz) && (ix < nx);
    const int y0 = A.ysplit ? (cy == 0 ? 0 : ny - 8) : A.org + cy * A.ychunk;
    const int y1 = A.ysplit ? y0 + 8 : min(A.yend, y0 + A.ychunk);
    const int cz = min(iz, nz - 1), cx = min(ix, nx - 1);
    const int col = cx * sx + cz;
    auto ldy = [&](const float* __restrict__ f, int y) { return gas(f)[max(0, min(ny - 1, y)) * sy + col]; };
    HaloMap<TX> H;
    H.init(lx, lz, x0, z0, nx, nz, sx);
    const bool inx = ix >= 4 && ix < nx - 4, inz = iz >= 4 && iz < nz - 4;
    // D1 closure rows along z and x: the row's six points counted from the face sit in
    // the staged plane (LDS column / row `zb|xb + k * zs|xs`) whenever the tile and its
    // 2-cell halo cover them; otherwise (a tile ending within 3 points of a face) the
    // row reads global memory. x is uniform per wave, z differs across lanes 0-3 / the
    // last 4.
    const bool zlo = iz < 4, xlo = ix < 4;
    const int zb = zlo ? 2 - z0 : nz + 1 - z0, zs = zlo ? 1 : -1, zr = zlo ? iz : nz - 1 - iz;
    const int xb = xlo ? 2 - x0 : nx + 1 - x0, xs = xlo ? 1 : -1, xr = xlo ? ix : nx - 1 - ix;
    const bool zlds = zlo ? (z0 <= 2) : (zb - 5 >= 0 && zb < TZ + 4);
    const bool xlds = xlo ? (x0 <= 2 && xb + 5 < TX + 4) : (xb - 5 >= 0 && xb < TX + 4);
    float q[3][6];  // y-2..y+2 and y+3 (loaded one iteration ahead of its use)
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int o = 0; o < 5; o++) q[c][o] = ldy(curr + c * n3, y0 - 2 + o);
    float hx[3] = {0.f, 0.f, 0.f}, hz[3] = {0.f, 0.f, 0.f};  // halos of the next plane to stage
    auto prefetch = [&](int y) {
        const int pl = min(y, ny - 1) * sy;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (H.has_x) hx[c] = gcur[c * n3 + pl + H.xsrc];
            if (H.has_z) hz[c] = gcur[c * n3 + pl + H.zsrc];
        }
    };
    prefetch(y0);
    // model values at this column of the current plane, loaded one iteration ahead:
    // lambda, mu [, eps, c13, c66]
    constexpr int NMQ = VTI ? 5 : 2;
    float mq[NMQ], mn[NMQ];
    auto mload = [&](int y, float* m) {
        const int pp = min(y, ny - 1) * sy + col;
        m[0] = lamf[pp];
        m[1] = muf[pp];
        if (VTI) { m[2] = epsf[pp]; m[3] = c13f[pp]; m[4] = c66f[pp]; }
    };
    mload(y0, mq);
    for (int iy = y0; iy < y1; iy++) {
#pragma unroll
        for (int c = 0; c < 3; c++) q[c][5] = ldy(curr + c * n3, iy + 3);
        mload(iy + 1, mn);
        const int pl = iy * sy;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            su[c][lx + 2][lz + 2] = q[c][2];
            if (H.has_x) su[c][H.xrow][H.xcol] = hx[c];
            if (H.has_z) su[c][H.zrow][H.zcol] = hz[c];
        }
        prefetch(iy + 1);
        __syncthreads();
        if (valid) {
            const int p = pl + col;
            const bool iny = iy >= 4 && iy < ny - 4;
            const bool near8 = ix < 8 || ix >= nx - 8 || iy < 8 || iy >= ny - 8 || iz < 8 || iz >= nz - 8;
            float dxu[3], dyu[3], dzu[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                dxu[c] = inx    ? D1C1 * (su[c][lx + 3][lz + 2] - su[c][lx + 1][lz + 2]) + D1C2 * (su[c][lx + 4][lz + 2] - su[c][lx][lz + 2])
                         : xlds ? d1_closure_at(xlo, xr, [&](int k) { return su[c][xb + k * xs][lz + 2]; })
                                : d1_closure(curr + c * n3, p, ix, nx, sx);
                dyu[c] = iny ? D1C1 * (q[c][3] - q[c][1]) + D1C2 * (q[c][4] - q[c][0])
                             : d1_closure(curr + c * n3, p, iy, ny, sy);
                dzu[c] = inz    ? D1C1 * (su[c][lx + 2][lz + 3] - su[c][lx + 2][lz + 1]) + D1C2 * (su[c][lx + 2][lz + 4] - su[c][lx + 2][lz])
                         : zlds ? d1_closure_at(zlo, zr, [&](int k) { return su[c][lx + 2][zb + k * zs]; })
                                : d1_closure(curr + c * n3, p, iz, nz, 1);
                dxu[c] = dxu[c] / dx;
                dyu[c] = dyu[c] / dy;
                dzu[c] = dzu[c] / dz;
            }
            const float mu = mq[1];
            int nT;
            if (!VTI) {
                aux[p] = mq[0] * (dxu[0] + dyu[1] + dzu[2]);
                nT = 1;
            } else {
                const float lam = mq[0], e2 = 1.f + 2.f * mq[2], c13 = mq[3], c66 = mq[4];
                const float c11c = e2 * lam, c12 = e2 * (lam + 2.f * mu) - 2.f * c66;
                aux[p] = c11c * dxu[0] + c12 * dyu[1] + c13 * dzu[2];
                aux[n3 + p] = c11c * dyu[1] + c12 * dxu[0] + c13 * dzu[2];
                aux[2 * n3 + p] = c13 * (dxu[0] + dyu[1]) + lam * dzu[2];
                nT = 3;
            }
            const float gxy = VTI ? mq[NMQ - 1] : mu;
            GAS float* __restrict__ F = aux + nT * n3;
            F[FX_YX * n3 + p] = dxu[1] * gxy;  // G * D_x u_y  (outer D_y, component x)
            F[FX_XY * n3 + p] = dyu[0] * gxy;  // G * D_y u_x  (outer D_x, component y)
            F[FX_ZY * n3 + p] = dyu[2] * mu;   // mu * D_y u_z (outer D_z, component y)
            F[FX_YZ * n3 + p] = dzu[1] * mu;   // mu * D_z u_y (outer D_y, component z)
            // the two in-plane fluxes are stored only where the shell and slab kernels
            // read them (within 8 cells of a face); the interior kernel recomputes them
            if (near8) {
                F[FX_ZX * n3 + p] = dxu[2] * mu;  // mu * D_x u_z (outer D_z, component x)
                F[FX_XZ * n3 + p] = dzu[0] * mu;  // mu * D_z u_x (outer D_x, component z)
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            q[c][0] = q[c][1]; q[c][1] = q[c][2]; q[c][2] = q[c][3]; q[c][3] = q[c][4]; q[c][4] = q[c][5];
        }
#pragma unroll
        for (int k = 0; k < NMQ; k++) mq[k] = mn[k];
    }
}

// b-field and factor of the D2 term of component c along axis a (we_op.cpp:647-657,
// 1290-1300): elastic mu (factor 2 along the component's own axis); VTI c11b for
// x-x / y-y, c66 across x-y, mu otherwise (factor 2 for z-z).
template <bool VTI>
__device__ __forceinline__ int d2_bfield(int c, int a) {
    if (VTI && c < 2 && a < 2) return a == c ? 2 : 1;
    return 0;
}
template <bool VTI>
__device__ __forceinline__ float d2_factor(int c, int a) {
    if (!VTI) return a == c ? 2.f : 1.f;
    return (c == 2 && a == 2) ? 2.f : 1.f;
}

// ------------------------------------------------------------------------------------
// Pass 2, interior points.
// LDS planes (x and z neighbours): u_x u_y u_z | b-fields | T_x [T_z] | F_xy F_zy.
// Register queues (y neighbours): u (3), b-fields, T_y, F_yx, F_yz.
// b-fields: elastic {mu}; VTI {mu, c66, c11b}.
// The in-plane fluxes F_zx = mu D_x u_z and F_xz = mu D_z u_x are recomputed from the
// staged u and mu planes (the u_x and u_z planes also stage the 2x2 corner cells of
// their halos), with the flux pass's arithmetic, so they make no round trip to HBM.
// Launched by elastic_interior_batch (batch_kernels.hpp) at 4 waves per SIMD.
template <int TX, bool VTI, class LFn>
__device__ __forceinline__ void elastic_interior_tile(const EArgs& __restrict__ A, const LFn& tile_id) {
    constexpr int TZ = 64;
    constexpr int NB = VTI ? 3 : 1;   // D2 coefficient fields
    constexpr int NT = VTI ? 3 : 1;   // normal-stress fields
    constexpr int PU = 0, PB = 3, PT = 3 + NB, PF = PT + (VTI ? 2 : 1);
    constexpr int NPL = PF + 2;
    constexpr int NQ = 3 + NB + 3;
    __shared__ float s[NPL][TX + 4][TZ + 4];
    const int nx = A.g.nx, ny = A.g.ny, nz = A.g.nz, sx = A.g.sx, sy = A.g.sy, n3 = A.g.n;
    const float dx = A.g.dx, dy = A.g.dy, dz = A.g.dz;
    const float* __restrict__ curr = A.curr;
    const GAS float* __restrict__ prev = gas(A.prev);
    GAS float* __restrict__ next = gas(A.next);
    const GAS float* __restrict__ rhof = gas(A.rho);
    const float* __restrict__ tz2 = A.tz2;
    const float* __restrict__ tx2 = A.tx2;
    const float* __restrict__ ty2 = A.ty2;
    const float dt2 = A.dt2;
    const bool ptap = A.prev_tapered != 0;
    const float* bf[3] = {A.mu, VTI ? A.c66 : A.mu, VTI ? A.c11b : A.mu};
    const float* Ff = A.aux + NT * n3;
    const int L = tile_id();
    const int tz = L % A.ntz, tx = (L / A.ntz) % A.ntx, cy = L / (A.ntz * A.ntx);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = A.org + tz * TZ, x0 = A.org + tx * TX, iz = z0 + lz, ix = x0 + lx;
    const int y0 = A.org + cy * A.ychunk, y1 = min(A.yend, y0 + A.ychunk);
    const int cz = min(iz, nz - 1), cx = min(ix, nx - 1);
    const int col = cx * sx + cz;
    const bool comp_xz = ix >= 6 && ix < nx - 6 && iz >= 6 && iz < nz - 6;
    auto ldy = [&](const float* __restrict__ f, int y) { return gas(f)[max(0, min(ny - 1, y)) * sy + col]; };
    HaloMap<TX> H;
    H.init(lx, lz, x0, z0, nx, nz, sx);

    // queue sources: u(3), b(NB), T_y, F_yx, F_yz
    const float* qsrc[NQ];
#pragma unroll
    for (int c = 0; c < 3; c++) qsrc[c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) qsrc[3 + k] = bf[k];
    qsrc[3 + NB] = A.aux + (VTI ? n3 : 0);
    qsrc[4 + NB] = Ff + FX_YX * n3;
    qsrc[5 + NB] = Ff + FX_YZ * n3;
    // LDS plane sources
    const float* psrc[NPL];
#pragma unroll
    for (int c = 0; c < 3; c++) psrc[PU + c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) psrc[PB + k] = bf[k];
    psrc[PT] = A.aux;
    if (VTI) psrc[PT + 1] = A.aux + 2 * n3;
    psrc[PF + 0] = Ff + FX_XY * n3;
    psrc[PF + 1] = Ff + FX_ZY * n3;
    // corner cells of the u_x and u_z planes (rows x0-2, x0-1, x0+TX, x0+TX+1 by columns
    // z0-2, z0-1, z0+64, z0+65), loaded by the first 16 threads of the last wave
    const int hc = (lx - (TX - 1)) * 64 + lz;
    const bool has_c = hc >= 0 && hc < 16;
    const int cr4 = (hc >> 2) & 3, cc4 = hc & 3;
    const int crow = cr4 < 2 ? cr4 : TX + cr4, ccol = cc4 < 2 ? cc4 : TZ + cc4;
    const int csrc = max(0, min(nx - 1, cr4 < 2 ? x0 - 2 + cr4 : x0 + TX + cr4 - 2)) * sx +
                     max(0, min(nz - 1, cc4 < 2 ? z0 - 2 + cc4 : z0 + TZ + cc4 - 2));
    float hcx = 0.f, hcz = 0.f;

    float q[NQ][6];  // y-2..y+2 and y+3 (loaded one iteration ahead of its use)
#pragma unroll
    for (int k = 0; k < NQ; k++)
#pragma unroll
        for (int o = 0; o < 5; o++) q[k][o] = ldy(qsrc[k], y0 - 2 + o);
    const float ih2[3] = {1.0f / (dx * dx), 1.0f / (dy * dy), 1.0f / (dz * dz)};
    // values of the next plane to stage (centres of the flux planes, halos of all
    // planes), loaded one iteration ahead of their use
    float pc[NPL], hx[NPL], hz[NPL];
#pragma unroll
    for (int k = 0; k < NPL; k++) pc[k] = hx[k] = hz[k] = 0.f;
    // planes differentiated along x / along z (only their halos in that direction are
    // read): u and mu along both; VTI c66 and c11b along x only (the D2 z terms use mu);
    // the normal stress along both (elastic) or T_x along x and T_z along z (VTI); F_xy
    // along x and F_zy along z
    auto along_x = [](int k) { return k <= PT || k == PF + 0; };
    auto along_z = [](int k) { return k < PB + 1 || (VTI ? k == PT + 1 : k == PT) || k == PF + 1; };
    auto prefetch = [&](int y) {
        const int pl = min(y, ny - 1) * sy;
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            const GAS float* __restrict__ f = gas(psrc[k]);
            if (k >= 3 + NB) pc[k] = f[pl + col];
            if (along_x(k) && H.has_x) hx[k] = f[pl + H.xsrc];
            if (along_z(k) && H.has_z) hz[k] = f[pl + H.zsrc];
        }
        if (has_c) {
            hcx = gas(curr)[pl + csrc];
            hcz = gas(curr)[2 * n3 + pl + csrc];
        }
    };
    prefetch(y0);
    // rho and the three prev values at this column of the current plane, loaded one
    // iteration ahead
    float rq[4], rn[4];
    auto rload = [&](int y, float* m) {
        const int pp = min(y, ny - 1) * sy + col;
        m[0] = rhof[pp];
#pragma unroll
        for (int c = 0; c < 3; c++) m[1 + c] = prev[c * n3 + pp];
    };
    rload(y0, rq);

    for (int iy = y0; iy < y1; iy++) {
#pragma unroll
        for (int k = 0; k < NQ; k++) q[k][5] = ldy(qsrc[k], iy + 3);
        rload(iy + 1, rn);
        const int pl = iy * sy;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            s[k][lx + 2][lz + 2] = (k < 3 + NB) ? q[k][2] : pc[k];
            if (along_x(k) && H.has_x) s[k][H.xrow][H.xcol] = hx[k];
            if (along_z(k) && H.has_z) s[k][H.zrow][H.zcol] = hz[k];
        }
        if (has_c) {
            s[PU + 0][crow][ccol] = hcx;
            s[PU + 2][crow][ccol] = hcz;
        }
        prefetch(iy + 1);
        __syncthreads();
        if (comp_xz && iy >= 6 && iy < ny - 6) {
            const int p = pl + col;
            auto SX = [&](int k, int o) { return s[k][lx + 2 + o][lz + 2]; };
            auto SZ = [&](int k, int o) { return s[k][lx + 2][lz + 2 + o]; };
            auto d1x = [&](int k) { return D1C1 * (SX(k, 1) - SX(k, -1)) + D1C2 * (SX(k, 2) - SX(k, -2)); };
            auto d1z = [&](int k) { return D1C1 * (SZ(k, 1) - SZ(k, -1)) + D1C2 * (SZ(k, 2) - SZ(k, -2)); };
            auto d1y = [&](int k) { return D1C1 * (q[k][3] - q[k][1]) + D1C2 * (q[k][4] - q[k][0]); };
            auto S2 = [&](int k, int ox, int oz) { return s[k][lx + 2 + ox][lz + 2 + oz]; };
            // F_zx at (x, z+oz) and F_xz at (x+ox, z), as the flux pass forms them
            auto fzx = [&](int oz) {
                float d = D1C1 * (S2(PU + 2, 1, oz) - S2(PU + 2, -1, oz)) + D1C2 * (S2(PU + 2, 2, oz) - S2(PU + 2, -2, oz));
                d = d / dx;
                return d * S2(PB, 0, oz);
            };
            auto fxz = [&](int ox) {
                float d = D1C1 * (S2(PU + 0, ox, 1) - S2(PU + 0, ox, -1)) + D1C2 * (S2(PU + 0, ox, 2) - S2(PU + 0, ox, -2));
                d = d / dz;
                return d * S2(PB, ox, 0);
            };
            auto d2 = [&](int c, int kb, int a) -> float {
                if (a == 0)
                    return d2_interior(SX(PU + c, -2), SX(PU + c, -1), SX(PU + c, 0), SX(PU + c, 1), SX(PU + c, 2),
                                       SX(PB + kb, -2), SX(PB + kb, -1), SX(PB + kb, 0), SX(PB + kb, 1), SX(PB + kb, 2));
                if (a == 1)
                    return d2_interior(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4], q[3 + kb][0], q[3 + kb][1],
                                       q[3 + kb][2], q[3 + kb][3], q[3 + kb][4]);
                return d2_interior(SZ(PU + c, -2), SZ(PU + c, -1), SZ(PU + c, 0), SZ(PU + c, 1), SZ(PU + c, 2),
                                   SZ(PB + kb, -2), SZ(PB + kb, -1), SZ(PB + kb, 0), SZ(PB + kb, 1), SZ(PB + kb, 2));
            };
            const float rho = rq[0];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float acc = (d2_factor<VTI>(c, 0) * ih2[0]) * d2(c, d2_bfield<VTI>(c, 0), 0);
                acc = acc + (d2_factor<VTI>(c, 1) * ih2[1]) * d2(c, d2_bfield<VTI>(c, 1), 1);
                acc = acc + (d2_factor<VTI>(c, 2) * ih2[2]) * d2(c, d2_bfield<VTI>(c, 2), 2);
                // normal stress and shear fluxes (we_op.cpp:660-688)
                if (c == 0) {
                    acc = acc + d1x(PT) / dx;
                    acc = acc + d1y(4 + NB) / dy;     // D_y F_yx
                    acc = acc + (D1C1 * (fzx(1) - fzx(-1)) + D1C2 * (fzx(2) - fzx(-2))) / dy;  // D_z F_zx (spacing dy as in the reference)
                } else if (c == 1) {
                    acc = acc + d1y(3 + NB) / dy;     // D_y T_y
                    acc = acc + d1x(PF + 0) / dx;     // D_x F_xy
                    acc = acc + d1z(PF + 1) / dz;     // D_z F_zy
                } else {
                    acc = acc + d1z(VTI ? PT + 1 : PT) / dz;
                    acc = acc + (D1C1 * (fxz(1) - fxz(-1)) + D1C2 * (fxz(2) - fxz(-2))) / dx;  // D_x F_xz
                    acc = acc + d1y(5 + NB) / dy;     // D_y F_yz
                }
                float pv = rq[1 + c];
                if (ptap) pv = taper_apply(pv, tz2, tx2, ty2, ix, iy, iz);
                const float nv = dt2 * acc / rho + 2.f * q[c][2] - pv;
                next[c * n3 + p] = taper_apply(nv, tz2, tx2, ty2, ix, iy, iz);
            }
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            q[k][0] = q[k][1]; q[k][1] = q[k][2]; q[k][2] = q[k][3]; q[k][3] = q[k][4]; q[k][4] = q[k][5];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) rq[k] = rn[k];
    }
}

// ------------------------------------------------------------------------------------
// Pass 2, boundary shell: one thread per point, every operand from global memory.
template <bool VTI>
struct ECoef {
    const GAS float *lam, *mu, *rho, *c13, *eps, *c66;
    __device__ ECoef(const float* l, const float* m, const float* r, const float* c, const float* e, const float* s)
        : lam(gas(l)), mu(gas(m)), rho(gas(r)), c13(gas(c)), eps(gas(e)), c66(gas(s)) {}
    // normal-traction coefficient of D_a u_a on a face with normal axis n (0 x, 1 y, 2 z)
    __device__ float cnorm(int n, int a, int p) const {
        if (!VTI) return lam[p];
        const float e2 = 1.f + 2.f * eps[p];
        if (n == 2) return a == 2 ? lam[p] : c13[p];
        if (a == n) return e2 * lam[p];
        return a == 2 ? c13[p] : e2 * (lam[p] + 2.f * mu[p]) - 2.f * c66[p];
    }
    __device__ float knn(int n, int p) const {
        if (!VTI || n == 2) return 2.f * mu[p];
        return 2.f * (1.f + 2.f * eps[p]) * mu[p];
    }
    __device__ float pimp(int n, int p) const {
        if (!VTI || n == 2) return 0.5f * sqrtf((lam[p] + 2.f * mu[p]) * rho[p]);
        return 0.5f * sqrtf((1.f + 2.f * eps[p]) * (lam[p] + 2.f * mu[p]) * rho[p]);
    }
    __device__ float gshear(int n, int c, int p) const {
        if (VTI && ((n == 0 && c == 1) || (n == 1 && c == 0))) return c66[p];
        return mu[p];
    }
};

__device__ __forceinline__ int ax_stride(const Geom& g, int a) { return a == 0 ? g.sx : (a == 1 ? g.sy : 1); }
__device__ __forceinline__ int ax_len(const Geom& g, int a) { return a == 0 ? g.nx : (a == 1 ? g.ny : g.nz); }
__device__ __forceinline__ float ax_h(const Geom& g, int a) { return a == 0 ? g.dx : (a == 1 ? g.dy : g.dz); }

// Divisor of the absorbing SATs' implicit part for component c at boundary point p.
template <bool VTI>
__device__ __forceinline__ float e_scale(const Geom& g, const int* bc, const ECoef<VTI>& M, int p, int ix, int iy, int iz,
                                         int c, float dt) {
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
        if (bc[f] != 2) continue;
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

// One shell point, displacement component c (a compile-time constant, so every axis
// and field choice below folds away and nothing is indexed at run time).
template <bool VTI, int c>
__device__ __forceinline__ void elastic_boundary_point(const EArgs& A, int ix, int iy, int iz) {
    const Geom& g = A.g;
    constexpr int NT = VTI ? 3 : 1;
    const int n3 = g.n;
    const float* __restrict__ curr = A.curr;
    const ECoef<VTI> M{A.lam, A.mu, A.rho, A.c13, A.eps, A.c66};
    auto BF = [&](int k) -> const float* { return k == 0 ? A.mu : (k == 1 ? (VTI ? A.c66 : A.mu) : (VTI ? A.c11b : A.mu)); };
    auto IDX = [&](int a) { return a == 0 ? ix : (a == 1 ? iy : iz); };
    const float* F = A.aux + NT * n3;
    const int p = g.at(ix, iy, iz);
    auto D = [&](const float* f, int a) { return d1_point(f, p, IDX(a), ax_len(g, a), ax_stride(g, a)); };
    const bool face = (iz == 0) | (iz == g.nz - 1) | (ix == 0) | (ix == g.nx - 1) | (iy == 0) | (iy == g.ny - 1);
    const float rho = M.rho[p];
    {
        const float* u = curr + c * n3;
        float acc = (d2_factor<VTI>(c, 0) / (g.dx * g.dx)) * d2_point(u, BF(d2_bfield<VTI>(c, 0)), p, ix, g.nx, g.sx);
        acc = acc + (d2_factor<VTI>(c, 1) / (g.dy * g.dy)) * d2_point(u, BF(d2_bfield<VTI>(c, 1)), p, iy, g.ny, g.sy);
        acc = acc + (d2_factor<VTI>(c, 2) / (g.dz * g.dz)) * d2_point(u, BF(d2_bfield<VTI>(c, 2)), p, iz, g.nz, 1);
        if (c == 0) {
            acc = acc + D(A.aux, 0) / g.dx;
            acc = acc + D(F + FX_YX * n3, 1) / g.dy;
            acc = acc + D(F + FX_ZX * n3, 2) / g.dy;  // spacing dy as in the reference
        } else if (c == 1) {
            acc = acc + D(A.aux + (VTI ? n3 : 0), 1) / g.dy;
            acc = acc + D(F + FX_XY * n3, 0) / g.dx;
            acc = acc + D(F + FX_ZY * n3, 2) / g.dz;
        } else {
            acc = acc + D(A.aux + (VTI ? 2 * n3 : 0), 2) / g.dz;
            acc = acc + D(F + FX_XZ * n3, 0) / g.dx;
            acc = acc + D(F + FX_YZ * n3, 1) / g.dy;
        }
        float pv = gld(A.prev, c * n3 + p);
        if (A.prev_tapered) pv = taper_apply(pv, A.tz2, A.tx2, A.ty2, ix, iy, iz);
        if (face) {
            // traction SATs of the faces through p (reference face order)
#pragma unroll
            for (int f = 0; f < 6; f++) {
                const int bc = A.bc[f];
                if (bc != 1 && bc != 2) continue;
                const int nax = f < 2 ? 2 : (f < 4 ? 0 : 1);
                const bool hi = f & 1;
                const int len = ax_len(g, nax);
                if (!(hi ? IDX(nax) == len - 1 : IDX(nax) == 0)) continue;
                const int t1 = nax == 2 ? 0 : (nax == 0 ? 1 : 0);
                const int t2 = nax == 2 ? 1 : 2;
                const float sg = hi ? 1.f : -1.f;
                const float hn = ax_h(g, nax);
                const int inward = hi ? -ax_stride(g, nax) : ax_stride(g, nax);
                const float adh = 1.0f / (hn * SAT_H0);
                float term;
                if (c == nax) {
                    const float tang = M.cnorm(nax, t1, p) * (D(curr + t1 * n3, t1) / ax_h(g, t1)) +
                                       M.cnorm(nax, t2, p) * (D(curr + t2 * n3, t2) / ax_h(g, t2)) +
                                       M.cnorm(nax, nax, p) * (D(curr + nax * n3, nax) / hn);
                    term = sg * tang + M.knn(nax, p) * sat_s(curr + nax * n3, p, inward) / hn;
                    if (bc == 2) term -= M.pimp(nax, p) * pv / A.dt;
                } else {
                    const float G = M.gshear(nax, c, p);
                    term = sg * G * (D(curr + nax * n3, c) / ax_h(g, c)) + G * sat_s(curr + c * n3, p, inward) / hn;
                    if (bc == 2) term -= 0.5f * sqrtf(M.mu[p] * M.rho[p]) * pv / A.dt;
                }
                acc -= adh * term;
            }
        }
        float nv = A.dt2 * acc / rho + 2.f * gld(u, p) - pv;
        if (face) nv = nv / e_scale<VTI>(g, A.bc, M, p, ix, iy, iz, c, A.dt);
        gst(A.next, c * n3 + p, taper_apply(nv, A.tz2, A.tx2, A.ty2, ix, iy, iz));
    }
}

// Shell workgroup b (256 points of the edge enumeration, edge_point), displacement
// component c. The slabs away from the edges are tiled by elastic_zslab_tile and
// elastic_side_tile.
template <bool VTI>
__device__ __forceinline__ void elastic_boundary_block(const EArgs& A, int b, int c, int npts) {
    int ix, iy, iz;
    const int t = b * 256 + threadIdx.x;
    if (t >= npts || !edge_point(t, 6, A.g.nx, A.g.ny, A.g.nz, ix, iy, iz)) return;
    if (c == 0) elastic_boundary_point<VTI, 0>(A, ix, iy, iz);
    else if (c == 1) elastic_boundary_point<VTI, 1>(A, ix, iy, iz);
    else elastic_boundary_point<VTI, 2>(A, ix, iy, iz);
}

// ------------------------------------------------------------------------------------
// Pass 2, z slabs: the points within 6 cells of the top or bottom face and at least 6
// cells inside the x and y faces (a third of the shell at 200^3, and its costliest part
// as point-wise work: a wave there covers only a few z columns). A y-marching tile of
// 8 z points (the slab's 6 plus the 2 further points its closure rows read) x TXZ x
// rows: the x-z plane of every operand is staged in LDS with 2-cell x halos, the y
// neighbours live in register queues (as in the interior kernel), so the z closure
// rows, the z-face traction SATs and the interior x / y stencils all read on-chip
// values. Arithmetic and order of operations are those of elastic_boundary_point.
// Tiling: ZSlabTiles (acoustic.hpp; the x tiles start at 6, the y chunks at 6 and stop
// at ny-6).

// LDS of the slab kernels (z-slab and side-slab tiles share it): [planes][36][8] or
// [planes][8][36] floats.
template <bool VTI>
constexpr int slab_lds_floats() { return ((VTI ? 3 : 1) + 3 + (VTI ? 2 : 1) + 4) * 36 * 8; }

template <int TXZ, bool VTI, class LFn>
__device__ __forceinline__ void elastic_zslab_tile(const EArgs& __restrict__ A, const ZSlabTiles Z, const LFn& tile_id,
                                                   float* smem) {
    static_assert(TXZ == 32, "the slab kernels share a [planes][36][8] LDS buffer");
    constexpr int NB = VTI ? 3 : 1;
    constexpr int NT = VTI ? 3 : 1;
    constexpr int PU = 0, PB = 3, PT = 3 + NB, PF = PT + (VTI ? 2 : 1);
    constexpr int NPL = PF + 4;
    constexpr int NQ = 3 + NB + 3;
    float(*s)[TXZ + 4][8] = reinterpret_cast<float(*)[TXZ + 4][8]>(smem);
    const Geom& g = A.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz, sx = g.sx, sy = g.sy, n3 = g.n;
    const float* __restrict__ curr = A.curr;
    const float* __restrict__ prev = A.prev;
    const float* bf[3] = {A.mu, VTI ? A.c66 : A.mu, VTI ? A.c11b : A.mu};
    const float* Ff = A.aux + NT * n3;
    const int L = tile_id();
    const int slab = L & 1, tx = (L >> 1) % Z.ntx, cy = (L >> 1) / Z.ntx;
    const int lz = threadIdx.x & 7, lx = threadIdx.x >> 3;
    const int zb = slab ? nz - 8 : 0;  // z of LDS column 0
    const int iz = zb + lz;
    const int x0 = 6 + tx * TXZ, ix = x0 + lx;
    const int y0 = 6 + cy * Z.ychunk, y1 = min(ny - 6, y0 + Z.ychunk);
    const int col = min(ix, nx - 1) * sx + iz;
    const bool act = ix < nx - 6 && (slab ? lz >= 2 : lz < 6);
    const bool hi = slab != 0;
    const int r = hi ? 7 - lz : lz;  // row counted from the face
    const bool face = r == 0;
    auto ldy = [&](const float* __restrict__ f, int y) { return gas(f)[max(0, min(ny - 1, y)) * sy + col]; };
    // x-halo value of this thread: threads 0..31 load rows x0-2, x0-1, x0+TXZ, x0+TXZ+1
    const int h = threadIdx.x;
    const bool has_h = h < 32;
    const int hr = (h >> 3) & 3;
    const int hrow = hr < 2 ? hr : TXZ + hr;
    const int hsrc = max(0, min(nx - 1, hr < 2 ? x0 - 2 + hr : x0 + TXZ + hr - 2)) * sx + zb + (h & 7);

    const float* qsrc[NQ];
#pragma unroll
    for (int c = 0; c < 3; c++) qsrc[c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) qsrc[3 + k] = bf[k];
    qsrc[3 + NB] = A.aux + (VTI ? n3 : 0);
    qsrc[4 + NB] = Ff + FX_YX * n3;
    qsrc[5 + NB] = Ff + FX_YZ * n3;
    const float* psrc[NPL];
#pragma unroll
    for (int c = 0; c < 3; c++) psrc[PU + c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) psrc[PB + k] = bf[k];
    psrc[PT] = A.aux;
    if (VTI) psrc[PT + 1] = A.aux + 2 * n3;
    psrc[PF + 0] = Ff + FX_ZX * n3;
    psrc[PF + 1] = Ff + FX_XY * n3;
    psrc[PF + 2] = Ff + FX_ZY * n3;
    psrc[PF + 3] = Ff + FX_XZ * n3;

    float q[NQ][6];
#pragma unroll
    for (int k = 0; k < NQ; k++)
#pragma unroll
        for (int o = 0; o < 5; o++) q[k][o] = ldy(qsrc[k], y0 - 2 + o);
    float pc[NPL], hv[NPL];
#pragma unroll
    for (int k = 0; k < NPL; k++) pc[k] = hv[k] = 0.f;
    // threads without a halo value reload their own point (a cache hit), which keeps the
    // loads unconditional and pc / hv in registers
    const int hoff = has_h ? hsrc : col;
    auto prefetch = [&](int y) {
        const int pl = min(y, ny - 1) * sy;
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            if (k >= 3 + NB) pc[k] = gas(psrc[k])[pl + col];
            hv[k] = gas(psrc[k])[pl + hoff];
        }
    };
    prefetch(y0);
    const ECoef<VTI> M{A.lam, A.mu, A.rho, A.c13, A.eps, A.c66};

    for (int iy = y0; iy < y1; iy++) {
#pragma unroll
        for (int k = 0; k < NQ; k++) q[k][5] = ldy(qsrc[k], iy + 3);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            s[k][lx + 2][lz] = (k < 3 + NB) ? q[k][2] : pc[k];
            if (has_h) s[k][hrow][h & 7] = hv[k];
        }
        prefetch(iy + 1);
        __syncthreads();
        if (act) {
            const int p = iy * sy + ix * sx + iz;
            auto SX = [&](int k, int o) { return s[k][lx + 2 + o][lz]; };
            auto FZ = [&](int k, int j) { return s[k][lx + 2][hi ? 7 - j : j]; };  // j-th point from the face
            auto d1x = [&](int k) { return D1C1 * (SX(k, 1) - SX(k, -1)) + D1C2 * (SX(k, 2) - SX(k, -2)); };
            auto d1y = [&](int k) { return D1C1 * (q[k][3] - q[k][1]) + D1C2 * (q[k][4] - q[k][0]); };
            auto d1z = [&](int k) {
                if (r >= 4)
                    return D1C1 * (s[k][lx + 2][lz + 1] - s[k][lx + 2][lz - 1]) +
                           D1C2 * (s[k][lx + 2][lz + 2] - s[k][lx + 2][lz - 2]);
                return d1_closure_at(!hi, r, [&](int j) { return FZ(k, j); });
            };
            auto d2 = [&](int c, int kb, int a) -> float {
                if (a == 0)
                    return d2_interior(SX(PU + c, -2), SX(PU + c, -1), SX(PU + c, 0), SX(PU + c, 1), SX(PU + c, 2),
                                       SX(PB + kb, -2), SX(PB + kb, -1), SX(PB + kb, 0), SX(PB + kb, 1), SX(PB + kb, 2));
                if (a == 1)
                    return d2_interior(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4], q[3 + kb][0], q[3 + kb][1],
                                       q[3 + kb][2], q[3 + kb][3], q[3 + kb][4]);
                return d2_rows(r, [&](int j) { return FZ(PU + c, j); }, [&](int k) { return FZ(PB + kb, k); });
            };
            auto sats = [&](int k) {  // outward S (times h) at the face point
                return c_sat_s[0] * FZ(k, 0) + c_sat_s[1] * FZ(k, 1) + c_sat_s[2] * FZ(k, 2) + c_sat_s[3] * FZ(k, 3);
            };
            const float rho = gld(A.rho, p);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float acc = (d2_factor<VTI>(c, 0) / (g.dx * g.dx)) * d2(c, d2_bfield<VTI>(c, 0), 0);
                acc = acc + (d2_factor<VTI>(c, 1) / (g.dy * g.dy)) * d2(c, d2_bfield<VTI>(c, 1), 1);
                acc = acc + (d2_factor<VTI>(c, 2) / (g.dz * g.dz)) * d2(c, d2_bfield<VTI>(c, 2), 2);
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
                float pv = gld(prev, c * n3 + p);
                if (A.prev_tapered) pv = taper_apply(pv, A.tz2, A.tx2, A.ty2, ix, iy, iz);
                const int bc = hi ? A.bc[1] : A.bc[0];  // constant indices (no stack copy of A)
                if (face && (bc == 1 || bc == 2)) {
                    // traction SAT of the top / bottom face (elastic_boundary_point, normal axis z)
                    const float sg = hi ? 1.f : -1.f;
                    const float adh = 1.0f / (g.dz * SAT_H0);
                    float term;
                    if (c == 2) {
                        const float tang = M.cnorm(2, 0, p) * (d1x(PU + 0) / g.dx) + M.cnorm(2, 1, p) * (d1y(1) / g.dy) +
                                           M.cnorm(2, 2, p) * (d1z(PU + 2) / g.dz);
                        term = sg * tang + M.knn(2, p) * sats(PU + 2) / g.dz;
                        if (bc == 2) term -= M.pimp(2, p) * pv / A.dt;
                    } else {
                        const float G = M.gshear(2, c, p);
                        const float duz = c == 0 ? d1x(PU + 2) / g.dx : d1y(2) / g.dy;
                        term = sg * G * duz + G * sats(PU + c) / g.dz;
                        if (bc == 2) term -= 0.5f * sqrtf(M.mu[p] * M.rho[p]) * pv / A.dt;
                    }
                    acc -= adh * term;
                }
                float nv = A.dt2 * acc / rho + 2.f * q[c][2] - pv;
                if (face && bc == 2) {
                    // e_scale of a z-face point away from the x and y faces: only this face
                    // absorbs (normal component: the P speed, tangential: the S speed)
                    const float v = c == 2 ? sqrtf((M.lam[p] + 2.f * M.mu[p]) / M.rho[p]) : sqrtf(M.mu[p] / M.rho[p]);
                    nv = nv / (1.f + (0.f + v * A.dt / (2.f * g.dz * SAT_H0)));
                }
                gst(A.next, c * n3 + p, taper_apply(nv, A.tz2, A.tx2, A.ty2, ix, iy, iz));
            }
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            q[k][0] = q[k][1]; q[k][1] = q[k][2]; q[k][2] = q[k][3]; q[k][3] = q[k][4]; q[k][4] = q[k][5];
        }
    }
}

// Source injection, all three displacement components (blockIdx.y):
// next += taper(dt^2 inj / rho / scale); mode 1: initial prev = 0.5 dt^2 inj(src[0]) / rho.
// ------------------------------------------------------------------------------------
// Pass 2, side slabs: the points within 6 cells of an x face (AX = 0) or a y face
// (AX = 1) and at least 6 cells inside the faces of the two other axes. A tile is 32 z
// lanes x the 8 points nearest the face along its normal axis (the slab's 6 rows plus
// the 2 further rows its closure rows read; a half-wave per row), marching along the
// remaining in-plane axis (y for x slabs, x for y slabs) with register queues: the
// normal-axis closure rows, the face's traction SAT and the z stencils read the staged
// x-z or y-z planes in LDS. Arithmetic and order of operations are those of
// elastic_boundary_point. Tiling (ZSlabTiles): z tiles of 32 from z = 6, march chunks of
// `schunk` from 6 to n - 6; tile L: bit 0 the low or high slab, then z tile, then chunk.
template <int AX, bool VTI>
__device__ __forceinline__ void elastic_side_tile(const EArgs& __restrict__ A, const ZSlabTiles Z, int L, float* smem) {
    constexpr int NB = VTI ? 3 : 1;
    constexpr int NT = VTI ? 3 : 1;
    constexpr int MX = 1 - AX;  // marching axis
    constexpr int PU = 0, PB = 3, PT = 3 + NB, PF = PT + (VTI ? 2 : 1);
    constexpr int NPL = PF + 4;
    constexpr int NQ = 3 + NB + 3;
    float(*s)[8][36] = reinterpret_cast<float(*)[8][36]>(smem);
    const Geom& g = A.g;
    const int nz = g.nz, n3 = g.n;
    const int sA = AX == 0 ? g.sx : g.sy, sM = AX == 0 ? g.sy : g.sx;
    const int nA = AX == 0 ? g.nx : g.ny, nM = AX == 0 ? g.ny : g.nx;
    const float hA = AX == 0 ? g.dx : g.dy, hM = AX == 0 ? g.dy : g.dx;
    const float* __restrict__ curr = A.curr;
    const float* __restrict__ prev = A.prev;
    const float* bf[3] = {A.mu, VTI ? A.c66 : A.mu, VTI ? A.c11b : A.mu};
    const float* Ff = A.aux + NT * n3;
    const int slab = L & 1, tz = (L >> 1) % Z.sntz, ch = (L >> 1) / Z.sntz;
    const bool hi = slab != 0;
    const int lz = threadIdx.x & 31, lr = threadIdx.x >> 5;  // z lane, LDS row
    const int z0 = 6 + tz * 32, iz = z0 + lz;
    const int base = hi ? nA - 8 : 0;  // normal-axis index of LDS row 0
    const int ia = base + lr;
    const int r = hi ? 7 - lr : lr;  // row counted from the face
    const bool act = iz < nz - 6 && r < 6;
    const bool face = r == 0;
    const int m0 = 6 + ch * Z.schunk, m1 = min(nM - 6, m0 + Z.schunk);
    const int col = ia * sA + min(iz, nz - 1);
    auto ldm = [&](const float* __restrict__ f, int m) { return gas(f)[max(0, min(nM - 1, m)) * sM + col]; };
    // z-halo value of this thread: threads 0..31 load rows 0..7 at z0-2, z0-1, z0+32, z0+33
    const int h = threadIdx.x;
    const bool has_h = h < 32;
    const int hrow = (h >> 2) & 7, hz4 = h & 3;
    const int hcol = hz4 < 2 ? hz4 : 32 + hz4;
    const int hsrc = (base + hrow) * sA + max(0, min(nz - 1, hz4 < 2 ? z0 - 2 + hz4 : z0 + 32 + hz4 - 2));
    // threads without a halo value reload their own point (a cache hit), which keeps the
    // loads unconditional
    const int hoff = has_h ? hsrc : col;

    // queues (derivatives along the march axis): u, b-fields, then the normal stress and
    // the two fluxes whose outer derivative runs along it
    const float* qsrc[NQ];
#pragma unroll
    for (int c = 0; c < 3; c++) qsrc[c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) qsrc[3 + k] = bf[k];
    // planes (derivatives along the normal axis and z): u, b-fields, the normal stress
    // of the normal axis [and T_z], then F_zx, the flux with outer derivative along the
    // normal axis for component x or y, F_zy, the one for component z
    const float* psrc[NPL];
#pragma unroll
    for (int c = 0; c < 3; c++) psrc[PU + c] = curr + c * n3;
#pragma unroll
    for (int k = 0; k < NB; k++) psrc[PB + k] = bf[k];
    if (AX == 0) {
        qsrc[3 + NB] = A.aux + (VTI ? n3 : 0);  // T_y
        qsrc[4 + NB] = Ff + FX_YX * n3;
        qsrc[5 + NB] = Ff + FX_YZ * n3;
        psrc[PT] = A.aux;                        // T_x
        psrc[PF + 1] = Ff + FX_XY * n3;
        psrc[PF + 3] = Ff + FX_XZ * n3;
    } else {
        qsrc[3 + NB] = A.aux;                    // T_x
        qsrc[4 + NB] = Ff + FX_XY * n3;
        qsrc[5 + NB] = Ff + FX_XZ * n3;
        psrc[PT] = A.aux + (VTI ? n3 : 0);       // T_y
        psrc[PF + 1] = Ff + FX_YX * n3;
        psrc[PF + 3] = Ff + FX_YZ * n3;
    }
    if (VTI) psrc[PT + 1] = A.aux + 2 * n3;      // T_z
    psrc[PF + 0] = Ff + FX_ZX * n3;
    psrc[PF + 2] = Ff + FX_ZY * n3;

    float q[NQ][6];
#pragma unroll
    for (int k = 0; k < NQ; k++)
#pragma unroll
        for (int o = 0; o < 5; o++) q[k][o] = ldm(qsrc[k], m0 - 2 + o);
    float pc[NPL], hv[NPL];
#pragma unroll
    for (int k = 0; k < NPL; k++) pc[k] = hv[k] = 0.f;
    auto prefetch = [&](int m) {
        const int pl = min(m, nM - 1) * sM;
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            if (k >= 3 + NB) pc[k] = gas(psrc[k])[pl + col];
            hv[k] = gas(psrc[k])[pl + hoff];
        }
    };
    prefetch(m0);
    const ECoef<VTI> M{A.lam, A.mu, A.rho, A.c13, A.eps, A.c66};
    const int bcf = AX == 0 ? (hi ? A.bc[3] : A.bc[2]) : (hi ? A.bc[5] : A.bc[4]);  // constant indices

    for (int im = m0; im < m1; im++) {
#pragma unroll
        for (int k = 0; k < NQ; k++) q[k][5] = ldm(qsrc[k], im + 3);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPL; k++) {
            s[k][lr][lz + 2] = (k < 3 + NB) ? q[k][2] : pc[k];
            if (has_h) s[k][hrow][hcol] = hv[k];
        }
        prefetch(im + 1);
        __syncthreads();
        if (act) {
            const int p = im * sM + col;
            const int ix = AX == 0 ? ia : im, iy = AX == 0 ? im : ia;
            auto FN = [&](int k, int j) { return s[k][hi ? 7 - j : j][lz + 2]; };  // j-th point from the face
            auto dA = [&](int k) {  // along the normal axis (times h)
                if (r >= 4)
                    return D1C1 * (s[k][lr + 1][lz + 2] - s[k][lr - 1][lz + 2]) +
                           D1C2 * (s[k][lr + 2][lz + 2] - s[k][lr - 2][lz + 2]);
                return d1_closure_at(!hi, r, [&](int j) { return FN(k, j); });
            };
            auto dZ = [&](int k) {
                return D1C1 * (s[k][lr][lz + 3] - s[k][lr][lz + 1]) + D1C2 * (s[k][lr][lz + 4] - s[k][lr][lz]);
            };
            auto dM = [&](int k) { return D1C1 * (q[k][3] - q[k][1]) + D1C2 * (q[k][4] - q[k][0]); };
            auto d2 = [&](int c, int kb, int a) -> float {
                if (a == AX) return d2_rows(r, [&](int j) { return FN(PU + c, j); }, [&](int k) { return FN(PB + kb, k); });
                if (a == MX)
                    return d2_interior(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4], q[3 + kb][0], q[3 + kb][1],
                                       q[3 + kb][2], q[3 + kb][3], q[3 + kb][4]);
                return d2_interior(s[PU + c][lr][lz], s[PU + c][lr][lz + 1], s[PU + c][lr][lz + 2], s[PU + c][lr][lz + 3],
                                   s[PU + c][lr][lz + 4], s[PB + kb][lr][lz], s[PB + kb][lr][lz + 1],
                                   s[PB + kb][lr][lz + 2], s[PB + kb][lr][lz + 3], s[PB + kb][lr][lz + 4]);
            };
            auto sats = [&](int k) {  // outward S (times h) at the face point
                return c_sat_s[0] * FN(k, 0) + c_sat_s[1] * FN(k, 1) + c_sat_s[2] * FN(k, 2) + c_sat_s[3] * FN(k, 3);
            };
            const float rho = gld(A.rho, p);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float acc = (d2_factor<VTI>(c, 0) / (g.dx * g.dx)) * d2(c, d2_bfield<VTI>(c, 0), 0);
                acc = acc + (d2_factor<VTI>(c, 1) / (g.dy * g.dy)) * d2(c, d2_bfield<VTI>(c, 1), 1);
                acc = acc + (d2_factor<VTI>(c, 2) / (g.dz * g.dz)) * d2(c, d2_bfield<VTI>(c, 2), 2);
                if (AX == 0) {
                    if (c == 0) {
                        acc = acc + dA(PT) / g.dx;          // D_x T_x
                        acc = acc + dM(4 + NB) / g.dy;      // D_y F_yx
                        acc = acc + dZ(PF + 0) / g.dy;      // D_z F_zx (spacing dy as in the reference)
                    } else if (c == 1) {
                        acc = acc + dM(3 + NB) / g.dy;      // D_y T_y
                        acc = acc + dA(PF + 1) / g.dx;      // D_x F_xy
                        acc = acc + dZ(PF + 2) / g.dz;      // D_z F_zy
                    } else {
                        acc = acc + dZ(VTI ? PT + 1 : PT) / g.dz;
                        acc = acc + dA(PF + 3) / g.dx;      // D_x F_xz
                        acc = acc + dM(5 + NB) / g.dy;      // D_y F_yz
                    }
                } else {
                    if (c == 0) {
                        acc = acc + dM(3 + NB) / g.dx;      // D_x T_x
                        acc = acc + dA(PF + 1) / g.dy;      // D_y F_yx
                        acc = acc + dZ(PF + 0) / g.dy;      // D_z F_zx (spacing dy as in the reference)
                    } else if (c == 1) {
                        acc = acc + dA(PT) / g.dy;          // D_y T_y
                        acc = acc + dM(4 + NB) / g.dx;      // D_x F_xy
                        acc = acc + dZ(PF + 2) / g.dz;      // D_z F_zy
                    } else {
                        acc = acc + dZ(VTI ? PT + 1 : PT) / g.dz;
                        acc = acc + dM(5 + NB) / g.dx;      // D_x F_xz
                        acc = acc + dA(PF + 3) / g.dy;      // D_y F_yz
                    }
                }
                float pv = gld(prev, c * n3 + p);
                if (A.prev_tapered) pv = taper_apply(pv, A.tz2, A.tx2, A.ty2, ix, iy, iz);
                if (face && (bcf == 1 || bcf == 2)) {
                    // traction SAT of this face (elastic_boundary_point, normal axis AX;
                    // tangential axes: the march axis and z)
                    const float sg = hi ? 1.f : -1.f;
                    const float adh = 1.0f / (hA * SAT_H0);
                    float term;
                    if (c == AX) {
                        const float tang = M.cnorm(AX, MX, p) * (dM(MX) / hM) + M.cnorm(AX, 2, p) * (dZ(PU + 2) / g.dz) +
                                           M.cnorm(AX, AX, p) * (dA(PU + AX) / hA);
                        term = sg * tang + M.knn(AX, p) * sats(PU + AX) / hA;
                        if (bcf == 2) term -= M.pimp(AX, p) * pv / A.dt;
                    } else {
                        const float G = M.gshear(AX, c, p);
                        const float du = c == 2 ? dZ(PU + AX) / g.dz : dM(AX) / hM;  // D_c u_AX
                        term = sg * G * du + G * sats(PU + c) / hA;
                        if (bcf == 2) term -= 0.5f * sqrtf(M.mu[p] * M.rho[p]) * pv / A.dt;
                    }
                    acc -= adh * term;
                }
                float nv = A.dt2 * acc / rho + 2.f * q[c][2] - pv;
                if (face && bcf == 2) {
                    // e_scale of an x- or y-face point away from the other faces: only this
                    // face absorbs (normal component: the P speed, (1 + 2 eps)-scaled for
                    // VTI; tangential: the S speed)
                    const float F = VTI ? 1.f + 2.f * M.eps[p] : 1.f;
                    const float v = c == AX ? sqrtf(F * (M.lam[p] + 2.f * M.mu[p]) / M.rho[p]) : sqrtf(M.mu[p] / M.rho[p]);
                    nv = nv / (1.f + (0.f + v * A.dt / (2.f * hA * SAT_H0)));
                }
                gst(A.next, c * n3 + p, taper_apply(nv, A.tz2, A.tx2, A.ty2, ix, iy, iz));
            }
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            q[k][0] = q[k][1]; q[k][1] = q[k][2]; q[k][2] = q[k][3]; q[k][3] = q[k][4]; q[k][4] = q[k][5];
        }
    }
}

template <bool VTI>
__device__ __forceinline__ void elastic_inject_one(float* __restrict__ out3, const EArgs& A, const InjTab& I,
                                                   const float* __restrict__ src, int nt, int it, int mode, int j, int comp) {
    float* __restrict__ out = out3 + comp * A.g.n;
    const Geom& g = A.g;
    const float dt = A.dt;
    const int e0 = I.tbeg[j], e1 = I.tbeg[j + 1];
    int r, a, bb, c, de, ix, iy, iz;
    inj_decode(I, I.tlst[e0], r, a, bb, c, de);
    inj_point(I, r, a, bb, c, de, ix, iy, iz);
    const int p = g.at(ix, iy, iz);
    float val = out[p];
    for (int e = e0; e < e1; e++) {
        inj_decode(I, I.tlst[e], r, a, bb, c, de);
        const float v = inj_value(I, src, nt, it, comp, r, a, bb, c, de);
        float add;
        if (mode == 1) {
            add = v * (0.5f * dt * dt / A.rho[p]);
        } else {
            const ECoef<VTI> M{A.lam, A.mu, A.rho, A.c13, A.eps, A.c66};
            add = dt * dt * v / A.rho[p];
            add = add / e_scale<VTI>(g, A.bc, M, p, ix, iy, iz, comp, dt);
            add = taper_apply(add, A.tz2, A.tx2, A.ty2, ix, iy, iz);
        }
        val += add;
    }
    out[p] = val;
}
template <bool VTI>
__global__ void elastic_inject(float* __restrict__ out3, EArgs A, InjTab I, const float* __restrict__ src, int nt, int it,
                               int mode) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < I.ntgt) elastic_inject_one<VTI>(out3, A, I, src, nt, it, mode, j, blockIdx.y);
}

// Imaging condition of nl_we_op_e / nl_we_op_vti ::compute_gradients
// (we_op.cpp:488-565, 1108-1216) with FP64 accumulation. `lag` is the adjoint field
// of the previous step (the reference uses the derivatives computed one step earlier);
// `adj` the current adjoint field (rho term). Gradients in the unpadded layout.
// Workgroups of 64 (z) x GTX (x) points of one y plane: the forward-field codes and the
// lagged adjoint field of the tile and its 2-cell x / z halos are staged in LDS once, so
// the interior x and z derivatives read on-chip values (closure rows and the y
// derivatives read memory, as d1_acc / d1_point do; same arithmetic).
constexpr int GTX = 4;  // 200^3 event: 472 us (8 x rows: 514, 16: 570; fewer workgroups hide less latency)
template <bool VTI, int FMT>
__global__ __launch_bounds__(64 * GTX) void elastic_grad(SnapView<FMT> fm, SnapView<FMT> fn, SnapView<FMT> fp,
                                                    const float* __restrict__ adj, const float* __restrict__ lag, EArgs A,
                                                    float sdt, int branch, double* __restrict__ gl, double* __restrict__ gm,
                                                    double* __restrict__ gr) {
    __shared__ float sf[3][GTX + 4][68], sl[3][GTX + 4][68];  // [component][x row from x0-2][z column from z0-2]
    const Geom g = A.g;
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = blockIdx.x * 64, x0 = blockIdx.y * GTX, iy = blockIdx.z;
    const int iz = z0 + lz, ix = x0 + lx;
    const int n3 = g.n;
    const int pl = iy * g.sy;
    const int p = pl + min(ix, g.nx - 1) * g.sx + min(iz, g.nz - 1);
    HaloMap<GTX> H;
    H.init(lx, lz, x0, z0, g.nx, g.nz, g.sx);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        sf[c][lx + 2][lz + 2] = fn.code(c * n3 + p);
        sl[c][lx + 2][lz + 2] = lag[c * n3 + p];
        if (H.has_x) {
            sf[c][H.xrow][H.xcol] = fn.code(c * n3 + pl + H.xsrc);
            sl[c][H.xrow][H.xcol] = lag[c * n3 + pl + H.xsrc];
        }
        if (H.has_z) {
            sf[c][H.zrow][H.zcol] = fn.code(c * n3 + pl + H.zsrc);
            sl[c][H.zrow][H.zcol] = lag[c * n3 + pl + H.zsrc];
        }
    }
    __syncthreads();
    if (iz >= g.nz || ix >= g.nx) return;
    const bool inx = ix >= 4 && ix < g.nx - 4, inz = iz >= 4 && iz < g.nz - 4;
    float F[3][3], Aj[3][3];  // [deriv axis][component]
    const float sn = fn.scale();
#pragma unroll
    for (int c = 0; c < 3; c++) {
        auto FN = [&](int i) { return fn.code(c * n3 + i); };
        const float fx = inx ? D1C1 * (sf[c][lx + 3][lz + 2] - sf[c][lx + 1][lz + 2]) +
                                   D1C2 * (sf[c][lx + 4][lz + 2] - sf[c][lx][lz + 2])
                             : d1_acc(FN, p, ix, g.nx, g.sx);
        const float fz = inz ? D1C1 * (sf[c][lx + 2][lz + 3] - sf[c][lx + 2][lz + 1]) +
                                   D1C2 * (sf[c][lx + 2][lz + 4] - sf[c][lx + 2][lz])
                             : d1_acc(FN, p, iz, g.nz, 1);
        const float ax = inx ? D1C1 * (sl[c][lx + 3][lz + 2] - sl[c][lx + 1][lz + 2]) +
                                   D1C2 * (sl[c][lx + 4][lz + 2] - sl[c][lx][lz + 2])
                             : d1_point(lag + c * n3, p, ix, g.nx, g.sx);
        const float az = inz ? D1C1 * (sl[c][lx + 2][lz + 3] - sl[c][lx + 2][lz + 1]) +
                                   D1C2 * (sl[c][lx + 2][lz + 4] - sl[c][lx + 2][lz])
                             : d1_point(lag + c * n3, p, iz, g.nz, 1);
        F[0][c] = fx * sn / g.dx;
        F[1][c] = d1_acc(FN, p, iy, g.ny, g.sy) * sn / g.dy;
        F[2][c] = fz * sn / g.dz;
        Aj[0][c] = ax / g.dx;
        Aj[1][c] = d1_point(lag + c * n3, p, iy, g.ny, g.sy) / g.dy;
        Aj[2][c] = az / g.dz;
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
        const float L = A.lam[p], U = A.mu[p], E = A.eps[p], C13 = A.c13[p], C66 = A.c66[p];
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
    const float sm = fm.scale(), sp = fp.p ? fp.scale() : 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int qc = c * n3 + p;
        const float fpv = fp.p ? fp.code(qc) * sp : 0.f;
        const float fnv = fn.code(qc) * sn, fmv = fm.code(qc) * sm;
        const float d2t = branch == 1 ? fpv - 2.f * fnv + fmv : (branch == 0 ? fpv - fnv : -fnv + fmv);
        rt += adj[qc] * d2t;
    }
    const long long o = ((long long)iy * g.nx + ix) * g.nz + iz;
    gl[o] += (double)gL;
    gm[o] += (double)gM;
    gr[o] += (double)(1.0f / sdt * rt);
}

}  // namespace fwi3d
