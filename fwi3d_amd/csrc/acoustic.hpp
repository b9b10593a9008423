// Acoustic variable-density kernels (gfx950).
//
// One time step of nl_we_op_a::propagate (src/cpp/we_op.cpp:1662-1791) is two
// launches that write disjoint points of `next` and only read curr/prev/model:
//   acoustic_interior  points 6+ cells inside every face: D2(1/rho) along x, y, z
//                      from an LDS x-z plane and a register queue along y, the
//                      leapfrog update and the cos^2 taper;
//   acoustic_boundary  the 6-cell shell: SBP closure rows, the face SATs, the update,
//                      the absorbing-boundary scaling and the taper.
// Source injection is a separate tiny launch that adds its already-updated
// contribution (the update is linear in the injected term); receiver extraction
// reads curr, which the step does not modify.
//
// Taper bookkeeping: the reference tapers curr and next at the end of every step
// (we_op.cpp:1780-1791), so a field is tapered once while it is `curr` and a second
// time before it becomes `prev`. Here the stored field carries the first taper and
// the second is applied when the field is read as prev (not at the first step,
// whose prev is the untapered initial injection).
#pragma once
#include "common.hpp"

namespace fwi3d {

struct AcuArgs {
    const float* curr;
    const float* prev;
    float* next;
    const float* K;  // bulk modulus
    const float* b;  // 1/rho
    const float* tz2;
    const float* tx2;
    const float* ty2;
    Geom g;
    int bc[6];
    float dt, dt2, alpha;  // alpha = 0.2508560249 / free_surface_stiffness
    int prev_tapered;
    // y-march tiling: tiles of 64 (z) x TX (x) x ychunk (y) from the corner (org, org, org);
    // the y march of the last chunk stops at yend
    int ychunk, ntz, ntx, nyc, org, yend;
};

// One workgroup of the interior y-march: tile tile_id() of the tiling in A (the
// single-shot kernel and the batched kernel share this body).
template <int TX, class LFn>
__device__ __forceinline__ void acoustic_interior_tile(const AcuArgs A, const LFn& tile_id) {
    constexpr int TZ = 64;
    __shared__ float su[TX + 4][TZ + 4];
    __shared__ float sb[TX + 4][TZ + 4];
    const GAS float* __restrict__ curr = gas(A.curr);
    const GAS float* __restrict__ prev = gas(A.prev);
    GAS float* __restrict__ next = gas(A.next);
    const GAS float* __restrict__ K = gas(A.K);
    const GAS float* __restrict__ bb = gas(A.b);
    const float* __restrict__ tz2 = A.tz2;
    const float* __restrict__ tx2 = A.tx2;
    const float* __restrict__ ty2 = A.ty2;
    const int nx = A.g.nx, ny = A.g.ny, nz = A.g.nz, sx = A.g.sx, sy = A.g.sy;
    const float idx2 = 1.0f / (A.g.dx * A.g.dx), idy2 = 1.0f / (A.g.dy * A.g.dy), idz2 = 1.0f / (A.g.dz * A.g.dz);
    const float dt2 = A.dt2;
    const bool ptap = A.prev_tapered != 0;

    const int L = tile_id();
    const int tz = L % A.ntz;
    const int tx = (L / A.ntz) % A.ntx;
    const int cy = L / (A.ntz * A.ntx);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = A.org + tz * TZ, x0 = A.org + tx * TX;
    const int iz = z0 + lz, ix = x0 + lx;
    const int y0 = A.org + cy * A.ychunk;
    const int y1 = min(A.yend, y0 + A.ychunk);
    // every load uses clamped coordinates: values outside the grid are never used
    // by an interior point, so no load needs a guard
    const int cz = min(iz, nz - 1), cx = min(ix, nx - 1);
    const int col = cx * sx + cz;
    const bool comp_xz = ix >= 6 && ix < nx - 6 && iz >= 6 && iz < nz - 6;
    auto ldy = [&](const GAS float* __restrict__ f, int y) { return f[max(0, min(ny - 1, y)) * sy + col]; };
    HaloMap<TX> H;
    H.init(lx, lz, x0, z0, nx, nz, sx);

    // y queues y-2..y+2 plus the plane y+3, loaded one iteration ahead of its use
    float um2 = ldy(curr, y0 - 2), um1 = ldy(curr, y0 - 1), u0 = ldy(curr, y0), up1 = ldy(curr, y0 + 1), up2 = ldy(curr, y0 + 2), up3;
    float bm2 = ldy(bb, y0 - 2), bm1 = ldy(bb, y0 - 1), b0 = ldy(bb, y0), bp1 = ldy(bb, y0 + 1), bp2 = ldy(bb, y0 + 2), bp3;
    // halo values of the plane to stage next, loaded one iteration ahead
    float hxu = 0.f, hxb = 0.f, hzu = 0.f, hzb = 0.f;
    auto prefetch = [&](int y) {
        const int pl = min(y, ny - 1) * sy;
        if (H.has_x) { hxu = curr[pl + H.xsrc]; hxb = bb[pl + H.xsrc]; }
        if (H.has_z) { hzu = curr[pl + H.zsrc]; hzb = bb[pl + H.zsrc]; }
    };
    prefetch(y0);

    for (int iy = y0; iy < y1; iy++) {
        up3 = ldy(curr, iy + 3);
        bp3 = ldy(bb, iy + 3);
        const int pl = iy * sy;
        __syncthreads();
        su[lx + 2][lz + 2] = u0;
        sb[lx + 2][lz + 2] = b0;
        if (H.has_x) { su[H.xrow][H.xcol] = hxu; sb[H.xrow][H.xcol] = hxb; }
        if (H.has_z) { su[H.zrow][H.zcol] = hzu; sb[H.zrow][H.zcol] = hzb; }
        prefetch(iy + 1);
        __syncthreads();
        if (comp_xz && iy >= 6 && iy < ny - 6) {
            const int p = pl + col;
            const float sx_ = d2_interior(su[lx][lz + 2], su[lx + 1][lz + 2], u0, su[lx + 3][lz + 2], su[lx + 4][lz + 2],
                                          sb[lx][lz + 2], sb[lx + 1][lz + 2], b0, sb[lx + 3][lz + 2], sb[lx + 4][lz + 2]);
            const float sy_ = d2_interior(um2, um1, u0, up1, up2, bm2, bm1, b0, bp1, bp2);
            const float sz_ = d2_interior(su[lx + 2][lz], su[lx + 2][lz + 1], u0, su[lx + 2][lz + 3], su[lx + 2][lz + 4],
                                          sb[lx + 2][lz], sb[lx + 2][lz + 1], b0, sb[lx + 2][lz + 3], sb[lx + 2][lz + 4]);
            float lap = idx2 * sx_;
            lap = lap + idy2 * sy_;
            lap = lap + idz2 * sz_;
            float pv = prev[p];
            if (ptap) pv = taper_apply(pv, tz2, tx2, ty2, ix, iy, iz);
            const float nv = dt2 * lap * K[p] + 2.f * u0 - pv;
            next[p] = taper_apply(nv, tz2, tx2, ty2, ix, iy, iz);
        }
        um2 = um1; um1 = u0; u0 = up1; up1 = up2; up2 = up3;
        bm2 = bm1; bm1 = b0; b0 = bp1; bp1 = bp2; bp2 = bp3;
    }
}

// Divisor of the absorbing SAT's implicit part at a boundary point (1 elsewhere).
__device__ __forceinline__ float acu_scale(const Geom& g, const int* bc, float Kp, float bp, int ix, int iy, int iz, float dt) {
    float sc = 0.f;
    bool any = false;
    const float c = sqrtf(Kp * bp);
    if (iz == 0 && bc[0] == 2) { sc += c * dt / (2.f * g.dz * SAT_H0); any = true; }
    if (iz == g.nz - 1 && bc[1] == 2) { sc += c * dt / (2.f * g.dz * SAT_H0); any = true; }
    if (ix == 0 && bc[2] == 2) { sc += c * dt / (2.f * g.dx * SAT_H0); any = true; }
    if (ix == g.nx - 1 && bc[3] == 2) { sc += c * dt / (2.f * g.dx * SAT_H0); any = true; }
    if (iy == 0 && bc[4] == 2) { sc += c * dt / (2.f * g.dy * SAT_H0); any = true; }
    if (iy == g.ny - 1 && bc[5] == 2) { sc += c * dt / (2.f * g.dy * SAT_H0); any = true; }
    return any ? 1.f + sc : 1.f;
}

// One boundary-shell point: closures, SATs (faces in the reference's order top,
// bottom, left, right, front, back), update, scaling, taper. Points of the shell
// enumeration from npts on are skipped (the acoustic step passes the whole shell).
__device__ __forceinline__ void acoustic_boundary_block(const AcuArgs& A, int b, int npts) {
    const Geom& g = A.g;
    int ix, iy, iz;
    const int t = b * 256 + threadIdx.y * blockDim.x + threadIdx.x;  // 256 x 1 or 64 x 4 threads
    if (t >= npts || !shell_point(t, 6, g.nx, g.ny, g.nz, ix, iy, iz)) return;
    const float* __restrict__ curr_ = A.curr;
    const GAS float* __restrict__ curr = gas(A.curr);
    const GAS float* __restrict__ bb = gas(A.b);
    const GAS float* __restrict__ K = gas(A.K);
    const int p = g.at(ix, iy, iz);
    float lap = (1.0f / (g.dx * g.dx)) * d2_point(curr_, A.b, p, ix, g.nx, g.sx);
    lap = lap + (1.0f / (g.dy * g.dy)) * d2_point(curr_, A.b, p, iy, g.ny, g.sy);
    lap = lap + (1.0f / (g.dz * g.dz)) * d2_point(curr_, A.b, p, iz, g.nz, 1);
    float pv = gld(A.prev, p);
    if (A.prev_tapered) pv = taper_apply(pv, A.tz2, A.tx2, A.ty2, ix, iy, iz);
    float sat = 0.f;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int bc = A.bc[f];
        if (bc == 0) continue;
        const int ax = f >> 1;  // 0: z, 1: x, 2: y
        const bool hi = f & 1;
        const int i = ax == 0 ? iz : (ax == 1 ? ix : iy);
        const int n = ax == 0 ? g.nz : (ax == 1 ? g.nx : g.ny);
        const int s = ax == 0 ? 1 : (ax == 1 ? g.sx : g.sy);
        const float h = ax == 0 ? g.dz : (ax == 1 ? g.dx : g.dy);
        const int k = hi ? n - 1 - i : i;
        if (k >= 4) continue;
        const int inward = hi ? -s : s;
        if (bc == 1) {
            const int pf = p - k * inward;
            const float v = bb[pf] * curr[pf];
            const float d2 = h * h;
            sat += 1.0f / (c_hnorm[k] * d2) * c_sat_s[k] * v;
            if (k == 0) sat -= v / (A.alpha * SAT_H0 * d2);
        } else if (k == 0) {
            const float sd = sat_s(curr_, p, inward);
            const float imp = (bc == 2) ? 0.5f / sqrtf(K[p] / bb[p]) : 0.f;
            const float adh = 1.0f / (SAT_H0 * h);
            sat -= adh * (bb[p] * sd / h - imp * pv / A.dt);
        }
    }
    lap += sat;
    float nv = A.dt2 * lap * K[p] + 2.f * curr[p] - pv;
    nv = nv / acu_scale(g, A.bc, K[p], bb[p], ix, iy, iz, A.dt);
    gst(A.next, p, taper_apply(nv, A.tz2, A.tx2, A.ty2, ix, iy, iz));
}

// Tiling of the slab kernel of the elastic / VTI shell (elastic.hpp): the y-marching
// z-slab tiles, then the side-slab tiles of the x and y slabs.
struct ZSlabTiles {
    int ntx, nyc, ychunk;          // z slabs: x tiles from 6, y chunks from 6
    int nzs;                       // z-slab workgroups (2 ntx nyc); the side slabs' follow
    int sntz, snc0, snc1, schunk;  // side slabs: 32-point z tiles from 6; march chunks of
                                   // schunk from 6 (x slabs along y, y slabs along x)
};

// ------------------------------------------------------------------------------------
// Point sources / receivers. Tables: ind [npts][npi], w [npts][2][nw] (row 0 inject,
// row 1 extract); kinds 0 (delta_m3, npi 3, nw 3), 1 (ddelta_m3, npi 3, nw 6) and 2
// (dipole_m3 DAS channels, npi 6, nw 6: entries 0-2 the end at -gl/2 along the fiber,
// 3-5 the end at +gl/2). Injection runs one thread per distinct grid point touched by
// the stencils: tbeg[ntgt+1] / tlst[nper npts] list, per target, the contributions in
// increasing t, i.e. in the order of the reference's serial loops (point, then y, x, z,
// and for a dipole end 1 before end 2: injector.cpp:106-115, 423-459), so overlapping
// stencils sum in a fixed order and every run gives the same bits:
//   kinds 0/1: t = 27 r + 9 a + 3 b + c;  kind 2: t = 54 r + 2 (9 a + 3 b + c) + e.
struct InjTab {
    const int* xi;
    const int* yi;
    const int* zi;
    const float* xw;
    const float* yw;
    const float* zw;
    const int* tbeg;
    const int* tlst;
    int npts, kind, nw, ntgt;
};

// Contribution t of the injection lists: point r, neighbour (a, b, c) along (y, x, z),
// dipole end e (0 for kinds 0/1).
__device__ __forceinline__ void inj_decode(const InjTab& I, int t, int& r, int& a, int& b, int& c, int& e) {
    int q;
    if (I.kind == 2) {
        r = t / 54;
        const int u = t - 54 * r;
        e = u & 1;
        q = u >> 1;
    } else {
        r = t / 27;
        e = 0;
        q = t - 27 * r;
    }
    a = q / 9;
    b = (q / 3) % 3;
    c = q % 3;
}
__device__ __forceinline__ void inj_point(const InjTab& I, int r, int a, int b, int c, int e, int& ix, int& iy, int& iz) {
    const int base = I.kind == 2 ? 6 * r + 3 * e : 3 * r;
    ix = I.xi[base + b];
    iy = I.yi[base + a];
    iz = I.zi[base + c];
}

// Injected value of displacement/pressure component `comp` at neighbour (a,b,c) of point
// r (of dipole end e: -w at end 1, +w at end 2).
__device__ __forceinline__ float inj_value(const InjTab& I, const float* __restrict__ src, int nt, int it, int comp, int r,
                                           int a, int bb, int c, int e = 0) {
    const float* xw = I.xw + r * 2 * I.nw;
    const float* yw = I.yw + r * 2 * I.nw;
    const float* zw = I.zw + r * 2 * I.nw;
    if (I.kind == 2) {
        const float v = xw[3 * e + bb] * yw[3 * e + a] * zw[3 * e + c] * src[(long long)(comp * I.npts + r) * nt + it];
        return e ? v : -v;
    }
    if (I.kind == 0) return yw[a] * xw[bb] * zw[c] * src[(long long)(comp * I.npts + r) * nt + it];
    int icx, icy, icz;
    if (comp == 0) { icx = 0; icy = 3; icz = 4; }
    else if (comp == 1) { icx = 3; icy = 1; icz = 5; }
    else { icx = 4; icy = 5; icz = 2; }
    return -xw[bb + 3] * yw[a] * zw[c] * src[(long long)(icx * I.npts + r) * nt + it] -
           xw[bb] * yw[a + 3] * zw[c] * src[(long long)(icy * I.npts + r) * nt + it] -
           xw[bb] * yw[a] * zw[c + 3] * src[(long long)(icz * I.npts + r) * nt + it];
}

// Acoustic: next[p] += taper( dt^2 * K * inj / scale ) at the 27 points of each source.
// mode 0: time-step injection into next; mode 1: initial prev = 0.5 K dt^2 inj(src[0]).
// Thread j of the injection handles target j (j < ntgt).
__device__ __forceinline__ void acoustic_inject_one(float* __restrict__ out, const AcuArgs& A, const InjTab& I,
                                                    const float* __restrict__ src, int nt, int it, int mode, int j) {
    const Geom& g = A.g;
    const float* __restrict__ K = A.K;
    const float dt = A.dt;
    const int e0 = I.tbeg[j], e1 = I.tbeg[j + 1];
    int r, a, bb, c, de, ix, iy, iz;
    inj_decode(I, I.tlst[e0], r, a, bb, c, de);
    inj_point(I, r, a, bb, c, de, ix, iy, iz);
    const int p = g.at(ix, iy, iz);
    float val = out[p];
    for (int e = e0; e < e1; e++) {
        inj_decode(I, I.tlst[e], r, a, bb, c, de);
        const float v = inj_value(I, src, nt, it, 0, r, a, bb, c, de);
        float add;
        if (mode == 1) {
            add = v * (0.5f * K[p] * dt * dt);
        } else {
            add = dt * dt * v * K[p];
            add = add / acu_scale(g, A.bc, K[p], A.b[p], ix, iy, iz, dt);
            add = taper_apply(add, A.tz2, A.tx2, A.ty2, ix, iy, iz);
        }
        val += add;
    }
    out[p] = val;
}
__global__ void acoustic_inject(float* __restrict__ out, AcuArgs A, InjTab I, const float* __restrict__ src, int nt, int it,
                                int mode) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < I.ntgt) acoustic_inject_one(out, A, I, src, nt, it, mode, j);
}

// rcv[comp][r][it] = Lagrange-weighted sum over the 27 points (delta_m3::extract; for a
// dipole the difference of the two ends' sums);
// component comp = blockIdx.y reads field u + comp * g.n.
__device__ __forceinline__ void extract_one(const float* __restrict__ u3, const Geom& g, const InjTab& I,
                                            float* __restrict__ rcv, int nt, int it, int r, int comp) {
    const GAS float* __restrict__ u = gas(u3 + comp * g.n);
    const GAS float* xw = gas(I.xw + r * 2 * I.nw + I.nw);
    const GAS float* yw = gas(I.yw + r * 2 * I.nw + I.nw);
    const GAS float* zw = gas(I.zw + r * 2 * I.nw + I.nw);
    float acc = 0.f;
    if (I.kind == 2) {  // dipole_m3::extract (injector.cpp:461-492): u(end 2) - u(end 1)
        const GAS int* xi = gas(I.xi + 6 * r);
        const GAS int* yi = gas(I.yi + 6 * r);
        const GAS int* zi = gas(I.zi + 6 * r);
        for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++)
                for (int c = 0; c < 3; c++)
                    acc += xw[3 + bb] * yw[3 + a] * zw[3 + c] * u[g.at(xi[3 + bb], yi[3 + a], zi[3 + c])] -
                           xw[bb] * yw[a] * zw[c] * u[g.at(xi[bb], yi[a], zi[c])];
        gas(rcv)[(long long)(comp * I.npts + r) * nt + it] = acc;
        return;
    }
    const GAS int* xi = gas(I.xi + 3 * r);
    const GAS int* yi = gas(I.yi + 3 * r);
    const GAS int* zi = gas(I.zi + 3 * r);
    for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++)
                acc += yw[a] * xw[bb] * zw[c] * u[g.at(xi[bb], yi[a], zi[c])];
    gas(rcv)[(long long)(comp * I.npts + r) * nt + it] = acc;
}

// Imaging condition of nl_we_op_a::compute_gradients (we_op.cpp:1555-1605), FP64
// accumulation. Snapshot views of slots itl-1, itl, itl+1 (a slot past the store has a
// null base and reads as zero); branch: 0 first sample, 1 middle, 2 last. Gradients
// use the unpadded reference layout.
template <int FMT>
__global__ void acoustic_grad(SnapView<FMT> fm, SnapView<FMT> fn, SnapView<FMT> fp, const float* __restrict__ adj,
                              const float* __restrict__ K, Geom g, float sdt, int branch, double* __restrict__ gK,
                              double* __restrict__ gB) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y * blockDim.y + threadIdx.y;
    const int iy = blockIdx.z;
    if (iz >= g.nz || ix >= g.nx) return;
    const int p = g.at(ix, iy, iz);
    const float sn = fn.scale();
    auto FN = [&](int i) { return fn.code(i); };
    const float fx = d1_acc(FN, p, ix, g.nx, g.sx) * sn / g.dx;
    const float fy = d1_acc(FN, p, iy, g.ny, g.sy) * sn / g.dy;
    const float fz = d1_acc(FN, p, iz, g.nz, 1) * sn / g.dz;
    const float ax = d1_point(adj, p, ix, g.nx, g.sx) / g.dx;
    const float ay = d1_point(adj, p, iy, g.ny, g.sy) / g.dy;
    const float az = d1_point(adj, p, iz, g.nz, 1) / g.dz;
    const float dot = ax * fx + ay * fy + az * fz;
    const float k = K[p];
    const float fnv = fn.code(p) * sn;
    const float fmv = fm.code(p) * fm.scale();
    const float fpv = fp.p ? fp.code(p) * fp.scale() : 0.f;
    float d2t;
    if (branch == 1) d2t = fpv - 2.f * fnv + fmv;
    else if (branch == 0) d2t = fpv - fnv;
    else d2t = -2.f * fnv + fmv;
    const float gk = 1.0f / (sdt * k * k) * adj[p] * d2t;
    const float gb = (branch == 0 ? 0.5f : 1.0f) * sdt * dot;
    const long long q = ((long long)iy * g.nx + ix) * g.nz + iz;
    gK[q] -= (double)gk;
    gB[q] += (double)gb;
}

}  // namespace fwi3d
