// Acoustic variable-density kernels (gfx950).
//
// One time step of nl_we_op_a::propagate (src/cpp/we_op.cpp:1662-1791) is one
// launch of acoustic_step: D2(1/rho) along x, y, z, the face SATs, the leapfrog
// update, the absorbing-boundary scaling and the cos^2 taper, fused. Source
// injection is a separate tiny launch that adds its already-updated contribution
// (the update is linear in the injected term); receiver extraction reads curr,
// which the step does not modify.
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
    const float* __restrict__ curr;
    const float* __restrict__ prev;
    float* __restrict__ next;
    const float* __restrict__ K;    // bulk modulus
    const float* __restrict__ b;    // 1/rho
    Geom g;
    Bound B;
    float dt, dt2, alpha;  // alpha = 0.2508560249 / free_surface_stiffness
    int prev_tapered;
    int ychunk, ntz, ntx, nyc;
};

// Divisor of the absorbing SAT's implicit part at a boundary point (1 elsewhere).
__device__ __forceinline__ float acu_scale(const Geom& g, const Bound& B, const float* __restrict__ K,
                                           const float* __restrict__ b, long long p, int ix, int iy, int iz, float dt) {
    const bool onb = (iz == 0) | (iz == g.nz - 1) | (ix == 0) | (ix == g.nx - 1) | (iy == 0) | (iy == g.ny - 1);
    if (!onb) return 1.f;
    float sc = 0.f;
    bool any = false;
    const float c = sqrtf(K[p] * b[p]);
    if (iz == 0 && B.bc[0] == 2) { sc += c * dt / (2.f * g.dz * SAT_H0); any = true; }
    if (iz == g.nz - 1 && B.bc[1] == 2) { sc += c * dt / (2.f * g.dz * SAT_H0); any = true; }
    if (ix == 0 && B.bc[2] == 2) { sc += c * dt / (2.f * g.dx * SAT_H0); any = true; }
    if (ix == g.nx - 1 && B.bc[3] == 2) { sc += c * dt / (2.f * g.dx * SAT_H0); any = true; }
    if (iy == 0 && B.bc[4] == 2) { sc += c * dt / (2.f * g.dy * SAT_H0); any = true; }
    if (iy == g.ny - 1 && B.bc[5] == 2) { sc += c * dt / (2.f * g.dy * SAT_H0); any = true; }
    return any ? 1.f + sc : 1.f;
}

// Sum of the face SAT terms at point p (within 4 cells of a face); pv is prev at p.
// Faces in the reference's order top, bottom, left, right, front, back.
__device__ float acu_sat(const AcuArgs& A, long long p, int ix, int iy, int iz, float pv) {
    const Geom& g = A.g;
    float acc = 0.f;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int bc = A.B.bc[f];
        if (bc == 0) continue;
        const int ax = f >> 1;  // 0: z, 1: x, 2: y
        const bool hi = f & 1;
        const int i = ax == 0 ? iz : (ax == 1 ? ix : iy);
        const int n = ax == 0 ? g.nz : (ax == 1 ? g.nx : g.ny);
        const long long s = ax == 0 ? 1 : (ax == 1 ? g.sx : g.sy);
        const float h = ax == 0 ? g.dz : (ax == 1 ? g.dx : g.dy);
        const int k = hi ? n - 1 - i : i;
        if (k >= 4) continue;
        const long long inward = hi ? -s : s;
        if (bc == 1) {
            const long long pf = p - (long long)k * inward;
            const float v = A.b[pf] * A.curr[pf];
            const float d2 = h * h;
            acc += 1.0f / (c_hnorm[k] * d2) * c_sat_s[k] * v;
            if (k == 0) acc -= v / (A.alpha * SAT_H0 * d2);
        } else if (k == 0) {
            const float sd = sat_s(A.curr, p, inward);
            const float imp = (bc == 2) ? 0.5f / sqrtf(A.K[p] / A.b[p]) : 0.f;
            const float adh = 1.0f / (SAT_H0 * h);
            acc -= adh * (A.b[p] * sd / h - imp * pv / A.dt);
        }
    }
    return acc;
}

// Map a workgroup to a tile so that the workgroups one XCD receives (hardware
// assigns workgroup i to XCD i % 8) are spatially contiguous.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int q = nblk >> 3, r = nblk & 7;
    const int x = bid & 7, k = bid >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

template <int TX>
__global__ __launch_bounds__(64 * TX) void acoustic_step(AcuArgs A) {
    constexpr int TZ = 64;
    __shared__ float su[TX + 4][TZ + 4];
    __shared__ float sb[TX + 4][TZ + 4];
    const Geom g = A.g;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = L % A.ntz;
    const int tx = (L / A.ntz) % A.ntx;
    const int cy = L / (A.ntz * A.ntx);
    const int lz = threadIdx.x, lx = threadIdx.y;
    const int z0 = tz * TZ, x0 = tx * TX;
    const int iz = z0 + lz, ix = x0 + lx;
    const bool valid = (iz < g.nz) && (ix < g.nx);
    const int y0 = cy * A.ychunk;
    const int y1 = min(g.ny, y0 + A.ychunk);
    const long long col = (long long)ix * g.sx + iz;

    auto ldc = [&](const float* __restrict__ f, int y) -> float {
        return (valid && y >= 0 && y < g.ny) ? f[(long long)y * g.sy + col] : 0.f;
    };
    auto ldp = [&](const float* __restrict__ f, int x, int y, int z) -> float {
        return (x >= 0 && x < g.nx && z >= 0 && z < g.nz) ? f[g.at(x, y, z)] : 0.f;
    };

    float um2 = ldc(A.curr, y0 - 2), um1 = ldc(A.curr, y0 - 1), u0 = ldc(A.curr, y0), up1 = ldc(A.curr, y0 + 1), up2;
    float bm2 = ldc(A.b, y0 - 2), bm1 = ldc(A.b, y0 - 1), b0 = ldc(A.b, y0), bp1 = ldc(A.b, y0 + 1), bp2;
    const float idx2 = 1.0f / (g.dx * g.dx), idy2 = 1.0f / (g.dy * g.dy), idz2 = 1.0f / (g.dz * g.dz);

    for (int iy = y0; iy < y1; iy++) {
        up2 = ldc(A.curr, iy + 2);
        bp2 = ldc(A.b, iy + 2);
        __syncthreads();
        su[lx + 2][lz + 2] = u0;
        sb[lx + 2][lz + 2] = b0;
        if (lx < 2) {
            su[lx][lz + 2] = ldp(A.curr, x0 - 2 + lx, iy, iz);
            sb[lx][lz + 2] = ldp(A.b, x0 - 2 + lx, iy, iz);
            su[TX + 2 + lx][lz + 2] = ldp(A.curr, x0 + TX + lx, iy, iz);
            sb[TX + 2 + lx][lz + 2] = ldp(A.b, x0 + TX + lx, iy, iz);
        }
        if (lz < 2) {
            su[lx + 2][lz] = ldp(A.curr, ix, iy, z0 - 2 + lz);
            sb[lx + 2][lz] = ldp(A.b, ix, iy, z0 - 2 + lz);
            su[lx + 2][TZ + 2 + lz] = ldp(A.curr, ix, iy, z0 + TZ + lz);
            sb[lx + 2][TZ + 2 + lz] = ldp(A.b, ix, iy, z0 + TZ + lz);
        }
        __syncthreads();
        if (valid) {
            const long long p = (long long)iy * g.sy + col;
            float sx_, sy_, sz_;
            if (ix >= 6 && ix < g.nx - 6)
                sx_ = d2_interior(su[lx][lz + 2], su[lx + 1][lz + 2], u0, su[lx + 3][lz + 2], su[lx + 4][lz + 2],
                                  sb[lx][lz + 2], sb[lx + 1][lz + 2], b0, sb[lx + 3][lz + 2], sb[lx + 4][lz + 2]);
            else
                sx_ = d2_point(A.curr, A.b, p, ix, g.nx, g.sx);
            if (iy >= 6 && iy < g.ny - 6)
                sy_ = d2_interior(um2, um1, u0, up1, up2, bm2, bm1, b0, bp1, bp2);
            else
                sy_ = d2_point(A.curr, A.b, p, iy, g.ny, g.sy);
            if (iz >= 6 && iz < g.nz - 6)
                sz_ = d2_interior(su[lx + 2][lz], su[lx + 2][lz + 1], u0, su[lx + 2][lz + 3], su[lx + 2][lz + 4],
                                  sb[lx + 2][lz], sb[lx + 2][lz + 1], b0, sb[lx + 2][lz + 3], sb[lx + 2][lz + 4]);
            else
                sz_ = d2_point(A.curr, A.b, p, iz, g.nz, 1);
            float lap = idx2 * sx_;
            lap = lap + idy2 * sy_;
            lap = lap + idz2 * sz_;
            float pv = A.prev[p];
            if (A.prev_tapered) pv = taper_apply(pv, A.B, ix, iy, iz);
            const bool near = (iz < 4) | (iz >= g.nz - 4) | (ix < 4) | (ix >= g.nx - 4) | (iy < 4) | (iy >= g.ny - 4);
            if (near) lap += acu_sat(A, p, ix, iy, iz, pv);
            float nv = A.dt2 * lap * A.K[p] + 2.f * u0 - pv;
            nv = nv / acu_scale(g, A.B, A.K, A.b, p, ix, iy, iz, A.dt);
            A.next[p] = taper_apply(nv, A.B, ix, iy, iz);
        }
        um2 = um1; um1 = u0; u0 = up1; up1 = up2;
        bm2 = bm1; bm1 = b0; b0 = bp1; bp1 = bp2;
    }
}

// ------------------------------------------------------------------------------------
// Point sources / receivers. Tables: ind [npts][3], w [npts][2][nw] (row 0 inject,
// row 1 extract). One thread per (point, 27 neighbours).
struct InjTab {
    const int* xi;
    const int* yi;
    const int* zi;
    const float* xw;
    const float* yw;
    const float* zw;
    int npts, kind, nw;
};

// Injected value of displacement/pressure component `comp` at neighbour (a,b,c) of point r.
__device__ __forceinline__ float inj_value(const InjTab& I, const float* __restrict__ src, int nt, int it, int comp, int r,
                                           int a, int bb, int c) {
    const float* xw = I.xw + (long long)r * 2 * I.nw;
    const float* yw = I.yw + (long long)r * 2 * I.nw;
    const float* zw = I.zw + (long long)r * 2 * I.nw;
    if (I.kind == 0) return yw[a] * xw[bb] * zw[c] * src[((long long)comp * I.npts + r) * nt + it];
    int icx, icy, icz;
    if (comp == 0) { icx = 0; icy = 3; icz = 4; }
    else if (comp == 1) { icx = 3; icy = 1; icz = 5; }
    else { icx = 4; icy = 5; icz = 2; }
    return -xw[bb + 3] * yw[a] * zw[c] * src[((long long)icx * I.npts + r) * nt + it] -
           xw[bb] * yw[a + 3] * zw[c] * src[((long long)icy * I.npts + r) * nt + it] -
           xw[bb] * yw[a] * zw[c + 3] * src[((long long)icz * I.npts + r) * nt + it];
}

// Acoustic: next[p] += taper( dt^2 * K * inj / scale ) at the 27 points of each source.
// mode 0: time-step injection into next; mode 1: initial prev = 0.5 K dt^2 inj(src[0]).
__global__ void acoustic_inject(float* __restrict__ out, const float* __restrict__ K, const float* __restrict__ b,
                                Geom g, Bound B, InjTab I, const float* __restrict__ src, int nt, int it, float dt,
                                int mode) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= I.npts * 27) return;
    const int r = t / 27, q = t % 27;
    const int a = q / 9, bb = (q / 3) % 3, c = q % 3;
    const int ix = I.xi[3 * r + bb], iy = I.yi[3 * r + a], iz = I.zi[3 * r + c];
    const long long p = g.at(ix, iy, iz);
    const float v = inj_value(I, src, nt, it, 0, r, a, bb, c);
    float add;
    if (mode == 1) {
        add = v * (0.5f * K[p] * dt * dt);
    } else {
        add = dt * dt * v * K[p];
        add = add / acu_scale(g, B, K, b, p, ix, iy, iz, dt);
        add = taper_apply(add, B, ix, iy, iz);
    }
    atomicAdd(out + p, add);
}

// rcv[comp][r][it] = Lagrange-weighted sum over the 27 points (delta_m3::extract).
__global__ void extract_points(const float* __restrict__ u, Geom g, InjTab I, float* __restrict__ rcv, int nt, int it,
                               int comp) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= I.npts) return;
    const float* xw = I.xw + (long long)r * 2 * I.nw + I.nw;
    const float* yw = I.yw + (long long)r * 2 * I.nw + I.nw;
    const float* zw = I.zw + (long long)r * 2 * I.nw + I.nw;
    float acc = 0.f;
    for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++)
                acc += yw[a] * xw[bb] * zw[c] * u[g.at(I.xi[3 * r + bb], I.yi[3 * r + a], I.zi[3 * r + c])];
    rcv[((long long)comp * I.npts + r) * nt + it] = acc;
}

// Imaging condition of nl_we_op_a::compute_gradients (we_op.cpp:1555-1605), FP64
// accumulation. fm may be null (first slot); branch: 0 first sample, 1 middle, 2 last.
__global__ void acoustic_grad(const float* __restrict__ fm, const float* __restrict__ fn, const float* __restrict__ fp,
                              const float* __restrict__ adj, const float* __restrict__ K, Geom g, float sdt, int branch,
                              double* __restrict__ gK, double* __restrict__ gB, long long gstride) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    const int iy = blockIdx.z;
    if (iz >= g.nz) return;
    const long long p = g.at(ix, iy, iz);
    const float fx = d1_point(fn, p, ix, g.nx, g.sx) / g.dx;
    const float fy = d1_point(fn, p, iy, g.ny, g.sy) / g.dy;
    const float fz = d1_point(fn, p, iz, g.nz, 1) / g.dz;
    const float ax = d1_point(adj, p, ix, g.nx, g.sx) / g.dx;
    const float ay = d1_point(adj, p, iy, g.ny, g.sy) / g.dy;
    const float az = d1_point(adj, p, iz, g.nz, 1) / g.dz;
    const float dot = ax * fx + ay * fy + az * fz;
    const float k = K[p];
    float d2t;
    const float fpv = fp ? fp[p] : 0.f;
    if (branch == 1) d2t = fpv - 2.f * fn[p] + fm[p];
    else if (branch == 0) d2t = fpv - fn[p];
    else d2t = -2.f * fn[p] + fm[p];
    const float gk = 1.0f / (sdt * k * k) * adj[p] * d2t;
    const float gb = (branch == 0 ? 0.5f : 1.0f) * sdt * dot;
    // gradient arrays use the unpadded reference layout
    const long long q = ((long long)iy * g.nx + ix) * g.nz + iz;
    gK[q] -= (double)gk;
    gB[q] += (double)gb;
    (void)gstride;
}

}  // namespace fwi3d
