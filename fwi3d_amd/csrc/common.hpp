// Device-side building blocks shared by the gfx950 kernels: grid geometry,
// SBP coefficients in constant memory, and per-point operator evaluations used on
// boundary rows (the interior paths of the y-marching kernels do not use these).
//
// All kernels of the engine live in one translation unit (engine.hip includes
// this header once), so the __constant__ tables below are defined exactly once.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "sbp_tables.hpp"

namespace fwi3d {

// Wavefield / model layout in HBM: [y][x][z] with the z row padded to a pitch `pz`
// that is a multiple of 64 floats (one 256-byte wavefront load per row segment).
// Offsets inside one field (and inside [3][field] arrays) are 32-bit: the host
// rejects grids with 3 * n >= 2^31.
struct Geom {
    int nx, ny, nz, pz;
    int sx, sy;  // element strides of x and y (z stride is 1)
    int n;       // padded elements per field: ny * sy
    float dx, dy, dz;
    __host__ __device__ int at(int ix, int iy, int iz) const { return iy * sy + ix * sx + iz; }
};

// Global-memory pointers in address space 1. The step kernels read their field pointers
// from the per-shot table in device memory, so the compiler cannot infer the address
// space of those pointers and would emit generic (FLAT) accesses, which also count
// against the LDS wait counter: every LDS barrier would wait for all prefetched global
// loads, and each access would need 64-bit address arithmetic. Casting to GAS pointers
// and indexing with 32-bit unsigned element offsets gives global loads with a scalar
// base and a 32-bit vector offset.
#define GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const GAS T* gas(const T* p) { return (const GAS T*)p; }
template <class T>
__device__ __forceinline__ GAS T* gas(T* p) { return (GAS T*)p; }
// (the byte offset is formed in 32 bits, so the access is base + zext(offset): a scalar
// base register pair and one 32-bit vector offset)
__device__ __forceinline__ float gld(const float* base, unsigned i) {
    return *(const GAS float*)((const GAS char*)gas(base) + (i << 2));
}
__device__ __forceinline__ void gst(float* base, unsigned i, float v) { *(GAS float*)((GAS char*)gas(base) + (i << 2)) = v; }

// S and H weights (the closure rows are compile-time constants, below).
__constant__ float c_sat_s[4];
__constant__ float c_hnorm[4];

#define D1C1 (2.0f / 3.0f)
#define D1C2 (-1.0f / 12.0f)
#define SAT_H0 (17.0f / 48.0f)

// D2 interior weights (spatial_operators.hpp:37).
#define D2C0 (-3.0f / 4.0f)
#define D2C1 (-5.0f / 6.0f)
#define D2C2 (-1.0f / 24.0f)
#define D2C3 (1.0f / 6.0f)
#define D2C4 (1.0f / 2.0f)
#define D2C5 (1.0f / 2.0f)
#define D2C6 (1.0f / 6.0f)
#define D2C7 (-1.0f / 8.0f)
#define D2C8 (1.0f / 6.0f)
#define D2C9 (-1.0f / 8.0f)

// Interior D2(b) (times h^2) from the 5 u and 5 b values around the point.
__device__ __forceinline__ float d2_interior(float um2, float um1, float u0, float up1, float up2, float bm2, float bm1,
                                             float b0, float bp1, float bp2) {
    return (D2C0 * b0 + D2C1 * (bm1 + bp1) + D2C2 * (bm2 + bp2)) * u0 +
           (D2C3 * bm2 + D2C4 * bm1 + D2C5 * b0 + D2C6 * bp1) * um1 +
           (D2C3 * bp2 + D2C4 * bp1 + D2C5 * b0 + D2C6 * bm1) * up1 +
           (D2C7 * bm2 + D2C8 * bm1 + D2C9 * b0) * um2 + (D2C7 * bp2 + D2C8 * bp1 + D2C9 * b0) * up2;
}

// ------------------------------------------------------------------------------------
// Closure rows with compile-time coefficients: row R keeps only its non-zero entries
// of the sparse tables, so a wave whose lanes sit on different rows runs each row's
// ~22 FMAs under its own mask instead of loading the coefficient table per lane.
// u(j) / b(k) return the field / coefficient at the j-th / k-th point from the face.
template <int E>
struct D2Term {
    static constexpr int r = kD2Closure[E].r, j = kD2Closure[E].j, k = kD2Closure[E].k;
    static constexpr float w = (float)((double)kD2Closure[E].num / (double)kD2Closure[E].den);
};
template <int R, int E, class U, class B>
__device__ __forceinline__ float d2_term(const U& u, const B& b) {
    if constexpr (D2Term<E>::r == R) return (D2Term<E>::w * b(D2Term<E>::k)) * u(D2Term<E>::j);
    else return 0.f;
}
template <int R, class U, class B, int... E>
__device__ __forceinline__ float d2_row_impl(const U& u, const B& b, std::integer_sequence<int, E...>) {
    return (0.f + ... + d2_term<R, E>(u, b));
}
// D2(b) (times h^2) of closure row r in [0, 6).
template <class U, class B>
__device__ __forceinline__ float d2_rows(int r, const U& u, const B& b) {
    using Seq = std::make_integer_sequence<int, 132>;
    switch (r) {
        case 0: return d2_row_impl<0>(u, b, Seq{});
        case 1: return d2_row_impl<1>(u, b, Seq{});
        case 2: return d2_row_impl<2>(u, b, Seq{});
        case 3: return d2_row_impl<3>(u, b, Seq{});
        case 4: return d2_row_impl<4>(u, b, Seq{});
        default: return d2_row_impl<5>(u, b, Seq{});
    }
}
template <int R, int J, class U>
__device__ __forceinline__ float d1_term(const U& u) {
    constexpr float w = (float)kD1Closure[R][J];
    if constexpr (kD1Closure[R][J] != 0.0) return w * u(J);
    else return 0.f;
}
template <int R, class U>
__device__ __forceinline__ float d1_row(const U& u) {
    return d1_term<R, 0>(u) + d1_term<R, 1>(u) + d1_term<R, 2>(u) + d1_term<R, 3>(u) + d1_term<R, 4>(u) + d1_term<R, 5>(u);
}
// D1 (times h) of closure row r in [0, 4), counted from the low face (negate for the
// high face, where u(j) runs inward from the last point).
template <class U>
__device__ __forceinline__ float d1_rows(int r, const U& u) {
    switch (r) {
        case 0: return d1_row<0>(u);
        case 1: return d1_row<1>(u);
        case 2: return d1_row<2>(u);
        default: return d1_row<3>(u);
    }
}

// D2(b) (times h^2) at axis index i (any row) reading global memory: the field and
// coefficient at point p, axis length n, element stride s. On boundary rows (i < 6 or
// i >= n-6) along z (s == 1) the lanes of a wave sit on different rows, so the eight
// u and b values are loaded once for all lanes and each row's terms then run from
// registers; along x / y (one row per wave) the row loads only its own operands.
__device__ __forceinline__ float d2_point(const float* __restrict__ u_, const float* __restrict__ b_, int p, int i, int n,
                                          int s) {
    const GAS float* __restrict__ u = gas(u_);
    const GAS float* __restrict__ b = gas(b_);
    if (i >= 6 && i < n - 6) {
        return d2_interior(u[p - 2 * s], u[p - s], u[p], u[p + s], u[p + 2 * s], b[p - 2 * s], b[p - s], b[p],
                           b[p + s], b[p + 2 * s]);
    }
    const bool lo = i < 6;
    const int r = lo ? i : n - 1 - i;
    const int base = lo ? p - i * s : p + (n - 1 - i) * s;
    const int step = lo ? s : -s;
    if (s == 1) {
        float uk[8], bk[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            uk[k] = u[base + k * step];
            bk[k] = b[base + k * step];
        }
        return d2_rows(r, [&](int j) { return uk[j]; }, [&](int k) { return bk[k]; });
    }
    return d2_rows(r, [&](int j) { return u[base + j * step]; }, [&](int k) { return b[base + k * step]; });
}
// D1 (times h) on a boundary row (i < 4 or i >= n-4), reading global memory (row
// forms as for d2_point).
__device__ __forceinline__ float d1_closure(const float* __restrict__ u_, int p, int i, int n, int s) {
    const GAS float* __restrict__ u = gas(u_);
    const bool lo = i < 4;
    const int r = lo ? i : n - 1 - i;
    const int base = lo ? p - i * s : p + (n - 1 - i) * s;
    const int step = lo ? s : -s;
    float v;
    if (s == 1) {
        float uk[6];
#pragma unroll
        for (int k = 0; k < 6; k++) uk[k] = u[base + k * step];
        v = d1_rows(r, [&](int j) { return uk[j]; });
    } else {
        v = d1_rows(r, [&](int j) { return u[base + j * step]; });
    }
    return lo ? v : -v;
}
// D1 (times h) of closure row r from values already staged on chip: a(k) returns the
// k-th point counted from the face (k < 6); lo as for d1_closure. Same arithmetic as
// d1_closure.
template <class A_>
__device__ __forceinline__ float d1_closure_at(bool lo, int r, const A_& a) {
    float uk[6];
#pragma unroll
    for (int k = 0; k < 6; k++) uk[k] = a(k);
    const float v = d1_rows(r, [&](int j) { return uk[j]; });
    return lo ? v : -v;
}
// D1 (times h) at axis index i reading global memory.
__device__ __forceinline__ float d1_point(const float* __restrict__ u_, int p, int i, int n, int s) {
    const GAS float* __restrict__ u = gas(u_);
    if (i >= 4 && i < n - 4) return D1C1 * (u[p + s] - u[p - s]) + D1C2 * (u[p + 2 * s] - u[p - 2 * s]);
    return d1_closure(u_, p, i, n, s);
}

// Outward boundary derivative S (times h) at a face point p, `inward` = +/- stride.
__device__ __forceinline__ float sat_s(const float* __restrict__ u_, int p, int inward) {
    const GAS float* __restrict__ u = gas(u_);
    return c_sat_s[0] * u[p] + c_sat_s[1] * u[p + inward] + c_sat_s[2] * u[p + 2 * inward] + c_sat_s[3] * u[p + 3 * inward];
}

// D1 (times h) at axis index i through an accessor f(offset) (snapshot views).
template <class F>
__device__ __forceinline__ float d1_acc(const F& f, int p, int i, int n, int s) {
    if (i >= 4 && i < n - 4) return D1C1 * (f(p + s) - f(p - s)) + D1C2 * (f(p + 2 * s) - f(p - 2 * s));
    const bool lo = i < 4;
    const int base = lo ? p - i * s : p + (n - 1 - i) * s;
    const int step = lo ? s : -s;
    const float v = d1_rows(lo ? i : n - 1 - i, [&](int j) { return f(base + j * step); });
    return lo ? v : -v;
}

// ------------------------------------------------------------------------------------
// Snapshot store. FP32, or 16-bit codes (BF16 / FP16) of u / s with one scale per
// snapshot, s = max|u| / 2^15: the largest value maps to 2^15, inside FP16's range,
// and small values stay normal down to max|u| * 2^-29. The max is reduced on the
// device as float bits (non-negative floats order like unsigned integers).
enum SnapFmt { SNAP_F32 = 0, SNAP_BF16 = 1, SNAP_F16 = 2 };
constexpr float kSnapTop = 32768.f;

__device__ __forceinline__ unsigned short bf16_rne(float v) {
    unsigned u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <int FMT>
struct SnapView {
    const void* p;           // slot base (null: a slot past the store, reads as zero)
    const unsigned* maxbits;  // the slot's max|u| as float bits (16-bit formats)
    __device__ float scale() const { return FMT == SNAP_F32 ? 1.f : __uint_as_float(*maxbits) / kSnapTop; }
    __device__ float code(int i) const {
        if (FMT == SNAP_F32) return ((const float*)p)[i];
        if (FMT == SNAP_BF16) return bf16_val(((const unsigned short*)p)[i]);
        return __half2float(((const __half*)p)[i]);
    }
};

// Per-step boundary description shared by the physics kernels.
struct Bound {
    int bc[6];  // top,bottom,left,right,front,back
    const float* tz2;  // squared cos taper factor per z index (1 outside the ramps)
    const float* tx2;
    const float* ty2;
};

// taper: ((v * fz) * fx) * fy, the order of the reference's six taper calls.
__device__ __forceinline__ float taper_apply(float v, const float* __restrict__ tz2, const float* __restrict__ tx2,
                                             const float* __restrict__ ty2, int ix, int iy, int iz) {
    return ((v * gld(tz2, iz)) * gld(tx2, ix)) * gld(ty2, iy);
}

// Workgroup -> tile remap so that the workgroups one XCD receives (hardware assigns
// workgroup i to XCD i % 8) are spatially contiguous.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int q = nblk >> 3, r = nblk & 7;
    const int x = bid & 7, k = bid >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Halo staging map of the y-marching kernels: a 64(z) x TX(x) tile staged in LDS
// as [TX+4][64+4] with 2-cell halos. The 4 x-halo rows (4 x 64 values) and the
// TX x 4 z-halo values are spread over the block's threads (thread h = lx*64+lz):
// every thread loads at most one x-halo and one z-halo value per plane.
template <int TX>
struct HaloMap {
    bool has_x, has_z;
    int xrow, xcol, xsrc;  // LDS row/column and in-plane global offset of the x-halo value
    int zrow, zcol, zsrc;  // same for the z-halo value
    __device__ void init(int lx, int lz, int x0, int z0, int nx, int nz, int sx) {
        const int h = lx * 64 + lz;
        has_x = h < 256;
        const int r4 = (h >> 6) & 3;
        xrow = r4 < 2 ? r4 : TX + r4;
        xcol = (h & 63) + 2;
        const int gx = r4 < 2 ? x0 - 2 + r4 : x0 + TX + r4 - 2;
        xsrc = max(0, min(nx - 1, gx)) * sx + min(z0 + (h & 63), nz - 1);
        has_z = h < 4 * TX;
        const int zr = (h >> 2) % TX, zc = h & 3;
        zrow = zr + 2;
        zcol = zc < 2 ? zc : 64 + zc;
        const int gz = zc < 2 ? z0 - 2 + zc : z0 + 64 + zc - 2;
        zsrc = min(x0 + zr, nx - 1) * sx + max(0, min(nz - 1, gz));
    }
};

// Boundary shell = points within `w` cells of any face. Enumerated as: the 2w
// y-planes at each end (all x, z), then inside them the 2w x-rows (all z), then the
// 2w z-points of the remaining columns. Returns false past the end.
__device__ __forceinline__ bool shell_point(int t, int w, int nx, int ny, int nz, int& ix, int& iy, int& iz) {
    const int a = 2 * w * nx * nz;
    if (t < a) {
        const int pl = t / (nx * nz), r = t - pl * (nx * nz);
        iy = pl < w ? pl : ny - 2 * w + pl;
        ix = r / nz;
        iz = r - ix * nz;
        return true;
    }
    t -= a;
    const int nyi = ny - 2 * w;
    const int b = nyi * 2 * w * nz;
    if (t < b) {
        const int pl = t / (2 * w * nz), r = t - pl * (2 * w * nz);
        iy = w + pl;
        const int xr = r / nz;
        ix = xr < w ? xr : nx - 2 * w + xr;
        iz = r - xr * nz;
        return true;
    }
    t -= b;
    const int nxi = nx - 2 * w;
    if (t < nyi * nxi * 2 * w) {
        const int col = t / (2 * w), k = t - col * (2 * w);
        iy = w + col / nxi;
        ix = w + col % nxi;
        iz = k < w ? k : nz - 2 * w + k;
        return true;
    }
    return false;
}

// Edges and corners of the shell: the points within w cells of faces of two or three
// different axes (what the slab kernels leave to the point-wise shell kernel).
// Enumerated as: in each of the 2w y-planes at the ends, the 2w x-rows (all z), then the
// 2w z-points of the columns between them; then, for the y planes between, the 2w z-points
// of the 2w x-rows. Returns false past the end.
__host__ __device__ inline int edge_count(int w, int nx, int ny, int nz) {
    return 2 * w * (2 * w * nz + (nx - 2 * w) * 2 * w) + (ny - 2 * w) * 4 * w * w;
}
__device__ __forceinline__ bool edge_point(int t, int w, int nx, int ny, int nz, int& ix, int& iy, int& iz) {
    const int per = 2 * w * nz + (nx - 2 * w) * 2 * w;
    if (t < 2 * w * per) {
        const int pl = t / per;
        int r = t - pl * per;
        iy = pl < w ? pl : ny - 2 * w + pl;
        if (r < 2 * w * nz) {
            const int xr = r / nz;
            ix = xr < w ? xr : nx - 2 * w + xr;
            iz = r - xr * nz;
        } else {
            r -= 2 * w * nz;
            const int xi = r / (2 * w), zr = r - xi * (2 * w);
            ix = w + xi;
            iz = zr < w ? zr : nz - 2 * w + zr;
        }
        return true;
    }
    t -= 2 * w * per;
    const int per2 = 4 * w * w;
    const int pl = t / per2;
    if (pl >= ny - 2 * w) return false;
    const int r = t - pl * per2;
    const int xr = r / (2 * w), zr = r - xr * (2 * w);
    iy = w + pl;
    ix = xr < w ? xr : nx - 2 * w + xr;
    iz = zr < w ? zr : nz - 2 * w + zr;
    return true;
}

__host__ __device__ inline int shell_count(int w, int nx, int ny, int nz) {
    return 2 * w * nx * nz + (ny - 2 * w) * 2 * w * nz + (ny - 2 * w) * (nx - 2 * w) * 2 * w;
}

}  // namespace fwi3d
