// Device-side building blocks shared by the gfx950 kernels: grid geometry,
// SBP coefficients in constant memory, and per-point operator evaluations used on
// boundary rows (the interior paths of the fused kernels do not use these).
//
// All kernels of the engine live in one translation unit (engine.hip includes
// this header once), so the __constant__ tables below are defined exactly once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fwi3d {

// Wavefield / model layout in HBM: [y][x][z] with the z row padded to a pitch `pz`
// that is a multiple of 64 floats (one 256-byte wavefront load per row segment).
struct Geom {
    int nx, ny, nz, pz;
    long long sx, sy;  // element strides of x and y (z stride is 1)
    long long n;       // padded elements per field: ny * sy
    float dx, dy, dz;
    __host__ __device__ long long at(int ix, int iy, int iz) const { return (long long)iy * sy + (long long)ix * sx + iz; }
};

// D2 closure C[row][u-point][b-point], D1 closure rows, S and H weights.
__constant__ float c_d2b[6][8][8];
__constant__ float c_d1b[4][6];
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

// D2(b) (times h^2) at axis index i (any row) reading global memory: the field and
// coefficient at point p, axis length n, element stride s.
__device__ __forceinline__ float d2_point(const float* __restrict__ u, const float* __restrict__ b, long long p, int i,
                                          int n, long long s) {
    if (i >= 6 && i < n - 6) {
        return d2_interior(u[p - 2 * s], u[p - s], u[p], u[p + s], u[p + 2 * s], b[p - 2 * s], b[p - s], b[p],
                           b[p + s], b[p + 2 * s]);
    }
    const bool lo = i < 6;
    const int r = lo ? i : n - 1 - i;
    // position of closure point k: lo -> k, hi -> n-1-k (relative to p: k-i or n-1-k-i)
    const long long base = lo ? p - (long long)i * s : p + (long long)(n - 1 - i) * s;
    const long long step = lo ? s : -s;
    float bk[8], uk[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        bk[k] = b[base + k * step];
        uk[k] = u[base + k * step];
    }
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float w = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) w += c_d2b[r][j][k] * bk[k];
        acc += w * uk[j];
    }
    return acc;
}

// D1 (times h) at axis index i reading global memory; optional coefficient `m`
// multiplies the field inside the derivative.
__device__ __forceinline__ float d1_point(const float* __restrict__ u, long long p, int i, int n, long long s) {
    if (i >= 4 && i < n - 4) return D1C1 * (u[p + s] - u[p - s]) + D1C2 * (u[p + 2 * s] - u[p - 2 * s]);
    const bool lo = i < 4;
    const int r = lo ? i : n - 1 - i;
    const long long base = lo ? p - (long long)i * s : p + (long long)(n - 1 - i) * s;
    const long long step = lo ? s : -s;
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 6; j++) acc += c_d1b[r][j] * u[base + j * step];
    return lo ? acc : -acc;
}

__device__ __forceinline__ float d1_point_m(const float* __restrict__ u, const float* __restrict__ m, long long p, int i,
                                            int n, long long s) {
    if (i >= 4 && i < n - 4)
        return D1C1 * (u[p + s] * m[p + s] - u[p - s] * m[p - s]) + D1C2 * (u[p + 2 * s] * m[p + 2 * s] - u[p - 2 * s] * m[p - 2 * s]);
    const bool lo = i < 4;
    const int r = lo ? i : n - 1 - i;
    const long long base = lo ? p - (long long)i * s : p + (long long)(n - 1 - i) * s;
    const long long step = lo ? s : -s;
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 6; j++) acc += c_d1b[r][j] * (u[base + j * step] * m[base + j * step]);
    return lo ? acc : -acc;
}

// Outward boundary derivative S (times h) at a face point p, `inward` = +/- stride.
__device__ __forceinline__ float sat_s(const float* __restrict__ u, long long p, long long inward) {
    return c_sat_s[0] * u[p] + c_sat_s[1] * u[p + inward] + c_sat_s[2] * u[p + 2 * inward] + c_sat_s[3] * u[p + 3 * inward];
}

// Per-step boundary description shared by the physics kernels.
struct Bound {
    int bc[6];  // top,bottom,left,right,front,back
    const float* tz2;  // squared cos taper factor per z index (1 outside the ramps)
    const float* tx2;
    const float* ty2;
};

// taper: ((v * fz) * fx) * fy, the order of the reference's six taper calls.
__device__ __forceinline__ float taper_apply(float v, const Bound& B, int ix, int iy, int iz) {
    return ((v * B.tz2[iz]) * B.tx2[ix]) * B.ty2[iy];
}

}  // namespace fwi3d
