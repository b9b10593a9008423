// Fused elastic / VTI time step (gfx950): one launch computes every point of the
// planes y in [6, ny-6) of one time step of nl_we_op_e::propagate
// (src/cpp/we_op.cpp:646-836) / nl_we_op_vti::propagate (we_op.cpp:1290-1497),
// including the x and z closure rows, the traction SATs of the x and z faces and the
// absorbing scaling there, with the normal stresses and shear fluxes formed on chip:
// no flux buffer travels through HBM for these planes. The 6 planes at each y face are
// left to the slab / edge kernels of elastic.hpp (they read a flux pass restricted to
// the 8 planes at each y face).
//
// Tile: 32 z points (one 128-byte line per row) x 16 x rows, 512 threads (a wave holds
// two x rows), marching along y. Each iteration stages plane o+2 of the three
// displacement components (with 2-cell halos) in a ring of three LDS planes, forms the
// fluxes that are differentiated along y (T_y, F_yx, F_yz) at plane o+2 into register
// queues, forms the fluxes differentiated along x and z at plane o for the tile and its
// halos into LDS, and then updates plane o. Halo fluxes are formed by halo threads: the
// first 128 threads own one point of the four x-halo rows, the next 64 one point of the
// four z-halo columns; each keeps the y queues of the components its fluxes
// differentiate along y at that point. The D1 stencils of the halo points reach two
// cells further (extx / extz).
//
// Tiling of an axis of n points into tiles of w (fused_tile): tiles at multiples of w
// whose outputs stop at n - w, then a last tile at n - w with the rest, so that the
// closure rows of each face and the points their stencils read lie inside one tile
// (tile 0 outputs at least [0, 8), the last tile [n-8, n)).
#pragma once
#include "elastic.hpp"

namespace fwi3d {

constexpr int FZ = 32;                 // z points per tile
constexpr int FW = 16;                 // x rows per tile
constexpr int FPZ = FZ + 4, FPX = FW + 4;
constexpr int FTHREADS = FZ * FW;      // 512

struct FusedTiles {
    int ntz, ntx, nyc, ychunk;
    int ya, yb;          // output planes [ya, yb) = [6, ny-6)
    int cnyc, cychunk;   // y chunks of the corner tiles (shorter: they are the costliest)
    int snyc, sychunk;   // y chunks of the side tiles (edge tiles with closures along one axis)
};

__host__ __device__ inline int fused_tile_count(int n, int w) {
    if (n <= w) return 1;
    if (n < w + 8) return 2;
    return (n - w + w - 1) / w + 1;
}
// Origin and output range [o0, o1) of tile t of nt.
__host__ __device__ inline void fused_tile(int n, int w, int t, int nt, int& org, int& o0, int& o1) {
    if (n <= w) {
        org = 0; o0 = 0; o1 = n;
        return;
    }
    if (t == nt - 1) {
        org = n - w; o0 = n < w + 8 ? 8 : n - w; o1 = n;
        return;
    }
    org = w * t; o0 = w * t;
    o1 = n < w + 8 ? 8 : (w * t + w < n - w ? w * t + w : n - w);
}

template <bool VTI>
struct FusedSmem {
    static constexpr int NB = VTI ? 3 : 1;
    float u[4][3][FPX][FPZ];  // ring slot (plane & 3), component, x row (from x0-2), z column (from z0-2)
    // buffers of parity o & 1 for plane o:
    float b[2][NB][FPX][FPZ];  // D2 coefficient fields: mu [, c66, c11b]
    float tx[2][FPX][FZ];      // T (elastic) / T_x (VTI), own and x-halo rows
    float tz[2][FW][FPZ];      // T / T_z, own and z-halo columns
    float fxy[2][FPX][FZ], fxz[2][FPX][FZ];
    float fzx[2][FW][FPZ], fzy[2][FW][FPZ];
};

__device__ __forceinline__ float d1c(float m2, float m1, float p1, float p2) {
    return D1C1 * (p1 - m1) + D1C2 * (p2 - m2);
}

// Closure rows with several outputs per row switch (edge tiles). The operands are loaded
// once into registers (u / b at the 8 points from the face for D2(b), the 6 points for
// D1), and a lane on closure row r then runs one switch case for all its outputs. D2(b)
// row r forms c_j = sum_k w_rjk b_k once per u point j (outputs that share a coefficient
// field share these sums) and then sum_j c_j u_j per output.
__host__ __device__ constexpr bool d2_uses(int r, int j) {
    for (int e = 0; e < 132; e++)
        if (kD2Closure[e].r == r && kD2Closure[e].j == j) return true;
    return false;
}
template <int R, int J, int E>
__device__ __forceinline__ float d2_cj(float acc, const float (&b)[8]) {
    if constexpr (E == 132) return acc;
    else if constexpr (kD2Closure[E].r == R && kD2Closure[E].j == J)
        return d2_cj<R, J, E + 1>(acc + D2Term<E>::w * b[kD2Closure[E].k], b);
    else return d2_cj<R, J, E + 1>(acc, b);
}
template <int R, int J, int NC>
__device__ __forceinline__ void d2_row_js(const float (&b)[NC][8], const float (&u)[NC][8], float (&out)[NC]) {
    if constexpr (J < 8) {
        if constexpr (d2_uses(R, J)) {
#pragma unroll
            for (int c = 0; c < NC; c++) out[c] += d2_cj<R, J, 0>(-0.0f, b[c]) * u[c][J];
        }
        d2_row_js<R, J + 1, NC>(b, u, out);
    }
}
// D2(b) (times h^2) of closure row r in [0, 6) for NC outputs.
template <int NC>
__device__ __forceinline__ void d2_rows_n(int r, const float (&b)[NC][8], const float (&u)[NC][8], float (&out)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = -0.0f;
    switch (r) {
        case 0: d2_row_js<0, 0, NC>(b, u, out); break;
        case 1: d2_row_js<1, 0, NC>(b, u, out); break;
        case 2: d2_row_js<2, 0, NC>(b, u, out); break;
        case 3: d2_row_js<3, 0, NC>(b, u, out); break;
        case 4: d2_row_js<4, 0, NC>(b, u, out); break;
        default: d2_row_js<5, 0, NC>(b, u, out); break;
    }
}
// D1 (times h) of closure row r in [0, 4) for NC outputs; lo as for d1_closure_at.
template <int NC>
__device__ __forceinline__ void d1_rows_n(int r, bool lo, const float (&u)[NC][6], float (&out)[NC]) {
    switch (r) {
        case 0:
#pragma unroll
            for (int c = 0; c < NC; c++) out[c] = d1_row<0>([&](int j) { return u[c][j]; });
            break;
        case 1:
#pragma unroll
            for (int c = 0; c < NC; c++) out[c] = d1_row<1>([&](int j) { return u[c][j]; });
            break;
        case 2:
#pragma unroll
            for (int c = 0; c < NC; c++) out[c] = d1_row<2>([&](int j) { return u[c][j]; });
            break;
        default:
#pragma unroll
            for (int c = 0; c < NC; c++) out[c] = d1_row<3>([&](int j) { return u[c][j]; });
            break;
    }
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = lo ? out[c] : -out[c];
}

// Tile classes: the tiles away from the x and z faces (the middle tiles of both axes,
// whose outputs, and the fluxes and stencils they read, are at least 14 points inside
// those faces) run a body without closure rows and face terms; the others (edge tiles)
// the full body. Edge tiles are the corner tiles (first or last tile along both x and
// z: closure rows of both axes) and the side tiles (the other tiles of the first and last
// x columns of tiles, then those of the first and last z tiles).
__host__ __device__ inline int fused_middle_tiles(int ntz, int ntx) {
    return (ntz < 3 || ntx < 3) ? 0 : (ntz - 2) * (ntx - 2);
}
__host__ __device__ inline int fused_corner_tiles(int ntz, int ntx) {
    return (ntz == 1 ? 1 : 2) * (ntx == 1 ? 1 : 2);
}
__host__ __device__ inline int fused_side_tiles(int ntz, int ntx) {
    return (ntx == 1 ? 1 : 2) * (ntz > 2 ? ntz - 2 : 0) + (ntz == 1 ? 1 : 2) * (ntx > 2 ? ntx - 2 : 0);
}
// Corner tile c: (tz, tx) packed as tz | tx << 16.
__device__ __forceinline__ int fused_corner_tile(int c, int ntz, int ntx) {
    const int nzc = ntz == 1 ? 1 : 2;
    return ((c % nzc) == 0 ? 0 : ntz - 1) | (((c / nzc) == 0 ? 0 : ntx - 1) << 16);
}
// Side tile e: the x-edge columns' tiles between the corners, then the z-edge rows'.
__device__ __forceinline__ int fused_side_tile(int e, int ntz, int ntx) {
    const int mz = ntz > 2 ? ntz - 2 : 0, nxs = (ntx == 1 ? 1 : 2) * mz;
    const int mz1 = mz > 0 ? mz : 1, mx = ntx > 2 ? ntx - 2 : 1, f = e - nxs;
    const int tz = e < nxs ? 1 + e % mz1 : (f / mx == 0 ? 0 : ntz - 1);
    const int tx = e < nxs ? (e / mz1 == 0 ? 0 : ntx - 1) : 1 + f % mx;
    return tz | (tx << 16);
}

// One workgroup: tile (tzi, txi), output planes [ya, yb). XE / ZE: the tile may hold x / z
// closure rows and face points (false only for middle tiles).
template <bool VTI, bool XE, bool ZE>
__device__ __forceinline__ void elastic_fused_tile(const EArgs& __restrict__ A, const FusedTiles& FT, int tzi, int txi,
                                                   int ya, int yb, FusedSmem<VTI>& S) {
    constexpr int NB = VTI ? 3 : 1;
    const Geom& g = A.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz, sx = g.sx, sy = g.sy, n3 = g.n;
    const float ihx = 1.0f / g.dx, ihy = 1.0f / g.dy, ihz = 1.0f / g.dz;
    const float idt = 1.0f / A.dt;
    int z0, zo0, zo1, x0, xo0, xo1;
    fused_tile(nz, FZ, tzi, FT.ntz, z0, zo0, zo1);
    fused_tile(nx, FW, txi, FT.ntx, x0, xo0, xo1);
    const int t = threadIdx.x;
    // Thread -> point of the tile. A z-edge tile holding the closure rows of one z face puts
    // each of those 6 z rows (x 16 x rows) on a 16-lane group, so a row's closures run
    // row-uniformly: rows 1-3 from the face on the first wave, rows 4-5 on the second, the
    // face row (traction SAT) on the third; the other 26 z rows x 16 x rows fill the other
    // groups. Every other tile runs 32 z lanes per x row (a wave holds two x rows).
    const int zface = !ZE ? 0 : ((zo0 == 0) == (zo1 == nz) ? 0 : (zo0 == 0 ? 1 : 2));  // 1 low, 2 high
    int lz, lx;
    if (zface == 0) {
        lz = t & (FZ - 1);
        lx = t >> 5;
    } else {
        const int grp = t >> 4;
        const int crow = grp < 3 ? grp + 1 : (grp == 4 ? 4 : (grp == 5 ? 5 : (grp == 8 ? 0 : -1)));
        if (crow >= 0) {
            lx = t & (FW - 1);
            lz = zface == 1 ? crow : FZ - 1 - crow;
        } else {
            // interior groups before this one (the closure groups are 0-2, 4, 5 and 8)
            const int nb = grp - min(grp, 3) - (grp > 4) - (grp > 5) - (grp > 8);
            const int u = nb * 16 + (t & 15);
            lx = u / (FZ - 6);
            lz = zface == 1 ? 6 + u % (FZ - 6) : u % (FZ - 6);
        }
    }
    const int ix = x0 + lx, iz = z0 + lz;
    const int cz = min(iz, nz - 1);
    const int col = min(ix, nx - 1) * sx + cz;
    const bool outp = ix >= xo0 && ix < xo1 && iz >= zo0 && iz < zo1;
    const float* __restrict__ curr = A.curr;
    const float* __restrict__ prev = A.prev;
    const float* __restrict__ lamf = A.lam;
    const float* __restrict__ muf = A.mu;
    const float* __restrict__ rhof = A.rho;
    const float* bfld[3] = {A.mu, VTI ? A.c66 : A.mu, VTI ? A.c11b : A.mu};
    auto clx = [&](int x) { return max(0, min(nx - 1, x)); };
    auto clz = [&](int z) { return max(0, min(nz - 1, z)); };
    auto ld = [&](const float* __restrict__ f, int y, int off) { return gld(f, (unsigned)(max(0, min(ny - 1, y)) * sy + off)); };

    // x / z closure geometry of the own point (rows counted from the face; LDS row /
    // column of the face point, clamped into the staged plane)
    const bool xlo = ix < nx - 1 - ix, zlo = iz < nz - 1 - iz;
    const int xr = xlo ? ix : nx - 1 - ix, zr = zlo ? iz : nz - 1 - iz;
    const int xs = xlo ? 1 : -1, zs = zlo ? 1 : -1;
    const int xb = max(0, min(FPX - 1, (xlo ? 0 : nx - 1) - x0 + 2));
    const int zb = max(0, min(FPZ - 1, (zlo ? 0 : nz - 1) - z0 + 2));
    auto xrow = [&](int k) { return max(0, min(FPX - 1, xb + k * xs)); };
    auto zcol = [&](int k) { return max(0, min(FPZ - 1, zb + k * zs)); };

    // halo role: 1 x-halo point, 2 z-halo point, 3 corner (staging only), 0 none. The
    // roles go to threads th < 208; th = t except in tiles whose closure rows sit on the
    // first waves (z-edge tiles, tiles at the low x face), where th = t + 256 (mod 512)
    // moves the halo work to the last waves.
    const bool rshift = zface != 0 || (XE && x0 == 0 && xo1 < nx);
    const int th = rshift ? (t + 256) & (FTHREADS - 1) : t;
    const int role = th < 128 ? 1 : (th < 192 ? 2 : (th < 208 ? 3 : 0));
    int hrow = 0, hcol = 0, hoff = col, eoff = col, hx = ix, hz = iz;
    int ca = 0, cb = 1, cc = 2;  // halo components: queued a, queued b, staged-only c
    // an x-halo point's z lane (its own, whatever the thread's own point)
    const int hzl = th & (FZ - 1), hzc = clz(z0 + hzl);
    if (role == 1) {
        const int hr = th >> 5;
        hrow = hr < 2 ? hr : FW + hr;
        hcol = hzl + 2;
        hx = hr < 2 ? x0 - 2 + hr : x0 + FW + hr - 2;
        hz = z0 + hzl;
        hoff = clx(hx) * sx + hzc;
        eoff = clx(hr < 2 ? x0 - 4 + hr : x0 + FW + hr) * sx + hzc;
    } else if (role == 2) {
        const int h = th - 128, zrw = h >> 2, c4 = h & 3;
        hrow = zrw + 2;
        hcol = c4 < 2 ? c4 : FZ + c4;
        hx = x0 + zrw;
        hz = c4 < 2 ? z0 - 2 + c4 : z0 + FZ + c4 - 2;
        hoff = clx(hx) * sx + clz(hz);
        eoff = clx(hx) * sx + clz(c4 < 2 ? z0 - 4 + c4 : z0 + FZ + c4);
        ca = 1; cb = 2; cc = 0;
    } else if (role == 3) {
        const int h = th - 192, cr = h >> 2, c4 = h & 3;
        hrow = cr < 2 ? cr : FW + cr;
        hcol = c4 < 2 ? c4 : FZ + c4;
        hx = cr < 2 ? x0 - 2 + cr : x0 + FW + cr - 2;
        hz = c4 < 2 ? z0 - 2 + c4 : z0 + FZ + c4 - 2;
        hoff = clx(hx) * sx + clz(hz);
    }
    // in-plane offsets of the halo values (component base folded in, so every load has
    // the uniform field pointer as its base): a, b, c and the ext value (u_z for a z-halo
    // point, u_x for an x-halo point)
    const int hoa = ca * n3 + hoff, hob = cb * n3 + hoff, hoc = cc * n3 + hoff;
    const int heo = (role == 2 ? 2 * n3 : 0) + eoff;

    // ---- register queues
    float qx[6], qz[6], qy[7];  // own u_x, u_z at o-2..o+3; u_y at o-2..o+4
    float qb[NB][6];            // own D2 coefficient fields at o-2..o+3
    float lq[3], eq[3], c13q[3];  // own lambda (VTI: eps, c13) at o..o+2
    float tq[5], fyxq[5], fyzq[5];  // T_y, F_yx, F_yz at o-2..o+2
    float txq[3], tzq[3];           // VTI: T_x, T_z at o..o+2
    float rq[4];                    // rho and prev (3) at o
    float ha[6], hb[6], hc;         // halo: components a, b at o-2..o+3; c at o+3
    float hm[2], hl[2];             // halo mu, lambda at o, o+1
    float ve[2][2];                 // halo ext values (one or two) at o, o+1
    float hv[2][4];                 // VTI halo c66, c11b, eps, c13 at o, o+1
    // ext values of a halo point: the D1 stencil of an x-halo point reaches two rows past
    // the staged plane (x0-4, x0-3 / x0+FW+2, x0+FW+3), that of a z-halo point two
    // columns; each halo thread loads the one or two it needs itself
    int eo0 = hoff, eo1 = hoff;
    if (role == 1) {
        const int hr = th >> 5;
        eo0 = clx(hr < 2 ? x0 - 4 + hr : x0 + FW + 2) * sx + hzc;
        eo1 = clx(hr == 0 ? x0 - 3 : x0 + FW + 3) * sx + hzc;
    } else if (role == 2) {
        const int zrw = (th - 128) >> 2, c4 = th & 3;
        eo0 = clx(x0 + zrw) * sx + clz(c4 < 2 ? z0 - 4 + c4 : z0 + FZ + 2) + 2 * n3;
        eo1 = clx(x0 + zrw) * sx + clz(c4 == 0 ? z0 - 3 : z0 + FZ + 3) + 2 * n3;
    }
    const int o0 = ya - 4;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        qx[k] = ld(curr, o0 - 2 + k, col);
        qz[k] = ld(curr + 2 * n3, o0 - 2 + k, col);
#pragma unroll
        for (int m = 0; m < NB; m++) qb[m][k] = ld(bfld[m], o0 - 2 + k, col);
    }
#pragma unroll
    for (int k = 0; k < 6; k++) qy[k] = ld(curr + n3, o0 - 2 + k, col);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        lq[k] = ld(lamf, o0 + k, col);
        eq[k] = VTI ? ld(A.eps, o0 + k, col) : 0.f;
        c13q[k] = VTI ? ld(A.c13, o0 + k, col) : 0.f;
    }
    lq[2] = eq[2] = c13q[2] = 0.f;
    // taper factors of the own point along z and x (the y factor is read per plane)
    const float tfz = gld(A.tz2, cz), tfx = gld(A.tx2, min(ix, nx - 1));
#pragma unroll
    for (int k = 0; k < 5; k++) tq[k] = fyxq[k] = fyzq[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) txq[k] = tzq[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) rq[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 6; k++) ha[k] = hb[k] = 0.f;
    hc = hm[0] = hm[1] = hl[0] = hl[1] = 0.f;
    ve[0][0] = ve[0][1] = ve[1][0] = ve[1][1] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) hv[0][k] = hv[1][k] = 0.f;
    if (role != 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) {
            ha[k] = ld(curr, o0 - 2 + k, hoa);
            hb[k] = ld(curr, o0 - 2 + k, hob);
        }
        hc = ld(curr, o0 + 2, hoc);
        hm[0] = ld(muf, o0, hoff);
        hl[0] = ld(lamf, o0, hoff);
        ve[0][0] = ld(curr, o0, eo0);
        ve[0][1] = ld(curr, o0, eo1);
        if (VTI) {
            hv[0][0] = ld(A.c66, o0, hoff);
            hv[0][1] = ld(A.c11b, o0, hoff);
            hv[0][2] = ld(A.eps, o0, hoff);
            hv[0][3] = ld(A.c13, o0, hoff);
        }
    }
    // plane o0+2 of the ring (the first plane the look-ahead of the loop reads)
    {
        const int sl = (o0 + 2) & 3;
        S.u[sl][0][lx + 2][lz + 2] = qx[4];
        S.u[sl][1][lx + 2][lz + 2] = qy[4];
        S.u[sl][2][lx + 2][lz + 2] = qz[4];
        if (role != 0) {
            S.u[sl][ca][hrow][hcol] = ha[4];
            S.u[sl][cb][hrow][hcol] = hb[4];
            S.u[sl][cc][hrow][hcol] = hc;
        }
    }
    __syncthreads();

    // One barrier per plane. Iteration o: (A) the fluxes differentiated along x and z at
    // plane o (own points and halos) and the D2 coefficient planes of plane o, into the
    // buffers of parity o; barrier; (C) the look-ahead fluxes of plane o+2, the update of
    // plane o, then plane o+3 staged into the ring (4 slots). A plane staged in (C) of
    // iteration o is read after the barrier of iteration o+1; a slot or a parity buffer is
    // rewritten only after a barrier that follows its last read.
    for (int o = o0; o < yb; o++) {
        const bool live = o >= ya;
        const int fb = o & 1;
        // ---- loads: plane o+3 (staged in (C)), u_y at o+4 and lambda at o+2 (look-ahead in
        // (C)), rho and prev at o (update in (C)); the halo values of the next iteration
        qx[5] = ld(curr, o + 3, col);
        qz[5] = ld(curr + 2 * n3, o + 3, col);
        qy[6] = ld(curr + n3, o + 4, col);
#pragma unroll
        for (int m = 0; m < NB; m++) qb[m][5] = ld(bfld[m], o + 3, col);
        lq[2] = ld(lamf, o + 2, col);
        if (VTI) {
            eq[2] = ld(A.eps, o + 2, col);
            c13q[2] = ld(A.c13, o + 2, col);
        }
        float tfy = 0.f;
        if (live) {
            tfy = gld(A.ty2, o);
            rq[0] = ld(rhof, o, col);
#pragma unroll
            for (int c = 0; c < 3; c++) rq[1 + c] = ld(prev + c * n3, o, col);
        }
        if (role != 0) {
            ha[5] = ld(curr, o + 3, hoa);
            hb[5] = ld(curr, o + 3, hob);
            hc = ld(curr, o + 3, hoc);
            hm[1] = ld(muf, o + 1, hoff);
            hl[1] = ld(lamf, o + 1, hoff);
            ve[1][0] = ld(curr, o + 1, eo0);
            ve[1][1] = ld(curr, o + 1, eo1);
            if (VTI) {
                hv[1][0] = ld(A.c66, o + 1, hoff);
                hv[1][1] = ld(A.c11b, o + 1, hoff);
                hv[1][2] = ld(A.eps, o + 1, hoff);
                hv[1][3] = ld(A.c13, o + 1, hoff);
            }
        }
        auto U = [&](int c, int r, int z) { return S.u[o & 3][c][r][z]; };
        auto Bf = [&](int m, int r, int z) { return S.b[fb][m][r][z]; };
        if (live) {
            // ---- (A) coefficient planes and the fluxes differentiated along x and z, plane o
#pragma unroll
            for (int m = 0; m < NB; m++) S.b[fb][m][lx + 2][lz + 2] = qb[m][2];
            if (role == 1 || role == 2) {
                S.b[fb][0][hrow][hcol] = hm[0];
                if (VTI) {
                    S.b[fb][1][hrow][hcol] = hv[0][0];
                    S.b[fb][2][hrow][hcol] = hv[0][1];
                }
            }
            {
                const float mu = qb[0][2];
                const float gxy = VTI ? qb[1][2] : mu;
                const float dyu0 = d1c(qx[0], qx[1], qx[3], qx[4]) * ihy;
                const float dyu2 = d1c(qz[0], qz[1], qz[3], qz[4]) * ihy;
                float dxu2, dzu0;
                if (!XE || xr >= 4) dxu2 = d1c(U(2, lx, lz + 2), U(2, lx + 1, lz + 2), U(2, lx + 3, lz + 2), U(2, lx + 4, lz + 2));
                else dxu2 = d1_closure_at(xlo, xr, [&](int k) { return U(2, xrow(k), lz + 2); });
                if (!ZE || zr >= 4) dzu0 = d1c(U(0, lx + 2, lz), U(0, lx + 2, lz + 1), U(0, lx + 2, lz + 3), U(0, lx + 2, lz + 4));
                else dzu0 = d1_closure_at(zlo, zr, [&](int k) { return U(0, lx + 2, zcol(k)); });
                // (the look-ahead of this iteration has not run yet: the queues end at o+1)
                S.tx[fb][lx + 2][lz] = VTI ? txq[1] : tq[3];
                S.tz[fb][lx][lz + 2] = VTI ? tzq[1] : tq[3];
                S.fxy[fb][lx + 2][lz] = dyu0 * gxy;
                S.fxz[fb][lx + 2][lz] = (dzu0 * ihz) * mu;
                S.fzx[fb][lx][lz + 2] = (dxu2 * ihx) * mu;
                S.fzy[fb][lx][lz + 2] = dyu2 * mu;
            }
            if (role == 1) {
                // x-halo point (hx, iz): the D1 along x reaches the ext values, the D1 along
                // z may use z closures
                const int hr = th >> 5;
                const float r0 = U(0, hrow - 2 < 0 ? 0 : hrow - 2, hzl + 2), r1 = U(0, hrow - 1 < 0 ? 0 : hrow - 1, hzl + 2);
                const float r3 = U(0, hrow + 1 >= FPX ? FPX - 1 : hrow + 1, hzl + 2);
                const float r4 = U(0, hrow + 2 >= FPX ? FPX - 1 : hrow + 2, hzl + 2);
                const float m2 = hr < 2 ? ve[0][0] : r0;
                const float m1 = hr == 0 ? ve[0][1] : r1;
                const float p1 = hr == 3 ? ve[0][0] : r3;
                const float p2 = hr == 3 ? ve[0][1] : (hr == 2 ? ve[0][0] : r4);
                const float dxu0 = d1c(m2, m1, p1, p2) * ihx;
                // z closure geometry of the halo point (z0 + hzl)
                const bool hzlo = hz < nz - 1 - hz;
                const int hzr = hzlo ? hz : nz - 1 - hz;
                const int hzb = max(0, min(FPZ - 1, (hzlo ? 0 : nz - 1) - z0 + 2));
                auto hzcol = [&](int k) { return max(0, min(FPZ - 1, hzb + k * (hzlo ? 1 : -1))); };
                float dzu2, dzu0;
                if (!ZE || hzr >= 4) {
                    dzu2 = d1c(U(2, hrow, hzl), U(2, hrow, hzl + 1), U(2, hrow, hzl + 3), U(2, hrow, hzl + 4));
                    dzu0 = d1c(U(0, hrow, hzl), U(0, hrow, hzl + 1), U(0, hrow, hzl + 3), U(0, hrow, hzl + 4));
                } else {
                    float uu[2][6], o[2];
#pragma unroll
                    for (int k = 0; k < 6; k++) {
                        const int zc = hzcol(k);
                        uu[0][k] = U(2, hrow, zc);
                        uu[1][k] = U(0, hrow, zc);
                    }
                    d1_rows_n<2>(hzr, hzlo, uu, o);
                    dzu2 = o[0];
                    dzu0 = o[1];
                }
                dzu2 *= ihz;
                dzu0 *= ihz;
                const float dyu0 = d1c(ha[0], ha[1], ha[3], ha[4]) * ihy;  // a = u_x
                const float dyu1 = d1c(hb[0], hb[1], hb[3], hb[4]) * ihy;  // b = u_y
                const float mu = hm[0], lam = hl[0];
                float Tx, gxy = mu;
                if (!VTI) {
                    Tx = lam * (dxu0 + dyu1 + dzu2);
                } else {
                    const float c66 = hv[0][0], e2 = 1.f + 2.f * hv[0][2], c13 = hv[0][3];
                    const float c11c = e2 * lam, c12 = e2 * (lam + 2.f * mu) - 2.f * c66;
                    Tx = c11c * dxu0 + c12 * dyu1 + c13 * dzu2;
                    gxy = c66;
                }
                S.tx[fb][hrow][hzl] = Tx;
                S.fxy[fb][hrow][hzl] = dyu0 * gxy;
                S.fxz[fb][hrow][hzl] = dzu0 * mu;
            } else if (role == 2) {
                // z-halo point (hx, hz): the D1 along z reaches the ext values, the D1 along x
                // may use x closures
                const int zrw = hrow - 2, c4 = th & 3;
                const int hxr = hx < nx - 1 - hx ? hx : nx - 1 - hx;
                const bool hxlo = hx < nx - 1 - hx;
                const int hxb = max(0, min(FPX - 1, (hxlo ? 0 : nx - 1) - x0 + 2));
                auto hxrow = [&](int k) { return max(0, min(FPX - 1, hxb + k * (hxlo ? 1 : -1))); };
                const float c0v = U(2, hrow, hcol - 2 < 0 ? 0 : hcol - 2), c1v = U(2, hrow, hcol - 1 < 0 ? 0 : hcol - 1);
                const float c3v = U(2, hrow, hcol + 1 >= FPZ ? FPZ - 1 : hcol + 1);
                const float c4v = U(2, hrow, hcol + 2 >= FPZ ? FPZ - 1 : hcol + 2);
                const float m2 = c4 < 2 ? ve[0][0] : c0v;
                const float m1 = c4 == 0 ? ve[0][1] : c1v;
                const float p1 = c4 == 3 ? ve[0][0] : c3v;
                const float p2 = c4 == 3 ? ve[0][1] : (c4 == 2 ? ve[0][0] : c4v);
                const float dzu2 = d1c(m2, m1, p1, p2) * ihz;
                float dxu0, dxu2;
                if (!XE || hxr >= 4) {
                    dxu0 = d1c(U(0, hrow - 2, hcol), U(0, hrow - 1, hcol), U(0, hrow + 1, hcol), U(0, hrow + 2, hcol));
                    dxu2 = d1c(U(2, hrow - 2, hcol), U(2, hrow - 1, hcol), U(2, hrow + 1, hcol), U(2, hrow + 2, hcol));
                } else {
                    float uu[2][6], o[2];
#pragma unroll
                    for (int k = 0; k < 6; k++) {
                        const int xc = hxrow(k);
                        uu[0][k] = U(0, xc, hcol);
                        uu[1][k] = U(2, xc, hcol);
                    }
                    d1_rows_n<2>(hxr, hxlo, uu, o);
                    dxu0 = o[0];
                    dxu2 = o[1];
                }
                dxu0 *= ihx;
                dxu2 *= ihx;
                const float dyu1 = d1c(ha[0], ha[1], ha[3], ha[4]) * ihy;  // a = u_y
                const float dyu2 = d1c(hb[0], hb[1], hb[3], hb[4]) * ihy;  // b = u_z
                const float mu = hm[0], lam = hl[0];
                float Tz;
                if (!VTI) Tz = lam * (dxu0 + dyu1 + dzu2);
                else Tz = hv[0][3] * (dxu0 + dyu1) + lam * dzu2;
                S.tz[fb][zrw][hcol] = Tz;
                S.fzx[fb][zrw][hcol] = dxu2 * mu;
                S.fzy[fb][zrw][hcol] = dyu2 * mu;
            }
        }
        __syncthreads();

        // ---- (C) look-ahead fluxes of plane o+2 (staged in iteration o-1), pushed before the
        // update of plane o reads the queues
        {
            auto UL = [&](int c, int r, int z) { return S.u[(o + 2) & 3][c][r][z]; };
            // D_x of u_x, u_y and D_z of u_y, u_z (times h) at the own point
            float ox2[2], oz2[2];
            if (!XE || xr >= 4) {
#pragma unroll
                for (int c = 0; c < 2; c++)
                    ox2[c] = d1c(UL(c, lx, lz + 2), UL(c, lx + 1, lz + 2), UL(c, lx + 3, lz + 2), UL(c, lx + 4, lz + 2));
            } else {
                float uu[2][6];
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    const int xc = xrow(k);
                    uu[0][k] = UL(0, xc, lz + 2);
                    uu[1][k] = UL(1, xc, lz + 2);
                }
                d1_rows_n<2>(xr, xlo, uu, ox2);
            }
            if (!ZE || zr >= 4) {
#pragma unroll
                for (int c = 0; c < 2; c++)
                    oz2[c] = d1c(UL(c + 1, lx + 2, lz), UL(c + 1, lx + 2, lz + 1), UL(c + 1, lx + 2, lz + 3), UL(c + 1, lx + 2, lz + 4));
            } else {
                float uu[2][6];
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    const int zc = zcol(k);
                    uu[0][k] = UL(1, lx + 2, zc);
                    uu[1][k] = UL(2, lx + 2, zc);
                }
                d1_rows_n<2>(zr, zlo, uu, oz2);
            }
            const float dxu0 = ox2[0] * ihx, dxu1 = ox2[1] * ihx;
            const float dzu1 = oz2[0] * ihz, dzu2 = oz2[1] * ihz;
            const float dyu1 = d1c(qy[2], qy[3], qy[5], qy[6]) * ihy;
            const float mu = qb[0][4];
            const float lam = lq[2];
            float Ty;
            if (!VTI) {
                Ty = lam * (dxu0 + dyu1 + dzu2);
            } else {
                const float e2 = 1.f + 2.f * eq[2], c13 = c13q[2], c66 = qb[1][4];
                const float c11c = e2 * lam, c12 = e2 * (lam + 2.f * mu) - 2.f * c66;
                Ty = c11c * dyu1 + c12 * dxu0 + c13 * dzu2;
                txq[0] = txq[1]; txq[1] = txq[2];
                tzq[0] = tzq[1]; tzq[1] = tzq[2];
                txq[2] = c11c * dxu0 + c12 * dyu1 + c13 * dzu2;
                tzq[2] = c13 * (dxu0 + dyu1) + lam * dzu2;
            }
            const float gxy = VTI ? qb[1][4] : mu;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                tq[k] = tq[k + 1]; fyxq[k] = fyxq[k + 1]; fyzq[k] = fyzq[k + 1];
            }
            tq[4] = Ty;
            fyxq[4] = dxu1 * gxy;
            fyzq[4] = dzu1 * mu;
        }


        if (live) {
            // ---- update of plane o
            if (outp) {
                const float ih2x = ihx * ihx, ih2y = ihy * ihy, ih2z = ihz * ihz;
                const int p = o * sy + ix * sx + iz;
                // D2(b) along x and z of the three components (times h^2); along z the
                // coefficient field is mu for every component
                float d2x[3], d2z[3];
                if (!XE || xr >= 6) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const int m = d2_bfield<VTI>(c, 0);
                        d2x[c] = d2_interior(U(c, lx, lz + 2), U(c, lx + 1, lz + 2), U(c, lx + 2, lz + 2), U(c, lx + 3, lz + 2),
                                             U(c, lx + 4, lz + 2), Bf(m, lx, lz + 2), Bf(m, lx + 1, lz + 2), Bf(m, lx + 2, lz + 2),
                                             Bf(m, lx + 3, lz + 2), Bf(m, lx + 4, lz + 2));
                    }
                } else {
                    float uu[3][8], bb[3][8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const int xc = xrow(k);
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            uu[c][k] = U(c, xc, lz + 2);
                            bb[c][k] = Bf(d2_bfield<VTI>(c, 0), xc, lz + 2);
                        }
                    }
                    d2_rows_n<3>(xr, bb, uu, d2x);
                }
                if (!ZE || zr >= 6) {
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        d2z[c] = d2_interior(U(c, lx + 2, lz), U(c, lx + 2, lz + 1), U(c, lx + 2, lz + 2), U(c, lx + 2, lz + 3),
                                             U(c, lx + 2, lz + 4), Bf(0, lx + 2, lz), Bf(0, lx + 2, lz + 1), Bf(0, lx + 2, lz + 2),
                                             Bf(0, lx + 2, lz + 3), Bf(0, lx + 2, lz + 4));
                } else {
                    float uu[3][8], bb[3][8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const int zc = zcol(k);
                        const float bk = Bf(0, lx + 2, zc);
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            uu[c][k] = U(c, lx + 2, zc);
                            bb[c][k] = bk;
                        }
                    }
                    d2_rows_n<3>(zr, bb, uu, d2z);
                }
                auto qc = [&](int c, int k) { return c == 0 ? qx[k] : (c == 1 ? qy[k] : qz[k]); };
                auto d2y = [&](int c, int m) {
                    return d2_interior(qc(c, 0), qc(c, 1), qc(c, 2), qc(c, 3), qc(c, 4), qb[m][0], qb[m][1], qb[m][2],
                                       qb[m][3], qb[m][4]);
                };
                // outer D1 (times h) at the own point of the staged flux planes: along x
                // T / T_x, F_xy, F_xz (components 0, 1, 2), along z F_zx, F_zy, T / T_z
                float ox[3], oz[3];
                {
                    const float(*PX[3])[FZ] = {S.tx[fb], S.fxy[fb], S.fxz[fb]};
                    const float(*PZ[3])[FPZ] = {S.fzx[fb], S.fzy[fb], S.tz[fb]};
                    if (!XE || xr >= 4) {
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            ox[c] = d1c(PX[c][lx][lz], PX[c][lx + 1][lz], PX[c][lx + 3][lz], PX[c][lx + 4][lz]);
                    } else {
                        float ff[3][6];
#pragma unroll
                        for (int k = 0; k < 6; k++) {
                            const int xc = xrow(k);
#pragma unroll
                            for (int c = 0; c < 3; c++) ff[c][k] = PX[c][xc][lz];
                        }
                        d1_rows_n<3>(xr, xlo, ff, ox);
                    }
                    if (!ZE || zr >= 4) {
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            oz[c] = d1c(PZ[c][lx][lz], PZ[c][lx][lz + 1], PZ[c][lx][lz + 3], PZ[c][lx][lz + 4]);
                    } else {
                        float ff[3][6];
#pragma unroll
                        for (int k = 0; k < 6; k++) {
                            const int zc = zcol(k);
#pragma unroll
                            for (int c = 0; c < 3; c++) ff[c][k] = PZ[c][lx][zc];
                        }
                        d1_rows_n<3>(zr, zlo, ff, oz);
                    }
                }
                auto oy = [&](const float* q) { return d1c(q[0], q[1], q[3], q[4]); };
                float acc[3], pv[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float a = (d2_factor<VTI>(c, 0) * ih2x) * d2x[c];
                    a = a + (d2_factor<VTI>(c, 1) * ih2y) * d2y(c, d2_bfield<VTI>(c, 1));
                    a = a + (d2_factor<VTI>(c, 2) * ih2z) * d2z[c];
                    if (c == 0) {
                        a = a + ox[0] * ihx;
                        a = a + oy(fyxq) * ihy;  // D_y F_yx
                        a = a + oz[0] * ihy;     // D_z F_zx (spacing dy as in the reference)
                    } else if (c == 1) {
                        a = a + oy(tq) * ihy;    // D_y T_y
                        a = a + ox[1] * ihx;     // D_x F_xy
                        a = a + oz[1] * ihz;     // D_z F_zy
                    } else {
                        a = a + oz[2] * ihz;
                        a = a + ox[2] * ihx;     // D_x F_xz
                        a = a + oy(fyzq) * ihy;  // D_y F_yz
                    }
                    acc[c] = a;
                    pv[c] = rq[1 + c];
                    if (A.prev_tapered) pv[c] = ((pv[c] * tfz) * tfx) * tfy;  // taper_apply's order
                }
                const float rho = rq[0], irho = 1.0f / rho;
                float sc[3] = {0.f, 0.f, 0.f};
                bool absorb = false;
                if ((XE && (ix == 0 || ix == nx - 1)) || (ZE && (iz == 0 || iz == nz - 1))) {
                    // traction SATs of the x / z faces through the point, in the reference's
                    // face order (top, bottom, left, right), and the absorbing scaling. D1 of u
                    // at the point (divided by h) and the outward S along x / z (times h):
                    float dX[3], dY[3], dZ[3], sX[3], sZ[3];
                    {
                        // the 6 values from the x face and from the z face of each component
                        // (only along the axes whose faces the tile may hold)
                        float ux[3][6], uz[3][6];
#pragma unroll
                        for (int k = 0; k < 6; k++) {
                            const int xc = xrow(k), zc = zcol(k);
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                ux[c][k] = XE ? U(c, xc, lz + 2) : 0.f;
                                uz[c][k] = ZE ? U(c, lx + 2, zc) : 0.f;
                            }
                        }
                        if (!XE || xr >= 4) {
#pragma unroll
                            for (int c = 0; c < 3; c++)
                                dX[c] = d1c(U(c, lx, lz + 2), U(c, lx + 1, lz + 2), U(c, lx + 3, lz + 2), U(c, lx + 4, lz + 2));
                        } else {
                            d1_rows_n<3>(xr, xlo, ux, dX);
                        }
                        if (!ZE || zr >= 4) {
#pragma unroll
                            for (int c = 0; c < 3; c++)
                                dZ[c] = d1c(U(c, lx + 2, lz), U(c, lx + 2, lz + 1), U(c, lx + 2, lz + 3), U(c, lx + 2, lz + 4));
                        } else {
                            d1_rows_n<3>(zr, zlo, uz, dZ);
                        }
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            dX[c] *= ihx;
                            dZ[c] *= ihz;
                            dY[c] = d1c(qc(c, 0), qc(c, 1), qc(c, 3), qc(c, 4)) * ihy;
                            sX[c] = c_sat_s[0] * ux[c][0] + c_sat_s[1] * ux[c][1] + c_sat_s[2] * ux[c][2] + c_sat_s[3] * ux[c][3];
                            sZ[c] = c_sat_s[0] * uz[c][0] + c_sat_s[1] * uz[c][1] + c_sat_s[2] * uz[c][2] + c_sat_s[3] * uz[c][3];
                        }
                    }
                    // the model at the point, from the queues (plane o)
                    const float lam = lq[0], mu = qb[0][2];
                    float e2 = 1.f, c13 = 0.f, c66 = mu;
                    if (VTI) {
                        e2 = 1.f + 2.f * eq[0];
                        c13 = c13q[0];
                        c66 = qb[1][2];
                    }
                    const int bcz0 = A.bc[0], bcz1 = A.bc[1], bcx0 = A.bc[2], bcx1 = A.bc[3];
                    for (int f = 0; f < 4; f++) {
                        if (f < 2 ? !ZE : !XE) continue;  // faces the tile cannot hold
                        const int bc = f == 0 ? bcz0 : (f == 1 ? bcz1 : (f == 2 ? bcx0 : bcx1));
                        if (bc != 1 && bc != 2) continue;
                        if (!(f == 0 ? iz == 0 : (f == 1 ? iz == nz - 1 : (f == 2 ? ix == 0 : ix == nx - 1)))) continue;
                        const bool zn = f < 2;  // normal axis z (else x)
                        const float sg = (f & 1) ? 1.f : -1.f;
                        const float hn = zn ? g.dz : g.dx;
                        const float adh = 1.0f / (hn * SAT_H0);
                        const float l2m = lam + 2.f * mu;
                        // normal component (z for a z face, x for an x face): the normal traction;
                        // z face: x, y, z terms; x face: y, z, x terms (the reference's order)
                        const int cn = zn ? 2 : 0;
                        float tang;
                        if (zn) tang = (VTI ? c13 : lam) * dX[0] + (VTI ? c13 : lam) * dY[1] + lam * dZ[2];
                        else tang = (VTI ? e2 * l2m - 2.f * c66 : lam) * dY[1] + (VTI ? c13 : lam) * dZ[2] + (VTI ? e2 * lam : lam) * dX[0];
                        const float knn = (VTI && !zn) ? 2.f * e2 * mu : 2.f * mu;
                        const float ihn = zn ? ihz : ihx;
                        float tn = sg * tang + knn * (zn ? sZ[cn] : sX[cn]) * ihn;
                        if (bc == 2) tn -= ((VTI && !zn) ? 0.5f * sqrtf(e2 * l2m * rho) : 0.5f * sqrtf(l2m * rho)) * pv[cn] * idt;
                        acc[cn] -= adh * tn;
                        // tangential components
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            if (c == cn) continue;
                            const float G = (VTI && !zn && c == 1) ? c66 : mu;
                            const float dn = c == 0 ? dX[cn] : (c == 1 ? dY[cn] : dZ[cn]);  // D_c of u_cn
                            float tt = sg * G * dn + G * (zn ? sZ[c] : sX[c]) * ihn;
                            if (bc == 2) tt -= 0.5f * sqrtf(mu * rho) * pv[c] * idt;
                            acc[c] -= adh * tt;
                        }
                        if (bc == 2) {
                            absorb = true;
                            const float F = (VTI && !zn) ? e2 : 1.f;
                            const float vp = sqrtf(F * l2m * irho), vs = sqrtf(mu * irho);
                            const float k = A.dt / (2.f * hn * SAT_H0);
#pragma unroll
                            for (int c = 0; c < 3; c++) sc[c] += (c == cn ? vp : vs) * k;
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float nv = A.dt2 * acc[c] * irho + 2.f * qc(c, 2) - pv[c];
                    if (absorb) nv = nv / (1.f + sc[c]);
                    gst(A.next, c * n3 + p, ((nv * tfz) * tfx) * tfy);
                }
            }
        }

        // ---- (C) stage plane o+3 into the ring
        {
            const int sl = (o + 3) & 3;
            S.u[sl][0][lx + 2][lz + 2] = qx[5];
            S.u[sl][1][lx + 2][lz + 2] = qy[5];
            S.u[sl][2][lx + 2][lz + 2] = qz[5];
            if (role != 0) {
                S.u[sl][ca][hrow][hcol] = ha[5];
                S.u[sl][cb][hrow][hcol] = hb[5];
                S.u[sl][cc][hrow][hcol] = hc;
            }
        }
        // ---- shift the queues
#pragma unroll
        for (int k = 0; k < 5; k++) {
            qx[k] = qx[k + 1];
            qz[k] = qz[k + 1];
#pragma unroll
            for (int m = 0; m < NB; m++) qb[m][k] = qb[m][k + 1];
            ha[k] = ha[k + 1];
            hb[k] = hb[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) qy[k] = qy[k + 1];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            lq[k] = lq[k + 1];
            eq[k] = eq[k + 1];
            c13q[k] = c13q[k + 1];
        }
        hm[0] = hm[1];
        hl[0] = hl[1];
        ve[0][0] = ve[1][0];
        ve[0][1] = ve[1][1];
#pragma unroll
        for (int k = 0; k < 4; k++) hv[0][k] = hv[1][k];
    }
}

}  // namespace fwi3d
