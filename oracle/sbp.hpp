// ORACLE — test infrastructure only. This is a CPU restatement of the reference
// (nmbader/fwi3d) hot path used to check the HIP implementation. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker, never as the thing measured or shipped.
//
// sbp.hpp: 4th-order summation-by-parts operators on a [y][x][z] grid (z fastest).
// Written per axis instead of per direction: one D1, one D2(b), one taper routine
// parameterised by axis stride.
//
// Follows:
//   D1 / mult_D1 ........ src/cpp/spatial_operators.cpp:179-406
//   D2(b) (Dxx/Dyy/Dzz_var) src/include/spatial_operators.hpp:34-232
//   H norm .............. src/cpp/spatial_operators.cpp:11-177
//   tapers .............. src/cpp/spatial_operators.cpp:408-506
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace orc {

enum Axis { AX_X = 0, AX_Y = 1, AX_Z = 2 };

struct Grid {
    int nx, ny, nz;
    size_t n() const { return (size_t)nx * ny * nz; }
    size_t at(int ix, int iy, int iz) const { return ((size_t)iy * nx + ix) * nz + iz; }
    int len(int a) const { return a == AX_X ? nx : (a == AX_Y ? ny : nz); }
    ptrdiff_t stride(int a) const { return a == AX_X ? nz : (a == AX_Y ? (ptrdiff_t)nx * nz : 1); }
    int coord(int a, int ix, int iy, int iz) const { return a == AX_X ? ix : (a == AX_Y ? iy : iz); }
};

// Diagonal H-norm weights of the (2,4) operator: 17/48, 59/48, 43/48, 49/48 at each end.
inline double hweight(int i, int n) {
    static const double w[4] = {17.0 / 48, 59.0 / 48, 43.0 / 48, 49.0 / 48};
    if (i < 4) return w[i];
    if (i >= n - 4) return w[n - 1 - i];
    return 1.0;
}

// First-derivative closure rows 0..3 over the first 6 points (mirrored and negated at the far end).
static const double D1_BND[4][6] = {
    {-24.0 / 17, 59.0 / 34, -4.0 / 17, -3.0 / 34, 0, 0},
    {-0.5, 0, 0.5, 0, 0, 0},
    {4.0 / 43, -59.0 / 86, 0, 59.0 / 86, -4.0 / 43, 0},
    {3.0 / 98, 0, -59.0 / 98, 0, 32.0 / 49, -4.0 / 49},
};
static const double D1_C1 = 2.0 / 3, D1_C2 = -1.0 / 12;

// Outward one-sided derivative used by the SAT terms: S = (11/6, -3, 3/2, -1/3) / h.
static const double SAT_S[4] = {11.0 / 6, -3.0, 1.5, -1.0 / 3};
static const double SAT_H0 = 17.0 / 48;

// Variable-coefficient second derivative D2(b): interior weights and the
// 6-row boundary closure C[row][u-point j][b-point k] (132 non-zeros).
static const double D2_C[10] = {-3.0 / 4, -5.0 / 6, -1.0 / 24, 1.0 / 6, 1.0 / 2,
                                1.0 / 2, 1.0 / 6, -1.0 / 8, 1.0 / 6, -1.0 / 8};
struct D2Entry { int r, j, k; long long num, den; };
static const D2Entry D2_BND_SPARSE[] = {
    {0,0,0,920,289},
    {0,0,1,-59,68},
    {0,0,2,-7549318,34157643},
    {0,0,3,-17440994,184112825},
    {0,1,0,-1740,289},
    {0,1,2,295314153,366719282},
    {0,1,3,262545878,1218454145},
    {0,2,0,1128,289},
    {0,2,1,59,68},
    {0,2,2,-19250923,26254840},
    {0,2,3,-12192537,324892213},
    {0,3,0,-308,289},
    {0,3,2,42283069,254173229},
    {0,3,3,-43013531,427521546},
    {0,4,2,-18700293,1355757959},
    {0,4,3,18700293,1355757959},
    {0,5,2,-3,833},
    {0,5,3,3,833},
    {1,0,0,12,17},
    {1,0,2,89562243,385991318},
    {1,0,3,16914691,272440171},
    {1,1,0,-59,68},
    {1,1,2,-47979680,48852831},
    {1,1,3,-18034913,120051851},
    {1,2,0,2,17},
    {1,2,2,299262497,373256703},
    {1,2,3,16156647,200473586},
    {1,3,0,3,68},
    {1,3,2,-14723969,177744748},
    {1,3,3,22633571,584543661},
    {1,4,2,46802031,1628311862},
    {1,4,3,-46802031,1628311862},
    {1,5,2,441,181507},
    {1,5,3,-441,181507},
    {2,0,0,-96,731},
    {2,0,1,59,172},
    {2,0,2,-47632751,164317197},
    {2,0,3,-5570723,375470930},
    {2,1,0,118,731},
    {2,1,2,54282598,49343777},
    {2,1,3,23802793,215253532},
    {2,2,0,-16,731},
    {2,2,1,-59,172},
    {2,2,2,-39119273,25083370},
    {2,2,3,-35971870,61324629},
    {2,2,4,-26254,557679},
    {2,3,0,-6,731},
    {2,3,2,360454121,368940022},
    {2,3,3,17254963,80047776},
    {2,3,4,1500708,7993399},
    {2,4,2,-18024731,79673021},
    {2,4,3,24178273,88099647},
    {2,4,4,-26254,185893},
    {2,5,2,-870707,620833782},
    {2,5,3,960119,1147305747},
    {2,5,4,13564,23980197},
    {3,0,0,-36,833},
    {3,0,2,54741803,948483020},
    {3,0,3,-13602043,389676498},
    {3,1,0,177,3332},
    {3,1,2,-35820026,359121865},
    {3,1,3,24921773,534548210},
    {3,2,0,-6,833},
    {3,2,2,813284372,948584067},
    {3,2,3,30057666,158897885},
    {3,2,4,1500708,9108757},
    {3,3,0,-9,3332},
    {3,3,2,-95056924,90903639},
    {3,3,3,-23417695,47008088},
    {3,3,4,-7476412,9108757},
    {3,3,5,-2,49},
    {3,4,2,23159719,99948527},
    {3,4,3,110687545,265515812},
    {3,4,4,4502124,9108757},
    {3,4,5,8,49},
    {3,5,2,-3671038,2687426923},
    {3,5,3,-1063649,8893843},
    {3,5,4,1473580,9108757},
    {3,5,5,-6,49},
    {4,0,2,-9437957,1931986386},
    {4,0,3,9437957,1931986386},
    {4,1,2,17289851,489388053},
    {4,1,3,-17289851,489388053},
    {4,2,2,-66355919,327412264},
    {4,2,3,24178273,98343792},
    {4,2,4,-564461,4461432},
    {4,3,2,17638343,74566894},
    {4,3,3,103749401,243793650},
    {4,3,4,375177,743572},
    {4,3,5,1,6},
    {4,4,2,-19321801,295845927},
    {4,4,3,-50677283,62943042},
    {4,4,4,-280535,371786},
    {4,4,5,-5,6},
    {4,4,6,-1,24},
    {4,5,2,5130905,5183662662},
    {4,5,3,35039615,213452232},
    {4,5,4,1118749,2230716},
    {4,5,5,1,2},
    {4,5,6,1,6},
    {4,6,4,-1,8},
    {4,6,5,1,6},
    {4,6,6,-1,8},
    {5,0,2,-1,784},
    {5,0,3,1,784},
    {5,1,2,8673,2904112},
    {5,1,3,-8673,2904112},
    {5,2,2,-403062,320810033},
    {5,2,3,960119,1280713392},
    {5,2,4,3391,6692148},
    {5,3,2,-1920494,1377228165},
    {5,3,3,-1063649,8712336},
    {5,3,4,368395,2230716},
    {5,3,5,-1,8},
    {5,4,2,5130905,5183662662},
    {5,4,3,35039615,213452232},
    {5,4,4,1118749,2230716},
    {5,4,5,1,2},
    {5,4,6,1,6},
    {5,5,2,-117380,2351569839},
    {5,5,3,-3290636,80044587},
    {5,5,4,-5580181,6692148},
    {5,5,5,-3,4},
    {5,5,6,-5,6},
    {5,5,7,-1,24},
    {5,6,4,1,6},
    {5,6,5,1,2},
    {5,6,6,1,2},
    {5,6,7,1,6},
    {5,7,5,-1,8},
    {5,7,6,1,6},
    {5,7,7,-1,8},
};

template <class T>
struct D2Table {
    T c[6][8][8];
    D2Table() {
        for (int r = 0; r < 6; r++)
            for (int j = 0; j < 8; j++)
                for (int k = 0; k < 8; k++) c[r][j][k] = 0;
        for (const auto& e : D2_BND_SPARSE) c[e.r][e.j][e.k] = (T)((double)e.num / (double)e.den);
    }
};

// Value of D1 (times h) along axis `a` at grid point p whose axis index is i.
// `u` is the field, `par` (optional) multiplies the field inside the derivative.
template <class T>
inline T d1_at(const T* u, const T* par, size_t p, int i, int n, ptrdiff_t s) {
    auto val = [&](ptrdiff_t off) -> T {
        size_t q = (size_t)((ptrdiff_t)p + off * s);
        return par ? u[q] * par[q] : u[q];
    };
    if (i < 4) {
        T acc = 0;
        for (int j = 0; j < 6; j++) acc += (T)D1_BND[i][j] * val(j - i);
        return acc;
    }
    if (i >= n - 4) {
        int r = n - 1 - i;
        T acc = 0;
        for (int j = 0; j < 6; j++) acc += (T)D1_BND[r][j] * val((n - 1 - j) - i);
        return -acc;
    }
    return (T)D1_C1 * (val(1) - val(-1)) + (T)D1_C2 * (val(2) - val(-2));
}

// out (+)= D1 in / h along axis a (mult_D when par != nullptr).
template <class T>
void d1(int a, bool add, const T* in, T* out, const Grid& g, T h, const T* par = nullptr) {
    const int n = g.len(a);
    const ptrdiff_t s = g.stride(a);
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; iy++)
        for (int ix = 0; ix < g.nx; ix++)
            for (int iz = 0; iz < g.nz; iz++) {
                size_t p = g.at(ix, iy, iz);
                T v = d1_at(in, par, p, g.coord(a, ix, iy, iz), n, s) / h;
                out[p] = add ? out[p] + v : v;
            }
}

// Value of D2(b) (times h^2) along axis a at point p with axis index i.
template <class T>
inline T d2_at(const T* u, const T* b, size_t p, int i, int n, ptrdiff_t s, const D2Table<T>& tb) {
    auto U = [&](ptrdiff_t off) { return u[(size_t)((ptrdiff_t)p + off * s)]; };
    auto B = [&](ptrdiff_t off) { return b[(size_t)((ptrdiff_t)p + off * s)]; };
    if (i < 6 || i >= n - 6) {
        const bool lo = i < 6;
        const int r = lo ? i : n - 1 - i;
        T acc = 0;
        for (int j = 0; j < 8; j++) {
            T w = 0;
            for (int k = 0; k < 8; k++) {
                T c = tb.c[r][j][k];
                if (c != 0) w += c * B(lo ? (k - i) : ((n - 1 - k) - i));
            }
            acc += w * U(lo ? (j - i) : ((n - 1 - j) - i));
        }
        return acc;
    }
    const double* c = D2_C;
    T b0 = B(0), bm1 = B(-1), bp1 = B(1), bm2 = B(-2), bp2 = B(2);
    return ((T)c[0] * b0 + (T)c[1] * (bm1 + bp1) + (T)c[2] * (bm2 + bp2)) * U(0)
         + ((T)c[3] * bm2 + (T)c[4] * bm1 + (T)c[5] * b0 + (T)c[6] * bp1) * U(-1)
         + ((T)c[3] * bp2 + (T)c[4] * bp1 + (T)c[5] * b0 + (T)c[6] * bm1) * U(1)
         + ((T)c[7] * bm2 + (T)c[8] * bm1 + (T)c[9] * b0) * U(-2)
         + ((T)c[7] * bp2 + (T)c[8] * bp1 + (T)c[9] * b0) * U(2);
}

// out (+)= a/h^2 D2(b) in along axis a_ (requires at least 12 points on the axis).
template <class T>
void d2(int a_, bool add, const T* in, T* out, const T* b, const Grid& g, T h, T a) {
    static const D2Table<T> tb;
    const int n = g.len(a_);
    if (n < 12) throw std::runtime_error("D2 needs at least 12 points per axis");
    const ptrdiff_t s = g.stride(a_);
    const T ad2 = a / (h * h);
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; iy++)
        for (int ix = 0; ix < g.nx; ix++)
            for (int iz = 0; iz < g.nz; iz++) {
                size_t p = g.at(ix, iy, iz);
                T v = ad2 * d2_at(in, b, p, g.coord(a_, ix, iy, iz), n, s, tb);
                out[p] = add ? out[p] + v : v;
            }
}

// out (+)= H_a in  (or H_a^-1 in) along one axis, H_a = h * diag(hweight).
template <class T>
void apply_h(int a, bool inv, bool add, const T* in, T* out, const Grid& g, T h) {
    const int n = g.len(a);
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; iy++)
        for (int ix = 0; ix < g.nx; ix++)
            for (int iz = 0; iz < g.nz; iz++) {
                size_t p = g.at(ix, iy, iz);
                int i = g.coord(a, ix, iy, iz);
                T w = (T)hweight(i, n);
                T v = inv ? in[p] / (w * h) : in[p] * w * h;
                out[p] = add ? out[p] + v : v;
            }
}

// Cos^2 taper factor of the ramp at the start of an axis (`lo` cells) and at the
// end (`hi` cells); the field is multiplied by f*f, f = cos(strength*pi/2*d/width)
// evaluated in double. Spatial_operators.cpp:408-506 (the start ramp uses the
// reversed branch with i+1).
inline double taper_lo(int i, int lo, double strength) {
    return std::cos(strength * 0.5 * M_PI * (double)(i + 1 - lo) / (double)(0 - lo));
}
inline double taper_hi(int i, int n, int hi, double strength) {
    return std::cos(strength * 0.5 * M_PI * (double)(i - (n - hi)) / (double)hi);
}

// Apply the six cos^2 tapers once, in the order z-top, z-bottom, x-left, x-right,
// y-front, y-back (tp[] = {top, bottom, left, right, front, back}).
template <class T>
void taper(T* u, int ncomp, const Grid& g, const int* tp, T strength) {
    const double a = (double)strength;
    const int nz = g.nz, nx = g.nx, ny = g.ny;
    for (int c = 0; c < ncomp; c++) {
        T* uc = u + (size_t)c * g.n();
#pragma omp parallel for
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++)
                for (int iz = 0; iz < nz; iz++) {
                    size_t p = g.at(ix, iy, iz);
                    T v = uc[p];
                    if (iz < tp[0]) { T f = (T)taper_lo(iz, tp[0], a); v *= f * f; }
                    if (iz >= nz - tp[1]) { T f = (T)taper_hi(iz, nz, tp[1], a); v *= f * f; }
                    if (ix < tp[2]) { T f = (T)taper_lo(ix, tp[2], a); v *= f * f; }
                    if (ix >= nx - tp[3]) { T f = (T)taper_hi(ix, nx, tp[3], a); v *= f * f; }
                    if (iy < tp[4]) { T f = (T)taper_lo(iy, tp[4], a); v *= f * f; }
                    if (iy >= ny - tp[5]) { T f = (T)taper_hi(iy, ny, tp[5], a); v *= f * f; }
                    uc[p] = v;
                }
    }
}

}  // namespace orc
