// ORACLE — test infrastructure only (see sbp.hpp header). CPU restatement of the
// reference's wave-equation operators, written face-generic: one SAT routine
// parameterised by (normal axis, side) instead of six face copies with nine
// region blocks each.
//
// Follows:
//   coefficient functors ........ src/cpp/we_op.cpp:469-486
//   elastic/VTI traction SATs ... src/include/spatial_operators.hpp:236-1871
//   elastic scale_boundaries .... src/include/spatial_operators.hpp:1875-2567
//   acoustic SATs ............... src/cpp/spatial_operators.cpp:508-622,
//                                 src/include/spatial_operators.hpp:2580-2716
//   acoustic scale_boundaries ... src/cpp/spatial_operators.cpp:624-1076
//   injectors ................... src/cpp/injector.cpp:3-311
//   propagate (a / e / vti) ..... src/cpp/we_op.cpp:1607-1827, 567-870, 1218-1531
//   compute_gradients ........... src/cpp/we_op.cpp:1555-1605, 488-565, 1108-1216
#pragma once
#include "sbp.hpp"
#include <cstring>
#include <string>

namespace orc {

enum Physics { PHYS_ACOUSTIC = 0, PHYS_ELASTIC = 1, PHYS_VTI = 2 };
enum FaceId { F_TOP = 0, F_BOTTOM = 1, F_LEFT = 2, F_RIGHT = 3, F_FRONT = 4, F_BACK = 5 };

struct Face {
    int n;   // normal axis
    int hi;  // 0: face at index 0, 1: face at index n-1
    int t1, t2;  // tangential axes in the order the traction terms list them
};
static const Face FACES[6] = {
    {AX_Z, 0, AX_X, AX_Y}, {AX_Z, 1, AX_X, AX_Y}, {AX_X, 0, AX_Y, AX_Z},
    {AX_X, 1, AX_Y, AX_Z}, {AX_Y, 0, AX_X, AX_Z}, {AX_Y, 1, AX_X, AX_Z},
};

template <class T>
struct PropParams {
    int nt = 0, sub = 0;
    T dt = 0;
    int bc[6] = {1, 1, 1, 1, 1, 1};     // top, bottom, left, right, front, back
    int taper[6] = {0, 0, 0, 0, 0, 0};  // same order
    T taper_strength = (T)0.05;
    T free_surface_stiffness = (T)1.05;
    // Test-only switch: 1 (the reference's behaviour) feeds the elastic/VTI imaging
    // condition the adjoint derivatives of the previous step (we_op.cpp:637-688); 0 takes
    // them from the current adjoint field, the lag-free formula the finite-difference
    // checks of the gradient need.
    int lag = 1;
};

// Iterate the points of one face plane: f(p, ix, iy, iz).
template <class F>
void for_face(const Grid& g, int face, F&& f) {
    const Face& fc = FACES[face];
    const int fixed = fc.hi ? g.len(fc.n) - 1 : 0;
    for (int iy = 0; iy < g.ny; iy++) {
        if (fc.n == AX_Y && iy != fixed) continue;
        for (int ix = 0; ix < g.nx; ix++) {
            if (fc.n == AX_X && ix != fixed) continue;
            for (int iz = 0; iz < g.nz; iz++) {
                if (fc.n == AX_Z && iz != fixed) continue;
                f(g.at(ix, iy, iz), ix, iy, iz);
            }
        }
    }
}

inline double h_of(int a, double dx, double dy, double dz) { return a == AX_X ? dx : (a == AX_Y ? dy : dz); }

// ----------------------------------------------------------------------------------
// Elastic / VTI model: mod = {lambda, mu, rho} or {lambda, mu, rho, c13, eps, c66}.
template <class T>
struct EMod {
    const T* m[6];
    bool vti;
    T lam(size_t i) const { return m[0][i]; }
    T mu(size_t i) const { return m[1][i]; }
    T rho(size_t i) const { return m[2][i]; }
    T c13(size_t i) const { return m[3][i]; }
    T eps(size_t i) const { return m[4][i]; }
    T c66(size_t i) const { return m[5][i]; }
    T mu2(size_t i) const { return 2 * m[1][i]; }
    T c11b(size_t i) const { return 2 * (1 + 2 * m[4][i]) * m[1][i]; }
    T c11c(size_t i) const { return (1 + 2 * m[4][i]) * m[0][i]; }
    T c12(size_t i) const { return (1 + 2 * m[4][i]) * (m[0][i] + 2 * m[1][i]) - 2 * m[5][i]; }
    T simp(size_t i) const { return (T)0.5 * std::sqrt(m[1][i] * m[2][i]); }
    T pimp(size_t i) const { return (T)0.5 * std::sqrt((m[0][i] + 2 * m[1][i]) * m[2][i]); }
    T pimp2(size_t i) const { return (T)0.5 * std::sqrt((1 + 2 * m[4][i]) * (m[0][i] + 2 * m[1][i]) * m[2][i]); }
    T eps2(size_t i) const { return 1 + 2 * m[4][i]; }

    // Coefficient of D_a u_a in the normal traction on a face with normal axis n
    // (the normal-normal stiffness part beyond it goes with the S term, knn()).
    T cnorm(int n, int a, size_t i) const {
        if (!vti) return lam(i);
        if (n == AX_Z) return a == AX_Z ? lam(i) : c13(i);
        if (a == n) return c11c(i);
        return a == AX_Z ? c13(i) : c12(i);
    }
    T knn(int n, size_t i) const { return (!vti || n == AX_Z) ? mu2(i) : c11b(i); }
    T pimp_face(int n, size_t i) const { return (!vti || n == AX_Z) ? pimp(i) : pimp2(i); }
    // Shear stiffness coupling normal axis n and tangential component c.
    T gshear(int n, int c, size_t i) const {
        if (vti && ((n == AX_X && c == AX_Y) || (n == AX_Y && c == AX_X))) return c66(i);
        return mu(i);
    }
};

// Component index of a displacement along axis a: u = (ux, uy, uz) -> axis x,y,z.
// (Components are stored [0]=x, [1]=y, [2]=z; axes are AX_X=0, AX_Y=1, AX_Z=2.)

// Traction SAT of one face for all three components (bc 1: free surface, 2: absorbing).
// next[c][p] -= 1/(h_n h0) * ( sigma*(tangential/normal D1 terms) + K S u / h_n - imp*prev/dt ).
template <class T>
void esat_face(int face, int bc, const T* const* curr, const T* const* prev, T* const* next,
               const EMod<T>& M, const Grid& g, T dx, T dy, T dz, T dt) {
    const Face& fc = FACES[face];
    const int n = fc.n;
    const T hn = (T)h_of(n, dx, dy, dz);
    const T adh = (T)1.0 / (hn * (T)SAT_H0);
    const T sg = fc.hi ? (T)1 : (T)-1;
    const int nn = g.len(n);
    const ptrdiff_t sn = g.stride(n);
    const ptrdiff_t inward = fc.hi ? -sn : sn;
    const bool absorb = (bc == 2);
    for_face(g, face, [&](size_t p, int ix, int iy, int iz) {
        auto D = [&](int a, const T* u) -> T {
            T h = (T)h_of(a, dx, dy, dz);
            return d1_at<T>(u, nullptr, p, g.coord(a, ix, iy, iz), g.len(a), g.stride(a)) / h;
        };
        auto S = [&](const T* u) -> T {
            T s = 0;
            for (int k = 0; k < 4; k++) s += (T)SAT_S[k] * u[(size_t)((ptrdiff_t)p + k * inward)];
            return s;
        };
        (void)nn;
        for (int c = 0; c < 3; c++) {
            T term;
            if (c == n) {
                T tang = M.cnorm(n, fc.t1, p) * D(fc.t1, curr[fc.t1]) + M.cnorm(n, fc.t2, p) * D(fc.t2, curr[fc.t2]) +
                         M.cnorm(n, n, p) * D(n, curr[n]);
                term = sg * tang + M.knn(n, p) * S(curr[n]) / hn;
                if (absorb) term -= M.pimp_face(n, p) * prev[c][p] / dt;
            } else {
                T G = M.gshear(n, c, p);
                term = sg * G * D(c, curr[n]) + G * S(curr[c]) / hn;
                if (absorb) term -= M.simp(p) * prev[c][p] / dt;
            }
            next[c][p] -= adh * term;
        }
    });
}

inline bool on_face(int face, const Grid& g, int ix, int iy, int iz) {
    const Face& fc = FACES[face];
    int i = g.coord(fc.n, ix, iy, iz);
    return fc.hi ? (i == g.len(fc.n) - 1) : (i == 0);
}

// Implicit half of the absorbing SATs: each boundary point of each component is
// divided by 1 + sum over the absorbing faces through it of v*dt/(2 h h0), with v
// the P velocity for the face-normal component (times sqrt(1+2eps) on x/y faces
// for VTI, except at the 8 corners) and the S velocity otherwise.
template <class T>
void escale(T* const* next, const EMod<T>& M, const Grid& g, T dx, T dy, T dz, T dt, const int* bc) {
    bool ab[6];
    for (int f = 0; f < 6; f++) ab[f] = (bc[f] == 2);
    const T h0 = (T)SAT_H0;
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; iy++)
        for (int ix = 0; ix < g.nx; ix++)
            for (int iz = 0; iz < g.nz; iz++) {
                bool bx = (ix == 0 || ix == g.nx - 1), by = (iy == 0 || iy == g.ny - 1), bz = (iz == 0 || iz == g.nz - 1);
                if (!(bx || by || bz)) continue;
                const bool corner = bx && by && bz;
                size_t p = g.at(ix, iy, iz);
                for (int c = 0; c < 3; c++) {
                    T sc = 0;
                    bool any = false;
                    for (int f = 0; f < 6; f++) {
                        if (!ab[f] || !on_face(f, g, ix, iy, iz)) continue;
                        any = true;
                        int n = FACES[f].n;
                        T h = (T)h_of(n, dx, dy, dz);
                        T v;
                        if (c == n) {
                            T F = (M.vti && n != AX_Z && !corner) ? M.eps2(p) : (T)1;
                            v = std::sqrt(F * (M.lam(p) + 2 * M.mu(p)) / M.rho(p));
                        } else {
                            v = std::sqrt(M.mu(p) / M.rho(p));
                        }
                        sc += v * dt / (2 * h * h0);
                    }
                    if (any) next[c][p] /= (1 + sc);
                }
            }
}

// ----------------------------------------------------------------------------------
// Acoustic model: mod = {K, 1/rho}.
template <class T>
void asat_face(int face, int bc, const T* curr, const T* prev, T* next, const T* K, const T* b, const Grid& g,
               T dx, T dy, T dz, T dt, T alpha) {
    const Face& fc = FACES[face];
    const int n = fc.n;
    const T hn = (T)h_of(n, dx, dy, dz);
    const ptrdiff_t sn = g.stride(n);
    const ptrdiff_t inward = fc.hi ? -sn : sn;
    if (bc == 1) {
        // weak Dirichlet (free surface): penalty over the first 4 points plus the face point
        const T d2 = hn * hn;
        for_face(g, face, [&](size_t p, int, int, int) {
            T v = b[p] * curr[p];
            for (int k = 0; k < 4; k++)
                next[(size_t)((ptrdiff_t)p + k * inward)] += (T)1.0 / ((T)hweight(k, 1 << 20) * d2) * (T)SAT_S[k] * v;
            next[p] -= v / (alpha * (T)SAT_H0 * d2);
        });
    } else if (bc == 2 || bc == 3) {
        const T adh = (T)1.0 / ((T)SAT_H0 * hn);
        for_face(g, face, [&](size_t p, int, int, int) {
            T s = 0;
            for (int k = 0; k < 4; k++) s += (T)SAT_S[k] * curr[(size_t)((ptrdiff_t)p + k * inward)];
            T imp = (bc == 2) ? (T)0.5 / std::sqrt(K[p] / b[p]) : (T)0;
            next[p] -= adh * (b[p] * s / hn - imp * prev[p] / dt);
        });
    }
}

template <class T>
void ascale(T* next, const T* K, const T* b, const Grid& g, T dx, T dy, T dz, T dt, const int* bc) {
    bool ab[6];
    for (int f = 0; f < 6; f++) ab[f] = (bc[f] == 2);
    const T h0 = (T)SAT_H0;
#pragma omp parallel for
    for (int iy = 0; iy < g.ny; iy++)
        for (int ix = 0; ix < g.nx; ix++)
            for (int iz = 0; iz < g.nz; iz++) {
                bool bx = (ix == 0 || ix == g.nx - 1), by = (iy == 0 || iy == g.ny - 1), bz = (iz == 0 || iz == g.nz - 1);
                if (!(bx || by || bz)) continue;
                size_t p = g.at(ix, iy, iz);
                T sc = 0;
                bool any = false;
                for (int f = 0; f < 6; f++) {
                    if (!ab[f] || !on_face(f, g, ix, iy, iz)) continue;
                    any = true;
                    T h = (T)h_of(FACES[f].n, dx, dy, dz);
                    sc += std::sqrt(K[p] * b[p]) * dt / (2 * h * h0);
                }
                if (any) next[p] /= (1 + sc);
            }
}

// ----------------------------------------------------------------------------------
// Point injectors. kind 0 (delta_m3): 3-point H-norm discrete delta per axis.
// kind 1 (ddelta_m3): adds derivative weights for moment-tensor sources.
// kind 2 (dipole_m3, DAS channels): two kind-0 stencils at the ends of the gauge,
// x -/+ gl/2 along the fiber; extraction is u(end 2) - u(end 1).
// Weight layout per point: [2][nw] with row 0 injection, row 1 extraction, nw = 3 or 6
// (kind 2: entries 0-2 end 1, 3-5 end 2); indices [npts][npi], npi = 3 or 6 (kind 2).
template <class T>
struct Injector {
    int npts = 0, kind = 0;
    std::vector<int> xi, yi, zi;    // [npts][npi]
    std::vector<T> xw, yw, zw;      // [npts][2][nw]
    int nw() const { return kind ? 6 : 3; }
    int npi() const { return kind == 2 ? 6 : 3; }
};

// Tables of dipole_m3::findIndicesWeights (src/cpp/injector.cpp:313-420): loc [npts][5]
// (x, y, z, dip, azimuth); the fiber direction is (cos dip cos az, cos dip sin az, sin dip).
// End positions, pivots and alpha are evaluated in double (the reference calls the
// double cos/sin), alpha stored in T, the weights from it as for delta_m3.
template <class T>
Injector<T> make_dipole(const T* loc, int npts, T gl, const Grid& g, T dx, T dy, T dz, T ox, T oy, T oz) {
    Injector<T> I;
    I.npts = npts;
    I.kind = 2;
    I.xi.assign(6 * npts, 0); I.yi.assign(6 * npts, 0); I.zi.assign(6 * npts, 0);
    I.xw.assign(12 * npts, 0); I.yw.assign(12 * npts, 0); I.zw.assign(12 * npts, 0);
    auto H = [](int i, int n) -> T { return (T)hweight(i, n); };
    auto axis = [&](double pos, T o, T d, int n, int* ind, T* w) {
        // pos: end position; fills ind[0..2] and w[0..2] (inject), w[6..8] (extract)
        int piv = (int)std::round((pos - o) / d);
        if (piv == 0) piv = 1;
        if (piv == n - 1) piv = n - 2;
        ind[0] = piv - 1; ind[1] = piv; ind[2] = piv + 1;
        const T a = (T)((pos - o) / d - piv);
        const double a2 = std::pow(a, 2);
        w[0] = (T)((a2 - a) / (2 * d * H(piv - 1, n)));
        w[1] = (T)((-a2 + 1) / (d * H(piv, n)));
        w[2] = (T)((a2 + a) / (2 * d * H(piv + 1, n)));
        w[6] = (T)((a2 - a) / 2);
        w[7] = (T)(-a2 + 1);
        w[8] = (T)((a2 + a) / 2);
    };
    for (int i = 0; i < npts; i++) {
        const T* L = loc + 5 * i;
        const double hx = gl / 2 * std::cos((double)L[3]) * std::cos((double)L[4]);
        const double hy = gl / 2 * std::cos((double)L[3]) * std::sin((double)L[4]);
        const double hz = gl / 2 * std::sin((double)L[3]);
        for (int e = 0; e < 2; e++) {
            const double sg = e ? 1.0 : -1.0;
            axis(L[0] + sg * hx, ox, dx, g.nx, &I.xi[6 * i + 3 * e], &I.xw[12 * i + 3 * e]);
            axis(L[1] + sg * hy, oy, dy, g.ny, &I.yi[6 * i + 3 * e], &I.yw[12 * i + 3 * e]);
            axis(L[2] + sg * hz, oz, dz, g.nz, &I.zi[6 * i + 3 * e], &I.zw[12 * i + 3 * e]);
        }
    }
    return I;
}

template <class T>
Injector<T> make_injector(int kind, const T* loc /*[npts][3] x,y,z*/, int npts, const Grid& g, T dx, T dy, T dz, T ox,
                          T oy, T oz) {
    // Arithmetic of findIndicesWeights (injector.cpp:28-80, 170-237) in data_t = T:
    // pivot and alpha in T, squares through std::pow(T,int) (double), H weights in T.
    Injector<T> I;
    I.npts = npts;
    I.kind = kind;
    const int nw = I.nw();
    I.xi.resize(3 * npts); I.yi.resize(3 * npts); I.zi.resize(3 * npts);
    I.xw.assign(2 * nw * npts, 0); I.yw.assign(2 * nw * npts, 0); I.zw.assign(2 * nw * npts, 0);
    auto H = [](int i, int n) -> T { return (T)hweight(i, n); };
    auto axis_setup = [&](T x, T o, T d, int n, int* ind, T* w) {
        int piv = (int)std::round((x - o) / d);
        if (piv == 0) piv = 1;
        if (piv == n - 1) piv = n - 2;
        ind[0] = piv - 1; ind[1] = piv; ind[2] = piv + 1;
        T a = (x - d * piv - o) / d;
        double a2 = std::pow(a, 2);
        w[0] = (T)((a2 - a) / (2 * d * H(piv - 1, n)));
        w[1] = (T)((-a2 + 1) / (d * H(piv, n)));
        w[2] = (T)((a2 + a) / (2 * d * H(piv + 1, n)));
        w[nw + 0] = (T)((a2 - a) / 2);
        w[nw + 1] = (T)(-a2 + 1);
        w[nw + 2] = (T)((a2 + a) / 2);
        if (nw == 6) {
            w[3] = (1 - 2 * a) / (2 * d * d * H(piv - 1, n));
            w[4] = 2 * a / (d * d * H(piv, n));
            w[5] = (-1 - 2 * a) / (2 * d * d * H(piv + 1, n));
            w[nw + 3] = (1 - 2 * a) / (2 * d);
            w[nw + 4] = 2 * a / (d);
            w[nw + 5] = (-1 - 2 * a) / (2 * d);
        }
    };
    for (int i = 0; i < npts; i++) {
        axis_setup(loc[3 * i + 0], ox, dx, g.nx, &I.xi[3 * i], &I.xw[2 * nw * i]);
        axis_setup(loc[3 * i + 1], oy, dy, g.ny, &I.yi[3 * i], &I.yw[2 * nw * i]);
        axis_setup(loc[3 * i + 2], oz, dz, g.nz, &I.zi[3 * i], &I.zw[2 * nw * i]);
    }
    return I;
}

// Add the source samples at time it of displacement component `comp` into `out`.
// src layout [ns][npts][nt] (ns = 1 or 3 for delta, 6 moment components for ddelta).
template <class T>
void inject(const Injector<T>& I, const T* src, int nt, int comp, int it, T* out, const Grid& g) {
    const int nw = I.nw();
    int icx = 0, icy = 0, icz = 0;
    if (I.kind == 1) {
        if (comp == 0) { icx = 0; icy = 3; icz = 4; }
        else if (comp == 1) { icx = 3; icy = 1; icz = 5; }
        else { icx = 4; icy = 5; icz = 2; }
    }
    if (I.kind == 2) {  // dipole_m3::inject (injector.cpp:423-459): -end 1, +end 2
        for (int r = 0; r < I.npts; r++) {
            const T* xw = &I.xw[12 * r];
            const T* yw = &I.yw[12 * r];
            const T* zw = &I.zw[12 * r];
            const T s = src[((size_t)comp * I.npts + r) * nt + it];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++) {
                        out[g.at(I.xi[6 * r + b], I.yi[6 * r + a], I.zi[6 * r + c])] -= xw[b] * yw[a] * zw[c] * s;
                        out[g.at(I.xi[6 * r + 3 + b], I.yi[6 * r + 3 + a], I.zi[6 * r + 3 + c])] +=
                            xw[3 + b] * yw[3 + a] * zw[3 + c] * s;
                    }
        }
        return;
    }
    for (int r = 0; r < I.npts; r++) {
        const T* xw = &I.xw[2 * nw * r];
        const T* yw = &I.yw[2 * nw * r];
        const T* zw = &I.zw[2 * nw * r];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++) {
                    size_t p = g.at(I.xi[3 * r + b], I.yi[3 * r + a], I.zi[3 * r + c]);
                    T v;
                    if (I.kind == 0) {
                        v = yw[a] * xw[b] * zw[c] * src[((size_t)comp * I.npts + r) * nt + it];
                    } else {
                        v = -xw[b + 3] * yw[a] * zw[c] * src[((size_t)icx * I.npts + r) * nt + it]
                            - xw[b] * yw[a + 3] * zw[c] * src[((size_t)icy * I.npts + r) * nt + it]
                            - xw[b] * yw[a] * zw[c + 3] * src[((size_t)icz * I.npts + r) * nt + it];
                    }
                    out[p] += v;
                }
    }
}

// rcv[comp][r][it] = sum of Lagrange-weighted field values (delta_m3 extraction).
template <class T>
void extract(const Injector<T>& I, const T* u, int nt, int comp, int it, T* rcv, const Grid& g) {
    const int nw = I.nw();
    if (I.kind == 2) {  // dipole_m3::extract (injector.cpp:461-492): u(end 2) - u(end 1)
        for (int r = 0; r < I.npts; r++) {
            const T* xw = &I.xw[12 * r + 6];
            const T* yw = &I.yw[12 * r + 6];
            const T* zw = &I.zw[12 * r + 6];
            T acc = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++)
                        acc += xw[3 + b] * yw[3 + a] * zw[3 + c] *
                                   u[g.at(I.xi[6 * r + 3 + b], I.yi[6 * r + 3 + a], I.zi[6 * r + 3 + c])] -
                               xw[b] * yw[a] * zw[c] * u[g.at(I.xi[6 * r + b], I.yi[6 * r + a], I.zi[6 * r + c])];
            rcv[((size_t)comp * I.npts + r) * nt + it] += acc;
        }
        return;
    }
    for (int r = 0; r < I.npts; r++) {
        const T* xw = &I.xw[2 * nw * r + nw];
        const T* yw = &I.yw[2 * nw * r + nw];
        const T* zw = &I.zw[2 * nw * r + nw];
        T acc = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 3; c++)
                    acc += yw[a] * xw[b] * zw[c] * u[g.at(I.xi[3 * r + b], I.yi[3 * r + a], I.zi[3 * r + c])];
        rcv[((size_t)comp * I.npts + r) * nt + it] += acc;
    }
}

// Snapshot slot holding forward step it = k*sub; 1 + nt/sub slots (zero-initialised).
inline int n_snap_slots(int nt, int sub) { return sub > 0 ? 1 + nt / sub : 0; }

// ----------------------------------------------------------------------------------
// Spatial operator of one acoustic step (we_op.cpp:1664-1766): the volume terms
// sum_a D2_a(1/rho) curr overwrite `next`; the SATs of the faces add to it (they read
// curr and, for absorbing faces, prev).
template <class T>
void acoustic_volume(const T* curr, T* next, const T* b, const Grid& g, T dx, T dy, T dz) {
    d2<T>(AX_X, false, curr, next, b, g, dx, 1);
    d2<T>(AX_Y, true, curr, next, b, g, dy, 1);
    d2<T>(AX_Z, true, curr, next, b, g, dz, 1);
}

template <class T>
void acoustic_sats(const int* bc, const T* curr, const T* prev, T* next, const T* K, const T* b, const Grid& g, T dx,
                   T dy, T dz, T dt, T alpha) {
    for (int f = 0; f < 6; f++) asat_face<T>(f, bc[f], curr, prev, next, K, b, g, dx, dy, dz, dt, alpha);
}

// Acoustic imaging condition of one event (we_op.cpp:1555-1605): snapshots itl-1, itl,
// itl+1 of the forward field and the adjoint field; tmp holds 6 fields.
template <class T>
void acoustic_imaging(const T* K, const T* snaps, int nslots, int itl, int nt, int sub, T dt, const T* adjf, T* tmp,
                      T* grad, const Grid& g, T dx, T dy, T dz) {
    const size_t N = g.n();
    const T sdt = (T)sub * dt;
    const T* fn = snaps + (size_t)itl * N;
    const T* fm = itl > 0 ? snaps + (size_t)(itl - 1) * N : nullptr;
    const T* fp = itl + 1 < nslots ? snaps + (size_t)(itl + 1) * N : nullptr;
    T* t = tmp;
    d1<T>(AX_X, false, fn, t + 0 * N, g, dx);
    d1<T>(AX_Y, false, fn, t + 1 * N, g, dy);
    d1<T>(AX_Z, false, fn, t + 2 * N, g, dz);
    d1<T>(AX_X, false, adjf, t + 3 * N, g, dx);
    d1<T>(AX_Y, false, adjf, t + 4 * N, g, dy);
    d1<T>(AX_Z, false, adjf, t + 5 * N, g, dz);
    T* gK = grad;
    T* gB = grad + N;
    const int last = nt / sub;
#pragma omp parallel for
    for (long long ii = 0; ii < (long long)N; ii++) {
        size_t i = (size_t)ii;
        T dot = t[3 * N + i] * t[i] + t[4 * N + i] * t[N + i] + t[5 * N + i] * t[2 * N + i];
        T k2 = sdt * K[i] * K[i];
        const T fpv = fp ? fp[i] : (T)0;
        if (itl > 0 && itl < last) {
            gK[i] -= (T)1.0 / k2 * adjf[i] * (fpv - 2 * fn[i] + fm[i]);
            gB[i] += sdt * dot;
        } else if (itl == 0) {
            gK[i] -= (T)1.0 / k2 * adjf[i] * (fpv - fn[i]);
            gB[i] += (T)0.5 * sdt * dot;
        } else {
            gK[i] -= (T)1.0 / k2 * adjf[i] * (-2 * fn[i] + fm[i]);
            gB[i] += sdt * dot;
        }
    }
}

// ----------------------------------------------------------------------------------
// Acoustic propagation. model = {K, 1/rho}; src [1][ns][nt]; rcv [1][nr][nt];
// grad [2][N] (gK, g(1/rho)) accumulated when non-null (adjoint); snaps [slots][N].
template <class T>
void propagate_acoustic(bool adj, const T* model, const T* src, T* rcv, const Injector<T>& inj,
                        const Injector<T>& ext, T* grad, T* snaps, const PropParams<T>& par, const Grid& g, T dx,
                        T dy, T dz) {
    (void)adj;
    const size_t N = g.n();
    const int nt = par.nt;
    const T dt = par.dt;
    std::vector<T> u(3 * N, 0);
    T* prev = u.data();
    T* curr = u.data() + N;
    T* next = u.data() + 2 * N;
    const T* K = model;
    const T* b = model + N;
    std::vector<T> tmp(grad ? 6 * N : 0, 0);
    const T alpha = (T)0.2508560249 / par.free_surface_stiffness;

    inject(inj, src, nt, 0, 0, prev, g);
    for (size_t i = 0; i < N; i++) prev[i] *= (T)0.5 * K[i] * dt * dt;

    auto grad_event = [&](int itl, const T* adjf) {
        acoustic_imaging<T>(K, snaps, n_snap_slots(nt, par.sub), itl, nt, par.sub, dt, adjf, tmp.data(), grad, g, dx,
                            dy, dz);
    };

    for (int it = 0; it < nt - 1; it++) {
        if (par.sub > 0 && it % par.sub == 0 && !grad) std::memcpy(snaps + (size_t)(it / par.sub) * N, curr, N * sizeof(T));
        if (!grad) extract(ext, curr, nt, 0, it, rcv, g);
        if (grad && par.sub > 0 && (nt - 1 - it) % par.sub == 0 && it != 0) grad_event((nt - 1 - it) / par.sub, curr);

        acoustic_volume<T>(curr, next, b, g, dx, dy, dz);
        inject(inj, src, nt, 0, it, next, g);
        acoustic_sats<T>(par.bc, curr, prev, next, K, b, g, dx, dy, dz, dt, alpha);
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            size_t i = (size_t)ii;
            next[i] = dt * dt * next[i] * K[i] + 2 * curr[i] - prev[i];
        }
        ascale<T>(next, K, b, g, dx, dy, dz, dt, par.bc);
        taper<T>(curr, 1, g, par.taper, par.taper_strength);
        taper<T>(next, 1, g, par.taper, par.taper_strength);
        T* bk = prev; prev = curr; curr = next; next = bk;
    }
    if (par.sub > 0 && (nt - 1) % par.sub == 0 && !grad) std::memcpy(snaps + (size_t)((nt - 1) / par.sub) * N, curr, N * sizeof(T));
    if (!grad) extract(ext, curr, nt, 0, nt - 1, rcv, g);
    if (grad && par.sub > 0) grad_event(0, curr);
}

// ----------------------------------------------------------------------------------
// Spatial operator of one elastic / VTI step (we_op.cpp:646-801, 1290-1465). The volume
// terms overwrite `next`: second-derivative terms D_aa(b_{c,a}) u_c, normal-stress terms
// D_c(sigma_cc'), mixed terms D_a(G D_c u_a). du[a*3 + c] receives D_a u_c of curr (the
// derivatives a lagged imaging event reads in the next step); tmp holds one field.
template <class T>
void elastic_volume(bool vti, const EMod<T>& M, const T* c11b, const T* curr, T* next, T* du, T* tmp, const Grid& g,
                    T dx, T dy, T dz) {
    const size_t N = g.n();
    auto DU = [&](int a, int c) { return du + (size_t)(a * 3 + c) * N; };
    const T* lam = M.m[0];
    const T* mu = M.m[1];
    // second-derivative terms: D_aa(b_{c,a}) u_c
    for (int c = 0; c < 3; c++) {
        for (int a = 0; a < 3; a++) {
            const T h = (T)h_of(a, dx, dy, dz);
            T* out = next + c * N;
            const T* in = curr + c * N;
            const bool add = (a != AX_X);
            if (!vti) {
                d2<T>(a, add, in, out, mu, g, h, a == c ? (T)2 : (T)1);
            } else {
                // VTI: x-x and y-y use c11b for the matching component, c66 across x/y
                if ((c == 0 && a == AX_X) || (c == 1 && a == AX_Y)) d2<T>(a, add, in, out, c11b, g, h, (T)1);
                else if ((c == 0 && a == AX_Y) || (c == 1 && a == AX_X)) d2<T>(a, add, in, out, M.m[5], g, h, 1);
                else d2<T>(a, add, in, out, mu, g, h, (a == AX_Z && c == 2) ? (T)2 : (T)1);
            }
        }
    }
    // normal-stress terms: next_c += D_c( sigma_cc' )
    d1<T>(AX_X, false, curr + 0 * N, DU(AX_X, 0), g, dx);
    d1<T>(AX_Y, false, curr + 1 * N, DU(AX_Y, 1), g, dy);
    d1<T>(AX_Z, false, curr + 2 * N, DU(AX_Z, 2), g, dz);
    {
        const T *ux = DU(AX_X, 0), *uy = DU(AX_Y, 1), *uz = DU(AX_Z, 2);
        T* t = tmp;
        if (!vti) {
            for (size_t i = 0; i < N; i++) t[i] = lam[i] * (ux[i] + uy[i] + uz[i]);
            d1<T>(AX_X, true, t, next + 0 * N, g, dx);
            d1<T>(AX_Y, true, t, next + 1 * N, g, dy);
            d1<T>(AX_Z, true, t, next + 2 * N, g, dz);
        } else {
            for (size_t i = 0; i < N; i++) t[i] = M.c11c(i) * ux[i] + M.c12(i) * uy[i] + M.c13(i) * uz[i];
            d1<T>(AX_X, true, t, next + 0 * N, g, dx);
            for (size_t i = 0; i < N; i++) t[i] = M.c11c(i) * uy[i] + M.c12(i) * ux[i] + M.c13(i) * uz[i];
            d1<T>(AX_Y, true, t, next + 1 * N, g, dy);
            for (size_t i = 0; i < N; i++) t[i] = M.c13(i) * (ux[i] + uy[i]) + lam[i] * uz[i];
            d1<T>(AX_Z, true, t, next + 2 * N, g, dz);
        }
    }
    // mixed terms: next_c += D_a( G * D_c u_a ), a != c.  The z-derivative of the
    // x-component term uses dy as its spacing (we_op.cpp:678, 1339).
    const T* gxy = vti ? M.m[5] : mu;
    d1<T>(AX_X, false, curr + 1 * N, DU(AX_X, 1), g, dx);
    d1<T>(AX_X, false, curr + 2 * N, DU(AX_X, 2), g, dx);
    d1<T>(AX_Y, true, DU(AX_X, 1), next + 0 * N, g, dy, gxy);
    d1<T>(AX_Z, true, DU(AX_X, 2), next + 0 * N, g, dy, mu);
    d1<T>(AX_Y, false, curr + 0 * N, DU(AX_Y, 0), g, dy);
    d1<T>(AX_Y, false, curr + 2 * N, DU(AX_Y, 2), g, dy);
    d1<T>(AX_X, true, DU(AX_Y, 0), next + 1 * N, g, dx, gxy);
    d1<T>(AX_Z, true, DU(AX_Y, 2), next + 1 * N, g, dz, mu);
    d1<T>(AX_Z, false, curr + 0 * N, DU(AX_Z, 0), g, dz);
    d1<T>(AX_Z, false, curr + 1 * N, DU(AX_Z, 1), g, dz);
    d1<T>(AX_X, true, DU(AX_Z, 0), next + 2 * N, g, dx, mu);
    d1<T>(AX_Y, true, DU(AX_Z, 1), next + 2 * N, g, dy, mu);
}

// Traction SATs of the faces with bc 1 (free surface) or 2 (absorbing), added to next.
template <class T>
void elastic_sats(const EMod<T>& M, const int* bc, const T* curr, const T* prev, T* next, const Grid& g, T dx, T dy,
                  T dz, T dt) {
    const size_t N = g.n();
    const T* cu[3] = {curr, curr + N, curr + 2 * N};
    const T* pv[3] = {prev, prev + N, prev + 2 * N};
    T* nx3[3] = {next, next + N, next + 2 * N};
    for (int f = 0; f < 6; f++)
        if (bc[f] == 1 || bc[f] == 2) esat_face<T>(f, bc[f], cu, pv, nx3, M, g, dx, dy, dz, dt);
}

// Elastic / VTI imaging condition of one event (we_op.cpp:488-565, 1108-1216): snapshots
// itl-1, itl, itl+1 of the forward field, the adjoint field adjf (rho term) and the
// adjoint derivatives adu[a*3 + c] = D_a u_c (lambda and mu terms; the reference passes
// those of the previous step). tmp holds 9 fields.
template <class T>
void elastic_imaging(bool vti, const EMod<T>& M, const T* snaps, int nslots, int itl, int nt, int sub, T dt,
                     const T* adjf, const T* adu, T* tmp, T* grad, const Grid& g, T dx, T dy, T dz) {
    const size_t N = g.n();
    const T sdt = (T)sub * dt;
    const T* fn = snaps + (size_t)itl * 3 * N;
    const T* fm = itl > 0 ? snaps + (size_t)(itl - 1) * 3 * N : nullptr;
    const T* fp = itl + 1 < nslots ? snaps + (size_t)(itl + 1) * 3 * N : nullptr;
    T* t = tmp;
    // forward-field derivatives: ux_x, uy_y, uz_z, uz_x, ux_z, uy_x, ux_y, uz_y, uy_z
    d1<T>(AX_X, false, fn + 0 * N, t + 0 * N, g, dx);
    d1<T>(AX_Y, false, fn + 1 * N, t + 1 * N, g, dy);
    d1<T>(AX_Z, false, fn + 2 * N, t + 2 * N, g, dz);
    d1<T>(AX_X, false, fn + 2 * N, t + 3 * N, g, dx);
    d1<T>(AX_Z, false, fn + 0 * N, t + 4 * N, g, dz);
    d1<T>(AX_X, false, fn + 1 * N, t + 5 * N, g, dx);
    d1<T>(AX_Y, false, fn + 0 * N, t + 6 * N, g, dy);
    d1<T>(AX_Y, false, fn + 2 * N, t + 7 * N, g, dy);
    d1<T>(AX_Z, false, fn + 1 * N, t + 8 * N, g, dz);
    T* g0 = grad;
    T* g1 = grad + N;
    T* g2 = grad + 2 * N;
    const int last = nt / sub;
    const int branch = (itl > 0 && itl < last) ? 1 : (itl == 0 ? 0 : 2);
    const T w = branch == 0 ? (T)0.5 : (T)1;
    auto DU = [&](int a, int c) { return adu + (size_t)(a * 3 + c) * N; };
    const T *ax_x = DU(AX_X, 0), *ay_y = DU(AX_Y, 1), *az_z = DU(AX_Z, 2);
    const T *az_x = DU(AX_Z, 0), *ax_z = DU(AX_X, 2), *ay_x = DU(AX_Y, 0), *ax_y = DU(AX_X, 1);
    const T *az_y = DU(AX_Z, 1), *ay_z = DU(AX_Y, 2);
#pragma omp parallel for
    for (long long ii = 0; ii < (long long)N; ii++) {
        size_t i = (size_t)ii;
        T rt = 0;
        for (int c = 0; c < 3; c++) {
            size_t q = c * N + i;
            const T fpv = fp ? fp[q] : (T)0;
            T d2t = branch == 1 ? fpv - 2 * fn[q] + fm[q] : (branch == 0 ? fpv - fn[q] : -fn[q] + fm[q]);
            rt += adjf[q] * d2t;
        }
        if (!vti) {
            g0[i] += w * sdt * (ax_x[i] + ay_y[i] + az_z[i]) * (t[i] + t[N + i] + t[2 * N + i]);
            g1[i] += w * sdt * ((az_x[i] + ax_z[i]) * (t[3 * N + i] + t[4 * N + i]) +
                                (ay_x[i] + ax_y[i]) * (t[5 * N + i] + t[6 * N + i]) +
                                (az_y[i] + ay_z[i]) * (t[7 * N + i] + t[8 * N + i]) + 2 * ax_x[i] * t[i] +
                                2 * ay_y[i] * t[N + i] + 2 * az_z[i] * t[2 * N + i]);
        } else {
            T L = M.lam(i), U = M.mu(i), E = M.eps(i);
            T del = ((M.c13(i) + U) * (M.c13(i) + U) - (L + U) * (L + U)) / (2 * (L + U) * (L + 2 * U));
            T gam = (M.c66(i) - U) / (2 * U);
            T v1 = std::sqrt(2 * (L + 2 * U) * (L + U) * del + (L + U) * (L + U));
            T v2 = ((1 + 2 * del) * L + (1 + 3 * del) * U) / v1;
            T v3 = ((1 + 3 * del) * L + (1 + 4 * del) * U) / v1 - 1;
            T v4 = (ax_x[i] + ay_y[i]) * t[2 * N + i] + az_z[i] * (t[i] + t[N + i]);
            g0[i] += w * sdt * ((1 + 2 * E) * (ax_x[i] + ay_y[i]) * (t[i] + t[N + i]) + az_z[i] * t[2 * N + i] + v2 * v4);
            g1[i] += w * sdt * (2 * (1 + 2 * E) * (ax_x[i] * t[i] + ay_y[i] * t[N + i]) + 2 * az_z[i] * t[2 * N + i] +
                                4 * (E - gam) * (ax_x[i] * t[N + i] + ax_x[i] * t[i]) +
                                (az_y[i] + ay_z[i]) * (t[7 * N + i] + t[8 * N + i]) +
                                (az_x[i] + ax_z[i]) * (t[3 * N + i] + t[4 * N + i]) +
                                (1 + 2 * gam) * (ay_x[i] + ax_y[i]) * (t[5 * N + i] + t[6 * N + i]) + v3 * v4);
        }
        g2[i] += (T)1.0 / sdt * rt;
    }
}

// The nine first derivatives D_a u_c of a 3-component field, du[a*3 + c].
template <class T>
void all_derivatives(const T* u, T* du, const Grid& g, T dx, T dy, T dz) {
    const size_t N = g.n();
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++)
            d1<T>(a, false, u + (size_t)c * N, du + (size_t)(a * 3 + c) * N, g, (T)h_of(a, dx, dy, dz));
}

template <class T>
EMod<T> emod(bool vti, const T* model, size_t N) {
    EMod<T> M;
    M.vti = vti;
    for (int k = 0; k < 6; k++) M.m[k] = (vti || k < 3) ? model + (size_t)k * N : nullptr;
    return M;
}

// ----------------------------------------------------------------------------------
// Elastic / VTI propagation. model = {lam, mu, rho[, c13, eps, c66]}; src [ns][nsrc][nt]
// (ns = 3 forces or 6 moment components); rcv [3][nr][nt]; grad [3][N] (glam, gmu, grho);
// snaps [slots][3][N].
template <class T>
void propagate_elastic(bool vti, bool adj, const T* model, const T* src, T* rcv, const Injector<T>& inj,
                       const Injector<T>& ext, T* grad, T* snaps, const PropParams<T>& par, const Grid& g, T dx,
                       T dy, T dz) {
    (void)adj;
    const size_t N = g.n();
    const int nt = par.nt;
    const T dt = par.dt;
    const EMod<T> M = emod<T>(vti, model, N);
    std::vector<T> u(9 * N, 0), du(9 * N, 0), tmp((grad ? 9 : 1) * N, 0);
    std::vector<T> du_now(grad && !par.lag ? 9 * N : 0);
    T* prev = u.data();
    T* curr = u.data() + 3 * N;
    T* next = u.data() + 6 * N;
    const T* rho = M.m[2];
    std::vector<T> c11b(vti ? N : 0);
    for (size_t i = 0; i < c11b.size(); i++) c11b[i] = M.c11b(i);

    for (int c = 0; c < 3; c++) inject(inj, src, nt, c, 0, prev + c * N, g);
    for (size_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) prev[c * N + i] *= (T)0.5 * dt * dt / rho[i];

    auto grad_event = [&](int itl, const T* adjf) {
        // lagged: du holds the derivatives of the previous step's field (computed in the
        // volume terms of that step); lag-free: the derivatives of adjf itself
        const T* adu = du.data();
        if (!par.lag) {
            all_derivatives<T>(adjf, du_now.data(), g, dx, dy, dz);
            adu = du_now.data();
        }
        elastic_imaging<T>(vti, M, snaps, n_snap_slots(nt, par.sub), itl, nt, par.sub, dt, adjf, adu, tmp.data(), grad,
                           g, dx, dy, dz);
    };

    for (int it = 0; it < nt - 1; it++) {
        if (par.sub > 0 && it % par.sub == 0 && !grad)
            std::memcpy(snaps + (size_t)(it / par.sub) * 3 * N, curr, 3 * N * sizeof(T));
        if (!grad)
            for (int c = 0; c < 3; c++) extract(ext, curr + c * N, nt, c, it, rcv, g);
        if (grad && par.sub > 0 && (nt - 1 - it) % par.sub == 0 && it != 0) grad_event((nt - 1 - it) / par.sub, curr);

        elastic_volume<T>(vti, M, c11b.data(), curr, next, du.data(), tmp.data(), g, dx, dy, dz);
        for (int c = 0; c < 3; c++) inject(inj, src, nt, c, it, next + c * N, g);
        elastic_sats<T>(M, par.bc, curr, prev, next, g, dx, dy, dz, dt);
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            size_t i = (size_t)ii;
            for (int c = 0; c < 3; c++) next[c * N + i] = dt * dt * next[c * N + i] / rho[i] + 2 * curr[c * N + i] - prev[c * N + i];
        }
        T* nx3[3] = {next, next + N, next + 2 * N};
        escale<T>(nx3, M, g, dx, dy, dz, dt, par.bc);
        taper<T>(curr, 3, g, par.taper, par.taper_strength);
        taper<T>(next, 3, g, par.taper, par.taper_strength);
        T* bk = prev; prev = curr; curr = next; next = bk;
    }
    if (par.sub > 0 && (nt - 1) % par.sub == 0 && !grad)
        std::memcpy(snaps + (size_t)((nt - 1) / par.sub) * 3 * N, curr, 3 * N * sizeof(T));
    if (!grad)
        for (int c = 0; c < 3; c++) extract(ext, curr + c * N, nt, c, nt - 1, rcv, g);
    if (grad && par.sub > 0) grad_event(0, curr);
}

}  // namespace orc
