// ORACLE — test infrastructure only. C-ABI wrappers around the reference's own
// SBP/SAT/taper operators (src/cpp/spatial_operators.cpp and the templates of
// src/include/spatial_operators.hpp), compiled in place from /root/reference by
// oracle/Makefile into oracle/_ref/ (never committed). Used to pin the oracle's
// restatement operator by operator; nothing in the product links it.
//
// Build twice: float (libref_ops_f.so, symbols *_f) and -DDOUBLE_PRECISION
// (libref_ops_d.so, symbols *_d).
#include "spatial_operators.hpp"

#ifdef REF_DOUBLE
#define SFX(name) name##_d
#else
#define SFX(name) name##_f
#endif

// Coefficient functors with the signature the reference templates expect. The
// formulas are the ones the reference uses (src/cpp/we_op.cpp:469-486); par is
// the converted model {lambda, mu, rho, c13, eps, c66} or {K, 1/rho}.
static inline data_t zero(const data_t**, int) { return 0; }
static inline data_t one(const data_t**, int) { return 1; }
static inline data_t lam(const data_t** p, int i) { return p[0][i]; }
static inline data_t mu(const data_t** p, int i) { return p[1][i]; }
static inline data_t c13(const data_t** p, int i) { return p[3][i]; }
static inline data_t c66(const data_t** p, int i) { return p[5][i]; }
static inline data_t mu2(const data_t** p, int i) { return 2 * p[1][i]; }
static inline data_t c11b(const data_t** p, int i) { return 2 * (1 + 2 * p[4][i]) * p[1][i]; }
static inline data_t c11c(const data_t** p, int i) { return (1 + 2 * p[4][i]) * p[0][i]; }
static inline data_t c12(const data_t** p, int i) { return (1 + 2 * p[4][i]) * (p[0][i] + 2 * p[1][i]) - 2 * p[5][i]; }
static inline data_t simp(const data_t** p, int i) { return 0.5 * sqrt(p[1][i] * p[2][i]); }
static inline data_t pimp(const data_t** p, int i) { return 0.5 * sqrt((p[0][i] + 2 * p[1][i]) * p[2][i]); }
static inline data_t pimp2(const data_t** p, int i) { return 0.5 * sqrt((1 + 2 * p[4][i]) * (p[0][i] + 2 * p[1][i]) * p[2][i]); }
static inline data_t aimp(const data_t** p, int i) { return 0.5 / sqrt(p[0][i] / p[1][i]); }
static inline data_t eps2(const data_t** p, int i) { return 1 + 2 * p[4][i]; }
static inline data_t bcoef(const data_t** p, int i) { return p[0][i]; }

extern "C" {

// axis: 0=x, 1=y, 2=z. par == NULL -> D1, else D1(par * in).
void SFX(ref_d1)(int axis, int add, const data_t* in, data_t* out, int nx, int ny, int nz, data_t h, const data_t* par) {
    if (par == nullptr) {
        if (axis == 0) Dx(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
        else if (axis == 1) Dy(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
        else Dz(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
    } else {
        if (axis == 0) mult_Dx(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par);
        else if (axis == 1) mult_Dy(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par);
        else mult_Dz(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par);
    }
}

void SFX(ref_d2)(int axis, int add, const data_t* in, data_t* out, const data_t* b, int nx, int ny, int nz, data_t h, data_t a) {
    const data_t* par[1] = {b};
    if (axis == 0) Dxx_var<bcoef>(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par, a);
    else if (axis == 1) Dyy_var<bcoef>(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par, a);
    else Dzz_var<bcoef>(add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz, par, a);
}

void SFX(ref_apply_h)(int axis, int inv, int add, const data_t* in, data_t* out, int nx, int ny, int nz, data_t h) {
    if (axis == 0) applyHx(inv, add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
    else if (axis == 1) applyHy(inv, add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
    else applyHz(inv, add, in, out, nx, ny, nz, h, 0, nx, 0, ny, 0, nz);
}

// taper[] = {top, bottom, left, right, front, back}; applied in that order.
void SFX(ref_taper)(data_t* u, int ncomp, int nx, int ny, int nz, const int* tp, data_t a) {
    taperz(u, nx, ny, nz, ncomp, 0, nx, 0, ny, tp[0], 0, 0, ncomp, a);
    taperz(u, nx, ny, nz, ncomp, 0, nx, 0, ny, nz - tp[1], nz, 0, ncomp, a);
    taperx(u, nx, ny, nz, ncomp, tp[2], 0, 0, ny, 0, nz, 0, ncomp, a);
    taperx(u, nx, ny, nz, ncomp, nx - tp[3], nx, 0, ny, 0, nz, 0, ncomp, a);
    tapery(u, nx, ny, nz, ncomp, 0, nx, tp[4], 0, 0, nz, 0, ncomp, a);
    tapery(u, nx, ny, nz, ncomp, 0, nx, ny - tp[5], ny, 0, nz, 0, ncomp, a);
}

// Acoustic SAT on one face (0 top,1 bottom,2 left,3 right,4 front,5 back);
// bc 1 Dirichlet (alpha), 2 absorbing, 3 rigid. model = {K, 1/rho}.
void SFX(ref_asat)(int face, int bc, const data_t* curr, const data_t* prev, data_t* next, const data_t* model,
                   int nx, int ny, int nz, data_t dx, data_t dy, data_t dz, data_t dt, data_t alpha) {
    const size_t n = (size_t)nx * ny * nz;
    const data_t* mod[2] = {model, model + n};
    const data_t* in1[1] = {curr};
    const data_t* in2[2] = {curr, prev};
    switch (face * 4 + bc) {
        case 0 * 4 + 1: asat_dirichlet_top(true, in1, next, nx, ny, nz, dx, dy, dz, 0, nx, 0, ny, mod, alpha); break;
        case 0 * 4 + 2: asat_neumann_top<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod); break;
        case 0 * 4 + 3: asat_neumann_top<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod); break;
        case 1 * 4 + 1: asat_dirichlet_bottom(true, in1, next, nx, ny, nz, dx, dy, dz, 0, nx, 0, ny, mod, alpha); break;
        case 1 * 4 + 2: asat_neumann_bottom<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod); break;
        case 1 * 4 + 3: asat_neumann_bottom<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod); break;
        case 2 * 4 + 1: asat_dirichlet_left(true, in1, next, nx, ny, nz, dx, dy, dz, 0, ny, 0, nz, mod, alpha); break;
        case 2 * 4 + 2: asat_neumann_left<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod); break;
        case 2 * 4 + 3: asat_neumann_left<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod); break;
        case 3 * 4 + 1: asat_dirichlet_right(true, in1, next, nx, ny, nz, dx, dy, dz, 0, ny, 0, nz, mod, alpha); break;
        case 3 * 4 + 2: asat_neumann_right<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod); break;
        case 3 * 4 + 3: asat_neumann_right<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod); break;
        case 4 * 4 + 1: asat_dirichlet_front(true, in1, next, nx, ny, nz, dx, dy, dz, 0, nx, 0, nz, mod, alpha); break;
        case 4 * 4 + 2: asat_neumann_front<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod); break;
        case 4 * 4 + 3: asat_neumann_front<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod); break;
        case 5 * 4 + 1: asat_dirichlet_back(true, in1, next, nx, ny, nz, dx, dy, dz, 0, nx, 0, nz, mod, alpha); break;
        case 5 * 4 + 2: asat_neumann_back<mu, aimp>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod); break;
        case 5 * 4 + 3: asat_neumann_back<mu, zero>(true, in2, next, nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod); break;
        default: break;
    }
}

void SFX(ref_ascale)(data_t* next, const data_t* model, int nx, int ny, int nz, data_t dx, data_t dy, data_t dz, data_t dt, const int* bc) {
    const size_t n = (size_t)nx * ny * nz;
    const data_t* mod[2] = {model, model + n};
    data_t* in[1] = {next};
    asat_scale_boundaries(in, nx, ny, nz, dx, dy, dz, 0, nx, 0, ny, 0, nz, mod, dt, bc[0] == 2, bc[1] == 2, bc[2] == 2,
                          bc[3] == 2, bc[4] == 2, bc[5] == 2);
}

}  // extern "C"

// Elastic (vti=0, model {lam,mu,rho}) or VTI (vti=1, model {lam,mu,rho,c13,eps,c66})
// traction SATs of one face for the three components; bc 1 free surface, 2 absorbing.
// The component wiring of each face is the one the reference's propagate uses.
template <expr F1n, expr F2n, expr F3n, expr F4n, expr F5n, expr G1, expr G2, expr Sx>
static void esat_face_t(int face, const data_t* const* c, const data_t* const* p, data_t* const* nx3, int nx, int ny, int nz,
                        data_t dx, data_t dy, data_t dz, data_t dt, const data_t** mod) {
    // F*n: normal-component functors; G1/G2: tangential stiffness for the two
    // tangential components (in the face's order); Sx: shear impedance (or zero).
    if (face == 0 || face == 1) {
        const data_t* a[4] = {c[2], c[1], c[0], p[0]};
        const data_t* b[4] = {c[2], c[2], c[1], p[1]};
        const data_t* n[4] = {c[0], c[1], c[2], p[2]};
        if (face == 0) {
            esat_neumann_top<G1, zero, zero, G1, Sx>(true, a, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
            esat_neumann_top<zero, G1, zero, G1, Sx>(true, b, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
            esat_neumann_top<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
        } else {
            esat_neumann_bottom<G1, zero, zero, G1, Sx>(true, a, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
            esat_neumann_bottom<zero, G1, zero, G1, Sx>(true, b, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
            esat_neumann_bottom<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, ny, mod);
        }
    } else if (face == 2 || face == 3) {
        const data_t* a[4] = {c[0], c[2], c[1], p[1]};
        const data_t* b[4] = {c[0], c[0], c[2], p[2]};
        const data_t* n[4] = {c[1], c[2], c[0], p[0]};
        if (face == 2) {
            esat_neumann_left<G1, zero, zero, G1, Sx>(true, a, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
            esat_neumann_left<zero, G2, zero, G2, Sx>(true, b, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
            esat_neumann_left<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
        } else {
            esat_neumann_right<G1, zero, zero, G1, Sx>(true, a, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
            esat_neumann_right<zero, G2, zero, G2, Sx>(true, b, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
            esat_neumann_right<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, ny, 0, nz, mod);
        }
    } else {
        const data_t* a[4] = {c[1], c[2], c[0], p[0]};
        const data_t* b[4] = {c[1], c[1], c[2], p[2]};
        const data_t* n[4] = {c[0], c[2], c[1], p[1]};
        if (face == 4) {
            esat_neumann_front<G1, zero, zero, G1, Sx>(true, a, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
            esat_neumann_front<zero, G2, zero, G2, Sx>(true, b, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
            esat_neumann_front<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
        } else {
            esat_neumann_back<G1, zero, zero, G1, Sx>(true, a, nx3[0], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
            esat_neumann_back<zero, G2, zero, G2, Sx>(true, b, nx3[2], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
            esat_neumann_back<F1n, F2n, F3n, F4n, F5n>(true, n, nx3[1], nx, ny, nz, dx, dy, dz, dt, 0, nx, 0, nz, mod);
        }
    }
}

extern "C" {

void SFX(ref_esat)(int face, int bc, int vti, const data_t* curr, const data_t* prev, data_t* next, const data_t* model,
                   int nx, int ny, int nz, data_t dx, data_t dy, data_t dz, data_t dt) {
    const size_t n = (size_t)nx * ny * nz;
    const data_t* mod[6] = {model, model + n, model + 2 * n, model + 3 * n, model + 4 * n, model + 5 * n};
    const data_t* c[3] = {curr, curr + n, curr + 2 * n};
    const data_t* p[3] = {prev, prev + n, prev + 2 * n};
    data_t* nx3[3] = {next, next + n, next + 2 * n};
    const bool z = (face <= 1);
    if (!vti) {
        if (bc == 1) esat_face_t<lam, lam, lam, mu2, zero, mu, mu, zero>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
        else if (bc == 2) esat_face_t<lam, lam, lam, mu2, pimp, mu, mu, simp>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
    } else if (z) {
        if (bc == 1) esat_face_t<c13, c13, lam, mu2, zero, mu, mu, zero>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
        else if (bc == 2) esat_face_t<c13, c13, lam, mu2, pimp, mu, mu, simp>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
    } else {
        if (bc == 1) esat_face_t<c12, c13, c11c, c11b, zero, c66, mu, zero>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
        else if (bc == 2) esat_face_t<c12, c13, c11c, c11b, pimp2, c66, mu, simp>(face, c, p, nx3, nx, ny, nz, dx, dy, dz, dt, mod);
    }
}

void SFX(ref_escale)(int vti, data_t* next, const data_t* model, int nx, int ny, int nz, data_t dx, data_t dy, data_t dz, data_t dt, const int* bc) {
    const size_t n = (size_t)nx * ny * nz;
    const data_t* mod[6] = {model, model + n, model + 2 * n, model + 3 * n, model + 4 * n, model + 5 * n};
    data_t* in[3] = {next, next + n, next + 2 * n};
    if (vti)
        esat_scale_boundaries<eps2>(in, nx, ny, nz, dx, dy, dz, 0, nx, 0, ny, 0, nz, mod, dt, bc[0] == 2, bc[1] == 2,
                                    bc[2] == 2, bc[3] == 2, bc[4] == 2, bc[5] == 2);
    else
        esat_scale_boundaries<one>(in, nx, ny, nz, dx, dy, dz, 0, nx, 0, ny, 0, nz, mod, dt, bc[0] == 2, bc[1] == 2,
                                   bc[2] == 2, bc[3] == 2, bc[4] == 2, bc[5] == 2);
}

}  // extern "C"
