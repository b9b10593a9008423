/*
 * fwi3d_hip.h — C ABI of the MI355X (gfx950) wave-propagation / adjoint-gradient
 * engine. Plain pointers and sizes only; implemented by the HIP sources in fwi3d_amd/csrc and
 * exported by fwi3d_amd/libfwi3d_amd.so.
 *
 * Layouts (the reference's, src/include/spatial_operators.hpp:5-8):
 *   3-D fields and models  [param][y][x][z], z fastest
 *   source / receiver traces [component][point][t], t fastest (src/cpp/injector.cpp:127)
 * Units: km, s, km/s, g/cc, GPa (src/include/param.hpp:28-31).
 *
 * Every function returns 0 on success and a negative fwi3d_status otherwise; the
 * message is available from fwi3d_last_error(). Where the reference calls
 * successCheck() -> exit(123) (src/include/misc.hpp:4-12) this library returns
 * FWI3D_EINVAL instead.
 */
#ifndef FWI3D_HIP_H
#define FWI3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum fwi3d_status {
    FWI3D_OK = 0,
    FWI3D_EINVAL = -1,   /* invalid argument (geometry, sizes, parameters) */
    FWI3D_EDEVICE = -2,  /* HIP runtime error */
    FWI3D_ENOMEM = -3,   /* device allocation failed */
    FWI3D_ESTATE = -4    /* call order (e.g. adjoint without a stored forward) */
};

enum fwi3d_physics { FWI3D_ACOUSTIC = 0, FWI3D_ELASTIC = 1, FWI3D_VTI = 2 };

enum fwi3d_snap_precision { FWI3D_SNAP_F32 = 0, FWI3D_SNAP_BF16 = 1, FWI3D_SNAP_F16 = 2 };

/* Propagation parameters: the subset of the reference's `param` that propagate()
 * reads (src/include/param.hpp:10-43), plus the snapshot precision. */
typedef struct {
    int nt;                       /* number of propagation time samples */
    int sub;                      /* snapshot / imaging interval in steps (0: none) */
    float dt;                     /* propagation time step (s) */
    int bc[6];                    /* top,bottom,left,right,front,back: 0 none, 1 free surface,
                                     2 locally absorbing, 3 rigid wall (acoustic only) */
    int taper[6];                 /* cos^2 taper width in cells, same face order */
    float taper_strength;         /* reference default 0.05 */
    float free_surface_stiffness; /* reference default 1.05 */
    int snap_precision;           /* fwi3d_snap_precision (reference: FP32 only) */
} fwi3d_param;

/* Point injector / extractor tables: the arrays of the reference's `injector`
 * (src/include/injector.hpp:8-16) as filled by findIndicesWeights
 * (src/cpp/injector.cpp:3-80, 142-237). nw = 3 (delta_m3) or 6 (ddelta_m3). */
typedef struct {
    int npts;
    int kind;                        /* 0 delta_m3, 1 ddelta_m3 (moment tensor) */
    const int *xind, *yind, *zind;   /* [npts][3] grid indices */
    const float *xw, *yw, *zw;       /* [npts][2][nw]: row 0 injection, row 1 extraction */
} fwi3d_injector;

typedef struct fwi3d_ctx fwi3d_ctx;

/* Context = one HIP device + one stream + the device-resident state of an operator
 * instance (wavefields, snapshot store). Replaces the per-instance state of
 * nl_we_op (src/include/we_op.hpp:18-24, `mutable zfpWrapper _zfp_wfld`). */
int fwi3d_create(fwi3d_ctx **ctx, int device);
int fwi3d_destroy(fwi3d_ctx *ctx);
const char *fwi3d_last_error(const fwi3d_ctx *ctx);

/* Compute the injector tables for points loc[npts][3] = (x, y, z) on the grid.
 * Replaces delta_m3 / ddelta_m3::findIndicesWeights (src/cpp/injector.cpp:3-80,
 * 142-237), in the same single-precision arithmetic. Output arrays: ind [npts][3],
 * w [npts][2][nw]. A point outside the grid gives FWI3D_EINVAL. */
int fwi3d_injector_tables(int kind, const float *loc, int npts, int nx, int ny, int nz, float dx, float dy,
                          float dz, float ox, float oy, float oz, int *xind, int *yind, int *zind, float *xw,
                          float *yw, float *zw);

/* One forward or adjoint propagation, host pointers. Replaces
 * nl_we_op_{a,e,vti}::propagate (src/include/we_op.hpp:93,127,163; src/cpp/we_op.cpp:
 * 567-870, 1218-1531, 1607-1827) with the same argument meaning:
 *   model  converted parameters: acoustic {K, 1/rho}, elastic {lambda, mu, rho},
 *          VTI {lambda, mu, rho, c13, eps, c66}; [np][ny][nx][nz]
 *   src    [ns][inj->npts][nt] (adjoint: the residual injected at the receivers)
 *   rcv    [nrc][ext->npts][nt], overwritten (forward only; may be NULL for adjoint)
 *   grad   NULL (forward) or [ng][ny][nx][nz] accumulated (+=) imaging condition of
 *          compute_gradients (acoustic {K, 1/rho}: ng=2; elastic/VTI {lambda,mu,rho}: ng=3)
 * A forward call with par->sub > 0 stores the snapshots in the context; an adjoint
 * call reads the ones of the last forward call on the same context. */
int fwi3d_propagate(fwi3d_ctx *ctx, int physics, int adj, const float *model, const float *src, float *rcv,
                    const fwi3d_injector *inj, const fwi3d_injector *ext, float *grad, const fwi3d_param *par,
                    int nx, int ny, int nz, float dx, float dy, float dz);

/* Same with device pointers (data already resident in HBM, reference layouts) and
 * an FP64 gradient accumulator grad64 [ng][ny][nx][nz] (+=). The injector tables
 * stay host arrays. Returns after the work has completed. */
int fwi3d_propagate_device(fwi3d_ctx *ctx, int physics, int adj, const float *d_model, const float *d_src,
                           float *d_rcv, const fwi3d_injector *inj, const fwi3d_injector *ext, double *d_grad64,
                           const fwi3d_param *par, int nx, int ny, int nz, float dx, float dy, float dz);

/* Kernel timing with HIP events on the context's stream, for roofline accounting
 * (no counterpart in the reference). enable=1 starts collecting (and resets), 0 stops.
 * read() returns the summed milliseconds and launch count since the last enable of
 * kernel class `kernel`: 0 = the interior time-step kernel, 1 = the elastic/VTI flux
 * pre-pass, 2 = the boundary-shell time-step kernel. */
int fwi3d_profile(fwi3d_ctx *ctx, int enable);
int fwi3d_profile_read(const fwi3d_ctx *ctx, int kernel, double *ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif
