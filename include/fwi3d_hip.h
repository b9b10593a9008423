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
 * (src/cpp/injector.cpp:3-80, 142-237, 313-420). nw = 3 (delta_m3) or 6 (ddelta_m3,
 * dipole_m3). A dipole (DAS channel) has two 3-point stencils per axis, entries 0-2 for
 * the end at -gl/2 along the fiber and 3-5 for the end at +gl/2; it extracts
 * u(end 2) - u(end 1) and injects -w at end 1, +w at end 2 (injector.cpp:423-492). */
typedef struct {
    int npts;
    int kind;                        /* 0 delta_m3, 1 ddelta_m3 (moment tensor), 2 dipole_m3 (DAS) */
    const int *xind, *yind, *zind;   /* [npts][3] grid indices ([npts][6] for kind 2) */
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

/* Dipole tables of DAS channels (kind 2): loc5 [npts][5] = x, y, z, dip, azimuth (rad;
 * fiber direction cos(dip)cos(az), cos(dip)sin(az), sin(dip)), gauge length gl > 0.
 * Replaces dipole_m3::findIndicesWeights (src/cpp/injector.cpp:313-420). Outputs:
 * ind [npts][6], w [npts][2][6]. An end outside the grid gives FWI3D_EINVAL. */
int fwi3d_dipole_tables(const float *loc5, int npts, float gl, int nx, int ny, int nz, float dx, float dy, float dz,
                        float ox, float oy, float oz, int *xind, int *yind, int *zind, float *xw, float *yw,
                        float *zw);

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

/* Allocation-free variant of fwi3d_propagate_device. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The wavefields, converted
 * model, flux buffer, taper and injector tables of the shot live in a static 128 MiB
 * device workspace that each GPU holds while the process runs; the call performs no
 * hipMalloc / hipFree and replays no graphs. Propagation without a snapshot store or
 * gradient only (par->sub must be 0; d_grad64 of the allocating entry point has no
 * counterpart here); a shot whose buffers exceed the workspace is rejected with
 * FWI3D_EINVAL. Every call on one GPU shares that workspace: calls that overlap on the
 * same GPU (other contexts, other threads) race and must be serialized by the caller.
 * Returns after the work has completed; results equal fwi3d_propagate_device's. */
int fwi3d_propagate_device_noalloc(fwi3d_ctx *ctx, int physics, int adj, const float *d_model, const float *d_src,
                                   float *d_rcv, const fwi3d_injector *inj, const fwi3d_injector *ext,
                                   const fwi3d_param *par, int nx, int ny, int nz, float dx, float dy, float dz);

/* Kernel timing with HIP events on the context's stream, for roofline accounting
 * (no counterpart in the reference). enable=1 starts collecting every class below (and
 * resets), enable=2 only whole time steps (class 4: fewer event records in the timed
 * work), 0 stops.
 * read() returns the summed milliseconds and launch count since the last enable of
 * kernel class `kernel`: 0 = the interior time-step kernel, 1 = the elastic/VTI flux
 * pre-pass, 2 = the boundary-shell time-step kernel (elastic/VTI: the shell's edges and
 * corners), 3 = the elastic/VTI slab kernel (the z slabs away from the x and y faces and
 * the x and y slabs away from the other faces), 4 = the whole time step. */
/* Tiling of the y-marching time-step kernels, per context (no counterpart in the
 * reference; the tile shapes an autotuner picks): FWI3D_TUNE_YCHUNK = y planes per
 * workgroup (0: default 8), FWI3D_TUNE_TX = x rows per workgroup (4 or 8).
 * FWI3D_TUNE_GRAPHS (default 1): time loops with no host work between steps (forward
 * propagation without snapshots, adjoint propagation without a gradient) replay a
 * captured hipGraph of 16 steps instead of launching each kernel.
 * FWI3D_TUNE_FUSED (default 1): elastic / VTI steps update the planes 6 cells inside the
 * y faces with the fused kernel (stresses and fluxes formed on chip; 0: the flux pass,
 * interior, slab and edge kernels over the whole grid). FWI3D_TUNE_FCHUNK: y planes per
 * workgroup of the fused kernel (0: chosen from the grid size). FWI3D_TUNE_OVERLAP
 * (default 3): the fused step's edge tiles run on a second stream of the context beside
 * its middle tiles, the y-face launches (their flux pass and slabs) after the middle
 * tiles (3); or the y-face launches run on the second stream, forked after the middle
 * tiles (2) or at the start of the step (1); or all on one stream (0). Results agree
 * to rounding between the two elastic forms and depend on none of the other settings.
 * FWI3D_TUNE_ECHUNK: y planes per workgroup of the fused kernel's side tiles (the tiles
 * at the x or z faces; 0: the middle tiles' chunk). */
enum fwi3d_tune_param {
    FWI3D_TUNE_YCHUNK = 0,
    FWI3D_TUNE_TX = 1,
    FWI3D_TUNE_GRAPHS = 2,
    FWI3D_TUNE_FUSED = 3,
    FWI3D_TUNE_FCHUNK = 4,
    FWI3D_TUNE_OVERLAP = 5,
    FWI3D_TUNE_ECHUNK = 6
};
int fwi3d_tune(fwi3d_ctx *ctx, int param, int value);
/* Autotuner (no counterpart in the reference): times `steps` time steps of the given
 * physics and grid on a synthetic model for every tile shape (TX 4 or 8, y chunk 4, 8,
 * 16 or 32, and for the fused elastic / VTI step the y chunk of its side tiles, about
 * 0.6 to 2.5 side workgroups per CU, and of its middle tiles, one chunk fewer or more
 * per tile column than the grid-size rule's; a chunk other than the rule's is kept when
 * it is at least 3 % faster) with HIP events, keeps the fastest for that
 * physics and grid on the context
 * (used by propagate and by batch admission in preference to fwi3d_tune's settings) and
 * returns it. */
int fwi3d_autotune(fwi3d_ctx *ctx, int physics, int nx, int ny, int nz, int steps, int *tx, int *ychunk);

/* The snapshot store of the last forward call with sub > 0 on the context, decoded to
 * FP32: out [slots][ncomp][ny][nx][nz] (slots = 1 + nt/sub, slot k = step k*sub), the
 * full wavefield of zfpWrapper::getFullWfld (src/include/zfpWrapper.hpp:118-128) that
 * WE_MODELING_3D.x writes with wavefield= (src/main/we_modeling.cpp:89-94). count: floats
 * available at out; out NULL: only *slots and *ncomp are returned. */
int fwi3d_snapshots_read(fwi3d_ctx *ctx, float *out, long long count, int *slots, int *ncomp);

int fwi3d_profile(fwi3d_ctx *ctx, int enable);
int fwi3d_profile_read(const fwi3d_ctx *ctx, int kernel, double *ms, long long *launches);

/* ---------------------------------------------------------------------------------
 * Shot operator: the nonlinear operator of one shot around propagate(), i.e.
 * nl_we_op_e::apply_forward / apply_jacobianT (src/include/we_op.hpp:50-51, 98-99;
 * src/cpp/we_op.cpp:872-1078), inherited by nl_we_op_a and nl_we_op_vti, plus the
 * per-shot part of nlls_fwi::compute_res_and_grad (src/cpp/optimization.cpp:5-197).
 * ------------------------------------------------------------------------------ */

/* The fields of the reference's `param` (src/include/param.hpp:8-71) that these
 * operators read or write. In/out fields are updated as analyzeModel does
 * (src/cpp/we_op.cpp:378-444), so passing the same struct to apply_forward and then
 * apply_jacobianT reproduces the reference's per-instance `_par`. */
typedef struct {
    int nmodels;                  /* 2 (vp, rho), 3 (vp, vs, rho), 6 (vp, vs, rho, delta, epsilon, gamma) */
    float dt;                     /* in/out: <0 CFL step rounded to divide the data dt, 0 CFL step,
                                     >0 min(dt, data dt); out: propagation dt */
    float courant;                /* in/out (default 0.5) */
    int sub;                      /* in/out: <0 round(data dt / dt), 0 no gradient, >0 imaging interval */
    int nt;                       /* out: propagation time samples */
    float vmin, vmax;             /* out: velocity extremes after clipping */
    int resampling;               /* 0 sinc (default), 1 linear */
    int sinc_half_length;         /* default 11 */
    int bc[6];                    /* top,bottom,left,right,front,back (default 1) */
    int taper[6];                 /* cos^2 taper widths (default 0) */
    float taper_strength;         /* default 0.05 */
    float free_surface_stiffness; /* default 1.05 (acoustic: at least 1) */
    int mt;                       /* 1: moment-tensor source (6 components) */
    int seismotype;               /* 0 displacement / pressure, 1 its time derivative */
    int soft_clip;                /* 1: skip the hard clip of elastic/VTI models */
    float vpmin, vpmax, vsmin, vsmax, rhomin, rhomax;
    float deltamin, deltamax, epsilonmin, epsilonmax, gammamin, gammamax;
    int snap_precision;           /* fwi3d_snap_precision */
    float gl;                     /* DAS gauge length (default 0: point receivers); > 0 (elastic,
                                     VTI): dipole channels, one strain component per channel */
} fwi3d_op_param;

/* Geometry of one shot: grid, one source position, receiver positions and the time
 * axis of the data (the domain/range hypercubes and par.sxyz/rxyz of the reference). */
typedef struct {
    int nx, ny, nz;
    float dx, dy, dz, ox, oy, oz;
    int nt_data;                  /* samples of source wavelets and data traces */
    float dt_data;
    const float *sxyz;            /* [3] source x, y, z (host) */
    int nrcv;
    const float *rxyz;            /* [nrcv][3] receiver x, y, z (host) */
    const float *rdipaz;          /* [nrcv][2] fiber dip and azimuth of DAS channels (host; read
                                     when gl > 0, else may be NULL) */
} fwi3d_shot;

/* Fill the defaults of src/include/param.hpp:8-71. */
void fwi3d_op_param_defaults(fwi3d_op_param *par);

/* analyzeModel (src/cpp/we_op.cpp:234-446) on a host model [nmodels][ny][nx][nz], in
 * place: hard clip, Thomsen-delta and vp >= sqrt(2) vs fixes, then the time axis
 * (par->dt, nt, sub, courant, vmin, vmax). */
int fwi3d_analyze_model(fwi3d_op_param *par, float *model, const fwi3d_shot *shot);

/* apply_forward (src/cpp/we_op.cpp:872-946): raw model [nmodels][ny][nx][nz] and source
 * wavelets src [nscomp][nt_data] (nscomp 1 acoustic, 3 force, 6 moment tensor) ->
 * data [nrcomp][nrcv][nt_data] (nrcomp 1 acoustic, 3 elastic), overwritten or added.
 * Stores the snapshots of the gradient when par->sub > 0 (after analysis). */
int fwi3d_apply_forward(fwi3d_ctx *ctx, fwi3d_op_param *par, const fwi3d_shot *shot, int add, const float *model,
                        const float *src, float *data);
/* apply_jacobianT (src/cpp/we_op.cpp:948-1078): data-space residual -> gradient w.r.t.
 * the raw (vp, vs, rho) or (vp, rho) parameters, grad [min(3,nmodels)][ny][nx][nz],
 * overwritten or added. Uses the snapshots of the last apply_forward on ctx. */
int fwi3d_apply_jacobianT(fwi3d_ctx *ctx, fwi3d_op_param *par, const fwi3d_shot *shot, int add, float *grad,
                          const float *model0, const float *data);

/* Device-pointer variants; the gradient is an FP64 accumulator. */
int fwi3d_apply_forward_device(fwi3d_ctx *ctx, fwi3d_op_param *par, const fwi3d_shot *shot, int add,
                               const float *d_model, const float *d_src, float *d_data);
int fwi3d_apply_jacobianT_device(fwi3d_ctx *ctx, fwi3d_op_param *par, const fwi3d_shot *shot, int add,
                                 double *d_grad64, const float *d_model0, const float *d_data);

/* One shot of nlls_fwi::compute_res_and_grad (src/cpp/optimization.cpp:35-188) without
 * weights, filters or normalisation: forward, residual r = syn - obs (copied to d_resid
 * when not NULL), the FP64 misfit contribution 0.5 * dt_data * sum_traces (sum_t r^2 -
 * (r_0^2 + r_{nt-1}^2)/2) (src/include/optimization.hpp:410-424, before the division by
 * dnorm), then apply_jacobianT(add=1) of Ht r into d_grad64. */
int fwi3d_fwi_shot_device(fwi3d_ctx *ctx, fwi3d_op_param *par, const fwi3d_shot *shot, const float *d_model,
                          const float *d_src, const float *d_obs, float *d_resid, double *d_grad64, double *misfit);

/* ---------------------------------------------------------------------------------
 * Optimiser: l-BFGS with the regular-Wolfe line search (lbfgs::run, updateSY,
 * computeHg, regular_wolfe::lineSearch/zoom: src/cpp/nlsolver.cpp:10-36, 185-423,
 * 1158-1359), on device vectors with FP64 dot products and FP64 curvature pairs.
 * ------------------------------------------------------------------------------ */

/* Objective callback: misfit f and gradient at d_model into d_grad (the caller's own
 * device buffers, those given to fwi3d_lbfgs_run), like nlls_fwi::res(), which
 * computes both (src/cpp/optimization.cpp:5-197). Return 0 on success. */
typedef int (*fwi3d_objective)(void *user, const float *d_model, float *d_grad, double *f);

/* Solver and line-search parameters (src/include/param.hpp:55-62 defaults). */
typedef struct {
    int niter;         /* maximum iterations */
    int m;             /* curvature pairs kept (lbfgs_m, default 5) */
    int max_trial;     /* line-search trials per iteration (default 10) */
    double threshold;  /* stop when the relative misfit decrease falls below it (default 0) */
    double c1, c2;     /* Wolfe constants (1e-4, 0.9) */
    double a0, a1;     /* first step a0/|g| or a1*|m|/|g| (1, 0) */
    double step_max;   /* 1e6 */
    int version;       /* trial-step rule after the first iteration (ls_version, 0) */
} fwi3d_lbfgs_param;

void fwi3d_lbfgs_param_defaults(fwi3d_lbfgs_param *p);

/* Minimise fn over the FP32 model d_model[n] (updated in place; d_grad holds the
 * gradient at the returned model). d_h0: optional positive diagonal initial inverse
 * Hessian (NULL: ones). f_history[niter+1] receives the misfit per accepted iterate,
 * counters[3] = {iterations, function evaluations, gradient evaluations}. */
int fwi3d_lbfgs_run(fwi3d_ctx *ctx, const fwi3d_lbfgs_param *par, long long n, float *d_model, float *d_grad,
                    const float *d_h0, fwi3d_objective fn, void *user, double *f_history, int *counters);

/* ---------------------------------------------------------------------------------
 * Batched forward modelling. Replaces the one-shot-at-a-time loops that call
 * apply_forward / propagate per shot (we_modeling.cpp:84-112; the shot loop of
 * nlls_fwi::compute_res_and_grad, optimization.cpp:33-190): shots of any physics and
 * grid size join a batch between time steps and advance together, each time step of the
 * whole batch being one launch per kernel class; a shot leaves after its last sample.
 * Per shot the arguments mean what they mean for fwi3d_propagate_device with adj = 0
 * (converted model, source, receiver buffer [ncomp][nrcv][nt]). With par->sub > 0 the
 * shot also fills a snapshot store (par->snap_precision), which outlives the shot until
 * an adjoint shot consumes it (fwi3d_batch_add_adjoint) or fwi3d_batch_release drops it.
 * The caller keeps d_model, d_src and d_rcv alive until fwi3d_batch_query reports the
 * shot done. Each shot's traces, and each adjoint shot's gradient, equal those of
 * fwi3d_propagate_device for the same inputs. Calls on one batch, and on the context it
 * was created on, must not overlap. */
typedef struct fwi3d_batch fwi3d_batch;
int fwi3d_batch_create(fwi3d_ctx *ctx, fwi3d_batch **out);
int fwi3d_batch_destroy(fwi3d_batch *b);
/* Device bytes one shot holds while in a batch (for admission by memory footprint);
 * -1 for invalid arguments. */
long long fwi3d_batch_footprint(int physics, int nx, int ny, int nz, const fwi3d_injector *inj,
                                const fwi3d_injector *ext);
/* Admit a shot; it starts at the batch's next step. *id identifies it in queries. */
int fwi3d_batch_add(fwi3d_batch *b, int physics, const float *d_model, const float *d_src, float *d_rcv,
                    const fwi3d_injector *inj, const fwi3d_injector *ext, const fwi3d_param *par, int nx, int ny,
                    int nz, float dx, float dy, float dz, int *id);
/* Device bytes of the snapshot store of a forward shot with imaging interval sub > 0
 * (held from its admission until its adjoint shot leaves); -1 for invalid arguments. */
long long fwi3d_batch_snap_footprint(int physics, int nx, int ny, int nz, int nt, int sub, int snap_precision);
/* Admit the adjoint shot of forward shot fwd_id (which must have left the batch):
 * fwi3d_propagate_device with adj = 1 -- d_src [ncomp][inj->npts][nt] injected at the
 * receivers inj, the imaging condition (+=) into d_grad64 [ng][ny][nx][nz] from fwd_id's
 * snapshots, par (nt, sub, snap_precision) and grid equal to the forward shot's. The
 * snapshot store is freed when the adjoint shot leaves. Replaces the adjoint
 * propagate(adj=true, grad) call of apply_jacobianT (src/cpp/we_op.cpp:1022-1031). */
int fwi3d_batch_add_adjoint(fwi3d_batch *b, int fwd_id, int physics, const float *d_model, const float *d_src,
                            const fwi3d_injector *inj, double *d_grad64, const fwi3d_param *par, int nx, int ny, int nz,
                            float dx, float dy, float dz, int *id);
/* Drop the snapshot store of forward shot fwd_id without an adjoint shot. */
int fwi3d_batch_release(fwi3d_batch *b, int fwd_id);
/* Enqueue up to nsteps time steps of every shot in the batch (asynchronous on the
 * context's stream); *nactive = shots still in the batch afterwards. */
int fwi3d_batch_step(fwi3d_batch *b, int nsteps, int *nactive);
/* Ids of the shots that left the batch during the last fwi3d_batch_step call (their
 * work is enqueued on the context's stream, so work the caller enqueues there afterwards
 * sees their results): *n ids, at most cap written to ids. */
int fwi3d_batch_retired(const fwi3d_batch *b, int *ids, int cap, int *n);
/* *state = 1 when shot id's traces are complete, 0 while it runs, -1 for an unknown id. */
int fwi3d_batch_query(fwi3d_batch *b, int id, int *state);
int fwi3d_batch_wait(fwi3d_batch *b);
int fwi3d_batch_status(const fwi3d_batch *b, int *nshots, long long *resident_bytes, int *step);
/* Online tile tuning of the batch (no counterpart in the reference): the next 4 x window
 * steps run with y chunks 4, 8, 16 and 32 in turn (window steps each, timed with HIP
 * events), then the fastest is kept for every shot, present and future. Results do not
 * depend on the tiling. fwi3d_batch_tiles returns the y chunk in force (0: per shot)
 * and, if cand_ms is not NULL, the measured milliseconds per step of the 4 candidates. */
int fwi3d_batch_autotune(fwi3d_batch *b, int window);
int fwi3d_batch_tiles(const fwi3d_batch *b, int *ychunk, float *cand_ms);

/* ---------------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, whose drivers run one operator per
 * shot to completion: we_modeling.cpp:84-112, optimization.cpp:33-190). A server thread
 * owns a batch on the context and keeps it full: queued requests are admitted in
 * arrival order as soon as their device footprint fits the HBM budget and one of
 * max_active slots is free, between two time steps of the shots already running
 * (continuous batching); model analysis and source resampling of new requests overlap
 * the running shots on a side stream. Each request computes what the one-shot entry
 * points compute for it alone: FWI3D_REQ_MODEL = fwi3d_apply_forward_device (add = 0),
 * FWI3D_REQ_GRADIENT = fwi3d_fwi_shot_device. While a server exists, the context belongs
 * to it: the caller makes no other calls on that context.
 * ------------------------------------------------------------------------------ */
enum fwi3d_request_kind { FWI3D_REQ_MODEL = 0, FWI3D_REQ_GRADIENT = 1 };

typedef struct {
    int kind;                  /* fwi3d_request_kind */
    fwi3d_op_param par;        /* operator parameters; physics from par.nmodels */
    fwi3d_shot shot;           /* geometry (the position arrays are copied at submission) */
    const float *d_model;      /* raw model [nmodels][ny][nx][nz] */
    const float *d_src;        /* source wavelets [nscomp][nt_data] */
    const float *d_obs;        /* GRADIENT: observed data [nrcomp][nrcv][nt_data] */
    float *d_data;             /* MODEL: data [nrcomp][nrcv][nt_data] (overwritten);
                                  GRADIENT: residual syn - obs, or NULL */
    double *d_grad64;          /* GRADIENT: FP64 gradient [min(3,nmodels)][ny][nx][nz] (+=) */
} fwi3d_request;

typedef struct {
    long long submitted, completed, failed;
    int queued, active;        /* requests waiting for admission / admitted and not done */
    long long hbm_in_use;      /* device bytes reserved by admitted requests */
    long long hbm_peak;
    int peak_active;
    long long steps;           /* batch time steps run */
} fwi3d_server_counters;

typedef struct fwi3d_server fwi3d_server;
/* hbm_budget: device bytes the admitted requests may hold together (their shots in the
 * batch, snapshot stores, converted models, traces and FP64 raw gradients). */
int fwi3d_server_create(fwi3d_ctx *ctx, long long hbm_budget, int max_active, fwi3d_server **out);
/* Queue a request (returns at once; device buffers stay owned by the caller and must
 * stay alive until the request is done). */
int fwi3d_server_submit(fwi3d_server *s, const fwi3d_request *r, int *id);
/* *state: 0 queued, 1 running, 2 done, -1 failed (fwi3d_server_error). misfit: the FP64
 * misfit contribution of a done gradient request; latency_s: submission to completion
 * (so far, if not done); queue_s: submission to admission. Any output may be NULL. */
int fwi3d_server_poll(fwi3d_server *s, int id, int *state, double *misfit, double *latency_s, double *queue_s);
/* Block until request id (id < 0: every submitted request) is done or failed. */
int fwi3d_server_wait(fwi3d_server *s, int id);
const char *fwi3d_server_error(fwi3d_server *s, int id);
/* The operator parameters of a done request after analyzeModel (dt, nt, sub, ...). */
int fwi3d_server_result(fwi3d_server *s, int id, fwi3d_op_param *par);
int fwi3d_server_stats(fwi3d_server *s, fwi3d_server_counters *st);
/* Finish every submitted request, stop the thread, free the server. */
int fwi3d_server_destroy(fwi3d_server *s);

#ifdef __cplusplus
}
#endif
#endif
