// Engine context and host-side helpers shared by the translation units of the
// library (engine.hip: propagation; operator.hip: the shot operator around it).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/fwi3d_hip.h"

namespace fwi3d {

struct Err {
    int code;
    std::string msg;
};

#define HIP_OK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            throw Err{FWI3D_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)};             \
    } while (0)

#define REQUIRE(cond, text)                                    \
    do {                                                       \
        if (!(cond)) throw Err{FWI3D_EINVAL, std::string(text)}; \
    } while (0)

// Growable device allocation. A borrowed buffer points into memory the library does not
// own (the static workspace of the allocation-free entry point): it is never freed or
// grown, and asking it for more than it holds is an error.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    void ensure(size_t nb) {
        if (nb <= bytes) return;
        if (borrowed)
            throw Err{FWI3D_ENOMEM, "workspace slice of " + std::to_string(bytes) + " bytes is too small for " +
                                        std::to_string(nb) + " bytes"};
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) throw Err{FWI3D_ENOMEM, "hipMalloc(" + std::to_string(nb) + " bytes) failed"};
        bytes = nb;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

// Message of the last error raised by a C-ABI call on the calling thread. Code that
// calls an entry point and turns its failure into an exception reads this, not the
// context's message, which another thread may overwrite at any time.
inline std::string& thread_error() {
    static thread_local std::string m;
    return m;
}

}  // namespace fwi3d

struct fwi3d_batch;
struct fwi3d_server;

struct fwi3d_ctx {
    int device = 0;
    std::vector<fwi3d_batch*> batches;    // live batches on this context (destroyed with it)
    std::vector<fwi3d_server*> servers;   // live servers (stopped and destroyed before the batches)
    // one-shot table of the single-shot time step (a BShot record, engine.hip): host
    // bytes kept until the next propagate call, and its device copy
    std::vector<char> shot1_host;
    fwi3d::DevBuf shot1_dev;
    hipStream_t stream = nullptr;
    // second stream of the fused elastic step (the y-face launches run on it beside the
    // fused tiles; default priority) and the fork / join events between the two
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // last error message on the context; a server thread and its clients both set it,
    // so it is only touched through fwi3d::set_error / fwi3d_last_error under err_mu
    std::mutex err_mu;
    std::string err;
    fwi3d::DevBuf fields, model, aux, snaps, tapers, injtab, exttab, tmpio;
    std::vector<int> injhost[2];  // host copies of the injection lists (source, receiver tables)
    // shot operator (operator.hip): converted model, reduction scalars, traces at the
    // propagation rate, raw FP64 gradient, synthetic data + misfit of an FWI shot
    fwi3d::DevBuf opmodel, opscal, optrace, opgrad, opsyn;
    fwi3d::DevBuf opdas;  // per-channel fiber trigonometry of DAS receivers
    // l-BFGS (lbfgs.hip): curvature pairs, directions, backups, reduction scratch
    fwi3d::DevBuf opt;
    // snapshot-store state (the per-instance state of the reference operator)
    fwi3d::DevBuf snapsc;  // per-slot max|u| (float bits) of 16-bit snapshot slots
    int snap_slots = 0, snap_ncomp = 0, snap_nt = 0, snap_sub = 0, snap_physics = -1, snap_fmt = 0;
    long long snap_n = 0;
    bool snap_valid = false;
    int snap_dims[3] = {0, 0, 0};
    // profiling: [0] interior time-step kernel, [1] flux pre-pass (elastic/VTI),
    // [2] boundary-shell time-step kernel, [3] slab kernel (elastic/VTI), [4] the whole
    // time step (its kernels, launched one after another)
    // tiling of the y-marching kernels (fwi3d_tune)
    int tune_ychunk = 0;  // y planes per workgroup (0: default 8)
    int tune_tx = 4;      // x rows per workgroup (4 or 8)
    // tiles chosen by fwi3d_autotune per physics and grid (they take precedence there)
    struct Tiles {
        int physics, nx, ny, nz, tx, ychunk;
        int echunk;  // side-tile y chunk of the fused elastic / VTI kernel (0: by grid size)
        int fchunk;  // its middle-tile y chunk (0: by grid size)
    };
    std::vector<Tiles> tuned;
    // time loops without host work between steps replay a captured graph of several steps
    // (fwi3d_tune FWI3D_TUNE_GRAPHS); the counter the graph's kernels read the step from
    bool use_graphs = true;
    // elastic / VTI steps: the fused kernel over the planes [6, ny-6) (fwi3d_tune
    // FWI3D_TUNE_FUSED, default on) and its y chunk (FWI3D_TUNE_FCHUNK, 0: by grid size)
    bool use_fused = true;
    int fused_chunk = 0;
    int fused_echunk = 0;  // y chunk of its side tiles (FWI3D_TUNE_ECHUNK, 0: fused_chunk's)
    // launches of the fused step on the second stream (FWI3D_TUNE_OVERLAP): 0 none, 1 the
    // y faces from the start of the step, 2 the y faces from the edge tiles on, 3 the edge
    // tiles from the start of the step
    int fused_overlap = 3;
    fwi3d::DevBuf stepctr;
    std::vector<hipGraphExec_t> graph_execs;  // instantiated in the current call, freed after it
    std::vector<hipGraph_t> graphs;
    bool prof = false;
    bool prof_steps_only = false;  // record only whole-step events (fwi3d_profile level 2)
    static constexpr int kProfClasses = 5;  // interior, flux, shell, z slabs, whole time step
    double prof_ms[kProfClasses] = {};
    long long prof_n[kProfClasses] = {};
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<int> ev_kind;
    std::vector<int> ev_count;  // time steps covered by each event pair
    int timed_count = 1;        // time steps the next timed launch covers
};

namespace fwi3d {

inline void set_error(fwi3d_ctx* c, const std::string& msg) {
    thread_error() = msg;
    if (!c) return;
    std::lock_guard<std::mutex> lk(c->err_mu);
    c->err = msg;
}

}  // namespace fwi3d
