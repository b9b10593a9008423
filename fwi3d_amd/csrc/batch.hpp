// Batched time stepping of many independent shots (forward modelling, uneven grids,
// any mix of the three physics). The reference models shots one after another: each
// pass of the shot loop of we_modeling.cpp:84-112 (and of compute_res_and_grad,
// optimization.cpp:33-190) builds an operator and runs propagate() to completion.
// Here every admitted shot is a set of workgroups inside the same launches: one time
// step of the whole batch is a fixed number of launches (flux, interior, shell and
// injection per physics present, one extraction), whatever the number of shots, and
// small grids that would leave most of the 256 CUs idle on their own share the chip.
//
// Device side (batch_kernels.hpp): a table of per-shot records (BShot) and per kernel
// class a map from launch block to shot with each shot's first block. The tables
// change only when shots join or leave. Shots join between steps and leave after their
// last extraction (continuous batching).
//
// Gradient shots run in the batch as two shots: a forward shot with a snapshot store
// (par->sub > 0), whose store outlives it, and, once it has left, an adjoint shot that
// reads that store and accumulates the imaging condition into an FP64 gradient
// (fwi3d_batch_add_adjoint). Snapshot writes and imaging events are per-shot launches
// placed between the batch's time steps exactly where propagate() places them.
//
// Included once, by engine.hip (all kernels share its translation unit).
#pragma once

// ======================================================================================
// Host side of the batch.

struct fwi3d_batch {
    enum Cls { C_AINT, C_ASHELL, C_AINJ, C_EFLUX, C_EINT, C_ESHELL, C_EZS, C_EINJ, C_VFLUX, C_VINT, C_VSHELL, C_VZS,
               C_VINJ, C_EXT, C_EFUSED, C_VFUSED, C_EFEDGE, C_VFEDGE, C_EYS, C_VYS, NCLS };
    struct Shot {
        int id = -1;
        fwi3d::BShot d{};
        int nblk[NCLS] = {};
        void* mem = nullptr;
        size_t bytes = 0;
        int mode = 0;         // 0 forward, 1 forward with snapshot store, 2 adjoint with gradient
        int sub = 0;          // imaging interval (modes 1, 2)
        float dt = 0.f;
        int store = -1;       // key (forward shot id) of the snapshot store (modes 1, 2)
        double* grad = nullptr;  // FP64 gradient accumulator (mode 2)
        std::shared_ptr<std::vector<int>> hinj, hext;  // host injection lists (read by async copies)
    };
    // Snapshot store of a gradient shot: written by the forward shot, read by the adjoint
    // shot, freed when the adjoint shot leaves (or on fwi3d_batch_release).
    struct Store {
        int key = -1;  // id of the forward shot
        void* mem = nullptr;
        size_t bytes = 0;
        SnapDesc S;
        int physics = 0, nt = 0, sub = 0, nx = 0, ny = 0, nz = 0;
        bool complete = false;  // the forward shot has left the batch
        bool taken = false;     // an adjoint shot reads it
    };
    std::vector<Store> stores;
    size_t store_bytes = 0;
    std::vector<int> last_retired;  // ids that left during the last fwi3d_batch_step call
    fwi3d_ctx* c = nullptr;
    int it = 0;       // global step counter
    int next_id = 0;
    int TX = 4;       // x rows per workgroup of every y-march launch of the batch
    std::vector<Shot> shots;                 // admitted, not yet retired
    std::vector<std::pair<int, hipEvent_t>> retired;  // id -> event after its last work (pending)
    std::vector<bool> done_ids;  // ids whose last work completed (their events are destroyed)
    bool dirty = true;
    fwi3d::DevBuf tab, maps;
    int cls_blocks[NCLS] = {};
    size_t cls_off[NCLS] = {};  // offset (ints) of the class's block->shot map; start array follows it
    std::vector<int> hmaps;
    std::vector<fwi3d::BShot> htab;
    size_t resident = 0;
    int force_yc = 0;  // y chunk of every shot (0: each shot's choice at admission)
    // online tile tuning: the candidate y chunks each run `tune_window` real steps of the
    // batch (results do not depend on the tiling), then the fastest is kept
    static constexpr int NCAND = 4;
    const int tune_cands[NCAND] = {4, 8, 16, 32};
    float tune_ms[NCAND] = {0, 0, 0, 0};
    int tune_window = 0, tune_idx = -1, tune_left = 0;
    hipEvent_t tev[2] = {nullptr, nullptr};
    // a captured graph of kGraphSteps steps of the current tables (forward-only batches),
    // replayed over the steps between joins and departures; dropped on every rebuild
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    fwi3d::DevBuf stepctr;
};

namespace {

using fwi3d::BShot;

int batch_group(int phys) { return phys == FWI3D_ACOUSTIC ? 0 : (phys == FWI3D_ELASTIC ? 1 : 2); }

// Device bytes of one batched shot: three wavefield sets, model, flux buffer, tapers,
// injector tables (256-byte aligned parts).
size_t batch_shot_bytes(int phys, const Geom& g, const fwi3d_injector* inj, const fwi3d_injector* ext) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const int ncomp = phys == FWI3D_ACOUSTIC ? 1 : 3;
    const int nmf = phys == FWI3D_ACOUSTIC ? 2 : (phys == FWI3D_ELASTIC ? 3 : 7);
    const int naux = phys == FWI3D_ACOUSTIC ? 0 : (phys == FWI3D_ELASTIC ? 7 : 9);
    const size_t n = (size_t)g.n;
    return al(3 * ncomp * n * 4) + al(nmf * n * 4) + al(naux * n * 4) + al((size_t)(g.nz + g.nx + g.ny) * 4) +
           al(inj_bytes(inj)) + al(inj_bytes(ext));
}

// Rebuild the device tables after shots joined or left.
void batch_drop_graph(fwi3d_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

void batch_rebuild(fwi3d_batch* b) {
    using B = fwi3d_batch;
    HIP_OK(hipStreamSynchronize(b->c->stream));  // the host copies below may still be in flight
    batch_drop_graph(b);  // its replays have completed
    const int ns = (int)b->shots.size();
    b->htab.resize(ns);
    for (int s = 0; s < ns; s++) b->htab[s] = b->shots[s].d;
    b->hmaps.clear();
    for (int k = 0; k < B::NCLS; k++) {
        b->cls_off[k] = b->hmaps.size();
        int total = 0;
        std::vector<int> start(ns, 0);
        for (int s = 0; s < ns; s++) {
            start[s] = total;
            for (int q = 0; q < b->shots[s].nblk[k]; q++) b->hmaps.push_back(s);
            total += b->shots[s].nblk[k];
        }
        b->cls_blocks[k] = total;
        b->hmaps.insert(b->hmaps.end(), start.begin(), start.end());
    }
    if (ns > 0) {
        b->tab.ensure(ns * sizeof(BShot));
        b->maps.ensure(b->hmaps.size() * sizeof(int) + 256);
        HIP_OK(hipMemcpyAsync(b->tab.p, b->htab.data(), ns * sizeof(BShot), hipMemcpyHostToDevice, b->c->stream));
        HIP_OK(hipMemcpyAsync(b->maps.p, b->hmaps.data(), b->hmaps.size() * sizeof(int), hipMemcpyHostToDevice,
                              b->c->stream));
    }
    b->dirty = false;
}

// Re-tile every shot of the batch with y chunk yc (block counts change, results do not).
void batch_retile(fwi3d_batch* b, int yc) {
    using B = fwi3d_batch;
    for (auto& sh : b->shots) {
        BShot& d = sh.d;
        if (d.phys == FWI3D_ACOUSTIC) {
            sh.nblk[B::C_AINT] = tile_grid_yc(d.a, d.g, b->TX, true, yc);
        } else if (!d.fused) {  // (the fused kernel keeps its own y chunk)
            const bool e = d.phys == FWI3D_ELASTIC;
            sh.nblk[e ? B::C_EFLUX : B::C_VFLUX] = tile_grid_yc(d.ef, d.g, b->TX, false, yc);
            sh.nblk[e ? B::C_EINT : B::C_VINT] = tile_grid_yc(d.ei, d.g, b->TX, true, yc);
        }
    }
    b->force_yc = yc;
    b->dirty = true;
}

fwi3d::BMap batch_map(const fwi3d_batch* b, int k) {
    const int* base = b->maps.as<int>() + b->cls_off[k];
    return fwi3d::BMap{base, base + b->cls_blocks[k]};
}

fwi3d_batch::Store* batch_store(fwi3d_batch* b, int key) {
    for (auto& st : b->stores)
        if (st.key == key) return &st;
    return nullptr;
}

// Per-shot work of gradient shots at the batch's current step, before its time-step
// kernels, in propagate()'s order: the forward shot stores slot k/sub when sub divides
// its local step k (k <= nt-1); the adjoint shot runs the imaging event of snapshot
// (nt-1-k)/sub when sub divides nt-1-k (0 < k < nt-1) and the event of snapshot 0 at its
// last step k = nt-1, with the current and the previous adjoint fields.
void batch_hooks(fwi3d_batch* b) {
    hipStream_t st = b->c->stream;
    for (auto& sh : b->shots) {
        if (sh.mode == 0) continue;
        const BShot& d = sh.d;
        const int k = b->it - d.t0, nt = d.nt;
        if (k < 0 || k > nt - 1) continue;
        const fwi3d_batch::Store* S = batch_store(b, sh.store);
        float* curr = d.F + (long long)((k + 1) % 3) * d.nf;
        const float* prev = d.F + (long long)(k % 3) * d.nf;
        if (sh.mode == 1) {
            if (k % sh.sub == 0) snap_put(st, S->S, k / sh.sub, curr);
            continue;
        }
        int itl = -1;
        if (k == nt - 1) itl = 0;
        else if (k != 0 && (nt - 1 - k) % sh.sub == 0) itl = (nt - 1 - k) / sh.sub;
        if (itl < 0) continue;
        const long long nunp = (long long)d.g.nx * d.g.ny * d.g.nz;
        if (d.phys == FWI3D_ACOUSTIC)
            grad_event_acoustic(st, S->S, itl, nt, sh.sub, sh.dt, curr, d.a.K, d.g, sh.grad, sh.grad + nunp);
        else
            grad_event_elastic(st, d.phys == FWI3D_VTI, S->S, itl, nt, sh.sub, sh.dt, curr, prev, d.ef, sh.grad, nunp);
    }
}

// One global step: extraction, then per physics the time step and the injection.
void batch_one_step(fwi3d_batch* b, StepIdx it) {
    using B = fwi3d_batch;
    using namespace fwi3d;
    hipStream_t st = b->c->stream;
    const BShot* T = b->tab.as<BShot>();
    auto go_on = [&](hipStream_t sq, int k, auto kern, dim3 blk) {
        if (b->cls_blocks[k] > 0) hipLaunchKernelGGL(kern, dim3(b->cls_blocks[k]), blk, 0, sq, T, batch_map(b, k), it);
    };
    auto go = [&](int k, auto kern, int threads, dim3 blk) {
        go_on(st, k, kern, blk);
        (void)threads;
    };
    go(B::C_EXT, extract_batch, 128, dim3(128));
    // When every elastic / VTI shot runs the fused step, its launches write disjoint points
    // of next and read no output of each other: as in the single-shot step, the fused edge
    // tiles run on the context's second stream beside the middle tiles (FWI3D_TUNE_OVERLAP
    // 3), or the y-face launches (flux pass of the 8 planes at each y face, then the y
    // slabs) beside the edge tiles (1, 2).
    fwi3d_ctx* c = b->c;
    bool ovl = c->fused_overlap != 0;
    for (const auto& sh : b->shots)
        if (sh.d.phys != FWI3D_ACOUSTIC && !sh.d.fused) ovl = false;
    auto yflux = [&](hipStream_t sq) {
        if (b->TX == 8) {
            go_on(sq, B::C_EFLUX, elastic_flux_batch<8, false>, dim3(64, 8));
            go_on(sq, B::C_VFLUX, elastic_flux_batch<8, true>, dim3(64, 8));
        } else {
            go_on(sq, B::C_EFLUX, elastic_flux_batch<4, false>, dim3(64, 4));
            go_on(sq, B::C_VFLUX, elastic_flux_batch<4, true>, dim3(64, 4));
        }
    };
    const bool edge_side = ovl && c->fused_overlap == 3;
    // flux passes, then the interiors, shells and z slabs (disjoint points of each shot's
    // next field), then the injections
    if (!ovl) yflux(st);
    if (edge_side) {
        HIP_OK(hipEventRecord(c->fork_ev, st));
        HIP_OK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
        go_on(c->side, B::C_EFEDGE, elastic_fused_batch<false, true>, dim3(FTHREADS));
        go_on(c->side, B::C_VFEDGE, elastic_fused_batch<true, true>, dim3(FTHREADS));
        HIP_OK(hipEventRecord(c->join_ev, c->side));
    }
    if (b->TX == 8) {
        go(B::C_AINT, acoustic_interior_batch<8>, 0, dim3(64, 8));
        go(B::C_EINT, elastic_interior_batch<8, false>, 0, dim3(64, 8));
        go(B::C_VINT, elastic_interior_batch<8, true>, 0, dim3(64, 8));
    } else {
        go(B::C_AINT, acoustic_interior_batch<4>, 0, dim3(64, 4));
        go(B::C_EINT, elastic_interior_batch<4, false>, 0, dim3(64, 4));
        go(B::C_VINT, elastic_interior_batch<4, true>, 0, dim3(64, 4));
    }
    go(B::C_EFUSED, elastic_fused_batch<false, false>, 0, dim3(FTHREADS));
    go(B::C_VFUSED, elastic_fused_batch<true, false>, 0, dim3(FTHREADS));
    if (edge_side) {
        yflux(st);
        go(B::C_EYS, elastic_yslabs_batch<false>, 0, dim3(256));
        go(B::C_VYS, elastic_yslabs_batch<true>, 0, dim3(256));
    } else {
        if (ovl) {
            HIP_OK(hipEventRecord(c->fork_ev, st));
            HIP_OK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
            yflux(c->side);
            go_on(c->side, B::C_EYS, elastic_yslabs_batch<false>, dim3(256));
            go_on(c->side, B::C_VYS, elastic_yslabs_batch<true>, dim3(256));
            HIP_OK(hipEventRecord(c->join_ev, c->side));
        }
        go(B::C_EFEDGE, elastic_fused_batch<false, true>, 0, dim3(FTHREADS));
        go(B::C_VFEDGE, elastic_fused_batch<true, true>, 0, dim3(FTHREADS));
    }
    go(B::C_ASHELL, acoustic_boundary_batch, 0, dim3(256));
    go(B::C_ESHELL, elastic_boundary_batch<false>, 0, dim3(256));
    go(B::C_VSHELL, elastic_boundary_batch<true>, 0, dim3(256));
    // slab tiles and shell edges share one launch per physics (class C_EZS / C_VZS; the
    // shell classes stay empty)
    go(B::C_EZS, elastic_edges_slabs_batch<false>, 0, dim3(256));
    go(B::C_VZS, elastic_edges_slabs_batch<true>, 0, dim3(256));
    if (ovl) {
        HIP_OK(hipStreamWaitEvent(st, c->join_ev, 0));
    } else {
        go(B::C_EYS, elastic_yslabs_batch<false>, 0, dim3(256));
        go(B::C_VYS, elastic_yslabs_batch<true>, 0, dim3(256));
    }
    go(B::C_AINJ, acoustic_inject_batch, 0, dim3(256));
    go(B::C_EINJ, elastic_inject_batch<false>, 0, dim3(256));
    go(B::C_VINJ, elastic_inject_batch<true>, 0, dim3(256));
    HIP_OK(hipGetLastError());
}

}  // namespace

extern "C" {

int fwi3d_batch_create(fwi3d_ctx* c, fwi3d_batch** out) {
    if (!c || !out) return FWI3D_EINVAL;
    *out = nullptr;
    fwi3d_batch* b = new fwi3d_batch();
    b->c = c;
    b->TX = c->tune_tx;
    c->batches.push_back(b);
    *out = b;
    return FWI3D_OK;
}

int fwi3d_batch_destroy(fwi3d_batch* b) {
    if (!b) return FWI3D_OK;
    auto& live = b->c->batches;
    live.erase(std::remove(live.begin(), live.end(), b), live.end());
    (void)hipSetDevice(b->c->device);
    (void)hipStreamSynchronize(b->c->stream);
    for (auto& s : b->shots)
        if (s.mem) (void)hipFreeAsync(s.mem, b->c->stream);
    for (auto& s : b->stores)
        if (s.mem) (void)hipFreeAsync(s.mem, b->c->stream);
    (void)hipStreamSynchronize(b->c->stream);
    for (auto& r : b->retired) (void)hipEventDestroy(r.second);
    for (auto& e : b->tev)
        if (e) (void)hipEventDestroy(e);
    batch_drop_graph(b);
    b->tab.release();
    b->maps.release();
    b->stepctr.release();
    delete b;
    return FWI3D_OK;
}

long long fwi3d_batch_footprint(int physics, int nx, int ny, int nz, const fwi3d_injector* inj,
                                const fwi3d_injector* ext) {
    if (!inj || !ext || nx < 12 || ny < 12 || nz < 12) return -1;
    if (physics != FWI3D_ACOUSTIC && physics != FWI3D_ELASTIC && physics != FWI3D_VTI) return -1;
    try {
        const Geom g = make_geom(nx, ny, nz, 1.f, 1.f, 1.f);
        return (long long)batch_shot_bytes(physics, g, inj, ext);
    } catch (const Err&) {
        return -1;
    }
}

}  // extern "C"

namespace {

// Snapshot slots of a gradient shot (slot k holds local step k*sub).
int batch_snap_slots(int nt, int sub) { return 1 + nt / sub; }

size_t batch_store_bytes(int physics, const Geom& g, int nt, int sub, int fmt) {
    const int slots = batch_snap_slots(nt, sub);
    const int ncomp = physics == FWI3D_ACOUSTIC ? 1 : 3;
    return (snap_store_bytes(slots, ncomp, g.n, fmt) + 255) / 256 * 256 + ((size_t)slots * 4 + 255) / 256 * 256;
}

// Admit a shot. mode 0: forward; 1: forward with a snapshot store (par->sub > 0);
// 2: adjoint reading the store of forward shot `key`, accumulating into d_grad.
int batch_add_impl(fwi3d_batch* b, int mode, int physics, const float* d_model, const float* d_src, float* d_rcv,
                   const fwi3d_injector* inj, const fwi3d_injector* ext, const fwi3d_param* par, int nx, int ny, int nz,
                   float dx, float dy, float dz, int key, double* d_grad, int* id) {
    if (!b) return FWI3D_EINVAL;
    fwi3d_ctx* c = b->c;
    void* mem = nullptr;
    void* smem = nullptr;
    try {
        using namespace fwi3d;
        validate(physics, par, nx, ny, nz, dx, dy, dz);
        REQUIRE(d_model && d_src, "null model or source");
        check_injector(inj, nx, ny, nz, "source");
        check_injector(ext, nx, ny, nz, "receiver");
        REQUIRE(d_rcv || ext->npts == 0, "null receiver buffer");
        fwi3d_batch::Store* src_store = nullptr;
        if (mode == 2) {
            REQUIRE(d_grad != nullptr, "null gradient accumulator");
            REQUIRE(par->sub > 0, "an adjoint shot needs sub > 0 (the imaging interval of its forward shot)");
            src_store = batch_store(b, key);
            REQUIRE(src_store != nullptr, "no snapshot store of forward shot " + std::to_string(key) + " in this batch");
            REQUIRE(src_store->complete, "forward shot " + std::to_string(key) + " has not left the batch yet");
            REQUIRE(!src_store->taken, "the snapshots of forward shot " + std::to_string(key) + " are already in use");
            REQUIRE(src_store->physics == physics && src_store->nt == par->nt && src_store->sub == par->sub &&
                        src_store->nx == nx && src_store->ny == ny && src_store->nz == nz &&
                        src_store->S.fmt == par->snap_precision,
                    "adjoint shot does not match its forward shot (physics, grid, nt, sub or snap_precision)");
        }
        HIP_OK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        const Geom g = make_geom(nx, ny, nz, dx, dy, dz);
        const size_t bytes = batch_shot_bytes(physics, g, inj, ext);
        if (hipMallocAsync(&mem, bytes, st) != hipSuccess || !mem)
            throw Err{FWI3D_ENOMEM, "hipMallocAsync(" + std::to_string(bytes) + " bytes) for a batched shot failed"};
        fwi3d_batch::Store store;
        if (mode == 1) {
            const int ncomp = physics == FWI3D_ACOUSTIC ? 1 : 3;
            const size_t sb = batch_store_bytes(physics, g, par->nt, par->sub, par->snap_precision);
            if (hipMallocAsync(&smem, sb, st) != hipSuccess || !smem)
                throw Err{FWI3D_ENOMEM, "hipMallocAsync(" + std::to_string(sb) + " bytes) for a snapshot store failed"};
            store.mem = smem;
            store.bytes = sb;
            store.S.slots = batch_snap_slots(par->nt, par->sub);
            store.S.ncomp = ncomp;
            store.S.n = g.n;
            store.S.fmt = par->snap_precision;
            store.S.data = (char*)smem;
            store.S.scale = (unsigned*)((char*)smem + (snap_store_bytes(store.S.slots, ncomp, g.n, store.S.fmt) + 255) /
                                                          256 * 256);
            store.physics = physics; store.nt = par->nt; store.sub = par->sub;
            store.nx = nx; store.ny = ny; store.nz = nz;
            // scale words start at zero (max-abs accumulation); slots past the last stored
            // step, (nt-1)/sub, are never written and read as zero
            HIP_OK(hipMemsetAsync(store.S.scale, 0, (size_t)store.S.slots * 4, st));
            const int first_unwritten = (par->nt - 1) / par->sub + 1;
            if (first_unwritten < store.S.slots) {
                const size_t se = store.S.slot_elems() * store.S.esz();
                HIP_OK(hipMemsetAsync(store.S.data + (size_t)first_unwritten * se, 0,
                                      (size_t)(store.S.slots - first_unwritten) * se, st));
            }
        }
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        const long long n = g.n;
        const int ncomp = physics == FWI3D_ACOUSTIC ? 1 : 3;
        const int npar = physics == FWI3D_ACOUSTIC ? 2 : (physics == FWI3D_ELASTIC ? 3 : 6);
        const int nmf = physics == FWI3D_ACOUSTIC ? 2 : (physics == FWI3D_ELASTIC ? 3 : 7);
        const int naux = physics == FWI3D_ACOUSTIC ? 0 : (physics == FWI3D_ELASTIC ? 7 : 9);
        char* p = (char*)mem;
        float* F = (float*)p;          p += al(3 * ncomp * n * 4);
        float* M = (float*)p;          p += al(nmf * n * 4);
        float* aux = (float*)p;        p += al(naux * n * 4);
        float* tp = (float*)p;         p += al((size_t)(g.nz + g.nx + g.ny) * 4);
        char* pinj = p;                p += al(inj_bytes(inj));
        char* pext = p;
        HIP_OK(hipMemsetAsync(F, 0, 3 * ncomp * n * 4, st));
        HIP_OK(hipMemsetAsync(M, 0, nmf * n * 4, st));
        to_padded_on(st, M, d_model, npar, g);
        const Tapers T = tapers_to(st, tp, par, g);
        fwi3d_batch::Shot s;
        s.id = b->next_id++;
        s.mem = mem;
        s.bytes = bytes;
        BShot& d = s.d;
        d.phys = physics;
        d.t0 = b->it;
        d.nt = par->nt;
        d.nf = ncomp * n;
        d.F = F;
        d.g = g;
        s.hinj = std::make_shared<std::vector<int>>();
        s.hext = std::make_shared<std::vector<int>>();
        d.I = upload_inj_to(st, pinj, inj, *s.hinj);
        d.E = upload_inj_to(st, pext, ext, *s.hext);
        d.src = d_src;
        d.rcv = d_rcv;
        const float dt = par->dt;
        int tx_tuned, yc, echunk, fchunk;  // the batch's launches share one TX; the y chunks are per shot
        tiles_for(c, physics, g, tx_tuned, yc, echunk, fchunk);
        if (b->force_yc > 0) yc = b->force_yc;
        const int TX = b->TX;
        const int nshell = shell_count(6, g.nx, g.ny, g.nz);
        d.nshell_blk = (nshell + 255) / 256;
        d.ninj_blk = (d.I.ntgt + 255) / 256;
        d.next_blk = (ext->npts + 127) / 128;
        using B = fwi3d_batch;
        const int grp = batch_group(physics);
        if (physics == FWI3D_ACOUSTIC) {
            AcuArgs& A = d.a;
            A.K = M; A.b = M + n; A.tz2 = T.tz2; A.tx2 = T.tx2; A.ty2 = T.ty2; A.g = g;
            for (int f = 0; f < 6; f++) A.bc[f] = par->bc[f];
            A.dt = dt; A.dt2 = dt * dt;
            A.alpha = 0.2508560249f / par->free_surface_stiffness;
            s.nblk[B::C_AINT] = tile_grid_yc(A, g, TX, true, yc);
            d.nshell_pts = nshell;
            s.nblk[B::C_ASHELL] = d.nshell_blk;
            s.nblk[B::C_AINJ] = d.ninj_blk;
            A.prev = F; A.curr = F + n; A.next = F + 2 * n;
            if (d.I.ntgt > 0)
                hipLaunchKernelGGL(acoustic_inject, dim3(d.ninj_blk), dim3(256), 0, st, F, A, d.I, d_src, d.nt, 0, 1);
        } else {
            const bool vti = physics == FWI3D_VTI;
            EArgs& A = d.ef;
            A.lam = M; A.mu = M + n; A.rho = M + 2 * n;
            if (vti) {
                A.c13 = M + 3 * n; A.eps = M + 4 * n; A.c66 = M + 5 * n; A.c11b = M + 6 * n;
                hipLaunchKernelGGL(vti_c11b_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M + 6 * n, A.eps,
                                   A.mu, n);
            }
            A.tz2 = T.tz2; A.tx2 = T.tx2; A.ty2 = T.ty2; A.g = g;
            for (int f = 0; f < 6; f++) A.bc[f] = par->bc[f];
            A.aux = aux;
            A.dt = dt; A.dt2 = dt * dt;
            d.ei = A;
            const int nf = tile_grid_yc(A, g, TX, false, yc);
            const int ni = tile_grid_yc(d.ei, g, TX, true, yc);
            s.nblk[grp == 1 ? B::C_EFLUX : B::C_VFLUX] = nf;
            s.nblk[grp == 1 ? B::C_EINT : B::C_VINT] = ni;
            d.nshell_pts = shell_edge_count(g);
            d.nshell_blk = (d.nshell_pts + 255) / 256;
            s.nblk[grp == 1 ? B::C_ESHELL : B::C_VSHELL] = 0;  // the edges run in the slab class
            s.nblk[grp == 1 ? B::C_EZS : B::C_VZS] = zslab_tiles(d.zs, g) + 3 * d.nshell_blk;
            if (c->use_fused) {
                // the single-shot path's plan with the context's fused y chunks, not the
                // autotuned ones: those are timed for one shot alone (a VTI shot's edge tiles
                // in 4 long chunks, fewer workgroups than CUs beside its middle tiles, made
                // the 32-shot batch 1.3x slower). The y chunks do not change a result bit.
                (void)fchunk;
                (void)echunk;
                const FusedPlan P = fused_plan(d, g, TX, yc, c->fused_chunk, c->fused_echunk);
                s.nblk[grp == 1 ? B::C_EFLUX : B::C_VFLUX] = P.nflux;
                s.nblk[grp == 1 ? B::C_EINT : B::C_VINT] = 0;
                s.nblk[grp == 1 ? B::C_EFUSED : B::C_VFUSED] = P.nfused;
                s.nblk[grp == 1 ? B::C_EFEDGE : B::C_VFEDGE] = P.nfedge;
                s.nblk[grp == 1 ? B::C_EZS : B::C_VZS] = 0;
                s.nblk[grp == 1 ? B::C_EYS : B::C_VYS] = P.nslab;
            }
            s.nblk[grp == 1 ? B::C_EINJ : B::C_VINJ] = 3 * d.ninj_blk;
            A.prev = F; A.curr = F + 3 * n; A.next = F + 6 * n;
            if (d.I.ntgt > 0) {
                if (vti)
                    hipLaunchKernelGGL(elastic_inject<true>, dim3(d.ninj_blk, 3), dim3(256), 0, st, F, A, d.I, d_src, d.nt, 0, 1);
                else
                    hipLaunchKernelGGL(elastic_inject<false>, dim3(d.ninj_blk, 3), dim3(256), 0, st, F, A, d.I, d_src, d.nt, 0, 1);
            }
        }
        s.nblk[B::C_EXT] = ncomp * d.next_blk;
        HIP_OK(hipGetLastError());
        s.mode = mode;
        s.sub = par->sub;
        s.dt = dt;
        s.grad = d_grad;
        if (mode == 1) {
            store.key = s.id;
            s.store = s.id;
            b->stores.push_back(store);
            b->store_bytes += store.bytes;
            smem = nullptr;  // owned by the store from here
        } else if (mode == 2) {
            s.store = key;
            src_store->taken = true;
        }
        mem = nullptr;  // owned by the shot from here
        b->shots.push_back(s);
        b->resident += bytes;
        b->dirty = true;
        if (id) *id = s.id;
        return FWI3D_OK;
    } catch (const Err& e) {
        if (mem) (void)hipFreeAsync(mem, c->stream);
        if (smem) (void)hipFreeAsync(smem, c->stream);
        return fail(c, e);
    }
}

// Drop the retired shots whose last work has completed: their events are destroyed and
// only their ids are kept (a bit each), so a long-lived batch (a server's) holds events
// for the shots still in flight only.
void batch_prune_retired(fwi3d_batch* b) {
    for (size_t k = 0; k < b->retired.size();) {
        if (hipEventQuery(b->retired[k].second) != hipSuccess) {  // not ready, or an error the query reports
            k++;
            continue;
        }
        const int id = b->retired[k].first;
        if ((size_t)id >= b->done_ids.size()) b->done_ids.resize((size_t)id + 1, false);
        b->done_ids[(size_t)id] = true;
        (void)hipEventDestroy(b->retired[k].second);
        b->retired.erase(b->retired.begin() + (long)k);
    }
}

}  // namespace

extern "C" {

int fwi3d_batch_add(fwi3d_batch* b, int physics, const float* d_model, const float* d_src, float* d_rcv,
                    const fwi3d_injector* inj, const fwi3d_injector* ext, const fwi3d_param* par, int nx, int ny,
                    int nz, float dx, float dy, float dz, int* id) {
    if (!par) return b ? fail(b->c, Err{FWI3D_EINVAL, "null parameters"}) : FWI3D_EINVAL;
    return batch_add_impl(b, par->sub > 0 ? 1 : 0, physics, d_model, d_src, d_rcv, inj, ext, par, nx, ny, nz, dx, dy,
                          dz, -1, nullptr, id);
}

int fwi3d_batch_add_adjoint(fwi3d_batch* b, int fwd_id, int physics, const float* d_model, const float* d_src,
                            const fwi3d_injector* inj, double* d_grad64, const fwi3d_param* par, int nx, int ny, int nz,
                            float dx, float dy, float dz, int* id) {
    fwi3d_injector none{};
    return batch_add_impl(b, 2, physics, d_model, d_src, nullptr, inj, &none, par, nx, ny, nz, dx, dy, dz, fwd_id,
                          d_grad64, id);
}

int fwi3d_batch_release(fwi3d_batch* b, int fwd_id) {
    if (!b) return FWI3D_EINVAL;
    auto& v = b->stores;
    for (size_t q = 0; q < v.size(); q++)
        if (v[q].key == fwd_id) {
            if (!v[q].complete || v[q].taken)
                return fail(b->c, Err{FWI3D_ESTATE, "the snapshots of shot " + std::to_string(fwd_id) +
                                                        " are still being written or read"});
            const hipError_t e = hipFreeAsync(v[q].mem, b->c->stream);
            b->store_bytes -= v[q].bytes;
            v.erase(v.begin() + q);
            if (e != hipSuccess) return fail(b->c, Err{FWI3D_EDEVICE, std::string("hipFreeAsync: ") + hipGetErrorString(e)});
            return FWI3D_OK;
        }
    return fail(b->c, Err{FWI3D_EINVAL, "no snapshot store of shot " + std::to_string(fwd_id)});
}

int fwi3d_batch_retired(const fwi3d_batch* b, int* ids, int cap, int* n) {
    if (!b || !n || cap < 0 || (cap > 0 && !ids)) return FWI3D_EINVAL;
    *n = (int)b->last_retired.size();
    for (int k = 0; k < *n && k < cap; k++) ids[k] = b->last_retired[k];
    return FWI3D_OK;
}

long long fwi3d_batch_snap_footprint(int physics, int nx, int ny, int nz, int nt, int sub, int snap_precision) {
    if (nx < 12 || ny < 12 || nz < 12 || nt < 2 || sub < 1 || snap_precision < 0 || snap_precision > 2) return -1;
    if (physics != FWI3D_ACOUSTIC && physics != FWI3D_ELASTIC && physics != FWI3D_VTI) return -1;
    try {
        const Geom g = make_geom(nx, ny, nz, 1.f, 1.f, 1.f);
        return (long long)batch_store_bytes(physics, g, nt, sub, snap_precision);
    } catch (const Err&) {
        return -1;
    }
}

// Batches whose fused elastic / VTI steps run on two streams and that hold 3 Mi cells or
// more launch their steps directly, as single shots of that size do: replayed from a
// captured graph, the two-stream step of the 32 VTI shots of the batched configuration
// took 7.01 against 6.44 ms launched directly.
bool batch_launch_directly(const fwi3d_batch* b) {
    if (b->c->fused_overlap == 0) return false;
    bool fused = false;
    long long cells = 0;
    for (const auto& sh : b->shots) {
        if (sh.d.phys != FWI3D_ACOUSTIC && !sh.d.fused) return false;
        fused = fused || sh.d.fused;
        cells += (long long)sh.d.g.nx * sh.d.g.ny * sh.d.g.nz;
    }
    return fused && cells >= (3LL << 20);
}

int fwi3d_batch_step(fwi3d_batch* b, int nsteps, int* nactive) {
    if (!b) return FWI3D_EINVAL;
    fwi3d_ctx* c = b->c;
    try {
        REQUIRE(nsteps >= 0, "nsteps must be >= 0");
        HIP_OK(hipSetDevice(c->device));
        b->last_retired.clear();
        batch_prune_retired(b);
        for (int q = 0; q < nsteps && !b->shots.empty(); q++) {
            const bool tuning = b->tune_idx >= 0;
            if (tuning && b->tune_left == b->tune_window) {  // a candidate's window starts
                batch_retile(b, b->tune_cands[b->tune_idx]);
                batch_rebuild(b);
                HIP_OK(hipEventRecord(b->tev[0], c->stream));
            }
            if (b->dirty) batch_rebuild(b);
            // a window of steps with no host work (no gradient shot, no tuning, no departure
            // before its last step) replays the captured graph; the last step of the window
            // may be a departure step, handled below
            int quiet = tuning || !c->use_graphs || batch_launch_directly(b) ? 0 : nsteps - q;
            for (const auto& sh : b->shots) {
                if (sh.mode != 0) quiet = 0;
                quiet = std::min(quiet, sh.d.t0 + sh.d.nt - 1 - b->it + 1);
            }
            const int K = kGraphSteps;
            if (quiet >= 2 * K) {
                if (!b->gexec) {
                    b->stepctr.ensure(256);
                    int* ctr = b->stepctr.as<int>();
                    try {
                        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                        for (int j = 0; j < K; j++) batch_one_step(b, StepIdx(ctr, j));
                        hipLaunchKernelGGL(advance_step, dim3(1), dim3(64), 0, c->stream, ctr, K);
                        HIP_OK(hipStreamEndCapture(c->stream, &b->graph));
                    } catch (...) {
                        hipGraph_t g2 = nullptr;
                        (void)hipStreamEndCapture(c->stream, &g2);
                        if (g2) (void)hipGraphDestroy(g2);
                        throw;
                    }
                    HIP_OK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
                }
                const int W = quiet / K;
                hipLaunchKernelGGL(set_step, dim3(1), dim3(64), 0, c->stream, b->stepctr.as<int>(), b->it);
                for (int w = 0; w < W; w++) HIP_OK(hipGraphLaunch(b->gexec, c->stream));
                b->it += W * K - 1;  // the window's last step
                q += W * K - 1;
            } else {
                batch_hooks(b);
                batch_one_step(b, StepIdx(b->it));
            }
            // retire the shots whose last extraction (local step nt-1) was this step
            bool left = false;
            for (size_t s = 0; s < b->shots.size();) {
                auto& sh = b->shots[s];
                if (b->it - sh.d.t0 >= sh.d.nt - 1) {
                    hipEvent_t ev;
                    HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                    HIP_OK(hipEventRecord(ev, c->stream));
                    HIP_OK(hipFreeAsync(sh.mem, c->stream));
                    b->resident -= sh.bytes;
                    if (sh.mode == 1) {
                        batch_store(b, sh.store)->complete = true;
                    } else if (sh.mode == 2) {  // the store's last reader has left
                        auto& v = b->stores;
                        for (size_t q = 0; q < v.size(); q++)
                            if (v[q].key == sh.store) {
                                HIP_OK(hipFreeAsync(v[q].mem, c->stream));
                                b->store_bytes -= v[q].bytes;
                                v.erase(v.begin() + q);
                                break;
                            }
                    }
                    b->retired.emplace_back(sh.id, ev);
                    b->last_retired.push_back(sh.id);
                    b->shots.erase(b->shots.begin() + s);
                    left = true;
                } else {
                    s++;
                }
            }
            if (left) b->dirty = true;
            b->it++;
            if (tuning && (--b->tune_left == 0 || b->shots.empty())) {  // the window ends
                HIP_OK(hipEventRecord(b->tev[1], c->stream));
                HIP_OK(hipEventSynchronize(b->tev[1]));
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, b->tev[0], b->tev[1]));
                b->tune_ms[b->tune_idx] = ms / (b->tune_window - b->tune_left);
                if (++b->tune_idx == fwi3d_batch::NCAND || b->shots.empty()) {
                    int best = 0;
                    for (int k = 1; k < b->tune_idx; k++)
                        if (b->tune_ms[k] < b->tune_ms[best]) best = k;
                    batch_retile(b, b->tune_cands[best]);
                    b->tune_idx = -1;
                } else {
                    b->tune_left = b->tune_window;
                }
            }
        }
        if (nactive) *nactive = (int)b->shots.size();
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_batch_autotune(fwi3d_batch* b, int window) {
    if (!b) return FWI3D_EINVAL;
    fwi3d_ctx* c = b->c;
    try {
        REQUIRE(window >= 1, "window must be >= 1 step");
        HIP_OK(hipSetDevice(c->device));
        for (auto& e : b->tev)
            if (!e) HIP_OK(hipEventCreate(&e));
        b->tune_window = window;
        b->tune_left = window;
        b->tune_idx = 0;
        for (auto& t : b->tune_ms) t = 0;
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_batch_tiles(const fwi3d_batch* b, int* ychunk, float* cand_ms) {
    if (!b) return FWI3D_EINVAL;
    if (ychunk) *ychunk = b->force_yc;
    if (cand_ms)
        for (int k = 0; k < fwi3d_batch::NCAND; k++) cand_ms[k] = b->tune_ms[k];
    return FWI3D_OK;
}

int fwi3d_batch_query(fwi3d_batch* b, int id, int* state) {
    if (!b || !state) return FWI3D_EINVAL;
    for (auto& s : b->shots)
        if (s.id == id) {
            *state = 0;
            return FWI3D_OK;
        }
    for (auto& r : b->retired)
        if (r.first == id) {
            const hipError_t e = hipEventQuery(r.second);
            if (e == hipSuccess) *state = 1;
            else if (e == hipErrorNotReady) *state = 0;
            else return fail(b->c, Err{FWI3D_EDEVICE, std::string("hipEventQuery: ") + hipGetErrorString(e)});
            if (e == hipSuccess) batch_prune_retired(b);
            return FWI3D_OK;
        }
    *state = (id >= 0 && (size_t)id < b->done_ids.size() && b->done_ids[(size_t)id]) ? 1 : -1;
    return FWI3D_OK;
}

int fwi3d_batch_wait(fwi3d_batch* b) {
    if (!b) return FWI3D_EINVAL;
    const hipError_t e = hipStreamSynchronize(b->c->stream);
    if (e != hipSuccess) return fail(b->c, Err{FWI3D_EDEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e)});
    return FWI3D_OK;
}

int fwi3d_batch_status(const fwi3d_batch* b, int* nshots, long long* resident_bytes, int* step) {
    if (!b) return FWI3D_EINVAL;
    if (nshots) *nshots = (int)b->shots.size();
    if (resident_bytes) *resident_bytes = (long long)(b->resident + b->store_bytes);
    if (step) *step = b->it;
    return FWI3D_OK;
}

}  // extern "C"
