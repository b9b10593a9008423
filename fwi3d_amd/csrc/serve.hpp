// Serving loop: shot-modelling and shot-gradient requests of any physics and grid size
// on one device, with continuous batching. The reference serves nothing: its drivers run
// one operator per shot to completion (we_modeling.cpp:84-112; the shot loop of
// nlls_fwi::compute_res_and_grad, optimization.cpp:33-190). Here a server thread owns a
// batch (batch.hpp) and keeps it full:
//
//   queue --(HBM footprint fits, slot free)--> admitted: analyzeModel/convert_model and
//   source resampling on a side stream, then the forward shot joins the batch
//   --(forward shot leaves)--> modelling: Dt / decimation to the data axis, done;
//                              gradient: residual, FP64 misfit, Ht, adjoint source, and
//                              the adjoint shot joins the batch
//   --(adjoint shot leaves)--> H quadrature and chain rule into the caller's FP64
//                              gradient, done.
//
// Requests that arrive while others run are admitted at the next time step as memory
// and slots free up; every admitted shot advances in the same launches as the others,
// and the preparation of new requests overlaps the batch's kernels on the side stream.
// Each request computes exactly what fwi3d_apply_forward_device /
// fwi3d_fwi_shot_device compute for it alone (same kernels, same order of operations).
//
// Included once, by operator.hip (it uses that file's kernels and host helpers).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

struct fwi3d_server {
    struct Req {
        int id = -1;
        fwi3d_request r{};
        std::vector<float> sxyz, rxyz;  // copies of the caller's position arrays
        fwi3d_op_param par{};           // normalized; time axis after analysis
        ShotDims d{};
        fwi3d_param pp{};
        HostInj si, ri;
        int state = 0;   // 0 queued, 1 running, 2 done, -1 failed
        int phase = 0;   // 0 waiting, 1 analysed, 2 forward in the batch, 3 adjoint in the batch, 4 finishing
        void* mconv = nullptr;  // converted model + reduction words (allocated for the analysis)
        void* mrest = nullptr;  // traces at the propagation rate, synthetic data, FP64 raw gradient, misfit
        size_t bconv = 0, brest = 0, reserved = 0;  // device bytes held / reserved for the request
        float *conv = nullptr, *srcp = nullptr, *rcvp = nullptr, *tmp = nullptr, *syn = nullptr;
        unsigned* red = nullptr;
        double *g64 = nullptr, *dmis = nullptr;
        int shot = -1;  // current shot id in the batch
        hipEvent_t done = nullptr;
        double t_submit = 0, t_admit = 0, t_done = 0;
        double misfit = 0;
        std::string err;
    };

    fwi3d_ctx* c = nullptr;
    fwi3d_batch* b = nullptr;
    hipStream_t prep = nullptr;  // side stream (default priority) for request preparation
    hipEvent_t prep_ev = nullptr;
    unsigned* keys_host = nullptr;  // pinned: velocity-extreme keys of the analysis
    long long budget = 0;
    int max_active = 0;
    int lead = 4;  // time steps the host may run ahead of the device

    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::string fatal;  // set when the device failed; every later request fails

    int next_id = 0;
    std::map<int, std::unique_ptr<Req>> reqs;  // every request until the server is destroyed
    std::deque<Req*> queue;                    // submitted, not admitted (server thread pops)
    std::vector<Req*> active;                  // admitted, not done (server thread only)
    std::map<int, Req*> by_shot;               // batch shot id -> request
    std::deque<hipEvent_t> inflight;           // events after recent steps (throttle)
    std::vector<hipEvent_t> ev_pool;

    // statistics (under mu)
    long long submitted = 0, completed = 0, failed = 0, steps = 0;
    long long in_use = 0, peak_in_use = 0;
    int peak_active = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

using Srv = fwi3d_server;
using SReq = fwi3d_server::Req;

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

hipEvent_t srv_event(Srv* s) {
    if (!s->ev_pool.empty()) {
        hipEvent_t e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

// Bytes of the request's own buffers after the analysis (traces at the propagation
// rate, synthetic data and raw gradient of a gradient request).
size_t req_rest_bytes(const SReq& q) {
    const int nt = q.par.nt;
    const size_t ntr = (size_t)q.d.ncomp * q.r.shot.nrcv;
    size_t b = al256((size_t)q.d.nscomp * nt * 4) + 2 * al256(std::max<size_t>(1, ntr * nt) * 4);
    if (q.r.kind == FWI3D_REQ_GRADIENT)
        b += al256(std::max<size_t>(1, (size_t)q.d.nrcomp * q.r.shot.nrcv * q.r.shot.nt_data) * 4) +
             al256((size_t)q.d.ng * q.d.N * 8) + 256;
    return b;
}

// Device bytes the request's shots hold in the batch (one shot at a time; a gradient
// request adds its snapshot store).
size_t req_batch_bytes(const SReq& q) {
    long long b = fwi3d_batch_footprint(q.d.physics, q.d.nx, q.d.ny, q.d.nz, &q.si.t, &q.ri.t);
    REQUIRE(b > 0, "request grid not admissible to a batch");
    if (q.r.kind == FWI3D_REQ_GRADIENT) {
        const long long sb = fwi3d_batch_snap_footprint(q.d.physics, q.d.nx, q.d.ny, q.d.nz, q.par.nt, q.par.sub,
                                                        q.par.snap_precision);
        REQUIRE(sb > 0, "invalid snapshot store");
        b += sb;
    }
    return (size_t)b;
}

void srv_fail_req(Srv* s, SReq* q, const std::string& msg) {
    // caller holds no lock; buffers are released in stream order
    if (q->mconv) (void)hipFreeAsync(q->mconv, s->c->stream);
    if (q->mrest) (void)hipFreeAsync(q->mrest, s->c->stream);
    q->mconv = q->mrest = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    s->in_use -= (long long)q->reserved;
    q->reserved = 0;
    q->err = msg;
    q->state = -1;
    q->t_done = s->now();
    s->failed++;
    s->cv_done.notify_all();
}

// analyzeModel + convert_model of the request on the side stream (the analysis needs
// the velocity extremes on the host for the time axis).
void srv_analyse(Srv* s, SReq* q) {
    hipStream_t ps = s->prep;
    q->bconv = al256((size_t)q->d.nm * q->d.N * 4) + 256;
    HIP_OK(hipMallocAsync(&q->mconv, q->bconv, ps));
    q->conv = (float*)q->mconv;
    q->red = (unsigned*)((char*)q->mconv + al256((size_t)q->d.nm * q->d.N * 4));
    const ClipBounds B = clip_bounds(&q->par);
    hipLaunchKernelGGL(key_init, dim3(1), dim3(64), 0, ps, q->red);
    const unsigned grid = (unsigned)std::min<long long>(4096, (q->d.N + kThreads - 1) / kThreads);
    const long long nxz = (long long)q->d.nx * q->d.nz;
    if (q->d.nm == 2)
        hipLaunchKernelGGL(prep_model<2>, dim3(grid), dim3(kThreads), 0, ps, q->r.d_model, q->conv, q->d.N, nxz, B, q->red);
    else if (q->d.nm == 3)
        hipLaunchKernelGGL(prep_model<3>, dim3(grid), dim3(kThreads), 0, ps, q->r.d_model, q->conv, q->d.N, nxz, B, q->red);
    else
        hipLaunchKernelGGL(prep_model<6>, dim3(grid), dim3(kThreads), 0, ps, q->r.d_model, q->conv, q->d.N, nxz, B, q->red);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(s->keys_host, q->red, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, ps));
    HIP_OK(hipStreamSynchronize(ps));
    time_axis(&q->par, unkey(s->keys_host[0]), unkey(s->keys_host[1]), &q->r.shot);
    if (q->r.kind == FWI3D_REQ_GRADIENT)
        REQUIRE(q->par.sub > 0, "a gradient request needs an imaging interval (sub != 0)");
    q->pp = prop_param(&q->par);
    q->phase = 1;
}

// Admission of an analysed request: its buffers, the source at the propagation rate
// (side stream), then its forward shot joins the batch.
void srv_start(Srv* s, SReq* q) {
    hipStream_t ps = s->prep;
    const int nt = q->par.nt;
    const fwi3d_shot* sh = &q->r.shot;
    q->brest = req_rest_bytes(*q);
    HIP_OK(hipMallocAsync(&q->mrest, q->brest, ps));
    char* p = (char*)q->mrest;
    const size_t ntr = (size_t)q->d.ncomp * sh->nrcv;
    q->srcp = (float*)p;  p += al256((size_t)q->d.nscomp * nt * 4);
    q->rcvp = (float*)p;  p += al256(std::max<size_t>(1, ntr * nt) * 4);
    q->tmp = (float*)p;   p += al256(std::max<size_t>(1, ntr * nt) * 4);
    if (q->r.kind == FWI3D_REQ_GRADIENT) {
        q->syn = (float*)p;  p += al256(std::max<size_t>(1, (size_t)q->d.nrcomp * sh->nrcv * sh->nt_data) * 4);
        q->g64 = (double*)p; p += al256((size_t)q->d.ng * q->d.N * 8);
        q->dmis = (double*)p;
        hipLaunchKernelGGL(zero_f64, dim3(1024), dim3(kThreads), 0, ps, q->g64, (long long)q->d.ng * q->d.N);
        hipLaunchKernelGGL(zero_f64, dim3(1), dim3(64), 0, ps, q->dmis, 1LL);
    }
    const Axis Td{sh->nt_data, sh->dt_data}, Tp{nt, q->par.dt};
    hipLaunchKernelGGL(resample_fwd, dim3(nblocks((long long)q->d.nscomp * nt)), dim3(kThreads), 0, ps, q->r.d_src, Td,
                       q->srcp, Tp, q->d.nscomp, q->par.resampling, q->par.sinc_half_length, 1.0f, 0);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(s->prep_ev, ps));
    HIP_OK(hipStreamWaitEvent(s->c->stream, s->prep_ev, 0));
    int id = -1;
    const int rc = fwi3d_batch_add(s->b, q->d.physics, q->conv, q->srcp, q->rcvp, &q->si.t, &q->ri.t, &q->pp, sh->nx,
                                   sh->ny, sh->nz, sh->dx, sh->dy, sh->dz, &id);
    if (rc != FWI3D_OK) throw Err{rc, fwi3d::thread_error()};
    q->shot = id;
    s->by_shot[id] = q;
    q->phase = 2;
}

// Work after the request's forward shot left the batch (main stream, in order).
void srv_after_forward(Srv* s, SReq* q) {
    hipStream_t st = s->c->stream;
    const fwi3d_shot* sh = &q->r.shot;
    const int nt = q->par.nt;
    const int ntr = q->d.nrcomp * sh->nrcv;
    const Axis Td{sh->nt_data, sh->dt_data}, Tp{nt, q->par.dt};
    const float alpha = q->par.dt / sh->dt_data;
    const float* r = q->rcvp;
    if (q->par.seismotype == 1 && ntr > 0) {
        hipLaunchKernelGGL(time_derivative, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, st, q->rcvp, q->tmp,
                           ntr, nt, q->par.dt, 0);
        r = q->tmp;
    }
    if (q->r.kind == FWI3D_REQ_MODEL) {
        if (ntr > 0)
            hipLaunchKernelGGL(resample_adj, dim3(nblocks((long long)ntr * sh->nt_data)), dim3(kThreads), 0, st,
                               q->r.d_data, Td, r, Tp, ntr, q->par.resampling, q->par.sinc_half_length, alpha, 0);
        HIP_OK(hipGetLastError());
        q->done = srv_event(s);
        HIP_OK(hipEventRecord(q->done, st));
        q->phase = 4;
        return;
    }
    // gradient: synthetic data, residual and misfit, Ht r, adjoint source (apply_jacobianT)
    const size_t ntd = (size_t)ntr * sh->nt_data;
    if (ntr > 0) {
        hipLaunchKernelGGL(resample_adj, dim3(nblocks((long long)ntd)), dim3(kThreads), 0, st, q->syn, Td, r, Tp, ntr,
                           q->par.resampling, q->par.sinc_half_length, alpha, 0);
        hipLaunchKernelGGL(residual_misfit, dim3(nblocks((long long)ntd)), dim3(kThreads), 0, st, q->syn, q->r.d_obs,
                           q->r.d_data, ntr, sh->nt_data, sh->dt_data, q->dmis);
        hipLaunchKernelGGL(resample_fwd, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, st, q->syn, Td, q->tmp,
                           Tp, ntr, q->par.resampling, q->par.sinc_half_length, 1.0f, 0);
        hipLaunchKernelGGL(adjoint_source, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, st, q->tmp, q->rcvp,
                           ntr, nt, q->par.dt, -alpha);
    }
    float* asrc = q->rcvp;
    if (q->par.seismotype == 1 && ntr > 0) {
        hipLaunchKernelGGL(time_derivative, dim3(nblocks((long long)ntr * nt)), dim3(kThreads), 0, st, q->rcvp, q->tmp,
                           ntr, nt, q->par.dt, 1);
        asrc = q->tmp;
    }
    HIP_OK(hipGetLastError());
    int id = -1;
    const int rc = fwi3d_batch_add_adjoint(s->b, q->shot, q->d.physics, q->conv, asrc, &q->ri.t, q->g64, &q->pp,
                                           sh->nx, sh->ny, sh->nz, sh->dx, sh->dy, sh->dz, &id);
    if (rc != FWI3D_OK) throw Err{rc, fwi3d::thread_error()};
    s->by_shot.erase(q->shot);
    q->shot = id;
    s->by_shot[id] = q;
    q->phase = 3;
}

// Work after the adjoint shot left: quadrature and chain rule into the caller's gradient.
void srv_after_adjoint(Srv* s, SReq* q) {
    hipStream_t st = s->c->stream;
    const fwi3d_shot* sh = &q->r.shot;
    hipLaunchKernelGGL(gradient_finish<double>, dim3(nblocks(q->d.N)), dim3(kThreads), 0, st, q->g64, q->conv,
                       q->r.d_grad64, q->d.nx, q->d.ny, q->d.nz, sh->dx, sh->dy, sh->dz, q->d.ng, 1, 1.0);
    HIP_OK(hipGetLastError());
    q->done = srv_event(s);
    HIP_OK(hipEventRecord(q->done, st));
    q->phase = 4;
}

// Requests whose last work has completed on the device.
void srv_collect(Srv* s, bool block_one) {
    for (size_t k = 0; k < s->active.size();) {
        SReq* q = s->active[k];
        if (q->phase != 4) { k++; continue; }
        hipError_t e = block_one ? hipEventSynchronize(q->done) : hipEventQuery(q->done);
        block_one = false;
        if (e == hipErrorNotReady) { k++; continue; }
        if (e != hipSuccess) throw Err{FWI3D_EDEVICE, std::string("request completion: ") + hipGetErrorString(e)};
        double mis = 0;
        if (q->r.kind == FWI3D_REQ_GRADIENT) {
            HIP_OK(hipMemcpy(&mis, q->dmis, sizeof(double), hipMemcpyDeviceToHost));
            mis = 0.5 * (double)q->r.shot.dt_data * mis;
        }
        HIP_OK(hipFreeAsync(q->mconv, s->c->stream));
        HIP_OK(hipFreeAsync(q->mrest, s->c->stream));
        q->mconv = q->mrest = nullptr;
        s->ev_pool.push_back(q->done);
        q->done = nullptr;
        s->by_shot.erase(q->shot);
        s->active.erase(s->active.begin() + k);
        std::lock_guard<std::mutex> lk(s->mu);
        q->misfit = mis;
        q->t_done = s->now();
        q->state = 2;
        s->in_use -= (long long)q->reserved;
        q->reserved = 0;
        s->completed++;
        s->cv_done.notify_all();
    }
}

// Admit queued requests in arrival order while their footprint fits the budget and a
// slot is free. A request larger than the whole budget fails.
void srv_admit(Srv* s) {
    for (;;) {
        SReq* q;
        long long in_use;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->queue.empty()) return;
            q = s->queue.front();
            in_use = s->in_use;
        }
        if ((int)s->active.size() >= s->max_active) return;
        const bool idle = s->active.empty();
        try {
            if (q->phase == 0) {
                const long long need = (long long)(al256((size_t)q->d.nm * q->d.N * 4) + 256);
                if (in_use + need > s->budget) {
                    if (idle) throw Err{FWI3D_ENOMEM, "request exceeds the server's HBM budget"};
                    return;
                }
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->in_use += need;
                    q->reserved = (size_t)need;
                }
                srv_analyse(s, q);
            }
            const long long rest = (long long)(req_rest_bytes(*q) + req_batch_bytes(*q));
            {
                std::lock_guard<std::mutex> lk(s->mu);
                in_use = s->in_use;
            }
            if (in_use + rest > s->budget) {
                if (idle) throw Err{FWI3D_ENOMEM, "request exceeds the server's HBM budget"};
                return;
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->in_use += rest;
                s->peak_in_use = std::max(s->peak_in_use, s->in_use);
                q->reserved += (size_t)rest;
                s->queue.pop_front();
                q->state = 1;
                q->t_admit = s->now();
            }
            s->active.push_back(q);
            s->peak_active = std::max(s->peak_active, (int)s->active.size());
            srv_start(s, q);
        } catch (const Err& e) {
            if (e.code == FWI3D_EDEVICE) throw;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (!s->queue.empty() && s->queue.front() == q) s->queue.pop_front();
            }
            s->active.erase(std::remove(s->active.begin(), s->active.end(), q), s->active.end());
            srv_fail_req(s, q, e.msg);
        }
    }
}

void srv_loop(Srv* s) {
    try {
        HIP_OK(hipSetDevice(s->c->device));
        std::vector<int> ids;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->active.empty(); });
                if (s->stop && s->queue.empty() && s->active.empty()) break;
            }
            srv_admit(s);
            int nshots = 0;
            (void)fwi3d_batch_status(s->b, &nshots, nullptr, nullptr);
            if (nshots > 0) {
                // one step per call, so that arriving requests join at the next step
                int nact = 0;
                if (fwi3d_batch_step(s->b, 1, &nact) != FWI3D_OK) throw Err{FWI3D_EDEVICE, fwi3d::thread_error()};
                int bstep = 0;
                (void)fwi3d_batch_status(s->b, nullptr, nullptr, &bstep);
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->steps = bstep;
                }
                int n = 0;
                fwi3d_batch_retired(s->b, nullptr, 0, &n);
                ids.resize(n);
                if (n > 0) fwi3d_batch_retired(s->b, ids.data(), n, &n);
                for (int id : ids) {
                    auto it = s->by_shot.find(id);
                    if (it == s->by_shot.end()) continue;
                    SReq* q = it->second;
                    if (q->phase == 2) srv_after_forward(s, q);
                    else if (q->phase == 3) srv_after_adjoint(s, q);
                }
                // keep the host at most `lead` steps ahead of the device
                hipEvent_t e = srv_event(s);
                HIP_OK(hipEventRecord(e, s->c->stream));
                s->inflight.push_back(e);
                while ((int)s->inflight.size() > s->lead) {
                    HIP_OK(hipEventSynchronize(s->inflight.front()));
                    s->ev_pool.push_back(s->inflight.front());
                    s->inflight.pop_front();
                }
                srv_collect(s, false);
            } else {
                srv_collect(s, true);  // only finishing work left: wait for one
            }
        }
        HIP_OK(hipStreamSynchronize(s->c->stream));
    } catch (const Err& e) {
        (void)hipStreamSynchronize(s->c->stream);
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = e.msg;
        auto fail_one = [&](SReq* q) {
            if (q->state == 2 || q->state == -1) return;
            q->state = -1;
            q->err = "server stopped: " + e.msg;
            q->t_done = s->now();
            s->failed++;
        };
        for (SReq* q : s->active) fail_one(q);
        for (SReq* q : s->queue) fail_one(q);
        s->active.clear();
        s->queue.clear();
        s->cv_done.notify_all();
    }
}

int srv_fail(Srv* s, const Err& e) {
    if (s) fwi3d::set_error(s->c, e.msg);
    return e.code;
}

}  // namespace

extern "C" {

int fwi3d_server_create(fwi3d_ctx* c, long long hbm_budget, int max_active, fwi3d_server** out) {
    if (!c || !out) return FWI3D_EINVAL;
    *out = nullptr;
    std::unique_ptr<Srv> s(new Srv());
    try {
        REQUIRE(hbm_budget > 0, "hbm_budget must be positive");
        REQUIRE(max_active >= 1, "max_active must be >= 1");
        HIP_OK(hipSetDevice(c->device));
        s->c = c;
        s->budget = hbm_budget;
        s->max_active = max_active;
        const int rc = fwi3d_batch_create(c, &s->b);
        if (rc != FWI3D_OK) throw Err{rc, fwi3d::thread_error()};
        HIP_OK(hipStreamCreateWithFlags(&s->prep, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&s->prep_ev, hipEventDisableTiming));
        HIP_OK(hipHostMalloc((void**)&s->keys_host, 64, hipHostMallocDefault));
        s->th = std::thread(srv_loop, s.get());
        c->servers.push_back(s.get());
    } catch (const Err& e) {
        if (s->b) fwi3d_batch_destroy(s->b);
        if (s->prep) (void)hipStreamDestroy(s->prep);
        if (s->prep_ev) (void)hipEventDestroy(s->prep_ev);
        if (s->keys_host) (void)hipHostFree(s->keys_host);
        fwi3d::set_error(c, e.msg);
        return e.code;
    }
    *out = s.release();
    return FWI3D_OK;
}

int fwi3d_server_submit(fwi3d_server* s, const fwi3d_request* r, int* id) {
    if (!s || !r) return FWI3D_EINVAL;
    try {
        REQUIRE(r->kind == FWI3D_REQ_MODEL || r->kind == FWI3D_REQ_GRADIENT, "unknown request kind");
        check_op(&r->par, &r->shot);
        REQUIRE(!(r->par.gl > 0) || r->par.nmodels == 2,
                "DAS receivers (gl > 0) are not served; use fwi3d_apply_forward_device / fwi3d_fwi_shot_device");
        REQUIRE(r->d_model && r->d_src, "null model or source");
        if (r->kind == FWI3D_REQ_MODEL) REQUIRE(r->d_data || r->shot.nrcv == 0, "null data buffer");
        else REQUIRE(r->d_grad64 && (r->d_obs || r->shot.nrcv == 0), "null gradient or observed data");
        std::unique_ptr<SReq> q(new SReq());
        q->r = *r;
        q->sxyz.assign(r->shot.sxyz, r->shot.sxyz + 3);
        q->rxyz.assign(r->shot.rxyz ? r->shot.rxyz : r->shot.sxyz, r->shot.rxyz ? r->shot.rxyz + 3 * r->shot.nrcv : r->shot.sxyz);
        q->r.shot.sxyz = q->sxyz.data();
        q->r.shot.rxyz = q->rxyz.data();
        q->par = normalized(&r->par);
        q->d = shot_dims(&q->par, &q->r.shot);
        q->si.build(q->par.mt ? 1 : 0, q->r.shot.sxyz, 1, &q->r.shot);
        q->ri.build(0, q->r.shot.rxyz, q->r.shot.nrcv, &q->r.shot);
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->fatal.empty()) throw Err{FWI3D_EDEVICE, "server stopped: " + s->fatal};
        REQUIRE(!s->stop, "server is shutting down");
        q->id = s->next_id++;
        q->t_submit = s->now();
        if (id) *id = q->id;
        s->queue.push_back(q.get());
        s->reqs[q->id] = std::move(q);
        s->submitted++;
        s->cv_work.notify_one();
        return FWI3D_OK;
    } catch (const Err& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        return srv_fail(s, e);
    }
}

int fwi3d_server_poll(fwi3d_server* s, int id, int* state, double* misfit, double* latency_s, double* queue_s) {
    if (!s || !state) return FWI3D_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail(s, Err{FWI3D_EINVAL, "unknown request id " + std::to_string(id)});
    const SReq* q = it->second.get();
    *state = q->state;
    if (misfit) *misfit = q->state == 2 ? q->misfit : 0.0;
    if (latency_s) *latency_s = (q->state == 2 || q->state == -1) ? q->t_done - q->t_submit : s->now() - q->t_submit;
    if (queue_s) *queue_s = q->state == 0 ? s->now() - q->t_submit : (q->t_admit > 0 ? q->t_admit - q->t_submit : 0.0);
    return FWI3D_OK;
}

int fwi3d_server_wait(fwi3d_server* s, int id) {
    if (!s) return FWI3D_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (id >= 0 && !s->reqs.count(id)) return srv_fail(s, Err{FWI3D_EINVAL, "unknown request id " + std::to_string(id)});
    s->cv_done.wait(lk, [&] {
        if (id >= 0) {
            const int st = s->reqs[id]->state;
            return st == 2 || st == -1;
        }
        return s->completed + s->failed == s->submitted;
    });
    if (id >= 0 && s->reqs[id]->state == -1) return srv_fail(s, Err{FWI3D_ESTATE, s->reqs[id]->err});
    return FWI3D_OK;
}

const char* fwi3d_server_error(fwi3d_server* s, int id) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return "unknown request id";
    return it->second->err.c_str();
}

int fwi3d_server_result(fwi3d_server* s, int id, fwi3d_op_param* par) {
    if (!s || !par) return FWI3D_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail(s, Err{FWI3D_EINVAL, "unknown request id " + std::to_string(id)});
    if (it->second->state != 2) return srv_fail(s, Err{FWI3D_ESTATE, "request " + std::to_string(id) + " is not done"});
    *par = it->second->par;
    return FWI3D_OK;
}

int fwi3d_server_stats(fwi3d_server* s, fwi3d_server_counters* st) {
    if (!s || !st) return FWI3D_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->queued = (int)s->queue.size();
    st->active = (int)(s->submitted - s->completed - s->failed) - st->queued;
    st->hbm_in_use = s->in_use;
    st->hbm_peak = s->peak_in_use;
    st->peak_active = s->peak_active;
    st->steps = s->steps;
    return FWI3D_OK;
}

int fwi3d_server_destroy(fwi3d_server* s) {
    if (!s) return FWI3D_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    auto& v = s->c->servers;
    v.erase(std::remove(v.begin(), v.end(), s), v.end());
    (void)hipSetDevice(s->c->device);
    (void)hipStreamSynchronize(s->c->stream);
    (void)hipStreamSynchronize(s->prep);
    fwi3d_batch_destroy(s->b);
    for (auto& e : s->ev_pool) (void)hipEventDestroy(e);
    for (auto& e : s->inflight) (void)hipEventDestroy(e);
    (void)hipEventDestroy(s->prep_ev);
    (void)hipStreamDestroy(s->prep);
    (void)hipHostFree(s->keys_host);
    const bool ok = s->fatal.empty();
    delete s;
    return ok ? FWI3D_OK : FWI3D_EDEVICE;
}

}  // extern "C"
