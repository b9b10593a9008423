// WE_MODELING_3D.x: 3-D seismic modelling of every shot of a survey, the drop-in
// replacement of the reference's executable (src/main/we_modeling.cpp:15-119) with the
// same command line and parfile keys (source=, model=, output=, wavefield=, parfile=,
// and every key of the run parameters), reading and writing the same hypercube files.
//
// The reference runs one operator per shot to completion (MPI ranks take shots
// round-robin). Here every shot is a modelling request to the engine's server on one
// GPU: the shots join one batch as HBM allows and advance together, and each shot's
// traces equal what apply_forward computes for it alone. With wavefield= and a single
// shot (the reference drops the snapshots when ns > 1) the shot runs through
// apply_forward on its own and its snapshot store is written as the wavefield file.
//
// Extra keys (not in the reference): device=<HIP device index> (default 0),
// snap_precision=f32|bf16|f16 (wavefield store; default f32).
#include <cstring>
#include <ctime>
#include <memory>

#include "driver.hpp"

using namespace fwi3d_cli;

int main(int argc, char** argv) {
    const auto args = args_of(argc, argv);
    RunParam par;
    read_parameters(args, par);
    par.compression_rate = 0;
    const int verbose = par.verbose;
    if (verbose > 0) {
        const time_t t = time(nullptr);
        std::fprintf(stderr, "\n====================\n%s\n====================\n", ctime(&t));
    }
    std::string source_file = "none", model_file = "none", wavefield_file = "none", output_file = "none", snap = "f32";
    int device = 0;
    get(args, "source", source_file);
    get(args, "model", model_file);
    get(args, "wavefield", wavefield_file);
    get(args, "output", output_file);
    get(args, "device", device);
    get(args, "snap_precision", snap);
    check(source_file != "none", "Source wavelet is not provided");
    check(model_file != "none", "Earth model is not provided");

    Cube src, model;
    try {
        src = read_cube(source_file);
        model = read_cube(model_file, 4);
    } catch (const std::exception& e) {
        die(e.what());
    }
    const Grid3 g = analyze_geometry(model, par);
    const Cube allsrc = analyze_wavelet(src, par);
    if (par.ns > 1) par.sub = 0;
    check(par.gl <= 0 || par.nmodels >= 3, "DAS receivers need an elastic model");

    const Axis T = allsrc.axes[0];
    const int nt = T.n;
    long ntr_total = 0;
    std::vector<long> ntr_cumul(par.ns, 0);
    for (int s = 0; s < par.ns; s++) {
        ntr_cumul[s] = ntr_total;
        ntr_total += (long)par.rxyz[s].size() * par.nrcomp;
    }
    Cube out;
    out.axes = {T, Axis{(int)ntr_total, 0.f, 1.f, ""}};
    out.alloc();

    hip_check(hipSetDevice(device), "hipSetDevice");
    fwi3d_ctx* c = nullptr;
    abi_check(fwi3d_create(&c, device), nullptr, "fwi3d_create");
    DevArray dmodel(model.v.size() * sizeof(float));
    dmodel.upload(model.v.data(), model.v.size() * sizeof(float));
    fwi3d_op_param base = op_param(par);
    base.snap_precision = snap == "bf16" ? FWI3D_SNAP_BF16 : (snap == "f16" ? FWI3D_SNAP_F16 : FWI3D_SNAP_F32);
    const size_t src_per_shot = (size_t)par.nscomp * nt;

    const bool want_wavefield = wavefield_file != "none" && par.sub > 0 && par.ns == 1;
    fwi3d_op_param q = base;  // after apply_forward: the time axis of the last shot
    if (want_wavefield || par.gl > 0) {
        // one shot after the other through apply_forward (DAS channels are not served);
        // with wavefield= (one shot) its snapshot store is written too
        DevArray ds(src_per_shot * sizeof(float)), dd(std::max<long>(1, ntr_total) * nt * sizeof(float));
        for (int s = 0; s < par.ns; s++) {
            if (verbose > 1) std::fprintf(stderr, "\nStart modeling shot %d\n", s);
            ShotGeo geo = shot_geometry(par, g, s, nt, T.d);
            q = base;
            if (!want_wavefield) q.sub = 0;
            ds.upload(allsrc.v.data() + s * src_per_shot, src_per_shot * sizeof(float));
            abi_check(fwi3d_apply_forward_device(c, &q, &geo.shot, 0, dmodel.as<float>(), ds.as<float>(),
                                                 dd.as<float>()),
                      c, "apply_forward");
            const size_t ntr = (size_t)par.rxyz[s].size() * par.nrcomp;
            dd.download(out.v.data() + ntr_cumul[s] * nt, ntr * nt * sizeof(float));
        }
    }
    if (want_wavefield) {
        int slots = 0, ncomp = 0;
        abi_check(fwi3d_snapshots_read(c, nullptr, 0, &slots, &ncomp), c, "snapshots");
        Cube wf;
        wf.axes = {model.axes[0], model.axes[1], model.axes[2], Axis{ncomp, 0.f, 1.f, ""},
                   Axis{slots, 0.f, q.dt * q.sub, ""}};
        wf.alloc();
        abi_check(fwi3d_snapshots_read(c, wf.v.data(), (long long)wf.v.size(), &slots, &ncomp), c, "snapshots");
        try {
            write_cube(wf, wavefield_file, par.format, par.datapath);
        } catch (const std::exception& e) {
            die(e.what());
        }
    } else if (!(par.gl > 0)) {
        // every shot a modelling request of the server (continuous batching)
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        fwi3d_server* srv = nullptr;
        abi_check(fwi3d_server_create(c, (long long)(0.85 * (double)free_b), std::max(1, par.ns), &srv), c,
                  "fwi3d_server_create");
        std::vector<std::unique_ptr<DevArray>> bufs;
        std::vector<ShotGeo> geos(par.ns);
        std::vector<int> ids(par.ns);
        for (int s = 0; s < par.ns; s++) {
            if (verbose > 1) std::fprintf(stderr, "\nStart modeling shot %d\n", s);
            geos[s] = shot_geometry(par, g, s, nt, T.d);
            const size_t ntr = (size_t)par.rxyz[s].size() * par.nrcomp;
            bufs.emplace_back(new DevArray(src_per_shot * sizeof(float)));
            bufs.back()->upload(allsrc.v.data() + s * src_per_shot, src_per_shot * sizeof(float));
            const float* dsrc = bufs.back()->as<float>();
            bufs.emplace_back(new DevArray(std::max<size_t>(1, ntr) * nt * sizeof(float)));
            fwi3d_request r{};
            r.kind = FWI3D_REQ_MODEL;
            r.par = base;
            r.par.sub = 0;
            r.shot = geos[s].shot;
            r.d_model = dmodel.as<float>();
            r.d_src = dsrc;
            r.d_data = bufs.back()->as<float>();
            if (fwi3d_server_submit(srv, &r, &ids[s]) != 0) die(std::string("shot submission failed: ") + fwi3d_last_error(c));
        }
        const int rc = fwi3d_server_wait(srv, -1);
        for (int s = 0; s < par.ns; s++) {
            int st = 0;
            (void)fwi3d_server_poll(srv, ids[s], &st, nullptr, nullptr, nullptr);
            if (st != 2) die("shot " + std::to_string(s) + " failed: " + fwi3d_server_error(srv, ids[s]));
            const size_t ntr = (size_t)par.rxyz[s].size() * par.nrcomp;
            bufs[2 * s + 1]->download(out.v.data() + ntr_cumul[s] * nt, ntr * nt * sizeof(float));
            if (verbose > 1) std::fprintf(stderr, "\nFinish modeling shot %d\n", s);
        }
        if (rc != 0) die("modelling failed");
        abi_check(fwi3d_server_destroy(srv), c, "fwi3d_server_destroy");
    }
    if (output_file != "none") {
        try {
            write_cube(out, output_file, par.format, par.datapath);
        } catch (const std::exception& e) {
            die(e.what());
        }
    }
    fwi3d_destroy(c);
    if (verbose > 0) {
        const time_t t = time(nullptr);
        std::fprintf(stderr, "\n====================\n%s\n====================\n", ctime(&t));
    }
    return 0;
}
