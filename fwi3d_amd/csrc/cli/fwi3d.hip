// FWI3D.x: full-waveform inversion driver, the drop-in replacement of the reference's
// executable (src/main/fwi3d.cpp:16-417) with the same command line and parfile keys
// (source=, model=, data=, output=, obj_func=, mask=, weights=,
// inverse_diagonal_hessian=, and every key of the run parameters) and hypercube files.
//
// Objective (nlls_fwi, src/include/optimization.hpp:320-425 and
// src/cpp/optimization.cpp:5-197): for every shot the synthetic data of apply_forward,
// data weights, trace normalisation, residual r = syn - obs, misfit
// 0.5 dt (sum r^2 - half the end samples) / dnorm, Ht r with the adjoint of the trace
// normalisation and the weights, apply_jacobianT into the gradient; the gradient is
// divided by dnorm and multiplied by the gradient mask. The optimiser is the engine's
// device l-BFGS with the regular Wolfe line search (lbfgs::run, nlsolver.cpp:1158-1359)
// with FP64 dot products; the propagation and imaging run on the GPU, the per-trace
// residual processing (O(traces x samples)) on the host, as in the reference.
//
// Supported: model_parameterization=1 (vp, vs, rho[, delta, epsilon, gamma]),
// nlsolver=lbfgs, lsearch=regular_wolfe, weights, normalize, normalize_obj_func, mask,
// inverse_diagonal_hessian. Other inversion options (B-splines, 1-D inversion, other
// parameterisations, soft clipping, regularisation, filters, envelopes, source-scale
// estimation, other solvers) stop the program with an error naming the option.
// Extra keys: device=<HIP device index>, snap_precision=f32|bf16|f16 (gradient snapshots).
#include <cmath>
#include <ctime>
#include <memory>

#include "driver.hpp"

using namespace fwi3d_cli;

namespace {

__global__ void grad_finish(const double* __restrict__ g64, const float* __restrict__ mask, float* __restrict__ g,
                            double inv_dnorm, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float v = (float)(g64[i] * inv_dnorm);
        if (mask) v *= mask[i];
        g[i] = v;
    }
}

struct Problem {
    RunParam par;
    Grid3 g;
    fwi3d_ctx* c = nullptr;
    int nt = 0;
    float dt = 0;
    std::vector<ShotGeo> geos;
    std::vector<long> ntr_cumul;
    std::vector<float> obs;      // processed observed data [traces][nt]
    std::vector<float> weights;  // empty: none
    std::vector<float> residual;
    std::unique_ptr<DevArray> dsrc, dsyn, dres, dgrad64, dmask;
    std::vector<float> allsrc;
    size_t src_per_shot = 0, ngrad = 0;
    double dnorm = 1.0;
    int snap_precision = FWI3D_SNAP_F32;
    long long evaluations = 0;
};

void normalize_traces(float* d, std::vector<float>& norms, int nt, long ntr) {
    norms.assign(ntr, 0.f);
    for (long i = 0; i < ntr; i++) {
        float s = 0;
        for (int t = 0; t < nt; t++) s += d[i * nt + t] * d[i * nt + t];
        if (s <= 1e-16f) s = 1;
        s = std::sqrt(s);
        norms[i] = s;
        for (int t = 0; t < nt; t++) d[i * nt + t] /= s;
    }
}

int objective(void* user, const float* d_model, float* d_grad, double* f) {
    Problem& P = *(Problem*)user;
    const RunParam& par = P.par;
    const int nt = P.nt;
    hip_check(hipMemset(P.dgrad64->p, 0, P.ngrad * sizeof(double)), "hipMemset");
    std::vector<float> syn, syn1, norms;
    for (int s = 0; s < par.ns; s++) {
        const long ntr = (long)par.rxyz[s].size() * par.nrcomp;
        fwi3d_op_param q = op_param(par);
        q.snap_precision = P.snap_precision;
        P.dsrc->upload(P.allsrc.data() + s * P.src_per_shot, P.src_per_shot * sizeof(float));
        abi_check(fwi3d_apply_forward_device(P.c, &q, &P.geos[s].shot, 0, d_model, P.dsrc->as<float>(),
                                             P.dsyn->as<float>()),
                  P.c, "apply_forward");
        syn.resize((size_t)ntr * nt);
        P.dsyn->download(syn.data(), syn.size() * sizeof(float));
        const float* w = P.weights.empty() ? nullptr : P.weights.data() + P.ntr_cumul[s] * nt;
        if (w)
            for (size_t i = 0; i < syn.size(); i++) syn[i] *= w[i];
        if (par.normalize) {
            syn1 = syn;
            normalize_traces(syn.data(), norms, nt, ntr);
        }
        float* r = P.residual.data() + P.ntr_cumul[s] * nt;
        const float* d = P.obs.data() + P.ntr_cumul[s] * nt;
        for (size_t i = 0; i < syn.size(); i++) r[i] = syn[i] - d[i];
        // adjoint source: Ht r, the adjoint of the normalisation, the weights
        std::vector<float> a(r, r + syn.size());
        for (long i = 0; i < ntr; i++) {
            float* x = a.data() + i * nt;
            x[0] *= 0.5f * P.dt;
            for (int t = 1; t < nt - 1; t++) x[t] *= P.dt;
            x[nt - 1] *= 0.5f * P.dt;
        }
        if (par.normalize)
            for (long i = 0; i < ntr; i++) {
                float* x = a.data() + i * nt;
                const float* u = syn1.data() + i * nt;
                float v = 0;
                for (int t = 0; t < nt; t++) v += x[t] * u[t];
                const float nn = norms[i];
                for (int t = 0; t < nt; t++) x[t] = x[t] / nn - v / (nn * nn * nn) * u[t];
            }
        if (w)
            for (size_t i = 0; i < a.size(); i++) a[i] *= w[i];
        P.dres->upload(a.data(), a.size() * sizeof(float));
        abi_check(fwi3d_apply_jacobianT_device(P.c, &q, &P.geos[s].shot, 1, P.dgrad64->as<double>(), d_model,
                                               P.dres->as<float>()),
                  P.c, "apply_jacobianT");
    }
    double sum = 0;
    const long ntr_total = (long)(P.residual.size() / nt);
    for (long i = 0; i < ntr_total; i++) {
        const float* x = P.residual.data() + i * nt;
        double s2 = 0;
        for (int t = 0; t < nt; t++) s2 += (double)x[t] * x[t];
        sum += s2 - 0.5 * ((double)x[0] * x[0] + (double)x[nt - 1] * x[nt - 1]);
    }
    *f = 0.5 * P.dt * sum / P.dnorm;
    hipLaunchKernelGGL(grad_finish, dim3(1024), dim3(256), 0, 0, P.dgrad64->as<double>(),
                       P.dmask ? P.dmask->as<float>() : nullptr, d_grad, 1.0 / P.dnorm, (long long)P.ngrad);
    hip_check(hipGetLastError(), "grad_finish");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    P.evaluations++;
    if (par.verbose > 1) std::fprintf(stderr, "evaluation %lld: f = %.9g\n", P.evaluations, *f);
    return 0;
}

void unsupported(bool used, const char* what) {
    if (used) die(std::string(what) + " is not supported by this build of FWI3D.x");
}

}  // namespace

int main(int argc, char** argv) {
    const auto args = args_of(argc, argv);
    Problem P;
    RunParam& par = P.par;
    read_parameters(args, par);
    if (par.verbose > 0) {
        const time_t t = time(nullptr);
        std::fprintf(stderr, "\n====================\n%s\n====================\n", ctime(&t));
    }
    std::string source_file = "none", model_file = "none", data_file = "none", output_file = "none",
                obj_func_file = "none", bsmodel_file = "none", ioutput_file = "none", snap = "f32";
    int device = 0;
    get(args, "source", source_file);
    get(args, "model", model_file);
    get(args, "bsmodel", bsmodel_file);
    get(args, "data", data_file);
    get(args, "output", output_file);
    get(args, "ioutput", ioutput_file);
    get(args, "obj_func", obj_func_file);
    get(args, "device", device);
    get(args, "snap_precision", snap);
    check(source_file != "none", "Source wavelet is not provided");
    check(model_file != "none", "Earth model is not provided");
    check(data_file != "none", "Data to be inverted is not provided");
    Cube src, data, model, mask, w, invh;
    try {
        src = read_cube(source_file);
        data = read_cube(data_file);
        model = read_cube(model_file, 4);
        if (par.mask_file != "none") mask = read_cube(par.mask_file);
        if (par.weights_file != "none") w = read_cube(par.weights_file);
        if (par.inverse_diagonal_hessian_file != "none") invh = read_cube(par.inverse_diagonal_hessian_file);
    } catch (const std::exception& e) {
        die(e.what());
    }
    check(src.axes[0].n == data.axes[0].n && src.axes[0].d == data.axes[0].d,
          "source wavelet(s) and data must have the same number of time samples and sampling rate");
    if (!mask.axes.empty()) check(mask.count() == model.count(), "Gradient mask must have the same number of samples as the model");
    if (!w.axes.empty()) check(w.count() == data.count(), "Data weights must have the same number of samples as the data");
    if (!invh.axes.empty())
        check(invh.count() == model.count(), "Inverse diagonal Hessian must have the same number of samples as the model");
    P.g = analyze_geometry(model, par);
    const Cube allsrc = analyze_wavelet(src, par);
    // analyzeNLInversion (src/cpp/param.cpp:255-289)
    par.isave = std::max(1, par.isave);
    if (par.sub == 0) par.sub = -1;
    unsupported(par.bsplines, "bsplines=1");
    unsupported(bsmodel_file != "none", "bsmodel=");
    unsupported(par.inversion1d, "inversion1d=1");
    unsupported(par.model_parameterization != 1 && par.nmodels >= 3, "model_parameterization other than 1");
    unsupported(par.soft_clip, "soft_clip=1");
    unsupported(par.regularization > -1, "regularization");
    unsupported(par.prior_file != "none", "prior=");
    unsupported(par.filter_file != "none", "filter=");
    unsupported(par.envelop != 0, "envelop");
    unsupported(par.scale_source_times > 0, "scale_source_times");
    unsupported(par.nlsolver != "lbfgs", ("nlsolver=" + par.nlsolver).c_str());
    unsupported(par.lsearch != "regular_wolfe", ("lsearch=" + par.lsearch).c_str());
    if (ioutput_file != "none") std::fprintf(stderr, "WARNING: ioutput= (intermediate models) is not written by this build\n");

    P.nt = allsrc.axes[0].n;
    P.dt = allsrc.axes[0].d;
    long ntr_total = 0;
    P.ntr_cumul.assign(par.ns, 0);
    for (int s = 0; s < par.ns; s++) {
        P.ntr_cumul[s] = ntr_total;
        ntr_total += (long)par.rxyz[s].size() * par.nrcomp;
    }
    check((long)data.count() == ntr_total * P.nt,
          "the data must hold " + std::to_string(ntr_total) + " traces of " + std::to_string(P.nt) + " samples");
    P.obs = data.v;
    P.residual.assign(P.obs.size(), 0.f);
    if (!w.axes.empty()) {
        P.weights = w.v;
        for (float x : P.weights) check(x >= 0, "Data weights  must be non-negative");
        for (size_t i = 0; i < P.obs.size(); i++) P.obs[i] *= P.weights[i];
    }
    if (par.normalize) {
        std::vector<float> norms;
        normalize_traces(P.obs.data(), norms, P.nt, ntr_total);
    }
    if (par.normalize_obj_func) {
        double s = 0;
        for (float x : P.obs) s += (double)x * x;
        P.dnorm = s * P.dt;
        if (P.dnorm < 1e-16) P.dnorm = 1;
    }
    P.allsrc = allsrc.v;
    P.src_per_shot = (size_t)par.nscomp * P.nt;
    P.snap_precision = snap == "bf16" ? FWI3D_SNAP_BF16 : (snap == "f16" ? FWI3D_SNAP_F16 : FWI3D_SNAP_F32);
    for (int s = 0; s < par.ns; s++) P.geos.push_back(shot_geometry(par, P.g, s, P.nt, P.dt));

    hip_check(hipSetDevice(device), "hipSetDevice");
    abi_check(fwi3d_create(&P.c, device), nullptr, "fwi3d_create");
    const long long nm = (long long)model.count();
    P.ngrad = (size_t)nm;  // gradient of every model parameter (the anisotropy part stays 0)
    long max_ntr = 1;
    for (int s = 0; s < par.ns; s++) max_ntr = std::max(max_ntr, (long)par.rxyz[s].size() * par.nrcomp);
    P.dsrc.reset(new DevArray(P.src_per_shot * sizeof(float)));
    P.dsyn.reset(new DevArray((size_t)max_ntr * P.nt * sizeof(float)));
    P.dres.reset(new DevArray((size_t)max_ntr * P.nt * sizeof(float)));
    P.dgrad64.reset(new DevArray(P.ngrad * sizeof(double)));
    if (!mask.axes.empty()) {
        for (float x : mask.v) check(x >= 0, "Gradient mask must be non-negative");
        P.dmask.reset(new DevArray(mask.v.size() * sizeof(float)));
        P.dmask->upload(mask.v.data(), mask.v.size() * sizeof(float));
    }
    DevArray dmodel(nm * sizeof(float)), dgrad(nm * sizeof(float));
    dmodel.upload(model.v.data(), nm * sizeof(float));
    std::unique_ptr<DevArray> dh0;
    if (!invh.axes.empty()) {
        dh0.reset(new DevArray(nm * sizeof(float)));
        dh0->upload(invh.v.data(), nm * sizeof(float));
    }
    fwi3d_lbfgs_param lp;
    fwi3d_lbfgs_param_defaults(&lp);
    lp.niter = par.niter;
    lp.m = par.lbfgs_m;
    lp.max_trial = par.max_trial;
    lp.threshold = par.threshold;
    lp.c1 = par.ls_c1; lp.c2 = par.ls_c2; lp.a0 = par.ls_a0; lp.a1 = par.ls_a1; lp.step_max = par.ls_max_step;
    lp.version = par.ls_version ? 1 : 0;
    std::vector<double> hist(par.niter + 1, 0.0);
    int counters[3] = {0, 0, 0};
    abi_check(fwi3d_lbfgs_run(P.c, &lp, nm, dmodel.as<float>(), dgrad.as<float>(), dh0 ? dh0->as<float>() : nullptr,
                              objective, &P, hist.data(), counters),
              P.c, "l-BFGS");
    if (par.verbose > 0)
        std::fprintf(stderr, "l-BFGS: %d iterations, %d function and %d gradient evaluations; f %.9g -> %.9g\n",
                     counters[0], counters[1], counters[2], hist[0], hist[counters[0]]);
    try {
        if (output_file != "none") {
            dmodel.download(model.v.data(), nm * sizeof(float));
            write_cube(model, output_file, par.format, par.datapath);
        }
        if (obj_func_file != "none") {
            Cube fc;
            fc.axes = {Axis{counters[0] + 1, 0.f, 1.f, ""}};
            fc.alloc();
            for (int i = 0; i <= counters[0]; i++) fc.v[i] = (float)hist[i];
            write_cube(fc, obj_func_file, par.format, par.datapath);
        }
    } catch (const std::exception& e) {
        die(e.what());
    }
    fwi3d_destroy(P.c);
    if (par.verbose > 0) {
        const time_t t = time(nullptr);
        std::fprintf(stderr, "\n====================\n%s\n====================\n", ctime(&t));
    }
    return 0;
}
