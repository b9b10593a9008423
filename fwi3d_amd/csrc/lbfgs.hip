// l-BFGS with the regular-Wolfe line search on device-resident vectors, FP64 dot
// products and FP64 curvature pairs. The model stays FP32 (the reference's data_t);
// the objective (misfit and gradient, e.g. fwi3d_fwi_shot_device over the shots) is
// a caller callback on the caller's own device buffers.
//
//   lbfgs::run / updateSY / computeHg ..... src/cpp/nlsolver.cpp:1158-1359
//   regular_wolfe::lineSearch / zoom ...... src/cpp/nlsolver.cpp:185-423
//   quadratic / cubic ..................... src/cpp/nlsolver.cpp:10-36
//   defaults .............................. src/include/param.hpp:55-62
//
// Dot products reduce in a fixed order (a fixed grid of per-block partial sums, then
// one block in index order), so a run is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fwi3d_hip.h"
#include "context.hpp"

using namespace fwi3d;

namespace {

constexpr int kThreads = 256;
constexpr int kParts = 1024;  // partial sums of a dot product
constexpr double kZero = 1e-10;
constexpr double kStol = 1e-3;

template <class A, class B>
__global__ __launch_bounds__(kThreads) void dot_parts(const A* __restrict__ a, const B* __restrict__ b, long long n,
                                                      double* __restrict__ parts) {
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)kParts * kThreads)
        s += (double)a[i] * (double)b[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double w[kThreads / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) parts[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

__global__ __launch_bounds__(kThreads) void dot_final(const double* __restrict__ parts, double* __restrict__ out) {
    // thread t sums parts t, t+256, t+512, t+768; then a fixed tree over the block
    __shared__ double w[kThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < kParts; i += kThreads) s += parts[i];
    w[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) w[threadIdx.x] += w[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = w[0];
}

// out = a - b in FP64 (the s and y pairs from FP32 models / gradients)
__global__ void diff64(double* __restrict__ out, const float* __restrict__ a, const float* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (double)a[i] - (double)b[i];
}
// y += c * x
__global__ void axpy64(double* __restrict__ y, const double* __restrict__ x, double c, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] += c * x[i];
}
// q = g (FP64 copy of the gradient)
__global__ void widen(double* __restrict__ q, const float* __restrict__ g, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        q[i] = (double)g[i];
}
// p = c * h0 * q  (h0 null: ones)
__global__ void hscale(double* __restrict__ p, const double* __restrict__ q, const float* __restrict__ h0, double c,
                       long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = c * (h0 ? (double)h0[i] : 1.0) * q[i];
}
// m = m0 + alpha * p
__global__ void trial_model(float* __restrict__ m, const float* __restrict__ m0, const double* __restrict__ p, double alpha,
                            long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m[i] = (float)((double)m0[i] + alpha * p[i]);
}

double quadratic(double f, double df, double stp0, double f0) {
    double ans;
    if (f0 <= f + df * stp0) ans = stp0 * 1.1;
    else ans = -df * stp0 * stp0 / (2 * (f0 - f - df * stp0));
    if ((ans / stp0 < kStol) || (std::abs(ans - stp0) / stp0) < kStol) ans = stp0 / 2.0;
    return ans;
}

double cubic(double f, double df, double stp0, double f0, double stp1, double f1) {
    const double x1 = f1 - f - df * stp1;
    const double x2 = f0 - f - df * stp0;
    const double x3 = stp0 * stp0 * stp1 * stp1 * (stp1 - stp0);
    const double a = (stp0 * stp0 * x1 - stp1 * stp1 * x2) / x3;
    const double b = (-stp0 * stp0 * stp0 * x1 + stp1 * stp1 * stp1 * x2) / x3;
    const double delta = b * b - 3 * a * df;
    double step, func, ans;
    if (f1 < f0) { step = stp1; func = f1; }
    else { step = stp0; func = f0; }
    if (std::abs(a) < kZero || delta < 0) ans = quadratic(f, df, step, func);
    else ans = (-b + std::sqrt(delta)) / (3 * a);
    if ((ans / step < kStol) || (std::abs(step - ans) / step) < kStol) ans = step / 2.0;
    else if (std::abs(stp0 + stp1 - step - ans) / (stp0 + stp1 - step) < kStol) ans = (stp0 + stp1 - step) / 2.0;
    return ans;
}

// Device vectors and reductions of one run.
struct Vec {
    fwi3d_ctx* c;
    long long n;
    double* parts;
    double* res;
    unsigned grid() const { return (unsigned)std::min<long long>(4096, (n + kThreads - 1) / kThreads); }
    template <class A, class B>
    double dot(const A* a, const B* b) {
        hipLaunchKernelGGL((dot_parts<A, B>), dim3(kParts), dim3(kThreads), 0, c->stream, a, b, n, parts);
        hipLaunchKernelGGL(dot_final, dim3(1), dim3(kThreads), 0, c->stream, parts, res);
        double v = 0;
        HIP_OK(hipMemcpyAsync(&v, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return v;
    }
};

struct Objective {
    fwi3d_objective fn;
    void* user;
    float* m;
    float* g;
    fwi3d_ctx* c;
    double f = 0;
    int feval = 0, geval = 0;
    void res() {  // misfit and gradient at the current model (nlls_fwi::res)
        HIP_OK(hipStreamSynchronize(c->stream));
        double v = 0;
        const int rc = fn(user, m, g, &v);
        REQUIRE(rc == 0, "the objective callback failed (status " + std::to_string(rc) + ")");
        REQUIRE(v == v, "the objective function is NaN");
        f = v;
    }
};

struct Wolfe {
    double c1, c2, a0, a1, stp_max;
    int version;
    double stp0 = 0, stp1 = 0, stp = 0, f = 0, df = 0, f0 = 0, f1 = 0;

    // m = m0 + alpha p, evaluate
    void trial(Objective& P, Vec& V, const float* m0, const double* p, double alpha) {
        hipLaunchKernelGGL(trial_model, dim3(V.grid()), dim3(kThreads), 0, V.c->stream, P.m, m0, p, alpha, V.n);
        P.res();
        P.feval++;
    }

    bool zoom(bool reverse, Objective& P, Vec& V, const float* m0, const double* p, int& trial_n, int max_trial,
              double& gnorm) {
        double alo = stp0, ahi = stp1, flo = f0, fhi = f1;
        if (reverse) { alo = stp1; ahi = stp0; flo = f1; fhi = f0; }
        while (true) {
            if (trial_n >= max_trial) return false;
            double alpha;
            if (alo == 0) alpha = quadratic(f, df, ahi, fhi);
            else if (ahi == 0) alpha = quadratic(f, df, alo, flo);
            else alpha = cubic(f, df, alo, flo, ahi, fhi);
            trial(P, V, m0, p, alpha);
            trial_n++;
            const double fv = P.f;
            f0 = f1;
            f1 = fv;
            if (fv <= kZero) { stp = alpha; f0 = f; f = fv; return false; }
            if ((f1 > f + c1 * alpha * df) || (f1 >= flo)) {
                ahi = alpha;
                fhi = f1;
            } else {
                P.geval++;
                gnorm = std::sqrt(V.dot(P.g, P.g));
                const double df1 = V.dot(P.g, p);
                if (df1 >= c2 * df) { stp = alpha; f0 = f; f = fv; return true; }
                if (df1 * (ahi - alo) >= 0) { ahi = alo; fhi = flo; }
                alo = alpha;
                flo = f1;
            }
        }
    }

    bool search(Objective& P, Vec& V, const float* m0, const double* p, int iter, int max_trial, double& gnorm) {
        const double df0 = df;
        f = P.f;
        df = V.dot(P.g, p);
        int trial_n = 0;
        stp0 = 0;
        if (iter == 0) {
            f0 = f;
            stp1 = a0 / gnorm;
            if (a1 > 0) stp1 = std::max(stp1, a1 * std::sqrt(V.dot(P.m, P.m)) / gnorm);
        } else {
            stp1 = stp * df0 / df;
            if (version == 1) stp1 = std::min(1.0, 1.00001 * (2 * (f - f0) / df));
        }
        double alpha = stp1;
        while (true) {
            if (trial_n >= max_trial) return false;
            trial(P, V, m0, p, alpha);
            trial_n++;
            const double fv = P.f;
            f1 = fv;
            if (fv <= kZero) { stp = alpha; f0 = f; f = fv; return false; }
            if ((f1 > f + c1 * alpha * df) || (f1 >= f0 && trial_n > 1)) return zoom(false, P, V, m0, p, trial_n, max_trial, gnorm);
            P.geval++;
            gnorm = std::sqrt(V.dot(P.g, P.g));
            const double df1 = V.dot(P.g, p);
            if (df1 >= c2 * df) { stp = alpha; f0 = f; f = fv; return true; }
            if (df1 >= 0) return zoom(true, P, V, m0, p, trial_n, max_trial, gnorm);
            stp0 = stp1;
            stp1 = std::min(quadratic(f, df, stp1, f1), stp_max);
            alpha = stp1;
            f0 = f1;
        }
    }
};

int fail(fwi3d_ctx* c, const Err& e) {
    fwi3d::set_error(c, e.msg);
    return e.code;
}

}  // namespace

extern "C" {

void fwi3d_lbfgs_param_defaults(fwi3d_lbfgs_param* p) {
    if (!p) return;
    p->niter = 0;
    p->m = 5;
    p->max_trial = 10;
    p->threshold = 0.0;
    p->c1 = 1e-4;
    p->c2 = 0.9;
    p->a0 = 1.0;
    p->a1 = 0.0;
    p->step_max = 1e6;
    p->version = 0;
}

int fwi3d_lbfgs_run(fwi3d_ctx* c, const fwi3d_lbfgs_param* lp, long long n, float* d_model, float* d_grad,
                    const float* d_h0, fwi3d_objective fn, void* user, double* f_history, int* counters) {
    if (!c) return FWI3D_EINVAL;
    try {
        REQUIRE(lp && d_model && d_grad && fn && n > 0, "null argument or empty model");
        REQUIRE(lp->m >= 1 && lp->niter >= 0 && lp->max_trial >= 1, "bad l-BFGS parameters");
        REQUIRE(lp->c1 > 0 && lp->c1 < 1 && lp->c2 > lp->c1 && lp->c2 < 1, "The constants c1, c2 must satisfy 0<c1<c2<1");
        HIP_OK(hipSetDevice(c->device));
        const int M = lp->m;
        // FP64: s[M], y[M], p, q; FP32: m0, g0; reduction scratch
        const size_t nd = (size_t)n * 8, nf = (size_t)n * 4;
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        c->opt.ensure((2 * M + 2) * al(nd) + 2 * al(nf) + al(kParts * 8) + 256);
        char* base = c->opt.as<char>();
        std::vector<double*> s(M), y(M);
        for (int i = 0; i < M; i++) { s[i] = (double*)base; base += al(nd); }
        for (int i = 0; i < M; i++) { y[i] = (double*)base; base += al(nd); }
        double* p = (double*)base; base += al(nd);
        double* q = (double*)base; base += al(nd);
        float* m0 = (float*)base; base += al(nf);
        float* g0 = (float*)base; base += al(nf);
        Vec V{c, n, (double*)base, nullptr};
        base += al(kParts * 8);
        V.res = (double*)base;
        const unsigned G = V.grid();

        Objective P{fn, user, d_model, d_grad, c};
        Wolfe W{lp->c1, lp->c2, lp->a0, lp->a1, lp->step_max, lp->version};
        std::vector<double> func, rho(M, 0.0);
        HIP_OK(hipMemcpyAsync(m0, d_model, nf, hipMemcpyDeviceToDevice, c->stream));
        P.res();
        P.feval++;
        double f = P.f;
        func.push_back(f);
        int k = 0;
        if (!(func[0] <= kZero || func[0] != func[0])) {
            double rate = 1, gnorm = 999;
            bool success = true;
            while (k < lp->niter && rate > lp->threshold && gnorm > kZero && f > kZero && success) {
                if (k == 0) {  // the regular-Wolfe search leaves the gradient at the accepted step
                    P.geval++;
                    gnorm = std::sqrt(V.dot(d_grad, d_grad));
                }
                // updateSY: newest pair last, the oldest recycled beyond M pairs
                if (k > 0) {
                    int slot = k - 1;
                    if (k > M) {
                        std::rotate(s.begin(), s.begin() + 1, s.end());
                        std::rotate(y.begin(), y.begin() + 1, y.end());
                        std::rotate(rho.begin(), rho.begin() + 1, rho.end());
                        slot = M - 1;
                    }
                    hipLaunchKernelGGL(diff64, dim3(G), dim3(kThreads), 0, c->stream, s[slot], d_model, m0, n);
                    hipLaunchKernelGGL(diff64, dim3(G), dim3(kThreads), 0, c->stream, y[slot], d_grad, g0, n);
                    rho[slot] = V.dot(s[slot], y[slot]);
                }
                // computeHg: two-loop recursion, then p = -H g
                const int mm = std::min(k, M);
                std::vector<double> a(mm, 0.0);
                double gamma = 1.0;
                hipLaunchKernelGGL(widen, dim3(G), dim3(kThreads), 0, c->stream, q, d_grad, n);
                for (int i = mm - 1; i >= 0; i--) {
                    a[i] = V.dot(s[i], q) / rho[i];
                    hipLaunchKernelGGL(axpy64, dim3(G), dim3(kThreads), 0, c->stream, q, y[i], -a[i], n);
                }
                if (mm > 0) gamma = rho[mm - 1] / V.dot(y[mm - 1], y[mm - 1]);
                hipLaunchKernelGGL(hscale, dim3(G), dim3(kThreads), 0, c->stream, p, q, d_h0, -gamma, n);
                for (int i = 0; i < mm; i++) {
                    // p holds -H g so far: b = y.(H g) / rho = -y.p / rho, and p -= (a - b) s
                    const double b = -V.dot(y[i], p) / rho[i];
                    hipLaunchKernelGGL(axpy64, dim3(G), dim3(kThreads), 0, c->stream, p, s[i], -(a[i] - b), n);
                }
                if (V.dot(p, d_grad) >= 0) break;  // not a descent direction
                if (k > 0) HIP_OK(hipMemcpyAsync(m0, d_model, nf, hipMemcpyDeviceToDevice, c->stream));
                HIP_OK(hipMemcpyAsync(g0, d_grad, nf, hipMemcpyDeviceToDevice, c->stream));
                success = W.search(P, V, m0, p, k, lp->max_trial, gnorm);
                if (!success) break;
                f = P.f;
                func.push_back(f);
                if (k > 0) rate = (func[k] - func[k + 1]) / func[k];
                k++;
            }
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        if (f_history)
            for (size_t i = 0; i < func.size() && i <= (size_t)lp->niter; i++) f_history[i] = func[i];
        if (counters) { counters[0] = k; counters[1] = P.feval; counters[2] = P.geval; }
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

}  // extern "C"
