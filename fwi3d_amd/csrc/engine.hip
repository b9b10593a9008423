// Engine: context, device-resident state and the time loop of propagate(), behind
// the C ABI of include/fwi3d_hip.h. The shot operator around propagate() is in
// operator.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fwi3d_hip.h"
#include "acoustic.hpp"
#include "context.hpp"
#include "common.hpp"
#include "elastic.hpp"
#include "batch_kernels.hpp"
#include "sbp_tables.hpp"
#include "snapshots.hpp"

using namespace fwi3d;

namespace {

bool g_tables_loaded[64] = {false};

// Workspace of the allocation-free entry point (fwi3d_propagate_device_noalloc): a
// file-scope device array, so the library allocates nothing at call time. Each device
// that loads the library holds its own copy until the process ends. In host code the
// array's name is only a stand-in: its device address, per device, comes from
// hipGetSymbolAddress at context creation (scratch_setup) and is the only handle used.
constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB
static __device__ __align__(256) char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {nullptr};

void scratch_setup(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= 64 || g_scratch_ptr[dev]) return;
    void* ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));  // the current device's copy
    g_scratch_ptr[dev] = ptr;
}
void* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (dev >= 0 && dev < 64) ? g_scratch_ptr[dev] : nullptr;
}

void load_tables(int dev) {
    if (dev >= 0 && dev < 64 && g_tables_loaded[dev]) return;
    float s[4], h[4];
    for (int k = 0; k < 4; k++) { s[k] = (float)kSatS[k]; h[k] = (float)kHnorm[k]; }
    HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sat_s), s, sizeof(s)));
    HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_hnorm), h, sizeof(h)));
    if (dev >= 0 && dev < 64) g_tables_loaded[dev] = true;
}

}  // namespace

namespace {

Geom make_geom(int nx, int ny, int nz, float dx, float dy, float dz) {
    Geom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.pz = (nz + 63) / 64 * 64;
    const long long n = (long long)ny * nx * g.pz;
    REQUIRE(9 * n < (1LL << 31), "grid too large for 32-bit field offsets (9 * padded cells must stay below 2^31)");
    g.sx = g.pz;
    g.sy = nx * g.pz;
    g.n = (int)n;
    g.dx = dx; g.dy = dy; g.dz = dz;
    return g;
}

// Squared cos taper factor per index along one axis, in the float arithmetic of the
// reference (val = (float)cos(double), then val*val): spatial_operators.cpp:408-506.
void taper_table(std::vector<float>& t, int n, int lo, int hi, float strength) {
    t.assign(n, 1.0f);
    const double a = (double)strength;
    for (int i = 0; i < lo && i < n; i++) {
        float v = (float)std::cos(a * 0.5 * M_PI * (double)(i + 1 - lo) / (double)(0 - lo));
        t[i] = v * v;
    }
    for (int i = std::max(0, n - hi); i < n; i++) {
        float v = (float)std::cos(a * 0.5 * M_PI * (double)(i - (n - hi)) / (double)hi);
        t[i] = v * v;
    }
}

// Copy a model/field set [k][ny][nx][nz] (unpadded, device) into the padded layout.
void to_padded_on(hipStream_t stream, float* dst, const float* src, int k, const Geom& g) {
    for (int q = 0; q < k; q++)
        HIP_OK(hipMemcpy2DAsync(dst + q * g.n, g.pz * sizeof(float), src + (size_t)q * g.nx * g.ny * g.nz,
                                g.nz * sizeof(float), g.nz * sizeof(float), (size_t)g.nx * g.ny, hipMemcpyDeviceToDevice,
                                stream));
}
void to_padded(fwi3d_ctx* c, float* dst, const float* src, int k, const Geom& g) { to_padded_on(c->stream, dst, src, k, g); }

// Indices per point and axis, and stencil contributions per point, of an injector kind
// (delta_m3 / ddelta_m3: 3 and 27; dipole_m3: 6 and 54).
inline int inj_npi(int kind) { return kind == 2 ? 6 : 3; }
inline int inj_nper(int kind) { return kind == 2 ? 54 : 27; }

// Device bytes of one injector's tables (ints, floats, then the per-target lists).
size_t inj_bytes(const fwi3d_injector* h) {
    const size_t nw = h->kind ? 6 : 3;
    return 3 * (size_t)inj_npi(h->kind) * h->npts * sizeof(int) + 3 * (2 * nw * h->npts) * sizeof(float) +
           (2 * (size_t)inj_nper(h->kind) * h->npts + 1) * sizeof(int) + 256;
}

// Per-target contribution lists of the injection (InjTab::tbeg / tlst) into hv:
// [tbeg (ntgt+1)][tlst (nper npts)]; returns ntgt. Targets in order of first appearance,
// each target's contributions in increasing t (t = 27 r + 9 a + 3 b + c, or for a dipole
// t = 54 r + 2 (9 a + 3 b + c) + e, the reference's loop order: InjTab).
int inj_targets(const fwi3d_injector* h, std::vector<int>& hv) {
    const int nper = inj_nper(h->kind), npi = inj_npi(h->kind);
    const int nc = nper * h->npts;
    std::vector<std::pair<long long, int>> key(nc);
    for (int t = 0; t < nc; t++) {
        const int r = t / nper;
        int q = t - nper * r, e = 0;
        if (h->kind == 2) { e = q & 1; q >>= 1; }
        const int base = npi * r + 3 * e;
        const long long iy = h->yind[base + q / 9], ix = h->xind[base + (q / 3) % 3], iz = h->zind[base + q % 3];
        key[t] = {(iy << 42) | (ix << 21) | iz, t};
    }
    std::vector<int> first(nc);
    {   // order targets by their first contribution
        std::vector<std::pair<long long, int>> srt(key);
        std::stable_sort(srt.begin(), srt.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (int k = 0; k < nc;) {
            int e = k;
            while (e < nc && srt[e].first == srt[k].first) e++;
            for (int q = k; q < e; q++) first[srt[q].second] = srt[k].second;
            k = e;
        }
    }
    std::vector<int> order(nc);
    for (int t = 0; t < nc; t++) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return first[x] < first[y]; });
    hv.assign((size_t)nc + 1, 0);
    std::vector<int> beg;
    beg.reserve(nc + 1);
    for (int k = 0; k < nc; k++) {
        if (k == 0 || first[order[k]] != first[order[k - 1]]) beg.push_back(k);
        hv[k] = order[k];
    }
    beg.push_back(nc);
    std::vector<int> out(beg);
    out.insert(out.end(), hv.begin(), hv.begin() + nc);
    hv.swap(out);
    return (int)beg.size() - 1;
}

// Upload injector tables to device memory at `base` (inj_bytes(h) bytes). hv keeps the
// host copy of the per-target lists alive until the stream has consumed it.
InjTab upload_inj_to(hipStream_t stream, char* base, const fwi3d_injector* h, std::vector<int>& hv) {
    InjTab t{};
    t.npts = h->npts;
    t.kind = h->kind;
    t.nw = h->kind ? 6 : 3;
    const size_t ni = (size_t)inj_npi(h->kind) * h->npts, nw = (size_t)2 * t.nw * h->npts;
    int* xi = (int*)base;
    int* yi = xi + ni;
    int* zi = yi + ni;
    float* xw = (float*)(zi + ni);
    float* yw = xw + nw;
    float* zw = yw + nw;
    int* tl = (int*)(zw + nw);
    t.ntgt = 0;
    if (h->npts > 0) {
        HIP_OK(hipMemcpyAsync(xi, h->xind, ni * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(yi, h->yind, ni * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(zi, h->zind, ni * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(xw, h->xw, nw * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(yw, h->yw, nw * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(zw, h->zw, nw * sizeof(float), hipMemcpyHostToDevice, stream));
        t.ntgt = inj_targets(h, hv);
        HIP_OK(hipMemcpyAsync(tl, hv.data(), hv.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    }
    t.xi = xi; t.yi = yi; t.zi = zi; t.xw = xw; t.yw = yw; t.zw = zw;
    t.tbeg = tl;
    t.tlst = tl + t.ntgt + 1;
    return t;
}

// Upload injector tables into the context buffer `buf`.
InjTab upload_inj(fwi3d_ctx* c, DevBuf& buf, const fwi3d_injector* h, std::vector<int>& hv) {
    buf.ensure(inj_bytes(h));
    return upload_inj_to(c->stream, buf.as<char>(), h, hv);
}

void check_injector(const fwi3d_injector* h, int nx, int ny, int nz, const char* what) {
    REQUIRE(h != nullptr, std::string(what) + " injector is null");
    REQUIRE(h->npts >= 0 && (h->kind == 0 || h->kind == 1 || h->kind == 2), std::string(what) + ": bad npts/kind");
    for (int i = 0; i < inj_npi(h->kind) * h->npts; i++) {
        REQUIRE(h->xind[i] >= 0 && h->xind[i] < nx && h->yind[i] >= 0 && h->yind[i] < ny && h->zind[i] >= 0 &&
                    h->zind[i] < nz,
                std::string(what) + " point outside the grid");
    }
}

// Event-bracketed launch of a step kernel when profiling. Events come from a pool
// that is created once and reused, so the timed region records but never creates.
template <class F>
void timed_on(fwi3d_ctx* c, int kind, hipStream_t st, F&& launch) {
    if (!c->prof || (c->prof_steps_only && kind != 4)) {
        launch();
        return;
    }
    if (c->ev_used + 2 > c->ev.size()) {
        for (int k = 0; k < 2; k++) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            c->ev.push_back(e);
        }
    }
    // reserve the pair first: the launch may time nested work (a whole step around its
    // kernels), which takes the pairs after this one
    const size_t i0 = c->ev_used;
    c->ev_used += 2;
    c->ev_kind.push_back(kind);
    c->ev_count.push_back(c->timed_count);
    HIP_OK(hipEventRecord(c->ev[i0], st));
    launch();
    HIP_OK(hipEventRecord(c->ev[i0 + 1], st));
}
template <class F>
void timed(fwi3d_ctx* c, int kind, F&& launch) {
    timed_on(c, kind, c->stream, launch);
}

void drain_events(fwi3d_ctx* c) {
    if (c->ev_used == 0) return;
    HIP_OK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        const int k = c->ev_kind[i / 2];
        c->prof_ms[k] += ms;
        c->prof_n[k] += c->ev_count[i / 2];
    }
    c->ev_used = 0;
    c->ev_kind.clear();
    c->ev_count.clear();
}

// Time steps [it0, it1) of a loop whose step(StepIdx) enqueues kernels only (no host
// work between steps): a graph of kGraphSteps steps, whose kernels read the step from a
// device counter that the graph's last node advances, is captured once and replayed;
// the remainder runs eagerly. The same kernels run either way, so results do not depend
// on the mode. The graphs live until release_graphs (after the call's last sync).
constexpr int kGraphSteps = 16;
template <class StepFn>
void run_steps(fwi3d_ctx* c, int it0, int it1, StepFn&& step) {
    const int K = kGraphSteps;
    const int W = (it1 - it0) / K;
    if (!c->use_graphs || W < 2) {
        for (int it = it0; it < it1; it++) step(StepIdx(it));
        return;
    }
    c->stepctr.ensure(256);
    int* ctr = c->stepctr.as<int>();
    const bool prof = c->prof;
    c->prof = false;  // no event records inside the capture
    hipGraph_t graph = nullptr;
    try {
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int j = 0; j < K; j++) step(StepIdx(ctr, j));
        hipLaunchKernelGGL(advance_step, dim3(1), dim3(64), 0, c->stream, ctr, K);
        HIP_OK(hipStreamEndCapture(c->stream, &graph));
    } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(c->stream, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        c->prof = prof;
        throw;
    }
    c->prof = prof;
    c->graphs.push_back(graph);
    hipGraphExec_t exec = nullptr;
    HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    c->graph_execs.push_back(exec);
    hipLaunchKernelGGL(set_step, dim3(1), dim3(64), 0, c->stream, ctr, it0);
    c->timed_count = K;
    for (int w = 0; w < W; w++) timed(c, 4, [&] { HIP_OK(hipGraphLaunch(exec, c->stream)); });
    c->timed_count = 1;
    for (int it = it0 + W * K; it < it1; it++) step(StepIdx(it));
}

// Free the graphs of the current call (its stream work has completed).
void release_graphs(fwi3d_ctx* c) {
    for (auto e : c->graph_execs) (void)hipGraphExecDestroy(e);
    for (auto g : c->graphs) (void)hipGraphDestroy(g);
    c->graph_execs.clear();
    c->graphs.clear();
}

// Upload a one-shot table for the time-step kernels of the single-shot path.
const BShot* upload_shot1(fwi3d_ctx* c, const BShot& d) {
    c->shot1_host.resize(sizeof(BShot));
    std::memcpy(c->shot1_host.data(), &d, sizeof(BShot));
    c->shot1_dev.ensure(sizeof(BShot));
    HIP_OK(hipMemcpyAsync(c->shot1_dev.p, c->shot1_host.data(), sizeof(BShot), hipMemcpyHostToDevice, c->stream));
    return c->shot1_dev.as<BShot>();
}

// --------------------------------------------------------------------------------------
struct Tapers {
    const float *tz2, *tx2, *ty2;
};

// Taper tables (nz + nx + ny floats) uploaded to `dtz`.
Tapers tapers_to(hipStream_t stream, float* dtz, const fwi3d_param* par, const Geom& g) {
    std::vector<float> tz, tx, ty;
    taper_table(tz, g.nz, par->taper[0], par->taper[1], par->taper_strength);
    taper_table(tx, g.nx, par->taper[2], par->taper[3], par->taper_strength);
    taper_table(ty, g.ny, par->taper[4], par->taper[5], par->taper_strength);
    float* dtx = dtz + tz.size();
    float* dty = dtx + tx.size();
    HIP_OK(hipMemcpyAsync(dtz, tz.data(), tz.size() * 4, hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dtx, tx.data(), tx.size() * 4, hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dty, ty.data(), ty.size() * 4, hipMemcpyHostToDevice, stream));
    return Tapers{dtz, dtx, dty};
}
Tapers setup_tapers(fwi3d_ctx* c, const fwi3d_param* par, const Geom& g) {
    c->tapers.ensure((size_t)(g.nz + g.nx + g.ny) * sizeof(float));
    return tapers_to(c->stream, c->tapers.as<float>(), par, g);
}

// Tiling of the y-marching kernels: 64(z) x TX(x) column tiles, y in chunks of
// `ychunk` planes (context tuning; default 8: many short chunks balance better than
// fewer long ones, measured on MI355X at 200^3).
template <class Args>
int tile_grid_yc(Args& A, const Geom& g, int TX, bool interior, int ychunk) {
    // interior: the points 6+ cells inside every face, [6, n-6) on each axis; otherwise
    // the whole grid (the flux pass)
    const int org = interior ? 6 : 0;
    const int ez = g.nz - 2 * org, ex = g.nx - 2 * org, ey = g.ny - 2 * org;
    A.org = org;
    A.yend = g.ny - org;
    A.ntz = (ez + 63) / 64;
    A.ntx = (ex + TX - 1) / TX;
    A.ychunk = std::max(4, std::min(std::max(ey, 1), ychunk > 0 ? ychunk : 8));
    A.nyc = (ey + A.ychunk - 1) / A.ychunk;
    return A.ntz * A.ntx * A.nyc;  // 0 when the interior is empty (an axis of 12 points)
}
// Slab kernel of the elastic / VTI shell. z slabs: x rows per workgroup and y planes per
// chunk; side (x and y) slabs: march planes per chunk.
constexpr int kZSlabTX = 32;
constexpr int kZSlabYChunk = 3;
constexpr int kSideChunk = 3;
// Its tiling: the z slabs inside the x and y slabs, then the x slabs and the y slabs
// inside the other two axes' slabs (0 workgroups where an axis has no points 6 cells
// inside its faces).
int zslab_tiles(ZSlabTiles& Z, const Geom& g) {
    Z.ychunk = kZSlabYChunk;
    Z.ntx = (g.nx - 12 + kZSlabTX - 1) / kZSlabTX;
    Z.nyc = (g.ny - 12 + Z.ychunk - 1) / Z.ychunk;
    Z.nzs = 2 * Z.ntx * Z.nyc;
    Z.schunk = kSideChunk;
    Z.sntz = (g.nz - 12 + 31) / 32;
    Z.snc0 = (g.ny - 12 + kSideChunk - 1) / kSideChunk;
    Z.snc1 = (g.nx - 12 + kSideChunk - 1) / kSideChunk;
    return Z.nzs + 2 * Z.sntz * (Z.snc0 + Z.snc1);
}
// Shell points left to the point-wise shell kernel once the slabs are tiled: the edges
// and corners (edge_point's enumeration).
int shell_edge_count(const Geom& g) { return edge_count(6, g.nx, g.ny, g.nz); }

// Elastic / VTI step with the fused kernel (fused.hpp): its tiling over the planes
// [6, ny-6); the flux pass (d.ef) over the 8 planes at each y face, or the whole grid
// when ny < 16; the slab launch over the y slabs: side tiles, then the y-slab points
// within 6 cells of an x or z face (the first part of edge_point's enumeration).
struct FusedPlan {
    int nfused, nfedge, nflux, nslab;  // fused middle / edge workgroups, flux pass, slab launch
};
// Side-tile y chunk giving about `target` side workgroups (at least 8 planes: each chunk
// spends 4 iterations filling its queues).
int fused_side_chunk(int ntz, int ntx, int nyo, int target) {
    const int sides = std::max(1, fused_side_tiles(ntz, ntx));
    const int want = std::max(1, (target + sides / 2) / sides);
    return std::max(8, (nyo + want - 1) / want);
}

FusedPlan fused_plan(BShot& d, const Geom& g, int TX, int ychunk, int fchunk, int echunk) {
    FusedPlan P{};
    d.fused = 1;
    FusedTiles& F = d.ft;
    F.ntz = fused_tile_count(g.nz, FZ);
    F.ntx = fused_tile_count(g.nx, FW);
    F.ya = 6;
    F.yb = g.ny - 6;
    const int nyo = std::max(0, F.yb - F.ya);
    const int tiles = F.ntz * F.ntx;
    int yc = fchunk;
    if (yc <= 0) {
        // at most 512 workgroups of middle tiles (two per CU, one round), chunks of at
        // least 8 planes (each chunk spends 4 iterations filling its queues)
        const int mid = fused_middle_tiles(F.ntz, F.ntx);
        const int want = std::max(1, 512 / std::max(1, mid > 0 ? mid : tiles));
        yc = std::max(8, (nyo + want - 1) / want);
    }
    F.ychunk = std::max(1, std::min(yc, std::max(nyo, 1)));
    F.nyc = nyo > 0 ? (nyo + F.ychunk - 1) / F.ychunk : 0;
    // Edge tiles run one 512-thread workgroup per CU. Side tiles in y chunks of `echunk`
    // planes (0: about 384 side workgroups, 1.5 per CU of the 256); corner tiles (the
    // costliest per plane) in chunks of half that length, at least 4 planes. Measured
    // (middle + edge launches): 200^3 230 against 265 us with the middle tiles' chunk
    // (21 planes) for the side tiles and a quarter of it for the corners; 256^3 363
    // against 468 us. Grids where other chunks win are left to fwi3d_autotune.
    const int sc = echunk > 0 ? echunk : fused_side_chunk(F.ntz, F.ntx, nyo, 384);
    F.sychunk = std::max(1, std::min(sc, std::max(nyo, 1)));
    F.snyc = nyo > 0 ? (nyo + F.sychunk - 1) / F.sychunk : 0;
    F.cychunk = std::max(1, std::min(std::max(4, F.sychunk / 2), std::max(nyo, 1)));
    F.cnyc = nyo > 0 ? (nyo + F.cychunk - 1) / F.cychunk : 0;
    P.nfused = fused_middle_tiles(F.ntz, F.ntx) * F.nyc;
    P.nfedge = fused_corner_tiles(F.ntz, F.ntx) * F.cnyc + fused_side_tiles(F.ntz, F.ntx) * F.snyc;
    EArgs& A = d.ef;
    if (g.ny >= 16) {
        A.org = 0;
        A.yend = g.ny;
        A.ntz = (g.nz + 63) / 64;
        A.ntx = (g.nx + TX - 1) / TX;
        A.ychunk = 8;
        A.nyc = 2;
        A.ysplit = 1;
        P.nflux = A.ntz * A.ntx * A.nyc;
    } else {
        A.ysplit = 0;
        P.nflux = tile_grid_yc(A, g, TX, false, ychunk);
    }
    zslab_tiles(d.zs, g);
    d.nshell_pts = 2 * 6 * (2 * 6 * g.nz + (g.nx - 2 * 6) * 2 * 6);
    d.nshell_blk = (d.nshell_pts + 255) / 256;
    P.nslab = 2 * d.zs.sntz * d.zs.snc1 + 3 * d.nshell_blk;
    return P;
}

// Tile shape for a physics and grid: the autotuned one if fwi3d_autotune measured this
// grid on the context, else the context's settings (fwi3d_tune). `echunk`: y chunk of
// the fused kernel's side tiles, `fchunk` of its middle tiles.
void tiles_for(const fwi3d_ctx* c, int physics, const Geom& g, int& TX, int& ychunk, int& echunk, int& fchunk) {
    TX = c->tune_tx;
    ychunk = c->tune_ychunk;
    echunk = c->fused_echunk;
    fchunk = c->fused_chunk;
    for (const auto& t : c->tuned)
        if (t.physics == physics && t.nx == g.nx && t.ny == g.ny && t.nz == g.nz) {
            TX = t.tx;
            ychunk = t.ychunk;
            if (t.echunk > 0) echunk = t.echunk;
            if (t.fchunk > 0) fchunk = t.fchunk;
        }
}

// ---- snapshot store: slot k holds forward step k*sub; unwritten slots read as zero.
// A store is `slots` slots of ncomp x n values (FP32, or 16-bit codes with one scale
// word per slot). The context owns the store of the single-shot path; a batch owns one
// per gradient shot.
struct SnapDesc {
    char* data = nullptr;
    unsigned* scale = nullptr;
    int slots = 0, ncomp = 0, fmt = FWI3D_SNAP_F32;
    long long n = 0;
    size_t slot_elems() const { return (size_t)ncomp * n; }
    size_t esz() const { return fmt == FWI3D_SNAP_F32 ? 4 : 2; }
};

size_t snap_store_bytes(int slots, int ncomp, long long n, int fmt) {
    return (size_t)slots * ncomp * n * (fmt == FWI3D_SNAP_F32 ? 4 : 2);
}

SnapDesc ctx_snap(const fwi3d_ctx* c) {
    SnapDesc s;
    s.data = c->snaps.as<char>();
    s.scale = c->snapsc.as<unsigned>();
    s.slots = c->snap_slots; s.ncomp = c->snap_ncomp; s.fmt = c->snap_fmt; s.n = c->snap_n;
    return s;
}

void snap_begin(fwi3d_ctx* c, int physics, int ncomp, const fwi3d_param* par, const Geom& g) {
    const int slots = 1 + par->nt / par->sub;
    const size_t bytes = snap_store_bytes(slots, ncomp, g.n, par->snap_precision);
    c->snaps.ensure(bytes);
    c->snapsc.ensure((size_t)slots * sizeof(unsigned));
    HIP_OK(hipMemsetAsync(c->snaps.p, 0, bytes, c->stream));
    HIP_OK(hipMemsetAsync(c->snapsc.p, 0, (size_t)slots * sizeof(unsigned), c->stream));
    c->snap_slots = slots; c->snap_ncomp = ncomp; c->snap_n = g.n; c->snap_nt = par->nt; c->snap_sub = par->sub;
    c->snap_physics = physics; c->snap_fmt = par->snap_precision;
    c->snap_dims[0] = g.nx; c->snap_dims[1] = g.ny; c->snap_dims[2] = g.nz;
    c->snap_valid = false;
}

void snap_check(const fwi3d_ctx* c, int physics, const fwi3d_param* par, const Geom& g) {
    REQUIRE(c->snap_valid && c->snap_physics == physics && c->snap_nt == par->nt && c->snap_sub == par->sub &&
                c->snap_dims[0] == g.nx && c->snap_dims[1] == g.ny && c->snap_dims[2] == g.nz,
            "adjoint propagation needs the snapshots of a matching forward propagation on this context");
    REQUIRE(c->snap_fmt == par->snap_precision, "adjoint snap_precision differs from the forward propagation's");
}

// Store field u (ncomp x n values) in slot k.
void snap_put(hipStream_t st, const SnapDesc& S, int k, const float* u) {
    const size_t ne = S.slot_elems();
    unsigned* mb = S.scale + k;
    if (S.fmt == FWI3D_SNAP_F32) {
        HIP_OK(hipMemcpyAsync((float*)S.data + (size_t)k * ne, u, ne * sizeof(float), hipMemcpyDeviceToDevice, st));
        return;
    }
    void* dst = S.data + (size_t)k * ne * 2;
    const unsigned grid = (unsigned)std::min<size_t>(2048, (ne / 4 + 255) / 256 + 1);
    hipLaunchKernelGGL(snap_maxabs, dim3(grid), dim3(256), 0, st, u, mb, (long long)ne);
    if (S.fmt == FWI3D_SNAP_BF16)
        hipLaunchKernelGGL(snap_store<SNAP_BF16>, dim3(2048), dim3(256), 0, st, u, dst, (const unsigned*)mb, (long long)ne);
    else
        hipLaunchKernelGGL(snap_store<SNAP_F16>, dim3(2048), dim3(256), 0, st, u, dst, (const unsigned*)mb, (long long)ne);
}

template <int FMT>
SnapView<FMT> snap_view(const SnapDesc& S, int k) {
    if (k < 0 || k >= S.slots) return SnapView<FMT>{nullptr, nullptr};
    return SnapView<FMT>{S.data + (size_t)k * S.slot_elems() * S.esz(), S.scale + k};
}

// Call f(std::integral_constant<int, FMT>) for a snapshot format.
template <class Fn>
void with_snap_fmt(int fmt, Fn&& f) {
    if (fmt == FWI3D_SNAP_BF16) f(std::integral_constant<int, SNAP_BF16>{});
    else if (fmt == FWI3D_SNAP_F16) f(std::integral_constant<int, SNAP_F16>{});
    else f(std::integral_constant<int, SNAP_F32>{});
}

// One imaging-condition event of compute_gradients (we_op.cpp:1555-1605) at snapshot
// index itl: the adjoint field adjf against forward slots itl-1, itl, itl+1.
void grad_event_acoustic(hipStream_t st, const SnapDesc& S, int itl, int nt, int sub, float dt, const float* adjf,
                         const float* M, const Geom& g, double* gK, double* gB) {
    const float sdt = (float)sub * dt;
    const int last = nt / sub;
    const int branch = (itl > 0 && itl < last) ? 1 : (itl == 0 ? 0 : 2);
    dim3 grid((g.nz + 63) / 64, (g.nx + 3) / 4, g.ny);  // 64 z x 4 x rows per workgroup
    with_snap_fmt(S.fmt, [&](auto tag) {
        constexpr int FMT = decltype(tag)::value;
        // a slot past the store reads as zero (the reference dereferences a null slot there)
        const SnapView<FMT> fn = snap_view<FMT>(S, itl), fm = itl > 0 ? snap_view<FMT>(S, itl - 1) : fn,
                            fp = snap_view<FMT>(S, itl + 1);
        hipLaunchKernelGGL(acoustic_grad<FMT>, grid, dim3(64, 4), 0, st, fm, fn, fp, adjf, M, g, sdt, branch, gK, gB);
    });
}

// Elastic / VTI event (we_op.cpp:488-565, 1108-1216); lagf is the previous adjoint field
// (the derivatives of the lambda/mu terms lag one step, as in the reference).
void grad_event_elastic(hipStream_t st, bool vti, const SnapDesc& S, int itl, int nt, int sub, float dt,
                        const float* adjf, const float* lagf, const EArgs& A, double* grad, long long nunp) {
    const float sdt = (float)sub * dt;
    const int last = nt / sub;
    const int branch = (itl > 0 && itl < last) ? 1 : (itl == 0 ? 0 : 2);
    const Geom& g = A.g;
    dim3 grid((g.nz + 63) / 64, (g.nx + GTX - 1) / GTX, g.ny);  // 64 z x GTX x rows per workgroup
    with_snap_fmt(S.fmt, [&](auto tag) {
        constexpr int FMT = decltype(tag)::value;
        const SnapView<FMT> fn = snap_view<FMT>(S, itl), fm = itl > 0 ? snap_view<FMT>(S, itl - 1) : fn,
                            fp = snap_view<FMT>(S, itl + 1);
        if (vti)
            hipLaunchKernelGGL((elastic_grad<true, FMT>), grid, dim3(64, GTX), 0, st, fm, fn, fp, adjf, lagf, A, sdt, branch,
                               grad, grad + nunp, grad + 2 * nunp);
        else
            hipLaunchKernelGGL((elastic_grad<false, FMT>), grid, dim3(64, GTX), 0, st, fm, fn, fp, adjf, lagf, A, sdt, branch,
                               grad, grad + nunp, grad + 2 * nunp);
    });
}

void propagate_acoustic(fwi3d_ctx* c, bool adj, const float* d_model, const float* d_src, float* d_rcv,
                        const fwi3d_injector* inj, const fwi3d_injector* ext, double* d_grad, const fwi3d_param* par,
                        const Geom& g) {
    const long long n = g.n;
    c->fields.ensure(3 * n * sizeof(float));
    c->model.ensure(2 * n * sizeof(float));
    float* F = c->fields.as<float>();
    float* M = c->model.as<float>();
    HIP_OK(hipMemsetAsync(F, 0, 3 * n * sizeof(float), c->stream));
    HIP_OK(hipMemsetAsync(M, 0, 2 * n * sizeof(float), c->stream));
    to_padded(c, M, d_model, 2, g);
    const Tapers tp = setup_tapers(c, par, g);

    InjTab I = upload_inj(c, c->injtab, inj, c->injhost[0]);
    InjTab E{};
    if (!adj) E = upload_inj(c, c->exttab, ext, c->injhost[1]);

    const int nt = par->nt, sub = par->sub;
    const float dt = par->dt;
    const bool fwd_snap = !adj && sub > 0;
    if (fwd_snap) snap_begin(c, FWI3D_ACOUSTIC, 1, par, g);
    if (adj && d_grad && sub > 0) snap_check(c, FWI3D_ACOUSTIC, par, g);
    float* prev = F;
    float* curr = F + n;
    float* next = F + 2 * n;

    int TX, ychunk, echunk, fchunk;
    tiles_for(c, FWI3D_ACOUSTIC, g, TX, ychunk, echunk, fchunk);
    AcuArgs A{};
    A.K = M; A.b = M + n; A.tz2 = tp.tz2; A.tx2 = tp.tx2; A.ty2 = tp.ty2; A.g = g;
    for (int f = 0; f < 6; f++) A.bc[f] = par->bc[f];
    A.dt = dt; A.dt2 = dt * dt;
    A.alpha = 0.2508560249f / par->free_surface_stiffness;
    const int nblk = tile_grid_yc(A, g, TX, true, ychunk);
    const int nshell = shell_count(6, g.nx, g.ny, g.nz);  // points within 6 cells of a face
    // one-shot table for the time-step kernels (rotation from the step index)
    BShot T1h{};
    const int nthr = 256;
    const int nbi = (I.ntgt + nthr - 1) / nthr;
    T1h.phys = FWI3D_ACOUSTIC; T1h.t0 = 0; T1h.nt = par->nt; T1h.nf = n; T1h.F = F; T1h.g = g; T1h.a = A;
    T1h.nshell_pts = nshell;
    T1h.nshell_blk = (nshell + 255) / 256;
    T1h.I = I; T1h.src = d_src; T1h.ninj_blk = nbi;
    T1h.E = E; T1h.rcv = d_rcv; T1h.next_blk = (E.npts + 127) / 128;
    const BShot* T1 = upload_shot1(c, T1h);

    if (I.ntgt > 0)
        hipLaunchKernelGGL(acoustic_inject, dim3(nbi), dim3(nthr), 0, c->stream, prev, A, I, d_src, nt, 0, 1);

    const long long nunp = (long long)g.nx * g.ny * g.nz;
    double* gK = d_grad;
    double* gB = d_grad ? d_grad + nunp : nullptr;
    auto grad_event = [&](int itl, const float* adjf) {
        grad_event_acoustic(c->stream, ctx_snap(c), itl, nt, sub, dt, adjf, (const float*)M, g, gK, gB);
    };
    // fields of local step k: prev = F[k % 3], curr = F[(k + 1) % 3] (the kernels' rotation)
    auto field = [&](int k) { return F + (long long)(k % 3) * n; };
    // one time step: receivers read curr, interior and shell write next, sources add to it
    // receivers are extracted by workgroups of the shell launch
    const int ext_blk = (!adj && E.npts > 0) ? extract_blocks256(E.npts, 1) : 0;
    // with 4-row tiles the interior, shell and receivers share one launch, except when
    // every kernel is timed on its own (fwi3d_profile level 1)
    const bool one_launch = TX == 4 && nblk > 0 && !(c->prof && !c->prof_steps_only);
    auto step = [&](StepIdx si) {
        timed(c, 4, [&] {
            if (one_launch) {
                hipLaunchKernelGGL(acoustic_step_ext, dim3(nblk + T1h.nshell_blk + ext_blk), dim3(64, 4), 0, c->stream,
                                   T1, si, nblk, T1h.nshell_blk);
                return;
            }
            if (nblk > 0) timed(c, 0, [&] {
                if (TX == 8) hipLaunchKernelGGL(acoustic_interior_batch<8>, dim3(nblk), dim3(64, 8), 0, c->stream, T1, BMap{}, si);
                else hipLaunchKernelGGL(acoustic_interior_batch<4>, dim3(nblk), dim3(64, 4), 0, c->stream, T1, BMap{}, si);
            });
            timed(c, 2, [&] {
                hipLaunchKernelGGL(acoustic_boundary_ext, dim3(T1h.nshell_blk + ext_blk), dim3(256), 0, c->stream, T1, si,
                                   T1h.nshell_blk);
            });
        });
        if (I.ntgt > 0)
            hipLaunchKernelGGL(acoustic_inject_batch, dim3(nbi), dim3(nthr), 0, c->stream, T1, BMap{}, si);
    };

    if (!fwd_snap && !(adj && d_grad && sub > 0)) {
        run_steps(c, 0, nt - 1, step);  // no host work between steps
    } else {
        for (int it = 0; it < nt - 1; it++) {
            if (fwd_snap && it % sub == 0) snap_put(c->stream, ctx_snap(c), it / sub, field(it + 1));
            if (adj && d_grad && sub > 0 && (nt - 1 - it) % sub == 0 && it != 0) grad_event((nt - 1 - it) / sub, field(it + 1));
            step(StepIdx(it));
        }
    }
    (void)prev; (void)curr; (void)next;
    if (fwd_snap && (nt - 1) % sub == 0) snap_put(c->stream, ctx_snap(c), (nt - 1) / sub, field(nt));
    if (!adj && E.npts > 0)
        hipLaunchKernelGGL(extract_batch, dim3(T1h.next_blk), dim3(128), 0, c->stream, T1, BMap{}, StepIdx(nt - 1));
    if (adj && d_grad && sub > 0) grad_event(0, field(nt));
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    release_graphs(c);
    if (fwd_snap) c->snap_valid = true;
}

// Decode one snapshot slot (padded layout) to FP32.
template <int FMT>
__global__ void snap_decode(SnapView<FMT> v, float* __restrict__ out, long long n) {
    const float sc = v.scale();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = v.code((int)i) * sc;
}

__global__ void vti_c11b_field(float* __restrict__ c11b, const float* __restrict__ eps, const float* __restrict__ mu,
                               long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c11b[i] = 2.f * (1.f + 2.f * eps[i]) * mu[i];
}

// One elastic/VTI time step: flux pre-pass, interior y-march, boundary shell.
// The one-shot table T carries the flux pass's tiling (whole grid, nflux workgroups)
// and the interior's (nint workgroups, 0 when the interior is empty).
// next_blk: receiver-extraction workgroups launched with the shell (0: none this step).
// One elastic/VTI time step with the fused kernel: fused planes, flux pass at the y
// faces, y slabs with the receivers of the step.
template <int TX, bool VTI>
void step_elastic_fused(fwi3d_ctx* c, const BShot* T, StepIdx it, const FusedPlan& P, int next_blk) {
    auto middle = [&](hipStream_t st) {
        if (P.nfused > 0)
            hipLaunchKernelGGL((elastic_fused_batch<VTI, false>), dim3(P.nfused), dim3(FTHREADS), 0, st, T, BMap{}, it);
    };
    auto edge = [&](hipStream_t st) {
        if (P.nfedge > 0)
            hipLaunchKernelGGL((elastic_fused_batch<VTI, true>), dim3(P.nfedge), dim3(FTHREADS), 0, st, T, BMap{}, it);
    };
    // the y faces: flux pass of their 8 planes, then their 6 slab planes and the receivers
    auto yfaces = [&](hipStream_t st) {
        hipLaunchKernelGGL((elastic_flux_batch<TX, VTI>), dim3(P.nflux), dim3(64, TX), 0, st, T, BMap{}, it);
        hipLaunchKernelGGL(elastic_yslabs_ext<VTI>, dim3(P.nslab + next_blk), dim3(256), 0, st, T, it, P.nslab);
    };
    const bool per_kernel = c->prof && !c->prof_steps_only;
    if (c->fused_overlap == 0 || per_kernel) {
        // one stream (and per-kernel timing: classes 0 fused tiles, 1 flux pass, 3 slabs)
        if (P.nfused + P.nfedge > 0)
            timed(c, 0, [&] {
                middle(c->stream);
                edge(c->stream);
            });
        timed(c, 1, [&] {
            hipLaunchKernelGGL((elastic_flux_batch<TX, VTI>), dim3(P.nflux), dim3(64, TX), 0, c->stream, T, BMap{}, it);
        });
        timed(c, 3, [&] {
            hipLaunchKernelGGL(elastic_yslabs_ext<VTI>, dim3(P.nslab + next_blk), dim3(256), 0, c->stream, T, it,
                               P.nslab);
        });
        return;
    }
    // The three launches write disjoint points of next and read none of each other's
    // outputs. Mode 3 (the default): the edge tiles on the second stream from the step's
    // start beside the middle tiles, then the y faces after the middle tiles (the edge
    // tiles, one workgroup per CU and latency-bound, take the CUs the bandwidth-bound
    // middle tiles leave; 257 against 271 us per step at 200^3, 443 against 464 at 256^3).
    // Modes 1 and 2: the y-face launches on the second stream, forked from the step's
    // start (1) or after the middle tiles (2). The step joins both streams at its end.
    if (c->fused_overlap == 3) {
        HIP_OK(hipEventRecord(c->fork_ev, c->stream));
        HIP_OK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
        middle(c->stream);
        edge(c->side);
        yfaces(c->stream);
        HIP_OK(hipEventRecord(c->join_ev, c->side));
        HIP_OK(hipStreamWaitEvent(c->stream, c->join_ev, 0));
        return;
    }
    if (c->fused_overlap == 2) middle(c->stream);
    HIP_OK(hipEventRecord(c->fork_ev, c->stream));
    HIP_OK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    yfaces(c->side);
    HIP_OK(hipEventRecord(c->join_ev, c->side));
    if (c->fused_overlap == 1) middle(c->stream);
    edge(c->stream);
    HIP_OK(hipStreamWaitEvent(c->stream, c->join_ev, 0));
}

template <int TX, bool VTI>
void step_elastic(fwi3d_ctx* c, const BShot* T, StepIdx it, int nflux, int nint, int nshell_blk, int nzslab,
                  int next_blk) {
    timed(c, 1, [&] {
        hipLaunchKernelGGL((elastic_flux_batch<TX, VTI>), dim3(nflux), dim3(64, TX), 0, c->stream, T, BMap{}, it);
    });
    if (nint > 0)
        timed(c, 0, [&] {
            hipLaunchKernelGGL((elastic_interior_batch<TX, VTI>), dim3(nint), dim3(64, TX), 0, c->stream, T, BMap{}, it);
        });
    if (nzslab > 0 && !(c->prof && !c->prof_steps_only)) {
        // slabs, edges and receivers in one launch (separate launches when every kernel is
        // timed on its own, fwi3d_profile level 1)
        hipLaunchKernelGGL(elastic_edges_slabs_ext<VTI>, dim3(nzslab + 3 * nshell_blk + next_blk), dim3(256), 0,
                           c->stream, T, it, nzslab);
        return;
    }
    timed(c, 2, [&] {
        hipLaunchKernelGGL(elastic_boundary_ext<VTI>, dim3(3 * nshell_blk + next_blk), dim3(256), 0, c->stream, T, it);
    });
    if (nzslab > 0)
        timed(c, 3, [&] {
            hipLaunchKernelGGL((elastic_zslab_batch<kZSlabTX, VTI>), dim3(nzslab), dim3(8 * kZSlabTX), 0, c->stream, T,
                               BMap{}, it);
        });
}

void propagate_elastic(fwi3d_ctx* c, bool vti, bool adj, const float* d_model, const float* d_src, float* d_rcv,
                       const fwi3d_injector* inj, const fwi3d_injector* ext, double* d_grad, const fwi3d_param* par,
                       const Geom& g) {
    const long long n = g.n;
    const int npar = vti ? 6 : 3;
    const int nmf = vti ? 7 : 3;  // device model fields (VTI adds c11b)
    const int nT = vti ? 3 : 1;
    c->fields.ensure(9 * n * sizeof(float));
    c->model.ensure(nmf * n * sizeof(float));
    c->aux.ensure((nT + 6) * n * sizeof(float));  // T (1 or 3), 6 shear fluxes
    float* F = c->fields.as<float>();
    float* Mf = c->model.as<float>();
    HIP_OK(hipMemsetAsync(F, 0, 9 * n * sizeof(float), c->stream));
    HIP_OK(hipMemsetAsync(Mf, 0, nmf * n * sizeof(float), c->stream));
    to_padded(c, Mf, d_model, npar, g);
    EArgs A{};
    A.lam = Mf; A.mu = Mf + n; A.rho = Mf + 2 * n;
    if (vti) {
        A.c13 = Mf + 3 * n; A.eps = Mf + 4 * n; A.c66 = Mf + 5 * n; A.c11b = Mf + 6 * n;
        hipLaunchKernelGGL(vti_c11b_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, Mf + 6 * n, A.eps, A.mu, n);
    }
    const Tapers tp = setup_tapers(c, par, g);
    A.tz2 = tp.tz2; A.tx2 = tp.tx2; A.ty2 = tp.ty2;
    A.g = g;
    for (int f = 0; f < 6; f++) A.bc[f] = par->bc[f];
    A.aux = c->aux.as<float>();
    InjTab I = upload_inj(c, c->injtab, inj, c->injhost[0]);
    InjTab E{};
    if (!adj) E = upload_inj(c, c->exttab, ext, c->injhost[1]);

    const int nt = par->nt, sub = par->sub;
    const float dt = par->dt;
    A.dt = dt; A.dt2 = dt * dt;
    const bool fwd_snap = !adj && sub > 0;
    const int phys = vti ? FWI3D_VTI : FWI3D_ELASTIC;
    if (fwd_snap) snap_begin(c, phys, 3, par, g);
    if (adj && d_grad && sub > 0) snap_check(c, phys, par, g);
    float* prev = F;
    float* curr = F + 3 * n;
    float* next = F + 6 * n;

    const int nthr = 256;
    const int nbi = (I.ntgt + nthr - 1) / nthr;
    auto inject = [&](float* out, int it, int mode) {
        if (I.ntgt == 0) return;
        if (vti)
            hipLaunchKernelGGL(elastic_inject<true>, dim3(nbi, 3), dim3(nthr), 0, c->stream, out, A, I, d_src, nt, it, mode);
        else
            hipLaunchKernelGGL(elastic_inject<false>, dim3(nbi, 3), dim3(nthr), 0, c->stream, out, A, I, d_src, nt, it, mode);
    };
    inject(prev, 0, 1);

    int TX, ychunk, echunk, fchunk;
    tiles_for(c, vti ? FWI3D_VTI : FWI3D_ELASTIC, g, TX, ychunk, echunk, fchunk);
    const int nblk = tile_grid_yc(A, g, TX, false, ychunk);
    EArgs AI = A;
    const int nint = tile_grid_yc(AI, g, TX, true, ychunk);
    // one-shot table for the time-step kernels (rotation from the step index)
    BShot T1h{};
    T1h.phys = vti ? FWI3D_VTI : FWI3D_ELASTIC; T1h.t0 = 0; T1h.nt = par->nt; T1h.nf = 3 * n; T1h.F = F; T1h.g = g;
    T1h.ef = A; T1h.ei = AI;
    T1h.nshell_pts = shell_edge_count(g);
    T1h.nshell_blk = (T1h.nshell_pts + 255) / 256;
    const int nzslab = zslab_tiles(T1h.zs, g);
    const bool fused = c->use_fused;
    FusedPlan FP{};
    if (fused) FP = fused_plan(T1h, g, TX, ychunk, fchunk, echunk);
    T1h.I = I; T1h.src = d_src; T1h.ninj_blk = nbi;
    T1h.E = E; T1h.rcv = d_rcv; T1h.next_blk = (E.npts + 127) / 128;
    const BShot* T1 = upload_shot1(c, T1h);

    const long long nunp = (long long)g.nx * g.ny * g.nz;
    auto grad_event = [&](int itl, const float* adjf, const float* lagf) {
        grad_event_elastic(c->stream, vti, ctx_snap(c), itl, nt, sub, dt, adjf, lagf, A, d_grad, nunp);
    };
    // fields of local step k: prev = F[k % 3], curr = F[(k + 1) % 3] (the kernels' rotation)
    auto field = [&](int k) { return F + (long long)(k % 3) * 3 * n; };
    // one time step: receivers read curr; flux, interior, shell and z slabs write next;
    // sources add to it (the batch kernels over the one-shot table: a batched shot runs
    // the same code)
    // receivers are extracted by workgroups of the shell launch
    const int ext_blk = (!adj && E.npts > 0) ? extract_blocks256(E.npts, 3) : 0;
    auto step = [&](StepIdx si) {
        timed(c, 4, [&] {
            if (fused) {
                if (vti) {
                    if (TX == 8) step_elastic_fused<8, true>(c, T1, si, FP, ext_blk);
                    else step_elastic_fused<4, true>(c, T1, si, FP, ext_blk);
                } else {
                    if (TX == 8) step_elastic_fused<8, false>(c, T1, si, FP, ext_blk);
                    else step_elastic_fused<4, false>(c, T1, si, FP, ext_blk);
                }
                return;
            }
            if (vti) {
                if (TX == 8) step_elastic<8, true>(c, T1, si, nblk, nint, T1h.nshell_blk, nzslab, ext_blk);
                else step_elastic<4, true>(c, T1, si, nblk, nint, T1h.nshell_blk, nzslab, ext_blk);
            } else {
                if (TX == 8) step_elastic<8, false>(c, T1, si, nblk, nint, T1h.nshell_blk, nzslab, ext_blk);
                else step_elastic<4, false>(c, T1, si, nblk, nint, T1h.nshell_blk, nzslab, ext_blk);
            }
        });
        if (I.ntgt > 0) {
            if (vti) hipLaunchKernelGGL(elastic_inject_batch<true>, dim3(3 * nbi), dim3(nthr), 0, c->stream, T1, BMap{}, si);
            else hipLaunchKernelGGL(elastic_inject_batch<false>, dim3(3 * nbi), dim3(nthr), 0, c->stream, T1, BMap{}, si);
        }
    };

    if (!fwd_snap && !(adj && d_grad && sub > 0)) {
        // no host work between steps: graph replay, except for fused two-stream steps of
        // 3 Mi cells or more, launched directly. Replayed from a graph against launched
        // directly, per step: 200^3 elastic 256 / 258 us (y faces on the second stream:
        // 288 / 271), VTI 96 x 192 x 192 (nz, nx, ny) 236 / 214 us; below the threshold
        // VTI 128^3 140 / 146 us, elastic 128^3 113 / 121 us.
        if (fused && c->fused_overlap != 0 && (long long)g.nx * g.ny * g.nz >= (3LL << 20))
            for (int it = 0; it < nt - 1; it++) step(StepIdx(it));
        else
            run_steps(c, 0, nt - 1, step);
    } else {
        for (int it = 0; it < nt - 1; it++) {
            if (fwd_snap && it % sub == 0) snap_put(c->stream, ctx_snap(c), it / sub, field(it + 1));
            if (adj && d_grad && sub > 0 && (nt - 1 - it) % sub == 0 && it != 0)
                grad_event((nt - 1 - it) / sub, field(it + 1), field(it));
            step(StepIdx(it));
        }
    }
    (void)prev; (void)curr; (void)next;
    if (fwd_snap && (nt - 1) % sub == 0) snap_put(c->stream, ctx_snap(c), (nt - 1) / sub, field(nt));
    if (!adj && E.npts > 0)
        hipLaunchKernelGGL(extract_batch, dim3(3 * T1h.next_blk), dim3(128), 0, c->stream, T1, BMap{}, StepIdx(nt - 1));
    if (adj && d_grad && sub > 0) grad_event(0, field(nt), field(nt - 1));
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    release_graphs(c);
    if (fwd_snap) c->snap_valid = true;
}

void validate(int physics, const fwi3d_param* par, int nx, int ny, int nz, float dx, float dy, float dz) {
    REQUIRE(physics == FWI3D_ACOUSTIC || physics == FWI3D_ELASTIC || physics == FWI3D_VTI, "unknown physics");
    REQUIRE(par != nullptr, "null parameters");
    REQUIRE(nx >= 12 && ny >= 12 && nz >= 12, "every grid axis needs at least 12 points (SBP closures)");
    REQUIRE(dx > 0 && dy > 0 && dz > 0, "grid spacing must be positive");
    REQUIRE(par->nt >= 2, "nt must be at least 2");
    REQUIRE(par->dt > 0, "dt must be positive");
    REQUIRE(par->sub >= 0, "sub must be >= 0");
    for (int f = 0; f < 6; f++) {
        REQUIRE(par->bc[f] >= 0 && par->bc[f] <= 3, "bc must be 0..3");
        REQUIRE(par->taper[f] >= 0, "taper widths must be >= 0");
    }
    REQUIRE(par->taper[0] + par->taper[1] <= nz && par->taper[2] + par->taper[3] <= nx && par->taper[4] + par->taper[5] <= ny,
            "opposite tapers overlap");
    REQUIRE(par->snap_precision >= 0 && par->snap_precision <= 2, "snap_precision must be 0..2");
    REQUIRE(physics == FWI3D_ACOUSTIC || par->free_surface_stiffness > 0 || true, "");
    if (physics == FWI3D_ACOUSTIC) REQUIRE(par->free_surface_stiffness > 0, "free_surface_stiffness must be positive");
}

int fail(fwi3d_ctx* c, const Err& e) {
    fwi3d::set_error(c, e.msg);
    return e.code;
}

}  // namespace

#include "batch.hpp"

// ======================================================================================
extern "C" {

int fwi3d_create(fwi3d_ctx** out, int device) {
    if (!out) return FWI3D_EINVAL;
    *out = nullptr;
    fwi3d_ctx* c = new fwi3d_ctx();
    try {
        int ndev = 0;
        HIP_OK(hipGetDeviceCount(&ndev));
        REQUIRE(device >= 0 && device < ndev, "device index out of range");
        c->device = device;
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
        load_tables(device);
        scratch_setup(device);
    } catch (const Err& e) {
        int code = e.code;
        delete c;
        return code;
    }
    *out = c;
    return FWI3D_OK;
}

int fwi3d_destroy(fwi3d_ctx* c) {
    if (!c) return FWI3D_OK;
    // servers first: each stops its thread, which steps a batch of this context
    while (!c->servers.empty()) fwi3d_server_destroy(c->servers.back());
    while (!c->batches.empty()) fwi3d_batch_destroy(c->batches.back());
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    c->fields.release(); c->model.release(); c->aux.release(); c->snaps.release(); c->snapsc.release();
    c->tapers.release(); c->injtab.release(); c->exttab.release(); c->tmpio.release();
    c->opmodel.release(); c->opscal.release(); c->optrace.release(); c->opgrad.release(); c->opsyn.release(); c->opdas.release();
    c->opt.release(); c->shot1_dev.release();
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FWI3D_OK;
}

const char* fwi3d_last_error(const fwi3d_ctx* c) {
    // a copy per calling thread: the context's message may be replaced by another
    // thread (a server's) while the caller still reads the returned string
    static thread_local std::string copy;
    if (!c) return "null context";
    std::lock_guard<std::mutex> lk(const_cast<fwi3d_ctx*>(c)->err_mu);
    copy = c->err;
    return copy.c_str();
}

int fwi3d_injector_tables(int kind, const float* loc, int npts, int nx, int ny, int nz, float dx, float dy, float dz,
                          float ox, float oy, float oz, int* xind, int* yind, int* zind, float* xw, float* yw,
                          float* zw) {
    if (!(kind == 0 || kind == 1) || npts < 0 || nx < 3 || ny < 3 || nz < 3) return FWI3D_EINVAL;
    const int nw = kind ? 6 : 3;
    // The arithmetic is the reference's single-precision one (injector.cpp:28-80):
    // pivot and alpha in float, squares in double (std::pow(float,int)), H weights
    // rounded to float. A position half-way between two nodes therefore picks the
    // same pivot as the reference.
    const float hc[4] = {(float)kHnorm[0], (float)kHnorm[1], (float)kHnorm[2], (float)kHnorm[3]};
    auto hw = [&](int i, int n) -> float {
        if (i < 4) return hc[i];
        if (i >= n - 4) return hc[n - 1 - i];
        return 1.0f;
    };
    auto one = [&](float x, float o, float d, int n, int* ind, float* w) -> bool {
        const float r = std::round((x - o) / d);
        if (!(r >= 0.f && r <= (float)(n - 1))) return false;
        int piv = (int)r;
        if (piv == 0) piv = 1;
        if (piv == n - 1) piv = n - 2;
        ind[0] = piv - 1; ind[1] = piv; ind[2] = piv + 1;
        const float a = (x - d * piv - o) / d;
        const double a2 = (double)a * (double)a;
        w[0] = (float)((a2 - a) / (2 * d * hw(piv - 1, n)));
        w[1] = (float)((-a2 + 1) / (d * hw(piv, n)));
        w[2] = (float)((a2 + a) / (2 * d * hw(piv + 1, n)));
        w[nw + 0] = (float)((a2 - a) / 2);
        w[nw + 1] = (float)(-a2 + 1);
        w[nw + 2] = (float)((a2 + a) / 2);
        if (nw == 6) {
            w[3] = (1 - 2 * a) / (2 * d * d * hw(piv - 1, n));
            w[4] = 2 * a / (d * d * hw(piv, n));
            w[5] = (-1 - 2 * a) / (2 * d * d * hw(piv + 1, n));
            w[nw + 3] = (1 - 2 * a) / (2 * d);
            w[nw + 4] = 2 * a / (d);
            w[nw + 5] = (-1 - 2 * a) / (2 * d);
        }
        return true;
    };
    for (int i = 0; i < npts; i++) {
        bool ok = one(loc[3 * i + 0], ox, dx, nx, xind + 3 * i, xw + 2 * nw * i) &&
                  one(loc[3 * i + 1], oy, dy, ny, yind + 3 * i, yw + 2 * nw * i) &&
                  one(loc[3 * i + 2], oz, dz, nz, zind + 3 * i, zw + 2 * nw * i);
        if (!ok) return FWI3D_EINVAL;
    }
    return FWI3D_OK;
}

int fwi3d_dipole_tables(const float* loc5, int npts, float gl, int nx, int ny, int nz, float dx, float dy, float dz,
                        float ox, float oy, float oz, int* xind, int* yind, int* zind, float* xw, float* yw, float* zw) {
    if (npts < 0 || nx < 3 || ny < 3 || nz < 3 || !(gl > 0)) return FWI3D_EINVAL;
    // dipole_m3::findIndicesWeights (injector.cpp:313-420): a delta_m3 stencil at each end
    // of the gauge, x -/+ gl/2 along the fiber. The reference evaluates the end positions
    // with the double cos/sin, so pivots and alpha come from double arithmetic, alpha is
    // stored in float, the weights as for delta_m3.
    const float hc[4] = {(float)kHnorm[0], (float)kHnorm[1], (float)kHnorm[2], (float)kHnorm[3]};
    auto hw = [&](int i, int n) -> float {
        if (i < 4) return hc[i];
        if (i >= n - 4) return hc[n - 1 - i];
        return 1.0f;
    };
    auto one = [&](double pos, float o, float d, int n, int* ind, float* w) -> bool {
        const double rr = std::round((pos - o) / d);
        if (!(rr >= 0.0 && rr <= (double)(n - 1))) return false;
        int piv = (int)rr;
        if (piv == 0) piv = 1;
        if (piv == n - 1) piv = n - 2;
        ind[0] = piv - 1; ind[1] = piv; ind[2] = piv + 1;
        const float a = (float)((pos - o) / d - piv);
        const double a2 = (double)a * (double)a;
        w[0] = (float)((a2 - a) / (2 * d * hw(piv - 1, n)));
        w[1] = (float)((-a2 + 1) / (d * hw(piv, n)));
        w[2] = (float)((a2 + a) / (2 * d * hw(piv + 1, n)));
        w[6] = (float)((a2 - a) / 2);
        w[7] = (float)(-a2 + 1);
        w[8] = (float)((a2 + a) / 2);
        return true;
    };
    for (int i = 0; i < npts; i++) {
        const float* L = loc5 + 5 * i;
        const double hx = gl / 2 * std::cos((double)L[3]) * std::cos((double)L[4]);
        const double hy = gl / 2 * std::cos((double)L[3]) * std::sin((double)L[4]);
        const double hz = gl / 2 * std::sin((double)L[3]);
        for (int e = 0; e < 2; e++) {
            const double sg = e ? 1.0 : -1.0;
            bool ok = one(L[0] + sg * hx, ox, dx, nx, xind + 6 * i + 3 * e, xw + 12 * i + 3 * e) &&
                      one(L[1] + sg * hy, oy, dy, ny, yind + 6 * i + 3 * e, yw + 12 * i + 3 * e) &&
                      one(L[2] + sg * hz, oz, dz, nz, zind + 6 * i + 3 * e, zw + 12 * i + 3 * e);
            if (!ok) return FWI3D_EINVAL;
        }
    }
    return FWI3D_OK;
}

int fwi3d_propagate_device(fwi3d_ctx* c, int physics, int adj, const float* d_model, const float* d_src, float* d_rcv,
                           const fwi3d_injector* inj, const fwi3d_injector* ext, double* d_grad, const fwi3d_param* par,
                           int nx, int ny, int nz, float dx, float dy, float dz) {
    if (!c) return FWI3D_EINVAL;
    try {
        validate(physics, par, nx, ny, nz, dx, dy, dz);
        REQUIRE(d_model && d_src, "null model or source");
        check_injector(inj, nx, ny, nz, "source");
        if (!adj) {
            check_injector(ext, nx, ny, nz, "receiver");
            REQUIRE(d_rcv || ext->npts == 0, "null receiver buffer");
        }
        HIP_OK(hipSetDevice(c->device));
        Geom g = make_geom(nx, ny, nz, dx, dy, dz);
        if (physics == FWI3D_ACOUSTIC)
            propagate_acoustic(c, adj, d_model, d_src, d_rcv, inj, ext, d_grad, par, g);
        else
            propagate_elastic(c, physics == FWI3D_VTI, adj, d_model, d_src, d_rcv, inj, ext, d_grad, par, g);
        drain_events(c);
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_propagate_device_noalloc(fwi3d_ctx* c, int physics, int adj, const float* d_model, const float* d_src,
                                   float* d_rcv, const fwi3d_injector* inj, const fwi3d_injector* ext,
                                   const fwi3d_param* par, int nx, int ny, int nz, float dx, float dy, float dz) {
    if (!c) return FWI3D_EINVAL;
    // the context's device buffers of the single-shot path, swapped for slices of the
    // static workspace for the duration of the call and restored afterwards
    DevBuf* bufs[7] = {&c->fields, &c->model, &c->aux, &c->tapers, &c->injtab, &c->exttab, &c->shot1_dev};
    DevBuf saved[7];
    bool swapped = false;
    const bool graphs0 = c->use_graphs, prof0 = c->prof;
    auto restore = [&] {
        if (swapped)
            for (int k = 0; k < 7; k++) *bufs[k] = saved[k];
        swapped = false;
        c->use_graphs = graphs0;
        c->prof = prof0;
    };
    try {
        validate(physics, par, nx, ny, nz, dx, dy, dz);
        REQUIRE(d_model && d_src, "null model or source");
        REQUIRE(par->sub == 0, "the allocation-free entry point keeps no snapshot store: sub must be 0 "
                               "(use fwi3d_propagate_device for imaging runs)");
        check_injector(inj, nx, ny, nz, "source");
        if (!adj) {
            check_injector(ext, nx, ny, nz, "receiver");
            REQUIRE(d_rcv || ext->npts == 0, "null receiver buffer");
        }
        HIP_OK(hipSetDevice(c->device));
        char* ws = (char*)scratch_ptr(c->device);
        REQUIRE(ws != nullptr, "no workspace address for this device");
        Geom g = make_geom(nx, ny, nz, dx, dy, dz);
        const size_t n = (size_t)g.n;
        const bool ac = physics == FWI3D_ACOUSTIC;
        const int ncomp = ac ? 1 : 3, nmf = ac ? 2 : (physics == FWI3D_ELASTIC ? 3 : 7);
        const int naux = ac ? 0 : (physics == FWI3D_ELASTIC ? 7 : 9);
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t sz[7] = {al(3 * ncomp * n * 4), al(nmf * n * 4), al(naux * n * 4),
                              al((size_t)(g.nz + g.nx + g.ny) * 4), al(inj_bytes(inj)),
                              adj ? 0 : al(inj_bytes(ext)), al(sizeof(BShot))};
        size_t total = 0;
        for (size_t x : sz) total += x;
        REQUIRE(total <= kScratchBytes, "this shot needs " + std::to_string(total) + " bytes of workspace; the "
                                        "allocation-free entry point holds at most " + std::to_string(kScratchBytes) +
                                        " bytes (use fwi3d_propagate_device)");
        size_t off = 0;
        for (int k = 0; k < 7; k++) {
            saved[k] = *bufs[k];
            DevBuf b;
            b.p = sz[k] ? (void*)(ws + off) : nullptr;
            b.bytes = sz[k];
            b.borrowed = true;
            *bufs[k] = b;
            off += sz[k];
        }
        swapped = true;
        c->use_graphs = false;  // graph instantiation may allocate
        c->prof = false;        // profiling creates events
        if (ac)
            propagate_acoustic(c, adj, d_model, d_src, d_rcv, inj, ext, nullptr, par, g);
        else
            propagate_elastic(c, physics == FWI3D_VTI, adj, d_model, d_src, d_rcv, inj, ext, nullptr, par, g);
        restore();
        return FWI3D_OK;
    } catch (const Err& e) {
        // the stream may still hold work on the workspace: wait before handing the
        // buffers back, so a later call cannot overlap this one's kernels
        (void)hipStreamSynchronize(c->stream);
        restore();
        return fail(c, e);
    }
}

int fwi3d_propagate(fwi3d_ctx* c, int physics, int adj, const float* model, const float* src, float* rcv,
                    const fwi3d_injector* inj, const fwi3d_injector* ext, float* grad, const fwi3d_param* par, int nx,
                    int ny, int nz, float dx, float dy, float dz) {
    if (!c) return FWI3D_EINVAL;
    try {
        validate(physics, par, nx, ny, nz, dx, dy, dz);
        REQUIRE(model && src && inj, "null model, source or injector");
        HIP_OK(hipSetDevice(c->device));
        const size_t N = (size_t)nx * ny * nz;
        const int np = physics == FWI3D_ACOUSTIC ? 2 : (physics == FWI3D_ELASTIC ? 3 : 6);
        const int ncomp = physics == FWI3D_ACOUSTIC ? 1 : 3;
        const int ns = physics == FWI3D_ACOUSTIC ? 1 : (inj->kind ? 6 : 3);
        const int ng = physics == FWI3D_ACOUSTIC ? 2 : 3;
        const size_t nsrc = (size_t)ns * inj->npts * par->nt;
        const size_t nrcv = adj ? 0 : (size_t)ncomp * (ext ? ext->npts : 0) * par->nt;
        const size_t bytes = np * N * 4 + nsrc * 4 + nrcv * 4 + (grad ? ng * N * 8 : 0) + 4 * 256;
        c->tmpio.ensure(bytes);
        char* base = c->tmpio.as<char>();
        auto align = [](size_t x) { return (x + 255) / 256 * 256; };
        float* dm = (float*)base;
        float* ds = (float*)(base + align(np * N * 4));
        float* dr = (float*)((char*)ds + align(nsrc * 4));
        double* dg = (double*)((char*)dr + align(nrcv * 4));
        HIP_OK(hipMemcpyAsync(dm, model, np * N * 4, hipMemcpyHostToDevice, c->stream));
        if (nsrc) HIP_OK(hipMemcpyAsync(ds, src, nsrc * 4, hipMemcpyHostToDevice, c->stream));
        std::vector<double> g64;
        if (grad) {
            g64.resize(ng * N);
            for (size_t i = 0; i < g64.size(); i++) g64[i] = grad[i];
            HIP_OK(hipMemcpyAsync(dg, g64.data(), ng * N * 8, hipMemcpyHostToDevice, c->stream));
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        int rc = fwi3d_propagate_device(c, physics, adj, dm, ds, nrcv ? dr : nullptr, inj, ext, grad ? dg : nullptr, par,
                                        nx, ny, nz, dx, dy, dz);
        if (rc != FWI3D_OK) return rc;
        if (nrcv) HIP_OK(hipMemcpy(rcv, dr, nrcv * 4, hipMemcpyDeviceToHost));
        if (grad) {
            HIP_OK(hipMemcpy(g64.data(), dg, ng * N * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < g64.size(); i++) grad[i] = (float)g64[i];
        }
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_tune(fwi3d_ctx* c, int param, int value) {
    if (!c) return FWI3D_EINVAL;
    if (param == FWI3D_TUNE_YCHUNK) {
        if (value < 0) return fail(c, Err{FWI3D_EINVAL, "ychunk must be >= 0 (0: default)"});
        c->tune_ychunk = value;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_TX) {
        if (value != 4 && value != 8) return fail(c, Err{FWI3D_EINVAL, "tile width TX must be 4 or 8"});
        c->tune_tx = value;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_GRAPHS) {
        if (value != 0 && value != 1) return fail(c, Err{FWI3D_EINVAL, "graphs must be 0 or 1"});
        c->use_graphs = value != 0;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_FUSED) {
        if (value != 0 && value != 1) return fail(c, Err{FWI3D_EINVAL, "fused must be 0 or 1"});
        c->use_fused = value != 0;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_FCHUNK) {
        if (value < 0) return fail(c, Err{FWI3D_EINVAL, "fused y chunk must be >= 0 (0: by grid size)"});
        c->fused_chunk = value;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_ECHUNK) {
        if (value < 0) return fail(c, Err{FWI3D_EINVAL, "edge y chunk must be >= 0 (0: the fused y chunk)"});
        c->fused_echunk = value;
        return FWI3D_OK;
    }
    if (param == FWI3D_TUNE_OVERLAP) {
        if (value < 0 || value > 3) return fail(c, Err{FWI3D_EINVAL, "overlap must be 0, 1, 2 or 3"});
        c->fused_overlap = value;
        return FWI3D_OK;
    }
    return fail(c, Err{FWI3D_EINVAL, "unknown tuning parameter"});
}

int fwi3d_autotune(fwi3d_ctx* c, int physics, int nx, int ny, int nz, int steps, int* tx_out, int* ychunk_out) {
    if (!c) return FWI3D_EINVAL;
    const bool prof0 = c->prof, steps_only0 = c->prof_steps_only;
    double ms0[fwi3d_ctx::kProfClasses];
    long long n0[fwi3d_ctx::kProfClasses];
    std::copy(c->prof_ms, c->prof_ms + fwi3d_ctx::kProfClasses, ms0);
    std::copy(c->prof_n, c->prof_n + fwi3d_ctx::kProfClasses, n0);
    const int tx0 = c->tune_tx, yc0 = c->tune_ychunk, ec0 = c->fused_echunk, fc0 = c->fused_chunk;
    auto restore = [&] {
        c->tune_tx = tx0; c->tune_ychunk = yc0; c->fused_echunk = ec0; c->fused_chunk = fc0;
        c->prof = prof0; c->prof_steps_only = steps_only0;
        std::copy(ms0, ms0 + fwi3d_ctx::kProfClasses, c->prof_ms);
        std::copy(n0, n0 + fwi3d_ctx::kProfClasses, c->prof_n);
    };
    try {
        REQUIRE(physics == FWI3D_ACOUSTIC || physics == FWI3D_ELASTIC || physics == FWI3D_VTI, "unknown physics");
        REQUIRE(nx >= 12 && ny >= 12 && nz >= 12, "every grid axis needs at least 12 points (SBP closures)");
        REQUIRE(steps >= 1 && steps <= 1000, "steps must be 1..1000");
        HIP_OK(hipSetDevice(c->device));
        drain_events(c);
        // a homogeneous synthetic model (vp 3, vs 1.7, rho 2.3, eps 0.1): the kernels'
        // timing does not depend on the values
        const size_t N = (size_t)nx * ny * nz;
        const float vp = 3.f, vs = 1.7f, rho = 2.3f, mu = rho * vs * vs, lam = rho * vp * vp - 2.f * mu;
        std::vector<float> vals;
        if (physics == FWI3D_ACOUSTIC) vals = {rho * vp * vp, 1.f / rho};
        else if (physics == FWI3D_ELASTIC) vals = {lam, mu, rho};
        else vals = {lam, mu, rho, lam, 0.1f, mu};
        DevBuf m;
        m.ensure(vals.size() * N * sizeof(float));
        std::vector<float> host(N);
        for (size_t q = 0; q < vals.size(); q++) {
            std::fill(host.begin(), host.end(), vals[q]);
            HIP_OK(hipMemcpy(m.as<float>() + q * N, host.data(), N * sizeof(float), hipMemcpyHostToDevice));
        }
        fwi3d_injector none{};
        fwi3d_param par{};
        par.nt = steps + 1;
        par.sub = 0;
        par.dt = 1e-4f;
        const int bc[6] = {1, 2, 2, 2, 2, 2};
        for (int f = 0; f < 6; f++) { par.bc[f] = bc[f]; par.taper[f] = 0; }
        par.taper_strength = 0.05f;
        par.free_surface_stiffness = 1.05f;
        par.snap_precision = FWI3D_SNAP_F32;
        // measure without a previous choice for this grid
        c->tuned.erase(std::remove_if(c->tuned.begin(), c->tuned.end(),
                                      [&](const fwi3d_ctx::Tiles& t) {
                                          return t.physics == physics && t.nx == nx && t.ny == ny && t.nz == nz;
                                      }),
                       c->tuned.end());
        auto run = [&](int tx, int yc) -> double {
            c->tune_tx = tx;
            c->tune_ychunk = yc;
            c->prof = true;
            c->prof_steps_only = true;
            for (int k = 0; k < fwi3d_ctx::kProfClasses; k++) { c->prof_ms[k] = 0; c->prof_n[k] = 0; }
            const int rc = fwi3d_propagate_device(c, physics, 0, m.as<float>(), m.as<float>(), nullptr, &none, &none,
                                                  nullptr, &par, nx, ny, nz, 0.01f, 0.01f, 0.01f);
            if (rc != FWI3D_OK) throw Err{rc, fwi3d::thread_error()};
            return c->prof_ms[4] / steps;  // whole time steps (the kernels overlap)
        };
        run(4, 8);  // warm-up
        int best_tx = 4, best_yc = 8, best_ec = 0, best_fc = 0;
        double best = 1e30;
        for (int tx : {4, 8})
            for (int yc : {4, 8, 16, 32}) {
                const double t = run(tx, yc);
                if (t < best) { best = t; best_tx = tx; best_yc = yc; }
            }
        if (physics != FWI3D_ACOUSTIC && c->use_fused && ny > 12) {
            // y chunk of the fused kernel's side tiles: about 0.6, 0.75, 1, 1.5, 2 or 2.5
            // side workgroups per CU (the edge tiles run one workgroup per CU; beside the
            // middle tiles, fewer edge workgroups than CUs can be fastest: 256^3 414
            // against 451 us per step with 160 side workgroups). A chunk other
            // than the grid-size rule's must be 3 % faster: the timings of 10-step runs
            // vary by a few per cent, and a noisy pick cost a batch of three grids 10 %.
            const int ntz = fused_tile_count(nz, FZ), ntx = fused_tile_count(nx, FW), nyo = ny - 12;
            int tried[6], ntried = 0;
            for (int target : {160, 192, 256, 384, 512, 640}) {
                const int ec = fused_side_chunk(ntz, ntx, nyo, target);
                if (std::find(tried, tried + ntried, ec) != tried + ntried) continue;
                tried[ntried++] = ec;
                c->fused_echunk = ec;
                const double t = run(best_tx, best_yc);
                if (t < 0.97 * best) { best = t; best_ec = ec; }
            }
            // middle-tile y chunk: one y chunk fewer or more per tile column than the
            // grid-size rule's (the middle tiles share the CUs with the edge tiles)
            c->fused_echunk = best_ec > 0 ? best_ec : ec0;
            if (fc0 <= 0) {
                const int mid = fused_middle_tiles(ntz, ntx);
                const int want = std::max(1, 512 / std::max(1, mid > 0 ? mid : ntz * ntx));
                const int rule = std::max(8, (nyo + want - 1) / want);
                for (int w : {want - 1, want + 1}) {
                    if (w < 1) continue;
                    const int fc = std::max(8, (nyo + w - 1) / w);
                    if (fc == rule) continue;
                    c->fused_chunk = fc;
                    const double t = run(best_tx, best_yc);
                    if (t < 0.97 * best) { best = t; best_fc = fc; }
                }
            }
        }
        restore();
        c->tuned.push_back(fwi3d_ctx::Tiles{physics, nx, ny, nz, best_tx, best_yc, best_ec, best_fc});
        if (tx_out) *tx_out = best_tx;
        if (ychunk_out) *ychunk_out = best_yc;
        m.release();
        return FWI3D_OK;
    } catch (const Err& e) {
        restore();
        return fail(c, e);
    }
}

int fwi3d_profile(fwi3d_ctx* c, int enable) {
    if (!c) return FWI3D_EINVAL;
    try {
        drain_events(c);
    } catch (const Err& e) {
        return fail(c, e);
    }
    c->prof = enable != 0;
    c->prof_steps_only = enable == 2;
    if (enable) {
        for (int k = 0; k < fwi3d_ctx::kProfClasses; k++) {
            c->prof_ms[k] = 0;
            c->prof_n[k] = 0;
        }
    }
    return FWI3D_OK;
}

int fwi3d_snapshots_read(fwi3d_ctx* c, float* out, long long count, int* slots, int* ncomp) {
    if (!c) return FWI3D_EINVAL;
    try {
        REQUIRE(c->snap_slots > 0 && c->snap_valid, "no snapshot store: run a forward call with sub > 0 first");
        if (slots) *slots = c->snap_slots;
        if (ncomp) *ncomp = c->snap_ncomp;
        if (!out) return FWI3D_OK;
        const int nx = c->snap_dims[0], ny = c->snap_dims[1], nz = c->snap_dims[2];
        const long long nunp = (long long)nx * ny * nz;
        REQUIRE(count >= (long long)c->snap_slots * c->snap_ncomp * nunp, "output array too small for the snapshot store");
        HIP_OK(hipSetDevice(c->device));
        const SnapDesc S = ctx_snap(c);
        const long long ne = (long long)S.slot_elems();
        const int pz = (int)(S.n / ((long long)nx * ny));  // padded z pitch of the store's fields
        c->tmpio.ensure((size_t)ne * sizeof(float));
        float* tmp = c->tmpio.as<float>();
        for (int k = 0; k < S.slots; k++) {
            with_snap_fmt(S.fmt, [&](auto tag) {
                constexpr int FMT = decltype(tag)::value;
                hipLaunchKernelGGL(snap_decode<FMT>, dim3(2048), dim3(256), 0, c->stream, snap_view<FMT>(S, k), tmp, ne);
            });
            HIP_OK(hipGetLastError());
            for (int q = 0; q < S.ncomp; q++)
                HIP_OK(hipMemcpy2DAsync(out + ((size_t)k * S.ncomp + q) * nunp, nz * sizeof(float), tmp + (size_t)q * S.n,
                                        pz * sizeof(float), nz * sizeof(float), (size_t)nx * ny, hipMemcpyDeviceToHost,
                                        c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
        }
        return FWI3D_OK;
    } catch (const Err& e) {
        return fail(c, e);
    }
}

int fwi3d_profile_read(const fwi3d_ctx* c, int kernel, double* ms, long long* n) {
    if (!c || kernel < 0 || kernel >= fwi3d_ctx::kProfClasses) return FWI3D_EINVAL;
    if (ms) *ms = c->prof_ms[kernel];
    if (n) *n = c->prof_n[kernel];
    return FWI3D_OK;
}

}  // extern "C"
