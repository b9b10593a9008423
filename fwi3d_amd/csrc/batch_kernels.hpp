// Time-step kernels over a table of shots. Each launch block finds its shot through a
// block -> shot map and runs the shared tile / point bodies of acoustic.hpp and
// elastic.hpp with that shot's arguments; the wavefield rotation (prev, curr, next) of
// a shot follows from the global step `it` and its admission step t0. The batch
// (batch.hpp) launches them over many shots; the single-shot path launches them with a
// one-entry table and no map, so a shot computes the same alone, in a batch, or through
// propagate().
#pragma once
#include "acoustic.hpp"
#include "elastic.hpp"
#include "fused.hpp"

namespace fwi3d {

struct BShot {
    int phys;       // FWI3D_ACOUSTIC, FWI3D_ELASTIC or FWI3D_VTI
    int t0, nt;     // global step of admission, time samples
    long long nf;   // elements of one wavefield set (components x padded field)
    float* F;       // three wavefield sets, rotated as prev / curr / next
    Geom g;
    AcuArgs a;      // acoustic: interior tiling
    EArgs ef, ei;   // elastic / VTI: flux-pass tiling, interior tiling
    InjTab I, E;    // source injector, receiver extractor
    const float* src;
    float* rcv;
    int nshell_blk;  // shell workgroups of one component (256 points each)
    int nshell_pts;  // shell points computed by the shell kernel (elastic / VTI: the edges and
                     // corners; with the fused kernel the edges of the y slabs only)
    ZSlabTiles zs;   // elastic / VTI: tiling of the slab kernel (z slabs, then the side slabs)
    int fused;       // elastic / VTI: the fused kernel updates the planes [6, ny-6), the slab
                     // launch only the y slabs (side tiles and edges), the flux pass ef
                     // covers the 8 planes at each y face
    FusedTiles ft;   // tiling of the fused kernel
    int ninj_blk;    // injection workgroups of one component (256 targets each)
    int next_blk;    // extraction workgroups of one component (128 receivers)
};

// Launch block -> shot index, and each shot's first block in the launch. A null map
// is a launch for the single shot T[0] (the single-shot path uses these kernels too).
struct BMap {
    const int* shot;
    const int* start;
    __device__ __forceinline__ int shot_of(int L) const { return shot ? shot[L] : 0; }
    __device__ __forceinline__ int first(int s) const { return start ? start[s] : 0; }
};

// Time-step index of a launch: `off`, plus the value at `base` when set (a device
// counter that a captured graph of several steps advances after each replay, so one
// graph serves every window of the time loop). An int converts to an eager index.
struct StepIdx {
    const int* base;
    int off;
    __host__ __device__ StepIdx(int v) : base(nullptr), off(v) {}
    __host__ __device__ StepIdx(const int* b, int o) : base(b), off(o) {}
    __device__ __forceinline__ int value() const { return base ? *base + off : off; }
};

// Advance the graph's step counter by k (the last node of a captured window).
__global__ void advance_step(int* __restrict__ base, int k) {
    if (blockIdx.x == 0 && threadIdx.x == 0) base[0] += k;
}
// Set the step counter (launched eagerly before the first replay).
__global__ void set_step(int* __restrict__ base, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) base[0] = v;
}

template <class Args>
__device__ __forceinline__ void bshot_rotate(Args& A, const BShot& S, int k) {
    A.prev = S.F + (long long)(k % 3) * S.nf;
    A.curr = S.F + (long long)((k + 1) % 3) * S.nf;
    A.next = S.F + (long long)((k + 2) % 3) * S.nf;
    A.prev_tapered = k > 0 ? 1 : 0;
}

// ---- time-step kernels: block -> (shot, local block); shots outside their stepping
// range (local step k in [0, nt-2]) exit at once.
// Occupancy: 8 waves per SIMD (forced, 58 VGPRs) measured faster than the compiler's
// 6 waves (75 VGPRs).
template <int TX>
__global__ __launch_bounds__(64 * TX) __attribute__((amdgpu_waves_per_eu(8))) void acoustic_interior_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    AcuArgs A = T[s].a;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s);
    acoustic_interior_tile<TX>(A, [&] { return l; });
}

__global__ __launch_bounds__(256) void acoustic_boundary_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    AcuArgs A = T[s].a;
    bshot_rotate(A, T[s], k);
    acoustic_boundary_block(A, L - m.first(s), T[s].nshell_pts);
}


// Receivers of the single shot T[0] at local step k (they read curr, which no kernel of
// the step writes), 256 per workgroup, component-major.
__device__ __forceinline__ void extract_block256(const BShot& S, int l, int k) {
    const int nb = (S.E.npts + 255) / 256;
    const int r = (l % nb) * 256 + threadIdx.y * blockDim.x + threadIdx.x;  // 256 x 1 or 64 x 4 threads
    if (r >= S.E.npts) return;
    const float* curr = S.F + (long long)((k + 1) % 3) * S.nf;
    extract_one(curr, S.g, S.E, S.rcv, S.nt, k, r, l / nb);
}
__host__ __device__ inline int extract_blocks256(int npts, int ncomp) { return ncomp * ((npts + 255) / 256); }

// Single-shot acoustic step in one launch of 64 x 4 threads per workgroup (the interior's
// TX = 4 tile): the nint interior tiles (XCD-remapped among themselves), then the nshell
// shell workgroups, then the receiver extraction of the same step. All three write
// disjoint outputs and read only curr, prev and the model.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void acoustic_step_ext(
    const BShot* __restrict__ T, StepIdx it, int nint, int nshell) {
    const int k = it.value() - T[0].t0;
    if (k < 0 || k > T[0].nt - 2) return;
    const int b = blockIdx.x;
    if (b >= nint + nshell) {
        extract_block256(T[0], b - nint - nshell, k);
        return;
    }
    AcuArgs A = T[0].a;
    bshot_rotate(A, T[0], k);
    if (b < nint) {
        const int L = xcd_remap(b, nint);
        acoustic_interior_tile<4>(A, [&] { return L; });
    } else {
        acoustic_boundary_block(A, b - nint, T[0].nshell_pts);
    }
}

// Single-shot path: the shell's nshell workgroups, then the receiver extraction of the
// same step in the workgroups past them (one launch fewer per step).
__global__ __launch_bounds__(256) void acoustic_boundary_ext(const BShot* __restrict__ T, StepIdx it, int nshell) {
    const int L = blockIdx.x;
    const int k = it.value() - T[0].t0;
    if (k < 0 || k > T[0].nt - 2) return;
    if (L >= nshell) {
        extract_block256(T[0], L - nshell, k);
        return;
    }
    AcuArgs A = T[0].a;
    bshot_rotate(A, T[0], k);
    acoustic_boundary_block(A, L, T[0].nshell_pts);
}

// Occupancy: elastic 5 waves per SIMD (96 VGPRs, 12 B/lane spilled) measured faster than
// the compiler's 4 waves without spills; VTI (more model fields per point) is faster at 4.
template <int TX, bool VTI>
__global__ __launch_bounds__(64 * TX) __attribute__((amdgpu_waves_per_eu(VTI ? 4 : 5))) void elastic_flux_batch(
    const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s);
    elastic_flux_tile<TX, VTI>(A, [&] { return l; });
}

// Occupancy: 4 waves per SIMD (<= 128 VGPRs) measured faster than the compiler's 3 waves
// (142-157 VGPRs); 5 waves spill heavily and are much slower.
template <int TX, bool VTI>
__global__ __launch_bounds__(64 * TX) __attribute__((amdgpu_waves_per_eu(4))) void elastic_interior_batch(
    const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ei;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s);
    elastic_interior_tile<TX, VTI>(A, [&] { return l; });
}

// Fused step over the planes [6, ny-6) (fused.hpp), 512 threads per workgroup: the
// middle tiles (EDGE false) or the edge tiles of each y chunk (EDGE true).
template <bool VTI, bool EDGE>
__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(VTI ? (EDGE ? 2 : 3) : (EDGE ? 2 : 4)))) void elastic_fused_batch(
    const BShot* __restrict__ T, BMap m, StepIdx it) {
    __shared__ FusedSmem<VTI> smem;
    // XCD-contiguous workgroups (neighbouring tiles share halo lines in one L2): measured
    // 125 against 232 us for the x-edge tiles at 200^3 with the dispatch order.
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const FusedTiles& FT = T[s].ft;
    const int l = L - m.first(s);
    if (EDGE) {
        // the side tiles first, then the corner tiles in their shorter y chunks
        const int per = fused_side_tiles(FT.ntz, FT.ntx), ns = per * FT.snyc;
        const int ncor = fused_corner_tiles(FT.ntz, FT.ntx);
        const bool side = l < ns;
        const int e = side ? l : l - ns, q = side ? per : ncor;
        const int tzx = side ? fused_side_tile(e % q, FT.ntz, FT.ntx) : fused_corner_tile(e % q, FT.ntz, FT.ntx);
        const int yc = side ? FT.sychunk : FT.cychunk, ya = FT.ya + (e / q) * yc;
        elastic_fused_tile<VTI, true, true>(A, FT, tzx & 0xffff, tzx >> 16, ya, min(FT.yb, ya + yc), smem);
    } else {
        const int mz = FT.ntz - 2, per = mz * (FT.ntx - 2);
        const int r = l % per, ya = FT.ya + (l / per) * FT.ychunk;
        elastic_fused_tile<VTI, false, false>(A, FT, 1 + r % mz, 1 + r / mz, ya, min(FT.yb, ya + FT.ychunk), smem);
    }
}

// Slab workgroup l of a shot whose planes [6, ny-6) the fused kernel updates: the y-slab
// side tiles (elastic_side_tile<1>), then the edges of the y slabs (3 components).
template <bool VTI>
__device__ __forceinline__ void elastic_yslab_block(const EArgs& A, const BShot& S, int l, float* smem) {
    const ZSlabTiles Z = S.zs;
    const int nside = 2 * Z.sntz * Z.snc1;
    if (l < nside) {
        elastic_side_tile<1, VTI>(A, Z, l, smem);
        return;
    }
    const int e = l - nside, nb = S.nshell_blk;
    elastic_boundary_block<VTI>(A, e % nb, e / nb, S.nshell_pts);
}

// Edges and corners of the shell, point-wise. Workgroups are not XCD-remapped: dispatch
// order gives every XCD the same mix of the rows along x and the z columns.
// Occupancy: 5 waves per SIMD measured faster than 6 and 7.
template <bool VTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void elastic_boundary_batch(
    const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s), nb = T[s].nshell_blk;
    elastic_boundary_block<VTI>(A, l % nb, l / nb, T[s].nshell_pts);
}

// Single-shot path: the shell's 3 nshell_blk workgroups, then the receiver extraction of
// the same step (as acoustic_boundary_ext).
template <bool VTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void elastic_boundary_ext(
    const BShot* __restrict__ T, StepIdx it) {
    const int L = blockIdx.x;
    const int k = it.value() - T[0].t0;
    if (k < 0 || k > T[0].nt - 2) return;
    const int nb = T[0].nshell_blk;
    if (L >= 3 * nb) {
        extract_block256(T[0], L - 3 * nb, k);
        return;
    }
    EArgs A = T[0].ef;
    bshot_rotate(A, T[0], k);
    elastic_boundary_block<VTI>(A, L % nb, L / nb, T[0].nshell_pts);
}

// Slabs of the shell away from its edges, 256 threads per workgroup: a shot's first
// zs.nzs workgroups tile the z slabs (elastic_zslab_tile, 8 x TXZ threads), the rest the
// x and then the y side slabs (elastic_side_tile, 32 x 8 threads); the branch is uniform
// per workgroup and the tiles share one LDS buffer.
template <int TXZ, bool VTI>
__global__ __launch_bounds__(8 * TXZ) void elastic_zslab_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    static_assert(8 * TXZ == 256, "z-slab and side-slab tiles both run 256 threads");
    __shared__ float smem[slab_lds_floats<VTI>()];
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s);
    const ZSlabTiles Z = T[s].zs;
    if (l < Z.nzs) {
        elastic_zslab_tile<TXZ, VTI>(A, Z, [&] { return l; }, smem);
    } else {
        const int r = l - Z.nzs, n0 = 2 * Z.sntz * Z.snc0;
        if (r < n0) elastic_side_tile<0, VTI>(A, Z, r, smem);
        else elastic_side_tile<1, VTI>(A, Z, r - n0, smem);
    }
}

// Single-shot path: the slab kernel's nslab workgroups (z slabs, then side slabs), then
// the shell's edge workgroups (3 components), then the receiver extraction of the same
// step, in one launch: the slab, edge and receiver work write disjoint outputs, and the
// edge workgroups fill the tail of the latency-bound slab workgroups.
template <bool VTI>
__global__ __launch_bounds__(256) void elastic_edges_slabs_ext(const BShot* __restrict__ T, StepIdx it, int nslab) {
    __shared__ float smem[slab_lds_floats<VTI>()];
    const int k = it.value() - T[0].t0;
    if (k < 0 || k > T[0].nt - 2) return;
    const int L = blockIdx.x;
    EArgs A = T[0].ef;
    bshot_rotate(A, T[0], k);
    if (L < nslab) {
        const ZSlabTiles Z = T[0].zs;
        if (L < Z.nzs) {
            elastic_zslab_tile<32, VTI>(A, Z, [&] { return L; }, smem);
        } else {
            const int r = L - Z.nzs, n0 = 2 * Z.sntz * Z.snc0;
            if (r < n0) elastic_side_tile<0, VTI>(A, Z, r, smem);
            else elastic_side_tile<1, VTI>(A, Z, r - n0, smem);
        }
        return;
    }
    const int e = L - nslab, nb = T[0].nshell_blk;
    if (e >= 3 * nb) {
        extract_block256(T[0], e - 3 * nb, k);
        return;
    }
    elastic_boundary_block<VTI>(A, e % nb, e / nb, T[0].nshell_pts);
}

// Batch form of the same: per shot, its slab workgroups, then its shell edge workgroups
// (receivers are a class of their own in a batch).
template <bool VTI>
__global__ __launch_bounds__(256) void elastic_edges_slabs_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    __shared__ float smem[slab_lds_floats<VTI>()];
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s);
    const ZSlabTiles Z = T[s].zs;
    const int nslab = Z.nzs + 2 * Z.sntz * (Z.snc0 + Z.snc1);
    if (l < nslab) {
        if (l < Z.nzs) {
            elastic_zslab_tile<32, VTI>(A, Z, [&] { return l; }, smem);
        } else {
            const int r = l - Z.nzs, n0 = 2 * Z.sntz * Z.snc0;
            if (r < n0) elastic_side_tile<0, VTI>(A, Z, r, smem);
            else elastic_side_tile<1, VTI>(A, Z, r - n0, smem);
        }
        return;
    }
    const int e = l - nslab, nb = T[s].nshell_blk;
    elastic_boundary_block<VTI>(A, e % nb, e / nb, T[s].nshell_pts);
}

// Single-shot path with the fused kernel: the y-slab side tiles and edges (nslab
// workgroups), then the receiver extraction of the same step. A kernel of its own, so
// its code holds only these tiles (the instruction cache is the limit of the slab code).
template <bool VTI>
__global__ __launch_bounds__(256) void elastic_yslabs_ext(const BShot* __restrict__ T, StepIdx it, int nslab) {
    __shared__ float smem[slab_lds_floats<VTI>()];
    const int k = it.value() - T[0].t0;
    if (k < 0 || k > T[0].nt - 2) return;
    const int L = blockIdx.x;
    if (L >= nslab) {
        extract_block256(T[0], L - nslab, k);
        return;
    }
    EArgs A = T[0].ef;
    bshot_rotate(A, T[0], k);
    elastic_yslab_block<VTI>(A, T[0], L, smem);
}

// Batch form (receivers are a class of their own in a batch).
template <bool VTI>
__global__ __launch_bounds__(256) void elastic_yslabs_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    __shared__ float smem[slab_lds_floats<VTI>()];
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    elastic_yslab_block<VTI>(A, T[s], L - m.first(s), smem);
}

// ---- source injection into `next` after the step (mode 0), receivers before it.
__global__ void acoustic_inject_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    AcuArgs A = T[s].a;
    bshot_rotate(A, T[s], k);
    const int t = (L - m.first(s)) * 256 + threadIdx.x;
    if (t < T[s].I.ntgt) acoustic_inject_one(A.next, A, T[s].I, T[s].src, T[s].nt, k, 0, t);
}

template <bool VTI>
__global__ void elastic_inject_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 2) return;
    EArgs A = T[s].ef;
    bshot_rotate(A, T[s], k);
    const int l = L - m.first(s), nb = T[s].ninj_blk;
    const int t = (l % nb) * 256 + threadIdx.x;
    if (t < T[s].I.ntgt) elastic_inject_one<VTI>(A.next, A, T[s].I, T[s].src, T[s].nt, k, 0, t, l / nb);
}

// Receivers of local step k in [0, nt-1] read the current field.
__global__ void extract_batch(const BShot* __restrict__ T, BMap m, StepIdx it) {
    const int L = blockIdx.x;
    const int s = m.shot_of(L);
    const int k = it.value() - T[s].t0;
    if (k < 0 || k > T[s].nt - 1) return;
    const int l = L - m.first(s), nb = T[s].next_blk;
    const int r = (l % nb) * 128 + threadIdx.x;
    if (r >= T[s].E.npts) return;
    const float* curr = T[s].F + (long long)((k + 1) % 3) * T[s].nf;
    extract_one(curr, T[s].g, T[s].E, T[s].rcv, T[s].nt, k, r, l / nb);
}

}  // namespace fwi3d
