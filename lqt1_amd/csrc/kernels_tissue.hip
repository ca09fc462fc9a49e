// Fused tissue step: diffusion term from V^n (the reference's calc_du,
// GPU3D.cu:370-385 and new_11.cu:352-417 in 3D form; a grid with nz == 1 is
// new_11.cu's 2D sheet, without z terms), then the TPORd update
// of the voxel with that term (Itot GPU3D.cu:425-1220 and new_u_v
// :1222-1230), one voxel per lane over the compacted tissue list.
//
// V lives in a dense zero-padded grid (one cell of padding on every side),
// so every neighbour read is a fixed offset from the voxel's grid index:
// consecutive lanes read consecutive z, which keeps all 19 stencil reads
// coalesced.  Ghost cells (the reference's mirror map, GPU3D.cu:1566-1604)
// take the new V of the voxel they mirror inside the step kernel itself;
// ghost_fill_kernel only refreshes them after the host writes state.  All
// other non-tissue cells hold -90 mV (GPU3D.cu:305).
#include "internal.hpp"

namespace lqt {

// Fibre direction f and divergence of D of one voxel (chunk-major slots
// 0-2 and 3-5 of the static voxel data).
template <class R>
struct Fibre {
    R f0, f1, f2, b0, b1, b2;
};

template <class R, bool ANISO, bool FLAT>
__device__ __forceinline__ void load_fibre(Fibre<R>& F, const TissueArgs<R>& A, uint32_t k) {
    if (!ANISO) return;
    const R* b = A.fd + (size_t)(k >> 6) * (6 * kChunk) + (k & 63);
    constexpr uint32_t w = kChunk * sizeof(R);
    F.f0 = ld_nt(b, 0);
    F.f1 = ld_nt(b, w);
    F.b0 = ld_nt(b, 3 * w);
    F.b1 = ld_nt(b, 4 * w);
    if (!FLAT) {
        F.f2 = ld_nt(b, 2 * w);
        F.b2 = ld_nt(b, 5 * w);
    }
}

// du = div(D grad V) at one voxel from its neighbours G(dx, dy, dz)
// (new_11.cu:352-417 generalised to 3D; 2D sheet without z terms;
// isotropic: GPU3D.cu:370-385).
template <class R, bool ANISO, bool FLAT, class NB>
__device__ __forceinline__ R diffusion(const TissueArgs<R>& A, const Fibre<R>& F, R vc, NB G) {
    const R vxp = G(1, 0, 0), vxm = G(-1, 0, 0);
    const R vyp = G(0, 1, 0), vym = G(0, -1, 0);
    if (FLAT) {
        const R vxx = (vxm + vxp - R(2) * vc) * A.inv_h2;
        const R vyy = (vym + vyp - R(2) * vc) * A.inv_h2;
        if (!ANISO) return A.dcross * (vxx + vyy);
        const R vxy = (G(1, 1, 0) + G(-1, -1, 0) - G(-1, 1, 0) - G(1, -1, 0)) * A.inv_4h2;
        const R dd = A.ddiff, dc = A.dcross;
        return (dc + dd * F.f0 * F.f0) * vxx + (dc + dd * F.f1 * F.f1) * vyy + R(2) * dd * F.f0 * F.f1 * vxy +
               (F.b0 * (vxp - vxm) + F.b1 * (vyp - vym)) * A.inv_2h;
    }
    const R vzp = G(0, 0, 1), vzm = G(0, 0, -1);
    if (!ANISO)
        return A.dcross * ((vxm + vxp - R(2) * vc) + (vym + vyp - R(2) * vc) + (vzm + vzp - R(2) * vc)) * A.inv_h2;
    const R vxx = (vxm + vxp - R(2) * vc) * A.inv_h2;
    const R vyy = (vym + vyp - R(2) * vc) * A.inv_h2;
    const R vzz = (vzm + vzp - R(2) * vc) * A.inv_h2;
    const R vxy = (G(1, 1, 0) + G(-1, -1, 0) - G(-1, 1, 0) - G(1, -1, 0)) * A.inv_4h2;
    const R vxz = (G(1, 0, 1) + G(-1, 0, -1) - G(-1, 0, 1) - G(1, 0, -1)) * A.inv_4h2;
    const R vyz = (G(0, 1, 1) + G(0, -1, -1) - G(0, -1, 1) - G(0, 1, -1)) * A.inv_4h2;
    const R dd = A.ddiff, dc = A.dcross;
    // D = Dc*I + (Df-Dc) f f^T ; du = sum_ij D_ij d_i d_j V + sum_j b_j d_j V
    return (dc + dd * F.f0 * F.f0) * vxx + (dc + dd * F.f1 * F.f1) * vyy + (dc + dd * F.f2 * F.f2) * vzz +
           R(2) * dd * (F.f0 * F.f1 * vxy + F.f0 * F.f2 * vxz + F.f1 * F.f2 * vyz) +
           (F.b0 * (vxp - vxm) + F.b1 * (vyp - vym) + F.b2 * (vzp - vzm)) * A.inv_2h;
}

// Stimulus current of voxel k on this step (GPU3D.cu:387-422).
template <class R>
__device__ __forceinline__ R stimulus(const TissueArgs<R>& A, uint32_t k) {
    R ist = R(0);
    if (A.stim_on) {
        const float off = A.stim_off[k];
        if (A.tc >= (double)off && A.tc < (double)off + A.stim_dur) ist = A.stim_strength;
    }
    if (A.s2_on && A.s2[k]) ist = A.s2_strength;
    return ist;
}

// V guard of GPU3D.cu:1796-1829, one atomic per wavefront.  Lanes with
// on == false take no part.
template <class R>
__device__ __forceinline__ void guard_v(const TissueArgs<R>& A, R vn, bool held, bool on) {
    const bool lo = on && vn < R(-90) && !held, hi = on && vn > R(200);
    const bool bad = on && (!(vn == vn) || !(vn - vn == R(0)));
    const unsigned long long mlo = __ballot(lo), mhi = __ballot(hi), mbad = __ballot(bad);
    if ((mlo | mhi | mbad) && (threadIdx.x & 63) == (__ffsll((long long)(mlo | mhi | mbad)) - 1)) {
        if (mlo) atomicAdd(&A.guard[0], (unsigned long long)__popcll(mlo));
        if (mhi) atomicAdd(&A.guard[1], (unsigned long long)__popcll(mhi));
        if (mbad) atomicAdd(&A.guard[2], (unsigned long long)__popcll(mbad));
    }
}

// The ghost mirrors of this wavefront's voxels (GPU3D.cu:1566-1604) take
// their new V: gc ghost cells per lane (0 on lanes without a voxel), listed
// from gout[gwave[chunk]] on in lane order.
template <class R>
__device__ __forceinline__ void write_ghosts(const TissueArgs<R>& A, uint32_t chunk, int gc, R vn) {
    if (!__ballot(gc != 0)) return;
    const int lane = threadIdx.x & 63;
    int x = gc;  // inclusive prefix sum over the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    const int first = A.gwave[chunk] + x - gc;
    for (int i = 0; i < gc; ++i) *(R*)((char*)A.vnew + (uint32_t)A.gout[first + i] * (uint32_t)sizeof(R)) = vn;
}

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X.
// This remap gives each XCD one contiguous range of tiles so that the
// neighbouring z-rows and x/y-planes a tile touches are reused from that
// XCD's own L2.  (Measured on the ventricle: 7 % fewer L2 misses, but the
// plain order, where all XCDs sweep one front and share the planes through
// the last-level cache, is faster; both are autotuner candidates.)
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    const int64_t per = nb >> 3, rem = nb & 7;
    const int64_t x = b & 7, i = b >> 3;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// FLAT: a 2D sheet (nz == 1, new_11.cu:352-417), no z terms.  B lanes per
// workgroup; XCD: the workgroup -> voxel-range map of xcd_tile (otherwise
// workgroup b takes voxels b*B .. b*B+B-1); W: at least W waves per SIMD
// for the register allocator (the FP32 anisotropic step needs 155
// registers, so three waves fit either way; FP64 spills at 2 and not at 1);
// LDS: a wavefront whose 64 voxels are one
// run along z stages the 9 columns around the run (z0-1 .. z0+64) in LDS
// and reads its 19-point stencil from there (3D anisotropic path).
template <class R, bool ANISO, bool FLAT, int B, bool XCD, int W, bool LDS>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(W))) void tissue_step_kernel(TissueArgs<R> A) {
    __shared__ R tile[LDS ? B / 64 : 1][9][66];
    // 32-bit byte offsets from uniform bases let the loads share address
    // registers (voxel and padded-grid counts are bounded on the host)
    const int64_t blk = XCD ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const uint32_t k = (uint32_t)(blk * B + threadIdx.x);
    if (k >= A.nvox) return;
    constexpr uint32_t S = sizeof(R);
    const uint32_t pb = (uint32_t)A.vpos[k] * S;  // byte offset of the voxel in the grid
    const R* __restrict__ vo = A.vold;
    const uint32_t sx = (uint32_t)A.sx * S, sy = (uint32_t)A.sy * S;
    auto V = [&](uint32_t off) -> R { return *(const R*)((const char*)vo + off); };
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    bool staged = false;
    if (LDS && ANISO && !FLAT) {
        // all 64 lanes present (every staged read is then the neighbour of a
        // tissue voxel or of its z-neighbour) and consecutive along z
        const uint32_t pb0 = __shfl(pb, 0, 64);
        const bool ok = pb == pb0 + (uint32_t)lane * S;
        staged = __ballot(1) == ~0ull && __ballot(!ok) == 0ull;
        if (staged) {
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                const int dx = c / 3 - 1, dy = c % 3 - 1;
                const uint32_t base = pb0 + (uint32_t)(dx * (int)sx + dy * (int)sy);
                tile[wv][c][lane + 1] = V(base + (uint32_t)lane * S);
            }
            if (lane < 18) {
                const int c = lane >> 1, dx = c / 3 - 1, dy = c % 3 - 1;
                const uint32_t base = pb0 + (uint32_t)(dx * (int)sx + dy * (int)sy);
                if (lane & 1) tile[wv][c][65] = V(base + 64u * S);
                else tile[wv][c][0] = V(base - S);
            }
            // the wavefront reads what its other lanes wrote
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    // neighbour (dx, dy, dz) of the voxel
    auto G = [&](int dx, int dy, int dz) -> R {
        if (LDS && staged) return tile[wv][(dx + 1) * 3 + (dy + 1)][lane + 1 + dz];
        return V(pb + (uint32_t)(dx * (int)sx + dy * (int)sy + dz * (int)S));
    };
    Fibre<R> F;
    load_fibre<R, ANISO, FLAT>(F, A, k);
    const R vc = G(0, 0, 0);
    const R du = diffusion<R, ANISO, FLAT>(A, F, vc, G);
    const R ist = stimulus(A, k);

    CellState<R> s;
    load_state_chunked(s, A.sr, A.sd, k, A.sref);
    // cls: bits 0-1 class, bits 2-6 ghost cells mirroring this voxel,
    // bit 7 V held at -100 mV (new_11.cu:1247-1248)
    const int c = A.cls[k];
    const bool held = c & 0x80;
    const Consts<R> K = *A.consts;  // uniform: scalar loads
    R vn = tpord_update<R, false>(s, vc, c & 3, K, A.dt, ist, du, nullptr);
    if (held) vn = R(-100);
    store_state_chunked(s, A.sr, A.sd, k, A.sref);
    *(R*)((char*)A.vnew + pb) = vn;
    // lanes past the last voxel have returned (they are the top lanes of the
    // last wavefront), so every lane here holds a voxel
    write_ghosts(A, k >> 6, (c >> 2) & 31, vn);
    guard_v(A, vn, held, true);
}

template <class R>
__global__ __launch_bounds__(256) void ghost_fill_kernel(R* __restrict__ g, const int32_t* __restrict__ dst,
                                                          const int32_t* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[dst[i]] = g[src[i]];
}

template <class R>
__global__ __launch_bounds__(256) void fill_kernel(R* __restrict__ p, int64_t n, R value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// Beat tracker of every voxel after a step that started at `time`: the
// rules of the strand tracker (activation at the first new V >= -30 mV,
// OneD.cc:422; 90 % repolarisation against the peak since activation and V
// at the window's first step).
template <class R>
__global__ __launch_bounds__(256) void track_kernel(int64_t nvox, const int32_t* __restrict__ vpos,
                                                     const R* __restrict__ grid, double time,
                                                     VoxTrack* __restrict__ tr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvox) return;
    const double v = (double)grid[vpos[k]];
    VoxTrack t = tr[k];
    if (!(t.flags & 1)) {
        t.flags |= 1;
        t.vrest = v;
    }
    if (!(t.flags & 2)) {
        if (v >= -30.0) {
            t.flags |= 2;
            t.act = time;
            t.vmax = v;
        }
    } else if (!(t.flags & 4)) {
        if (v > t.vmax) t.vmax = v;
        if (v <= t.vmax - 0.9 * (t.vmax - t.vrest)) {
            t.flags |= 4;
            t.rep = time;
        }
    }
    tr[k] = t;
}

__global__ __launch_bounds__(256) void track_reset_kernel(int64_t nvox, VoxTrack* __restrict__ tr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvox) return;
    VoxTrack t;
    t.act = -1.0;
    t.rep = -1.0;
    t.vrest = 0.0;
    t.vmax = 0.0;
    t.flags = 0;
    t.pad = 0;
    tr[k] = t;
}

// Streaming copy (16 B per lane, nontemporal loads and stores, the
// fastest of the simple forms measured on MI355X): the reference point of
// the HBM counters and of achievable bandwidth.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_kernel(f32x4* __restrict__ dst, const f32x4* __restrict__ src,
                                                   int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

void launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t st) {
    const int64_t n = bytes / 16;
    const int B = 256;
    copy_kernel<<<(n + B - 1) / B, B, 0, st>>>((f32x4*)dst, (const f32x4*)src, n);
}

const TissueVariant kTissueVariants[kNumTissueVariants] = {
    {256, 0, 2, 0}, {256, 1, 2, 0}, {128, 1, 2, 0}, {512, 1, 2, 0},
    {256, 1, 1, 0}, {128, 1, 1, 0}, {256, 0, 2, 1}, {256, 1, 2, 1},
};

template <class R, bool ANISO, bool FLAT, int B, bool XCD, int W, bool LDS = false>
static void launch_v(const TissueArgs<R>& a, int64_t nvox, hipStream_t st) {
    tissue_step_kernel<R, ANISO, FLAT, B, XCD, W, LDS><<<(nvox + B - 1) / B, B, 0, st>>>(a);
}

template <class R, bool ANISO>
static void launch_a(bool flat, int variant, const TissueArgs<R>& a, int64_t nvox, hipStream_t st) {
    if (flat) {  // 2D sheets run the default shape only
        launch_v<R, ANISO, true, 256, false, 2>(a, nvox, st);
        return;
    }
    switch (variant) {
        case 1: launch_v<R, ANISO, false, 256, true, 2>(a, nvox, st); break;
        case 2: launch_v<R, ANISO, false, 128, true, 2>(a, nvox, st); break;
        case 3: launch_v<R, ANISO, false, 512, true, 2>(a, nvox, st); break;
        case 4: launch_v<R, ANISO, false, 256, true, 1>(a, nvox, st); break;
        case 5: launch_v<R, ANISO, false, 128, true, 1>(a, nvox, st); break;
        case 6: launch_v<R, ANISO, false, 256, false, 2, true>(a, nvox, st); break;
        case 7: launch_v<R, ANISO, false, 256, true, 2, true>(a, nvox, st); break;
        default: launch_v<R, ANISO, false, 256, false, 2>(a, nvox, st); break;
    }
}

void launch_tissue_step(bool fp32, bool aniso, bool flat, int variant, const void* args, int64_t nvox,
                        hipStream_t st) {
    if (fp32) {
        const TissueArgs<float>& a = *(const TissueArgs<float>*)args;
        if (aniso) launch_a<float, true>(flat, variant, a, nvox, st);
        else launch_a<float, false>(flat, variant, a, nvox, st);
    } else {
        const TissueArgs<double>& a = *(const TissueArgs<double>*)args;
        if (aniso) launch_a<double, true>(flat, variant, a, nvox, st);
        else launch_a<double, false>(flat, variant, a, nvox, st);
    }
}

void launch_ghost_fill(bool fp32, void* grid, const int32_t* dst, const int32_t* src, int64_t n,
                       hipStream_t st) {
    if (n == 0) return;
    const int B = 256;
    const int64_t G = (n + B - 1) / B;
    if (fp32) ghost_fill_kernel<float><<<G, B, 0, st>>>((float*)grid, dst, src, n);
    else ghost_fill_kernel<double><<<G, B, 0, st>>>((double*)grid, dst, src, n);
}

void launch_track(bool fp32, int64_t nvox, const int32_t* vpos, const void* grid, double time, VoxTrack* tr,
                  hipStream_t st) {
    if (nvox == 0) return;
    const int B = 256;
    const int64_t G = (nvox + B - 1) / B;
    if (fp32) track_kernel<float><<<G, B, 0, st>>>(nvox, vpos, (const float*)grid, time, tr);
    else track_kernel<double><<<G, B, 0, st>>>(nvox, vpos, (const double*)grid, time, tr);
}

void launch_track_reset(int64_t nvox, VoxTrack* tr, hipStream_t st) {
    if (nvox == 0) return;
    const int B = 256;
    track_reset_kernel<<<(nvox + B - 1) / B, B, 0, st>>>(nvox, tr);
}

void launch_fill(bool fp32, void* p, int64_t n, double value, hipStream_t st) {
    if (n == 0) return;
    const int B = 256;
    const int64_t G = (n + B - 1) / B;
    if (fp32) fill_kernel<float><<<G, B, 0, st>>>((float*)p, n, (float)value);
    else fill_kernel<double><<<G, B, 0, st>>>((double*)p, n, value);
}

}  // namespace lqt
