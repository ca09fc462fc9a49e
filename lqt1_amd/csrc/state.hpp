// Structure-of-arrays storage of CellState and the host-side folding of the
// run parameters into per-class constants.
#pragma once

#include <stdint.h>

#include "tpord_model.hpp"

// R-typed slots in storage order (the order of CellState).
#define LQT_RFIELDS(X)                                                                              \
    X(sm) X(sh) X(sj) X(mL) X(hL) X(hLp) X(a) X(iF) X(iS) X(ap) X(iFp) X(iSp) X(d) X(ff) X(fs)     \
    X(fcaf) X(fcas) X(jca) X(nca) X(ffp) X(fcafp) X(xrf) X(xrs) X(xs1) X(xs2) X(xk1) X(Jrelnp)    \
    X(Jrelp) X(nass) X(kss) X(cai) X(cass)
#define LQT_DFIELDS(X) X(nai) X(ki) X(cansr) X(cajsr) X(CaMKt)

namespace lqt {

enum RSlot {
#define LQT_ENUM(n) RS_##n,
    LQT_RFIELDS(LQT_ENUM)
#undef LQT_ENUM
    RS_COUNT
};
enum DSlot {
#define LQT_ENUM(n) DS_##n,
    LQT_DFIELDS(LQT_ENUM)
#undef LQT_ENUM
    DS_COUNT
};
static_assert(RS_COUNT == NR, "R slot count");
static_assert(DS_COUNT == ND, "D slot count");

// SoA view: slot j of cell k is r[j*pitch + k] / d[j*pitch + k].
template <class R>
struct SoA {
    R* r;
    double* d;
    int64_t pitch;
};

// The state is streamed: each value is read once and written once per step
// and the whole state of a large tissue is far larger than the caches, so
// the accesses carry the nontemporal hint.  Addresses are a uniform slot
// base plus a 32-bit byte offset, which the compiler turns into one shared
// offset register for all 37 slots (the voxel count stays below 2^29).
#ifndef LQT_STATE_NT
#define LQT_STATE_NT 1
#endif
template <class T>
__device__ __forceinline__ T ld_nt(const T* base, uint32_t byte_off) {
#if LQT_STATE_NT
    return __builtin_nontemporal_load((const T*)((const char*)base + byte_off));
#else
    return *(const T*)((const char*)base + byte_off);
#endif
}
template <class T>
__device__ __forceinline__ void st_nt(T v, T* base, uint32_t byte_off) {
#if LQT_STATE_NT
    __builtin_nontemporal_store(v, (T*)((char*)base + byte_off));
#else
    *(T*)((char*)base + byte_off) = v;
#endif
}

template <class R>
__device__ __forceinline__ void load_state(CellState<R>& s, const SoA<R>& a, uint32_t k) {
    const uint32_t orr = k * (uint32_t)sizeof(R), od = k * 8u;
#define LQT_LD(n) s.n = ld_nt(a.r + (size_t)RS_##n * a.pitch, orr);
    LQT_RFIELDS(LQT_LD)
#undef LQT_LD
#define LQT_LD(n) s.n = ld_nt(a.d + (size_t)DS_##n * a.pitch, od);
    LQT_DFIELDS(LQT_LD)
#undef LQT_LD
}

template <class R>
__device__ __forceinline__ void store_state(const CellState<R>& s, const SoA<R>& a, uint32_t k) {
    const uint32_t orr = k * (uint32_t)sizeof(R), od = k * 8u;
#define LQT_ST(n) st_nt(s.n, a.r + (size_t)RS_##n * a.pitch, orr);
    LQT_RFIELDS(LQT_ST)
#undef LQT_ST
#define LQT_ST(n) st_nt(s.n, a.d + (size_t)DS_##n * a.pitch, od);
    LQT_DFIELDS(LQT_ST)
#undef LQT_ST
}

// The gates that tp_group_b advances; a cell split over two lanes keeps the
// current ones of these on its group-B lane and all other slots on its
// group-A lane.
__host__ __device__ constexpr bool b_gate(int slot) {
    return slot == RS_d || slot == RS_ff || slot == RS_fs || slot == RS_fcaf || slot == RS_fcas || slot == RS_jca ||
           slot == RS_nca || slot == RS_ffp || slot == RS_fcafp || slot == RS_xrf || slot == RS_xrs ||
           slot == RS_xk1 || slot == RS_Jrelnp || slot == RS_Jrelp;
}

// store_state for one lane of a split cell: half 0 (group A) writes every
// slot but the group-B gates, half 1 writes only those.
template <class R>
__device__ __forceinline__ void store_state_half(const CellState<R>& s, const SoA<R>& a, uint32_t k, int half) {
    const uint32_t orr = k * (uint32_t)sizeof(R), od = k * 8u;
#define LQT_ST(n) \
    if (b_gate(RS_##n) == (half != 0)) st_nt(s.n, a.r + (size_t)RS_##n * a.pitch, orr);
    LQT_RFIELDS(LQT_ST)
#undef LQT_ST
    if (half == 0) {
#define LQT_ST(n) st_nt(s.n, a.d + (size_t)DS_##n * a.pitch, od);
        LQT_DFIELDS(LQT_ST)
#undef LQT_ST
    }
}

// The quarter of a cell split over four lanes that advances gate slot
// `slot`: 0 the gates of tp_group_a part 1 (INa, INaL, IKs), 1 those of part
// 2 (Ito), 2 those of tp_group_b part 1 (ICaL, Jrel), 3 those of part 2 (IKr,
// IK1).  Every other slot (the concentrations) belongs to quarter 0.
__host__ __device__ constexpr int quarter_of(int slot) {
    return (slot == RS_a || slot == RS_iF || slot == RS_iS || slot == RS_ap || slot == RS_iFp || slot == RS_iSp) ? 1
           : (slot == RS_xrf || slot == RS_xrs || slot == RS_xk1)                                              ? 3
           : b_gate(slot)                                                                                       ? 2
                                                                                                                : 0;
}

// store_state for one lane of a cell split over four lanes: quarter q writes
// the slots it owns.
template <class R>
__device__ __forceinline__ void store_state_quarter(const CellState<R>& s, const SoA<R>& a, uint32_t k, int q) {
    const uint32_t orr = k * (uint32_t)sizeof(R), od = k * 8u;
#define LQT_ST(n) \
    if (quarter_of(RS_##n) == q) st_nt(s.n, a.r + (size_t)RS_##n * a.pitch, orr);
    LQT_RFIELDS(LQT_ST)
#undef LQT_ST
    if (q == 0) {
#define LQT_ST(n) st_nt(s.n, a.d + (size_t)DS_##n * a.pitch, od);
        LQT_DFIELDS(LQT_ST)
#undef LQT_ST
    }
}

// Chunk-major storage (the tissue's state and static voxel data): voxel k
// belongs to chunk k / 64, the voxels of one wavefront, and a chunk keeps
// its slots one after another, 64 values each.  A wavefront then reads and
// writes its whole state as one contiguous block, and every slot is a
// constant offset from one per-lane address.
constexpr int kChunk = 64;
__host__ __device__ inline int64_t chunk_slot(int64_t k, int j, int nslot) {
    return ((k >> 6) * nslot + j) * kChunk + (k & 63);
}

// The tissue's stored slow accumulators (nai, ki, cansr, cajsr, CaMKt):
// double in the FP64 path; in the FP32 path float offsets from reference
// values of the tissue (SlowRef, set when the state is written from the
// host).  The update runs in double in registers either way; storing the
// offset keeps a step's change, which for ki (about 1e-6 mM per step at
// 145 mM) is far below one FP32 ULP of the value itself (1.5e-5 mM) but
// many ULPs of an offset that stays within a few mM of zero.
template <class R>
struct SlowStore {
    using type = double;
};
template <>
struct SlowStore<float> {
    using type = float;
};
struct SlowRef {
    double x[ND];
};

template <class R>
__device__ __forceinline__ void load_state_chunked(CellState<R>& s, const R* r, const typename SlowStore<R>::type* d,
                                                   uint32_t k, const SlowRef& ref) {
    using S = typename SlowStore<R>::type;
    const R* rb = r + (size_t)(k >> 6) * (NR * kChunk) + (k & 63);
    const S* db = d + (size_t)(k >> 6) * (ND * kChunk) + (k & 63);
#define LQT_LD(n) s.n = ld_nt(rb, (uint32_t)(RS_##n * kChunk * sizeof(R)));
    LQT_RFIELDS(LQT_LD)
#undef LQT_LD
    if constexpr (sizeof(S) == 4) {
#define LQT_LD(n) s.n = ref.x[DS_##n] + (double)ld_nt(db, (uint32_t)(DS_##n * kChunk * sizeof(S)));
        LQT_DFIELDS(LQT_LD)
#undef LQT_LD
    } else {
        (void)ref;
#define LQT_LD(n) s.n = ld_nt(db, (uint32_t)(DS_##n * kChunk * sizeof(S)));
        LQT_DFIELDS(LQT_LD)
#undef LQT_LD
    }
}

template <class R>
__device__ __forceinline__ void store_state_chunked(const CellState<R>& s, R* r, typename SlowStore<R>::type* d,
                                                    uint32_t k, const SlowRef& ref) {
    using S = typename SlowStore<R>::type;
    R* rb = r + (size_t)(k >> 6) * (NR * kChunk) + (k & 63);
    S* db = d + (size_t)(k >> 6) * (ND * kChunk) + (k & 63);
#define LQT_ST(n) st_nt(s.n, rb, (uint32_t)(RS_##n * kChunk * sizeof(R)));
    LQT_RFIELDS(LQT_ST)
#undef LQT_ST
    if constexpr (sizeof(S) == 4) {
#define LQT_ST(n) st_nt((float)(s.n - ref.x[DS_##n]), db, (uint32_t)(DS_##n * kChunk * sizeof(S)));
        LQT_DFIELDS(LQT_ST)
#undef LQT_ST
    } else {
        (void)ref;
#define LQT_ST(n) st_nt(s.n, db, (uint32_t)(DS_##n * kChunk * sizeof(S)));
        LQT_DFIELDS(LQT_ST)
#undef LQT_ST
    }
}

// Location of each of the reference's 40 states (TPORd.cc:200-252 order)
// in storage: kind 0 = V, 1 = R slot, 2 = D slot, 3 = constant 1 (the dead
// hsp/jp gates, never updated by TPORd.cc).
struct Ref40 {
    int kind, slot;
};
static const Ref40 kRef40[40] = {
    {0, 0},         {2, DS_nai},   {1, RS_nass},  {2, DS_ki},     {1, RS_kss},    {1, RS_cai},
    {1, RS_cass},   {2, DS_cansr}, {2, DS_cajsr}, {1, RS_sm},     {1, RS_sh},     {1, RS_sj},
    {3, 0},         {3, 0},        {1, RS_mL},    {1, RS_hL},     {1, RS_hLp},    {1, RS_a},
    {1, RS_iF},     {1, RS_iS},    {1, RS_ap},    {1, RS_iFp},    {1, RS_iSp},    {1, RS_d},
    {1, RS_ff},     {1, RS_fs},    {1, RS_fcaf},  {1, RS_fcas},   {1, RS_jca},    {1, RS_nca},
    {1, RS_ffp},    {1, RS_fcafp}, {1, RS_xrf},   {1, RS_xrs},    {1, RS_xs1},    {1, RS_xs2},
    {1, RS_xk1},    {1, RS_Jrelnp},{1, RS_Jrelp}, {2, DS_CaMKt},
};

// Initial values of TPORd::init (TPORd.cc:142-198) in the 40-line order.
static const double kInit40[40] = {
    -87.5, 7, 7, 145, 145, 1.0e-4, 1.0e-4, 1.2, 1.2, 0, 1, 1, 1, 1, 0, 1, 1, 0, 1, 1,
    0,     1, 1, 0,   1,   1,      1,      1,   1,   0, 1, 1, 0, 0, 0, 0, 1, 0, 0, 0,
};

// Cell class of an enum CellType value: only EPI and MCELL are tested by
// the model (TPORd.cc:523-1011); every other type behaves as ENDO.
__host__ __device__ inline int cell_class(int celltype) {
    return celltype == 2 ? CLS_EPI : (celltype == 4 ? CLS_M : CLS_ENDO);
}

}  // namespace lqt
