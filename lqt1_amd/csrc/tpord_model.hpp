// TPORd (TP06 INa + ORd) cell update for gfx950 kernels, one cell per lane.
//
// Follows the update order of SingleCell/TPORd.cc:379-1021 (gates are
// Rush-Larsen-advanced from V^n before each current; V is advanced by
// forward Euler from the currents; then CaMK, fluxes and concentrations),
// with algebra rearranged for the GPU:
//   * every exp((v+a)/b) is exp(s*v + o) with s, o folded at compile time,
//     which is one FMA plus v_exp_f32 in the FP32 instantiation;
//   * tau = 1/(alpha+beta) gates use exp(-dt*(alpha+beta)) directly;
//   * the GHK terms (TPORd.cc:618-620,917,920) use x/expm1(x), which is the
//     same function without the 0/0 at V = 0;
//   * integer powers are products; pow(x,1.4) is exp2(1.4*log2 x).
// Conductances that depend only on the cell class and the run parameters
// (IKs scale, ISO, gene-therapy dose) are folded on the host into Consts.
//
// R is the arithmetic type of gates, currents and fast concentrations.
// The slow concentrations nai, ki, cansr, cajsr and CaMKt are held in double
// in both instantiations: at ki = 145 mM one float ULP is 1.5e-5 mM while a
// step changes ki by about 1e-6 mM, so a float accumulator would drop it.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace lqt {

enum CellClass { CLS_ENDO = 0, CLS_EPI = 1, CLS_M = 2 };

// Per-parameter-set constants (host-derived, see derive_consts()).
template <class R>
struct Consts {
    R gks[3];        // final GKs per class (TPORd.cc:693-733)
    R pca[3];        // final PCa per class (TPORd.cc:625-645)
    R dss_shift;     // ISO1 shift of dss (12 mV) or 0 (TPORd.cc:583)
    R fss_shift;     // ISO1 shift of fss (8 mV) or 0 (TPORd.cc:589)
    int clamp;       // CLAMP: V held (TPORd.cc:929-933)
};

// Per-class entry by selects, not by an indexed load: with the constants in
// registers this stays in registers for every lane's class.
template <class R>
__host__ __device__ __forceinline__ R pick_class(const R (&a)[3], int cls) {
    return cls == CLS_EPI ? a[CLS_EPI] : (cls == CLS_M ? a[CLS_M] : a[CLS_ENDO]);
}

// Stored state of one cell, apart from V (kept by the caller) and the two
// dead ORd INa gates hsp/jp, which TPORd never updates.
template <class R>
struct CellState {
    R sm, sh, sj, mL, hL, hLp, a, iF, iS, ap, iFp, iSp, d, ff, fs, fcaf, fcas, jca, nca, ffp,
        fcafp, xrf, xrs, xs1, xs2, xk1, Jrelnp, Jrelp;  // 28 Rush-Larsen states
    R nass, kss, cai, cass;                              // fast concentrations
    double nai, ki, cansr, cajsr, CaMKt;                 // slow accumulators
};

// Number of R-typed and double-typed slots of CellState in SoA storage.
constexpr int NR = 32;
constexpr int ND = 5;

struct Diag {
    double Iion, dV_dt, ICaL, IKs;
};

template <class R>
struct Math;

template <>
struct Math<float> {
    // exp(x) on v_exp_f32 (exp2) after the log2(e) scaling.
    static __device__ __forceinline__ float ex(float x) {
        return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
    }
    // exp(s*v + o) with the log2(e) factor folded into s and o.
    template <int DUMMY = 0>
    static __device__ __forceinline__ float exl(float v, double s, double o) {
        return __builtin_amdgcn_exp2f(fmaf(v, (float)(s * 1.4426950408889634), (float)(o * 1.4426950408889634)));
    }
    // Rush-Larsen decay exp(-dt * rate) as exp2(rate * ndt) with
    // ndt = -dt * log2(e) formed once per update.
    static __device__ __forceinline__ float decay_scale(float dt) { return -dt * 1.4426950408889634f; }
    static __device__ __forceinline__ float decay(float ndt, float rate) { return __builtin_amdgcn_exp2f(rate * ndt); }
    // Rush-Larsen step of a gate whose time constant is at least 0.8 ms at
    // every V, for dt <= 0.1 ms: x + (x_inf - x)(1 - exp(-z)), z = dt*rate
    // <= 0.125, with 1 - exp(-z) from its series to z^5 (truncation below
    // 4e-8 relative).  Forming 1 - exp(-z) from a rounded exp(-z) near 1
    // would keep only about log2(z / 6e-8) bits of it (8 bits for tau =
    // 1000 ms), a systematic error in the slow time constants.
    static __device__ __forceinline__ float rl_slow(float x, float xinf, float dts, float rate) {
        const float z = rate * dts;
        const float em1 = z * (1.0f - z * (0.5f - z * (1.0f / 6.0f - z * (1.0f / 24.0f - z * (1.0f / 120.0f)))));
        return x + (xinf - x) * em1;
    }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
    static __device__ __forceinline__ float pow14(float x) {
        return __builtin_amdgcn_exp2f(1.4f * __builtin_amdgcn_logf(x));
    }
    static __device__ __forceinline__ float sqrtc(double c) { return (float)::sqrt(c); }
    // x / expm1(x): Taylor series of expm1 below |x| < 0.25, exp otherwise.
    static __device__ __forceinline__ float x_over_expm1(float x, float ex_x) {
        float p = x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f + x * (1.0f / 120.0f + x * (1.0f / 720.0f))))));
        float em = fabsf(x) < 0.25f ? p : ex_x - 1.0f;
        return x == 0.0f ? 1.0f : x * __builtin_amdgcn_rcpf(em);
    }
};

template <>
struct Math<double> {
    static __device__ __forceinline__ double ex(double x) { return ::exp(x); }
    template <int DUMMY = 0>
    static __device__ __forceinline__ double exl(double v, double s, double o) { return ::exp(v * s + o); }
    // exp(-dt * rate): (-dt) * rate rounds exactly as -(dt * rate)
    static __device__ __forceinline__ double decay_scale(double dt) { return -dt; }
    static __device__ __forceinline__ double decay(double ndt, double rate) { return ::exp(rate * ndt); }
    // in double the exponential form keeps full precision: the same
    // expression as every other gate, x_inf - (x_inf - x) exp(-dt*rate)
    static __device__ __forceinline__ double rl_slow(double x, double xinf, double dts, double rate) {
        return xinf - (xinf - x) * ::exp(rate * -dts);
    }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double pow14(double x) { return ::pow(x, 1.4); }
    static __device__ __forceinline__ double sqrtc(double c) { return ::sqrt(c); }
    static __device__ __forceinline__ double x_over_expm1(double x, double) {
        double em = ::expm1(x);
        return x == 0.0 ? 1.0 : x / em;
    }
};

// Physical constants of TPORd.cc:381-421.
namespace k {
constexpr double nao = 140.0, cao = 1.8, ko = 5.4;
constexpr double BSRmax = 0.047, KmBSR = 0.00087, BSLmax = 1.124, KmBSL = 0.0087;
constexpr double cmdnmax = 0.05, kmcmdn = 0.00238, trpnmax = 0.07, kmtrpn = 0.0005;
constexpr double csqnmax = 10.0, kmcsqn = 0.8;
constexpr double aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015, KmCaMK = 0.15;
constexpr double Rg = 8314.0, T = 310.0, F = 96485.0;
constexpr double RTF = Rg * T / F;
constexpr double FRT = F / (Rg * T);
constexpr double vcell = 1000 * 3.14 * 0.0011 * 0.0011 * 0.01;
constexpr double Ageo = 2 * 3.14 * 0.0011 * 0.0011 + 2 * 3.14 * 0.0011 * 0.01;
constexpr double Acap = 2 * Ageo;
constexpr double vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell, vjsr = 0.0048 * vcell, vss = 0.02 * vcell;
// INaK constants (TPORd.cc:854-880) that do not depend on the state.
constexpr double Kko = 0.3582, Kki = 0.5, MgADP = 0.05, MgATP = 9.8, Kmgatp = 1.698e-7;
constexpr double H = 1.0e-7, eP = 4.2, Khp = 1.698e-7, Knap = 224.0, Kxkur = 292.0;
constexpr double b1 = 182.4 * MgADP;
constexpr double a2 = 687.2;
constexpr double kok2 = (1.0 + ko / Kko) * (1.0 + ko / Kko);
constexpr double a3num = 1899.0 * (ko / Kko) * (ko / Kko);
constexpr double b3c = 79300.0 * H / (1.0 + MgATP / Kmgatp);
constexpr double a4 = (639.0 * MgATP / Kmgatp) / (1.0 + MgATP / Kmgatp);
// NCX constants (TPORd.cc:753-852) that do not depend on the state.
constexpr double kna1 = 15.0, kna2 = 5.0, kna3 = 88.12, kasymm = 12.5;
constexpr double wna = 6.0e4, wca = 6.0e4, wnaca = 5.0e3, kcaon = 1.5e6, kcaoff = 5.0e3;
constexpr double h10 = kasymm + 1.0 + nao / kna1 * (1.0 + nao / kna2);
constexpr double h11 = nao * nao / (h10 * kna1 * kna2);
constexpr double h12 = 1.0 / h10;
constexpr double k1 = h12 * cao * kcaon;
}  // namespace k

// One NCX compartment (TPORd.cc:765-803 and 814-850): returns
// zna*JncxNa + zca*JncxCa given hna = exp(qna*vfrt), inv_hna, hca.
template <class R>
__device__ __forceinline__ R ncx(R na, R ca, R hna, R inv_hna, R inv_hca) {
    using M = Math<R>;
    const R h1 = R(1) + na * R(1.0 / k::kna3) * (R(1) + hna);
    const R inv_h1 = M::rcp(h1);
    const R h2 = na * hna * R(1.0 / k::kna3) * inv_h1;
    const R h3 = inv_h1;
    const R h4 = R(1) + na * R(1.0 / k::kna1) * (R(1) + na * R(1.0 / k::kna2));
    const R inv_h4 = M::rcp(h4);
    const R h5 = na * na * inv_h4 * R(1.0 / (k::kna1 * k::kna2));
    const R h6 = inv_h4;
    const R h7 = R(1) + R(k::nao / k::kna3) * (R(1) + inv_hna);
    const R inv_h7 = M::rcp(h7);
    const R h8 = R(k::nao / k::kna3) * inv_hna * inv_h7;
    const R h9 = inv_h7;
    const R k1 = R(k::k1);
    const R k2 = R(k::kcaoff);
    const R k3p = h9 * R(k::wca);
    const R k3pp = h8 * R(k::wnaca);
    const R k3 = k3p + k3pp;
    const R k4p = h3 * R(k::wca) * inv_hca;
    const R k4pp = h2 * R(k::wnaca);
    const R k4 = k4p + k4pp;
    const R k5 = R(k::kcaoff);
    const R k6 = h6 * ca * R(k::kcaon);
    const R k7 = h5 * h2 * R(k::wna);
    const R k8 = h8 * R(k::h11 * k::wna);
    const R x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
    const R x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
    const R x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
    const R x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
    const R inv_sum = M::rcp(x1 + x2 + x3 + x4);
    const R E1 = x1 * inv_sum, E2 = x2 * inv_sum, E3 = x3 * inv_sum, E4 = x4 * inv_sum;
    const R JncxNa = R(3) * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
    const R JncxCa = E2 * k2 - E1 * k1;
    return JncxNa + R(2) * JncxCa;
}

// The update of one cell in four parts.  tp_pre forms what both current
// groups read from the step's start state; tp_group_a and tp_group_b advance
// two disjoint sets of gates and return their currents; tp_finish sums the
// currents in the reference's order and advances V, CaMK and the
// concentrations.  Run one after the other (tpord_update) they are the
// single-lane update; run on two lanes of different wavefronts that swap
// their groups' currents, the two lanes hold the same V and concentrations
// bit for bit, every expression being the same.
template <class R>
struct Pre {
    R nai, ki, cansr, cajsr, CaMKt, nass, kss, cai, cass;
    R ENa, EK, EKs, CaMKb, fCaMKp, vfrt, ndt, e1, e2, r1, r2;
};

// Group A: INa, INaL, Ito, IKs and INaK, with the gates sm sh sj mL hL hLp
// a iF iS ap iFp iSp xs1 xs2.
template <class R>
struct GroupA {
    R INa, INaL, Ito, IKs, INaK;
};

// Group B: ICaL/ICaNa/ICaK, IKr, IK1, INaCa, IpCa, the background currents
// and Jrel, with the gates d ff fs fcaf fcas jca nca ffp fcafp xrf xrs xk1
// Jrelnp Jrelp.  The split strand kernel runs the groups on different
// wavefronts that meet at a barrier every step; this grouping is the one
// with the shortest measured step (1.79 us for a 100-cell strand on gfx950),
// although group B issues more (about 2,340 cycles per step before the
// barrier against 2,110 for group A, at 4 cycles per wave64 VALU instruction
// and 8 per transcendental one): with IK1 in group A the two issue about
// 2,230 each and the step takes 1.86 us, one wavefront per SIMD being bound
// by its dependency chains rather than by issue.
template <class R>
struct GroupB {
    R ICaL, ICaNa, ICaK, IKr, IK1, INaCa_i, INaCa_ss, IpCa, ICab, Jrel, INab, IKb;
};

// WHO: 0 for both groups, 1 for group A alone (without the ICaL/ICab term
// r2), 2 for group B alone (without ENa and EKs); what a group reads is
// computed the same way in every form.
template <class R, int WHO = 0>
__device__ __forceinline__ Pre<R> tp_pre(const CellState<R>& s, R v, R dt) {
    using M = Math<R>;
    const R nai = R(s.nai), ki = R(s.ki), cansr = R(s.cansr), cajsr = R(s.cajsr), CaMKt = R(s.CaMKt);
    const R nass = s.nass, kss = s.kss, cai = s.cai, cass = s.cass;

    // reversal potentials (:425-427)
    const R ENa = WHO == 2 ? R(0) : R(k::RTF) * M::log(R(k::nao) * M::rcp(nai));
    const R EK = R(k::RTF) * M::log(R(k::ko) * M::rcp(ki));
    const R EKs = WHO == 2 ? R(0) : R(k::RTF) * M::log(R(k::ko + 0.01833 * k::nao) * M::rcp(ki + R(0.01833) * nai));

    // CaMK (:430-431); the same CaMKb is used again at :939-940
    const R CaMKb = R(k::CaMKo) * (R(1) - CaMKt) * M::rcp(R(1) + R(k::KmCaM) * M::rcp(cass));
    const R CaMKa = CaMKb + CaMKt;
    const R fCaMKp = M::rcp(R(1) + R(k::KmCaMK) * M::rcp(CaMKa));  // fINaLp = fItop = fICaLp = fJrelp = fJupp
    const R vfrt = v * R(k::FRT);
    const R ndt = M::decay_scale(dt);
    // the GHK exponentials of ICaL/ICaNa/ICaK shared with INab/ICab
    const R e1 = M::ex(vfrt);
    const R e2 = e1 * e1;
    const R r1 = M::x_over_expm1(vfrt, e1);            // vfrt/(exp(vfrt)-1)
    const R r2 = WHO == 1 ? R(0) : M::x_over_expm1(R(2) * vfrt, e2);  // 2vfrt/(exp(2vfrt)-1)
    return Pre<R>{nai, ki, cansr, cajsr, CaMKt, nass, kss, cai, cass, ENa, EK, EKs, CaMKb, fCaMKp, vfrt, ndt,
                  e1, e2, r1, r2};
}

// PART: 0 the whole group; 1 INa, INaL and IKs with their gates; 2 Ito and
// INaK with Ito's gates (the currents of the other part are returned as 0).
template <class R, int PART = 0>
__device__ __forceinline__ GroupA<R> tp_group_a(CellState<R>& s, R v, int cls, const Consts<R>& c, R dt,
                                                const Pre<R>& P) {
    using M = Math<R>;
    const bool epi = (cls == CLS_EPI);
    const bool mid = (cls == CLS_M);
    const R nai = P.nai, ki = P.ki, cansr = P.cansr, cajsr = P.cajsr, CaMKt = P.CaMKt;
    const R nass = P.nass, kss = P.kss, cai = P.cai, cass = P.cass;
    const R ENa = P.ENa, EK = P.EK, EKs = P.EKs, CaMKb = P.CaMKb, fCaMKp = P.fCaMKp;
    const R vfrt = P.vfrt, ndt = P.ndt, e1 = P.e1, e2 = P.e2, r1 = P.r1, r2 = P.r2;
    (void)epi; (void)mid; (void)nai; (void)ki; (void)cansr; (void)cajsr; (void)CaMKt; (void)nass; (void)kss;
    (void)cai; (void)cass; (void)ENa; (void)EK; (void)EKs; (void)CaMKb; (void)fCaMKp; (void)vfrt; (void)ndt;
    (void)e1; (void)e2; (void)r1; (void)r2;
    R INa = R(0), INaL = R(0), Ito = R(0), IKs = R(0), INaK = R(0);

    if constexpr (PART != 2) {
        // ---- TP06 INa (:465-504)
        {
            const R AM = M::rcp(R(1) + M::exl(v, -1.0 / 5.0, -60.0 / 5.0));
            const R BM = R(0.1) * M::rcp(R(1) + M::exl(v, 1.0 / 5.0, 35.0 / 5.0)) +
                         R(0.10) * M::rcp(R(1) + M::exl(v, 1.0 / 200.0, -50.0 / 200.0));
            const R TAU_M = AM * BM;
            const R em = R(1) + M::exl(v, -1.0 / 9.03, -56.86 / 9.03);
            const R M_INF = M::rcp(em * em);
            R rate_h, rate_j;  // 1/TAU_H, 1/TAU_J
            if (v >= R(-40)) {
                rate_h = R(0.77) * M::rcp(R(0.13) * (R(1) + M::exl(v, -1.0 / 11.1, -10.66 / 11.1)));
                rate_j = R(0.6) * M::exl(v, 0.057, 0.0) * M::rcp(R(1) + M::exl(v, -0.1, -3.2));
            } else {
                rate_h = R(0.057) * M::exl(v, -1.0 / 6.8, -80.0 / 6.8) +
                         (R(2.7) * M::exl(v, 0.079, 0.0) + R(3.1e5) * M::exl(v, 0.3485, 0.0));
                const R AJ = ((R(-2.5428e4) * M::exl(v, 0.2444, 0.0) - R(6.948e-6) * M::exl(v, -0.04391, 0.0)) *
                              (v + R(37.78))) *
                             M::rcp(R(1) + M::exl(v, 0.311, 0.311 * 79.23));
                const R BJ = R(0.02424) * M::exl(v, -0.01052, 0.0) * M::rcp(R(1) + M::exl(v, -0.1378, -0.1378 * 40.14));
                rate_j = AJ + BJ;
            }
            const R eh = R(1) + M::exl(v, 1.0 / 7.43, 71.55 / 7.43);
            const R H_INF = M::rcp(eh * eh);
            s.sm = M_INF - (M_INF - s.sm) * M::decay(ndt, M::rcp(TAU_M));
            s.sh = H_INF - (H_INF - s.sh) * M::decay(ndt, rate_h);
            s.sj = H_INF - (H_INF - s.sj) * M::decay(ndt, rate_j);
        }
        INa = R(14.838) * s.sm * s.sm * s.sm * s.sh * s.sj * (v - ENa);

        // ---- INaL (:509-528)
        {
            const R mLss = M::rcp(R(1) + M::exl(v, -1.0 / 5.264, -42.85 / 5.264));
            const R tm_rate = R(6.765) * M::exl(v, 1.0 / 34.77, 11.64 / 34.77) + R(8.552) * M::exl(v, -1.0 / 5.955, -77.42 / 5.955);
            s.mL = mLss - (mLss - s.mL) * M::decay(ndt, tm_rate);
            const R hLss = M::rcp(R(1) + M::exl(v, 1.0 / 7.488, 87.61 / 7.488));
            s.hL = M::rl_slow(s.hL, hLss, dt, R(1.0 / 200.0));
            const R hLssp = M::rcp(R(1) + M::exl(v, 1.0 / 7.488, 93.81 / 7.488));
            s.hLp = M::rl_slow(s.hLp, hLssp, dt, R(1.0 / 600.0));
        }
        const R GNaL = epi ? R(0.0075 * 0.6) : R(0.0075);
        INaL = GNaL * (v - ENa) * s.mL * ((R(1) - fCaMKp) * s.hL + fCaMKp * s.hLp);
    }

    // ---- Ito (:530-571)
    if constexpr (PART != 1) {
        const R ass = M::rcp(R(1) + M::exl(v, -1.0 / 14.82, 14.34 / 14.82));
        const R ta = R(1.0515) * M::rcp(M::rcp(R(1.2089) * (R(1) + M::exl(v, -1.0 / 29.3814, 18.4099 / 29.3814))) +
                                        R(3.5) * M::rcp(R(1) + M::exl(v, 1.0 / 29.3814, 100.0 / 29.3814)));
        const R inv_ta = M::rcp(ta);
        const R e_ta = M::decay(ndt, inv_ta);
        s.a = ass - (ass - s.a) * e_ta;
        const R iss = M::rcp(R(1) + M::exl(v, 1.0 / 5.711, 43.94 / 5.711));
        R tiF = R(4.562) + M::rcp(R(0.3933) * M::exl(v, -1.0 / 100.0, -1.0) + R(0.08004) * M::exl(v, 1.0 / 16.59, 50.0 / 16.59));
        R tiS = R(23.62) + M::rcp(R(0.001416) * M::exl(v, -1.0 / 59.05, -96.52 / 59.05) +
                                  R(1.780e-8) * M::exl(v, 1.0 / 8.079, 114.1 / 8.079));
        if (epi) {
            const R delta_epi = R(1) - R(0.95) * M::rcp(R(1) + M::exl(v, 1.0 / 5.0, 70.0 / 5.0));
            tiF *= delta_epi;
            tiS *= delta_epi;
        }
        const R AiF = M::rcp(R(1) + M::exl(v, 1.0 / 151.2, -213.6 / 151.2));
        const R AiS = R(1) - AiF;
        s.iF = iss - (iss - s.iF) * M::decay(ndt, M::rcp(tiF));
        s.iS = M::rl_slow(s.iS, iss, dt, M::rcp(tiS));
        const R i_ = AiF * s.iF + AiS * s.iS;
        const R assp = M::rcp(R(1) + M::exl(v, -1.0 / 14.82, 24.34 / 14.82));
        s.ap = assp - (assp - s.ap) * e_ta;
        const R dti_develop = R(1.354) + R(1.0e-4) * M::rcp(M::exl(v, 1.0 / 15.89, -167.4 / 15.89) +
                                                             M::exl(v, -1.0 / 0.2154, 12.23 / 0.2154));
        const R dti_recover = R(1) - R(0.5) * M::rcp(R(1) + M::exl(v, 1.0 / 20.0, 70.0 / 20.0));
        const R dd = dti_develop * dti_recover;
        s.iFp = iss - (iss - s.iFp) * M::decay(ndt, M::rcp(dd * tiF));
        s.iSp = M::rl_slow(s.iSp, iss, dt, M::rcp(dd * tiS));
        const R ip = AiF * s.iFp + AiS * s.iSp;
        const R Gto = (epi || mid) ? R(0.08) : R(0.02);
        Ito = Gto * (v - EK) * ((R(1) - fCaMKp) * s.a * i_ + fCaMKp * s.ap * ip);
    }

    // ---- IKs (:680-735)
    if constexpr (PART != 2) {
        const R xs1ss = M::rcp(R(1) + M::exl(v, -1.0 / 8.932, -11.60 / 8.932));
        const R txs1 = R(817.3) + M::rcp(R(2.326e-4) * M::exl(v, 1.0 / 17.80, 48.28 / 17.80) +
                                         R(0.001292) * M::exl(v, -1.0 / 230.0, -210.0 / 230.0));
        s.xs1 = M::rl_slow(s.xs1, xs1ss, dt, M::rcp(txs1));
        const R rate_xs2 = R(0.01) * M::exl(v, 1.0 / 20.0, -50.0 / 20.0) + R(0.0193) * M::exl(v, -1.0 / 31.0, -66.54 / 31.0);
        s.xs2 = xs1ss - (xs1ss - s.xs2) * M::decay(ndt, rate_xs2);
        const R KsCa = R(1) + R(0.6) * M::rcp(R(1) + M::pow14(R(3.8e-5) * M::rcp(cai)));
        IKs = pick_class(c.gks, cls) * KsCa * s.xs1 * s.xs2 * (v - EKs);
    }

    // ---- INaK (:854-906)
    if constexpr (PART != 1) {
        const R Knai = R(9.073) * M::exl(v, -0.1550 * k::FRT / 3.0, 0.0);
        const R Knao = R(27.78) * M::exl(v, (1.0 + 0.1550) * k::FRT / 3.0, 0.0);
        const R P = R(k::eP) * M::rcp(R(1.0 + k::H / k::Khp) + nai * R(1.0 / k::Knap) + ki * R(1.0 / k::Kxkur));
        const R ni = nai * M::rcp(Knai);
        const R ni1 = R(1) + ni;
        const R kk = R(1) + ki * R(1.0 / k::Kki);
        const R deni = M::rcp(ni1 * ni1 * ni1 + kk * kk - R(1));
        const R no = R(k::nao) * M::rcp(Knao);
        const R no1 = R(1) + no;
        const R deno = M::rcp(no1 * no1 * no1 + R(k::kok2) - R(1));
        const R a1 = R(949.5) * ni * ni * ni * deni;
        const R b1 = R(k::b1);
        const R a2 = R(k::a2);
        const R b2 = R(39.4) * no * no * no * deno;
        const R a3 = R(k::a3num) * deno;
        const R b3 = R(k::b3c) * P;
        const R a4 = R(k::a4);
        const R kr = ki * R(1.0 / k::Kki);
        const R b4 = R(40.0) * kr * kr * deni;
        const R x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
        const R x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
        const R x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
        const R x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
        const R inv_sum = M::rcp(x1 + x2 + x3 + x4);
        const R JnakNa = R(3) * (x1 * a3 - x2 * b3) * inv_sum;
        const R JnakK = R(2) * (x4 * b1 - x3 * a1) * inv_sum;
        const R Pnak = epi ? R(30 * 0.9) : (mid ? R(30 * 0.7) : R(30));
        INaK = Pnak * (JnakNa + JnakK);
    }

    return GroupA<R>{INa, INaL, Ito, IKs, INaK};
}

// PART: 0 the whole group; 1 ICaL/ICaNa/ICaK and Jrel (whose steady state
// follows ICaL) with their gates; 2 IKr, IK1, INaCa, IpCa and the background
// currents with the gates of IKr and IK1.
template <class R, int PART = 0>
__device__ __forceinline__ GroupB<R> tp_group_b(CellState<R>& s, R v, int cls, const Consts<R>& c, R dt,
                                                const Pre<R>& P) {
    using M = Math<R>;
    const bool epi = (cls == CLS_EPI);
    const bool mid = (cls == CLS_M);
    const R nai = P.nai, ki = P.ki, cansr = P.cansr, cajsr = P.cajsr, CaMKt = P.CaMKt;
    const R nass = P.nass, kss = P.kss, cai = P.cai, cass = P.cass;
    const R ENa = P.ENa, EK = P.EK, EKs = P.EKs, CaMKb = P.CaMKb, fCaMKp = P.fCaMKp;
    const R vfrt = P.vfrt, ndt = P.ndt, e1 = P.e1, e2 = P.e2, r1 = P.r1, r2 = P.r2;
    (void)epi; (void)mid; (void)nai; (void)ki; (void)cansr; (void)cajsr; (void)CaMKt; (void)nass; (void)kss;
    (void)cai; (void)cass; (void)ENa; (void)EK; (void)EKs; (void)CaMKb; (void)fCaMKp; (void)vfrt; (void)ndt;
    (void)e1; (void)e2; (void)r1; (void)r2;
    R ICaL = R(0), ICaNa = R(0), ICaK = R(0), IKr = R(0), IK1 = R(0), INaCa_i = R(0), INaCa_ss = R(0);
    R IpCa = R(0), ICab = R(0), Jrel = R(0), INab = R(0), IKb = R(0);

    // ---- ICaL, ICaNa, ICaK (:573-652)
    if constexpr (PART != 2) {
        const R vs = v + c.dss_shift;
        const R dss = M::rcp(R(1) + M::exl(vs, -1.0 / 4.230, -3.940 / 4.230));
        const R td = R(0.6) + M::rcp(M::exl(v, -0.05, -0.3) + M::exl(v, 0.09, 1.26));
        s.d = dss - (dss - s.d) * M::decay(ndt, M::rcp(td));
        const R fss = M::rcp(R(1) + M::exl(v + c.fss_shift, 1.0 / 3.696, 19.58 / 3.696));
        const R tff = R(7) + M::rcp(R(0.0045) * M::exl(v, -0.1, -2.0) + R(0.0045) * M::exl(v, 0.1, 2.0));
        const R tfs = R(1000) + M::rcp(R(0.000035) * M::exl(v, -0.25, -1.25) + R(0.000035) * M::exl(v, 1.0 / 6.0, 5.0 / 6.0));
        const R Aff = R(0.6), Afs = R(0.4);
        const R inv_tff = M::rcp(tff);
        s.ff = M::rl_slow(s.ff, fss, dt, inv_tff);
        s.fs = M::rl_slow(s.fs, fss, dt, M::rcp(tfs));
        const R f = Aff * s.ff + Afs * s.fs;
        const R fcass = fss;
        const R tfcaf = R(7) + M::rcp(R(0.04) * M::exl(v, -1.0 / 7.0, 4.0 / 7.0) + R(0.04) * M::exl(v, 1.0 / 7.0, -4.0 / 7.0));
        const R tfcas = R(100) + M::rcp(R(0.00012) * M::exl(v, -1.0 / 3.0, 0.0) + R(0.00012) * M::exl(v, 1.0 / 7.0, 0.0));
        const R Afcaf = R(0.3) + R(0.6) * M::rcp(R(1) + M::exl(v, 0.1, -1.0));
        const R Afcas = R(1) - Afcaf;
        const R inv_tfcaf = M::rcp(tfcaf);
        s.fcaf = M::rl_slow(s.fcaf, fcass, dt, inv_tfcaf);
        s.fcas = M::rl_slow(s.fcas, fcass, dt, M::rcp(tfcas));
        const R fca = Afcaf * s.fcaf + Afcas * s.fcas;
        s.jca = M::rl_slow(s.jca, fcass, dt, R(1.0 / 75.0));
        s.ffp = M::rl_slow(s.ffp, fss, dt * R(1.0 / 2.5), inv_tff);
        const R fp = Aff * s.ffp + Afs * s.fs;
        s.fcafp = M::rl_slow(s.fcafp, fcass, dt * R(1.0 / 2.5), inv_tfcaf);
        const R fcap = Afcaf * s.fcafp + Afcas * s.fcas;
        // nca (:613-617): anca*k2n/km2n relaxes with rate km2n = jca
        const R km2n = s.jca;
        const R q = R(1) + R(0.002) * M::rcp(cass);
        const R q2 = q * q;
        const R anca = M::rcp(R(1000) * M::rcp(km2n) + q2 * q2);
        const R ninf = anca * R(1000) * M::rcp(km2n);
        s.nca = M::rl_slow(s.nca, ninf, dt, km2n);
        const R nca = s.nca, jca = s.jca;
        const R PhiCaL = R(2.0 * k::F) * r2 * (cass * e2 - R(0.341 * k::cao));
        const R PhiCaNa = R(k::F) * r1 * (R(0.75) * nass * e1 - R(0.75 * k::nao));
        const R PhiCaK = R(k::F) * r1 * (R(0.75) * kss * e1 - R(0.75 * k::ko));
        const R PCa = pick_class(c.pca, cls);
        const R gate_np = s.d * (f * (R(1) - nca) + jca * fca * nca);
        const R gate_p = s.d * (fp * (R(1) - nca) + jca * fcap * nca);
        // PCap = 1.1 PCa; PCaNa = 0.00125 PCa; PCaK = 3.574e-4 PCa (:638-642)
        const R mix = (R(1) - fCaMKp) * gate_np + fCaMKp * R(1.1) * gate_p;
        ICaL = PCa * PhiCaL * mix;
        ICaNa = R(0.00125) * PCa * PhiCaNa * mix;
        ICaK = R(3.574e-4) * PCa * PhiCaK * mix;
    }

    if constexpr (PART != 1) {
        // ---- IKr (:654-678)
        {
            const R xrss = M::rcp(R(1) + M::exl(v, -1.0 / 6.789, -8.337 / 6.789));
            const R txrf = R(12.98) + M::rcp(R(0.3652) * M::exl(v, 1.0 / 3.869, -31.66 / 3.869) +
                                             R(4.123e-5) * M::exl(v, -1.0 / 20.38, 47.78 / 20.38));
            const R txrs = R(1.865) + M::rcp(R(0.06629) * M::exl(v, 1.0 / 7.355, -34.70 / 7.355) +
                                             R(1.128e-5) * M::exl(v, -1.0 / 25.94, 29.74 / 25.94));
            const R Axrf = M::rcp(R(1) + M::exl(v, 1.0 / 38.21, 54.81 / 38.21));
            s.xrf = M::rl_slow(s.xrf, xrss, dt, M::rcp(txrf));
            s.xrs = M::rl_slow(s.xrs, xrss, dt, M::rcp(txrs));
            const R xr = Axrf * s.xrf + (R(1) - Axrf) * s.xrs;
            const R rkr = M::rcp((R(1) + M::exl(v, 1.0 / 75.0, 55.0 / 75.0)) * (R(1) + M::exl(v, 1.0 / 30.0, -10.0 / 30.0)));
            const R GKr = epi ? R(0.046 * 1.3) : (mid ? R(0.046 * 0.8) : R(0.046));
            IKr = GKr * xr * rkr * (v - EK);  // sqrt(ko/5.4) == 1
        }

        // ---- IK1 (:738-751)
        {
            const R xk1ss = M::rcp(R(1) + M::exl(v, -1.0 / (1.5692 * k::ko + 3.8115), -(2.5538 * k::ko + 144.59) / (1.5692 * k::ko + 3.8115)));
            const R rate_k1 = (M::exl(v, -1.0 / 20.36, -127.2 / 20.36) + M::exl(v, 1.0 / 69.33, 236.8 / 69.33)) * R(1.0 / 122.2);
            s.xk1 = xk1ss - (xk1ss - s.xk1) * M::decay(ndt, rate_k1);
            const R rk1 = M::rcp(R(1) + M::exl(v, 1.0 / 9.493, (105.8 - 2.6 * k::ko) / 9.493));
            const R GK1 = epi ? R(0.1908 * 1.2) : (mid ? R(0.1908 * 1.3) : R(0.1908));
            IK1 = GK1 * M::sqrtc(k::ko) * rk1 * s.xk1 * (v - EK);
        }

        // ---- INaCa (:753-852)
        {
            const R hca = M::exl(v, 0.1670 * k::FRT, 0.0);
            const R hna = M::exl(v, 0.5224 * k::FRT, 0.0);
            const R inv_hna = M::rcp(hna);
            const R inv_hca = M::rcp(hca);
            const R Gncx = epi ? R(0.0008 * 1.1) : (mid ? R(0.0008 * 1.4) : R(0.0008));
            const R ri = R(150.0e-6) * M::rcp(cai);
            const R allo_i = M::rcp(R(1) + ri * ri);
            INaCa_i = R(0.8) * Gncx * allo_i * ncx<R>(nai, cai, hna, inv_hna, inv_hca);
            const R rs = R(150.0e-6) * M::rcp(cass);
            const R allo_ss = M::rcp(R(1) + rs * rs);
            INaCa_ss = R(0.2) * Gncx * allo_ss * ncx<R>(nass, cass, hna, inv_hna, inv_hca);
        }

        // background (:908-923)
        const R xkb = M::rcp(R(1) + M::exl(v, -1.0 / 18.34, 14.48 / 18.34));
        const R GKb = epi ? R(0.003 * 0.6) : R(0.003);
        IKb = GKb * xkb * (v - EK);
        INab = R(3.75e-10 * k::F) * r1 * (nai * e1 - R(k::nao));
        ICab = R(2.5e-8 * 2.0 * k::F) * r2 * (cai * e2 - R(0.341 * k::cao));
        IpCa = R(0.0005) * cai * M::rcp(R(0.0005) + cai);
    }

    if constexpr (PART != 2) {
        // Jrel (:946-979), RL gates Jrelnp/Jrelp driven by ICaL
        const R rj = R(1.5) * M::rcp(cajsr);
        const R rj2 = rj * rj;
        const R rj4 = rj2 * rj2;
        const R jrel_den = M::rcp(R(1) + rj4 * rj4);
        const R mfac = mid ? R(1.7) : R(1);
        const R Jrel_inf = R(0.5 * 4.75) * (-ICaL) * jrel_den * mfac;
        const R q_rel = M::rcp(R(1) + R(0.0123) * M::rcp(cajsr));
        R tau_rel = R(4.75) * q_rel;
        tau_rel = tau_rel < R(0.005) ? R(0.005) : tau_rel;
        s.Jrelnp = Jrel_inf - (Jrel_inf - s.Jrelnp) * M::decay(ndt, M::rcp(tau_rel));
        const R Jrel_infp = R(0.5 * 1.25 * 4.75) * (-ICaL) * jrel_den * mfac;
        R tau_relp = R(1.25 * 4.75) * q_rel;
        tau_relp = tau_relp < R(0.005) ? R(0.005) : tau_relp;
        s.Jrelp = Jrel_infp - (Jrel_infp - s.Jrelp) * M::decay(ndt, M::rcp(tau_relp));
        Jrel = (R(1) - fCaMKp) * s.Jrelnp + fCaMKp * s.Jrelp;
    }
    return GroupB<R>{ICaL, ICaNa, ICaK, IKr, IK1, INaCa_i, INaCa_ss, IpCa, ICab, Jrel, INab, IKb};
}

template <class R, bool WANT_DIAG>
__device__ __forceinline__ R tp_finish(CellState<R>& s, R v, int cls, const Consts<R>& c, R dt, R Ist, R dVgap,
                                       const Pre<R>& P, const GroupA<R>& GA, const GroupB<R>& GB, Diag* diag) {
    using M = Math<R>;
    const bool epi = (cls == CLS_EPI);
    const bool mid = (cls == CLS_M);
    const R nai = P.nai, ki = P.ki, cansr = P.cansr, cajsr = P.cajsr, CaMKt = P.CaMKt;
    const R nass = P.nass, kss = P.kss, cai = P.cai, cass = P.cass;
    const R ENa = P.ENa, EK = P.EK, EKs = P.EKs, CaMKb = P.CaMKb, fCaMKp = P.fCaMKp;
    const R vfrt = P.vfrt, ndt = P.ndt, e1 = P.e1, e2 = P.e2, r1 = P.r1, r2 = P.r2;
    (void)epi; (void)mid; (void)nai; (void)ki; (void)cansr; (void)cajsr; (void)CaMKt; (void)nass; (void)kss;
    (void)cai; (void)cass; (void)ENa; (void)EK; (void)EKs; (void)CaMKb; (void)fCaMKp; (void)vfrt; (void)ndt;
    (void)e1; (void)e2; (void)r1; (void)r2;
    const R INa = GA.INa, INaL = GA.INaL, Ito = GA.Ito, IKs = GA.IKs, INaK = GA.INaK;
    const R ICaL = GB.ICaL, ICaNa = GB.ICaNa, ICaK = GB.ICaK, IKr = GB.IKr, IK1 = GB.IK1, INaCa_i = GB.INaCa_i,
            INaCa_ss = GB.INaCa_ss, IpCa = GB.IpCa, ICab = GB.ICab, Jrel = GB.Jrel, INab = GB.INab, IKb = GB.IKb;

    // ---- voltage (:927-935)
    const R INaCa = INaCa_i + INaCa_ss;
    const R Iion = INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa + INaK + INab + IKb + IpCa + ICab;
    R dV_dt = -(Iion + Ist);
    R vnew;
    if (c.clamp) {
        dV_dt = R(0);
        vnew = v;
    } else {
        vnew = v + (dV_dt + dVgap) * dt;
    }
    if (WANT_DIAG) {
        diag->Iion = Iion;
        diag->dV_dt = dV_dt;
        diag->ICaL = ICaL;
        diag->IKs = IKs;
    }

    // ---- CaMK, fluxes, buffers, concentrations (:939-1020)
    const R dCaMKt = R(k::aCaMK) * CaMKb * (CaMKb + CaMKt) - R(k::bCaMK) * CaMKt;
    const R JdiffNa = (nass - nai) * R(0.5);
    const R JdiffK = (kss - ki) * R(0.5);
    const R Jdiff = (cass - cai) * R(5.0);


    const R upf = epi ? R(1.3) : R(1);
    const R Jupnp = upf * R(0.004375) * cai * M::rcp(cai + R(0.00092));
    const R Jupp = upf * R(2.75 * 0.004375) * cai * M::rcp(cai + R(0.00092 - 0.00017));
    const R Jleak = R(0.0039375 / 15.0) * cansr;
    const R Jup = (R(1) - fCaMKp) * Jupnp + fCaMKp * Jupp - Jleak;
    const R Jtr = (cansr - cajsr) * R(0.01);

    const R AF_myo = R(k::Acap / (k::F * k::vmyo));
    const R AF_ss = R(k::Acap / (k::F * k::vss));
    const R dnai = -(INa + INaL + R(3) * INaCa_i + R(3) * INaK + INab) * AF_myo + JdiffNa * R(k::vss / k::vmyo);
    const R dnass = -(ICaNa + R(3) * INaCa_ss) * AF_ss - JdiffNa;
    const R dki = -(Ito + IKr + IKs + IK1 + IKb + Ist - R(2) * INaK) * AF_myo + JdiffK * R(k::vss / k::vmyo);
    const R dkss = -ICaK * AF_ss - JdiffK;
    const R cmf = epi ? R(1.3 * k::cmdnmax * k::kmcmdn) : R(k::cmdnmax * k::kmcmdn);
    const R ci1 = M::rcp(R(k::kmcmdn) + cai);
    const R ci2 = M::rcp(R(k::kmtrpn) + cai);
    const R Bcai = M::rcp(R(1) + cmf * ci1 * ci1 + R(k::trpnmax * k::kmtrpn) * ci2 * ci2);
    const R dcai = Bcai * (-(IpCa + ICab - R(2) * INaCa_i) * R(k::Acap / (2.0 * k::F * k::vmyo)) -
                           Jup * R(k::vnsr / k::vmyo) + Jdiff * R(k::vss / k::vmyo));
    const R cs1 = M::rcp(R(k::KmBSR) + cass);
    const R cs2 = M::rcp(R(k::KmBSL) + cass);
    const R Bcass = M::rcp(R(1) + R(k::BSRmax * k::KmBSR) * cs1 * cs1 + R(k::BSLmax * k::KmBSL) * cs2 * cs2);
    const R dcass = Bcass * (-(ICaL - R(2) * INaCa_ss) * R(k::Acap / (2.0 * k::F * k::vss)) +
                             Jrel * R(k::vjsr / k::vss) - Jdiff);
    const R dcansr = Jup - Jtr * R(k::vjsr / k::vnsr);
    const R cj = M::rcp(R(k::kmcsqn) + cajsr);
    const R Bcajsr = M::rcp(R(1) + R(k::csqnmax * k::kmcsqn) * cj * cj);
    const R dcajsr = Bcajsr * (Jtr - Jrel);

    s.CaMKt += double(dt * dCaMKt);
    s.nai += double(dt * dnai);
    s.nass = nass + dt * dnass;
    s.ki += double(dt * dki);
    s.kss = kss + dt * dkss;
    s.cai = cai + dt * dcai;
    s.cass = cass + dt * dcass;
    s.cansr += double(dt * dcansr);
    s.cajsr += double(dt * dcajsr);
    return vnew;
}

// One dt of TPORd::update() for one cell.  v is V^n; returns V^{n+1}.
template <class R, bool WANT_DIAG>
__device__ __forceinline__ R tpord_update(CellState<R>& s, R v, int cls, const Consts<R>& c, R dt,
                                          R Ist, R dVgap, Diag* diag) {
    const Pre<R> P = tp_pre(s, v, dt);
    const GroupA<R> GA = tp_group_a(s, v, cls, c, dt, P);
    const GroupB<R> GB = tp_group_b(s, v, cls, c, dt, P);
    return tp_finish<R, WANT_DIAG>(s, v, cls, c, dt, Ist, dVgap, P, GA, GB, diag);
}

}  // namespace lqt
