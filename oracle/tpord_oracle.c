/*
 * FP64 CPU oracle of the TPORd cell model and its drivers.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  Each function follows the reference file:line named beside it,
 * in the same order of floating-point operations, so that with
 * `-O2 -ffp-contract=off` it agrees bit for bit with the reference's own
 * CPU build on x86-64 (checked by tests/test_oracle_vs_ref.py).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void or_params_default(or_params *p)
{
    /* TPORd.cc:30-36 (CON + ISO1), :45 (sca = 1), :629 (PCa x1.6) */
    p->iks_scale = 1.0;
    p->iso = 1;
    p->pca_iso = 1.6;
    p->suprep = 0;
    p->dose_fmol = 150.0;
    p->sca = 1.0;
    p->clamp = 0;
}

/* TPORd.cc:142-198 */
void or_cell_init(double *s, double v0)
{
    for (int i = 0; i < OR_NSTATE; ++i) s[i] = 0.0;
    s[OR_V] = v0;
    s[OR_NAI] = 7;      s[OR_NASS] = 7;
    s[OR_KI] = 145;     s[OR_KSS] = 145;
    s[OR_CAI] = 1.0e-4; s[OR_CASS] = 1.0e-4;
    s[OR_CANSR] = 1.2;  s[OR_CAJSR] = 1.2;
    s[OR_SM] = 0; s[OR_SH] = 1; s[OR_SJ] = 1;
    s[OR_HSP] = 1; s[OR_JP] = 1;
    s[OR_ML] = 0; s[OR_HL] = 1; s[OR_HLP] = 1;
    s[OR_A] = 0; s[OR_IF] = 1; s[OR_IS] = 1; s[OR_AP] = 0; s[OR_IFP] = 1; s[OR_ISP] = 1;
    s[OR_D] = 0; s[OR_FF] = 1; s[OR_FS] = 1; s[OR_FCAF] = 1; s[OR_FCAS] = 1;
    s[OR_JCA] = 1; s[OR_NCA] = 0; s[OR_FFP] = 1; s[OR_FCAFP] = 1;
    s[OR_XRF] = 0; s[OR_XRS] = 0; s[OR_XS1] = 0; s[OR_XS2] = 0; s[OR_XK1] = 1;
    s[OR_JRELNP] = 0; s[OR_JRELP] = 0; s[OR_CAMKT] = 0;
}

/* Rush-Larsen step x_inf - (x_inf - x) exp(-dt/tau), as written throughout
 * TPORd.cc:465-981. */
static inline double rl(double xinf, double x, double dt, double tau)
{
    return xinf - (xinf - x) * exp(-dt / tau);
}

/* NCX of TPORd.cc:753-852 for one compartment (na = nai or nass, ca = cai
 * or cass). */
static double ncx_flux(double na, double ca, double hna, double hca, double nao,
                       double cao, double *out_E)
{
    const double kna1 = 15.0, kna2 = 5.0, kna3 = 88.12, kasymm = 12.5;
    const double wna = 6.0e4, wca = 6.0e4, wnaca = 5.0e3;
    const double kcaon = 1.5e6, kcaoff = 5.0e3;
    double h1 = 1 + na / kna3 * (1 + hna);
    double h2 = (na * hna) / (kna3 * h1);
    double h3 = 1.0 / h1;
    double h4 = 1.0 + na / kna1 * (1 + na / kna2);
    double h5 = na * na / (h4 * kna1 * kna2);
    double h6 = 1.0 / h4;
    double h7 = 1.0 + nao / kna3 * (1.0 + 1.0 / hna);
    double h8 = nao / (kna3 * hna * h7);
    double h9 = 1.0 / h7;
    double h10 = kasymm + 1.0 + nao / kna1 * (1.0 + nao / kna2);
    double h11 = nao * nao / (h10 * kna1 * kna2);
    double h12 = 1.0 / h10;
    double k1 = h12 * cao * kcaon;
    double k2 = kcaoff;
    double k3p = h9 * wca;
    double k3pp = h8 * wnaca;
    double k3 = k3p + k3pp;
    double k4p = h3 * wca / hca;
    double k4pp = h2 * wnaca;
    double k4 = k4p + k4pp;
    double k5 = kcaoff;
    double k6 = h6 * ca * kcaon;
    double k7 = h5 * h2 * wna;
    double k8 = h8 * h11 * wna;
    double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
    double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
    double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
    double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
    double E1 = x1 / (x1 + x2 + x3 + x4);
    double E2 = x2 / (x1 + x2 + x3 + x4);
    double E3 = x3 / (x1 + x2 + x3 + x4);
    double E4 = x4 / (x1 + x2 + x3 + x4);
    out_E[0] = 3.0 * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp; /* JncxNa */
    out_E[1] = E2 * k2 - E1 * k1;                                 /* JncxCa */
    return 0.0;
}

/* TPORd.cc:379-1021, one dt. */
void or_cell_update(double *s, int celltype, const or_params *p, double dt,
                    double Ist, double dVgap_dt, or_diag *diag)
{
    const int epi = (celltype == OR_EPI);
    const int mid = (celltype == OR_MCELL);

    double v = s[OR_V];
    double nai = s[OR_NAI], nass = s[OR_NASS], ki = s[OR_KI], kss = s[OR_KSS];
    double cai = s[OR_CAI], cass = s[OR_CASS], cansr = s[OR_CANSR], cajsr = s[OR_CAJSR];
    double CaMKt = s[OR_CAMKT];

    /* constants, TPORd.cc:381-421 */
    const double nao = 140.0, cao = 1.8, ko = 5.4;
    const double BSRmax = 0.047, KmBSR = 0.00087, BSLmax = 1.124, KmBSL = 0.0087;
    const double cmdnmax = 0.05, kmcmdn = 0.00238, trpnmax = 0.07, kmtrpn = 0.0005;
    const double csqnmax = 10.0, kmcsqn = 0.8;
    const double aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015, KmCaMK = 0.15;
    const double R = 8314.0, T = 310.0, F = 96485.0;
    const double L = 0.01, rad = 0.0011;
    const double vcell = 1000 * 3.14 * rad * rad * L;
    const double Ageo = 2 * 3.14 * rad * rad + 2 * 3.14 * rad * L;
    const double Acap = 2 * Ageo;
    const double vmyo = 0.68 * vcell;
    const double vnsr = 0.0552 * vcell;
    const double vjsr = 0.0048 * vcell;
    const double vss = 0.02 * vcell;

    /* reversal potentials, :425-427 */
    double ENa = (R * T / F) * log(nao / nai);
    double EK = (R * T / F) * log(ko / ki);
    double EKs = (R * T / F) * log((ko + 0.01833 * nao) / (ki + 0.01833 * nai));

    double CaMKb = CaMKo * (1.0 - CaMKt) / (1.0 + KmCaM / cass);
    double CaMKa = CaMKb + CaMKt;
    double vffrt = v * F * F / (R * T);
    double vfrt = v * F / (R * T);

    /* TP06 INa, :465-504 */
    double AM = 1. / (1. + exp((-60. - v) / 5.));
    double BM = 0.1 / (1. + exp((v + 35.) / 5.)) + 0.10 / (1. + exp((v - 50.) / 200.));
    double TAU_M = AM * BM;
    double M_INF = 1. / ((1. + exp((-56.86 - v) / 9.03)) * (1. + exp((-56.86 - v) / 9.03)));
    double TAU_H, TAU_J;
    if (v >= -40.) {
        double AH_1 = 0.;
        double BH_1 = (0.77 / (0.13 * (1. + exp(-(v + 10.66) / 11.1))));
        TAU_H = 1.0 / (AH_1 + BH_1);
    } else {
        double AH_2 = (0.057 * exp(-(v + 80.) / 6.8));
        double BH_2 = (2.7 * exp(0.079 * v) + (3.1e5) * exp(0.3485 * v));
        TAU_H = 1.0 / (AH_2 + BH_2);
    }
    double H_INF = 1. / ((1. + exp((v + 71.55) / 7.43)) * (1. + exp((v + 71.55) / 7.43)));
    if (v >= -40.) {
        double AJ_1 = 0.;
        double BJ_1 = (0.6 * exp((0.057) * v) / (1. + exp(-0.1 * (v + 32.))));
        TAU_J = 1.0 / (AJ_1 + BJ_1);
    } else {
        double AJ_2 = (((-2.5428e4) * exp(0.2444 * v) - (6.948e-6) * exp(-0.04391 * v)) * (v + 37.78) /
                       (1. + exp(0.311 * (v + 79.23))));
        double BJ_2 = (0.02424 * exp(-0.01052 * v) / (1. + exp(-0.1378 * (v + 40.14))));
        TAU_J = 1.0 / (AJ_2 + BJ_2);
    }
    double J_INF = H_INF;
    s[OR_SM] = M_INF - (M_INF - s[OR_SM]) * exp(-dt / TAU_M);
    s[OR_SH] = H_INF - (H_INF - s[OR_SH]) * exp(-dt / TAU_H);
    s[OR_SJ] = J_INF - (J_INF - s[OR_SJ]) * exp(-dt / TAU_J);
    double sm = s[OR_SM], sh = s[OR_SH], sj = s[OR_SJ];
    double GNa = 14.838;
    double INa = GNa * sm * sm * sm * sh * sj * (v - ENa);

    /* INaL, :509-528 */
    double mLss = 1.0 / (1.0 + exp((-(v + 42.85)) / 5.264));
    double tm = 1.0 / (6.765 * exp((v + 11.64) / 34.77) + 8.552 * exp(-(v + 77.42) / 5.955));
    double tmL = tm;
    s[OR_ML] = rl(mLss, s[OR_ML], dt, tmL);
    double hLss = 1.0 / (1.0 + exp((v + 87.61) / 7.488));
    double thL = 200.0;
    s[OR_HL] = rl(hLss, s[OR_HL], dt, thL);
    double hLssp = 1.0 / (1.0 + exp((v + 93.81) / 7.488));
    double thLp = 3.0 * thL;
    s[OR_HLP] = rl(hLssp, s[OR_HLP], dt, thLp);
    double GNaL = 0.0075;
    if (epi) GNaL *= 0.6;
    double fINaLp = (1.0 / (1.0 + KmCaMK / CaMKa));
    double INaL = GNaL * (v - ENa) * s[OR_ML] * ((1.0 - fINaLp) * s[OR_HL] + fINaLp * s[OR_HLP]);

    /* Ito, :530-571 */
    double ass = 1.0 / (1.0 + exp((-(v - 14.34)) / 14.82));
    double ta = 1.0515 / (1.0 / (1.2089 * (1.0 + exp(-(v - 18.4099) / 29.3814))) +
                          3.5 / (1.0 + exp((v + 100.0) / 29.3814)));
    s[OR_A] = rl(ass, s[OR_A], dt, ta);
    double iss = 1.0 / (1.0 + exp((v + 43.94) / 5.711));
    double delta_epi = epi ? 1.0 - (0.95 / (1.0 + exp((v + 70.0) / 5.0))) : 1.0;
    double tiF = 4.562 + 1 / (0.3933 * exp((-(v + 100.0)) / 100.0) + 0.08004 * exp((v + 50.0) / 16.59));
    double tiS = 23.62 + 1 / (0.001416 * exp((-(v + 96.52)) / 59.05) + 1.780e-8 * exp((v + 114.1) / 8.079));
    tiF *= delta_epi;
    tiS *= delta_epi;
    double AiF = 1.0 / (1.0 + exp((v - 213.6) / 151.2));
    double AiS = 1.0 - AiF;
    s[OR_IF] = rl(iss, s[OR_IF], dt, tiF);
    s[OR_IS] = rl(iss, s[OR_IS], dt, tiS);
    double i_ = AiF * s[OR_IF] + AiS * s[OR_IS];
    double assp = 1.0 / (1.0 + exp((-(v - 24.34)) / 14.82));
    s[OR_AP] = rl(assp, s[OR_AP], dt, ta);
    double dti_develop = 1.354 + 1.0e-4 / (exp((v - 167.4) / 15.89) + exp(-(v - 12.23) / 0.2154));
    double dti_recover = 1.0 - 0.5 / (1.0 + exp((v + 70.0) / 20.0));
    double tiFp = dti_develop * dti_recover * tiF;
    double tiSp = dti_develop * dti_recover * tiS;
    s[OR_IFP] = rl(iss, s[OR_IFP], dt, tiFp);
    s[OR_ISP] = rl(iss, s[OR_ISP], dt, tiSp);
    double ip = AiF * s[OR_IFP] + AiS * s[OR_ISP];
    double Gto = 0.02;
    if (epi) Gto *= 4.0;
    if (mid) Gto *= 4.0;
    double fItop = (1.0 / (1.0 + KmCaMK / CaMKa));
    double Ito = Gto * (v - EK) * ((1.0 - fItop) * s[OR_A] * i_ + fItop * s[OR_AP] * ip);

    /* ICaL, ICaNa, ICaK, :573-652 */
    double dss = 1.0 / (1.0 + exp((-(v + 3.940)) / 4.230));
    if (p->iso) dss = 1.0 / (1.0 + exp((-(v + 3.940 + 12)) / 4.230));
    double td = 0.6 + 1.0 / (exp(-0.05 * (v + 6.0)) + exp(0.09 * (v + 14.0)));
    s[OR_D] = rl(dss, s[OR_D], dt, td);
    double fss = 1.0 / (1.0 + exp((v + 19.58) / 3.696));
    if (p->iso) fss = 1.0 / (1.0 + exp((v + 19.58 + 8.0) / 3.696));
    double tff = 7.0 + 1.0 / (0.0045 * exp(-(v + 20.0) / 10.0) + 0.0045 * exp((v + 20.0) / 10.0));
    double tfs = 1000.0 + 1.0 / (0.000035 * exp(-(v + 5.0) / 4.0) + 0.000035 * exp((v + 5.0) / 6.0));
    double Aff = 0.6;
    double Afs = 1.0 - Aff;
    s[OR_FF] = rl(fss, s[OR_FF], dt, tff);
    s[OR_FS] = rl(fss, s[OR_FS], dt, tfs);
    double f = Aff * s[OR_FF] + Afs * s[OR_FS];
    double fcass = fss;
    double tfcaf = 7.0 + 1.0 / (0.04 * exp(-(v - 4.0) / 7.0) + 0.04 * exp((v - 4.0) / 7.0));
    double tfcas = 100.0 + 1.0 / (0.00012 * exp(-v / 3.0) + 0.00012 * exp(v / 7.0));
    double Afcaf = 0.3 + 0.6 / (1.0 + exp((v - 10.0) / 10.0));
    double Afcas = 1.0 - Afcaf;
    s[OR_FCAF] = rl(fcass, s[OR_FCAF], dt, tfcaf);
    s[OR_FCAS] = rl(fcass, s[OR_FCAS], dt, tfcas);
    double fca = Afcaf * s[OR_FCAF] + Afcas * s[OR_FCAS];
    double tjca = 75.0;
    s[OR_JCA] = rl(fcass, s[OR_JCA], dt, tjca);
    double jca = s[OR_JCA];
    double tffp = 2.5 * tff;
    s[OR_FFP] = rl(fss, s[OR_FFP], dt, tffp);
    double fp = Aff * s[OR_FFP] + Afs * s[OR_FS];
    double tfcafp = 2.5 * tfcaf;
    s[OR_FCAFP] = rl(fcass, s[OR_FCAFP], dt, tfcafp);
    double fcap = Afcaf * s[OR_FCAFP] + Afcas * s[OR_FCAS];
    double Kmn = 0.002;
    double k2n = 1000.0;
    double km2n = jca * 1.0;
    double anca = 1.0 / (k2n / km2n + pow(1.0 + Kmn / cass, 4.0));
    s[OR_NCA] = anca * k2n / km2n - (anca * k2n / km2n - s[OR_NCA]) * exp(-km2n * dt);
    double nca = s[OR_NCA];
    double PhiCaL = 4.0 * vffrt * (cass * exp(2.0 * vfrt) - 0.341 * cao) / (exp(2.0 * vfrt) - 1.0);
    double PhiCaNa = 1.0 * vffrt * (0.75 * nass * exp(1.0 * vfrt) - 0.75 * nao) / (exp(1.0 * vfrt) - 1.0);
    double PhiCaK = 1.0 * vffrt * (0.75 * kss * exp(1.0 * vfrt) - 0.75 * ko) / (exp(1.0 * vfrt) - 1.0);
    double zca = 2.0;
    double PCa = 0.0001;
    if (p->iso) PCa *= p->pca_iso;
    if (epi) PCa *= 1.2;
    if (mid) PCa *= 2.5;
    double PCap = 1.1 * PCa;
    double PCaNa = 0.00125 * PCa;
    double PCaK = 3.574e-4 * PCa;
    double PCaNap = 0.00125 * PCap;
    double PCaKp = 3.574e-4 * PCap;
    double fICaLp = (1.0 / (1.0 + KmCaMK / CaMKa));
    double d = s[OR_D];
    double ICaL = (1.0 - fICaLp) * PCa * PhiCaL * d * (f * (1.0 - nca) + jca * fca * nca) +
                  fICaLp * PCap * PhiCaL * d * (fp * (1.0 - nca) + jca * fcap * nca);
    double ICaNa = (1.0 - fICaLp) * PCaNa * PhiCaNa * d * (f * (1.0 - nca) + jca * fca * nca) +
                   fICaLp * PCaNap * PhiCaNa * d * (fp * (1.0 - nca) + jca * fcap * nca);
    double ICaK = (1.0 - fICaLp) * PCaK * PhiCaK * d * (f * (1.0 - nca) + jca * fca * nca) +
                  fICaLp * PCaKp * PhiCaK * d * (fp * (1.0 - nca) + jca * fcap * nca);

    /* IKr, :654-678 */
    double xrss = 1.0 / (1.0 + exp((-(v + 8.337)) / 6.789));
    double txrf = 12.98 + 1.0 / (0.3652 * exp((v - 31.66) / 3.869) + 4.123e-5 * exp((-(v - 47.78)) / 20.38));
    double txrs = 1.865 + 1.0 / (0.06629 * exp((v - 34.70) / 7.355) + 1.128e-5 * exp((-(v - 29.74)) / 25.94));
    double Axrf = 1.0 / (1.0 + exp((v + 54.81) / 38.21));
    double Axrs = 1.0 - Axrf;
    s[OR_XRF] = rl(xrss, s[OR_XRF], dt, txrf);
    s[OR_XRS] = rl(xrss, s[OR_XRS], dt, txrs);
    double xr = Axrf * s[OR_XRF] + Axrs * s[OR_XRS];
    double rkr = 1.0 / (1.0 + exp((v + 55.0) / 75.0)) * 1.0 / (1.0 + exp((v - 10.0) / 30.0));
    double GKr = 0.046;
    if (epi) GKr *= 1.3;
    if (mid) GKr *= 0.8;
    double IKr = GKr * sqrt(ko / 5.4) * xr * rkr * (v - EK);

    /* IKs, :680-735 */
    double xs1ss = 1.0 / (1.0 + exp((-(v + 11.60)) / 8.932));
    double txs1 = 817.3 + 1.0 / (2.326e-4 * exp((v + 48.28) / 17.80) + 0.001292 * exp((-(v + 210.0)) / 230.0));
    s[OR_XS1] = rl(xs1ss, s[OR_XS1], dt, txs1);
    double xs2ss = xs1ss;
    double txs2 = 1.0 / (0.01 * exp((v - 50.0) / 20.0) + 0.0193 * exp((-(v + 66.54)) / 31.0));
    s[OR_XS2] = rl(xs2ss, s[OR_XS2], dt, txs2);
    double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-5 / cai, 1.4));
    double GKs = 0.0034 * p->sca;
    if (epi) GKs *= 1.4;
    if (p->iks_scale != 1.0) GKs = p->iks_scale * GKs;
    if (p->suprep) {
        double GKs_sup = GKs;
        double GKs_rep = epi ? 1.4 * 0.0034 : 0.0034;
        double dose = p->dose_fmol;
        double frac_wt = 1.0;
        if (dose >= 25)
            frac_wt = 3.129 * pow(dose, -0.560);
        else
            frac_wt = pow(0.974, dose);
        double frac_shIMM = 2.018 * (1.0 - exp(-2.0 * dose * 0.0059)) / (1.0 + exp(-2.0 * dose * 0.0059));
        GKs = frac_wt * GKs_sup + frac_shIMM * GKs_rep;
    }
    if (p->iso) GKs *= 3.2;
    double IKs = GKs * KsCa * s[OR_XS1] * s[OR_XS2] * (v - EKs);

    /* IK1, :738-751 */
    double xk1ss = 1.0 / (1.0 + exp(-(v + 2.5538 * ko + 144.59) / (1.5692 * ko + 3.8115)));
    double txk1 = 122.2 / (exp((-(v + 127.2)) / 20.36) + exp((v + 236.8) / 69.33));
    s[OR_XK1] = rl(xk1ss, s[OR_XK1], dt, txk1);
    double rk1 = 1.0 / (1.0 + exp((v + 105.8 - 2.6 * ko) / 9.493));
    double GK1 = 0.1908;
    if (epi) GK1 *= 1.2;
    if (mid) GK1 *= 1.3;
    double IK1 = GK1 * sqrt(ko) * rk1 * s[OR_XK1] * (v - EK);

    /* INaCa, :753-852 */
    const double qna = 0.5224, qca = 0.1670;
    double hca = exp((qca * v * F) / (R * T));
    double hna = exp((qna * v * F) / (R * T));
    double J[2];
    ncx_flux(nai, cai, hna, hca, nao, cao, J);
    double KmCaAct = 150.0e-6;
    double allo = 1.0 / (1.0 + pow(KmCaAct / cai, 2.0));
    double zna = 1.0;
    double Gncx = 0.0008;
    if (epi) Gncx *= 1.1;
    if (mid) Gncx *= 1.4;
    double INaCa_i = 0.8 * Gncx * allo * (zna * J[0] + zca * J[1]);
    ncx_flux(nass, cass, hna, hca, nao, cao, J);
    allo = 1.0 / (1.0 + pow(KmCaAct / cass, 2.0));
    double INaCa_ss = 0.2 * Gncx * allo * (zna * J[0] + zca * J[1]);
    double INaCa = INaCa_i + INaCa_ss;

    /* INaK, :854-906 */
    double k1p = 949.5, k1m = 182.4, k2p = 687.2, k2m = 39.4, k3p = 1899.0, k3m = 79300.0;
    double k4p = 639.0, k4m = 40.0;
    double Knai0 = 9.073, Knao0 = 27.78, delta = -0.1550;
    double Knai = Knai0 * exp((delta * v * F) / (3.0 * R * T));
    double Knao = Knao0 * exp(((1.0 - delta) * v * F) / (3.0 * R * T));
    double Kki = 0.5, Kko = 0.3582, MgADP = 0.05, MgATP = 9.8, Kmgatp = 1.698e-7;
    double H = 1.0e-7, eP = 4.2, Khp = 1.698e-7, Knap = 224.0, Kxkur = 292.0;
    double P = eP / (1.0 + H / Khp + nai / Knap + ki / Kxkur);
    double a1 = (k1p * pow(nai / Knai, 3.0)) / (pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / Kki, 2.0) - 1.0);
    double b1 = k1m * MgADP;
    double a2 = k2p;
    double b2 = (k2m * pow(nao / Knao, 3.0)) / (pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / Kko, 2.0) - 1.0);
    double a3 = (k3p * pow(ko / Kko, 2.0)) / (pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / Kko, 2.0) - 1.0);
    double b3 = (k3m * P * H) / (1.0 + MgATP / Kmgatp);
    double a4 = (k4p * MgATP / Kmgatp) / (1.0 + MgATP / Kmgatp);
    double b4 = (k4m * pow(ki / Kki, 2.0)) / (pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / Kki, 2.0) - 1.0);
    double x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    double x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    double x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    double x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    double E1 = x1 / (x1 + x2 + x3 + x4);
    double E2 = x2 / (x1 + x2 + x3 + x4);
    double E3 = x3 / (x1 + x2 + x3 + x4);
    double E4 = x4 / (x1 + x2 + x3 + x4);
    double zk = 1.0;
    double JnakNa = 3.0 * (E1 * a3 - E2 * b3);
    double JnakK = 2.0 * (E4 * b1 - E3 * a1);
    double Pnak = 30;
    if (epi) Pnak *= 0.9;
    if (mid) Pnak *= 0.7;
    double INaK = Pnak * (zna * JnakNa + zk * JnakK);

    /* background currents, :908-923 */
    double xkb = 1.0 / (1.0 + exp(-(v - 14.48) / 18.34));
    double GKb = 0.003;
    if (epi) GKb *= 0.6;
    double IKb = GKb * xkb * (v - EK);
    double PNab = 3.75e-10;
    double INab = PNab * vffrt * (nai * exp(vfrt) - nao) / (exp(vfrt) - 1.0);
    double PCab = 2.5e-8;
    double ICab = PCab * 4.0 * vffrt * (cai * exp(2.0 * vfrt) - 0.341 * cao) / (exp(2.0 * vfrt) - 1.0);
    double GpCa = 0.0005;
    double IpCa = GpCa * cai / (0.0005 + cai);

    /* voltage, :927-935 */
    double Iion = INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa + INaK + INab + IKb + IpCa + ICab;
    double dV_dt = -1.0 * (Iion + Ist);
    if (p->clamp) {
        dV_dt = 0;
        dVgap_dt = 0;
    }
    v += (dV_dt + dVgap_dt) * dt;

    /* CaMK, fluxes, buffers and concentrations, :939-1020 */
    CaMKb = CaMKo * (1.0 - CaMKt) / (1.0 + KmCaM / cass);
    CaMKa = CaMKb + CaMKt;
    CaMKt += dt * (aCaMK * CaMKb * (CaMKb + CaMKt) - bCaMK * CaMKt);

    double JdiffNa = (nass - nai) / 2.0;
    double JdiffK = (kss - ki) / 2.0;
    double Jdiff = (cass - cai) / 0.2;

    double bt = 4.75;
    double a_rel = 0.5 * bt;
    double Jrel_inf = a_rel * (-ICaL) / (1.0 + pow(1.5 / cajsr, 8.0));
    if (mid) Jrel_inf *= 1.7;
    double tau_rel = bt / (1.0 + 0.0123 / cajsr);
    if (tau_rel < 0.005) tau_rel = 0.005;
    s[OR_JRELNP] = rl(Jrel_inf, s[OR_JRELNP], dt, tau_rel);
    double btp = 1.25 * bt;
    double a_relp = 0.5 * btp;
    double Jrel_infp = a_relp * (-ICaL) / (1.0 + pow(1.5 / cajsr, 8.0));
    if (mid) Jrel_infp *= 1.7;
    double tau_relp = btp / (1.0 + 0.0123 / cajsr);
    if (tau_relp < 0.005) tau_relp = 0.005;
    s[OR_JRELP] = rl(Jrel_infp, s[OR_JRELP], dt, tau_relp);
    double fJrelp = (1.0 / (1.0 + KmCaMK / CaMKa));
    double Jrel = (1.0 - fJrelp) * s[OR_JRELNP] + fJrelp * s[OR_JRELP];

    double Jupnp = 0.004375 * cai / (cai + 0.00092);
    double Jupp = 2.75 * 0.004375 * cai / (cai + 0.00092 - 0.00017);
    if (epi) {
        Jupnp *= 1.3;
        Jupp *= 1.3;
    }
    double fJupp = (1.0 / (1.0 + KmCaMK / CaMKa));
    double Jleak = 0.0039375 * cansr / 15.0;
    double Jup = (1.0 - fJupp) * Jupnp + fJupp * Jupp - Jleak;
    double Jtr = (cansr - cajsr) / 100.0;

    nai += dt * (-(INa + INaL + 3.0 * INaCa_i + 3.0 * INaK + INab) * Acap / (F * vmyo) + JdiffNa * vss / vmyo);
    nass += dt * (-(ICaNa + 3.0 * INaCa_ss) * Acap / (F * vss) - JdiffNa);
    ki += dt * (-(Ito + IKr + IKs + IK1 + IKb + Ist - 2.0 * INaK) * Acap / (F * vmyo) + JdiffK * vss / vmyo);
    kss += dt * (-(ICaK) * Acap / (F * vss) - JdiffK);
    double Bcai;
    if (epi)
        Bcai = 1.0 / (1.0 + 1.3 * cmdnmax * kmcmdn / pow(kmcmdn + cai, 2.0) + trpnmax * kmtrpn / pow(kmtrpn + cai, 2.0));
    else
        Bcai = 1.0 / (1.0 + cmdnmax * kmcmdn / pow(kmcmdn + cai, 2.0) + trpnmax * kmtrpn / pow(kmtrpn + cai, 2.0));
    cai += dt * (Bcai * (-(IpCa + ICab - 2.0 * INaCa_i) * Acap / (2.0 * F * vmyo) - Jup * vnsr / vmyo + Jdiff * vss / vmyo));
    double Bcass = 1.0 / (1.0 + BSRmax * KmBSR / pow(KmBSR + cass, 2.0) + BSLmax * KmBSL / pow(KmBSL + cass, 2.0));
    cass += dt * (Bcass * (-(ICaL - 2.0 * INaCa_ss) * Acap / (2.0 * F * vss) + Jrel * vjsr / vss - Jdiff));
    cansr += dt * (Jup - Jtr * vjsr / vnsr);
    double Bcajsr = 1.0 / (1.0 + csqnmax * kmcsqn / pow(kmcsqn + cajsr, 2.0));
    cajsr += dt * (Bcajsr * (Jtr - Jrel));

    s[OR_V] = v;
    s[OR_NAI] = nai; s[OR_NASS] = nass; s[OR_KI] = ki; s[OR_KSS] = kss;
    s[OR_CAI] = cai; s[OR_CASS] = cass; s[OR_CANSR] = cansr; s[OR_CAJSR] = cajsr;
    s[OR_CAMKT] = CaMKt;

    if (diag) {
        diag->Iion = Iion;
        diag->dV_dt = dV_dt;
        diag->ICaL = ICaL;
        diag->IKs = IKs;
        diag->IKr = IKr;
        diag->IK1 = IK1;
        diag->INa = INa;
    }
}

void or_cells_update(int n, double *s, const uint8_t *celltype, const or_params *p,
                     double dt, const double *ist, const double *dvgap)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        or_cell_update(s + (size_t)i * OR_NSTATE, celltype[i], p, dt,
                       ist ? ist[i] : 0.0, dvgap ? dvgap[i] : 0.0, 0);
}

/* SingleCell.cc:38-287 (protocol :95-110, statistics :144-215).  The
 * statistics locals that the reference declares inside the loop body are
 * kept here as variables that persist across iterations, which is what the
 * reference's optimised build does with them. */
int64_t or_single_cell(int celltype, const or_params *p, double bcl, int nbeats,
                       double dt, double stimStrength, double stimDuration,
                       double stimStart, double *stats,
                       double *trace, int trace_every, int64_t trace_cap)
{
    double s[OR_NSTATE];
    or_cell_init(s, -87.5);
    or_diag dg;
    double numS1 = nbeats;
    double stopTime = numS1 * bcl;
    double apd20 = 0, apd25 = 0, apd50 = 0, apd75 = 0, apd80 = 0, apd90 = 0;
    double dvdt_max = 0, rest_potential = 0, amplitude = 0, overshoot = 0;
    double oldV = 0, repo20 = 0, repo25 = 0, repo50 = 0, repo75 = 0, repo80 = 0, repo90 = 0;
    int peakfound = 0;
    double time;
    int64_t step;
    int64_t ntrace = 0;
    double absdvdt = 0.0;
    for (time = 0.0, step = 0; time < stopTime; time += dt, step++) {
        double ist;
        if (time - floor(time / bcl) * bcl >= stimStart && time - floor(time / bcl) * bcl < stimStart + stimDuration)
            ist = stimStrength;
        else
            ist = 0.0;
        or_cell_update(s, celltype, p, dt, ist, 0.0, &dg);
        absdvdt = fabs(dg.dV_dt);
        double V = s[OR_V];
        if (trace && trace_every > 0 && step % trace_every == 0 && ntrace < trace_cap) trace[ntrace++] = V;

        if (floor((time - stimStart) / bcl) == numS1 - 2 && floor((time + dt - stimStart) / bcl) == numS1 - 1) {
            rest_potential = V;
            oldV = V;
            peakfound = 0;
            dvdt_max = -1;
        }
        if (floor((time - stimStart) / bcl) == numS1 - 1) {
            if (V > oldV) {
                oldV = V;
                if (absdvdt > dvdt_max) dvdt_max = absdvdt;
            } else if (V <= oldV && !peakfound) {
                peakfound = 1;
                overshoot = oldV;
                amplitude = overshoot - rest_potential;
                repo20 = overshoot - 0.20 * amplitude;
                repo25 = overshoot - 0.25 * amplitude;
                repo50 = overshoot - 0.50 * amplitude;
                repo75 = overshoot - 0.75 * amplitude;
                repo80 = overshoot - 0.80 * amplitude;
                repo90 = overshoot - 0.90 * amplitude;
                oldV = V;
            } else if (V <= oldV && peakfound) {
                double tc = time - floor(time / bcl) * bcl - stimStart - stimDuration;
                if (oldV >= repo20 && V <= repo20) apd20 = tc;
                else if (oldV >= repo25 && V <= repo25) apd25 = tc;
                else if (oldV >= repo50 && V <= repo50) apd50 = tc;
                else if (oldV >= repo75 && V <= repo75) apd75 = tc;
                else if (oldV >= repo80 && V <= repo80) apd80 = tc;
                else if (oldV >= repo90 && V <= repo90) apd90 = tc;
                oldV = V;
            }
        }
    }
    if (stats) {
        stats[0] = rest_potential; stats[1] = dvdt_max; stats[2] = overshoot; stats[3] = amplitude;
        stats[4] = apd20; stats[5] = apd25; stats[6] = apd50; stats[7] = apd75; stats[8] = apd80; stats[9] = apd90;
    }
    return step;
}

void or_track_init(or_track *t)
{
    t->act = -1.0;
    t->rep = -1.0;
    t->vrest = 0.0;
    t->vmax = -1e30;
    t->vmin = 1e30;
    t->flags = 0;
}

void or_track_step(or_track *t, double time, double v)
{
    if (!(t->flags & 1)) {
        t->flags |= 1;
        t->vrest = v;
        /* still depolarised from the previous beat: it failed to repolarise */
        if (v > -60.0) t->flags |= 16;
    }
    if (!(t->flags & 2)) {
        if (v >= -30.0) {
            t->flags |= 2;
            t->act = time;
            t->vmax = v;
        }
        return;
    }
    if (t->flags & 4) return;
    if (v > t->vmax) t->vmax = v;
    double level = t->vmax - 0.9 * (t->vmax - t->vrest);
    if (v <= level) {
        t->flags |= 4;
        t->rep = time;
    } else if (time >= t->act + 100.0) {
        if (v < t->vmin) t->vmin = v;
        if (v > t->vmin + 2.0) t->flags |= 8;
    }
}

int or_track_result(const or_track *t, double *act, double *apd)
{
    if (act) *act = (t->flags & 2) ? t->act : -1.0;
    if (apd) *apd = (t->flags & 4) ? t->rep - t->act : -1.0;
    int e = 0;
    if (t->flags & 8) e |= 1;
    if (((t->flags & 2) && !(t->flags & 4)) || (t->flags & 16)) e |= 2;
    return e;
}

/* OneD.cc:315-436 */
int64_t or_strand_run2(int n, const uint8_t *celltype, const double *coeff,
                       const or_params *p, double dx, double dt, double bcl,
                       double stopTime, double stimStrength, double stimDuration,
                       double stimStart, int nstim_cells,
                       double *vrec, double *rec_t, int rec_every, int64_t rec_cap,
                       int64_t *nrows, int act_beat, double *act, double *apd, int *ead,
                       double *final_state)
{
    double *s = (double *)malloc(sizeof(double) * OR_NSTATE * (size_t)n);
    double *oldV = (double *)malloc(sizeof(double) * (size_t)n);
    double *dcoeff = (double *)malloc(sizeof(double) * (size_t)n);
    or_track *tr = (or_track *)malloc(sizeof(or_track) * (size_t)n);
    for (int i = 0; i < n; ++i) or_cell_init(s + (size_t)i * OR_NSTATE, -87.5);
    for (int i = 0; i < n; ++i) or_track_init(&tr[i]);
    /* OneD.cc:121-131 */
    for (int i = 0; i < n; ++i) {
        if (n == 1) dcoeff[i] = 0.0;
        else if (i == 0) dcoeff[i] = (coeff[i + 1] - coeff[i]) / dx;
        else if (i == n - 1) dcoeff[i] = (coeff[i] - coeff[i - 1]) / dx;
        else dcoeff[i] = (coeff[i + 1] - coeff[i - 1]) / (2.0 * dx);
    }
    int64_t rows = 0;
    double time;
    int64_t step;
    for (time = 0.0, step = 0; time <= stopTime; time += dt, step++) {
        for (int i = 0; i < n; ++i) oldV[i] = s[(size_t)i * OR_NSTATE + OR_V];
        int stim_on = (time - floor(time / bcl) * bcl >= stimStart &&
                       time - floor(time / bcl) * bcl < stimStart + stimDuration);
#pragma omp parallel for schedule(static) if (n >= 512)
        for (int i = 0; i < n; ++i) {
            double ist = (stim_on && i < nstim_cells) ? stimStrength : 0.0;
            double first_order, second_order;
            if (n == 1) {
                first_order = 0.0;
                second_order = 0.0;
            } else if (i == 0) {
                first_order = (oldV[i + 1] - oldV[i]) / (1.0 * dx);
                second_order = (oldV[i + 1] + oldV[i] - 2.0 * oldV[i]) / (dx * dx);
            } else if (i < n - 1) {
                first_order = (oldV[i + 1] - oldV[i - 1]) / (2.0 * dx);
                second_order = (oldV[i + 1] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            } else {
                first_order = (oldV[i] - oldV[i - 1]) / (1.0 * dx);
                second_order = (oldV[i] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            }
            double dVgap_dt = dcoeff[i] * first_order + coeff[i] * second_order;
            or_cell_update(s + (size_t)i * OR_NSTATE, celltype[i], p, dt, ist, dVgap_dt, 0);
        }
        if (vrec && rec_every > 0 && step % rec_every == 0 && rows < rec_cap) {
            for (int i = 0; i < n; ++i) vrec[rows * n + i] = s[(size_t)i * OR_NSTATE + OR_V];
            if (rec_t) rec_t[rows] = time;
            rows++;
        }
        if (floor(time / bcl) == act_beat)
            for (int i = 0; i < n; ++i) or_track_step(&tr[i], time, s[(size_t)i * OR_NSTATE + OR_V]);
    }
    for (int i = 0; i < n; ++i) {
        double a, d;
        int e = or_track_result(&tr[i], &a, &d);
        if (act) act[i] = a;
        if (apd) apd[i] = d;
        if (ead) ead[i] = e;
    }
    if (nrows) *nrows = rows;
    if (final_state) memcpy(final_state, s, sizeof(double) * OR_NSTATE * (size_t)n);
    free(s);
    free(oldV);
    free(dcoeff);
    free(tr);
    return step;
}

/* OneD_VW.cc:239-398, one S2 candidate.  The S1 test and the diffusion
 * term are those of OneD.cc (OneD_VW.cc:264-279,306-338); S2 is the
 * `else if` of :291-301, so it is looked at only while the S1 window is
 * closed; the peaks are :390-394, taken after the update with the time of
 * the step.  The caller owns the state so that a run can pause (at the
 * first step with time >= pause_at) and resume, or be copied to fork
 * candidates that share their S1 prefix. */
int64_t or_strand_vw(int n, const uint8_t *celltype, const double *coeff,
                     const or_params *p, double dx, double dt, double bcl,
                     double stopTime, double stimStrength, double stimDuration,
                     double stimStart, int nstim_cells, const or_s2 *s2,
                     double *s, double *time_io, int64_t *step_io, double pause_at,
                     double *vmax, double *vrec, double *rec_t, int rec_every,
                     int rec_beat, int64_t rec_cap, int64_t *nrows)
{
    double *oldV = (double *)malloc(sizeof(double) * (size_t)n);
    double *dcoeff = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        if (n == 1) dcoeff[i] = 0.0;
        else if (i == 0) dcoeff[i] = (coeff[i + 1] - coeff[i]) / dx;
        else if (i == n - 1) dcoeff[i] = (coeff[i] - coeff[i - 1]) / dx;
        else dcoeff[i] = (coeff[i + 1] - coeff[i - 1]) / (2.0 * dx);
    }
    int64_t rows = nrows ? *nrows : 0;
    double time = *time_io;
    int64_t step = *step_io;
    for (; time <= stopTime && time < pause_at; time += dt, step++) {
        for (int i = 0; i < n; ++i) oldV[i] = s[(size_t)i * OR_NSTATE + OR_V];
        int s1_on = (time - floor(time / bcl) * bcl >= stimStart &&
                     time - floor(time / bcl) * bcl < stimStart + stimDuration);
        int s2_on = !s1_on && s2 && time - s2->base >= s2->start &&
                    time - s2->base < s2->start + s2->duration;
#pragma omp parallel for schedule(static) if (n >= 512)
        for (int i = 0; i < n; ++i) {
            double ist = 0.0;
            if (s1_on) {
                if (i < nstim_cells) ist = stimStrength;
            } else if (s2_on && i >= s2->first && i < s2->first + s2->count) {
                ist = s2->strength;
            }
            double first_order, second_order;
            if (n == 1) {
                first_order = 0.0;
                second_order = 0.0;
            } else if (i == 0) {
                first_order = (oldV[i + 1] - oldV[i]) / (1.0 * dx);
                second_order = (oldV[i + 1] + oldV[i] - 2.0 * oldV[i]) / (dx * dx);
            } else if (i < n - 1) {
                first_order = (oldV[i + 1] - oldV[i - 1]) / (2.0 * dx);
                second_order = (oldV[i + 1] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            } else {
                first_order = (oldV[i] - oldV[i - 1]) / (1.0 * dx);
                second_order = (oldV[i] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            }
            double dVgap_dt = dcoeff[i] * first_order + coeff[i] * second_order;
            or_cell_update(s + (size_t)i * OR_NSTATE, celltype[i], p, dt, ist, dVgap_dt, 0);
        }
        if (vrec && rec_every > 0 && (rec_beat < 0 || floor(time / bcl) == rec_beat) &&
            step % rec_every == 0 && rows < rec_cap) {
            for (int i = 0; i < n; ++i) vrec[rows * n + i] = s[(size_t)i * OR_NSTATE + OR_V];
            if (rec_t) rec_t[rows] = time;
            rows++;
        }
        if (vmax && s2 && time - s2->base >= s2->start) {
            double ve = s[OR_V], vl = s[(size_t)(n - 1) * OR_NSTATE + OR_V];
            if (vl >= vmax[1]) vmax[1] = vl;
            if (ve >= vmax[0]) vmax[0] = ve;
        }
    }
    *time_io = time;
    *step_io = step;
    if (nrows) *nrows = rows;
    free(oldV);
    free(dcoeff);
    return step;
}

int64_t or_strand_run(int n, const uint8_t *celltype, const double *coeff,
                      const or_params *p, double dx, double dt, double bcl,
                      double stopTime, double stimStrength, double stimDuration,
                      double stimStart, int nstim_cells,
                      double *vrec, double *rec_t, int rec_every, int64_t rec_cap,
                      int64_t *nrows, int act_beat, double *act, double *final_state)
{
    return or_strand_run2(n, celltype, coeff, p, dx, dt, bcl, stopTime, stimStrength, stimDuration,
                          stimStart, nstim_cells, vrec, rec_t, rec_every, rec_cap, nrows, act_beat,
                          act, 0, 0, final_state);
}

/* PlotScript/plot1DECG_compare.py:89-105 */
void or_pseudo_ecg(int64_t T, int L, const double *v, double dx, double x0, double *phi)
{
    for (int64_t t = 0; t < T; ++t) {
        const double *row = v + t * L;
        double fai = 0.0;
        for (int i = 0; i < L; ++i) {
            double r = (i * dx - x0) * (i * dx - x0);
            if (i == 0) fai += -(row[i + 1] - row[i]) / r;
            else if (i == L - 1) fai += -(row[i] - row[i - 1]) / r;
            else fai += -(row[i + 1] - row[i - 1]) / (2 * r);
        }
        phi[t] = fai;
    }
}

/* ---------------------------------------------------------------------- */
/* tissue                                                                 */

/* GPU3D.cu:1301-1333 (compaction) and :1487-1604 (ghost mirror), restated
 * for an nx*ny*nz grid.  The 26 neighbours are numbered as in the
 * reference: gg[1..9] at x-1, gg[10..17] at x, gg[18..26] at x+1, each
 * block in (y, z) lexicographic order skipping the centre. */
/* The 2D sheet of new_11.cu:1535-1604 (nz == 1): 8 neighbours numbered
 *   | 1 2 3 |
 *   | 4 s 5 |      (x to the right, y down)
 *   | 6 7 8 |
 * il (x) and ir (y) summed in the reference's order, the tissue neighbour
 * nearest to (il, ir)/max(|il|,|ir|) searched y outer, x inner.  The
 * reference reads past the grid edge for x+1 == X and y+1 == Y (its tests
 * are `x+1 >= 0`, `y+1 >= 0`); here such neighbours count as non-tissue,
 * which is what the reference sees whenever the tissue keeps a one-cell
 * margin. */
static void ghost_map_2d(int nx, int ny, int32_t *g)
{
    const double root2 = sqrt(2.0);
    int gg[9];
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            if (g[(int64_t)x * ny + y] != 0) continue;
            static const int ox[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
            static const int oy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};
            int isbound = 0;
            for (int i = 1; i <= 8; ++i) {
                int xx = x + ox[i], yy = y + oy[i];
                int val = (xx >= 0 && xx < nx && yy >= 0 && yy < ny) ? g[(int64_t)xx * ny + yy] : 0;
                gg[i] = val > 0 ? 1 : 0;
                isbound += gg[i];
            }
            if (isbound == 0) continue;
            double il = -gg[1] / root2 + gg[3] / root2 - gg[4] + gg[5] - gg[6] / root2 + gg[8] / root2;
            double ir = -gg[1] / root2 - gg[2] - gg[3] / root2 + gg[6] / root2 + gg[7] + gg[8] / root2;
            double imax = fabs(il);
            if (fabs(ir) > imax) imax = fabs(ir);
            double nxv = il / imax, nyv = ir / imax;
            if (imax - 0 < 0.0001) { nxv = 0; nyv = 0; }
            double dislow = 1000;
            int bx = 0, by = 0;
            for (int yc = -1; yc <= 1; ++yc)
                for (int xc = -1; xc <= 1; ++xc) {
                    double dis = sqrt(1.0 * (xc - nxv) * (xc - nxv) + (yc - nyv) * (yc - nyv));
                    int xx = x + xc, yy = y + yc;
                    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny)
                        if (g[(int64_t)xx * ny + yy] > 0 && dis < dislow) {
                            bx = xc; by = yc;
                            dislow = dis;
                        }
                }
            g[(int64_t)x * ny + y] = -g[(int64_t)(x + bx) * ny + (y + by)];
        }
}

int64_t or_ghost_map(int nx, int ny, int nz, const uint8_t *mask, int32_t *g)
{
    const int64_t DX = (int64_t)ny * nz, DY = nz;
    const int64_t ncell = (int64_t)nx * DX;
    int64_t num = 0;
    for (int64_t c = 0; c < ncell; ++c) g[c] = mask[c] ? (int32_t)(++num) : 0;
    if (nz == 1) {
        ghost_map_2d(nx, ny, g);
        return num;
    }
    const double root2 = sqrt(2.0), root3 = sqrt(3.0);
    int gg[27];
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                int64_t c = x * DX + y * DY + z;
                if (g[c] != 0) continue;
                /* neighbours whose value is still a tissue index count as
                 * tissue; ghosts assigned earlier in the sweep are negative
                 * and do not (GPU3D.cu:1541-1553) */
                int k = 1;
                int isbound = 0;
                for (int dxo = -1; dxo <= 1; ++dxo)
                    for (int dyo = -1; dyo <= 1; ++dyo)
                        for (int dzo = -1; dzo <= 1; ++dzo) {
                            if (dxo == 0 && dyo == 0 && dzo == 0) continue;
                            int xx = x + dxo, yy = y + dyo, zz = z + dzo;
                            int val = 0;
                            if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                                val = g[xx * DX + yy * DY + zz];
                            gg[k] = val > 0 ? 1 : 0;
                            isbound += gg[k];
                            ++k;
                        }
                if (isbound == 0) continue;
                double ic = (gg[3] / root3) - (gg[1] / root3) + (gg[6] / root2) + (gg[9] / root3) - (gg[7] / root3) -
                            (gg[4] / root2) + (gg[12] / root2) - (gg[10] / root2) + gg[14] + (gg[17] / root2) -
                            (gg[15] / root2) - gg[13] + (gg[20] / root3) - (gg[18] / root3) + (gg[23] / root2) +
                            (gg[26] / root3) - (gg[24] / root3) - (gg[21] / root2);
                double ir = (gg[9] / root3) - (gg[2] / root2) - (gg[3] / root3) - (gg[1] / root3) + (gg[8] / root2) +
                            (gg[7] / root3) + (gg[17] / root2) - gg[11] - (gg[12] / root2) - (gg[10] / root2) + gg[16] +
                            (gg[15] / root2) + (gg[26] / root3) - (gg[19] / root2) - (gg[20] / root3) -
                            (gg[18] / root3) + (gg[25] / root2) + (gg[24] / root3);
                double il = (gg[18] / root3) + (gg[19] / root2) + (gg[20] / root3) + (gg[21] / root2) + gg[22] +
                            (gg[23] / root2) + (gg[24] / root3) + (gg[25] / root2) + (gg[26] / root3) -
                            (gg[1] / root3) - (gg[2] / root2) - (gg[3] / root3) - (gg[4] / root2) - gg[5] -
                            (gg[6] / root2) - (gg[7] / root3) - (gg[8] / root2) - (gg[9] / root3);
                double imax = fabs(ic);
                if (fabs(ir) > imax) imax = fabs(ir);
                if (fabs(il) > imax) imax = fabs(il);
                double nxv = il / imax, nyv = ir / imax, nzv = ic / imax;
                if (imax - 0 < 0.0001) { nxv = 0; nyv = 0; nzv = 0; }
                double dislow = 1000;
                int bx = 0, by = 0, bz = 0;
                for (int zc = -1; zc <= 1; ++zc)
                    for (int yc = -1; yc <= 1; ++yc)
                        for (int xc = -1; xc <= 1; ++xc) {
                            double dis = sqrt(1.0 * (xc - nxv) * (xc - nxv) + (yc - nyv) * (yc - nyv) +
                                              (zc - nzv) * (zc - nzv));
                            int xx = x + xc, yy = y + yc, zz = z + zc;
                            if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                                if (g[xx * DX + yy * DY + zz] > 0 && dis < dislow) {
                                    bx = xc; by = yc; bz = zc;
                                    dislow = dis;
                                }
                        }
                g[c] = -g[(x + bx) * DX + (y + by) * DY + (z + bz)];
            }
    return num;
}

static void decode(int64_t c, int ny, int nz, int *x, int *y, int *z)
{
    *z = (int)(c % nz);
    int64_t r = c / nz;
    *y = (int)(r % ny);
    *x = (int)(r / ny);
}

void or_tissue_tensor(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                      const float *fibre, double d_fibre, double d_cross,
                      double dx, double *dten, double *drift)
{
    const int64_t ncell = (int64_t)nx * ny * nz;
    const int64_t strides[3] = {(int64_t)ny * nz, nz, 1};
    const double DD = d_fibre - d_cross;
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)nvox);
    for (int64_t c = 0; c < ncell; ++c)
        if (g[c] > 0) pos[g[c] - 1] = c;
    for (int64_t k = 0; k < nvox; ++k) {
        double f0 = fibre[3 * k], f1 = fibre[3 * k + 1], f2 = fibre[3 * k + 2];
        double *D = dten + 6 * k;
        D[0] = d_cross + DD * f0 * f0;
        D[1] = d_cross + DD * f1 * f1;
        D[2] = d_cross + DD * f2 * f2;
        D[3] = DD * f0 * f1;
        D[4] = DD * f0 * f2;
        D[5] = DD * f1 * f2;
    }
    /* map (i, j) -> component of the symmetric tensor */
    static const int comp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    for (int64_t k = 0; k < nvox; ++k) {
        int x, y, z;
        decode(pos[k], ny, nz, &x, &y, &z);
        int co[3] = {x, y, z};
        int dims[3] = {nx, ny, nz};
        double b[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
            int32_t gp = 0, gm = 0;
            if (co[i] + 1 < dims[i]) gp = g[pos[k] + strides[i]];
            if (co[i] - 1 >= 0) gm = g[pos[k] - strides[i]];
            for (int j = 0; j < 3; ++j) {
                int cj = comp[i][j];
                double d;
                if (gp > 0 && gm > 0) d = (dten[6 * (gp - 1) + cj] - dten[6 * (gm - 1) + cj]) / (2 * dx);
                else if (gp > 0) d = (dten[6 * (gp - 1) + cj] - dten[6 * k + cj]) / (dx);
                else if (gm > 0) d = (dten[6 * k + cj] - dten[6 * (gm - 1) + cj]) / (dx);
                else d = 0;
                b[j] += d;
            }
        }
        drift[3 * k] = b[0];
        drift[3 * k + 1] = b[1];
        drift[3 * k + 2] = b[2];
    }
    free(pos);
}

static inline double vat(const int32_t *g, const double *v, int64_t c)
{
    int32_t t = g[c];
    if (t > 0) return v[t - 1];
    if (t < 0) return v[-t - 1];
    return -90.0;
}

void or_tissue_du(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                  const double *dten, const double *drift, double dx,
                  const double *v, double *du)
{
    const int64_t ncell = (int64_t)nx * ny * nz;
    const int64_t SX = (int64_t)ny * nz, SY = nz, SZ = 1;
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)nvox);
    for (int64_t c = 0; c < ncell; ++c)
        if (g[c] > 0) pos[g[c] - 1] = c;
    const double h2 = dx * dx;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nvox; ++k) {
        int64_t c = pos[k];
        int x, y, z;
        decode(c, ny, nz, &x, &y, &z);
        /* neighbour value with the grid border treated as outside */
#define NB(ox, oy, oz)                                                                             \
    ((x + (ox) >= 0 && x + (ox) < nx && y + (oy) >= 0 && y + (oy) < ny && z + (oz) >= 0 &&        \
      z + (oz) < nz)                                                                               \
         ? vat(g, v, c + (ox) * SX + (oy) * SY + (oz) * SZ)                                        \
         : -90.0)
        double vc = v[k];
        double vxp = NB(1, 0, 0), vxm = NB(-1, 0, 0);
        double vyp = NB(0, 1, 0), vym = NB(0, -1, 0);
        double vxx = (vxm + vxp - 2 * vc) / h2;
        double vyy = (vym + vyp - 2 * vc) / h2;
        double vxy = (NB(1, 1, 0) + NB(-1, -1, 0) - NB(-1, 1, 0) - NB(1, -1, 0)) / (4 * h2);
        double vx = (vxp - vxm) / (2 * dx);
        double vy = (vyp - vym) / (2 * dx);
        const double *D = dten + 6 * k;
        const double *b = drift + 3 * k;
        if (nz == 1) {
            /* the 2D sheet of new_11.cu:352-417: no z terms */
            du[k] = D[0] * vxx + D[1] * vyy + 2 * D[3] * vxy + b[0] * vx + b[1] * vy;
            continue;
        }
        double vzp = NB(0, 0, 1), vzm = NB(0, 0, -1);
        double vzz = (vzm + vzp - 2 * vc) / h2;
        double vxz = (NB(1, 0, 1) + NB(-1, 0, -1) - NB(-1, 0, 1) - NB(1, 0, -1)) / (4 * h2);
        double vyz = (NB(0, 1, 1) + NB(0, -1, -1) - NB(0, -1, 1) - NB(0, 1, -1)) / (4 * h2);
        double vz = (vzp - vzm) / (2 * dx);
#undef NB
        du[k] = D[0] * vxx + D[1] * vyy + D[2] * vzz + 2 * (D[3] * vxy + D[4] * vxz + D[5] * vyz) +
                b[0] * vx + b[1] * vy + b[2] * vz;
    }
    free(pos);
}

void or_tissue_step(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                    const double *dten, const double *drift, double dx, double dt,
                    const uint8_t *celltype, const or_params *p,
                    double *s, const double *ist, double *vbuf, double *du)
{
    for (int64_t k = 0; k < nvox; ++k) vbuf[k] = s[(size_t)k * OR_NSTATE + OR_V];
    or_tissue_du(nx, ny, nz, g, nvox, dten, drift, dx, vbuf, du);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nvox; ++k) {
        or_cell_update(s + (size_t)k * OR_NSTATE, celltype[k] & 0x7f, p, dt, ist ? ist[k] : 0.0, du[k], 0);
        /* the type-0 voxels of new_11.cu:1247-1248 */
        if (celltype[k] & 0x80) s[(size_t)k * OR_NSTATE + OR_V] = -100.0;
    }
}
