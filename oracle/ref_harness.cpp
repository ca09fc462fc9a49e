// C entry points around the reference's own TPORd class, compiled straight
// from /root/reference/SingleCell/TPORd.cc by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY: used to pin the plain-C oracle,
// never by the product.
//
// The standard headers TPORd.cc pulls in are included first so that the
// access override below only reaches the model class itself; that gives the
// harness the full-precision state instead of the 10-decimal text of
// TPORd::outputAllStates (TPORd.cc:200-252).
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <ctime>
#include <fstream>
#include <iostream>
#include <typeinfo>
#include <omp.h>

#define private public
#include "SingleCell/TPORd.cc"
#undef private

extern "C" {

void *ref_cell_new(int celltype)
{
    return new TPORd(static_cast<CellType>(celltype));
}

void ref_cell_free(void *c) { delete static_cast<TPORd *>(c); }

void ref_cell_set_dt(void *c, double dt) { static_cast<TPORd *>(c)->setDt(dt); }

// One Cell::update() with the stimulus and gap current that the drivers set
// before it (OneD.cc:335-391).
void ref_cell_step(void *c, double ist, double dvgap)
{
    TPORd *t = static_cast<TPORd *>(c);
    t->setIstim(ist);
    t->setDVgap_dt(dvgap);
    t->update();
}

double ref_cell_dvdt(void *c) { return static_cast<TPORd *>(c)->getDvdt(); }
double ref_cell_ical(void *c) { return static_cast<TPORd *>(c)->getICaL(); }
double ref_cell_iks(void *c) { return static_cast<TPORd *>(c)->getIKs(); }

// All 40 states in the order of TPORd::outputAllStates.
void ref_cell_get(void *c, double *o)
{
    TPORd *t = static_cast<TPORd *>(c);
    const double vals[40] = {t->v, t->nai, t->nass, t->ki, t->kss, t->cai, t->cass, t->cansr,
                             t->cajsr, t->sm, t->sh, t->sj, t->hsp, t->jp, t->mL, t->hL,
                             t->hLp, t->a, t->iF, t->iS, t->ap, t->iFp, t->iSp, t->d,
                             t->ff, t->fs, t->fcaf, t->fcas, t->jca, t->nca, t->ffp, t->fcafp,
                             t->xrf, t->xrs, t->xs1, t->xs2, t->xk1, t->Jrelnp, t->Jrelp, t->CaMKt};
    memcpy(o, vals, sizeof(vals));
}

void ref_cell_set(void *c, const double *i)
{
    TPORd *t = static_cast<TPORd *>(c);
    double *dst[40] = {&t->v, &t->nai, &t->nass, &t->ki, &t->kss, &t->cai, &t->cass, &t->cansr,
                       &t->cajsr, &t->sm, &t->sh, &t->sj, &t->hsp, &t->jp, &t->mL, &t->hL,
                       &t->hLp, &t->a, &t->iF, &t->iS, &t->ap, &t->iFp, &t->iSp, &t->d,
                       &t->ff, &t->fs, &t->fcaf, &t->fcas, &t->jca, &t->nca, &t->ffp, &t->fcafp,
                       &t->xrf, &t->xrs, &t->xs1, &t->xs2, &t->xk1, &t->Jrelnp, &t->Jrelp, &t->CaMKt};
    for (int k = 0; k < 40; ++k) *dst[k] = i[k];
}

// A whole strand of n cells stepped like OneD.cc:315-398 for nsteps steps
// (uniform coefficient, stimulus on the first nstim cells while stim_on is
// set for the step).  Kept here so the strand pin runs the reference's
// class, not the oracle's.
void ref_strand_steps(void **cells, int n, double coeff, double dx, double dt,
                      long nsteps, double t0, double bcl, double stim_start,
                      double stim_dur, double stim_strength, int nstim, double *time_io)
{
    double *oldV = new double[n];
    double time = t0;
    for (long s = 0; s < nsteps; ++s, time += dt) {
        for (int i = 0; i < n; ++i) oldV[i] = static_cast<TPORd *>(cells[i])->getV();
        for (int i = 0; i < n; ++i) {
            TPORd *t = static_cast<TPORd *>(cells[i]);
            double ist = 0.0;
            if (time - floor(time / bcl) * bcl >= stim_start &&
                time - floor(time / bcl) * bcl < stim_start + stim_dur && i < nstim)
                ist = stim_strength;
            double second;
            if (i == 0) second = (oldV[i + 1] + oldV[i] - 2.0 * oldV[i]) / (dx * dx);
            else if (i < n - 1) second = (oldV[i + 1] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            else second = (oldV[i] + oldV[i - 1] - 2.0 * oldV[i]) / (dx * dx);
            t->setIstim(ist);
            t->setDVgap_dt(0.0 * 0.0 + coeff * second);
            t->update();
        }
    }
    *time_io = time;
    delete[] oldV;
}

}  // extern "C"
