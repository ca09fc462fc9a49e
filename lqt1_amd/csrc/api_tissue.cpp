// C ABI: the tissue handle.  Geometry set-up restates GPU3D.cu:1263-1335
// (compaction), :1487-1604 (ghost mirror map) and new_11.cu:1496-1499 /
// :381-407 (diffusion tensor and its one-sided divergence) in 3D; the time
// loop is GPU3D.cu:1724-1842 with the stimulus of :1757-1762.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"

using namespace lqt;

struct lqt_tissue {
    int nx = 0, ny = 0, nz = 0;
    int px = 0, py = 0, pz = 0;  // padded dims
    double dx = 0, dt = 0;
    bool fp32 = true, aniso = false;
    double d_fibre = 0, d_cross = 0;
    int64_t nvox = 0, nghost = 0, ncell_pad = 0, pitch = 0;
    std::vector<int32_t> h_vpos;   // padded index of voxel k
    std::vector<int64_t> h_cell;   // unpadded grid index of voxel k
    std::vector<int32_t> h_gmap;   // unpadded grid -> k+1 (0 outside)
    std::vector<uint8_t> h_label;  // per voxel, enum CellType
    lqt_params params{};
    // device
    int32_t* vpos = nullptr;
    uint8_t* cls = nullptr;
    void* grid[2] = {nullptr, nullptr};
    int cur = 0;
    void* r = nullptr;
    // slow accumulators: double (FP64), or float offsets from dref (FP32)
    void* d = nullptr;
    double dref[ND] = {0, 0, 0, 0, 0};
    void* fd = nullptr;  // fibre and drift, chunk-major (6 slots)
    int32_t* gdst = nullptr;   // ghost cells (padded index) and their sources,
    int32_t* gsrc = nullptr;   // for refreshes outside the step kernel
    int32_t* gout = nullptr;   // ghost cells listed by the voxel they mirror
    int32_t* gwave = nullptr;  // first gout entry of each wavefront's voxels
    void* consts = nullptr;
    float* stim_off = nullptr;
    uint8_t* s2 = nullptr;
    unsigned long long* guard = nullptr;
    // per-voxel activation / APD90 trackers, updated after each step that
    // starts in [trk_from, trk_to)
    VoxTrack* track = nullptr;
    bool trk_on = false;
    double trk_from = 0, trk_to = 0;
    // stimulus
    lqt_stim_protocol proto{};
    int64_t nsites = 0, ns2 = 0;
    double off_min = 0, off_max = 0;
    // time
    double time = 0;
    int64_t step = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;
    double us_step = 0, us_ghost = 0;
    int variant = 0;  // launch shape of the fused step (kTissueVariants)
};

static size_t rsz(const lqt_tissue* t) { return t->fp32 ? sizeof(float) : sizeof(double); }
static size_t dsz(const lqt_tissue* t) { return t->fp32 ? sizeof(float) : sizeof(double); }

// Stored form of slow slot j of value x, and back (FP32: offset from dref).
static void slow_put(const lqt_tissue* t, int j, double x, void* out) {
    if (t->fp32) *(float*)out = (float)(x - t->dref[j]);
    else *(double*)out = x;
}
static double slow_get(const lqt_tissue* t, int j, const void* in) {
    return t->fp32 ? t->dref[j] + (double)*(const float*)in : *(const double*)in;
}


// Ghost map of GPU3D.cu:1487-1604: every non-tissue cell with a tissue cell
// among its 26 neighbours mirrors the tissue neighbour closest to the
// outward normal estimated from the neighbour pattern.  Only tissue cells
// count in the pattern and as candidates, so the result does not depend on
// the sweep order.  The normal components are summed term by term in the
// reference's order so that exact ties between candidates resolve the same
// way.
namespace {
// neighbour n (1..26) of the reference numbering is at
// (x + NX[n], y + NY[n], z + NZ[n])
constexpr int NXo[27] = {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr int NYo[27] = {0, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, -1, -1, 0, 0, 1, 1, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1};
constexpr int NZo[27] = {0, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1};
// signed terms (+-n, weight class 1/2/3 = 1, 1/sqrt2, 1/sqrt3) of the z, y
// and x components, in the order written at GPU3D.cu:1565-1571
struct Term {
    int8_t n, w;
};
constexpr Term IC[18] = {{3, 3},   {-1, 3},  {6, 2},  {9, 3},   {-7, 3},  {-4, 2},  {12, 2},  {-10, 2}, {14, 1},
                         {17, 2},  {-15, 2}, {-13, 1}, {20, 3}, {-18, 3}, {23, 2},  {26, 3},  {-24, 3}, {-21, 2}};
constexpr Term IR[18] = {{9, 3},   {-2, 2},  {-3, 3}, {-1, 3},  {8, 2},   {7, 3},   {17, 2},  {-11, 1}, {-12, 2},
                         {-10, 2}, {16, 1},  {15, 2}, {26, 3},  {-19, 2}, {-20, 3}, {-18, 3}, {25, 2},  {24, 3}};
constexpr Term IL[18] = {{18, 3}, {19, 2}, {20, 3}, {21, 2}, {22, 1}, {23, 2}, {24, 3}, {25, 2}, {26, 3},
                         {-1, 3}, {-2, 2}, {-3, 3}, {-4, 2}, {-5, 1}, {-6, 2}, {-7, 3}, {-8, 2}, {-9, 3}};
}  // namespace

// The 2D sheet (nz == 1) of new_11.cu:1535-1604: 8 neighbours, the x and
// y components of the tissue direction summed in the reference's order, the
// nearest tissue neighbour searched y outer, x inner.  Cells beyond the grid
// edge count as non-tissue (the reference reads past the edge for x+1 == X
// and y+1 == Y, which is harmless there only when the tissue keeps a margin).
static int64_t ghost_map_2d(int nx, int ny, std::vector<int32_t>& g) {
    static const int ox[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
    static const int oy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};
    const double r2 = sqrt(2.0);
    auto tis = [&](int x, int y) -> int {
        return (x >= 0 && x < nx && y >= 0 && y < ny) ? g[(int64_t)x * ny + y] > 0 : 0;
    };
    int64_t ng = 0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            if (g[(int64_t)x * ny + y] != 0) continue;
            int gg[9];
            int nb = 0;
            for (int i = 1; i <= 8; ++i) nb += (gg[i] = tis(x + ox[i], y + oy[i]));
            if (!nb) continue;
            const double il = -gg[1] / r2 + gg[3] / r2 - gg[4] + gg[5] - gg[6] / r2 + gg[8] / r2;
            const double ir = -gg[1] / r2 - gg[2] - gg[3] / r2 + gg[6] / r2 + gg[7] + gg[8] / r2;
            double imax = fabs(il);
            if (fabs(ir) > imax) imax = fabs(ir);
            double nxv = il / imax, nyv = ir / imax;
            if (imax - 0 < 0.0001) nxv = nyv = 0;
            double best = 1000;
            int bx = 0, by = 0;
            for (int oyc = -1; oyc <= 1; ++oyc)
                for (int oxc = -1; oxc <= 1; ++oxc) {
                    const double dis = sqrt(1.0 * (oxc - nxv) * (oxc - nxv) + (oyc - nyv) * (oyc - nyv));
                    if (tis(x + oxc, y + oyc) && dis < best) {
                        best = dis;
                        bx = oxc;
                        by = oyc;
                    }
                }
            g[(int64_t)x * ny + y] = -g[(int64_t)(x + bx) * ny + (y + by)];
            ++ng;
        }
    return ng;
}

// g[c] = k+1 (tissue voxel k), -(k+1) (ghost mirroring voxel k), 0 otherwise.
static void ghost_map(int nx, int ny, int nz, const uint8_t* label, std::vector<int32_t>& g, int64_t* nvox,
                      int64_t* nghost) {
    const int64_t SX = (int64_t)ny * nz, SY = nz, ncell = (int64_t)nx * SX;
    g.assign(ncell, 0);
    int64_t num = 0;
    for (int64_t c = 0; c < ncell; ++c)
        if (label[c] != LQT_NONE) g[c] = (int32_t)(++num);
    if (nz == 1) {
        const int64_t ng = ghost_map_2d(nx, ny, g);
        if (nvox) *nvox = num;
        if (nghost) *nghost = ng;
        return;
    }
    const double wt[4] = {0.0, 1.0, sqrt(2.0), sqrt(3.0)};
    auto tis = [&](int x, int y, int z) -> int {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return 0;
        return g[x * SX + y * SY + z] > 0;
    };
    int64_t ng = 0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                if (g[x * SX + y * SY + z] != 0) continue;
                int gg[27];
                int nb = 0;
                for (int n = 1; n <= 26; ++n) nb += (gg[n] = tis(x + NXo[n], y + NYo[n], z + NZo[n]));
                if (!nb) continue;
                auto sum = [&](const Term* T) {
                    double a = 0.0;
                    for (int i = 0; i < 18; ++i) {
                        const int n = T[i].n > 0 ? T[i].n : -T[i].n;
                        const double term = T[i].w == 1 ? (double)gg[n] : gg[n] / wt[T[i].w];
                        a = T[i].n > 0 ? a + term : a - term;
                    }
                    return a;
                };
                const double ic = sum(IC), ir = sum(IR), il = sum(IL);
                double imax = fabs(ic);
                if (fabs(ir) > imax) imax = fabs(ir);
                if (fabs(il) > imax) imax = fabs(il);
                double nxv = il / imax, nyv = ir / imax, nzv = ic / imax;
                if (imax - 0 < 0.0001) nxv = nyv = nzv = 0;
                double best = 1000;
                int bx = 0, by = 0, bz = 0;
                for (int oz = -1; oz <= 1; ++oz)
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int ox = -1; ox <= 1; ++ox) {
                            const double dis = sqrt(1.0 * (ox - nxv) * (ox - nxv) + (oy - nyv) * (oy - nyv) +
                                                    (oz - nzv) * (oz - nzv));
                            if (tis(x + ox, y + oy, z + oz) && dis < best) {
                                best = dis;
                                bx = ox;
                                by = oy;
                                bz = oz;
                            }
                        }
                g[x * SX + y * SY + z] = -g[(x + bx) * SX + (y + by) * SY + (z + bz)];
                ++ng;
            }
    if (nvox) *nvox = num;
    if (nghost) *nghost = ng;
}

// Padded destination / source grid indices of every ghost cell, and the
// tissue voxel each one mirrors.
static void build_ghosts(const lqt_tissue* t, const std::vector<int32_t>& g, std::vector<int32_t>& dst,
                         std::vector<int32_t>& src, std::vector<int32_t>& srck) {
    const int64_t ncell = (int64_t)t->nx * t->ny * t->nz;
    for (int64_t c = 0; c < ncell; ++c) {
        if (g[c] >= 0) continue;
        const int x = (int)(c / ((int64_t)t->ny * t->nz)), y = (int)((c / t->nz) % t->ny), z = (int)(c % t->nz);
        dst.push_back((int32_t)(((int64_t)(x + 1) * t->py + (y + 1)) * t->pz + (z + 1)));
        src.push_back(t->h_vpos[-g[c] - 1]);
        srck.push_back(-g[c] - 1);
    }
}

// The step kernel refreshes the ghosts itself: voxel k writes its new V to
// the ghost cells that mirror it.  Their destinations are listed by voxel
// (gout); voxel k finds its first one at gwave[k/64] plus the ghosts of the
// lower lanes of its wavefront (a prefix sum over the counts that bits 2-6
// of cls[k] carry, at most 26 for the 26 neighbours).
static void ghosts_by_voxel(int64_t nvox, const std::vector<int32_t>& dst, const std::vector<int32_t>& srck,
                            std::vector<uint8_t>& cl, std::vector<int32_t>& gout, std::vector<int32_t>& gwave) {
    std::vector<int32_t> cnt(nvox + 1, 0);
    for (int32_t k : srck) cnt[k + 1] += 1;
    for (int64_t k = 0; k < nvox; ++k) cl[k] = (uint8_t)(cl[k] | (cnt[k + 1] << 2));
    for (int64_t k = 0; k < nvox; ++k) cnt[k + 1] += cnt[k];  // cnt[k] = first ghost of voxel k
    gout.assign(dst.size(), 0);
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (size_t i = 0; i < dst.size(); ++i) gout[fill[srck[i]]++] = dst[i];
    const int64_t nw = (nvox + 63) / 64;
    gwave.resize(nw);
    for (int64_t w = 0; w < nw; ++w) gwave[w] = cnt[64 * w];
}

template <class R>
static void fill_args(lqt_tissue* t, TissueArgs<R>& a) {
    a.nvox = t->nvox;
    a.vpos = t->vpos;
    a.cls = t->cls;
    a.vold = (const R*)t->grid[t->cur];
    a.vnew = (R*)t->grid[t->cur ^ 1];
    a.sx = t->py * t->pz;
    a.sy = t->pz;
    a.fd = (const R*)t->fd;
    a.dcross = (R)t->d_cross;
    a.ddiff = (R)(t->d_fibre - t->d_cross);
    a.inv_h2 = (R)(1.0 / (t->dx * t->dx));
    a.inv_4h2 = (R)(1.0 / (4.0 * t->dx * t->dx));
    a.inv_2h = (R)(1.0 / (2.0 * t->dx));
    a.sr = (R*)t->r;
    a.sd = (typename SlowStore<R>::type*)t->d;
    for (int j = 0; j < ND; ++j) a.sref.x[j] = t->dref[j];
    a.consts = (const Consts<R>*)t->consts;
    a.dt = (R)t->dt;
    a.stim_on = 0;
    a.tc = 0;
    a.stim_dur = t->proto.duration;
    a.stim_strength = (R)t->proto.strength;
    a.stim_off = t->stim_off;
    a.s2_on = 0;
    a.s2_strength = (R)t->proto.s2_strength;
    a.s2 = t->s2;
    a.guard = t->guard;
    a.gout = t->gout;
    a.gwave = t->gwave;
}

static lqt_status tissue_upload_state(lqt_tissue* t, const std::vector<double>& s40);

extern "C" {

lqt_status lqt_tissue_create(const lqt_geometry* geo, const lqt_params* params, lqt_precision prec, double dt,
                             lqt_tissue** out) {
    LQT_CHECK(geo && params && out && geo->label, LQT_ERR_ARG, "lqt_tissue_create: null argument");
    LQT_CHECK(geo->nx > 0 && geo->ny > 0 && geo->nz > 0 && geo->dx > 0 && dt > 0, LQT_ERR_ARG,
              "lqt_tissue_create: bad dimensions or step");
    LQT_CHECK(prec == LQT_FP64 || dt <= LQT_FP32_MAX_DT, LQT_ERR_ARG, "lqt_tissue_create: FP32 needs dt <= %g ms",
              LQT_FP32_MAX_DT);
    const int64_t ncell = (int64_t)geo->nx * geo->ny * geo->nz;
    const int64_t npad = (int64_t)(geo->nx + 2) * (geo->ny + 2) * (geo->nz + 2);
    LQT_CHECK(npad < (int64_t)INT32_MAX, LQT_ERR_ARG, "lqt_tissue_create: grid too large for 32-bit indices");
    // the step kernel addresses the V grid by 32-bit byte offsets
    LQT_CHECK(npad * (prec == LQT_FP32 ? 4 : 8) < ((int64_t)1 << 32), LQT_ERR_ARG,
              "lqt_tissue_create: padded grid of %lld cells exceeds the %s limit of 2^%d", (long long)npad,
              prec == LQT_FP32 ? "FP32" : "FP64", prec == LQT_FP32 ? 30 : 29);
    int ndev = 0;
    LQT_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, LQT_ERR_HIP,
              "lqt_tissue_create: no HIP device available");
    lqt_tissue* t = new lqt_tissue();
    t->nx = geo->nx;
    t->ny = geo->ny;
    t->nz = geo->nz;
    t->px = geo->nx + 2;
    t->py = geo->ny + 2;
    t->pz = geo->nz + 2;
    t->dx = geo->dx;
    t->dt = dt;
    t->fp32 = prec == LQT_FP32;
    t->d_fibre = geo->d_fibre;
    t->d_cross = geo->d_cross;
    t->aniso = geo->fibre != nullptr && geo->d_fibre != geo->d_cross;
    t->params = *params;
    t->ncell_pad = npad;
    // compaction in (x, y, z) order (GPU3D.cu:1307-1330)
    t->h_gmap.assign(ncell, 0);
    for (int64_t c = 0; c < ncell; ++c) {
        if (geo->label[c] != LQT_NONE) {
            t->h_cell.push_back(c);
            t->h_gmap[c] = (int32_t)t->h_cell.size();
        }
    }
    t->nvox = (int64_t)t->h_cell.size();
    if (t->nvox == 0) {
        delete t;
        return set_error(LQT_ERR_ARG, "lqt_tissue_create: geometry has no tissue voxel");
    }
    // voxel slots per array: whole 64-voxel chunks, rounded up to 256
    t->pitch = (t->nvox + 255) / 256 * 256;
    const int64_t SX = (int64_t)t->ny * t->nz, SY = t->nz;
    t->h_vpos.resize(t->nvox);
    t->h_label.resize(t->nvox);
    std::vector<uint8_t> cl(t->nvox);
    for (int64_t k = 0; k < t->nvox; ++k) {
        const int64_t c = t->h_cell[k];
        const int x = (int)(c / SX), y = (int)((c / SY) % t->ny), z = (int)(c % t->nz);
        t->h_vpos[k] = (int32_t)(((int64_t)(x + 1) * t->py + (y + 1)) * t->pz + (z + 1));
        t->h_label[k] = geo->label[c];
        cl[k] = (uint8_t)(cell_class(geo->label[c] & 0x7f) | (geo->label[c] & LQT_HOLD));
    }
    std::vector<int32_t> gsigned, gd, gs, gk, gout, gwave;
    ghost_map(t->nx, t->ny, t->nz, geo->label, gsigned, nullptr, nullptr);
    build_ghosts(t, gsigned, gd, gs, gk);
    gsigned.clear();
    gsigned.shrink_to_fit();
    t->nghost = (int64_t)gd.size();
    ghosts_by_voxel(t->nvox, gd, gk, cl, gout, gwave);

    // tensor divergence b_j = sum_i d_i D_ij with the one-sided rule of
    // new_11.cu:381-407 (tissue neighbours only)
    std::vector<double> fibre_h, drift_h;
    if (t->aniso) {
        const double DD = geo->d_fibre - geo->d_cross;
        std::vector<double> D((size_t)t->nvox * 6);
        fibre_h.resize((size_t)t->nvox * 3);
        for (int64_t k = 0; k < t->nvox; ++k) {
            const float* f = geo->fibre + 3 * t->h_cell[k];
            const double f0 = f[0], f1 = f[1], f2 = f[2];
            fibre_h[3 * k] = f0;
            fibre_h[3 * k + 1] = f1;
            fibre_h[3 * k + 2] = f2;
            double* Dk = &D[6 * k];
            Dk[0] = geo->d_cross + DD * f0 * f0;
            Dk[1] = geo->d_cross + DD * f1 * f1;
            Dk[2] = geo->d_cross + DD * f2 * f2;
            Dk[3] = DD * f0 * f1;
            Dk[4] = DD * f0 * f2;
            Dk[5] = DD * f1 * f2;
        }
        static const int comp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
        const int64_t strides[3] = {SX, SY, 1};
        const int dims[3] = {t->nx, t->ny, t->nz};
        drift_h.assign((size_t)t->nvox * 3, 0.0);
        for (int64_t k = 0; k < t->nvox; ++k) {
            const int64_t c = t->h_cell[k];
            const int co[3] = {(int)(c / SX), (int)((c / SY) % t->ny), (int)(c % t->nz)};
            for (int i = 0; i < 3; ++i) {
                const int32_t gp = co[i] + 1 < dims[i] ? t->h_gmap[c + strides[i]] : 0;
                const int32_t gm = co[i] - 1 >= 0 ? t->h_gmap[c - strides[i]] : 0;
                for (int j = 0; j < 3; ++j) {
                    const int cj = comp[i][j];
                    double dv = 0;
                    if (gp > 0 && gm > 0) dv = (D[6 * (gp - 1) + cj] - D[6 * (gm - 1) + cj]) / (2 * t->dx);
                    else if (gp > 0) dv = (D[6 * (gp - 1) + cj] - D[6 * k + cj]) / t->dx;
                    else if (gm > 0) dv = (D[6 * k + cj] - D[6 * (gm - 1) + cj]) / t->dx;
                    drift_h[3 * k + j] += dv;
                }
            }
        }
    }

    lqt_status st = [&]() -> lqt_status {
        LQT_HIP(hipStreamCreate(&t->stream));
        const size_t rs = rsz(t);
        lqt_status s2;
        if ((s2 = upload(&t->vpos, t->h_vpos.data(), t->nvox))) return s2;
        if ((s2 = upload(&t->cls, cl.data(), t->nvox))) return s2;
        if ((s2 = upload(&t->gdst, gd.data(), gd.size()))) return s2;
        if ((s2 = upload(&t->gsrc, gs.data(), gs.size()))) return s2;
        if ((s2 = upload(&t->gout, gout.data(), gout.size()))) return s2;
        if ((s2 = upload(&t->gwave, gwave.data(), gwave.size()))) return s2;
        LQT_HIP(hipMalloc(&t->grid[0], rs * npad));
        LQT_HIP(hipMalloc(&t->grid[1], rs * npad));
        LQT_HIP(hipMalloc(&t->r, rs * NR * t->pitch));
        LQT_HIP(hipMalloc(&t->d, dsz(t) * ND * t->pitch));
        LQT_HIP(hipMalloc((void**)&t->guard, sizeof(unsigned long long) * 3));
        LQT_HIP(hipMemset(t->guard, 0, sizeof(unsigned long long) * 3));
        if (t->aniso) {
            // chunk-major: slots 0-2 fibre f, 3-5 divergence of D
            std::vector<unsigned char> fb(rs * 6 * t->pitch, 0);
            for (int64_t k = 0; k < t->nvox; ++k)
                for (int j = 0; j < 6; ++j) {
                    const int64_t idx = chunk_slot(k, j, 6);
                    const double x = j < 3 ? fibre_h[3 * k + j] : drift_h[3 * k + j - 3];
                    if (t->fp32) ((float*)fb.data())[idx] = (float)x;
                    else ((double*)fb.data())[idx] = x;
                }
            if ((s2 = upload((unsigned char**)&t->fd, fb.data(), fb.size()))) return s2;
        }
        const Consts<double> cd = derive_consts(*params);
        if (t->fp32) {
            const Consts<float> cf = to_float(cd);
            if ((s2 = upload((Consts<float>**)&t->consts, &cf, 1))) return s2;
        } else {
            if ((s2 = upload((Consts<double>**)&t->consts, &cd, 1))) return s2;
        }
        std::vector<double> s40((size_t)t->nvox * 40);
        for (int64_t k = 0; k < t->nvox; ++k) {
            memcpy(&s40[(size_t)k * 40], kInit40, sizeof(kInit40));
            s40[(size_t)k * 40] = params->v0;
        }
        return tissue_upload_state(t, s40);
    }();
    if (st) {
        lqt_tissue_destroy(t);
        return st;
    }
    *out = t;
    return LQT_OK;
}

lqt_status lqt_tissue_destroy(lqt_tissue* t) {
    if (!t) return LQT_OK;
    if (t->stream) hipStreamSynchronize(t->stream);
    hipFree(t->vpos);
    hipFree(t->cls);
    hipFree(t->grid[0]);
    hipFree(t->grid[1]);
    hipFree(t->r);
    hipFree(t->d);
    hipFree(t->fd);
    hipFree(t->gdst);
    hipFree(t->gsrc);
    hipFree(t->gout);
    hipFree(t->gwave);
    hipFree(t->consts);
    hipFree(t->stim_off);
    hipFree(t->s2);
    hipFree(t->guard);
    hipFree(t->track);
    for (hipEvent_t e : t->ev) hipEventDestroy(e);
    if (t->stream) hipStreamDestroy(t->stream);
    delete t;
    return LQT_OK;
}

lqt_status lqt_geometry_ghost_map(int nx, int ny, int nz, const uint8_t* label, int32_t* g, int64_t* nvox,
                                  int64_t* nghost) {
    LQT_CHECK(nx > 0 && ny > 0 && nz > 0 && label && g, LQT_ERR_ARG, "lqt_geometry_ghost_map: bad argument");
    std::vector<int32_t> m;
    ghost_map(nx, ny, nz, label, m, nvox, nghost);
    memcpy(g, m.data(), sizeof(int32_t) * m.size());
    return LQT_OK;
}

lqt_status lqt_tissue_info(const lqt_tissue* t, int64_t* nvox, int64_t* nghost) {
    LQT_CHECK(t, LQT_ERR_ARG, "lqt_tissue_info: null handle");
    if (nvox) *nvox = t->nvox;
    if (nghost) *nghost = t->nghost;
    return LQT_OK;
}

lqt_status lqt_tissue_set_stimulus(lqt_tissue* t, const lqt_stim_protocol* p, int64_t nsites,
                                   const int64_t* site_cells, const float* site_offset_ms, int64_t ns2,
                                   const int64_t* s2_cells) {
    LQT_CHECK(t && p && p->bcl > 0 && p->duration >= 0 && p->window >= 0, LQT_ERR_ARG,
              "lqt_tissue_set_stimulus: bad protocol");
    LQT_CHECK(nsites >= 0 && ns2 >= 0 && (nsites == 0 || site_cells) && (ns2 == 0 || s2_cells), LQT_ERR_ARG,
              "lqt_tissue_set_stimulus: bad site arrays");
    const int64_t ncell = (int64_t)t->nx * t->ny * t->nz;
    std::vector<float> off(t->nvox, NAN);
    std::vector<uint8_t> s2(t->nvox, 0);
    double mn = 1e300, mx = -1e300;
    int64_t used = 0;
    for (int64_t i = 0; i < nsites; ++i) {
        const int64_t c = site_cells[i];
        LQT_CHECK(c >= 0 && c < ncell, LQT_ERR_ARG, "stimulus site %lld outside the grid", (long long)c);
        const int32_t g = t->h_gmap[c];
        if (g <= 0) continue;  // GPU3D.cu:1704: sites outside tissue are skipped
        const float o = site_offset_ms ? site_offset_ms[i] : 0.0f;
        off[g - 1] = o;
        mn = std::min(mn, (double)o);
        mx = std::max(mx, (double)o);
        ++used;
    }
    int64_t used2 = 0;
    for (int64_t i = 0; i < ns2; ++i) {
        const int64_t c = s2_cells[i];
        LQT_CHECK(c >= 0 && c < ncell, LQT_ERR_ARG, "S2 site %lld outside the grid", (long long)c);
        const int32_t g = t->h_gmap[c];
        if (g <= 0) continue;
        s2[g - 1] = 1;
        ++used2;
    }
    hipFree(t->stim_off);
    hipFree(t->s2);
    t->stim_off = nullptr;
    t->s2 = nullptr;
    lqt_status s;
    if ((s = upload(&t->stim_off, off.data(), off.size()))) return s;
    if ((s = upload(&t->s2, s2.data(), s2.size()))) return s;
    t->proto = *p;
    t->nsites = used;
    t->ns2 = used2;
    t->off_min = used ? mn : 0;
    t->off_max = used ? mx : 0;
    return LQT_OK;
}

// The stimulation profile's sites inside an nx x ny x nz grid, in file
// order, with their offsets in ms (host only).
static lqt_status read_stimulus(const char* path, int nx, int ny, int nz, std::vector<int64_t>& cells,
                                std::vector<float>& offs) {
    FILE* f = fopen(path, "r");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    int x, y, z;
    float a4, a5, a6, tact;
    char tag[64];
    // "x y z f f f t_act str" (GPU3D.cu:1700); t_act in seconds
    while (fscanf(f, "%d %d %d %f %f %f %f %63s", &x, &y, &z, &a4, &a5, &a6, &tact, tag) >= 7) {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) continue;
        cells.push_back(((int64_t)x * ny + y) * nz + z);
        offs.push_back(1000.0f * tact);
    }
    fclose(f);
    return LQT_OK;
}

lqt_status lqt_stimulus_read_file(const char* path, int nx, int ny, int nz, int64_t cap, int64_t* cells,
                                  float* offsets_ms, int64_t* n) {
    LQT_CHECK(path && n && nx > 0 && ny > 0 && nz > 0 && cap >= 0, LQT_ERR_ARG, "lqt_stimulus_read_file: bad argument");
    std::vector<int64_t> c;
    std::vector<float> o;
    lqt_status s = read_stimulus(path, nx, ny, nz, c, o);
    if (s) return s;
    *n = (int64_t)c.size();
    if (cells || offsets_ms) {
        LQT_CHECK((int64_t)c.size() <= cap, LQT_ERR_ARG, "lqt_stimulus_read_file: %s holds %lld sites, cap is %lld",
                  path, (long long)c.size(), (long long)cap);
        for (size_t i = 0; i < c.size(); ++i) {
            if (cells) cells[i] = c[i];
            if (offsets_ms) offsets_ms[i] = o[i];
        }
    }
    return LQT_OK;
}

lqt_status lqt_tissue_set_stimulus_file(lqt_tissue* t, const lqt_stim_protocol* p, const char* path) {
    LQT_CHECK(t && p && path, LQT_ERR_ARG, "lqt_tissue_set_stimulus_file: null argument");
    std::vector<int64_t> cells;
    std::vector<float> offs;
    lqt_status s = read_stimulus(path, t->nx, t->ny, t->nz, cells, offs);
    if (s) return s;
    return lqt_tissue_set_stimulus(t, p, (int64_t)cells.size(), cells.data(), offs.data(), 0, nullptr);
}

// nsteps steps enqueued on stream s.  With `timed`, one pair of events on s
// brackets all of them (the mean device time per step is read after the
// caller's synchronisation); without it nothing waits for the device.
static lqt_status step_enqueue(lqt_tissue* t, int64_t nsteps, hipStream_t s, bool timed) {
    timed = timed && nsteps > 0;
    if (timed) {
        while (t->ev.size() < 2) {
            hipEvent_t e;
            LQT_HIP(hipEventCreate(&e));
            t->ev.push_back(e);
        }
        LQT_HIP(hipEventRecord(t->ev[0], s));
    }
    for (int64_t i = 0; i < nsteps; ++i) {
        const double tc = t->time - floor(t->time / t->proto.bcl) * t->proto.bcl;
        const int stim_on = t->nsites > 0 && tc >= t->off_min && tc < t->off_max + t->proto.duration &&
                            (t->proto.window <= 0 || tc < t->proto.window);
        const int s2_on = t->ns2 > 0 && t->proto.s2_duration > 0 && t->time >= t->proto.s2_time &&
                          t->time < t->proto.s2_time + t->proto.s2_duration;
        if (t->fp32) {
            TissueArgs<float> a;
            fill_args(t, a);
            a.stim_on = stim_on;
            a.tc = tc;
            a.s2_on = s2_on;
            launch_tissue_step(true, t->aniso, t->nz == 1, t->variant, &a, t->nvox, s);
        } else {
            TissueArgs<double> a;
            fill_args(t, a);
            a.stim_on = stim_on;
            a.tc = tc;
            a.s2_on = s2_on;
            launch_tissue_step(false, t->aniso, t->nz == 1, t->variant, &a, t->nvox, s);
        }
        if (t->trk_on && t->time >= t->trk_from && t->time < t->trk_to)
            launch_track(t->fp32, t->nvox, t->vpos, t->grid[t->cur ^ 1], t->time, t->track, s);
        LQT_HIP(hipGetLastError());
        t->cur ^= 1;
        t->time += t->dt;  // GPU3D.cu:1840 double accumulator
        t->step += 1;
    }
    if (timed) LQT_HIP(hipEventRecord(t->ev[1], s));
    return LQT_OK;
}

lqt_status lqt_tissue_step_async(lqt_tissue* t, int64_t nsteps, void* stream) {
    LQT_CHECK(t && nsteps >= 0, LQT_ERR_ARG, "lqt_tissue_step_async: bad argument");
    return step_enqueue(t, nsteps, stream ? (hipStream_t)stream : t->stream, false);
}

lqt_status lqt_tissue_step(lqt_tissue* t, int64_t nsteps) {
    LQT_CHECK(t && nsteps >= 0, LQT_ERR_ARG, "lqt_tissue_step: bad argument");
    lqt_status s = step_enqueue(t, nsteps, t->stream, true);
    if (s) return s;
    LQT_HIP(hipStreamSynchronize(t->stream));
    if (nsteps > 0) {
        float ms = 0;
        LQT_HIP(hipEventElapsedTime(&ms, t->ev[0], t->ev[1]));
        t->us_step = 1000.0 * ms / nsteps;
        t->us_ghost = 0.0;  // the step kernel refreshes the ghosts itself
    }
    unsigned long long g[3];
    LQT_HIP(hipMemcpy(g, t->guard, sizeof(g), hipMemcpyDeviceToHost));
    if (g[2]) return set_error(LQT_ERR_RANGE, "non-finite V in %llu voxel-steps (t = %.3f ms)", g[2], t->time);
    return LQT_OK;
}

lqt_status lqt_tissue_time(const lqt_tissue* t, double* time_ms, int64_t* step) {
    LQT_CHECK(t, LQT_ERR_ARG, "lqt_tissue_time: null handle");
    if (time_ms) *time_ms = t->time;
    if (step) *step = t->step;
    return LQT_OK;
}

lqt_status lqt_tissue_read_v(lqt_tissue* t, double* out) {
    LQT_CHECK(t && out, LQT_ERR_ARG, "lqt_tissue_read_v: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    const size_t rs = rsz(t);
    std::vector<unsigned char> g(rs * t->ncell_pad);
    LQT_HIP(hipMemcpy(g.data(), t->grid[t->cur], g.size(), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < t->nvox; ++k) {
        const int64_t p = t->h_vpos[k];
        out[k] = t->fp32 ? (double)((float*)g.data())[p] : ((double*)g.data())[p];
    }
    return LQT_OK;
}

lqt_status lqt_tissue_read_state(lqt_tissue* t, int64_t k, double* out) {
    LQT_CHECK(t && out && k >= 0 && k < t->nvox, LQT_ERR_ARG, "lqt_tissue_read_state: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    const size_t rs = rsz(t);
    for (int j = 0; j < 40; ++j) {
        const Ref40 m = kRef40[j];
        double val = 1.0;
        unsigned char buf[8];
        if (m.kind == 0) {
            LQT_HIP(hipMemcpy(buf, (unsigned char*)t->grid[t->cur] + rs * t->h_vpos[k], rs, hipMemcpyDeviceToHost));
            val = t->fp32 ? (double)*(float*)buf : *(double*)buf;
        } else if (m.kind == 1) {
            LQT_HIP(hipMemcpy(buf, (unsigned char*)t->r + rs * chunk_slot(k, m.slot, NR), rs, hipMemcpyDeviceToHost));
            val = t->fp32 ? (double)*(float*)buf : *(double*)buf;
        } else if (m.kind == 2) {
            LQT_HIP(hipMemcpy(buf, (unsigned char*)t->d + dsz(t) * chunk_slot(k, m.slot, ND), dsz(t),
                              hipMemcpyDeviceToHost));
            val = slow_get(t, m.slot, buf);
        }
        out[j] = val;
    }
    return LQT_OK;
}

lqt_status lqt_tissue_write_state(lqt_tissue* t, int64_t k, const double* in) {
    LQT_CHECK(t && in && k >= 0 && k < t->nvox, LQT_ERR_ARG, "lqt_tissue_write_state: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    const size_t rs = rsz(t);
    for (int j = 0; j < 40; ++j) {
        const Ref40 m = kRef40[j];
        unsigned char buf[8];
        if (t->fp32) *(float*)buf = (float)in[j];
        else *(double*)buf = in[j];
        if (m.kind == 0) {
            LQT_HIP(hipMemcpy((unsigned char*)t->grid[t->cur] + rs * t->h_vpos[k], buf, rs, hipMemcpyHostToDevice));
        } else if (m.kind == 1) {
            LQT_HIP(hipMemcpy((unsigned char*)t->r + rs * chunk_slot(k, m.slot, NR), buf, rs, hipMemcpyHostToDevice));
        } else if (m.kind == 2) {
            slow_put(t, m.slot, in[j], buf);
            LQT_HIP(hipMemcpy((unsigned char*)t->d + dsz(t) * chunk_slot(k, m.slot, ND), buf, dsz(t),
                              hipMemcpyHostToDevice));
        }
    }
    // refresh the ghost mirrors of the current grid
    launch_ghost_fill(t->fp32, t->grid[t->cur], t->gdst, t->gsrc, t->nghost, t->stream);
    LQT_HIP(hipGetLastError());
    LQT_HIP(hipStreamSynchronize(t->stream));
    return LQT_OK;
}

lqt_status lqt_tissue_v_guard(lqt_tissue* t, int64_t counts[3]) {
    LQT_CHECK(t && counts, LQT_ERR_ARG, "lqt_tissue_v_guard: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    unsigned long long g[3];
    LQT_HIP(hipMemcpy(g, t->guard, sizeof(g), hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; ++i) counts[i] = (int64_t)g[i];
    return LQT_OK;
}

lqt_status lqt_tissue_variants(int* count) {
    LQT_CHECK(count, LQT_ERR_ARG, "lqt_tissue_variants: null argument");
    *count = kNumTissueVariants;
    return LQT_OK;
}

lqt_status lqt_tissue_variant_info(int i, int* block, int* xcd_map, int* waves, int* lds_stage) {
    LQT_CHECK(i >= 0 && i < kNumTissueVariants, LQT_ERR_ARG, "lqt_tissue_variant_info: no variant %d", i);
    if (block) *block = kTissueVariants[i].block;
    if (xcd_map) *xcd_map = kTissueVariants[i].xcd;
    if (waves) *waves = kTissueVariants[i].waves;
    if (lds_stage) *lds_stage = kTissueVariants[i].lds;
    return LQT_OK;
}

lqt_status lqt_tissue_set_variant(lqt_tissue* t, int i) {
    LQT_CHECK(t && i >= 0 && i < kNumTissueVariants, LQT_ERR_ARG, "lqt_tissue_set_variant: bad argument");
    t->variant = i;
    return LQT_OK;
}

lqt_status lqt_tissue_get_variant(const lqt_tissue* t, int* i) {
    LQT_CHECK(t && i, LQT_ERR_ARG, "lqt_tissue_get_variant: bad argument");
    *i = t->variant;
    return LQT_OK;
}

// Every launch shape runs trial_steps steps from the same state; the state,
// V grids, guard counters and clock are copied aside first and put back
// after each trial, so the simulation is left exactly where it was.
lqt_status lqt_tissue_autotune(lqt_tissue* t, int trial_steps, int* best, double* us) {
    LQT_CHECK(t && trial_steps > 0, LQT_ERR_ARG, "lqt_tissue_autotune: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    const size_t rs = rsz(t);
    const size_t nr = rs * NR * t->pitch, nd = dsz(t) * ND * t->pitch, ng = rs * t->ncell_pad;
    void *sr = nullptr, *sd = nullptr, *sg0 = nullptr, *sg1 = nullptr, *sgu = nullptr;
    auto release = [&]() {
        hipFree(sr);
        hipFree(sd);
        hipFree(sg0);
        hipFree(sg1);
        hipFree(sgu);
    };
    if (hipMalloc(&sr, nr) != hipSuccess || hipMalloc(&sd, nd) != hipSuccess || hipMalloc(&sg0, ng) != hipSuccess ||
        hipMalloc(&sg1, ng) != hipSuccess || hipMalloc(&sgu, 3 * sizeof(unsigned long long)) != hipSuccess) {
        release();
        return set_error(LQT_ERR_HIP, "lqt_tissue_autotune: no device memory for the state snapshot");
    }
    const double time0 = t->time;
    const int64_t step0 = t->step;
    const int cur0 = t->cur, keep = t->variant;
    auto copy_all = [&](bool save) -> hipError_t {
        hipError_t e = hipSuccess;
        void* pairs[5][2] = {{sr, t->r}, {sd, t->d}, {sg0, t->grid[0]}, {sg1, t->grid[1]}, {sgu, t->guard}};
        const size_t sz[5] = {nr, nd, ng, ng, 3 * sizeof(unsigned long long)};
        for (int i = 0; i < 5 && e == hipSuccess; ++i)
            e = save ? hipMemcpy(pairs[i][0], pairs[i][1], sz[i], hipMemcpyDeviceToDevice)
                     : hipMemcpy(pairs[i][1], pairs[i][0], sz[i], hipMemcpyDeviceToDevice);
        return e;
    };
    if (copy_all(true) != hipSuccess) {
        release();
        return set_error(LQT_ERR_HIP, "lqt_tissue_autotune: snapshot copy failed");
    }
    // Each shape runs trial_steps steps from the snapshot, timed as a whole
    // by one event pair; the shapes run twice, in opposite orders, after a
    // warm-up, so that the clocks' ramp and drift do not favour any of them.
    // The trackers are held off: the trials are not part of the simulation.
    const bool trk = t->trk_on;
    t->trk_on = false;
    hipEvent_t ea = nullptr, eb = nullptr;
    lqt_status st = LQT_OK;
    if (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess)
        st = set_error(LQT_ERR_HIP, "lqt_tissue_autotune: event creation failed");
    std::vector<double> acc(kNumTissueVariants, 0.0);
    auto restore = [&]() {
        t->time = time0;
        t->step = step0;
        t->cur = cur0;
        if (copy_all(false) != hipSuccess && !st) st = set_error(LQT_ERR_HIP, "lqt_tissue_autotune: restore failed");
    };
    if (!st) {
        t->variant = keep;
        st = lqt_tissue_step(t, trial_steps);  // warm-up
        restore();
    }
    for (int pass = 0; pass < 2 && !st; ++pass) {
        for (int i = 0; i < kNumTissueVariants && !st; ++i) {
            const int v = pass == 0 ? i : kNumTissueVariants - 1 - i;
            t->variant = v;
            st = lqt_tissue_step(t, 1);  // first launch of this shape, untimed
            if (!st && hipEventRecord(ea, t->stream) != hipSuccess) st = set_error(LQT_ERR_HIP, "event record failed");
            if (!st) st = lqt_tissue_step_async(t, trial_steps, nullptr);
            if (!st && (hipEventRecord(eb, t->stream) != hipSuccess || hipEventSynchronize(eb) != hipSuccess))
                st = set_error(LQT_ERR_HIP, "lqt_tissue_autotune: timing failed");
            float ms = 0.f;
            if (!st && hipEventElapsedTime(&ms, ea, eb) != hipSuccess)
                st = set_error(LQT_ERR_HIP, "lqt_tissue_autotune: timing failed");
            acc[v] += 1000.0 * ms / trial_steps;
            restore();
        }
    }
    if (ea) hipEventDestroy(ea);
    if (eb) hipEventDestroy(eb);
    t->trk_on = trk;
    int ibest = keep;
    for (int v = 0; v < kNumTissueVariants; ++v) {
        acc[v] *= 0.5;
        if (us) us[v] = acc[v];
        if (acc[v] < acc[ibest]) ibest = v;
    }
    release();
    t->variant = st ? keep : ibest;
    if (best) *best = t->variant;
    return st;
}

lqt_status lqt_tissue_track(lqt_tissue* t, double t_from, double t_to) {
    LQT_CHECK(t, LQT_ERR_ARG, "lqt_tissue_track: null handle");
    LQT_HIP(hipStreamSynchronize(t->stream));
    if (!(t_to > t_from)) {
        t->trk_on = false;
        return LQT_OK;
    }
    if (!t->track) LQT_HIP(hipMalloc((void**)&t->track, sizeof(VoxTrack) * std::max<int64_t>(1, t->nvox)));
    launch_track_reset(t->nvox, t->track, t->stream);
    LQT_HIP(hipGetLastError());
    LQT_HIP(hipStreamSynchronize(t->stream));
    t->trk_on = true;
    t->trk_from = t_from;
    t->trk_to = t_to;
    return LQT_OK;
}

lqt_status lqt_tissue_track_results(lqt_tissue* t, double* act, double* apd90) {
    LQT_CHECK(t && (act || apd90), LQT_ERR_ARG, "lqt_tissue_track_results: bad argument");
    LQT_CHECK(t->track, LQT_ERR_STATE, "lqt_tissue_track_results: call lqt_tissue_track first");
    LQT_HIP(hipStreamSynchronize(t->stream));
    std::vector<VoxTrack> h(t->nvox);
    LQT_HIP(hipMemcpy(h.data(), t->track, sizeof(VoxTrack) * t->nvox, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < t->nvox; ++k) {
        const VoxTrack& v = h[k];
        if (act) act[k] = (v.flags & 2) ? v.act : -1.0;
        if (apd90) apd90[k] = (v.flags & 4) ? v.rep - v.act : -1.0;
    }
    return LQT_OK;
}

lqt_status lqt_tissue_last_kernel_us(const lqt_tissue* t, double* us_step_kernel, double* us_ghost_kernel) {
    LQT_CHECK(t, LQT_ERR_ARG, "lqt_tissue_last_kernel_us: null handle");
    if (us_step_kernel) *us_step_kernel = t->us_step;
    if (us_ghost_kernel) *us_ghost_kernel = t->us_ghost;
    return LQT_OK;
}

lqt_status lqt_tissue_bytes_per_step(const lqt_tissue* t, double* bytes) {
    LQT_CHECK(t && bytes, LQT_ERR_ARG, "lqt_tissue_bytes_per_step: bad argument");
    const double rs = (double)rsz(t);
    // per voxel: V read + write, NR R-slots and ND slow slots read + written,
    // class byte and grid index, fibre + drift (anisotropic)
    double per = 2.0 * (rs + NR * rs + ND * (double)dsz(t)) + 1.0 + 4.0;
    if (t->aniso) per += (t->nz == 1 ? 4.0 : 6.0) * rs;  // fibre + drift (x, y only in 2D)
    // plus the ghost refresh: destination index read and V written per
    // ghost cell, one list offset per wavefront
    *bytes = per * (double)t->nvox + (4.0 + rs) * (double)t->nghost + 4.0 * (double)((t->nvox + 63) / 64);
    return LQT_OK;
}

}  // extern "C"

static lqt_status tissue_upload_state(lqt_tissue* t, const std::vector<double>& s40) {
    const size_t rs = rsz(t);
    std::vector<unsigned char> r(rs * NR * t->pitch, 0);
    std::vector<unsigned char> d(dsz(t) * ND * t->pitch, 0);
    std::vector<unsigned char> g(rs * t->ncell_pad);
    // FP32: the offsets' reference values, midway between each slow slot's
    // smallest and largest value over the voxels (the initial values, when
    // every voxel starts from the same state)
    for (int j = 0; j < ND; ++j) t->dref[j] = 0.0;
    if (t->fp32 && t->nvox > 0) {
        for (int j = 0; j < 40; ++j) {
            const Ref40 m = kRef40[j];
            if (m.kind != 2) continue;
            double lo = s40[j], hi = s40[j];
            for (int64_t k = 1; k < t->nvox; ++k) {
                lo = std::min(lo, s40[(size_t)k * 40 + j]);
                hi = std::max(hi, s40[(size_t)k * 40 + j]);
            }
            t->dref[m.slot] = 0.5 * (lo + hi);
        }
    }
    for (int64_t i = 0; i < t->ncell_pad; ++i) {
        if (t->fp32) ((float*)g.data())[i] = -90.0f;
        else ((double*)g.data())[i] = -90.0;
    }
    for (int64_t k = 0; k < t->nvox; ++k) {
        for (int j = 0; j < 40; ++j) {
            const Ref40 m = kRef40[j];
            const double val = s40[(size_t)k * 40 + j];
            if (m.kind == 0) {
                if (t->fp32) ((float*)g.data())[t->h_vpos[k]] = (float)val;
                else ((double*)g.data())[t->h_vpos[k]] = val;
            } else if (m.kind == 1) {
                const int64_t idx = chunk_slot(k, m.slot, NR);
                if (t->fp32) ((float*)r.data())[idx] = (float)val;
                else ((double*)r.data())[idx] = val;
            } else if (m.kind == 2) {
                slow_put(t, m.slot, val, d.data() + dsz(t) * (size_t)chunk_slot(k, m.slot, ND));
            }
        }
    }
    LQT_HIP(hipMemcpy(t->r, r.data(), r.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->d, d.data(), d.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->grid[t->cur], g.data(), g.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->grid[t->cur ^ 1], g.data(), g.size(), hipMemcpyHostToDevice));
    launch_ghost_fill(t->fp32, t->grid[t->cur], t->gdst, t->gsrc, t->nghost, t->stream);
    launch_ghost_fill(t->fp32, t->grid[t->cur ^ 1], t->gdst, t->gsrc, t->nghost, t->stream);
    LQT_HIP(hipGetLastError());
    LQT_HIP(hipStreamSynchronize(t->stream));
    return LQT_OK;
}

extern "C" {

// Labels 1-20 of the reference's geometry files (GPU3D.cu:865-871 and the
// label tests of Itot): 7-10 and 17-20 are epicardial, 4-6 and 14-16 are
// mid-myocardial (M) cells, every other label is endocardial.
static uint8_t label20_to_celltype(int l) {
    if ((l >= 7 && l <= 10) || (l >= 17 && l <= 20)) return LQT_EPI;
    if ((l >= 4 && l <= 6) || (l >= 14 && l <= 16)) return LQT_MCELL;
    return LQT_ENDO;
}

lqt_status lqt_geometry_read_file(const char* path, int nx, int ny, int nz, uint8_t* label, float* fibre,
                                  int64_t* nvox) {
    LQT_CHECK(path && label && nx > 0 && ny > 0 && nz > 0, LQT_ERR_ARG, "lqt_geometry_read_file: bad argument");
    const int64_t ncell = (int64_t)nx * ny * nz;
    for (int64_t c = 0; c < ncell; ++c) label[c] = LQT_NONE;
    if (fibre)
        for (int64_t c = 0; c < 3 * ncell; ++c) fibre[c] = 0.0f;
    FILE* f = fopen(path, "r");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    int x, y, z, l;
    float f0, f1, f2;
    int64_t nread = 0;
    // one voxel per line: "x y z fx fy fz label" (GPU3D.cu:1276-1283)
    while (fscanf(f, "%d %d %d %f %f %f %d", &x, &y, &z, &f0, &f1, &f2, &l) == 7) {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) {
            fclose(f);
            return set_error(LQT_ERR_ARG, "%s: voxel (%d,%d,%d) outside the %dx%dx%d grid", path, x, y, z, nx, ny, nz);
        }
        const int64_t c = ((int64_t)x * ny + y) * nz + z;
        label[c] = label20_to_celltype(l);
        if (fibre) {
            fibre[3 * c] = f0;
            fibre[3 * c + 1] = f1;
            fibre[3 * c + 2] = f2;
        }
        ++nread;
    }
    fclose(f);
    LQT_CHECK(nread > 0, LQT_ERR_IO, "%s: no voxel read", path);
    if (nvox) {
        int64_t k = 0;
        for (int64_t c = 0; c < ncell; ++c) k += label[c] != LQT_NONE;
        *nvox = k;
    }
    return LQT_OK;
}

lqt_status lqt_tissue_create_from_file(const char* path, int nx, int ny, int nz, double dx, double d_fibre,
                                       double d_cross, const lqt_params* params, lqt_precision prec, double dt,
                                       lqt_tissue** out) {
    LQT_CHECK(path && nx > 0 && ny > 0 && nz > 0, LQT_ERR_ARG, "lqt_tissue_create_from_file: bad argument");
    const int64_t ncell = (int64_t)nx * ny * nz;
    std::vector<uint8_t> label(ncell);
    std::vector<float> fib((size_t)ncell * 3);
    lqt_status s = lqt_geometry_read_file(path, nx, ny, nz, label.data(), fib.data(), nullptr);
    if (s) return s;
    lqt_geometry g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx = dx;
    g.label = label.data();
    g.fibre = fib.data();
    g.d_fibre = d_fibre;
    g.d_cross = d_cross;
    return lqt_tissue_create(&g, params, prec, dt, out);
}

lqt_status lqt_tissue_write_vtk(lqt_tissue* t, const char* path, int stride) {
    LQT_CHECK(t && path && stride > 0, LQT_ERR_ARG, "lqt_tissue_write_vtk: bad argument");
    std::vector<double> v(t->nvox);
    lqt_status s = lqt_tissue_read_v(t, v.data());
    if (s) return s;
    FILE* f = fopen(path, "w");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    const int cx = (t->nx + stride - 1) / stride, cy = (t->ny + stride - 1) / stride, cz = (t->nz + stride - 1) / stride;
    // legacy VTK STRUCTURED_POINTS as GPU3D.cu:123-163 writes it (x fastest,
    // -100 for non-tissue); the DIMENSIONS line gives the counts written
    fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n");
    fprintf(f, "DIMENSIONS %d %d %d\nSPACING 1 1 1\nORIGIN 0 0 0\nPOINT_DATA %lld\n", cx, cy, cz,
            (long long)cx * cy * cz);
    fprintf(f, "SCALARS ImageFile float 1\nLOOKUP_TABLE default\n");
    const int64_t SX = (int64_t)t->ny * t->nz, SY = t->nz;
    for (int z = 0; z < t->nz; z += stride) {
        for (int y = 0; y < t->ny; y += stride) {
            for (int x = 0; x < t->nx; x += stride) {
                const int32_t g = t->h_gmap[x * SX + y * SY + z];
                if (g > 0) fprintf(f, "%g ", v[g - 1]);
                else fprintf(f, "-100 ");
            }
            fprintf(f, "\n");
        }
        fprintf(f, "\n");
    }
    fclose(f);
    return LQT_OK;
}

// Checkpoint: magic, header (dims, voxels, precision, step), time, the slow
// slots' reference values, then the state arrays as stored (chunk-major)
// and the current V grid.  Checkpoints of the earlier format (slow slots
// stored as double in both precisions, no reference values) are read too.
static const char kMagic[8] = {'L', 'Q', 'T', 'T', 'I', 'S', '3', 0};
static const char kMagic2[8] = {'L', 'Q', 'T', 'T', 'I', 'S', '2', 0};

lqt_status lqt_tissue_save(lqt_tissue* t, const char* path) {
    LQT_CHECK(t && path, LQT_ERR_ARG, "lqt_tissue_save: bad argument");
    LQT_HIP(hipStreamSynchronize(t->stream));
    const size_t rs = rsz(t);
    std::vector<unsigned char> r(rs * NR * t->pitch), d(dsz(t) * ND * t->pitch), g(rs * t->ncell_pad);
    LQT_HIP(hipMemcpy(r.data(), t->r, r.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(d.data(), t->d, d.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(g.data(), t->grid[t->cur], g.size(), hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    const int64_t hdr[6] = {t->nx, t->ny, t->nz, t->nvox, (int64_t)t->fp32, t->step};
    bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
              fwrite(&t->time, sizeof(double), 1, f) == 1 && fwrite(t->dref, sizeof(double), ND, f) == (size_t)ND &&
              fwrite(r.data(), 1, r.size(), f) == r.size() && fwrite(d.data(), 1, d.size(), f) == d.size() &&
              fwrite(g.data(), 1, g.size(), f) == g.size();
    fclose(f);
    LQT_CHECK(ok, LQT_ERR_IO, "short write to %s", path);
    return LQT_OK;
}

lqt_status lqt_tissue_load(lqt_tissue* t, const char* path) {
    LQT_CHECK(t && path, LQT_ERR_ARG, "lqt_tissue_load: bad argument");
    FILE* f = fopen(path, "rb");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    char magic[8];
    int64_t hdr[6];
    double time;
    bool ok = fread(magic, 1, 8, f) == 8 && (!memcmp(magic, kMagic, 8) || !memcmp(magic, kMagic2, 8)) &&
              fread(hdr, sizeof(hdr), 1, f) == 1 && fread(&time, sizeof(double), 1, f) == 1;
    if (!ok || hdr[0] != t->nx || hdr[1] != t->ny || hdr[2] != t->nz || hdr[3] != t->nvox || hdr[4] != (int64_t)t->fp32) {
        fclose(f);
        return set_error(LQT_ERR_IO, "%s: not a checkpoint of this geometry and precision", path);
    }
    const bool v2 = !memcmp(magic, kMagic2, 8);
    double dref[ND] = {0, 0, 0, 0, 0};
    const size_t rs = rsz(t), ds = v2 ? sizeof(double) : dsz(t);
    std::vector<unsigned char> r(rs * NR * t->pitch), d(ds * ND * t->pitch), g(rs * t->ncell_pad);
    ok = (v2 || fread(dref, sizeof(double), ND, f) == (size_t)ND) && fread(r.data(), 1, r.size(), f) == r.size() &&
         fread(d.data(), 1, d.size(), f) == d.size() && fread(g.data(), 1, g.size(), f) == g.size();
    fclose(f);
    LQT_CHECK(ok, LQT_ERR_IO, "%s: truncated checkpoint", path);
    if (v2 && t->fp32) {
        // slow slots stored as double: re-express them as offsets
        const double* dd = (const double*)d.data();
        for (int j = 0; j < ND; ++j) {
            double lo = 0, hi = 0;
            for (int64_t k = 0; k < t->nvox; ++k) {
                const double x = dd[chunk_slot(k, j, ND)];
                lo = k ? std::min(lo, x) : x;
                hi = k ? std::max(hi, x) : x;
            }
            dref[j] = 0.5 * (lo + hi);
        }
        std::vector<unsigned char> df(sizeof(float) * ND * t->pitch, 0);
        for (int64_t k = 0; k < t->nvox; ++k)
            for (int j = 0; j < ND; ++j) {
                const int64_t i = chunk_slot(k, j, ND);
                ((float*)df.data())[i] = (float)(dd[i] - dref[j]);
            }
        d.swap(df);
    }
    LQT_HIP(hipStreamSynchronize(t->stream));
    LQT_HIP(hipMemcpy(t->r, r.data(), r.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->d, d.data(), d.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->grid[t->cur], g.data(), g.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(t->grid[t->cur ^ 1], g.data(), g.size(), hipMemcpyHostToDevice));
    for (int j = 0; j < ND; ++j) t->dref[j] = dref[j];
    t->time = time;
    t->step = hdr[5];
    return LQT_OK;
}

}  // extern "C"
