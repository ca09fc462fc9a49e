// C ABI: the dose-sweep surrogate, an MLP n_in -> H -> H -> n_out with GELU,
// BF16 operands on the matrix cores with FP32 accumulation and Adam on FP32
// master weights.  The reference has no learned component.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "common.hpp"
#include "mlp.hpp"

using namespace lqt;

namespace {
constexpr int K1 = 32;  // padded input width (n_in <= 32)
constexpr int N3 = 32;  // padded output width (n_out <= 32)
}  // namespace

struct lqt_mlp {
    int n_in = 0, H = 0, n_out = 0, n_reg = 0, n_bce = 0;
    int dev = 0;  // device the handle's memory lives on
    hipStream_t stream = nullptr;
    // fp32 masters (padded), Adam moments, gradients
    float *w1 = nullptr, *b1 = nullptr, *w2 = nullptr, *b2 = nullptr, *w3 = nullptr, *b3 = nullptr;
    float* mom = nullptr;  // 2 x total
    float* grad = nullptr; // total
    size_t off[6] = {0}, cnt[6] = {0}, total = 0;
    // bf16 weights and transposes
    __bf16 *w1b = nullptr, *w1t = nullptr, *w2b = nullptr, *w2t = nullptr, *w3b = nullptr, *w3t = nullptr;
    // batch buffers (B rows)
    int Bcap = 0;
    __bf16 *xb = nullptr, *xt = nullptr, *h1b = nullptr, *h1t = nullptr, *h2b = nullptr, *h2t = nullptr;
    __bf16 *gob = nullptr, *got = nullptr, *g2b = nullptr, *g2t = nullptr, *g1b = nullptr, *g1t = nullptr;
    float *h1pre = nullptr, *h2pre = nullptr, *out = nullptr, *yb = nullptr, *mb = nullptr, *loss = nullptr;
    int* ctr = nullptr;  // Adam step, batch index
    // training set resident on the device
    float *X = nullptr, *Y = nullptr, *Mk = nullptr;
    int* idx = nullptr;
    float* loss_log = nullptr;
    int64_t nX = 0, nidx = 0, nlog = 0;
    // inference input
    float* xin = nullptr;
    int64_t nxin = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_B = 0, graph_m = 0;
    float lr = 1e-3f;
};

static lqt_status ensure_batch(lqt_mlp* m, int B) {
    if (B <= m->Bcap) return LQT_OK;
    __bf16** bf[] = {&m->xb, &m->xt, &m->h1b, &m->h1t, &m->h2b, &m->h2t, &m->gob, &m->got, &m->g2b, &m->g2t, &m->g1b, &m->g1t};
    for (auto p : bf) hipFree(*p);
    float** fl[] = {&m->h1pre, &m->h2pre, &m->out, &m->yb, &m->mb};
    for (auto p : fl) hipFree(*p);
    const int H = m->H;
    const size_t S = sizeof(__bf16);
    LQT_HIP(hipMalloc((void**)&m->xb, S * B * K1));
    LQT_HIP(hipMalloc((void**)&m->xt, S * B * K1));
    LQT_HIP(hipMalloc((void**)&m->h1b, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->h1t, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->h2b, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->h2t, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->gob, S * B * N3));
    LQT_HIP(hipMalloc((void**)&m->got, S * B * N3));
    LQT_HIP(hipMalloc((void**)&m->g2b, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->g2t, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->g1b, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->g1t, S * B * H));
    LQT_HIP(hipMalloc((void**)&m->h1pre, sizeof(float) * B * H));
    LQT_HIP(hipMalloc((void**)&m->h2pre, sizeof(float) * B * H));
    LQT_HIP(hipMalloc((void**)&m->out, sizeof(float) * B * N3));
    LQT_HIP(hipMalloc((void**)&m->yb, sizeof(float) * B * N3));
    LQT_HIP(hipMalloc((void**)&m->mb, sizeof(float) * B * N3));
    m->Bcap = B;
    return LQT_OK;
}

// Buffers of one forward pass: the packed input, the hidden activations
// (bf16, and optionally their transposes and pre-activation fp32 copies for
// the backward pass) and the padded fp32 output.
struct FwdBufs {
    const __bf16* xb;
    __bf16 *h1b, *h2b, *h1t, *h2t;  // h1t/h2t may be null
    float *h1pre, *h2pre;           // may be null
    float* out;
};

// Forward pass of B rows already in b.xb (m_valid real rows).
static void forward_into(const lqt_mlp* m, const FwdBufs& b, int B, int mv, hipStream_t st) {
    const int H = m->H;
    GemmEpi e{};
    e.bias = m->b1;
    e.act = 1;
    e.c32 = b.h1pre;
    e.ldc32 = H;
    e.cb = b.h1b;
    e.ldcb = H;
    e.cbt = b.h1t;
    e.ldcbt = B;
    e.n_valid = H;
    e.m_valid = mv;
    launch_gemm_bt(b.xb, K1, m->w1b, K1, B, H, K1, e, st);
    e.bias = m->b2;
    e.c32 = b.h2pre;
    e.cb = b.h2b;
    e.cbt = b.h2t;
    launch_gemm_bt(b.h1b, H, m->w2b, H, B, H, H, e, st);
    GemmEpi o{};
    o.bias = m->b3;
    o.act = 0;
    o.c32 = b.out;
    o.ldc32 = N3;
    o.n_valid = N3;
    o.m_valid = mv;
    launch_gemm_bt(b.h2b, H, m->w3b, H, B, N3, H, o, st);
}

static void forward(lqt_mlp* m, int B, int mv, bool keep_t) {
    const FwdBufs b{m->xb, m->h1b, m->h2b, keep_t ? m->h1t : nullptr, keep_t ? m->h2t : nullptr,
                    m->h1pre, m->h2pre, m->out};
    forward_into(m, b, B, mv, m->stream);
}

// Loss, backward pass and (if adam) the parameter update for the batch in
// xb/xt/yb/mb.  Gradients land in m->grad.
static void backward(lqt_mlp* m, int B, int mv, bool adam) {
    const int H = m->H;
    float* g = m->grad;
    launch_mlp_loss(m->out, N3, m->yb, m->mb, N3, m->n_reg, m->n_bce, B, mv, m->gob, N3, m->got, B, m->loss, m->stream);
    // layer 3: dW3 = dOut^T . H2, db3, dH2pre = (dOut . W3) * GELU'(H2pre)
    GemmEpi e{};
    e.c32 = g + m->off[4];
    e.ldc32 = H;
    e.n_valid = H;
    e.m_valid = N3;
    launch_gemm_bt(m->got, B, m->h2t, B, N3, H, B, e, m->stream);
    launch_rowsum(m->got, B, B, N3, g + m->off[5], m->stream);
    GemmEpi d{};
    d.pre = m->h2pre;
    d.ldpre = H;
    d.cb = m->g2b;
    d.ldcb = H;
    d.cbt = m->g2t;
    d.ldcbt = B;
    d.n_valid = H;
    d.m_valid = mv;
    launch_gemm_bt(m->gob, N3, m->w3t, N3, B, H, N3, d, m->stream);
    // layer 2
    e.c32 = g + m->off[2];
    e.ldc32 = H;
    e.m_valid = H;
    launch_gemm_bt(m->g2t, B, m->h1t, B, H, H, B, e, m->stream);
    launch_rowsum(m->g2t, B, B, H, g + m->off[3], m->stream);
    d.pre = m->h1pre;
    d.cb = m->g1b;
    d.cbt = m->g1t;
    launch_gemm_bt(m->g2b, H, m->w2t, H, B, H, H, d, m->stream);
    // layer 1
    e.c32 = g + m->off[0];
    e.ldc32 = K1;
    e.n_valid = K1;
    e.m_valid = H;
    launch_gemm_bt(m->g1t, B, m->xt, B, H, K1, B, e, m->stream);
    launch_rowsum(m->g1t, B, B, H, g + m->off[1], m->stream);
    if (!adam) return;
    float* P[6] = {m->w1, m->b1, m->w2, m->b2, m->w3, m->b3};
    const int cols[6] = {K1, 1, H, 1, H, 1};
    __bf16* wb[6] = {m->w1b, nullptr, m->w2b, nullptr, m->w3b, nullptr};
    __bf16* wt[6] = {m->w1t, nullptr, m->w2t, nullptr, m->w3t, nullptr};
    const int ldw[6] = {K1, 0, H, 0, H, 0};
    const int ldt[6] = {H, 0, H, 0, N3, 0};
    for (int i = 0; i < 6; ++i)
        launch_adam(P[i], g + m->off[i], m->mom + m->off[i], m->mom + m->total + m->off[i], (int)m->cnt[i], m->ctr,
                    m->lr, 0.9f, 0.999f, 1e-8f, cols[i], wb[i], ldw[i], wt[i], ldt[i], m->stream);
    launch_mlp_step(m->ctr, m->loss_log, m->loss, m->stream);
}

static lqt_status upload_bf16_weights(lqt_mlp* m) {
    // host copies of the masters -> bf16 row-major and transposed
    const int H = m->H;
    std::vector<float> w1(m->cnt[0]), w2(m->cnt[2]), w3(m->cnt[4]);
    LQT_HIP(hipMemcpy(w1.data(), m->w1, sizeof(float) * w1.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(w2.data(), m->w2, sizeof(float) * w2.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(w3.data(), m->w3, sizeof(float) * w3.size(), hipMemcpyDeviceToHost));
    auto put = [&](const std::vector<float>& w, int rows, int cols, __bf16* wb, __bf16* wt) -> lqt_status {
        std::vector<__bf16> a((size_t)rows * cols), t((size_t)rows * cols);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                a[(size_t)r * cols + c] = (__bf16)w[(size_t)r * cols + c];
                t[(size_t)c * rows + r] = (__bf16)w[(size_t)r * cols + c];
            }
        LQT_HIP(hipMemcpy(wb, a.data(), sizeof(__bf16) * a.size(), hipMemcpyHostToDevice));
        LQT_HIP(hipMemcpy(wt, t.data(), sizeof(__bf16) * t.size(), hipMemcpyHostToDevice));
        return LQT_OK;
    };
    lqt_status s;
    if ((s = put(w1, H, K1, m->w1b, m->w1t))) return s;
    if ((s = put(w2, H, H, m->w2b, m->w2t))) return s;
    return put(w3, N3, H, m->w3b, m->w3t);
}

// ---------------------------------------------------------------------------
// Scratch of lqt_mlp_forward_noalloc, the forward pass that allocates
// nothing: the packed bf16 input, the two bf16 hidden activations and the
// padded fp32 output of one call live in this file-scope device array.
// Every device the entry point runs on holds its own copy until the process
// ends.  On the host the array's name is only a stand-in: its device address
// on each device comes from hipGetSymbolAddress, called once per device
// with that device current and kept in g_scratch_ptr; only that address is
// ever given to a kernel.
static __device__ __align__(256) unsigned char scratch[(size_t)8 << 20];

namespace {
constexpr size_t kScratchBytes = (size_t)8 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Bytes of scratch one call of B padded rows needs at hidden width H.
size_t scratch_need(int64_t B, int H) {
    return align256((size_t)B * K1 * sizeof(__bf16)) + 2 * align256((size_t)B * H * sizeof(__bf16)) +
           align256((size_t)B * N3 * sizeof(float));
}

// Largest padded row count whose buffers fit the scratch at hidden width H.
int64_t scratch_rows(int H) {
    int64_t B = (int64_t)(kScratchBytes / (sizeof(__bf16) * (K1 + 2 * (size_t)H) + sizeof(float) * N3)) / 32 * 32;
    while (B > 0 && scratch_need(B, H) > kScratchBytes) B -= 32;
    return B;
}

// Device address of the scratch array on device dev (the current device).
lqt_status scratch_base(int dev, unsigned char** out) {
    LQT_CHECK(dev >= 0 && dev < kMaxDevices, LQT_ERR_ARG, "scratch: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        LQT_CHECK(e == hipSuccess && p, LQT_ERR_HIP, "scratch: hipGetSymbolAddress on device %d failed: %s", dev,
                  hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return LQT_OK;
}
}  // namespace

extern "C" {

lqt_status lqt_mlp_create(int n_in, int hidden, int n_reg, int n_bce, uint64_t seed, lqt_mlp** out) {
    LQT_CHECK(out && n_in > 0 && n_in <= K1 && hidden >= 32 && hidden % 32 == 0 && n_reg >= 0 && n_bce >= 0 &&
                  n_reg + n_bce > 0 && n_reg + n_bce <= N3,
              LQT_ERR_ARG, "lqt_mlp_create: need 0 < n_in <= 32, hidden a multiple of 32, 0 < n_out <= 32");
    int ndev = 0;
    LQT_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, LQT_ERR_HIP, "lqt_mlp_create: no HIP device available");
    lqt_mlp* m = new lqt_mlp();
    m->n_in = n_in;
    m->H = hidden;
    m->n_reg = n_reg;
    m->n_bce = n_bce;
    m->n_out = n_reg + n_bce;
    const int H = hidden;
    if (hipGetDevice(&m->dev) != hipSuccess) m->dev = 0;
    const size_t cnt[6] = {(size_t)H * K1, (size_t)H, (size_t)H * H, (size_t)H, (size_t)N3 * H, (size_t)N3};
    size_t o = 0;
    for (int i = 0; i < 6; ++i) {
        m->off[i] = o;
        m->cnt[i] = cnt[i];
        o += (cnt[i] + 63) / 64 * 64;
    }
    m->total = o;
    lqt_status st = [&]() -> lqt_status {
        LQT_HIP(hipStreamCreate(&m->stream));
        float* params = nullptr;
        LQT_HIP(hipMalloc((void**)&params, sizeof(float) * m->total));
        m->w1 = params + m->off[0];
        m->b1 = params + m->off[1];
        m->w2 = params + m->off[2];
        m->b2 = params + m->off[3];
        m->w3 = params + m->off[4];
        m->b3 = params + m->off[5];
        LQT_HIP(hipMalloc((void**)&m->mom, sizeof(float) * 2 * m->total));
        LQT_HIP(hipMemset(m->mom, 0, sizeof(float) * 2 * m->total));
        LQT_HIP(hipMalloc((void**)&m->grad, sizeof(float) * m->total));
        LQT_HIP(hipMemset(m->grad, 0, sizeof(float) * m->total));
        LQT_HIP(hipMalloc((void**)&m->w1b, sizeof(__bf16) * H * K1));
        LQT_HIP(hipMalloc((void**)&m->w1t, sizeof(__bf16) * H * K1));
        LQT_HIP(hipMalloc((void**)&m->w2b, sizeof(__bf16) * H * H));
        LQT_HIP(hipMalloc((void**)&m->w2t, sizeof(__bf16) * H * H));
        LQT_HIP(hipMalloc((void**)&m->w3b, sizeof(__bf16) * N3 * H));
        LQT_HIP(hipMalloc((void**)&m->w3t, sizeof(__bf16) * N3 * H));
        LQT_HIP(hipMalloc((void**)&m->loss, sizeof(float)));
        LQT_HIP(hipMalloc((void**)&m->ctr, sizeof(int) * 2));
        const int one[2] = {1, 0};
        LQT_HIP(hipMemcpy(m->ctr, one, sizeof(one), hipMemcpyHostToDevice));
        // He-style initialisation from the seed; padding stays zero
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        std::vector<float> hp(m->total, 0.0f);
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < n_in; ++c) hp[m->off[0] + (size_t)r * K1 + c] = nd(rng) * sqrtf(2.0f / n_in);
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < H; ++c) hp[m->off[2] + (size_t)r * H + c] = nd(rng) * sqrtf(2.0f / H);
        for (int r = 0; r < m->n_out; ++r)
            for (int c = 0; c < H; ++c) hp[m->off[4] + (size_t)r * H + c] = nd(rng) * sqrtf(1.0f / H);
        LQT_HIP(hipMemcpy(params, hp.data(), sizeof(float) * m->total, hipMemcpyHostToDevice));
        return upload_bf16_weights(m);
    }();
    if (st) {
        lqt_mlp_destroy(m);
        return st;
    }
    *out = m;
    return LQT_OK;
}

lqt_status lqt_mlp_destroy(lqt_mlp* m) {
    if (!m) return LQT_OK;
    if (m->stream) hipStreamSynchronize(m->stream);
    if (m->exec) hipGraphExecDestroy(m->exec);
    if (m->graph) hipGraphDestroy(m->graph);
    hipFree(m->w1);
    void* p[] = {m->mom, m->grad, m->w1b, m->w1t, m->w2b, m->w2t, m->w3b, m->w3t, m->xb, m->xt, m->h1b, m->h1t,
                 m->h2b, m->h2t, m->gob, m->got, m->g2b, m->g2t, m->g1b, m->g1t, m->h1pre, m->h2pre, m->out, m->yb,
                 m->mb, m->loss, m->ctr, m->X, m->Y, m->Mk, m->idx, m->loss_log, m->xin};
    for (void* q : p) hipFree(q);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return LQT_OK;
}

// Master parameters, unpadded: w1[H*n_in], b1[H], w2[H*H], b2[H], w3[n_out*H], b3[n_out].
lqt_status lqt_mlp_get_params(lqt_mlp* m, float* w1, float* b1, float* w2, float* b2, float* w3, float* b3) {
    LQT_CHECK(m, LQT_ERR_ARG, "lqt_mlp_get_params: null handle");
    LQT_HIP(hipStreamSynchronize(m->stream));
    std::vector<float> hp(m->total);
    LQT_HIP(hipMemcpy(hp.data(), m->w1, sizeof(float) * m->total, hipMemcpyDeviceToHost));
    const int H = m->H;
    if (w1)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < m->n_in; ++c) w1[(size_t)r * m->n_in + c] = hp[m->off[0] + (size_t)r * K1 + c];
    if (b1) memcpy(b1, &hp[m->off[1]], sizeof(float) * H);
    if (w2) memcpy(w2, &hp[m->off[2]], sizeof(float) * H * H);
    if (b2) memcpy(b2, &hp[m->off[3]], sizeof(float) * H);
    if (w3) memcpy(w3, &hp[m->off[4]], sizeof(float) * m->n_out * H);
    if (b3) memcpy(b3, &hp[m->off[5]], sizeof(float) * m->n_out);
    return LQT_OK;
}

lqt_status lqt_mlp_set_params(lqt_mlp* m, const float* w1, const float* b1, const float* w2, const float* b2,
                              const float* w3, const float* b3) {
    LQT_CHECK(m && w1 && b1 && w2 && b2 && w3 && b3, LQT_ERR_ARG, "lqt_mlp_set_params: bad argument");
    LQT_HIP(hipStreamSynchronize(m->stream));
    const int H = m->H;
    std::vector<float> hp(m->total, 0.0f);
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < m->n_in; ++c) hp[m->off[0] + (size_t)r * K1 + c] = w1[(size_t)r * m->n_in + c];
    memcpy(&hp[m->off[1]], b1, sizeof(float) * H);
    memcpy(&hp[m->off[2]], w2, sizeof(float) * H * H);
    memcpy(&hp[m->off[3]], b2, sizeof(float) * H);
    memcpy(&hp[m->off[4]], w3, sizeof(float) * m->n_out * H);
    memcpy(&hp[m->off[5]], b3, sizeof(float) * m->n_out);
    LQT_HIP(hipMemcpy(m->w1, hp.data(), sizeof(float) * m->total, hipMemcpyHostToDevice));
    return upload_bf16_weights(m);
}

lqt_status lqt_mlp_forward(lqt_mlp* m, int64_t n, const float* x, float* y) {
    LQT_CHECK(m && x && y && n > 0, LQT_ERR_ARG, "lqt_mlp_forward: bad argument");
    const int64_t chunk = 1 << 16;
    if (m->nxin < std::min(n, chunk)) {
        hipFree(m->xin);
        m->xin = nullptr;
        m->nxin = std::min(n, chunk);
        LQT_HIP(hipMalloc((void**)&m->xin, sizeof(float) * m->nxin * m->n_in));
    }
    std::vector<float> out;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int rows = (int)std::min(chunk, n - r0);
        const int B = (rows + 31) / 32 * 32;
        lqt_status s = ensure_batch(m, B);
        if (s) return s;
        LQT_HIP(hipMemcpyAsync(m->xin, x + r0 * m->n_in, sizeof(float) * rows * m->n_in, hipMemcpyHostToDevice,
                               m->stream));
        launch_pack(m->xin, m->n_in, rows, m->xb, K1, m->stream);
        forward(m, B, rows, false);
        LQT_HIP(hipGetLastError());
        out.resize((size_t)B * N3);
        LQT_HIP(hipMemcpyAsync(out.data(), m->out, sizeof(float) * B * N3, hipMemcpyDeviceToHost, m->stream));
        LQT_HIP(hipStreamSynchronize(m->stream));
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < m->n_out; ++j) y[(r0 + i) * m->n_out + j] = out[(size_t)i * N3 + j];
    }
    return LQT_OK;
}

static lqt_status upload_set(lqt_mlp* m, int64_t n, const float* x, const float* y, const float* mask) {
    if (m->nX < n) {
        hipFree(m->X);
        hipFree(m->Y);
        hipFree(m->Mk);
        m->X = m->Y = m->Mk = nullptr;
        LQT_HIP(hipMalloc((void**)&m->X, sizeof(float) * n * m->n_in));
        LQT_HIP(hipMalloc((void**)&m->Y, sizeof(float) * n * N3));
        LQT_HIP(hipMalloc((void**)&m->Mk, sizeof(float) * n * N3));
        m->nX = n;
    }
    std::vector<float> yp((size_t)n * N3, 0.0f), mp((size_t)n * N3, 0.0f);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < m->n_out; ++j) {
            yp[(size_t)i * N3 + j] = y[i * m->n_out + j];
            mp[(size_t)i * N3 + j] = mask ? mask[i * m->n_out + j] : 1.0f;
        }
    LQT_HIP(hipMemcpy(m->X, x, sizeof(float) * n * m->n_in, hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(m->Y, yp.data(), sizeof(float) * yp.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(m->Mk, mp.data(), sizeof(float) * mp.size(), hipMemcpyHostToDevice));
    return LQT_OK;
}

// Loss and gradients of one batch of n rows (no update): the gradient check
// entry.  Gradients are returned unpadded like lqt_mlp_get_params.
lqt_status lqt_mlp_grads(lqt_mlp* m, int n, const float* x, const float* y, const float* mask, float* loss,
                         float* gw1, float* gb1, float* gw2, float* gb2, float* gw3, float* gb3) {
    LQT_CHECK(m && x && y && n > 0 && n <= 65536, LQT_ERR_ARG, "lqt_mlp_grads: bad argument");
    lqt_status s = upload_set(m, n, x, y, mask);
    if (s) return s;
    const int B = (n + 31) / 32 * 32;
    if ((s = ensure_batch(m, B))) return s;
    std::vector<int> id(n);
    std::iota(id.begin(), id.end(), 0);
    hipFree(m->idx);
    m->idx = nullptr;
    if ((s = upload(&m->idx, id.data(), n))) return s;
    m->nidx = n;
    int* bctr = nullptr;
    LQT_HIP(hipMalloc((void**)&bctr, sizeof(int)));
    LQT_HIP(hipMemset(bctr, 0, sizeof(int)));
    launch_gather(m->X, m->n_in, m->Y, m->Mk, N3, m->idx, bctr, B, n, m->xb, K1, m->xt, B, m->yb, m->mb, m->loss,
                  m->stream);
    forward(m, B, n, true);
    backward(m, B, n, false);
    LQT_HIP(hipGetLastError());
    LQT_HIP(hipStreamSynchronize(m->stream));
    hipFree(bctr);
    if (loss) LQT_HIP(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> g(m->total);
    LQT_HIP(hipMemcpy(g.data(), m->grad, sizeof(float) * m->total, hipMemcpyDeviceToHost));
    const int H = m->H;
    if (gw1)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < m->n_in; ++c) gw1[(size_t)r * m->n_in + c] = g[m->off[0] + (size_t)r * K1 + c];
    if (gb1) memcpy(gb1, &g[m->off[1]], sizeof(float) * H);
    if (gw2) memcpy(gw2, &g[m->off[2]], sizeof(float) * H * H);
    if (gb2) memcpy(gb2, &g[m->off[3]], sizeof(float) * H);
    if (gw3) memcpy(gw3, &g[m->off[4]], sizeof(float) * m->n_out * H);
    if (gb3) memcpy(gb3, &g[m->off[5]], sizeof(float) * m->n_out);
    return LQT_OK;
}

// Adam training: epochs passes over the n rows in shuffled batches of
// `batch` rows (the last partial batch of an epoch is dropped).  The step of
// one batch (gather, forward, loss, backward, Adam) is captured once in a
// HIP graph and replayed.  loss_out[epochs*nb] receives every batch loss.
lqt_status lqt_mlp_train(lqt_mlp* m, int64_t n, const float* x, const float* y, const float* mask, int epochs,
                         int batch, double lr, uint64_t seed, float* loss_out) {
    LQT_CHECK(m && x && y && n > 0 && epochs > 0 && batch >= 32 && batch <= 65536 && lr > 0, LQT_ERR_ARG,
              "lqt_mlp_train: bad argument");
    const int nb = (int)(n / batch);
    LQT_CHECK(nb > 0, LQT_ERR_ARG, "lqt_mlp_train: fewer rows than one batch");
    lqt_status s = upload_set(m, n, x, y, mask);
    if (s) return s;
    const int B = (batch + 31) / 32 * 32;
    if ((s = ensure_batch(m, B))) return s;
    const int64_t steps = (int64_t)epochs * nb;
    std::vector<int> id((size_t)steps * batch);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        memcpy(&id[(size_t)e * nb * batch], perm.data(), sizeof(int) * (size_t)nb * batch);
    }
    hipFree(m->idx);
    m->idx = nullptr;
    if ((s = upload(&m->idx, id.data(), id.size()))) return s;
    if (m->nlog < steps) {
        hipFree(m->loss_log);
        m->loss_log = nullptr;
        LQT_HIP(hipMalloc((void**)&m->loss_log, sizeof(float) * steps));
        m->nlog = steps;
    }
    m->lr = (float)lr;
    // the graph bakes in the buffers and sizes; rebuild when they change
    if (m->exec) {
        hipGraphExecDestroy(m->exec);
        hipGraphDestroy(m->graph);
        m->exec = nullptr;
        m->graph = nullptr;
    }
    launch_set_int(m->ctr + 1, 0, m->stream);
    LQT_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    launch_gather(m->X, m->n_in, m->Y, m->Mk, N3, m->idx, m->ctr + 1, B, batch, m->xb, K1, m->xt, B, m->yb, m->mb,
                  m->loss, m->stream);
    forward(m, B, batch, true);
    backward(m, B, batch, true);
    hipGraph_t g = nullptr;
    LQT_HIP(hipStreamEndCapture(m->stream, &g));
    m->graph = g;
    LQT_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    // At most kInFlight replays (about 20 kernels each) are queued ahead of
    // the device: the host waits for the stream after every kInFlight
    // launches instead of queueing all epochs * nb replays at once.
    constexpr int64_t kInFlight = 64;
    for (int64_t i = 0; i < steps; ++i) {
        LQT_HIP(hipGraphLaunch(m->exec, m->stream));
        if ((i + 1) % kInFlight == 0) LQT_HIP(hipStreamSynchronize(m->stream));
    }
    LQT_HIP(hipStreamSynchronize(m->stream));
    if (loss_out) LQT_HIP(hipMemcpy(loss_out, m->loss_log, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return LQT_OK;
}

// Forward pass on device buffers that allocates nothing: x[n*n_in] and
// y[n*n_out] are device memory on the handle's device, which must be
// current; the intermediate buffers are the file-scope scratch array, so
// n is limited to scratch_rows(hidden) rows (11,904 at hidden 128).  The
// work is enqueued on `stream` and not waited for.  Calls that overlap on
// one device share the scratch and race.  The kernels are those of
// lqt_mlp_forward, so the outputs are the same bits.
lqt_status lqt_mlp_forward_noalloc(lqt_mlp* m, int64_t n, const float* x, float* y, void* stream) {
    LQT_CHECK(m && x && y && n > 0, LQT_ERR_ARG, "lqt_mlp_forward_noalloc: bad argument");
    const int64_t nmax = scratch_rows(m->H);
    LQT_CHECK(n <= nmax, LQT_ERR_ARG,
              "lqt_mlp_forward_noalloc: %lld rows exceed the scratch limit of %lld rows at hidden width %d "
              "(%zu bytes)",
              (long long)n, (long long)nmax, m->H, kScratchBytes);
    int dev = -1;
    LQT_HIP(hipGetDevice(&dev));
    LQT_CHECK(dev == m->dev, LQT_ERR_ARG,
              "lqt_mlp_forward_noalloc: the handle lives on device %d, make it current (device %d is)", m->dev, dev);
    unsigned char* base = nullptr;
    lqt_status s = scratch_base(dev, &base);
    if (s) return s;
    const int rows = (int)n;
    const int B = (rows + 31) / 32 * 32;
    __bf16* xb = (__bf16*)base;
    __bf16* h1b = (__bf16*)(base + align256((size_t)B * K1 * sizeof(__bf16)));
    __bf16* h2b = (__bf16*)((unsigned char*)h1b + align256((size_t)B * m->H * sizeof(__bf16)));
    float* out = (float*)((unsigned char*)h2b + align256((size_t)B * m->H * sizeof(__bf16)));
    hipStream_t st = (hipStream_t)stream;
    launch_pack(x, m->n_in, rows, xb, K1, st);
    forward_into(m, FwdBufs{xb, h1b, h2b, nullptr, nullptr, nullptr, nullptr, out}, B, rows, st);
    launch_unpack(out, N3, rows, m->n_out, y, st);
    LQT_HIP(hipGetLastError());
    return LQT_OK;
}

}  // extern "C"
