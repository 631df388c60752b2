// mlp_abi.hip -- host side of the surrogate C-ABI (hadros_mlp_*).
//
// The network: n_in -> width (SiLU) x depth -> n_out, FP32 master weights
// with Adam moments, BF16 weight copies for the MFMA GEMMs of mlp.hip.
// Inputs and targets are standardised with vectors set by the caller
// (hadros_mlp_set_norm); hadros_mlp_forward returns de-standardised outputs.
// The first layer's input is padded to 16 features (zero columns, whose
// weight rows stay exactly zero).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"
#include "mlp.h"

using namespace hd;

namespace {

#define MC(expr)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return set_error(HADROS_ERR_DEVICE, hipGetErrorString(e_));    \
    } while (0)

template <class T>
int dal(T** p, size_t n)
{
    *p = nullptr;
    if (!n) return HADROS_OK;
    MC(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
    return HADROS_OK;
}
template <class T>
void dfr(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// splitmix64 -> uniform in [0, 1)
struct Rng {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

constexpr int PAD_IN = 16;   // first-layer input width (n_in <= 16)

}  // namespace

struct hadros_mlp {
    int n_in, width, depth, n_out, tb, ib;
    int L;                     // layers = depth + 1
    std::vector<int> in, out;  // per layer (padded input for layer 0)
    std::vector<size_t> woff, boff;   // offsets into the flat internal parameter array
    size_t n_int = 0;          // internal parameter count (padded W1)
    long step = 0;
    float* P = nullptr;        // fp32 masters (internal layout)
    float* G = nullptr;        // gradients
    float* Mo = nullptr;       // Adam first moment
    float* Vo = nullptr;       // Adam second moment
    __bf16* W16 = nullptr;     // bf16 copies of all parameters (biases unused)
    float *xm = nullptr, *xi = nullptr, *ym = nullptr, *yi = nullptr, *ys = nullptr, *zero = nullptr, *one = nullptr;
    // training buffers (tb rows)
    __bf16* Xb = nullptr;
    std::vector<__bf16*> H;    // [L-1] activations
    std::vector<float*> Z;     // [L-1] pre-activations
    float* Y = nullptr;        // [tb x n_out]
    __bf16 *dY = nullptr, *Ga = nullptr, *Gb = nullptr;
    float* part = nullptr;     // split-K partials / column-sum partials
    size_t part_n = 0;
    float* lpart = nullptr;    // loss partials
    // inference buffers (ib rows)
    __bf16 *Xi = nullptr, *Ha = nullptr, *Hb = nullptr;
    // fused inference: weights in MFMA fragment order, refreshed after every
    // parameter change (sw_dirty)
    bool fused_ok = false;
    int infer_path = 0;        // 0: fused when the shape allows, 1: per-layer GEMMs
    __bf16* Wsw = nullptr;
    std::vector<long> swo;     // per layer: first 16-byte fragment
    size_t n_sw = 0;           // bf16 elements
    bool sw_dirty = true;
};

namespace {

constexpr int SPLITS = 16;
constexpr int COLSUM_ROWS = 32;   // 256 row chunks of a 8192-row batch: one workgroup each

void layout(hadros_mlp* m)
{
    m->L = m->depth + 1;
    m->in.assign(m->L, m->width);
    m->out.assign(m->L, m->width);
    m->in[0] = PAD_IN;
    m->out[m->L - 1] = m->n_out;
    size_t o = 0;
    for (int l = 0; l < m->L; ++l) {
        m->woff.push_back(o);
        o += size_t(m->in[l]) * m->out[l];
        m->boff.push_back(o);
        o += size_t(m->out[l]);
    }
    m->n_int = o;
    // swizzled copies: K padded to 32, N to 16
    m->fused_ok = m->width == 256 && m->n_out <= 16 && m->L <= FUSED_MAX_LAYERS;
    long f = 0;
    for (int l = 0; l < m->L; ++l) {
        m->swo.push_back(f);
        f += long((m->in[l] + 31) / 32) * ((m->out[l] + 15) / 16) * 64;
    }
    m->n_sw = size_t(f) * 8;
}

int refresh_swizzle(hadros_mlp* m, hipStream_t s)
{
    if (!m->sw_dirty) return HADROS_OK;
    for (int l = 0; l < m->L; ++l)
        MC(mlp_swizzle(m->W16 + m->woff[l], m->in[l], m->out[l], (m->in[l] + 31) / 32, (m->out[l] + 15) / 16,
                       m->Wsw + size_t(m->swo[l]) * 8, s));
    m->sw_dirty = false;
    return HADROS_OK;
}

// fused inference of n rows: mode 0 de-standardised outputs, mode 1 survivals
int infer_fused(hadros_mlp* m, const float* x, int64_t n, float* y, int mode, hipStream_t s)
{
    int rc = refresh_swizzle(m, s);
    if (rc) return rc;
    FusedArgs a{};
    a.x = x;
    a.M = long(n);
    a.n_in = m->n_in;
    a.L = m->L;
    a.n_out = m->n_out;
    a.xm = m->xm;
    a.xi = m->xi;
    a.w = m->Wsw;
    for (int l = 0; l < m->L; ++l) {
        a.wo[l] = m->swo[l];
        a.bo[l] = long(m->boff[l]);
    }
    a.bias = m->P;
    a.ys = m->ys;
    a.ym = m->ym;
    a.y = y;
    MC(mlp_fused(a, mode, s));
    return HADROS_OK;
}

size_t n_public(const hadros_mlp* m) { return m->n_int - size_t(PAD_IN - m->n_in) * m->out[0]; }

// public flat layout: W1 (n_in x width), b1, W2, b2, ..., W_L, b_L
template <class F>
void map_public(const hadros_mlp* m, F&& f)   // f(public_index, internal_index)
{
    size_t q = 0;
    for (int l = 0; l < m->L; ++l) {
        const int rows = (l == 0) ? m->n_in : m->in[l];
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < m->out[l]; ++c) f(q++, m->woff[l] + size_t(r) * m->out[l] + c);
        for (int c = 0; c < m->out[l]; ++c) f(q++, m->boff[l] + c);
    }
}

int forward_train(hadros_mlp* m, const float* x, long n, hipStream_t s)
{
    MC(mlp_pack(x, n, m->n_in, PAD_IN, m->xm, m->xi, m->Xb, s));
    const __bf16* A = m->Xb;
    for (int l = 0; l < m->L; ++l) {
        GemmArgs g{};
        g.A = A;
        g.B = m->W16 + m->woff[l];
        g.M = int(n);
        g.N = m->out[l];
        g.K = m->in[l];
        g.lda = m->in[l];
        g.ldb = m->out[l];
        g.ldc = m->out[l];
        g.bias = m->P + m->boff[l];
        if (l < m->L - 1) {
            g.Cf = m->Z[l];
            g.Cb = m->H[l];
            MC(mlp_gemm(g, false, false, EPI_FWD_SILU, 1, s));
            A = m->H[l];
        } else {
            g.Cf = m->Y;
            g.scale = m->one;
            g.shift = m->zero;
            MC(mlp_gemm(g, false, false, EPI_FWD_LINEAR, 1, s));
        }
    }
    return HADROS_OK;
}

int backward(hadros_mlp* m, long n, hipStream_t s)
{
    const __bf16* Gc = m->dY;
    for (int l = m->L - 1; l >= 0; --l) {
        const __bf16* Ain = (l == 0) ? m->Xb : m->H[l - 1];
        // dW_l = A_in^T . G  (split over the batch, deterministic reduction)
        GemmArgs g{};
        g.A = Ain;
        g.B = Gc;
        g.M = m->in[l];
        g.N = m->out[l];
        g.K = int(n);
        g.lda = m->in[l];
        g.ldb = m->out[l];
        g.ldc = m->out[l];
        g.Cf = m->part;
        MC(mlp_gemm(g, true, false, EPI_F32, SPLITS, s));
        MC(mlp_sum_parts(m->part, SPLITS, long(m->in[l]) * m->out[l], m->G + m->woff[l], s));
        // db_l = column sums of G
        const int chunks = int((n + COLSUM_ROWS - 1) / COLSUM_ROWS);
        MC(mlp_colsum(Gc, int(n), m->out[l], COLSUM_ROWS, m->part, s));
        MC(mlp_sum_parts(m->part, chunks, m->out[l], m->G + m->boff[l], s));
        if (l == 0) break;
        // G_prev = (G . W_l^T) * silu'(Z_{l-1})
        __bf16* Gn = (Gc == m->Ga) ? m->Gb : m->Ga;
        GemmArgs h{};
        h.A = Gc;
        h.B = m->W16 + m->woff[l];
        h.M = int(n);
        h.N = m->in[l];
        h.K = m->out[l];
        h.lda = m->out[l];
        h.ldb = m->out[l];
        h.ldc = m->in[l];
        h.aux = m->Z[l - 1];
        h.Cb = Gn;
        MC(mlp_gemm(h, false, true, EPI_BWD_SILU, 1, s));
        Gc = Gn;
    }
    return HADROS_OK;
}

int loss_and_grad(hadros_mlp* m, const float* x, const float* y, long n, float* loss, hipStream_t s)
{
    if (n <= 0 || n > m->tb) return set_error(HADROS_ERR_ARG, "batch rows must be in [1, train_batch]");
    int rc = forward_train(m, x, n, s);
    if (rc) return rc;
    int np = 0;
    MC(mlp_mse(m->Y, y, n, m->n_out, m->ym, m->yi, m->dY, m->lpart, &np, s));
    rc = backward(m, n, s);
    if (rc) return rc;
    if (loss) {
        std::vector<float> h(static_cast<size_t>(np));
        MC(hipMemcpyAsync(h.data(), m->lpart, h.size() * 4, hipMemcpyDeviceToHost, s));
        MC(hipStreamSynchronize(s));
        double t = 0.0;
        for (float v : h) t += v;
        *loss = float(t / double(n * m->n_out));
    }
    return HADROS_OK;
}

}  // namespace

extern "C" {

int hadros_mlp_destroy(hadros_mlp* m);

int hadros_mlp_create(int n_in, int width, int depth, int n_out, int train_batch, int infer_batch, uint64_t seed,
                      hadros_mlp** out)
{
    if (!out || n_in < 1 || n_in > PAD_IN || width < 16 || width % 16 || depth < 1 || n_out < 1 || train_batch < 1 ||
        infer_batch < 1)
        return set_error(HADROS_ERR_ARG, "bad MLP shape (n_in <= 16, width multiple of 16, depth >= 1)");
    int rc = device_ready();
    if (rc) return rc;
    hadros_mlp* m = new hadros_mlp();
    m->n_in = n_in;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    m->tb = train_batch;
    m->ib = infer_batch;
    layout(m);
    m->H.assign(size_t(m->L - 1), nullptr);
    m->Z.assign(size_t(m->L - 1), nullptr);
    size_t maxw = 0;
    for (int l = 0; l < m->L; ++l) maxw = std::max(maxw, size_t(m->in[l]) * m->out[l]);
    const size_t chunks = size_t((train_batch + COLSUM_ROWS - 1) / COLSUM_ROWS);
    m->part_n = std::max(maxw * SPLITS, chunks * size_t(std::max(width, n_out)));
    const size_t tb = size_t(train_batch), ib = size_t(infer_batch);
    if ((rc = dal(&m->P, m->n_int)) || (rc = dal(&m->G, m->n_int)) || (rc = dal(&m->Mo, m->n_int)) ||
        (rc = dal(&m->Vo, m->n_int)) || (rc = dal(&m->W16, m->n_int)) || (rc = dal(&m->xm, PAD_IN)) ||
        (rc = dal(&m->xi, PAD_IN)) || (rc = dal(&m->ym, size_t(n_out))) || (rc = dal(&m->yi, size_t(n_out))) ||
        (rc = dal(&m->ys, size_t(n_out))) || (rc = dal(&m->zero, size_t(n_out))) || (rc = dal(&m->one, size_t(n_out))) ||
        (rc = dal(&m->Xb, tb * PAD_IN)) || (rc = dal(&m->Y, tb * n_out)) || (rc = dal(&m->dY, tb * n_out)) ||
        (rc = dal(&m->Ga, tb * width)) || (rc = dal(&m->Gb, tb * width)) || (rc = dal(&m->part, m->part_n)) ||
        (rc = dal(&m->lpart, (tb * n_out + 255) / 256)) || (rc = dal(&m->Xi, ib * PAD_IN)) ||
        (rc = dal(&m->Ha, ib * width)) || (rc = dal(&m->Hb, ib * width)) || (rc = dal(&m->Wsw, m->n_sw))) {
        hadros_mlp_destroy(m);
        return rc;
    }
    for (int l = 0; l < m->L - 1; ++l)
        if ((rc = dal(&m->H[l], tb * width)) || (rc = dal(&m->Z[l], tb * width))) {
            hadros_mlp_destroy(m);
            return rc;
        }
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases; padded rows zero
    std::vector<float> p(m->n_int, 0.0f);
    Rng rng{seed};
    for (int l = 0; l < m->L; ++l) {
        const int rows = (l == 0) ? n_in : m->in[l];
        const double bound = std::sqrt(6.0 / double(rows));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < m->out[l]; ++c)
                p[m->woff[l] + size_t(r) * m->out[l] + c] = float((2.0 * rng.next() - 1.0) * bound);
    }
    std::vector<float> zeros(m->n_int, 0.0f), ones(size_t(std::max(PAD_IN, n_out)), 1.0f);
    MC(hipMemcpy(m->P, p.data(), m->n_int * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->Mo, zeros.data(), m->n_int * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->Vo, zeros.data(), m->n_int * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->xm, zeros.data(), PAD_IN * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->xi, ones.data(), PAD_IN * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->ym, zeros.data(), size_t(n_out) * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->zero, zeros.data(), size_t(n_out) * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->yi, ones.data(), size_t(n_out) * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->ys, ones.data(), size_t(n_out) * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->one, ones.data(), size_t(n_out) * 4, hipMemcpyHostToDevice));
    MC(mlp_to_bf16(m->P, long(m->n_int), m->W16, nullptr));
    m->sw_dirty = true;
    MC(hipDeviceSynchronize());
    *out = m;
    return HADROS_OK;
}

int hadros_mlp_n_params(hadros_mlp* m, int64_t* n)
{
    if (!m || !n) return set_error(HADROS_ERR_ARG, "NULL argument");
    *n = int64_t(n_public(m));
    return HADROS_OK;
}

int hadros_mlp_get_params(hadros_mlp* m, float* host, int64_t n)
{
    if (!m || !host || size_t(n) != n_public(m)) return set_error(HADROS_ERR_ARG, "parameter count mismatch");
    std::vector<float> p(m->n_int);
    MC(hipMemcpy(p.data(), m->P, m->n_int * 4, hipMemcpyDeviceToHost));
    map_public(m, [&](size_t q, size_t i) { host[q] = p[i]; });
    return HADROS_OK;
}

int hadros_mlp_set_params(hadros_mlp* m, const float* host, int64_t n)
{
    if (!m || !host || size_t(n) != n_public(m)) return set_error(HADROS_ERR_ARG, "parameter count mismatch");
    std::vector<float> p(m->n_int, 0.0f);
    map_public(m, [&](size_t q, size_t i) { p[i] = host[q]; });
    MC(hipMemcpy(m->P, p.data(), m->n_int * 4, hipMemcpyHostToDevice));
    MC(mlp_to_bf16(m->P, long(m->n_int), m->W16, nullptr));
    m->sw_dirty = true;
    MC(hipDeviceSynchronize());
    return HADROS_OK;
}

int hadros_mlp_set_norm(hadros_mlp* m, const float* x_mean, const float* x_std, const float* y_mean, const float* y_std)
{
    if (!m || !x_mean || !x_std || !y_mean || !y_std) return set_error(HADROS_ERR_ARG, "NULL argument");
    std::vector<float> xm(PAD_IN, 0.0f), xi(PAD_IN, 1.0f), ym(size_t(m->n_out)), yi(size_t(m->n_out)),
        ys(size_t(m->n_out));
    for (int c = 0; c < m->n_in; ++c) {
        if (!(x_std[c] > 0.0f)) return set_error(HADROS_ERR_ARG, "feature std must be positive");
        xm[c] = x_mean[c];
        xi[c] = 1.0f / x_std[c];
    }
    for (int c = 0; c < m->n_out; ++c) {
        if (!(y_std[c] > 0.0f)) return set_error(HADROS_ERR_ARG, "target std must be positive");
        ym[c] = y_mean[c];
        yi[c] = 1.0f / y_std[c];
        ys[c] = y_std[c];
    }
    MC(hipMemcpy(m->xm, xm.data(), PAD_IN * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->xi, xi.data(), PAD_IN * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->ym, ym.data(), ym.size() * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->yi, yi.data(), yi.size() * 4, hipMemcpyHostToDevice));
    MC(hipMemcpy(m->ys, ys.data(), ys.size() * 4, hipMemcpyHostToDevice));
    return HADROS_OK;
}

int hadros_mlp_forward(hadros_mlp* m, const float* x, int64_t n, float* y, void* stream)
{
    if (!m || (n > 0 && (!x || !y)) || n < 0) return set_error(HADROS_ERR_ARG, "bad forward arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (m->fused_ok && m->infer_path == 0) return infer_fused(m, x, n, y, 0, s);
    for (int64_t r0 = 0; r0 < n; r0 += m->ib) {
        const long rows = long(std::min<int64_t>(m->ib, n - r0));
        MC(mlp_pack(x + r0 * m->n_in, rows, m->n_in, PAD_IN, m->xm, m->xi, m->Xi, s));
        const __bf16* A = m->Xi;
        for (int l = 0; l < m->L; ++l) {
            GemmArgs g{};
            g.A = A;
            g.B = m->W16 + m->woff[l];
            g.M = int(rows);
            g.N = m->out[l];
            g.K = m->in[l];
            g.lda = m->in[l];
            g.ldb = m->out[l];
            g.ldc = m->out[l];
            g.bias = m->P + m->boff[l];
            if (l < m->L - 1) {
                __bf16* Hn = (A == m->Ha) ? m->Hb : m->Ha;
                g.Cb = Hn;
                MC(mlp_gemm(g, false, false, EPI_FWD_SILU, 1, s));
                A = Hn;
            } else {
                g.Cf = y + r0 * m->n_out;
                g.scale = m->ys;
                g.shift = m->ym;
                MC(mlp_gemm(g, false, false, EPI_FWD_LINEAR, 1, s));
            }
        }
    }
    return HADROS_OK;
}

int hadros_mlp_survival(hadros_mlp* m, const float* x, int64_t n, float* p, void* stream)
{
    if (!m || (n > 0 && (!x || !p)) || n < 0) return set_error(HADROS_ERR_ARG, "bad survival arguments");
    if (!m->fused_ok || m->n_out % 2)
        return set_error(HADROS_ERR_ARG, "survival output needs width 256, an even n_out <= 16 and <= 7 hidden layers");
    return infer_fused(m, x, n, p, 1, reinterpret_cast<hipStream_t>(stream));
}

int hadros_mlp_set_infer_path(hadros_mlp* m, int path)
{
    if (!m || path < 0 || path > 1) return set_error(HADROS_ERR_ARG, "path must be 0 (fused) or 1 (per-layer GEMMs)");
    if (path == 0 && !m->fused_ok) return set_error(HADROS_ERR_ARG, "this network shape has no fused path");
    m->infer_path = path;
    return HADROS_OK;
}

int hadros_mlp_grad(hadros_mlp* m, const float* x, const float* y, int64_t n, float* grad, float* loss)
{
    if (!m) return set_error(HADROS_ERR_ARG, "NULL mlp");
    int rc = loss_and_grad(m, x, y, long(n), loss, nullptr);
    if (rc) return rc;
    if (grad) {
        std::vector<float> g(m->n_int);
        MC(hipMemcpy(g.data(), m->G, m->n_int * 4, hipMemcpyDeviceToHost));
        map_public(m, [&](size_t q, size_t i) { grad[q] = g[i]; });
    }
    return HADROS_OK;
}

int hadros_mlp_train_step(hadros_mlp* m, const float* x, const float* y, int64_t n, float lr, float* loss)
{
    if (!m) return set_error(HADROS_ERR_ARG, "NULL mlp");
    int rc = loss_and_grad(m, x, y, long(n), loss, nullptr);
    if (rc) return rc;
    ++m->step;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, float(m->step)), c2 = 1.0f - std::pow(b2, float(m->step));
    MC(mlp_adam(m->P, m->G, m->Mo, m->Vo, long(m->n_int), lr, b1, b2, eps, c1, c2, m->W16, nullptr));
    m->sw_dirty = true;
    return HADROS_OK;
}

int hadros_mlp_destroy(hadros_mlp* m)
{
    if (!m) return HADROS_OK;
    dfr(m->P);
    dfr(m->G);
    dfr(m->Mo);
    dfr(m->Vo);
    dfr(m->W16);
    dfr(m->Wsw);
    dfr(m->xm);
    dfr(m->xi);
    dfr(m->ym);
    dfr(m->yi);
    dfr(m->ys);
    dfr(m->zero);
    dfr(m->one);
    dfr(m->Xb);
    for (auto& h : m->H) dfr(h);
    for (auto& z : m->Z) dfr(z);
    dfr(m->Y);
    dfr(m->dY);
    dfr(m->Ga);
    dfr(m->Gb);
    dfr(m->part);
    dfr(m->lpart);
    dfr(m->Xi);
    dfr(m->Ha);
    dfr(m->Hb);
    delete m;
    return HADROS_OK;
}

}  // extern "C"
