// nn_engine.hip -- host side of the learned-preconditioner ABI (fwi2d_nn.h):
// buffers, the forward / backward schedule of the 8-layer encoder-decoder
// and the Adam update, all on one HIP stream.
#include "fwi2d.h"
#include "fwi2d_nn.h"
#include "nn_kernels.h"
#include "status.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

using fwi::fail;
using fwi::nn::bf16;

namespace {

#define NN_TRY(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(FWI_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

inline int up8(int c) { return (c + 7) & ~7; }
inline int up16(int c) { return (c + 15) & ~15; }
inline int up32(int c) { return (c + 31) & ~31; }

struct Layer {
    int cin_real;    // external input channels (3 for the first layer)
    int cin;         // padded input channels (multiple of 8)
    int cout;
    int coutp;       // cout rounded up to 16 (MFMA N)
    int cout8;       // cout rounded up to 8 (channel stride of its gradient)
    int kp;          // forward K = 9*cin rounded up to 32
    int cinp16;      // cin rounded up to 16 (backward-data N)
    int kpb;         // backward-data K = 9*cout8 rounded up to 32
    long woff, boff; // offsets in the master parameter array
    int h, w;        // spatial size this layer runs at
    bf16 *wf = nullptr, *wb = nullptr;
};

}  // namespace

struct fwi_nn {
    int B = 0, H = 0, W = 0;
    hipStream_t stream = nullptr;
    Layer L[8];
    long nparams = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    int t = 0;
    // activations
    bf16 *x0 = nullptr, *a1 = nullptr, *a2 = nullptr, *p2 = nullptr, *a3 = nullptr,
         *a4 = nullptr, *p4 = nullptr, *a5 = nullptr, *s5 = nullptr, *a6 = nullptr,
         *s6 = nullptr, *a7 = nullptr;
    float *y = nullptr, *target = nullptr, *xin = nullptr;
    // gradients
    bf16 *dy8 = nullptr, *dz7 = nullptr, *ds6 = nullptr, *dz6 = nullptr, *ds5 = nullptr,
         *dz5 = nullptr, *dp4 = nullptr, *dz4 = nullptr, *dz3 = nullptr, *dp2 = nullptr,
         *dz2 = nullptr, *dz1 = nullptr;
    float *part = nullptr, *partb = nullptr;
    int maxchunks = 512;         // blocks of the weight-gradient kernel (partial sums)
    float *loss_part = nullptr, *loss = nullptr;
    int nloss = 512;
    std::vector<void *> allocs;
};

namespace {

template <typename T>
int nn_alloc(fwi_nn *n, T **p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(FWI_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    n->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, count * sizeof(T), n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) return fail(FWI_ERR_HIP, "zero-fill failed: %s", hipGetErrorString(e));
    return FWI_OK;
}

#define TRY(expr)                    \
    do {                             \
        int r_ = (expr);             \
        if (r_ != FWI_OK) return r_; \
    } while (0)

uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

int pack_all(fwi_nn *n)
{
    for (Layer &l : n->L)
        NN_TRY(fwi::nn::pack_weights(n->params + l.woff, l.cout, l.cin, l.wf, l.coutp, l.kp, l.wb,
                                     l.cinp16, l.cout8, l.kpb, n->stream));
    return FWI_OK;
}

// y = conv(x) with the layer's forward weights.
int conv_forward(fwi_nn *n, const Layer &l, int nb, const bf16 *x, bf16 *y, float *yf, bool relu)
{
    fwi::nn::ConvFwd a{};
    a.x = x;
    a.cin = l.cin;
    a.wgt = l.wf;
    a.kp = l.kp;
    a.coutp = l.coutp;
    a.cout = l.cout;
    a.cout_stride = l.cout;
    a.bias = n->params + l.boff;
    a.relu = relu ? 1 : 0;
    a.y = y;
    a.yf = yf;
    a.nb = nb;
    a.h = l.h;
    a.w = l.w;
    NN_TRY(fwi::nn::conv_fwd(a, n->stream));
    return FWI_OK;
}

// dx = conv^T(dy) (+ addsrc), zeroed where mask <= 0.
int conv_backward_data(fwi_nn *n, const Layer &l, int nb, const bf16 *dy, const bf16 *addsrc,
                       const bf16 *mask, bf16 *dx)
{
    fwi::nn::ConvFwd a{};
    a.x = dy;
    a.cin = l.cout8;
    a.wgt = l.wb;
    a.kp = l.kpb;
    a.coutp = l.cinp16;
    a.cout = l.cin;
    a.cout_stride = l.cin;
    a.addsrc = addsrc;
    a.mask = mask;
    a.y = dx;
    a.nb = nb;
    a.h = l.h;
    a.w = l.w;
    NN_TRY(fwi::nn::conv_fwd(a, n->stream));
    return FWI_OK;
}

int conv_backward_weights(fwi_nn *n, const Layer &l, int nb, const bf16 *x, const bf16 *dy)
{
    const long ntiles = (long)nb * l.h * ((l.w + 31) / 32);
    const int nchunk = (int)std::min<long>(n->maxchunks, ntiles);
    fwi::nn::ConvWgrad a{};
    a.x = x;
    a.cin = l.cin;
    a.cin16 = up16(l.cin);
    a.dy = dy;
    a.coutp = l.coutp;
    a.dy_stride = l.cout8;
    a.part = n->part;
    a.partb = n->partb;
    a.nb = nb;
    a.h = l.h;
    a.w = l.w;
    NN_TRY(fwi::nn::conv_wgrad(a, nchunk, n->stream));
    NN_TRY(fwi::nn::wgrad_reduce(n->part, n->partb, nchunk, l.cout, l.cin, a.cin16, l.coutp,
                                 n->grads + l.woff, n->grads + l.boff, n->stream));
    return FWI_OK;
}

int forward(fwi_nn *n, int nb, const float *x_host)
{
    const long nin = (long)nb * 3 * n->H * n->W;
    NN_TRY(hipMemcpyAsync(n->xin, x_host, nin * 4, hipMemcpyHostToDevice, n->stream));
    NN_TRY(fwi::nn::pack_input(n->xin, nb, 3, n->H, n->W, n->x0, 8, n->stream));
    const int H = n->H, W = n->W;
    Layer *L = n->L;
    TRY(conv_forward(n, L[0], nb, n->x0, n->a1, nullptr, true));
    TRY(conv_forward(n, L[1], nb, n->a1, n->a2, nullptr, true));
    NN_TRY(fwi::nn::pool_fwd(n->a2, nb, H, W, 32, n->p2, n->stream));
    TRY(conv_forward(n, L[2], nb, n->p2, n->a3, nullptr, true));
    TRY(conv_forward(n, L[3], nb, n->a3, n->a4, nullptr, true));
    NN_TRY(fwi::nn::pool_fwd(n->a4, nb, H / 2, W / 2, 64, n->p4, n->stream));
    TRY(conv_forward(n, L[4], nb, n->p4, n->a5, nullptr, true));
    NN_TRY(fwi::nn::up_add_fwd(n->a5, n->a3, nb, H / 2, W / 2, 64, n->s5, n->stream));
    TRY(conv_forward(n, L[5], nb, n->s5, n->a6, nullptr, true));
    NN_TRY(fwi::nn::up_add_fwd(n->a6, n->a1, nb, H, W, 32, n->s6, n->stream));
    TRY(conv_forward(n, L[6], nb, n->s6, n->a7, nullptr, true));
    TRY(conv_forward(n, L[7], nb, n->a7, nullptr, n->y, false));
    return FWI_OK;
}

// Loss and parameter gradients (n->grads) for the batch in n->x0 / n->target.
int backward(fwi_nn *n, int nb, float *loss_out)
{
    const int H = n->H, W = n->W;
    Layer *L = n->L;
    const long npix = (long)nb * H * W;
    NN_TRY(fwi::nn::mse(n->y, n->target, npix, n->dy8, n->loss_part, n->nloss, n->stream));
    NN_TRY(fwi::nn::sum_f32(n->loss_part, n->nloss, n->loss, n->stream));
    // c8: a7 -> y
    TRY(conv_backward_weights(n, L[7], nb, n->a7, n->dy8));
    TRY(conv_backward_data(n, L[7], nb, n->dy8, nullptr, n->a7, n->dz7));
    // c7: s6 -> a7
    TRY(conv_backward_weights(n, L[6], nb, n->s6, n->dz7));
    TRY(conv_backward_data(n, L[6], nb, n->dz7, nullptr, nullptr, n->ds6));
    // s6 = up(a6) + a1
    NN_TRY(fwi::nn::up_bwd(n->ds6, n->a6, nb, H, W, 32, n->dz6, n->stream));
    // c6: s5 -> a6
    TRY(conv_backward_weights(n, L[5], nb, n->s5, n->dz6));
    TRY(conv_backward_data(n, L[5], nb, n->dz6, nullptr, nullptr, n->ds5));
    // s5 = up(a5) + a3
    NN_TRY(fwi::nn::up_bwd(n->ds5, n->a5, nb, H / 2, W / 2, 64, n->dz5, n->stream));
    // c5: p4 -> a5
    TRY(conv_backward_weights(n, L[4], nb, n->p4, n->dz5));
    TRY(conv_backward_data(n, L[4], nb, n->dz5, nullptr, nullptr, n->dp4));
    NN_TRY(fwi::nn::pool_bwd(n->dp4, n->a4, nb, H / 2, W / 2, 64, n->dz4, n->stream));
    // c4: a3 -> a4; a3 also feeds s5
    TRY(conv_backward_weights(n, L[3], nb, n->a3, n->dz4));
    TRY(conv_backward_data(n, L[3], nb, n->dz4, n->ds5, n->a3, n->dz3));
    // c3: p2 -> a3
    TRY(conv_backward_weights(n, L[2], nb, n->p2, n->dz3));
    TRY(conv_backward_data(n, L[2], nb, n->dz3, nullptr, nullptr, n->dp2));
    NN_TRY(fwi::nn::pool_bwd(n->dp2, n->a2, nb, H, W, 32, n->dz2, n->stream));
    // c2: a1 -> a2; a1 also feeds s6
    TRY(conv_backward_weights(n, L[1], nb, n->a1, n->dz2));
    TRY(conv_backward_data(n, L[1], nb, n->dz2, n->ds6, n->a1, n->dz1));
    // c1: x0 -> a1
    TRY(conv_backward_weights(n, L[0], nb, n->x0, n->dz1));
    if (loss_out) {
        float h = 0.f;
        NN_TRY(hipMemcpyAsync(&h, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
        NN_TRY(hipStreamSynchronize(n->stream));
        *loss_out = h / (float)npix;
    }
    return FWI_OK;
}

int check_nb(fwi_nn *n, int nb)
{
    if (!n) return fail(FWI_ERR_ARG, "null network");
    if (nb < 1 || nb > n->B) return fail(FWI_ERR_ARG, "batch %d outside [1, %d]", nb, n->B);
    return FWI_OK;
}

// external W[cout][3][3][cin_real] <-> master [cout][9][cin]
void to_external(const fwi_nn *n, const std::vector<float> &master, float *out)
{
    long o = 0;
    for (const Layer &l : n->L) {
        for (int co = 0; co < l.cout; ++co)
            for (int tap = 0; tap < 9; ++tap)
                for (int ci = 0; ci < l.cin_real; ++ci)
                    out[o++] = master[l.woff + ((long)co * 9 + tap) * l.cin + ci];
        for (int co = 0; co < l.cout; ++co) out[o++] = master[l.boff + co];
    }
}

void from_external(const fwi_nn *n, const float *in, std::vector<float> &master)
{
    master.assign(n->nparams, 0.f);
    long o = 0;
    for (const Layer &l : n->L) {
        for (int co = 0; co < l.cout; ++co)
            for (int tap = 0; tap < 9; ++tap)
                for (int ci = 0; ci < l.cin_real; ++ci)
                    master[l.woff + ((long)co * 9 + tap) * l.cin + ci] = in[o++];
        for (int co = 0; co < l.cout; ++co) master[l.boff + co] = in[o++];
    }
}

long external_count(const fwi_nn *n)
{
    long c = 0;
    for (const Layer &l : n->L) c += (long)l.cout * 9 * l.cin_real + l.cout;
    return c;
}

}  // namespace

extern "C" {

int fwi_nn_destroy(fwi_nn *n)
{
    if (!n) return FWI_OK;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return FWI_OK;
}

int fwi_nn_create(int32_t batch, int32_t height, int32_t width, uint64_t seed, fwi_nn **out)
{
    if (!out) return fail(FWI_ERR_ARG, "null argument");
    *out = nullptr;
    if (batch < 1 || height < 4 || width < 4 || height % 4 || width % 4)
        return fail(FWI_ERR_ARG, "bad network shape %d x %d x %d (H, W multiples of 4)", batch,
                    height, width);
    fwi_nn *n = new fwi_nn();
    n->B = batch;
    n->H = height;
    n->W = width;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return fail(FWI_ERR_HIP, "hipStreamCreate failed");
    }
    const int H = height, W = width;
    const int spec[8][4] = {   // cin_real, cout, h divisor, w divisor
        {3, 32, 1, 1}, {32, 32, 1, 1}, {32, 64, 2, 2}, {64, 64, 2, 2},
        {64, 64, 4, 4}, {64, 32, 2, 2}, {32, 32, 1, 1}, {32, 1, 1, 1}};
    long off = 0;
    for (int k = 0; k < 8; ++k) {
        Layer &l = n->L[k];
        l.cin_real = spec[k][0];
        l.cin = up8(l.cin_real);
        l.cout = spec[k][1];
        l.coutp = up16(l.cout);
        l.cout8 = up8(l.cout);
        l.kp = up32(9 * l.cin);
        l.cinp16 = up16(l.cin);
        l.kpb = up32(9 * l.cout8);
        l.h = H / spec[k][2];
        l.w = W / spec[k][3];
        l.woff = off;
        off += (long)l.cout * 9 * l.cin;
        l.boff = off;
        off += l.cout;
    }
    n->nparams = off;
    const long P = (long)batch * H * W;
    int r = FWI_OK;
    auto A = [&](auto **p, long count) {
        if (r == FWI_OK) r = nn_alloc(n, p, (size_t)count);
    };
    A(&n->params, off); A(&n->grads, off); A(&n->m, off); A(&n->v, off);
    for (Layer &l : n->L) {
        A(&l.wf, (long)l.coutp * l.kp);
        A(&l.wb, (long)l.cinp16 * l.kpb);
    }
    A(&n->x0, P * 8); A(&n->a1, P * 32); A(&n->a2, P * 32); A(&n->p2, P / 4 * 32);
    A(&n->a3, P / 4 * 64); A(&n->a4, P / 4 * 64); A(&n->p4, P / 16 * 64); A(&n->a5, P / 16 * 64);
    A(&n->s5, P / 4 * 64); A(&n->a6, P / 4 * 32); A(&n->s6, P * 32); A(&n->a7, P * 32);
    A(&n->y, P); A(&n->target, P); A(&n->xin, P * 3);
    A(&n->dy8, P * 8); A(&n->dz7, P * 32); A(&n->ds6, P * 32); A(&n->dz6, P / 4 * 32);
    A(&n->ds5, P / 4 * 64); A(&n->dz5, P / 16 * 64); A(&n->dp4, P / 16 * 64);
    A(&n->dz4, P / 4 * 64); A(&n->dz3, P / 4 * 64); A(&n->dp2, P / 4 * 32);
    A(&n->dz2, P * 32); A(&n->dz1, P * 32);
    A(&n->part, (long)n->maxchunks * 64 * 9 * 64);
    A(&n->partb, (long)n->maxchunks * 64);
    A(&n->loss_part, n->nloss);
    A(&n->loss, 1);
    if (r != FWI_OK) {
        fwi_nn_destroy(n);
        return r;
    }
    // seeded He-uniform weights: U(-a, a), a = sqrt(6 / fan_in); last layer x 0.1
    std::vector<float> master(off, 0.f);
    for (int k = 0; k < 8; ++k) {
        const Layer &l = n->L[k];
        const double a = std::sqrt(6.0 / (9.0 * l.cin_real)) * (k == 7 ? 0.1 : 1.0);
        for (int co = 0; co < l.cout; ++co)
            for (int tap = 0; tap < 9; ++tap)
                for (int ci = 0; ci < l.cin_real; ++ci) {
                    const long idx = l.woff + ((long)co * 9 + tap) * l.cin + ci;
                    const uint64_t h = mix64(seed ^ mix64((uint64_t)idx + 0x9E3779B97F4A7C15ULL));
                    const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
                    master[idx] = (float)((2.0 * u - 1.0) * a);
                }
    }
    hipError_t e = hipMemcpy(n->params, master.data(), off * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fwi_nn_destroy(n);
        return fail(FWI_ERR_HIP, "parameter upload: %s", hipGetErrorString(e));
    }
    r = pack_all(n);
    if (r == FWI_OK && hipStreamSynchronize(n->stream) != hipSuccess)
        r = fail(FWI_ERR_HIP, "weight packing failed");
    if (r != FWI_OK) {
        fwi_nn_destroy(n);
        return r;
    }
    *out = n;
    return FWI_OK;
}

int fwi_nn_get_shape(fwi_nn *n, int32_t *batch, int32_t *height, int32_t *width)
{
    if (!n) return fail(FWI_ERR_ARG, "null network");
    if (batch) *batch = n->B;
    if (height) *height = n->H;
    if (width) *width = n->W;
    return FWI_OK;
}

int fwi_nn_num_params(fwi_nn *n, int64_t *count)
{
    if (!n || !count) return fail(FWI_ERR_ARG, "null argument");
    *count = external_count(n);
    return FWI_OK;
}

int fwi_nn_get_params(fwi_nn *n, float *params)
{
    if (!n || !params) return fail(FWI_ERR_ARG, "null argument");
    std::vector<float> master(n->nparams);
    NN_TRY(hipStreamSynchronize(n->stream));
    NN_TRY(hipMemcpy(master.data(), n->params, n->nparams * 4, hipMemcpyDeviceToHost));
    to_external(n, master, params);
    return FWI_OK;
}

int fwi_nn_set_params(fwi_nn *n, const float *params)
{
    if (!n || !params) return fail(FWI_ERR_ARG, "null argument");
    std::vector<float> master;
    from_external(n, params, master);
    NN_TRY(hipStreamSynchronize(n->stream));
    NN_TRY(hipMemcpy(n->params, master.data(), n->nparams * 4, hipMemcpyHostToDevice));
    TRY(pack_all(n));
    NN_TRY(hipStreamSynchronize(n->stream));
    return FWI_OK;
}

int fwi_nn_get_optimizer(fwi_nn *n, float *m, float *v, int64_t *step)
{
    if (!n || !m || !v || !step) return fail(FWI_ERR_ARG, "null argument");
    std::vector<float> master(n->nparams);
    NN_TRY(hipStreamSynchronize(n->stream));
    NN_TRY(hipMemcpy(master.data(), n->m, n->nparams * 4, hipMemcpyDeviceToHost));
    to_external(n, master, m);
    NN_TRY(hipMemcpy(master.data(), n->v, n->nparams * 4, hipMemcpyDeviceToHost));
    to_external(n, master, v);
    *step = n->t;
    return FWI_OK;
}

int fwi_nn_set_optimizer(fwi_nn *n, const float *m, const float *v, int64_t step)
{
    if (!n || !m || !v) return fail(FWI_ERR_ARG, "null argument");
    if (step < 0 || step > 0x7fffffff)
        return fail(FWI_ERR_ARG, "Adam step count %lld out of range", (long long)step);
    const long count = external_count(n);
    for (long i = 0; i < count; ++i)
        if (!std::isfinite(m[i]) || !std::isfinite(v[i]) || v[i] < 0.f)
            return fail(FWI_ERR_ARG, "Adam moment %ld is not finite (or second moment < 0)", i);
    std::vector<float> mm, vv;
    from_external(n, m, mm);
    from_external(n, v, vv);
    NN_TRY(hipStreamSynchronize(n->stream));
    NN_TRY(hipMemcpy(n->m, mm.data(), n->nparams * 4, hipMemcpyHostToDevice));
    NN_TRY(hipMemcpy(n->v, vv.data(), n->nparams * 4, hipMemcpyHostToDevice));
    n->t = (int)step;
    return FWI_OK;
}

int fwi_nn_forward(fwi_nn *n, int32_t nb, const float *x, float *y)
{
    TRY(check_nb(n, nb));
    if (!x || !y) return fail(FWI_ERR_ARG, "null argument");
    TRY(forward(n, nb, x));
    const long npix = (long)nb * n->H * n->W;
    NN_TRY(hipMemcpyAsync(y, n->y, npix * 4, hipMemcpyDeviceToHost, n->stream));
    NN_TRY(hipStreamSynchronize(n->stream));
    return FWI_OK;
}

int fwi_nn_gradient(fwi_nn *n, int32_t nb, const float *x, const float *target, float *loss,
                    float *grads)
{
    TRY(check_nb(n, nb));
    if (!x || !target) return fail(FWI_ERR_ARG, "null argument");
    const long npix = (long)nb * n->H * n->W;
    NN_TRY(hipMemcpyAsync(n->target, target, npix * 4, hipMemcpyHostToDevice, n->stream));
    TRY(forward(n, nb, x));
    TRY(backward(n, nb, loss));
    if (grads) {
        std::vector<float> master(n->nparams);
        NN_TRY(hipMemcpyAsync(master.data(), n->grads, n->nparams * 4, hipMemcpyDeviceToHost,
                              n->stream));
        NN_TRY(hipStreamSynchronize(n->stream));
        to_external(n, master, grads);
    }
    NN_TRY(hipStreamSynchronize(n->stream));
    return FWI_OK;
}

int fwi_nn_train_step(fwi_nn *n, int32_t nb, const float *x, const float *target, float lr,
                      float *loss)
{
    TRY(check_nb(n, nb));
    if (!x || !target) return fail(FWI_ERR_ARG, "null argument");
    const long npix = (long)nb * n->H * n->W;
    NN_TRY(hipMemcpyAsync(n->target, target, npix * 4, hipMemcpyHostToDevice, n->stream));
    TRY(forward(n, nb, x));
    TRY(backward(n, nb, loss));
    n->t += 1;
    NN_TRY(fwi::nn::adam(n->params, n->grads, n->m, n->v, n->nparams, lr, 0.9f, 0.999f, 1e-8f,
                         n->t, n->stream));
    TRY(pack_all(n));
    NN_TRY(hipStreamSynchronize(n->stream));
    return FWI_OK;
}

int fwi_nn_mfma_probe(const float *a, const float *b, float *d)
{
    if (!a || !b || !d) return fail(FWI_ERR_ARG, "null argument");
    float *da = nullptr, *db = nullptr, *dd = nullptr;
    hipError_t e = hipMalloc((void **)&da, 16 * 32 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&db, 32 * 16 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dd, 16 * 16 * 4);
    if (e == hipSuccess) e = hipMemcpy(da, a, 16 * 32 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, 32 * 16 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = fwi::nn::mfma_probe(da, db, dd, nullptr);
    if (e == hipSuccess) e = hipMemcpy(d, dd, 16 * 16 * 4, hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dd) (void)hipFree(dd);
    if (e != hipSuccess) return fail(FWI_ERR_HIP, "mfma probe: %s", hipGetErrorString(e));
    return FWI_OK;
}

}  // extern "C"
