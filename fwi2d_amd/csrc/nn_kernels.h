// nn_kernels.h -- launch interface of the learned-preconditioner kernels
// (device code in nn.hip).  Activations are NHWC bf16 with the channel count
// padded to a multiple of 8; MFMA tiles are 16 x 16 x 32 (bf16 in, fp32 acc).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fwi {
namespace nn {

typedef __bf16 bf16;

// 3x3 same-padding convolution as an implicit GEMM on MFMA:
// y[p][co] = epilogue(sum_{tap,ci} x[p + off(tap)][ci] w[co][tap*cin + ci] + bias[co]).
// Epilogue order: + bias, + addsrc[p][co], relu, then zero where mask[p][co] <= 0.
struct ConvFwd {
    const bf16 *x;        // [npix][cin]
    int cin;              // multiple of 8
    const bf16 *wgt;      // [coutp][kp], kp = ceil(9*cin/32)*32, zero padded
    int kp;
    int coutp;            // output channels rounded up to 16 (<= 64)
    int cout;             // channels written
    int cout_stride;      // channel stride of y / yf / addsrc / mask
    const float *bias;    // [cout] or null
    const bf16 *addsrc;   // or null
    const bf16 *mask;     // or null
    int relu;
    bf16 *y;              // bf16 output, or
    float *yf;            // fp32 output
    int nb, h, w;         // images, rows (slow), columns (fast)
};
hipError_t conv_fwd(const ConvFwd &a, hipStream_t s);

// Weight gradient partials: part[chunk][co][tap][ci16] = sum over the chunk's
// pixels of dy[p][co] x[p + off(tap)][ci]; partb[chunk][co] = sum dy[p][co].
struct ConvWgrad {
    const bf16 *x;
    int cin, cin16;       // cin: channel stride of x (multiple of 8); cin16: 16, 32 or 64
    const bf16 *dy;
    int coutp, dy_stride; // coutp: 16, 32 or 64; dy_stride: channel stride of dy
    float *part, *partb;  // per-block partial sums [block][coutp][9][cin16], [block][coutp]
    int nb, h, w;
};
// nblocks blocks share the 32-pixel row tiles; each writes its partial sums.
hipError_t conv_wgrad(const ConvWgrad &a, int nblocks, hipStream_t s);
// dw[co][tap][ci] (padded master layout cin) = sum_chunk part; db[co] likewise.
hipError_t wgrad_reduce(const float *part, const float *partb, int nchunk, int cout, int cin,
                        int cin16, int coutp, float *dw, float *db, hipStream_t s);

// master fp32 [cout][9][cin] -> forward bf16 [coutp][kp] and backward
// (flipped, transposed) bf16 [cinp16][kpb] with kpb = ceil(9*cout8/32)*32.
hipError_t pack_weights(const float *wm, int cout, int cin, bf16 *wf, int coutp, int kp,
                        bf16 *wb, int cinp16, int cout8, int kpb, hipStream_t s);

// x[b][c][h][w] fp32 (c < creal) -> NHWC bf16 with cpad channels.
hipError_t pack_input(const float *x, int nb, int creal, int h, int w, bf16 *out, int cpad,
                      hipStream_t s);
// 2x2 average pool (fwd) and its adjoint with an optional relu mask of the
// pool's input (bwd): din = 0.25 dout, zeroed where mask <= 0.
hipError_t pool_fwd(const bf16 *in, int nb, int h, int w, int c, bf16 *out, hipStream_t s);
hipError_t pool_bwd(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c, bf16 *din,
                    hipStream_t s);
// nearest 2x upsample plus a skip tensor (fwd), and the adjoint of the
// upsample with an optional relu mask of its input (bwd).
hipError_t up_add_fwd(const bf16 *in, const bf16 *skip, int nb, int h, int w, int c, bf16 *out,
                      hipStream_t s);
hipError_t up_bwd(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c, bf16 *din,
                  hipStream_t s);
// MSE: loss partials and dy = 2 (y - t) / n into channel 0 of an 8-channel bf16 tensor.
hipError_t mse(const float *y, const float *t, long n, bf16 *dy8, float *partials, int nblocks,
               hipStream_t s);
hipError_t sum_f32(const float *partials, int n, float *out, hipStream_t s);
// Adam on n parameters (step index t >= 1).
hipError_t adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                float b2, float eps, int t, hipStream_t s);
// One 16x16x32 MFMA of row-major a[16][32], b[32][16] -> d[16][16] (layout check).
hipError_t mfma_probe(const float *a, const float *b, float *d, hipStream_t s);

}  // namespace nn
}  // namespace fwi
