// nn.hip -- gfx950 kernels of the learned preconditioner (a small conv
// encoder-decoder).  Convolutions are implicit GEMMs on the CDNA4 matrix
// cores: v_mfma_f32_16x16x32_bf16, bf16 operands, fp32 accumulation.
//
// Operand layout of one 16x16x32 MFMA (wave64): lane l holds row l%16 of A
// and column l%16 of B, each with the 8 consecutive k values 8*(l/16) .. +7;
// the result gives lane l the 4 values D[4*(l/16) + e][l%16], e = 0..3.
// mfma_probe() checks this layout against a plain matrix product.
#include "nn_kernels.h"

namespace fwi {
namespace nn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
    return z;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

inline int blocks_for(long n, int bs, int cap = 65536)
{
    long b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// ---------------------------------------------------------------- conv fwd

// Block: 4 waves; the block's weights [coutp][kp] sit in LDS when they fit
// in 64 KiB (otherwise the B operands are read from global memory, where
// they stay L2-resident); each wave computes 32 output pixels x coutp
// channels per tile (2 x coutp/16 MFMA accumulators) and walks the K
// dimension (tap, ci) in chunks of 32.
constexpr int kLdsWeightBytes = 65536;

template <bool LDS>
__global__ __launch_bounds__(256) void conv_fwd_kernel(ConvFwd a)
{
    extern __shared__ bf16 sw[];
    const bf16 *wsrc = a.wgt;
    if constexpr (LDS) {
        const int nw = a.coutp * a.kp;
        for (int i = threadIdx.x * 8; i < nw; i += 256 * 8)
            *reinterpret_cast<bf16x8 *>(sw + i) = ld8(a.wgt + i);
        __syncthreads();
        wsrc = sw;
    }

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int hw = a.h * a.w;
    const int npix = a.nb * hw;
    const int ntile = (npix + 31) / 32;
    const int nblk = a.coutp / 16;
    const int nkc = a.kp / 32;
    for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
        f32x4 acc[2][4];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int n = 0; n < 4; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        int pn[2], ph[2], pw[2];
        bool pv[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int p = tile * 32 + m * 16 + r16;
            pv[m] = p < npix;
            const int q = pv[m] ? p : 0;
            pn[m] = q / hw;
            const int rem = q - pn[m] * hw;
            ph[m] = rem / a.w;
            pw[m] = rem - ph[m] * a.w;
        }
        for (int kc = 0; kc < nkc; ++kc) {
            const int k0 = kc * 32 + g * 8;
            const int tap = k0 / a.cin, ci = k0 - tap * a.cin;
            bf16x8 A[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                A[m] = zero8();
                if (pv[m] && tap < 9) {
                    const int hh = ph[m] + tap / 3 - 1, ww = pw[m] + tap % 3 - 1;
                    if (hh >= 0 && hh < a.h && ww >= 0 && ww < a.w)
                        A[m] = ld8(a.x + ((long)(pn[m] * a.h + hh) * a.w + ww) * a.cin + ci);
                }
            }
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n < nblk) {
                    const bf16x8 B = *reinterpret_cast<const bf16x8 *>(wsrc + (n * 16 + r16) * a.kp + k0);
                    acc[0][n] = mfma(A[0], B, acc[0][n]);
                    acc[1][n] = mfma(A[1], B, acc[1][n]);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n >= nblk) continue;
                const int co = n * 16 + r16;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const long p = (long)tile * 32 + m * 16 + 4 * g + e;
                    if (p >= npix || co >= a.cout) continue;
                    float v = acc[m][n][e];
                    if (a.bias) v += a.bias[co];
                    const long idx = p * a.cout_stride + co;
                    if (a.addsrc) v += (float)a.addsrc[idx];
                    if (a.relu) v = fmaxf(v, 0.f);
                    if (a.mask && !((float)a.mask[idx] > 0.f)) v = 0.f;
                    if (a.yf) a.yf[idx] = v;
                    else a.y[idx] = (bf16)v;
                }
            }
    }
}

// ---------------------------------------------------------------- conv wgrad

// dW[co][tap][ci] = sum_p dy[p][co] x[p + offset(tap)][ci] as MFMAs with the
// pixels as the K dimension.  A block walks tiles of 32 consecutive pixels of
// one image row.  Per tile it stages dy (32 x coutp) transposed, and the
// input rows h-1 .. h+1, columns w0-1 .. w0+32, transposed and stored once
// per column shift, so that every MFMA operand (8 consecutive pixels of one
// channel) is an aligned 16-byte LDS read.  Wave w owns output channels
// 16w .. 16w+15 and accumulates all 9 taps x all input-channel blocks over
// its tiles; the per-block partial sums are reduced by wgrad_reduce.
constexpr int kPixRow = 40;   // LDS row: 32 pixels + 8 padding

template <int CIB>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(ConvWgrad a)
{
    constexpr int CI = CIB * 16;
    __shared__ __attribute__((aligned(16))) bf16 sdy[64 * kPixRow];            // [co][p]
    __shared__ __attribute__((aligned(16))) bf16 sx[3 * CI * 3 * kPixRow];     // [shift][ci][row][p]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int nwt = (a.w + 31) >> 5;
    const long ntiles = (long)a.nb * a.h * nwt;
    const bool has_m = wave * 16 < a.coutp;
    f32x4 acc[9][CIB];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CIB; ++c) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;

    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int n = (int)(tile / ((long)a.h * nwt));
        const int rem = (int)(tile - (long)n * a.h * nwt);
        const int h = rem / nwt, w0 = (rem - h * nwt) * 32;
        // dy: 8 channels of one pixel per thread (pixel index fastest)
        for (int e = tid; e < 32 * (a.coutp >> 3); e += 256) {
            const int p = e & 31, c8 = (e >> 5) * 8;
            bf16x8 v = zero8();
            if (w0 + p < a.w && c8 < a.dy_stride)
                v = ld8(a.dy + ((long)(n * a.h + h) * a.w + w0 + p) * a.dy_stride + c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) sdy[(c8 + k) * kPixRow + p] = v[k];
        }
        // x: sx[s][ci][r][p] = x[row h-1+r][column w0-1+p+s][ci]
        for (int e = tid; e < 3 * 34 * (CI >> 3); e += 256) {
            const int col = e % 34, t2 = e / 34;
            const int r = t2 % 3, c8 = (t2 / 3) * 8;
            const int hh = h - 1 + r, ww = w0 - 1 + col;
            bf16x8 v = zero8();
            if (hh >= 0 && hh < a.h && ww >= 0 && ww < a.w && c8 < a.cin)
                v = ld8(a.x + ((long)(n * a.h + hh) * a.w + ww) * a.cin + c8);
#pragma unroll
            for (int sh = 0; sh < 3; ++sh) {
                const int p = col - sh;
                if (p >= 0 && p < 32) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) sx[((sh * CI + c8 + k) * 3 + r) * kPixRow + p] = v[k];
                }
            }
        }
        __syncthreads();
        if (has_m) {
            const bf16x8 A =
                *reinterpret_cast<const bf16x8 *>(sdy + (wave * 16 + r16) * kPixRow + g * 8);
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int r = tap / 3, sh = tap % 3;
#pragma unroll
                for (int cb = 0; cb < CIB; ++cb) {
                    const bf16x8 B = *reinterpret_cast<const bf16x8 *>(
                        sx + ((sh * CI + cb * 16 + r16) * 3 + r) * kPixRow + g * 8);
                    acc[tap][cb] = mfma(A, B, acc[tap][cb]);
                }
            }
        }
        if (tid < a.coutp)
            for (int p = 0; p < 32; ++p) bsum += (float)sdy[tid * kPixRow + p];
        __syncthreads();
    }
    if (has_m) {
        float *part = a.part + (long)blockIdx.x * a.coutp * 9 * CI;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
#pragma unroll
            for (int cb = 0; cb < CIB; ++cb)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int co = wave * 16 + 4 * g + e;
                    const int ci = cb * 16 + r16;
                    part[((long)co * 9 + tap) * CI + ci] = acc[tap][cb][e];
                }
    }
    if (tid < a.coutp) a.partb[(long)blockIdx.x * a.coutp + tid] = bsum;
}

// dw / db = sum over the blocks' partial sums.  A block covers 32
// consecutive outputs (weights, then biases) with 8 threads each, which sum
// every 8th partial; the 8 sums are added in a fixed order (deterministic).
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float *part, const float *partb,
                                                           int nchunk, int cout, int cin,
                                                           int cin16, int coutp, float *dw,
                                                           float *db)
{
    __shared__ float red[8][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const long n = (long)cout * 9 * cin;
    const long idx = (long)blockIdx.x * 32 + tx;
    float s = 0.f;
    if (idx < n) {
        const int ci = (int)(idx % cin);
        const int tap = (int)((idx / cin) % 9);
        const int co = (int)(idx / (9L * cin));
        const long src = ((long)co * 9 + tap) * cin16 + ci;
        const long stride = (long)coutp * 9 * cin16;
        for (int k = ty; k < nchunk; k += 8) s += part[(long)k * stride + src];
    } else if (idx < n + cout) {
        const int co = (int)(idx - n);
        for (int k = ty; k < nchunk; k += 8) s += partb[(long)k * coutp + co];
    }
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && idx < n + cout) {
        float t = red[0][tx];
#pragma unroll
        for (int j = 1; j < 8; ++j) t += red[j][tx];
        if (idx < n) dw[idx] = t;
        else db[idx - n] = t;
    }
}

// ---------------------------------------------------------------- weights

__global__ void pack_weights_kernel(const float *wm, int cout, int cin, bf16 *wf, int coutp,
                                    int kp, bf16 *wb, int cinp16, int cout8, int kpb)
{
    const long nf = (long)coutp * kp, nbk = (long)cinp16 * kpb;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < nf + nbk;
         idx += (long)gridDim.x * blockDim.x) {
        if (idx < nf) {                       // forward: wf[co][tap*cin + ci]
            const int co = (int)(idx / kp), k = (int)(idx % kp);
            const int tap = k / cin, ci = k % cin;
            float v = 0.f;
            if (co < cout && tap < 9) v = wm[((long)co * 9 + tap) * cin + ci];
            wf[idx] = (bf16)v;
        } else {                              // backward data: wb[ci][tap'*cout8 + co]
            const long j = idx - nf;
            const int ci = (int)(j / kpb), k = (int)(j % kpb);
            const int tapb = k / cout8, co = k % cout8;
            float v = 0.f;
            if (ci < cin && tapb < 9 && co < cout) v = wm[((long)co * 9 + (8 - tapb)) * cin + ci];
            wb[j] = (bf16)v;
        }
    }
}

// ---------------------------------------------------------------- elementwise

__global__ void pack_input_kernel(const float *x, int nb, int creal, int h, int w, bf16 *out,
                                  int cpad)
{
    const long n = (long)nb * h * w * cpad;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int c = (int)(idx % cpad);
        const long pix = idx / cpad;
        const int ww = (int)(pix % w);
        const int hh = (int)((pix / w) % h);
        const int b = (int)(pix / ((long)w * h));
        float v = 0.f;
        if (c < creal) v = x[(((long)b * creal + c) * h + hh) * w + ww];
        out[idx] = (bf16)v;
    }
}

__global__ void pool_fwd_kernel(const bf16 *in, int nb, int h, int w, int c, bf16 *out)
{
    const int h2 = h / 2, w2 = w / 2;
    const long n = (long)nb * h2 * w2 * c;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(idx % c);
        const long pix = idx / c;
        const int x2 = (int)(pix % w2), y2 = (int)((pix / w2) % h2);
        const int b = (int)(pix / ((long)w2 * h2));
        const long base = (((long)b * h + 2 * y2) * w + 2 * x2) * c + ch;
        const float s = (float)in[base] + (float)in[base + c] + (float)in[base + (long)w * c] +
                        (float)in[base + (long)w * c + c];
        out[idx] = (bf16)(0.25f * s);
    }
}

__global__ void pool_bwd_kernel(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c,
                                bf16 *din)
{
    const long n = (long)nb * h * w * c;
    const int h2 = h / 2, w2 = w / 2;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(idx % c);
        const long pix = idx / c;
        const int xx = (int)(pix % w), yy = (int)((pix / w) % h);
        const int b = (int)(pix / ((long)w * h));
        float v = 0.25f * (float)dout[(((long)b * h2 + yy / 2) * w2 + xx / 2) * c + ch];
        if (mask && !((float)mask[idx] > 0.f)) v = 0.f;
        din[idx] = (bf16)v;
    }
}

__global__ void up_add_fwd_kernel(const bf16 *in, const bf16 *skip, int nb, int h, int w, int c,
                                  bf16 *out)
{
    // in: [nb][h/2][w/2][c]; out, skip: [nb][h][w][c]
    const long n = (long)nb * h * w * c;
    const int h2 = h / 2, w2 = w / 2;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(idx % c);
        const long pix = idx / c;
        const int xx = (int)(pix % w), yy = (int)((pix / w) % h);
        const int b = (int)(pix / ((long)w * h));
        const float v = (float)in[(((long)b * h2 + yy / 2) * w2 + xx / 2) * c + ch] +
                        (float)skip[idx];
        out[idx] = (bf16)v;
    }
}

__global__ void up_bwd_kernel(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c,
                              bf16 *din)
{
    // dout: [nb][h][w][c]; din, mask: [nb][h/2][w/2][c]
    const int h2 = h / 2, w2 = w / 2;
    const long n = (long)nb * h2 * w2 * c;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(idx % c);
        const long pix = idx / c;
        const int x2 = (int)(pix % w2), y2 = (int)((pix / w2) % h2);
        const int b = (int)(pix / ((long)w2 * h2));
        const long base = (((long)b * h + 2 * y2) * w + 2 * x2) * c + ch;
        float v = (float)dout[base] + (float)dout[base + c] + (float)dout[base + (long)w * c] +
                  (float)dout[base + (long)w * c + c];
        if (mask && !((float)mask[idx] > 0.f)) v = 0.f;
        din[idx] = (bf16)v;
    }
}

__global__ __launch_bounds__(256) void mse_kernel(const float *y, const float *t, long n,
                                                  bf16 *dy8, float *partials)
{
    __shared__ float red[256];
    float s = 0.f;
    const float scale = 2.0f / (float)n;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        s += d * d;
        dy8[i * 8] = (bf16)(scale * d);
#pragma unroll
        for (int c = 1; c < 8; ++c) dy8[i * 8 + c] = (bf16)0.0f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_f32_kernel(const float *p, int n, float *out)
{
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void adam_kernel(float *w, const float *g, float *m, float *v, long n, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / c1, vh = vi / c2;
        w[i] -= lr * mh / (sqrtf(vh) + eps);
    }
}

__global__ void mfma_probe_kernel(const float *a, const float *b, float *d)
{
    const int lane = threadIdx.x, g = lane >> 4, r16 = lane & 15;
    bf16x8 A, B;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        A[e] = (bf16)a[r16 * 32 + g * 8 + e];      // A[row r16][k]
        B[e] = (bf16)b[(g * 8 + e) * 16 + r16];    // B[k][col r16]
    }
    f32x4 acc = mfma(A, B, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int e = 0; e < 4; ++e) d[(4 * g + e) * 16 + r16] = acc[e];
}

}  // namespace

hipError_t conv_fwd(const ConvFwd &a, hipStream_t s)
{
    const long npix = (long)a.nb * a.h * a.w;
    const int ntile = (int)((npix + 31) / 32);
    const int nblocks = blocks_for(ntile, 4, 2048);
    const size_t lds = (size_t)a.coutp * a.kp * sizeof(bf16);
    if (lds <= (size_t)kLdsWeightBytes)
        hipLaunchKernelGGL(conv_fwd_kernel<true>, dim3(nblocks), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL(conv_fwd_kernel<false>, dim3(nblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t conv_wgrad(const ConvWgrad &a, int nblocks, hipStream_t s)
{
    switch (a.cin16) {
    case 16: hipLaunchKernelGGL(conv_wgrad_kernel<1>, dim3(nblocks), dim3(256), 0, s, a); break;
    case 32: hipLaunchKernelGGL(conv_wgrad_kernel<2>, dim3(nblocks), dim3(256), 0, s, a); break;
    case 64: hipLaunchKernelGGL(conv_wgrad_kernel<4>, dim3(nblocks), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t wgrad_reduce(const float *part, const float *partb, int nchunk, int cout, int cin,
                        int cin16, int coutp, float *dw, float *db, hipStream_t s)
{
    const long n = (long)cout * 9 * cin + cout;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, s,
                       part, partb, nchunk, cout, cin, cin16, coutp, dw, db);
    return hipGetLastError();
}

hipError_t pack_weights(const float *wm, int cout, int cin, bf16 *wf, int coutp, int kp,
                        bf16 *wb, int cinp16, int cout8, int kpb, hipStream_t s)
{
    const long n = (long)coutp * kp + (long)cinp16 * kpb;
    hipLaunchKernelGGL(pack_weights_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, wm, cout,
                       cin, wf, coutp, kp, wb, cinp16, cout8, kpb);
    return hipGetLastError();
}

hipError_t pack_input(const float *x, int nb, int creal, int h, int w, bf16 *out, int cpad,
                      hipStream_t s)
{
    const long n = (long)nb * h * w * cpad;
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, nb, creal,
                       h, w, out, cpad);
    return hipGetLastError();
}

hipError_t pool_fwd(const bf16 *in, int nb, int h, int w, int c, bf16 *out, hipStream_t s)
{
    const long n = (long)nb * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(pool_fwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, nb, h, w,
                       c, out);
    return hipGetLastError();
}

hipError_t pool_bwd(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c, bf16 *din,
                    hipStream_t s)
{
    const long n = (long)nb * h * w * c;
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, dout, mask, nb,
                       h, w, c, din);
    return hipGetLastError();
}

hipError_t up_add_fwd(const bf16 *in, const bf16 *skip, int nb, int h, int w, int c, bf16 *out,
                      hipStream_t s)
{
    const long n = (long)nb * h * w * c;
    hipLaunchKernelGGL(up_add_fwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, skip, nb,
                       h, w, c, out);
    return hipGetLastError();
}

hipError_t up_bwd(const bf16 *dout, const bf16 *mask, int nb, int h, int w, int c, bf16 *din,
                  hipStream_t s)
{
    const long n = (long)nb * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(up_bwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, dout, mask, nb,
                       h, w, c, din);
    return hipGetLastError();
}

hipError_t mse(const float *y, const float *t, long n, bf16 *dy8, float *partials, int nblocks,
               hipStream_t s)
{
    hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(256), 0, s, y, t, n, dy8, partials);
    return hipGetLastError();
}

hipError_t sum_f32(const float *partials, int n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_f32_kernel, dim3(1), dim3(256), 0, s, partials, n, out);
    return hipGetLastError();
}

hipError_t adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                float b2, float eps, int t, hipStream_t s)
{
    const float c1 = 1.f - powf(b1, (float)t);
    const float c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, w, g, m, v, n, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t mfma_probe(const float *a, const float *b, float *d, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, a, b, d);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace fwi
