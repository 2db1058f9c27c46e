/*
 * hd_cnn.hip -- the flow surrogate's network on gfx950: 3x3 convolutions as
 * implicit GEMMs on the BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
 * accumulation), their data and weight gradients, masked-MSE loss, Adam, and
 * the packing between a hydist context's FP64 state and the network's
 * images (include/hydist_cnn.h).
 *
 * MFMA operand layout (16x16x32, wave64): lane l holds A[m = l%16][k =
 * 8(l/16) .. 8(l/16)+7], B[k = 8(l/16) .. +7][n = l%16], and D[m = 4(l/16) +
 * r][n = l%16] for r = 0..3 (checked by hd_cnn_mfma_check against a host
 * product).
 *
 * Forward conv: a workgroup of four waves covers four image rows of 64
 * pixels; a wave owns one row: 4 M-blocks (16 pixels) x COUT/16 N-blocks of
 * accumulators.  K runs over the 9 taps x CIN channels in chunks of 32; the
 * A fragment of a pixel is 8 consecutive channels (one 16-byte load, NHWC),
 * the B fragment 8 consecutive input channels of one output channel (one
 * 16-byte load, weights [tap][cout][cin]).  The data gradient is the same
 * kernel on the gradient image with the tap-flipped, transposed weights.
 * Weight gradient: K = pixels; 32-pixel chunks of the gradient and of the
 * three input rows are staged in LDS and read transposed.
 */
#include "hd_internal.h"
#include "../../include/hydist_cnn.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* hd_api.hip: device view and stream of a context (same library) */
const HdDev* hd_ctx_device_view(hd_ctx* c, hipStream_t* stream, int* state);

namespace {

constexpr int kCinPad = 32;   /* input channels after padding */
constexpr int kWidth = 64;    /* hidden channels */
constexpr int kOutPad = 32;   /* output channels after padding */

enum { EPI_RELU = 0, EPI_RESRELU = 1, EPI_LINEAR = 2, EPI_BWD_SET = 3, EPI_BWD_ADD = 4 };

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* Y = conv3x3(X) (+ bias) with epilogue EPI; images [nimg][H][W][C]. */
/* With a tile list (tl != nullptr) workgroup k covers tile tl[k] for k <
 * *tn and does nothing beyond; without one, workgroup k covers tile k.
 * The tile's input window (rows h0 - 1 .. h0 + 4, columns w0 - 1 .. w0 + 64,
 * zero outside the image) is staged in LDS once, pixel pitch padded by 16
 * bytes against bank conflicts, and the nine taps' A fragments are read
 * from there; each tap's weights ([cout][cin]) pass through a double-buffered
 * LDS slot, the next tap's loaded into registers while the current one is
 * multiplied. */
template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(256) k_conv(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                              const float* __restrict__ bias, int nimg, int H, int W,
                                              const bf16* __restrict__ res, bf16* __restrict__ Y,
                                              bf16* __restrict__ R, float* __restrict__ Yf,
                                              const int* __restrict__ tl, const int* __restrict__ tn)
{
    constexpr int NB = COUT / 16;
    constexpr int CP = CIN + 8;      /* padded channels per pixel in LDS */
    constexpr int WR = 6, WC = 66;   /* window rows and columns */
    constexpr int C8 = CIN / 8;      /* 16-byte chunks per pixel */
    constexpr int TW = COUT * CIN;   /* weights of one tap */
    constexpr int NW = (TW / 8 + 255) / 256;
    __shared__ bf16 sx[WR * WC * CP];
    __shared__ bf16 sw[2][TW];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tilesW = (W + 63) >> 6, tilesH = (H + 3) >> 2;
    int t = blockIdx.x;
    if (tl) {
        if ((int)blockIdx.x >= *tn) return;
        t = tl[blockIdx.x];
    }
    const int tw = t % tilesW;
    t /= tilesW;
    const int th = t % tilesH;
    const int b = t / tilesH;
    if (b >= nimg) return;               /* uniform over the workgroup */
    const int h0 = th * 4, w0 = tw * 64;
    const size_t img = (size_t)b * H * W;
    for (int k = threadIdx.x; k < WR * WC * C8; k += 256) {
        const int r = k / (WC * C8), rem = k - r * (WC * C8);
        const int c = rem / C8, c8 = rem - c * C8;
        const int hh = h0 - 1 + r, ww = w0 - 1 + c;
        bf16x8 v = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = *(const bf16x8*)(X + (img + (size_t)hh * W + ww) * CIN + c8 * 8);
        *(bf16x8*)(sx + (r * WC + c) * CP + c8 * 8) = v;
    }
    for (int k = threadIdx.x; k < TW / 8; k += 256) *(bf16x8*)(sw[0] + k * 8) = *(const bf16x8*)(Wt + k * 8);
    __syncthreads();
    const int h = h0 + wv;
    const bool live = h < H;   /* every wave stays for the barriers */
    f32x4 acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = 8 * (l >> 4);
    for (int tap = 0; tap < 9; ++tap) {
        const int r = wv + tap / 3, dx = tap % 3 - 1;
        const bf16* swt = sw[tap & 1];
        bf16x8 nw[NW];
        if (tap < 8) {
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const int k = threadIdx.x + 256 * q;
                if (k < TW / 8) nw[q] = *(const bf16x8*)(Wt + (size_t)(tap + 1) * TW + k * 8);
            }
        }
#pragma unroll
        for (int c0 = 0; c0 < CIN; c0 += 32) {
            bf16x8 bf[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + (l & 15);
                bf[nb] = *(const bf16x8*)(swt + co * CIN + c0 + kq);
            }
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const int c = mb * 16 + (l & 15) + dx + 1;
                const bf16x8 af = *(const bf16x8*)(sx + (r * WC + c) * CP + c0 + kq);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma(af, bf[nb], acc[mb][nb]);
            }
        }
        if (tap < 8) {
            /* the other slot was last read in the previous tap, which every
             * wave finished before the barrier that ended it */
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const int k = threadIdx.x + 256 * q;
                if (k < TW / 8) *(bf16x8*)(sw[(tap + 1) & 1] + k * 8) = nw[q];
            }
            __syncthreads();
        }
    }
    if (!live) return;
    const int w0e = w0;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int w = w0e + mb * 16 + 4 * (l >> 4) + r;
            if (w >= W) continue;
            const size_t p = img + (size_t)h * W + w;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + (l & 15);
                float y = acc[mb][nb][r];
                const size_t q = p * COUT + co;
                if (EPI == EPI_RELU) {
                    Y[q] = (bf16)fmaxf(y + bias[co], 0.f);
                } else if (EPI == EPI_RESRELU) {
                    float rr = fmaxf(y + bias[co], 0.f);
                    R[q] = (bf16)rr;
                    Y[q] = (bf16)((float)res[q] + rr);
                } else if (EPI == EPI_LINEAR) {
                    Yf[q] = y + bias[co];
                } else if (EPI == EPI_BWD_SET) {
                    Yf[q] = y;
                } else {
                    Yf[q] += y;
                }
            }
        }
}

/* dW[tap][co][ci] += sum over pixels p of dZ[p][co] * X[p + off(tap)][ci];
 * chunks of 32 pixels of one image row, workgroups stride over chunks. */
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(const bf16* __restrict__ X, const bf16* __restrict__ dZ, int nimg,
                                               int H, int W, float* __restrict__ dw)
{
    constexpr int NBI = CIN / 16, MB = COUT / 16, PAIRS = 9 * NBI, PPW = (PAIRS + 3) / 4;
    __shared__ __attribute__((aligned(16))) bf16 sZ[32][COUT];
    __shared__ __attribute__((aligned(16))) bf16 sX[3][34][CIN];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int cpr = (W + 31) >> 5;
    const long nchunks = (long)nimg * H * cpr;
    f32x4 acc[PPW][MB];
#pragma unroll
    for (int q = 0; q < PPW; ++q)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[q][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kq = 8 * (l >> 4);
    for (long ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int cw = (int)(ch % cpr);
        const long rest = ch / cpr;
        const int h = (int)(rest % H);
        const int b = (int)(rest / H);
        const int p0 = cw * 32;
        const size_t img = (size_t)b * H * W;
        __syncthreads();
        for (int v = threadIdx.x; v < 32 * COUT / 8; v += 256) {
            const int k = v / (COUT / 8), c8 = (v % (COUT / 8)) * 8;
            bf16x8 z = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (p0 + k < W) z = *(const bf16x8*)(dZ + (img + (size_t)h * W + p0 + k) * COUT + c8);
            *(bf16x8*)&sZ[k][c8] = z;
        }
        for (int v = threadIdx.x; v < 3 * 34 * CIN / 8; v += 256) {
            const int c8 = (v % (CIN / 8)) * 8;
            const int q = v / (CIN / 8);
            const int k = q % 34, r = q / 34;
            const int hh = h + r - 1, ww = p0 + k - 1;
            bf16x8 x = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (hh >= 0 && hh < H && ww >= 0 && ww < W) x = *(const bf16x8*)(X + (img + (size_t)hh * W + ww) * CIN + c8);
            *(bf16x8*)&sX[r][k][c8] = x;
        }
        __syncthreads();
        bf16x8 af[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int e = 0; e < 8; ++e) af[mb][e] = sZ[kq + e][mb * 16 + (l & 15)];
#pragma unroll
        for (int q = 0; q < PPW; ++q) {
            const int pair = wv + 4 * q;
            if (pair < PAIRS) {
                const int tap = pair / NBI, nb = pair % NBI;
                const int r = tap / 3, dx = tap % 3 - 1;
                bf16x8 bfr;
#pragma unroll
                for (int e = 0; e < 8; ++e) bfr[e] = sX[r][kq + e + dx + 1][nb * 16 + (l & 15)];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) acc[q][mb] = mfma(af[mb], bfr, acc[q][mb]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PPW; ++q) {
        const int pair = wv + 4 * q;
        if (pair >= PAIRS) continue;
        const int tap = pair / NBI, nb = pair % NBI;
        const int ci = nb * 16 + (l & 15);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = mb * 16 + 4 * (l >> 4) + r;
                atomicAdd(&dw[((size_t)tap * COUT + co) * CIN + ci], acc[q][mb][r]);
            }
    }
}

/* db[co] += sum over pixels of dZ[p][co] (C channels, 256 threads: 4 pixel
 * groups of 64 channels each, channels beyond C idle). */
template <int C>
__global__ void __launch_bounds__(256) k_bias_grad(const bf16* __restrict__ dZ, long npix, float* __restrict__ db)
{
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    float s = 0.f;
    if (c < C)
        for (long p = blockIdx.x * 4L + g; p < npix; p += gridDim.x * 4L) s += (float)dZ[p * C + c];
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && c < C) atomicAdd(&db[c], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

/* dZ = bf16(dA * (gate > 0)) over npix * 64 elements */
__global__ void __launch_bounds__(256) k_relu_grad(const float* __restrict__ dA, const bf16* __restrict__ gate,
                                                   long n, bf16* __restrict__ dZ)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += gridDim.x * 256L)
        dZ[i] = (float)gate[i] > 0.f ? (bf16)dA[i] : (bf16)0.f;
}

/* prediction = base + out[0..2]; masked squared error and its gradient
 * dOut (bf16, kOutPad channels) = 2 m (pred - tgt) * inv_norm. */
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ out, const float* __restrict__ base,
                                              const float* __restrict__ tgt, const float* __restrict__ msk,
                                              long npix, float inv_norm, hd_cnn_scales sc, float* __restrict__ pred,
                                              bf16* __restrict__ dOut, double* __restrict__ loss)
{
    __shared__ double red[256];
    double s = 0;
    const float dsc[3] = {sc.d_u, sc.d_v, sc.d_z};
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L) {
        const float m = msk[p];
        for (int c = 0; c < kOutPad; ++c) {
            bf16 g = (bf16)0.f;
            if (c < 3) {
                const float pr = base[p * 3 + c] + dsc[c] * out[p * kOutPad + c];
                pred[p * 3 + c] = pr;
                if (m != 0.f) {   /* pixels outside the mask neither add loss nor gradient */
                    const float d = (pr - tgt[p * 3 + c]) / dsc[c];   /* increment error in d units */
                    s += (double)(m * d * d);
                    g = (bf16)(2.f * m * d * inv_norm);
                }
            }
            if (dOut) dOut[p * kOutPad + c] = g;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * (double)inv_norm);
}

/* Adam on n parameters; writes the BF16 copy (and, for weights, the
 * tap-flipped transposed BF16 copy used by the data gradient). */
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                              float bc1, float bc2)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += gridDim.x * 256L) {
        const float gi = g[i];
        const float mi = 0.9f * m[i] + 0.1f * gi;
        const float vi = 0.999f * v[i] + 0.001f * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
    }
}

/* w16[tap][co][ci] = bf16(w32), wT16[8 - tap][ci][co] = bf16(w32) */
__global__ void __launch_bounds__(256) k_wcast(const float* __restrict__ w32, int cin, int cout,
                                               bf16* __restrict__ w16, bf16* __restrict__ wT16)
{
    const long n = 9L * cin * cout;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += gridDim.x * 256L) {
        const int ci = (int)(i % cin);
        const long q = i / cin;
        const int co = (int)(q % cout), tap = (int)(q / cout);
        const bf16 x = (bf16)w32[i];
        w16[i] = x;
        wT16[((size_t)(8 - tap) * cin + ci) * cout + co] = x;
    }
}

/* FP32 host images -> BF16 input (channels padded) and the FP32 base u, v, z */
__global__ void __launch_bounds__(256) k_load_x(const float* __restrict__ x8, long npix, bf16* __restrict__ x0,
                                                float* __restrict__ base)
{
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L) {
        for (int c = 0; c < kCinPad; ++c) x0[p * kCinPad + c] = (bf16)(c < HD_CNN_CIN ? x8[p * HD_CNN_CIN + c] : 0.f);
        for (int c = 0; c < 3; ++c) base[p * 3 + c] = x8[p * HD_CNN_CIN + c];
    }
}

/* Sample store (8 BF16 channels per pixel) -> batch input image */
__global__ void __launch_bounds__(256) k_expand_x(const bf16* __restrict__ xs, long npix, bf16* __restrict__ x0)
{
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L)
        for (int c = 0; c < kCinPad; ++c) x0[p * kCinPad + c] = c < HD_CNN_CIN ? xs[p * HD_CNN_CIN + c] : (bf16)0.f;
}

/* A context's state -> sample input (8 BF16 channels), FP32 base, mask */
__global__ void __launch_bounds__(256) k_pack(const HdDev* __restrict__ dv, hd_cnn_scales s, float q, float zb,
                                              bf16* __restrict__ xs, float* __restrict__ base, float* __restrict__ msk)
{
    const HdDevG& d = reinterpret_cast<const HdDevG*>(dv)[0];
    const long npix = (long)(d.N + 3) * (d.M + 3);
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L) {
        /* state channels only on wet cells: elsewhere the solver's fields
         * carry values no equation reads (they may even be non-finite) */
        const bool wet = d.khouot[p] == 0;
        const float u = wet ? (float)(d.u[p] / s.s_u) : 0.f;
        const float v = wet ? (float)(d.v[p] / s.s_v) : 0.f;
        const float z = wet ? (float)(d.z[p] / s.s_z) : 0.f;
        const float bed = wet ? (float)(d.dH[p] / s.s_b) : 0.f;
        const float ch[HD_CNN_CIN] = {u, v, z, bed, wet ? 1.f : 0.f, q / s.s_q, zb / s.s_z, (float)(d.h[p] / s.s_h)};
        /* the network sees every channel clipped to [-8, 8]; the residual
         * base keeps the unclipped u, v, z */
        for (int c = 0; c < HD_CNN_CIN; ++c) xs[p * HD_CNN_CIN + c] = (bf16)fminf(fmaxf(ch[c], -8.f), 8.f);
        base[p * 3 + 0] = u;
        base[p * 3 + 1] = v;
        base[p * 3 + 2] = z;
        if (msk) msk[p] = wet ? 1.f : 0.f;
    }
}

/* A context's u, v, z -> normalised FP32 target */
__global__ void __launch_bounds__(256) k_target(const HdDev* __restrict__ dv, hd_cnn_scales s, float* __restrict__ tgt)
{
    const HdDevG& d = reinterpret_cast<const HdDevG*>(dv)[0];
    const long npix = (long)(d.N + 3) * (d.M + 3);
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L) {
        const bool wet = d.khouot[p] == 0;
        tgt[p * 3 + 0] = wet ? (float)(d.u[p] / s.s_u) : 0.f;
        tgt[p * 3 + 1] = wet ? (float)(d.v[p] / s.s_v) : 0.f;
        tgt[p * 3 + 2] = wet ? (float)(d.z[p] / s.s_z) : 0.f;
    }
}

/* Network output (normalised increments) -> the context's wet cells, FP64 */
__global__ void __launch_bounds__(256) k_unpack(const HdDev* __restrict__ dv, hd_cnn_scales s,
                                                const float* __restrict__ out)
{
    const HdDevG& d = reinterpret_cast<const HdDevG*>(dv)[0];
    const int N = d.N, M = d.M, W = d.W;
    const long npix = (long)(N + 3) * W;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npix; p += gridDim.x * 256L) {
        const int i = (int)(p / W), j = (int)(p % W);
        if (i < 1 || i > N || j < 1 || j > M || d.khouot[p] != 0) continue;
        d.u[p] += (double)s.s_u * (double)s.d_u * (double)out[p * kOutPad + 0];
        d.v[p] += (double)s.s_v * (double)s.d_v * (double)out[p * kOutPad + 1];
        d.z[p] += (double)s.s_z * (double)s.d_z * (double)out[p * kOutPad + 2];
    }
}

/* The forward tiles (4 rows x 64 columns of the image) that lie within
 * `rad` cells (Chebyshev distance) of a wet cell of [1,N]x[1,M]: one wave
 * per tile scans the tile grown by rad on every side; the list order is
 * arbitrary (each tile's outputs do not depend on it).  *count must be zero
 * on entry. */
__global__ void __launch_bounds__(64) k_active_tiles(const HdDev* __restrict__ dv, int rad, int* __restrict__ list,
                                                     int* __restrict__ count)
{
    const HdDevG& d = reinterpret_cast<const HdDevG*>(dv)[0];
    const int N = d.N, M = d.M, W = d.W, H = N + 3;
    const int tilesW = (W + 63) >> 6;
    const int t = blockIdx.x, tw = t % tilesW, th = t / tilesW;
    if (th * 4 >= H) return;
    const int i0 = max(th * 4 - rad, 1), i1 = min(th * 4 + 3 + rad, N);
    const int j0 = max(tw * 64 - rad, 1), j1 = min(tw * 64 + 63 + rad, M);
    bool any = false;
    if (i0 <= i1 && j0 <= j1) {
        const int wj = j1 - j0 + 1;
        const long n = (long)(i1 - i0 + 1) * wj;
        for (long k = threadIdx.x; k < n && !any; k += 64) {
            const int i = i0 + (int)(k / wj), j = j0 + (int)(k % wj);
            any = d.khouot[(long)i * W + j] == 0;
        }
    }
    if (__any(any) && threadIdx.x == 0) list[atomicAdd(count, 1)] = t;
}

__global__ void k_zero_int(int* p) { *p = 0; }

__global__ void k_mfma_check(const bf16* A, const bf16* B, float* D)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
        a[e] = A[(l % 16) * 32 + 8 * (l / 16) + e];
        b[e] = B[(8 * (l / 16) + e) * 16 + l % 16];
    }
    f32x4 c = f32x4{0.f, 0.f, 0.f, 0.f};
    c = mfma(a, b, c);
    for (int r = 0; r < 4; ++r) D[(4 * (l / 16) + r) * 16 + l % 16] = c[r];
}

struct Layer {
    int cin = 0, cout = 0;
    float *w32 = nullptr, *b32 = nullptr, *dw = nullptr, *db = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    bf16 *w16 = nullptr, *wT16 = nullptr;
    size_t nw() const { return (size_t)9 * cin * cout; }
};

inline int nblk(long n) { long b = (n + 255) / 256; return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096); }

} // namespace

struct hd_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    int H = 0, W = 0, B = 0, L = 0;
    long P = 0;   /* pixels per image */
    std::vector<Layer> layer;
    hd_cnn_scales sc{1, 1, 1, 1, 1, 1, 1, 1, 1};
    std::vector<void*> allocs;
    bf16* x0 = nullptr;                 /* [B][P][32] */
    float* base = nullptr;              /* [B][P][3] */
    float *tgt = nullptr, *msk = nullptr, *pred = nullptr;
    std::vector<bf16*> act;             /* act[1..L-1]: [B][P][64] */
    std::vector<bf16*> rel;             /* rel[1..L-2]: ReLU outputs of the residual layers */
    float* out = nullptr;               /* [B][P][32] */
    float* dA = nullptr;                /* [B][P][64] */
    bf16* dZ = nullptr;                 /* [B][P][64] */
    double* loss = nullptr;
    int nb = 0;                         /* images in the batch */
    float nmask = 0;
    int step = 0;
    /* sample store */
    int nsamp = 0;
    bf16* sx = nullptr;                 /* [n][P][8] */
    /* tile-sparse inference (hd_cnn_predict_into_ctx) */
    bool sparse = true;
    int* tlist = nullptr;               /* [tiles of one image] */
    int* tcount = nullptr;              /* [1] */
    const int* tl = nullptr;            /* the list the next forward uses (nullptr: all tiles) */
    float *sbase = nullptr, *stgt = nullptr, *smsk = nullptr;
    std::string err;
};

static std::string g_cnn_err;

static int cfail(hd_cnn* n, int code, const std::string& m)
{
    if (n) n->err = m;
    g_cnn_err = m;
    return code;
}

#define CHK(n, x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return cfail((n), HD_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

/* zero-filled device allocation; the fill runs on the network's
 * (non-blocking) stream and has finished when this returns, so it is
 * ordered before everything queued on that stream afterwards */
static int calloc_dev(hd_cnn* n, void** p, size_t bytes)
{
    if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) return cfail(n, HD_ENOMEM, "hipMalloc (network)");
    n->allocs.push_back(*p);
    if (hipMemsetAsync(*p, 0, bytes ? bytes : 8, n->stream) != hipSuccess ||
        hipStreamSynchronize(n->stream) != hipSuccess)
        return cfail(n, HD_EHIP, "hipMemsetAsync");
    return HD_OK;
}

/* splitmix64 -> uniform [-1, 1) */
static double rnd(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

static int refresh_layer(hd_cnn* n, int l)
{
    Layer& y = n->layer[l];
    hipLaunchKernelGGL(k_wcast, dim3(nblk((long)y.nw())), dim3(256), 0, n->stream, y.w32, y.cin, y.cout, y.w16, y.wT16);
    CHK(n, hipGetLastError());
    return HD_OK;
}

template <int CIN, int COUT, int EPI>
static void conv(hd_cnn* n, const bf16* X, const Layer& y, bool transposed, int nimg, const bf16* res, bf16* Y,
                 bf16* R, float* Yf)
{
    int tiles = ((n->W + 63) / 64) * ((n->H + 3) / 4) * nimg;
    hipLaunchKernelGGL((k_conv<CIN, COUT, EPI>), dim3(tiles), dim3(256), 0, n->stream, X, transposed ? y.wT16 : y.w16,
                       y.b32, nimg, n->H, n->W, res, Y, R, Yf, n->tl, n->tcount);
}

template <int CIN, int COUT>
static void wgrad(hd_cnn* n, const bf16* X, const bf16* dZ, const Layer& y, int nimg)
{
    long chunks = (long)nimg * n->H * ((n->W + 31) / 32);
    int grid = (int)(chunks < 1024 ? chunks : 1024);
    hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(grid), dim3(256), 0, n->stream, X, dZ, nimg, n->H, n->W, y.dw);
    long npix = (long)nimg * n->P;
    hipLaunchKernelGGL((k_bias_grad<COUT>), dim3(nblk(npix / 4 + 1) < 1024 ? nblk(npix / 4 + 1) : 1024), dim3(256), 0,
                       n->stream, dZ, npix, y.db);
}

static int forward(hd_cnn* n, int nimg)
{
    const int L = n->L;
    conv<kCinPad, kWidth, EPI_RELU>(n, n->x0, n->layer[0], false, nimg, nullptr, n->act[1], nullptr, nullptr);
    for (int l = 1; l <= L - 2; ++l)
        conv<kWidth, kWidth, EPI_RESRELU>(n, n->act[l], n->layer[l], false, nimg, n->act[l], n->act[l + 1], n->rel[l],
                                          nullptr);
    conv<kWidth, kOutPad, EPI_LINEAR>(n, n->act[L - 1], n->layer[L - 1], false, nimg, nullptr, nullptr, nullptr,
                                      n->out);
    CHK(n, hipGetLastError());
    return HD_OK;
}

static int loss_and_grad(hd_cnn* n, int nimg, bool grad, double* loss)
{
    long npix = (long)nimg * n->P;
    CHK(n, hipMemsetAsync(n->loss, 0, sizeof(double), n->stream));
    float inv = n->nmask > 0 ? 1.f / (3.f * n->nmask) : 0.f;
    hipLaunchKernelGGL(k_loss, dim3(nblk(npix) < 2048 ? nblk(npix) : 2048), dim3(256), 0, n->stream, n->out, n->base,
                       n->tgt, n->msk, npix, inv, n->sc, n->pred, grad ? n->dZ : nullptr, n->loss);
    CHK(n, hipGetLastError());
    if (loss) {
        CHK(n, hipMemcpyAsync(loss, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
        CHK(n, hipStreamSynchronize(n->stream));
    }
    return HD_OK;
}

static int backward(hd_cnn* n, int nimg)
{
    const int L = n->L;
    long npix = (long)nimg * n->P;
    for (auto& y : n->layer) {
        CHK(n, hipMemsetAsync(y.dw, 0, y.nw() * 4, n->stream));
        CHK(n, hipMemsetAsync(y.db, 0, (size_t)y.cout * 4, n->stream));
    }
    /* output layer: dZ holds dOut (kOutPad channels) */
    wgrad<kWidth, kOutPad>(n, n->act[L - 1], n->dZ, n->layer[L - 1], nimg);
    conv<kOutPad, kWidth, EPI_BWD_SET>(n, n->dZ, n->layer[L - 1], true, nimg, nullptr, nullptr, nullptr, n->dA);
    for (int l = L - 2; l >= 1; --l) {
        hipLaunchKernelGGL(k_relu_grad, dim3(nblk(npix * kWidth) < 4096 ? nblk(npix * kWidth) : 4096), dim3(256), 0,
                           n->stream, n->dA, n->rel[l], npix * kWidth, n->dZ);
        wgrad<kWidth, kWidth>(n, n->act[l], n->dZ, n->layer[l], nimg);
        conv<kWidth, kWidth, EPI_BWD_ADD>(n, n->dZ, n->layer[l], true, nimg, nullptr, nullptr, nullptr, n->dA);
    }
    hipLaunchKernelGGL(k_relu_grad, dim3(4096), dim3(256), 0, n->stream, n->dA, n->act[1], npix * kWidth, n->dZ);
    wgrad<kCinPad, kWidth>(n, n->x0, n->dZ, n->layer[0], nimg);
    CHK(n, hipGetLastError());
    return HD_OK;
}

static int adam(hd_cnn* n, float lr)
{
    n->step += 1;
    float bc1 = 1.f - std::pow(0.9f, (float)n->step), bc2 = 1.f - std::pow(0.999f, (float)n->step);
    for (int l = 0; l < n->L; ++l) {
        Layer& y = n->layer[l];
        hipLaunchKernelGGL(k_adam, dim3(nblk((long)y.nw())), dim3(256), 0, n->stream, y.w32, y.dw, y.mw, y.vw,
                           (long)y.nw(), lr, bc1, bc2);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, n->stream, y.b32, y.db, y.mb, y.vb, (long)y.cout, lr, bc1,
                           bc2);
        int rc = refresh_layer(n, l);
        if (rc) return rc;
    }
    CHK(n, hipGetLastError());
    return HD_OK;
}

extern "C" {

const char* hd_cnn_last_error(const hd_cnn* n) { return n ? n->err.c_str() : g_cnn_err.c_str(); }

int hd_cnn_create(int device, int H, int W, int batch, int layers, uint64_t seed, hd_cnn** out)
{
    if (!out || H < 1 || W < 1 || batch < 1 || layers < 3) return cfail(nullptr, HD_EINVAL, "hd_cnn_create: bad arguments");
    if (hipSetDevice(device) != hipSuccess) return cfail(nullptr, HD_EHIP, "hipSetDevice");
    hd_cnn* n = new hd_cnn();
    n->device = device;
    n->H = H; n->W = W; n->B = batch; n->L = layers;
    n->P = (long)H * W;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return cfail(nullptr, HD_EHIP, "hipStreamCreate");
    }
    int rc;
    const size_t PB = (size_t)n->P * batch;
    n->layer.resize(layers);
    uint64_t s = seed;
    for (int l = 0; l < layers; ++l) {
        Layer& y = n->layer[l];
        y.cin = l == 0 ? kCinPad : kWidth;
        y.cout = l == layers - 1 ? kOutPad : kWidth;
        size_t nw = y.nw();
        void** f32[] = {(void**)&y.w32, (void**)&y.dw, (void**)&y.mw, (void**)&y.vw};
        for (auto p : f32)
            if ((rc = calloc_dev(n, p, nw * 4))) { hd_cnn_destroy(n); return rc; }
        void** fb[] = {(void**)&y.b32, (void**)&y.db, (void**)&y.mb, (void**)&y.vb};
        for (auto p : fb)
            if ((rc = calloc_dev(n, p, (size_t)y.cout * 4))) { hd_cnn_destroy(n); return rc; }
        if ((rc = calloc_dev(n, (void**)&y.w16, nw * 2)) || (rc = calloc_dev(n, (void**)&y.wT16, nw * 2))) {
            hd_cnn_destroy(n);
            return rc;
        }
        /* He-uniform over the real fan-in; padded input channels and padded
         * outputs stay zero; the last layer starts small (near-identity
         * residual prediction) */
        int cin_real = l == 0 ? HD_CNN_CIN : kWidth;
        int cout_real = l == layers - 1 ? HD_CNN_COUT : kWidth;
        double a = std::sqrt(6.0 / (9.0 * cin_real)) * (l == layers - 1 ? 0.1 : (l == 0 ? 1.0 : 0.5));
        std::vector<float> w(nw, 0.f);
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < cout_real; ++co)
                for (int ci = 0; ci < cin_real; ++ci) w[((size_t)tap * y.cout + co) * y.cin + ci] = (float)(a * rnd(s));
        if (hipMemcpy(y.w32, w.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess) {
            hd_cnn_destroy(n);
            return cfail(nullptr, HD_EHIP, "hipMemcpy weights");
        }
        if ((rc = refresh_layer(n, l))) { hd_cnn_destroy(n); return rc; }
    }
    n->act.assign(layers, nullptr);
    n->rel.assign(layers, nullptr);
    for (int l = 1; l <= layers - 1; ++l)
        if ((rc = calloc_dev(n, (void**)&n->act[l], PB * kWidth * 2))) { hd_cnn_destroy(n); return rc; }
    for (int l = 1; l <= layers - 2; ++l)
        if ((rc = calloc_dev(n, (void**)&n->rel[l], PB * kWidth * 2))) { hd_cnn_destroy(n); return rc; }
    if ((rc = calloc_dev(n, (void**)&n->x0, PB * kCinPad * 2)) || (rc = calloc_dev(n, (void**)&n->base, PB * 3 * 4)) ||
        (rc = calloc_dev(n, (void**)&n->tgt, PB * 3 * 4)) || (rc = calloc_dev(n, (void**)&n->msk, PB * 4)) ||
        (rc = calloc_dev(n, (void**)&n->pred, PB * 3 * 4)) || (rc = calloc_dev(n, (void**)&n->out, PB * kOutPad * 4)) ||
        (rc = calloc_dev(n, (void**)&n->dA, PB * kWidth * 4)) || (rc = calloc_dev(n, (void**)&n->dZ, PB * kWidth * 2)) ||
        (rc = calloc_dev(n, (void**)&n->loss, 8)) ||
        (rc = calloc_dev(n, (void**)&n->tlist, (size_t)((W + 63) / 64) * ((H + 3) / 4) * 4)) ||
        (rc = calloc_dev(n, (void**)&n->tcount, 4))) {
        hd_cnn_destroy(n);
        return rc;
    }
    if (hipStreamSynchronize(n->stream) != hipSuccess) { hd_cnn_destroy(n); return cfail(nullptr, HD_EHIP, "sync"); }
    *out = n;
    return HD_OK;
}

int hd_cnn_destroy(hd_cnn* n)
{
    if (!n) return HD_OK;
    hipSetDevice(n->device);
    if (n->stream) hipStreamSynchronize(n->stream);
    for (void* p : n->allocs) hipFree(p);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
    return HD_OK;
}

int hd_cnn_set_scales(hd_cnn* n, const hd_cnn_scales* s)
{
    if (!n || !s || !(s->d_u > 0) || !(s->d_v > 0) || !(s->d_z > 0)) return HD_EINVAL;
    n->sc = *s;
    return HD_OK;
}

int hd_cnn_set_sparse(hd_cnn* n, int on)
{
    if (!n) return HD_EINVAL;
    n->sparse = on != 0;
    return HD_OK;
}

int hd_cnn_layer_dims(const hd_cnn* n, int l, int* cin, int* cout)
{
    if (!n || l < 0 || l >= n->L) return HD_EINVAL;
    if (cin) *cin = n->layer[l].cin;
    if (cout) *cout = n->layer[l].cout;
    return HD_OK;
}

int hd_cnn_get_layer(hd_cnn* n, int l, float* host_w, float* host_b)
{
    if (!n || l < 0 || l >= n->L) return HD_EINVAL;
    const Layer& y = n->layer[l];
    CHK(n, hipStreamSynchronize(n->stream));
    if (host_w) CHK(n, hipMemcpy(host_w, y.w32, y.nw() * 4, hipMemcpyDeviceToHost));
    if (host_b) CHK(n, hipMemcpy(host_b, y.b32, (size_t)y.cout * 4, hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_cnn_set_layer(hd_cnn* n, int l, const float* host_w, const float* host_b)
{
    if (!n || l < 0 || l >= n->L) return HD_EINVAL;
    Layer& y = n->layer[l];
    CHK(n, hipStreamSynchronize(n->stream));
    if (host_w) CHK(n, hipMemcpy(y.w32, host_w, y.nw() * 4, hipMemcpyHostToDevice));
    if (host_b) CHK(n, hipMemcpy(y.b32, host_b, (size_t)y.cout * 4, hipMemcpyHostToDevice));
    CHK(n, hipMemsetAsync(y.mw, 0, y.nw() * 4, n->stream));
    CHK(n, hipMemsetAsync(y.vw, 0, y.nw() * 4, n->stream));
    CHK(n, hipMemsetAsync(y.mb, 0, (size_t)y.cout * 4, n->stream));
    CHK(n, hipMemsetAsync(y.vb, 0, (size_t)y.cout * 4, n->stream));
    int rc = refresh_layer(n, l);
    if (rc) return rc;
    CHK(n, hipStreamSynchronize(n->stream));
    return HD_OK;
}

int hd_cnn_get_grad(hd_cnn* n, int l, float* host_dw, float* host_db)
{
    if (!n || l < 0 || l >= n->L) return HD_EINVAL;
    const Layer& y = n->layer[l];
    CHK(n, hipStreamSynchronize(n->stream));
    if (host_dw) CHK(n, hipMemcpy(host_dw, y.dw, y.nw() * 4, hipMemcpyDeviceToHost));
    if (host_db) CHK(n, hipMemcpy(host_db, y.db, (size_t)y.cout * 4, hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_cnn_load_batch(hd_cnn* n, int b, const float* host_x, const float* host_y, const float* host_m)
{
    if (!n || b < 1 || b > n->B || !host_x) return cfail(n, HD_EINVAL, "hd_cnn_load_batch: bad arguments");
    long npix = (long)b * n->P;
    float* tmp = nullptr;
    CHK(n, hipMalloc((void**)&tmp, (size_t)npix * HD_CNN_CIN * 4));
    hipError_t e = hipMemcpy(tmp, host_x, (size_t)npix * HD_CNN_CIN * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_load_x, dim3(nblk(npix)), dim3(256), 0, n->stream, tmp, npix, n->x0, n->base);
        e = hipStreamSynchronize(n->stream);
    }
    hipFree(tmp);
    if (e != hipSuccess) return cfail(n, HD_EHIP, hipGetErrorString(e));
    if (host_y) CHK(n, hipMemcpy(n->tgt, host_y, (size_t)npix * 3 * 4, hipMemcpyHostToDevice));
    double ms = 0;
    if (host_m) {
        CHK(n, hipMemcpy(n->msk, host_m, (size_t)npix * 4, hipMemcpyHostToDevice));
        for (long p = 0; p < npix; ++p) ms += host_m[p];
    }
    n->nmask = (float)ms;
    n->nb = b;
    return HD_OK;
}

int hd_cnn_forward(hd_cnn* n, float* host_out)
{
    if (!n || n->nb < 1) return cfail(n, HD_ESTATE, "hd_cnn_forward: no batch loaded");
    int rc = forward(n, n->nb);
    if (rc) return rc;
    if ((rc = loss_and_grad(n, n->nb, false, nullptr))) return rc;
    if (host_out) {
        CHK(n, hipStreamSynchronize(n->stream));
        CHK(n, hipMemcpy(host_out, n->pred, (size_t)n->nb * n->P * 3 * 4, hipMemcpyDeviceToHost));
    }
    return HD_OK;
}

int hd_cnn_backward(hd_cnn* n, double* loss)
{
    if (!n || n->nb < 1) return cfail(n, HD_ESTATE, "hd_cnn_backward: no batch loaded");
    int rc = loss_and_grad(n, n->nb, true, loss);
    if (rc) return rc;
    if ((rc = backward(n, n->nb))) return rc;
    CHK(n, hipStreamSynchronize(n->stream));
    return HD_OK;
}

int hd_cnn_adam(hd_cnn* n, float lr)
{
    if (!n) return HD_EINVAL;
    int rc = adam(n, lr);
    if (rc) return rc;
    CHK(n, hipStreamSynchronize(n->stream));
    return HD_OK;
}

int hd_cnn_train_step(hd_cnn* n, float lr, double* loss)
{
    int rc = hd_cnn_forward(n, nullptr);
    if (!rc) rc = hd_cnn_backward(n, loss);
    if (!rc) rc = hd_cnn_adam(n, lr);
    return rc;
}

int hd_cnn_data_alloc(hd_cnn* n, int count)
{
    if (!n || count < 1 || n->nsamp) return cfail(n, HD_EINVAL, "hd_cnn_data_alloc: bad arguments");
    size_t np = (size_t)n->P * count;
    int rc;
    if ((rc = calloc_dev(n, (void**)&n->sx, np * HD_CNN_CIN * 2)) || (rc = calloc_dev(n, (void**)&n->sbase, np * 3 * 4)) ||
        (rc = calloc_dev(n, (void**)&n->stgt, np * 3 * 4)) || (rc = calloc_dev(n, (void**)&n->smsk, np * 4)))
        return rc;
    n->nsamp = count;
    return HD_OK;
}

int hd_cnn_capture(hd_cnn* n, int idx, int role, hd_ctx* c, float q, float zb)
{
    if (!n || !c || idx < 0 || idx >= n->nsamp || (role != 0 && role != 1))
        return cfail(n, HD_EINVAL, "hd_cnn_capture: bad arguments");
    hipStream_t cs;
    int st;
    const HdDev* dv = hd_ctx_device_view(c, &cs, &st);
    HdDev hd;
    CHK(n, hipMemcpy(&hd, dv, sizeof(HdDev), hipMemcpyDeviceToHost));
    if ((long)(hd.N + 3) * (hd.M + 3) != n->P || hd.N + 3 != n->H) return cfail(n, HD_ESHAPE, "grid mismatch");
    CHK(n, hipStreamSynchronize(cs));
    size_t o = (size_t)idx * n->P;
    if (role == 0)
        hipLaunchKernelGGL(k_pack, dim3(nblk(n->P)), dim3(256), 0, n->stream, dv, n->sc, q, zb, n->sx + o * HD_CNN_CIN,
                           n->sbase + o * 3, n->smsk + o);
    else
        hipLaunchKernelGGL(k_target, dim3(nblk(n->P)), dim3(256), 0, n->stream, dv, n->sc, n->stgt + o * 3);
    CHK(n, hipGetLastError());
    CHK(n, hipStreamSynchronize(n->stream));
    return HD_OK;
}

static int gather(hd_cnn* n, const int* idx, int b)
{
    if (!idx || b < 1 || b > n->B) return cfail(n, HD_EINVAL, "sample list");
    double ms = 0;
    std::vector<float> hm;
    for (int k = 0; k < b; ++k) {
        if (idx[k] < 0 || idx[k] >= n->nsamp) return cfail(n, HD_EINVAL, "sample index out of range");
        size_t o = (size_t)idx[k] * n->P, d = (size_t)k * n->P;
        hipLaunchKernelGGL(k_expand_x, dim3(nblk(n->P)), dim3(256), 0, n->stream, n->sx + o * HD_CNN_CIN, n->P,
                           n->x0 + d * kCinPad);
        CHK(n, hipMemcpyAsync(n->base + d * 3, n->sbase + o * 3, (size_t)n->P * 12, hipMemcpyDeviceToDevice, n->stream));
        CHK(n, hipMemcpyAsync(n->tgt + d * 3, n->stgt + o * 3, (size_t)n->P * 12, hipMemcpyDeviceToDevice, n->stream));
        CHK(n, hipMemcpyAsync(n->msk + d, n->smsk + o, (size_t)n->P * 4, hipMemcpyDeviceToDevice, n->stream));
    }
    hm.resize((size_t)b * n->P);
    CHK(n, hipMemcpyAsync(hm.data(), n->msk, hm.size() * 4, hipMemcpyDeviceToHost, n->stream));
    CHK(n, hipStreamSynchronize(n->stream));
    for (float m : hm) ms += m;
    n->nmask = (float)ms;
    n->nb = b;
    return HD_OK;
}

int hd_cnn_train_samples(hd_cnn* n, const int* host_idx, int b, float lr, double* loss)
{
    if (!n) return HD_EINVAL;
    int rc = gather(n, host_idx, b);
    if (!rc) rc = hd_cnn_train_step(n, lr, loss);
    return rc;
}

int hd_cnn_eval_samples(hd_cnn* n, const int* host_idx, int b, double* loss)
{
    if (!n) return HD_EINVAL;
    int rc = gather(n, host_idx, b);
    if (!rc) rc = forward(n, b);
    if (!rc) rc = loss_and_grad(n, b, false, loss);
    return rc;
}

int hd_cnn_predict_into_ctx(hd_cnn* n, hd_ctx* c, float q, float zb)
{
    if (!n || !c) return HD_EINVAL;
    hipStream_t cs;
    int st;
    const HdDev* dv = hd_ctx_device_view(c, &cs, &st);
    HdDev hd;
    CHK(n, hipMemcpy(&hd, dv, sizeof(HdDev), hipMemcpyDeviceToHost));
    if ((long)(hd.N + 3) * (hd.M + 3) != n->P || hd.N + 3 != n->H) return cfail(n, HD_ESHAPE, "grid mismatch");
    if (st != 0) return cfail(n, HD_ESTATE, "context between half-steps");
    CHK(n, hipStreamSynchronize(cs));
    bf16* xs = (bf16*)n->dZ;   /* scratch for the 8-channel sample (the gradient buffer is free here) */
    hipLaunchKernelGGL(k_pack, dim3(nblk(n->P)), dim3(256), 0, n->stream, dv, n->sc, q, zb, xs, n->base, nullptr);
    hipLaunchKernelGGL(k_expand_x, dim3(nblk(n->P)), dim3(256), 0, n->stream, xs, n->P, n->x0);
    if (n->sparse) {
        /* Only the tiles within L - 1 cells of a wet cell: the first layer
         * reads the whole input image, and a layer's value at a cell is the
         * dense one wherever its 3x3 neighbours' values were, so with every
         * cell within L - 1 of a wet cell computed the network's output at
         * the wet cells is that of the dense pass, bit for bit. */
        int tiles = ((n->W + 63) / 64) * ((n->H + 3) / 4);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, n->stream, n->tcount);
        hipLaunchKernelGGL(k_active_tiles, dim3(tiles), dim3(64), 0, n->stream, dv, n->L - 1, n->tlist, n->tcount);
        n->tl = n->tlist;
    }
    int rc = forward(n, 1);
    n->tl = nullptr;
    if (rc) return rc;
    hipLaunchKernelGGL(k_unpack, dim3(nblk(n->P)), dim3(256), 0, n->stream, dv, n->sc, n->out);
    CHK(n, hipGetLastError());
    CHK(n, hipStreamSynchronize(n->stream));
    return hd_flow_refresh(c);
}

int hd_cnn_mfma_check(int device, const float* host_a, const float* host_b, float* host_d)
{
    if (!host_a || !host_b || !host_d) return HD_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return cfail(nullptr, HD_EHIP, "hipSetDevice");
    std::vector<bf16> a(16 * 32), b(32 * 16);
    for (int i = 0; i < 16 * 32; ++i) a[i] = (bf16)host_a[i];
    for (int i = 0; i < 32 * 16; ++i) b[i] = (bf16)host_b[i];
    bf16 *da = nullptr, *db = nullptr;
    float* dd = nullptr;
    int rc = HD_OK;
    if (hipMalloc((void**)&da, a.size() * 2) != hipSuccess || hipMalloc((void**)&db, b.size() * 2) != hipSuccess ||
        hipMalloc((void**)&dd, 256 * 4) != hipSuccess)
        rc = cfail(nullptr, HD_ENOMEM, "hipMalloc");
    if (!rc && (hipMemcpy(da, a.data(), a.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(db, b.data(), b.size() * 2, hipMemcpyHostToDevice) != hipSuccess))
        rc = cfail(nullptr, HD_EHIP, "hipMemcpy");
    if (!rc) {
        hipLaunchKernelGGL(k_mfma_check, dim3(1), dim3(64), 0, 0, da, db, dd);
        if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(host_d, dd, 256 * 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = cfail(nullptr, HD_EHIP, "mfma check");
    }
    if (da) hipFree(da);
    if (db) hipFree(db);
    if (dd) hipFree(dd);
    return rc;
}

} // extern "C"
