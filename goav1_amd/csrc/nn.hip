// Learned proposal for the encoder search: a small CNN over 64x64 luma
// (source, reconstruction) tiles, trained and run on gfx950 with BF16 MFMA and
// FP32 accumulation.
//
// Network (46k parameters), activations NHWC:
//   conv1 3x3 stride 2, pad 1:  2 -> 16 channels, 64x64 -> 32x32, ReLU
//   conv2 3x3 stride 2, pad 1: 16 -> 32 channels, 32x32 -> 16x16, ReLU
//   conv3 3x3 stride 2, pad 1: 32 -> 64 channels, 16x16 ->  8x8,  ReLU
//   global average pool -> 64
//   fc1 64 -> 128, ReLU
//   heads 128 -> 131: 64 luma-strength logits, 64 chroma-strength logits,
//   3 loop-restoration type logits
// Every layer is a GEMM on the matrix cores: a convolution is im2col (patches
// of 9*Cin in tap-major, channel-minor order) times the [9*Cin][Cout] weight;
// the backward pass is two GEMMs per layer (weight gradient = patches^T x dY,
// patch gradient = dY x W^T) plus a gather (col2im) back to the input layout.
//
// GEMM: C[M][N] = A[M][K] * B[K][N], A and B BF16 with arbitrary strides (so
// transposed operands need no copy), FP32 accumulate, 64x64 workgroup tile,
// K step 32 staged in LDS, each of the 4 waves owning a 32x32 sub-tile as
// 2x2 v_mfma_f32_16x16x16bf16 blocks; epilogue: + bias, ReLU, FP32 and/or
// BF16 output.
#include "pf_internal.h"
#include "nn_internal.h"

namespace pf {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

#define GT 64     // workgroup tile (M and N)
#define GK 32     // K step
#define APAD 8

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g)
{
    __shared__ uint16_t As[GT][GK + APAD];
    __shared__ uint16_t Bs[GK][GT + APAD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (v4f){ 0.f, 0.f, 0.f, 0.f };
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < (GT * GK) / 256; e++) {
            const int idx = t + 256 * e;
            const int r = idx / GK, c = idx % GK;          // A tile: row r, k c
            const int m = m0 + r, k = k0 + c;
            As[r][c] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.sam + (int64_t)k * g.sak] : (uint16_t)0;
            const int kr = idx / GT, nc = idx % GT;        // B tile: k kr, col nc
            const int kb = k0 + kr, n = n0 + nc;
            Bs[kr][nc] = (kb < g.K && n < g.N) ? g.B[(int64_t)kb * g.sbk + (int64_t)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 16) {
            const int kq = ks + 4 * (lane >> 4);
            v4s af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const uint16_t *p = &As[wm + a * 16 + (lane & 15)][kq];
                af[a] = (v4s){ (short)p[0], (short)p[1], (short)p[2], (short)p[3] };
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int col = wn + b * 16 + (lane & 15);
                bfr[b] = (v4s){ (short)Bs[kq][col], (short)Bs[kq + 1][col], (short)Bs[kq + 2][col], (short)Bs[kq + 3][col] };
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
    }
    // D lane layout: column lane % 16, rows 4 * (lane / 16) + r
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + b * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[a][b][r];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = fmaxf(v, 0.f);
                if (g.C) g.C[(int64_t)m * g.ldc + n] = v;
                if (g.Cb) g.Cb[(int64_t)m * g.ldc + n] = f2bf(v);
            }
}

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s)
{
    dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

// ---- im2col / col2im for 3x3 stride 2 pad 1 ----------------------------------
// cols[(n*Ho + y)*Wo + x][(ky*3 + kx)*C + c] = X[n][2y + ky - 1][2x + kx - 1][c]
__global__ void im2col_kernel(const uint16_t *X, uint16_t *cols, int N, int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, K = 9 * C;
    const int64_t total = (int64_t)N * Ho * Wo * K;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % K);
        const int64_t row = i / K;
        const int x = (int)(row % Wo), y = (int)((row / Wo) % Ho), n = (int)(row / ((int64_t)Wo * Ho));
        const int tap = k / C, c = k % C, ky = tap / 3, kx = tap % 3;
        const int iy = 2 * y + ky - 1, ix = 2 * x + kx - 1;
        cols[i] = (iy >= 0 && iy < H && ix >= 0 && ix < W) ? X[(((int64_t)n * H + iy) * W + ix) * C + c] : (uint16_t)0;
    }
}

// dX[n][iy][ix][c] = sum over the (ky, kx) whose window covers (iy, ix) of
// dcols, times the ReLU mask of the forward activation A (A > 0).
__global__ void col2im_relu_kernel(const float *dcols, const uint16_t *A, uint16_t *dX, float *dXf, int N, int H,
                                   int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, K = 9 * C;
    const int64_t total = (int64_t)N * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int ix = (int)(p % W), iy = (int)((p / W) % H), n = (int)(p / ((int64_t)W * H));
        float s = 0.f;
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
            const int ty = iy + 1 - ky;
            if (ty < 0 || (ty & 1) || (ty >> 1) >= Ho) continue;
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int tx = ix + 1 - kx;
                if (tx < 0 || (tx & 1) || (tx >> 1) >= Wo) continue;
                s += dcols[(((int64_t)n * Ho + (ty >> 1)) * Wo + (tx >> 1)) * K + (ky * 3 + kx) * C + c];
            }
        }
        if (A && !(bf2f(A[i]) > 0.f)) s = 0.f;
        if (dX) dX[i] = f2bf(s);
        if (dXf) dXf[i] = s;
    }
}

// ---- pooling, heads, loss ----------------------------------------------------
__global__ void avgpool_kernel(const uint16_t *X, uint16_t *Y, float *Yf, int N, int HW, int C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * C) return;
    const int n = i / C, c = i % C;
    float s = 0.f;
    for (int p = 0; p < HW; p++) s += bf2f(X[((int64_t)n * HW + p) * C + c]);
    s /= HW;
    Y[i] = f2bf(s);
    if (Yf) Yf[i] = s;
}

// dX[n][p][c] = dY[n][c] / HW * (A > 0)
__global__ void avgpool_back_kernel(const float *dY, const uint16_t *A, uint16_t *dX, int N, int HW, int C)
{
    const int64_t total = (int64_t)N * HW * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int n = (int)(i / ((int64_t)HW * C));
        float v = dY[n * C + c] / HW;
        if (!(bf2f(A[i]) > 0.f)) v = 0.f;
        dX[i] = f2bf(v);
    }
}

// Softmax cross-entropy over the three heads against target distributions;
// dlogits = (softmax - target) / N (FP32 and BF16), loss accumulated.
__global__ void softmax_ce_kernel(const float *logits, const float *target, float *dlog, uint16_t *dlogb, float *loss,
                                  int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float *z = logits + (int64_t)n * NN_OUT;
    const float *tg = target + (int64_t)n * NN_OUT;
    const int off[3] = { 0, 64, 128 }, len[3] = { 64, 64, 3 };
    float l = 0.f;
    for (int h = 0; h < 3; h++) {
        float mx = -1e30f;
        for (int i = 0; i < len[h]; i++) mx = fmaxf(mx, z[off[h] + i]);
        float se = 0.f;
        for (int i = 0; i < len[h]; i++) se += __expf(z[off[h] + i] - mx);
        const float lse = mx + __logf(se);
        for (int i = 0; i < len[h]; i++) {
            const float p = __expf(z[off[h] + i] - lse);
            const float d = (p - tg[off[h] + i]) / N;
            dlog[(int64_t)n * NN_OUT + off[h] + i] = d;
            dlogb[(int64_t)n * NN_OUT + off[h] + i] = f2bf(d);
            l -= tg[off[h] + i] * (z[off[h] + i] - lse);
        }
    }
    atomicAdd(loss, l / N);
}

// ReLU backward in place on the FP32 gradient, plus its BF16 copy.
__global__ void relu_mask_kernel(float *d, const uint16_t *A, uint16_t *out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = bf2f(A[i]) > 0.f ? d[i] : 0.f;
        d[i] = v;
        out[i] = f2bf(v);
    }
}

__global__ void bf16_colsum_kernel(const uint16_t *dY, float *db, int M, int N)
{
    __shared__ float part[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += 256) s += bf2f(dY[(int64_t)m * N + n]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = part[0];
}

__global__ void zero_f32_kernel(float *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

// bias gradient: column sums of dY [M][N]
__global__ void colsum_kernel(const float *dY, float *db, int M, int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; m++) s += dY[(int64_t)m * N + n];
    db[n] = s;
}

__global__ void f32_to_bf16_kernel(const float *x, uint16_t *y, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// Adam on FP32 master weights; the BF16 copy the GEMMs read is refreshed.
__global__ void adam_kernel(float *w, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2, float wd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

// ---- tiles from frames --------------------------------------------------------
// tiles[fb][y][x][0] = recon luma / maxv, tiles[fb][y][x][1] = (source - recon)
// * 16 / maxv (the coding error, amplified so BF16 keeps its low bits); NHWC
// with C = 2; samples past the frame repeat the edge.
template <typename T>
__device__ __forceinline__ float luma_at(const PlaneView &p, int y, int x)
{
    y = min(max(y, 0), p.h - 1);
    x = min(max(x, 0), p.w - 1);
    return (float)row_ptr<T>(p, y)[x];
}

__global__ void tiles_kernel(PlaneView S, PlaneView R, int bd, int nhfb, uint16_t *tiles)
{
    const int fb = blockIdx.x;
    const int fx = (fb % nhfb) * 64, fy = (fb / nhfb) * 64;
    const float inv = 1.f / (float)((1 << bd) - 1);
    for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
        const int y = i >> 6, x = i & 63;
        float s, r;
        if (bd == 8) { s = luma_at<uint8_t>(S, fy + y, fx + x); r = luma_at<uint8_t>(R, fy + y, fx + x); }
        else { s = luma_at<uint16_t>(S, fy + y, fx + x); r = luma_at<uint16_t>(R, fy + y, fx + x); }
        tiles[((int64_t)fb * 4096 + i) * 2 + 0] = f2bf(r * inv);
        tiles[((int64_t)fb * 4096 + i) * 2 + 1] = f2bf((s - r) * 16.f * inv);
    }
}

hipError_t launch_tiles(const PlaneView &S, const PlaneView &R, int bd, int nhfb, int nfb, uint16_t *tiles, hipStream_t s)
{
    hipLaunchKernelGGL(tiles_kernel, dim3(nfb), dim3(256), 0, s, S, R, bd, nhfb, tiles);
    return hipGetLastError();
}

// ---- fused inference ----------------------------------------------------------
// One workgroup per 64x64 filter block runs the whole forward pass from the
// frame: the tile is built in LDS, the three convolutions are MFMA GEMMs whose
// A fragments are gathered straight from the LDS activations (the patch
// matrix is never written), then pooling, fc1 and the heads; only the 131
// logits leave the CU.  Convolution weights come as pre-packed B fragments
// (one 8-byte load per lane), fc weights from the BF16 parameter copy.
// Rounding points are those of the layer-by-layer pass: BF16 tile, BF16
// activations after bias and ReLU, BF16 pool and fc1, FP32 logits.  LDS:
// the tile (16 KB) is overwritten by conv2's output, conv1's (32 KB) by
// conv3's.
//
// Fragment layout per conv layer: [k-step][16-column block][lane][4] with
// lane l holding B[16 ks + 4 (l / 16) + e][16 cb + l % 16], zero past K.
constexpr int kFragOff[3] = { 0, 512, 5120 };

__global__ void pack_frag_kernel(const uint16_t *Pb, int64_t w0, int64_t w1, int64_t w2, uint16_t *frag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NN_FRAG_ELEMS) return;
    const int l = i < kFragOff[1] ? 0 : i < kFragOff[2] ? 1 : 2;
    const int o = i - kFragOff[l];
    const int K = l == 0 ? 18 : l == 1 ? 144 : 288, N = l == 0 ? 16 : l == 1 ? 32 : 64, CB = N / 16;
    const int e = o & 3, lane = (o >> 2) & 63, blk = o >> 8, cb = blk % CB, ks = blk / CB;
    const int k = 16 * ks + 4 * (lane >> 4) + e, n = 16 * cb + (lane & 15);
    const int64_t woff = l == 0 ? w0 : l == 1 ? w1 : w2;
    frag[i] = k < K ? Pb[woff + (int64_t)k * N + n] : (uint16_t)0;
}

hipError_t nn_pack_frag(const NNWeights &w, uint16_t *frag, hipStream_t s)
{
    hipLaunchKernelGGL(pack_frag_kernel, dim3((NN_FRAG_ELEMS + 255) / 256), dim3(256), 0, s, w.wb[0], (int64_t)0,
                       w.woff[1] - w.woff[0], w.woff[2] - w.woff[0], frag);
    return hipGetLastError();
}

// Four consecutive BF16 values (8-byte aligned) as an MFMA operand.
__device__ __forceinline__ v4s ld_frag(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return (v4s){ (short)(u.x & 0xffff), (short)(u.x >> 16), (short)(u.y & 0xffff), (short)(u.y >> 16) };
}

template <bool PROPOSE, typename T>
__global__ __launch_bounds__(256) void nn_infer_kernel(PlaneView S, PlaneView R, int bd, int nhfb, NNInfer p,
                                                       float *logits, NNPropose pr)
{
    __shared__ __align__(16) uint16_t bufA[64 * 64 * 2];    // tile [64][64][2], then conv2 out [256][32]
    __shared__ __align__(16) uint16_t bufB[32 * 32 * 16];   // conv1 out [1024][16], then conv3 out [64][64]
    __shared__ float pool[64], fc1[128], fcp[4][NN_OUT];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, g = lane >> 4, l16 = lane & 15;
    const int fb = blockIdx.x, fx = (fb % nhfb) * 64, fy = (fb / nhfb) * 64;
    const float inv = 1.f / (float)((1 << bd) - 1);
    uint32_t *tile = reinterpret_cast<uint32_t *>(bufA);
    {
        // every load of the lane in flight before the first conversion
        T sv[16], rv[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = t + k * 256;
            const int y = min(fy + (i >> 6), R.h - 1), x = min(fx + (i & 63), R.w - 1);
            sv[k] = row_ptr<T>(S, y)[x];
            rv[k] = row_ptr<T>(R, y)[x];
        }
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const float s = sv[k], r = rv[k];
            tile[t + k * 256] = (uint32_t)f2bf(r * inv) | ((uint32_t)f2bf((s - r) * 16.f * inv) << 16);
        }
    }
    __syncthreads();
    // conv1: M 1024 (64 row blocks, one output half-row each), N 16, K 18 in two k-steps
    {
        const v4s b0 = ld_frag(p.frag + kFragOff[0] + lane * 4), b1 = ld_frag(p.frag + kFragOff[0] + (64 + lane) * 4);
        const float bias = p.bias[0][l16];
        auto tap = [&](int oy, int ox, int tp) -> uint32_t {
            const int iy = 2 * oy + tp / 3 - 1, ix = 2 * ox + tp % 3 - 1;
            return (iy >= 0 && iy < 64 && ix >= 0 && ix < 64) ? tile[iy * 64 + ix] : 0u;
        };
        for (int rb = w; rb < 64; rb += 4) {
            const int oy = rb >> 1, ox = (rb & 1) * 16 + l16;
            const uint32_t t0 = tap(oy, ox, 2 * g), t1 = tap(oy, ox, 2 * g + 1), t8 = g == 0 ? tap(oy, ox, 8) : 0u;
            const v4s a0 = { (short)(t0 & 0xffff), (short)(t0 >> 16), (short)(t1 & 0xffff), (short)(t1 >> 16) };
            const v4s a1 = { (short)(t8 & 0xffff), (short)(t8 >> 16), 0, 0 };
            v4f acc = { 0.f, 0.f, 0.f, 0.f };
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) bufB[(rb * 16 + 4 * g + r) * 16 + l16] = f2bf(fmaxf(acc[r] + bias, 0.f));
        }
    }
    __syncthreads();
    // conv2: M 256 (16 row blocks = output rows), N 32, K 144 (one tap per k-step); wave w: rows 4w..4w+3
    {
        v4s bf[9][2];
#pragma unroll
        for (int ks = 0; ks < 9; ks++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) bf[ks][cb] = ld_frag(p.frag + kFragOff[1] + ((ks * 2 + cb) * 64 + lane) * 4);
        v4f acc[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++) acc[a][0] = acc[a][1] = (v4f){ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 9; ks++) {
            const int ky = ks / 3, kx = ks % 3, ix = 2 * l16 + kx - 1;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int iy = 2 * (4 * w + a) + ky - 1;
                const v4s av = (iy >= 0 && iy < 32 && ix >= 0) ? ld_frag(&bufB[(iy * 32 + ix) * 16 + 4 * g])
                                                                : (v4s){ 0, 0, 0, 0 };
                acc[a][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bf[ks][0], acc[a][0], 0, 0, 0);
                acc[a][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bf[ks][1], acc[a][1], 0, 0, 0);
            }
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const float bias = p.bias[1][cb * 16 + l16];
#pragma unroll
                for (int r = 0; r < 4; r++)
                    bufA[((4 * w + a) * 16 + 4 * g + r) * 32 + cb * 16 + l16] = f2bf(fmaxf(acc[a][cb][r] + bias, 0.f));
            }
    }
    __syncthreads();
    // conv3: M 64 (4 row blocks of two output rows), N 64 (wave w: columns 16w..16w+15), K 288
    {
        v4s bf[18];
#pragma unroll
        for (int ks = 0; ks < 18; ks++) bf[ks] = ld_frag(p.frag + kFragOff[2] + ((ks * 4 + w) * 64 + lane) * 4);
        v4f acc[4];
#pragma unroll
        for (int a = 0; a < 4; a++) acc[a] = (v4f){ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 18; ks++) {
            const int tp = ks >> 1, h = ks & 1, ky = tp / 3, kx = tp % 3;
            const int ix = 2 * (l16 & 7) + kx - 1;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int iy = 2 * (2 * a + (l16 >> 3)) + ky - 1;
                const v4s av = (iy >= 0 && iy < 16 && ix >= 0) ? ld_frag(&bufA[(iy * 16 + ix) * 32 + 16 * h + 4 * g])
                                                                : (v4s){ 0, 0, 0, 0 };
                acc[a] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bf[ks], acc[a], 0, 0, 0);
            }
        }
        const float bias = p.bias[2][16 * w + l16];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int r = 0; r < 4; r++) bufB[(a * 16 + 4 * g + r) * 64 + 16 * w + l16] = f2bf(fmaxf(acc[a][r] + bias, 0.f));
    }
    __syncthreads();
    if (t < 64) {   // average pool, in pixel order as the layer-by-layer pass
        float sum = 0.f;
#pragma unroll 16
        for (int q = 0; q < 64; q++) sum += bf2f(bufB[q * 64 + t]);
        pool[t] = bf2f(f2bf(sum / 64));
    }
    __syncthreads();
    // fc1 and the heads: the inputs split in parts summed by separate threads
    // (every weight load of a part in flight at once), then the parts added
    // in order
    {
        const int o = t & 127, h = t >> 7;   // fc1: two halves of the 64 inputs
        float sum = 0.f;
#pragma unroll
        for (int c = 0; c < 32; c++) sum += pool[32 * h + c] * bf2f(p.w3[(32 * h + c) * 128 + o]);
        fcp[h][o] = sum;
    }
    __syncthreads();
    if (t < 128) fc1[t] = bf2f(f2bf(fmaxf(fcp[0][t] + fcp[1][t] + p.b3[t], 0.f)));
    __syncthreads();
    for (int it = t; it < 4 * NN_OUT; it += 256) {   // heads: four quarters of the 128 inputs
        const int o = it % NN_OUT, q = it / NN_OUT;
        float sum = 0.f;
#pragma unroll
        for (int j = 0; j < 32; j++) sum += fc1[32 * q + j] * bf2f(p.w4[(32 * q + j) * NN_OUT + o]);
        fcp[q][o] = sum;
    }
    __syncthreads();
    float lg = 0.f;
    if (t < NN_OUT) {
        lg = fcp[0][t] + fcp[1][t] + fcp[2][t] + fcp[3][t] + p.b4[t];
        logits[(int64_t)fb * NN_OUT + t] = lg;
    }
    if constexpr (PROPOSE) {
        // waves 0 and 1 hold the luma and chroma strength logits: softmax per wave
        const bool active = fb_active(pr.noskip, fb, nhfb, pr.mi_cols, pr.mi_rows);
        float *pb = pr.probs + (size_t)fb * NN_PROP_STRIDE;
        if (w < 2) {
            float mx = lg;
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            const float e = __expf(lg - mx);
            float se = e;
            for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
            const float pv = e / se;
            pb[t] = active ? pv : 0.f;
            if (w == 0) {
                float pm = pv;
                for (int o = 32; o > 0; o >>= 1) pm = fmaxf(pm, __shfl_xor(pm, o, 64));
                if (t == 0) {
                    pb[128] = active ? pm : 0.f;
                    pb[129] = active ? 1.f : 0.f;
                }
            }
        }
    }
}

// Shortlist from the per-block probabilities: the column sums in two passes
// with a fixed order (so the ranking does not depend on timing): workgroup g
// sums the blocks 64 g .. 64 g + 63 of every column, then one workgroup adds
// the partial sums in g order and takes the NCAND strengths with the largest
// sums per plane (ascending strength order) and the mean largest luma
// probability.
constexpr int PROP_CHUNK = 64;
__global__ __launch_bounds__(256) void propose_partial_kernel(const float *probs, int nfb, float *partial)
{
    const int g = blockIdx.x, f0 = g * PROP_CHUNK, f1 = min(nfb, f0 + PROP_CHUNK);
    for (int c = threadIdx.x; c < NN_PROP_STRIDE; c += 256) {
        float acc = 0.f;
#pragma unroll 16
        for (int f = f0; f < f1; f++) acc += probs[(size_t)f * NN_PROP_STRIDE + c];
        partial[(size_t)g * NN_PROP_STRIDE + c] = acc;
    }
}

__global__ __launch_bounds__(1024) void propose_kernel(const float *partial, int ng, int *cand, float *conf)
{
    __shared__ float sums[NN_PROP_STRIDE];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < NN_PROP_STRIDE) {
        float v = 0.f;
#pragma unroll 16
        for (int g = 0; g < ng; g++) v += partial[(size_t)g * NN_PROP_STRIDE + t];
        sums[t] = v;
    }
    __syncthreads();
    if (w < 2) {
        // rank = strengths strictly ahead (larger sum, or equal and earlier)
        const float v = sums[t];
        int rank = 0;
        for (int j = 0; j < 64; j++) {
            const float u = sums[w * 64 + j];
            rank += (u > v) || (u == v && j < lane);
        }
        const bool in = rank < NCAND;
        const uint64_t vote = __ballot(in);
        if (in) cand[w * NCAND + __popcll(vote & ((1ull << lane) - 1))] = lane;
    }
    if (t == 0) *conf = sums[129] > 0.f ? sums[128] / sums[129] : 1.f;
}

hipError_t nn_infer(const NNInfer &p, const PlaneView &S, const PlaneView &R, int bd, int nhfb, int nfb, float *logits,
                    const NNPropose *prop, hipStream_t s)
{
    const NNPropose none{};
    const NNPropose &pp = prop ? *prop : none;
    if (prop && bd == 8)
        hipLaunchKernelGGL((nn_infer_kernel<true, uint8_t>), dim3(nfb), dim3(256), 0, s, S, R, bd, nhfb, p, logits, pp);
    else if (prop)
        hipLaunchKernelGGL((nn_infer_kernel<true, uint16_t>), dim3(nfb), dim3(256), 0, s, S, R, bd, nhfb, p, logits, pp);
    else if (bd == 8)
        hipLaunchKernelGGL((nn_infer_kernel<false, uint8_t>), dim3(nfb), dim3(256), 0, s, S, R, bd, nhfb, p, logits, pp);
    else
        hipLaunchKernelGGL((nn_infer_kernel<false, uint16_t>), dim3(nfb), dim3(256), 0, s, S, R, bd, nhfb, p, logits, pp);
    return hipGetLastError();
}

hipError_t nn_propose(const NNPropose &prop, int nfb, hipStream_t s)
{
    const int ng = (nfb + PROP_CHUNK - 1) / PROP_CHUNK;
    hipLaunchKernelGGL(propose_partial_kernel, dim3(ng), dim3(256), 0, s, prop.probs, nfb, prop.partial);
    hipLaunchKernelGGL(propose_kernel, dim3(1), dim3(1024), 0, s, prop.partial, ng, prop.cand, prop.conf);
    return hipGetLastError();
}

// ---- network passes -------------------------------------------------------------
static int grid_for(int64_t n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }

static GemmArgs gemm(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn, int M,
                     int N, int K)
{
    GemmArgs g{};
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    g.M = M; g.N = N; g.K = K; g.ldc = N;
    return g;
}

hipError_t nn_forward(const NNWeights &w, const NNWork &ws, const uint16_t *tiles, int n, float *logits, hipStream_t s)
{
    static const int Hs[3] = { 64, 32, 16 }, Cin[3] = { 2, 16, 32 }, Cout[3] = { 16, 32, 64 };
    const uint16_t *x = tiles;
    for (int l = 0; l < 3; l++) {
        const int H = Hs[l], Ho = H / 2, K = 9 * Cin[l];
        const int64_t rows = (int64_t)n * Ho * Ho;
        hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(rows * K)), dim3(256), 0, s, x, ws.cols[l], n, H, H, Cin[l]);
        GemmArgs g = gemm(ws.cols[l], K, 1, w.wb[l], Cout[l], 1, (int)rows, Cout[l], K);
        g.bias = w.b[l];
        g.relu = 1;
        g.Cb = ws.act[l];
        hipError_t e = launch_gemm(g, s);
        if (e != hipSuccess) return e;
        x = ws.act[l];
    }
    hipLaunchKernelGGL(avgpool_kernel, dim3((n * 64 + 255) / 256), dim3(256), 0, s, ws.act[2], ws.pool, (float *)nullptr, n,
                       64, 64);
    GemmArgs g = gemm(ws.pool, 64, 1, w.wb[3], 128, 1, n, 128, 64);
    g.bias = w.b[3];
    g.relu = 1;
    g.Cb = ws.fc1;
    hipError_t e = launch_gemm(g, s);
    if (e != hipSuccess) return e;
    g = gemm(ws.fc1, 128, 1, w.wb[4], NN_OUT, 1, n, NN_OUT, 128);
    g.bias = w.b[4];
    g.C = logits;
    return launch_gemm(g, s);
}

hipError_t nn_backward(const NNWeights &w, const NNWork &ws, const uint16_t *tiles, int n, const float *target,
                       float *grads, float *loss, hipStream_t s)
{
    static const int Hs[3] = { 64, 32, 16 }, Cin[3] = { 2, 16, 32 }, Cout[3] = { 16, 32, 64 };
    hipError_t e = nn_forward(w, ws, tiles, n, ws.logits, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(softmax_ce_kernel, dim3((n + 63) / 64), dim3(64), 0, s, ws.logits, target, ws.dlog, ws.dlogb,
                       loss, n);
    // heads: dW4 = fc1^T dlog, db4 = colsum, dfc1 = dlog W4^T (masked by fc1 > 0)
    GemmArgs g = gemm(ws.fc1, 1, 128, ws.dlogb, NN_OUT, 1, 128, NN_OUT, n);
    g.C = grads + w.woff[4];
    if ((e = launch_gemm(g, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, s, ws.dlog, grads + w.boff[4], n, NN_OUT);
    g = gemm(ws.dlogb, NN_OUT, 1, w.wb[4], 1, NN_OUT, n, 128, NN_OUT);
    g.C = ws.dfc1f;
    if ((e = launch_gemm(g, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for((int64_t)n * 128)), dim3(256), 0, s, ws.dfc1f, ws.fc1, ws.dfc1b,
                       (int64_t)n * 128);
    // fc1: dW3 = pool^T dfc1, db3, dpool = dfc1 W3^T
    g = gemm(ws.pool, 1, 64, ws.dfc1b, 128, 1, 64, 128, n);
    g.C = grads + w.woff[3];
    if ((e = launch_gemm(g, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, s, ws.dfc1f, grads + w.boff[3], n, 128);
    g = gemm(ws.dfc1b, 128, 1, w.wb[3], 1, 128, n, 64, 128);
    g.C = ws.dpoolf;
    if ((e = launch_gemm(g, s)) != hipSuccess) return e;
    // pool back into conv3's output gradient (masked by ReLU)
    hipLaunchKernelGGL(avgpool_back_kernel, dim3(grid_for((int64_t)n * 64 * 64)), dim3(256), 0, s, ws.dpoolf, ws.act[2],
                       ws.dact[2], n, 64, 64);
    for (int l = 2; l >= 0; l--) {
        const int H = Hs[l], Ho = H / 2, K = 9 * Cin[l];
        const int rows = n * Ho * Ho;
        // weight gradient: cols^T [K][rows] x dY [rows][Cout]
        g = gemm(ws.cols[l], 1, K, ws.dact[l], Cout[l], 1, K, Cout[l], rows);
        g.C = grads + w.woff[l];
        if ((e = launch_gemm(g, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(bf16_colsum_kernel, dim3(Cout[l]), dim3(256), 0, s, ws.dact[l], grads + w.boff[l], rows,
                           Cout[l]);
        if (l == 0) break;
        // patch gradient: dY [rows][Cout] x W^T [Cout][K]
        g = gemm(ws.dact[l], Cout[l], 1, w.wb[l], 1, Cout[l], rows, K, Cout[l]);
        g.C = ws.dcols;
        if ((e = launch_gemm(g, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(col2im_relu_kernel, dim3(grid_for((int64_t)n * H * H * Cin[l])), dim3(256), 0, s, ws.dcols,
                           ws.act[l - 1], ws.dact[l - 1], (float *)nullptr, n, H, H, Cin[l]);
    }
    return hipGetLastError();
}

hipError_t nn_adam(float *w, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, int step, float wd,
                   hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, wb, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2,
                       wd);
    return hipGetLastError();
}

hipError_t nn_zero(float *p, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_f32_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Workspace carve-up for a batch of up to max_batch tiles (256-byte aligned).
static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

size_t nn_work_bytes(int B)
{
    NNWork ws;
    nn_work_carve(nullptr, B, ws);
    return (size_t)(uintptr_t)ws.dcols + al((size_t)B * 256 * 144 * 4);
}

void nn_work_carve(char *base, int B, NNWork &ws)
{
    size_t o = 0;
    auto take = [&](size_t bytes) { char *p = base + o; o += al(bytes); return p; };
    const size_t colsz[3] = { (size_t)1024 * 18, (size_t)256 * 144, (size_t)64 * 288 };
    const size_t actsz[3] = { (size_t)1024 * 16, (size_t)256 * 32, (size_t)64 * 64 };
    for (int l = 0; l < 3; l++) ws.cols[l] = (uint16_t *)take(B * colsz[l] * 2);
    for (int l = 0; l < 3; l++) ws.act[l] = (uint16_t *)take(B * actsz[l] * 2);
    for (int l = 0; l < 3; l++) ws.dact[l] = (uint16_t *)take(B * actsz[l] * 2);
    ws.pool = (uint16_t *)take((size_t)B * 64 * 2);
    ws.fc1 = (uint16_t *)take((size_t)B * 128 * 2);
    ws.logits = (float *)take((size_t)B * NN_OUT * 4);
    ws.dlog = (float *)take((size_t)B * NN_OUT * 4);
    ws.dlogb = (uint16_t *)take((size_t)B * NN_OUT * 2);
    ws.dfc1f = (float *)take((size_t)B * 128 * 4);
    ws.dfc1b = (uint16_t *)take((size_t)B * 128 * 2);
    ws.dpoolf = (float *)take((size_t)B * 64 * 4);
    ws.dcols = (float *)(base + o);   // B * 256 * 144 FP32, the largest patch gradient
}

}  // namespace pf
