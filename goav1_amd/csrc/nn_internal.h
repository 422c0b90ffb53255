// Internal declarations of the learned-proposal network (nn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

constexpr int NN_OUT = 131;        // 64 luma + 64 chroma strength logits, 3 LR-type logits
constexpr int NN_LAYERS = 5;       // conv1..3, fc1, heads
constexpr int NCAND = AV1PF_SHORTLIST;   // shortlisted strengths per plane in the learned search

// Parameter layout of the flat FP32 vector: per layer W [K][N] row-major then b [N].
struct NNShape {
    int K[NN_LAYERS], N[NN_LAYERS];
    int64_t woff[NN_LAYERS], boff[NN_LAYERS], total;
};

inline NNShape nn_shape()
{
    NNShape s{};
    const int K[NN_LAYERS] = { 18, 144, 288, 64, 128 }, N[NN_LAYERS] = { 16, 32, 64, 128, NN_OUT };
    int64_t o = 0;
    for (int l = 0; l < NN_LAYERS; l++) {
        s.K[l] = K[l];
        s.N[l] = N[l];
        s.woff[l] = o;
        o += (int64_t)K[l] * N[l];
        s.boff[l] = o;
        o += N[l];
    }
    s.total = o;
    return s;
}

struct NNWeights {
    const uint16_t *wb[NN_LAYERS];   // BF16 weights (GEMM operands)
    const float *b[NN_LAYERS];       // FP32 biases
    int64_t woff[NN_LAYERS], boff[NN_LAYERS];
};

struct NNWork {
    uint16_t *cols[3];   // im2col patches per conv layer (kept for the backward pass)
    uint16_t *act[3];    // conv outputs after ReLU (BF16, NHWC)
    uint16_t *pool, *fc1;
    float *logits, *dlog, *dfc1f, *dpoolf, *dcols;
    uint16_t *dlogb, *dfc1b, *dact[3];
};

struct GemmArgs {
    const uint16_t *A; int64_t sam, sak;
    const uint16_t *B; int64_t sbk, sbn;
    int M, N, K;
    const float *bias;
    int relu;
    float *C;          // FP32 output (or null)
    uint16_t *Cb;      // BF16 output (or null)
    int64_t ldc;
};

// A filter block takes part in the CDEF search when one of its 8x8 blocks
// inside the frame is not skipped (noskip bit (r / 2) * 8 + c / 2).
__device__ __forceinline__ bool fb_active(const uint64_t *noskip, int fb, int nhfb, int mi_cols, int mi_rows)
{
    const int fbr = fb / nhfb, fbc = fb % nhfb;
    const int nhb = min(16, mi_cols - 16 * fbc), nvb = min(16, mi_rows - 16 * fbr);
    uint64_t valid = 0;
    for (int r = 0; r < nvb; r += 2)
        for (int c = 0; c < nhb; c += 2) valid |= 1ull << ((r >> 1) * 8 + (c >> 1));
    return ((noskip ? noskip[fb] : ~0ull) & valid) != 0;
}

// Fused inference: conv B fragments (packed from the BF16 weights), biases,
// fc weights.
constexpr int NN_FRAG_ELEMS = 23552;
struct NNInfer {
    const uint16_t *frag;
    const float *bias[3];
    const uint16_t *w3, *w4;
    const float *b3, *b4;
};
// Shortlist proposal of the learned search: the inference writes, per
// active filter block, the softmax of the luma and chroma strength logits;
// a one-workgroup pass sums them over the blocks and keeps the NCAND
// strengths with the largest sums per plane.
constexpr int NN_PROP_STRIDE = 130;   // 64 + 64 probabilities, largest luma probability, active flag
struct NNPropose {
    const uint64_t *noskip;
    int mi_cols, mi_rows;
    float *probs;            // [nfb][NN_PROP_STRIDE] scratch
    float *partial;          // [ceil(nfb / 64)][NN_PROP_STRIDE] scratch: column sums of 64 blocks each
    int *cand;               // [2 * NCAND], ascending per plane
    float *conf;             // mean largest luma probability over the active blocks
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);
hipError_t nn_pack_frag(const NNWeights &w, uint16_t *frag, hipStream_t s);
hipError_t nn_infer(const NNInfer &p, const PlaneView &S, const PlaneView &R, int bd, int nhfb, int nfb, float *logits,
                    const NNPropose *prop, hipStream_t s);
// The frame shortlist from the probabilities nn_infer wrote with a proposal
// (per-strength sums in a fixed order, then the NCAND largest per plane).
hipError_t nn_propose(const NNPropose &prop, int nfb, hipStream_t s);
hipError_t launch_tiles(const PlaneView &S, const PlaneView &R, int bd, int nhfb, int nfb, uint16_t *tiles, hipStream_t s);
hipError_t nn_forward(const NNWeights &w, const NNWork &ws, const uint16_t *tiles, int n, float *logits, hipStream_t s);
hipError_t nn_backward(const NNWeights &w, const NNWork &ws, const uint16_t *tiles, int n, const float *target,
                       float *grads, float *loss, hipStream_t s);
hipError_t nn_adam(float *w, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, int step, float wd,
                   hipStream_t s);
hipError_t nn_zero(float *p, int64_t n, hipStream_t s);
size_t nn_work_bytes(int max_batch);
// Exact SSE of the shortlisted strengths and the one-launch selection over
// the shortlist pairs.
hipError_t launch_search_shortlist(const FrameView &src, const FrameView &rec, const uint64_t *noskip, int damping,
                                   uint64_t lambda, const int *cand, uint64_t *mse_sub, av1pf_cdef_decision *dec,
                                   uint8_t *fb_index, void *small_state, hipStream_t s);
// Device state of the shortlist selection (zeroed once): per-level results
// and the ticket.
size_t small_state_bytes();
void nn_work_carve(char *base, int max_batch, NNWork &ws);

}  // namespace pf
