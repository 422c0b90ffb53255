// Host half of the encoder CDEF strength search: from the per-filter-block MSE
// table the device search writes, pick the frame's strength set and each
// block's index (cdef_search.go:780-945, finish / jointStrengthSearchDual /
// searchOneDual), and the full-search lambda (cdef_search.go:274-299).
//
// The table is small (2 x nfb x 64 uint64, 240 filter blocks at 720p) and the
// selection is a greedy pass over it with data-dependent control flow, so it
// stays on the host, as in the reference encoder.
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "av1pf.h"

namespace {

constexpr int kTrials = 64;                 // strength pairs of search level 1
constexpr uint64_t kHuge = uint64_t(1) << 63;

// Strength-list position -> packed strength (primary * 4 + secondary code):
// positions 0..15 are the primaries with secondary 0, then three secondary
// codes per primary (cdef_search.go:72-110, level 1).
int packed_strength(int gi) { return gi < 16 ? 4 * gi : 4 * ((gi - 16) / 3) + (gi - 16) % 3 + 1; }

class StrengthSelector {
public:
    StrengthSelector(const uint64_t *mse, int nfb, const uint8_t *skip_fb) : mse_(mse), nfb_(nfb), gain_(kTrials * kTrials)
    {
        for (int fb = 0; fb < nfb; fb++)
            if (!skip_fb[fb]) active_.push_back(fb);
    }

    int active() const { return (int)active_.size(); }
    uint64_t luma(int fb, int g) const { return mse_[(size_t)fb * kTrials + g]; }
    uint64_t chroma(int fb, int g) const { return mse_[((size_t)nfb_ + fb) * kTrials + g]; }

    // Choose pair `slot` with pairs [0, slot) fixed; returns the resulting
    // total distortion (searchOneDual).
    uint64_t choose(int *ly, int *lc, int slot)
    {
        if (active_.empty()) {
            ly[slot] = lc[slot] = 0;
            return 0;
        }
        if (slot == 0) {   // luma and chroma separately: the lowest column sums
            uint64_t sy[kTrials] = {}, sc[kTrials] = {};
            for (int fb : active_)
                for (int g = 0; g < kTrials; g++) {
                    sy[g] += luma(fb, g);
                    sc[g] += chroma(fb, g);
                }
            int by = 0, bc = 0;
            for (int g = 1; g < kTrials; g++) {
                if (sy[g] < sy[by]) by = g;
                if (sc[g] < sc[bc]) bc = g;
            }
            ly[0] = by;
            lc[0] = bc;
            return sy[by] + sc[bc];
        }
        // gain of adding pair (j, k): for every block, how far it would lower
        // the block's best cost among the pairs already chosen
        std::fill(gain_.begin(), gain_.end(), 0);
        uint64_t base = 0;
        for (int fb : active_) {
            uint64_t best = luma(fb, ly[0]) + chroma(fb, lc[0]);
            for (int g = 1; g < slot; g++) {
                const uint64_t c = luma(fb, ly[g]) + chroma(fb, lc[g]);
                if (c < best) best = c;
            }
            base += best;
            for (int j = 0; j < kTrials; j++) {
                const uint64_t yj = luma(fb, j);
                if (yj >= best) continue;
                const uint64_t room = best - yj;
                uint64_t *row = &gain_[(size_t)j * kTrials];
                for (int k = 0; k < kTrials; k++) {
                    const uint64_t ck = chroma(fb, k);
                    if (ck < room) row[k] += room - ck;
                }
            }
        }
        int bj = 0, bk = 0;
        uint64_t bg = gain_[0];
        for (int j = 0; j < kTrials; j++)
            for (int k = 0; k < kTrials; k++)
                if (gain_[(size_t)j * kTrials + k] > bg) {
                    bg = gain_[(size_t)j * kTrials + k];
                    bj = j;
                    bk = k;
                }
        ly[slot] = bj;
        lc[slot] = bk;
        return base - bg;
    }

private:
    const uint64_t *mse_;
    int nfb_;
    std::vector<int> active_;
    std::vector<uint64_t> gain_;
};

}  // namespace

extern "C" {

uint64_t av1pf_cdef_lambda(int dc_quant, int key_frame, int gf_update, int tl_factor)
{
    const double q = dc_quant;
    const double mult = key_frame ? 3.3 : (gf_update ? 3.25 : 3.2);
    int64_t rdmult = (int64_t)((mult + 0.0015 * q) * q * q);
    if (key_frame) tl_factor = 150;
    rdmult = (rdmult * tl_factor) >> 7;
    if (key_frame) rdmult = (rdmult * 100) >> 7;
    if (rdmult < 1) rdmult = 1;
    return (uint64_t)rdmult;
}

int av1pf_cdef_search_finish(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda,
                             uint8_t y_strength[8], uint8_t uv_strength[8], uint8_t *fb_index, int *cdef_bits)
{
    if (!mse || nfb < 0 || (nfb && (!skip_fb || !fb_index)) || !y_strength || !uv_strength || !cdef_bits)
        return AV1PF_ERR_INVALID_ARG;
    StrengthSelector sel(mse, nfb, skip_fb);
    const int nsb = sel.active();
    uint64_t best_cost = kHuge;
    int best_bits = 0;
    int ys[8] = {}, us[8] = {};
    for (int bits = 0; bits <= 3; bits++) {
        const int nb = 1 << bits;
        int ly[8] = {}, lc[8] = {};
        uint64_t dist = kHuge;
        for (int k = 0; k < nb; k++) dist = sel.choose(ly, lc, k);
        // refinement: drop the oldest pair and choose the last slot again
        for (int it = 0; it < 4 * nb; it++) {
            memmove(ly, ly + 1, sizeof(int) * (nb - 1));
            memmove(lc, lc + 1, sizeof(int) * (nb - 1));
            dist = sel.choose(ly, lc, nb - 1);
        }
        const int64_t rate = (int64_t)(nsb * bits + nb * 12) << 9;
        const uint64_t cost = (uint64_t)((rate * (int64_t)lambda + (1 << 8)) >> 9) + ((dist * 16) << 7);
        if (cost < best_cost) {
            best_cost = cost;
            best_bits = bits;
            memcpy(ys, ly, sizeof(ys));
            memcpy(us, lc, sizeof(us));
        }
    }
    const int nb = 1 << best_bits;
    for (int fb = 0; fb < nfb; fb++) {
        fb_index[fb] = 0;
        if (skip_fb[fb]) continue;
        uint64_t bm = kHuge;
        for (int g = 0; g < nb; g++) {
            const uint64_t c = sel.luma(fb, ys[g]) + sel.chroma(fb, us[g]);
            if (c < bm) {
                bm = c;
                fb_index[fb] = (uint8_t)g;
            }
        }
    }
    for (int i = 0; i < 8; i++) {
        y_strength[i] = i < nb ? (uint8_t)packed_strength(ys[i]) : 0;
        uv_strength[i] = i < nb ? (uint8_t)packed_strength(us[i]) : 0;
    }
    *cdef_bits = best_bits;
    return AV1PF_OK;
}

}  // extern "C"
