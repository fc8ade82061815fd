// Learned offset proposal: host interface of the small 3D CNN that predicts,
// per 32x8x8 filter tile, the distribution of weight mass over the search
// offsets, and of the top-K selection that turns it into per-tile offset
// lists with a fallback to the full search.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace inlm {

constexpr int kMaxProposalOffsets = 2048;  // the top-K sort holds 2048 keys in LDS

class ProposalNet {
public:
    struct Impl;
    ProposalNet();
    ~ProposalNet();
    ProposalNet(const ProposalNet &) = delete;
    ProposalNet &operator=(const ProposalNet &) = delete;

    // ns search offsets (<= kMaxProposalOffsets); He-uniform weights from a
    // SplitMix64 stream seeded with `seed`
    hipError_t init(int ns, int device, uint64_t seed);
    long num_params() const;
    int num_offsets() const;
    // parameter vector layout: w1[27][16] b1[16], then three times
    // w[27 tap][16 cin][16 cout] b[16], then Wh[16][ns] bh[ns]
    hipError_t set_params(const float *host, long n);
    hipError_t get_params(float *host, long n) const;
    hipError_t get_grads(float *host, long n) const;

    // vol: device FP32 volume (x fastest), shape {nx, ny, nz}; the network
    // sees (vol - mean) / std of the volume.  Always computes the per-tile
    // probabilities.  With `mass` (device, [ntiles][ns] weight mass per tile
    // and offset) also the cross-entropy loss; with backward also the
    // gradients.
    hipError_t run(const float *vol, const int shape[3], const float *mass, bool backward, float *loss_host);
    hipError_t adam(float lr);
    const float *probs() const;  // device [ntiles][ns]

private:
    Impl *d;
};

// Per tile: the K most probable offsets in ascending search order (lists,
// [ntiles][K]) and counts[t] = K, or counts[t] = ns (full search) when the
// predicted mass of those K is below `threshold` or the tile is drawn for the
// audit (probability audit_frac).  stats (4 floats, accumulated):
// fallback tiles, audited tiles, sum of predicted captured mass, sum of
// visited offsets.
hipError_t launch_topk(const float *probs, int ntiles, int ns, int K, float threshold, float audit_frac,
                       uint64_t seed, int *lists, int *counts, float *stats, hipStream_t st);

}  // namespace inlm
