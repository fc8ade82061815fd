// Learned offset proposal: a small 3D CNN that scores, per 32x8x8 filter
// tile, how much weight mass each search offset carries, trained on the GPU
// with these kernels.
//
// Network (per voxel, "same" zero padding, ReLU after every conv):
//   conv1 1->16 (3^3)  conv2..4 16->16 (3^3)  mean-pool per tile -> 16 features
//   head 16 -> N_s logits, softmax over the search offsets.
// Loss: cross-entropy of the softmax against the tile's measured weight-mass
// distribution over offsets (mass / total), averaged over tiles.
//
// Activations are channels-last BF16 with a one-voxel zero halo:
// act[((z+1)*(ny+2) + (y+1))*(nx+2) + (x+1)][16].  Convolutions 2..4 (forward
// and the gradient with respect to their input) are implicit GEMMs on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation: a wavefront computes 16
// consecutive voxels x 16 output channels, K = 27 taps x 16 channels in 14
// chunks of two taps.  The weight gradients (a reduction over all voxels) and
// the first layer (one input channel) run on the vector ALUs in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "inlm_kernels.hpp"
#include "inlm_learn.hpp"

namespace inlm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int C = 16;      // channels
constexpr int NTAP = 27;
constexpr int PTX = 32, PTY = 8, PTZ = 8;  // proposal tile = filter tile

struct ActGeom {
    int nx, ny, nz;       // unpadded
    long sx, sy;          // padded strides in voxels: x 1, y (nx+2), z (nx+2)(ny+2)
    __device__ __host__ long vox(int x, int y, int z) const { return (long)(z + 1) * sy + (long)(y + 1) * sx + (x + 1); }
};

__device__ __host__ inline long tap_off(const ActGeom &g, int tap)
{
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    return (long)dz * g.sy + (long)dy * g.sx + dx;
}

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }

// ---------------------------------------------------------------------------
// layer 1: input (FP32, padded with zero halo, one channel) -> 16 channels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv1_fwd(const float *__restrict__ in, const float *__restrict__ w,
                                                 const float *__restrict__ b, __bf16 *__restrict__ out, ActGeom g)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((long)g.nx * g.ny));
    const long v = g.vox(x, y, z);
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = b[c];
    for (int t = 0; t < NTAP; t++) {
        const float xv = in[v + tap_off(g, t)];
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] = __builtin_fmaf(w[t * C + c], xv, acc[c]);
    }
    __bf16 o[C];
#pragma unroll
    for (int c = 0; c < C; c++) o[c] = (__bf16)__builtin_fmaxf(acc[c], 0.f);
    __builtin_memcpy(out + v * C, o, sizeof o);
}

// ---------------------------------------------------------------------------
// 16 -> 16 channel 3^3 convolution on MFMA.  wk is [27 tap][16 n][16 k] BF16
// (n = output column of the GEMM, k = input channel).  mode 0: forward,
// out = relu(acc + bias).  mode 1: input gradient, out = acc * (mask > 0)
// with mask the forward input activation of this layer.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv16_mfma(const __bf16 *__restrict__ X, const __bf16 *__restrict__ wk,
                                                   const float *__restrict__ bias, const __bf16 *__restrict__ mask,
                                                   __bf16 *__restrict__ out, ActGeom g, int mode)
{
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nxb = (g.nx + 15) / 16;
    const long nwave = (long)nxb * g.ny * g.nz;
    if (wave >= nwave) return;  // whole wavefront leaves together
    const int xb = (int)(wave % nxb);
    const int y = (int)((wave / nxb) % g.ny);
    const int z = (int)(wave / ((long)nxb * g.ny));
    const int x0 = xb * 16;
    const int m = lane & 15, q = lane >> 4;
    const long vm = g.vox(x0 + m, y, z);  // A row of this lane (may be past nx: reads halo/guard, result dropped)
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 14; c++) {
        const int tap = 2 * c + (q >> 1);
        const int k0 = 8 * (q & 1);
        bf16x8 a, b;
        if (tap < NTAP) {
            __builtin_memcpy(&a, X + (vm + tap_off(g, tap)) * C + k0, 16);
            __builtin_memcpy(&b, wk + ((long)tap * C + m) * C + k0, 16);  // B[k][n]: lane n = m reads its row
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                a[e] = (__bf16)0.f;
                b[e] = (__bf16)0.f;
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // C[row 4q+i][col m]: voxel x0 + 4q + i, channel m
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int xv = x0 + 4 * q + i;
        if (xv >= g.nx) continue;
        const long v = g.vox(xv, y, z);
        float r;
        if (mode == 0) r = __builtin_fmaxf(acc[i] + bias[m], 0.f);
        else r = bf2f(mask[v * C + m]) > 0.f ? acc[i] : 0.f;
        out[v * C + m] = (__bf16)r;
    }
}

// ---------------------------------------------------------------------------
// Weight gradients on the vector ALUs (FP32): dW[tap][k][n] += sum_v
// X[v + tap][k] * G[v][n] and dB[n] += sum_v G[v][n].  A block of 256 threads
// = (k, n) pairs takes a chunk of 64 voxels of one row at a time through LDS.
// Cin is 16 (BF16 activations) or 1 (the FP32 input, layer 1).  Each of the
// kWgradBlocks blocks writes its partial sums (part[block][tap][k][n], then
// part[block][NTAP CIN C + n] for the bias) and wgrad_reduce adds them in
// block order, so training repeats bit for bit.
// ---------------------------------------------------------------------------
constexpr int kWgradBlocks = 2048;
constexpr int kWgradStride = NTAP * C * C + C;  // floats per block's partials
template <int CIN>
__global__ __launch_bounds__(256) void conv_wgrad(const void *__restrict__ Xv, const __bf16 *__restrict__ G,
                                                  float *__restrict__ part, ActGeom g, int rows_per_block)
{
    constexpr int CH = 64;                   // voxels per chunk along x
    __shared__ float xs[NTAP][CH][CIN];     // X at the 27 shifted positions
    __shared__ float gs[CH][C];
    const int tid = threadIdx.x;
    const int k = tid / C, n = tid % C;     // valid for CIN = 16; CIN = 1 uses k = 0 and tid < 16
    float acc[NTAP];
#pragma unroll
    for (int t = 0; t < NTAP; t++) acc[t] = 0.f;
    float accb = 0.f;
    const long nrows = (long)g.ny * g.nz;
    const long r0 = (long)blockIdx.x * rows_per_block;
    for (long r = r0; r < r0 + rows_per_block && r < nrows; r++) {
        const int y = (int)(r % g.ny), z = (int)(r / g.ny);
        for (int x0 = 0; x0 < g.nx; x0 += CH) {
            const int nv = g.nx - x0 < CH ? g.nx - x0 : CH;
            __syncthreads();
            for (int i = tid; i < CH * C; i += 256) {
                const int xv = i / C, c = i % C;
                gs[xv][c] = xv < nv ? bf2f(G[g.vox(x0 + xv, y, z) * C + c]) : 0.f;
            }
            for (int i = tid; i < NTAP * CH * CIN; i += 256) {
                const int t = i / (CH * CIN), rem = i % (CH * CIN), xv = rem / CIN, c = rem % CIN;
                float val = 0.f;
                if (xv < nv) {
                    const long v = g.vox(x0 + xv, y, z) + tap_off(g, t);
                    if constexpr (CIN == 1) val = static_cast<const float *>(Xv)[v];
                    else val = bf2f(static_cast<const __bf16 *>(Xv)[v * C + c]);
                }
                xs[t][xv][c] = val;
            }
            __syncthreads();
            if (CIN == C || tid < C) {
                for (int xv = 0; xv < nv; xv++) {
                    const float gv = gs[xv][n];
                    accb += gv;
#pragma unroll
                    for (int t = 0; t < NTAP; t++) acc[t] = __builtin_fmaf(xs[t][xv][CIN == 1 ? 0 : k], gv, acc[t]);
                }
            }
        }
    }
    float *pb = part + (long)blockIdx.x * kWgradStride;
    if (CIN == C || tid < C) {
        for (int t = 0; t < NTAP; t++) pb[(t * CIN + (CIN == 1 ? 0 : k)) * C + n] = acc[t];
        if (CIN == 1 || k == 0) pb[NTAP * CIN * C + n] = accb;
    }
}

// dW[e] += sum over blocks b (in order) of part[b][e], e < NTAP CIN C, and
// dB likewise from the bias partials
__global__ __launch_bounds__(256) void wgrad_reduce(const float *__restrict__ part, int nblocks, int nw,
                                                    float *__restrict__ dW, float *__restrict__ dB)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nw + C) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += part[(long)b * kWgradStride + e];
    if (e < nw) dW[e] += s;
    else dB[e - nw] += s;
}

// ---------------------------------------------------------------------------
// Mean pool per 32x8x8 tile (tiles x-fastest, partial tiles at the edges).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pool_fwd(const __bf16 *__restrict__ A, float *__restrict__ feat, ActGeom g,
                                                int ntx, int nty)
{
    const int t = blockIdx.x;
    const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    __shared__ float part[256][C + 1];
    float s[C];
#pragma unroll
    for (int c = 0; c < C; c++) s[c] = 0.f;
    for (int i = threadIdx.x; i < PTX * PTY * PTZ; i += 256) {
        const int x = bx * PTX + i % PTX, y = by * PTY + (i / PTX) % PTY, z = bz * PTZ + i / (PTX * PTY);
        if (x < g.nx && y < g.ny && z < g.nz) {
            __bf16 v[C];
            __builtin_memcpy(v, A + g.vox(x, y, z) * C, sizeof v);
#pragma unroll
            for (int c = 0; c < C; c++) s[c] += bf2f(v[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; c++) part[threadIdx.x][c] = s[c];
    __syncthreads();
    if (threadIdx.x < C) {
        float tot = 0.f;
        for (int i = 0; i < 256; i++) tot += part[i][threadIdx.x];
        const int nxv = min(PTX, g.nx - bx * PTX), nyv = min(PTY, g.ny - by * PTY), nzv = min(PTZ, g.nz - bz * PTZ);
        feat[(long)t * C + threadIdx.x] = tot / (float)(nxv * nyv * nzv);
    }
}

// gradient of the pool (times the ReLU mask of the pooled activation)
__global__ __launch_bounds__(256) void pool_bwd(const float *__restrict__ dfeat, const __bf16 *__restrict__ A,
                                                __bf16 *__restrict__ G, ActGeom g, int ntx, int nty)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((long)g.nx * g.ny));
    const int bx = x / PTX, by = y / PTY, bz = z / PTZ;
    const int t = (bz * nty + by) * ntx + bx;
    const int nxv = min(PTX, g.nx - bx * PTX), nyv = min(PTY, g.ny - by * PTY), nzv = min(PTZ, g.nz - bz * PTZ);
    const float inv = 1.f / (float)(nxv * nyv * nzv);
    const long v = g.vox(x, y, z);
    __bf16 a[C], o[C];
    __builtin_memcpy(a, A + v * C, sizeof a);
#pragma unroll
    for (int c = 0; c < C; c++) o[c] = (__bf16)(bf2f(a[c]) > 0.f ? dfeat[(long)t * C + c] * inv : 0.f);
    __builtin_memcpy(G + v * C, o, sizeof o);
}

// ---------------------------------------------------------------------------
// Head: logits = bh + feat . Wh, softmax over the N_s offsets (one block per
// tile).  With a target mass it writes the tile's cross-entropy and
// dlogits = (q - p) / ntiles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float block_reduce(float v, float *sh, bool is_max)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float u = __shfl_xor(v, o, 64);
        v = is_max ? __builtin_fmaxf(v, u) : v + u;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sh[0];
    for (int w = 1; w < 4; w++) r = is_max ? __builtin_fmaxf(r, sh[w]) : r + sh[w];
    return r;
}

__global__ __launch_bounds__(256) void head_fwd(const float *__restrict__ feat, const float *__restrict__ Wh,
                                                const float *__restrict__ bh, int ns, const float *__restrict__ mass,
                                                float *__restrict__ probs, float *__restrict__ dlogits,
                                                float *__restrict__ tile_loss, float inv_tiles)
{
    extern __shared__ float lg[];  // ns logits
    __shared__ float sh[4];
    const int t = blockIdx.x;
    float f[C];
#pragma unroll
    for (int c = 0; c < C; c++) f[c] = feat[(long)t * C + c];
    float mx = -3.0e38f;
    for (int s = threadIdx.x; s < ns; s += 256) {
        float l = bh[s];
#pragma unroll
        for (int c = 0; c < C; c++) l = __builtin_fmaf(f[c], Wh[(long)c * ns + s], l);
        lg[s] = l;
        mx = __builtin_fmaxf(mx, l);
    }
    mx = block_reduce(mx, sh, true);
    float se = 0.f;
    for (int s = threadIdx.x; s < ns; s += 256) se += expf(lg[s] - mx);
    se = block_reduce(se, sh, false);
    const float lse = mx + logf(se);
    float tm = 0.f;
    if (mass)
        for (int s = threadIdx.x; s < ns; s += 256) tm += mass[(long)t * ns + s];
    if (mass) tm = block_reduce(tm, sh, false);
    float ce = 0.f;
    for (int s = threadIdx.x; s < ns; s += 256) {
        const float lq = lg[s] - lse;
        const float qv = expf(lq);
        if (probs) probs[(long)t * ns + s] = qv;
        if (mass) {
            const float p = tm > 0.f ? mass[(long)t * ns + s] / tm : 1.f / (float)ns;
            ce -= p * lq;
            dlogits[(long)t * ns + s] = (qv - p) * inv_tiles;
        }
    }
    if (mass) {
        ce = block_reduce(ce, sh, false);
        if (threadIdx.x == 0) tile_loss[t] = ce * inv_tiles;
    }
}

// dWh[c][s] = sum_t feat[t][c] dl[t][s]; dbh[s] = sum_t dl[t][s]
// loss = sum of the tiles' terms, in a fixed order (strided per thread,
// then the block's fixed reduction tree)
__global__ __launch_bounds__(256) void loss_reduce(const float *__restrict__ tile_loss, int ntiles,
                                                   float *__restrict__ loss)
{
    __shared__ float sh[4];
    float v = 0.f;
    for (int t = threadIdx.x; t < ntiles; t += 256) v += tile_loss[t];
    v = block_reduce(v, sh, false);
    if (threadIdx.x == 0) loss[0] = v;
}

__global__ __launch_bounds__(256) void head_bwd_w(const float *__restrict__ feat, const float *__restrict__ dl,
                                                  int ntiles, int ns, float *__restrict__ dWh, float *__restrict__ dbh)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= ns) return;
    float acc[C], ab = 0.f;
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    for (int t = 0; t < ntiles; t++) {
        const float d = dl[(long)t * ns + s];
        ab += d;
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] = __builtin_fmaf(feat[(long)t * C + c], d, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < C; c++) dWh[(long)c * ns + s] += acc[c];
    dbh[s] += ab;
}

// dfeat[t][c] = sum_s dl[t][s] Wh[c][s]  (one block per tile)
__global__ __launch_bounds__(256) void head_bwd_x(const float *__restrict__ dl, const float *__restrict__ Wh, int ns,
                                                  float *__restrict__ dfeat)
{
    __shared__ float sh[4];
    const int t = blockIdx.x;
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    for (int s = threadIdx.x; s < ns; s += 256) {
        const float d = dl[(long)t * ns + s];
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] = __builtin_fmaf(d, Wh[(long)c * ns + s], acc[c]);
    }
    for (int c = 0; c < C; c++) {
        const float r = block_reduce(acc[c], sh, false);
        if (threadIdx.x == 0) dfeat[(long)t * C + c] = r;
    }
}

__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, const float *__restrict__ gr,
                                                   float *__restrict__ m1, float *__restrict__ m2, long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float b = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// BF16 copies of the conv weights for the MFMA kernel: forward layout
// [tap][cout][cin] and input-gradient layout [tap][cin][cout] with the taps
// mirrored (tap -> 26 - tap).
__global__ __launch_bounds__(256) void pack_weights(const float *__restrict__ w, __bf16 *__restrict__ wf,
                                                    __bf16 *__restrict__ wb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;  // over [tap][cin][cout]
    if (i >= NTAP * C * C) return;
    const int t = i / (C * C), ci = (i / C) % C, co = i % C;
    const __bf16 v = (__bf16)w[i];
    wf[((long)t * C + co) * C + ci] = v;
    wb[((long)(NTAP - 1 - t) * C + ci) * C + co] = v;
}

// volume (float, unpadded) -> normalised FP32 input with zero halo
__global__ __launch_bounds__(256) void prep_input(const float *__restrict__ vol, float *__restrict__ in, ActGeom g,
                                                  float shift, float scale)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((long)g.nx * g.ny));
    in[g.vox(x, y, z)] = (vol[i] - shift) * scale;
}

// Top-K offsets per tile by predicted probability (bitonic sort of up to
// 2048 (probability, index) pairs in LDS), then the fallback decision: the
// tile keeps its K offsets (sorted by search index, i.e. in reference order)
// when their predicted mass reaches `threshold` and the tile is not drawn
// for the audit; otherwise it visits the full search.
__global__ __launch_bounds__(256) void topk_kernel(const float *__restrict__ probs, int ns, int K, float threshold,
                                                   float audit_frac, uint64_t seed, int *__restrict__ lists,
                                                   int *__restrict__ counts, float *__restrict__ stats)
{
    __shared__ float key[2048];
    __shared__ int idx[2048];
    __shared__ float sh[4];
    const int t = blockIdx.x;
    for (int i = threadIdx.x; i < 2048; i += 256) {
        // probabilities lie in [0, 1]; a NaN from a diverged network ranks
        // as 0 and the padding below every real key
        const float pv = i < ns ? probs[(long)t * ns + i] : -1.f;
        key[i] = pv == pv ? pv : 0.f;
        idx[i] = i;
    }
    __syncthreads();
    for (int size = 2; size <= 2048; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < 1024; i += 256) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool desc = (lo & size) == 0;  // sort descending overall
                const bool swap = desc ? key[lo] < key[hi] : key[lo] > key[hi];
                if (swap) {
                    const float tk = key[lo];
                    key[lo] = key[hi];
                    key[hi] = tk;
                    const int ti = idx[lo];
                    idx[lo] = idx[hi];
                    idx[hi] = ti;
                }
            }
            __syncthreads();
        }
    float cap = 0.f;
    for (int i = threadIdx.x; i < K; i += 256) cap += key[i];
    cap = block_reduce(cap, sh, false);
    // audit draw: SplitMix64 of (seed, tile)
    uint64_t zz = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
    zz ^= zz >> 31;
    const bool audit = (double)(zz >> 11) * (1.0 / 9007199254740992.0) < (double)audit_frac;
    const bool keep = cap >= threshold && !audit && K < ns;
    if (threadIdx.x == 0) {
        counts[t] = keep ? K : ns;
        atomicAdd(&stats[0], keep ? 0.f : 1.f);
        atomicAdd(&stats[1], audit ? 1.f : 0.f);
        atomicAdd(&stats[2], cap);
        atomicAdd(&stats[3], (float)(keep ? K : ns));
    }
    __syncthreads();
    if (keep) {
        // the K kept indices in ascending order (rank by counting)
        for (int i = threadIdx.x; i < K; i += 256) {
            int rank = 0;
            const int me = idx[i];
            for (int j = 0; j < K; j++) rank += idx[j] < me;
            lists[(long)t * K + rank] = me;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct ProposalNet::Impl {
    int ns = 0;
    int device = 0;
    hipStream_t st = nullptr;
    // parameters, FP32 master copies, in one buffer:
    // w1[27][16] b1[16] | w2[27][16][16] b2[16] | w3 ... | w4 ... | Wh[16][ns] bh[ns]
    long n_params = 0;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    long off_w1 = 0, off_b1 = 0, off_w[3] = {0, 0, 0}, off_b[3] = {0, 0, 0}, off_wh = 0, off_bh = 0;
    __bf16 *wf[3] = {nullptr, nullptr, nullptr}, *wb[3] = {nullptr, nullptr, nullptr};
    // activations: act[0..1] always (inference ping-pongs between them);
    // act[2..3] and the two gradient planes only for training
    long cap_vox = 0, cap_train = 0;
    float *in = nullptr;
    __bf16 *act[4] = {nullptr, nullptr, nullptr, nullptr};
    __bf16 *gbuf[2] = {nullptr, nullptr};
    long cap_tiles = 0;
    float *feat = nullptr, *dfeat = nullptr, *dl = nullptr, *probs = nullptr, *loss = nullptr;
    float *tile_loss = nullptr;  // per tile, summed by loss_reduce
    double *moments = nullptr;
    float *wpart = nullptr;  // weight-gradient partials, kWgradBlocks x kWgradStride
    int steps = 0;
};

namespace {

hipError_t ensure_acts(ProposalNet::Impl &m, const int shape[3], bool training)
{
    // padded voxels + a guard: a wavefront's 16-voxel rows run up to 15
    // voxels past the last row of the last slice
    const long pv = (long)(shape[0] + 2) * (shape[1] + 2) * (shape[2] + 2) + 64;
    const long tiles = (long)((shape[0] + PTX - 1) / PTX) * ((shape[1] + PTY - 1) / PTY) * ((shape[2] + PTZ - 1) / PTZ);
    const size_t ab = (size_t)pv * C * sizeof(__bf16);
    hipError_t e;
    if (pv > m.cap_vox) {
        (void)hipFree(m.in);
        (void)hipFree(m.act[0]);
        (void)hipFree(m.act[1]);
        m.in = nullptr;
        m.act[0] = m.act[1] = nullptr;
        m.cap_vox = 0;
        if ((e = hipMalloc(&m.in, pv * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc(&m.act[0], ab)) != hipSuccess) return e;
        if ((e = hipMalloc(&m.act[1], ab)) != hipSuccess) return e;
        m.cap_vox = pv;
    }
    if (training && pv > m.cap_train) {
        for (int i = 2; i < 4; i++) {
            (void)hipFree(m.act[i]);
            m.act[i] = nullptr;
        }
        for (auto &a : m.gbuf) {
            (void)hipFree(a);
            a = nullptr;
        }
        m.cap_train = 0;
        for (int i = 2; i < 4; i++)
            if ((e = hipMalloc(&m.act[i], ab)) != hipSuccess) return e;
        for (auto &a : m.gbuf)
            if ((e = hipMalloc(&a, ab)) != hipSuccess) return e;
        m.cap_train = pv;
    }
    // the zero halos of this shape: clear the used extent
    if ((e = hipMemsetAsync(m.in, 0, pv * sizeof(float), m.st)) != hipSuccess) return e;
    for (int i = 0; i < (training ? 4 : 2); i++)
        if ((e = hipMemsetAsync(m.act[i], 0, ab, m.st)) != hipSuccess) return e;
    if (training)
        for (auto &a : m.gbuf)
            if ((e = hipMemsetAsync(a, 0, ab, m.st)) != hipSuccess) return e;
    if (tiles > m.cap_tiles) {
        (void)hipFree(m.feat);
        (void)hipFree(m.dfeat);
        (void)hipFree(m.dl);
        (void)hipFree(m.probs);
        (void)hipFree(m.tile_loss);
        m.feat = m.dfeat = m.dl = m.probs = m.tile_loss = nullptr;
        m.cap_tiles = 0;
        if ((e = hipMalloc(&m.feat, tiles * C * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc(&m.dfeat, tiles * C * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc(&m.dl, tiles * (long)m.ns * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc(&m.probs, tiles * (long)m.ns * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc(&m.tile_loss, tiles * sizeof(float))) != hipSuccess) return e;
        m.cap_tiles = tiles;
    }
    return hipSuccess;
}

}  // namespace

ProposalNet::ProposalNet() : d(new Impl()) {}

ProposalNet::~ProposalNet()
{
    (void)hipSetDevice(d->device);
    if (d->st) (void)hipStreamSynchronize(d->st);
    for (float *p : {d->params, d->grads, d->m1, d->m2, d->in, d->feat, d->dfeat, d->dl, d->probs, d->loss,
                     d->tile_loss})
        (void)hipFree(p);
    for (int l = 0; l < 3; l++) {
        (void)hipFree(d->wf[l]);
        (void)hipFree(d->wb[l]);
    }
    for (auto &a : d->act) (void)hipFree(a);
    for (auto &a : d->gbuf) (void)hipFree(a);
    (void)hipFree(d->moments);
    (void)hipFree(d->wpart);
    if (d->st) (void)hipStreamDestroy(d->st);
    delete d;
}

long ProposalNet::num_params() const { return d->n_params; }
int ProposalNet::num_offsets() const { return d->ns; }
const float *ProposalNet::probs() const { return d->probs; }

hipError_t ProposalNet::init(int ns, int device, uint64_t seed)
{
    if (ns < 1 || ns > kMaxProposalOffsets) return hipErrorInvalidValue;
    d->ns = ns;
    d->device = device;
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return e;
    if ((e = hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking)) != hipSuccess) return e;
    long o = 0;
    d->off_w1 = o;
    o += NTAP * C;
    d->off_b1 = o;
    o += C;
    for (int l = 0; l < 3; l++) {
        d->off_w[l] = o;
        o += (long)NTAP * C * C;
        d->off_b[l] = o;
        o += C;
    }
    d->off_wh = o;
    o += (long)C * ns;
    d->off_bh = o;
    o += ns;
    d->n_params = o;
    const size_t bytes = o * sizeof(float);
    if ((e = hipMalloc(&d->params, bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d->grads, bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d->m1, bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d->m2, bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d->loss, sizeof(float))) != hipSuccess) return e;
    if ((e = hipMalloc(&d->moments, kMomentDoubles * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&d->wpart, (size_t)kWgradBlocks * kWgradStride * sizeof(float))) != hipSuccess) return e;
    for (int l = 0; l < 3; l++) {
        if ((e = hipMalloc(&d->wf[l], NTAP * C * C * sizeof(__bf16))) != hipSuccess) return e;
        if ((e = hipMalloc(&d->wb[l], NTAP * C * C * sizeof(__bf16))) != hipSuccess) return e;
    }
    // He-uniform initialisation from a SplitMix64 stream; zero biases; a
    // small head so the first predictions are near uniform
    std::vector<float> h(o, 0.f);
    uint64_t s = seed;
    auto rnd = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const double a1 = std::sqrt(6.0 / NTAP), a2 = std::sqrt(6.0 / (NTAP * C)), ah = std::sqrt(6.0 / C) * 0.1;
    for (int i = 0; i < NTAP * C; i++) h[d->off_w1 + i] = (float)(a1 * rnd());
    for (int l = 0; l < 3; l++)
        for (long i = 0; i < (long)NTAP * C * C; i++) h[d->off_w[l] + i] = (float)(a2 * rnd());
    for (long i = 0; i < (long)C * ns; i++) h[d->off_wh + i] = (float)(ah * rnd());
    if ((e = hipMemcpy(d->params, h.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemset(d->m1, 0, bytes)) != hipSuccess) return e;
    if ((e = hipMemset(d->m2, 0, bytes)) != hipSuccess) return e;
    return hipDeviceSynchronize();
}

hipError_t ProposalNet::set_params(const float *host, long n)
{
    if (n != d->n_params) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipSetDevice(d->device)) != hipSuccess) return e;
    if ((e = hipMemcpy(d->params, host, n * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemset(d->m1, 0, n * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMemset(d->m2, 0, n * sizeof(float))) != hipSuccess) return e;
    d->steps = 0;
    return hipDeviceSynchronize();
}

hipError_t ProposalNet::get_params(float *host, long n) const
{
    if (n != d->n_params) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipSetDevice(d->device)) != hipSuccess) return e;
    return hipMemcpy(host, d->params, n * sizeof(float), hipMemcpyDeviceToHost);
}

hipError_t ProposalNet::get_grads(float *host, long n) const
{
    if (n != d->n_params) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipSetDevice(d->device)) != hipSuccess) return e;
    return hipMemcpy(host, d->grads, n * sizeof(float), hipMemcpyDeviceToHost);
}

hipError_t ProposalNet::run(const float *vol, const int shape[3], const float *mass, bool backward, float *loss_host)
{
    Impl &m = *d;
    hipError_t e;
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1) return hipErrorInvalidValue;
    if ((e = hipSetDevice(m.device)) != hipSuccess) return e;
    const bool training = mass && backward;
    if ((e = ensure_acts(m, shape, training)) != hipSuccess) return e;
    ActGeom g;
    g.nx = shape[0];
    g.ny = shape[1];
    g.nz = shape[2];
    g.sx = g.nx + 2;
    g.sy = (long)(g.nx + 2) * (g.ny + 2);
    const long nv = (long)g.nx * g.ny * g.nz;
    const unsigned vb = (unsigned)((nv + 255) / 256);
    const int ntx = (g.nx + PTX - 1) / PTX, nty = (g.ny + PTY - 1) / PTY, ntz = (g.nz + PTZ - 1) / PTZ;
    const int ntiles = ntx * nty * ntz;
    const long nwave = (long)((g.nx + 15) / 16) * g.ny * g.nz;
    const unsigned cb = (unsigned)((nwave + 3) / 4);

    // input normalisation: (v - mean) / std of this volume (FP64 moments)
    if ((e = launch_moments(vol, nv, m.moments, m.st)) != hipSuccess) return e;
    double mom[2];
    if ((e = hipMemcpyAsync(mom, m.moments, sizeof mom, hipMemcpyDeviceToHost, m.st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(m.st)) != hipSuccess) return e;
    const double mean = mom[0] / (double)nv;
    const double var = std::max(mom[1] / (double)nv - mean * mean, 0.0);
    const float shift = (float)mean, scale = var > 0 ? (float)(1.0 / std::sqrt(var)) : 1.f;

    const float *P = m.params;
    for (int l = 0; l < 3; l++)
        hipLaunchKernelGGL(pack_weights, dim3((NTAP * C * C + 255) / 256), dim3(256), 0, m.st, P + m.off_w[l],
                           m.wf[l], m.wb[l]);
    hipLaunchKernelGGL(prep_input, dim3(vb), dim3(256), 0, m.st, vol, m.in, g, shift, scale);
    hipLaunchKernelGGL(conv1_fwd, dim3(vb), dim3(256), 0, m.st, m.in, P + m.off_w1, P + m.off_b1, m.act[0], g);
    // training keeps every layer's activation for the backward pass;
    // inference alternates between two planes
    auto A = [&](int l) { return training ? m.act[l] : m.act[l & 1]; };
    for (int l = 0; l < 3; l++)
        hipLaunchKernelGGL(conv16_mfma, dim3(cb), dim3(256), 0, m.st, A(l), m.wf[l], P + m.off_b[l],
                           (const __bf16 *)nullptr, A(l + 1), g, 0);
    hipLaunchKernelGGL(pool_fwd, dim3(ntiles), dim3(256), 0, m.st, A(3), m.feat, g, ntx, nty);
    hipLaunchKernelGGL(head_fwd, dim3(ntiles), dim3(256), m.ns * sizeof(float), m.st, m.feat, P + m.off_wh,
                       P + m.off_bh, m.ns, mass, m.probs, mass ? m.dl : nullptr, m.tile_loss, 1.f / (float)ntiles);
    if (mass) hipLaunchKernelGGL(loss_reduce, dim3(1), dim3(256), 0, m.st, m.tile_loss, ntiles, m.loss);
    else if ((e = hipMemsetAsync(m.loss, 0, sizeof(float), m.st)) != hipSuccess) return e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (training) {
        float *G = m.grads;
        if ((e = hipMemsetAsync(G, 0, m.n_params * sizeof(float), m.st)) != hipSuccess) return e;
        hipLaunchKernelGGL(head_bwd_w, dim3((m.ns + 255) / 256), dim3(256), 0, m.st, m.feat, m.dl, ntiles, m.ns,
                           G + m.off_wh, G + m.off_bh);
        hipLaunchKernelGGL(head_bwd_x, dim3(ntiles), dim3(256), 0, m.st, m.dl, P + m.off_wh, m.ns, m.dfeat);
        // gradient at the pre-activation of the last convolution
        hipLaunchKernelGGL(pool_bwd, dim3(vb), dim3(256), 0, m.st, m.dfeat, m.act[3], m.gbuf[0], g, ntx, nty);
        // a fixed number of blocks over the rows, each writing its partials
        const long nrows = g.ny * (long)g.nz;
        const int rpb = (int)((nrows + kWgradBlocks - 1) / kWgradBlocks);
        const unsigned wgb = (unsigned)((nrows + rpb - 1) / rpb);
        auto reduce = [&](int cin, float *dW, float *dB) {
            const int nw = NTAP * cin * C;
            hipLaunchKernelGGL(wgrad_reduce, dim3((unsigned)((nw + C + 255) / 256)), dim3(256), 0, m.st, m.wpart,
                               (int)wgb, nw, dW, dB);
        };
        int cur = 0;
        for (int l = 2; l >= 0; l--) {
            // convolution l+2: input act[l], gradient at its output gbuf[cur]
            hipLaunchKernelGGL(conv_wgrad<C>, dim3(wgb), dim3(256), 0, m.st, (const void *)m.act[l], m.gbuf[cur],
                               m.wpart, g, rpb);
            reduce(C, G + m.off_w[l], G + m.off_b[l]);
            hipLaunchKernelGGL(conv16_mfma, dim3(cb), dim3(256), 0, m.st, m.gbuf[cur], m.wb[l],
                               (const float *)nullptr, m.act[l], m.gbuf[cur ^ 1], g, 1);
            cur ^= 1;
        }
        hipLaunchKernelGGL(conv_wgrad<1>, dim3(wgb), dim3(256), 0, m.st, (const void *)m.in, m.gbuf[cur], m.wpart,
                           g, rpb);
        reduce(1, G + m.off_w1, G + m.off_b1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (loss_host) {
        if ((e = hipMemcpyAsync(loss_host, m.loss, sizeof(float), hipMemcpyDeviceToHost, m.st)) != hipSuccess)
            return e;
    }
    return hipStreamSynchronize(m.st);
}

hipError_t ProposalNet::adam(float lr)
{
    Impl &m = *d;
    hipError_t e;
    if ((e = hipSetDevice(m.device)) != hipSuccess) return e;
    m.steps++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)m.steps)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)m.steps)));
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((m.n_params + 255) / 256)), dim3(256), 0, m.st, m.params,
                       m.grads, m.m1, m.m2, m.n_params, lr, (float)b1, (float)b2, 1e-8f, c1, c2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipStreamSynchronize(m.st);
}

hipError_t launch_topk(const float *probs, int ntiles, int ns, int K, float threshold, float audit_frac,
                       uint64_t seed, int *lists, int *counts, float *stats, hipStream_t st)
{
    if (ns < 1 || ns > kMaxProposalOffsets || K < 1 || K > ns) return hipErrorInvalidValue;
    hipLaunchKernelGGL(topk_kernel, dim3(ntiles), dim3(256), 0, st, probs, ns, K, threshold, audit_frac, seed, lists,
                       counts, stats);
    return hipGetLastError();
}

}  // namespace inlm
