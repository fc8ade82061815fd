// Matrix-core (MFMA) form of the iterNLM filter for gfx950.
//
// For a search offset s the patch distance is a small 3D convolution of the
// non-negative field D_s(y) = (P(y+s) - P(y))^2 with the patch weights:
//   d_s(x) = sum_rows sum_dx W(dx,dy,dz) D_s(x + (dx,dy,dz)).
// For one patch row (dz,dy) and a block of 16 outputs along x the row sum is
// a banded Toeplitz product, so a 16x16 block of outputs (16 consecutive x by
// 16 (y,z) rows) takes one v_mfma_f32_16x16x32 per patch row:
//   A[m][k] = D_s[row m + (dy,dz)][x0 - 8 + k]    (16-bit D plane in LDS)
//   B[k][n] = W(x0 - 8 + k - (x0 + n), dy, dz)     (band of the row's weights)
//   C[m][n] += A . B                               (FP32 accumulate)
// D is a sum of non-negative terms (no cancellation), so FP16 D planes keep
// the distance to ~1e-4 relative (the BF16 storage form widens its values
// to FP16 before the difference; the FP16_FP8 form scales FP8 planes per
// tile).  The weighted sums stay FP32 and the final normalisation is FP64.
// Volumes are stored normalised, v_s = (v - shift) * scale, which leaves
// every weight unchanged once the per-slice multiplier is divided by scale^2.
//
// This row form (filter_mfma: every patch radius, the FP8-plane and FP32_FAST
// forms, the pruned per-tile lists): a workgroup (8 wavefronts) owns a
// 32x8x8 output tile and stages D_s over the tile plus the patch halo in LDS
// (double buffered, one barrier per offset); wavefront w computes the 16x16
// block x-half (w & 1), z pair (w >> 1).  The full search at patch radius 2
// in the 16-bit forms runs the block form (filter_mfma_b4, further below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "inlm_kernels.hpp"

namespace inlm {
namespace {

constexpr int NTM = 512;
constexpr int MTX = 32, MTY = 8, MTZ = 8;
// D-plane layout in LDS: rows of XW = 48 16-bit elements (columns [0,4) and
// [44,48) stay zero), 16 rows per slice whatever the patch radius.  With a
// 96-byte row pitch and a 1536-byte slice pitch the 16-lane groups in which
// gfx950 serves a ds_read_b128 (64 banks, 256 bytes per clock) hit 16
// distinct 16-byte bank groups for the A-operand pattern (2 slices x 8 rows
// x 4 column blocks), so the reads run conflict-free; the former 112-byte
// pitch with 12 rows per slice took three times as many LDS cycles per read.
constexpr int XW = 48;
constexpr int RYP = 16;
// FP8 planes (1-byte elements) use 48-byte rows and 24 rows per slice: rows
// m and m + 8 of an operand then fall 8 bank groups apart, and a slot's 16
// rows hit 16 distinct 16-byte bank groups.
constexpr int XW8 = 48;
constexpr int RYP8 = 24;
constexpr float kCut = -3.56648f;

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));
// OCP FP8 e4m3 D-plane element
typedef unsigned char f8_t;
// FP8-plane plans: at most this many MFMAs per offset and band operands
constexpr int kF8MaxM = 8, kF8MaxSig = 8;
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Patch rows of the isotropic ellipsoid of radius R and their weight
// patterns: a row's weights as a function of dx depend only on (dy^2, dz^2),
// so rows sharing that key share one B operand.  When the weights depend on
// dx^2+dy^2+dz^2 alone (no z anisotropy), patterns of equal dy^2+dz^2 carry
// the same band too and share a B operand class.
//
// One MFMA serves one patch row, or two rows of one tap each (half-width 0):
// a row of one tap needs only the 16 columns of its outputs, so two of them
// fill the 32-deep K of one MFMA.  The MFMA operations, their B operands
// (single-row bands and paired taps, one per pattern or per class) and the
// distinct row distances of the pairs are tabulated here, for the per-pattern
// (op*) and the per-class (iop*) operand sets.
struct MRow {
    int dz, dy, j, pat;
};
constexpr int kMaxRowOps = 32;
template <int R>
struct MRows {
    MRow r[4 * R * R + 4 * R + 1];
    int n, npat, ncls;
    int pdy[kMfmaMaxPatterns], pdz[kMfmaMaxPatterns], pj[kMfmaMaxPatterns];
    int cls[kMfmaMaxPatterns], crep[kMfmaMaxPatterns], ckey[kMfmaMaxPatterns];
    // MFMA operations: row a, row b (-1: single), index of the pair's row
    // distance (-1: single)
    int nops;
    int opa[kMaxRowOps], opb[kMaxRowOps], opd[kMaxRowOps];
    int npd;
    int pdelta[4];  // LDS element distance from row a to row b of a pair
    // B operands: kind (0 band of a row, 1 paired single taps), pattern;
    // operand of each operation; per pattern (o*) and per class (i*)
    int nob, obkind[kMfmaMaxPatterns], obpat[kMfmaMaxPatterns], obop[kMaxRowOps];
    int nib, ibkind[kMfmaMaxPatterns], ibpat[kMfmaMaxPatterns], ibop[kMaxRowOps];
    static constexpr int lds_off(const MRow &w) { return (w.dz * RYP + w.dy) * XW; }
    static constexpr bool inside(int x, int y, int z)
    {
        const float rr = (float)R;
        const float fx = (float)x / rr, fy = (float)y / rr, fz = (float)z / rr;
        float e = fx * fx;
        e = e + fy * fy;
        e = e + fz * fz;
        return e <= 1.f;
    }
    constexpr MRows()
        : r(), n(0), npat(0), ncls(0), pdy(), pdz(), pj(), cls(), crep(), ckey(), nops(0), opa(), opb(), opd(),
          npd(0), pdelta(), nob(0), obkind(), obpat(), obop(), nib(0), ibkind(), ibpat(), ibop()
    {
        for (int z = -R; z <= R; ++z)
            for (int y = -R; y <= R; ++y) {
                int j = -1;
                for (int x = 0; x <= R; ++x)
                    if (inside(x, y, z)) j = x;
                if (j < 0) continue;
                int p = -1;
                for (int q = 0; q < npat; q++)
                    if (pdy[q] == (y < 0 ? -y : y) && pdz[q] == (z < 0 ? -z : z)) p = q;
                if (p < 0) {
                    p = npat++;
                    pdy[p] = y < 0 ? -y : y;
                    pdz[p] = z < 0 ? -z : z;
                    pj[p] = j;
                }
                r[n++] = MRow{z, y, j, p};
            }
        for (int p = 0; p < npat; p++) {
            const int key = pdy[p] * pdy[p] + pdz[p] * pdz[p];
            int c = -1;
            for (int q = 0; q < ncls; q++)
                if (ckey[q] == key) c = q;
            if (c < 0) {
                c = ncls++;
                ckey[c] = key;
                crep[c] = p;
            }
            cls[p] = c;
        }
        // operations in row order; a one-tap row takes the next unused
        // one-tap row of its pattern as its partner
        bool used[4 * R * R + 4 * R + 1] = {};
        for (int i = 0; i < n; i++) {
            if (used[i]) continue;
            used[i] = true;
            int b = -1;
            if (r[i].j == 0)
                for (int k = i + 1; k < n && b < 0; k++)
                    if (!used[k] && r[k].j == 0 && r[k].pat == r[i].pat) b = k;
            int d = -1;
            if (b >= 0) {
                used[b] = true;
                const int delta = lds_off(r[b]) - lds_off(r[i]);
                for (int k = 0; k < npd; k++)
                    if (pdelta[k] == delta) d = k;
                if (d < 0) {
                    d = npd++;
                    pdelta[d] = delta;
                }
            }
            opa[nops] = i;
            opb[nops] = b;
            opd[nops] = d;
            nops++;
        }
        for (int o = 0; o < nops; o++) {
            const int kind = opb[o] >= 0 ? 1 : 0, p = r[opa[o]].pat;
            int x = -1, y = -1;
            for (int k = 0; k < nob; k++)
                if (obkind[k] == kind && obpat[k] == p) x = k;
            if (x < 0) {
                x = nob++;
                obkind[x] = kind;
                obpat[x] = p;
            }
            for (int k = 0; k < nib; k++)
                if (ibkind[k] == kind && cls[ibpat[k]] == cls[p]) y = k;
            if (y < 0) {
                y = nib++;
                ibkind[y] = kind;
                ibpat[y] = crep[cls[p]];
            }
            obop[o] = x;
            ibop[o] = y;
        }
    }
};
template <int R>
struct MRowsOf {
    static constexpr MRows<R> value{};
    static_assert(value.npat <= kMfmaMaxPatterns && value.nob <= kMfmaMaxPatterns, "too many row patterns");
    static_assert(value.nops <= kMaxRowOps && value.npd <= 4, "too many MFMA operations");
};

// FP8 planes: one v_mfma_f32_16x16x128_f8f6f4 sums four 32-byte K slots,
// slot s taken by lane group s (lanes 16s .. 16s+15) of both operands.  A
// slot holds one item of the 16-bit plan: a patch row, read over input
// columns x0-4 .. x0+27, or a pair of one-tap rows, x0 .. x0+15 of each.
// Items are ordered by how often their band key occurs and laid out two
// slots at a time over pairs of MFMAs, so MFMAs whose slots carry the same
// keys share one band operand (at r = 2: three MFMAs, two band operands).
template <int R, bool ISO>
struct F8Plan {
    static constexpr int kMaxItems = 4 * R * R + 4 * R + 1;
    static constexpr int kMaxM = (kMaxItems + 3) / 4;
    int ni;
    int ia[kMaxItems], ib[kMaxItems], key[kMaxItems];
    int nm;
    int slot[kMaxM][4];  // item of each MFMA slot, -1 for an empty slot
    bool haspair[kMaxM];
    int nsig;
    int sig[kMaxM];   // band operand of each MFMA
    int sigm[kMaxM];  // the MFMA whose slots define each band operand
    constexpr F8Plan() : ni(0), ia(), ib(), key(), nm(0), slot(), haspair(), nsig(0), sig(), sigm()
    {
        constexpr const MRows<R> &rows = MRowsOf<R>::value;
        for (int o = 0; o < rows.nops; o++) {
            const int p = rows.r[rows.opa[o]].pat;
            ia[ni] = rows.opa[o];
            ib[ni] = rows.opb[o];
            key[ni] = (rows.opb[o] >= 0 ? 1000 : 0) + (ISO ? rows.cls[p] : p);
            ni++;
        }
        int order[kMaxItems] = {}, cnt[kMaxItems] = {};
        for (int i = 0; i < ni; i++) {
            order[i] = i;
            for (int k = 0; k < ni; k++) cnt[i] += key[k] == key[i] ? 1 : 0;
        }
        // insertion sort by (descending count, key, dy^2, dy, dz of the
        // first row): consecutive items share an MFMA in lane groups 0/1 or
        // 2/3, which LDS serves in the same cycles; rows with equal dy whose
        // dz differ by 2 lie a multiple of 16 bank groups apart (48-byte
        // rows, 24 rows per slice), so such pairs read without conflicts
        auto rank = [&](int it) {
            const MRow &w = rows.r[ia[it]];
            return ((w.dy * w.dy) * 16 + (w.dy + 8)) * 16 + (w.dz + 8);
        };
        for (int i = 1; i < ni; i++) {
            const int v = order[i];
            int j = i - 1;
            while (j >= 0) {
                const int u = order[j];
                const bool before = cnt[v] != cnt[u]   ? cnt[v] > cnt[u]
                                    : key[v] != key[u] ? key[v] < key[u]
                                    : rank(v) != rank(u) ? rank(v) < rank(u)
                                                         : v < u;
                if (!before) break;
                order[j + 1] = u;
                j--;
            }
            order[j + 1] = v;
        }
        nm = (ni + 3) / 4;
        for (int o = 0; o < kMaxM; o++)
            for (int q = 0; q < 4; q++) slot[o][q] = -1;
        int pos = 0;
        for (int b = 0; 2 * b + 1 < nm; b++)
            for (int h = 0; h < 2; h++)
                for (int m = 0; m < 2; m++)
                    for (int q = 0; q < 2; q++)
                        if (pos < ni) slot[2 * b + m][2 * h + q] = order[pos++];
        if (nm % 2 != 0)
            for (int q = 0; q < 4; q++)
                if (pos < ni) slot[nm - 1][q] = order[pos++];
        for (int o = 0; o < nm; o++) {
            haspair[o] = false;
            for (int q = 0; q < 4; q++) haspair[o] = haspair[o] || (slot[o][q] >= 0 && ib[slot[o][q]] >= 0);
            int found = -1;
            for (int g = 0; g < nsig && found < 0; g++) {
                bool same = true;
                for (int q = 0; q < 4; q++) {
                    const int u = slot[o][q], v = slot[sigm[g]][q];
                    same = same && (u < 0 ? -1 : key[u]) == (v < 0 ? -1 : key[v]);
                }
                if (same) found = g;
            }
            if (found < 0) {
                found = nsig++;
                sigm[found] = o;
            }
            sig[o] = found;
        }
    }
};
template <int R, bool ISO>
struct F8PlanOf {
    static constexpr F8Plan<R, ISO> value{};
    static_assert(value.ni <= F8Plan<R, ISO>::kMaxItems && value.nm <= F8Plan<R, ISO>::kMaxM, "FP8 plan too large");
    static_assert(value.nm <= kF8MaxM && value.nsig <= kF8MaxSig, "FP8 plan exceeds the argument tables");
};

template <int I, int N, class F>
__device__ __forceinline__ void mstatic_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        mstatic_for<I + 1, N>(f);
    }
}

// 16-bit D plane element type and the matching MFMA.
template <class TD>
struct DPlane;
template <>
struct DPlane<_Float16> {
    static constexpr int XWE = XW, RYPE = RYP;  // LDS row pitch and rows per slice, in elements
    typedef half8 vec8;
    static __device__ __forceinline__ _Float16 from(float v) { return (_Float16)v; }
    static __device__ __forceinline__ floatx4 mfma(vec8 a, vec8 b, floatx4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
// FP8 e4m3 plane: 32 bytes per lane and operand, v_mfma_f32_16x16x128_f8f6f4
// (format 0 = FP8 e4m3 for both operands, no block scales)
template <>
struct DPlane<f8_t> {
    static constexpr int XWE = XW8, RYPE = RYP8;
    typedef v8i vec8;
    static __device__ __forceinline__ floatx4 mfma(vec8 a, vec8 b, floatx4 c)
    {
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
    }
};

template <class TS>
struct MfmaArgs {
    const TS *R;
    const TS *P;
    const TS *R_end;   // one past the allocations of R and P (buffer ranges)
    const TS *P_end;
    const float *raw;  // unpadded original
    float *out;
    const float *sigma;
    const SearchEntry *search;       // the search table; lists index it
    const SearchEntry *search_full;  // the same offsets in full-search order
    int n_search;
    float beta;
    float mscale;      // 1 / scale^2: distances in stored units -> data units
    double shift;      // data = shift + stored * inv_scale
    double inv_scale;
    FilterGeom g;
    int ntx, nty, ntz;
    float band[kMfmaMaxPatterns][8];  // band[p][dx + R]: weight of pattern p at dx
    // optional per-tile offset lists: tile t visits lists[t*K .. t*K+counts[t])
    // (indices into search, ascending) when counts[t] < n_search, else all
    const int *lists;
    const int *counts;
    int K;
    int min_lin;  // smallest linear search offset (<= 0): the tile's buffer ranges start there
    float one;    // 1.0 (a runtime value the compiler cannot fold)
    // FP8 planes: band operand g, lane group q: a 48-byte Toeplitz template
    // of FP8 e4m3 codes (the weights times wscale) that lane n reads from
    // byte 16 - n; and the LDS byte offsets, relative to the lane's own row,
    // of the two 16-byte halves of slot q of MFMA o
    unsigned char f8tmpl[kF8MaxSig][4][48];
    int f8offa[kF8MaxM][4], f8offb[kF8MaxM][4];
    float wscale;
    float wmin;  // the smallest patch weight (scaled 16-bit planes)
    // 16-bit storage: R and P are each followed, `shifted` elements further
    // on, by a copy shifted by one element (copy[j] = volume[j + 1])
    long shifted;
};

// Staging geometry.  A region row (slice Z0-R+rz, row Y0-R+ry) is staged
// as 10 units of 4 columns, LDS columns [4, 44).  Every load is one 16-byte
// (or, for the last 16-bit units, 8-byte) buffer load at a 4-byte-aligned
// address: on gfx950 such a load costs the texture address unit the same
// cycles whether it carries 8 or 16 bytes, while a 2-byte-aligned one costs
// two to four times as many.  16-bit storage therefore stages chunks of 8
// columns (two units) in S8 slots of NTM chunks and any remainder as 4-column
// chunks in S4 slots; FP32 storage stages 4-column chunks (16 bytes).
template <int R, class TS, class TD>
struct MTile {
    static constexpr int RY = MTY + 2 * R, RZ = MTZ + 2 * R;
    static constexpr int UPR = 10;  // 4-column units per region row
    static constexpr int NUNIT = RZ * RY * UPR;
    static constexpr bool WIDE = sizeof(TS) == 2;
    static constexpr int S8A = WIDE ? NUNIT / (2 * NTM) : 0;
    static constexpr int REM = NUNIT - 2 * NTM * S8A;
    // a remainder larger than one slot of 4-column chunks takes another slot
    // of 8-column chunks
    static constexpr bool REM8 = WIDE && REM > NTM;
    static constexpr int S8 = S8A + (REM8 ? 1 : 0);
    static constexpr int N8 = REM8 ? NUNIT / 2 : S8A * NTM;
    static constexpr int N4 = NUNIT - 2 * N8;
    static constexpr int S4 = (N4 + NTM - 1) / NTM;
    static constexpr int CPT = S8 + S4;
    // staging slots past N8 (N4) write to their own 8 (4) elements behind the
    // plane, so every thread stages its chunks with no per-chunk branch
    static constexpr int PLANE = RZ * DPlane<TD>::RYPE * DPlane<TD>::XWE;
    static constexpr int SPARE8 = PLANE;
    static constexpr int SPARE4 = PLANE + 8 * (S8 * NTM - N8);
    static constexpr int LDS = SPARE4 + 4 * (S4 * NTM - N4);  // elements per buffer
    static_assert(RY <= DPlane<TD>::RYPE, "patch radius too large for the padded slice");
    static_assert(UPR % 2 == 0, "8-column chunks must not straddle region rows");
};

__device__ __forceinline__ int mxcd_swizzle(int wg, int nwg)
{
    if ((nwg & 7) != 0) return wg;
    return (wg & 7) * (nwg >> 3) + (wg >> 3);
}

// Weight of one distance, from the MFMA accumulator d and a per-lane scale.
// WFN 0 is the reference's Pade [3/3] with its cutoff (iternlm_cpu.h:125-126)
// as a degree-7 minimax polynomial in t = a / h + 1 with a = d * m and
// h = -cutoff / 2, which maps the kept range (cutoff, 0] onto (-1, 1]
// (largest deviation from the rational form 3.7e-6, weights lie in
// [0.028, 1]); it needs no reciprocal, a quarter-rate instruction.  The scale
// is m / h, so t is one fused multiply-add, and the cutoff is t > -1 (a NaN
// or overflowed distance gives weight zero).  WFN 1 is exp(a) = 2^(a log2 e)
// with the scale m log2 e.
// vector instructions per output and offset in the weight step (distance
// scale, polynomial, cutoff test and select, the three sums)
constexpr int kWeightValu = 13;
// per staged unit of four values (difference, square, clamp; packed for FP16;
// and the conversion for FP8 planes)
template <class TS, class TD>
constexpr int kStageValu = std::is_same<TD, f8_t>::value ? 6 : std::is_same<TS, __bf16>::value ? 12 : sizeof(TS) == 2 ? 6 : 12;
// per pair of outputs in the packed FP16 weight (conversion, scale,
// polynomial, step, the sums)
constexpr int kWeightValuPacked = 13;
constexpr float kPadePoly[8] = {0.1679810881614685f,  0.30033349990844727f,  0.2661350965499878f,
                                0.1601015031337738f,  0.06991086900234222f,  0.024921204894781113f,
                                0.008454217575490475f, 0.002158894669264555f};
// 16-bit storage forms: the weights of two outputs per instruction in packed
// FP16.  u = a / (2h) + 1 = d m / 3.56648 + 1, clamped to [0, 1] (u = 0 at
// and beyond the cutoff).  The weight is computed divided by its value at
// the cutoff, c0 = 0.0248 (the sums and the maximum weight all scale alike,
// so the result is unchanged): w / c0 = Q(u) u + s(u), with Q a degree-4
// polynomial with FP16 coefficients and s a step that is 1 for u >= 2^-14
// and 0 at u = 0.  Dividing by c0 turns the step's product into the last
// Horner addend, one instruction fewer per pair.  The coefficients minimise
// the largest deviation of the FP16 evaluation over every FP16 u in
// [2^-13, 1]: on a uniform grid of a (a > -0.999 h) the weight deviates by at
// most 1.7e-3 from the rational form (3.3e-4 rms), most of it from rounding
// u to FP16.  A degree-7 weight (Q of degree 6) was within 1.6e-3 (2.5e-4
// rms) and cost two instructions more per pair, for no measurable change in
// PSNR.
constexpr _Float16 kPadePolyQ5[5] = {(_Float16)5.2734375f, (_Float16)-0.86474609375f, (_Float16)37.53125f,
                                     (_Float16)-40.03125f, (_Float16)37.40625f};
typedef _Float16 half2w __attribute__((ext_vector_type(2)));
typedef unsigned short ushort2w __attribute__((ext_vector_type(2)));
// the weights (divided by c0) of two pairs of outputs, the two polynomial
// chains in lockstep
__device__ __forceinline__ void packed_weights(half2w d0, half2w d1, half2w ms, half2w &w0, half2w &w1)
{
    constexpr int DEG = 5;
    constexpr const _Float16 *Q = kPadePolyQ5;
    const half2w zero = {(_Float16)0.f, (_Float16)0.f}, one = {(_Float16)1.f, (_Float16)1.f};
    half2w u0 = __builtin_elementwise_fma(d0, ms, one), u1 = __builtin_elementwise_fma(d1, ms, one);
    u0 = __builtin_elementwise_min(__builtin_elementwise_max(u0, zero), one);
    u1 = __builtin_elementwise_min(__builtin_elementwise_max(u1, zero), one);
    half2w p0 = {Q[DEG - 1], Q[DEG - 1]}, p1 = p0;
#pragma unroll
    for (int j = DEG - 2; j >= 0; j--) {
        const half2w c = {Q[j], Q[j]};
        p0 = __builtin_elementwise_fma(p0, u0, c);
        p1 = __builtin_elementwise_fma(p1, u1, c);
    }
    const half2w big = {(_Float16)16384.f, (_Float16)16384.f};
    half2w s0 = u0 * big, s1 = u1 * big;
    s0 = __builtin_elementwise_min(__builtin_elementwise_max(s0, zero), one);
    s1 = __builtin_elementwise_min(__builtin_elementwise_max(s1, zero), one);
    w0 = __builtin_elementwise_fma(p0, u0, s0);
    w1 = __builtin_elementwise_fma(p1, u1, s1);
}

template <int WFN>
__host__ __device__ constexpr float weight_scale_factor()
{
    return WFN == 0 ? 2.f / 3.56648f : 1.4426950408889634f;
}
template <int WFN>
__device__ __forceinline__ float fast_weight(float d, float ms)
{
    if constexpr (WFN == 0) {
        const float t = __builtin_fmaf(d, ms, 1.f);
        float p = kPadePoly[7];
#pragma unroll
        for (int j = 6; j >= 0; j--) p = __builtin_fmaf(p, t, kPadePoly[j]);
        return t > -1.f ? p : 0.f;
    } else {
        return __builtin_amdgcn_exp2f(d * ms);
    }
}

// Raw buffer loads: a uniform 128-bit resource (base, byte range) plus a
// 32-bit per-lane byte offset.  Reads past the range return zero, so a load
// can never leave the allocation the resource describes.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer, 32-bit data format
// `bytes` is the distance from base to the end of the allocation; callers
// only form bases inside their allocation (a tile origin plus an offset of
// the padded volume, or the guard in front of it), so it is positive, and
// the unsigned test keeps the clamp on the scalar unit.
__device__ __forceinline__ rsrc_t make_rsrc(const void *base, long bytes)
{
    const unsigned long b = (unsigned long)bytes;
    const int n = (b >> 31) != 0 ? 0x7fffffff : (int)b;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, n, kRsrcWord3);
}
template <class TS>
__device__ __forceinline__ void load4(const TS *p, float (&v)[4])
{
    if constexpr (sizeof(TS) == 4) {
        float4 q;
        __builtin_memcpy(&q, p, 16);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        TS q[4];
        __builtin_memcpy(q, p, 8);
#pragma unroll
        for (int e = 0; e < 4; e++) v[e] = (float)q[e];
    }
}

// Staging of D = (P(y+s) - P(y))^2 into the 16-bit D plane, W columns per
// chunk (W * sizeof(TS) is 8 or 16 bytes: one buffer load).  Generic form:
// FP32 arithmetic, clamped below the FP16 maximum when the plane is FP16 (an
// infinite D would turn the zero band weights of its neighbours into NaN).
template <int W, class TS>
__device__ __forceinline__ void load_chunk(void *c, rsrc_t r, unsigned off, unsigned soff)
{
    // the builtins return GCC-style vectors; copy them out whole (element
    // subscripts on them are not lowered correctly by this compiler)
    static_assert(W * sizeof(TS) == 8 || W * sizeof(TS) == 16, "one 8- or 16-byte load per chunk");
    if constexpr (W * sizeof(TS) == 16) {
        const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0);
        __builtin_memcpy(c, &q, 16);
    } else {
        const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
        __builtin_memcpy(c, &q, 8);
    }
}
// 8-byte-aligned LDS stores of a chunk's W 16-bit D values
template <int W>
__device__ __forceinline__ void store_dplane(void *dst, const void *src)
{
#pragma unroll
    for (int h = 0; h < W / 4; h++) __builtin_memcpy(static_cast<char *>(dst) + 8 * h, static_cast<const char *>(src) + 8 * h, 8);
}

template <int W, class TS, class TD>
struct Stager {
    typedef TS chunk[W];  // stored values, widened to FP32 when staged
    __device__ __forceinline__ void load(chunk &c, const TS *p) const { __builtin_memcpy(c, p, sizeof(chunk)); }
    __device__ __forceinline__ void load(chunk &c, rsrc_t r, unsigned off, unsigned soff) const
    {
        load_chunk<W, TS>(&c, r, off, soff);
    }
    // pn: the chunk at the shifted position; b: the chunk at the tile's own
    // position (from LDS)
    __device__ __forceinline__ void store(const chunk &pn, const chunk &b, TD *dst) const
    {
        TD dv[W];
#pragma unroll
        for (int e = 0; e < W; e++) {
            const float df = (float)pn[e] - (float)b[e];
            float dd = df * df;
            if constexpr (std::is_same<TD, _Float16>::value) dd = __builtin_fminf(dd, 65504.f);
            dv[e] = DPlane<TD>::from(dd);
        }
        store_dplane<W>(dst, dv);
    }
};

// FP16 storage and FP16 plane: packed half arithmetic, two values per
// instruction (difference, square, clamp).
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
template <int W>
struct Stager<W, _Float16, _Float16> {
    typedef half2v chunk[W / 2];
    __device__ __forceinline__ void load(chunk &c, const _Float16 *p) const { __builtin_memcpy(&c[0], p, sizeof(chunk)); }
    __device__ __forceinline__ void load(chunk &c, rsrc_t r, unsigned off, unsigned soff) const
    {
        load_chunk<W, _Float16>(&c[0], r, off, soff);
    }
    __device__ __forceinline__ void store(const chunk &pn, const chunk &b, _Float16 *dst) const
    {
        const half2v big = {(_Float16)65504.f, (_Float16)65504.f};
        half2v out[W / 2];
#pragma unroll
        for (int h = 0; h < W / 2; h++) {
            const half2v d = pn[h] - b[h];
            out[h] = __builtin_elementwise_min(d * d, big);
        }
        store_dplane<W>(dst, out);
    }
    // scaled plane D * f^2 clamped to [0, 1]: one FMA (b = -P(y) f) and one
    // multiply with the instruction's clamp per pair
    half2v f;
    __device__ __forceinline__ void own_scaled(chunk &b) const
    {
#pragma unroll
        for (int h = 0; h < W / 2; h++) b[h] = -(b[h] * f);
    }
    __device__ __forceinline__ void scaled(const chunk &pn, const chunk &b, half2v (&out)[W / 2]) const
    {
        const half2v zero = {(_Float16)0.f, (_Float16)0.f}, one = {(_Float16)1.f, (_Float16)1.f};
#pragma unroll
        for (int h = 0; h < W / 2; h++) {
            const half2v d = __builtin_elementwise_fma(pn[h], f, b[h]);
            out[h] = __builtin_elementwise_min(__builtin_elementwise_max(d * d, zero), one);
        }
    }
    __device__ __forceinline__ void store_scaled(const chunk &pn, const chunk &b, _Float16 *dst) const
    {
        half2v out[W / 2];
        scaled(pn, b, out);
        store_dplane<W>(dst, out);
    }
};

// BF16 storage and FP16 plane: each pair of stored values is widened to
// FP32 (a shift and a mask) and packed to FP16, exactly for the normalised
// data's range; the difference, square and clamp then run as in the FP16
// form.  The tile's own chunk is converted once (own) and kept as FP16.
template <int W>
struct Stager<W, __bf16, _Float16> {
    typedef unsigned chunk[W / 2];  // two BF16 values per word (FP16 pairs after own)
    __device__ __forceinline__ void load(chunk &c, const __bf16 *p) const { __builtin_memcpy(&c[0], p, sizeof(chunk)); }
    __device__ __forceinline__ void load(chunk &c, rsrc_t r, unsigned off, unsigned soff) const
    {
        load_chunk<W, __bf16>(&c[0], r, off, soff);
    }
    static __device__ __forceinline__ half2v widen(unsigned v)
    {
        const float lo = __builtin_bit_cast(float, v << 16), hi = __builtin_bit_cast(float, v & 0xffff0000u);
        return __builtin_bit_cast(half2v, __builtin_amdgcn_cvt_pkrtz(lo, hi));
    }
    __device__ __forceinline__ void store(const chunk &pn, const chunk &b, _Float16 *dst) const
    {
        const half2v big = {(_Float16)65504.f, (_Float16)65504.f};
        half2v out[W / 2];
#pragma unroll
        for (int h = 0; h < W / 2; h++) {
            const half2v d = widen(pn[h]) - __builtin_bit_cast(half2v, b[h]);
            out[h] = __builtin_elementwise_min(d * d, big);
        }
        store_dplane<W>(dst, out);
    }
    __device__ __forceinline__ void own(chunk &b) const
    {
#pragma unroll
        for (int h = 0; h < W / 2; h++) b[h] = __builtin_bit_cast(unsigned, widen(b[h]));
    }
    half2v f;
    __device__ __forceinline__ void own_scaled(chunk &b) const
    {
#pragma unroll
        for (int h = 0; h < W / 2; h++) b[h] = __builtin_bit_cast(unsigned, -(widen(b[h]) * f));
    }
    __device__ __forceinline__ void scaled(const chunk &pn, const chunk &b, half2v (&out)[W / 2]) const
    {
        const half2v zero = {(_Float16)0.f, (_Float16)0.f}, one = {(_Float16)1.f, (_Float16)1.f};
#pragma unroll
        for (int h = 0; h < W / 2; h++) {
            const half2v d = __builtin_elementwise_fma(widen(pn[h]), f, __builtin_bit_cast(half2v, b[h]));
            out[h] = __builtin_elementwise_min(__builtin_elementwise_max(d * d, zero), one);
        }
    }
    __device__ __forceinline__ void store_scaled(const chunk &pn, const chunk &b, _Float16 *dst) const
    {
        half2v out[W / 2];
        scaled(pn, b, out);
        store_dplane<W>(dst, out);
    }
};

// FP16 storage and FP8 plane.  With f = 2^m per tile the plane holds
// D * S, S = 256 f^2: the staging forms (P(y+s) f - P(y) f) with one packed
// FMA (the tile's own chunk is kept as -P f), squares it clamped to [0, 1]
// (the instruction's clamp: larger values become 1), and converts two values
// per instruction times 256 (the conversion divides by its scale operand,
// 1/256): no value exceeds the FP8 maximum of 448, which the conversion does
// not saturate.
typedef short short2v __attribute__((ext_vector_type(2)));
template <int W>
struct Stager<W, _Float16, f8_t> {
    typedef half2v chunk[W / 2];
    half2v f;  // the tile's 2^m
    __device__ __forceinline__ void load(chunk &c, const _Float16 *p) const { __builtin_memcpy(&c[0], p, sizeof(chunk)); }
    __device__ __forceinline__ void load(chunk &c, rsrc_t r, unsigned off, unsigned soff) const
    {
        load_chunk<W, _Float16>(&c[0], r, off, soff);
    }
    // b: the tile's own chunk times -f
    __device__ __forceinline__ void store(const chunk &pn, const chunk &b, f8_t *dst) const
    {
        const half2v zero = {(_Float16)0.f, (_Float16)0.f}, one = {(_Float16)1.f, (_Float16)1.f};
        unsigned words[W / 4];
#pragma unroll
        for (int h = 0; h < W / 4; h++) {
            const half2v d0 = __builtin_elementwise_fma(pn[2 * h], f, b[2 * h]);
            const half2v d1 = __builtin_elementwise_fma(pn[2 * h + 1], f, b[2 * h + 1]);
            const half2v e0 = __builtin_elementwise_min(__builtin_elementwise_max(d0 * d0, zero), one);
            const half2v e1 = __builtin_elementwise_min(__builtin_elementwise_max(d1 * d1, zero), one);
            // the conversion keeps the other half of its destination; the
            // first one writes over e0's own register (its high half is
            // replaced by the second), which saves a register clear
            short2v r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(__builtin_bit_cast(short2v, e0), e0, 1.f / 256.f, false);
            r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(r, e1, 1.f / 256.f, true);
            __builtin_memcpy(&words[h], &r, 4);
        }
        // chunk columns are multiples of W in the row (W-byte aligned)
        __builtin_memcpy(__builtin_assume_aligned(dst, W), words, W);
    }
    // the tile's own chunk as kept in LDS
    __device__ __forceinline__ void own(chunk &b) const
    {
#pragma unroll
        for (int h = 0; h < W / 2; h++) b[h] = -(b[h] * f);
    }
};

// 3 workgroups of 8 waves per CU (6 waves per SIMD) fit the LDS for R <= 2;
// the register budget is capped to match.  FP32 storage holds twice the
// staging registers and runs 2 workgroups per CU.
template <int R, class TS>
constexpr int mfma_waves_per_eu() { return R <= 2 && sizeof(TS) == 2 ? 6 : 4; }

template <int R, class TS, class TD, int WFN, bool PRUNED, bool ISO>
__global__ __launch_bounds__(NTM) __attribute__((amdgpu_waves_per_eu(mfma_waves_per_eu<R, TS>(), mfma_waves_per_eu<R, TS>())))
void filter_mfma(MfmaArgs<TS> a)
{
    using T = MTile<R, TS, TD>;
    using DP = DPlane<TD>;
    typedef typename DP::vec8 vec8;
    constexpr const MRows<R> &rows = MRowsOf<R>::value;
    constexpr bool kF8 = std::is_same<TD, f8_t>::value;
    constexpr const F8Plan<R, ISO> &pl = F8PlanOf<R, ISO>::value;
    __shared__ __attribute__((aligned(16))) TD lds[2][T::LDS];

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wg = mxcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = wg % a.ntx;
    const int rem = wg / a.ntx;
    const int by = rem % a.nty;
    const int bz = rem / a.nty;
    const int X0 = bx * MTX, Y0 = by * MTY, Z0 = bz * MTZ;
    const FilterGeom g = a.g;
    const long sp = (long)g.nxp * g.nyp;

    // Zero both buffers: the guard columns are never staged and must hold
    // zeros (they meet zero weights in the band, and 0 * NaN would poison).
    for (int i = t; i < (int)(sizeof(lds) / 4); i += NTM) reinterpret_cast<unsigned *>(&lds[0][0])[i] = 0u;

    // Band operands, one per pattern (per pattern class when ISO) and kind.
    // Lane l holds entry k = 8*(l/16) + i of output column n = l % 16.  A
    // single row's band reads input column x0 - 8 + k, so dx = k - 8 - n; a
    // pair of one-tap rows reads columns x0 .. x0+15 of its first row for
    // k < 16 and of its second row for k >= 16, so output n takes k = n and
    // k = n + 16.
    // FP8 planes: band operand g, lane group q holds slot q of an MFMA that
    // uses it, bytes k = 0..31: a single row reads input column x0 - 4 + k
    // (dx = k - 4 - n), a pair reads x0 + k of its first row (k < 16) and
    // x0 + k - 16 of its second; the bytes are FP8 codes of the weights
    // times the common scale a.wscale.
    const int n = lane & 15, q = lane >> 4;
    constexpr int NB = kF8 ? pl.nsig : ISO ? rows.nib : rows.nob;
    vec8 bop[NB];
    if constexpr (!kF8) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int p = ISO ? rows.ibpat[b] : rows.obpat[b];
            const bool pair = (ISO ? rows.ibkind[b] : rows.obkind[b]) != 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int k = 8 * q + i, dx = k - 8 - n;
                const float w = pair ? ((k & 15) == n ? a.band[p][R] : 0.f)
                                     : ((dx >= -rows.pj[p] && dx <= rows.pj[p]) ? a.band[p][dx + R] : 0.f);
                bop[b][i] = DP::from(w);
            }
        }
    } else {
        // the template row of this lane's group, read from byte 16 - n
        // (4-byte loads at any byte offset: the argument block is read as
        // bytes and packed)
        const unsigned char *tb = &a.f8tmpl[0][0][0];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const unsigned char *row = tb + (b * 4 + q) * 48 + 16 - n;
#pragma unroll
            for (int wd = 0; wd < 8; wd++) {
                unsigned w = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) w |= (unsigned)row[4 * wd + e] << (8 * e);
                bop[b][wd] = (int)w;
            }
        }
    }

    // FP8 planes hold D * S with S = 256 f^2, f = 2^m per tile.  One
    // saturated tap (its square clamped to 1, times 256) adds at least
    // 256 * (15/16) * wmin * wscale to the accumulator (an FP8 e4m3 code of
    // a weight lies within 1/16 of it), so 4^m <= (15/32) wmin |mm| / 3.56648
    // with |mm| the smallest over the tile's slices puts twice the largest
    // cutoff distance of the tile below what that tap alone adds: a patch
    // pair with a saturated tap keeps weight exactly zero at every patch
    // radius and under z-varying sigma.  The cutoff then lands between 0.25
    // and 1 times 120 wmin in plane units (2.4 at r = 3 and constant
    // sigma), inside the FP8 normal range [2^-6, 448].  f is staged as a
    // normal FP16 number, so 4^m stays within [4^-12, 4^8]; at the lower
    // clamp (a noise level thousands of standard deviations of the
    // standardised volume) a tap saturates only at a difference of 4096
    // standard deviations.  The band holds the weights times a.wscale, so
    // the accumulator is d * S * wscale.
    float f8f = 1.f;
    if constexpr (kF8) {
        float smax = 0.f;
#pragma unroll
        for (int zz = 0; zz < MTZ; zz++) smax = __builtin_fmaxf(smax, a.sigma[Z0 + zz < g.nz ? Z0 + zz : g.nz - 1]);
        const float x = __builtin_fabsf((-1.f / (smax * smax * a.beta)) * a.mscale) * a.wmin * (15.f / 32.f / 3.56648f);
        // exponent of x from its bits (0 and subnormals give -127, infinity
        // and NaN 128; both clamp), S / 256 = 4^m <= x
        int e = (int)((__builtin_bit_cast(unsigned, x) >> 23) & 0xffu) - 127;
        e = e < -24 ? -24 : e > 16 ? 16 : e;
        const int m = e >> 1;
        f8f = __builtin_bit_cast(float, (unsigned)(127 + m) << 23);
    }

    // Scaled FP16 planes (FP16 and BF16 storage, Pade weights): the plane
    // holds D * f^2, f = 2^m per tile, clamped to [0, 1] by the squaring
    // instruction.  f^2 <= wmin / (4 * 3.56648 / |mm|) with |mm| the smallest
    // of the tile's slices: a clamped tap alone then adds at least four
    // times the largest cutoff distance of the tile, so every patch with a
    // clamped tap stays beyond the cutoff and its weight stays exactly zero.
    constexpr bool kSC = WFN == 0 && std::is_same<TD, _Float16>::value && sizeof(TS) == 2;
    float pf = 1.f;
    if constexpr (kSC) {
        float smax = 0.f;
#pragma unroll
        for (int zz = 0; zz < MTZ; zz++) smax = __builtin_fmaxf(smax, a.sigma[Z0 + zz < g.nz ? Z0 + zz : g.nz - 1]);
        const float x = __builtin_fabsf((-1.f / (smax * smax * a.beta)) * a.mscale) * a.wmin * (1.f / (4.f * 3.56648f));
        int e = (int)((__builtin_bit_cast(unsigned, x) >> 23) & 0xffu) - 127;
        e = e < -24 ? -24 : e > 16 ? 16 : e;
        const int m = e >> 1;
        pf = __builtin_bit_cast(float, (unsigned)(127 + m) << 23);
    }

    // Staging chunks: 8 or 4 columns of one region row.  Region row (rz, ry)
    // is slice Z0-R+rz, row Y0-R+ry; LDS column c holds x = X0 - 8 + c
    // (16-bit planes) or x = X0 - 4 + c (FP8 planes).
    const TS *Pt = a.P + ((long)(Z0 - R + g.pz) * sp + (long)(Y0 - R + g.py) * g.nxp + (X0 - 4 + g.px));
    constexpr int W8 = sizeof(TS) == 2 ? 8 : 4;
    using ST8 = Stager<W8, TS, TD>;
    using ST4 = Stager<4, TS, TD>;
    typedef typename ST8::chunk chunk8;
    typedef typename ST4::chunk chunk4;
    constexpr int S8 = T::S8, S4 = T::S4;
    // the chunks at the tile's own positions stay fixed over the offsets.
    // Instances with isotropic patches up to radius 2 keep them in registers
    // (their LDS reads sat on each step's critical path: 5% of the FP8-plane
    // pass, 3% of the BF16 and FP32_FAST passes); the other instances, which
    // would spill in the loop, keep them in LDS, one slot per (thread, chunk)
    constexpr bool kP0Reg = WFN == 0 && ISO && R <= 2 && (kF8 ? (!PRUNED || R < 2) : std::is_same<TD, _Float16>::value);
    constexpr int P0L8 = kP0Reg ? 1 : (S8 > 0 ? S8 * NTM : 1), P0L4 = kP0Reg ? 1 : (S4 > 0 ? S4 * NTM : 1);
    __shared__ __attribute__((aligned(16))) chunk8 p0s8[P0L8];
    __shared__ __attribute__((aligned(16))) chunk4 p0s4[P0L4];
    chunk8 p0r8[kP0Reg && S8 > 0 ? S8 : 1];
    chunk4 p0r4[kP0Reg && S4 > 0 ? S4 : 1];
    ST8 stg8;
    ST4 stg4;
    if constexpr (kF8 || kSC) {
        const _Float16 fh = (_Float16)(kF8 ? f8f : pf);
        stg8.f = half2v{fh, fh};
        stg4.f = half2v{fh, fh};
    }
    int coff8[S8 > 0 ? S8 : 1], loff8[S8 > 0 ? S8 : 1], coff4[S4 > 0 ? S4 : 1], loff4[S4 > 0 ? S4 : 1];
    // unit u (4 columns) of the region: its volume offset from Pt and its
    // LDS element offset
    auto place = [&](int u, int &co, int &lo) __attribute__((always_inline)) {
        const int row = u / T::UPR, col = 4 * (u % T::UPR);
        const int ry = row % T::RY, rz = row / T::RY;
        co = (int)(rz * sp + (long)ry * g.nxp + col);
        // 16-bit planes: LDS column c holds x = X0 - 8 + c; FP8 planes:
        // x = X0 - 4 + c (8-byte aligned chunk stores, 16-byte aligned row windows)
        lo = (rz * DP::RYPE + ry) * DP::XWE + (kF8 ? 0 : 4) + col;
    };
#pragma unroll
    for (int c = 0; c < S8; c++) {
        const int cs = t + c * NTM;
        if (cs < T::N8) {
            place(2 * cs, coff8[c], loff8[c]);
        } else {  // spare slots re-read chunk 0
            coff8[c] = 0;
            loff8[c] = T::SPARE8 + 8 * (cs - T::N8);
        }
        chunk8 b;
        stg8.load(b, Pt + coff8[c]);
        if constexpr (kSC)
            stg8.own_scaled(b);
        else if constexpr (kF8 || std::is_same<TS, __bf16>::value)
            stg8.own(b);
        if constexpr (kP0Reg)
            __builtin_memcpy(&p0r8[c], b, sizeof(b));
        else
            __builtin_memcpy(&p0s8[cs], b, sizeof(b));
    }
#pragma unroll
    for (int c = 0; c < S4; c++) {
        const int cs = t + c * NTM;
        if (cs < T::N4) {
            place(2 * T::N8 + cs, coff4[c], loff4[c]);
        } else {
            coff4[c] = 0;
            loff4[c] = T::SPARE4 + 4 * (cs - T::N4);
        }
        chunk4 b;
        stg4.load(b, Pt + coff4[c]);
        if constexpr (kSC)
            stg4.own_scaled(b);
        else if constexpr (kF8 || std::is_same<TS, __bf16>::value)
            stg4.own(b);
        if constexpr (kP0Reg)
            __builtin_memcpy(&p0r4[c], b, sizeof(b));
        else
            __builtin_memcpy(&p0s4[cs], b, sizeof(b));
    }

    // This lane's outputs.  The MFMA forms the transposed block,
    // C^T = B^T A^T (band rows times D-plane columns): the A and B operands
    // of the 16x16x32 MFMA share one lane layout, so the registers are the
    // same, and lane l then holds row n = l % 16 (slice pair member n/8, row
    // n%8) at the four consecutive columns 16*xh + 4*q + i.
    const int xh = wave & 1, gz = wave >> 1;
    const int xb = X0 + 16 * xh + 4 * q;
    const int z = Z0 + 2 * gz + (n >> 3);
    const int y = Y0 + (n & 7);
    // R is read through the tile origin (uniform) plus a 32-bit byte offset
    // per lane, four consecutive values per load.
    const TS *Rt = a.R + ((long)(Z0 + g.pz) * sp + (long)(Y0 + g.py) * g.nxp + (X0 + g.px));
    const unsigned rbyte = (unsigned)(((long)(z - Z0) * sp + (long)(y - Y0) * g.nxp + (xb - X0)) * sizeof(TS));
    const TS *Rv = reinterpret_cast<const TS *>(reinterpret_cast<const char *>(Rt) + rbyte);
    const int zc = z < g.nz ? z : g.nz - 1;
    const float sig = a.sigma[zc];
    const float mm = (-1.f / (sig * sig * a.beta)) * a.mscale;
    // D operand: lane l reads row n = l % 16, columns 16*xh + 8*q .. +7.
    const int aoff = ((2 * gz + (n >> 3) + R) * RYP + ((n & 7) + R)) * XW + 16 * xh + 8 * q;
    // FP8 planes: lane l reads 32 bytes of row n of the item in slot q of
    // each MFMA: a single row from LDS column 16*xh (x0 - 4), a pair 16
    // bytes from column 16*xh + 4 (x0) of each of its two rows; an empty
    // slot reads the lane's own row against zero band bytes.
    constexpr int NM8 = kF8 ? pl.nm : 1;
    int f8a[NM8], f8b[NM8];
    if constexpr (kF8) {
        const int mrow = ((2 * gz + (n >> 3) + R) * RYP8 + ((n & 7) + R)) * XW8 + 16 * xh;
#pragma unroll
        for (int o = 0; o < NM8; o++) {
            f8a[o] = mrow + a.f8offa[o][q];
            f8b[o] = mrow + a.f8offb[o][q];
        }
    }
    // for a pair of one-tap rows: columns x0 + 8 (q % 2) of the first row
    // (q < 2) or of the second (q >= 2)
    int poff[rows.npd > 0 ? rows.npd : 1];
#pragma unroll
    for (int d = 0; d < rows.npd; d++) poff[d] = aoff + 8 - 16 * (q >> 1) + (q >> 1) * rows.pdelta[d];

    // wmax is kept as the bit pattern of a non-negative float: integer max
    // orders those like floats and needs no NaN canonicalisation
    float fv[4] = {0.f, 0.f, 0.f, 0.f}, fw[4] = {0.f, 0.f, 0.f, 0.f};
    unsigned wmax_bits[4] = {0u, 0u, 0u, 0u};
    float dinv = 1.f;  // FP8 planes: 1 / (S * wscale)
    if constexpr (kF8) dinv = 1.f / (256.f * f8f * f8f * a.wscale);
    if constexpr (kSC) dinv = 1.f / (pf * pf);
    const float ms = mm * weight_scale_factor<WFN>() * dinv;
    constexpr bool kPacked = (std::is_same<TS, _Float16>::value || std::is_same<TS, __bf16>::value) && WFN == 0;
    ushort2w wmax2[2] = {{0, 0}, {0, 0}};
    const _Float16 msu1 = (_Float16)(mm * (1.f / 3.56648f) * dinv);
    const half2w msu = {msu1, msu1};

    // Buffer resources per volume and tile, starting at the smallest search
    // offset: a search offset is then the uniform (scalar) part of the
    // address and the per-lane part stays fixed.  Every address lies inside
    // the allocation; reads past its end would return zero.  With 16-bit
    // storage the row and slice pitches are even, so every load of one
    // offset has the parity of px + lin in its element index; an odd one
    // reads the shifted copy one element earlier, so that every load address
    // is 4-byte aligned.
    const TS *Pb = Pt + a.min_lin;
    const TS *Rb = Rt + a.min_lin;
    const TS *Pb1 = Pb + (a.shifted - 1);
    const TS *Rb1 = Rb + (a.shifted - 1);
    auto records = [](long bytes) __attribute__((always_inline)) {
        const unsigned long b = (unsigned long)bytes;
        return (b >> 31) != 0 ? 0x7fffffff : (int)b;
    };
    const int nP = records((long)(a.P_end - Pb) * (long)sizeof(TS));
    const int nR = records((long)(a.R_end - Rb) * (long)sizeof(TS));
    auto odd_of = [&](int lin) __attribute__((always_inline)) {
        if constexpr (sizeof(TS) == 2) return ((lin + g.px) & 1) != 0;
        return false;
    };
    auto rsrc_p = [&](int lin) __attribute__((always_inline)) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<TS *>(odd_of(lin) ? Pb1 : Pb), (short)0, nP, kRsrcWord3);
    };
    auto rsrc_r = [&](int lin) __attribute__((always_inline)) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<TS *>(odd_of(lin) ? Rb1 : Rb), (short)0, nR, kRsrcWord3);
    };
    // the smallest offset and the runtime 1.0, read here: an argument first
    // read inside the loop would make every step wait for all outstanding LDS
    // work (scalar argument loads and LDS share one wait counter)
    const int min_lin = a.min_lin;
    const float one = a.one;
    asm volatile("" ::"s"(min_lin), "s"(one));
    auto soff_of = [&](int lin) __attribute__((always_inline)) {
        return (unsigned)((lin - min_lin) * (int)sizeof(TS));
    };
    chunk8 pn8[S8 > 0 ? S8 : 1];
    chunk4 pn4[S4 > 0 ? S4 : 1];
    auto load_chunks = [&](int lin) __attribute__((always_inline)) {
        const rsrc_t r = rsrc_p(lin);
        const unsigned so = soff_of(lin);
#pragma unroll
        for (int c = 0; c < S8; c++) stg8.load(pn8[c], r, (unsigned)(coff8[c] * (int)sizeof(TS)), so);
#pragma unroll
        for (int c = 0; c < S4; c++) stg4.load(pn4[c], r, (unsigned)(coff4[c] * (int)sizeof(TS)), so);
    };
    auto stage = [&](TD *buf) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < S8; c++) {
            if constexpr (kSC)
                stg8.store_scaled(pn8[c], kP0Reg ? p0r8[c] : p0s8[t + c * NTM], buf + loff8[c]);
            else
                stg8.store(pn8[c], kP0Reg ? p0r8[c] : p0s8[t + c * NTM], buf + loff8[c]);
        }
#pragma unroll
        for (int c = 0; c < S4; c++) {
            if constexpr (kSC)
                stg4.store_scaled(pn4[c], kP0Reg ? p0r4[c] : p0s4[t + c * NTM], buf + loff4[c]);
            else
                stg4.store(pn4[c], kP0Reg ? p0r4[c] : p0s4[t + c * NTM], buf + loff4[c]);
        }
    };
    // R at this lane's four outputs shifted by one search offset, in the
    // stored type
    struct RVals {
        TS v[4];
    };
    auto load_r = [&](RVals &rv, int lin) __attribute__((always_inline)) {
        const rsrc_t rR = rsrc_r(lin);
        const unsigned so = soff_of(lin);
        if constexpr (sizeof(TS) == 4) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rR, rbyte, so, 0);
            __builtin_memcpy(rv.v, &v, 16);
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rR, rbyte, so, 0);
            __builtin_memcpy(rv.v, &v, 8);
        }
    };
    // the weights of one offset's distances and their sums
    auto accumulate = [&](const floatx4 &d, const RVals &rv) __attribute__((always_inline)) {
        if constexpr (kPacked) {
            half2w w[2];
            packed_weights(__builtin_bit_cast(half2w, __builtin_amdgcn_cvt_pkrtz(d[0], d[1])),
                           __builtin_bit_cast(half2w, __builtin_amdgcn_cvt_pkrtz(d[2], d[3])), msu, w[0], w[1]);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float wi = (float)w[i >> 1][i & 1];
                fv[i] = __builtin_fmaf(wi, (float)rv.v[i], fv[i]);
                // times a runtime 1.0, so the sum stays one mixed-precision FMA
                fw[i] = __builtin_fmaf(wi, one, fw[i]);
            }
            // weights are non-negative: their FP16 bit patterns order like them
#pragma unroll
            for (int j = 0; j < 2; j++)
                wmax2[j] = __builtin_elementwise_max(wmax2[j], __builtin_bit_cast(ushort2w, w[j]));
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float w = fast_weight<WFN>(d[i], ms);
                fv[i] = __builtin_fmaf(w, (float)rv.v[i], fv[i]);
                fw[i] = fw[i] + w;
                wmax_bits[i] = __builtin_elementwise_max(wmax_bits[i], __builtin_bit_cast(unsigned, w));
            }
        }
    };

    // offsets this tile visits: the whole search, or its list
    int nS = a.n_search;
    const int *lst = nullptr;
    if constexpr (PRUNED) {
        const int cnt = a.counts[wg];
        if (cnt >= 1 && cnt < nS) {
            nS = cnt;
            lst = a.lists + (long)wg * a.K;
        }
    }
    auto lin_of = [&](int k) __attribute__((always_inline)) {
        if constexpr (PRUNED) {
            if (lst) {
                const unsigned i = (unsigned)lst[k];
                return a.search[i < (unsigned)a.n_search ? i : 0u].lin;
            }
        }
        // the whole search (also a pruned tile's fallback) in one order, so
        // a tile that falls back sums exactly like the full-search kernel
        return a.search_full[k].lin;
    };

    // Software pipeline over the offsets (nS >= 1).  Step k runs the
    // matrix-core chain of offset k right after the barrier, with the weights
    // of offset k-1 and the staging arithmetic of offset k+1's D plane (from
    // chunks loaded one step earlier) interleaved between its MFMAs, then
    // writes that plane to the idle buffer and loads the chunks of offset
    // k+2.  The distances of offset k are first read in step k+1, behind its
    // barrier and its first MFMA (and behind the wait states below the loop
    // for the last offset).  Steps come in pairs that swap the two distance
    // registers and the two LDS buffers; when nS is odd the last step
    // repeats offset nS-1 and its distances are dropped.
    __syncthreads();  // zeroed LDS visible before the first staging
    auto lin_at = [&](int k) __attribute__((always_inline)) { return lin_of(k < nS ? k : nS - 1); };
    load_chunks(lin_at(0));
    stage(lds[0]);
    load_chunks(lin_at(1));
    __syncthreads();

    // no offset before the first: an infinite distance has weight zero
    floatx4 dA = {0.f, 0.f, 0.f, 0.f};
    floatx4 dB = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
    RVals rA, rB;
#pragma unroll
    for (int i = 0; i < 4; i++) rA.v[i] = rB.v[i] = TS(0.f);
    // The linear offsets of steps kb .. kb+63, one per lane (the search
    // table or list read once per block, not per step): a block runs steps
    // kb .. kb+61, which look up to two steps ahead.
    constexpr int kBlock = 62;
    int linv = 0, kb = 0;
    auto step = [&](int k, const TD *rd, TD *st, floatx4 &dout, RVals &rout, const floatx4 &din,
                    const RVals &rin) __attribute__((always_inline)) {
        const int l0 = __builtin_amdgcn_readlane(linv, k - kb);
        const int l2 = __builtin_amdgcn_readlane(linv, k + 2 - kb);
        load_r(rout, l0);
        const TD *L = rd + aoff;
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        if constexpr (!kF8) {
            // A operands in flight: KD reads ahead of the MFMA that consumes
            // them, so the chain waits on LDS latency once rather than per row
            constexpr int NR = rows.nops;
            constexpr int KD = NR < 2 ? NR : 2;
            auto a_of = [&](auto oi) __attribute__((always_inline)) {
                constexpr int o = decltype(oi)::value;
                constexpr const MRows<R> &rw = MRowsOf<R>::value;
                constexpr int off = MRows<R>::lds_off(rw.r[rw.opa[o]]);
                if constexpr (rw.opb[o] >= 0)
                    return *reinterpret_cast<const vec8 *>(rd + poff[rw.opd[o]] + off);
                else
                    return *reinterpret_cast<const vec8 *>(L + off);
            };
            vec8 ab[KD];
            mstatic_for<0, KD>([&](auto ri) __attribute__((always_inline)) { ab[decltype(ri)::value] = a_of(ri); });
            mstatic_for<0, NR>([&](auto ri) __attribute__((always_inline)) {
                constexpr int r = decltype(ri)::value;
                constexpr int b = ISO ? MRowsOf<R>::value.ibop[r] : MRowsOf<R>::value.obop[r];
                acc = DP::mfma(bop[b], ab[r % KD], acc);
                if constexpr (r + KD < NR) ab[r % KD] = a_of(std::integral_constant<int, r + KD>{});
            });
            accumulate(din, rin);
            stage(st);
            // order: the R load, KD reads, then per MFMA the next A read and
            // a share of the vector work (none right after the first MFMA);
            // the staging writes and the next loads follow the chain
            constexpr int NV = ((kPacked ? 2 * kWeightValuPacked : 4 * kWeightValu) +
                                kStageValu<TS, TD> * (2 * T::S8 + T::S4) + NR - 2) /
                               (NR - 1);
            __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x0100, KD, 0);
            mstatic_for<0, NR>([&](auto ri) __attribute__((always_inline)) {
                __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);
                if constexpr (decltype(ri)::value + KD < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);
                if constexpr (decltype(ri)::value >= 1) __builtin_amdgcn_sched_group_barrier(0x0002, NV, 0);
            });
        } else {
            // FP8: NM8 MFMAs of four slots; a D operand is two 16-byte
            // halves (4-byte aligned when the MFMA holds a pair)
            constexpr int KD = NM8 < 2 ? NM8 : 2;
            auto d_of = [&](auto oi) __attribute__((always_inline)) {
                constexpr int o = decltype(oi)::value;
                vec8 d;
                if constexpr (F8PlanOf<R, ISO>::value.haspair[o]) {
                    __builtin_memcpy(&d, __builtin_assume_aligned(rd + f8a[o], 4), 16);
                    __builtin_memcpy(reinterpret_cast<char *>(&d) + 16, __builtin_assume_aligned(rd + f8b[o], 4), 16);
                } else {
                    __builtin_memcpy(&d, __builtin_assume_aligned(rd + f8a[o], 16), 16);
                    __builtin_memcpy(reinterpret_cast<char *>(&d) + 16, __builtin_assume_aligned(rd + f8b[o], 16), 16);
                }
                return d;
            };
            vec8 ab[KD];
            mstatic_for<0, KD>([&](auto ri) __attribute__((always_inline)) { ab[decltype(ri)::value] = d_of(ri); });
            if constexpr (kP0Reg) stage(st);
            accumulate(din, rin);
            mstatic_for<0, NM8>([&](auto ri) __attribute__((always_inline)) {
                constexpr int r = decltype(ri)::value;
                acc = DP::mfma(bop[F8PlanOf<R, ISO>::value.sig[r]], ab[r % KD], acc);
                if constexpr (r + KD < NM8) ab[r % KD] = d_of(std::integral_constant<int, r + KD>{});
            });
            if constexpr (!kP0Reg) stage(st);
            // order: the first KD operand reads; while they land, the staging
            // of the next plane when its inputs are all in registers (its two
            // LDS writes then overlap the MFMA chain instead of preceding the
            // barrier) and the weights of the previous offset; then the first
            // MFMA.  With three MFMAs per offset there is little else to cover
            // the LDS latency.
            constexpr int NVA = kPacked ? 2 * kWeightValuPacked : 4 * kWeightValu;
            constexpr int NVS = kStageValu<TS, TD> * (2 * T::S8 + T::S4);
            __builtin_amdgcn_sched_group_barrier(0x0100, 2 * KD, 0);
            if constexpr (kP0Reg) {
                __builtin_amdgcn_sched_group_barrier(0x0002, NVS, 0);
                __builtin_amdgcn_sched_group_barrier(0x0200, T::S8 + T::S4, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x0002, NVA, 0);
            __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);
        }
        // the loads below overwrite the chunks the staging has just read
        __builtin_amdgcn_sched_barrier(0);
        load_chunks(l2);
        __syncthreads();
        dout = acc;
    };
    for (kb = 0; kb < nS; kb += kBlock) {
        linv = lin_at(kb + lane);
        const int ke = kb + kBlock < nS ? kb + kBlock : nS;
        for (int k = kb; k < ke; k += 2) {
            step(k, lds[0], lds[1], dA, rA, dB, rB);
            step(k + 1, lds[1], lds[0], dB, rB, dA, rA);
        }
    }
    // the last offset's distances: wait states for the matrix-core result
    // (the compiler counts them across the loop exit too; these make sure)
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
    if ((nS & 1) == 0) accumulate(dB, rB);

    float wmax[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        wmax[i] = kPacked ? (float)__builtin_bit_cast(half2w, wmax2[i >> 1])[i & 1]
                          : __builtin_bit_cast(float, wmax_bits[i]);
    if (y < g.ny && z < g.nz) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int x = xb + i;
            if (x < g.nx) {
                const long o = ((long)z * g.ny + y) * g.nx + x;
                float res = a.raw[o];
                if (wmax[i] > 0.f) {
                    const double r0 = (double)(float)Rv[i];
                    const double v = ((double)fv[i] + (double)wmax[i] * r0) / ((double)fw[i] + (double)wmax[i]);
                    res = (float)(a.shift + v * a.inv_scale);
                }
                a.out[o] = res;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block form for patch radius 2 (FP16 and BF16 storage, FP16 D planes).
//
// The patch distance of a 4x4 block of outputs in (x, y) at one z is, per
// patch plane dz, a 2D correlation of that plane of D with the plane's
// weights.  Every output of a block reads the plane's rows y0-2 .. y0+5 at
// columns x0-2 .. x0+5 (a disk of radius 2 around a 4x4 square fits in 8
// columns), so one v_mfma_f32_16x16x32_f16 covers four such 8-column rows
// (one per lane group) of 16 blocks:
//   A[m][k]  m = output (bx, by) of a block, k = 8 q + c: the weight of
//            input column x0-2+c of lane group q's row, (dx, dy, dz) =
//            (c-2-bx, row-by, dz), zero outside the patch
//   B[k][n]  the D plane at block n, lane group q's row, column x0-2+c
//   C[m][n] += A . B
// Seven MFMAs cover the 33 taps: plane dz=0 rows -2..5 (two), planes
// dz=+-1 rows -1..2 (one each, equal weights), their rows 3 and 4 together
// (one), planes dz=+-2 rows 0..3 (one each, equal weights).  The row form
// (filter_mfma) takes eleven, since its 16-wide rows leave 12 of 32 K slots
// unused for a 5-tap row and 14 for a 3-tap row; the seven operand reads
// here replace eleven.
//
// A workgroup of 8 wavefronts owns a 32x16x8 output tile and stages D over
// the tile and its patch halo, 36x20x12 (2.1 staged values per output), in
// LDS (double buffered, one barrier per offset).  Wave w computes slice
// Z0 + w in two halves of 8 rows, h = 0, 1, with one MFMA chain each (the
// two chains interleaved): its 16 blocks per half are 8 x-blocks by 2
// y-blocks, lane l reads block n = l % 16 (x-block n % 8, y-block n / 8)
// and holds, after the MFMA, outputs x0..x0+3 of row 8 h + by, by = l / 16.
// The lanes of one row then cover 32 consecutive x, so a wave's R load
// touches 8 rows: a layout whose wave spans 4 columns and 64 rows took 2.3
// times as long, bound by the cache lines each load instruction touches.
// The 32x8x8 tile (one half per wave, 36x12x12 staged, 2.5 per output) took
// 9-12% longer: the staging loads are the kernel's first limit.
//
// LDS layout: a region row is 8 windows of 8 columns, window k holding
// region columns 4k .. 4k+7, so every block's window (region column 4 xb)
// is window xb, 16-byte aligned.  Rows are 128 bytes apart and slices 2576
// bytes (161 16-byte units, 1 mod 16).  gfx950 serves a 16-byte read in
// 16-lane groups ({0-3,12-15,20-27}, {4-11,16-19,28-31} and the same + 32)
// and a 16-byte write in 8 consecutive lanes per cycle (measured: an
// 8-cycle ds_write_b128 against a 4-cycle read): every read of the plan
// hits 16 distinct 16-byte bank groups per cycle, every staging write 8.
// A layout that wrote the shifted half-windows with 8-byte stores spent half
// its LDS cycles in bank conflicts, 7% slower than the row form.
//
// Staging: 15 slots of 64 lanes each cover a 4x4 block of region rows
// (ry in 4 (s % 5) .. +3, rz in 4 (s / 5) .. +3); lane 4 j + c computes the D
// values of columns 8c .. 8c+7 of row (ry0 + j / 4, rz0 + j % 4), and lanes
// with c = 3 also of columns 32-35.  It writes window 2c (its own 8 columns)
// and window 2c+1 (its upper four and the next lane's lower four, moved by
// DPP; columns 32-35 for c = 3).  The two rows of 8 consecutive lanes
// (j = 2i, 2i+1) are one slice apart, an odd number of 16-byte units, so
// their chunks interleave in the banks; an earlier mapping with the two rows
// two slices apart wrote with 2-way conflicts (2% slower).  Wave w runs
// slot w and, unless w = (workgroup index + 7) mod 8, slot
// 8 + (w - workgroup index) mod 8.
//
// Run moves (FP16 form): slot w's chunks are loaded only at every third
// offset of a run of the full search (one dz and dy, dx stepping by 2) and
// moved one dword along the row in between (the cg = 3 lanes load columns
// 32-39 for that), which skips about half of slot w's loads at search
// radius 7: 3-4% faster.  Moving the second slot as well, or moving in the
// BF16 form, puts a spill reload into the loop and was slower (3% and 7%).
// ---------------------------------------------------------------------------
// operand reads in flight ahead of the MFMA that consumes them
#ifndef INLM_B4_KD
#define INLM_B4_KD 2
#endif
namespace b4 {
constexpr int TX = 32, TY = 16, TZ = 8;
constexpr int NH = TY / 8;  // 32x8 output halves per wavefront
constexpr int RX = TX + 4, RY = TY + 4, RZ = TZ + 4;  // staged region
constexpr int ROW = 64;                               // LDS row pitch, halves (8 windows)
constexpr int SLICE = RY * ROW + 8;                   // LDS slice pitch, halves (97 16-byte units)
constexpr int LDS = RZ * SLICE;                       // halves per buffer
constexpr int NSLOT = (RY / 4) * (RZ / 4);
static_assert(RX == 36 && RY % 4 == 0 && RZ % 4 == 0 && NSLOT > 8 && NSLOT <= 16, "staging slots");
constexpr int NOPS = 7, NCLS = 5;
// plane and row of the K group q of MFMA o; o = 4 holds rows 3, 4 of
// plane +1 (q < 2) and of plane -1 (q >= 2)
__host__ __device__ constexpr int op_dz(int o, int q)
{
    return o <= 1 ? 0 : o == 2 ? 1 : o == 3 ? -1 : o == 4 ? (q < 2 ? 1 : -1) : o == 5 ? 2 : -2;
}
__host__ __device__ constexpr int op_row(int o, int q)
{
    return o == 0 ? q - 2 : o == 1 ? q + 2 : o <= 3 ? q - 1 : o == 4 ? (q < 2 ? q + 3 : q + 1) : q;
}
// operations with equal weights share an A operand (the weights depend on |dz|)
constexpr int kCls[NOPS] = {0, 1, 2, 2, 3, 4, 4};
constexpr int kClsOp[NCLS] = {0, 1, 2, 4, 5};
// LDS offset of MFMA o relative to a lane's row of K group 0 (uniform except o = 4)
constexpr int kOff[NOPS] = {-2 * ROW, 2 * ROW, SLICE - ROW, -SLICE - ROW, 0, 2 * SLICE, -2 * SLICE};
}  // namespace b4

template <class TS>
__device__ __forceinline__ float b4_weight(const MfmaArgs<TS> &a, int dx, int dy, int dz)
{
    constexpr const MRows<2> &rows = MRowsOf<2>::value;
    const int ady = dy < 0 ? -dy : dy, adz = dz < 0 ? -dz : dz;
    float w = 0.f;
#pragma unroll
    for (int p = 0; p < rows.npat; p++)
        if (rows.pdy[p] == ady && rows.pdz[p] == adz && dx >= -rows.pj[p] && dx <= rows.pj[p]) w = a.band[p][dx + 2];
    return w;
}

// Its loop holds up to 128 registers: two workgroups per CU (4 waves per SIMD)
template <class TS>
__global__ __launch_bounds__(NTM) __attribute__((amdgpu_waves_per_eu(4, 4))) void filter_mfma_b4(MfmaArgs<TS> a)
{
    using ST8 = Stager<8, TS, _Float16>;
    using ST4 = Stager<4, TS, _Float16>;
    typedef typename ST8::chunk chunk8;
    typedef typename ST4::chunk chunk4;
    __shared__ __attribute__((aligned(16))) _Float16 lds[2][b4::LDS];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wg = mxcd_swizzle(blockIdx.x, gridDim.x);
    const int bxi = wg % a.ntx;
    const int rem = wg / a.ntx;
    const int byi = rem % a.nty;
    const int bzi = rem / a.nty;
    const int X0 = bxi * b4::TX, Y0 = byi * b4::TY, Z0 = bzi * b4::TZ;
    const FilterGeom g = a.g;
    const long sp = (long)g.nxp * g.nyp;

    for (int i = t; i < (int)(sizeof(lds) / 4); i += NTM) reinterpret_cast<unsigned *>(&lds[0][0])[i] = 0u;

    // A operands: lane l holds row m = l % 16 (output bx = m & 3, by = m >> 2)
    // of K group q = l / 16
    const int q = lane >> 4;
    half8 aop[b4::NCLS];
    {
        const int m = lane & 15, mx = m & 3, my = m >> 2;
#pragma unroll
        for (int c = 0; c < b4::NCLS; c++) {
            const int o = b4::kClsOp[c];
            const int dz = b4::op_dz(o, q), row = b4::op_row(o, q);
#pragma unroll
            for (int i = 0; i < 8; i++) aop[c][i] = (_Float16)b4_weight(a, i - 2 - mx, row - my, dz);
        }
    }

    // Scaled FP16 planes: D * f^2 clamped to [0, 1] by the squaring
    // instruction, f = 2^m per tile with f^2 <= wmin / (4 * 3.56648 / |mm|)
    // and |mm| the smallest over the tile's slices: a clamped tap alone adds
    // four times the tile's largest cutoff distance, so a patch with a
    // clamped tap keeps weight exactly zero.
    float pf;
    {
        float smax = 0.f;
#pragma unroll
        for (int zz = 0; zz < b4::TZ; zz++) smax = __builtin_fmaxf(smax, a.sigma[Z0 + zz < g.nz ? Z0 + zz : g.nz - 1]);
        const float x = __builtin_fabsf((-1.f / (smax * smax * a.beta)) * a.mscale) * a.wmin * (1.f / (4.f * 3.56648f));
        int e = (int)((__builtin_bit_cast(unsigned, x) >> 23) & 0xffu) - 127;
        e = e < -24 ? -24 : e > 16 ? 16 : e;
        const int mexp = e >> 1;
        pf = __builtin_bit_cast(float, (unsigned)(127 + mexp) << 23);
    }

    // Staging (region row (ry, rz) = slice Z0-2+rz, row Y0-2+ry; region
    // column c = x X0-2+c): this lane's rows in its slot(s) and column group
    const TS *Pt = a.P + ((long)(Z0 - 2 + g.pz) * sp + (long)(Y0 - 2 + g.py) * g.nxp + (X0 - 2 + g.px));
    ST8 stg8;
    ST4 stg4;
    const _Float16 fh = (_Float16)pf;
    stg8.f = half2v{fh, fh};
    stg4.f = half2v{fh, fh};
    const int cg = lane & 3;
    // slots 8-14 go to seven of the waves; the one left out rotates with the
    // workgroup, so that the workgroups sharing a CU idle different SIMDs
    const int jx = (wave + 8 - (int)(blockIdx.x & 7)) & 7;
    const bool two = jx < b4::NSLOT - 8;
    auto slot_place = [&](int sl, int &coff, int &lrow) __attribute__((always_inline)) {
        const int j = lane >> 2;
        const int ry = 4 * (sl % (b4::RY / 4)) + (j >> 2);
        const int rz = 4 * (sl / (b4::RY / 4)) + (j & 3);
        coff = (int)(rz * sp + (long)ry * g.nxp + 8 * cg);
        lrow = rz * b4::SLICE + ry * b4::ROW + 16 * cg;
    };
    int coffA, lrowA, coffX = 0, lrowX = 0;
    slot_place(wave, coffA, lrowA);
    if (two) slot_place(8 + jx, coffX, lrowX);

    // this lane's block and outputs
    const int n = lane & 15, xb = n & 7, yb = n >> 3;
    const int x0 = X0 + 4 * xb;
    const int y = Y0 + 4 * yb + q;
    const int z = Z0 + wave;
    const TS *Rt = a.R + ((long)(Z0 + g.pz) * sp + (long)(Y0 + g.py) * g.nxp + (X0 + g.px));
    const unsigned rbyte = (unsigned)(((long)(z - Z0) * sp + (long)(y - Y0) * g.nxp + (x0 - X0)) * sizeof(TS));
    const TS *Rv = reinterpret_cast<const TS *>(reinterpret_cast<const char *>(Rt) + rbyte);
    const int zc = z < g.nz ? z : g.nz - 1;
    const float sig = a.sigma[zc];
    const float mm = (-1.f / (sig * sig * a.beta)) * a.mscale;
    const unsigned rhalf = (unsigned)(8 * g.nxp * (int)sizeof(TS));
    // D operand: block n's window xb (region columns 4 xb .. 4 xb + 7) of K
    // group q's row; MFMA o adds kOff[o], o = 4 its own per-group offset
    const int bbase = (wave + 2) * b4::SLICE + (4 * yb + 2 + q) * b4::ROW + 8 * xb;
    const int b4base = bbase + (q < 2 ? b4::SLICE + 3 * b4::ROW : -b4::SLICE + b4::ROW);

    float fv[b4::NH][4], fw[b4::NH][4];
    ushort2w wmax2[b4::NH][2];
#pragma unroll
    for (int h = 0; h < b4::NH; h++) {
#pragma unroll
        for (int i = 0; i < 4; i++) fv[h][i] = fw[h][i] = 0.f;
        wmax2[h][0] = wmax2[h][1] = ushort2w{0, 0};
    }
    const float dinv = 1.f / (pf * pf);
    const _Float16 msu1 = (_Float16)(mm * (1.f / 3.56648f) * dinv);
    const half2w msu = {msu1, msu1};

    // Buffer resources per volume and tile from the smallest search offset;
    // an odd (px + lin) reads the copy shifted by one element, one element
    // earlier, so every load is 4-byte aligned.
    const TS *Pb = Pt + a.min_lin;
    const TS *Rb = Rt + a.min_lin;
    const TS *Pb1 = Pb + (a.shifted - 1);
    const TS *Rb1 = Rb + (a.shifted - 1);
    auto records = [](long bytes) __attribute__((always_inline)) {
        const unsigned long b = (unsigned long)bytes;
        return (b >> 31) != 0 ? 0x7fffffff : (int)b;
    };
    const int nP = records((long)(a.P_end - Pb) * (long)sizeof(TS));
    const int nR = records((long)(a.R_end - Rb) * (long)sizeof(TS));
    auto odd_of = [&](int lin) __attribute__((always_inline)) { return ((lin + g.px) & 1) != 0; };
    auto rsrc_p = [&](int lin) __attribute__((always_inline)) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<TS *>(odd_of(lin) ? Pb1 : Pb), (short)0, nP, kRsrcWord3);
    };
    auto rsrc_r = [&](int lin) __attribute__((always_inline)) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<TS *>(odd_of(lin) ? Rb1 : Rb), (short)0, nR, kRsrcWord3);
    };
    const int min_lin = a.min_lin;
    const float one = a.one;
    asm volatile("" ::"s"(min_lin), "s"(one));
    auto soff_of = [&](int lin) __attribute__((always_inline)) {
        return (unsigned)((lin - min_lin) * (int)sizeof(TS));
    };
    // per slot: 8 columns, and the row's tail on lanes with cg = 3 (at 16
    // bytes past their 8-column chunk): columns 32-35, or 32-39 for slot A of
    // the FP16 form, which moves its chunks along runs of the search instead
    // of loading them at every offset (below)
    constexpr bool kMove = std::is_same<TS, _Float16>::value;
    chunk8 pnA, pnX;
    std::conditional_t<kMove, chunk8, chunk4> pqA = {};
    chunk4 pqX = {};
    auto load_chunks = [&](int lin, bool slot_a) __attribute__((always_inline)) {
        const rsrc_t r = rsrc_p(lin);
        const unsigned so = soff_of(lin);
        const unsigned oA = (unsigned)(coffA * (int)sizeof(TS)), oX = (unsigned)(coffX * (int)sizeof(TS));
        if (slot_a) {
            stg8.load(pnA, r, oA, so);
            if constexpr (kMove) {
                if (cg == 3) stg8.load(pqA, r, oA + 8 * (unsigned)sizeof(TS), so);
            } else {
                if (cg == 3) stg4.load(pqA, r, oA + 8 * (unsigned)sizeof(TS), so);
            }
        }
        if (two) {
            stg8.load(pnX, r, oX, so);
            if (cg == 3) stg4.load(pqX, r, oX + 8 * (unsigned)sizeof(TS), so);
        }
    };
    // the chunks at the tile's own positions (the P(y) side of every
    // difference, search offset 0), scaled by -f, in registers
    chunk8 p0A, p0X;
    chunk4 q0A, q0X;
    load_chunks(0, true);
    __builtin_memcpy(&p0A, &pnA, sizeof(p0A));
    __builtin_memcpy(&q0A, &pqA, sizeof(q0A));
    stg8.own_scaled(p0A);
    stg4.own_scaled(q0A);
    if (two) {
        __builtin_memcpy(&p0X, &pnX, sizeof(p0X));
        __builtin_memcpy(&q0X, &pqX, sizeof(q0X));
        stg8.own_scaled(p0X);
        stg4.own_scaled(q0X);
    }
    auto st16 = [](_Float16 *dst, const half2v *v) __attribute__((always_inline)) {
        __builtin_memcpy(__builtin_assume_aligned(dst, 16), v, 16);
    };
    // window 2 cg (own 8 columns) and window 2 cg + 1 (upper four, then the
    // next lane's lower four by DPP row shift, or columns 32-35 for cg = 3)
    auto store_slot = [&](_Float16 *buf, const chunk8 &pn, const chunk8 &p0, const chunk4 &pq, const chunk4 &q0,
                          int lrow) __attribute__((always_inline)) {
        half2v d[4], e[4];
        stg8.scaled(pn, p0, d);
        e[0] = d[2];
        e[1] = d[3];
#pragma unroll
        for (int h = 0; h < 2; h++)  // row_shl:1, lane i reads lane i + 1 of its 16-lane row
            e[2 + h] = __builtin_bit_cast(half2v, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, d[h]), 0x101, 0xf, 0xf,
                                                                            true));
        if (cg == 3) {
            half2v d4[2];
            stg4.scaled(pq, q0, d4);
            e[2] = d4[0];
            e[3] = d4[1];
        }
        st16(buf + lrow, d);
        st16(buf + lrow + 8, e);
    };
    auto stage = [&](_Float16 *buf) __attribute__((always_inline)) {
        chunk4 pq;
        __builtin_memcpy(&pq, &pqA, sizeof(pq));
        store_slot(buf, pnA, p0A, pq, q0A, lrowA);
        if (two) store_slot(buf, pnX, p0X, pqX, q0X, lrowX);
    };
    // Slot A's chunks of an offset from those of the offset before it when
    // the two differ by dx + 2 (same dz and dy): each region row moves by one
    // dword, so dwords 1-3 move down and the last comes from the next lane of
    // the row (DPP row shift) or, for cg = 3, from the tail, which holds
    // columns 32-39 of the loaded offset and so serves two such moves.
    auto move_a = [&]() __attribute__((always_inline)) {
        if constexpr (kMove) {
            unsigned w[4], tl[4];
            __builtin_memcpy(w, &pnA, 16);
            __builtin_memcpy(tl, &pqA, 16);
            const unsigned from_next = (unsigned)__builtin_amdgcn_mov_dpp((int)w[0], 0x101, 0xf, 0xf, true);
            w[0] = w[1];
            w[1] = w[2];
            w[2] = w[3];
            w[3] = cg == 3 ? tl[0] : from_next;
            tl[0] = tl[1];
            tl[1] = tl[2];
            tl[2] = tl[3];
            __builtin_memcpy(&pnA, w, 16);
            __builtin_memcpy(&pqA, tl, 16);
        }
    };
    struct RVals {
        TS v[b4::NH][4];
    };
    auto load_r = [&](RVals &rv, int lin) __attribute__((always_inline)) {
        const rsrc_t r = rsrc_r(lin);
        const unsigned so = soff_of(lin);
#pragma unroll
        for (int h = 0; h < b4::NH; h++) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, rbyte + h * rhalf, so, 0);
            __builtin_memcpy(rv.v[h], &v, 8);
        }
    };
    typedef floatx4 DVals[b4::NH];
    auto accumulate = [&](const DVals &dd, const RVals &rv) __attribute__((always_inline)) {
#pragma unroll
        for (int h = 0; h < b4::NH; h++) {
            const floatx4 &d = dd[h];
            half2w w[2];
            packed_weights(__builtin_bit_cast(half2w, __builtin_amdgcn_cvt_pkrtz(d[0], d[1])),
                           __builtin_bit_cast(half2w, __builtin_amdgcn_cvt_pkrtz(d[2], d[3])), msu, w[0], w[1]);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float wi = (float)w[i >> 1][i & 1];
                fv[h][i] = __builtin_fmaf(wi, (float)rv.v[h][i], fv[h][i]);
                // times a runtime 1.0, so the sum stays one mixed-precision FMA
                fw[h][i] = __builtin_fmaf(wi, one, fw[h][i]);
            }
#pragma unroll
            for (int j = 0; j < 2; j++)
                wmax2[h][j] = __builtin_elementwise_max(wmax2[h][j], __builtin_bit_cast(ushort2w, w[j]));
        }
    };

    const int nS = a.n_search;
    auto lin_at = [&](int k) __attribute__((always_inline)) { return a.search_full[k < nS ? k : nS - 1].lin; };
    // per offset of the full search, after the table: 0 to load slot A's
    // chunks, s > 0 to move those of the offset before it (kStageMovesMax)
    const int *moves = reinterpret_cast<const int *>(a.search_full + nS);
    auto loads_a = [&](int mv) __attribute__((always_inline)) { return !kMove || mv == 0; };

    // Software pipeline as in filter_mfma: step k runs the MFMA chain of
    // offset k right after the barrier, with the weights of offset k-1 and
    // the staging arithmetic of offset k+1 interleaved, writes that plane to
    // the idle buffer and loads the chunks of offset k+2.
    __syncthreads();
    load_chunks(lin_at(0), true);
    stage(lds[0]);
    load_chunks(lin_at(1), loads_a(nS > 1 ? moves[1] : 0));
    __syncthreads();

    DVals dA, dB;
    RVals rA, rB;
#pragma unroll
    for (int h = 0; h < b4::NH; h++) {
        dA[h] = floatx4{0.f, 0.f, 0.f, 0.f};
        dB[h] = floatx4{__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
#pragma unroll
        for (int i = 0; i < 4; i++) rA.v[h][i] = rB.v[h][i] = TS(0.f);
    }
    constexpr int kBlock = 62;
    // per lane of a 62-step block: offset kb + lane packed as lin * 4 + moves;
    // a step reads the entry of offset k+2 and carries those of k and k+1
    int pkv = 0, kb = 0;
    auto packed_at = [&](int k) __attribute__((always_inline)) {
        return lin_at(k) * 4 + (kMove && k < nS ? moves[k] : 0);
    };
    int pc0 = packed_at(0), pc1 = packed_at(1);
    auto step = [&](int k, const _Float16 *rd, _Float16 *st, DVals &dout, RVals &rout, const DVals &din,
                    const RVals &rin) __attribute__((always_inline)) {
        const int p2 = __builtin_amdgcn_readlane(pkv, k + 2 - kb);
        const int l0 = pc0 >> 2;
        const int l2 = p2 >> 2;
        const int m1 = pc1 & 3;
        const int m2 = p2 & 3;
        pc0 = pc1;
        pc1 = p2;
        load_r(rout, l0);
        DVals acc;
#pragma unroll
        for (int h = 0; h < b4::NH; h++) acc[h] = floatx4{0.f, 0.f, 0.f, 0.f};
        // the halves' chains interleaved: operation r is MFMA r / NH of half r % NH
        constexpr int NR = b4::NOPS * b4::NH;
        constexpr int KD = INLM_B4_KD;
        auto b_of = [&](auto ri) __attribute__((always_inline)) {
            constexpr int o = decltype(ri)::value / b4::NH, h = decltype(ri)::value % b4::NH;
            if constexpr (o == 4)
                return *reinterpret_cast<const half8 *>(__builtin_assume_aligned(rd + b4base + h * 8 * b4::ROW, 16));
            else
                return *reinterpret_cast<const half8 *>(
                    __builtin_assume_aligned(rd + bbase + h * 8 * b4::ROW + b4::kOff[o], 16));
        };
        half8 ab[KD];
        mstatic_for<0, KD>([&](auto ri) __attribute__((always_inline)) { ab[decltype(ri)::value] = b_of(ri); });
        mstatic_for<0, NR>([&](auto ri) __attribute__((always_inline)) {
            constexpr int r = decltype(ri)::value;
            constexpr int o = r / b4::NH, h = r % b4::NH;
            acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aop[b4::kCls[o]], ab[r % KD], acc[h], 0, 0, 0);
            if constexpr (r + KD < NR) ab[r % KD] = b_of(std::integral_constant<int, r + KD>{});
        });
        accumulate(din, rin);
        if (!loads_a(m1)) move_a();
        stage(st);
        constexpr int NV = (2 * b4::NH * kWeightValuPacked + kStageValu<TS, _Float16> * 3 * 2 + NR - 2) / (NR - 1);
        __builtin_amdgcn_sched_group_barrier(0x0020, b4::NH, 0);
        __builtin_amdgcn_sched_group_barrier(0x0100, KD, 0);
        mstatic_for<0, NR>([&](auto ri) __attribute__((always_inline)) {
            __builtin_amdgcn_sched_group_barrier(0x0008, 1, 0);
            if constexpr (decltype(ri)::value + KD < NR) __builtin_amdgcn_sched_group_barrier(0x0100, 1, 0);
            if constexpr (decltype(ri)::value >= 1) __builtin_amdgcn_sched_group_barrier(0x0002, NV, 0);
        });
        __builtin_amdgcn_sched_barrier(0);
        load_chunks(l2, loads_a(m2));
        __syncthreads();
#pragma unroll
        for (int h = 0; h < b4::NH; h++) dout[h] = acc[h];
    };
    for (kb = 0; kb < nS; kb += kBlock) {
        // the lane index again (mbcnt of a full mask), so that the table
        // addresses are formed here rather than held in registers (or
        // spilled) through the inner loop; volatile, so it is not hoisted
        int ln;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(ln));
        pkv = packed_at(kb + ln);
        const int ke = kb + kBlock < nS ? kb + kBlock : nS;
        for (int k = kb; k < ke; k += 2) {
            step(k, lds[0], lds[1], dA, rA, dB, rB);
            step(k + 1, lds[1], lds[0], dB, rB, dA, rA);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
    if ((nS & 1) == 0) accumulate(dB, rB);

#pragma unroll
    for (int h = 0; h < b4::NH; h++) {
        const int yh = y + 8 * h;
        float wmax[4];
#pragma unroll
        for (int i = 0; i < 4; i++) wmax[i] = (float)__builtin_bit_cast(half2w, wmax2[h][i >> 1])[i & 1];
        if (yh < g.ny && z < g.nz) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = x0 + i;
                if (x < g.nx) {
                    const long o = ((long)z * g.ny + yh) * g.nx + x;
                    float res = a.raw[o];
                    if (wmax[i] > 0.f) {
                        const double r0 = (double)(float)Rv[(long)8 * h * g.nxp + i];
                        const double v =
                            ((double)fv[h][i] + (double)wmax[i] * r0) / ((double)fw[h][i] + (double)wmax[i]);
                        res = (float)(a.shift + v * a.inv_scale);
                    }
                    a.out[o] = res;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Normalising reflect-pad: padded v_s = (v - shift) * scale in the storage
// type, mirror without repeating the edge (iternlm_prepare.cpp:6-48).
// ---------------------------------------------------------------------------
template <class TS>
__global__ __launch_bounds__(256) void pad_normalise_kernel(const float *__restrict__ in, TS *__restrict__ out,
                                                            long shifted, FilterGeom g, float shift, float scale)
{
    const long nn = (long)g.nxp * g.nyp * g.nzp;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nn) return;
    const int x1 = (int)(idx % g.nxp);
    const int y1 = (int)((idx / g.nxp) % g.nyp);
    const int z1 = (int)(idx / ((long)g.nxp * g.nyp));
    int xx = x1 - g.px, yy = y1 - g.py, zz = z1 - g.pz;
    while (xx < 0 || xx >= g.nx) xx = (xx < 0) ? -xx : 2 * g.nx - xx - 2;
    while (yy < 0 || yy >= g.ny) yy = (yy < 0) ? -yy : 2 * g.ny - yy - 2;
    while (zz < 0 || zz >= g.nz) zz = (zz < 0) ? -zz : 2 * g.nz - zz - 2;
    const float v = in[((long)zz * g.ny + yy) * g.nx + xx];
    const TS o = (TS)((v - shift) * scale);
    out[idx] = o;
    if (shifted > 0) out[shifted + idx - 1] = o;  // the copy shifted by one element
}

// Sum and sum of squares of a float volume in FP64, deterministic: each of
// kMomentBlocks workgroups writes its partial pair to acc[2 + 2 b], and one
// workgroup then sums the partials in a fixed order into acc[0], acc[1], so
// the normalisation (and every 16-bit form's output) repeats bit for bit.
__global__ __launch_bounds__(256) void moments_kernel(const float *__restrict__ in, long n, double *acc)
{
    double s = 0.0, s2 = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = in[i];
        s += v;
        s2 += v * v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    __shared__ double red[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = s;
        red[1][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        acc[2 + 2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        acc[3 + 2 * blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

__global__ __launch_bounds__(256) void moments_final_kernel(double *acc)
{
    // thread t sums partials t, t + 256, ... in order; then a fixed tree
    double s = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < kMomentBlocks; b += 256) {
        s += acc[2 + 2 * b];
        s2 += acc[3 + 2 * b];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    __shared__ double red[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = s;
        red[1][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        acc[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        acc[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

template <int R>
bool mfma_bands_r(const int *dxyz, const float *w, int ntap, float (*band)[8])
{
    constexpr const MRows<R> &rows = MRowsOf<R>::value;
    bool seen[kMfmaMaxPatterns][8] = {};
    for (int p = 0; p < kMfmaMaxPatterns; p++)
        for (int i = 0; i < 8; i++) band[p][i] = 0.f;
    for (int i = 0; i < ntap; i++) {
        const int dx = dxyz[3 * i], dy = dxyz[3 * i + 1], dz = dxyz[3 * i + 2];
        int pat = -1;
        for (int r = 0; r < rows.n; r++)
            if (rows.r[r].dy == dy && rows.r[r].dz == dz) pat = rows.r[r].pat;
        if (pat < 0 || dx < -rows.pj[pat] || dx > rows.pj[pat]) return false;
        if (!seen[pat][dx + R]) {
            band[pat][dx + R] = w[i];
            seen[pat][dx + R] = true;
        } else if (band[pat][dx + R] != w[i])
            return false;
    }
    // every band position of every pattern must be covered
    for (int p = 0; p < rows.npat; p++)
        for (int dx = -rows.pj[p]; dx <= rows.pj[p]; dx++)
            if (!seen[p][dx + R]) return false;
    int count = 0;
    for (int r = 0; r < rows.n; r++) count += 2 * rows.r[r].j + 1;
    return count == ntap;
}

// OCP FP8 e4m3 (bias 7, no infinities, 448 the largest finite value):
// round to nearest even, values >= 0.
unsigned char e4m3_code(float x)
{
    if (!(x > 0.f)) return 0;
    if (x >= 448.f) return 0x7e;
    int e = 0;
    (void)std::frexp(x, &e);
    int E = e - 1;  // x = 1.f * 2^E
    if (E < -6) {   // subnormal: multiples of 2^-9
        const long r = std::lrint(std::ldexp((double)x, 9));
        return (unsigned char)(r >= 8 ? 0x08 : r);
    }
    long m = std::lrint((std::ldexp((double)x, -E) - 1.0) * 8.0);
    if (m == 8) {
        m = 0;
        E++;
    }
    const int code = ((E + 7) << 3) | (int)m;
    return (unsigned char)(code > 0x7e ? 0x7e : code);
}
float e4m3_value(unsigned char c)
{
    const int E = (c >> 3) & 15, m = c & 7;
    return E == 0 ? std::ldexp((float)m, -9) : std::ldexp(1.f + m / 8.f, E - 7);
}
// The common factor in [1, 2), a multiple of 1/512, that brings the band
// weights closest to FP8 values (largest relative rounding error over the
// weights), and the codes.
float fp8_bands(const float (*band)[8], int npat, unsigned char (*codes)[8])
{
    // one entry per host thread: a handle launches with the same bands every
    // iteration
    thread_local float last_band[kMfmaMaxPatterns][8];
    thread_local int last_npat = -1;
    thread_local float last_c = 1.f;
    float best = 1.f, err_best = 1e30f;
    if (npat == last_npat && std::memcmp(last_band, band, sizeof(float) * 8 * (size_t)npat) == 0) {
        best = last_c;
        err_best = 0.f;
    }
    for (int i = 0; i < 512 && err_best > 0.f; i++) {
        const float c = 1.f + i / 512.f;
        float err = 0.f;
        for (int p = 0; p < npat; p++)
            for (int k = 0; k < 8; k++)
                if (band[p][k] > 0.f) {
                    const float v = band[p][k] * c;
                    err = std::max(err, std::fabs(e4m3_value(e4m3_code(v)) / v - 1.f));
                }
        if (err < err_best) {
            err_best = err;
            best = c;
        }
    }
    for (int p = 0; p < kMfmaMaxPatterns; p++)
        for (int k = 0; k < 8; k++) codes[p][k] = p < npat ? e4m3_code(band[p][k] * best) : 0;
    std::memcpy(last_band, band, sizeof(float) * 8 * (size_t)npat);
    last_npat = npat;
    last_c = best;
    return best;
}

// FP8 planes: per band operand and lane group the Toeplitz template (byte
// 16 - n + k of the template is byte k of lane n: a single row of half-width
// j puts its codes at 20 + dx, a pair of one-tap rows its two codes at 16
// and 32), and per MFMA and lane group the LDS offsets of the two 16-byte
// halves relative to the lane's own row (a single row from its window start
// at x0 - 4, a pair from x0 in each of its rows; an empty slot reads the
// lane's own row).
template <int R, bool ISO, class A>
void fp8_tables_plan(const unsigned char (*codes)[8], A &a)
{
    constexpr const MRows<R> &rows = MRowsOf<R>::value;
    constexpr const F8Plan<R, ISO> &pl = F8PlanOf<R, ISO>::value;
    for (int g = 0; g < pl.nsig; g++)
        for (int q = 0; q < 4; q++) {
            unsigned char *t = a.f8tmpl[g][q];
            const int it = pl.slot[pl.sigm[g]][q];
            if (it < 0) continue;
            const int pa = rows.r[pl.ia[it]].pat, pp = ISO ? rows.crep[rows.cls[pa]] : pa;
            if (pl.ib[it] >= 0) {
                const int pb = rows.r[pl.ib[it]].pat, qb = ISO ? rows.crep[rows.cls[pb]] : pb;
                t[16] = codes[pp][R];
                t[32] = codes[qb][R];
            } else {
                for (int dx = -rows.pj[pp]; dx <= rows.pj[pp]; dx++) t[20 + dx] = codes[pp][dx + R];
            }
        }
    auto roff = [&](int r) { return (rows.r[r].dz * RYP8 + rows.r[r].dy) * XW8; };
    for (int o = 0; o < pl.nm; o++)
        for (int q = 0; q < 4; q++) {
            const int it = pl.slot[o][q];
            int oa = 0, ob = 16;
            if (it >= 0 && pl.ib[it] >= 0) {
                oa = 4 + roff(pl.ia[it]);
                ob = 4 + roff(pl.ib[it]);
            } else if (it >= 0) {
                oa = roff(pl.ia[it]);
                ob = oa + 16;
            }
            a.f8offa[o][q] = oa;
            a.f8offb[o][q] = ob;
        }
}

template <int R, class TS, class TD, int WFN>
hipError_t launch_mfma_t(const void *Rs, const void *Ps, size_t elems, long shifted, const float *raw, float *out,
                         const float *sigma, float beta, float mscale, double shift, double inv_scale,
                         const SearchEntry *search, const SearchEntry *search_full, int n_search, int min_lin,
                         const TapTable &taps, const FilterGeom &g, const OffsetLists *ol, bool row_form,
                         hipStream_t st)
{
    if constexpr (sizeof(TS) == 2) {
        // 16-bit loads stay 4-byte aligned only with an even row pitch and
        // the shifted copies behind the volumes
        if ((g.nxp & 1) != 0 || shifted < (long)elems || (shifted & 1) != 0) return hipErrorInvalidValue;
    } else if (shifted != 0) {
        return hipErrorInvalidValue;
    }
    MfmaArgs<TS> a;
    a.shifted = shifted;
    a.min_lin = min_lin;
    a.one = 1.f;
    a.lists = ol ? ol->lists : nullptr;
    a.counts = ol ? ol->counts : nullptr;
    a.K = ol ? ol->K : 0;
    a.R = static_cast<const TS *>(Rs);
    a.P = static_cast<const TS *>(Ps);
    a.R_end = a.R + elems;
    a.P_end = a.P + elems;
    a.raw = raw;
    a.out = out;
    a.sigma = sigma;
    a.search = search;
    a.search_full = search_full;
    a.n_search = n_search;
    a.beta = beta;
    a.mscale = mscale;
    a.shift = shift;
    a.inv_scale = inv_scale;
    a.g = g;
    a.ntx = (g.nx + MTX - 1) / MTX;
    a.nty = (g.ny + MTY - 1) / MTY;
    a.ntz = (g.nz + MTZ - 1) / MTZ;
    if (!mfma_bands_r<R>(taps.dxyz, taps.w, taps.n, a.band)) return hipErrorInvalidValue;
    a.wscale = 1.f;
    a.wmin = 1.f;
    for (int p = 0; p < MRowsOf<R>::value.npat; p++)
        for (int dx = -MRowsOf<R>::value.pj[p]; dx <= MRowsOf<R>::value.pj[p]; dx++)
            if (a.band[p][dx + R] > 0.f) a.wmin = std::min(a.wmin, a.band[p][dx + R]);
    std::memset(a.f8tmpl, 0, sizeof(a.f8tmpl));
    std::memset(a.f8offa, 0, sizeof(a.f8offa));
    std::memset(a.f8offb, 0, sizeof(a.f8offb));
    unsigned char codes[kMfmaMaxPatterns][8] = {};
    if constexpr (std::is_same<TD, f8_t>::value) a.wscale = fp8_bands(a.band, MRowsOf<R>::value.npat, codes);
    const long nwg = (long)a.ntx * a.nty * a.ntz;
    if (ol && (ol->ntiles != nwg || !ol->lists || !ol->counts || ol->K < 1)) return hipErrorInvalidValue;
    // patterns of one class carry identical bands when the weights are
    // isotropic: then one B operand serves the class
    constexpr const MRows<R> &rows = MRowsOf<R>::value;
    bool iso = true;
    for (int p = 0; p < rows.npat; p++)
        for (int i = 0; i < 8; i++) iso = iso && a.band[p][i] == a.band[rows.crep[rows.cls[p]]][i];
    if constexpr (std::is_same<TD, f8_t>::value) {
        // the tables follow the plan of the kernel launched below
        if (iso) fp8_tables_plan<R, true>(codes, a);
        else fp8_tables_plan<R, false>(codes, a);
    }
    // the block form: full search at patch radius 2, 16-bit storage, FP16
    // planes, Pade weights; it packs lin * 4 + moves into one int, so |lin|
    // stays below 2^29 (the search space is symmetric: the largest lin is
    // -min_lin); beyond that the row form runs
    if constexpr (R == 2 && sizeof(TS) == 2 && std::is_same<TD, _Float16>::value && WFN == 0) {
        if (!ol && !row_form && -(long)min_lin < (1L << 29)) {
            a.ntx = (g.nx + b4::TX - 1) / b4::TX;
            a.nty = (g.ny + b4::TY - 1) / b4::TY;
            a.ntz = (g.nz + b4::TZ - 1) / b4::TZ;
            const long nb = (long)a.ntx * a.nty * a.ntz;
            hipLaunchKernelGGL((filter_mfma_b4<TS>), dim3((unsigned)nb), dim3(NTM), 0, st, a);
            return hipGetLastError();
        }
    }
    const dim3 grid((unsigned)nwg), block(NTM);
    if (ol && iso) hipLaunchKernelGGL((filter_mfma<R, TS, TD, WFN, true, true>), grid, block, 0, st, a);
    else if (ol) hipLaunchKernelGGL((filter_mfma<R, TS, TD, WFN, true, false>), grid, block, 0, st, a);
    else if (iso) hipLaunchKernelGGL((filter_mfma<R, TS, TD, WFN, false, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((filter_mfma<R, TS, TD, WFN, false, false>), grid, block, 0, st, a);
    return hipGetLastError();
}

template <int R, int WFN>
hipError_t launch_mfma_form(int form, const void *Rs, const void *Ps, size_t elems, long shifted, const float *raw,
                            float *out,
                            const float *sigma, float beta, float mscale, double shift, double inv_scale,
                            const SearchEntry *search, const SearchEntry *search_full, int n_search, int min_lin,
                            const TapTable &taps, const FilterGeom &g, const OffsetLists *ol, bool row_form,
                            hipStream_t st)
{
    switch (form) {
    case FORM_FP32_FAST:
        return launch_mfma_t<R, float, _Float16, WFN>(Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift,
                                                      inv_scale, search, search_full, n_search, min_lin, taps, g, ol, row_form, st);
    case FORM_FP16:
        return launch_mfma_t<R, _Float16, _Float16, WFN>(Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift,
                                                         inv_scale, search, search_full, n_search, min_lin, taps, g, ol, row_form, st);
    case FORM_FP16_FP8:
        return launch_mfma_t<R, _Float16, f8_t, WFN>(Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift,
                                                     inv_scale, search, search_full, n_search, min_lin, taps, g, ol, row_form, st);
    case FORM_BF16:
        return launch_mfma_t<R, __bf16, _Float16, WFN>(Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift,
                                                     inv_scale, search, search_full, n_search, min_lin, taps, g, ol, row_form, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool mfma_taps_supported(int rp, const int *dxyz, const float *w, int n)
{
    float band[kMfmaMaxPatterns][8];
    switch (rp) {
    case 1: return mfma_bands_r<1>(dxyz, w, n, band);
    case 2: return mfma_bands_r<2>(dxyz, w, n, band);
    case 3: return mfma_bands_r<3>(dxyz, w, n, band);
    default: return false;
    }
}

hipError_t launch_filter_mfma(int form, int rp, int weight_fn, const void *Rs, const void *Ps, size_t elems,
                              long shifted, const float *raw, float *out, const float *sigma, float beta, float mscale,
                              double shift, double inv_scale, const SearchEntry *search,
                              const SearchEntry *search_full, int n_search, int min_lin, const TapTable &taps,
                              const FilterGeom &g, hipStream_t st, const OffsetLists *ol, bool row_form)
{
    if (min_lin > 0) return hipErrorInvalidValue;
#define INLM_MF(RR)                                                                                               \
    return weight_fn == 0 ? launch_mfma_form<RR, 0>(form, Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift, inv_scale, \
                                                    search, search_full, n_search, min_lin, taps, g, ol, row_form, st)                       \
                          : launch_mfma_form<RR, 1>(form, Rs, Ps, elems, shifted, raw, out, sigma, beta, mscale, shift, inv_scale, \
                                                    search, search_full, n_search, min_lin, taps, g, ol, row_form, st)
    switch (rp) {
    case 1: INLM_MF(1);
    case 2: INLM_MF(2);
    case 3: INLM_MF(3);
    default: return hipErrorInvalidValue;
    }
#undef INLM_MF
}

hipError_t launch_pad_normalise(int form, const float *in, void *out, long shifted, const FilterGeom &g, float shift,
                                float scale, hipStream_t st)
{
    const long nn = (long)g.nxp * g.nyp * g.nzp;
    const unsigned blocks = (unsigned)((nn + 255) / 256);
    switch (form) {
    case FORM_FP32_FAST:
        if (shifted != 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pad_normalise_kernel<float>, dim3(blocks), dim3(256), 0, st, in, (float *)out, 0L, g,
                           shift, scale);
        break;
    case FORM_FP16:
    case FORM_FP16_FP8:
        hipLaunchKernelGGL(pad_normalise_kernel<_Float16>, dim3(blocks), dim3(256), 0, st, in, (_Float16 *)out,
                           shifted, g, shift, scale);
        break;
    case FORM_BF16:
        hipLaunchKernelGGL(pad_normalise_kernel<__bf16>, dim3(blocks), dim3(256), 0, st, in, (__bf16 *)out, shifted,
                           g, shift, scale);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_moments(const float *in, long n, double *acc, hipStream_t st)
{
    hipLaunchKernelGGL(moments_kernel, dim3(kMomentBlocks), dim3(256), 0, st, in, n, acc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, st, acc);
    return hipGetLastError();
}

}  // namespace inlm
