// iterNLM filter, padding and noise kernels for gfx950 (MI355X).
//
// The filter pass computes, per voxel x of slice z (sigma = sigma[z],
// m = -1/(sigma^2*beta)):
//   d_s(x) = sum_p W_p * (P[x+s+p] - P[x+p])^2       (taps p in reference order)
//   w_s    = pade(d_s*m) if d_s*m > -3.56648 else 0   (or expf)
//   num += w_s*R[x+s]; den += w_s; wmax = max(wmax, w_s)
//   out    = (num + wmax*R[x]) / (den + wmax)          (R[x] when wmax == 0)
// which restates Denoiser/iternlm_cpu.cpp:758-843 and the specialised
// Denoiser/iternlm_cpu_kernels.cpp:7-1108 of the reference.
//
// The tiled kernels are offset-major: a workgroup owns a 32x8x8 output tile;
// for each search offset s it stages D_s(y) = (P(y+s) - P(y))^2 over the tile
// plus the patch halo in LDS (double buffered, one barrier per offset) and
// every thread sums the stencil for a run of 8 voxels along x from LDS
// windows.  D_s(x+p) is the very float the reference computes as tmp*tmp for
// tap p, and every voxel adds its taps in the reference order with separate
// multiply and add (this file is compiled with -ffp-contract=off), so the
// result is bit-identical to the reference CPU path.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "inlm_kernels.hpp"

namespace inlm {
namespace {

constexpr int NT = 256;   // threads per workgroup of the helper kernels
constexpr int NTF = 512;  // threads per workgroup of the tiled filter (8 wavefronts)
constexpr int TX = 32, TY = 8, TZ = 8, VX = 4;  // tile; each thread sums a 4-voxel x-run
constexpr float kPadeCut = -3.56648f;

// ---------------------------------------------------------------------------
// Compile-time patch rows of an isotropic ellipsoid of radius R, in the
// reference scan order z -> y -> x (iternlm_prepare.cpp:152-203).  Each row
// (dz,dy) holds the taps dx = -j..j; tap0 is the reference index of its first
// tap (index 0 is the centre, which leads the list).  Every tap also carries a
// weight class: the centre is class 0, other taps are grouped by
// (dx^2+dy^2, dz^2), which fixes 1/(2d+1)^2 for any z anisotropy, so the
// kernel needs one scalar weight per class instead of one per tap.
// ---------------------------------------------------------------------------
struct Row {
    int dz, dy, j, tap0;
};
constexpr int kMaxClasses = 48;
template <int R>
struct Rows {
    Row r[4 * R * R + 4 * R + 1];
    int n, ntaps, nclass;
    int key_rho[kMaxClasses], key_z[kMaxClasses];
    static constexpr bool inside(int x, int y, int z)
    {
        const float rr = (float)R;
        const float fx = (float)x / rr, fy = (float)y / rr, fz = (float)z / rr;
        float e = fx * fx;
        e = e + fy * fy;
        e = e + fz * fz;
        return e <= 1.f;
    }
    constexpr int class_of(int dx, int dy, int dz) const
    {
        if (dx == 0 && dy == 0 && dz == 0) return 0;
        for (int c = 1; c < nclass; c++)
            if (key_rho[c] == dx * dx + dy * dy && key_z[c] == dz * dz) return c;
        return -1;
    }
    constexpr Rows() : r(), n(0), ntaps(0), nclass(1), key_rho(), key_z()
    {
        int tap = 1;
        for (int z = -R; z <= R; ++z)
            for (int y = -R; y <= R; ++y) {
                int j = -1;
                for (int x = 0; x <= R; ++x)
                    if (inside(x, y, z)) j = x;
                if (j < 0) continue;
                r[n] = Row{z, y, j, tap};
                n++;
                tap += 2 * j + 1 - ((z == 0 && y == 0) ? 1 : 0);
                for (int x = -j; x <= j; ++x) {
                    if (x == 0 && y == 0 && z == 0) continue;
                    if (class_of(x, y, z) < 0) {
                        key_rho[nclass] = x * x + y * y;
                        key_z[nclass] = z * z;
                        nclass++;
                    }
                }
            }
        ntaps = tap;
    }
};
template <int R>
struct RowsOf {
    static constexpr Rows<R> value{};
    static_assert(value.nclass <= kMaxClasses, "too many weight classes");
};

// Compiler fence that also pins the running sums: the row's arithmetic must
// complete before it and no memory access moves across it.
__device__ __forceinline__ void pin(float (&d)[4])
{
    asm volatile("" : "+v"(d[0]), "+v"(d[1]), "+v"(d[2]), "+v"(d[3])::"memory");
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// LDS layout of the parity kernel: rows of 40 floats.  gfx950 serves a
// ds_read_b128 in four groups of 16 lanes ({0-3, 12-15, 20-27},
// {4-11, 16-19, 28-31} and the same +32), each group one 256-byte pass over
// 64 banks.  A lane reads the 16 bytes at row ty, float 4*xr, i.e. 16-byte
// bank group (10*ty + xr) mod 16; with a 40-float pitch every bank group
// occurs exactly four times over the 8x8 (xr, ty) cells of a wavefront, and
// kLaneCell hands each hardware group 16 cells with 16 distinct bank groups,
// so the window reads are conflict-free for every row shift.
constexpr int kPitch = 40;
__constant__ const unsigned char kLaneCell[64] = {
    0,  1,  2,  3,  14, 15, 28, 29, 16, 17, 18, 19, 4,  5,  6,  7,  32, 33, 22, 23, 20, 21,
    8,  9,  10, 11, 12, 13, 36, 37, 24, 25, 26, 27, 40, 41, 52, 53, 54, 55, 42, 43, 56, 57,
    30, 31, 44, 45, 58, 59, 60, 61, 46, 47, 34, 35, 48, 49, 38, 39, 62, 63, 50, 51};  // xr | ty << 3

template <int R>
struct Tile {
    static constexpr int RX = TX + 2 * R, RY = TY + 2 * R, RZ = TZ + 2 * R;
    static constexpr int PITCH = kPitch;  // LDS row pitch in floats
    static_assert(RX <= kPitch, "patch radius too large for the LDS row");
    static constexpr int CPR = PITCH / 4;        // 4-float chunks per row
    static constexpr int NCHUNK = RZ * RY * CPR;
    static constexpr int CPT = (NCHUNK + NTF - 1) / NTF;  // chunks staged per thread
    static constexpr int LDS = RZ * RY * PITCH;         // floats per buffer
    static constexpr int WIN = ((VX + 2 * R) + 3) & ~3; // floats read per row window
};

struct TiledArgs {
    const float *R;
    const float *P;
    float *out;
    const float *sigma;
    const SearchEntry *search;
    int n_search;
    float beta;
    FilterGeom g;
    int ntx, nty, ntz;
    float wc[kMaxClasses];  // weight per class (class 0 = centre)
    float *mass;            // optional: [tile][offset] sum of the weights over the tile's voxels
};

// Correctly rounded n/d for the operands the Pade weight produces
// (n in [0, 120], d in [120, 600]: no overflow, underflow or special value):
// a reciprocal refined by one Newton step, then two remainder corrections
// with exact fused remainders.  This is the IEEE division sequence without the
// range scaling and special-value fixup that these operands never need;
// filter_selftest_div checks it against IEEE division on every float in the
// weight's domain.
__device__ __forceinline__ float div_pade(float n, float d)
{
    float y = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = n * y;
    float r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(r, y, q);
}

// (120 + 60a + 12a^2 + a^3) / (120 - 60a + 12a^2 - a^3), evaluated left to
// right with every product and sum rounded, as Denoiser/iternlm_cpu.h:125
// writes it, and zero at or below the cutoff of iternlm_cpu.h:126.
template <bool IEEE_DIV>
__device__ __forceinline__ float pade_cut(float a)
{
    const float t60 = 60.f * a;
    const float t12 = (12.f * a) * a;
    const float t3 = (a * a) * a;
    float num = 120.f + t60;
    num = num + t12;
    num = num + t3;
    float den = 120.f - t60;
    den = den + t12;
    den = den - t3;
    const float q = IEEE_DIV ? num / den : div_pade(num, den);
    return (a > kPadeCut) ? q : 0.0f;
}

template <int WFN>
__device__ __forceinline__ float weight_of(float a)
{
    if constexpr (WFN == 0) return pade_cut<false>(a);
    else return expf(a);
}

// Workgroup id -> tile, so that the tiles one XCD runs are contiguous (the
// hardware deals workgroups to the 8 XCDs round-robin, and neighbouring tiles
// share their halo in that XCD's L2).
__device__ __forceinline__ int xcd_swizzle(int wg, int nwg)
{
    if ((nwg & 7) != 0) return wg;
    return (wg & 7) * (nwg >> 3) + (wg >> 3);
}

// Occupancy is bounded by LDS to 3 workgroups of 8 waves per CU (R <= 2) or
// 2 (R = 3), i.e. 6 or 4 waves per SIMD; the register budget is capped to
// match so that the scheduler does not trade occupancy for hoisted loads.
template <int R>
constexpr int waves_per_eu() { return R <= 2 ? 6 : 4; }

template <int R, int WFN, bool MASS>
__global__ __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(waves_per_eu<R>(), waves_per_eu<R>()))) void
filter_tiled_fp32(TiledArgs a)
{
    using T = Tile<R>;
    __shared__ __attribute__((aligned(16))) float lds[2][T::LDS];
    __shared__ float red[2][NTF / 64];

    const int t = threadIdx.x;
    const int wg = xcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = wg % a.ntx;
    const int rem = wg / a.ntx;
    const int by = rem % a.nty;
    const int bz = rem / a.nty;
    const int X0 = bx * TX, Y0 = by * TY, Z0 = bz * TZ;
    const FilterGeom g = a.g;
    const long sp = (long)g.nxp * g.nyp;

    // Staging: region origin (padded coordinates) is the tile origin minus R.
    const float *Pbase = a.P + ((long)(Z0 + g.pz - R) * sp + (long)(Y0 + g.py - R) * g.nxp + (X0 + g.px - R));
    int coff[T::CPT];
    float p0[T::CPT][4];
#pragma unroll
    for (int c = 0; c < T::CPT; c++) {
        int ci = t + c * NTF;
        if (ci >= T::NCHUNK) ci = 0;  // duplicate of chunk 0, never written
        const int cx = ci % T::CPR, row = ci / T::CPR;
        const int ry = row % T::RY, rz = row / T::RY;
        coff[c] = (int)(rz * sp + (long)ry * g.nxp + 4 * cx);
#pragma unroll
        for (int e = 0; e < 4; e++) p0[c][e] = Pbase[coff[c] + e];
    }

    // This thread's 4-voxel run.
    const int cell = kLaneCell[t & 63];
    const int xr = cell & 7, ty = cell >> 3, tz = t >> 6;
    const int x0 = X0 + VX * xr, y = Y0 + ty, z = Z0 + tz;
    const long vox = (long)(z + g.pz) * sp + (long)(y + g.py) * g.nxp + (x0 + g.px);
    const float *Rv = a.R + vox;
    const int zc = z < g.nz ? z : g.nz - 1;
    const float sig = a.sigma[zc];
    const float m = -1.f / (sig * sig * a.beta);
    const bool valid = y < g.ny && z < g.nz;

    // wmax as the bits of a non-negative float: the integer max orders those
    // like floats (== (wt > wmax ? wt : wmax)) without NaN canonicalisation
    float fv[VX], fw[VX];
    unsigned wmax_bits[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        fv[v] = fw[v] = 0.f;
        wmax_bits[v] = 0u;
    }

    const int ldsrow = ((tz + R) * T::RY + (ty + R)) * T::PITCH + VX * xr;
    const int nS = a.n_search;

    auto load_window = [](const float *src, float (&win)[T::WIN]) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < T::WIN / 4; i++) {
            typedef float f4v __attribute__((ext_vector_type(4)));
            f4v q = *reinterpret_cast<const f4v *>(src + 4 * i);
            // whole 16-byte reads (one ds_read_b128 with an immediate
            // offset), not the narrowed pieces the compiler would pick
            asm("" : "+v"(q));
            win[4 * i] = q[0];
            win[4 * i + 1] = q[1];
            win[4 * i + 2] = q[2];
            win[4 * i + 3] = q[3];
        }
    };
    auto stage = [&](float (&pn)[T::CPT][4], float *buf) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < T::CPT; c++) {
            const int ci = t + c * NTF;
            if (T::NCHUNK % NTF == 0 || ci < T::NCHUNK) {
                float4 dv;
                float q;
                q = pn[c][0] - p0[c][0]; dv.x = q * q;
                q = pn[c][1] - p0[c][1]; dv.y = q * q;
                q = pn[c][2] - p0[c][2]; dv.z = q * q;
                q = pn[c][3] - p0[c][3]; dv.w = q * q;
                *reinterpret_cast<float4 *>(buf + 4 * ci) = dv;
            }
        }
    };
    auto load_chunks = [&](float (&pn)[T::CPT][4], int lin) __attribute__((always_inline)) {
        const float *base = Pbase + lin;
#pragma unroll
        for (int c = 0; c < T::CPT; c++)
#pragma unroll
            for (int e = 0; e < 4; e++) pn[c][e] = base[coff[c] + e];
    };

    {
        float pn[T::CPT][4];
        load_chunks(pn, a.search[0].lin);
        stage(pn, lds[0]);
    }
    __syncthreads();

    for (int k = 0; k < nS; k++) {
        const int lin = a.search[k].lin;
        const bool more = (k + 1) < nS;
        float pn[T::CPT][4];
        if (more) load_chunks(pn, a.search[k + 1].lin);
        float rs[VX];
#pragma unroll
        for (int v = 0; v < VX; v++) rs[v] = Rv[lin + v];

        const float *L = lds[k & 1] + ldsrow;
        float d[VX];
        {
            // the centre tap leads the sum (iternlm_prepare.cpp:176-177)
            float win[T::WIN];
            load_window(L, win);
            const float w0 = a.wc[0];
#pragma unroll
            for (int v = 0; v < VX; v++) d[v] = win[R + v] * w0;
        }
        // Rows in reference order; the window of row i+1 is read while row i
        // is summed.  A compiler barrier per row keeps the optimiser from
        // hoisting every window above the arithmetic (which exhausts the
        // register budget and spills).
        constexpr int NR = RowsOf<R>::value.n;
        float win[2][T::WIN];
        load_window(L + (RowsOf<R>::value.r[0].dz * T::RY + RowsOf<R>::value.r[0].dy) * T::PITCH, win[0]);
        static_for<0, NR>([&](auto ri) __attribute__((always_inline)) {
            constexpr int i = decltype(ri)::value;
            constexpr Row rw = RowsOf<R>::value.r[i];
            if constexpr (i + 1 < NR) {
                constexpr Row nx = RowsOf<R>::value.r[i + 1];
                load_window(L + (nx.dz * T::RY + nx.dy) * T::PITCH, win[(i + 1) & 1]);
            }
            static_for<0, 2 * rw.j + 1>([&](auto kk) __attribute__((always_inline)) {
                constexpr int dx = decltype(kk)::value - rw.j;
                if constexpr (!(rw.dz == 0 && rw.dy == 0 && dx == 0)) {
                    constexpr int cls = RowsOf<R>::value.class_of(dx, rw.dy, rw.dz);
                    const float wt = a.wc[cls];
#pragma unroll
                    for (int v = 0; v < VX; v++) d[v] = d[v] + win[i & 1][R + v + dx] * wt;
                }
            });
            pin(d);
        });
        float sw = 0.f;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            const float wt = weight_of<WFN>(d[v] * m);
            fv[v] = fv[v] + wt * rs[v];
            fw[v] = fw[v] + wt;
            wmax_bits[v] = __builtin_elementwise_max(wmax_bits[v], __builtin_bit_cast(unsigned, wt));
            if constexpr (MASS) sw += (valid && x0 + v < g.nx) ? wt : 0.f;
        }
        if constexpr (MASS) {
            // weight mass of this offset over the tile: wave sums into LDS,
            // thread 0 adds the 8 wave sums after the barrier below
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) sw += __shfl_xor(sw, o, 64);
            if ((t & 63) == 0) red[k & 1][t >> 6] = sw;
        }
        if (more) stage(pn, lds[(k + 1) & 1]);
        __syncthreads();
        if constexpr (MASS) {
            if (t == 0) {
                float tot = 0.f;
#pragma unroll
                for (int w = 0; w < NTF / 64; w++) tot += red[k & 1][w];
                a.mass[(long)wg * nS + k] = tot;
            }
        }
    }

    float wmax[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) wmax[v] = __builtin_bit_cast(float, wmax_bits[v]);
    if (y < g.ny && z < g.nz) {
        float *o = a.out + ((long)z * g.ny + y) * g.nx + x0;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            if (x0 + v < g.nx) {
                const float r0 = Rv[v];
                float res;
                if (wmax[v] > 0.0f) {
                    const float num = fv[v] + wmax[v] * r0;
                    const float den = fw[v] + wmax[v];
                    res = num / den;
                } else
                    res = r0;
                o[v] = res;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Per-voxel kernel for any tap table (non-isotropic patches, anisotropic
// weights): one thread per voxel, gathers straight from global memory, the
// same operation sequence as the tiled kernel.
// ---------------------------------------------------------------------------
struct VoxelArgs {
    const float *R;
    const float *P;
    float *out;
    const float *sigma;
    const SearchEntry *search;
    int n_search;
    float beta;
    FilterGeom g;
    TapTable taps;
};

template <int WFN>
__global__ __launch_bounds__(NT) void filter_pervoxel_fp32(VoxelArgs a)
{
    const FilterGeom g = a.g;
    const long nvox = (long)g.nx * g.ny * g.nz;
    const long idx = (long)blockIdx.x * NT + threadIdx.x;
    if (idx >= nvox) return;
    const int x = (int)(idx % g.nx);
    const int y = (int)((idx / g.nx) % g.ny);
    const int z = (int)(idx / ((long)g.nx * g.ny));
    const long sp = (long)g.nxp * g.nyp;
    const long i0 = (long)(z + g.pz) * sp + (long)(y + g.py) * g.nxp + (x + g.px);
    const float sig = a.sigma[z];
    const float m = -1.f / (sig * sig * a.beta);
    float fv = 0.f, fw = 0.f, wmax = 0.f;
    for (int k = 0; k < a.n_search; k++) {
        const long i1 = i0 + a.search[k].lin;
        float d = 0.f;
        for (int p = 0; p < a.taps.n; p++) {
            const float q = a.P[i1 + a.taps.lin[p]] - a.P[i0 + a.taps.lin[p]];
            d = d + (q * q) * a.taps.w[p];
        }
        const float wt = weight_of<WFN>(d * m);
        fv = fv + wt * a.R[i1];
        fw = fw + wt;
        wmax = __builtin_fmaxf(wmax, wt);
    }
    const float r0 = a.R[i0];
    float res = r0;
    if (wmax > 0.0f) {
        const float num = fv + wmax * r0;
        const float den = fw + wmax;
        res = num / den;
    }
    a.out[idx] = res;
}

// ---------------------------------------------------------------------------
// Three-channel filter of the colour path (iternlm_cpu.cpp:932-1075): one
// thread per output pixel, the three channels' distances in patch order,
// each scaled by its channel's multiplier.  Joint weights (the default) take
// the mean of the three scaled distances and weight all channels alike,
// the centre counted once with the largest weight; independent channels
// weigh each channel by its own distance and, as the reference does, leave
// the centre out.  Every product and sum rounds as the reference's float code
// does (no contraction, IEEE division).
// ---------------------------------------------------------------------------
struct RgbArgs {
    const float *R[3];  // padded originals
    const float *P[3];  // padded previous iterates
    float *out[3];      // unpadded results
    const SearchEntry *search;
    int n_search;
    float mult[3];  // -1 / (sigma_c^2 beta)
    int independent;
    FilterGeom g;
    TapTable taps;
};

template <int WFN>
__global__ __launch_bounds__(NT) void filter_rgb_fp32(RgbArgs a)
{
    const FilterGeom g = a.g;
    const long nvox = (long)g.nx * g.ny * g.nz;
    const long idx = (long)blockIdx.x * NT + threadIdx.x;
    if (idx >= nvox) return;
    const int x = (int)(idx % g.nx);
    const int y = (int)((idx / g.nx) % g.ny);
    const int z = (int)(idx / ((long)g.nx * g.ny));
    const long sp = (long)g.nxp * g.nyp;
    const long i0 = (long)(z + g.pz) * sp + (long)(y + g.py) * g.nxp + (x + g.px);
    const float *P0 = a.P[0], *P1 = a.P[1], *P2 = a.P[2];
    float fv0 = 0.f, fv1 = 0.f, fv2 = 0.f, fw0 = 0.f, fw1 = 0.f, fw2 = 0.f, wm0 = 0.f, wm1 = 0.f, wm2 = 0.f;
    for (int k = 0; k < a.n_search; k++) {
        const long i1 = i0 + a.search[k].lin;
        float d0 = 0.f, d1 = 0.f, d2 = 0.f;
        for (int p = 0; p < a.taps.n; p++) {
            const int pp = a.taps.lin[p];
            const float dw = a.taps.w[p];
            float t = P0[i1 + pp] - P0[i0 + pp];
            d0 = d0 + (t * t) * dw;
            t = P1[i1 + pp] - P1[i0 + pp];
            d1 = d1 + (t * t) * dw;
            t = P2[i1 + pp] - P2[i0 + pp];
            d2 = d2 + (t * t) * dw;
        }
        d0 = d0 * a.mult[0];
        d1 = d1 * a.mult[1];
        d2 = d2 * a.mult[2];
        if (!a.independent) {
            const float ds = ((d0 + d1) + d2) / 3.f;
            const float w = weight_of<WFN>(ds);
            fv0 = fv0 + w * a.R[0][i1];
            fv1 = fv1 + w * a.R[1][i1];
            fv2 = fv2 + w * a.R[2][i1];
            fw0 = fw0 + w;
            wm0 = w > wm0 ? w : wm0;
        } else {
            const float w0 = weight_of<WFN>(d0), w1 = weight_of<WFN>(d1), w2 = weight_of<WFN>(d2);
            fv0 = fv0 + w0 * a.R[0][i1];
            fv1 = fv1 + w1 * a.R[1][i1];
            fv2 = fv2 + w2 * a.R[2][i1];
            fw0 = fw0 + w0;
            fw1 = fw1 + w1;
            fw2 = fw2 + w2;
            wm0 = w0 > wm0 ? w0 : wm0;
            wm1 = w1 > wm1 ? w1 : wm1;
            wm2 = w2 > wm2 ? w2 : wm2;
        }
    }
    const float r0 = a.R[0][i0], r1 = a.R[1][i0], r2 = a.R[2][i0];
    float o0 = r0, o1 = r1, o2 = r2;
    if (!a.independent) {
        if (wm0 > 0.f) {
            fv0 = fv0 + wm0 * r0;
            fv1 = fv1 + wm0 * r1;
            fv2 = fv2 + wm0 * r2;
            fw0 = fw0 + wm0;
            o0 = fv0 / fw0;
            o1 = fv1 / fw0;
            o2 = fv2 / fw0;
        }
    } else {
        if (wm0 > 0.f) o0 = fv0 / fw0;
        if (wm1 > 0.f) o1 = fv1 / fw1;
        if (wm2 > 0.f) o2 = fv2 / fw2;
    }
    a.out[0][idx] = o0;
    a.out[1][idx] = o1;
    a.out[2][idx] = o2;
}

// ---------------------------------------------------------------------------
// Reflective padding (iternlm_prepare.cpp:6-48): mirror without repeating the
// edge, iterated until in range.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void pad_reflect_kernel(const float *__restrict__ in, float *__restrict__ out,
                                                         FilterGeom g)
{
    const long n = (long)g.nxp * g.nyp * g.nzp;
    const long idx = (long)blockIdx.x * NT + threadIdx.x;
    if (idx >= n) return;
    const int x1 = (int)(idx % g.nxp);
    const int y1 = (int)((idx / g.nxp) % g.nyp);
    const int z1 = (int)(idx / ((long)g.nxp * g.nyp));
    int x = x1 - g.px, y = y1 - g.py, z = z1 - g.pz;
    while (x < 0 || x >= g.nx) x = (x < 0) ? -x : 2 * g.nx - x - 2;
    while (y < 0 || y >= g.ny) y = (y < 0) ? -y : 2 * g.ny - y - 2;
    while (z < 0 || z >= g.nz) z = (z < 0) ? -z : 2 * g.nz - z - 2;
    out[idx] = in[((long)z * g.ny + y) * g.nx + x];
}

// ---------------------------------------------------------------------------
// Noise: one wavefront per sample computes the population standard deviation
// of its patchsize x patchsize in-slice patch (noiselevel.cpp:235-283) with
// FP64 two-pass accumulation and a wavefront butterfly reduction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(NT) void patchstd_kernel(const float *__restrict__ origin, long row_pitch,
                                                      long slice_pitch, const int64_t *__restrict__ samples,
                                                      long n, int nx, long nslice, int ps, double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const long s = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (s >= n) return;  // whole wavefront leaves together
    const long idx = samples[s];
    const int z0 = (int)(idx / nslice);
    const int y0 = (int)((idx - z0 * nslice) / nx);
    const int x0 = (int)(idx - z0 * nslice - (long)y0 * nx);
    const float *base = origin + (long)z0 * slice_pitch + (long)y0 * row_pitch + x0;
    const int plen = ps * ps;
    double sum = 0.0;
    for (int i = lane; i < plen; i += 64) sum += (double)base[(long)(i / ps) * row_pitch + (i % ps)];
    const double mean = wave_sum(sum) / (double)plen;
    double sq = 0.0;
    for (int i = lane; i < plen; i += 64) {
        const double dv = (double)base[(long)(i / ps) * row_pitch + (i % ps)] - mean;
        sq += dv * dv;
    }
    const double var = wave_sum(sq) / (double)plen;
    if (lane == 0) out[s] = sqrt(var);
}

// Every negative float from -0 down to the cutoff, as bit patterns
// 0x80000000 + [first, first + count): the Pade weight with the shortened
// division against the same weight with IEEE division.
__global__ __launch_bounds__(NT) void selftest_div_kernel(unsigned first, unsigned count,
                                                          unsigned long long *mismatch, unsigned *example)
{
    const unsigned stride = gridDim.x * NT;
    unsigned bad = 0;
    for (unsigned i = blockIdx.x * NT + threadIdx.x; i < count; i += stride) {
        const float a = __uint_as_float(0x80000000u + first + i);
        const float x = pade_cut<false>(a), y = pade_cut<true>(a);
        if (__float_as_uint(x) != __float_as_uint(y)) {
            bad++;
            *example = 0x80000000u + first + i;
        }
    }
    if (bad) atomicAdd(mismatch, (unsigned long long)bad);
}

}  // namespace

hipError_t launch_selftest_div(unsigned first, unsigned count, unsigned long long *mismatch, unsigned *example,
                               hipStream_t st)
{
    hipLaunchKernelGGL(selftest_div_kernel, dim3(8192), dim3(NT), 0, st, first, count, mismatch, example);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <int R>
static bool taps_match_r(const int *dxyz, int n)
{
    constexpr const Rows<R> &rows = RowsOf<R>::value;
    if (n != rows.ntaps) return false;
    if (dxyz[0] != 0 || dxyz[1] != 0 || dxyz[2] != 0) return false;
    for (int ri = 0; ri < rows.n; ri++) {
        const Row rw = rows.r[ri];
        int tap = rw.tap0;
        for (int dx = -rw.j; dx <= rw.j; dx++) {
            if (rw.dz == 0 && rw.dy == 0 && dx == 0) continue;
            if (dxyz[3 * tap] != dx || dxyz[3 * tap + 1] != rw.dy || dxyz[3 * tap + 2] != rw.dz) return false;
            tap++;
        }
    }
    return true;
}

// Per-class weights of the tiled kernel; false when two taps of one class
// carry different weights (then the per-voxel kernel must run).
template <int R>
static bool class_weights_r(const int *dxyz, const float *w, int n, float *wc)
{
    constexpr const Rows<R> &rows = RowsOf<R>::value;
    bool seen[kMaxClasses] = {};
    for (int i = 0; i < n; i++) {
        const int c = rows.class_of(dxyz[3 * i], dxyz[3 * i + 1], dxyz[3 * i + 2]);
        if (c < 0) return false;
        if (!seen[c]) {
            wc[c] = w[i];
            seen[c] = true;
        } else if (__builtin_memcmp(&wc[c], &w[i], sizeof(float)) != 0)
            return false;
    }
    return true;
}

bool tiled_taps_match(int rp, const int *dxyz, const float *w, int n)
{
    float wc[kMaxClasses];
    switch (rp) {
    case 1: return taps_match_r<1>(dxyz, n) && class_weights_r<1>(dxyz, w, n, wc);
    case 2: return taps_match_r<2>(dxyz, n) && class_weights_r<2>(dxyz, w, n, wc);
    case 3: return taps_match_r<3>(dxyz, n) && class_weights_r<3>(dxyz, w, n, wc);
    default: return false;
    }
}

hipError_t launch_pad_reflect(const float *in, float *out, const FilterGeom &g, hipStream_t st)
{
    const long n = (long)g.nxp * g.nyp * g.nzp;
    const unsigned blocks = (unsigned)((n + NT - 1) / NT);
    hipLaunchKernelGGL(pad_reflect_kernel, dim3(blocks), dim3(NT), 0, st, in, out, g);
    return hipGetLastError();
}

template <int R, int WFN>
static hipError_t launch_tiled(const float *Rv, const float *P, float *out, const float *sigma, float beta,
                               const SearchEntry *search, int n_search, const TapTable &taps, const FilterGeom &g,
                               float *mass, hipStream_t st)
{
    TiledArgs a;
    a.mass = mass;
    a.R = Rv;
    a.P = P;
    a.out = out;
    a.sigma = sigma;
    a.search = search;
    a.n_search = n_search;
    a.beta = beta;
    a.g = g;
    a.ntx = (g.nx + TX - 1) / TX;
    a.nty = (g.ny + TY - 1) / TY;
    a.ntz = (g.nz + TZ - 1) / TZ;
    for (int i = 0; i < kMaxClasses; i++) a.wc[i] = 0.f;
    if (!class_weights_r<R>(taps.dxyz, taps.w, taps.n, a.wc)) return hipErrorInvalidValue;
    const long nwg = (long)a.ntx * a.nty * a.ntz;
    if (mass)
        hipLaunchKernelGGL((filter_tiled_fp32<R, WFN, true>), dim3((unsigned)nwg), dim3(NTF), 0, st, a);
    else
        hipLaunchKernelGGL((filter_tiled_fp32<R, WFN, false>), dim3((unsigned)nwg), dim3(NTF), 0, st, a);
    return hipGetLastError();
}

template <int WFN>
static hipError_t launch_pervoxel(const float *Rv, const float *P, float *out, const float *sigma, float beta,
                                  const SearchEntry *search, int n_search, const TapTable &taps,
                                  const FilterGeom &g, hipStream_t st)
{
    VoxelArgs a;
    a.R = Rv;
    a.P = P;
    a.out = out;
    a.sigma = sigma;
    a.search = search;
    a.n_search = n_search;
    a.beta = beta;
    a.g = g;
    a.taps = taps;
    const long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL((filter_pervoxel_fp32<WFN>), dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_filter_rgb(int weight_fn, const float *const R[3], const float *const P[3], float *const out[3],
                             const float mult[3], int independent, const SearchEntry *search, int n_search,
                             const TapTable &taps, const FilterGeom &g, hipStream_t st)
{
    RgbArgs a;
    for (int c = 0; c < 3; c++) {
        a.R[c] = R[c];
        a.P[c] = P[c];
        a.out[c] = out[c];
        a.mult[c] = mult[c];
    }
    a.search = search;
    a.n_search = n_search;
    a.independent = independent;
    a.g = g;
    a.taps = taps;
    const long n = (long)g.nx * g.ny * g.nz;
    const dim3 grid((unsigned)((n + NT - 1) / NT)), block(NT);
    if (weight_fn == 0) hipLaunchKernelGGL((filter_rgb_fp32<0>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((filter_rgb_fp32<1>), grid, block, 0, st, a);
    return hipGetLastError();
}

hipError_t launch_filter(int form, int rp, int weight_fn, const float *R, const float *P, float *out,
                         const float *sigma, float beta, const SearchEntry *search, int n_search,
                         const TapTable &taps, const FilterGeom &g, hipStream_t st, float *mass)
{
    if (form != FORM_FP32) return hipErrorInvalidValue;
#define INLM_TILED(RR)                                                                                   \
    return weight_fn == 0 ? launch_tiled<RR, 0>(R, P, out, sigma, beta, search, n_search, taps, g, mass, st) \
                          : launch_tiled<RR, 1>(R, P, out, sigma, beta, search, n_search, taps, g, mass, st)
    switch (rp) {
    case 1: INLM_TILED(1);
    case 2: INLM_TILED(2);
    case 3: INLM_TILED(3);
    default: break;
    }
#undef INLM_TILED
    if (mass) return hipErrorInvalidValue;  // weight mass needs the tiled kernel
    return weight_fn == 0 ? launch_pervoxel<0>(R, P, out, sigma, beta, search, n_search, taps, g, st)
                          : launch_pervoxel<1>(R, P, out, sigma, beta, search, n_search, taps, g, st);
}

hipError_t launch_patchstd(const float *origin, int64_t row_pitch, int64_t slice_pitch, const int64_t *samples,
                           int64_t n, int nx, int64_t nslice, int patchsize, double *out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + (NT / 64) - 1) / (NT / 64));
    hipLaunchKernelGGL(patchstd_kernel, dim3(blocks), dim3(NT), 0, st, origin, (long)row_pitch, (long)slice_pitch,
                       samples, (long)n, nx, (long)nslice, patchsize, out);
    return hipGetLastError();
}

}  // namespace inlm
