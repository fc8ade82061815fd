// Host tables and the FP64 noise clustering.  Compiled with
// -ffp-contract=off: the ellipsoid membership test and the weights must round
// exactly as the reference's float32 code does (a fused multiply-add changes
// which offsets fall inside the ellipsoid, e.g. 502 instead of 514 at r=5).
#include "inlm_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace inlm {

Geometry make_geometry(const int shape[3], const int rs[3], const int rp[3], int nslices)
{
    Geometry g;
    g.nx = shape[0];
    g.ny = shape[1];
    g.nz = shape[2];
    g.px = rs[0] + rp[0];
    g.py = rs[1] + rp[1];
    g.pz = std::min(nslices / 2, rs[2]) + rp[2];
    g.nxp = g.nx + 2 * g.px;
    g.nyp = g.ny + 2 * g.py;
    g.nzp = g.nz + 2 * g.pz;
    return g;
}

// Float32 ellipsoid test of iternlm_prepare.cpp:139/:191/:239, evaluated
// left to right with every product rounded.
static bool inside(int x, int y, int z, float rx, float ry, float rz)
{
    volatile float fx = (float)x / rx, fy = (float)y / ry, fz = (float)z / rz;
    float e = fx * fx;
    e = e + fy * fy;
    e = e + fz * fz;
    return e <= 1.f;
}

static std::vector<Offset3> ellipsoid(const int r[3], int nslices, bool centre_first)
{
    const float rx = (float)std::max(r[0], 1), ry = (float)std::max(r[1], 1), rz = (float)std::max(r[2], 1);
    std::vector<Offset3> v;
    if (centre_first) v.push_back({0, 0, 0});
    for (int z = -r[2]; z <= r[2]; z++)
        for (int y = -r[1]; y <= r[1]; y++)
            for (int x = -r[0]; x <= r[0]; x++) {
                if (std::abs(z) > nslices / 2) continue;
                if (x == 0 && y == 0 && z == 0) continue;
                if (inside(x, y, z, rx, ry, rz)) v.push_back({x, y, z});
            }
    return v;
}

std::vector<Offset3> search_offsets(const int rs[3], int nslices) { return ellipsoid(rs, nslices, false); }
std::vector<Offset3> patch_offsets(const int rp[3], int nslices) { return ellipsoid(rp, nslices, true); }

std::vector<float> computed_weights(const int rp[3], int nslices, float z_aniso, int nz_padded)
{
    std::vector<Offset3> taps = patch_offsets(rp, nslices);
    std::vector<float> w(taps.size());
    const float sq_an = z_aniso * z_aniso;
    float maxw = 0.f;
    for (size_t i = 1; i < taps.size(); i++) {
        const Offset3 &t = taps[i];
        // sqrtf((xp*xp)+(yp*yp)+(zp*zp)*sq_anisotropy): integer x/y terms,
        // float z term (iternlm_prepare.cpp:242-243).
        float zz = (float)(t.dz * t.dz) * sq_an;
        float ed = std::sqrt((float)(t.dx * t.dx + t.dy * t.dy) + zz);
        float q = 2.f * ed + 1.f;
        w[i] = 1.f / (q * q);
        if (w[i] > maxw) maxw = w[i];
    }
    w[0] = maxw;
    const bool unit = std::max(rp[0], 1) == 1 && std::max(rp[1], 1) == 1 && std::max(rp[2], 1) == 1;
    if (unit && nz_padded > 1)
        for (float &x : w) x = 1.f / 7.f;
    return w;
}

bool cpu_weight_table(const int rp[3], int nslices, float z_aniso, int nz_padded, bool blockwise,
                      std::vector<float> &out, std::string &err)
{
    const int a = rp[0], b = rp[1], c = rp[2];
    // 0 generic (computed), 1 literal table, 2 the r=1 kernel (1/7).
    int kind = 0;
    if (a == 1 && b == 1 && c == 1) kind = 2;
    else if (a == 1 && b == 1 && (c == 2 || c == 3)) kind = 1;
    else if (a == 2 && b == 2 && (c == 1 || c == 2)) kind = 1;
    else if (!blockwise && a == 3 && b == 3 && c >= 1 && c <= 3) kind = 1;
    else if (blockwise && a == 2 && b == 2 && c == 3) {
        // iternlm_cpu.cpp:543-544: block mode sends (2,2,3) to the r=3
        // kernel, which reads 115 taps from a shorter table.
        err = "patch radius (2,2,3) in block mode has undefined behaviour in the reference CPU path";
        return false;
    }
    if (kind == 2) {
        out.assign(patch_offsets(rp, nslices).size(), 1.f / 7.f);
        return true;
    }
    // The literal tables hold isotropic weights only.  With z anisotropy the
    // reference GPU class applies the computed table at every radius but
    // (1,1,3) (iternlm_gpu.cu:1470, :1700-1720: its generic kernel runs after,
    // and overwrites, every specialised one except patch113's), so -aniso
    // takes effect here as it does there.
    if (kind == 1 && z_aniso != 1.f && !(a == 1 && b == 1 && c == 3)) kind = 0;
    if (kind == 0) {
        out = computed_weights(rp, nslices, z_aniso, nz_padded);
        return true;
    }
    // Literal tables: the computed weight (isotropic voxels) streamed with
    // six significant digits, as the reference's kernel-code generator printed
    // them (iternlm_gpu.cu:1516-1517), parsed back as a float.
    std::vector<float> w = computed_weights(rp, nslices, 1.f, nz_padded);
    out.resize(w.size());
    char buf[64];
    for (size_t i = 0; i < w.size(); i++) {
        std::snprintf(buf, sizeof buf, "%g", (double)w[i]);
        out[i] = std::strtof(buf, nullptr);
    }
    return true;
}

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

bool draw_samples(const int shape[3], int64_t n, int patchsize, uint64_t seed, int64_t *out, int mask_diameter)
{
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx <= patchsize || ny <= patchsize || nz < 1 || n < 0) return false;
    // Circular field of view (noiselevel.cpp:204-224): the corner (x, y) of
    // a sample is redrawn while (x - nx/2)^2 + (y - ny/2)^2 exceeds
    // (diameter/2 - sqrt(2 ps^2))^2, the square taken in float as the
    // reference does (a negative radius therefore squares to a positive one).
    float maxr2 = 0.f;
    const int cx = nx / 2, cy = ny / 2;
    auto r2 = [&](int x, int y) { return (float)((x - cx) * (x - cx) + (y - cy) * (y - cy)); };
    if (mask_diameter > 0) {
        maxr2 = (float)((double)(mask_diameter / 2.f) - std::sqrt((double)(2 * patchsize * patchsize)));
        maxr2 *= maxr2;
        // the reference loops for ever when no corner of the sampling range
        // lies inside the circle; the corner closest to the centre decides
        const int bx = std::min(cx, nx - patchsize - 1), by = std::min(cy, ny - patchsize - 1);
        if (r2(bx, by) > maxr2) return false;
    }
    uint64_t st = seed;
    for (int64_t i = 0; i < n; i++) {
        int x = (int)(splitmix64(st) % (uint64_t)(nx - patchsize));
        int y = (int)(splitmix64(st) % (uint64_t)(ny - patchsize));
        const int z = (int)(splitmix64(st) % (uint64_t)nz);
        if (mask_diameter > 0)
            while (r2(x, y) > maxr2) {
                x = (int)(splitmix64(st) % (uint64_t)(nx - patchsize));
                y = (int)(splitmix64(st) % (uint64_t)(ny - patchsize));
            }
        out[i] = (int64_t)z * nx * ny + (int64_t)y * nx + x;
    }
    std::sort(out, out + n);
    return true;
}

void create2clusters(const double *values, int n, int fs, int ls, double &mean, double &stdev)
{
    if (fs < 0) fs = 0;
    if (ls < 0) ls = n - 1;
    double mn = 1e300, mx = -1e300;
    for (int i = fs; i <= ls; i++) {
        mn = std::min(mn, values[i]);
        mx = std::max(mx, values[i]);
    }
    double m1 = mn + 0.25 * (mx - mn), m2 = mx - 0.25 * (mx - mn);
    std::vector<signed char> asg((size_t)std::max(ls - fs + 1, 1));
    double c1 = 0, c2 = 0, s1 = 0, s2 = 0;
    for (int i = fs; i <= ls; i++) {
        const double v = values[i];
        signed char &a = asg[i - fs];
        if (v == 0.0) a = -1;
        else if (std::fabs(v - m1) < std::fabs(v - m2)) { c1 += v; s1 += 1; a = 0; }
        else { c2 += v; s2 += 1; a = 1; }
    }
    if (s1 > 0) m1 = c1 / s1;
    if (s2 > 0) m2 = c2 / s2;
    if (s1 == 0 && s2 == 0) {
        mean = 0;
        stdev = 0;
        return;
    }
    auto variances = [&](double &v1, double &v2) {
        v1 = v2 = 0;
        for (int i = fs; i <= ls; i++) {
            const double v = values[i];
            if (asg[i - fs] == 0) v1 += (v - m1) * (v - m1);
            else if (asg[i - fs] == 1) v2 += (v - m2) * (v - m2);
        }
        v1 /= std::max(1.0, s1);
        v2 /= std::max(1.0, s2);
    };
    double v1, v2;
    variances(v1, v2);
    const double PI = 3.14159265358979323846;
    bool change = true;
    int iters = 0;
    while (change && v1 > 0 && v2 > 0) {
        change = false;
        iters++;
        c1 = c2 = s1 = s2 = 0;
        for (int i = fs; i <= ls; i++) {
            const double v = values[i];
            if (v == 0.0) continue;
            double p1 = 1. / std::sqrt(v1 * 2 * PI) * std::exp(-((v - m1) * (v - m1) / (2 * v1)));
            double p2 = 1. / std::sqrt(v2 * 2 * PI) * std::exp(-((v - m2) * (v - m2) / (2 * v2)));
            signed char &a = asg[i - fs];
            if (p2 > p1) {
                if (a != 1) { change = true; a = 1; }
                c2 += v; s2 += 1;
            } else {
                if (a != 0) { change = true; a = 0; }
                c1 += v; s1 += 1;
            }
        }
        if (s1 > 0) m1 = c1 / s1;
        if (s2 > 0) m2 = c2 / s2;
        variances(v1, v2);
        if (iters == 1000) break;
    }
    if (m1 <= m2 || v2 <= 0.0) { mean = m1; stdev = std::sqrt(v1); }
    else { mean = m2; stdev = std::sqrt(v2); }
}

void zwindow_clusters(const int64_t *samples, const double *values, int64_t n, const int shape[3], int windowsize,
                      std::vector<double> &mean, std::vector<double> &stdev)
{
    const int64_t nslice = (int64_t)shape[0] * shape[1];
    const int nz = shape[2];
    mean.assign(nz, 0.0);
    stdev.assign(nz, 0.0);
    if (n <= 0) return;
    int64_t startpos = 0, endpos = 0;
    double max_mean = -1.0;
    bool empty = false;
    for (int pos = 0; pos < nz; pos++) {
        // Window of noiselevel.cpp:541-548: slices in (pos-w/2, pos+w/2]
        // plus the first sample beyond (the reference's endpos stops on it).
        while (startpos < n - 1 && samples[startpos] / nslice <= (pos - windowsize / 2)) startpos++;
        while (endpos < n - 1 && samples[endpos] / nslice <= (pos + windowsize / 2)) endpos++;
        create2clusters(values, (int)n, (int)startpos, (int)endpos, mean[pos], stdev[pos]);
        max_mean = std::max(max_mean, mean[pos]);
        if (mean[pos] == 0.0) empty = true;
    }
    // Fill empty slices from their neighbours (noiselevel.cpp:572-603).
    while (empty && max_mean > 0.0) {
        empty = false;
        for (int i = 0; i < nz; i++) {
            if (mean[i] == 0.0) {
                double nm = 0, ns = 0;
                int c = 0;
                if (i > 0 && mean[i - 1] > 0) { nm += mean[i - 1]; ns += stdev[i - 1]; c++; }
                if (i < nz - 1 && mean[i + 1] > 0) { nm += mean[i + 1]; ns += stdev[i + 1]; c++; }
                if (c > 0) { mean[i] = nm / c; stdev[i] = ns / c; }
            }
            if (mean[i] == 0.0) empty = true;
        }
    }
}

}  // namespace inlm
