// Volume I/O and the on-disk output contract of the reference's command-line
// program: TIFF slice sequences in and out, the per-iteration output
// directories, the noise CSV files, the resume check and the log file.
//
// TIFF is read and written here directly (baseline, uncompressed, one sample
// per pixel, strips), without libtiff:
//   * reading: classic TIFF in either byte order; 8- and 16-bit unsigned
//     samples become float, 32-bit samples are taken as IEEE floats
//     (Get2DTifImage_32bitPointer, hdcommunication.cpp:441-551); a single
//     file holds one slice per IFD, or an ImageJ stack ("images=N" in the
//     description, N slices stored back to back after the first strip);
//   * writing: one file per slice, name + 4-digit slice number, float32
//     (SAMPLEFORMAT_IEEEFP) or uint16 with min(max(0, round(x)), 65535)
//     (Save2DTifImage_32bit / _as16bit, hdcommunication.cpp:1356-1486), rows
//     per strip max(1, 2048 / width).
#include <dirent.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "inlm_capi.h"
#include "inlm_internal.hpp"

namespace {

using inlm::set_global_error;

inlm_status io_fail(const std::string &msg)
{
    set_global_error(msg);
    return INLM_IO_ERROR;
}

bool has_ending(const std::string &s, const std::string &e)
{
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

bool is_tif_name(const std::string &s) { return has_ending(s, ".tif") || has_ending(s, ".tiff"); }

bool file_exists(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool dir_exists(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool make_path(const std::string &p)
{
    if (p.empty() || dir_exists(p)) return true;
    std::string cur;
    std::stringstream ss(p);
    std::string part;
    if (p[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (!dir_exists(cur) && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

std::string zfill(int v, int n)
{
    std::string s = std::to_string(v);
    while ((int)s.size() < n) s = "0" + s;
    return s;
}

// sorted *.tif / *.tiff of a directory (GetFilelist, hdcommunication.cpp:1788-1810)
bool list_tifs(const std::string &dir, std::vector<std::string> &out)
{
    DIR *dp = opendir(dir.c_str());
    if (!dp) return false;
    std::string base = dir;
    if (!base.empty() && base.back() != '/') base += "/";
    while (struct dirent *e = readdir(dp)) {
        const std::string n = e->d_name;
        if (is_tif_name(n)) out.push_back(base + n);
    }
    closedir(dp);
    std::sort(out.begin(), out.end());
    return true;
}

// ---------------------------------------------------------------------------
// TIFF reading
// ---------------------------------------------------------------------------
struct TiffImage {
    uint32_t width = 0, height = 0;
    uint16_t bps = 1, spp = 1, compression = 1, planar = 1, sample_format = 1;
    uint32_t rows_per_strip = 0xffffffffu;
    std::vector<uint64_t> offsets, counts;
    std::string description;
    bool tiled = false;
};

class TiffFile {
public:
    explicit TiffFile(const std::string &path) : path_(path) { f_ = std::fopen(path.c_str(), "rb"); }
    ~TiffFile()
    {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr; }
    std::string err;

    // parses every IFD of the file
    bool parse(std::vector<TiffImage> &ifds)
    {
        uint8_t h[8];
        if (!read_at(0, h, 8)) return fail("file shorter than a TIFF header");
        if (h[0] == 'I' && h[1] == 'I') big_ = false;
        else if (h[0] == 'M' && h[1] == 'M') big_ = true;
        else return fail("not a TIFF file (byte-order mark)");
        const uint16_t magic = u16(h + 2);
        if (magic == 43) return fail("BigTIFF is not supported");
        if (magic != 42) return fail("not a TIFF file (magic number)");
        uint64_t off = u32(h + 4);
        int guard = 0;
        while (off != 0) {
            if (++guard > 1000000) return fail("IFD chain does not end");
            uint8_t nb[2];
            if (!read_at(off, nb, 2)) return fail("IFD outside the file");
            const uint16_t n = u16(nb);
            std::vector<uint8_t> ent((size_t)n * 12 + 4);
            if (!read_at(off + 2, ent.data(), ent.size())) return fail("IFD outside the file");
            TiffImage im;
            for (int i = 0; i < n; i++) {
                const uint8_t *e = ent.data() + 12 * i;
                const uint16_t tag = u16(e), type = u16(e + 2);
                const uint32_t count = u32(e + 4);
                std::vector<uint64_t> vals;
                if (tag == 270 && type == 2) {
                    std::string s(count, '\0');
                    if (count <= 4) std::memcpy(&s[0], e + 8, count);
                    else if (!read_at(u32(e + 8), reinterpret_cast<uint8_t *>(&s[0]), count))
                        return fail("description outside the file");
                    im.description = s;
                    continue;
                }
                if (type != 3 && type != 4) continue;  // only SHORT / LONG fields are needed
                if (!values(e, type, count, vals)) return fail("field outside the file");
                if (vals.empty()) continue;
                switch (tag) {
                case 256: im.width = (uint32_t)vals[0]; break;
                case 257: im.height = (uint32_t)vals[0]; break;
                case 258: im.bps = (uint16_t)vals[0]; break;
                case 259: im.compression = (uint16_t)vals[0]; break;
                case 273: im.offsets = vals; break;
                case 277: im.spp = (uint16_t)vals[0]; break;
                case 278: im.rows_per_strip = (uint32_t)vals[0]; break;
                case 279: im.counts = vals; break;
                case 284: im.planar = (uint16_t)vals[0]; break;
                case 322:
                case 323:
                case 324:
                case 325: im.tiled = true; break;
                case 339: im.sample_format = (uint16_t)vals[0]; break;
                default: break;
                }
            }
            ifds.push_back(im);
            off = u32(ent.data() + (size_t)n * 12);
        }
        if (ifds.empty()) return fail("no image in the file");
        return true;
    }

    bool check(const TiffImage &im)
    {
        if (im.width == 0 || im.height == 0) return fail("image without width or height");
        if (im.tiled) return fail("tiled TIFF is not supported");
        if (im.compression != 1) return fail("compressed TIFF is not supported");
        if (im.spp != 1) return fail("images with more than one sample per pixel (RGB) are not supported");
        if (im.bps != 8 && im.bps != 16 && im.bps != 32)
            return fail("only 8-, 16- and 32-bit samples are supported");
        if (im.offsets.empty() || im.offsets.size() != im.counts.size()) return fail("missing strip tables");
        return true;
    }

    // rows of one image as float into out (width*height values); slice k of
    // an ImageJ stack (k > 0) starts k slice sizes after the first strip
    bool read_image(const TiffImage &im, float *out, uint64_t stack_index = 0)
    {
        const size_t bpp = im.bps / 8;
        const uint64_t slice_bytes = (uint64_t)im.width * im.height * bpp;
        std::vector<uint8_t> raw(slice_bytes);
        if (stack_index > 0) {
            if (!read_at(im.offsets[0] + stack_index * slice_bytes, raw.data(), slice_bytes))
                return fail("stack slice outside the file");
        } else {
            const uint64_t rps = std::min<uint64_t>(im.rows_per_strip, im.height);
            uint64_t done = 0;
            for (size_t s = 0; s < im.offsets.size() && done < slice_bytes; s++) {
                const uint64_t want = std::min<uint64_t>(rps * im.width * bpp, slice_bytes - done);
                if (!read_at(im.offsets[s], raw.data() + done, want)) return fail("strip outside the file");
                done += want;
            }
            if (done < slice_bytes) return fail("strips hold fewer rows than the image");
        }
        const uint64_t n = (uint64_t)im.width * im.height;
        if (im.bps == 8) {
            for (uint64_t i = 0; i < n; i++) out[i] = (float)raw[i];
        } else if (im.bps == 16) {
            for (uint64_t i = 0; i < n; i++) out[i] = (float)u16(&raw[2 * i]);
        } else {
            for (uint64_t i = 0; i < n; i++) {
                const uint32_t b = u32(&raw[4 * i]);
                std::memcpy(&out[i], &b, 4);
            }
        }
        return true;
    }

    // an RGB image (3 samples per pixel, a 4th alpha sample ignored, as
    // hdcommunication.cpp:708-711), 8 or 16 bits, interleaved; out holds
    // the R, G and B planes of width*height floats each
    bool check_rgb(const TiffImage &im)
    {
        if (im.width == 0 || im.height == 0) return fail("image without width or height");
        if (im.tiled) return fail("tiled TIFF is not supported");
        if (im.compression != 1) return fail("compressed TIFF is not supported");
        if (im.spp != 3 && im.spp != 4) return fail("an RGB image needs 3 or 4 samples per pixel");
        if (im.planar != 1) return fail("only interleaved (contiguous) RGB samples are supported");
        if (im.bps != 8 && im.bps != 16) return fail("only 8- and 16-bit RGB samples are supported");
        if (im.offsets.empty() || im.offsets.size() != im.counts.size()) return fail("missing strip tables");
        return true;
    }
    bool read_rgb(const TiffImage &im, float *out)
    {
        const size_t bpp = im.bps / 8, spp = im.spp;
        const uint64_t n = (uint64_t)im.width * im.height, bytes = n * spp * bpp;
        std::vector<uint8_t> raw(bytes);
        const uint64_t rps = std::min<uint64_t>(im.rows_per_strip, im.height);
        uint64_t done = 0;
        for (size_t s = 0; s < im.offsets.size() && done < bytes; s++) {
            const uint64_t want = std::min<uint64_t>(rps * im.width * spp * bpp, bytes - done);
            if (!read_at(im.offsets[s], raw.data() + done, want)) return fail("strip outside the file");
            done += want;
        }
        if (done < bytes) return fail("strips hold fewer rows than the image");
        for (uint64_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                const uint8_t *q = &raw[(i * spp + c) * bpp];
                out[c * n + i] = bpp == 1 ? (float)q[0] : (float)u16(q);
            }
        return true;
    }

private:
    std::string path_;
    FILE *f_ = nullptr;
    bool big_ = false;

    bool fail(const std::string &m)
    {
        err = path_ + ": " + m;
        return false;
    }
    bool read_at(uint64_t off, uint8_t *dst, uint64_t n)
    {
        if (fseeko(f_, (off_t)off, SEEK_SET) != 0) return false;
        return std::fread(dst, 1, n, f_) == n;
    }
    uint16_t u16(const uint8_t *p) const { return big_ ? (uint16_t)(p[0] << 8 | p[1]) : (uint16_t)(p[1] << 8 | p[0]); }
    uint32_t u32(const uint8_t *p) const
    {
        return big_ ? ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3])
                    : ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]);
    }
    bool values(const uint8_t *e, uint16_t type, uint32_t count, std::vector<uint64_t> &out)
    {
        const size_t sz = type == 3 ? 2 : 4;
        std::vector<uint8_t> buf((size_t)count * sz);
        if (count * sz <= 4) std::memcpy(buf.data(), e + 8, count * sz);
        else if (!read_at(u32(e + 8), buf.data(), buf.size())) return false;
        out.resize(count);
        for (uint32_t i = 0; i < count; i++) out[i] = sz == 2 ? u16(&buf[2 * i]) : u32(&buf[4 * i]);
        return true;
    }
};

// One input: a directory of slice files, or one file (IFD per slice, or an
// ImageJ stack).
struct StackSource {
    std::vector<std::string> files;  // directory mode
    std::string single;               // single-file mode
    int nx = 0, ny = 0, nz = 0;
    bool imagej = false;
    std::string err;

    bool open(const std::string &path)
    {
        if (is_tif_name(path)) {
            if (!file_exists(path)) return (err = "input file not found: " + path), false;
            single = path;
            TiffFile tf(path);
            if (!tf.ok()) return (err = "cannot open " + path), false;
            std::vector<TiffImage> ifds;
            if (!tf.parse(ifds) || !tf.check(ifds[0])) return (err = tf.err), false;
            nx = (int)ifds[0].width;
            ny = (int)ifds[0].height;
            nz = (int)ifds.size();
            const size_t p = ifds[0].description.find("images=");
            if (ifds.size() == 1 && p != std::string::npos) {
                const int n = std::atoi(ifds[0].description.c_str() + p + 7);
                if (n > 1) {
                    nz = n;
                    imagej = true;
                }
            }
            return true;
        }
        if (!dir_exists(path)) return (err = "input directory not found: " + path), false;
        if (!list_tifs(path, files)) return (err = "cannot list " + path), false;
        if (files.empty()) return (err = "no tif file in " + path), false;
        TiffFile tf(files[0]);
        std::vector<TiffImage> ifds;
        if (!tf.ok() || !tf.parse(ifds) || !tf.check(ifds[0])) return (err = tf.ok() ? tf.err : "cannot open " + files[0]), false;
        nx = (int)ifds[0].width;
        ny = (int)ifds[0].height;
        nz = (int)files.size();
        return true;
    }

    // slices [first, last] into out (nx*ny*(last-first+1) floats)
    bool read(int first, int last, float *out)
    {
        const uint64_t ns = (uint64_t)nx * ny;
        if (!single.empty()) {
            TiffFile tf(single);
            std::vector<TiffImage> ifds;
            if (!tf.ok() || !tf.parse(ifds)) return (err = tf.err), false;
            for (int z = first; z <= last; z++) {
                const TiffImage &im = imagej ? ifds[0] : ifds[z];
                if (!tf.check(im)) return (err = tf.err), false;
                if ((int)im.width != nx || (int)im.height != ny) return (err = single + ": slices differ in size"), false;
                if (!tf.read_image(im, out + (uint64_t)(z - first) * ns, imagej ? (uint64_t)z : 0))
                    return (err = tf.err), false;
            }
            return true;
        }
        bool ok = true;
        std::string msg;
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = first; z <= last; z++) {
            TiffFile tf(files[z]);
            std::vector<TiffImage> ifds;
            bool good = tf.ok() && tf.parse(ifds) && tf.check(ifds[0]);
            if (good && ((int)ifds[0].width != nx || (int)ifds[0].height != ny)) {
                good = false;
                tf.err = files[z] + ": slice size differs from the first slice";
            }
            if (good) good = tf.read_image(ifds[0], out + (uint64_t)(z - first) * ns);
            if (!good) {
#pragma omp critical
                {
                    ok = false;
                    msg = tf.ok() ? tf.err : "cannot open " + files[z];
                }
            }
        }
        if (!ok) err = msg;
        return ok;
    }
};

// ---------------------------------------------------------------------------
// TIFF writing (little-endian, uncompressed, strips)
// ---------------------------------------------------------------------------
void put16(std::vector<uint8_t> &b, uint16_t v)
{
    b.push_back((uint8_t)(v & 0xff));
    b.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t> &b, uint32_t v)
{
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}

// uint16 sample of the 16-bit writer: std::min(std::max(0.f, round(x)), 65535.f)
uint16_t to_u16(float x)
{
    const float r = (float)std::round(x);
    const float lo = std::max(0.f, r);  // NaN -> 0, as std::max(0, NaN) returns its first argument
    return (uint16_t)std::min(lo, 65535.f);
}

bool write_slice(const std::string &file, const float *img, int width, int height, bool as16)
{
    const uint32_t bps = as16 ? 16 : 32, bpp = bps / 8;
    const uint32_t rps = std::max(1, 2048 / width);
    const uint32_t nstrips = (height + rps - 1) / rps;
    const uint64_t data_bytes = (uint64_t)width * height * bpp;
    if (data_bytes + 4096 + 8ull * nstrips > 0xffffffffull) return false;  // classic TIFF offsets are 32-bit
    std::vector<uint8_t> b;
    b.reserve(data_bytes + 256 + 8 * nstrips);
    b.push_back('I');
    b.push_back('I');
    put16(b, 42);
    const uint32_t ifd_off = (uint32_t)(8 + data_bytes + (data_bytes & 1));
    put32(b, ifd_off);
    // pixel data right after the header, as libtiff lays it out
    for (int y = 0; y < height; y++) {
        const float *row = img + (uint64_t)y * width;
        for (int x = 0; x < width; x++) {
            if (as16) put16(b, to_u16(row[x]));
            else {
                uint32_t u;
                std::memcpy(&u, &row[x], 4);
                put32(b, u);
            }
        }
    }
    if (data_bytes & 1) b.push_back(0);
    const int nent = 11;
    const uint32_t after_ifd = ifd_off + 2 + 12 * nent + 4;
    const uint32_t off_tab = after_ifd, cnt_tab = after_ifd + (nstrips > 1 ? 4 * nstrips : 0);
    auto entry = [&](uint16_t tag, uint16_t type, uint32_t count, uint32_t value) {
        put16(b, tag);
        put16(b, type);
        put32(b, count);
        if (type == 3 && count == 1) {
            put16(b, (uint16_t)value);
            put16(b, 0);
        } else
            put32(b, value);
    };
    put16(b, nent);
    entry(256, 4, 1, (uint32_t)width);
    entry(257, 4, 1, (uint32_t)height);
    entry(258, 3, 1, bps);
    entry(259, 3, 1, 1);  // no compression
    entry(262, 3, 1, 1);  // min-is-black
    entry(273, 4, nstrips, nstrips > 1 ? off_tab : 8);
    entry(277, 3, 1, 1);
    entry(278, 4, 1, rps);
    entry(279, 4, nstrips, nstrips > 1 ? cnt_tab : (uint32_t)data_bytes);
    entry(284, 3, 1, 1);                // contiguous
    entry(339, 3, 1, as16 ? 1u : 3u);   // unsigned integer / IEEE float
    put32(b, 0);                        // no next IFD
    if (nstrips > 1) {
        const uint64_t strip_bytes = (uint64_t)rps * width * bpp;
        for (uint32_t s = 0; s < nstrips; s++) put32(b, (uint32_t)(8 + s * strip_bytes));
        for (uint32_t s = 0; s < nstrips; s++)
            put32(b, (uint32_t)std::min<uint64_t>(strip_bytes, data_bytes - s * strip_bytes));
    }
    FILE *f = std::fopen(file.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return std::fclose(f) == 0 && ok;
}

// An 8-bit RGB image as Save2DTifImage_RGB writes it (hdcommunication.cpp:
// 1487-1560): interleaved samples, rows per strip 2048 / width, no
// compression.  rgb holds width*height*3 bytes.
bool write_rgb(const std::string &file, const uint8_t *rgb, int width, int height)
{
    const uint32_t rps = std::max(1, 2048 / width);
    const uint32_t nstrips = (height + rps - 1) / rps;
    const uint64_t data_bytes = (uint64_t)width * height * 3;
    if (data_bytes + 4096 + 8ull * nstrips > 0xffffffffull) return false;
    std::vector<uint8_t> b;
    b.reserve(data_bytes + 256 + 8 * nstrips);
    b.push_back('I');
    b.push_back('I');
    put16(b, 42);
    const uint32_t ifd_off = (uint32_t)(8 + data_bytes + (data_bytes & 1));
    put32(b, ifd_off);
    b.insert(b.end(), rgb, rgb + data_bytes);
    if (data_bytes & 1) b.push_back(0);
    const int nent = 11;
    const uint32_t after_ifd = ifd_off + 2 + 12 * nent + 4;
    const uint32_t bps_tab = after_ifd, fmt_tab = after_ifd + 6;
    const uint32_t off_tab = after_ifd + 12, cnt_tab = off_tab + (nstrips > 1 ? 4 * nstrips : 0);
    auto entry = [&](uint16_t tag, uint16_t type, uint32_t count, uint32_t value) {
        put16(b, tag);
        put16(b, type);
        put32(b, count);
        if (type == 3 && count == 1) {
            put16(b, (uint16_t)value);
            put16(b, 0);
        } else
            put32(b, value);
    };
    put16(b, nent);
    entry(256, 4, 1, (uint32_t)width);
    entry(257, 4, 1, (uint32_t)height);
    entry(258, 3, 3, bps_tab);  // 8, 8, 8
    entry(259, 3, 1, 1);        // no compression
    entry(262, 3, 1, 2);        // RGB
    entry(273, 4, nstrips, nstrips > 1 ? off_tab : 8);
    entry(277, 3, 1, 3);
    entry(278, 4, 1, rps);
    entry(279, 4, nstrips, nstrips > 1 ? cnt_tab : (uint32_t)data_bytes);
    entry(284, 3, 1, 1);         // contiguous
    entry(339, 3, 3, fmt_tab);   // unsigned integer x 3
    put32(b, 0);
    put16(b, 8);
    put16(b, 8);
    put16(b, 8);
    put16(b, 1);
    put16(b, 1);
    put16(b, 1);
    if (nstrips > 1) {
        const uint64_t strip_bytes = (uint64_t)rps * width * 3;
        for (uint32_t s = 0; s < nstrips; s++) put32(b, (uint32_t)(8 + s * strip_bytes));
        for (uint32_t s = 0; s < nstrips; s++)
            put32(b, (uint32_t)std::min<uint64_t>(strip_bytes, data_bytes - s * strip_bytes));
    }
    FILE *f = std::fopen(file.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return std::fclose(f) == 0 && ok;
}

// float channels to the 8-bit output of color_denoising.h:72-77: the value
// capped at 255 and truncated toward zero (the channels are non-negative
// there; a negative value is written as 0 here)
void to_rgb8(const float *planes, size_t n, uint8_t *rgb)
{
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            const float v = std::min(255.f, planes[c * n + i]);
            rgb[3 * i + c] = v > 0.f ? (uint8_t)v : (uint8_t)0;
        }
}

bool read_rgb_file(const std::string &file, int shape[2], std::vector<float> *planes, std::string &err)
{
    TiffFile f(file);
    if (!f.ok()) {
        err = file + ": cannot open";
        return false;
    }
    std::vector<TiffImage> ifds;
    if (!f.parse(ifds) || !f.check_rgb(ifds[0])) {
        err = f.err;
        return false;
    }
    shape[0] = (int)ifds[0].width;
    shape[1] = (int)ifds[0].height;
    if (!planes) return true;
    planes->resize((size_t)3 * shape[0] * shape[1]);
    if (!f.read_rgb(ifds[0], planes->data())) {
        err = f.err;
        return false;
    }
    return true;
}

bool write_stack(const std::string &dir, const std::string &name, const float *vol, int nx, int ny, int nz, int z0,
                 bool as16, std::string &err)
{
    if (!make_path(dir)) {
        err = "cannot create " + dir;
        return false;
    }
    bool ok = true;
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < nz; z++) {
        const std::string f = dir + "/" + name + zfill(z + z0, 4) + ".tif";
        if (!write_slice(f, vol + (uint64_t)z * nx * ny, nx, ny, as16)) {
#pragma omp critical
            {
                ok = false;
                err = "cannot write " + f;
            }
        }
    }
    return ok;
}

// iternlm_cpu.cpp:51-85: an iteration's output is complete when every slice
// file exists and none differs in size from the largest seen before it
bool iteration_complete(const std::string &dir, int first, int last)
{
    long long expected = 0;
    for (int i = first; i <= last; i++) {
        const std::string f = dir + "/denoised" + zfill(i, 4) + ".tif";
        struct stat st;
        if (stat(f.c_str(), &st) != 0) return false;
        const long long sz = (long long)st.st_size;
        if (sz > expected) expected = sz;
        else if (sz != expected) return false;
    }
    return true;
}

// slices first..last of a finished iteration, found by name (denoisedNNNN)
bool read_iteration(const std::string &dir, int first, int last, int nx, int ny, float *out, std::string &err)
{
    StackSource src;
    if (!src.open(dir)) {
        err = src.err;
        return false;
    }
    if (src.nx != nx || src.ny != ny) {
        err = dir + ": slices of another size";
        return false;
    }
    const std::string want = "denoised" + zfill(first, 4) + ".tif";
    int fs = -1;
    for (int i = 0; i < (int)src.files.size(); i++)
        if (has_ending(src.files[i], "/" + want)) fs = i;
    if (fs < 0 || fs + (last - first) >= (int)src.files.size()) {
        err = dir + ": slice files missing";
        return false;
    }
    if (!src.read(fs, fs + (last - first), out)) {
        err = src.err;
        return false;
    }
    return true;
}

bool write_noise_csv(const std::string &dir, const std::string &name, const float *s, int n)
{
    if (!make_path(dir)) return false;
    std::ofstream o(dir + "/" + name + ".csv");
    if (!o) return false;
    for (int i = 0; i < n; i++) o << i << "," << s[i] << "\n";  // default stream format: 6 significant digits
    return (bool)o;
}

// ReadNoisefile (hdcommunication.cpp:1747-1786): the second column of each
// line, accepted only when there is one line per slice
bool read_noise_csv(const std::string &file, int nz, float *s)
{
    std::ifstream in(file);
    if (!in) return false;
    std::vector<double> y;
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string a, b;
        std::getline(ss, a, ',');
        std::getline(ss, b, ',');
        y.push_back(std::atof(b.c_str()));
    }
    if ((int)y.size() != nz) return false;
    for (int i = 0; i < nz; i++) s[i] = (float)y[i];
    return true;
}

}  // namespace

extern "C" {

inlm_status inlm_tiff_stack_shape(const char *path, int32_t shape[3])
{
    if (!path || !shape) return io_fail("null argument");
    StackSource src;
    if (!src.open(path)) return io_fail(src.err);
    shape[0] = src.nx;
    shape[1] = src.ny;
    shape[2] = src.nz;
    return INLM_OK;
}

inlm_status inlm_tiff_read_stack(const char *path, int32_t first, int32_t last, float *out)
{
    if (!path || !out) return io_fail("null argument");
    StackSource src;
    if (!src.open(path)) return io_fail(src.err);
    if (first < 0 || last >= src.nz || first > last) {
        set_global_error("slice range outside the stack");
        return INLM_BAD_ARG;
    }
    if (!src.read(first, last, out)) return io_fail(src.err);
    return INLM_OK;
}

inlm_status inlm_tiff_write_stack(const char *dir, const char *name, const float *vol, const int32_t shape[3],
                                  int32_t z0, int32_t as16bit)
{
    if (!dir || !name || !vol || !shape) return io_fail("null argument");
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1 || z0 < 0) {
        set_global_error("shape must be positive and z0 non-negative");
        return INLM_BAD_ARG;
    }
    std::string err;
    if (!write_stack(dir, name, vol, shape[0], shape[1], shape[2], z0, as16bit != 0, err)) return io_fail(err);
    return INLM_OK;
}

inlm_status inlm_noise_csv_write(const char *dir, const char *name, const float *sigma, int32_t n)
{
    if (!dir || !name || (!sigma && n > 0)) return io_fail("null argument");
    if (!write_noise_csv(dir, name, sigma, n)) return io_fail(std::string("cannot write ") + dir + "/" + name + ".csv");
    return INLM_OK;
}

inlm_status inlm_noise_csv_read(const char *file, int32_t nz, float *sigma)
{
    if (!file || !sigma) return io_fail("null argument");
    if (!read_noise_csv(file, nz, sigma)) return io_fail(std::string(file) + ": missing or not one value per slice");
    return INLM_OK;
}

int32_t inlm_iteration_on_disk(const char *dir, int32_t first, int32_t last)
{
    if (!dir || first < 0 || last < first) return 0;
    return iteration_complete(dir, first, last) ? 1 : 0;
}

void inlm_default_io_options(inlm_io_options *o)
{
    std::memset(o, 0, sizeof(*o));
    o->firstslice = -1;
    o->lastslice = -1;
    o->save_16bit = 0;
    o->resume = 0;
    o->cleanup = 0;
    o->verbose = 0;
}

inlm_status inlm_denoise_files(const inlm_params *pin, int32_t device, const char *inpath_c, const char *outpath_c,
                               const inlm_io_options *oin)
{
    if (!pin || !inpath_c) return io_fail("null argument");
    inlm_io_options o;
    if (oin) o = *oin;
    else inlm_default_io_options(&o);
    inlm_params p = *pin;
    const std::string inpath = inpath_c;
    // main.cpp:213-218: output next to the input unless given; the log file
    // goes to the parent of the output
    std::string rootpath = inpath.substr(0, inpath.rfind('/', inpath.size() >= 2 ? inpath.size() - 2 : 0) + 1);
    std::string outpath = outpath_c ? outpath_c : "";
    if (outpath.empty()) outpath = rootpath + "/denoised/";
    else rootpath = outpath.substr(0, outpath.rfind('/', outpath.size() >= 2 ? outpath.size() - 2 : 0) + 1);

    StackSource src;
    if (!src.open(inpath)) return io_fail(src.err);
    int first = o.firstslice, last = o.lastslice;
    if (first != -1 || last != -1) {
        first = std::max(0, first);
        last = last < 0 ? src.nz - 1 : std::min(src.nz - 1, last);
    } else {
        first = 0;
        last = src.nz - 1;
    }
    if (first > last) {
        set_global_error("empty slice range");
        return INLM_BAD_ARG;
    }
    const int nx = src.nx, ny = src.ny, nz = last - first + 1;
    const int32_t shape[3] = {nx, ny, nz};
    const size_t nvox = (size_t)nx * ny * nz;
    std::vector<float> instack(nvox), output(nvox);
    if (!src.read(first, last, instack.data())) return io_fail(src.err);
    if (o.verbose)
        std::printf("iterNLM on %s: %d x %d x %d voxels\n", inpath.c_str(), nx, ny, nz);

    inlm_handle *h = nullptr;
    inlm_status s = inlm_create(&p, shape, device, &h);
    if (s != INLM_OK) return s;
    auto bail = [&](inlm_status st) {
        const std::string m = inlm_last_error(h);
        inlm_destroy(h);
        set_global_error(m);
        return st;
    };
    if ((s = inlm_load_raw(h, instack.data(), 0)) != INLM_OK) return bail(s);

    std::vector<int64_t> samples;
    if (p.noise_mode != INLM_NOISE_MANUAL) {
        samples.resize((size_t)p.n_samples);
        if ((s = inlm_draw_samples_mask(shape, p.n_samples, p.patchsize, p.circular_mask_diameter, p.sample_seed,
                                        samples.data())) != INLM_OK)
            return bail(s);
    }
    const int64_t nsmp = (int64_t)samples.size();
    const bool save_csv = !o.cleanup || o.verbose;
    std::vector<float> s0(nz), s1(nz), tmp(nz);
    // main.cpp:325-351: initial noise level (from the CSV when resuming)
    bool have = o.resume && read_noise_csv(outpath + "/noise_vs_slice00.csv", nz, s0.data());
    if (!have) {
        if ((s = inlm_estimate_noise(h, 0, samples.data(), nsmp, s0.data())) != INLM_OK) return bail(s);
        if (save_csv && !write_noise_csv(outpath, "noise_vs_slice00", s0.data(), nz))
            return bail(io_fail("cannot write the noise file"));
    }
    if ((s = inlm_set_sigma(h, s0.data(), nz)) != INLM_OK) return bail(s);
    if (o.verbose) std::printf("initial noise level (slice %d): %g\n", nz / 2, s0[nz / 2]);

    auto iter_dir = [&](int iter) {
        if (iter == p.maxiterations && o.cleanup) return outpath;
        return outpath + "/denoise-iteration" + zfill(iter, 2) + "/";
    };
    auto run = [&](int iter) -> inlm_status {
        const std::string dir = iter_dir(iter);
        inlm_status r;
        if (o.resume && iteration_complete(dir, first + 0, last)) {
            // a finished iteration is read back and becomes the result
            std::string err;
            if (!read_iteration(dir, first, last, nx, ny, output.data(), err)) return io_fail(err);
            if ((r = inlm_set_result(h, iter, output.data(), 0)) != INLM_OK) return r;
            if (o.verbose) std::printf("iteration %d read in from disk\n", iter);
            return INLM_OK;
        }
        if ((r = inlm_run_iteration(h, iter, nullptr)) != INLM_OK) return r;
        if (o.verbose) std::printf("iteration %d: filter %.1f ms\n", iter, inlm_last_filter_ms(h));
        if (!o.cleanup || iter == p.maxiterations) {
            if ((r = inlm_get_result(h, output.data(), 0)) != INLM_OK) return r;
            std::string err;
            if (!write_stack(dir, "denoised", output.data(), nx, ny, nz, first, o.save_16bit != 0, err))
                return io_fail(err);
        }
        return INLM_OK;
    };
    auto reestimate = [&](int idx) -> inlm_status {
        const std::string name = "noise_vs_slice" + zfill(idx, 2);
        bool got = o.resume && read_noise_csv(outpath + "/" + name + ".csv", nz, tmp.data());
        if (!got) {
            inlm_status r = inlm_estimate_noise(h, 1, samples.data(), nsmp, tmp.data());
            if (r != INLM_OK) return r;
            if (save_csv && !write_noise_csv(outpath, name, tmp.data(), nz)) return io_fail("cannot write the noise file");
        }
        for (int i = 0; i < nz; i++) s1[i] = p.alpha * s0[i] + (1.f - p.alpha) * tmp[i];
        return inlm_set_sigma(h, s1.data(), nz);
    };

    if ((s = run(1)) != INLM_OK) return bail(s);
    if (p.maxiterations > 1 && (s = reestimate(1)) != INLM_OK) return bail(s);
    for (int iter = 2; iter <= p.maxiterations; iter++) {
        if ((s = run(iter)) != INLM_OK) return bail(s);
        if (p.continuous_estimate && p.maxiterations > iter + 1 && (s = reestimate(iter)) != INLM_OK) return bail(s);
    }
    inlm_destroy(h);

    // main.cpp:548-561: intermediate iterations removed with cleanup
    if (o.cleanup)
        for (int iter = 1; iter < p.maxiterations; iter++) {
            const std::string d = outpath + "/denoise-iteration" + zfill(iter, 2);
            std::vector<std::string> fl;
            if (dir_exists(d) && list_tifs(d, fl)) {
                for (const auto &f : fl) std::remove(f.c_str());
                rmdir(d.c_str());
            }
        }
    // main.cpp:563-591: the appended log file
    {
        std::ofstream lf(rootpath + "/logfile.txt", std::ios::app);
        const time_t now = time(nullptr);
        lf << ctime(&now);
        lf << "ran iterNLM (MI355X) for " << inpath << "\n";
        lf << "-------------------------------------------------------------------------\n";
        const char *mode = p.noise_mode == INLM_NOISE_ZADAPTIVE   ? "z-adaptive"
                           : p.noise_mode == INLM_NOISE_GLOBAL    ? "global"
                           : p.noise_mode == INLM_NOISE_SEMIMANUAL ? "semimanual"
                                                                  : "manual";
        if (p.noise_mode == INLM_NOISE_ZADAPTIVE) {
            lf << "    noise estimate: z-adaptive (moving window " << p.zwindow << ")\n";
            lf << "    sigma0(slice " << nz / 2 << "): " << s0[nz / 2] << "\n";
            if (p.maxiterations > 1) lf << "    sigma1(slice " << nz / 2 << "): " << s1[nz / 2] << "\n";
        } else {
            lf << "    noise estimate: " << mode << "\n";
            lf << "    sigma0: " << s0[nz / 2] << "\n";
            if (p.maxiterations > 1) lf << "    sigma1: " << s1[nz / 2] << "\n";
        }
        lf << "    alpha:  " << p.alpha << "\n";
        lf << "    iterations: " << p.maxiterations << "\n";
        lf << "    slices: " << p.nslices << "\n";
        lf << "    patch space:  (" << p.radius_patchspace[0] << "," << p.radius_patchspace[1] << ","
           << p.radius_patchspace[2] << ")\n";
        lf << "    search space: (" << p.radius_searchspace[0] << "," << p.radius_searchspace[1] << ","
           << p.radius_searchspace[2] << ")\n";
        lf << "-------------------------------------------------------------------------\n\n";
    }
    return INLM_OK;
}

inlm_status inlm_rgb_tiff_shape(const char *file, int32_t shape[2])
{
    if (!file || !shape) return io_fail("null argument");
    int sh[2];
    std::string err;
    if (!read_rgb_file(file, sh, nullptr, err)) return io_fail(err);
    shape[0] = sh[0];
    shape[1] = sh[1];
    return INLM_OK;
}

inlm_status inlm_rgb_tiff_read(const char *file, float *planes)
{
    if (!file || !planes) return io_fail("null argument");
    int sh[2];
    std::vector<float> v;
    std::string err;
    if (!read_rgb_file(file, sh, &v, err)) return io_fail(err);
    std::memcpy(planes, v.data(), v.size() * sizeof(float));
    return INLM_OK;
}

inlm_status inlm_rgb_tiff_write(const char *file, const float *planes, const int32_t shape[2])
{
    if (!file || !planes || !shape || shape[0] < 1 || shape[1] < 1) return io_fail("bad argument");
    const size_t n = (size_t)shape[0] * shape[1];
    std::vector<uint8_t> rgb(3 * n);
    to_rgb8(planes, n, rgb.data());
    if (!write_rgb(file, rgb.data(), shape[0], shape[1])) return io_fail(std::string(file) + ": cannot write");
    return INLM_OK;
}

inlm_status inlm_denoise_rgb_files(const inlm_params *p, const inlm_rgb_options *copt, int32_t device,
                                   const char *inpath_c, const char *outpath_c, const inlm_io_options *oin)
{
    if (!p || !inpath_c) return io_fail("null argument");
    inlm_io_options o;
    if (oin) o = *oin;
    else inlm_default_io_options(&o);
    inlm_rgb_options c;
    if (copt) c = *copt;
    else inlm_default_rgb_options(&c);
    const std::string inpath = inpath_c;
    // main.cpp:213-218, as for volumes
    std::string rootpath = inpath.substr(0, inpath.rfind('/', inpath.size() >= 2 ? inpath.size() - 2 : 0) + 1);
    std::string outpath = outpath_c ? outpath_c : "";
    if (outpath.empty()) outpath = rootpath + "/denoised/";
    std::vector<std::string> files;
    if (dir_exists(inpath)) {
        if (!list_tifs(inpath, files) || files.empty()) return io_fail(inpath + ": no TIFF images");
    } else if (file_exists(inpath))
        files.push_back(inpath);
    else
        return io_fail(inpath + ": no such file or folder");
    if (!make_path(outpath)) return io_fail(outpath + ": cannot create");
    const std::string iterdir = outpath + "/denoise_iterations/";
    // color_denoising.h:19-150, one image after the other
    for (const std::string &file : files) {
        const size_t slash = file.rfind('/'), dot = file.rfind('.');
        const std::string name = file.substr(slash + 1, dot - slash - 1);
        int sh[2];
        std::vector<float> planes;
        std::string err;
        if (!read_rgb_file(file, sh, &planes, err)) return io_fail(err);
        const size_t n = (size_t)sh[0] * sh[1];
        std::vector<float> out(3 * n);
        std::vector<uint8_t> rgb(3 * n);
        const int32_t shape[2] = {sh[0], sh[1]};
        if (o.verbose) std::printf("image: %s (%d x %d)\n", name.c_str(), sh[0], sh[1]);
        auto on_iteration = [&](int iter, float *res) -> bool {
            if (o.cleanup) return true;
            // written through the inverse transform, then transformed back:
            // that rounding reaches the next iteration, as in the reference
            if (c.stabilize_variance) inlm::inverse_anscombe(res, 3 * n);
            to_rgb8(res, n, rgb.data());
            if (c.stabilize_variance) inlm::anscombe(res, 3 * n);
            if (!make_path(iterdir)) return false;
            return write_rgb(iterdir + name + "_iter" + zfill(iter, 2) + ".tif", rgb.data(), sh[0], sh[1]);
        };
        const inlm_status st = inlm::denoise_rgb(p, &c, shape, device, planes.data(), out.data(), nullptr, on_iteration);
        if (st != INLM_OK) return st;
        to_rgb8(out.data(), n, rgb.data());
        if (!write_rgb(outpath + "/" + name + "_denoised.tif", rgb.data(), sh[0], sh[1]))
            return io_fail(outpath + ": cannot write " + name + "_denoised.tif");
    }
    return INLM_OK;
}

}  // extern "C"
