// ff_input.cpp -- the input side of the reference's fusion loop, host code.
//
// * TUM RGB-D association files and their poses: onlinefusionviewer_main.cpp:
//   266-346 reads, per line, "stamp tx ty tz qx qy qz qw stamp depth
//   [stamp rgb]" into floats and strings with operator>>, makes the rotation
//   with Eigen (Quaterniond(qw, qx, qy, qz).toRotationMatrix(), :322), and
//   composes every pose of a further association file with the last pose of
//   the previous one (startpos, :337-341).  Image names get the association
//   file's directory as prefix.
// * PNG images as cv::imread reads them (onlinefusionviewer.cpp:934-935):
//   the depth image unchanged (16-bit grayscale, IMREAD_UNCHANGED), the color
//   image as 8-bit BGR (IMREAD_COLOR: gray replicated, alpha dropped, 16-bit
//   samples reduced to their high byte).  OpenCV is not used; this is a PNG
//   reader of its own (ISO/IEC 15948: chunks, CRC, zlib stream, the five row
//   filters) over zlib's inflate.  Interlaced images and bit depths below 8
//   are rejected.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ff_abi.h"

namespace {

thread_local std::string g_err;

int fail(const std::string &what)
{
    g_err = what;
    return FF_E_ARG;
}

bool read_file(const char *path, std::vector<unsigned char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    unsigned char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    return true;
}

uint32_t be32(const unsigned char *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, channels = 0;
    std::vector<unsigned char> pix;      // unfiltered rows, w * channels * depth / 8 bytes each
    std::vector<unsigned char> palette;  // RGB triples (color type 3)
};

int channels_of(int ctype)
{
    switch (ctype) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
    default: return 0;
    }
}

// Paeth predictor (PNG specification, filter type 4)
inline int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

int decode_png(const char *path, Png &img, bool header_only = false)
{
    std::vector<unsigned char> f;
    if (!read_file(path, f)) return fail(std::string("cannot open ") + path);
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (f.size() < 8 || memcmp(f.data(), sig, 8)) return fail(std::string(path) + ": not a PNG file");
    size_t o = 8;
    std::vector<unsigned char> z;
    bool ihdr = false, iend = false;
    while (o + 12 <= f.size() && !iend) {
        const uint32_t len = be32(&f[o]);
        if (o + 12 + (size_t)len > f.size()) return fail(std::string(path) + ": truncated chunk");
        const unsigned char *type = &f[o + 4], *data = &f[o + 8];
        const uint32_t crc = be32(&f[o + 8 + len]);
        if ((uint32_t)crc32(0L, type, 4 + len) != crc) return fail(std::string(path) + ": chunk CRC mismatch");
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(std::string(path) + ": bad IHDR");
            img.w = be32(data);
            img.h = be32(data + 4);
            img.depth = data[8];
            img.ctype = data[9];
            if (data[10] != 0 || data[11] != 0) return fail(std::string(path) + ": unknown compression or filter method");
            if (data[12] != 0) return fail(std::string(path) + ": interlaced PNG not supported");
            img.channels = channels_of(img.ctype);
            if (!img.channels || (img.depth != 8 && img.depth != 16) || (img.ctype == 3 && img.depth != 8))
                return fail(std::string(path) + ": unsupported color type / bit depth");
            if (!img.w || !img.h || img.w > (1u << 15) || img.h > (1u << 15))
                return fail(std::string(path) + ": unsupported image size");
            ihdr = true;
            if (header_only) return FF_OK;
        } else if (!memcmp(type, "PLTE", 4)) {
            img.palette.assign(data, data + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            z.insert(z.end(), data, data + len);
        } else if (!memcmp(type, "IEND", 4)) {
            iend = true;
        }
        o += 12 + (size_t)len;
    }
    if (!ihdr) return fail(std::string(path) + ": no IHDR");
    const size_t bpp = (size_t)img.channels * img.depth / 8;  // bytes per pixel
    const size_t stride = bpp * img.w;
    std::vector<unsigned char> raw((stride + 1) * img.h);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return fail("inflateInit failed");
    zs.next_in = z.data();
    zs.avail_in = (uInt)z.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR) || got != raw.size())
        return fail(std::string(path) + ": corrupt or short image data");
    img.pix.assign(stride * img.h, 0);
    for (uint32_t y = 0; y < img.h; y++) {
        const unsigned char ft = raw[y * (stride + 1)];
        const unsigned char *src = &raw[y * (stride + 1) + 1];
        unsigned char *cur = &img.pix[y * stride];
        const unsigned char *up = y ? &img.pix[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return fail(std::string(path) + ": bad row filter");
            }
            cur[i] = (unsigned char)v;
        }
    }
    return FF_OK;
}

inline uint16_t sample(const Png &img, size_t p, int ch)
{
    const size_t i = p * img.channels + ch;
    return img.depth == 16 ? (uint16_t)(img.pix[2 * i] << 8 | img.pix[2 * i + 1]) : img.pix[i];
}

// ---- association files ---------------------------------------------------

// operator>> on a line: skips white space, then takes the longest prefix that
// reads as a decimal float (num_get's grammar); a failed extraction stores 0
// and fails every later extraction of the line, which then leaves its target
// unchanged
struct LineReader {
    const char *p;
    bool ok = true;
    explicit LineReader(const char *s) : p(s) {}
    void ws()
    {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') p++;
    }
    void num(float &out)
    {
        if (!ok) return;
        ws();
        const char *q = p;
        if (*q == '+' || *q == '-') q++;
        const char *d0 = q;
        while (*q >= '0' && *q <= '9') q++;
        bool digits = q > d0;
        if (*q == '.') {
            q++;
            const char *f0 = q;
            while (*q >= '0' && *q <= '9') q++;
            digits = digits || q > f0;
        }
        if (digits && (*q == 'e' || *q == 'E')) {
            const char *e = q + 1;
            if (*e == '+' || *e == '-') e++;
            if (*e >= '0' && *e <= '9') {
                while (*e >= '0' && *e <= '9') e++;
                q = e;
            }
        }
        if (!digits) {
            out = 0.0f;
            ok = false;
            return;
        }
        out = strtof(std::string(p, q).c_str(), nullptr);
        p = q;
    }
    void str(std::string &out)
    {
        if (!ok) return;
        ws();
        const char *q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n' && *q != '\v' && *q != '\f') q++;
        if (q == p) {
            ok = false;
            return;
        }
        out.assign(p, q);
        p = q;
    }
};

// Eigen's Quaternion::toRotationMatrix (Eigen 3, Geometry/Quaternion.h), in
// its order of operations, without normalisation
void quat_to_rot(double w, double x, double y, double z, double R[9])
{
    const double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x;
    const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0] = 1.0 - (tyy + tzz); R[1] = txy - twz;         R[2] = txz + twy;
    R[3] = txy + twz;         R[4] = 1.0 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy;         R[7] = tyz + twx;         R[8] = 1.0 - (txx + tyy);
}

// 4x4 extrinsic product A * B (each: R row-major, t), sums over k in order
void compose(const double Ra[9], const double ta[3], const double Rb[9], const double tb[3], double R[9], double t[3])
{
    double r[9], u[3];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += Ra[3 * i + k] * Rb[3 * k + j];
            r[3 * i + j] = s;  // row 3 of B is (0, 0, 0, 1): its term adds 0
        }
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += Ra[3 * i + k] * tb[k];
        u[i] = s + ta[i];
    }
    memcpy(R, r, sizeof r);
    memcpy(t, u, sizeof u);
}

}  // namespace

extern "C" {

const char *ff_input_last_error(void) { return g_err.c_str(); }

int ff_png_size(const char *path, int *width, int *height, int *channels, int *bitdepth)
{
    if (!path) return FF_E_ARG;
    Png img;
    const int r = decode_png(path, img, true);
    if (r) return r;
    if (width) *width = (int)img.w;
    if (height) *height = (int)img.h;
    if (channels) *channels = img.ctype == 3 ? 3 : img.channels;
    if (bitdepth) *bitdepth = img.depth;
    return FF_OK;
}

int ff_png_read_depth(const char *path, uint16_t *out, size_t n)
{
    if (!path || !out) return FF_E_ARG;
    Png img;
    const int r = decode_png(path, img);
    if (r) return r;
    if (img.ctype != 0) return fail(std::string(path) + ": a depth image must be single-channel grayscale");
    const size_t np = (size_t)img.w * img.h;
    if (n < np) return fail("ff_png_read_depth: output buffer too small");
    for (size_t p = 0; p < np; p++) out[p] = sample(img, p, 0);
    return FF_OK;
}

int ff_png_read_bgr(const char *path, uint8_t *out, size_t n)
{
    if (!path || !out) return FF_E_ARG;
    Png img;
    const int r = decode_png(path, img);
    if (r) return r;
    const size_t np = (size_t)img.w * img.h;
    if (n < 3 * np) return fail("ff_png_read_bgr: output buffer too small");
    const int sh = img.depth == 16 ? 8 : 0;
    for (size_t p = 0; p < np; p++) {
        uint8_t rgb[3];
        if (img.ctype == 3) {
            const size_t k = 3 * (size_t)img.pix[p];
            if (k + 2 >= img.palette.size()) return fail(std::string(path) + ": palette index out of range");
            rgb[0] = img.palette[k]; rgb[1] = img.palette[k + 1]; rgb[2] = img.palette[k + 2];
        } else if (img.ctype == 0 || img.ctype == 4) {
            rgb[0] = rgb[1] = rgb[2] = (uint8_t)(sample(img, p, 0) >> sh);
        } else {
            for (int ch = 0; ch < 3; ch++) rgb[ch] = (uint8_t)(sample(img, p, ch) >> sh);
        }
        out[3 * p] = rgb[2];
        out[3 * p + 1] = rgb[1];
        out[3 * p + 2] = rgb[0];
    }
    return FF_OK;
}

int ff_tum_read(const char *const *assoc_paths, int n_files, int use_color, ff_tum_frame *out, size_t cap)
{
    if (!assoc_paths || n_files < 0 || (cap && !out)) return FF_E_ARG;
    size_t count = 0;
    double startR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, startT[3] = {0, 0, 0};
    for (int fi = 0; fi < n_files; fi++) {
        const std::string path = assoc_paths[fi];
        const std::string prefix = path.substr(0, path.find_last_of('/') + 1);
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) return fail("cannot open association file " + path);
        std::string text;
        char buf[1 << 16];
        size_t k;
        while ((k = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
        fclose(f);
        // the reader's variables live across the lines of one file
        float stamp = 0, t1 = 0, t2 = 0, t3 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0;
        std::string depthname, rgbname;
        double lastR[9], lastT[3];
        bool any = false;
        size_t pos = 0;
        while (pos <= text.size()) {
            size_t nl = text.find('\n', pos);
            if (nl == std::string::npos) nl = text.size();
            const std::string line = text.substr(pos, nl - pos);
            pos = nl + 1;
            if (line.empty()) {
                if (nl == text.size()) break;
                continue;
            }
            LineReader rd(line.c_str());
            rd.num(stamp);
            rd.num(t1); rd.num(t2); rd.num(t3);
            rd.num(q1); rd.num(q2); rd.num(q3); rd.num(q4);
            rd.num(stamp);
            rd.str(depthname);
            if (use_color) {
                rd.num(stamp);
                rd.str(rgbname);
            }
            double R[9], T[3] = {(double)t1, (double)t2, (double)t3}, Rc[9], Tc[3];
            quat_to_rot((double)q4, (double)q1, (double)q2, (double)q3, R);
            compose(startR, startT, R, T, Rc, Tc);
            if (count < cap) {
                ff_tum_frame &fr = out[count];
                memcpy(fr.R, Rc, sizeof Rc);
                memcpy(fr.t, Tc, sizeof Tc);
                snprintf(fr.depth, sizeof fr.depth, "%s%s", prefix.c_str(), depthname.c_str());
                snprintf(fr.rgb, sizeof fr.rgb, "%s%s", use_color ? prefix.c_str() : "", use_color ? rgbname.c_str() : "");
            }
            memcpy(lastR, Rc, sizeof Rc);
            memcpy(lastT, Tc, sizeof Tc);
            any = true;
            count++;
            if (nl == text.size()) break;
        }
        if (any) {
            memcpy(startR, lastR, sizeof lastR);
            memcpy(startT, lastT, sizeof lastT);
        }
    }
    return (int)count;
}

}  // extern "C"
