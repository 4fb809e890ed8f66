/*
 * image_io.cpp -- on-disk image formats for HDRIs and film output, with no
 * third-party library (the reference uses stb_image and FreeImage,
 * src/delegate/imageutil.cpp:13-139).
 *
 *   Radiance RGBE (.hdr)  read: flat or new-style run-length scanlines,
 *                         orientation "-Y H +X W", value = c * 2^(e - 136),
 *                         e = 0 -> black, alpha 1 (the conventions of the
 *                         reference's loader, imageutil.cpp:98-139);
 *                         write: run-length scanlines, Ward's encoding
 *                         (frexp mantissa * 256 / max component).
 *   PFM (.pfm)            read/write: "PF", little-endian float32 RGB,
 *                         scale -1, rows stored bottom to top.
 *   OpenEXR (.exr)        write: single-part scanline file, no compression,
 *                         FLOAT channels A, B, G, R, one line per block,
 *                         line 0 = top row (the reference writes RGBA float
 *                         EXR through FreeImage, imageutil.cpp:31-79).
 *
 * Images in memory are RGBA float32, row 0 at the top.
 */
#include "../../include/clhip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace
{
thread_local std::string g_io_error;

int io_fail(const std::string& msg)
{
    g_io_error = msg;
    return CLHIP_E_INVALID;
}

struct File
{
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f)
            std::fclose(f);
    }
};

bool read_line(FILE* f, std::string& line)
{
    line.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF)
    {
        if (c == '\n')
            return true;
        line.push_back((char)c);
        if (line.size() > 4096)
            return false;
    }
    return !line.empty();
}

void rgbe_to_float(const unsigned char* e, float* out)
{
    if (e[3] == 0)
    {
        out[0] = out[1] = out[2] = 0.0f;
    }
    else
    {
        const float f = std::ldexp(1.0f, (int)e[3] - 136);
        out[0]        = (float)e[0] * f;
        out[1]        = (float)e[1] * f;
        out[2]        = (float)e[2] * f;
    }
    out[3] = 1.0f;
}

void float_to_rgbe(const float* c, unsigned char* e)
{
    float v = std::fmax(c[0], std::fmax(c[1], c[2]));
    if (!(v >= 1e-32f))
    {
        e[0] = e[1] = e[2] = e[3] = 0;
        return;
    }
    int         ex  = 0;
    const float m   = std::frexp(v, &ex);
    const float scl = m * 256.0f / v;
    for (int k = 0; k < 3; ++k)
    {
        const float x = std::fmax(c[k], 0.0f) * scl;
        e[k]          = (unsigned char)std::fmin(x, 255.0f);
    }
    e[3] = (unsigned char)(ex + 128);
}

/* read the header and resolution line; leaves f at the pixel data */
int hdr_header(FILE* f, uint32_t* w, uint32_t* h)
{
    std::string line;
    if (!read_line(f, line) || (line.rfind("#?RADIANCE", 0) != 0 && line.rfind("#?RGBE", 0) != 0))
        return io_fail("not a Radiance HDR file (no #?RADIANCE / #?RGBE signature)");
    bool fmt_ok = true;
    for (;;)
    {
        if (!read_line(f, line))
            return io_fail("truncated HDR header");
        if (line.empty())
            break;
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe")
            fmt_ok = false;
    }
    if (!fmt_ok)
        return io_fail("unsupported HDR FORMAT (only 32-bit_rle_rgbe)");
    if (!read_line(f, line))
        return io_fail("missing HDR resolution line");
    int hh = 0, ww = 0;
    if (std::sscanf(line.c_str(), "-Y %d +X %d", &hh, &ww) != 2 || hh <= 0 || ww <= 0)
        return io_fail("unsupported HDR orientation (only \"-Y H +X W\"): " + line);
    *w = (uint32_t)ww;
    *h = (uint32_t)hh;
    return CLHIP_OK;
}

int hdr_pixels(FILE* f, uint32_t w, uint32_t h, float* rgba)
{
    std::vector<unsigned char> line(4 * (size_t)w);
    for (uint32_t y = 0; y < h; ++y)
    {
        unsigned char head[4];
        if (std::fread(head, 1, 4, f) != 4)
            return io_fail("truncated HDR pixel data");
        const bool rle = w >= 8 && w < 32768 && head[0] == 2 && head[1] == 2 && (head[2] & 0x80) == 0;
        if (rle)
        {
            if (((uint32_t)head[2] << 8 | head[3]) != w)
                return io_fail("HDR scanline width mismatch");
            for (int k = 0; k < 4; ++k)
            {
                uint32_t x = 0;
                while (x < w)
                {
                    int c = std::fgetc(f);
                    if (c == EOF)
                        return io_fail("truncated HDR run");
                    if (c > 128)
                    {
                        const int n = c - 128;
                        const int v = std::fgetc(f);
                        if (v == EOF || x + (uint32_t)n > w)
                            return io_fail("bad HDR run");
                        for (int i = 0; i < n; ++i)
                            line[4 * (size_t)(x++) + k] = (unsigned char)v;
                    }
                    else
                    {
                        if (c == 0 || x + (uint32_t)c > w)
                            return io_fail("bad HDR literal run");
                        for (int i = 0; i < c; ++i)
                        {
                            const int v = std::fgetc(f);
                            if (v == EOF)
                                return io_fail("truncated HDR literal");
                            line[4 * (size_t)(x++) + k] = (unsigned char)v;
                        }
                    }
                }
            }
        }
        else
        {
            /* flat quads; the first one was read as the "header" */
            std::memcpy(line.data(), head, 4);
            if (w > 1 && std::fread(line.data() + 4, 1, 4 * ((size_t)w - 1), f) != 4 * ((size_t)w - 1))
                return io_fail("truncated flat HDR scanline");
        }
        for (uint32_t x = 0; x < w; ++x)
            rgbe_to_float(&line[4 * (size_t)x], rgba + 4 * ((size_t)y * w + x));
    }
    return CLHIP_OK;
}

void put_u32(std::vector<unsigned char>& b, uint32_t v)
{
    for (int k = 0; k < 4; ++k)
        b.push_back((unsigned char)(v >> (8 * k)));
}
void put_f32(std::vector<unsigned char>& b, float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    put_u32(b, u);
}
void put_str(std::vector<unsigned char>& b, const char* s)
{
    b.insert(b.end(), s, s + std::strlen(s) + 1);
}
void put_attr(std::vector<unsigned char>& b, const char* name, const char* type, const std::vector<unsigned char>& v)
{
    put_str(b, name);
    put_str(b, type);
    put_u32(b, (uint32_t)v.size());
    b.insert(b.end(), v.begin(), v.end());
}
} // namespace

extern "C" {

const char* clhip_image_last_error(void) { return g_io_error.c_str(); }

int clhip_image_load_hdr(const char* path, float* rgba, size_t bytes, uint32_t* w, uint32_t* h)
{
    if (!path || !w || !h)
        return io_fail("clhip_image_load_hdr: bad arguments");
    File f(path, "rb");
    if (!f.f)
        return io_fail(std::string("cannot open ") + path);
    int rc = hdr_header(f.f, w, h);
    if (rc || !rgba)
        return rc;
    if (bytes < (size_t)*w * *h * 16)
        return io_fail("clhip_image_load_hdr: output buffer too small");
    return hdr_pixels(f.f, *w, *h, rgba);
}

int clhip_image_save_hdr(const char* path, const float* rgba, uint32_t w, uint32_t h)
{
    if (!path || !rgba || !w || !h)
        return io_fail("clhip_image_save_hdr: bad arguments");
    File f(path, "wb");
    if (!f.f)
        return io_fail(std::string("cannot create ") + path);
    std::fprintf(f.f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    std::vector<unsigned char> q(4 * (size_t)w), out;
    for (uint32_t y = 0; y < h; ++y)
    {
        for (uint32_t x = 0; x < w; ++x)
            float_to_rgbe(rgba + 4 * ((size_t)y * w + x), &q[4 * (size_t)x]);
        out.clear();
        if (w < 8 || w >= 32768)
            out = q;
        else
        {
            out.push_back(2);
            out.push_back(2);
            out.push_back((unsigned char)(w >> 8));
            out.push_back((unsigned char)(w & 255));
            for (int k = 0; k < 4; ++k)
            {
                uint32_t x = 0;
                while (x < w)
                {
                    /* a run of >= 3 equal bytes, else a literal up to the next such run */
                    uint32_t r = 1;
                    while (x + r < w && r < 127 && q[4 * (size_t)(x + r) + k] == q[4 * (size_t)x + k])
                        ++r;
                    if (r >= 3)
                    {
                        out.push_back((unsigned char)(128 + r));
                        out.push_back(q[4 * (size_t)x + k]);
                        x += r;
                        continue;
                    }
                    uint32_t n = 0;
                    while (x + n < w && n < 128)
                    {
                        if (x + n + 2 < w && q[4 * (size_t)(x + n) + k] == q[4 * (size_t)(x + n + 1) + k] &&
                            q[4 * (size_t)(x + n) + k] == q[4 * (size_t)(x + n + 2) + k])
                            break;
                        ++n;
                    }
                    out.push_back((unsigned char)n);
                    for (uint32_t i = 0; i < n; ++i)
                        out.push_back(q[4 * (size_t)(x + i) + k]);
                    x += n;
                }
            }
        }
        if (std::fwrite(out.data(), 1, out.size(), f.f) != out.size())
            return io_fail("write failed");
    }
    return CLHIP_OK;
}

int clhip_image_save_pfm(const char* path, const float* rgba, uint32_t w, uint32_t h)
{
    if (!path || !rgba || !w || !h)
        return io_fail("clhip_image_save_pfm: bad arguments");
    File f(path, "wb");
    if (!f.f)
        return io_fail(std::string("cannot create ") + path);
    std::fprintf(f.f, "PF\n%u %u\n-1.0\n", w, h);
    std::vector<float> row(3 * (size_t)w);
    for (uint32_t y = h; y-- > 0;)
    {
        for (uint32_t x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k)
                row[3 * (size_t)x + k] = rgba[4 * ((size_t)y * w + x) + k];
        if (std::fwrite(row.data(), sizeof(float), row.size(), f.f) != row.size())
            return io_fail("write failed");
    }
    return CLHIP_OK;
}

int clhip_image_load_pfm(const char* path, float* rgba, size_t bytes, uint32_t* w, uint32_t* h)
{
    if (!path || !w || !h)
        return io_fail("clhip_image_load_pfm: bad arguments");
    File f(path, "rb");
    if (!f.f)
        return io_fail(std::string("cannot open ") + path);
    char   tag[3] = {0, 0, 0};
    int    ww = 0, hh = 0;
    double scale = 0.0;
    if (std::fscanf(f.f, "%2s %d %d %lf", tag, &ww, &hh, &scale) != 4 || std::strcmp(tag, "PF") != 0 || ww <= 0 ||
        hh <= 0 || scale >= 0.0)
        return io_fail("only little-endian colour PFM (\"PF\", negative scale) is supported");
    std::fgetc(f.f); /* the single whitespace after the scale */
    *w = (uint32_t)ww;
    *h = (uint32_t)hh;
    if (!rgba)
        return CLHIP_OK;
    if (bytes < (size_t)*w * *h * 16)
        return io_fail("clhip_image_load_pfm: output buffer too small");
    std::vector<float> row(3 * (size_t)*w);
    for (uint32_t y = *h; y-- > 0;)
    {
        if (std::fread(row.data(), sizeof(float), row.size(), f.f) != row.size())
            return io_fail("truncated PFM data");
        for (uint32_t x = 0; x < *w; ++x)
        {
            float* p = rgba + 4 * ((size_t)y * *w + x);
            p[0]     = row[3 * (size_t)x];
            p[1]     = row[3 * (size_t)x + 1];
            p[2]     = row[3 * (size_t)x + 2];
            p[3]     = 1.0f;
        }
    }
    return CLHIP_OK;
}

int clhip_image_save_exr(const char* path, const float* rgba, uint32_t w, uint32_t h)
{
    if (!path || !rgba || !w || !h || w > (1u << 30) / 16 || h > (1u << 30))
        return io_fail("clhip_image_save_exr: bad arguments");
    std::vector<unsigned char> b;
    put_u32(b, 20000630u); /* magic 76 2f 31 01 */
    put_u32(b, 2u);        /* version 2, single-part scanline */
    std::vector<unsigned char> v;
    for (const char* ch : {"A", "B", "G", "R"})
    {
        put_str(v, ch);
        put_u32(v, 2u);             /* FLOAT */
        v.insert(v.end(), 4, 0);    /* pLinear + reserved */
        put_u32(v, 1u);             /* x sampling */
        put_u32(v, 1u);             /* y sampling */
    }
    v.push_back(0);
    put_attr(b, "channels", "chlist", v);
    put_attr(b, "compression", "compression", {0});
    v.clear();
    put_u32(v, 0u);
    put_u32(v, 0u);
    put_u32(v, w - 1);
    put_u32(v, h - 1);
    put_attr(b, "dataWindow", "box2i", v);
    put_attr(b, "displayWindow", "box2i", v);
    put_attr(b, "lineOrder", "lineOrder", {0});
    v.clear();
    put_f32(v, 1.0f);
    put_attr(b, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32(v, 0.0f);
    put_f32(v, 0.0f);
    put_attr(b, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32(v, 1.0f);
    put_attr(b, "screenWindowWidth", "float", v);
    b.push_back(0); /* end of header */
    const uint64_t line_bytes = 8 + 16 * (uint64_t)w;
    const uint64_t first      = b.size() + 8 * (uint64_t)h;
    for (uint32_t y = 0; y < h; ++y)
    {
        const uint64_t off = first + y * line_bytes;
        put_u32(b, (uint32_t)off);
        put_u32(b, (uint32_t)(off >> 32));
    }
    File f(path, "wb");
    if (!f.f)
        return io_fail(std::string("cannot create ") + path);
    if (std::fwrite(b.data(), 1, b.size(), f.f) != b.size())
        return io_fail("write failed");
    std::vector<unsigned char> line;
    for (uint32_t y = 0; y < h; ++y)
    {
        line.clear();
        put_u32(line, y);
        put_u32(line, 16u * w);
        for (int ch : {3, 2, 1, 0}) /* A, B, G, R */
            for (uint32_t x = 0; x < w; ++x)
                put_f32(line, rgba[4 * ((size_t)y * w + x) + ch]);
        if (std::fwrite(line.data(), 1, line.size(), f.f) != line.size())
            return io_fail("write failed");
    }
    return CLHIP_OK;
}

} // extern "C"
