// image.cpp -- the reference's image export (OutImage.cpp): a value map or a
// material map of a plane port, as 8-bit RGB pixels and as a PNG file.
// Host code only.  The PNG container is written here; zlib supplies the
// deflate stream and the chunk CRC.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <zlib.h>

#include "cpsfdtd.h"

namespace {

thread_local std::string g_image_error;

int fail(const std::string &msg)
{
    g_image_error = msg;
    return -1;
}

// Pixel maps larger than this per side are rejected: at 16384 x 16384 RGB
// the image data (805 MB, plus deflate's worst-case growth) stays inside one
// PNG chunk, whose length must be below 2^31, and inside zlib's 32-bit CRC
// length.
constexpr int kMaxSide = 1 << 14;

bool bad_dims(int nx, int ny) { return nx <= 0 || ny <= 0 || nx > kMaxSide || ny > kMaxSide; }

// Colour of each material ID, the table of OutImage.cpp:44-53 (ten entries).
constexpr int kNumMaterialColours = 10;
const unsigned char kMaterialColour[kNumMaterialColours][3] = {
    {254, 254, 254}, {254, 0, 0},   {0, 0, 254},   {0, 254, 0},   {0, 0, 0},
    {254, 254, 0},   {0, 254, 254}, {254, 0, 254}, {127, 0, 254}, {0, 127, 254},
};

void put_be32(std::vector<unsigned char> &b, uint32_t v)
{
    b.push_back((unsigned char)(v >> 24));
    b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8));
    b.push_back((unsigned char)v);
}

// One PNG chunk: length, type, data, CRC-32 over type and data.
void put_chunk(std::vector<unsigned char> &out, const char *type, const unsigned char *data, size_t n)
{
    put_be32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), out.data() + at, (uInt)(4 + n));
    put_be32(out, (uint32_t)crc);
}

}  // namespace

extern "C" {

const char *cfd_image_last_error(void) { return g_image_error.c_str(); }

// OutImage::plot_save / plot (OutImage.cpp:140-202, 76-138; thresholds
// OutImage.cpp:30-31).  Pixel (x, y) =
// map[x * ny + y] (the reference's dD[x][y], x along the image width).  At or
// below cmin: black; at or above cmax: (254, 254, 254); between, with
// t = (int)((v - cmin) * 254 / (cmax - cmin)), a third = (cmax + cmin) / 3 and
// two_third = (cmax + third) / 2 as the reference defines them:
// v <= third -> (t, 0, 0), v <= two_third -> (254, t, 0), else (254, 254, t).
// A NaN takes none of these branches and leaves its pixel as it was.
int cfd_image_colormap(const double *map, int nx, int ny, double cmin, double cmax, unsigned char *rgb)
{
    if (!map || !rgb) return fail("cfd_image_colormap: null buffer");
    if (bad_dims(nx, ny)) return fail("cfd_image_colormap: bad map size");
    const double third = (cmax + cmin) / 3.0;
    const double two_third = (cmax + third) / 2.0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const double v = map[(size_t)x * ny + y];
            unsigned char *p = rgb + ((size_t)y * nx + x) * 3;
            if (v <= cmin) {
                p[0] = p[1] = p[2] = 0;
            } else if (v >= cmax) {
                p[0] = p[1] = p[2] = 254;
            } else if (v == v) {
                const unsigned char t = (unsigned char)(int)((v - cmin) * 254.0 / (cmax - cmin));
                if (v <= third) {
                    p[0] = t; p[1] = 0; p[2] = 0;
                } else if (v <= two_third) {
                    p[0] = 254; p[1] = t; p[2] = 0;
                } else {
                    p[0] = 254; p[1] = 254; p[2] = t;
                }
            }
        }
    return 0;
}

// OutImage::structure_plot_save (OutImage.cpp:204-217): pixel (x, y) gets the
// colour of material ids[x * ny + y].  IDs outside the ten-entry table are an
// error (the reference reads past its table).
int cfd_image_materials(const short *ids, int nx, int ny, unsigned char *rgb)
{
    if (!ids || !rgb) return fail("cfd_image_materials: null buffer");
    if (bad_dims(nx, ny)) return fail("cfd_image_materials: bad map size");
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int id = ids[(size_t)x * ny + y];
            if (id < 0 || id >= kNumMaterialColours)
                return fail("cfd_image_materials: material ID " + std::to_string(id) + " has no colour (IDs 0-9 have)");
            unsigned char *p = rgb + ((size_t)y * nx + x) * 3;
            for (int c = 0; c < 3; ++c) p[c] = kMaterialColour[id][c];
        }
    return 0;
}

// The PNG the reference's CImg::save_png writes for these maps: 8-bit RGB,
// width x height, rows top to bottom, rgb[(y * width + x) * 3 + c].  Each row
// is stored with filter type 0; IDAT is one zlib stream.
int cfd_image_write_png(const char *path, int width, int height, const unsigned char *rgb)
{
    if (!path || !rgb) return fail("cfd_image_write_png: null argument");
    if (bad_dims(width, height)) return fail("cfd_image_write_png: bad image size");
    const size_t row = (size_t)width * 3 + 1;
    std::vector<unsigned char> raw(row * height);
    for (int y = 0; y < height; ++y) {
        raw[y * row] = 0;
        const unsigned char *src = rgb + (size_t)y * width * 3;
        std::copy(src, src + (size_t)width * 3, raw.begin() + y * row + 1);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return fail("cfd_image_write_png: deflate failed");

    static const unsigned char kSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> out(kSignature, kSignature + 8);
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    for (unsigned char b : {8, 2, 0, 0, 0}) ihdr.push_back(b);  // depth 8, RGB, deflate, filter 0, no interlace
    put_chunk(out, "IHDR", ihdr.data(), ihdr.size());
    put_chunk(out, "IDAT", z.data(), zlen);
    put_chunk(out, "IEND", nullptr, 0);

    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(std::string("cfd_image_write_png: cannot open ") + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) return fail(std::string("cfd_image_write_png: write failed: ") + path);
    return 0;
}

}  // extern "C"
