// es_png.cpp — images of output_type "all" and a small PNG writer.
//
// The pixels follow Logger::WriteImageSection (src/Logger.cpp:278-318).  The
// file is a valid PNG with the same pixels as the reference's stb_image_write
// output; its IDAT holds stored (uncompressed) deflate blocks, so the bytes
// differ from stb's compressed stream while every decoder returns the same image.
#include "es_png.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

namespace es_host {

namespace {

// The reference converts double to unsigned char with a plain cast; its build
// (x64) truncates to a 32-bit integer and keeps the low byte, which is what
// values outside [0, 256) then give.
uint8_t to_u8(double x)
{
    int32_t v;
    if (std::isfinite(x) && x > -2147483649.0 && x < 2147483648.0)
        v = (int32_t)x;
    else
        v = std::numeric_limits<int32_t>::min();
    return (uint8_t)(v & 0xFF);
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t crc = 0xFFFFFFFFu)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
    return crc;
}

void put32(std::vector<uint8_t>& o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data)
{
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put32(out, crc32(td.data(), td.size()) ^ 0xFFFFFFFFu);
}

}  // namespace

void write_image_section(std::vector<uint8_t>& pixels, const double* values, int dim, int image_id, bool colored)
{
    const int image_width = dim * 3 * 3;  // bytes per row: 3 panels x RGB
    const int offset = image_id * dim * 3;
    double min_element = std::numeric_limits<double>::infinity();
    double max_element = -std::numeric_limits<double>::infinity();
    const size_t n = (size_t)dim * dim;
    for (size_t i = 0; i < n; i++) {
        const double k = values[i];
        if (k < min_element) min_element = k;
        if (k > max_element) max_element = k;
    }
    for (int j = 0; j < dim; j++)
        for (int i = 0; i < dim; i++) {
            const size_t idx = (size_t)j * image_width + (size_t)i * 3 + offset;
            const double v = values[(size_t)j * dim + i];
            if (!colored) {
                const uint8_t k = to_u8((v - min_element) / max_element * 255.0);
                pixels[idx] = k;
                pixels[idx + 1] = k;
                pixels[idx + 2] = k;
            } else if (v < 0) {
                pixels[idx] = 0;
                pixels[idx + 1] = 0;
                pixels[idx + 2] = to_u8(v / min_element * 255.0);
            } else {
                pixels[idx] = to_u8(v / max_element * 255.0);
                pixels[idx + 1] = 0;
                pixels[idx + 2] = 0;
            }
        }
}

bool write_png_rgb(const std::string& path, int width, int height, const std::vector<uint8_t>& pixels)
{
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)width);
    put32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, RGB, deflate, adaptive filters, no interlace
    chunk(png, "IHDR", ihdr);
    // raw scanlines, filter type 0 (None)
    const size_t row = (size_t)width * 3;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * height);
    for (int y = 0; y < height; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), pixels.begin() + (ptrdiff_t)(y * row), pixels.begin() + (ptrdiff_t)((y + 1) * row));
    }
    // zlib stream of stored deflate blocks (RFC 1950 / 1951)
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (ptrdiff_t)pos, raw.begin() + (ptrdiff_t)(pos + len));
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(png.data()), (std::streamsize)png.size());
    return (bool)f;
}

bool log_images(const std::string& path, int dim, const double* lifetimes, const double* sigma_xx,
                const double* sigma_xy)
{
    const int image_width = dim * 3, image_height = dim;
    std::vector<uint8_t> pixels((size_t)image_width * image_height * 3);
    write_image_section(pixels, lifetimes, dim, 0, false);
    write_image_section(pixels, sigma_xx, dim, 1, false);
    write_image_section(pixels, sigma_xy, dim, 2, true);
    return write_png_rgb(path, image_width, image_height, pixels);
}

}  // namespace es_host
