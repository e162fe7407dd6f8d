// mrh_image.cpp — image output (the EXR path of ImageIO::WriteImage,
// ImageIO/ImageIOI.h:83-91, ImageType::EXR) and PFM, written directly: the
// library carries no image-format dependency.
//
// EXR: OpenEXR 2.0 single-part scanline file, no compression, one scanline
// per block, channels B, G, R (sorted by name, as the format requires) in
// FP32, line order increasing y from the top row.  PFM: "PF" colour map,
// little-endian (scale -1), rows bottom to top as the format stores them.
// Input rows run bottom to top (the tracer's image layout).
#include "mrh.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

void put_bytes(std::vector<unsigned char>& out, const void* p, size_t n)
{
    const unsigned char* b = (const unsigned char*)p;
    out.insert(out.end(), b, b + n);
}
void put_i32(std::vector<unsigned char>& out, int32_t v) { put_bytes(out, &v, 4); }
void put_f32(std::vector<unsigned char>& out, float v) { put_bytes(out, &v, 4); }
void put_str(std::vector<unsigned char>& out, const char* s) { put_bytes(out, s, std::strlen(s) + 1); }

// attribute: name, type name, byte size, then the value
void attr_head(std::vector<unsigned char>& out, const char* name, const char* type, int32_t size)
{
    put_str(out, name);
    put_str(out, type);
    put_i32(out, size);
}

int32_t write_file(const char* path, const std::vector<unsigned char>& data)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return MRH_ERR_INVALID;
    const size_t n = std::fwrite(data.data(), 1, data.size(), f);
    const int c = std::fclose(f);
    return (n == data.size() && c == 0) ? MRH_OK : MRH_ERR_INVALID;
}

int32_t write_exr(const char* path, uint32_t w, uint32_t h, const float* rgb)
{
    std::vector<unsigned char> out;
    put_i32(out, 20000630);   // magic
    put_i32(out, 2);          // version 2, single-part scanline
    // channels: B, G, R, each FLOAT (2), not linear, sampling 1x1
    const char* names[3] = {"B", "G", "R"};
    attr_head(out, "channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* n : names) {
        put_str(out, n);
        put_i32(out, 2);
        const unsigned char lin_res[4] = {0, 0, 0, 0};
        put_bytes(out, lin_res, 4);
        put_i32(out, 1);
        put_i32(out, 1);
    }
    out.push_back(0);
    attr_head(out, "compression", "compression", 1);
    out.push_back(0);   // NO_COMPRESSION
    const int32_t box[4] = {0, 0, (int32_t)w - 1, (int32_t)h - 1};
    attr_head(out, "dataWindow", "box2i", 16);
    for (int32_t v : box) put_i32(out, v);
    attr_head(out, "displayWindow", "box2i", 16);
    for (int32_t v : box) put_i32(out, v);
    attr_head(out, "lineOrder", "lineOrder", 1);
    out.push_back(0);   // INCREASING_Y
    attr_head(out, "pixelAspectRatio", "float", 4);
    put_f32(out, 1.0f);
    attr_head(out, "screenWindowCenter", "v2f", 8);
    put_f32(out, 0.0f);
    put_f32(out, 0.0f);
    attr_head(out, "screenWindowWidth", "float", 4);
    put_f32(out, 1.0f);
    out.push_back(0);   // end of header
    // offset table, then one block per scanline: y, byte count, B row, G row, R row
    const uint64_t table = out.size();
    const uint64_t block = 8 + (uint64_t)w * 12;
    for (uint32_t y = 0; y < h; y++) {
        const uint64_t off = table + (uint64_t)h * 8 + (uint64_t)y * block;
        put_bytes(out, &off, 8);
    }
    for (uint32_t y = 0; y < h; y++) {
        put_i32(out, (int32_t)y);
        put_i32(out, (int32_t)(w * 12));
        const float* row = rgb + (size_t)(h - 1 - y) * w * 3;   // EXR's first line is the top row
        for (int c = 2; c >= 0; c--)
            for (uint32_t x = 0; x < w; x++) put_f32(out, row[3 * (size_t)x + c]);
    }
    return write_file(path, out);
}

int32_t write_pfm(const char* path, uint32_t w, uint32_t h, const float* rgb)
{
    std::vector<unsigned char> out;
    const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    put_bytes(out, head.data(), head.size());
    put_bytes(out, rgb, (size_t)w * h * 12);   // bottom row first, little endian
    return write_file(path, out);
}

}  // namespace

extern "C" int32_t mrh_image_write(const char* path, uint32_t type, uint32_t width, uint32_t height,
                                   const float* rgb)
{
    if (!path || !rgb || width == 0 || height == 0 || (uint64_t)width * height > (1ull << 31))
        return MRH_ERR_INVALID;
    if (type == MRH_IMAGE_EXR) return write_exr(path, width, height, rgb);
    if (type == MRH_IMAGE_PFM) return write_pfm(path, width, height, rgb);
    return MRH_ERR_INVALID;
}
