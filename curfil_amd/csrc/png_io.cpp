// PNG reading and writing over zlib: the image formats of the reference's
// datasets (VIGRA importImage / exportImage, image.cpp:109-121, 396-442,
// 582-645, 736-760): 8-bit RGB colour and ground-truth images, 16-bit
// grayscale depth images in millimetres.
//
// Reading covers every colour type and bit depth of the PNG specification
// (grayscale, RGB, palette, with or without alpha; 1-16 bits), interlaced or
// not, and reports the samples as they are stored (palette indices expanded
// to RGB).  Writing emits non-interlaced 8-bit RGB or 16-bit grayscale.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cf_host.h"

namespace cf {
namespace {

uint32_t be32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(static_cast<uint8_t>(x >> 24));
    v.push_back(static_cast<uint8_t>(x >> 16));
    v.push_back(static_cast<uint8_t>(x >> 8));
    v.push_back(static_cast<uint8_t>(x));
}

const uint8_t kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};

std::vector<uint8_t> readAll(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(CF_ERR_ARG, "cannot open '" + path + "'");
    std::vector<uint8_t> data;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + n);
    std::fclose(f);
    return data;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// reverses the per-row filters of one (sub-)image in place; rows are
// 1 filter byte + rowBytes; bpp = bytes per complete pixel (at least 1)
void unfilter(uint8_t* data, size_t rows, size_t rowBytes, size_t bpp, const std::string& path) {
    std::vector<uint8_t> zero(rowBytes, 0);
    const uint8_t* prev = zero.data();
    for (size_t r = 0; r < rows; r++) {
        uint8_t* row = data + r * (rowBytes + 1);
        const uint8_t type = row[0];
        uint8_t* cur = row + 1;
        switch (type) {
            case 0: break;
            case 1:
                for (size_t i = bpp; i < rowBytes; i++) cur[i] = static_cast<uint8_t>(cur[i] + cur[i - bpp]);
                break;
            case 2:
                for (size_t i = 0; i < rowBytes; i++) cur[i] = static_cast<uint8_t>(cur[i] + prev[i]);
                break;
            case 3:
                for (size_t i = 0; i < rowBytes; i++) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    cur[i] = static_cast<uint8_t>(cur[i] + ((left + prev[i]) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < rowBytes; i++) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    const int upLeft = i >= bpp ? prev[i - bpp] : 0;
                    cur[i] = static_cast<uint8_t>(cur[i] + paeth(left, prev[i], upLeft));
                }
                break;
            default: throw Error(CF_ERR_ARG, "'" + path + "': bad PNG filter type");
        }
        prev = cur;
    }
}

// sample k of a row of `depth`-bit samples
uint16_t sampleAt(const uint8_t* row, size_t k, int depth) {
    switch (depth) {
        case 16: return static_cast<uint16_t>((row[2 * k] << 8) | row[2 * k + 1]);
        case 8: return row[k];
        default: {
            const size_t bit = k * depth;
            const int shift = 8 - depth - static_cast<int>(bit & 7);
            return static_cast<uint16_t>((row[bit >> 3] >> shift) & ((1 << depth) - 1));
        }
    }
}

}  // namespace

PngImage readPng(const std::string& path) {
    const std::vector<uint8_t> file = readAll(path);
    if (file.size() < 8 || std::memcmp(file.data(), kSignature, 8) != 0) {
        throw Error(CF_ERR_ARG, "'" + path + "' is not a PNG file");
    }
    size_t pos = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, colorType = -1, interlace = 0;
    std::vector<uint8_t> idat, palette;
    bool seenHeader = false, seenEnd = false;
    while (pos + 12 <= file.size() && !seenEnd) {
        const uint32_t len = be32(&file[pos]);
        if (len > file.size() - pos - 12) throw Error(CF_ERR_ARG, "'" + path + "': truncated PNG chunk");
        const uint8_t* type = &file[pos + 4];
        const uint8_t* body = &file[pos + 8];
        const uint32_t crc = be32(&file[pos + 8 + len]);
        if (crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) throw Error(CF_ERR_ARG, "'" + path + "': PNG CRC error");
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) throw Error(CF_ERR_ARG, "'" + path + "': bad IHDR");
            W = be32(body);
            H = be32(body + 4);
            depth = body[8];
            colorType = body[9];
            interlace = body[12];
            if (body[10] != 0 || body[11] != 0 || interlace > 1) {
                throw Error(CF_ERR_ARG, "'" + path + "': unknown PNG compression, filter or interlace method");
            }
            seenHeader = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            palette.assign(body, body + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            seenEnd = true;
        } else if (!(type[0] & 0x20)) {
            throw Error(CF_ERR_ARG, "'" + path + "': unknown critical PNG chunk");
        }
        pos += 12 + len;
    }
    if (!seenHeader || !seenEnd) throw Error(CF_ERR_ARG, "'" + path + "': incomplete PNG file");
    if (W == 0 || H == 0 || W > 32767 || H > 32767) throw Error(CF_ERR_ARG, "'" + path + "': image size out of range");
    int channels;
    switch (colorType) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: throw Error(CF_ERR_ARG, "'" + path + "': bad PNG colour type");
    }
    const bool okDepth = (colorType == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
            (colorType == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
            ((colorType == 2 || colorType == 4 || colorType == 6) && (depth == 8 || depth == 16));
    if (!okDepth) throw Error(CF_ERR_ARG, "'" + path + "': bad PNG bit depth");
    if (colorType == 3 && (palette.empty() || palette.size() % 3 != 0)) {
        throw Error(CF_ERR_ARG, "'" + path + "': palette image without a palette");
    }

    // the (sub-)images: one, or the seven Adam7 passes
    struct Pass { int x0, y0, dx, dy; };
    const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass whole = {0, 0, 1, 1};
    const int numPasses = interlace ? 7 : 1;
    size_t rawSize = 0;
    std::vector<size_t> passW(numPasses), passH(numPasses), passRow(numPasses);
    for (int p = 0; p < numPasses; p++) {
        const Pass& ps = interlace ? adam7[p] : whole;
        passW[p] = W > static_cast<uint32_t>(ps.x0) ? (W - ps.x0 + ps.dx - 1) / ps.dx : 0;
        passH[p] = H > static_cast<uint32_t>(ps.y0) ? (H - ps.y0 + ps.dy - 1) / ps.dy : 0;
        passRow[p] = (passW[p] * channels * depth + 7) / 8;
        if (passW[p] && passH[p]) rawSize += passH[p] * (passRow[p] + 1);
    }
    std::vector<uint8_t> raw(rawSize);
    uLongf got = static_cast<uLongf>(rawSize);
    const int zr = uncompress(raw.data(), &got, idat.data(), static_cast<uLong>(idat.size()));
    if (zr != Z_OK || got != rawSize) throw Error(CF_ERR_ARG, "'" + path + "': corrupt PNG image data");

    PngImage img;
    img.width = static_cast<int>(W);
    img.height = static_cast<int>(H);
    img.bitDepth = depth == 16 ? 16 : 8;
    img.channels = colorType == 3 ? 3 : channels;
    img.colorType = colorType;
    img.samples.assign(static_cast<size_t>(W) * H * img.channels, 0);
    const size_t bpp = std::max<size_t>(1, static_cast<size_t>(channels) * depth / 8);
    size_t off = 0;
    for (int p = 0; p < numPasses; p++) {
        if (!passW[p] || !passH[p]) continue;
        const Pass& ps = interlace ? adam7[p] : whole;
        uint8_t* data = raw.data() + off;
        unfilter(data, passH[p], passRow[p], bpp, path);
        for (size_t r = 0; r < passH[p]; r++) {
            const uint8_t* row = data + r * (passRow[p] + 1) + 1;
            const size_t y = ps.y0 + r * ps.dy;
            for (size_t c = 0; c < passW[p]; c++) {
                const size_t x = ps.x0 + c * ps.dx;
                uint16_t* out = &img.samples[(y * W + x) * img.channels];
                if (colorType == 3) {
                    const size_t idx = sampleAt(row, c, depth);
                    if (3 * idx + 2 >= palette.size()) throw Error(CF_ERR_ARG, "'" + path + "': palette index out of range");
                    for (int k = 0; k < 3; k++) out[k] = palette[3 * idx + k];
                } else {
                    for (int k = 0; k < channels; k++) out[k] = sampleAt(row, c * channels + k, depth);
                }
            }
        }
        off += passH[p] * (passRow[p] + 1);
    }
    return img;
}

void writePng(const std::string& path, int W, int H, int channels, int bitDepth, const void* pixels) {
    if (W <= 0 || H <= 0) throw Error(CF_ERR_ARG, "image size must be positive");
    if (!((channels == 3 && bitDepth == 8) || (channels == 1 && (bitDepth == 8 || bitDepth == 16)))) {
        throw Error(CF_ERR_ARG, "PNG output is 8-bit RGB or 8/16-bit grayscale");
    }
    const size_t rowBytes = static_cast<size_t>(W) * channels * bitDepth / 8;
    std::vector<uint8_t> raw;
    raw.reserve((rowBytes + 1) * H);
    for (int y = 0; y < H; y++) {
        raw.push_back(0);   // filter type None
        if (bitDepth == 16) {
            const uint16_t* row = static_cast<const uint16_t*>(pixels) + static_cast<size_t>(y) * W;
            for (int x = 0; x < W; x++) {
                raw.push_back(static_cast<uint8_t>(row[x] >> 8));
                raw.push_back(static_cast<uint8_t>(row[x]));
            }
        } else {
            const uint8_t* row = static_cast<const uint8_t*>(pixels) + static_cast<size_t>(y) * rowBytes;
            raw.insert(raw.end(), row, row + rowBytes);
        }
    }
    uLongf zsize = compressBound(static_cast<uLong>(raw.size()));
    std::vector<uint8_t> z(zsize);
    if (compress2(z.data(), &zsize, raw.data(), static_cast<uLong>(raw.size()), 6) != Z_OK) {
        throw Error(CF_ERR_RUNTIME, "PNG compression failed");
    }
    z.resize(zsize);
    std::vector<uint8_t> out(kSignature, kSignature + 8);
    auto chunk = [&](const char* type, const std::vector<uint8_t>& body) {
        put32(out, static_cast<uint32_t>(body.size()));
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), body.begin(), body.end());
        put32(out, static_cast<uint32_t>(crc32(crc32(0L, Z_NULL, 0), &out[start], static_cast<uInt>(body.size() + 4))));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, static_cast<uint32_t>(W));
    put32(ihdr, static_cast<uint32_t>(H));
    ihdr.push_back(static_cast<uint8_t>(bitDepth));
    ihdr.push_back(channels == 3 ? 2 : 0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(CF_ERR_ARG, "cannot create '" + path + "'");
    const size_t n = std::fwrite(out.data(), 1, out.size(), f);
    const int rc = std::fclose(f);
    if (n != out.size() || rc != 0) throw Error(CF_ERR_RUNTIME, "cannot write '" + path + "'");
}

}  // namespace cf
