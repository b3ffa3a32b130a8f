// 8-bit TIFF file output: the "file"/"tiff" display driver's integer branch
// (tools/displays/file/display.cpp:227-350, WriteTIFF), written without
// libtiff.  Baseline TIFF 6.0, little-endian, one IFD, contiguous samples,
// strips of libtiff's default height (TIFFDefaultStripSize with no request:
// as many rows as fit in 8 KiB, at least one), uncompressed or LZW (the
// driver's default compression, display.cpp:73).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/aqsis_hider.h"

namespace rh {
void set_last_error(const std::string& msg);
}

namespace {

enum { T_ASCII = 2, T_SHORT = 3, T_LONG = 4, T_RATIONAL = 5, T_DOUBLE = 12 };

struct Entry {
    uint16_t tag, type;
    uint32_t count;
    std::vector<uint8_t> data;  // little-endian values
};

void put16(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)x);
    v.push_back((uint8_t)(x >> 8));
}
void put32(std::vector<uint8_t>& v, uint32_t x) {
    for (int i = 0; i < 4; ++i) v.push_back((uint8_t)(x >> (8 * i)));
}
void set32(std::vector<uint8_t>& v, size_t at, uint32_t x) {
    for (int i = 0; i < 4; ++i) v[at + i] = (uint8_t)(x >> (8 * i));
}

Entry shorts(uint16_t tag, const std::vector<uint32_t>& xs) {
    Entry e{tag, T_SHORT, (uint32_t)xs.size(), {}};
    for (uint32_t x : xs) put16(e.data, x);
    return e;
}
Entry longs(uint16_t tag, const std::vector<uint32_t>& xs) {
    Entry e{tag, T_LONG, (uint32_t)xs.size(), {}};
    for (uint32_t x : xs) put32(e.data, x);
    return e;
}
Entry rational(uint16_t tag, uint32_t num, uint32_t den) {
    Entry e{tag, T_RATIONAL, 1, {}};
    put32(e.data, num);
    put32(e.data, den);
    return e;
}
Entry ascii(uint16_t tag, const std::string& s) {
    Entry e{tag, T_ASCII, (uint32_t)s.size() + 1, {}};
    e.data.assign(s.begin(), s.end());
    e.data.push_back(0);
    return e;
}

// TIFF LZW (TIFF 6.0 section 13): codes packed most significant bit first,
// 256 = Clear, 257 = EndOfInformation, new strings from 258, code width
// 9..12 bits.  The width grows as soon as the next free code no longer fits
// (the decoder, one string behind, switches one code early), the table is
// reset with a Clear (at the old width) when the next free code would be
// 4094, and after the strip's last string the free code is advanced once
// more before EndOfInformation is written, so that both sides agree on its
// width.  Each strip starts with a Clear.
class LzwWriter {
public:
    explicit LzwWriter(std::vector<uint8_t>& out) : out_(out) {}

    void encode(const uint8_t* p, size_t n) {
        reset_table();
        put(256);
        if (n == 0) { put(257); flush(); return; }
        int prefix = p[0];
        for (size_t i = 1; i < n; ++i) {
            const uint8_t c = p[i];
            const uint32_t key = ((uint32_t)prefix << 8) | c;
            int code = find(key);
            if (code >= 0) { prefix = code; continue; }
            put((uint32_t)prefix);
            insert(key, next_++);
            if (next_ == 4094) {
                put(256);
                reset_table();
            } else if (next_ > maxcode_) {
                ++nbits_;
                maxcode_ = (1 << nbits_) - 1;
            }
            prefix = c;
        }
        put((uint32_t)prefix);
        ++next_;
        if (next_ == 4094) {
            put(256);
            nbits_ = 9;
        } else if (next_ > maxcode_) {
            ++nbits_;
        }
        put(257);
        flush();
    }

private:
    static constexpr int kHash = 1 << 13;  // open addressing, load <= 0.5
    std::vector<uint8_t>& out_;
    std::vector<uint32_t> keys_ = std::vector<uint32_t>(kHash);
    std::vector<int16_t> codes_ = std::vector<int16_t>(kHash);
    int next_ = 258, nbits_ = 9, maxcode_ = 511;
    uint64_t acc_ = 0;
    int nacc_ = 0;

    void reset_table() {
        std::fill(codes_.begin(), codes_.end(), (int16_t)-1);
        next_ = 258;
        nbits_ = 9;
        maxcode_ = 511;
    }
    static uint32_t slot(uint32_t key) { return (key * 2654435761u) >> 19; }
    int find(uint32_t key) const {
        for (uint32_t h = slot(key);; h = (h + 1) & (kHash - 1)) {
            if (codes_[h] < 0) return -1;
            if (keys_[h] == key) return codes_[h];
        }
    }
    void insert(uint32_t key, int code) {
        uint32_t h = slot(key);
        while (codes_[h] >= 0) h = (h + 1) & (kHash - 1);
        keys_[h] = key;
        codes_[h] = (int16_t)code;
    }
    void put(uint32_t code) {
        acc_ = (acc_ << nbits_) | code;
        nacc_ += nbits_;
        while (nacc_ >= 8) {
            out_.push_back((uint8_t)(acc_ >> (nacc_ - 8)));
            nacc_ -= 8;
        }
        acc_ &= (1ull << nacc_) - 1;
    }
    void flush() {
        if (nacc_ > 0) out_.push_back((uint8_t)(acc_ << (8 - nacc_)));
        acc_ = 0;
        nacc_ = 0;
    }
};

}  // namespace

namespace {

Entry one_double(uint16_t tag, double x) {
    Entry e{tag, T_DOUBLE, 1, std::vector<uint8_t>(8)};
    std::memcpy(e.data.data(), &x, 8);   // little-endian host
    return e;
}

// One image of `channels` samples of `bits` bits per pixel.  sample_format
// 0 writes the 8-bit branch's tags (no SampleFormat, no StoNits); 1 (uint)
// or 3 (IEEE float) the other branch's.
rh_status write_tiff(const char* path, const uint8_t* pixels, int32_t width, int32_t height, int32_t channels,
                     int bits, int sample_format, int32_t compression, const int32_t* origin,
                     const int32_t* full_size, const char* description) {
    if (!path || !pixels) { rh::set_last_error("null argument"); return RH_ERR_INVALID; }
    if (width <= 0 || height <= 0 || (channels != 3 && channels != 4)) {
        rh::set_last_error("rh_tiff_write: width and height must be positive, channels 3 (rgb) or 4 (rgba)");
        return RH_ERR_INVALID;
    }
    if (compression != RH_TIFF_NONE && compression != RH_TIFF_LZW) {
        rh::set_last_error("rh_tiff_write: compression must be RH_TIFF_NONE or RH_TIFF_LZW");
        return RH_ERR_INVALID;
    }
    const int32_t ox = origin ? origin[0] : 0, oy = origin ? origin[1] : 0;
    const int32_t fw = full_size ? full_size[0] : width, fh = full_size ? full_size[1] : height;
    if (ox < 0 || oy < 0 || fw <= 0 || fh <= 0) {
        rh::set_last_error("rh_tiff_write: origin must be non-negative and the full size positive");
        return RH_ERR_INVALID;
    }
    const size_t line = (size_t)width * channels * (bits / 8);
    if ((uint64_t)line * height > 0xF0000000ull) {
        rh::set_last_error("rh_tiff_write: image exceeds a classic TIFF file");
        return RH_ERR_CAPACITY;
    }
    const uint32_t rowsPerStrip = line >= 8192 ? 1u : (uint32_t)(8192 / line);
    const uint32_t nstrips = ((uint32_t)height + rowsPerStrip - 1) / rowsPerStrip;

    std::vector<uint8_t> file;
    file.reserve(16 + line * height + 1024);
    file.insert(file.end(), {'I', 'I'});
    put16(file, 42);
    put32(file, 0);  // IFD offset, set below
    std::vector<uint32_t> offsets(nstrips), counts(nstrips);
    for (uint32_t s = 0; s < nstrips; ++s) {
        const uint32_t r0 = s * rowsPerStrip;
        const uint32_t rows = std::min<uint32_t>(rowsPerStrip, (uint32_t)height - r0);
        const uint8_t* src = pixels + (size_t)r0 * line;
        offsets[s] = (uint32_t)file.size();
        if (compression == RH_TIFF_LZW) {
            LzwWriter(file).encode(src, (size_t)rows * line);
        } else {
            file.insert(file.end(), src, src + (size_t)rows * line);
        }
        counts[s] = (uint32_t)file.size() - offsets[s];
        if (file.size() & 1) file.push_back(0);
    }

    char datetime[64];
    std::time_t now = std::time(nullptr);
    std::tm tmv{};
    localtime_r(&now, &tmv);
    std::snprintf(datetime, sizeof datetime, "%04d:%02d:%02d %02d:%02d:%02d", 1900 + tmv.tm_year, tmv.tm_mon + 1,
                  tmv.tm_mday, tmv.tm_hour, tmv.tm_min, tmv.tm_sec);

    // Tags in ascending order, as an IFD requires.
    std::vector<Entry> tags;
    tags.push_back(longs(256, {(uint32_t)width}));                       // ImageWidth
    tags.push_back(longs(257, {(uint32_t)height}));                      // ImageLength
    tags.push_back(shorts(258, std::vector<uint32_t>(channels, (uint32_t)bits)));   // BitsPerSample
    tags.push_back(shorts(259, {(uint32_t)compression}));                // Compression
    tags.push_back(shorts(262, {2}));                                    // Photometric RGB
    tags.push_back(ascii(270, description ? description : "aqsis_amd REYES hider"));
    tags.push_back(longs(273, offsets));                                 // StripOffsets
    tags.push_back(shorts(274, {1}));                                    // Orientation top-left
    tags.push_back(shorts(277, {(uint32_t)channels}));                   // SamplesPerPixel
    tags.push_back(longs(278, {rowsPerStrip}));                          // RowsPerStrip
    tags.push_back(longs(279, counts));                                  // StripByteCounts
    tags.push_back(rational(282, 1, 1));                                 // XResolution
    tags.push_back(rational(283, 1, 1));                                 // YResolution
    tags.push_back(shorts(284, {1}));                                    // PlanarConfig contiguous
    tags.push_back(rational(286, (uint32_t)ox, 1));                      // XPosition
    tags.push_back(rational(287, (uint32_t)oy, 1));                      // YPosition
    tags.push_back(shorts(296, {1}));                                    // ResolutionUnit none
    tags.push_back(ascii(305, "aqsis_amd"));                             // Software
    tags.push_back(ascii(306, datetime));                                // DateTime
    if (channels == 4) tags.push_back(shorts(338, {1}));                 // ExtraSamples associated alpha
    if (sample_format) tags.push_back(shorts(339, std::vector<uint32_t>(channels, (uint32_t)sample_format)));
    tags.push_back(longs(33300, {(uint32_t)fw}));                        // PIXAR_IMAGEFULLWIDTH
    tags.push_back(longs(33301, {(uint32_t)fh}));                        // PIXAR_IMAGEFULLLENGTH
    if (sample_format) tags.push_back(one_double(37439, 1.0));           // StoNits

    // Values longer than four bytes go before the IFD.
    std::vector<uint32_t> where(tags.size(), 0);
    for (size_t i = 0; i < tags.size(); ++i) {
        if (tags[i].data.size() <= 4) continue;
        where[i] = (uint32_t)file.size();
        file.insert(file.end(), tags[i].data.begin(), tags[i].data.end());
        if (file.size() & 1) file.push_back(0);
    }
    const uint32_t ifd = (uint32_t)file.size();
    set32(file, 4, ifd);
    put16(file, (uint32_t)tags.size());
    for (size_t i = 0; i < tags.size(); ++i) {
        put16(file, tags[i].tag);
        put16(file, tags[i].type);
        put32(file, tags[i].count);
        if (tags[i].data.size() <= 4) {
            std::vector<uint8_t> v = tags[i].data;
            v.resize(4, 0);
            file.insert(file.end(), v.begin(), v.end());
        } else {
            put32(file, where[i]);
        }
    }
    put32(file, 0);  // no further IFD

    FILE* f = std::fopen(path, "wb");
    if (!f) { rh::set_last_error(std::string("rh_tiff_write: cannot open ") + path); return RH_ERR_INVALID; }
    const size_t wrote = std::fwrite(file.data(), 1, file.size(), f);
    const int closed = std::fclose(f);
    if (wrote != file.size() || closed != 0) {
        rh::set_last_error(std::string("rh_tiff_write: write failed on ") + path);
        return RH_ERR_INVALID;
    }
    return RH_OK;
}

}  // namespace

extern "C" rh_status rh_tiff_write(const char* path, const uint8_t* pixels, int32_t width, int32_t height,
                                   int32_t channels, int32_t compression, const int32_t* origin,
                                   const int32_t* full_size, const char* description) {
    return write_tiff(path, pixels, width, height, channels, 8, 0, compression, origin, full_size, description);
}

// WriteTIFF's second branch (display.cpp:347-411): 16-bit uint for
// PkDspyUnsigned16, else 32-bit IEEE float samples (the 32-bit unsigned
// values keep the float tag, as the driver writes them).
extern "C" rh_status rh_tiff_write_typed(const char* path, const void* pixels, int32_t width, int32_t height,
                                         int32_t channels, int32_t type, int32_t compression, const int32_t* origin,
                                         const int32_t* full_size, const char* description) {
    const uint8_t* p = static_cast<const uint8_t*>(pixels);
    switch (type) {
    case RH_DSPY_UNSIGNED8:
        return write_tiff(path, p, width, height, channels, 8, 0, compression, origin, full_size, description);
    case RH_DSPY_UNSIGNED16:
        return write_tiff(path, p, width, height, channels, 16, 1, compression, origin, full_size, description);
    case RH_DSPY_FLOAT32: case RH_DSPY_UNSIGNED32:
        return write_tiff(path, p, width, height, channels, 32, 3, compression, origin, full_size, description);
    default:
        rh::set_last_error("rh_tiff_write_typed: type must be RH_DSPY_UNSIGNED8, RH_DSPY_UNSIGNED16, "
                           "RH_DSPY_UNSIGNED32 or RH_DSPY_FLOAT32 (the file driver widens signed types)");
        return RH_ERR_INVALID;
    }
}
