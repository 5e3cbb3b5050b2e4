// TIFF stacks, .svo octrees and PNG frames for the xenodon CLI.
#include "volume_io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

namespace xdcli {

namespace {

std::string fmt_s(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt_s(const char* f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

// ---------------------------------------------------------------- TIFF ---

enum : uint16_t {
    kWidth = 256, kLength = 257, kBits = 258, kCompression = 259, kPhotometric = 262, kStripOffsets = 273,
    kOrientation = 274, kSamples = 277, kRowsPerStrip = 278, kStripCounts = 279, kPlanar = 284, kPredictor = 317,
    kColorMap = 320, kTileWidth = 322, kTileLength = 323, kTileOffsets = 324, kTileCounts = 325, kExtraSamples = 338,
    kSampleFormat = 339
};

class TiffFile {
    std::ifstream f_;
    uint64_t size_ = 0;
    bool be_ = false, big_ = false;

public:
    uint64_t first_ifd = 0;

    explicit TiffFile(const std::string& path) : f_(path, std::ios::binary) {
        if (!f_) throw Error("Failed to open");
        f_.seekg(0, std::ios::end);
        size_ = (uint64_t)f_.tellg();
        uint8_t h[16] = {};
        read(0, h, std::min<uint64_t>(16, size_));
        if (size_ < 8 || !((h[0] == 'I' && h[1] == 'I') || (h[0] == 'M' && h[1] == 'M')))
            throw Error("Failed to open");
        be_ = h[0] == 'M';
        const uint16_t version = u16(h + 2);
        if (version == 42) {
            first_ifd = u32(h + 4);
        } else if (version == 43 && size_ >= 16 && u16(h + 4) == 8) {
            big_ = true;
            first_ifd = u64(h + 8);
        } else {
            throw Error("Failed to open");
        }
    }
    bool big() const { return big_; }
    uint64_t size() const { return size_; }
    void read(uint64_t off, void* dst, uint64_t n) {
        if (off > size_ || n > size_ - off) throw Error("TIFF data lies outside the file");
        f_.clear();
        f_.seekg((std::streamoff)off);
        f_.read(static_cast<char*>(dst), (std::streamsize)n);
        if (!f_) throw Error("TIFF read failed");
    }
    uint16_t u16(const uint8_t* p) const { return be_ ? (uint16_t)(p[0] << 8 | p[1]) : (uint16_t)(p[1] << 8 | p[0]); }
    uint32_t u32(const uint8_t* p) const {
        return be_ ? ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3])
                   : ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]);
    }
    uint64_t u64(const uint8_t* p) const {
        return be_ ? ((uint64_t)u32(p) << 32 | u32(p + 4)) : ((uint64_t)u32(p + 4) << 32 | u32(p));
    }
};

size_t type_size(uint16_t t) {
    switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
    }
}

struct Ifd {
    std::map<uint16_t, std::vector<uint64_t>> tags;
    uint64_t next = 0;
    bool has(uint16_t t) const { return tags.count(t) != 0; }
    uint64_t get(uint16_t t, uint64_t def) const {
        auto it = tags.find(t);
        return it == tags.end() || it->second.empty() ? def : it->second[0];
    }
    const std::vector<uint64_t>& arr(uint16_t t) const {
        static const std::vector<uint64_t> none;
        auto it = tags.find(t);
        return it == tags.end() ? none : it->second;
    }
};

Ifd read_ifd(TiffFile& tf, uint64_t off) {
    Ifd d;
    const size_t cnt_size = tf.big() ? 8 : 2, ent_size = tf.big() ? 20 : 12, val_size = tf.big() ? 8 : 4;
    uint8_t b[8];
    tf.read(off, b, cnt_size);
    const uint64_t n = tf.big() ? tf.u64(b) : tf.u16(b);
    if (n == 0 || n > 4096) throw Error("Invalid TIFF directory");
    std::vector<uint8_t> ents(n * ent_size);
    tf.read(off + cnt_size, ents.data(), ents.size());
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* e = ents.data() + i * ent_size;
        const uint16_t tag = tf.u16(e), type = tf.u16(e + 2);
        const uint64_t count = tf.big() ? tf.u64(e + 4) : tf.u32(e + 4);
        const size_t ts = type_size(type);
        // only the integer tags used here are decoded
        if (!(type == 1 || type == 3 || type == 4 || type == 16) || ts == 0) continue;
        if (count > (1ull << 28)) throw Error("Invalid TIFF directory");
        std::vector<uint8_t> raw(count * ts);
        const uint8_t* vp = e + 4 + (tf.big() ? 8 : 4);
        if (raw.size() <= val_size) {
            memcpy(raw.data(), vp, raw.size());
        } else {
            const uint64_t voff = tf.big() ? tf.u64(vp) : tf.u32(vp);
            tf.read(voff, raw.data(), raw.size());
        }
        std::vector<uint64_t> vals(count);
        for (uint64_t k = 0; k < count; ++k) {
            const uint8_t* p = raw.data() + k * ts;
            vals[k] = ts == 1 ? p[0] : ts == 2 ? tf.u16(p) : ts == 4 ? tf.u32(p) : tf.u64(p);
        }
        d.tags[tag] = std::move(vals);
    }
    uint8_t nb[8];
    tf.read(off + cnt_size + n * ent_size, nb, val_size);
    d.next = tf.big() ? tf.u64(nb) : tf.u32(nb);
    return d;
}

void unpack_packbits(const std::vector<uint8_t>& in, uint8_t* out, size_t n) {
    size_t i = 0, o = 0;
    while (o < n && i < in.size()) {
        const int8_t c = (int8_t)in[i++];
        if (c >= 0) {
            const size_t k = std::min<size_t>((size_t)c + 1, std::min(n - o, in.size() - i));
            memcpy(out + o, in.data() + i, k);
            i += k;
            o += k;
        } else if (c != -128) {
            if (i >= in.size()) break;
            const size_t k = std::min<size_t>((size_t)(1 - c), n - o);
            memset(out + o, in[i++], k);
            o += k;
        }
    }
}

void unpack_deflate(const std::vector<uint8_t>& in, uint8_t* out, size_t n) {
    z_stream z{};
    if (inflateInit(&z) != Z_OK) throw Error("zlib initialisation failed");
    z.next_in = const_cast<Bytef*>(in.data());
    z.avail_in = (uInt)in.size();
    z.next_out = out;
    z.avail_out = (uInt)n;
    const int r = inflate(&z, Z_FINISH);
    inflateEnd(&z);
    if (r != Z_STREAM_END && r != Z_OK && r != Z_BUF_ERROR) throw Error("Corrupt Deflate data in TIFF");
}

// TIFF LZW: MSB-first codes of 9..12 bits, 256 = clear, 257 = end, code
// width grows one code early (when the next free code is 2^width - 1).
void unpack_lzw(const std::vector<uint8_t>& in, uint8_t* out, size_t n) {
    std::vector<uint16_t> prefix(4096);
    std::vector<uint8_t> suffix(4096), first(4096);
    std::vector<uint16_t> len(4096);
    for (int k = 0; k < 256; ++k) {
        prefix[k] = 0xFFFF;
        suffix[k] = first[k] = (uint8_t)k;
        len[k] = 1;
    }
    size_t bitpos = 0, o = 0;
    int width = 9, next = 258, prev = -1;
    auto read_code = [&]() -> int {
        if (bitpos + width > in.size() * 8) return 257;
        uint32_t v = 0;
        for (int b = 0; b < width; ++b, ++bitpos) v = (v << 1) | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
        return (int)v;
    };
    auto emit = [&](int code) {
        const size_t l = len[code];
        if (o + l > n) {  // truncate at the chunk size
            std::vector<uint8_t> tmp(l);
            for (int c = code, k = (int)l - 1; k >= 0; --k, c = prefix[c]) tmp[k] = suffix[c];
            memcpy(out + o, tmp.data(), n - o);
            o = n;
            return;
        }
        for (int c = code, k = (int)l - 1; k >= 0; --k, c = prefix[c]) out[o + k] = suffix[c];
        o += l;
    };
    while (o < n) {
        const int code = read_code();
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255) throw Error("Corrupt LZW data in TIFF");
            emit(code);
            prev = code;
            continue;
        }
        if (code > next || next >= 4096) throw Error("Corrupt LZW data in TIFF");
        const uint8_t f = code < next ? first[code] : first[prev];
        prefix[next] = (uint16_t)prev;
        suffix[next] = f;
        first[next] = first[prev];
        len[next] = (uint16_t)(len[prev] + 1);
        ++next;
        emit(code);
        prev = code;
        if (next == (1 << width) - 1 && width < 12) ++width;
    }
}

struct Page {
    uint32_t w = 0, h = 0;
    uint32_t spp = 1, bits = 8, compression = 1, photometric = 1, planar = 1, predictor = 1, orientation = 1;
    int alpha = -1;           // sample index of alpha, or -1
    bool unassoc = false;     // alpha is unassociated: premultiply
    bool tiled = false;
    uint32_t cw = 0, ch = 0;  // chunk (strip or tile) width / height
    std::vector<uint64_t> offsets, counts;
    std::vector<uint64_t> cmap;
};

Page parse_page(const Ifd& d, uint32_t index) {
    Page p;
    p.w = (uint32_t)d.get(kWidth, 0);
    p.h = (uint32_t)d.get(kLength, 0);
    p.spp = (uint32_t)d.get(kSamples, 1);
    p.compression = (uint32_t)d.get(kCompression, 1);
    p.photometric = (uint32_t)d.get(kPhotometric, p.spp >= 3 ? 2 : 1);
    p.planar = (uint32_t)d.get(kPlanar, 1);
    p.predictor = (uint32_t)d.get(kPredictor, 1);
    p.orientation = (uint32_t)d.get(kOrientation, 1);
    for (uint64_t b : d.arr(kBits))
        if (b != 8) throw Error(fmt_s("Layer %u has %llu bits per sample (only 8 is supported)", index,
                                      (unsigned long long)b));
    if (d.get(kSampleFormat, 1) != 1) throw Error(fmt_s("Layer %u is not unsigned integer data", index));
    if (p.spp == 0 || p.spp > 16) throw Error(fmt_s("Layer %u has %u samples per pixel", index, p.spp));
    if (p.compression != 1 && p.compression != 5 && p.compression != 8 && p.compression != 32946 &&
        p.compression != 32773)
        throw Error(fmt_s("Layer %u uses TIFF compression %u (supported: none, LZW, Deflate, PackBits)", index,
                          p.compression));
    if (p.predictor != 1 && p.predictor != 2) throw Error(fmt_s("Layer %u uses TIFF predictor %u", index, p.predictor));
    if (p.orientation < 1 || p.orientation > 4)
        throw Error(fmt_s("Layer %u has TIFF orientation %u (supported: 1-4)", index, p.orientation));
    const auto& extra = d.arr(kExtraSamples);
    const uint32_t colour = p.photometric == 2 ? 3u : 1u;
    if (p.photometric > 3) throw Error(fmt_s("Layer %u has photometric interpretation %u", index, p.photometric));
    if (p.spp < colour) throw Error(fmt_s("Layer %u has too few samples per pixel", index));
    if (p.spp > colour && !extra.empty() && (extra[0] == 1 || extra[0] == 2)) {
        p.alpha = (int)colour;
        p.unassoc = extra[0] == 2;
    }
    if (p.photometric == 3) {
        p.cmap = d.arr(kColorMap);
        if (p.cmap.size() != 3 * 256) throw Error(fmt_s("Layer %u has no valid colour map", index));
    }
    p.tiled = d.has(kTileWidth);
    if (p.tiled) {
        p.cw = (uint32_t)d.get(kTileWidth, 0);
        p.ch = (uint32_t)d.get(kTileLength, 0);
        p.offsets = d.arr(kTileOffsets);
        p.counts = d.arr(kTileCounts);
    } else {
        p.cw = p.w;
        p.ch = (uint32_t)std::min<uint64_t>(d.get(kRowsPerStrip, p.h), p.h);
        p.offsets = d.arr(kStripOffsets);
        p.counts = d.arr(kStripCounts);
    }
    if (p.cw == 0 || p.ch == 0 || p.offsets.empty() || p.offsets.size() != p.counts.size())
        throw Error(fmt_s("Layer %u has no valid image data layout", index));
    return p;
}

// Decodes a page into interleaved 8-bit samples (w * h * spp), top row first.
std::vector<uint8_t> decode_page(TiffFile& tf, const Page& p) {
    const uint32_t planes = p.planar == 2 ? p.spp : 1, per = p.planar == 2 ? 1 : p.spp;
    const uint64_t across = (p.w + p.cw - 1) / p.cw, down = (p.h + p.ch - 1) / p.ch;
    if (p.offsets.size() < across * down * planes) throw Error("TIFF layer has too few strips or tiles");
    std::vector<uint8_t> img((size_t)p.w * p.h * p.spp);
    const size_t row_bytes = (size_t)p.cw * per;
    std::vector<uint8_t> chunk(row_bytes * p.ch), raw;
    for (uint32_t pl = 0; pl < planes; ++pl)
        for (uint64_t cy = 0; cy < down; ++cy)
            for (uint64_t cx = 0; cx < across; ++cx) {
                const uint64_t idx = (pl * down + cy) * across + cx;
                raw.resize(p.counts[idx]);
                if (!raw.empty()) tf.read(p.offsets[idx], raw.data(), raw.size());
                std::fill(chunk.begin(), chunk.end(), 0);
                switch (p.compression) {
                case 1: memcpy(chunk.data(), raw.data(), std::min(raw.size(), chunk.size())); break;
                case 5: unpack_lzw(raw, chunk.data(), chunk.size()); break;
                case 8: case 32946: unpack_deflate(raw, chunk.data(), chunk.size()); break;
                default: unpack_packbits(raw, chunk.data(), chunk.size()); break;
                }
                if (p.predictor == 2)
                    for (uint32_t r = 0; r < p.ch; ++r) {
                        uint8_t* row = chunk.data() + r * row_bytes;
                        for (size_t k = per; k < row_bytes; ++k) row[k] = (uint8_t)(row[k] + row[k - per]);
                    }
                for (uint32_t r = 0; r < p.ch; ++r) {
                    const uint64_t y = cy * p.ch + r;
                    if (y >= p.h) break;
                    for (uint32_t c = 0; c < p.cw; ++c) {
                        const uint64_t x = cx * p.cw + c;
                        if (x >= p.w) break;
                        const uint8_t* src = chunk.data() + r * row_bytes + (size_t)c * per;
                        uint8_t* dst = img.data() + ((size_t)y * p.w + x) * p.spp;
                        if (planes == 1)
                            memcpy(dst, src, p.spp);
                        else
                            dst[pl] = src[0];
                    }
                }
            }
    return img;
}

// TIFFReadRGBAImage's pixel for one decoded pixel (packed ABGR = RGBA bytes).
inline uint32_t rgba_of(const Page& p, const uint8_t* s, bool cmap16) {
    uint32_t r, g, b, a = 255;
    if (p.photometric == 2) {
        r = s[0], g = s[1], b = s[2];
    } else if (p.photometric == 3) {
        const uint64_t* m = p.cmap.data();
        auto cv = [&](uint64_t v) { return (uint32_t)(cmap16 ? (v >> 8) & 0xFF : v & 0xFF); };
        r = cv(m[s[0]]), g = cv(m[256 + s[0]]), b = cv(m[512 + s[0]]);
    } else {
        r = g = b = p.photometric == 0 ? 255u - s[0] : s[0];
    }
    if (p.alpha >= 0) {
        a = s[p.alpha];
        if (p.unassoc && p.photometric == 2) {  // libtiff premultiplies RGB only; grey keeps its value
            r = (r * a + 127) / 255;
            g = (g * a + 127) / 255;
            b = (b * a + 127) / 255;
        }
    }
    return r | g << 8 | b << 16 | a << 24;
}

template <typename T>
void put_le(std::vector<uint8_t>& v, T x) {
    for (size_t i = 0; i < sizeof(T); ++i) v.push_back((uint8_t)(x >> (8 * i)));
}

}  // namespace

GridData load_tiff(const std::string& path, bool print_size) {
    TiffFile tf(path);
    std::vector<Ifd> dirs;
    uint64_t off = tf.first_ifd;
    uint32_t width = 0, height = 0;
    while (off != 0) {
        if (dirs.size() >= 65535) throw Error("TIFF has more than 65535 layers");
        dirs.push_back(read_ifd(tf, off));
        const Ifd& d = dirs.back();
        const uint32_t lw = (uint32_t)d.get(kWidth, 0), lh = (uint32_t)d.get(kLength, 0);
        const size_t depth = dirs.size() - 1;
        if (lw == 0 || lh == 0) throw Error(fmt_s("Layer %zu has invalid dimensions (%ux%u)", depth, lw, lh));
        if (width == 0) {
            width = lw;
            height = lh;
        } else if (lw != width || lh != height) {
            throw Error(fmt_s("Dimensions of layer %zu differ from previous dimensions (%ux%u, previously %ux%u)",
                              depth, lw, lh, width, height));
        }
        off = d.next;
        if (off >= tf.size()) off = 0;
    }
    GridData g;
    g.dim[0] = width;
    g.dim[1] = height;
    g.dim[2] = dirs.size();
    const uint64_t layer = (uint64_t)width * height;
    if (print_size)
        printf("%ux%ux%zu = %llu pixels\n", width, height, dirs.size(), (unsigned long long)(layer * dirs.size()));
    g.texels.resize(layer * dirs.size());
    for (uint32_t z = 0; z < dirs.size(); ++z) {
        const Page p = parse_page(dirs[z], z);
        const std::vector<uint8_t> img = decode_page(tf, p);
        bool cmap16 = false;
        for (uint64_t v : p.cmap) cmap16 |= v >= 256;
        const bool flip_v = p.orientation == 1 || p.orientation == 2;  // raster origin at the lower left
        const bool flip_h = p.orientation == 2 || p.orientation == 3;
        uint32_t* out = g.texels.data() + layer * z;
        for (uint32_t y = 0; y < height; ++y) {
            const uint32_t sy = flip_v ? height - 1 - y : y;
            for (uint32_t x = 0; x < width; ++x) {
                const uint32_t sx = flip_h ? width - 1 - x : x;
                out[(uint64_t)y * width + x] = rgba_of(p, img.data() + ((size_t)sy * width + sx) * p.spp, cmap16);
            }
        }
    }
    return g;
}

void save_tiff(const std::string& path, const GridData& g) {
    const uint64_t W = g.dim[0], H = g.dim[1], Z = g.dim[2];
    const uint64_t page = W * H * 4;
    const bool big = page * Z + Z * 512 + 64 > 0xFFFFFFFFull;
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error("Failed to open");
    std::vector<uint8_t> hdr{'I', 'I'};
    if (big) {
        put_le<uint16_t>(hdr, 43);
        put_le<uint16_t>(hdr, 8);
        put_le<uint16_t>(hdr, 0);
        put_le<uint64_t>(hdr, 0);  // patched below
    } else {
        put_le<uint16_t>(hdr, 42);
        put_le<uint32_t>(hdr, 0);
    }
    f.write((const char*)hdr.data(), (std::streamsize)hdr.size());
    uint64_t pos = hdr.size();
    uint64_t link_at = big ? 8 : 4;  // where the next-IFD offset is to be written
    for (uint64_t z = 0; z < Z; ++z) {
        // pixel data, then the directory that points to it
        // rows bottom-up: TIFFReadRGBAImage's raster (origin lower left) is
        // then the grid layer itself, so load_tiff(save_tiff(g)) == g
        const uint64_t data_off = pos;
        for (uint64_t r = 0; r < H; ++r)
            f.write((const char*)(g.texels.data() + W * H * z + W * (H - 1 - r)), (std::streamsize)(W * 4));
        pos += page;
        if (pos & 1) {
            f.put(0);
            ++pos;
        }
        const uint64_t bits_off = pos;
        const uint16_t bits[4] = {8, 8, 8, 8};
        f.write((const char*)bits, 8);
        pos += 8;
        const uint64_t ifd_off = pos;
        std::vector<uint8_t> d;
        struct E { uint16_t tag, type; uint64_t count, value; };
        const E ents[] = {{kWidth, 4, 1, W}, {kLength, 4, 1, H}, {kBits, 3, 4, bits_off}, {kCompression, 3, 1, 1},
                          {kPhotometric, 3, 1, 2}, {kStripOffsets, (uint16_t)(big ? 16 : 4), 1, data_off},
                          {kOrientation, 3, 1, 1}, {kSamples, 3, 1, 4}, {kRowsPerStrip, 4, 1, H},
                          {kStripCounts, (uint16_t)(big ? 16 : 4), 1, page}, {kPlanar, 3, 1, 1},
                          {kExtraSamples, 3, 1, 2}};
        const size_t n = sizeof(ents) / sizeof(ents[0]);
        if (big) put_le<uint64_t>(d, n); else put_le<uint16_t>(d, (uint16_t)n);
        for (const E& e : ents) {
            put_le<uint16_t>(d, e.tag);
            put_le<uint16_t>(d, e.type);
            if (big) put_le<uint64_t>(d, e.count); else put_le<uint32_t>(d, (uint32_t)e.count);
            const size_t vs = big ? 8 : 4;
            std::vector<uint8_t> v;
            if (e.tag == kBits && big) for (int k = 0; k < 4; ++k) put_le<uint16_t>(v, 8);  // fits inline
            else if (e.type == 3 && e.count == 1) put_le<uint16_t>(v, (uint16_t)e.value);
            else if (e.type == 16) put_le<uint64_t>(v, e.value);
            else put_le<uint32_t>(v, (uint32_t)e.value);
            v.resize(vs, 0);
            d.insert(d.end(), v.begin(), v.end());
        }
        const uint64_t next_at = ifd_off + d.size();
        if (big) put_le<uint64_t>(d, 0); else put_le<uint32_t>(d, 0);
        f.write((const char*)d.data(), (std::streamsize)d.size());
        pos += d.size();
        // link the previous directory (or the header) to this one
        f.seekp((std::streamoff)link_at);
        std::vector<uint8_t> l;
        if (big) put_le<uint64_t>(l, ifd_off); else put_le<uint32_t>(l, (uint32_t)ifd_off);
        f.write((const char*)l.data(), (std::streamsize)l.size());
        f.seekp((std::streamoff)pos);
        link_at = next_at;
    }
    if (!f) throw Error("Failed to write");
}

// ----------------------------------------------------------------- SVO ---

static const char kSvoId[8] = {'X', 'N', 'D', 'N', '-', 'S', 'V', 'O'};

SvoData load_svo(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error("Failed to open");
    in.seekg(0, std::ios::end);
    const uint64_t size = (uint64_t)in.tellg();
    in.seekg(0);
    uint8_t h[24] = {};
    in.read((char*)h, 24);
    if (size < 8 || memcmp(h, kSvoId, 8) != 0) throw Error("Invalid format id");
    if (size < 24) throw Error("File size does not match number of nodes");
    uint64_t side = 0, n = 0;
    for (int i = 7; i >= 0; --i) {
        side = side << 8 | h[8 + i];
        n = n << 8 | h[16 + i];
    }
    if (n > (size - 24) / 40 || size - 24 != n * 40) throw Error("File size does not match number of nodes");
    SvoData s;
    s.side = side;
    s.words.resize(n * 10);
    std::vector<uint8_t> buf(std::min<uint64_t>(n * 40, 1ull << 26));
    uint64_t done = 0;
    while (done < n * 40) {
        const uint64_t k = std::min<uint64_t>(buf.size(), n * 40 - done);
        in.read((char*)buf.data(), (std::streamsize)k);
        if (!in) throw Error("File size does not match number of nodes");
        for (uint64_t i = 0; i < k; i += 4)
            s.words[(done + i) / 4] = (uint32_t)buf[i] | (uint32_t)buf[i + 1] << 8 | (uint32_t)buf[i + 2] << 16 |
                                      (uint32_t)buf[i + 3] << 24;
        done += k;
    }
    return s;
}

void save_svo(const std::string& path, const SvoData& s) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw Error("Failed to open");
    std::vector<uint8_t> h(kSvoId, kSvoId + 8);
    put_le<uint64_t>(h, s.side);
    put_le<uint64_t>(h, s.node_count());
    out.write((const char*)h.data(), (std::streamsize)h.size());
    std::vector<uint8_t> buf;
    buf.reserve(1 << 22);
    for (size_t i = 0; i < s.words.size(); ++i) {
        put_le<uint32_t>(buf, s.words[i]);
        if (buf.size() >= (1u << 22)) {
            out.write((const char*)buf.data(), (std::streamsize)buf.size());
            buf.clear();
        }
    }
    out.write((const char*)buf.data(), (std::streamsize)buf.size());
    if (!out) throw Error("Failed to write");
}

// ----------------------------------------------------------------- PNG ---

namespace {
void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    for (int i = 3; i >= 0; --i) v.push_back((uint8_t)(x >> (8 * i)));
}
void chunk(std::ofstream& f, const char* type, const std::vector<uint8_t>& data) {
    std::vector<uint8_t> c;
    put_be32(c, (uint32_t)data.size());
    c.insert(c.end(), type, type + 4);
    c.insert(c.end(), data.begin(), data.end());
    put_be32(c, (uint32_t)crc32(0, c.data() + 4, (uInt)(c.size() - 4)));
    f.write((const char*)c.data(), (std::streamsize)c.size());
}
}  // namespace

void save_png(const std::string& path, const uint32_t* rgba, uint32_t w, uint32_t h) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error("Failed to open");
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    f.write((const char*)sig, 8);
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, adaptive filters, no interlace
    chunk(f, "IHDR", ihdr);
    // rows with the "up" filter (a good default for rendered frames)
    const size_t stride = (size_t)w * 4;
    std::vector<uint8_t> raw((stride + 1) * h);
    const uint8_t* px = reinterpret_cast<const uint8_t*>(rgba);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = raw.data() + y * (stride + 1);
        row[0] = y ? 2 : 0;
        const uint8_t* cur = px + y * stride;
        if (y)
            for (size_t k = 0; k < stride; ++k) row[1 + k] = (uint8_t)(cur[k] - cur[k - stride]);
        else
            memcpy(row + 1, cur, stride);
    }
    uLongf n = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(n);
    if (compress2(z.data(), &n, raw.data(), (uLong)raw.size(), 3) != Z_OK) throw Error("PNG compression failed");
    z.resize(n);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    if (!f) throw Error("Failed to write");
}

}  // namespace xdcli
