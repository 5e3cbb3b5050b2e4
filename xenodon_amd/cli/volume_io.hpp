// Volume and image files of the xenodon CLI: TIFF stacks (read and write),
// .svo octrees (read and write) and RGBA8 PNG frames (write).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace xdcli {

// The reference's Error (src/core/Error.h): a message the CLI prints as
// "Error: <message>".
struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// An RGBA8 grid, x fastest, then y, then z (src/model/Grid.h:50-56).
struct GridData {
    uint64_t dim[3] = {0, 0, 0};
    std::vector<uint32_t> texels;  // R in the low byte (src/model/Pixel.h:18-21)
};

// Grid::load_tiff (src/model/Grid.cpp:27-79): one page per z layer, each read
// with the semantics of libtiff's TIFFReadRGBAImage -- the raster's origin is
// the lower-left corner (grid row y = image row height-1-y for top-left
// pages), RGB pages get A = 255, grayscale pages R = G = B, palette pages use
// the colour map, and unassociated alpha of RGB pages is premultiplied
// ((c * a + 127) / 255; grey pages keep their value).
// Baseline TIFF and BigTIFF, either byte order, strips or tiles, chunky or
// planar, 8-bit samples, uncompressed, PackBits, LZW or Deflate (with or
// without horizontal predictor).  Throws Error with the reference's messages
// where the reference has one.
GridData load_tiff(const std::string& path, bool print_size = true);

// Writes an uncompressed little-endian RGBA8 TIFF stack (one top-left page per
// z layer, rows stored bottom-up so that load_tiff returns the same grid,
// unassociated alpha); BigTIFF when the file would exceed 4 GiB.
void save_tiff(const std::string& path, const GridData& grid);

struct SvoData {
    uint64_t side = 0;
    std::vector<uint32_t> words;  // node_count * 10 (children[8], color, is_leaf_depth)
    uint64_t node_count() const { return words.size() / 10; }
};

// Octree::load_svo / save_svo (src/model/Octree.cpp:50-114): "XNDN-SVO", u64
// side, u64 node count, then 10 little-endian u32 per node; the remaining size
// must match the count exactly.
SvoData load_svo(const std::string& path);
void save_svo(const std::string& path, const SvoData& svo);

// RGBA8 PNG (8-bit truecolour with alpha), rows top to bottom.
void save_png(const std::string& path, const uint32_t* rgba, uint32_t width, uint32_t height);

}  // namespace xdcli
