// LZ4 frame codec for the depth-map files (raw.npz / clean.npz are an
// np.savez archive compressed with lz4.frame, opensfm/dataset.py:1416-1452;
// the reference pins python-lz4 >= 4.4.5 over lz4-c 1.10.0).  Written from
// the published LZ4 Frame Format and LZ4 Block Format descriptions and the
// xxHash32 specification.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace lz4f {

// xxHash32 of n bytes.
uint32_t xxh32(const uint8_t* p, size_t n, uint32_t seed);

// One frame: independent blocks of at most 64 KB << (2 * (block_max_id - 4))
// (block_max_id 4..7 = 64 KB, 256 KB, 1 MB, 4 MB), the content size in the
// header, an optional content checksum.  Blocks that do not compress are
// stored uncompressed.
std::vector<uint8_t> compress_frame(const uint8_t* src, size_t n, int block_max_id = 7,
                                    bool content_checksum = false);

// Decodes one or more concatenated frames (skippable frames are skipped):
// linked or independent blocks, block / content checksums and the content
// size are verified.  Returns false with a message on malformed input.
bool decompress_frames(const uint8_t* src, size_t n, std::vector<uint8_t>& out, std::string& err);

}  // namespace lz4f
