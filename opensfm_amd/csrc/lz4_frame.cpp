// LZ4 frame codec (see lz4_frame.h).  Host code of the depth-map file I/O.
#include "lz4_frame.h"

#include <algorithm>
#include <cstring>

namespace lz4f {

namespace {

constexpr uint32_t kFrameMagic = 0x184D2204u;
constexpr uint32_t kSkippableMagicMask = 0xFFFFFFF0u;
constexpr uint32_t kSkippableMagic = 0x184D2A50u;
constexpr int kMinMatch = 4;
constexpr int kLastLiterals = 5;   // the last 5 bytes of a block are literals
constexpr int kMatchFindLimit = 12; // no match starts within the last 12 bytes
constexpr int kHashLog = 16;

constexpr uint32_t P1 = 0x9E3779B1u, P2 = 0x85EBCA77u, P3 = 0xC2B2AE3Du, P4 = 0x27D4EB2Fu, P5 = 0x165667B1u;

inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
inline uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);  // little-endian host
  return v;
}
inline void wr32(std::vector<uint8_t>& o, uint32_t v) {
  for (int i = 0; i < 4; ++i) o.push_back(static_cast<uint8_t>(v >> (8 * i)));
}
inline void wr64(std::vector<uint8_t>& o, uint64_t v) {
  for (int i = 0; i < 8; ++i) o.push_back(static_cast<uint8_t>(v >> (8 * i)));
}

size_t block_max_bytes(int id) { return size_t(1) << (8 + 2 * id); }  // 4 -> 64 KB ... 7 -> 4 MB

void put_length(std::vector<uint8_t>& o, size_t rem) {
  while (rem >= 255) {
    o.push_back(255);
    rem -= 255;
  }
  o.push_back(static_cast<uint8_t>(rem));
}

void emit_sequence(std::vector<uint8_t>& o, const uint8_t* lit, size_t nlit, size_t offset, size_t mlen) {
  const size_t ml = mlen - kMinMatch;
  o.push_back(static_cast<uint8_t>((std::min<size_t>(nlit, 15) << 4) | std::min<size_t>(ml, 15)));
  if (nlit >= 15) put_length(o, nlit - 15);
  o.insert(o.end(), lit, lit + nlit);
  o.push_back(static_cast<uint8_t>(offset & 0xFF));
  o.push_back(static_cast<uint8_t>(offset >> 8));
  if (ml >= 15) put_length(o, ml - 15);
}

void emit_last_literals(std::vector<uint8_t>& o, const uint8_t* lit, size_t nlit) {
  o.push_back(static_cast<uint8_t>(std::min<size_t>(nlit, 15) << 4));
  if (nlit >= 15) put_length(o, nlit - 15);
  o.insert(o.end(), lit, lit + nlit);
}

// Greedy single-pass LZ4 block compression (hash of 4 bytes -> last
// position): the block references nothing outside itself.
std::vector<uint8_t> compress_block(const uint8_t* src, size_t n) {
  std::vector<uint8_t> o;
  o.reserve(n + n / 255 + 16);
  size_t anchor = 0;
  if (n >= static_cast<size_t>(kMatchFindLimit) + 1) {
    std::vector<int64_t> table(size_t(1) << kHashLog, -1);
    const size_t match_start_limit = n - kMatchFindLimit;  // a match starts at ip <= n - 12
    const size_t match_end_limit = n - kLastLiterals;      // and ends at or before n - 5
    size_t ip = 0;
    while (ip <= match_start_limit) {
      const uint32_t seq = rd32(src + ip);
      const uint32_t h = (seq * 2654435761u) >> (32 - kHashLog);
      const int64_t ref = table[h];
      table[h] = static_cast<int64_t>(ip);
      if (ref >= 0 && ip - static_cast<size_t>(ref) <= 65535 && rd32(src + ref) == seq) {
        size_t r = static_cast<size_t>(ref), s = ip;
        size_t mlen = kMinMatch;
        while (s + mlen < match_end_limit && src[r + mlen] == src[s + mlen]) ++mlen;
        while (s > anchor && r > 0 && src[s - 1] == src[r - 1]) {  // extend backwards
          --s;
          --r;
          ++mlen;
        }
        emit_sequence(o, src + anchor, s - anchor, s - r, mlen);
        ip = s + mlen;
        anchor = ip;
        if (ip >= 2 && ip - 2 <= match_start_limit) {
          const uint32_t h2 = (rd32(src + ip - 2) * 2654435761u) >> (32 - kHashLog);
          table[h2] = static_cast<int64_t>(ip - 2);
        }
      } else {
        ++ip;
      }
    }
  }
  emit_last_literals(o, src + anchor, n - anchor);
  return o;
}

// Decodes one block appended to out; earlier output (previous linked
// blocks) is the dictionary a match may reach into.
bool decode_block(const uint8_t* p, size_t n, std::vector<uint8_t>& out, size_t max_out, std::string& err) {
  const size_t start = out.size();
  size_t i = 0;
  while (true) {
    if (i >= n) {
      err = "lz4: truncated block";
      return false;
    }
    const uint8_t token = p[i++];
    size_t nlit = token >> 4;
    if (nlit == 15) {
      uint8_t b;
      do {
        if (i >= n) { err = "lz4: truncated literal length"; return false; }
        b = p[i++];
        nlit += b;
      } while (b == 255);
    }
    if (nlit > n - i) { err = "lz4: literals past the block end"; return false; }
    if (out.size() - start + nlit > max_out) { err = "lz4: block larger than its maximum size"; return false; }
    out.insert(out.end(), p + i, p + i + nlit);
    i += nlit;
    if (i == n) return true;  // the last sequence has no match
    if (n - i < 2) { err = "lz4: truncated offset"; return false; }
    const size_t offset = static_cast<size_t>(p[i]) | (static_cast<size_t>(p[i + 1]) << 8);
    i += 2;
    if (offset == 0 || offset > out.size()) { err = "lz4: match offset outside the decoded data"; return false; }
    size_t mlen = token & 15;
    if (mlen == 15) {
      uint8_t b;
      do {
        if (i >= n) { err = "lz4: truncated match length"; return false; }
        b = p[i++];
        mlen += b;
      } while (b == 255);
    }
    mlen += kMinMatch;
    if (out.size() - start + mlen > max_out) { err = "lz4: block larger than its maximum size"; return false; }
    size_t from = out.size() - offset;
    for (size_t k = 0; k < mlen; ++k) out.push_back(out[from + k]);  // byte-wise: overlapping copies repeat
  }
}

}  // namespace

uint32_t xxh32(const uint8_t* p, size_t n, uint32_t seed) {
  const uint8_t* end = p + n;
  uint32_t acc;
  if (n >= 16) {
    uint32_t v1 = seed + P1 + P2, v2 = seed + P2, v3 = seed, v4 = seed - P1;
    const uint8_t* limit = end - 16;
    do {
      v1 = rotl(v1 + rd32(p) * P2, 13) * P1;
      v2 = rotl(v2 + rd32(p + 4) * P2, 13) * P1;
      v3 = rotl(v3 + rd32(p + 8) * P2, 13) * P1;
      v4 = rotl(v4 + rd32(p + 12) * P2, 13) * P1;
      p += 16;
    } while (p <= limit);
    acc = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
  } else {
    acc = seed + P5;
  }
  acc += static_cast<uint32_t>(n);
  while (p + 4 <= end) {
    acc = rotl(acc + rd32(p) * P3, 17) * P4;
    p += 4;
  }
  while (p < end) {
    acc = rotl(acc + (*p) * P5, 11) * P1;
    ++p;
  }
  acc ^= acc >> 15;
  acc *= P2;
  acc ^= acc >> 13;
  acc *= P3;
  acc ^= acc >> 16;
  return acc;
}

std::vector<uint8_t> compress_frame(const uint8_t* src, size_t n, int block_max_id, bool content_checksum) {
  block_max_id = std::min(std::max(block_max_id, 4), 7);
  std::vector<uint8_t> o;
  o.reserve(n + n / 255 + 64);
  wr32(o, kFrameMagic);
  const size_t desc = o.size();
  // FLG: version 01, independent blocks, content size, [content checksum]
  o.push_back(static_cast<uint8_t>(0x40 | 0x20 | 0x08 | (content_checksum ? 0x04 : 0x00)));
  o.push_back(static_cast<uint8_t>(block_max_id << 4));  // BD
  wr64(o, static_cast<uint64_t>(n));
  o.push_back(static_cast<uint8_t>((xxh32(o.data() + desc, o.size() - desc, 0) >> 8) & 0xFF));  // HC
  const size_t bmax = block_max_bytes(block_max_id);
  for (size_t off = 0; off < n; off += bmax) {
    const size_t len = std::min(bmax, n - off);
    std::vector<uint8_t> c = compress_block(src + off, len);
    if (c.size() < len) {
      wr32(o, static_cast<uint32_t>(c.size()));
      o.insert(o.end(), c.begin(), c.end());
    } else {
      wr32(o, static_cast<uint32_t>(len) | 0x80000000u);  // stored uncompressed
      o.insert(o.end(), src + off, src + off + len);
    }
  }
  wr32(o, 0);  // end mark
  if (content_checksum) wr32(o, xxh32(src, n, 0));
  return o;
}

bool decompress_frames(const uint8_t* src, size_t n, std::vector<uint8_t>& out, std::string& err) {
  out.clear();
  size_t i = 0;
  bool any = false;
  while (i < n) {
    if (n - i < 4) { err = "lz4: truncated frame"; return false; }
    const uint32_t magic = rd32(src + i);
    if ((magic & kSkippableMagicMask) == kSkippableMagic) {
      if (n - i < 8) { err = "lz4: truncated skippable frame"; return false; }
      const size_t len = rd32(src + i + 4);
      if (len > n - i - 8) { err = "lz4: truncated skippable frame"; return false; }
      i += 8 + len;
      continue;
    }
    if (magic != kFrameMagic) { err = "lz4: not an LZ4 frame (bad magic number)"; return false; }
    i += 4;
    const size_t desc = i;
    if (n - i < 3) { err = "lz4: truncated frame descriptor"; return false; }
    const uint8_t flg = src[i], bd = src[i + 1];
    if ((flg >> 6) != 1) { err = "lz4: unsupported frame version"; return false; }
    if (flg & 0x02) { err = "lz4: reserved frame flag set"; return false; }
    const bool indep = flg & 0x20, bsum = flg & 0x10, csize = flg & 0x08, csum = flg & 0x04, dict = flg & 0x01;
    const int bid = (bd >> 4) & 7;
    if (bid < 4 || (bd & 0x8F)) { err = "lz4: bad block descriptor"; return false; }
    i += 2;
    uint64_t content_size = 0;
    if (csize) {
      if (n - i < 8) { err = "lz4: truncated content size"; return false; }
      for (int k = 0; k < 8; ++k) content_size |= static_cast<uint64_t>(src[i + k]) << (8 * k);
      i += 8;
    }
    if (dict) { err = "lz4: frames with a dictionary are not supported"; return false; }
    if (i >= n) { err = "lz4: truncated header checksum"; return false; }
    if (src[i] != static_cast<uint8_t>((xxh32(src + desc, i - desc, 0) >> 8) & 0xFF)) {
      err = "lz4: header checksum mismatch";
      return false;
    }
    ++i;
    const size_t frame_start = out.size();
    const size_t bmax = block_max_bytes(bid);
    if (csize && content_size < (uint64_t(1) << 40)) out.reserve(out.size() + static_cast<size_t>(content_size));
    while (true) {
      if (n - i < 4) { err = "lz4: truncated block header"; return false; }
      const uint32_t bs = rd32(src + i);
      i += 4;
      if (bs == 0) break;  // end mark
      const size_t len = bs & 0x7FFFFFFFu;
      if (len > n - i) { err = "lz4: truncated block"; return false; }
      if (bsum) {
        if (n - i - len < 4) { err = "lz4: truncated block checksum"; return false; }
        if (rd32(src + i + len) != xxh32(src + i, len, 0)) { err = "lz4: block checksum mismatch"; return false; }
      }
      if (bs & 0x80000000u) {
        if (len > bmax) { err = "lz4: block larger than its maximum size"; return false; }
        out.insert(out.end(), src + i, src + i + len);
      } else if (indep) {
        // matches may not reach before this block: decode against a view
        // whose history starts at the block
        std::vector<uint8_t> blk;
        if (!decode_block(src + i, len, blk, bmax, err)) return false;
        out.insert(out.end(), blk.begin(), blk.end());
      } else {
        if (!decode_block(src + i, len, out, bmax, err)) return false;
      }
      i += len + (bsum ? 4 : 0);
    }
    if (csize && out.size() - frame_start != content_size) { err = "lz4: content size mismatch"; return false; }
    if (csum) {
      if (n - i < 4) { err = "lz4: truncated content checksum"; return false; }
      if (rd32(src + i) != xxh32(out.data() + frame_start, out.size() - frame_start, 0)) {
        err = "lz4: content checksum mismatch";
        return false;
      }
      i += 4;
    }
    any = true;
  }
  if (!any) { err = "lz4: no frame"; return false; }
  return true;
}

}  // namespace lz4f
