// formats.cpp — text and on-disk formats of include/rtx_formats.h (host only).
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rtx_formats.h"

namespace {

uint64_t rd64(const uint8_t* p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

int32_t rd32s(const uint8_t* p) {
  int32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

bool inside(uint64_t off, uint64_t len, uint64_t bytes) { return off <= bytes && len <= bytes - off; }

// Walks the nodes reachable from the header at `hdr` without entering
// instance leaves' BLAS: returns the end of the furthest node and, for a top
// level, the BLAS header offsets its instance leaves name.
rtx_status walk_level(const uint8_t* A, uint64_t bytes, uint64_t hdr, uint64_t* end,
                      std::vector<uint64_t>* blas_headers) {
  if (hdr % 64 || !inside(hdr, 64, bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
  uint64_t hi = hdr + 64;
  const uint64_t root = hdr + rd64(A + hdr);
  if (root % 64 || !inside(root, 64, bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
  std::vector<uint64_t> stack{root};
  size_t visited = 0;
  const size_t limit = bytes / 64 + 1;
  while (!stack.empty()) {
    if (++visited > limit) return RTX_ERR_BAD_NODE_TYPE;  // a cycle
    const uint64_t n = stack.back();
    stack.pop_back();
    hi = std::max(hi, n + 64);
    int64_t child = (int64_t)n + (int64_t)rd32s(A + n + 12) * 64;
    for (int i = 0; i < 6; ++i) {
      const uint8_t info = A[n + 22 + i];
      const uint32_t size = info & 3u, type = (info >> 2) & 15u;
      if (size == 0) continue;
      if (child < 0 || !inside((uint64_t)child, (uint64_t)size * 64, bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
      const uint64_t c = (uint64_t)child;
      hi = std::max(hi, c + (uint64_t)size * 64);
      if (type == 0) {
        stack.push_back(c);
      } else if (type == 1) {
        if (!blas_headers) return RTX_ERR_BAD_NODE_TYPE;  // instance leaf below the top level
        if (!inside(c, 128, bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
        blas_headers->push_back(c + rd64(A + c + 64));
      } else if (type != 3 && type != 4) {
        return RTX_ERR_BAD_NODE_TYPE;
      }
      child += (int64_t)size * 64;
    }
  }
  *end = hi;
  return RTX_OK;
}

bool write_file(const std::string& path, const uint8_t* p, uint64_t n, const char* mode) {
  FILE* f = std::fopen(path.c_str(), mode);
  if (!f) return false;
  const bool ok = n == 0 || std::fwrite(p, 1, n, f) == n;
  return std::fclose(f) == 0 && ok;
}

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)n);
  const bool ok = n == 0 || std::fread(out.data(), 1, (size_t)n, f) == (size_t)n;
  std::fclose(f);
  return ok;
}

struct DumpMeta {
  int64_t desc_size = 0, type = 0, max_back = 0, min_back = 0, min_front = 0, max_front = 0;
  int64_t back_buf = 0, front_buf = 0, have_back = 0, have_front = 0;
  int64_t lo = 0, hi = 0;  // arena extent relative to the TLAS
};

rtx_status read_meta(const char* prefix, DumpMeta& m) {
  std::vector<uint8_t> txt;
  if (!read_file(std::string(prefix) + ".asmetadata", txt)) return RTX_ERR_INVALID_ARGUMENT;
  txt.push_back(0);
  long long v[10];
  if (std::sscanf((const char*)txt.data(), "%lld,%lld,%lld,%lld,%lld,%lld,%lld,%lld,%lld,%lld", &v[0], &v[1],
                  &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9]) != 10)
    return RTX_ERR_INVALID_ARGUMENT;
  m.desc_size = v[0];
  m.type = v[1];
  m.max_back = v[2];
  m.min_back = v[3];
  m.min_front = v[4];
  m.max_front = v[5];
  m.back_buf = v[6];
  m.front_buf = v[7];
  m.have_back = v[8];
  m.have_front = v[9];
  if (m.desc_size < 64 || m.back_buf < 0 || m.front_buf < 0) return RTX_ERR_INVALID_ARGUMENT;
  if (m.have_back && !(m.max_back <= m.min_back && m.min_back < 0)) return RTX_ERR_INVALID_ARGUMENT;
  if (m.have_front && !(0 <= m.min_front && m.min_front <= m.max_front)) return RTX_ERR_INVALID_ARGUMENT;
  int64_t lo = 0, hi = m.desc_size;
  if (m.have_back) {
    lo = std::min(lo, m.max_back);
    hi = std::max(hi, m.min_back + m.back_buf);
  }
  if (m.have_front) hi = std::max(hi, m.max_front + m.front_buf);
  lo = -((-lo + 63) / 64) * 64;  // the TLAS lands on a 64-byte boundary
  hi = lo + ((hi - lo + 63) / 64) * 64;
  if (hi - lo > (int64_t)1 << 40) return RTX_ERR_INVALID_ARGUMENT;
  m.lo = lo;
  m.hi = hi;
  return RTX_OK;
}

}  // namespace

extern "C" {

rtx_status rtx_format_treelet_trace(const rtx_record* records, const uint64_t* offsets, uint32_t n_rays,
                                    const uint64_t* roots, uint32_t n_treelets, uint64_t base_address,
                                    uint32_t first_ray_id, char* out, size_t capacity, size_t* length) {
  if (!offsets || !length || (n_rays && !records && offsets[n_rays] > 0) || (!roots && n_treelets)) {
    return RTX_ERR_INVALID_ARGUMENT;
  }
  size_t len = 0;
  char buf[32];
  auto put = [&](const char* s, size_t n) {
    if (out && len < capacity) std::memcpy(out + len, s, std::min(n, capacity - len));
    len += n;
  };
  for (uint32_t i = 0; i < n_rays; ++i) {
    if (offsets[i + 1] < offsets[i]) return RTX_ERR_INVALID_ARGUMENT;
    int k = std::snprintf(buf, sizeof buf, "RayID,%u", first_ray_id + i);
    put(buf, (size_t)k);
    for (uint64_t r = offsets[i]; r < offsets[i + 1]; ++r) {
      const uint32_t t = records[r].treelet;
      if (t == RTX_NO_TREELET || t >= n_treelets) return RTX_ERR_NO_PARTITION;
      k = std::snprintf(buf, sizeof buf, ",0x%x", (unsigned)(uint32_t)(base_address + roots[t]));
      put(buf, (size_t)k);
    }
    put("\n", 1);
  }
  if (out && len < capacity) out[len] = 0;
  *length = len;
  return RTX_OK;
}

rtx_status rtx_as_dump_write(const void* arena_host, uint64_t arena_bytes, uint64_t tlas_offset, const char* prefix) {
  if (!arena_host || !prefix || arena_bytes < 64) return RTX_ERR_INVALID_ARGUMENT;
  const uint8_t* A = (const uint8_t*)arena_host;
  uint64_t top_end = 0;
  std::vector<uint64_t> blas;
  rtx_status st = walk_level(A, arena_bytes, tlas_offset, &top_end, &blas);
  if (st != RTX_OK) return st;
  std::sort(blas.begin(), blas.end());
  blas.erase(std::unique(blas.begin(), blas.end()), blas.end());
  // findOffsetBounds (vulkan_ray_tracing.cc:4901-4944) over the BLAS headers
  int64_t max_back = 0, min_back = 0, min_front = 0, max_front = 0, back_buf = 0, front_buf = 0;
  bool have_back = false, have_front = false;
  for (uint64_t b : blas) {
    uint64_t end = 0;
    st = walk_level(A, arena_bytes, b, &end, nullptr);
    if (st != RTX_OK) return st;
    const int64_t off = (int64_t)b - (int64_t)tlas_offset;
    const int64_t rel_end = (int64_t)end - (int64_t)tlas_offset;
    if (off < 0) {
      if (!have_back) max_back = min_back = off;
      max_back = std::min(max_back, off);
      min_back = std::max(min_back, off);
      have_back = true;
      back_buf = std::max(back_buf, rel_end);  // made relative to min_back below
    } else {
      if (!have_front) min_front = max_front = off;
      min_front = std::min(min_front, off);
      max_front = std::max(max_front, off);
      have_front = true;
      front_buf = std::max(front_buf, rel_end);
    }
  }
  if (have_back) back_buf = std::max<int64_t>(0, back_buf - min_back);
  if (have_front) front_buf = std::max<int64_t>(0, front_buf - max_front);
  const int64_t desc_size = (int64_t)(top_end - tlas_offset);
  const std::string pre(prefix);
  if (!write_file(pre + ".asmain", A + tlas_offset, (uint64_t)desc_size, "wb")) return RTX_ERR_INVALID_ARGUMENT;
  if (have_back) {
    const uint64_t from = (uint64_t)((int64_t)tlas_offset + max_back);
    const uint64_t n = (uint64_t)(min_back - max_back + back_buf);
    if (!inside(from, n, arena_bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
    if (!write_file(pre + ".asback", A + from, n, "wb")) return RTX_ERR_INVALID_ARGUMENT;
  }
  if (have_front) {
    const uint64_t from = tlas_offset + (uint64_t)min_front;
    const uint64_t n = (uint64_t)(max_front - min_front + front_buf);
    if (!inside(from, n, arena_bytes)) return RTX_ERR_OFFSET_OUT_OF_ARENA;
    if (!write_file(pre + ".asfront", A + from, n, "wb")) return RTX_ERR_INVALID_ARGUMENT;
  }
  char meta[256];
  const int k = std::snprintf(meta, sizeof meta, "%" PRId64 ",%d,%" PRId64 ",%" PRId64 ",%" PRId64 ",%" PRId64
                                                  ",%" PRId64 ",%" PRId64 ",%d,%d",
                              desc_size, 1000150000, max_back, min_back, min_front, max_front, back_buf, front_buf,
                              have_back ? 1 : 0, have_front ? 1 : 0);
  if (!write_file(pre + ".asmetadata", (const uint8_t*)meta, (uint64_t)k, "w")) return RTX_ERR_INVALID_ARGUMENT;
  return RTX_OK;
}

rtx_status rtx_as_dump_info(const char* prefix, uint64_t* arena_bytes, uint64_t* tlas_offset) {
  if (!prefix || !arena_bytes || !tlas_offset) return RTX_ERR_INVALID_ARGUMENT;
  DumpMeta m;
  rtx_status st = read_meta(prefix, m);
  if (st != RTX_OK) return st;
  *arena_bytes = (uint64_t)(m.hi - m.lo);
  *tlas_offset = (uint64_t)(-m.lo);
  return RTX_OK;
}

rtx_status rtx_as_dump_read(const char* prefix, void* arena_host, uint64_t capacity) {
  if (!prefix || !arena_host) return RTX_ERR_INVALID_ARGUMENT;
  DumpMeta m;
  rtx_status st = read_meta(prefix, m);
  if (st != RTX_OK) return st;
  const uint64_t bytes = (uint64_t)(m.hi - m.lo);
  if (capacity < bytes) return RTX_ERR_INVALID_ARGUMENT;
  uint8_t* A = (uint8_t*)arena_host;
  std::memset(A, 0, bytes);
  const std::string pre(prefix);
  std::vector<uint8_t> f;
  auto place = [&](const char* ext, int64_t at, int64_t n) -> rtx_status {
    if (!read_file(pre + ext, f) || (int64_t)f.size() != n) return RTX_ERR_INVALID_ARGUMENT;
    const int64_t pos = at - m.lo;
    if (pos < 0 || pos + n > (int64_t)bytes) return RTX_ERR_OFFSET_OUT_OF_ARENA;
    if (n) std::memcpy(A + pos, f.data(), (size_t)n);
    return RTX_OK;
  };
  if (m.have_back && (st = place(".asback", m.max_back, m.min_back - m.max_back + m.back_buf)) != RTX_OK) return st;
  if ((st = place(".asmain", 0, m.desc_size)) != RTX_OK) return st;
  if (m.have_front && (st = place(".asfront", m.min_front, m.max_front - m.min_front + m.front_buf)) != RTX_OK)
    return st;
  return RTX_OK;
}

}  // extern "C"
