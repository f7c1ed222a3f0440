// .fs3d world files: a 64-byte little-endian header and the raw grid (host side).
//
// The reference keeps no simulation state on disk (its assets are OBJ/MTL/PNG
// data); this is the format the library checkpoints a world to and resumes it
// from.  Layout (all fields little-endian, the host's order on x86-64):
//    0 "FS3DGRID" | 8 u32 version | 12 u32 header bytes | 16 i32 nx, ny, nz
//   28 u32 flags | 32 u64 tick | 40 u64 rule hash | 48 u64 state hash
//   56 u64 payload bytes | 64 payload v[z][y][x]
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "fs3d_internal.h"

namespace fs3d {

namespace {

constexpr char kMagic[8] = {'F', 'S', '3', 'D', 'G', 'R', 'I', 'D'};
constexpr uint32_t kHeaderBytes = 64;

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t dims[3];
  uint32_t flags;
  uint64_t tick;
  uint64_t rule_hash;
  uint64_t state_hash;
  uint64_t payload;
};
static_assert(sizeof(Header) == kHeaderBytes, "header layout");

struct File {
  std::FILE *f = nullptr;
  explicit File(std::FILE *p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

u64 rule_hash() {
  u64 h = 0;
  for (int v = 0; v < 4; v++)
    for (u32 s = 0; s < 65536u; s++)
      h += mix64((((u64)v << 16 | s) << 16) | rule_state_packed((u16)s, v));
  return h;
}

int file_write(const char *path, const int32_t dims[3], u64 tick, u64 state_hash, const u8 *grid, size_t nbytes,
               int (*fill)(void *ctx, size_t off, u8 *buf, size_t n), void *ctx) {
  if (!path || !dims) return set_error(FS3D_ERR_INVALID, "path/dims is NULL");
  for (int a = 0; a < 3; a++)
    if (dims[a] < 1) return set_error(FS3D_ERR_INVALID, "dims must be >= 1");
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (nbytes != n) return set_error(FS3D_ERR_INVALID, "grid has %zu bytes, dims need %zu", nbytes, n);
  Header h;
  std::memcpy(h.magic, kMagic, 8);
  h.version = FS3D_FILE_VERSION;
  h.header_bytes = kHeaderBytes;
  std::memcpy(h.dims, dims, sizeof h.dims);
  h.flags = 0;
  h.tick = tick;
  h.rule_hash = rule_hash();
  h.state_hash = state_hash;
  h.payload = n;
  File f(std::fopen(path, "wb"));
  if (!f.f) return set_error(FS3D_ERR_INVALID, "cannot open %s for writing", path);
  if (std::fwrite(&h, sizeof h, 1, f.f) != 1) return set_error(FS3D_ERR_INVALID, "write failed: %s", path);
  if (grid) {
    if (n && std::fwrite(grid, 1, n, f.f) != n) return set_error(FS3D_ERR_INVALID, "write failed: %s", path);
  } else {
    // streamed: fill() provides the payload chunk by chunk
    std::vector<u8> buf(std::min<size_t>(n, (size_t)64 << 20));
    for (size_t off = 0; off < n;) {
      const size_t c = std::min(buf.size(), n - off);
      if (int rc = fill(ctx, off, buf.data(), c)) return rc;
      if (std::fwrite(buf.data(), 1, c, f.f) != c) return set_error(FS3D_ERR_INVALID, "write failed: %s", path);
      off += c;
    }
  }
  if (std::fclose(f.f) != 0) {
    f.f = nullptr;
    return set_error(FS3D_ERR_INVALID, "close failed: %s", path);
  }
  f.f = nullptr;
  return FS3D_OK;
}

int file_info(const char *path, int32_t dims[3], u64 *tick, u64 *state_hash, u64 *rule) {
  if (!path) return set_error(FS3D_ERR_INVALID, "path is NULL");
  File f(std::fopen(path, "rb"));
  if (!f.f) return set_error(FS3D_ERR_INVALID, "cannot open %s", path);
  Header h;
  if (std::fread(&h, sizeof h, 1, f.f) != 1) return set_error(FS3D_ERR_INVALID, "%s: truncated header", path);
  if (std::memcmp(h.magic, kMagic, 8) != 0) return set_error(FS3D_ERR_INVALID, "%s: not an fs3d world file", path);
  if (h.version != FS3D_FILE_VERSION || h.header_bytes != kHeaderBytes)
    return set_error(FS3D_ERR_INVALID, "%s: unsupported format version %u", path, h.version);
  for (int a = 0; a < 3; a++)
    if (h.dims[a] < 1) return set_error(FS3D_ERR_INVALID, "%s: bad dims", path);
  if (h.payload != (u64)h.dims[0] * h.dims[1] * h.dims[2])
    return set_error(FS3D_ERR_INVALID, "%s: payload size does not match dims", path);
  // the file holds exactly the header and the payload
  if (fseeko(f.f, 0, SEEK_END) != 0) return set_error(FS3D_ERR_INVALID, "cannot seek %s", path);
  const off_t len = ftello(f.f);
  if (len < 0 || (u64)len != kHeaderBytes + h.payload)
    return set_error(FS3D_ERR_INVALID, "%s: %lld bytes, header and payload need %llu", path, (long long)len,
                     (unsigned long long)(kHeaderBytes + h.payload));
  if (dims) std::memcpy(dims, h.dims, sizeof h.dims);
  if (tick) *tick = h.tick;
  if (state_hash) *state_hash = h.state_hash;
  if (rule) *rule = h.rule_hash;
  return FS3D_OK;
}

int file_read(const char *path, u8 *grid, size_t nbytes, int (*sink)(void *ctx, size_t off, const u8 *buf, size_t n),
              void *ctx) {
  int32_t dims[3];
  u64 rule = 0;
  if (int rc = file_info(path, dims, nullptr, nullptr, &rule)) return rc;
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (nbytes != n) return set_error(FS3D_ERR_INVALID, "%s holds %zu bytes, buffer has %zu", path, n, nbytes);
  if (rule != rule_hash())
    return set_error(FS3D_ERR_STATE, "%s was written under a different tick rule (rule hash %016llx)", path,
                     (unsigned long long)rule);
  File f(std::fopen(path, "rb"));
  if (!f.f || std::fseek(f.f, kHeaderBytes, SEEK_SET) != 0) return set_error(FS3D_ERR_INVALID, "cannot read %s", path);
  if (grid) {
    if (n && std::fread(grid, 1, n, f.f) != n) return set_error(FS3D_ERR_INVALID, "%s: truncated payload", path);
  } else {
    std::vector<u8> buf(std::min<size_t>(n, (size_t)64 << 20));
    for (size_t off = 0; off < n;) {
      const size_t c = std::min(buf.size(), n - off);
      if (std::fread(buf.data(), 1, c, f.f) != c) return set_error(FS3D_ERR_INVALID, "%s: truncated payload", path);
      if (int rc = sink(ctx, off, buf.data(), c)) return rc;
      off += c;
    }
  }
  return FS3D_OK;
}

}  // namespace fs3d

using namespace fs3d;

extern "C" {

int fs3d_rule_hash(uint64_t *hash) {
  if (!hash) return set_error(FS3D_ERR_INVALID, "hash is NULL");
  *hash = rule_hash();
  return clear_error();
}

int fs3d_file_write(const char *path, const int32_t dims[3], uint64_t tick, uint64_t state_hash, const uint8_t *grid,
                    size_t nbytes) {
  if (!grid) return set_error(FS3D_ERR_INVALID, "grid is NULL");
  if (int rc = file_write(path, dims, tick, state_hash, grid, nbytes, nullptr, nullptr)) return rc;
  return clear_error();
}

int fs3d_file_info(const char *path, int32_t dims[3], uint64_t *tick, uint64_t *state_hash) {
  if (int rc = file_info(path, dims, tick, state_hash, nullptr)) return rc;
  return clear_error();
}

int fs3d_file_read(const char *path, uint8_t *grid, size_t nbytes) {
  if (!grid) return set_error(FS3D_ERR_INVALID, "grid is NULL");
  if (int rc = file_read(path, grid, nbytes, nullptr, nullptr)) return rc;
  return clear_error();
}

}  // extern "C"
