// Readers and writers of the reference's on-disk table and index files.
//
// Table file (joins/hashJoin_gpu/main_street/TableBuild.cpp:113-141 and the
// sort_index variant): an int32 row count n, then n TUPLEs {int32 key (row
// id), int32 val (join attribute)}.  The joins/join_ori data files are the
// same tuples without the count (EO_REF_FILE_RAW).  Index file
// (joins/innerJoin_gpu/sort_index/main_street/TableBuild.cpp:180-218,
// createIndex at :77-95): n BUCKETs {int32 val, int32 adr} sorted by val, no
// header.  In memory the library's eo_tuple is {key = val, payload = row id /
// adr}, so reading and writing swap the two fields of a TUPLE and keep those
// of a BUCKET.  Host-only code: no device is needed.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"

namespace eo {
namespace {

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

eo_status open_file(const char *path, const char *mode, File *fp) {
  fp->f = std::fopen(path, mode);
  if (!fp->f) {
    set_last_error(std::string("cannot open ") + path);
    return EO_EINVAL;
  }
  return EO_OK;
}

eo_status file_size(FILE *f, uint64_t *bytes) {
  if (std::fseek(f, 0, SEEK_END) != 0) return EO_EINVAL;
  const long sz = std::ftell(f);
  if (sz < 0 || std::fseek(f, 0, SEEK_SET) != 0) return EO_EINVAL;
  *bytes = (uint64_t)sz;
  return EO_OK;
}

// Reads n 8-byte records into out (as eo_tuple), swapping the two fields when asked.
eo_status read_records(FILE *f, eo_tuple *out, uint64_t n, bool swap) {
  const uint64_t chunk = 1u << 20;
  for (uint64_t i = 0; i < n; i += chunk) {
    const uint64_t m = n - i < chunk ? n - i : chunk;
    if (std::fread(out + i, sizeof(eo_tuple), m, f) != m) {
      set_last_error("short read");
      return EO_EINVAL;
    }
    if (swap)
      for (uint64_t j = i; j < i + m; ++j) {
        const int32_t k = out[j].key;
        out[j].key = out[j].payload;
        out[j].payload = k;
      }
  }
  return EO_OK;
}

eo_status write_records(FILE *f, const eo_tuple *t, uint64_t n, bool swap) {
  std::vector<eo_tuple> buf;
  const uint64_t chunk = 1u << 20;
  for (uint64_t i = 0; i < n; i += chunk) {
    const uint64_t m = n - i < chunk ? n - i : chunk;
    const eo_tuple *src = t + i;
    if (swap) {
      buf.resize(m);
      for (uint64_t j = 0; j < m; ++j) buf[j] = eo_tuple{t[i + j].payload, t[i + j].key};
      src = buf.data();
    }
    if (std::fwrite(src, sizeof(eo_tuple), m, f) != m) {
      set_last_error("short write");
      return EO_EINVAL;
    }
  }
  return EO_OK;
}

}  // namespace
}  // namespace eo

using namespace eo;

extern "C" {

eo_status eo_ref_table_read(const char *path, int flags, eo_tuple *out, uint64_t cap, uint64_t *n) {
  if (!path || !n || (cap && !out) || (flags & ~EO_REF_FILE_RAW)) return EO_EINVAL;
  File fp;
  eo_status st = open_file(path, "rb", &fp);
  if (st != EO_OK) return st;
  uint64_t bytes = 0;
  if ((st = file_size(fp.f, &bytes)) != EO_OK) return st;
  uint64_t rows = 0;
  if (flags & EO_REF_FILE_RAW) {
    if (bytes % sizeof(eo_tuple)) {
      set_last_error("file size is not a multiple of 8 bytes");
      return EO_EINVAL;
    }
    rows = bytes / sizeof(eo_tuple);
  } else {
    int32_t hdr = 0;
    if (bytes < 4 || std::fread(&hdr, sizeof(hdr), 1, fp.f) != 1 || hdr < 0 ||
        (uint64_t)hdr * sizeof(eo_tuple) + 4 != bytes) {
      set_last_error("not a table file: the int32 row count does not match the file size");
      return EO_EINVAL;
    }
    rows = (uint64_t)hdr;
  }
  *n = rows;
  if (!out) return EO_OK;
  if (cap < rows) return EO_ECAPACITY;
  return read_records(fp.f, out, rows, true);
}

eo_status eo_ref_table_write(const char *path, int flags, const eo_tuple *t, uint64_t n) {
  if (!path || (n && !t) || (flags & ~EO_REF_FILE_RAW) || n > (uint64_t)INT32_MAX) return EO_EINVAL;
  File fp;
  eo_status st = open_file(path, "wb", &fp);
  if (st != EO_OK) return st;
  if (!(flags & EO_REF_FILE_RAW)) {
    const int32_t hdr = (int32_t)n;
    if (std::fwrite(&hdr, sizeof(hdr), 1, fp.f) != 1) return EO_EINVAL;
  }
  st = write_records(fp.f, t, n, true);
  if (st != EO_OK) return st;
  const int rc = std::fclose(fp.f);
  fp.f = nullptr;
  return rc == 0 ? EO_OK : EO_EINVAL;
}

eo_status eo_ref_index_read(const char *path, eo_tuple *out, uint64_t cap, uint64_t *n) {
  if (!path || !n || (cap && !out)) return EO_EINVAL;
  File fp;
  eo_status st = open_file(path, "rb", &fp);
  if (st != EO_OK) return st;
  uint64_t bytes = 0;
  if ((st = file_size(fp.f, &bytes)) != EO_OK) return st;
  if (bytes % sizeof(eo_tuple)) {
    set_last_error("index file size is not a multiple of 8 bytes");
    return EO_EINVAL;
  }
  *n = bytes / sizeof(eo_tuple);
  if (!out) return EO_OK;
  if (cap < *n) return EO_ECAPACITY;
  return read_records(fp.f, out, *n, false);
}

eo_status eo_ref_index_write(const char *path, const eo_tuple *sorted, uint64_t n) {
  if (!path || (n && !sorted)) return EO_EINVAL;
  File fp;
  eo_status st = open_file(path, "wb", &fp);
  if (st != EO_OK) return st;
  st = write_records(fp.f, sorted, n, false);
  if (st != EO_OK) return st;
  const int rc = std::fclose(fp.f);
  fp.f = nullptr;
  return rc == 0 ? EO_OK : EO_EINVAL;
}

}  // extern "C"
