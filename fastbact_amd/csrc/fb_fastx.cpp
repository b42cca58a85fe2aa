// fb_fastx.cpp -- FASTQ / FASTA reader for the ingest front end
// (open_fastx, src/ingest/trim_merge.rs:233-252).
//
// The whole file is read into memory (gzip through zlib when the name ends
// in ".gz", every member of a multi-member file), then split into records:
//   FASTQ: "@header", the sequence line, "+..." and the quality line; the
//          quality must be as long as the sequence;
//   FASTA: ">header" and the sequence lines up to the next '>', joined.
// Line ends may be "\n" or "\r\n".  Records keep their bytes as written
// (case, N); the 2-bit packing maps them later.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <thread>
#include <vector>

#include "fastbact_gpu.h"
#include "fb_host.hpp"

struct fb_fastx {
  std::vector<uint8_t> seq, qual, ids;
  std::vector<uint64_t> soff, qoff, ioff;
  std::vector<uint32_t> slen, qlen, ilen;
  void add_id(const uint8_t* b, const uint8_t* e) {  // the header line after '@' / '>'
    ioff.push_back(ids.size());
    ilen.push_back((uint32_t)(e - b));
    ids.insert(ids.end(), b, e);
  }
};

static thread_local std::string g_fastx_err;

static int fastx_fail(const std::string& msg) {
  g_fastx_err = msg;
  return FB_ERR_INVALID;
}

static bool ends_with(const std::string& s, const char* suf) {
  const size_t n = strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

static int read_all(const std::string& path, std::vector<uint8_t>* buf) {
  if (ends_with(path, ".gz")) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return fastx_fail("Failed to open file \"" + path + "\"");
    gzbuffer(f, 256 * 1024);
    std::vector<uint8_t> chunk(1 << 20);
    for (;;) {
      const int n = gzread(f, chunk.data(), (unsigned)chunk.size());
      if (n < 0) {
        int e = 0;
        const std::string m = gzerror(f, &e);
        gzclose(f);
        return fastx_fail("Failed to parse gzip FASTX: " + m);
      }
      if (n == 0) break;
      buf->insert(buf->end(), chunk.data(), chunk.data() + n);
    }
    gzclose(f);
    return FB_OK;
  }
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return fastx_fail("Failed to open file \"" + path + "\"");
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf->resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = buf->empty() ? 0 : fread(buf->data(), 1, buf->size(), f);
  fclose(f);
  if (got != buf->size()) return fastx_fail("Failed to read file \"" + path + "\"");
  return FB_OK;
}

// One line [p, e) without its "\r\n" / "\n"; advances p past the line end.
struct LineReader {
  const uint8_t* p;
  const uint8_t* end;
  bool next(const uint8_t** b, const uint8_t** e) {
    if (p >= end) return false;
    const uint8_t* nl = static_cast<const uint8_t*>(memchr(p, '\n', (size_t)(end - p)));
    const uint8_t* le = nl ? nl : end;
    *b = p;
    *e = (le > p && le[-1] == '\r') ? le - 1 : le;
    p = nl ? nl + 1 : end;
    return true;
  }
};

static int parse(const std::vector<uint8_t>& buf, fb_fastx* x, const std::string& path) {
  if (buf.empty()) return fastx_fail("Failed to parse FASTX \"" + path + "\": the file is empty");
  const uint8_t kind = buf[0];
  if (kind != '@' && kind != '>')
    return fastx_fail("Failed to parse FASTX \"" + path + "\": unknown format (first byte is not '@' or '>')");
  LineReader lr{buf.data(), buf.data() + buf.size()};
  const uint8_t *b = nullptr, *e = nullptr;
  uint64_t rec = 0;
  if (kind == '@') {
    while (lr.next(&b, &e)) {
      if (b == e && lr.p >= lr.end) break;  // a final empty line
      if (b == e || *b != '@')
        return fastx_fail("Corrupted record " + std::to_string(rec + 1) + " in \"" + path + "\": expected '@'");
      const uint8_t *sb, *se, *pb, *pe, *qb, *qe;
      if (!lr.next(&sb, &se) || !lr.next(&pb, &pe) || pb == pe || *pb != '+' || !lr.next(&qb, &qe))
        return fastx_fail("Corrupted record " + std::to_string(rec + 1) + " in \"" + path + "\": truncated");
      if (se - sb != qe - qb)
        return fastx_fail("Corrupted record " + std::to_string(rec + 1) + " in \"" + path +
                          "\": sequence and quality lengths differ");
      x->add_id(b + 1, e);
      x->soff.push_back(x->seq.size());
      x->slen.push_back((uint32_t)(se - sb));
      x->seq.insert(x->seq.end(), sb, se);
      x->qoff.push_back(x->qual.size());
      x->qlen.push_back((uint32_t)(qe - qb));
      x->qual.insert(x->qual.end(), qb, qe);
      ++rec;
    }
  } else {
    bool open = false;
    while (lr.next(&b, &e)) {
      if (b < e && *b == '>') {
        x->add_id(b + 1, e);
        x->soff.push_back(x->seq.size());
        x->slen.push_back(0);
        x->qoff.push_back(x->qual.size());
        x->qlen.push_back(0);
        open = true;
        continue;
      }
      if (!open) return fastx_fail("Corrupted record in \"" + path + "\": expected '>'");
      x->seq.insert(x->seq.end(), b, e);
      x->slen.back() += (uint32_t)(e - b);
    }
  }
  return FB_OK;
}

extern "C" int fb_fastx_read(const char* path, fb_fastx** out) {
  if (!path || !out) return fastx_fail("null argument");
  *out = nullptr;
  std::vector<uint8_t> buf;
  if (int r = read_all(path, &buf)) return r;
  fb_fastx* x = new fb_fastx();
  if (int r = parse(buf, x, path)) {
    delete x;
    return r;
  }
  *out = x;
  return FB_OK;
}

extern "C" const char* fb_fastx_last_error(void) { return g_fastx_err.c_str(); }
extern "C" void fb_fastx_destroy(fb_fastx* x) { delete x; }
extern "C" uint64_t fb_fastx_count(const fb_fastx* x) { return x ? x->slen.size() : 0; }

extern "C" int fb_fastx_view(const fb_fastx* x, fb_records_view* v) {
  if (!x || !v) return FB_ERR_INVALID;
  v->seq = x->seq.data();
  v->seq_offsets = x->soff.data();
  v->seq_lengths = x->slen.data();
  v->qual = x->qual.data();
  v->qual_offsets = x->qoff.data();
  v->qual_lengths = x->qlen.data();
  v->n = x->slen.size();
  return FB_OK;
}

namespace fb {

int fastx_read_pair(const char* r1, const char* r2, fb_fastx** a, fb_fastx** b, std::string* err) {
  *a = *b = nullptr;
  int rc[2] = {FB_OK, FB_OK};
  std::string e[2];
  std::thread t([&] {
    rc[1] = fb_fastx_read(r2, b);
    if (rc[1]) e[1] = fb_fastx_last_error();
  });
  rc[0] = fb_fastx_read(r1, a);
  if (rc[0]) e[0] = fb_fastx_last_error();
  t.join();
  for (int i = 0; i < 2; ++i)
    if (rc[i]) {
      fb_fastx_destroy(*a);
      fb_fastx_destroy(*b);
      *a = *b = nullptr;
      *err = e[i];
      return rc[i];
    }
  return FB_OK;
}

// extract_read_id_prefix (estimator.rs:89-99): up to the first ASCII
// whitespace, less a trailing "/1" or "/2".
static std::string id_prefix(const uint8_t* p, uint32_t n) {
  uint32_t e = 0;
  while (e < n && p[e] != ' ' && p[e] != '\t' && p[e] != '\n' && p[e] != '\r' && p[e] != '\f') ++e;
  if (e >= 2 && p[e - 2] == '/' && (p[e - 1] == '1' || p[e - 1] == '2')) e -= 2;
  return std::string(reinterpret_cast<const char*>(p), e);
}

int fastx_count(const fb_fastx* a, const fb_fastx* b, fb_count_stats* out, std::string* err) {
  const uint64_t n1 = a->slen.size(), n2 = b->slen.size(), n = n1 < n2 ? n1 : n2;
  for (uint64_t i = 0; i < n; ++i) {
    const std::string p1 = id_prefix(a->ids.data() + a->ioff[i], a->ilen[i]);
    const std::string p2 = id_prefix(b->ids.data() + b->ioff[i], b->ilen[i]);
    if (p1 != p2) {
      *err = "FASTQ paired-end synchronization error at pair #" + std::to_string(i + 1) + ": R1 ID '" +
             std::string(reinterpret_cast<const char*>(a->ids.data() + a->ioff[i]), a->ilen[i]) +
             "' does not match R2 ID '" +
             std::string(reinterpret_cast<const char*>(b->ids.data() + b->ioff[i]), b->ilen[i]) + "'";
      return FB_ERR_INVALID;
    }
  }
  if (n1 > n2) {
    *err = "FASTQ pairing mismatch: R1 has more reads than R2 (R2 ended prematurely after " + std::to_string(n) +
           " pairs)";
    return FB_ERR_INVALID;
  }
  if (n2 > n1) {
    *err = "FASTQ pairing mismatch: R2 has more reads than R1 (R1 ended prematurely after " + std::to_string(n) +
           " pairs)";
    return FB_ERR_INVALID;
  }
  const int r = fb_count_stats_from_lengths(a->slen.data(), b->slen.data(), n, out);
  if (r) *err = "Input FASTQ files contain 0 valid read pairs";
  return r;
}

}  // namespace fb

extern "C" int fb_count_fastx(const char* r1_path, const char* r2_path, fb_count_stats* out) {
  if (!r1_path || !r2_path || !out) return fastx_fail("null argument");
  fb_fastx *a = nullptr, *b = nullptr;
  std::string err;
  if (int r = fb::fastx_read_pair(r1_path, r2_path, &a, &b, &err)) return fastx_fail(err), r;
  const int r = fb::fastx_count(a, b, out, &err);
  fb_fastx_destroy(a);
  fb_fastx_destroy(b);
  if (r) fastx_fail(err);
  return r;
}
