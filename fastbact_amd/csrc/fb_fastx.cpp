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
#include <vector>

#include "fastbact_gpu.h"

struct fb_fastx {
  std::vector<uint8_t> seq, qual;
  std::vector<uint64_t> soff, qoff;
  std::vector<uint32_t> slen, qlen;
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
