// fb_runtime.hip -- context, device memory, objects and the C ABI of
// include/fastbact_gpu.h.  One HIP stream per context; device buffers come
// from a per-context pool so repeated passes reuse memory instead of calling
// hipMalloc/hipFree.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <tuple>
#include <condition_variable>
#include <map>
#include <thread>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fastbact_gpu.h"
#include "fb_host.hpp"
#include "fb_internal.hpp"

using namespace fb;

// ---------------------------------------------------------------------------
// Context and pool

struct PoolBuf {
  void* p;
  size_t bytes;
};

struct fb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;  // 0 = unlimited
  uint64_t held = 0;    // bytes allocated from HIP (in use + pooled)
  uint64_t pooled = 0;  // bytes of the free buffers in the pool
  uint64_t pool_cap = 0;  // most free bytes the pool keeps (the longest-unused go first); 0 = no limit
  std::string err;
  std::vector<PoolBuf> pool;  // free buffers
  hipEvent_t ev[12] = {};
  unsigned char* scratch = nullptr;  // this device's copy of `scratch` below
};

// Static device scratch of the entry points that allocate nothing
// (fb_cms_add_and_estimate_static).  Each device holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (static_scratch_of), never
// through the array's name.  Every call on a device shares that copy.
constexpr uint64_t kStaticMaxKeys = 1ull << 18;            // keys per call
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;   // 16 MiB: 8 MiB of arrays, 8 MiB of sort space
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static std::mutex g_scratch_mu;
static std::map<int, unsigned char*> g_scratch;  // device -> address of its copy

// The current device's copy of `scratch`, looked up once per device; called
// with that device current (context creation).
static hipError_t static_scratch_of(int device, unsigned char** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it == g_scratch.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = g_scratch.emplace(device, static_cast<unsigned char*>(p)).first;
  }
  *out = it->second;
  return hipSuccess;
}

static int set_err(fb_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIP_TRY(ctx, expr)                                                                              \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return set_err((ctx), FB_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                         __FILE__, __LINE__);                                           \
  } while (0)

static int ctx_current(fb_ctx* c) {
  HIP_TRY(c, hipSetDevice(c->device));
  return FB_OK;
}

// Takes a pooled buffer of at least `bytes` (and at most 2x) or allocates.
static int pool_get(fb_ctx* c, size_t bytes, void** out) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  size_t best = (size_t)-1;
  for (size_t i = 0; i < c->pool.size(); ++i) {
    const PoolBuf& b = c->pool[i];
    if (b.bytes >= bytes && b.bytes <= 2 * bytes && (best == (size_t)-1 || b.bytes < c->pool[best].bytes)) best = i;
  }
  if (best != (size_t)-1) {
    *out = c->pool[best].p;
    c->pooled -= c->pool[best].bytes;
    c->pool.erase(c->pool.begin() + best);
    return FB_OK;
  }
  // free pooled buffers first, the longest-unused first, if the budget would
  // be exceeded
  size_t drop = 0;
  uint64_t held = c->held;
  while (c->budget && held + bytes > c->budget && drop < c->pool.size()) {
    (void)hipFree(c->pool[drop].p);
    held -= c->pool[drop].bytes;
    ++drop;
  }
  if (drop) {
    c->pooled -= c->held - held;
    c->pool.erase(c->pool.begin(), c->pool.begin() + drop);
    c->held = held;
  }
  if (c->budget && c->held + bytes > c->budget)
    return set_err(c, FB_ERR_OOM, "HBM budget exceeded: need %zu bytes, %llu of %llu held", bytes,
                   (unsigned long long)c->held, (unsigned long long)c->budget);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // release the pool and retry once
    for (auto& b : c->pool) {
      (void)hipFree(b.p);
      c->held -= b.bytes;
    }
    c->pool.clear();
    c->pooled = 0;
    (void)hipGetLastError();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return set_err(c, FB_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
  }
  c->held += bytes;
  *out = p;
  return FB_OK;
}
static void pool_put(fb_ctx* c, void* p, size_t bytes) {
  if (!p) return;
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  c->pool.push_back({p, bytes});
  c->pooled += bytes;
  size_t drop = 0;
  while (c->pool_cap && c->pooled > c->pool_cap && drop + 1 < c->pool.size()) {
    (void)hipFree(c->pool[drop].p);
    c->held -= c->pool[drop].bytes;
    c->pooled -= c->pool[drop].bytes;
    ++drop;
  }
  if (drop) c->pool.erase(c->pool.begin(), c->pool.begin() + drop);
}

// RAII device buffer from the pool.
struct DBuf {
  fb_ctx* c = nullptr;
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  int get(fb_ctx* ctx, size_t n) {
    release();
    c = ctx;
    bytes = n;
    return pool_get(ctx, n, &p);
  }
  void release() {
    if (p && c) pool_put(c, p, bytes);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

extern "C" int fb_ctx_create(int device, uint64_t budget, fb_ctx** out) {
  if (!out) return FB_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return FB_ERR_NODEV;
  }
  if (device < 0 || device >= n) return FB_ERR_INVALID;
  fb_ctx* c = new fb_ctx();
  c->device = device;
  c->budget = budget;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FB_ERR_HIP;
  }
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return FB_ERR_HIP;
    }
  }
  if (static_scratch_of(device, &c->scratch) != hipSuccess) {
    (void)hipGetLastError();
    fb_ctx_destroy(c);
    return FB_ERR_HIP;
  }
  *out = c;
  return FB_OK;
}

extern "C" void fb_ctx_destroy(fb_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->pool) (void)hipFree(b.p);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* fb_last_error(const fb_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int fb_ctx_synchronize(fb_ctx* c) {
  if (!c) return FB_ERR_INVALID;
  if (int r = ctx_current(c)) return r;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

extern "C" uint64_t fb_ctx_device_bytes(const fb_ctx* c) { return c ? c->held : 0; }

extern "C" int fb_ctx_trim(fb_ctx* c) {
  if (!c) return FB_ERR_INVALID;
  if (int r = ctx_current(c)) return r;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto& b : c->pool) {
    (void)hipFree(b.p);
    c->held -= std::min(c->held, (uint64_t)b.bytes);
  }
  c->pool.clear();
  c->pooled = 0;
  return FB_OK;
}

extern "C" int fb_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return FB_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return FB_ERR_NODEV;
  }
  size_t fr = 0, tot = 0;
  if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return FB_ERR_HIP;
  }
  *free_bytes = fr;
  *total_bytes = tot;
  return FB_OK;
}

extern "C" int fb_diag_copy(fb_ctx* c, uint64_t bytes) {
  if (!c) return FB_ERR_INVALID;
  if (int r = ctx_current(c)) return r;
  const uint64_t n16 = (bytes + 15) / 16;
  if (!n16) return FB_OK;
  DBuf src, dst;
  if (int r = src.get(c, n16 * 16)) return r;
  if (int r = dst.get(c, n16 * 16)) return r;
  hipStream_t s = c->stream;
  HIP_TRY(c, launch_fill_u32(src.as<uint32_t>(), n16 * 4, 0x5a5a5a5au, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  HIP_TRY(c, launch_copy_calibration(src.p, dst.p, n16, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return FB_OK;
}

extern "C" int fb_diag_atomic_rate(fb_ctx* c, uint32_t log2_counters, uint64_t n_updates, int mode, float* ms) {
  if (!c || !ms || log2_counters < 10 || log2_counters > 30 || mode < 0 || mode > 3)
    return set_err(c, FB_ERR_INVALID, "bad argument");
  if (int r = ctx_current(c)) return r;
  const uint64_t n = 1ull << log2_counters;
  DBuf table, sink;
  if (int r = table.get(c, n * 4)) return r;
  if (int r = sink.get(c, 4)) return r;
  hipStream_t s = c->stream;
  HIP_TRY(c, launch_fill_u32(table.as<uint32_t>(), n, 0, s));
  // one launch to bring the table into whatever cache level holds it, then the timed one
  HIP_TRY(c, launch_atomic_rate(table.as<uint32_t>(), log2_counters, n_updates, mode, sink.as<uint32_t>(), s));
  HIP_TRY(c, hipEventRecord(c->ev[0], s));
  HIP_TRY(c, launch_atomic_rate(table.as<uint32_t>(), log2_counters, n_updates, mode, sink.as<uint32_t>(), s));
  HIP_TRY(c, hipEventRecord(c->ev[1], s));
  HIP_TRY(c, hipEventSynchronize(c->ev[1]));
  HIP_TRY(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
  return FB_OK;
}

// ---------------------------------------------------------------------------
// ReadBank

struct fb_bank {
  fb_ctx* ctx = nullptr;
  DBuf data, offs, lens;
  uint64_t n_reads = 0;
  uint64_t total_bases = 0;
  std::vector<uint64_t> len_hist;  // bases summed per read length (index = length), last bin open
  uint32_t max_len = 0;
  // the three arenas (merged, pe_forward, pe_reverse) inside the device list:
  // first read, read count, first word, words, and BitSeqArena::total_bases
  // (bitseq.rs:185-192: last offset + last length, arena-local)
  uint64_t ar_first[3] = {0, 0, 0}, ar_reads[3] = {0, 0, 0};
  uint64_t ar_word0[3] = {0, 0, 0}, ar_words[3] = {0, 0, 0}, ar_bases[3] = {0, 0, 0};
  DevReads view() const {
    DevReads r;
    r.data = data.as<uint64_t>();
    r.offs = offs.as<uint64_t>();
    r.lens = lens.as<uint32_t>();
    r.n_reads = n_reads;
    return r;
  }
};

// Concatenates arenas word-aligned into one device read list (offsets become
// global base indices) with kArenaPadWords zero words after the data.
static int upload_reads(fb_ctx* c, const fb_arena_view* arenas, int n_arenas, DBuf& data, DBuf& offs, DBuf& lens,
                        uint64_t* n_reads_out, std::vector<uint64_t>* hist, uint32_t* max_len, uint64_t* bases) {
  uint64_t words = 0, nr = 0;
  for (int a = 0; a < n_arenas; ++a) {
    const fb_arena_view& v = arenas[a];
    if (v.n_reads && (!v.offsets || !v.lengths || !v.data))
      return set_err(c, FB_ERR_INVALID, "arena %d: null pointer with %llu reads", a, (unsigned long long)v.n_reads);
    // every read must lie inside data
    for (uint64_t i = 0; i < v.n_reads; ++i) {
      uint64_t end = v.offsets[i] + v.lengths[i];
      if (v.lengths[i] && (end + 31) / 32 > v.n_words)
        return set_err(c, FB_ERR_INVALID, "arena %d read %llu ends past data (%llu words)", a, (unsigned long long)i,
                       (unsigned long long)v.n_words);
    }
    words += v.n_words;
    nr += v.n_reads;
  }
  if (int r = data.get(c, (words + kArenaPadWords) * 8)) return r;
  if (int r = offs.get(c, (nr ? nr : 1) * 8)) return r;
  if (int r = lens.get(c, (nr ? nr : 1) * 4)) return r;
  std::vector<uint64_t> h_offs(nr);
  std::vector<uint32_t> h_lens(nr);
  uint64_t wpos = 0, rpos = 0, tot = 0;
  uint32_t mx = 0;
  for (int a = 0; a < n_arenas; ++a) {
    const fb_arena_view& v = arenas[a];
    if (v.n_words) HIP_TRY(c, hipMemcpyAsync(data.as<uint64_t>() + wpos, v.data, v.n_words * 8, hipMemcpyHostToDevice, c->stream));
    for (uint64_t i = 0; i < v.n_reads; ++i) {
      h_offs[rpos] = v.offsets[i] + wpos * 32;
      h_lens[rpos] = v.lengths[i];
      tot += v.lengths[i];
      mx = std::max(mx, v.lengths[i]);
      ++rpos;
    }
    wpos += v.n_words;
  }
  HIP_TRY(c, hipMemsetAsync(data.as<uint64_t>() + wpos, 0, kArenaPadWords * 8, c->stream));
  if (nr) {
    HIP_TRY(c, hipMemcpyAsync(offs.p, h_offs.data(), nr * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(lens.p, h_lens.data(), nr * 4, hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (hist) {
    hist->assign((size_t)mx + 1, 0);
    for (uint64_t i = 0; i < nr; ++i) (*hist)[h_lens[i]] += h_lens[i];
  }
  if (max_len) *max_len = mx;
  if (bases) *bases = tot;
  *n_reads_out = nr;
  return FB_OK;
}

extern "C" int fb_bank_upload(fb_ctx* c, const fb_arena_view arenas[3], fb_bank** out) {
  if (!c || !arenas || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (int r = ctx_current(c)) return r;
  fb_bank* b = new fb_bank();
  b->ctx = c;
  int r = upload_reads(c, arenas, 3, b->data, b->offs, b->lens, &b->n_reads, &b->len_hist, &b->max_len, &b->total_bases);
  if (r) {
    delete b;
    return r;
  }
  uint64_t rpos = 0, wpos = 0;
  for (int a = 0; a < 3; ++a) {
    const fb_arena_view& v = arenas[a];
    b->ar_first[a] = rpos;
    b->ar_reads[a] = v.n_reads;
    b->ar_word0[a] = wpos;
    b->ar_words[a] = v.n_words;
    b->ar_bases[a] = v.n_reads ? v.offsets[v.n_reads - 1] + v.lengths[v.n_reads - 1] : 0;
    rpos += v.n_reads;
    wpos += v.n_words;
  }
  // ReadBank::total_bases (bitseq.rs:373-375) sums the arenas' last offset +
  // last length, which equals the sum of lengths for push_seq-built arenas
  b->total_bases = b->ar_bases[0] + b->ar_bases[1] + b->ar_bases[2];
  *out = b;
  return FB_OK;
}

extern "C" int fb_bank_arena_info(const fb_bank* b, int arena, uint64_t* n_reads, uint64_t* total_bases,
                                  uint64_t* n_words) {
  if (!b || arena < 0 || arena > 2) return FB_ERR_INVALID;
  if (n_reads) *n_reads = b->ar_reads[arena];
  if (total_bases) *total_bases = b->ar_bases[arena];
  if (n_words) *n_words = b->ar_words[arena];
  return FB_OK;
}

extern "C" int fb_bank_download_arena(const fb_bank* b, int arena, uint64_t* data, uint64_t* offsets,
                                      uint32_t* lengths) {
  if (!b || arena < 0 || arena > 2) return FB_ERR_INVALID;
  fb_ctx* c = b->ctx;
  if (int r = ctx_current(c)) return r;
  const uint64_t nr = b->ar_reads[arena], nw = b->ar_words[arena];
  if (data && nw)
    HIP_TRY(c, hipMemcpyAsync(data, b->data.as<uint64_t>() + b->ar_word0[arena], nw * 8, hipMemcpyDeviceToHost,
                              c->stream));
  if (lengths && nr)
    HIP_TRY(c, hipMemcpyAsync(lengths, b->lens.as<uint32_t>() + b->ar_first[arena], nr * 4, hipMemcpyDeviceToHost,
                              c->stream));
  if (offsets && nr)
    HIP_TRY(c, hipMemcpyAsync(offsets, b->offs.as<uint64_t>() + b->ar_first[arena], nr * 8, hipMemcpyDeviceToHost,
                              c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (offsets)  // device offsets are global base indices of the list
    for (uint64_t i = 0; i < nr; ++i) offsets[i] -= b->ar_word0[arena] * 32;
  return FB_OK;
}

extern "C" void fb_bank_destroy(fb_bank* b) { delete b; }
extern "C" uint64_t fb_bank_total_bases(const fb_bank* b) { return b ? b->total_bases : 0; }
extern "C" uint64_t fb_bank_total_reads(const fb_bank* b) { return b ? b->n_reads : 0; }
// Windows of a k pass: sum over reads with len >= k of len - k + 1.
static uint64_t bank_windows(const fb_bank* b, uint32_t k) {
  uint64_t w = 0;
  for (size_t len = k; len < b->len_hist.size(); ++len)
    if (b->len_hist[len]) w += (b->len_hist[len] / len) * (len - k + 1);
  return w;
}

extern "C" uint64_t fb_bank_bases_for_k(const fb_bank* b, uint32_t k) {
  if (!b) return 0;
  uint64_t s = 0;
  for (size_t len = k; len < b->len_hist.size(); ++len) s += b->len_hist[len];
  return s;
}

// ---------------------------------------------------------------------------
// CountMinSketch

struct fb_cms {
  fb_ctx* ctx = nullptr;
  DBuf counts;
  uint64_t width = 0;
  uint32_t log2w = 0;
  int bits = 4;
  DevCms dev() const {
    DevCms d;
    d.counts = counts.as<uint32_t>();
    d.log2w = log2w;
    return d;
  }
  uint32_t cap() const { return bits == 4 ? 15u : 65535u; }
};

static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

static int cms_init(fb_ctx* c, fb_cms* m, uint64_t width, int bits) {
  if (bits != 4 && bits != 16) return set_err(c, FB_ERR_INVALID, "counter_bits must be 4 or 16 (got %d)", bits);
  if (width == 0 || width > (1ull << 32)) return set_err(c, FB_ERR_INVALID, "CMS width %llu out of range", (unsigned long long)width);
  m->ctx = c;
  m->width = next_pow2(width);
  m->log2w = (uint32_t)__builtin_ctzll(m->width);
  m->bits = bits;
  if (int r = m->counts.get(c, m->width * 4 * 4)) return r;
  HIP_TRY(c, launch_fill_u32(m->counts.as<uint32_t>(), m->width * 4, 0, c->stream));
  return FB_OK;
}

extern "C" int fb_cms_create(fb_ctx* c, uint64_t width, int bits, fb_cms** out) {
  if (!c || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (int r = ctx_current(c)) return r;
  fb_cms* m = new fb_cms();
  if (int r = cms_init(c, m, width, bits)) {
    delete m;
    return r;
  }
  *out = m;
  return FB_OK;
}
extern "C" void fb_cms_destroy(fb_cms* m) { delete m; }
extern "C" uint64_t fb_cms_width(const fb_cms* m) { return m ? m->width : 0; }
extern "C" int fb_cms_clear(fb_cms* m) {
  if (!m) return FB_ERR_INVALID;
  if (int r = ctx_current(m->ctx)) return r;
  HIP_TRY(m->ctx, launch_fill_u32(m->counts.as<uint32_t>(), m->width * 4, 0, m->ctx->stream));
  return FB_OK;
}

// Copies a host array to a pooled device buffer.
static int to_device(fb_ctx* c, DBuf& d, const void* src, size_t bytes) {
  if (int r = d.get(c, bytes ? bytes : 8)) return r;
  if (bytes) HIP_TRY(c, hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  return FB_OK;
}

// Total adds must stay below 2^32 so the exact u32 cells cannot wrap.
static int check_add_budget(fb_ctx* c, uint64_t adds) {
  if (adds >= 0xFFFFFFFFull)
    return set_err(c, FB_ERR_INVALID, "%llu sketch additions exceed the 2^32-1 exact-counter limit of one pass",
                   (unsigned long long)adds);
  return FB_OK;
}

extern "C" int fb_cms_add_keys(fb_cms* m, const uint64_t* keys, uint64_t n) {
  if (!m || (n && !keys)) return FB_ERR_INVALID;
  fb_ctx* c = m->ctx;
  if (int r = ctx_current(c)) return r;
  if (int r = check_add_budget(c, n)) return r;
  DBuf d;
  if (int r = to_device(c, d, keys, n * 8)) return r;
  HIP_TRY(c, launch_cms_add_keys(m->dev(), d.as<uint64_t>(), n, 1, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

extern "C" int fb_cms_query(fb_cms* m, const uint64_t* keys, uint64_t n, uint32_t* out) {
  if (!m || (n && (!keys || !out))) return FB_ERR_INVALID;
  fb_ctx* c = m->ctx;
  if (int r = ctx_current(c)) return r;
  DBuf d, o;
  if (int r = to_device(c, d, keys, n * 8)) return r;
  if (int r = o.get(c, (n ? n : 1) * 4)) return r;
  HIP_TRY(c, launch_cms_query_keys(m->dev(), d.as<uint64_t>(), n, m->cap(), o.as<uint32_t>(), c->stream));
  if (n) HIP_TRY(c, hipMemcpyAsync(out, o.p, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

extern "C" int fb_cms_add_and_estimate(fb_cms* m, const uint64_t* keys, uint64_t n, uint32_t* out) {
  if (!m || (n && (!keys || !out))) return FB_ERR_INVALID;
  fb_ctx* c = m->ctx;
  if (int r = ctx_current(c)) return r;
  if (!n) return FB_OK;
  if (n > 0xFFFFFFFFull) return set_err(c, FB_ERR_INVALID, "at most 2^32 - 1 keys per call");
  if (int r = check_add_budget(c, n)) return r;
  hipStream_t s = c->stream;
  DBuf d, o, cells, idx, scells, sidx, tmp;
  if (int r = to_device(c, d, keys, n * 8)) return r;
  if (int r = o.get(c, n * 4)) return r;
  if (int r = cells.get(c, n * 4)) return r;
  if (int r = idx.get(c, n * 4)) return r;
  if (int r = scells.get(c, n * 4)) return r;
  if (int r = sidx.get(c, n * 4)) return r;
  const uint32_t end_bit = m->log2w + 2;  // cells < 4 * width
  size_t tb = 0;
  HIP_TRY(c, sort_u32_pairs(nullptr, nullptr, nullptr, nullptr, n, end_bit, nullptr, &tb, s));
  if (int r = tmp.get(c, tb)) return r;
  // every row's estimates read the counts before the batch, then the batch is added
  for (uint32_t row = 0; row < 4; ++row) {
    HIP_TRY(c, launch_cms_row_cells(m->dev(), d.as<uint64_t>(), n, row, cells.as<uint32_t>(), idx.as<uint32_t>(), s));
    size_t t2 = tb;
    HIP_TRY(c, sort_u32_pairs(cells.as<uint32_t>(), scells.as<uint32_t>(), idx.as<uint32_t>(), sidx.as<uint32_t>(), n,
                              end_bit, tmp.p, &t2, s));
    HIP_TRY(c, launch_cms_row_estimate(m->counts.as<uint32_t>(), scells.as<uint32_t>(), sidx.as<uint32_t>(), n, row,
                                       m->cap(), o.as<uint32_t>(), s));
  }
  HIP_TRY(c, launch_cms_add_keys(m->dev(), d.as<uint64_t>(), n, 1, s));
  HIP_TRY(c, hipMemcpyAsync(out, o.p, n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return FB_OK;
}

// fb_cms_add_and_estimate with every buffer of the call in the device's
// static scratch: keys, estimates, per-row cells and indices, their sorted
// copies and the sort's temporary space.  No allocation happens.
extern "C" int fb_cms_add_and_estimate_static(fb_cms* m, const uint64_t* keys, uint64_t n, uint32_t* out) {
  if (!m || (n && (!keys || !out))) return FB_ERR_INVALID;
  fb_ctx* c = m->ctx;
  if (n > kStaticMaxKeys)
    return set_err(c, FB_ERR_INVALID, "at most %llu keys per call (got %llu)", (unsigned long long)kStaticMaxKeys,
                   (unsigned long long)n);
  if (!c->scratch) return set_err(c, FB_ERR_INVALID, "no static scratch on this device");
  if (int r = ctx_current(c)) return r;
  if (!n) return FB_OK;
  if (int r = check_add_budget(c, n)) return r;
  hipStream_t s = c->stream;
  // layout: keys | estimates | cells | idx | sorted cells | sorted idx | sort space
  unsigned char* p = c->scratch;
  uint64_t* d_keys = reinterpret_cast<uint64_t*>(p);
  uint32_t* d_out = reinterpret_cast<uint32_t*>(p + kStaticMaxKeys * 8);
  uint32_t* cells = d_out + kStaticMaxKeys;
  uint32_t* idx = cells + kStaticMaxKeys;
  uint32_t* scells = idx + kStaticMaxKeys;
  uint32_t* sidx = scells + kStaticMaxKeys;
  unsigned char* tmp = reinterpret_cast<unsigned char*>(sidx + kStaticMaxKeys);
  const size_t tmp_room = kStaticScratchBytes - (size_t)(tmp - p);
  const uint32_t end_bit = m->log2w + 2;
  size_t tb = 0;
  HIP_TRY(c, sort_u32_pairs(nullptr, nullptr, nullptr, nullptr, n, end_bit, nullptr, &tb, s));
  if (tb > tmp_room)
    return set_err(c, FB_ERR_INVALID, "sort space %zu bytes exceeds the static scratch (%zu)", tb, tmp_room);
  HIP_TRY(c, hipMemcpyAsync(d_keys, keys, n * 8, hipMemcpyHostToDevice, s));
  for (uint32_t row = 0; row < 4; ++row) {
    HIP_TRY(c, launch_cms_row_cells(m->dev(), d_keys, n, row, cells, idx, s));
    size_t t2 = tb;
    HIP_TRY(c, sort_u32_pairs(cells, scells, idx, sidx, n, end_bit, tmp, &t2, s));
    HIP_TRY(c, launch_cms_row_estimate(m->counts.as<uint32_t>(), scells, sidx, n, row, m->cap(), d_out, s));
  }
  HIP_TRY(c, launch_cms_add_keys(m->dev(), d_keys, n, 1, s));
  HIP_TRY(c, hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return FB_OK;
}

extern "C" int fb_nthash_windows(fb_ctx* c, const uint8_t* seq, uint64_t len, uint32_t k, uint64_t* out) {
  if (!c || k == 0 || (len && !seq)) return set_err(c, FB_ERR_INVALID, "bad argument");
  if (len < k) return FB_OK;  // no window
  if (!out) return set_err(c, FB_ERR_INVALID, "null output");
  if (int r = ctx_current(c)) return r;
  const uint64_t nw = len - k + 1;
  DBuf d, o;
  if (int r = to_device(c, d, seq, len)) return r;
  if (int r = o.get(c, nw * 8)) return r;
  HIP_TRY(c, launch_nthash_windows(d.as<uint8_t>(), nw, k, o.as<uint64_t>(), c->stream));
  HIP_TRY(c, hipMemcpyAsync(out, o.p, nw * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

static uint32_t default_split(uint32_t k, uint32_t max_len) {
  // about 48 windows per thread
  uint32_t win = max_len >= k ? max_len - k + 1 : 1;
  uint32_t s = (win + 47) / 48;
  return std::max(1u, std::min(s, 8u));
}

extern "C" int fb_cms_add_bank(fb_cms* m, const fb_bank* b, uint32_t k) {
  if (!m || !b) return FB_ERR_INVALID;
  fb_ctx* c = m->ctx;
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  if (int r = ctx_current(c)) return r;
  if (int r = check_add_budget(c, fb_bank_bases_for_k(b, k))) return r;
  Pass1Args a{};
  a.reads = b->view();
  a.k = k;
  a.split = default_split(k, b->max_len);
  a.weight = 1;
  a.cms = m->dev();
  HIP_TRY(c, launch_pass1(a, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

extern "C" uint64_t fb_cms_table_bytes(const fb_cms* m) {
  if (!m) return 0;
  return m->bits == 4 ? m->width * 2 : m->width * 8;
}

static int cms_download(fb_ctx* c, const DevCms& d, uint64_t width, int bits, void* out, uint64_t bytes) {
  const uint64_t need = bits == 4 ? width * 2 : width * 8;
  if (!out || bytes < need) return set_err(c, FB_ERR_INVALID, "CMS download needs %llu bytes", (unsigned long long)need);
  DBuf tmp;
  if (int r = tmp.get(c, need)) return r;
  if (bits == 4) HIP_TRY(c, launch_pack_nibble(d.counts, width * 4, tmp.as<uint8_t>(), c->stream));
  else HIP_TRY(c, launch_pack_u16(d.counts, width * 4, tmp.as<uint16_t>(), c->stream));
  HIP_TRY(c, hipMemcpyAsync(out, tmp.p, need, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}

extern "C" int fb_cms_download(fb_cms* m, void* out, uint64_t bytes) {
  if (!m) return FB_ERR_INVALID;
  if (int r = ctx_current(m->ctx)) return r;
  return cms_download(m->ctx, m->dev(), m->width, m->bits, out, bytes);
}

// ---------------------------------------------------------------------------
// HyperLogLog

struct fb_hll {
  fb_ctx* ctx = nullptr;
  DBuf regs;
  int p = 14;
  DevHll dev() const {
    DevHll d;
    d.regs = regs.as<uint32_t>();
    d.p = (uint32_t)p;
    return d;
  }
};

static int hll_init(fb_ctx* c, fb_hll* h, int p) {
  if (p < 4 || p > 18) return set_err(c, FB_ERR_INVALID, "HLL precision %d outside 4..18", p);
  h->ctx = c;
  h->p = p;
  if (int r = h->regs.get(c, ((size_t)1 << p) * 4)) return r;
  HIP_TRY(c, launch_fill_u32(h->regs.as<uint32_t>(), (uint64_t)1 << p, 0, c->stream));
  return FB_OK;
}

extern "C" int fb_hll_create(fb_ctx* c, int p, fb_hll** out) {
  if (!c || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (int r = ctx_current(c)) return r;
  fb_hll* h = new fb_hll();
  if (int r = hll_init(c, h, p)) {
    delete h;
    return r;
  }
  *out = h;
  return FB_OK;
}
extern "C" void fb_hll_destroy(fb_hll* h) { delete h; }
extern "C" int fb_hll_clear(fb_hll* h) {
  if (!h) return FB_ERR_INVALID;
  if (int r = ctx_current(h->ctx)) return r;
  HIP_TRY(h->ctx, launch_fill_u32(h->regs.as<uint32_t>(), (uint64_t)1 << h->p, 0, h->ctx->stream));
  return FB_OK;
}
extern "C" int fb_hll_add_hashes(fb_hll* h, const uint64_t* hashes, uint64_t n) {
  if (!h || (n && !hashes)) return FB_ERR_INVALID;
  fb_ctx* c = h->ctx;
  if (int r = ctx_current(c)) return r;
  DBuf d;
  if (int r = to_device(c, d, hashes, n * 8)) return r;
  HIP_TRY(c, launch_hll_add_hashes(h->dev(), d.as<uint64_t>(), n, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" int fb_hll_add_canonical_hashes(fb_hll* h, const uint64_t* hashes, uint64_t n) {
  if (!h || (n && !hashes)) return FB_ERR_INVALID;
  fb_ctx* c = h->ctx;
  if (int r = ctx_current(c)) return r;
  DBuf d;
  if (int r = to_device(c, d, hashes, n * 8)) return r;
  HIP_TRY(c, launch_hll_add_canonical(h->dev(), d.as<uint64_t>(), n, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" int fb_hll_merge(fb_hll* dst, const fb_hll* src) {
  if (!dst || !src) return FB_ERR_INVALID;
  fb_ctx* c = dst->ctx;
  if (dst->p != src->p)
    return set_err(c, FB_ERR_INVALID, "HLL precisions differ (%d and %d)", dst->p, src->p);
  if (int r = ctx_current(c)) return r;
  if (src->ctx != c) HIP_TRY(c, hipStreamSynchronize(src->ctx->stream));
  HIP_TRY(c, launch_hll_merge(dst->regs.as<uint32_t>(), src->regs.as<uint32_t>(), (uint64_t)1 << dst->p, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" int fb_hll_add_bank(fb_hll* h, const fb_bank* b, uint32_t k) {
  if (!h || !b) return FB_ERR_INVALID;
  fb_ctx* c = h->ctx;
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  if (int r = ctx_current(c)) return r;
  Pass1Args a{};
  a.reads = b->view();
  a.k = k;
  a.split = default_split(k, b->max_len);
  a.weight = 1;
  a.hll = h->dev();
  HIP_TRY(c, launch_pass1(a, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
static int hll_download(fb_ctx* c, const DevHll& d, uint8_t* out) {
  if (!out) return FB_ERR_INVALID;
  const uint64_t m = (uint64_t)1 << d.p;
  DBuf tmp;
  if (int r = tmp.get(c, m)) return r;
  HIP_TRY(c, launch_pack_u8(d.regs, m, tmp.as<uint8_t>(), c->stream));
  HIP_TRY(c, hipMemcpyAsync(out, tmp.p, m, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" int fb_hll_registers(fb_hll* h, uint8_t* out) {
  if (!h) return FB_ERR_INVALID;
  if (int r = ctx_current(h->ctx)) return r;
  return hll_download(h->ctx, h->dev(), out);
}
extern "C" int fb_hll_estimate(fb_hll* h, double* out) {
  if (!h || !out) return FB_ERR_INVALID;
  std::vector<uint8_t> regs((size_t)1 << h->p);
  if (int r = fb_hll_registers(h, regs.data())) return r;
  *out = fb_hll_estimate_registers(regs.data(), h->p);
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Solid set

struct fb_solid {
  fb_ctx* ctx = nullptr;
  uint32_t k = 0;
  uint32_t nref = 1;
  uint64_t n = 0;
  DBuf keys, cov, mask;
  fb_cms cms;
  fb_hll hll;
  bool has_hll = false;
  float ms[4] = {0, 0, 0, 0};
  // pass-1 scatter, pass-1 count, pass-2 gate+count, pass-2 scatter, pass-2 bin tables
  float kms[5] = {0, 0, 0, 0, 0};
  // windows, read records, claimed keys, table slots, pass-2 route, LDS bins,
  // keys per LDS bin table, estimated keys, LDS-route records, reads >= k,
  // pass-2 rounds
  uint64_t stats[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool sorted = false;  // records sorted by k-mer (the parity form)
  int p2_gate = 0;      // as fb_solid_pass2_gate
  int p2_layout = 0;    // as fb_solid_pass2_layout
};

extern "C" void fb_build_params_default(fb_build_params* p, uint32_t k) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->k = k;
  p->target_solid_cov = -1;
  p->counter_bits = 4;
  p->sort = 1;
}

static uint32_t ref_words(uint32_t k) { return k <= 31 ? 1u : (k <= 63 ? 2u : 4u); }
// Solid-table slot layout: kw key words of 62 bits, then the counters.
static uint32_t key_words_for_k(uint32_t k) { return (k + 30) / 31; }

// Entries per (bin, XCD slice) region when per_bin entries land in a bin on
// average from n_wg workgroups: the even share of one slice plus 10%, and room
// for two more workgroups' shares (few workgroups split unevenly over the
// XCDs).  A full region spills (pass 1) or falls back (pass 2), so the size
// only matters for speed.
// Pass-1 regions are filled in line-sized chunks, each workgroup leaving at
// most one partly used chunk per bin: add one chunk per workgroup of the
// slice, and keep regions a whole number of chunks (lines).
static uint64_t bin_slice_cap(double per_bin, uint64_t n_wg) {
  const double wgs = (double)std::max<uint64_t>(n_wg, 1);
  const uint64_t c = (uint64_t)(per_bin * (1.1 / kBinSlices + 2.0 / wgs)) + 64 +
                     (n_wg / kBinSlices + 2) * kP1Chunk;
  return (c + kP1Chunk - 1) / kP1Chunk * kP1Chunk;
}
// Region positions are kept in 26 bits by the scatter kernel.
constexpr uint64_t kP1MaxCap = (1ull << 26) - 128;
static uint32_t table_stride(uint32_t kw) { return kw + 5; }

extern "C" void fb_solid_out_bytes(uint64_t cap, uint32_t nref, uint64_t out[3]) {
  if (!out) return;
  out[0] = cap * nref * 8;
  out[1] = cap * 2;
  out[2] = cap;
}
extern "C" int fb_solid_out_fits(uint64_t cap, uint32_t nref, const uint64_t have[3]) {
  if (!have) return 0;
  uint64_t need[3];
  fb_solid_out_bytes(cap, nref, need);
  return have[0] >= need[0] && have[1] >= need[1] && have[2] >= need[2];
}
// Solid outputs of an isolate that keeps them across the k ladder: grown by
// bytes at this k (an entry is nref words wide), never by entry count.
static int solid_out_grow(fb_ctx* c, DBuf& keys, DBuf& cov, DBuf& mask, uint64_t cap, uint32_t nref) {
  uint64_t need[3];
  fb_solid_out_bytes(std::max<uint64_t>(cap, 1), nref, need);
  DBuf* b[3] = {&keys, &cov, &mask};
  for (int i = 0; i < 3; ++i)
    if (!b[i]->p || b[i]->bytes < need[i])
      if (int r = b[i]->get(c, need[i])) return r;
  return FB_OK;
}
// The guard before every launch that writes cap solid entries of nref words.
static int solid_out_check(fb_ctx* c, const DBuf& keys, const DBuf& cov, const DBuf& mask, uint64_t cap,
                           uint32_t nref) {
  const uint64_t have[3] = {keys.p ? keys.bytes : 0, cov.p ? cov.bytes : 0, mask.p ? mask.bytes : 0};
  if (fb_solid_out_fits(cap, nref, have)) return FB_OK;
  return set_err(c, FB_ERR_INVALID,
                 "internal: solid output buffers (%llu/%llu/%llu bytes) cannot hold %llu entries of %u key words",
                 (unsigned long long)have[0], (unsigned long long)have[1], (unsigned long long)have[2],
                 (unsigned long long)cap, nref);
}

// Keys the ungated pass-2 tables hold: keys with at least f windows (the
// kept keys, plus the record filter's few false positives).  Row-0 cells
// whose count reaches T hold them, where T = max(f, a quarter of the mean
// count of the cells that pass the gate) leaves out the cells that collect
// only error k-mers (on the 5 Mbp isolate at k 21 the mean is about 36);
// linear counting over those cells (at least a quarter of the count at f),
// plus a tenth.  hist[c]: row-0 cells with
// min(count, 255) = c.  The estimate only sizes the tables: keys beyond it
// go to the global table or, past the output buffer, to the global route.
double fb::est_kept_keys(const uint32_t* hist, uint64_t width, uint64_t windows, uint32_t t, uint32_t f) {
  uint64_t occ_t = 0;
  double sum_t = 0.0;
  for (uint32_t c = std::max(std::min(t, 255u), 1u); c < 256; ++c) {
    occ_t += hist[c];
    sum_t += (double)c * hist[c];
  }
  const double mean = occ_t ? sum_t / (double)occ_t : 0.0;
  const double W = (double)width;
  auto linear = [&](uint32_t T) {  // distinct keys in cells with count >= T
    uint64_t occ = 0;
    for (uint32_t c = T; c < 256; ++c) occ += hist[c];
    return occ >= width ? (double)windows : -W * log(1.0 - (double)occ / W);
  };
  // with f = 1 (edge gate 1) every key with a record is kept: all of them count
  const uint32_t fl = std::max(f, 1u);
  const uint32_t T = fl == 1u ? 1u : std::min<uint32_t>(255u, std::max<uint32_t>(fl, (uint32_t)(mean / 4.0)));
  const double n = std::max(linear(T), 0.25 * linear(fl));
  return std::min(1.1 * n, (double)windows + 1.0);
}
// Distinct keys of the pass (linear counting over the occupied row-0 cells).
static double est_distinct_keys(const uint32_t* hist, uint64_t width, uint64_t windows) {
  uint64_t occ = 0;
  for (uint32_t c = 1; c < 256; ++c) occ += hist[c];
  const double W = (double)width;
  const double n = occ >= width ? (double)windows : -W * log(1.0 - (double)occ / W);
  return std::min(1.1 * n, (double)windows + 1.0);
}

extern "C" int fb_build_solid(fb_ctx* c, const fb_bank* bank, const uint8_t* contigs, const uint32_t* contig_lens,
                              uint64_t n_contigs, const fb_build_params* prm, fb_solid** out) {
  if (!c || !bank || !prm || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  const uint32_t k = prm->k;
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  if (prm->counter_bits != 4 && prm->counter_bits != 16)
    return set_err(c, FB_ERR_INVALID, "counter_bits must be 4 or 16 (got %d)", prm->counter_bits);
  if (prm->hll_p != 0 && (prm->hll_p < 4 || prm->hll_p > 18))
    return set_err(c, FB_ERR_INVALID, "hll_p %d outside 4..18", prm->hll_p);
  if (prm->target_solid_cov > 255) return set_err(c, FB_ERR_INVALID, "target_solid_cov is a u8");
  if (n_contigs && (!contigs || !contig_lens)) return set_err(c, FB_ERR_INVALID, "contigs: null pointer");
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;

  // build.rs:74-84: genome size source and CMS width
  uint64_t est_g = prm->genome_size ? std::max<uint64_t>(prm->genome_size, 50000)
                                    : std::max<uint64_t>(bank->total_bases / 60, 50000);
  uint64_t width = prm->cms_width ? prm->cms_width : fb_cms_width_for_genome_size(est_g);
  const uint32_t t_solid = prm->target_solid_cov >= 0 ? (uint32_t)prm->target_solid_cov : 2u;
  const uint32_t gate = prm->target_solid_cov >= 0 ? (uint32_t)prm->target_solid_cov : 3u;

  uint64_t contig_bases = 0;
  uint32_t contig_max = 0;
  for (uint64_t i = 0; i < n_contigs; ++i) {
    contig_bases += contig_lens[i];
    contig_max = std::max(contig_max, contig_lens[i]);
  }
  // contigs run to many kb: about 64 windows per thread share
  const uint32_t contig_split = std::max<uint32_t>(1u, std::min<uint32_t>(4096u, (contig_max + 63) / 64));
  if (int r = check_add_budget(c, fb_bank_bases_for_k(bank, k) + 10 * contig_bases)) return r;

  fb_solid* so = new fb_solid();
  so->ctx = c;
  so->k = k;
  so->nref = ref_words(k);
  auto fail = [&](int r) {
    delete so;
    return r;
  };
  if (int r = cms_init(c, &so->cms, width, prm->counter_bits)) return fail(r);
  if (prm->hll_p) {
    if (int r = hll_init(c, &so->hll, prm->hll_p)) return fail(r);
    so->has_hll = true;
  }
  // prior contigs as a packed read list (weights applied by the kernels)
  DBuf cdata, coffs, clens;
  uint64_t c_reads = 0;
  if (n_contigs) {
    std::vector<uint64_t> words((contig_bases + 31) / 32 + 1, 0), offs(n_contigs);
    fb_pack_ascii(contigs, contig_lens, n_contigs, words.data(), offs.data());
    fb_arena_view v{words.data(), (uint64_t)words.size(), offs.data(), contig_lens, n_contigs};
    if (int r = upload_reads(c, &v, 1, cdata, coffs, clens, &c_reads, nullptr, nullptr, nullptr)) return fail(r);
  }
  DevReads cview;
  cview.data = cdata.as<uint64_t>();
  cview.offs = coffs.as<uint64_t>();
  cview.lens = clens.as<uint32_t>();
  cview.n_reads = c_reads;
  const uint32_t split = prm->split ? prm->split : default_split(k, bank->max_len);

  // pass 1: LDS-binned Count-Min slices or direct global atomics
  const uint64_t ncnt = so->cms.width * 4;
  const uint64_t windows = bank_windows(bank, k);
  uint32_t sb = kBinLog2Cells;
  while (sb > 6 && (1ull << sb) > ncnt) --sb;
  const uint64_t nbins = std::max<uint64_t>(1, ncnt >> sb);
  // placement 3: the LDS route with the per-window pass-2 gate (the ungated
  // form is the default wherever it applies)
  const bool lds_placement = prm->placement == 2 || prm->placement == 3;
  const bool binned = lds_placement || (prm->placement == 0 && nbins <= 8192 && windows >= (1u << 20));
  if (lds_placement && nbins > 8192)
    return fail(set_err(c, FB_ERR_INVALID, "LDS placement needs at most 8192 bins (width %llu)",
                        (unsigned long long)so->cms.width));
  DBuf bits, bdata, bcur;
  if (int r = bits.get(c, ((ncnt + 31) / 32) * 4)) return fail(r);
  BinArgs ba;
  if (binned) {
    const uint64_t thr1 = bank->n_reads * (uint64_t)split;
    // without an explicit chunk count the scatter grid is the resident blocks, striding over the shares
    ba.grid = prm->chunks1 ? 0u : bin_scatter_resident(k, so->has_hll, (uint32_t)nbins);
    const uint64_t n_wg = ba.grid ? ba.grid : scatter_blocks(thr1, prm->chunks1, kP1BlockSize);
    ba.cap = std::min<uint64_t>(bin_slice_cap(4.0 * (double)windows / (double)nbins, n_wg), kP1MaxCap);
    ba.sb = sb;
    ba.nbins = (uint32_t)nbins;
    if (int r = bdata.get(c, nbins * kBinSlices * ba.cap * 2)) return fail(r);
    if (int r = bcur.get(c, nbins * kBinSlices * 4)) return fail(r);
    HIP_TRY(c, launch_fill_u32(bcur.as<uint32_t>(), nbins * kBinSlices, 0, s));
  }
  HIP_TRY(c, hipEventRecord(c->ev[0], s));
  if (c_reads) {  // prior contigs: every k-mer added 10 times (build.rs:119-138)
    Pass1Args pc{};
    pc.reads = cview;
    pc.k = k;
    pc.split = contig_split;
    pc.weight = 10;
    pc.cms = so->cms.dev();
    HIP_TRY(c, launch_pass1(pc, s));
  }
  if (binned) {
    ba.reads = bank->view();
    ba.k = k;
    ba.split = split;
    ba.chunks = prm->chunks1 ? prm->chunks1 : 1u;
    ba.log2w = so->cms.log2w;
    ba.n_cells = ncnt;
    ba.counts = so->cms.counts.as<uint32_t>();
    if (so->has_hll) ba.hll = so->hll.dev();
    ba.data = bdata.as<uint16_t>();
    ba.cursor = bcur.as<uint32_t>();
    ba.bits = bits.as<uint32_t>();
    ba.t_solid = t_solid;
    ba.cap_value = so->cms.cap();
    HIP_TRY(c, launch_bin_pass1(ba, s, c->ev[8]));
  } else {
    Pass1Args p1{};
    p1.reads = bank->view();
    p1.k = k;
    p1.split = split;
    p1.weight = 1;
    p1.cms = so->cms.dev();
    if (so->has_hll) p1.hll = so->hll.dev();
    HIP_TRY(c, launch_pass1(p1, s));
  }
  HIP_TRY(c, hipEventRecord(c->ev[1], s));
  if (!binned)  // threshold bitmap of the pass-2 gate
    HIP_TRY(c, launch_solid_bits(so->cms.counts.as<uint32_t>(), ncnt, t_solid, so->cms.cap(), bits.as<uint32_t>(), s));
  bdata.release();
  bcur.release();

  // pass 2 into the solid table; grows and re-runs on overflow
  DevSolidTable tab;
  tab.kw = key_words_for_k(k);
  tab.stride = table_stride(tab.kw);
  tab.max_probe = 4096;
  // Capacity from the sketch itself: occ = row-0 cells whose count passes the
  // gate; linear counting N = -W ln(1 - occ/W) estimates the distinct k-mers
  // landing in such cells, an upper-side estimate of the keys pass 2 inserts.
  // The ungated pass-2 form (records of every window pair, the gate applied
  // per kept key) needs the counters' cap >= t_solid, so that a k-mer with
  // t_solid windows passes the gate, and with prior contigs t_solid <= 10, so
  // that every contig k-mer (ten adds per window) does.
  const bool ungated = prm->placement != 3 && t_solid <= so->cms.cap() && (c_reads == 0 || t_solid <= 10);
  const uint32_t filt = ungated && gate >= 2 ? std::min(gate, kP2MaxFilt) : 0u;
  uint64_t cap;
  double n_est_keys = 0.0, n_distinct = 0.0;
  {
    DBuf pc;
    if (int r = pc.get(c, 8 + 256 * 4)) return fail(r);
    HIP_TRY(c, launch_fill_u32(pc.as<uint32_t>(), 2 + 256, 0, s));
    HIP_TRY(c, launch_popcount(bits.as<uint32_t>(), so->cms.width / 32 ? so->cms.width / 32 : 1,
                               pc.as<unsigned long long>(), s));
    if (ungated) HIP_TRY(c, launch_count_hist(so->cms.counts.as<uint32_t>(), so->cms.width, pc.as<uint32_t>() + 2, s));
    uint32_t hbuf[2 + 256];
    HIP_TRY(c, hipMemcpyAsync(hbuf, pc.p, sizeof(hbuf), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    unsigned long long occ = 0;
    memcpy(&occ, hbuf, 8);
    const double W = (double)so->cms.width;
    double n_est = occ >= so->cms.width ? (double)windows : -W * log(1.0 - (double)occ / W);
    n_est = std::min(n_est, (double)windows + 1.0);
    cap = next_pow2(std::max<uint64_t>(1ull << 16, (uint64_t)(2.0 * n_est) + 1024));
    n_est_keys = ungated ? est_kept_keys(hbuf + 2, so->cms.width, windows, t_solid, std::max(filt, 1u)) : n_est;
    n_distinct = ungated ? est_distinct_keys(hbuf + 2, so->cms.width, windows) : n_est;
  }
  DBuf tbuf, status;
  if (int r = status.get(c, 16)) return fail(r);
  tab.status = status.as<uint32_t>();
  uint32_t n_out = 0;
  uint64_t rec_sum = 0, claimed = 0;
  uint64_t route = 0;
  bool done = false;
  // LDS-privatised pass 2+3 (reads, then the prior contigs' windows)
  const uint32_t kw = tab.kw;
  const uint32_t lds_cap = p2_lds_cap(kw, filt);
  // bins over all rounds: the estimated keys at half the LDS table capacity
  // (0.65 for the ungated form, whose estimate carries its own margin)
  const double load = ungated ? kP2UngatedLoad : 0.5;
  uint64_t nb2 = 1;
  while ((double)nb2 * load * (double)lds_cap < n_est_keys) nb2 <<= 1;
  const uint64_t max_bins = (uint64_t)p2_max_bins() << kP2MaxLog2Rounds;
  if (prm->bins2) {
    if (prm->bins2 > max_bins || (prm->bins2 & (prm->bins2 - 1)))
      return fail(set_err(c, FB_ERR_INVALID, "bins2 must be a power of two <= %llu (got %u)",
                          (unsigned long long)max_bins, prm->bins2));
    nb2 = prm->bins2;
  }
  // rounds of at most p2_max_bins() bins each
  uint32_t log2_rounds = 0;
  while ((nb2 >> log2_rounds) > p2_max_bins()) ++log2_rounds;
  const uint64_t nbr = nb2 >> log2_rounds;
  so->stats[5] = nb2;
  so->stats[6] = lds_cap;
  so->stats[7] = (uint64_t)n_est_keys;
  const uint32_t pairs_max = bank->max_len > k ? bank->max_len - k : 1;
  const uint32_t split2 = std::max<uint32_t>(split, (pairs_max + 62) / 63);
  // prior contigs follow the reads in the share space, at most 63 pairs a share
  const uint32_t cpairs_max = contig_max > k ? contig_max - k : 1;
  const uint32_t csplit2 = (cpairs_max + 62) / 63;
  uint64_t contig_windows = 0;
  for (uint64_t i = 0; i < n_contigs; ++i)
    if (contig_lens[i] >= k) contig_windows += contig_lens[i] - k + 1;
  const uint64_t threads2 = bank->n_reads * (uint64_t)split2 + c_reads * (uint64_t)csplit2;
  // a window at a share boundary can give one record to each of two threads
  const uint64_t rec_cap = windows + contig_windows + threads2;
  if (prm->placement != 1 && nb2 <= max_bins && windows > 0 && rec_cap < (1ull << 32) &&
      c_reads * (uint64_t)csplit2 <= (1ull << 26)) {
    const uint32_t cflag = c_reads ? 1u : 0u;
    const uint32_t rw = (2 * k + 8 + cflag + 63) / 64;
    // keys beyond what the LDS tables are expected to hold, plus slack
    const double excess = std::max(0.0, n_est_keys - (double)nb2 * load * (double)lds_cap);
    // fewer ungated bins than the estimate asks for (an explicit bins2) also
    // overload their record filters, which then let most keys through: the
    // spill table is sized for all of them
    const double spilled = ungated && excess > 0.0 ? n_distinct : 2.0 * excess + n_est_keys / 4;
    const uint64_t spill_cap = next_pow2(std::max<uint64_t>(1ull << 16, (uint64_t)spilled));
    // solid keys from the bins: at most the keys they hold; sized from the
    // estimate, with a count past the end sending the build to the global route
    const uint64_t lds_out = std::min<uint64_t>(nb2 * lds_cap, (uint64_t)(1.25 * n_est_keys) + 65536);
    const uint64_t out_cap = std::min<uint64_t>(lds_out + spill_cap, 0xFFFFFFFFull);
    const uint32_t chunks2 = prm->chunks2 ? prm->chunks2 : scatter_chunks(threads2);
    const uint64_t n_blocks2 = scatter_blocks(threads2, chunks2);
    // Without the count walk (ungated, one round, the buffered scatter's
    // bins): per (bin, slice) regions of the expected records plus a fifth,
    // and two chunks for each workgroup of the slice's XCD; a region that
    // fills up sends the pass to a second run with exact runs.
    uint64_t region_cap = 0;
    if (ungated && log2_rounds == 0 && p2_scatter_buffered(k, (uint32_t)nbr) && prm->placement != 3) {
      const uint64_t CR = (uint64_t)kP2ChunkUnits * (kP2BufWords / rw);
      const double per = (double)rec_cap / (double)(nbr * kBinSlices);
      uint64_t rc = (uint64_t)(1.2 * per) + (p2_chunk_grid() / kBinSlices + 1) * 2 * CR;
      rc = (rc + CR - 1) / CR * CR;
      if (nbr * kBinSlices * rc < (1ull << 32)) region_cap = rc;
    }
    DBuf recs, gm, runs, bh, st2, p3s;
    int r = 0;
    if (!r) r = bh.get(c, std::max<uint64_t>(n_blocks2 * nbr, 1) * 4);
    // the bin kernel reads a word past the last record
    if (!r) r = recs.get(c, std::max<uint64_t>(rec_cap, nbr * kBinSlices * region_cap) * rw * 8 + 64);
    if (!r && !ungated) r = gm.get(c, std::max<uint64_t>(threads2, 1) * 8);
    if (!r) r = runs.get(c, (nbr * kBinSlices + 1) * 4);
    if (!r) r = st2.get(c, 64);
    if (!r) r = p3s.get(c, pass3_scratch_bytes());
    if (!r) r = tbuf.get(c, spill_cap * tab.stride * 8);
    if (!r) r = so->keys.get(c, out_cap * so->nref * 8);
    if (!r) r = so->cov.get(c, out_cap * 2);
    if (!r) r = so->mask.get(c, out_cap);
    if (!r) r = solid_out_check(c, so->keys, so->cov, so->mask, out_cap, so->nref);
    if (r) return fail(r);
    tab.cap = spill_cap;
    tab.slots = tbuf.as<uint64_t>();
    P2BinArgs pa;
    pa.reads = bank->view();
    pa.contigs = cview;
    pa.csplit = csplit2;
    pa.cflag = cflag;
    pa.k = k;
    pa.split = split2;
    pa.chunks = chunks2;
    pa.blk_hist = bh.as<uint32_t>();
    pa.n_blocks = (uint32_t)n_blocks2;
    pa.solid_bits = bits.as<uint32_t>();
    pa.log2w = so->cms.log2w;
    pa.nbins = (uint32_t)nbr;
    pa.log2_nbins = (uint32_t)__builtin_ctzll(nbr);
    pa.log2_rounds = log2_rounds;
    pa.gate_masks = gm.as<uint64_t>();
    pa.rw = rw;
    pa.runs = runs.as<uint32_t>();
    pa.recs = recs.as<uint64_t>();
    pa.rec_cap = rec_cap;
    pa.status = st2.as<uint32_t>();            // [0] record buffer overflow, [1] records
    pa.out_count = st2.as<uint32_t>() + 2;     // records emitted from LDS
    pa.out_cap = (uint32_t)out_cap;
    pa.stats = reinterpret_cast<unsigned long long*>(st2.as<uint32_t>() + 4);
    pa.lds_cap = lds_cap;
    pa.min_trans_cov = gate;
    pa.ungated = ungated ? 1u : 0u;
    pa.t_solid = t_solid;
    pa.filt = filt;
    pa.table = tab;
    pa.nref = so->nref;
    pa.out_keys = so->keys.as<uint64_t>();
    pa.out_cov = so->cov.as<uint16_t>();
    pa.out_mask = so->mask.as<uint8_t>();
    pa.region_cap = region_cap;
    uint32_t hb[8] = {0}, sp[4] = {0}, ts[2] = {0};
    for (;;) {
      HIP_TRY(c, launch_table_clear(tab, s));
      HIP_TRY(c, launch_fill_u32(st2.as<uint32_t>(), 16, 0, s));
      HIP_TRY(c, hipEventRecord(c->ev[2], s));
      HIP_TRY(c, launch_p2_binned(pa, s, c->ev[9], c->ev[10]));
      HIP_TRY(c, hipEventRecord(c->ev[3], s));
      HIP_TRY(c, hipEventRecord(c->ev[4], s));
      HIP_TRY(c, launch_pass3(tab, k, so->nref, gate, so->keys.as<uint64_t>(), so->cov.as<uint16_t>(),
                              so->mask.as<uint8_t>(), p3s.p, s, pa.out_count, (uint32_t)out_cap));
      HIP_TRY(c, hipMemcpyAsync(hb, st2.p, 32, hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipMemcpyAsync(sp, p3s.p, 16, hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipMemcpyAsync(ts, tab.status, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipStreamSynchronize(s));
      so->p2_layout = pa.region_cap ? 1 : (region_cap ? 2 : 0);
      if (!(hb[0] && pa.region_cap)) break;
      pa.region_cap = 0;  // a chunked region filled up: again with exact runs
    }
    const bool out_full = (uint64_t)hb[2] + sp[0] > out_cap;
    if (!hb[0] && !ts[0] && !out_full) {
      uint64_t bins_rec = 0, bins_keys = 0, sp_rec = 0;
      memcpy(&bins_rec, hb + 4, 8);
      memcpy(&bins_keys, hb + 6, 8);
      memcpy(&sp_rec, sp + 2, 8);
      n_out = hb[2] + sp[0];
      rec_sum = bins_rec + sp_rec;
      claimed = bins_keys + ts[1];
      cap = spill_cap;
      route = 1;
      so->stats[8] = hb[1];
      so->stats[10] = 1ull << log2_rounds;
      so->p2_gate = ungated ? 1 : 0;
      done = true;
    } else {
      route = hb[0] ? 2 : (ts[0] ? 3 : 4);
    }
  }
  if (!done) {
    uint32_t st[2] = {0, 0};
    for (int attempt = 0;; ++attempt) {
      tab.cap = cap;
      if (int r = tbuf.get(c, cap * tab.stride * 8)) return fail(r);
      tab.slots = tbuf.as<uint64_t>();
      HIP_TRY(c, launch_table_clear(tab, s));
      HIP_TRY(c, hipEventRecord(c->ev[2], s));
      Pass2Args p2{};
      p2.reads = bank->view();
      p2.k = k;
      p2.split = split;
      p2.solid_bits = bits.as<uint32_t>();
      p2.log2w = so->cms.log2w;
      p2.table = tab;
      p2.min_trans_cov = gate;
      p2.contig_mode = 0;
      HIP_TRY(c, launch_pass2(p2, s));
      if (c_reads) {
        Pass2Args pc = p2;
        pc.reads = cview;
        pc.split = contig_split;
        pc.contig_mode = 1;
        HIP_TRY(c, launch_pass2(pc, s));
      }
      HIP_TRY(c, hipEventRecord(c->ev[3], s));
      HIP_TRY(c, hipMemcpyAsync(st, tab.status, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipStreamSynchronize(s));
      if (!st[0]) break;
      if (attempt >= 6)
        return fail(set_err(c, FB_ERR_OVERFLOW, "solid table overflow at %llu slots", (unsigned long long)cap));
      cap *= 2;
    }
    // pass 3
    claimed = st[1];
    const uint64_t maxn = std::max<uint64_t>(claimed, 1);
    if (int r = so->keys.get(c, maxn * so->nref * 8)) return fail(r);
    if (int r = so->cov.get(c, maxn * 2)) return fail(r);
    if (int r = so->mask.get(c, maxn)) return fail(r);
    if (int r = solid_out_check(c, so->keys, so->cov, so->mask, maxn, so->nref)) return fail(r);
    DBuf cnt;
    if (int r = cnt.get(c, pass3_scratch_bytes())) return fail(r);
    HIP_TRY(c, hipEventRecord(c->ev[4], s));
    HIP_TRY(c, launch_pass3(tab, k, so->nref, gate, so->keys.as<uint64_t>(), so->cov.as<uint16_t>(),
                            so->mask.as<uint8_t>(), cnt.p, s));
    uint32_t cbuf[4] = {0, 0, 0, 0};
    HIP_TRY(c, hipMemcpyAsync(cbuf, cnt.p, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    n_out = cbuf[0];
    memcpy(&rec_sum, cbuf + 2, 8);
  }
  bits.release();
  so->stats[0] = windows;
  so->stats[1] = rec_sum;
  so->stats[2] = claimed;
  so->stats[3] = cap;
  so->stats[4] = route;
  for (size_t len = k; len < bank->len_hist.size(); ++len) so->stats[9] += bank->len_hist[len] / len;
  tbuf.release();
  so->n = n_out;
  so->sorted = prm->sort != 0;
  if (prm->sort && n_out > 1) {
    size_t tb = 0;
    HIP_TRY(c, sort_solid(nullptr, nullptr, nullptr, n_out, so->nref, k, nullptr, &tb, s));
    DBuf tmp;
    if (int r = tmp.get(c, tb)) return fail(r);
    HIP_TRY(c, sort_solid(so->keys.as<uint64_t>(), so->cov.as<uint16_t>(), so->mask.as<uint8_t>(), n_out, so->nref, k,
                          tmp.p, &tb, s, true));
  }
  HIP_TRY(c, hipEventRecord(c->ev[5], s));
  HIP_TRY(c, hipEventSynchronize(c->ev[5]));
  (void)hipEventElapsedTime(&so->ms[0], c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&so->ms[1], c->ev[2], c->ev[3]);
  (void)hipEventElapsedTime(&so->ms[2], c->ev[4], c->ev[5]);
  (void)hipEventElapsedTime(&so->ms[3], c->ev[0], c->ev[5]);
  if (binned) {
    (void)hipEventElapsedTime(&so->kms[0], c->ev[0], c->ev[8]);
    (void)hipEventElapsedTime(&so->kms[1], c->ev[8], c->ev[1]);
  }
  if (route == 1) {
    (void)hipEventElapsedTime(&so->kms[2], c->ev[2], c->ev[9]);
    (void)hipEventElapsedTime(&so->kms[3], c->ev[9], c->ev[10]);
    (void)hipEventElapsedTime(&so->kms[4], c->ev[10], c->ev[3]);
  }
  *out = so;
  return FB_OK;
}

extern "C" void fb_solid_destroy(fb_solid* s) { delete s; }
extern "C" uint64_t fb_solid_count(const fb_solid* s) { return s ? s->n : 0; }
extern "C" int fb_solid_key_words(const fb_solid* s) { return s ? (int)s->nref : 0; }
extern "C" int fb_solid_download(const fb_solid* s, uint64_t* keys, uint16_t* cov, uint8_t* mask) {
  if (!s) return FB_ERR_INVALID;
  fb_ctx* c = s->ctx;
  if (int r = ctx_current(c)) return r;
  if (s->n) {
    if (keys) HIP_TRY(c, hipMemcpyAsync(keys, s->keys.p, s->n * s->nref * 8, hipMemcpyDeviceToHost, c->stream));
    if (cov) HIP_TRY(c, hipMemcpyAsync(cov, s->cov.p, s->n * 2, hipMemcpyDeviceToHost, c->stream));
    if (mask) HIP_TRY(c, hipMemcpyAsync(mask, s->mask.p, s->n, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" uint64_t fb_solid_cms_bytes(const fb_solid* s) { return s ? fb_cms_table_bytes(&s->cms) : 0; }
extern "C" int fb_solid_cms_download(const fb_solid* s, void* out, uint64_t bytes) {
  if (!s) return FB_ERR_INVALID;
  if (int r = ctx_current(s->ctx)) return r;
  return cms_download(s->ctx, s->cms.dev(), s->cms.width, s->cms.bits, out, bytes);
}
extern "C" int fb_solid_hll_registers(const fb_solid* s, uint8_t* out) {
  if (!s || !s->has_hll) return FB_ERR_INVALID;
  if (int r = ctx_current(s->ctx)) return r;
  return hll_download(s->ctx, s->hll.dev(), out);
}
extern "C" int fb_solid_timings(const fb_solid* s, float out[4]) {
  if (!s || !out) return FB_ERR_INVALID;
  memcpy(out, s->ms, sizeof s->ms);
  return FB_OK;
}

extern "C" int fb_solid_kernel_timings(const fb_solid* s, float out[5]) {
  if (!s || !out) return FB_ERR_INVALID;
  memcpy(out, s->kms, sizeof s->kms);
  return FB_OK;
}
extern "C" int fb_solid_stats(const fb_solid* s, uint64_t out[11]) {
  if (!s || !out) return FB_ERR_INVALID;
  memcpy(out, s->stats, sizeof s->stats);
  return FB_OK;
}
extern "C" int fb_solid_pass2_gate(const fb_solid* s) { return s ? s->p2_gate : FB_ERR_INVALID; }
extern "C" int fb_solid_pass2_layout(const fb_solid* s) { return s ? s->p2_layout : FB_ERR_INVALID; }

// ---------------------------------------------------------------------------
// Genome estimator

extern "C" int fb_estimate_genome(fb_ctx* c, const fb_arena_view* r1, const fb_arena_view* r2,
                                  const fb_count_stats* st, uint64_t max_pairs, double target_cov,
                                  fb_genome_estimate* out) {
  if (!c || !r1 || !r2 || !st || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  memset(out, 0, sizeof *out);
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  // estimator.rs:305: sample_target = min(max_pairs, 150000, exact pairs);
  // the loop also ends when either reader runs dry.
  uint64_t n = std::min<uint64_t>(std::min<uint64_t>(max_pairs, 150000), st->exact_read_pairs);
  n = std::min<uint64_t>(n, std::min(r1->n_reads, r2->n_reads));
  if (n == 0) return set_err(c, FB_ERR_EMPTY, "Input FASTQ files contain 0 valid read pairs");
  const uint32_t K = 21;
  // interleaved read list R1[0], R2[0], R1[1], ... and occurrence bases
  fb_arena_view parts[2] = {*r1, *r2};
  parts[0].n_reads = n;
  parts[1].n_reads = n;
  DBuf data, offs0, lens0;
  uint64_t nr = 0;
  if (int r = upload_reads(c, parts, 2, data, offs0, lens0, &nr, nullptr, nullptr, nullptr)) return r;
  std::vector<uint64_t> h_offs(2 * n);
  std::vector<uint32_t> h_lens(2 * n), occ(2 * n);
  std::vector<uint64_t> m_prefix(n + 1, 0);  // windows of pairs [0, i)
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i) {
    h_offs[2 * i] = r1->offsets[i];
    h_offs[2 * i + 1] = r2->offsets[i] + r1->n_words * 32;
    h_lens[2 * i] = r1->lengths[i];
    h_lens[2 * i + 1] = r2->lengths[i];
    for (int j = 0; j < 2; ++j) {
      occ[2 * i + j] = (uint32_t)t;
      const uint32_t len = h_lens[2 * i + j];
      if (len >= K) t += len - K + 1;
    }
    m_prefix[i + 1] = t;
    if (t >= 0xFFFFFFFFull) return set_err(c, FB_ERR_INVALID, "estimator sample exceeds 2^32 k-mers");
  }
  DBuf offs, lens, occd, first_ins, ftot, fsol;
  if (int r = to_device(c, offs, h_offs.data(), h_offs.size() * 8)) return r;
  if (int r = to_device(c, lens, h_lens.data(), h_lens.size() * 4)) return r;
  if (int r = to_device(c, occd, occ.data(), occ.size() * 4)) return r;
  const uint64_t ncell = 4ull << 23, nreg = (1u << 14) * 64u;
  if (int r = first_ins.get(c, ncell * 4)) return r;
  if (int r = ftot.get(c, nreg * 4)) return r;
  if (int r = fsol.get(c, nreg * 4)) return r;
  HIP_TRY(c, launch_fill_u32(first_ins.as<uint32_t>(), ncell, 0xFFFFFFFFu, s));
  HIP_TRY(c, launch_fill_u32(ftot.as<uint32_t>(), nreg, 0xFFFFFFFFu, s));
  HIP_TRY(c, launch_fill_u32(fsol.as<uint32_t>(), nreg, 0xFFFFFFFFu, s));
  EstArgs a{};
  a.reads.data = data.as<uint64_t>();
  a.reads.offs = offs.as<uint64_t>();
  a.reads.lens = lens.as<uint32_t>();
  a.reads.n_reads = 2 * n;
  a.occ_base = occd.as<uint32_t>();
  a.n_reads = (uint32_t)(2 * n);
  a.first_ins = first_ins.as<uint32_t>();
  a.first_total = ftot.as<uint32_t>();
  a.first_solid = fsol.as<uint32_t>();
  HIP_TRY(c, launch_est_first_insert(a, s));
  HIP_TRY(c, launch_est_solid(a, s));
  std::vector<uint32_t> ft(nreg), fs(nreg);
  HIP_TRY(c, hipMemcpyAsync(ft.data(), ftot.p, nreg * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(fs.data(), fsol.p, nreg * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));

  // Register timelines: events (pair, register, rank) sorted by pair.
  struct Ev {
    uint32_t pair, idx, rho;
  };
  auto events = [&](const std::vector<uint32_t>& f) {
    std::vector<Ev> e;
    for (uint32_t idx = 0; idx < (1u << 14); ++idx)
      for (uint32_t rho = 1; rho < 64; ++rho)
        if (f[idx * 64 + rho] != 0xFFFFFFFFu) e.push_back({f[idx * 64 + rho], idx, rho});
    std::sort(e.begin(), e.end(), [](const Ev& x, const Ev& y) { return x.pair < y.pair; });
    return e;
  };
  std::vector<Ev> et = events(ft), es = events(fs);
  std::vector<uint8_t> regt(1u << 14, 0), regs(1u << 14, 0);
  size_t it = 0, is = 0;
  auto advance = [](const std::vector<Ev>& e, size_t& i, std::vector<uint8_t>& reg, uint64_t upto) {
    while (i < e.size() && e[i].pair < upto) {
      if (e[i].rho > reg[e[i].idx]) reg[e[i].idx] = (uint8_t)e[i].rho;
      ++i;
    }
  };
  uint64_t stop = n;
  for (uint64_t cp = 6000; cp <= n; cp += 2000) {  // pair_count >= 5000 && pair_count % 2000 == 0
    advance(et, it, regt, cp);
    advance(es, is, regs, cp);
    uint64_t u1 = f64_to_u64(round(fb_hll_estimate_registers(regt.data(), 14)));
    uint64_t u2 = f64_to_u64(round(fb_hll_estimate_registers(regs.data(), 14)));
    if (checkpoint_stop(u1, u2, m_prefix[cp])) {
      stop = cp;
      break;
    }
  }
  advance(et, it, regt, stop);
  advance(es, is, regs, stop);
  uint64_t u1 = f64_to_u64(round(fb_hll_estimate_registers(regt.data(), 14)));
  uint64_t u2 = f64_to_u64(round(fb_hll_estimate_registers(regs.data(), 14)));
  estimate_finish(u1, u2, m_prefix[stop], st, target_cov, out);
  out->pairs_sampled = stop;
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Batched form: all isolates of a batch in each launch.  A pass at k is two
// captured HIP graphs, replayed while the layout holds: (1) the sketch clears,
// pass 1 on LDS bins and the gate occupancy of every isolate; (2) after the
// host has sized each isolate's pass 2 from that occupancy, pass 2 and the
// fused pass 3 on the LDS route (in rounds) and pass 3 of the spill tables.
// An isolate whose LDS route overflows its record, spill or output buffer
// re-runs pass 2 alone on the global table, which grows until it fits.

struct TuneKey {
  uint32_t kind;        // 0 batch, 1 single isolate
  uint32_t limbs;       // 64-bit limbs of the k-mer: 1, 2, 3 or 4
  uint32_t bits;        // counter bits
  uint32_t log2w;       // sketch width
  uint32_t log2_bases;  // bases per isolate, rounded to a power of two
  uint32_t log2_n;      // isolates, rounded to a power of two
  bool operator<(const TuneKey& o) const {
    return std::tie(kind, limbs, bits, log2w, log2_bases, log2_n) <
           std::tie(o.kind, o.limbs, o.bits, o.log2w, o.log2_bases, o.log2_n);
  }
};
struct TuneChoice {
  uint32_t placement = 0, split = 0, chunks1 = 0, chunks2 = 0;
  float ms = 0.0f;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune;

struct BatchIso {
  const fb_bank* bank = nullptr;
  uint64_t est_g = 0;
  fb_cms cms;
  fb_hll hll;
  bool has_hll = false;
  // gate bits; pass-1 bins, then pass-2 records; bin cursors; gate occupancy
  DBuf bits, scratch, bcur, occ;
  DBuf gm, bh, runs, st2, p3, slots, keys, cov, mask;
  // pass-1 layout
  uint32_t split = 1, sb = kBinLog2Cells, nbins = 1, chunks1 = 1;
  uint64_t bins_cap = 0, windows = 0, threads1 = 0;
  bool binned = false;
  // pass-2 layout of the LDS route
  bool lds = false;
  uint32_t split2 = 1, chunks2 = 1, log2_rounds = 0;
  uint64_t nb2 = 1, nbr = 1, n_blocks2 = 0, rec_cap = 0, spill_cap = 0, out_cap = 0;
  double est_keys = 0.0;    // keys the LDS tables of pass 2 are sized for
  double est_gated = 0.0;   // keys passing the per-window gate (the global route's table)
  // results of the last run
  uint64_t n_out = 0;
  uint32_t route = 0;  // as fb_solid_stats[4]
};

struct fb_batch {
  fb_ctx* ctx = nullptr;
  fb_build_params prm{};
  std::vector<std::unique_ptr<BatchIso>> iso;
  uint32_t k = 0;
  uint32_t nref = 1;
  // pass 1
  DBuf d_p1, d_bin, d_fill1, d_blk1, d_blkc;
  MultiLaunch m1, mc;
  uint32_t n_fill1 = 0, max_nbins = 1, max_sb = 6;
  bool binned = false;
  bool buffered = false;  // binned with the per-bin buffer kernel
  // pass 2 (LDS isolates only, in lds_idx order)
  std::vector<uint32_t> lds_idx;
  DBuf d_p2, d_p3, d_tabs, d_fill2, d_runs_fill, d_blkg, d_blks, d_blkb;
  MultiLaunch mg, ms2, mb;
  uint32_t n_fill2 = 0, max_nbr = 1, max_rounds = 1, lds_cap = 0, filt = 0;
  // A graph reads the per-isolate pointers from the argument arrays, which
  // are re-uploaded before every launch; what it holds fixed are the launch
  // shapes, recorded here per k.
  std::map<uint32_t, hipGraphExec_t> g1, g2;
  std::map<uint32_t, hipGraph_t> g1_def, g2_def;
  std::map<uint32_t, std::vector<uint64_t>> sig1, sig2;
  bool graphs = true;   // false: launch directly (membership changes every run)
  std::map<uint32_t, TuneChoice> tuned;  // per k, from fb_batch_autotune
  size_t cap_iso = 0;   // isolates the argument arrays hold
  float ms[4] = {0, 0, 0, 0};
  void drop(std::map<uint32_t, hipGraphExec_t>& ge, std::map<uint32_t, hipGraph_t>& gd, uint32_t kk) {
    auto e = ge.find(kk);
    if (e != ge.end()) {
      (void)hipGraphExecDestroy(e->second);
      ge.erase(e);
    }
    auto d = gd.find(kk);
    if (d != gd.end()) {
      (void)hipGraphDestroy(d->second);
      gd.erase(d);
    }
  }
  ~fb_batch() {
    for (auto& g : g1) (void)hipGraphExecDestroy(g.second);
    for (auto& g : g2) (void)hipGraphExecDestroy(g.second);
    for (auto& g : g1_def) (void)hipGraphDestroy(g.second);
    for (auto& g : g2_def) (void)hipGraphDestroy(g.second);
  }
};

static uint64_t batch_cms_width(const fb_build_params* p, uint64_t est_g) {
  return next_pow2(p->cms_width ? p->cms_width : fb_cms_width_for_genome_size(est_g));
}

// Pass-2 sizing of the LDS route from the estimated keys (as fb_build_solid).
struct P2Plan {
  uint64_t nb2 = 1, nbr = 1, spill_cap = 0, out_cap = 0, rec_cap = 0, n_blocks2 = 0, threads2 = 0;
  uint32_t log2_rounds = 0, split2 = 1, chunks2 = 1, lds_cap = 0;
  bool ok = false;
};
// lds_cap: LDS table entries of one bin; round_bins: bins one round takes.
// load: the LDS tables' target fill at the estimate (0.5; kP2UngatedLoad for the
// ungated form, whose estimate has a margin of its own); n_distinct (the
// ungated form): the spill table's size when bins2 forces fewer bins than
// the estimate asks for, since their record filters then let most keys in.
static P2Plan plan_pass2(double n_est, uint64_t windows, uint64_t n_reads, uint32_t max_len, uint32_t k,
                         uint32_t split, uint32_t bins2, uint32_t lds_cap, uint32_t round_bins, uint32_t chunks2,
                         double load = 0.5, double n_distinct = 0.0) {
  P2Plan q;
  q.lds_cap = lds_cap;
  while ((double)q.nb2 * load * (double)q.lds_cap < n_est) q.nb2 <<= 1;
  const uint64_t max_bins = (uint64_t)round_bins << kP2MaxLog2Rounds;
  if (bins2) q.nb2 = bins2;
  while ((q.nb2 >> q.log2_rounds) > round_bins) ++q.log2_rounds;
  q.nbr = q.nb2 >> q.log2_rounds;
  const uint32_t pairs_max = max_len > k ? max_len - k : 1;
  q.split2 = std::max<uint32_t>(split, (pairs_max + 62) / 63);
  q.threads2 = n_reads * (uint64_t)q.split2;
  q.rec_cap = windows + q.threads2;
  q.chunks2 = chunks2 ? chunks2 : scatter_chunks(q.threads2);
  q.n_blocks2 = scatter_blocks(q.threads2, q.chunks2);
  const double excess = std::max(0.0, n_est - (double)q.nb2 * load * (double)q.lds_cap);
  const double spilled = n_distinct > 0.0 && excess > 0.0 ? n_distinct : 2.0 * excess + n_est / 4;
  q.spill_cap = next_pow2(std::max<uint64_t>(1ull << 16, (uint64_t)spilled));
  const uint64_t lds_out = std::min<uint64_t>(q.nb2 * q.lds_cap, (uint64_t)(1.25 * n_est) + 65536);
  q.out_cap = std::min<uint64_t>(lds_out + q.spill_cap, 0xFFFFFFFFull);
  q.ok = q.nb2 <= max_bins && windows > 0 && q.rec_cap < (1ull << 32);
  return q;
}

extern "C" int fb_plan_pass2(double est_keys, uint64_t windows, uint64_t n_reads, uint32_t max_len, uint32_t k,
                             fb_pass2_plan* out) {
  if (!out || k < 1 || k > 127) return FB_ERR_INVALID;
  const uint32_t kw = key_words_for_k(k);
  const uint32_t lds_cap = ((kP2LdsBytes - 1024u) / p2_entry_bytes(kw)) & ~1u;  // whole 2-slot buckets
  const P2Plan q = plan_pass2(est_keys, windows, n_reads, max_len, k, default_split(k, max_len), 0, lds_cap, 16384u, 0);
  out->nb2 = q.nb2;
  out->nbr = q.nbr;
  out->spill_cap = q.spill_cap;
  out->out_cap = q.out_cap;
  out->rec_cap = q.rec_cap;
  out->n_blocks2 = q.n_blocks2;
  out->threads2 = q.threads2;
  out->log2_rounds = q.log2_rounds;
  out->split2 = q.split2;
  out->chunks2 = q.chunks2;
  out->lds_cap = q.lds_cap;
  out->ok = q.ok ? 1u : 0u;
  return FB_OK;
}

// Linear-counting estimate of the distinct keys landing in gate-passing
// row-0 cells: an upper-side estimate of the keys pass 2 inserts.
static double est_keys_from_occ(uint64_t occ, uint64_t width, uint64_t windows) {
  const double W = (double)width;
  double n = occ >= width ? (double)windows : -W * log(1.0 - (double)occ / W);
  return std::min(n, (double)windows + 1.0);
}

// Device bytes of one isolate of a batch at k, by buffer.  The buffers of an
// isolate kept across a k ladder only grow, so the ladder holds, per buffer,
// its largest size over the ladder's k (serve_footprint).
struct IsoBytes {
  uint64_t fixed = 0;    // counts, gate bits, HLL registers, the bank
  uint64_t scratch = 0;  // pass-1 bins, then pass-2 records
  uint64_t masks = 0;    // gate masks
  uint64_t runs = 0;     // per-block run matrix
  uint64_t spill = 0;    // spill table
  uint64_t out = 0;      // solid set
  uint64_t total() const { return fixed + scratch + masks + runs + spill + out; }
  void take_max(const IsoBytes& o) {
    fixed = std::max(fixed, o.fixed);
    scratch = std::max(scratch, o.scratch);
    masks = std::max(masks, o.masks);
    runs = std::max(runs, o.runs);
    spill = std::max(spill, o.spill);
    out = std::max(out, o.out);
  }
};

static IsoBytes iso_bytes(uint64_t genome_size, uint64_t bases, uint64_t reads, const fb_build_params* p,
                          uint32_t k) {
  IsoBytes b;
  k = std::max<uint32_t>(k, 1);
  const uint64_t est_g = std::max<uint64_t>(genome_size, 50000);
  const uint64_t W = batch_cms_width(p, est_g);
  const uint64_t windows = bases;  // upper bound of the windows at any k
  const uint32_t max_len = reads ? (uint32_t)std::min<uint64_t>(bases / reads + 1, 1u << 20) : 1u;
  // the key estimate of pass 2 reaches the window count when the sketch is
  // overloaded (every row-0 cell passes the gate), so the footprint takes that
  const double n_est = (double)windows;
  // host-only sizing: the LDS of a workgroup with the full 160 KB granted
  const uint32_t kw = key_words_for_k(k);
  const uint32_t lds_cap = ((kP2LdsBytes - 1024u) / p2_entry_bytes(kw)) & ~1u;  // whole 2-slot buckets
  const P2Plan q = plan_pass2(n_est, windows, reads, max_len, k, default_split(k, max_len), p->bins2, lds_cap, 16384u,
                              p->chunks2);
  const uint32_t rw = (2 * k + 8 + 63) / 64;
  b.fixed = W * 4 * 4 + (W * 4 + 31) / 32 * 4;               // exact counts + gate bits
  b.fixed += p->hll_p ? (4ull << p->hll_p) : 0;               // HLL registers
  b.fixed += bases / 4 + reads * 12 + 4096;                   // the bank itself
  const uint64_t thr1 = reads * (uint64_t)default_split(k, max_len);
  const uint64_t n_wg1 = scatter_blocks(thr1, p->chunks1 ? p->chunks1 : p1_chunks(thr1), kP1BlockSize);
  const uint64_t nbins1 = std::max<uint64_t>(1, (W * 4) >> kBinLog2Cells);
  const uint64_t bins = (uint64_t)(4.0 * (double)windows * 1.1 * 2) + 65536 + nbins1 * (n_wg1 + 16) * kP1Chunk * 2;
  b.scratch = std::max<uint64_t>(bins, q.rec_cap * rw * 8 + 64);  // pass-1 bins, then pass-2 records
  b.masks = q.threads2 * 8;
  b.runs = q.n_blocks2 * q.nbr * 4;
  b.spill = q.spill_cap * table_stride(key_words_for_k(k)) * 8;
  b.out = q.out_cap * (ref_words(k) * 8 + 3);
  return b;
}

extern "C" uint64_t fb_batch_footprint(uint64_t genome_size, uint64_t bases, uint64_t reads,
                                       const fb_build_params* p, uint32_t k) {
  if (!p) return 0;
  return iso_bytes(genome_size, bases, reads, p, k).total();
}

// Sketches and fixed buffers of one isolate of a batch.
static int batch_iso_init(fb_ctx* c, BatchIso* it, const fb_bank* bank, uint64_t genome_size,
                          const fb_build_params* prm) {
  it->bank = bank;
  it->est_g = std::max<uint64_t>(genome_size, 50000);
  if (int r = cms_init(c, &it->cms, batch_cms_width(prm, it->est_g), prm->counter_bits)) return r;
  if (prm->hll_p) {
    if (int r = hll_init(c, &it->hll, prm->hll_p)) return r;
    it->has_hll = true;
  }
  if (int r = it->bits.get(c, ((it->cms.width * 4 + 31) / 32) * 4)) return r;
  if (int r = it->occ.get(c, 8 + 256 * 4)) return r;  // gate occupancy (u64), row-0 count histogram
  if (int r = it->p3.get(c, pass3_scratch_bytes())) return r;
  if (int r = it->st2.get(c, 64)) return r;
  return FB_OK;
}

// Argument arrays for up to n isolates.
static int batch_reserve(fb_batch* b, size_t n) {
  if (n <= b->cap_iso) return FB_OK;
  fb_ctx* c = b->ctx;
  const size_t nn = std::max<size_t>(n, 2 * b->cap_iso);
  if (int r = b->d_p1.get(c, nn * sizeof(Pass1Args))) return r;
  if (int r = b->d_bin.get(c, nn * sizeof(BinArgs))) return r;
  if (int r = b->d_fill1.get(c, nn * 4 * sizeof(FillRange))) return r;
  if (int r = b->d_blk1.get(c, (nn + 1) * 4)) return r;
  if (int r = b->d_blkc.get(c, (nn + 1) * 4)) return r;
  if (int r = b->d_p2.get(c, nn * sizeof(P2BinArgs))) return r;
  if (int r = b->d_p3.get(c, nn * sizeof(P3Args))) return r;
  if (int r = b->d_tabs.get(c, nn * sizeof(DevSolidTable))) return r;
  if (int r = b->d_fill2.get(c, nn * sizeof(FillRange))) return r;
  if (int r = b->d_runs_fill.get(c, nn * sizeof(FillRange))) return r;
  if (int r = b->d_blkg.get(c, (nn + 1) * 4)) return r;
  if (int r = b->d_blks.get(c, (nn + 1) * 4)) return r;
  if (int r = b->d_blkb.get(c, (nn + 1) * 4)) return r;
  b->cap_iso = nn;
  // the arrays moved: every captured graph is stale
  for (auto& g : b->g1) (void)hipGraphExecDestroy(g.second);
  for (auto& g : b->g2) (void)hipGraphExecDestroy(g.second);
  for (auto& g : b->g1_def) (void)hipGraphDestroy(g.second);
  for (auto& g : b->g2_def) (void)hipGraphDestroy(g.second);
  b->g1.clear();
  b->g2.clear();
  b->g1_def.clear();
  b->g2_def.clear();
  return FB_OK;
}

extern "C" int fb_batch_create(fb_ctx* c, const fb_bank* const* banks, const uint64_t* genome_sizes, uint32_t n,
                               const fb_build_params* prm, fb_batch** out) {
  if (!c || !banks || !genome_sizes || !prm || !out || n == 0) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (prm->counter_bits != 4 && prm->counter_bits != 16)
    return set_err(c, FB_ERR_INVALID, "counter_bits must be 4 or 16 (got %d)", prm->counter_bits);
  if (prm->hll_p != 0 && (prm->hll_p < 4 || prm->hll_p > 18))
    return set_err(c, FB_ERR_INVALID, "hll_p %d outside 4..18", prm->hll_p);
  if (prm->bins2 && (prm->bins2 & (prm->bins2 - 1)))
    return set_err(c, FB_ERR_INVALID, "bins2 must be a power of two (got %u)", prm->bins2);
  if (int r = ctx_current(c)) return r;
  auto b = std::make_unique<fb_batch>();
  b->ctx = c;
  b->prm = *prm;
  for (uint32_t i = 0; i < n; ++i) {
    if (!banks[i]) return set_err(c, FB_ERR_INVALID, "bank %u is null", i);
    auto it = std::make_unique<BatchIso>();
    if (int r = batch_iso_init(c, it.get(), banks[i], genome_sizes[i], prm)) return r;
    b->iso.push_back(std::move(it));
  }
  if (int r = batch_reserve(b.get(), n)) return r;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *out = b.release();
  return FB_OK;
}

extern "C" void fb_batch_destroy(fb_batch* b) { delete b; }
extern "C" uint32_t fb_batch_size(const fb_batch* b) { return b ? (uint32_t)b->iso.size() : 0; }
extern "C" int fb_batch_key_words(const fb_batch* b) { return b ? (int)b->nref : 0; }
extern "C" uint64_t fb_batch_count(const fb_batch* b, uint32_t i) {
  return (b && i < b->iso.size()) ? b->iso[i]->n_out : 0;
}

// ---------------------------------------------------------------------------
// Autotuner: launch shapes (pass-1 placement, threads per read, thread
// shares per scatter block in passes 1 and 2) picked by timing candidates.
// Every candidate computes the same sets (the shapes only change which
// thread handles which windows); the tuner checks that it does.  Choices are
// kept per workload shape for the process and can be saved and loaded.

static uint32_t ceil_log2(uint64_t x) {
  uint32_t l = 0;
  while ((1ull << l) < x && l < 63) ++l;
  return l;
}

static TuneKey batch_tune_key(const fb_batch* b, uint32_t k) {
  uint64_t bases = 0, w = 1;
  for (auto& it : b->iso) {
    bases += fb_bank_bases_for_k(it->bank, k);
    w = std::max<uint64_t>(w, it->cms.width);
  }
  const uint64_t n = std::max<size_t>(b->iso.size(), 1);
  return TuneKey{0, (2 * k + 63) / 64, (uint32_t)b->prm.counter_bits, ceil_log2(w), ceil_log2(bases / n + 1),
                 ceil_log2(n)};
}

// The batch's parameters at k: explicit fields win, then the batch's own
// tuned choice, then a tuned choice of the same workload shape.
static fb_build_params batch_params(const fb_batch* b, uint32_t k) {
  fb_build_params p = b->prm;
  TuneChoice t;
  bool have = false;
  auto own = b->tuned.find(k);
  if (own != b->tuned.end()) {
    t = own->second;
    have = true;
  } else {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(batch_tune_key(b, k));
    if (it != g_tune.end()) {
      t = it->second;
      have = true;
    }
  }
  if (have) {
    if (!p.placement) p.placement = t.placement;
    if (!p.split) p.split = t.split;
    if (!p.chunks1) p.chunks1 = t.chunks1;
    if (!p.chunks2) p.chunks2 = t.chunks2;
  }
  return p;
}

// Pass-1 layout and argument arrays; *changed when a launch shape differs
// from the one the graph of this k was captured with.
static int batch_prepare1(fb_batch* b, uint32_t k, uint32_t t_solid, bool* changed) {
  fb_ctx* c = b->ctx;
  const uint32_t n = (uint32_t)b->iso.size();
  std::vector<Pass1Args> p1(n);
  std::vector<BinArgs> ba(n);
  std::vector<FillRange> fills;
  std::vector<uint32_t> blk1(n + 1, 0), blkc(n + 1, 0);
  b->max_nbins = 1;
  b->max_sb = 6;
  const fb_build_params prm = batch_params(b, k);
  // one pass-1 placement for the whole batch: LDS bins unless asked for
  // global atomics or a sketch has more bins than the scatter kernel holds
  bool binned = prm.placement != 1;
  for (auto& it : b->iso) {
    uint32_t sb = kBinLog2Cells;
    while (sb > 6 && (1ull << sb) > it->cms.width * 4) --sb;
    if (((it->cms.width * 4) >> sb) > 8192) binned = false;
  }
  if (prm.placement == 2 && !binned)
    return set_err(c, FB_ERR_INVALID, "LDS placement needs at most 8192 pass-1 bins per isolate");
  b->binned = binned;
  // the per-bin buffer kernel when every isolate's bins fit it: each isolate
  // gets threads / (block * chunks1) blocks striding over its thread shares
  bool buffered = binned;
  uint32_t most_bins = 1;
  for (auto& it : b->iso) {
    uint32_t sb = kBinLog2Cells;
    while (sb > 6 && (1ull << sb) > it->cms.width * 4) --sb;
    most_bins = std::max<uint32_t>(most_bins, (uint32_t)std::max<uint64_t>(1, (it->cms.width * 4) >> sb));
  }
  // one kernel serves the whole batch: buffered when the most bins of any
  // isolate take it (an isolate with fewer bins runs it with its own count)
  if (!bin_buffered(most_bins)) buffered = false;
  b->buffered = buffered;
  const uint32_t blk_threads = buffered ? kP1BufBlock : kP1BlockSize;
  for (uint32_t i = 0; i < n; ++i) {
    BatchIso& it = *b->iso[i];
    const uint64_t ncnt = it.cms.width * 4;
    it.windows = bank_windows(it.bank, k);
    it.split = prm.split ? prm.split : default_split(k, it.bank->max_len);
    it.threads1 = it.bank->n_reads * (uint64_t)it.split;
    uint32_t sb = kBinLog2Cells;
    while (sb > 6 && (1ull << sb) > ncnt) --sb;
    const uint64_t nbins = std::max<uint64_t>(1, ncnt >> sb);
    it.sb = sb;
    it.nbins = (uint32_t)nbins;
    it.binned = binned;
    it.chunks1 = prm.chunks1 ? prm.chunks1 : p1_chunks(it.threads1);
    if (binned) {
      const uint64_t n_wg = scatter_blocks(it.threads1, it.chunks1, blk_threads);
      const uint64_t cap =
          std::min<uint64_t>(bin_slice_cap(4.0 * (double)it.windows / (double)nbins, n_wg), kP1MaxCap);
      const uint64_t need = nbins * kBinSlices * cap * 2;
      if (need > it.scratch.bytes || !it.scratch.p)
        if (int r = it.scratch.get(c, need)) return r;
      it.bins_cap = cap;
      if (!it.bcur.p)
        if (int r = it.bcur.get(c, 8192 * kBinSlices * 4)) return r;
      b->max_nbins = std::max<uint32_t>(b->max_nbins, it.nbins);
      b->max_sb = std::max<uint32_t>(b->max_sb, sb);
    }
    DevReads rv = it.bank->view();
    fills.push_back({it.cms.counts.as<uint32_t>(), ncnt, 0u});
    fills.push_back({it.occ.as<uint32_t>(), 2 + 256, 0u});
    if (it.has_hll) fills.push_back({it.hll.regs.as<uint32_t>(), (uint64_t)1 << it.hll.p, 0u});
    if (binned) fills.push_back({it.bcur.as<uint32_t>(), nbins * kBinSlices, 0u});
    Pass1Args& a1 = p1[i];
    a1.reads = rv;
    a1.k = k;
    a1.split = it.split;
    a1.weight = 1;
    a1.cms = it.cms.dev();
    if (it.has_hll) a1.hll = it.hll.dev();
    BinArgs& bi = ba[i];
    bi.reads = rv;
    bi.k = k;
    bi.split = it.split;
    bi.chunks = it.chunks1;
    bi.log2w = it.cms.log2w;
    bi.n_cells = ncnt;
    bi.counts = it.cms.counts.as<uint32_t>();
    if (it.has_hll) bi.hll = it.hll.dev();
    bi.sb = sb;
    bi.nbins = it.nbins;
    bi.cap = it.bins_cap;
    bi.data = it.scratch.as<uint16_t>();
    bi.cursor = it.bcur.as<uint32_t>();
    bi.bits = it.bits.as<uint32_t>();
    bi.t_solid = t_solid;
    bi.cap_value = it.cms.cap();
    bi.grid = buffered ? (uint32_t)scatter_blocks(it.threads1, it.chunks1, blk_threads) : 0u;
    blk1[i + 1] = blk1[i] + (uint32_t)(binned ? scatter_blocks(it.threads1, it.chunks1, blk_threads)
                                              : (it.threads1 + kPassBlock - 1) / kPassBlock);
    blkc[i + 1] = blkc[i] + (binned ? it.nbins : 0u);
  }
  if (fills.size() > 4 * (size_t)n) return set_err(c, FB_ERR_INVALID, "internal: fill ranges");
  hipStream_t s = c->stream;
  HIP_TRY(c, hipMemcpyAsync(b->d_p1.p, p1.data(), n * sizeof(Pass1Args), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_bin.p, ba.data(), n * sizeof(BinArgs), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_fill1.p, fills.data(), fills.size() * sizeof(FillRange), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_blk1.p, blk1.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_blkc.p, blkc.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipStreamSynchronize(s));  // host vectors go out of scope
  b->m1 = MultiLaunch{binned ? b->d_bin.p : b->d_p1.p, b->d_blk1.as<uint32_t>(), n, blk1[n]};
  b->mc = MultiLaunch{b->d_bin.p, b->d_blkc.as<uint32_t>(), n, blkc[n]};
  b->n_fill1 = (uint32_t)fills.size();
  const std::vector<uint64_t> sig = {binned, buffered, b->n_fill1, blk1[n], blkc[n], b->max_nbins, b->max_sb};
  if (b->sig1[k] != sig) *changed = true;
  b->sig1[k] = sig;
  return FB_OK;
}

// Graph 1: clears, pass 1, gate bits, gate occupancy of row 0.
static int batch_enqueue1(fb_batch* b, uint32_t k, uint32_t t_solid) {
  fb_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  const bool hll = b->prm.hll_p != 0;
  HIP_TRY(c, launch_fill_ranges(b->d_fill1.as<FillRange>(), b->n_fill1, s));
  if (b->binned) {
    HIP_TRY(c, launch_bin_scatter_multi(b->m1, k, hll, b->max_nbins, b->buffered, s));
    HIP_TRY(c, launch_bin_count_multi(b->mc, b->max_sb, s));
  } else {
    HIP_TRY(c, launch_pass1_multi(b->m1, k, true, hll, s));
    for (auto& it : b->iso)
      HIP_TRY(c, launch_solid_bits(it->cms.counts.as<uint32_t>(), it->cms.width * 4, t_solid, it->cms.cap(),
                                   it->bits.as<uint32_t>(), s));
  }
  for (auto& it : b->iso) {
    HIP_TRY(c, launch_popcount(it->bits.as<uint32_t>(), it->cms.width / 32 ? it->cms.width / 32 : 1,
                               it->occ.as<unsigned long long>(), s));
    HIP_TRY(c, launch_count_hist(it->cms.counts.as<uint32_t>(), it->cms.width, it->occ.as<uint32_t>() + 2, s));
  }
  return FB_OK;
}

// Pass-2 layout from the gate occupancies; *changed as for pass 1.
// occ: per isolate the gate occupancy (u64) and the row-0 count histogram
// (256 u32), as graph 1 leaves them.
static int batch_prepare2(fb_batch* b, uint32_t k, uint32_t t_solid, uint32_t gate, const std::vector<uint32_t>& occ,
                          bool* changed) {
  fb_ctx* c = b->ctx;
  const fb_build_params prm = batch_params(b, k);
  const uint32_t n = (uint32_t)b->iso.size();
  const uint32_t kw = key_words_for_k(k), stride = table_stride(kw), nref = ref_words(k);
  const uint32_t rw = (2 * k + 8 + 63) / 64;
  std::vector<uint32_t> lds_idx;
  std::vector<P2BinArgs> p2;
  std::vector<P3Args> p3;
  std::vector<DevSolidTable> tabs;
  std::vector<FillRange> fill2, runs_fill;
  std::vector<uint32_t> blkg(1, 0), blks(1, 0), blkb(1, 0);
  b->max_nbr = 1;
  b->max_rounds = 1;
  // the ungated pass-2 form wherever it applies (as fb_build_solid; the
  // counters' cap is the batch's)
  const uint32_t cap_value = prm.counter_bits == 4 ? 15u : 65535u;
  const bool ungated = prm.placement != 3 && t_solid <= cap_value;
  const uint32_t filt = ungated && gate >= 2 ? std::min(gate, kP2MaxFilt) : 0u;
  b->filt = filt;
  for (uint32_t i = 0; i < n; ++i) {
    BatchIso& it = *b->iso[i];
    const uint32_t* oi = occ.data() + (size_t)i * (2 + 256);
    uint64_t occ_i = 0;
    memcpy(&occ_i, oi, 8);
    it.est_gated = est_keys_from_occ(occ_i, it.cms.width, it.windows);
    it.est_keys = ungated ? est_kept_keys(oi + 2, it.cms.width, it.windows, t_solid, std::max(filt, 1u))
                          : it.est_gated;
    const double n_distinct = ungated ? est_distinct_keys(oi + 2, it.cms.width, it.windows) : it.est_keys;
    const P2Plan q = plan_pass2(it.est_keys, it.windows, it.bank->n_reads, it.bank->max_len, k, it.split,
                                prm.bins2, p2_lds_cap(kw, filt), p2_max_bins(), prm.chunks2, ungated ? kP2UngatedLoad : 0.5,
                                ungated ? n_distinct : 0.0);
    const bool lds = q.ok && prm.placement != 1;
    it.lds = lds;
    if (!lds) continue;
    it.nb2 = q.nb2;
    it.nbr = q.nbr;
    it.log2_rounds = q.log2_rounds;
    it.split2 = q.split2;
    it.chunks2 = q.chunks2;
    it.n_blocks2 = q.n_blocks2;
    it.rec_cap = q.rec_cap;
    b->lds_cap = q.lds_cap;
    auto grow = [&](DBuf& d, uint64_t bytes) -> int { return (d.p && d.bytes >= bytes) ? FB_OK : d.get(c, bytes); };
    int r = grow(it.scratch, q.rec_cap * rw * 8 + 64);  // the bin kernel reads a word past the last record
    if (!r && !ungated) r = grow(it.gm, std::max<uint64_t>(q.threads2, 1) * 8);
    if (!r) r = grow(it.bh, std::max<uint64_t>(q.n_blocks2 * q.nbr, 1) * 4);
    if (!r) r = grow(it.runs, (q.nbr * kBinSlices + 1) * 4);
    if (!r) r = grow(it.slots, q.spill_cap * stride * 8);
    it.spill_cap = q.spill_cap;
    if (!r) r = solid_out_grow(c, it.keys, it.cov, it.mask, q.out_cap, nref);
    if (!r) r = solid_out_check(c, it.keys, it.cov, it.mask, q.out_cap, nref);
    if (r) return r;
    it.out_cap = q.out_cap;
    DevSolidTable t;
    t.slots = it.slots.as<uint64_t>();
    t.cap = it.spill_cap;
    t.kw = kw;
    t.stride = stride;
    t.max_probe = 4096;
    t.status = it.st2.as<uint32_t>() + 8;  // [8] spill overflow, [9] claimed
    P2BinArgs pa;
    pa.reads = it.bank->view();
    pa.k = k;
    pa.split = it.split2;
    pa.chunks = it.chunks2;
    pa.blk_hist = it.bh.as<uint32_t>();
    pa.n_blocks = (uint32_t)it.n_blocks2;
    pa.solid_bits = it.bits.as<uint32_t>();
    pa.log2w = it.cms.log2w;
    pa.nbins = (uint32_t)it.nbr;
    pa.log2_nbins = (uint32_t)__builtin_ctzll(it.nbr);
    pa.log2_rounds = it.log2_rounds;
    pa.gate_masks = ungated ? nullptr : it.gm.as<uint64_t>();
    pa.rw = rw;
    pa.runs = it.runs.as<uint32_t>();
    pa.recs = it.scratch.as<uint64_t>();
    pa.rec_cap = it.rec_cap;
    pa.ungated = ungated ? 1u : 0u;
    pa.t_solid = t_solid;
    pa.filt = filt;
    pa.status = it.st2.as<uint32_t>();          // [0] record overflow, [1] records
    pa.out_count = it.st2.as<uint32_t>() + 2;   // [2] keys emitted from LDS
    pa.stats = reinterpret_cast<unsigned long long*>(it.st2.as<uint32_t>() + 4);  // [4..7]
    pa.lds_cap = q.lds_cap;
    pa.min_trans_cov = gate;
    pa.table = t;
    pa.nref = nref;
    pa.out_keys = it.keys.as<uint64_t>();
    pa.out_cov = it.cov.as<uint16_t>();
    pa.out_mask = it.mask.as<uint8_t>();
    pa.out_cap = (uint32_t)it.out_cap;
    p2.push_back(pa);
    P3Args q3;
    q3.t = t;
    q3.nref = nref;
    q3.gate = gate;
    q3.out_keys = pa.out_keys;
    q3.out_cov = pa.out_cov;
    q3.out_mask = pa.out_mask;
    q3.scratch = it.p3.p;
    q3.out_base = pa.out_count;
    q3.out_cap = pa.out_cap;
    p3.push_back(q3);
    tabs.push_back(t);
    fill2.push_back({it.st2.as<uint32_t>(), 8, 0u});  // [8..9] are cleared with the table
    runs_fill.push_back({it.runs.as<uint32_t>(), it.nbr * kBinSlices + 1, 0u});
    const uint64_t thr2 = it.bank->n_reads * (uint64_t)it.split2;
    // the gate kernels take no blocks of an ungated isolate
    blkg.push_back(blkg.back() + (ungated ? 0u : (uint32_t)((thr2 + kP2GateBlockSize - 1) / kP2GateBlockSize)));
    blks.push_back(blks.back() + (uint32_t)it.n_blocks2);
    blkb.push_back(blkb.back() + (uint32_t)it.nbr);
    b->max_nbr = std::max<uint32_t>(b->max_nbr, (uint32_t)it.nbr);
    b->max_rounds = std::max<uint32_t>(b->max_rounds, 1u << it.log2_rounds);
    lds_idx.push_back(i);
  }
  b->lds_idx = lds_idx;
  const uint32_t m = (uint32_t)lds_idx.size();
  b->n_fill2 = m;
  b->nref = nref;
  const std::vector<uint64_t> sig = {m,         blkg.back(),   blks.back(), blkb.back(),
                                     b->max_nbr, b->max_rounds, b->lds_cap,  b->filt};
  if (b->sig2[k] != sig) *changed = true;
  b->sig2[k] = sig;
  if (!m) return FB_OK;
  hipStream_t s = c->stream;
  HIP_TRY(c, hipMemcpyAsync(b->d_p2.p, p2.data(), m * sizeof(P2BinArgs), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_p3.p, p3.data(), m * sizeof(P3Args), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_tabs.p, tabs.data(), m * sizeof(DevSolidTable), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_fill2.p, fill2.data(), m * sizeof(FillRange), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_runs_fill.p, runs_fill.data(), m * sizeof(FillRange), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_blkg.p, blkg.data(), (m + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_blks.p, blks.data(), (m + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(b->d_blkb.p, blkb.data(), (m + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  b->mg = MultiLaunch{b->d_p2.p, b->d_blkg.as<uint32_t>(), m, blkg[m]};
  b->ms2 = MultiLaunch{b->d_p2.p, b->d_blks.as<uint32_t>(), m, blks[m]};
  b->mb = MultiLaunch{b->d_p2.p, b->d_blkb.as<uint32_t>(), m, blkb[m]};
  return FB_OK;
}

// Graph 2: passes 2 and 3 of the LDS isolates.
static int batch_enqueue2(fb_batch* b, uint32_t k) {
  fb_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  const uint32_t m = (uint32_t)b->lds_idx.size();
  HIP_TRY(c, launch_fill_ranges(b->d_fill2.as<FillRange>(), m, s));
  HIP_TRY(c, launch_table_clear_multi(b->d_tabs.as<DevSolidTable>(), m, s));
  HIP_TRY(c, launch_p2_binned_multi(b->mg, b->ms2, b->mb, b->d_runs_fill.as<FillRange>(), m, k, b->max_nbr,
                                    b->max_rounds, b->lds_cap, b->filt, s));
  HIP_TRY(c, launch_pass3_multi(b->d_p3.as<P3Args>(), m, key_words_for_k(k), s));
  return FB_OK;
}

// Captures (when absent) and launches the graph of one phase.
template <class F>
static int run_graph(fb_batch* b, std::map<uint32_t, hipGraphExec_t>& ge, std::map<uint32_t, hipGraph_t>& gd,
                     uint32_t k, bool changed, F&& enqueue) {
  fb_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  if (!b->graphs) return enqueue();
  if (changed) b->drop(ge, gd, k);
  if (ge.find(k) == ge.end()) {
    hipGraph_t graph = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = enqueue();
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (r) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    HIP_TRY(c, e);
    hipGraphExec_t exec = nullptr;
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(graph);
      HIP_TRY(c, ei);
    }
    ge[k] = exec;
    gd[k] = graph;
  }
  HIP_TRY(c, hipGraphLaunch(ge[k], s));
  return FB_OK;
}

// Pass 2 and 3 of one isolate on the global table, grown until it fits.
static int batch_global_route(fb_ctx* c, BatchIso& it, uint32_t k, uint32_t gate) {
  hipStream_t s = c->stream;
  const uint32_t kw = key_words_for_k(k), stride = table_stride(kw), nref = ref_words(k);
  uint64_t cap = next_pow2(std::max<uint64_t>(1ull << 16, (uint64_t)(2.0 * it.est_gated) + 1024));
  DBuf tbuf, status;
  if (int r = status.get(c, 16)) return r;
  DevSolidTable t;
  t.kw = kw;
  t.stride = stride;
  t.max_probe = 4096;
  t.status = status.as<uint32_t>();
  uint32_t st[2] = {0, 0};
  for (int attempt = 0;; ++attempt) {
    t.cap = cap;
    if (int r = tbuf.get(c, cap * stride * 8)) return r;
    t.slots = tbuf.as<uint64_t>();
    HIP_TRY(c, launch_table_clear(t, s));
    Pass2Args a2{};
    a2.reads = it.bank->view();
    a2.k = k;
    a2.split = it.split;
    a2.solid_bits = it.bits.as<uint32_t>();
    a2.log2w = it.cms.log2w;
    a2.table = t;
    a2.min_trans_cov = gate;
    HIP_TRY(c, launch_pass2(a2, s));
    HIP_TRY(c, hipMemcpyAsync(st, t.status, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    if (!st[0]) break;
    if (attempt >= 6) return set_err(c, FB_ERR_OVERFLOW, "solid table overflow at %llu slots", (unsigned long long)cap);
    cap *= 2;
  }
  const uint64_t maxn = std::max<uint64_t>(st[1], 1);
  if (int r = solid_out_grow(c, it.keys, it.cov, it.mask, maxn, nref)) return r;
  if (int r = solid_out_check(c, it.keys, it.cov, it.mask, maxn, nref)) return r;
  it.out_cap = maxn;
  HIP_TRY(c, launch_pass3(t, k, nref, gate, it.keys.as<uint64_t>(), it.cov.as<uint16_t>(), it.mask.as<uint8_t>(),
                          it.p3.p, s));
  uint32_t cnt[4] = {0, 0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(cnt, it.p3.p, 16, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  it.n_out = cnt[0];
  return FB_OK;
}

extern "C" int fb_batch_run(fb_batch* b, uint32_t k, int32_t target) {
  if (!b) return FB_ERR_INVALID;
  fb_ctx* c = b->ctx;
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  if (target > 255) return set_err(c, FB_ERR_INVALID, "target_solid_cov is a u8");
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  const uint32_t t_solid = target >= 0 ? (uint32_t)target : 2u;
  const uint32_t gate = target >= 0 ? (uint32_t)target : 3u;
  const uint32_t n = (uint32_t)b->iso.size();
  for (auto& it : b->iso)
    if (int r = check_add_budget(c, fb_bank_bases_for_k(it->bank, k))) return r;
  if (int r = batch_reserve(b, n)) return r;
  HIP_TRY(c, hipEventRecord(c->ev[0], s));
  bool changed = false;
  if (int r = batch_prepare1(b, k, t_solid, &changed)) return r;
  if (int r = run_graph(b, b->g1, b->g1_def, k, changed, [&] { return batch_enqueue1(b, k, t_solid); })) return r;
  std::vector<uint32_t> occ((size_t)n * (2 + 256), 0);
  for (uint32_t i = 0; i < n; ++i)
    HIP_TRY(c, hipMemcpyAsync(occ.data() + (size_t)i * (2 + 256), b->iso[i]->occ.p, (2 + 256) * 4,
                              hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  bool changed2 = false;
  if (int r = batch_prepare2(b, k, t_solid, gate, occ, &changed2)) return r;
  if (!b->lds_idx.empty())
    if (int r = run_graph(b, b->g2, b->g2_def, k, changed2, [&] { return batch_enqueue2(b, k); })) return r;
  HIP_TRY(c, hipEventRecord(c->ev[1], s));
  // results of the LDS isolates; the others, and any that overflowed, take the global route
  std::vector<uint32_t> st(16 * (size_t)n, 0), p3c(4 * (size_t)n, 0);
  for (uint32_t i : b->lds_idx) {
    HIP_TRY(c, hipMemcpyAsync(st.data() + 16 * i, b->iso[i]->st2.p, 64, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(p3c.data() + 4 * i, b->iso[i]->p3.p, 16, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(c, hipStreamSynchronize(s));
  HIP_TRY(c, hipEventRecord(c->ev[2], s));
  for (uint32_t i = 0; i < n; ++i) {
    BatchIso& it = *b->iso[i];
    if (it.lds) {
      const uint32_t* q = st.data() + 16 * i;
      const uint64_t n_out = (uint64_t)q[2] + p3c[4 * i];
      if (!q[0] && !q[8] && n_out <= it.out_cap) {
        it.n_out = n_out;
        it.route = 1;
        continue;
      }
      it.route = q[0] ? 2 : (q[8] ? 3 : 4);
    } else {
      it.route = 0;
    }
    if (it.windows == 0 && !it.lds) {
      it.n_out = 0;
      continue;
    }
    if (int r = batch_global_route(c, it, k, gate)) return r;
  }
  HIP_TRY(c, hipEventRecord(c->ev[3], s));
  HIP_TRY(c, hipEventRecord(c->ev[4], s));
  if (b->prm.sort) {
    size_t need = 0;
    for (auto& it : b->iso) {
      size_t tb = 0;
      HIP_TRY(c, sort_solid(nullptr, nullptr, nullptr, (uint32_t)it->n_out, ref_words(k), k, nullptr, &tb, s));
      need = std::max(need, tb);
    }
    DBuf tmp;
    if (int r = tmp.get(c, need)) return r;
    for (auto& it : b->iso) {
      if (it->n_out > 1) {
        size_t tb = need;
        HIP_TRY(c, sort_solid(it->keys.as<uint64_t>(), it->cov.as<uint16_t>(), it->mask.as<uint8_t>(),
                              (uint32_t)it->n_out, ref_words(k), k, tmp.p, &tb, s, true));
      }
    }
  }
  HIP_TRY(c, hipEventRecord(c->ev[5], s));
  HIP_TRY(c, hipEventSynchronize(c->ev[5]));
  b->k = k;
  b->nref = ref_words(k);
  (void)hipEventElapsedTime(&b->ms[0], c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&b->ms[1], c->ev[2], c->ev[3]);
  (void)hipEventElapsedTime(&b->ms[2], c->ev[4], c->ev[5]);
  (void)hipEventElapsedTime(&b->ms[3], c->ev[0], c->ev[5]);
  return FB_OK;
}

extern "C" int fb_batch_download(const fb_batch* b, uint32_t i, uint64_t* keys, uint16_t* cov, uint8_t* mask) {
  if (!b || i >= b->iso.size()) return FB_ERR_INVALID;
  fb_ctx* c = b->ctx;
  if (int r = ctx_current(c)) return r;
  const BatchIso& it = *b->iso[i];
  if (it.n_out) {
    if (keys) HIP_TRY(c, hipMemcpyAsync(keys, it.keys.p, it.n_out * b->nref * 8, hipMemcpyDeviceToHost, c->stream));
    if (cov) HIP_TRY(c, hipMemcpyAsync(cov, it.cov.p, it.n_out * 2, hipMemcpyDeviceToHost, c->stream));
    if (mask) HIP_TRY(c, hipMemcpyAsync(mask, it.mask.p, it.n_out, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FB_OK;
}
extern "C" uint64_t fb_batch_cms_bytes(const fb_batch* b, uint32_t i) {
  return (b && i < b->iso.size()) ? fb_cms_table_bytes(&b->iso[i]->cms) : 0;
}
extern "C" int fb_batch_cms_download(const fb_batch* b, uint32_t i, void* out, uint64_t bytes) {
  if (!b || i >= b->iso.size()) return FB_ERR_INVALID;
  if (int r = ctx_current(b->ctx)) return r;
  const BatchIso& it = *b->iso[i];
  return cms_download(b->ctx, it.cms.dev(), it.cms.width, it.cms.bits, out, bytes);
}
extern "C" int fb_batch_hll_registers(const fb_batch* b, uint32_t i, uint8_t* out) {
  if (!b || i >= b->iso.size() || !b->iso[i]->has_hll) return FB_ERR_INVALID;
  if (int r = ctx_current(b->ctx)) return r;
  return hll_download(b->ctx, b->iso[i]->hll.dev(), out);
}
extern "C" int fb_batch_timings(const fb_batch* b, float out[4]) {
  if (!b || !out) return FB_ERR_INVALID;
  memcpy(out, b->ms, sizeof b->ms);
  return FB_OK;
}
extern "C" int fb_batch_routes(const fb_batch* b, uint32_t* out) {
  if (!b || !out) return FB_ERR_INVALID;
  for (size_t i = 0; i < b->iso.size(); ++i) out[i] = b->iso[i]->route;
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Serving: a request loop with continuous batching.  One worker thread owns
// the server's context.  Every iteration it admits queued requests whose
// device footprint fits the budget (fb::Admission), uploads their reads and
// sizes their sketches (from the genome estimate when asked), then runs one
// batched pass for every k that in-flight requests are at: requests at
// different stages of their k ladder share the iteration, new requests join
// at the next one, and a request that finishes its last pass leaves at once,
// freeing its memory for the queue.  A request's sketches and buffers stay on
// the device across its passes.

// Page-locked host blocks for served solid sets, kept and reused across
// requests.  A download into page-locked memory runs at the host link's
// rate, where one into pageable memory goes through the driver's staging
// copies (and a std::vector would first zero what it is about to receive).
// A block returns to the pool when its request is released; the pool keeps
// at most kCacheBytes of free blocks and frees the rest.
class PinnedPool {
 public:
  static constexpr size_t kGrain = 4u << 20;
  static constexpr size_t kCacheBytes = 8ull << 30;
  ~PinnedPool() {
    for (auto& kv : free_) (void)hipHostFree(kv.second);
  }
  // A block of at least bytes (nullptr when none can be pinned); *got = its size.
  void* get(size_t bytes, size_t* got) {
    bytes = (bytes + kGrain - 1) / kGrain * kGrain;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = free_.lower_bound(bytes);
      if (it != free_.end() && it->first <= bytes + bytes / 2) {
        *got = it->first;
        void* p = it->second;
        cached_ -= it->first;
        free_.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    *got = bytes;
    return p;
  }
  void put(void* p, size_t bytes) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (cached_ + bytes <= kCacheBytes) {
        free_.emplace(bytes, p);
        cached_ += bytes;
        return;
      }
    }
    (void)hipHostFree(p);
  }

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  size_t cached_ = 0;
};

// One pass's solid set on the host: n entries of nref key words, cov and
// mask, in one pinned block (keys, then cov, then mask) or, when no block
// could be pinned, in vectors.
struct ServePass {
  PinnedPool* pool = nullptr;
  void* blk = nullptr;
  size_t blk_bytes = 0;
  uint64_t n = 0;
  uint32_t nref = 1;
  std::vector<uint64_t> keys_v;
  std::vector<uint16_t> cov_v;
  std::vector<uint8_t> mask_v;
  std::vector<uint8_t> cms, hll;  // with keep_sketches
  uint32_t route = 0;

  ServePass() = default;
  ServePass(const ServePass&) = delete;
  ServePass& operator=(const ServePass&) = delete;
  ServePass(ServePass&& o) noexcept { *this = std::move(o); }
  ServePass& operator=(ServePass&& o) noexcept {
    if (this != &o) {
      release();
      pool = o.pool;
      blk = o.blk;
      blk_bytes = o.blk_bytes;
      n = o.n;
      nref = o.nref;
      keys_v = std::move(o.keys_v);
      cov_v = std::move(o.cov_v);
      mask_v = std::move(o.mask_v);
      cms = std::move(o.cms);
      hll = std::move(o.hll);
      route = o.route;
      o.blk = nullptr;
      o.blk_bytes = 0;
    }
    return *this;
  }
  ~ServePass() { release(); }
  void release() {
    if (blk) pool->put(blk, blk_bytes);
    blk = nullptr;
    blk_bytes = 0;
  }
  // Room for n entries of nref words.
  void alloc(PinnedPool* pl, uint64_t entries, uint32_t words) {
    n = entries;
    nref = words;
    const size_t kb = (size_t)n * nref * 8, bytes = kb + (size_t)n * 3;
    if (pl && bytes) blk = pl->get(bytes, &blk_bytes);
    if (blk) {
      pool = pl;
    } else {
      keys_v.resize((size_t)n * nref);
      cov_v.resize(n);
      mask_v.resize(n);
    }
  }
  uint64_t* keys() { return blk ? static_cast<uint64_t*>(blk) : keys_v.data(); }
  uint16_t* cov() {
    return blk ? reinterpret_cast<uint16_t*>(static_cast<char*>(blk) + (size_t)n * nref * 8) : cov_v.data();
  }
  uint8_t* mask() {
    return blk ? reinterpret_cast<uint8_t*>(static_cast<char*>(blk) + (size_t)n * nref * 8 + (size_t)n * 2)
               : mask_v.data();
  }
};

struct ServeReq {
  uint64_t id = 0;
  // host copies of R1 / R2 until the reads are on the device
  std::vector<uint64_t> data[2], offs[2];
  std::vector<uint32_t> lens[2];
  uint64_t genome_hint = 0, bases = 0, reads = 0, footprint = 0;
  uint32_t max_len = 0;
  fb_bank* bank = nullptr;
  std::unique_ptr<BatchIso> iso;
  uint32_t stage = 0;
  int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed
  int status = FB_OK;
  std::string err;
  fb_genome_estimate est{};
  bool has_est = false;
  uint64_t genome_size = 0;
  std::vector<ServePass> passes;
  double t_submit = 0, t_admit = 0, t_done = 0;
  ~ServeReq() {
    iso.reset();
    if (bank) fb_bank_destroy(bank);
  }
};

struct fb_server {
  PinnedPool pinned;  // declared first: destroyed after the requests that hold its blocks
  fb_ctx* ctx = nullptr;
  fb_serve_params prm{};
  fb_build_params bprm{};
  std::unique_ptr<fb::Admission> adm;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<uint64_t, std::unique_ptr<ServeReq>> reqs;
  std::vector<ServeReq*> running;  // worker-owned
  uint64_t next_id = 1;
  bool stop = false;
  std::string err;
  std::chrono::steady_clock::time_point t0;
  std::map<uint32_t, std::unique_ptr<fb_batch>> batches;  // per pass index, membership set every iteration
  fb_serve_stats st{};
  uint32_t waiters = 0;  // fb_server_wait calls in progress (under mu)
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

extern "C" void fb_serve_params_default(fb_serve_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->n_passes = 3;
  p->k[0] = 21;
  p->k[1] = 55;
  p->k[2] = 77;
  p->target_solid_cov[0] = -1;
  p->target_solid_cov[1] = 4;
  p->target_solid_cov[2] = 4;
  p->counter_bits = 16;
  p->hll_p = 14;
  p->estimate = 1;
  p->est_max_pairs = 250000;
  p->target_cov = 100.0;
}

// Genome size a request is sized for before its estimate exists.
static uint64_t serve_size_hint(const fb_server* sv, const ServeReq& r) {
  if (r.genome_hint) return r.genome_hint;
  // at >= 20x coverage the genome is at most bases / 20
  return std::max<uint64_t>(r.bases / (sv->prm.estimate ? 20 : 60), 50000);
}

// A request's buffers only grow along its k ladder: the footprint is, per
// buffer, the largest size over the ladder.
static uint64_t serve_footprint(const fb_server* sv, const ServeReq& r) {
  IsoBytes m;
  for (uint32_t j = 0; j < sv->prm.n_passes; ++j)
    m.take_max(iso_bytes(serve_size_hint(sv, r), r.bases, r.reads, &sv->bprm, sv->prm.k[j]));
  uint64_t f = m.total();
  if (sv->prm.estimate) f += (4ull << 23) * 4 + (3ull << 20) * 8;  // estimator cells and registers
  return f;
}

static void serve_fail(fb_server* sv, ServeReq* r, int code, const std::string& msg) {
  r->iso.reset();
  if (r->bank) {
    fb_bank_destroy(r->bank);
    r->bank = nullptr;
  }
  std::lock_guard<std::mutex> g(sv->mu);
  (void)sv->adm->retire(r->id);
  r->status = code;
  r->err = msg;
  r->state = 3;
  r->t_done = sv->now();
  ++sv->st.failed;
  sv->cv_done.notify_all();
}

// Uploads an admitted request's reads, estimates its genome and creates its sketches.
static int serve_start(fb_server* sv, ServeReq* r) {
  fb_ctx* c = sv->ctx;
  fb_arena_view v[3] = {};
  for (int a = 0; a < 2; ++a) {
    v[a + 1].data = r->data[a].data();
    v[a + 1].n_words = r->data[a].size();
    v[a + 1].offsets = r->offs[a].data();
    v[a + 1].lengths = r->lens[a].data();
    v[a + 1].n_reads = r->lens[a].size();
  }
  if (int e = fb_bank_upload(c, v, &r->bank)) return e;
  r->genome_size = r->genome_hint;
  if (sv->prm.estimate) {
    const uint64_t np = std::min(r->lens[0].size(), r->lens[1].size());
    fb_count_stats cs{};
    if (int e = fb_count_stats_from_lengths(r->lens[0].data(), r->lens[1].data(), np, &cs))
      return set_err(c, e, "Input FASTQ files contain 0 valid read pairs");
    if (int e = fb_estimate_genome(c, &v[1], &v[2], &cs, sv->prm.est_max_pairs, sv->prm.target_cov, &r->est)) return e;
    r->has_est = true;
    if (!r->genome_size) r->genome_size = r->est.genome_size;
  }
  if (!r->genome_size) r->genome_size = std::max<uint64_t>(r->bank->total_bases / 60, 50000);  // build.rs:74-83
  for (int a = 0; a < 2; ++a) {  // the host copies are no longer needed
    std::vector<uint64_t>().swap(r->data[a]);
    std::vector<uint64_t>().swap(r->offs[a]);
  }
  r->iso = std::make_unique<BatchIso>();
  return batch_iso_init(c, r->iso.get(), r->bank, r->genome_size, &sv->bprm);
}

// One batched pass for the requests g, all at pass index j of the ladder.
static int serve_pass(fb_server* sv, uint32_t j, std::vector<ServeReq*>& g) {
  fb_ctx* c = sv->ctx;
  const uint32_t k = sv->prm.k[j];
  const int32_t target = sv->prm.target_solid_cov[j];
  auto& slot = sv->batches[j];
  if (!slot) {
    slot = std::make_unique<fb_batch>();
    slot->ctx = c;
    slot->prm = sv->bprm;
    slot->graphs = false;
  }
  fb_batch* b = slot.get();
  b->iso.clear();
  for (ServeReq* r : g) b->iso.push_back(std::move(r->iso));
  const double t0 = sv->now();
  int rc = fb_batch_run(b, k, target);
  const double t1 = sv->now();
  // results are read while the batch still holds the isolates
  for (size_t i = 0; i < g.size() && !rc; ++i) {
    const BatchIso& it = *b->iso[i];
    ServePass p;
    p.alloc(&sv->pinned, it.n_out, b->nref);
    p.route = it.route;
    rc = fb_batch_download(b, (uint32_t)i, p.keys(), p.cov(), p.mask());
    if (!rc && sv->prm.keep_sketches) {
      p.cms.resize(fb_cms_table_bytes(&it.cms));
      rc = cms_download(c, it.cms.dev(), it.cms.width, it.cms.bits, p.cms.data(), p.cms.size());
      if (!rc && it.has_hll) {
        p.hll.resize((size_t)1 << it.hll.p);
        rc = hll_download(c, it.hll.dev(), p.hll.data());
      }
    }
    if (!rc) g[i]->passes.push_back(std::move(p));
  }
  for (size_t i = 0; i < g.size(); ++i) g[i]->iso = std::move(b->iso[i]);
  b->iso.clear();
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->st.pass_s += t1 - t0;
  sv->st.fetch_s += sv->now() - t1;
  return rc;
}

static void serve_loop(fb_server* sv) {
  (void)ctx_current(sv->ctx);
  for (;;) {
    std::vector<ServeReq*> admitted;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] { return sv->stop || sv->adm->queued() > 0 || !sv->running.empty(); });
      if (sv->stop && sv->running.empty()) break;
      if (!sv->stop) {
        std::vector<uint64_t> ids;
        sv->adm->admit(ids);
        for (uint64_t id : ids) {
          ServeReq* r = sv->reqs[id].get();
          r->state = 1;
          r->t_admit = sv->now();
          admitted.push_back(r);
        }
      }
    }
    for (ServeReq* r : admitted) {
      const double t0 = sv->now();
      const int e = serve_start(sv, r);
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->st.start_s += sv->now() - t0;
      }
      if (e) {
        serve_fail(sv, r, e, sv->ctx->err);
        continue;
      }
      sv->running.push_back(r);
    }
    if (sv->running.empty()) {
      // admitted requests may still be blocked by budget only while others
      // run, so an empty running set with a waiting queue cannot happen
      continue;
    }
    const double t_it = sv->now();
    // one batched pass per ladder position that in-flight requests are at
    std::map<uint32_t, std::vector<ServeReq*>> groups;
    for (ServeReq* r : sv->running) groups[r->stage].push_back(r);
    // A request published as finished (state 2 or 3) may be released, and so
    // freed, by a client at once: the worker takes it out of `running` first
    // and touches it no more after publishing it.
    std::vector<std::pair<ServeReq*, int>> failed;
    std::string fail_msg;
    for (auto& kv : groups) {
      if (int e = serve_pass(sv, kv.first, kv.second)) {
        fail_msg = sv->ctx->err;
        for (ServeReq* r : kv.second) failed.emplace_back(r, e);
        continue;
      }
      for (ServeReq* r : kv.second) ++r->stage;
    }
    if (!failed.empty()) {
      std::vector<ServeReq*> keep;
      for (ServeReq* r : sv->running) {
        bool f = false;
        for (auto& fr : failed) f = f || fr.first == r;
        if (!f) keep.push_back(r);
      }
      sv->running.swap(keep);
      for (auto& fr : failed) serve_fail(sv, fr.first, fr.second, fail_msg);
    }
    {
      std::lock_guard<std::mutex> g(sv->mu);
      ++sv->st.iterations;
      sv->st.pass_runs += sv->running.size();
      sv->st.max_batch = std::max<uint64_t>(sv->st.max_batch, sv->running.size());
      sv->st.max_groups = std::max<uint64_t>(sv->st.max_groups, groups.size());
      sv->st.busy_s += sv->now() - t_it;
    }
    // retire finished requests
    std::vector<ServeReq*> still;
    for (ServeReq* r : sv->running) {
      if (r->stage < sv->prm.n_passes) {
        still.push_back(r);
        continue;
      }
      r->iso.reset();
      fb_bank_destroy(r->bank);
      r->bank = nullptr;
      std::lock_guard<std::mutex> g(sv->mu);
      (void)sv->adm->retire(r->id);
      r->state = 2;
      r->t_done = sv->now();
      ++sv->st.completed;
      sv->cv_done.notify_all();
    }
    sv->running.swap(still);
  }
}

extern "C" int fb_server_create(int device, const fb_serve_params* p, fb_server** out) {
  if (!p || !out) return FB_ERR_INVALID;
  *out = nullptr;
  if (p->n_passes < 1 || p->n_passes > 4) return FB_ERR_INVALID;
  for (uint32_t j = 0; j < p->n_passes; ++j)
    if (p->k[j] < 1 || p->k[j] > 127 || p->target_solid_cov[j] > 255) return FB_ERR_INVALID;
  if (p->counter_bits != 4 && p->counter_bits != 16) return FB_ERR_INVALID;
  if (p->hll_p != 0 && (p->hll_p < 4 || p->hll_p > 18)) return FB_ERR_INVALID;
  auto sv = std::make_unique<fb_server>();
  sv->prm = *p;
  if (int r = fb_ctx_create(device, p->ctx_budget, &sv->ctx)) return r;
  fb_build_params_default(&sv->bprm, p->k[0]);
  sv->bprm.cms_width = p->cms_width;
  sv->bprm.counter_bits = p->counter_bits;
  sv->bprm.hll_p = p->hll_p;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    fb_ctx_destroy(sv->ctx);
    return FB_ERR_HIP;
  }
  // The context caches the buffers of finished requests for the next ones,
  // at most a fifth of the device's free memory, the longest-unused freed
  // first.  Uncached beyond that, it filled the device until an allocation
  // failed and then freed its whole cache at once (seconds of hipFree and
  // hipMalloc in one pass, the p99 of the served stream).
  sv->ctx->pool_cap = (uint64_t)((double)fr * 0.2);
  const uint64_t budget = p->hbm_budget ? p->hbm_budget : (uint64_t)((double)fr * 0.8);
  sv->st.budget = budget;
  sv->adm = std::make_unique<fb::Admission>(budget, p->max_inflight ? p->max_inflight : 64);
  sv->t0 = std::chrono::steady_clock::now();
  fb_server* raw = sv.get();
  sv->worker = std::thread(serve_loop, raw);
  *out = sv.release();
  return FB_OK;
}

extern "C" void fb_server_destroy(fb_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
    // queued requests are refused; in-flight ones finish
    for (auto& kv : sv->reqs)
      if (kv.second->state == 0) {
        kv.second->state = 3;
        kv.second->status = FB_ERR_BUSY;
        kv.second->err = "server stopped";
      }
  }
  sv->cv_done.notify_all();
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  {
    // the waiters of queued and in-flight requests have all been woken
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->waiters == 0; });
  }
  sv->reqs.clear();
  sv->batches.clear();
  fb_ctx_destroy(sv->ctx);
  delete sv;
}

extern "C" const char* fb_server_last_error(const fb_server* sv) { return sv ? sv->err.c_str() : "null server"; }

extern "C" int fb_server_submit(fb_server* sv, const fb_arena_view* r1, const fb_arena_view* r2, uint64_t genome_size,
                                uint64_t* ticket) {
  if (!sv || !r1 || !r2 || !ticket) return FB_ERR_INVALID;
  auto r = std::make_unique<ServeReq>();
  const fb_arena_view* src[2] = {r1, r2};
  for (int a = 0; a < 2; ++a) {
    const fb_arena_view& v = *src[a];
    if (v.n_reads && (!v.data || !v.offsets || !v.lengths)) {
      sv->err = "null arena pointer";
      return FB_ERR_INVALID;
    }
    for (uint64_t i = 0; i < v.n_reads; ++i)
      if (v.lengths[i] && (v.offsets[i] + v.lengths[i] + 31) / 32 > v.n_words) {
        sv->err = "read ends past its arena";
        return FB_ERR_INVALID;
      }
    r->data[a].assign(v.data, v.data + v.n_words);
    r->offs[a].assign(v.offsets, v.offsets + v.n_reads);
    r->lens[a].assign(v.lengths, v.lengths + v.n_reads);
    for (uint64_t i = 0; i < v.n_reads; ++i) {
      r->bases += v.lengths[i];
      r->max_len = std::max(r->max_len, v.lengths[i]);
    }
    r->reads += v.n_reads;
  }
  r->genome_hint = genome_size;
  r->footprint = serve_footprint(sv, *r);
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stop) return FB_ERR_BUSY;
  if (sv->adm->queued() >= (sv->prm.max_queue ? sv->prm.max_queue : 4096)) {
    sv->err = "serving queue full";
    return FB_ERR_BUSY;
  }
  r->id = sv->next_id++;
  if (sv->adm->submit(r->id, r->footprint) != FB_OK) {
    sv->err = "request needs more device memory than the whole budget";
    return FB_ERR_OOM;
  }
  r->t_submit = sv->now();
  *ticket = r->id;
  ++sv->st.submitted;
  sv->reqs[r->id] = std::move(r);
  sv->cv_work.notify_all();
  return FB_OK;
}

extern "C" int fb_server_wait(fb_server* sv, uint64_t ticket, double timeout_s, fb_serve_result* out) {
  if (!sv || !out) return FB_ERR_INVALID;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return FB_ERR_INVALID;
  ServeReq* r = it->second.get();
  auto done = [&] { return r->state >= 2; };
  // counted while waiting, so that fb_server_destroy frees nothing under a waiter
  struct Waiting {
    fb_server* sv;
    explicit Waiting(fb_server* s) : sv(s) { ++sv->waiters; }
    ~Waiting() {
      if (--sv->waiters == 0 && sv->stop) sv->cv_done.notify_all();
    }
  } waiting(sv);
  if (timeout_s < 0) sv->cv_done.wait(lk, done);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return FB_ERR_TIMEOUT;
  memset(out, 0, sizeof *out);
  out->status = r->status;
  snprintf(out->error, sizeof out->error, "%s", r->err.c_str());
  out->n_passes = (uint32_t)r->passes.size();
  out->genome_size = r->genome_size;
  out->has_estimate = r->has_est ? 1u : 0u;
  out->estimate = r->est;
  for (size_t j = 0; j < r->passes.size() && j < 4; ++j) {
    out->n_solid[j] = r->passes[j].n;
    out->route[j] = r->passes[j].route;
  }
  out->bases = r->bases;
  out->footprint = r->footprint;
  out->t_submit = r->t_submit;
  out->t_admit = r->t_admit;
  out->t_done = r->t_done;
  return r->status;
}

static ServeReq* serve_done_req(fb_server* sv, uint64_t ticket, uint32_t pass) {
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end() || it->second->state != 2 || pass >= it->second->passes.size()) return nullptr;
  return it->second.get();
}

extern "C" int fb_server_download(fb_server* sv, uint64_t ticket, uint32_t pass, uint64_t* keys, uint16_t* cov,
                                  uint8_t* mask) {
  if (!sv) return FB_ERR_INVALID;
  std::lock_guard<std::mutex> g(sv->mu);
  ServeReq* r = serve_done_req(sv, ticket, pass);
  if (!r) return FB_ERR_INVALID;
  ServePass& p = r->passes[pass];
  if (keys) memcpy(keys, p.keys(), (size_t)p.n * p.nref * 8);
  if (cov) memcpy(cov, p.cov(), (size_t)p.n * 2);
  if (mask) memcpy(mask, p.mask(), (size_t)p.n);
  return FB_OK;
}

extern "C" int fb_server_sketches(fb_server* sv, uint64_t ticket, uint32_t pass, void* cms, uint64_t cms_bytes,
                                  uint8_t* hll) {
  if (!sv) return FB_ERR_INVALID;
  std::lock_guard<std::mutex> g(sv->mu);
  ServeReq* r = serve_done_req(sv, ticket, pass);
  if (!r) return FB_ERR_INVALID;
  const ServePass& p = r->passes[pass];
  if (p.cms.empty()) return FB_ERR_INVALID;  // the server was created without keep_sketches
  if (cms) {
    if (cms_bytes < p.cms.size()) return FB_ERR_INVALID;
    memcpy(cms, p.cms.data(), p.cms.size());
  }
  if (hll && !p.hll.empty()) memcpy(hll, p.hll.data(), p.hll.size());
  return FB_OK;
}

extern "C" int fb_server_release(fb_server* sv, uint64_t ticket) {
  if (!sv) return FB_ERR_INVALID;
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end() || it->second->state < 2) return FB_ERR_INVALID;
  sv->reqs.erase(it);
  return FB_OK;
}

extern "C" int fb_server_stats(fb_server* sv, fb_serve_stats* out) {
  if (!sv || !out) return FB_ERR_INVALID;
  std::lock_guard<std::mutex> g(sv->mu);
  *out = sv->st;
  out->queued = sv->adm->queued();
  out->inflight = sv->adm->inflight();
  out->inflight_bytes = sv->adm->used();
  out->device_bytes = sv->ctx->held;
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Autotuner entry points

extern "C" int fb_batch_autotune(fb_batch* b, uint32_t k, int32_t target, fb_tune_result* out) {
  if (!b || !out) return FB_ERR_INVALID;
  fb_ctx* c = b->ctx;
  memset(out, 0, sizeof *out);
  const size_t n = b->iso.size();
  // one candidate: a warm-up run (allocations, graph capture) then a timed
  // run; a warm-up already twice as slow as the best ends the candidate
  std::vector<uint64_t> ref_counts;
  auto trial = [&](const TuneChoice& t, float best, float* ms) -> int {
    b->tuned[k] = t;
    if (int r = fb_batch_run(b, k, target)) return r;
    float warm = b->ms[3];
    *ms = warm;
    if (best <= 0.0f || warm < 2.0f * best) {
      if (int r = fb_batch_run(b, k, target)) return r;
      *ms = b->ms[3];
    }
    ++out->trials;
    std::vector<uint64_t> counts(n);
    for (size_t i = 0; i < n; ++i) counts[i] = b->iso[i]->n_out;
    if (ref_counts.empty()) ref_counts = counts;
    else if (counts != ref_counts)
      return set_err(c, FB_ERR_INVALID, "autotune: launch shapes gave different solid sets at k=%u", k);
    return FB_OK;
  };
  TuneChoice best;
  float best_ms = 0.0f;
  if (int r = trial(best, 0.0f, &best_ms)) return r;
  out->default_ms = best_ms;
  auto sweep = [&](uint32_t TuneChoice::*field, std::initializer_list<uint32_t> values) -> int {
    for (uint32_t v : values) {
      TuneChoice t = best;
      t.*field = v;
      float ms = 0.0f;
      if (int r = trial(t, best_ms, &ms)) return r;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    return FB_OK;
  };
  const fb_build_params& p = b->prm;
  if (!p.placement)
    if (int r = sweep(&TuneChoice::placement, {1u, 2u})) return r;
  if (!p.split)
    if (int r = sweep(&TuneChoice::split, {1u, 2u, 3u, 4u, 6u})) return r;
  if (!p.chunks1 && best.placement != 1)
    if (int r = sweep(&TuneChoice::chunks1, {2u, 4u, 8u, 16u, 32u, 64u})) return r;
  if (!p.chunks2 && best.placement != 1)
    if (int r = sweep(&TuneChoice::chunks2, {1u, 2u, 4u, 8u, 16u})) return r;
  best.ms = best_ms;
  b->tuned[k] = best;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[batch_tune_key(b, k)] = best;
  }
  out->placement = best.placement;
  out->split = best.split;
  out->chunks1 = best.chunks1;
  out->chunks2 = best.chunks2;
  out->best_ms = best_ms;
  // leave the batch's results those of the choice
  return fb_batch_run(b, k, target);
}

extern "C" int fb_autotune_save(const char* path) {
  if (!path) return FB_ERR_INVALID;
  FILE* f = fopen(path, "w");
  if (!f) return FB_ERR_INVALID;
  std::lock_guard<std::mutex> g(g_tune_mu);
  fprintf(f, "# kind limbs bits log2w log2_bases log2_n placement split chunks1 chunks2 ms\n");
  for (auto& kv : g_tune)
    fprintf(f, "%u %u %u %u %u %u %u %u %u %u %.4f\n", kv.first.kind, kv.first.limbs, kv.first.bits, kv.first.log2w,
            kv.first.log2_bases, kv.first.log2_n, kv.second.placement, kv.second.split, kv.second.chunks1,
            kv.second.chunks2, (double)kv.second.ms);
  return fclose(f) == 0 ? FB_OK : FB_ERR_INVALID;
}

extern "C" int fb_autotune_load(const char* path) {
  if (!path) return FB_ERR_INVALID;
  FILE* f = fopen(path, "r");
  if (!f) return FB_ERR_INVALID;
  char line[256];
  std::map<TuneKey, TuneChoice> got;
  int rc = FB_OK;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    TuneKey key{};
    TuneChoice t;
    double ms = 0.0;
    if (sscanf(line, "%u %u %u %u %u %u %u %u %u %u %lf", &key.kind, &key.limbs, &key.bits, &key.log2w,
               &key.log2_bases, &key.log2_n, &t.placement, &t.split, &t.chunks1, &t.chunks2, &ms) != 11 ||
        t.placement > 2 || t.split > 64 || t.chunks1 > 64 || t.chunks2 > 64) {
      rc = FB_ERR_INVALID;
      break;
    }
    t.ms = (float)ms;
    got[key] = t;
  }
  fclose(f);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (auto& kv : got) g_tune[kv.first] = kv.second;
  return FB_OK;
}

extern "C" void fb_autotune_clear(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune.clear();
}


// ---------------------------------------------------------------------------
// Unitig compaction (build.rs:330-597) of a sorted solid set: walks and
// overlap edges (build.rs:552-594) on the device (fb_compact.hip).

struct fb_graph {
  fb_ctx* ctx = nullptr;
  uint32_t k = 0;
  uint64_t n_lin = 0;
  std::vector<uint8_t> seq;
  std::vector<uint64_t> off;
  std::vector<uint32_t> len;
  std::vector<float> cov;
  std::vector<uint8_t> circ;
  std::vector<uint32_t> e_src, e_dst;
  std::vector<uint8_t> e_ov;
  std::vector<uint16_t> e_sup;
  // per node: flags (bit 0 is_deleted, bit 1 is_circular), edges_fwd.len(), edges_rev.len()
  std::vector<uint8_t> flags;
  std::vector<uint32_t> nf, nr;
  float ms[3] = {0, 0, 0};  // device walks, downloads, device edges
};

// The overlap edges of the U unitigs of a (sequences, offsets, lengths and
// circular flags on the device), into g's edge lists; tot: U * 2 edge counts
// per end (node u: forward end 2u, reverse end 2u + 1) on the host.
static int compact_edges(fb_ctx* c, const CgArgs& a, uint64_t U, uint32_t k, fb_graph* g,
                         std::vector<uint32_t>& tot_h) {
  hipStream_t s = c->stream;
  tot_h.assign(2 * U, 0);
  if (!U) return FB_OK;
  if (U >= (1ull << 30)) return set_err(c, FB_ERR_INVALID, "edges take fewer than 2^30 unitigs");
  const uint64_t N2 = 2 * U;
  DBuf hk, hv, sk[2], sv[2], cnt, tot, base, pos, tmp;
  int r = 0;
  if (!r) r = hk.get(c, N2 * 8);
  if (!r) r = hv.get(c, N2 * 4);
  for (int p = 0; p < 2 && !r; ++p) {
    r = sk[p].get(c, N2 * 8);
    if (!r) r = sv[p].get(c, N2 * 4);
  }
  if (!r) r = cnt.get(c, N2 * 4);
  if (!r) r = tot.get(c, N2 * 4);
  if (!r) r = base.get(c, N2 * 4);
  if (!r) r = pos.get(c, N2 * 4);
  if (r) return r;
  size_t tb = 0, t1 = 0;
  HIP_TRY(c, cg_sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, N2, nullptr, &tb, s));
  HIP_TRY(c, cg_exclusive_sum_u32(nullptr, nullptr, N2, nullptr, &t1, s));
  tb = std::max(tb, t1);
  if (int e = tmp.get(c, tb + 256)) return e;
  CgEdgeArgs e;
  e.seq = a.seq;
  e.off = a.u_off;
  e.len = a.u_len;
  e.circ = a.u_circ;
  e.U = (uint32_t)U;
  e.k = k;
  e.hkey = hk.as<uint64_t>();
  e.hval = hv.as<uint32_t>();
  e.cnt = cnt.as<uint32_t>();
  e.pos = pos.as<uint32_t>();
  const int npass = k > 1 ? 2 : 1;  // overlaps k and k - 1 (none of 0)
  for (int p = 0; p < npass; ++p) {
    e.ov = p == 0 ? k : k - 1;
    e.merge = p == 1;
    HIP_TRY(c, launch_cg_prefix(e, s));
    size_t t = tb;
    HIP_TRY(c, cg_sort_pairs_u64(hk.as<uint64_t>(), sk[p].as<uint64_t>(), hv.as<uint32_t>(), sv[p].as<uint32_t>(), N2,
                                 tmp.p, &t, s));
    e.skey = sk[p].as<uint64_t>();
    e.sval = sv[p].as<uint32_t>();
    HIP_TRY(c, launch_cg_end_count(e, p == 0, s));
  }
  HIP_TRY(c, launch_cg_end_total(cnt.as<uint32_t>(), tot.as<uint32_t>(), (uint32_t)N2, s));
  {
    size_t t = tb;
    HIP_TRY(c, cg_exclusive_sum_u32(tot.as<uint32_t>(), base.as<uint32_t>(), N2, tmp.p, &t, s));
  }
  uint32_t last_base = 0;
  HIP_TRY(c, hipMemcpyAsync(&last_base, base.as<uint32_t>() + N2 - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(tot_h.data(), tot.p, N2 * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const uint64_t E = (uint64_t)last_base + tot_h[N2 - 1];
  g->e_src.resize(E);
  g->e_dst.resize(E);
  g->e_ov.resize(E);
  g->e_sup.resize(E);
  if (!E) return FB_OK;
  DBuf es, ed, eo, eu;
  if (!r) r = es.get(c, E * 4);
  if (!r) r = ed.get(c, E * 4);
  if (!r) r = eo.get(c, E);
  if (!r) r = eu.get(c, E * 2);
  if (r) return r;
  e.base = base.as<uint32_t>();
  e.e_src = es.as<uint32_t>();
  e.e_dst = ed.as<uint32_t>();
  e.e_ov = eo.as<uint8_t>();
  e.e_sup = eu.as<uint16_t>();
  for (int p = 0; p < npass; ++p) {
    e.ov = p == 0 ? k : k - 1;
    e.merge = p == 1;
    e.skey = sk[p].as<uint64_t>();
    e.sval = sv[p].as<uint32_t>();
    HIP_TRY(c, launch_cg_end_write(e, p == 0, s));
  }
  HIP_TRY(c, hipMemcpyAsync(g->e_src.data(), es.p, E * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(g->e_dst.data(), ed.p, E * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(g->e_ov.data(), eo.p, E, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(g->e_sup.data(), eu.p, E * 2, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return FB_OK;
}

static int compact_device(fb_ctx* c, const uint64_t* keys, const uint16_t* cov, const uint8_t* mask, uint64_t n,
                          uint32_t k, fb_graph* g) {
  hipStream_t s = c->stream;
  g->ctx = c;
  g->k = k;
  if (n == 0) return FB_OK;
  // start edges (up to 4 per directed k-mer) and cycle labels share 31-bit ids
  if (n >= (1ull << 28)) return set_err(c, FB_ERR_INVALID, "compaction takes fewer than 2^28 solid k-mers");
  const uint64_t N = 2 * n;
  CgArgs a;
  a.keys = keys;
  a.cov = cov;
  a.mask = mask;
  a.n = (uint32_t)n;
  a.k = k;
  a.nref = ref_words(k);
  DBuf kind, fbase, lbase, obase, nstart, sbase, succ, bnext, jb[4], jflag, ccov, clen, cyc_len, cid, coff, tmp;
  int r = 0;
  if (!r) r = kind.get(c, N);
  if (!r) r = fbase.get(c, N);
  if (!r) r = lbase.get(c, N);
  if (!r) r = obase.get(c, N);
  if (!r) r = nstart.get(c, N * 4);
  if (!r) r = sbase.get(c, N * 4);
  if (!r) r = succ.get(c, N * 4);
  if (!r) r = bnext.get(c, N * 16);
  for (int j = 0; j < 4 && !r; ++j) r = jb[j].get(c, N * 16);
  if (!r) r = jflag.get(c, 128 * 4);
  if (!r) r = ccov.get(c, N * 8);
  if (!r) r = clen.get(c, N * 4);
  if (!r) r = cyc_len.get(c, N * 4);
  if (!r) r = cid.get(c, N * 4);
  if (!r) r = coff.get(c, N * 8);
  if (r) return r;
  a.kind = kind.as<uint8_t>();
  a.fbase = fbase.as<uint8_t>();
  a.lbase = lbase.as<uint8_t>();
  a.obase = obase.as<uint8_t>();
  a.nstart = nstart.as<uint32_t>();
  a.sbase = sbase.as<uint32_t>();
  a.succ = succ.as<uint32_t>();
  a.bnext = bnext.as<uint32_t>();
  for (int j = 0; j < 4; ++j) a.js[j] = jb[j].as<uint4>();
  a.jflag = jflag.as<uint32_t>();
  a.ccov = ccov.as<unsigned long long>();
  a.clen = clen.as<uint32_t>();
  a.cyc_len = cyc_len.as<uint32_t>();
  a.cid = cid.as<uint32_t>();
  a.coff = coff.as<uint64_t>();
  HIP_TRY(c, hipEventRecord(c->ev[0], s));
  HIP_TRY(c, launch_cg_links(a, s));
  // scan scratch: the largest of the scans below
  size_t tb = 0, t1 = 0;
  HIP_TRY(c, cg_exclusive_sum_u32(nullptr, nullptr, N, nullptr, &t1, s));
  tb = std::max(tb, t1);
  HIP_TRY(c, cg_exclusive_sum_u64(nullptr, nullptr, N, nullptr, &t1, s));
  tb = std::max(tb, t1);
  if (int e = tmp.get(c, tb + 256)) return e;
  size_t tbs = tb;
  HIP_TRY(c, cg_exclusive_sum_u32(a.nstart, a.sbase, N, tmp.p, &tbs, s));
  uint32_t tail[2] = {0, 0};
  HIP_TRY(c, hipMemcpyAsync(&tail[0], a.sbase + N - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&tail[1], a.nstart + N - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  a.n_se = (uint64_t)tail[0] + tail[1];
  const uint64_t nse = std::max<uint64_t>(a.n_se, 1);
  DBuf csum, last, walked, slen, uid, soff;
  if (!r) r = csum.get(c, nse * 8);
  if (!r) r = last.get(c, nse * 8);
  if (!r) r = walked.get(c, nse * 4);
  if (!r) r = slen.get(c, nse * 4);
  if (!r) r = uid.get(c, nse * 4);
  if (!r) r = soff.get(c, nse * 8);
  if (r) return r;
  a.csum = csum.as<unsigned long long>();
  a.last = last.as<unsigned long long>();
  a.walked = walked.as<uint32_t>();
  a.slen = slen.as<uint32_t>();
  a.uid = uid.as<uint32_t>();
  a.soff = soff.as<uint64_t>();
  HIP_TRY(c, launch_cg_pred(a, s));
  uint32_t sp = 0, sc = 2;
  HIP_TRY(c, launch_cg_jump(a, 0u, 1u, &sp, s));
  HIP_TRY(c, launch_fill_u32(a.clen, N, 0u, s));
  HIP_TRY(c, launch_fill_u32(reinterpret_cast<uint32_t*>(a.ccov), 2 * N, 0u, s));
  HIP_TRY(c, launch_cg_accumulate(a, sp, s));
  HIP_TRY(c, launch_cg_cycle_ranks(a, sp, &sc, s));
  // unitig ids and offsets
  uint32_t lin[4] = {0, 0, 0, 0};
  uint64_t lb[2] = {0, 0};
  if (a.n_se) {
    size_t t = tb;
    HIP_TRY(c, cg_exclusive_sum_u32(a.walked, a.uid, a.n_se, tmp.p, &t, s));
    t = tb;
    HIP_TRY(c, cg_exclusive_sum_u64(a.slen, a.soff, a.n_se, tmp.p, &t, s));
    HIP_TRY(c, hipMemcpyAsync(&lin[0], a.uid + a.n_se - 1, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&lin[1], a.walked + a.n_se - 1, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&lb[0], a.soff + a.n_se - 1, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&lin[2], a.slen + a.n_se - 1, 4, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(c, launch_cg_cycle_flags(a, a.nstart, s));  // nstart is free again: cycle flags
  {
    size_t t = tb;
    HIP_TRY(c, cg_exclusive_sum_u32(a.nstart, a.cid, N, tmp.p, &t, s));
    t = tb;
    HIP_TRY(c, cg_exclusive_sum_u64(a.cyc_len, a.coff, N, tmp.p, &t, s));
  }
  uint32_t ct[3] = {0, 0, 0};
  uint64_t cb = 0;
  HIP_TRY(c, hipMemcpyAsync(&ct[0], a.cid + N - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&ct[1], a.nstart + N - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&ct[2], a.cyc_len + N - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&cb, a.coff + N - 1, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const uint64_t n_lin = (uint64_t)lin[0] + lin[1];
  const uint64_t lin_bases = lb[0] + lin[2];
  const uint64_t n_circ = (uint64_t)ct[0] + ct[1];
  const uint64_t circ_bases = cb + ct[2];
  const uint64_t U = n_lin + n_circ, B = lin_bases + circ_bases;
  a.n_lin = (uint32_t)n_lin;
  a.circ_base = lin_bases;
  DBuf seq, uoff, ulen, ucov, ucirc, flip;
  if (!r) r = seq.get(c, std::max<uint64_t>(B, 1));
  if (!r) r = uoff.get(c, std::max<uint64_t>(U, 1) * 8);
  if (!r) r = ulen.get(c, std::max<uint64_t>(U, 1) * 4);
  if (!r) r = ucov.get(c, std::max<uint64_t>(U, 1) * 4);
  if (!r) r = ucirc.get(c, std::max<uint64_t>(U, 1));
  if (!r) r = flip.get(c, std::max<uint64_t>(n_lin, 1));
  if (r) return r;
  a.seq = seq.as<uint8_t>();
  a.u_off = uoff.as<uint64_t>();
  a.u_len = ulen.as<uint32_t>();
  a.u_cov = ucov.as<float>();
  a.u_circ = ucirc.as<uint8_t>();
  HIP_TRY(c, launch_cg_table(a, s));
  HIP_TRY(c, launch_cg_write(a, sp, sc, s));
  HIP_TRY(c, launch_cg_flip(a.seq, a.u_off, a.u_len, a.n_lin, flip.as<uint8_t>(), lin_bases, s));
  HIP_TRY(c, hipEventRecord(c->ev[1], s));
  std::vector<uint32_t> tot;
  if (int e = compact_edges(c, a, U, k, g, tot)) return e;
  HIP_TRY(c, hipEventRecord(c->ev[3], s));
  g->n_lin = n_lin;
  g->seq.resize(B);
  g->off.resize(U);
  g->len.resize(U);
  g->cov.resize(U);
  g->circ.resize(U);
  if (B) HIP_TRY(c, hipMemcpyAsync(g->seq.data(), a.seq, B, hipMemcpyDeviceToHost, s));
  if (U) {
    HIP_TRY(c, hipMemcpyAsync(g->off.data(), a.u_off, U * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(g->len.data(), a.u_len, U * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(g->cov.data(), a.u_cov, U * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(g->circ.data(), a.u_circ, U, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(c, hipEventRecord(c->ev[2], s));
  HIP_TRY(c, hipEventSynchronize(c->ev[2]));
  (void)hipEventElapsedTime(&g->ms[0], c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&g->ms[2], c->ev[1], c->ev[3]);
  (void)hipEventElapsedTime(&g->ms[1], c->ev[3], c->ev[2]);
  g->flags.assign(U, 0);
  g->nf.assign(U, 0);
  g->nr.assign(U, 0);
  for (uint64_t u = 0; u < U; ++u) {
    g->flags[u] = g->circ[u] ? 2 : 0;
    g->nf[u] = tot[2 * u];
    g->nr[u] = tot[2 * u + 1];
  }
  return FB_OK;
}

extern "C" int fb_compact_solid(const fb_solid* so, fb_graph** out) {
  if (!so || !out) return FB_ERR_INVALID;
  *out = nullptr;
  fb_ctx* c = so->ctx;
  if (!so->sorted) return set_err(c, FB_ERR_INVALID, "compaction needs the solid set sorted (params sort = 1)");
  if (int r = ctx_current(c)) return r;
  auto g = std::make_unique<fb_graph>();
  if (int r = compact_device(c, so->keys.as<uint64_t>(), so->cov.as<uint16_t>(), so->mask.as<uint8_t>(), so->n,
                             so->k, g.get()))
    return r;
  *out = g.release();
  return FB_OK;
}

extern "C" int fb_compact_arrays(fb_ctx* c, const uint64_t* keys, const uint16_t* cov, const uint8_t* mask,
                                 uint64_t n, uint32_t k, fb_graph** out) {
  if (!c || !out || (n && (!keys || !cov || !mask))) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  const uint32_t nref = ref_words(k);
  // sorted, distinct and within 2k bits: the lookups assume it
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t* x = keys + i * nref;
    for (uint32_t w = 0; w < nref; ++w) {
      const uint32_t bit = 64u * (nref - 1 - w);  // lowest bit of word w
      const uint64_t lim = bit >= 2 * k ? 0ull : (2 * k - bit >= 64 ? ~0ull : (1ull << (2 * k - bit)) - 1ull);
      if (x[w] & ~lim) return set_err(c, FB_ERR_INVALID, "key %llu has bits beyond 2k", (unsigned long long)i);
    }
    if (i && !std::lexicographical_compare(x - nref, x, x, x + nref))
      return set_err(c, FB_ERR_INVALID, "keys must be strictly ascending (key %llu)", (unsigned long long)i);
  }
  if (int r = ctx_current(c)) return r;
  DBuf dk, dc, dm;
  if (int r = to_device(c, dk, keys, n * nref * 8)) return r;
  if (int r = to_device(c, dc, cov, n * 2)) return r;
  if (int r = to_device(c, dm, mask, n)) return r;
  auto g = std::make_unique<fb_graph>();
  if (int r = compact_device(c, dk.as<uint64_t>(), dc.as<uint16_t>(), dm.as<uint8_t>(), n, k, g.get())) return r;
  *out = g.release();
  return FB_OK;
}

extern "C" void fb_graph_destroy(fb_graph* g) { delete g; }
extern "C" uint64_t fb_graph_unitig_count(const fb_graph* g) { return g ? g->len.size() : 0; }
extern "C" uint64_t fb_graph_base_count(const fb_graph* g) { return g ? g->seq.size() : 0; }
extern "C" uint64_t fb_graph_edge_count(const fb_graph* g) { return g ? g->e_src.size() : 0; }
extern "C" int fb_graph_unitigs(const fb_graph* g, uint8_t* seq, uint64_t* offsets, uint32_t* lengths, float* cov,
                                uint8_t* circular) {
  if (!g) return FB_ERR_INVALID;
  if (seq && !g->seq.empty()) memcpy(seq, g->seq.data(), g->seq.size());
  if (offsets && !g->off.empty()) memcpy(offsets, g->off.data(), g->off.size() * 8);
  if (lengths && !g->len.empty()) memcpy(lengths, g->len.data(), g->len.size() * 4);
  if (cov && !g->cov.empty()) memcpy(cov, g->cov.data(), g->cov.size() * 4);
  if (circular && !g->circ.empty()) memcpy(circular, g->circ.data(), g->circ.size());
  return FB_OK;
}
extern "C" int fb_graph_edges(const fb_graph* g, uint32_t* src, uint32_t* dst, uint8_t* ov, uint16_t* support) {
  if (!g) return FB_ERR_INVALID;
  const size_t n = g->e_src.size();
  if (!n) return FB_OK;
  if (src) memcpy(src, g->e_src.data(), n * 4);
  if (dst) memcpy(dst, g->e_dst.data(), n * 4);
  if (ov) memcpy(ov, g->e_ov.data(), n);
  if (support) memcpy(support, g->e_sup.data(), n * 2);
  return FB_OK;
}
extern "C" int fb_graph_timings(const fb_graph* g, float out[3]) {
  if (!g || !out) return FB_ERR_INVALID;
  memcpy(out, g->ms, sizeof g->ms);
  return FB_OK;
}


// ---------------------------------------------------------------------------
// filter_active_reads (local_multik.rs:259-333) on the device (fb_filter.hip)
// and the multi-k upgrade driver (upgrade_multik_ext, local_multik.rs:38-147).

extern "C" int fb_graph_view_of(const fb_graph* g, fb_graph_view* v) {
  if (!g || !v) return FB_ERR_INVALID;
  v->k = g->k;
  v->n_nodes = g->len.size();
  v->seq = g->seq.data();
  v->offsets = g->off.data();
  v->lengths = g->len.data();
  v->cov = g->cov.data();
  v->flags = g->flags.data();
  v->n_edges_fwd = g->nf.data();
  v->n_edges_rev = g->nr.data();
  return FB_OK;
}

static uint8_t view_flags(const fb_graph_view* g, uint64_t i) { return g->flags ? g->flags[i] : 0; }
static uint32_t view_nf(const fb_graph_view* g, uint64_t i) { return g->n_edges_fwd ? g->n_edges_fwd[i] : 0; }
static uint32_t view_nr(const fb_graph_view* g, uint64_t i) { return g->n_edges_rev ? g->n_edges_rev[i] : 0; }

static int check_view(fb_ctx* c, const fb_graph_view* g) {
  if (!g) return set_err(c, FB_ERR_INVALID, "null graph");
  if (g->k < 1) return set_err(c, FB_ERR_INVALID, "graph k must be >= 1");
  if (g->n_nodes && (!g->seq || !g->offsets || !g->lengths)) return set_err(c, FB_ERR_INVALID, "graph: null arrays");
  return FB_OK;
}

// Kept reads, filtered list: kept[0] merged reads, kept[1] pairs.
static int filter_active(fb_ctx* c, const fb_bank* b, const fb_graph_view* g, fb_bank** out, uint8_t* keep_m,
                         uint8_t* keep_p, uint64_t kept[2]) {
  if (int r = check_view(c, g)) return r;
  if (b->n_reads >= (1ull << 31)) return set_err(c, FB_ERR_INVALID, "filter_active_reads takes < 2^31 reads");
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  const uint32_t k = std::min<uint32_t>(g->k, 31);  // local_multik.rs:261
  const uint64_t win = 350;                          // boundary_window

  // windows of the boundary set, as ranges over the kept unitig sequences
  std::vector<uint8_t> useq;
  std::vector<uint64_t> rstart, rpre(1, 0);
  auto range = [&](uint64_t at, uint64_t n) {
    if (!n) return;
    rstart.push_back(at);
    rpre.push_back(rpre.back() + n);
  };
  for (uint64_t i = 0; i < g->n_nodes; ++i) {
    const uint8_t fl = view_flags(g, i);
    const uint64_t L = g->lengths[i];
    if ((fl & 1) || L < k) continue;
    const uint64_t base = useq.size();
    useq.insert(useq.end(), g->seq + g->offsets[i], g->seq + g->offsets[i] + L);
    if (L < 2000 || (fl & 2) || view_nf(g, i) > 1 || view_nr(g, i) > 1) {
      range(base, L - k + 1);
    } else {
      range(base, std::min<uint64_t>(win, L - k + 1));
      const uint64_t rs = std::min<uint64_t>(L > win ? L - win : 0, L - k);
      range(base + rs, L - k - rs + 1);
    }
  }
  if (rstart.size() >= (1ull << 32)) return set_err(c, FB_ERR_INVALID, "too many unitigs");
  FaArgs a;
  a.k = k;
  a.n_ranges = (uint32_t)rstart.size();
  a.n_windows = rpre.back();
  uint32_t log2cap = 10;
  while ((1ull << log2cap) < 2 * a.n_windows) ++log2cap;
  a.log2cap = log2cap;
  const uint64_t nm = b->ar_reads[0], np = std::min(b->ar_reads[1], b->ar_reads[2]);
  a.reads = b->view();
  for (int i = 0; i < 3; ++i) a.first[i] = b->ar_first[i];
  a.n_merged = nm;
  a.n_pairs = np;
  const uint64_t n = nm + np;

  DBuf dseq, drs, drp, dset, dflag, dkeep, dpos, dsrc, dolen, dopos, dtmp, dhist;
  if (int r = to_device(c, dseq, useq.data(), useq.size())) return r;
  if (int r = to_device(c, drs, rstart.data(), rstart.size() * 8)) return r;
  if (int r = to_device(c, drp, rpre.data(), rpre.size() * 8)) return r;
  if (int r = dset.get(c, (1ull << log2cap) * 8)) return r;
  if (int r = dflag.get(c, 256)) return r;
  if (int r = dkeep.get(c, (n + 1) * 4)) return r;
  if (int r = dpos.get(c, (n + 1) * 4)) return r;
  a.useq = dseq.as<uint8_t>();
  a.rstart = drs.as<uint64_t>();
  a.rpre = drp.as<uint64_t>();
  a.set = dset.as<uint64_t>();
  a.set_flag = dflag.as<uint32_t>();
  a.keep = dkeep.as<uint32_t>();
  a.pos = dpos.as<uint32_t>();
  HIP_TRY(c, launch_fa_set_clear(a, s));
  HIP_TRY(c, launch_fa_set_insert(a, s));
  HIP_TRY(c, launch_fa_keep(a, s));
  size_t tb = 0, tb2 = 0;
  HIP_TRY(c, cg_exclusive_sum_u32(nullptr, nullptr, n + 1, nullptr, &tb, s));
  HIP_TRY(c, cg_exclusive_sum_u64(nullptr, nullptr, n + 1 + np, nullptr, &tb2, s));
  if (int r = dtmp.get(c, std::max<size_t>(std::max(tb, tb2), 16))) return r;
  tb = dtmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u32(a.keep, a.pos, n + 1, dtmp.p, &tb, s));
  uint32_t cnt[2] = {0, 0};
  HIP_TRY(c, hipMemcpyAsync(&cnt[0], a.pos + nm, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&cnt[1], a.pos + n, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const uint64_t nkm = cnt[0], nkp = cnt[1] - cnt[0], n_out = nkm + 2 * nkp;
  kept[0] = nkm;
  kept[1] = nkp;

  if (int r = dsrc.get(c, std::max<uint64_t>(n_out, 1) * 4)) return r;
  if (int r = dolen.get(c, (n_out + 1) * 4)) return r;
  if (int r = dopos.get(c, (n_out + 1) * 8)) return r;
  a.src = dsrc.as<uint32_t>();
  a.olen = dolen.as<uint32_t>();
  a.opos = dopos.as<uint64_t>();
  HIP_TRY(c, launch_fill_u32(a.olen, n_out + 1, 0, s));
  HIP_TRY(c, launch_fa_list(a, s));
  tb2 = dtmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u64(a.olen, a.opos, n_out + 1, dtmp.p, &tb2, s));
  uint64_t bnd[3] = {0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(&bnd[0], a.opos + nkm, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&bnd[1], a.opos + nkm + nkp, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&bnd[2], a.opos + n_out, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const uint64_t B[3] = {bnd[0], bnd[1] - bnd[0], bnd[2] - bnd[1]};
  a.rfirst[0] = 0;
  a.rfirst[1] = nkm;
  a.rfirst[2] = nkm + nkp;
  a.rfirst[3] = n_out;
  a.word0[0] = 0;
  for (int i = 0; i < 3; ++i) a.word0[i + 1] = a.word0[i] + (B[i] + 31) / 32;

  auto nb = std::make_unique<fb_bank>();
  nb->ctx = c;
  if (int r = nb->data.get(c, (a.word0[3] + kArenaPadWords) * 8)) return r;
  if (int r = nb->offs.get(c, std::max<uint64_t>(n_out, 1) * 8)) return r;
  if (int r = nb->lens.get(c, std::max<uint64_t>(n_out, 1) * 4)) return r;
  a.hist_n = b->max_len + 1;
  if (int r = dhist.get(c, (uint64_t)a.hist_n * 4)) return r;
  a.odata = nb->data.as<uint64_t>();
  a.ooffs = nb->offs.as<uint64_t>();
  a.olens = nb->lens.as<uint32_t>();
  a.hist = dhist.as<uint32_t>();
  HIP_TRY(c, launch_fill_u32(a.hist, a.hist_n, 0, s));
  HIP_TRY(c, launch_fa_offsets(a, s));
  HIP_TRY(c, launch_fa_gather(a, s));
  HIP_TRY(c, hipMemsetAsync(a.odata + a.word0[3], 0, kArenaPadWords * 8, s));
  std::vector<uint32_t> hist(a.hist_n), hk;
  HIP_TRY(c, hipMemcpyAsync(hist.data(), a.hist, (uint64_t)a.hist_n * 4, hipMemcpyDeviceToHost, s));
  if (keep_m || keep_p) {
    hk.resize(n);
    if (n) HIP_TRY(c, hipMemcpyAsync(hk.data(), a.keep, n * 4, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(c, hipStreamSynchronize(s));
  for (uint64_t i = 0; keep_m && i < nm; ++i) keep_m[i] = (uint8_t)hk[i];
  for (uint64_t i = 0; keep_p && i < np; ++i) keep_p[i] = (uint8_t)hk[nm + i];

  nb->n_reads = n_out;
  nb->len_hist.assign(a.hist_n, 0);
  nb->max_len = 0;
  for (uint32_t L = 0; L < a.hist_n; ++L)
    if (hist[L]) {
      nb->len_hist[L] = (uint64_t)hist[L] * L;
      nb->max_len = L;
    }
  const uint64_t nr_out[3] = {nkm, nkp, nkp};
  for (int i = 0; i < 3; ++i) {
    nb->ar_first[i] = a.rfirst[i];
    nb->ar_reads[i] = nr_out[i];
    nb->ar_word0[i] = a.word0[i];
    nb->ar_words[i] = a.word0[i + 1] - a.word0[i];
    nb->ar_bases[i] = B[i];
  }
  nb->total_bases = B[0] + B[1] + B[2];
  *out = nb.release();
  return FB_OK;
}

extern "C" int fb_filter_active_reads(fb_ctx* c, const fb_bank* bank, const fb_graph_view* g, fb_bank** out,
                                      uint8_t* keep_merged, uint8_t* keep_pairs) {
  if (!c || !bank || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (bank->ctx != c) return set_err(c, FB_ERR_INVALID, "the bank belongs to another context");
  uint64_t kept[2];
  return filter_active(c, bank, g, out, keep_merged, keep_pairs, kept);
}

// Appends an old circular unitig to the new graph as upgrade_multik_ext
// does (local_multik.rs:117-128): circular, its cov, one forward self edge
// of overlap k2 - 1 and support 10.
static void graph_push_circular(fb_graph* g, const uint8_t* seq, uint32_t len, float cov, uint32_t k2) {
  const uint32_t id = (uint32_t)g->len.size();
  g->off.push_back(g->seq.size());
  g->seq.insert(g->seq.end(), seq, seq + len);
  g->len.push_back(len);
  g->cov.push_back(cov);
  g->circ.push_back(1);
  g->flags.push_back(2);
  g->nf.push_back(1);
  g->nr.push_back(0);
  g->e_src.push_back(id << 1);
  g->e_dst.push_back(id << 1);
  g->e_ov.push_back((uint8_t)(k2 - 1));
  g->e_sup.push_back(10);
}

extern "C" int fb_upgrade_multik(fb_ctx* c, const fb_bank* bank, const fb_graph_view* prev, uint32_t k2,
                                 uint64_t genome_size, const fb_build_params* params, fb_multik_result* res,
                                 fb_graph** out_graph, fb_solid** out_solid, fb_bank** out_active) {
  if (!c || !bank || !res || !out_graph) return set_err(c, FB_ERR_INVALID, "null argument");
  *out_graph = nullptr;
  if (out_solid) *out_solid = nullptr;
  if (out_active) *out_active = nullptr;
  if (bank->ctx != c) return set_err(c, FB_ERR_INVALID, "the bank belongs to another context");
  if (int r = check_view(c, prev)) return r;
  if (k2 > 127) return set_err(c, FB_ERR_INVALID, "k2=%u above 127", k2);
  memset(res, 0, sizeof *res);
  // inspect_graph (diag.rs:112-133): active unitigs and edges
  for (uint64_t i = 0; i < prev->n_nodes; ++i)
    if (!(view_flags(prev, i) & 1)) {
      res->active_unitigs_before++;
      res->active_edges_before += view_nf(prev, i) + view_nr(prev, i);
    }
  res->active_unitigs_after = res->active_unitigs_before;
  res->active_edges_after = res->active_edges_before;
  if (k2 <= prev->k) {
    res->early_stop = FB_MULTIK_K2_LE_PREV_K;
    return FB_OK;
  }
  bool branching = false;  // local_multik.rs:60-68
  for (uint64_t i = 0; i < prev->n_nodes && !branching; ++i) {
    const uint8_t fl = view_flags(prev, i);
    if (!(fl & 1) && !(fl & 2) && (view_nf(prev, i) > 1 || view_nr(prev, i) > 1)) branching = true;
  }
  if (!branching) {
    res->early_stop = FB_MULTIK_NO_BRANCHING_REGIONS;
    return FB_OK;
  }
  // prior contigs: every active linear unitig (local_multik.rs:84-90)
  std::vector<uint8_t> contigs;
  std::vector<uint32_t> clens;
  for (uint64_t i = 0; i < prev->n_nodes; ++i)
    if (!(view_flags(prev, i) & 3)) {
      contigs.insert(contigs.end(), prev->seq + prev->offsets[i], prev->seq + prev->offsets[i] + prev->lengths[i]);
      clens.push_back(prev->lengths[i]);
    }
  res->prior_contigs = clens.size();
  fb_bank* active = nullptr;
  uint64_t kept[2] = {0, 0};
  if (int r = filter_active(c, bank, prev, &active, nullptr, nullptr, kept)) return r;
  std::unique_ptr<fb_bank> active_own(active);
  res->kept_merged = kept[0];
  res->kept_pairs = kept[1];
  // local_multik.rs:95-102: genome size of the original bank
  if (genome_size) {
    res->genome_size_used = std::max<uint64_t>(genome_size, 50000);
    res->genome_size_source = FB_GENOME_EXPLICIT_PASSED;
  } else {
    res->genome_size_used = std::max<uint64_t>(bank->total_bases / 60, 50000);
    res->genome_size_source = FB_GENOME_FALLBACK_READ_BASES_DIV_60;
  }
  fb_build_params p;
  if (params) p = *params;
  else fb_build_params_default(&p, k2);
  p.k = k2;
  p.target_solid_cov = 4;
  p.genome_size = res->genome_size_used;
  p.sort = 1;
  res->cms_width = p.cms_width ? next_pow2(p.cms_width) : fb_cms_width_for_genome_size(res->genome_size_used);
  res->cms_allocated_bytes = 4 * res->cms_width / 2;  // sketch.rs:171-174
  fb_solid* solid = nullptr;
  if (int r = fb_build_solid(c, active, contigs.data(), clens.data(), clens.size(), &p, &solid)) return r;
  std::unique_ptr<fb_solid> solid_own(solid);
  fb_graph* g = nullptr;
  if (int r = fb_compact_solid(solid, &g)) return r;
  std::unique_ptr<fb_graph> g_own(g);
  // local_multik.rs:107-129: keep the old circular unitigs the new graph lacks
  for (uint64_t i = 0; i < prev->n_nodes; ++i) {
    const uint8_t fl = view_flags(prev, i);
    if ((fl & 1) || !(fl & 2)) continue;
    const int64_t old_len = prev->lengths[i];
    bool exists = false;
    for (uint64_t j = 0; j < g->len.size() && !exists; ++j)
      exists = g->circ[j] && std::llabs((int64_t)g->len[j] - old_len) <= (int64_t)k2 + 10;
    if (!exists) {
      graph_push_circular(g, prev->seq + prev->offsets[i], prev->lengths[i], prev->cov ? prev->cov[i] : 0.0f, k2);
      res->transferred_circular++;
    }
  }
  res->executed = 1;
  res->active_unitigs_after = g->len.size();
  res->active_edges_after = g->e_src.size();
  *out_graph = g_own.release();
  if (out_solid) *out_solid = solid_own.release();
  if (out_active) *out_active = active_own.release();
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Ingest front end (trim_merge.rs:17-417): records to a device bank
// (fb_ingest.hip), and the FASTQ reader on two host threads.

extern "C" void fb_ingest_params_default(fb_ingest_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->subsample_factor = 1.0;
  p->min_len = 30;  // main.rs:215-224
  p->min_qual = 20;
  p->enable_trim = 1;
  p->enable_merge = 1;
}

// Uploads one mate file's records; returns the device view.
static int upload_records(fb_ctx* c, const fb_records_view* v, DBuf* bufs, DevRecords* d) {
  uint64_t sb = 0, qb = 0;
  for (uint64_t i = 0; i < v->n; ++i) {
    sb = std::max<uint64_t>(sb, v->seq_offsets[i] + v->seq_lengths[i]);
    qb = std::max<uint64_t>(qb, v->qual_offsets[i] + v->qual_lengths[i]);
  }
  if (int r = to_device(c, bufs[0], v->seq, sb)) return r;
  if (int r = to_device(c, bufs[1], v->seq_offsets, v->n * 8)) return r;
  if (int r = to_device(c, bufs[2], v->seq_lengths, v->n * 4)) return r;
  if (int r = to_device(c, bufs[3], v->qual, qb)) return r;
  if (int r = to_device(c, bufs[4], v->qual_offsets, v->n * 8)) return r;
  if (int r = to_device(c, bufs[5], v->qual_lengths, v->n * 4)) return r;
  d->seq = bufs[0].as<uint8_t>();
  d->seq_off = bufs[1].as<uint64_t>();
  d->seq_len = bufs[2].as<uint32_t>();
  d->qual = bufs[3].as<uint8_t>();
  d->qual_off = bufs[4].as<uint64_t>();
  d->qual_len = bufs[5].as<uint32_t>();
  return FB_OK;
}

static bool records_ok(const fb_records_view* v) {
  return v && (v->n == 0 || (v->seq_offsets && v->seq_lengths && v->qual_offsets && v->qual_lengths));
}

extern "C" int fb_ingest_records(fb_ctx* c, const fb_records_view* r1, const fb_records_view* r2,
                                 const fb_ingest_params* p, fb_bank** out, fb_ingest_stats* st) {
  if (!c || !p || !out || !records_ok(r1) || !records_ok(r2)) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (r1->n != r2->n) return set_err(c, FB_ERR_INVALID, "Mismatched paired-end FASTQ mate counts between R1 and R2");
  const uint64_t n = r1->n;
  if (n >= (1ull << 31)) return set_err(c, FB_ERR_INVALID, "ingest takes < 2^31 pairs per call");
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  fb_ingest_stats hs;
  memset(&hs, 0, sizeof hs);
  hs.input_read_pairs = n;
  for (uint64_t i = 0; i < n; ++i) hs.bases_before_trim += (uint64_t)r1->seq_lengths[i] + r2->seq_lengths[i];
  HIP_TRY(c, hipEventRecord(c->ev[0], s));

  IngArgs a;
  DBuf b1[6], b2[6];
  if (int r = upload_records(c, r1, b1, &a.r1)) return r;
  if (int r = upload_records(c, r2, b2, &a.r2)) return r;
  a.n = n;
  a.subsample = p->subsample_factor < 0.999 ? 1u : 0u;
  const double thr = p->subsample_factor * 18446744073709551616.0;  // (factor * u64::MAX as f64) as u64
  a.keep_threshold = thr <= 0.0 ? 0ull : (thr >= 18446744073709551616.0 ? ~0ull : (uint64_t)thr);
  a.seed = p->seed;
  a.enable_trim = p->enable_trim ? 1u : 0u;
  a.enable_merge = p->enable_merge ? 1u : 0u;
  a.min_len = p->min_len;
  a.min_qual = p->min_qual;
  DBuf dstat, dsum, dm, du, dpm, dpu, dt1, dt2, dov, dtmp, dsrc, dolen, dopos, dhist;
  if (int r = dstat.get(c, 256)) return r;
  if (int r = dsum.get(c, 64)) return r;
  for (DBuf* d : {&dm, &du, &dpm, &dpu})
    if (int r = d->get(c, (n + 1) * 4)) return r;
  for (DBuf* d : {&dt1, &dt2, &dov})
    if (int r = d->get(c, std::max<uint64_t>(n, 1) * 4)) return r;
  a.status = dstat.as<uint32_t>();
  a.stats = dsum.as<unsigned long long>();
  a.merged = dm.as<uint32_t>();
  a.unmerged = du.as<uint32_t>();
  a.pos_m = dpm.as<uint32_t>();
  a.pos_u = dpu.as<uint32_t>();
  a.t1 = dt1.as<uint32_t>();
  a.t2 = dt2.as<uint32_t>();
  a.ov = dov.as<uint32_t>();
  HIP_TRY(c, launch_fill_u32(a.status, 1, 0, s));
  HIP_TRY(c, launch_fill_u32(dsum.as<uint32_t>(), 10, 0, s));
  HIP_TRY(c, launch_ing_pairs(a, s));
  size_t tb = 0, tb2 = 0;
  HIP_TRY(c, cg_exclusive_sum_u32(nullptr, nullptr, n + 1, nullptr, &tb, s));
  HIP_TRY(c, cg_exclusive_sum_u64(nullptr, nullptr, 2 * n + 1, nullptr, &tb2, s));
  if (int r = dtmp.get(c, std::max<size_t>(std::max(tb, tb2), 16))) return r;
  tb = dtmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u32(a.merged, a.pos_m, n + 1, dtmp.p, &tb, s));
  tb = dtmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u32(a.unmerged, a.pos_u, n + 1, dtmp.p, &tb, s));
  uint32_t cnt[3] = {0, 0, 0};
  unsigned long long sums[5] = {0, 0, 0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(&cnt[0], a.pos_m + n, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&cnt[1], a.pos_u + n, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&cnt[2], a.status, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(sums, a.stats, sizeof sums, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (cnt[2]) return set_err(c, FB_ERR_INVALID, "a quality string is shorter than its adapter-clipped read");
  hs.retained_read_pairs = sums[0];
  hs.bases_after_trim = sums[1];
  hs.merged_pairs_count = sums[2];
  hs.merged_bases = sums[3];
  hs.unmerged_pairs_count = sums[4];
  if (st) *st = hs;
  const uint64_t nm = cnt[0], nu = cnt[1], n_out = nm + 2 * nu;
  if (n_out == 0)
    return set_err(c, FB_ERR_EMPTY, "No valid reads survived trimming and quality filtering (check input FASTQ files)");

  if (int r = dsrc.get(c, n_out * 4)) return r;
  if (int r = dolen.get(c, (n_out + 1) * 4)) return r;
  if (int r = dopos.get(c, (n_out + 1) * 8)) return r;
  a.src = dsrc.as<uint32_t>();
  a.olen = dolen.as<uint32_t>();
  a.opos = dopos.as<uint64_t>();
  HIP_TRY(c, launch_fill_u32(a.olen, n_out + 1, 0, s));
  HIP_TRY(c, launch_ing_list(a, s));
  tb2 = dtmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u64(a.olen, a.opos, n_out + 1, dtmp.p, &tb2, s));
  uint64_t bnd[3] = {0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(&bnd[0], a.opos + nm, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&bnd[1], a.opos + nm + nu, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&bnd[2], a.opos + n_out, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const uint64_t B[3] = {bnd[0], bnd[1] - bnd[0], bnd[2] - bnd[1]};
  a.rfirst[0] = 0;
  a.rfirst[1] = nm;
  a.rfirst[2] = nm + nu;
  a.rfirst[3] = n_out;
  a.word0[0] = 0;
  for (int i = 0; i < 3; ++i) a.word0[i + 1] = a.word0[i] + (B[i] + 31) / 32;
  uint32_t max_len = 0;
  for (uint64_t i = 0; i < n; ++i) max_len = std::max(max_len, r1->seq_lengths[i] + r2->seq_lengths[i]);

  auto nb = std::make_unique<fb_bank>();
  nb->ctx = c;
  if (int r = nb->data.get(c, (a.word0[3] + kArenaPadWords) * 8)) return r;
  if (int r = nb->offs.get(c, n_out * 8)) return r;
  if (int r = nb->lens.get(c, n_out * 4)) return r;
  a.hist_n = max_len + 1;
  if (int r = dhist.get(c, (uint64_t)a.hist_n * 4)) return r;
  a.odata = nb->data.as<uint64_t>();
  a.ooffs = nb->offs.as<uint64_t>();
  a.olens = nb->lens.as<uint32_t>();
  a.hist = dhist.as<uint32_t>();
  HIP_TRY(c, launch_fill_u32(a.hist, a.hist_n, 0, s));
  HIP_TRY(c, launch_ing_offsets(a, s));
  HIP_TRY(c, launch_ing_gather(a, s));
  HIP_TRY(c, hipMemsetAsync(a.odata + a.word0[3], 0, kArenaPadWords * 8, s));
  std::vector<uint32_t> hist(a.hist_n);
  HIP_TRY(c, hipMemcpyAsync(hist.data(), a.hist, (uint64_t)a.hist_n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipEventRecord(c->ev[1], s));
  HIP_TRY(c, hipStreamSynchronize(s));
  (void)hipEventElapsedTime(&hs.device_ms, c->ev[0], c->ev[1]);
  if (st) *st = hs;
  nb->n_reads = n_out;
  nb->len_hist.assign(a.hist_n, 0);
  for (uint32_t L = 0; L < a.hist_n; ++L)
    if (hist[L]) {
      nb->len_hist[L] = (uint64_t)hist[L] * L;
      nb->max_len = L;
    }
  const uint64_t nr_out[3] = {nm, nu, nu};
  for (int i = 0; i < 3; ++i) {
    nb->ar_first[i] = a.rfirst[i];
    nb->ar_reads[i] = nr_out[i];
    nb->ar_word0[i] = a.word0[i];
    nb->ar_words[i] = a.word0[i + 1] - a.word0[i];
    nb->ar_bases[i] = B[i];
  }
  nb->total_bases = B[0] + B[1] + B[2];
  *out = nb.release();
  return FB_OK;
}

extern "C" int fb_ingest_files(fb_ctx* c, const char* r1_path, const char* r2_path, const fb_ingest_params* p,
                               fb_bank** out, fb_ingest_stats* st) {
  if (!c || !r1_path || !r2_path || !p || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  fb_fastx* x[2] = {nullptr, nullptr};
  int rc[2] = {FB_OK, FB_OK};
  std::string err[2];
  std::thread t([&] {
    rc[1] = fb_fastx_read(r2_path, &x[1]);
    if (rc[1]) err[1] = fb_fastx_last_error();
  });
  rc[0] = fb_fastx_read(r1_path, &x[0]);
  if (rc[0]) err[0] = fb_fastx_last_error();
  t.join();
  std::unique_ptr<fb_fastx, void (*)(fb_fastx*)> own0(x[0], fb_fastx_destroy), own1(x[1], fb_fastx_destroy);
  for (int i = 0; i < 2; ++i)
    if (rc[i]) return set_err(c, rc[i], "%s", err[i].c_str());
  const float parse_ms =
      (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  fb_ingest_stats hs;
  memset(&hs, 0, sizeof hs);
  fb_records_view v[2];
  fb_fastx_view(x[0], &v[0]);
  fb_fastx_view(x[1], &v[1]);
  const int r = fb_ingest_records(c, &v[0], &v[1], p, out, &hs);
  hs.parse_ms = parse_ms;
  if (st) *st = hs;
  return r;
}

// estimate_genome on two read files (estimator.rs:286-440): count_fastx_exact
// over every pair, then the sketch loop over the first pairs, packed to 2 bits
// (lossless for the estimator: its ntHash maps non-ACGT bytes as the packing does).
extern "C" int fb_estimate_genome_files(fb_ctx* c, const char* r1_path, const char* r2_path, uint64_t max_pairs,
                                        double target_cov, fb_genome_estimate* out) {
  if (!c || !r1_path || !r2_path || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  fb_fastx *a = nullptr, *b = nullptr;
  std::string err;
  if (int r = fb::fastx_read_pair(r1_path, r2_path, &a, &b, &err)) return set_err(c, r, "%s", err.c_str());
  std::unique_ptr<fb_fastx, void (*)(fb_fastx*)> own0(a, fb_fastx_destroy), own1(b, fb_fastx_destroy);
  fb_count_stats cs;
  if (int r = fb::fastx_count(a, b, &cs, &err)) return set_err(c, r, "%s", err.c_str());
  const uint64_t n = std::min<uint64_t>(std::min<uint64_t>(max_pairs, 150000), cs.exact_read_pairs);
  fb_records_view v[2];
  fb_fastx_view(a, &v[0]);
  fb_fastx_view(b, &v[1]);
  std::vector<uint64_t> words[2], offs[2];
  std::vector<uint32_t> lens[2];
  fb_arena_view av[2];
  for (int m = 0; m < 2; ++m) {
    uint64_t bases = 0;
    lens[m].assign(v[m].seq_lengths, v[m].seq_lengths + n);
    for (uint64_t i = 0; i < n; ++i) bases += lens[m][i];
    std::vector<uint8_t> cat;
    cat.reserve(bases);
    for (uint64_t i = 0; i < n; ++i)
      cat.insert(cat.end(), v[m].seq + v[m].seq_offsets[i], v[m].seq + v[m].seq_offsets[i] + lens[m][i]);
    words[m].assign((bases + 31) / 32 + 1, 0);
    offs[m].assign(n, 0);
    fb_pack_ascii(cat.data(), lens[m].data(), n, words[m].data(), offs[m].data());
    av[m] = fb_arena_view{words[m].data(), (uint64_t)words[m].size(), offs[m].data(), lens[m].data(), n};
  }
  return fb_estimate_genome(c, &av[0], &av[1], &cs, max_pairs, target_cov, out);
}

// ---------------------------------------------------------------------------
// PE mapper (mapper.rs:89-279): index and pair mapping on the device
// (fb_mapper.hip), the insert-size distribution on the host.

struct fb_pair_map {
  std::vector<fb_pair_link> links;
  std::vector<uint32_t> sizes;
  fb_insert_size_dist dist;
};

// InsertSizeDist (mapper.rs:255-276): median and MAD by selection, which
// picks the same elements as the reference's full sorts.
static void insert_dist(const std::vector<uint32_t>& sizes, fb_insert_size_dist* o) {
  o->n_sizes = sizes.size();
  if (sizes.size() < 20) {
    o->median = 450.0;
    o->mad = 50.0;
    o->min_insert = 250;
    o->max_insert = 700;
    return;
  }
  std::vector<uint32_t> s(sizes);
  const size_t h = s.size() / 2;
  std::nth_element(s.begin(), s.begin() + h, s.end());
  const double med = (double)s[h];
  std::vector<double> d(s.size());
  for (size_t i = 0; i < s.size(); ++i) d[i] = fabs((double)s[i] - med);
  std::nth_element(d.begin(), d.begin() + h, d.end());
  const double mad = std::max(d[h], 20.0);
  o->median = med;
  o->mad = mad;
  o->min_insert = (uint64_t)std::max(med - 3.5 * mad, 50.0);
  o->max_insert = (uint64_t)(med + 3.5 * mad);
}

extern "C" int fb_map_paired_reads(fb_ctx* c, const fb_graph_view* g, const fb_bank* b, uint32_t k,
                                   fb_pair_map** out) {
  if (!c || !b || !out) return set_err(c, FB_ERR_INVALID, "null argument");
  *out = nullptr;
  if (int r = check_view(c, g)) return r;
  if (b->ctx != c) return set_err(c, FB_ERR_INVALID, "the bank belongs to another context");
  if (k < 1 || k > 127) return set_err(c, FB_ERR_INVALID, "k=%u outside 1..127", k);
  if (g->n_nodes >= (1ull << 31)) return set_err(c, FB_ERR_INVALID, "too many unitigs");
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  // every unitig packed as one read (inactive ones too, so ids stay unitig ids)
  const uint64_t U = g->n_nodes;
  std::vector<uint8_t> cat;
  std::vector<uint32_t> lens(U);
  uint64_t bases = 0;
  for (uint64_t i = 0; i < U; ++i) bases += lens[i] = g->lengths[i];
  cat.reserve(bases);
  for (uint64_t i = 0; i < U; ++i) cat.insert(cat.end(), g->seq + g->offsets[i], g->seq + g->offsets[i] + lens[i]);
  std::vector<uint64_t> words((bases + 31) / 32 + 1, 0), offs(U);
  if (U) fb_pack_ascii(cat.data(), lens.data(), U, words.data(), offs.data());
  // index_terminal_kmers windows (mapper.rs:99-131)
  std::vector<uint32_t> ru, rp0;
  std::vector<uint64_t> rpre(1, 0);
  auto range = [&](uint32_t u, uint32_t p0, uint64_t n) {
    if (!n) return;
    ru.push_back(u);
    rp0.push_back(p0);
    rpre.push_back(rpre.back() + n);
  };
  for (uint64_t i = 0; i < U; ++i) {
    if ((view_flags(g, i) & 1) || lens[i] < k) continue;
    const uint64_t t = lens[i] - k + 1;
    if (t <= 2000) {
      range((uint32_t)i, 0, t);
    } else {
      range((uint32_t)i, 0, 1000);
      range((uint32_t)i, (uint32_t)(t - 1000), 1000);
    }
  }
  MapArgs a;
  a.k = k;
  DBuf ud, uo, ul, dru, drp, dpre, dkeys, dcnt, dval, dspec;
  fb_arena_view uv{words.data(), (uint64_t)words.size(), offs.data(), lens.data(), U};
  uint64_t unr = 0;
  if (int r = upload_reads(c, &uv, 1, ud, uo, ul, &unr, nullptr, nullptr, nullptr)) return r;
  a.useq.data = ud.as<uint64_t>();
  a.useq.offs = uo.as<uint64_t>();
  a.useq.lens = ul.as<uint32_t>();
  a.useq.n_reads = U;
  a.ulen = ul.as<uint32_t>();
  if (int r = to_device(c, dru, ru.data(), ru.size() * 4)) return r;
  if (int r = to_device(c, drp, rp0.data(), rp0.size() * 4)) return r;
  if (int r = to_device(c, dpre, rpre.data(), rpre.size() * 8)) return r;
  a.rng_u = dru.as<uint32_t>();
  a.rng_p0 = drp.as<uint32_t>();
  a.rpre = dpre.as<uint64_t>();
  a.n_ranges = (uint32_t)ru.size();
  a.n_windows = rpre.back();
  uint32_t log2cap = 10;
  while ((1ull << log2cap) < 2 * a.n_windows) ++log2cap;
  a.log2cap = log2cap;
  if (int r = dkeys.get(c, (1ull << log2cap) * 8)) return r;
  if (int r = dcnt.get(c, (1ull << log2cap) * 4)) return r;
  if (int r = dval.get(c, (1ull << log2cap) * 8)) return r;
  if (int r = dspec.get(c, 256)) return r;
  a.keys = dkeys.as<uint64_t>();
  a.cnt = dcnt.as<uint32_t>();
  a.val = dval.as<uint64_t>();
  a.special_cnt = dspec.as<uint32_t>();
  a.special_val = reinterpret_cast<uint64_t*>(dspec.as<uint8_t>() + 64);
  HIP_TRY(c, launch_map_clear(a, s));
  HIP_TRY(c, launch_map_insert(a, s));
  // map_paired_reads (mapper.rs:138-253)
  const uint64_t n = std::min(b->ar_reads[1], b->ar_reads[2]);
  if (n >= (1ull << 31)) return set_err(c, FB_ERR_INVALID, "mapping takes < 2^31 pairs");
  a.reads = b->view();
  a.first_f = b->ar_first[1];
  a.first_r = b->ar_first[2];
  a.n_pairs = n;
  DBuf fs, fl, ps, pl, sp, u1, u2, d1, d2, tmp, dsz, dln;
  for (DBuf* d : {&fs, &fl, &ps, &pl})
    if (int r = d->get(c, (n + 1) * 4)) return r;
  for (DBuf* d : {&sp, &u1, &u2, &d1, &d2})
    if (int r = d->get(c, std::max<uint64_t>(n, 1) * 4)) return r;
  a.f_size = fs.as<uint32_t>();
  a.f_link = fl.as<uint32_t>();
  a.pos_size = ps.as<uint32_t>();
  a.pos_link = pl.as<uint32_t>();
  a.span = sp.as<uint32_t>();
  a.lu1 = u1.as<uint32_t>();
  a.lu2 = u2.as<uint32_t>();
  a.ld1 = d1.as<uint32_t>();
  a.ld2 = d2.as<uint32_t>();
  HIP_TRY(c, launch_map_pairs(a, s));
  size_t tb = 0;
  HIP_TRY(c, cg_exclusive_sum_u32(nullptr, nullptr, n + 1, nullptr, &tb, s));
  if (int r = tmp.get(c, std::max<size_t>(tb, 16))) return r;
  tb = tmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u32(a.f_size, a.pos_size, n + 1, tmp.p, &tb, s));
  tb = tmp.bytes;
  HIP_TRY(c, cg_exclusive_sum_u32(a.f_link, a.pos_link, n + 1, tmp.p, &tb, s));
  uint32_t cnt[2] = {0, 0};
  HIP_TRY(c, hipMemcpyAsync(&cnt[0], a.pos_size + n, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&cnt[1], a.pos_link + n, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (int r = dsz.get(c, std::max<uint64_t>(cnt[0], 1) * 4)) return r;
  if (int r = dln.get(c, std::max<uint64_t>(cnt[1], 1) * 2 * sizeof(PairLinkRec))) return r;
  a.sizes = dsz.as<uint32_t>();
  a.links = dln.as<PairLinkRec>();
  HIP_TRY(c, launch_map_compact(a, s));
  auto m = std::make_unique<fb_pair_map>();
  m->sizes.resize(cnt[0]);
  m->links.resize(2ull * cnt[1]);
  static_assert(sizeof(PairLinkRec) == sizeof(fb_pair_link), "link record layout");
  if (cnt[0]) HIP_TRY(c, hipMemcpyAsync(m->sizes.data(), a.sizes, cnt[0] * 4ull, hipMemcpyDeviceToHost, s));
  if (cnt[1])
    HIP_TRY(c, hipMemcpyAsync(m->links.data(), a.links, 2ull * cnt[1] * sizeof(PairLinkRec), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  insert_dist(m->sizes, &m->dist);
  *out = m.release();
  return FB_OK;
}

extern "C" uint64_t fb_pair_map_link_count(const fb_pair_map* m) { return m ? m->links.size() : 0; }
extern "C" int fb_pair_map_links(const fb_pair_map* m, fb_pair_link* out) {
  if (!m || (!out && !m->links.empty())) return FB_ERR_INVALID;
  if (!m->links.empty()) memcpy(out, m->links.data(), m->links.size() * sizeof(fb_pair_link));
  return FB_OK;
}
extern "C" int fb_pair_map_sizes(const fb_pair_map* m, uint32_t* out) {
  if (!m || (!out && !m->sizes.empty())) return FB_ERR_INVALID;
  if (!m->sizes.empty()) memcpy(out, m->sizes.data(), m->sizes.size() * 4);
  return FB_OK;
}
extern "C" int fb_pair_map_dist(const fb_pair_map* m, fb_insert_size_dist* out) {
  if (!m || !out) return FB_ERR_INVALID;
  *out = m->dist;
  return FB_OK;
}
extern "C" void fb_pair_map_destroy(fb_pair_map* m) { delete m; }

// ---------------------------------------------------------------------------
// Consensus polishing (consensus.rs:64-271): seeds, anchors, pileups and the
// consensus on the device (fb_polish.hip).

extern "C" int fb_polish_contigs(fb_ctx* c, const fb_graph_view* g, const fb_bank* b, uint32_t min_output_len,
                                 uint8_t* out_seq, fb_polish_stats* st) {
  if (!c || !b || !st) return set_err(c, FB_ERR_INVALID, "null argument");
  if (int r = check_view(c, g)) return r;
  if (b->ctx != c) return set_err(c, FB_ERR_INVALID, "the bank belongs to another context");
  if (g->n_nodes >= (1ull << 31)) return set_err(c, FB_ERR_INVALID, "too many unitigs");
  const uint64_t U = g->n_nodes;
  uint64_t total = 0;
  for (uint64_t i = 0; i < U; ++i) total += g->lengths[i];
  if (total && !out_seq) return set_err(c, FB_ERR_INVALID, "null output");
  memset(st, 0, sizeof *st);
  if (int r = ctx_current(c)) return r;
  hipStream_t s = c->stream;
  const uint32_t K = 31;
  // nodes as one contiguous ASCII buffer in node order
  std::vector<uint8_t> cat;
  cat.reserve(total);
  std::vector<uint64_t> soff(U);
  std::vector<uint32_t> lens(U);
  for (uint64_t i = 0; i < U; ++i) {
    soff[i] = cat.size();
    lens[i] = g->lengths[i];
    cat.insert(cat.end(), g->seq + g->offsets[i], g->seq + g->offsets[i] + lens[i]);
  }
  std::vector<uint32_t> elig;
  std::vector<uint64_t> poff(U + 1, 0);
  for (uint64_t i = 0; i < U; ++i) {
    const bool e = !(view_flags(g, i) & 1) && lens[i] >= min_output_len;
    poff[i + 1] = poff[i] + (e ? lens[i] : 0);
    if (e) elig.push_back((uint32_t)i);
  }
  auto ranges = [&](uint32_t step, std::vector<uint32_t>* ru, std::vector<uint64_t>* pre) {
    pre->assign(1, 0);
    for (uint32_t u : elig)
      if (lens[u] >= K) {
        ru->push_back(u);
        pre->push_back(pre->back() + (lens[u] - K) / step + 1);
      }
  };
  std::vector<uint32_t> su, au;
  std::vector<uint64_t> spre, apre;
  ranges(25, &su, &spre);
  ranges(1, &au, &apre);
  PolArgs a;
  DBuf dseq, dsoff, dlen, dsu, dspre, dau, dapre, dkeys, dcnt, dval, dspec, dpile, dpoff, delig, dout, dstat;
  if (int r = to_device(c, dseq, cat.data(), cat.size())) return r;
  if (int r = to_device(c, dsoff, soff.data(), U * 8)) return r;
  if (int r = to_device(c, dlen, lens.data(), U * 4)) return r;
  if (int r = to_device(c, dsu, su.data(), su.size() * 4)) return r;
  if (int r = to_device(c, dspre, spre.data(), spre.size() * 8)) return r;
  if (int r = to_device(c, dau, au.data(), au.size() * 4)) return r;
  if (int r = to_device(c, dapre, apre.data(), apre.size() * 8)) return r;
  a.seq = dseq.as<uint8_t>();
  a.soff = dsoff.as<uint64_t>();
  a.ulen = dlen.as<uint32_t>();
  uint32_t log2cap = 10;
  while ((1ull << log2cap) < 2 * spre.back()) ++log2cap;
  a.log2cap = log2cap;
  if (int r = dkeys.get(c, (1ull << log2cap) * 8)) return r;
  if (int r = dcnt.get(c, (1ull << log2cap) * 4)) return r;
  if (int r = dval.get(c, (1ull << log2cap) * 8)) return r;
  if (int r = dspec.get(c, 256)) return r;
  a.keys = dkeys.as<uint64_t>();
  a.cnt = dcnt.as<uint32_t>();
  a.val = dval.as<uint64_t>();
  a.special = dspec.as<uint32_t>();
  a.special_val = reinterpret_cast<uint64_t*>(dspec.as<uint8_t>() + 64);
  HIP_TRY(c, launch_pol_clear(a, s));
  a.rng_u = dsu.as<uint32_t>();
  a.rpre = dspre.as<uint64_t>();
  a.n_ranges = (uint32_t)su.size();
  a.rng_step = 25;
  a.n_windows = spre.back();
  HIP_TRY(c, launch_pol_sample(a, s));
  a.rng_u = dau.as<uint32_t>();
  a.rpre = dapre.as<uint64_t>();
  a.n_ranges = (uint32_t)au.size();
  a.rng_step = 1;
  a.n_windows = apre.back();
  HIP_TRY(c, launch_pol_count(a, s));
  // anchors and pileups: merged reads, then pe_forward (main.rs:466-469)
  if (int r = dpile.get(c, std::max<uint64_t>(poff[U], 1) * 16)) return r;
  if (int r = to_device(c, dpoff, poff.data(), (U + 1) * 8)) return r;
  if (int r = to_device(c, delig, elig.data(), elig.size() * 4)) return r;
  if (int r = dstat.get(c, 64)) return r;
  a.pile = dpile.as<uint32_t>();
  a.poff = dpoff.as<uint64_t>();
  a.elig = delig.as<uint32_t>();
  a.n_elig = (uint32_t)elig.size();
  a.n_pile = poff[U];
  a.stats = dstat.as<unsigned long long>();
  HIP_TRY(c, launch_fill_u32(a.pile, std::max<uint64_t>(poff[U], 1) * 4, 0, s));
  HIP_TRY(c, launch_fill_u32(dstat.as<uint32_t>(), 16, 0, s));
  a.reads = b->view();
  a.first0 = b->ar_first[0];
  a.n_first = b->ar_reads[0];
  a.first1 = b->ar_first[1];
  a.n_reads = b->ar_reads[0] + b->ar_reads[1];
  HIP_TRY(c, launch_pol_reads(a, s));
  if (int r = to_device(c, dout, cat.data(), cat.size())) return r;
  a.out = dout.as<uint8_t>();
  HIP_TRY(c, launch_pol_consensus(a, s));
  unsigned long long sums[6];
  HIP_TRY(c, hipMemcpyAsync(sums, a.stats, sizeof sums, hipMemcpyDeviceToHost, s));
  if (total) HIP_TRY(c, hipMemcpyAsync(out_seq, a.out, total, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  st->fwd_mapped_reads = sums[0];
  st->rev_mapped_reads = sums[1];
  st->fwd_pileup_bases = sums[2];
  st->rev_pileup_bases = sums[3];
  st->rejected_anchors = sums[4];
  st->bases_modified = sums[5];
  return FB_OK;
}
