// fb_host.hpp -- host-side helpers shared by the runtime.
#pragma once
#include <stdint.h>

#include "fastbact_gpu.h"

#include <deque>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace fb {

// Admission of serving requests by device-memory footprint.  Requests wait
// in arrival order; the head is admitted while its bytes fit the budget next
// to the admitted ones and fewer than max_inflight are admitted.  A request
// larger than the whole budget is refused at submission.  The head is never
// overtaken, so a large request waits for memory instead of starving.
class Admission {
 public:
  Admission(uint64_t budget, uint32_t max_inflight) : budget_(budget), max_inflight_(max_inflight) {}
  int submit(uint64_t id, uint64_t bytes);  // FB_OK or FB_ERR_OOM
  // Admitted ids in arrival order (appended to out).
  void admit(std::vector<uint64_t>& out);
  int retire(uint64_t id);                  // FB_OK or FB_ERR_INVALID (not admitted)
  uint64_t used() const { return used_; }
  uint64_t budget() const { return budget_; }
  size_t queued() const { return queue_.size(); }
  size_t inflight() const { return inflight_.size(); }

 private:
  uint64_t budget_, used_ = 0;
  uint32_t max_inflight_;
  std::deque<std::pair<uint64_t, uint64_t>> queue_;
  std::map<uint64_t, uint64_t> inflight_;
};

uint64_t f64_to_u64(double x);
// The early-stop test of the estimator's sketch loop (estimator.rs:366-375).
bool checkpoint_stop(uint64_t u1, uint64_t u2, uint64_t m_sample);
// estimator.rs:385-439: Poisson-truncated genome size, coverage, confidence,
// subsample factor and k ladder.
void estimate_finish(uint64_t u_ge_1, uint64_t u_ge_2, uint64_t m_sample, const fb_count_stats* st,
                     double target_cov, fb_genome_estimate* o);
// Both mate files read on two host threads; *err names the failing file's error.
int fastx_read_pair(const char* r1, const char* r2, fb_fastx** a, fb_fastx** b, std::string* err);
// count_fastx_exact (estimator.rs:167-283) on two read files.
int fastx_count(const fb_fastx* a, const fb_fastx* b, fb_count_stats* out, std::string* err);
}  // namespace fb
