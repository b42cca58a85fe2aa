// fb_host.hpp -- host-side helpers shared by the runtime.
#pragma once
#include <stdint.h>

#include "fastbact_gpu.h"

namespace fb {
uint64_t f64_to_u64(double x);
// The early-stop test of the estimator's sketch loop (estimator.rs:366-375).
bool checkpoint_stop(uint64_t u1, uint64_t u2, uint64_t m_sample);
// estimator.rs:385-439: Poisson-truncated genome size, coverage, confidence,
// subsample factor and k ladder.
void estimate_finish(uint64_t u_ge_1, uint64_t u_ge_2, uint64_t m_sample, const fb_count_stats* st,
                     double target_cov, fb_genome_estimate* o);
}  // namespace fb
