/*
 * fastbact_synth.h -- seeded synthetic bacterial isolates for tests and the
 * benchmark.  Follows the method of the reference's generators
 * (src/bin/generate_benchmark.rs:14-158: seeded random chromosome with
 * identical repeat copies written over it; src/bin/experiment_a.rs:14-31:
 * GC-controlled random bases) with its own code and its own RNG streams.
 * Reads are paired-end: R1 from the fragment's 5' end, R2 the reverse
 * complement of its 3' end, fragment strand chosen at random.
 */
#ifndef FASTBACT_SYNTH_H
#define FASTBACT_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t seed;
  uint64_t genome_len;
  double gc;                   /* GC fraction of random bases */
  uint32_t operon_copies;      /* identical copies of one operon-like block */
  uint32_t operon_len;
  uint32_t tandem_arrays;      /* tandem repeat arrays */
  uint32_t tandem_unit_min, tandem_unit_max;
  uint32_t tandem_copies_min, tandem_copies_max;
  double coverage;             /* pairs = round(coverage * genome_len / (2 * read_len)) */
  uint32_t read_len;
  double insert_mean, insert_sd;
  double sub_rate;             /* per-base substitution probability */
  double indel_rate;           /* per-base probability of a 1-base insertion or deletion */
} fb_synth_params;

void fb_synth_params_default(fb_synth_params* p);
/* Writes genome_len ASCII bases. */
int fb_synth_genome(const fb_synth_params* p, uint8_t* out);
uint64_t fb_synth_n_pairs(const fb_synth_params* p);
/* Simulates n_pairs pairs of read_len bases each from the genome:
 * r1/r2 receive n_pairs*read_len ASCII bases. */
int fb_synth_reads(const fb_synth_params* p, const uint8_t* genome, uint64_t genome_len, uint64_t n_pairs,
                   uint8_t* r1, uint8_t* r2);

#ifdef __cplusplus
}
#endif
#endif
