"""fastbact_amd: MI355X-native (gfx950) k-mer spectrum path of fastBact.

Rolling canonical k-mers over 2-bit packed reads, the Count-Min Sketch,
HyperLogLog registers, the genome estimator and the solid-k-mer set, as HIP
kernels behind the C ABI of include/fastbact_gpu.h.
"""
from .api import (  # noqa: F401
    Batch,
    BitSeqArena,
    Context,
    CountMinSketch,
    DeviceBank,
    FastBactError,
    GenomeEstimate,
    HyperLogLog,
    ReadBank,
    SolidSet,
    batch_footprint,
    build_solid_handle,
    build_solid_kmers,
    compact_arrays,
    compact_unitigs,
    compute_dynamic_kmers,
    count_fastx,
    count_stats,
    estimate_genome,
    estimate_genome_files,
    filter_active_reads,
    auto_subsample_factor,
    hll_estimate_registers,
    nthash_windows,
    ingest_arrays,
    ingest_files,
    ingest_records,
    map_paired_reads,
    polish_contigs,
    splitmix64,
    synth_isolate,
    read_fastx,
    synth_params,
    upgrade_multik,
)

from .serve import Admission, Server, ServeOutcome, serve_params  # noqa: F401,E402

__version__ = "0.1.0"
