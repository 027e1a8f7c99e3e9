// Measurement of the GPU itself (no counterpart in the reference).
#pragma once

namespace sr {

// Attainable HBM read bandwidth: a plain streaming read of `bytes` (>= 1 MiB) timed
// over `reps` passes after two warm-up passes; gbs = bytes / mean pass time in GB/s.
int measure_stream_read(long bytes, int reps, double *gbs);
// Attainable HBM write bandwidth: 16-byte stores, non-temporal or ordinary.
int measure_stream_write(long bytes, int reps, int nontemporal, double *gbs);
// Attainable vector-instruction issue rate (G wave-instructions/s) with every SIMD full.
int measure_valu_issue(int reps, double *ginstr);

}  // namespace sr
