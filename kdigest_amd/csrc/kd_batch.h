// kd_batch.h — device state and launchers of the per-sample batched path
// (many digests, add_at replay and merge_into tree).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kd {

// Scalar state of one digest (kdigest/tdigest.mojo:116-151) plus the sizes of
// its device-resident centroid and buffer arrays.
struct BatchState {
  double stored_total, log_scale, clock, vmin, vmax;
  int nc, nb, err, pad;
};

int batch_capacity_centroids();
int batch_capacity_buffer();

void launch_replay(hipStream_t s, int scale, double compression, double lambda, int limit,
                   const int64_t* order, int64_t nd, const int64_t* off, const double* vals,
                   const double* times, const double* wts, BatchState* st, double* gcm,
                   double* gcw, double* gbm, double* gbw);
void launch_flush(hipStream_t s, int scale, double compression, double lambda, int64_t nd,
                  BatchState* st, double* gcm, double* gcw, double* gbm, double* gbw);
void launch_merge_pairs(hipStream_t s, int scale, double compression, double lambda, int64_t nd,
                        int64_t stride, BatchState* st, double* gcm, double* gcw, double* gbm,
                        double* gbw);

}  // namespace kd
