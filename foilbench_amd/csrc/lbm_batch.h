// Batched lattice stepping: many D2Q9 lattices advanced by one launch per step, recorded in a
// HIP graph per chunk of steps, with lattices joining and leaving at chunk boundaries.
#pragma once

#include <vector>

#include "fb_common.h"

namespace fb {

struct LbmBatch;

struct LbmDone {
  int64_t id = 0;
  int status = FB_OK;
  int steps = 0;
  fb_lbm_stats stats{};
};

// chunk: lattice steps per graph launch (even, so that a lattice in mid-run sees the same
// buffers at every chunk start and the arguments need no upload).
int lbm_batch_create(Ctx* c, int prec, int chunk, int use_graph, LbmBatch** out);
void lbm_batch_destroy(LbmBatch* b);
// Adds a lattice of the batch's precision whose own stream has no pending work (the caller
// synchronised it). It runs on the batch's stream until it leaves.
int lbm_batch_add(LbmBatch* b, fb_lbm* lattice, const fb_lbm_params& p, int steps, int64_t id);
int lbm_batch_members(const LbmBatch* b);
// Precision of a lattice handle (FB_F64 ... FB_F16_MIXED).
int lbm_handle_precision(const fb_lbm* lattice);
// Waits for the work queued on the lattice's own stream.
hipError_t lbm_handle_sync(fb_lbm* lattice);
int lbm_batch_prec(const LbmBatch* b);
int64_t lbm_batch_cells(const LbmBatch* b);
// One chunk for every member, then a synchronisation; members that reach their step count
// leave the batch (their own context restored) and are appended to `done`.
int lbm_batch_run_chunk(LbmBatch* b, std::vector<LbmDone>& done);

}  // namespace fb
