// kh_batch.h -- batches of runs of one grid stepped by shared launches
// (kh_run.hip), used by the serving loop (kh_server.hip).
//
// A batch owns `capacity` equal device arenas in one allocation. A run is
// created in a free slot (its buffers carved from that slot's arena), takes
// its first steps alone on its own stream (the initial block, the Euler
// step), then joins the batch: from then on every pass of the fluid part of
// the step is one launch over all joined runs (grid z = slot), the particle
// launches stay per run, and the whole batch step is one captured HIP graph
// per membership. Only runs of plain spheres (no squirmers, rigid bodies or
// shear) on the fused step path join.
#pragma once

#include <stddef.h>

#include "../../include/kapsel_hip.h"

struct kh_batch;

namespace kh {

// arena bytes per slot for runs of P's class with up to np_cap particles
size_t batch_stride(const kh_params &P, int np_cap);
// P's grid, precision, dx, rho, eta, dt, FIX_CELL, a and a_xi define the class
int  batch_create(const kh_params &P, int np_cap, int capacity, kh_batch **out);
bool batch_accepts(const kh_batch *b, const kh_params &P, int np);
int  batch_free_slot(const kh_batch *b);  // -1: none
int  batch_run_create(kh_batch *b, int slot, const kh_params *P, int np, kh_run **out);
// r (created in the batch, ts >= 1, its own stream drained here) steps with the batch from now on
int batch_join(kh_batch *b, kh_run *r);
// r leaves (the batch's stream drained here) and its slot is free once r is destroyed
int batch_leave(kh_batch *b, kh_run *r);
int batch_members(const kh_batch *b);
int batch_step(kh_batch *b, int nsteps);
int batch_query(kh_batch *b);  // 1 idle, 0 busy, < 0 error
int batch_destroy(kh_batch *b);

}  // namespace kh
