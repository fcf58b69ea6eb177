// Hook between the solver loop and the learned coarse correction.
#pragma once

#include <memory>

#include "flgpu_internal.h"

namespace flgpu {

// Adds the network's correction to the preconditioned vector q (both of the
// system's precision and length) for the residual r it was computed from.
struct LearnedHook {
  virtual ~LearnedHook() {}
  // q holds the cycle's output z1 for the residual r.  The network's correction is added to q
  // (on the device) when its energy-optimal step adds at least min_gain of the energy of the
  // cycle's own step, z1 . r; the reduction slot `flag` receives 1.0 when it was added, else
  // 0.0.  Slots 11..13 hold the step's dots.
  virtual void correct(void *r, void *q, double min_gain, int flag) = 0;
};

// The hook of the network attached to mg (error if none is attached).
std::unique_ptr<LearnedHook> learned_hook(flgpu_mg *mg, const flgpu_mat *A);
void mg_apply_ptr(flgpu_mg *mg, int vprec, void *x, void *b, void *r);
// the context a network was created on
flgpu_ctx *unet_ctx(const flgpu_unet *u);
// network attached to a hierarchy (flgpu_mg_attach_unet)
void mg_unet(flgpu_mg *mg, flgpu_unet **u, const flgpu_vec **kappa, int *n);
// lattice hierarchies: the level with n cells per axis (-1 if none), the top level, the
// cycle's precision, a level's rows, and the cycle's transfers between levels l and l - 1
int mg_level_cells(const flgpu_mg *mg, int n);
int mg_top_level(const flgpu_mg *mg);
int mg_precision(const flgpu_mg *mg);
long mg_level_rows(const flgpu_mg *mg, int l);
void mg_restrict(flgpu_mg *mg, int l, const void *fine, void *coarse);
void mg_prolong(flgpu_mg *mg, int l, const void *coarse, void *fine);

}  // namespace flgpu
