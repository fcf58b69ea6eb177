// Hook between the solver loop and the learned coarse correction.
#pragma once

#include "flgpu_internal.h"

namespace flgpu {

// Adds the network's correction for residual r to the preconditioned vector
// q (both of the system's precision and length).
struct LearnedHook {
  virtual ~LearnedHook() {}
  virtual void correct(void *r, void *q) = 0;
};

// The hook of the network attached to mg (error if none is attached).
LearnedHook *learned_hook(flgpu_mg *mg, const flgpu_mat *A);
void mg_apply_ptr(flgpu_mg *mg, int vprec, void *x, void *b, void *r);

}  // namespace flgpu
