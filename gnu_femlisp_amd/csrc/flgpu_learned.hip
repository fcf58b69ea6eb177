// Learned coarse correction: attachment point in the solver loop.
#include "flgpu_learned.h"

namespace flgpu {

LearnedHook *learned_hook(flgpu_mg *mg, const flgpu_mat *A) {
  (void)mg;
  (void)A;
  FLGPU_REQUIRE(false, "no learned network is attached to the hierarchy");
  return nullptr;
}

}  // namespace flgpu
