// Host thread count (threads.hpp:19-22 surface). In this build the step loop
// runs on the GPU; host threads only parse catalogues, sort event lists and
// run the breakup model.

#pragma once

namespace kessler {

// requested > 0: use that many OpenMP threads; otherwise the OpenMP default
// (OMP_NUM_THREADS). Returns the count in effect.
int configure_threads(int requested);

}  // namespace kessler
