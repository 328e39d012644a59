#include "threads.hpp"

#include <omp.h>

namespace kessler {

int configure_threads(int requested) {
    if (requested > 0) omp_set_num_threads(requested);
    return omp_get_max_threads();
}

}  // namespace kessler
