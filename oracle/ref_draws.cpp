// Test infrastructure: a C entry point over the reference's own
// initial-velocity generator (lib/Driver/InitialVelocityDraws.h, a
// self-contained C++17 header), compiled where the header lies into
// oracle/_ref by `make -C oracle ref`.  The tests compare the oracle's
// normals (orc_initial_normals) with these word for word.  Nothing outside
// tests/ loads it.
#include <cstdint>

#include "InitialVelocityDraws.h"

extern "C" {

// Three normals per particle for the particles order[0..count), in that
// order (each seeks to its counter 6 id first, System.cpp:1298-1311):
// out[3 * order[k] + c].
void ref_initial_normals(uint64_t seed, int64_t count, const int64_t *order, double *out) {
  mdir::driver::InitialVelocityDraws draws(seed);
  for (int64_t k = 0; k < count; ++k) {
    const int64_t id = order[k];
    draws.particle((std::size_t)id);
    for (int c = 0; c < 3; ++c) out[3 * id + c] = draws.nextNormal();
  }
}
}
