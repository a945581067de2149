// Fused per-window kernels (filled in once the staged pipeline is parity-green).
#include "engine.h"

namespace ampcor {

bool fused_supported(const ampcor_params &, const ampcor_sizes &) { return false; }

void Engine::run_fused(const RunArgs &) {
    throw Error(AMPCOR_E_UNSUPPORTED, "fused kernels not built");
}

}  // namespace ampcor
