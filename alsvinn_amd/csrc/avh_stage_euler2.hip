// Stage kernels for euler2 (hll and hll3 fluxes).
#include "avh_stage_inst.hpp"
namespace avh {
Launchers get_launchers_euler2(int recon, int flux, int dim) {
    if (flux == 1) return pick_dim<EulerEq<2>, HLLFlux>(recon, dim);
    if (flux == 2) return pick_dim<EulerEq<2>, HLL3Flux>(recon, dim);
    return Launchers();
}
}  // namespace avh
