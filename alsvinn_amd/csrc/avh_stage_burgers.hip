// Stage kernels for burgers (godunov flux).
#include "avh_stage_inst.hpp"
namespace avh {
Launchers get_launchers_burgers(int recon, int flux, int dim) {
    if (flux != 0) return Launchers();
    return pick_dim<BurgersEq, GodunovFlux>(recon, dim);
}
}  // namespace avh
