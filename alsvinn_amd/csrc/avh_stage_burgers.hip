// Stage kernels for burgers (godunov flux).
#include "avh_stage_inst.hpp"
namespace avh {
StageLauncher get_launcher_burgers(int recon, int flux, int dim) {
    if (flux != 0) return nullptr;
    return pick_dim<BurgersEq, GodunovFlux>(recon, dim);
}
}  // namespace avh
