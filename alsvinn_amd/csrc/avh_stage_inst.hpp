// Explicit instantiation of the stage launchers of one equation for every
// reconstruction, flux and grid dimension the equation supports.
#pragma once
#include "avh_stage.hpp"

namespace avh {

template <class EQ, class FLUX, int DIM>
StageLauncher pick_recon(int recon) {
    switch (recon) {
    case RECON_NONE: return &launch_stage<EQ, RECON_NONE, FLUX, DIM>;
    case RECON_WENO2: return &launch_stage<EQ, RECON_WENO2, FLUX, DIM>;
    case RECON_ENO2: return &launch_stage<EQ, RECON_ENO2, FLUX, DIM>;
    case RECON_ENO3: return &launch_stage<EQ, RECON_ENO3, FLUX, DIM>;
    }
    return nullptr;
}

template <class EQ, class FLUX>
StageLauncher pick_dim(int recon, int dim) {
    if (dim == 1) return pick_recon<EQ, FLUX, 1>(recon);
    if constexpr (EQ::NSD >= 2) {
        if (dim == 2) return pick_recon<EQ, FLUX, 2>(recon);
    }
    if constexpr (EQ::NSD >= 3) {
        if (dim == 3) return pick_recon<EQ, FLUX, 3>(recon);
    }
    return nullptr;
}

// flux ids follow avh_config: 0 godunov, 1 hll, 2 hll3
StageLauncher get_launcher_burgers(int recon, int flux, int dim);
StageLauncher get_launcher_euler1(int recon, int flux, int dim);
StageLauncher get_launcher_euler2(int recon, int flux, int dim);
StageLauncher get_launcher_euler3(int recon, int flux, int dim);

}  // namespace avh
