// Explicit instantiation of the stage launchers of one equation for every
// reconstruction, flux and grid dimension the equation supports.
#pragma once
#include "avh_stage.hpp"
#include "avh_stage3d.hpp"

namespace avh {

// Host-side launcher type: one per (equation, reconstruction, flux, dim).
typedef hipError_t (*StageLauncher)(const StageParams& P, hipStream_t stream);

// 1D: 256-cell tiles; 2D: 32x16 tiles; 3D: 32x8 column tiles marching over
// 32 planes.
template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_stage(const StageParams& P, hipStream_t stream) {
    if constexpr (DIM == 3) {
        typedef MarchKernel3<EQ, RECON, FLUX, 32, 8, 32> K;
        const unsigned int nb = ((P.n[0] + 31) / 32) * ((P.n[1] + 7) / 8) * ((P.n[2] + 31) / 32);
        hipLaunchKernelGGL(march_kernel3<K>, dim3(nb), dim3(K::NT), 0, stream, P);
    } else {
        typedef typename std::conditional<DIM == 1, StageKernel<EQ, RECON, FLUX, 1, 256, 1, 1, 256>,
                                          StageKernel<EQ, RECON, FLUX, 2, 32, 16, 1, 256>>::type K;
        const unsigned int nb = ((P.n[0] + K::TXC - 1) / K::TXC) * ((P.n[1] + K::TYC - 1) / K::TYC);
        hipLaunchKernelGGL(stage_kernel<K>, dim3(nb), dim3(256), 0, stream, P);
    }
    return hipGetLastError();
}

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
