// Explicit instantiation of the stage launchers of one equation for every
// reconstruction, flux and grid dimension the equation supports.
#pragma once
#include "avh_batch.hpp"
#include "avh_stage.hpp"
#include "avh_stage3d.hpp"

namespace avh {

// Host-side launcher types: one per (equation, reconstruction, flux, dim).
typedef hipError_t (*StageLauncher)(const StageParams& P, hipStream_t stream);
typedef hipError_t (*BatchLauncher)(const BatchParams& B, unsigned int nblocks, hipStream_t stream);

// tile shapes of the 1D / 2D stage kernel (cells)
constexpr int TILE1_X = 256;
constexpr int TILE2_X = 32;
constexpr int TILE2_Y = 16;

inline unsigned int stage_tiles(const int* n, int dim) {
    if (dim == 1) return (unsigned int)((n[0] + TILE1_X - 1) / TILE1_X);
    return (unsigned int)(((n[0] + TILE2_X - 1) / TILE2_X) * ((n[1] + TILE2_Y - 1) / TILE2_Y));
}

template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_batch_stage(const BatchParams& B, unsigned int nblocks, hipStream_t stream) {
    if constexpr (DIM == 3) {
        return hipErrorInvalidValue;
    } else {
        typedef typename std::conditional<DIM == 1, StageKernel<EQ, RECON, FLUX, 1, TILE1_X, 1, 1, 256>,
                                          StageKernel<EQ, RECON, FLUX, 2, TILE2_X, TILE2_Y, 1, 256>>::type K;
        hipLaunchKernelGGL(batch_stage_kernel<K>, dim3(nblocks), dim3(256), 0, stream, B);
        return hipGetLastError();
    }
}

// 1D: 256-cell tiles; 2D: 32x16 tiles; 3D: 32x8 column tiles marching over
// 32 planes.
template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_stage(const StageParams& P, hipStream_t stream) {
    if constexpr (DIM == 3) {
        typedef MarchKernel3<EQ, RECON, FLUX, 32, 8, 32> K;
        const unsigned int nb = ((P.n[0] + 31) / 32) * ((P.n[1] + 7) / 8) * ((P.n[2] + 31) / 32);
        hipLaunchKernelGGL(march_kernel3<K>, dim3(nb), dim3(K::NT), 0, stream, P);
    } else {
        typedef typename std::conditional<DIM == 1, StageKernel<EQ, RECON, FLUX, 1, TILE1_X, 1, 1, 256>,
                                          StageKernel<EQ, RECON, FLUX, 2, TILE2_X, TILE2_Y, 1, 256>>::type K;
        const unsigned int nb = ((P.n[0] + K::TXC - 1) / K::TXC) * ((P.n[1] + K::TYC - 1) / K::TYC);
        hipLaunchKernelGGL(stage_kernel<K>, dim3(nb), dim3(256), 0, stream, P);
    }
    return hipGetLastError();
}

struct Launchers {
    StageLauncher stage = nullptr;
    BatchLauncher batch = nullptr;
};

template <class EQ, class FLUX, int DIM>
Launchers pick_recon(int recon) {
    Launchers L;
    switch (recon) {
    case RECON_NONE:
        L.stage = &launch_stage<EQ, RECON_NONE, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_NONE, FLUX, DIM>;
        break;
    case RECON_WENO2:
        L.stage = &launch_stage<EQ, RECON_WENO2, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_WENO2, FLUX, DIM>;
        break;
    case RECON_ENO2:
        L.stage = &launch_stage<EQ, RECON_ENO2, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_ENO2, FLUX, DIM>;
        break;
    case RECON_ENO3:
        L.stage = &launch_stage<EQ, RECON_ENO3, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_ENO3, FLUX, DIM>;
        break;
    }
    if (DIM == 3) L.batch = nullptr;
    return L;
}

template <class EQ, class FLUX>
Launchers pick_dim(int recon, int dim) {
    if (dim == 1) return pick_recon<EQ, FLUX, 1>(recon);
    if constexpr (EQ::NSD >= 2) {
        if (dim == 2) return pick_recon<EQ, FLUX, 2>(recon);
    }
    if constexpr (EQ::NSD >= 3) {
        if (dim == 3) return pick_recon<EQ, FLUX, 3>(recon);
    }
    return Launchers();
}

// flux ids follow avh_config: 0 godunov, 1 hll, 2 hll3
Launchers get_launchers_burgers(int recon, int flux, int dim);
Launchers get_launchers_euler1(int recon, int flux, int dim);
Launchers get_launchers_euler2(int recon, int flux, int dim);
Launchers get_launchers_euler3(int recon, int flux, int dim);

}  // namespace avh
