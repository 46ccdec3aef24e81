// Explicit instantiation of the stage launchers of one equation for every
// reconstruction, flux and grid dimension the equation supports.
#pragma once
#include "avh_batch.hpp"
#include "avh_stage.hpp"
#include "avh_stage3d.hpp"

namespace avh {

// Host-side launcher types: one per (equation, reconstruction, flux, dim);
// `shape` selects one of the tile shapes below.
typedef hipError_t (*StageLauncher)(const StageParams& P, int shape, hipStream_t stream);
typedef hipError_t (*BatchLauncher)(const BatchParams& B, unsigned int nblocks, int shape, hipStream_t stream);

// Tile shapes the stage kernels are built for; the autotuner times them on
// the solver's own grid and keeps the fastest.  Every shape computes each
// interface flux and each cell update with the same operations in the same
// order, so the choice never changes a result bit.
//   1D: (cells, threads) = (256, 256) or (64, 64): small grids get more
//       workgroups from the smaller tile.
//   2D: 32x16 (LDS 40.8 KB, 4 waves/SIMD for euler2), 32x8 (23 KB, 5
//       waves/SIMD) or 64x4 cells, 256 threads.
//   3D: 32x8 columns marching over 32 planes, 256 threads.
constexpr int NSHAPES1 = 2, NSHAPES2 = 3, NSHAPES3 = 1;
constexpr int SHAPE1_TX[NSHAPES1] = {256, 64};
constexpr int SHAPE2_TX[NSHAPES2] = {32, 32, 64};
constexpr int SHAPE2_TY[NSHAPES2] = {16, 8, 4};

inline int num_shapes(int dim) { return dim == 1 ? NSHAPES1 : (dim == 2 ? NSHAPES2 : NSHAPES3); }

template <class EQ, int RECON, class FLUX, int DIM, int S>
struct KernelFor;
template <class EQ, int RECON, class FLUX, int S>
struct KernelFor<EQ, RECON, FLUX, 1, S> {
    typedef StageKernel<EQ, RECON, FLUX, 1, SHAPE1_TX[S], 1, 1, SHAPE1_TX[S]> type;
};
template <class EQ, int RECON, class FLUX, int S>
struct KernelFor<EQ, RECON, FLUX, 2, S> {
    typedef StageKernel<EQ, RECON, FLUX, 2, SHAPE2_TX[S], SHAPE2_TY[S], 1, 256> type;
};

// stage tiles (workgroups) covering one grid of the 1D / 2D kernels
inline unsigned int stage_tiles(const int* n, int dim, int shape) {
    if (dim == 1) return (unsigned int)((n[0] + SHAPE1_TX[shape] - 1) / SHAPE1_TX[shape]);
    return (unsigned int)(((n[0] + SHAPE2_TX[shape] - 1) / SHAPE2_TX[shape]) *
                          ((n[1] + SHAPE2_TY[shape] - 1) / SHAPE2_TY[shape]));
}

template <class K>
hipError_t launch_batch_k(const BatchParams& B, unsigned int nblocks, hipStream_t stream) {
    hipLaunchKernelGGL(batch_stage_kernel<K>, dim3(nblocks), dim3(K::NTHREADS), 0, stream, B);
    return hipGetLastError();
}

template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_batch_stage(const BatchParams& B, unsigned int nblocks, int shape, hipStream_t stream) {
    if constexpr (DIM == 3) {
        return hipErrorInvalidValue;
    } else if constexpr (DIM == 1) {
        switch (shape) {
        case 0: return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 1, 0>::type>(B, nblocks, stream);
        case 1: return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 1, 1>::type>(B, nblocks, stream);
        }
        return hipErrorInvalidValue;
    } else {
        switch (shape) {
        case 0: return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 2, 0>::type>(B, nblocks, stream);
        case 1: return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 2, 1>::type>(B, nblocks, stream);
        case 2: return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 2, 2>::type>(B, nblocks, stream);
        }
        return hipErrorInvalidValue;
    }
}

template <class K>
hipError_t launch_stage_k(const StageParams& P, hipStream_t stream) {
    hipLaunchKernelGGL(stage_kernel<K>, dim3(K::tiles(P.n)), dim3(K::NTHREADS), 0, stream, P);
    return hipGetLastError();
}

template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_stage(const StageParams& P, int shape, hipStream_t stream) {
    if constexpr (DIM == 3) {
        if (shape != 0) return hipErrorInvalidValue;
        typedef MarchKernel3<EQ, RECON, FLUX, 32, 8, 32> K;
        const unsigned int nb = ((P.n[0] + 31) / 32) * ((P.n[1] + 7) / 8) * ((P.n[2] + 31) / 32);
        hipLaunchKernelGGL(march_kernel3<K>, dim3(nb), dim3(K::NT), 0, stream, P);
        return hipGetLastError();
    } else if constexpr (DIM == 1) {
        switch (shape) {
        case 0: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 1, 0>::type>(P, stream);
        case 1: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 1, 1>::type>(P, stream);
        }
        return hipErrorInvalidValue;
    } else {
        switch (shape) {
        case 0: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 2, 0>::type>(P, stream);
        case 1: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 2, 1>::type>(P, stream);
        case 2: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 2, 2>::type>(P, stream);
        }
        return hipErrorInvalidValue;
    }
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
