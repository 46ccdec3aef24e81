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
// Flux-only stage over a box of a reference Volume (P.n = box extent)
typedef hipError_t (*VolumeLauncher)(const StageParams& P, const VolumeParams& V, hipStream_t stream);

// Tile shapes the stage kernels are built for; the autotuner times them on
// the solver's own grid and keeps the fastest.  Every shape computes each
// interface flux and each cell update with the same operations in the same
// order, so the choice never changes a result bit.
//   1D: (cells, threads) = (256, 256) or (64, 64): small grids get more
//       workgroups from the smaller tile.
//   2D: 32x16, 16x15, 64x4, 32x15 or 24x20 cells, 256 threads (LDS 40.8,
//       20.7, 28.0, 38.6 and 38.0 KB for euler2 with WENO2).  A
//       flux pass hands interface items to threads in turns of 256, and a
//       turn costs a wave its full instruction stream however few of its
//       lanes have an item, so shapes whose interface counts sit just under
//       a multiple of 256 waste the least: 32x15 (495 x- / 512 y-interfaces),
//       24x20 (500 / 504) and 16x15 (255 / 256), against 32x16 (528 / 544).
//   3D: columns marching over 32 planes, 256 threads: 32x8 (264 / 288
//       plane interfaces), 16x15 or 32x7 (at most 256 each: one turn; the
//       threads past the columns only work on the planes); then 16x15 over
//       32 and over 16 planes and 32x7 over 32 planes with two barriers
//       per plane (double-buffered plane tile); then 16x15 over 32 planes,
//       two barriers, fluxes under SharedOps with the PlainOps redo.
constexpr int NSHAPES1 = 2, NSHAPES2 = 5, NSHAPES3 = 7;
constexpr int SHAPE1_TX[NSHAPES1] = {256, 64};
constexpr int SHAPE2_TX[NSHAPES2] = {32, 16, 64, 32, 24};
constexpr int SHAPE2_TY[NSHAPES2] = {16, 15, 4, 15, 20};
constexpr int SHAPE3_TX[NSHAPES3] = {32, 16, 32, 16, 16, 32, 16};
constexpr int SHAPE3_TY[NSHAPES3] = {8, 15, 7, 15, 15, 7, 15};
constexpr int SHAPE3_LZ[NSHAPES3] = {32, 32, 32, 32, 16, 32, 32};
constexpr bool SHAPE3_TWO_BARRIERS[NSHAPES3] = {false, false, false, true, true, true, true};
constexpr bool SHAPE3_SHARED[NSHAPES3] = {false, false, false, false, false, false, true};

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
// 2D shapes whose RK3 stages are each compiled for their mode (the shapes
// the autotuner takes for the benchmark grids and the batched set); those
// kernels index cells with 32-bit values, so they serve grids of fewer than
// 2^29 cells (per sample in a batch)
constexpr bool SHAPE2_BY_MODE[NSHAPES2] = {false, false, false, true, true};
constexpr long long MODE_KERNEL_MAX_CELLS = 1ll << 29;
template <class EQ, int RECON, class FLUX, int S, int MODE>
struct KernelForMode {
    typedef StageKernel<EQ, RECON, FLUX, 2, SHAPE2_TX[S], SHAPE2_TY[S], 1, 256, MODE> type;
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
#define AVH_BATCH_MODES(S)                                                                                  \
    case S:                                                                                                 \
        if constexpr (SHAPE2_BY_MODE[S]) {                                                                  \
            if (!small) {                                                                                   \
            } else if (B.mode == MODE_L)                                                                    \
                return launch_batch_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_L>::type>(B, nblocks, \
                                                                                               stream);     \
            else if (B.mode == MODE_RK3_2)                                                                  \
                return launch_batch_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_RK3_2>::type>(        \
                    B, nblocks, stream);                                                                    \
            else if (B.mode == MODE_RK3_3)                                                                  \
                return launch_batch_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_RK3_3>::type>(        \
                    B, nblocks, stream);                                                                    \
        }                                                                                                   \
        return launch_batch_k<typename KernelFor<EQ, RECON, FLUX, 2, S>::type>(B, nblocks, stream);
        bool small = true;
        for (int i = 0; i < B.ngroups; ++i) small = small && B.g[i].ncells < MODE_KERNEL_MAX_CELLS;
        switch (shape) {
            AVH_BATCH_MODES(0)
            AVH_BATCH_MODES(1)
            AVH_BATCH_MODES(2)
            AVH_BATCH_MODES(3)
            AVH_BATCH_MODES(4)
        }
#undef AVH_BATCH_MODES
        return hipErrorInvalidValue;
    }
}

template <class K>
hipError_t launch_stage_k(const StageParams& P, hipStream_t stream) {
    hipLaunchKernelGGL(stage_kernel<K>, dim3(K::tiles(P.n)), dim3(K::NTHREADS), 0, stream, P);
    return hipGetLastError();
}

template <class K>
hipError_t launch_march_k(const StageParams& P, hipStream_t stream) {
    hipLaunchKernelGGL(march_kernel3<K>, dim3(K::tiles(P.n)), dim3(K::NT), 0, stream, P);
    return hipGetLastError();
}

template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_stage(const StageParams& P, int shape, hipStream_t stream) {
    if constexpr (DIM == 3) {
        // shapes without SharedOps use the compiler's sqrt and divide throughout
#define AVH_MARCH_CASE(S)                                                                                   \
    case S:                                                                                                 \
        return launch_march_k<MarchKernel3<EQ, RECON,                                                       \
                                           typename std::conditional<SHAPE3_SHARED[S], FLUX,                \
                                                                     typename FLUX::Plain>::type,           \
                                           SHAPE3_TX[S], SHAPE3_TY[S], SHAPE3_LZ[S], 256,                   \
                                           SHAPE3_TWO_BARRIERS[S]>>(P, stream);
        switch (shape) {
            AVH_MARCH_CASE(0)
            AVH_MARCH_CASE(1)
            AVH_MARCH_CASE(2)
            AVH_MARCH_CASE(3)
            AVH_MARCH_CASE(4)
            AVH_MARCH_CASE(5)
        case 6: {
            // the RK3 stages of the shape a 3D run normally takes, each compiled
            // for its mode; other modes take the kernel that reads P.mode
            typedef typename std::conditional<SHAPE3_SHARED[6], FLUX, typename FLUX::Plain>::type F6;
            if (P.ncells >= MODE_KERNEL_MAX_CELLS)  // the per-mode kernels index with 32-bit values
                return launch_march_k<MarchKernel3<EQ, RECON, F6, SHAPE3_TX[6], SHAPE3_TY[6], SHAPE3_LZ[6], 256,
                                                   SHAPE3_TWO_BARRIERS[6]>>(P, stream);
            if (P.mode == MODE_L)
                return launch_march_k<MarchKernel3<EQ, RECON, F6, SHAPE3_TX[6], SHAPE3_TY[6], SHAPE3_LZ[6], 256,
                                                   SHAPE3_TWO_BARRIERS[6], MODE_L>>(P, stream);
            if (P.mode == MODE_RK3_2)
                return launch_march_k<MarchKernel3<EQ, RECON, F6, SHAPE3_TX[6], SHAPE3_TY[6], SHAPE3_LZ[6], 256,
                                                   SHAPE3_TWO_BARRIERS[6], MODE_RK3_2>>(P, stream);
            if (P.mode == MODE_RK3_3)
                return launch_march_k<MarchKernel3<EQ, RECON, F6, SHAPE3_TX[6], SHAPE3_TY[6], SHAPE3_LZ[6], 256,
                                                   SHAPE3_TWO_BARRIERS[6], MODE_RK3_3>>(P, stream);
            return launch_march_k<MarchKernel3<EQ, RECON, F6, SHAPE3_TX[6], SHAPE3_TY[6], SHAPE3_LZ[6], 256,
                                               SHAPE3_TWO_BARRIERS[6]>>(P, stream);
        }
        }
#undef AVH_MARCH_CASE
        return hipErrorInvalidValue;
    } else if constexpr (DIM == 1) {
        switch (shape) {
        case 0: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 1, 0>::type>(P, stream);
        case 1: return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 1, 1>::type>(P, stream);
        }
        return hipErrorInvalidValue;
    } else {
#define AVH_STAGE_MODES(S)                                                                                  \
    case S:                                                                                                 \
        if constexpr (SHAPE2_BY_MODE[S]) {                                                                  \
            if (P.ncells >= MODE_KERNEL_MAX_CELLS) {                                                        \
            } else if (P.mode == MODE_L)                                                                    \
                return launch_stage_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_L>::type>(P, stream); \
            else if (P.mode == MODE_RK3_2)                                                                  \
                return launch_stage_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_RK3_2>::type>(P,      \
                                                                                                 stream);   \
            else if (P.mode == MODE_RK3_3)                                                                  \
                return launch_stage_k<typename KernelForMode<EQ, RECON, FLUX, S, MODE_RK3_3>::type>(P,      \
                                                                                                 stream);   \
        }                                                                                                   \
        return launch_stage_k<typename KernelFor<EQ, RECON, FLUX, 2, S>::type>(P, stream);
        switch (shape) {
            AVH_STAGE_MODES(0)
            AVH_STAGE_MODES(1)
            AVH_STAGE_MODES(2)
            AVH_STAGE_MODES(3)
            AVH_STAGE_MODES(4)
        }
#undef AVH_STAGE_MODES
        return hipErrorInvalidValue;
    }
}

// The reference-Volume kernels: one tile per dimension -- 1D 256 cells, 2D
// 32x15, 3D 8x4x4 bricks on 128 threads (the ghost-width-3 halo of a brick
// still fits one workgroup's LDS for euler3) -- since this path serves the
// reference's unfused per-call interfaces.
template <class EQ, int RECON, class FLUX, int DIM>
struct VolumeKernelFor;
template <class EQ, int RECON, class FLUX>
struct VolumeKernelFor<EQ, RECON, FLUX, 1> {
    typedef StageKernel<EQ, RECON, FLUX, 1, 256, 1, 1, 256> type;
};
template <class EQ, int RECON, class FLUX>
struct VolumeKernelFor<EQ, RECON, FLUX, 2> {
    typedef StageKernel<EQ, RECON, FLUX, 2, 32, 15, 1, 256> type;
};
template <class EQ, int RECON, class FLUX>
struct VolumeKernelFor<EQ, RECON, FLUX, 3> {
    typedef StageKernel<EQ, RECON, typename FLUX::Plain, 3, 8, 4, 4, 128> type;
};

template <class EQ, int RECON, class FLUX, int DIM>
hipError_t launch_volume_stage(const StageParams& P, const VolumeParams& V, hipStream_t stream) {
    typedef typename VolumeKernelFor<EQ, RECON, FLUX, DIM>::type K;
    hipLaunchKernelGGL(volume_stage_kernel<K>, dim3(K::tiles(P.n)), dim3(K::NTHREADS), 0, stream, P, V);
    return hipGetLastError();
}

struct Launchers {
    StageLauncher stage = nullptr;
    BatchLauncher batch = nullptr;
    VolumeLauncher volume = nullptr;
};

template <class EQ, class FLUX, int DIM>
Launchers pick_recon(int recon) {
    Launchers L;
    switch (recon) {
    case RECON_NONE:
        L.stage = &launch_stage<EQ, RECON_NONE, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_NONE, FLUX, DIM>;
        L.volume = &launch_volume_stage<EQ, RECON_NONE, FLUX, DIM>;
        break;
    case RECON_WENO2:
        L.stage = &launch_stage<EQ, RECON_WENO2, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_WENO2, FLUX, DIM>;
        L.volume = &launch_volume_stage<EQ, RECON_WENO2, FLUX, DIM>;
        break;
    case RECON_ENO2:
        L.stage = &launch_stage<EQ, RECON_ENO2, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_ENO2, FLUX, DIM>;
        L.volume = &launch_volume_stage<EQ, RECON_ENO2, FLUX, DIM>;
        break;
    case RECON_ENO3:
        L.stage = &launch_stage<EQ, RECON_ENO3, FLUX, DIM>;
        L.batch = &launch_batch_stage<EQ, RECON_ENO3, FLUX, DIM>;
        L.volume = &launch_volume_stage<EQ, RECON_ENO3, FLUX, DIM>;
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
