// Fused FP64 finite-volume stage kernel.
//
// One launch = one RK stage over the whole grid.  Each workgroup owns a
// TX x TY x TZ tile of interior cells:
//   1. stage the tile plus a GHOST-wide halo of the stage input into LDS;
//      halo cells outside the domain are read through the boundary map
//      (periodic wrap / neumann mirror), which is the ghost fill of the
//      reference done at load time; in the on-the-fly modes the stage input
//      is formed while loading as U1 = L0*s + U0;
//   2. per direction: one work item per interface, reconstructing the right
//      face of the left cell and the left face of the right cell and
//      evaluating the numerical flux; fluxes go to LDS, the interface wave
//      speeds to a per-thread max;
//   3. each thread accumulates the net flux of its cells in registers,
//      out = 0 + ((-F[i+1/2]) + F[i-1/2]) per direction, exactly as
//      NumericalFluxCPU::computeNetFlux does on a zeroed output;
//   4. the RK stage combination of the integrator, the constraint check of
//      the final stage, and the write of the result.
// The CFL wave speed is reduced with wavefront shuffles and one atomicMax per
// direction per workgroup (non-negative doubles order like their bit
// patterns); the launches after the first stage derive dt from those maxima
// themselves, and a one-wave kernel closes the step (time, save counters,
// loop test), so a whole time step runs without a host round trip and
// without grid-wide fences.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "avh_physics.hpp"

#pragma clang fp contract(off)

namespace avh {

constexpr int MAX_ADJ = 4;

// Device-resident step bookkeeping.  Written by the host between launches,
// by atomics (wave-speed maxima, violation flag) and by finish_step_kernel.
struct Control {
    unsigned long long ws_bits[3];  // running max wave speed per direction (double bits)
    double t;                        // current time (TimestepInformation)
    double dt;                       // dt of the step in flight
    double scale;                    // dt / dx (RungeKutta*: output *= dt/cellLengths.x)
    double last_dt;
    double t_end;
    double cfl;
    double cell[3];
    double adj_interval[MAX_ADJ];
    double adj_start[MAX_ADJ];
    long long adj_saved[MAX_ADJ];  // FixedIntervalWriter::numberSaved
    long long steps;
    int n_adj;
    int active;         // 1 while the step in flight must run
    int violated;       // sticky constraint failure
    int saves_pending;  // fixed-interval save times reached, not yet acknowledged
    int pause_on_save;
    int block_violation;
    // workgroups of a finishing stage launch that are done (low 32 bits) and
    // how many of them found a constraint violation (high 32 bits); the last
    // one runs the end of the step and sets it back to 0
    unsigned long long arrive;
};

enum StageMode {
    MODE_FLUX_ONLY = 0,  // out = net flux, wave speeds accumulated, nothing else
    MODE_L = 1,          // first stage: out = L(U0), last block computes dt
    MODE_RK2_2 = 2,      // in = U1;      out = ((L*s + U1) + U0) * 0.5, ends step
    MODE_RK3_2 = 3,      // in = L0*s+U0; out = ((L*s + U1) * (1/3) + U0) * (3/4)
    MODE_RK3_3 = 4       // in = U2;      out = ((L*s + U2) * 2 + U0) * (1/3), ends step
};

enum { BC_PERIODIC = 0, BC_NEUMANN = 1 };

struct StageParams {
    const double* in;  // stage input U_s, or L0 in the on-the-fly modes
    const double* u0;  // U0 (on-the-fly input and RK combination)
    double* out;       // stage output (may equal u0 in MODE_RK2_2 / MODE_RK3_3:
                       // u0 is then read only at the cell being written)
    Control* ctrl;
    long long ncells;  // interior cells; variable stride of every buffer
    int n[3];
    int bc;
    int mode;
    double gamma;
    // 1: the launch is the step's last stage and its last workgroup runs the
    // end of the step (end_step); 0: the caller runs finish_step_kernel
    int finish = 0;
};

// The reference's Volume (alsfvm/src/volume/Volume.cpp:44-60): one array per
// conserved variable including its ghost cells, x fastest.  A stage kernel
// instantiated for it covers a box of the interior (P.n = the box's extent)
// and reads the halo from the ghost cells exactly as the caller filled them
// (boundary conditions or a neighbour's cells), as the reference's
// NumericalFlux::computeFlux does.
struct VolumeParams {
    const double* in[5];
    double* out[5];
    int tot[3];    // cells per axis including ghosts
    int ghost[3];  // ghost cells per axis of the volume
    int org[3];    // first cell of the box, interior coordinates
};

__device__ __forceinline__ long long volume_index(const VolumeParams& V, int x, int y, int z) {
    return ((long long)(z + V.org[2] + V.ghost[2]) * V.tot[1] + (y + V.org[1] + V.ghost[1])) * V.tot[0] +
           (x + V.org[0] + V.ghost[0]);
}

// Element v of cell gi of an SoA buffer of nc cells per variable.  With
// OFF32 the cell index is a 32-bit value (the caller guarantees nc < 2^29,
// so byte offsets stay below 2^32) and the variable's base a uniform value
// the compiler cannot re-associate with the lane offset: the access is then
// a global load / store with a scalar base and a 32-bit lane offset, with no
// 64-bit address arithmetic per lane.
typedef __attribute__((address_space(1))) char global_char;
__device__ __forceinline__ global_char* uniform_base(const double* p) {
    const unsigned long long a = (unsigned long long)p;
    const unsigned int lo = __builtin_amdgcn_readfirstlane((unsigned int)a);
    const unsigned int hi = __builtin_amdgcn_readfirstlane((unsigned int)(a >> 32));
    return (global_char*)(((unsigned long long)hi << 32) | lo);
}
template <bool OFF32, class Idx>
__device__ __forceinline__ double soa_ld(const double* base, long long nc, int v, Idx gi) {
    if constexpr (OFF32) {
        return *(const __attribute__((address_space(1))) double*)(uniform_base(base + v * nc) + gi * 8u);
    } else {
        return base[v * nc + gi];
    }
}
template <bool OFF32, class Idx>
__device__ __forceinline__ void soa_st(double* base, long long nc, int v, Idx gi, double x) {
    if constexpr (OFF32) {
        *(__attribute__((address_space(1))) double*)(uniform_base(base + v * nc) + gi * 8u) = x;
    } else {
        base[v * nc + gi] = x;
    }
}

// Boundary map of a coordinate in [-G, n+G) onto the interior: periodic wrap
// (alsfvm/include/alsfvm/boundary/Periodic.hpp:24-42) or neumann mirror
// (alsfvm/include/alsfvm/boundary/Neumann.hpp:24-35).  Applied per axis this
// reproduces the reference's x -> y -> z fill order, corners included.
__device__ __forceinline__ int bc_map(int g, int n, int bc) {
    if (g < 0) return bc == BC_PERIODIC ? g + n : -1 - g;
    if (g >= n) return bc == BC_PERIODIC ? g - n : 2 * n - 1 - g;
    return g;
}

__device__ __forceinline__ double ws_max(double acc, double v) {
    // std::max(acc, v): NaN candidates leave acc unchanged
    return (acc < v) ? v : acc;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = ws_max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ unsigned int tile_id_xcd(unsigned int b, unsigned int nb) {
    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
    // contiguous run of tiles so tiles sharing halos share an L2.
    const unsigned int nfull = nb - nb % 8u;
    if (b >= nfull) return b;
    return (b % 8u) * (nfull / 8u) + b / 8u;
}

// largest wave speed recorded along direction d
__host__ __device__ __forceinline__ double wave_speed(const Control* c, int d) {
    double w;
    __builtin_memcpy(&w, &c->ws_bits[d], sizeof w);
    return w;
}

__host__ __device__ __forceinline__ void clear_wave_speeds(Control* c) {
    for (int d = 0; d < 3; ++d) c->ws_bits[d] = 0ull;
}

// dt of the step in flight: cfl / sum_d(ws_d / dx_d), then the
// FixedIntervalWriter clips (alsfvm/src/integrator/Integrator.cpp:20-66,
// alsfvm/src/io/FixedIntervalWriter.cpp:45-53).  ws_d are the maxima the
// first stage accumulated; every launch after it evaluates this same
// expression, so all of them use the identical dt without a grid-wide
// barrier or fence.
__device__ __forceinline__ double step_dt(const Control* c) {
    double total = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (c->cell[d] == 0) continue;
        total += wave_speed(c, d) / c->cell[d];
    }
    double dt = c->cfl / total;
    for (int i = 0; i < c->n_adj; ++i) {
        if (c->adj_saved[i] > 0) {
            const double next = (double)c->adj_saved[i] * c->adj_interval[i] + c->adj_start[i];
            const double cand = next - c->t;
            dt = (cand < dt) ? cand : dt;
        }
    }
    return dt;
}

// End of a step, one thread: t += dt, constraint flag, then the writers'
// save counters (alsfvm/src/simulator/Simulator.cpp:81-85, 180-214;
// alsfvm/src/io/FixedIntervalWriter.cpp:33-43) and the loop test
// `t >= endTime` (Simulator.cpp:77-79); clears the wave-speed maxima for the
// next step.
__device__ __forceinline__ void end_step(Control* c, bool violation) {
    const double dt = step_dt(c);
    c->dt = dt;
    c->scale = dt / c->cell[0];
    c->t = c->t + dt;
    c->steps = c->steps + 1;
    c->last_dt = dt;
    if (violation) {
        c->violated = 1;
    } else {
        for (int i = 0; i < c->n_adj; ++i) {
            if (c->t >= (double)c->adj_saved[i] * c->adj_interval[i] + c->adj_start[i]) {
                c->adj_saved[i] = c->adj_saved[i] + 1;
                c->saves_pending = c->saves_pending + 1;
            }
        }
    }
    c->block_violation = 0;
    clear_wave_speeds(c);
    const bool more = !(c->t >= c->t_end) && !c->violated && !(c->pause_on_save && c->saves_pending > 0);
    c->active = more ? 1 : 0;
}

// Thread 0 of every workgroup of a step's last stage reports here once the
// workgroup's cells are written (`bad`: a cell broke the constraints); the
// last workgroup to arrive ends the step.  The count and the violations
// travel in one atomic, so no fence is needed: the end of the step reads
// only what earlier launches wrote (the first stage's wave speeds, the
// previous step's bookkeeping).
__device__ __forceinline__ void arrive_final(Control* c, bool bad) {
    const unsigned long long add = 1ull + (bad ? (1ull << 32) : 0ull);
    const unsigned long long prev = atomicAdd(&c->arrive, add);
    if ((unsigned int)prev == gridDim.x - 1) {
        c->arrive = 0ull;
        end_step(c, ((prev + add) >> 32) != 0ull);
    }
}

// MODE_CT >= 0 compiles the kernel for that StageMode alone (P.mode must
// equal it): the other modes' branches and registers drop out.
template <class EQ, int RECON, class FLUX, int DIM, int TX, int TY, int TZ, int NT, int MODE_CT = -1>
struct StageKernel {
    // the kernels compiled per stage mode index cells with 32-bit values
    // (their launchers take them only for grids of fewer than 2^29 cells)
    static constexpr bool OFF32 = MODE_CT >= 0;
    typedef typename std::conditional<OFF32, unsigned int, long long>::type Idx;
    static constexpr int NV = EQ::NV;
    static constexpr int G = ReconTraits<RECON>::GHOST;
    static constexpr int HX = G;
    static constexpr int HY = DIM >= 2 ? G : 0;
    static constexpr int HZ = DIM >= 3 ? G : 0;
    static constexpr int LX = TX + 2 * HX;
    static constexpr int LY = TY + 2 * HY;
    static constexpr int LZ = TZ + 2 * HZ;
    static constexpr int LT = LX * LY * LZ;  // tile cells incl. halo
    static constexpr int NI_X = (TX + 1) * TY * TZ;
    static constexpr int NI_Y = DIM >= 2 ? TX * (TY + 1) * TZ : 0;
    static constexpr int NI_Z = DIM >= 3 ? TX * TY * (TZ + 1) : 0;
    static constexpr int NI_MAX = NI_X > NI_Y ? (NI_X > NI_Z ? NI_X : NI_Z) : (NI_Y > NI_Z ? NI_Y : NI_Z);
    static constexpr int TC = TX * TY * TZ;
    static constexpr int TXC = TX, TYC = TY, TZC = TZ;
    static constexpr int NTHREADS = NT;
    static constexpr int CPT = (TC + NT - 1) / NT;  // cells per thread in phase 3
    static constexpr int LDS_BYTES = (NV * LT + NV * NI_MAX) * 8;

    __device__ __forceinline__ static int lidx(int cx, int cy, int cz) {
        return ((cz + HZ) * LY + (cy + HY)) * LX + (cx + HX);
    }

    // reconstruct both sides of the interface between cell ia and ia + SD
    // (LDS indices) and evaluate the numerical flux under OPS
    template <class OPS, int DIR>
    __device__ __forceinline__ static double interface(const double* su, int ia, int SD, double gamma,
                                                       double* F, bool& ok) {
        double qa[NV], qb[NV];
        recon_face_ops<RECON, NV, true, OPS>(su, LT, ia, SD, qa, ok);        // right face of a
        recon_face_ops<RECON, NV, false, OPS>(su, LT, ia + SD, SD, qb, ok);  // left face of b
        const typename EQ::All aL = EQ::template make_all_ops<OPS>(qa, gamma, ok);
        const typename EQ::All aR = EQ::template make_all_ops<OPS>(qb, gamma, ok);
        return FLUX::template flux_ops<OPS, DIR, EQ>(aL, aR, gamma, F, ok);
    }

    // cell a = (cx, cy, cz) left of interface `it` along DIR, and whether the
    // interface lies inside the grid
    template <int DIR>
    __device__ __forceinline__ static bool interface_cell(int it, int ox, int oy, int oz, const int* n, int& ia) {
        int cx, cy, cz, k;
        if (DIR == 0) {
            k = it % (TX + 1);
            const int r = it / (TX + 1);
            cy = r % TY;
            cz = r / TY;
            cx = k - 1;
        } else if (DIR == 1) {
            cx = it % TX;
            const int r = it / TX;
            k = r % (TY + 1);
            cz = r / (TY + 1);
            cy = k - 1;
        } else {
            cx = it % TX;
            cy = (it / TX) % TY;
            k = it / (TX * TY);
            cz = k - 1;
        }
        ia = lidx(cx, cy, cz);
        const int ga = (DIR == 0 ? ox + cx : (DIR == 1 ? oy + cy : oz + cz));
        return (ga < n[DIR]) && (oy + (DIR == 1 ? 0 : cy) < n[1]) && (ox + (DIR == 0 ? 0 : cx) < n[0]) &&
               (oz + (DIR == 2 ? 0 : cz) < n[2]);
    }

    // Fluxes of every interface along DIR into sf.  A FAST flux evaluates each
    // interface with shared reciprocals (SharedOps); the interfaces whose
    // operands fall outside the range where that is exact are redone with
    // PlainOps in a second loop, after the first one's registers are free.
    template <int DIR>
    __device__ __forceinline__ static void flux_pass(const double* su, double* sf, int ox, int oy, int oz,
                                                     const int* n, double gamma, double& ws) {
        constexpr int NI = DIR == 0 ? NI_X : (DIR == 1 ? NI_Y : NI_Z);
        constexpr int SD = DIR == 0 ? 1 : (DIR == 1 ? LX : LX * LY);
        static_assert((NI + NT - 1) / NT <= 32, "one redo bit per iteration");
        typedef typename OpsFor<FLUX::FAST>::type Ops;
        unsigned int redo = 0;
        int m = 0;
        for (int it = threadIdx.x; it < NI; it += NT, ++m) {
            int ia;
            if (!interface_cell<DIR>(it, ox, oy, oz, n, ia)) continue;
            double F[NV];
            bool ok = true;
            const double w = interface<Ops, DIR>(su, ia, SD, gamma, F, ok);
            if (ok) {
                ws = ws_max(ws, w);
            } else {
                redo |= 1u << m;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) sf[v * NI_MAX + it] = F[v];
        }
        if (FLUX::FAST && redo != 0) {
            m = 0;
            for (int it = threadIdx.x; it < NI; it += NT, ++m) {
                int ia;
                if (!((redo >> m) & 1u) || !interface_cell<DIR>(it, ox, oy, oz, n, ia)) continue;
                double F[NV];
                bool ok = true;
                ws = ws_max(ws, interface<PlainOps, DIR>(su, ia, SD, gamma, F, ok));
#pragma unroll
                for (int v = 0; v < NV; ++v) sf[v * NI_MAX + it] = F[v];
            }
        }
    }

    template <int DIR>
    __device__ __forceinline__ static void accumulate(const double* sf, double (&acc)[CPT][NV]) {
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int c = threadIdx.x + m * NT;
            if (TC % NT != 0 && c >= TC) continue;
            const int cx = c % TX;
            const int cy = (c / TX) % TY;
            const int cz = c / (TX * TY);
            int ip;
            if (DIR == 0) ip = (cz * TY + cy) * (TX + 1) + cx + 1;
            else if (DIR == 1) ip = (cz * (TY + 1) + cy + 1) * TX + cx;
            else ip = ((cz + 1) * TY + cy) * TX + cx;
            constexpr int STEP = DIR == 0 ? 1 : (DIR == 1 ? TX : TX * TY);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double fp = sf[v * NI_MAX + ip];
                const double fm = sf[v * NI_MAX + ip - STEP];
                // temp = -F; out[i] += temp[i] + (-1)*temp[i-1]
                acc[m][v] = acc[m][v] + ((-fp) + fm);
            }
        }
    }

    __device__ static void run(const StageParams& P) { run_tile(P, tile_id_xcd(blockIdx.x, gridDim.x)); }
    __device__ static void run_volume(const StageParams& P, const VolumeParams& V) {
        run_tile<true>(P, tile_id_xcd(blockIdx.x, gridDim.x), &V);
    }

    // number of tiles covering a grid of n cells
    __host__ __device__ static unsigned int tiles(const int* n) {
        return (unsigned int)(((n[0] + TX - 1) / TX) * ((n[1] + TY - 1) / TY) * ((n[2] + TZ - 1) / TZ));
    }

    // VOL: the cells live in a reference Volume (V) and P.n is the extent of
    // the box the launch covers; otherwise in the interior-only buffers of P
    // with the boundary map for the halo.
    template <bool VOL = false>
    __device__ static void run_tile(const StageParams& P, const unsigned int tile, const VolumeParams* V = nullptr) {
        __shared__ double lds[NV * LT + NV * NI_MAX];
        __shared__ double s_ws[3][NT / 64];
        double* su = lds;
        double* sf = lds + NV * LT;

        Control* ctrl = P.ctrl;
        const int mode = MODE_CT >= 0 ? MODE_CT : P.mode;
        if (!ctrl->active && mode != MODE_FLUX_ONLY) return;  // step past the end: no-op

        const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
        const int ntx = (n0 + TX - 1) / TX, nty = (n1 + TY - 1) / TY;
        const int tx = tile % ntx;
        const int ty = (tile / ntx) % nty;
        const int tz = tile / (ntx * nty);
        const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
        const long long nc = P.ncells;
        const bool otf = (mode == MODE_RK3_2);
        const double s = (mode == MODE_FLUX_ONLY || mode == MODE_L) ? 0.0 : step_dt(ctrl) / ctrl->cell[0];

        // 1. stage input tile + halo into LDS through the boundary map.  Each
        // thread issues the loads of all its points before writing any of
        // them, so that their global-memory latencies overlap; points past
        // the tile repeat the last one and are not written.
        constexpr int SPT = (LT + NT - 1) / NT;  // staged points per thread
        Idx sg[SPT];
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            int p = threadIdx.x + k * NT;
            p = p < LT ? p : LT - 1;
            const int lx = p % LX;
            const int ly = (p / LX) % LY;
            const int lz = p / (LX * LY);
            int gx = ox + lx - HX, gy = oy + ly - HY, gz = oz + lz - HZ;
            gx = gx < n0 + HX ? gx : n0 + HX - 1;
            gy = gy < n1 + HY ? gy : n1 + HY - 1;
            gz = gz < n2 + HZ ? gz : n2 + HZ - 1;
            if constexpr (VOL) {
                sg[k] = volume_index(*V, gx, DIM >= 2 ? gy : 0, DIM >= 3 ? gz : 0);
            } else {
                gx = bc_map(gx, n0, P.bc);
                gy = DIM >= 2 ? bc_map(gy, n1, P.bc) : 0;
                gz = DIM >= 3 ? bc_map(gz, n2, P.bc) : 0;
                sg[k] = ((Idx)gz * n1 + gy) * n0 + gx;
            }
        }
        double sv[SPT][NV];
#pragma unroll
        for (int k = 0; k < SPT; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (VOL) sv[k][v] = V->in[v][sg[k]];
                else sv[k][v] = soa_ld<OFF32>(P.in, nc, v, sg[k]);
            }
        if (otf) {
#pragma unroll
            for (int k = 0; k < SPT; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) sv[k][v] = sv[k][v] * s + soa_ld<OFF32>(P.u0, nc, v, sg[k]);
        }
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int p = threadIdx.x + k * NT;
            if (LT % NT != 0 && p >= LT) continue;
#pragma unroll
            for (int v = 0; v < NV; ++v) su[v * LT + p] = sv[k][v];
        }
        __syncthreads();

        double acc[CPT][NV];
#pragma unroll
        for (int m = 0; m < CPT; ++m)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[m][v] = 0.0;
        double ws[3] = {0.0, 0.0, 0.0};

        // 2./3. per direction: interface fluxes into LDS, then accumulate
        flux_pass<0>(su, sf, ox, oy, oz, P.n, P.gamma, ws[0]);
        __syncthreads();
        accumulate<0>(sf, acc);
        if (DIM >= 2) {
            __syncthreads();
            flux_pass<DIM >= 2 ? 1 : 0>(su, sf, ox, oy, oz, P.n, P.gamma, ws[1]);
            __syncthreads();
            accumulate<DIM >= 2 ? 1 : 0>(sf, acc);
        }
        if (DIM >= 3) {
            __syncthreads();
            flux_pass<DIM >= 3 ? 2 : 0>(su, sf, ox, oy, oz, P.n, P.gamma, ws[2]);
            __syncthreads();
            accumulate<DIM >= 3 ? 2 : 0>(sf, acc);
        }

        // 4. RK combination and write-back; the U0 loads of all the thread's
        // cells are issued first (cells outside the grid load element 0 and
        // write nothing)
        bool bad = false;
        const bool final_stage = (mode == MODE_RK2_2 || mode == MODE_RK3_3);
        const bool combine = !(mode == MODE_FLUX_ONLY || mode == MODE_L);
        Idx gc[CPT];
        bool inside[CPT];
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int c = threadIdx.x + m * NT;
            const int cx = c % TX;
            const int cy = (c / TX) % TY;
            const int cz = c / (TX * TY);
            const int gx = ox + cx, gy = oy + cy, gz = oz + cz;
            inside[m] = (TC % NT == 0 || c < TC) && gx < n0 && gy < n1 && gz < n2;
            if constexpr (VOL) gc[m] = inside[m] ? volume_index(*V, gx, gy, gz) : 0;
            else gc[m] = inside[m] ? ((Idx)gz * n1 + gy) * n0 + gx : 0;
        }
        const double ca = mode == MODE_RK3_2 ? 1. / 3. : (mode == MODE_RK3_3 ? 2.0 : 1.0);
        const double cb = mode == MODE_RK2_2 ? 0.5 : (mode == MODE_RK3_2 ? 3. / 4. : 1. / 3.);
        double u0c[CPT][NV];
        if (combine) {
#pragma unroll
            for (int m = 0; m < CPT; ++m)
#pragma unroll
                for (int v = 0; v < NV; ++v) u0c[m][v] = soa_ld<OFF32>(P.u0, nc, v, gc[m]);
        }
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            if (!inside[m]) continue;
            const int c = threadIdx.x + m * NT;
            const int cx = c % TX;
            const int cy = (c / TX) % TY;
            const int cz = c / (TX * TY);
            const Idx gi = gc[m];
            const int li = lidx(cx, cy, cz);
            double r[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double l = acc[m][v];
                if (!combine) {
                    r[v] = l;
                } else {
                    // ((L*s + U) * ca + U0) * cb; the factor 1 of the RK2
                    // stage is exact, so every mode keeps its own rounding
                    const double us = su[v * LT + li];
                    double o = l * s;
                    o = o + us;
                    o = o * ca;
                    o = o + u0c[m][v];
                    o = o * cb;
                    r[v] = o;
                }
            }
            if (final_stage) bad = bad || !EQ::obeys_fast(r, P.gamma);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (VOL) V->out[v][gi] = r[v];
                else soa_st<OFF32>(P.out, nc, v, gi, r[v]);
            }
        }

        if (mode == MODE_FLUX_ONLY || mode == MODE_L) {
            // wave speed: wavefront reduction, then one atomicMax per direction
            // per workgroup (non-negative doubles order like their bit patterns)
            const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double w = wave_max(ws[d]);
                if (lane == 0) s_ws[d][wid] = w;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                double w = 0.0;
#pragma unroll
                for (int i = 0; i < NT / 64; ++i) w = ws_max(w, s_ws[threadIdx.x][i]);
                if (w > 0.0)
                    atomicMax(&ctrl->ws_bits[threadIdx.x], (unsigned long long)__double_as_longlong(w));
            }
        } else if (final_stage) {
            const int any_bad = __syncthreads_or(bad);
            if (threadIdx.x == 0) {
                if (P.finish) arrive_final(ctrl, any_bad != 0);
                else if (any_bad) atomicOr(&ctrl->block_violation, 1);
            }
        }
    }
};

template <class K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void stage_kernel(StageParams P) {
    K::run(P);
}

// A stage over a box of a reference Volume (flux-only mode)
template <class K>
__global__ __launch_bounds__(256) void volume_stage_kernel(StageParams P, VolumeParams V) {
    K::run_volume(P, V);
}

}  // namespace avh
