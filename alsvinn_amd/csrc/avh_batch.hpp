// Batched Monte Carlo: every sample of every resolution group advances in the
// same launches.
//
// A stage launch covers all groups with 8 * per_xcd blocks.  Workgroups are
// dealt to the 8 XCDs round-robin (block b runs on XCD b % 8), and XCD x
// takes every 8th sample -- samples x, x + 8, x + 16, ... of every group --
// tile by tile: the tiles of a sample share one L2, and every XCD holds an
// equal share of every resolution, so a group that finishes early leaves no
// XCD idle.  Element-wise launches run blocks [first_block, first_block +
// nsamples * tiles) per group, sample-major.
// Each sample owns a Control block, so dt, time, save counters, the
// constraint flag and the loop test are per sample, exactly as the
// reference's Runner advances each sample with its own Simulator
// (alsuq/src/run/Runner.cpp:39-90).  Finished samples turn into no-ops.
//
// At a statistics save time (FixedIntervalStatistics,
// alsuq/src/stats/FixedIntervalStatistics.cpp:64-86) the end-of-step kernel
// lists the samples that reached it, in sample order, and the statistics
// kernel folds them into per-save mean and M2 arrays (Welford), for the
// conserved variables (MeanVariance, alsuq/src/stats/MeanVariance.cpp:30-67)
// and for rho, u, p.
#pragma once
#include "avh_stage.hpp"

#pragma clang fp contract(off)

namespace avh {

constexpr int MAX_GROUPS = 4;

struct GroupDesc {
    const double* in;
    const double* u0;
    double* out;
    Control* ctrl;             // nsamples consecutive Control blocks
    long long stride;          // doubles from one sample's state to the next
    long long ncells;
    int n[3];
    unsigned int tiles;        // tiles per sample
    unsigned int first_block;  // first block of this group in a launch
    int nsamples;
};

struct BatchParams {
    GroupDesc g[MAX_GROUPS];
    int ngroups;
    int bc;
    int mode;
    double gamma;
    unsigned int per_xcd;  // stage blocks per XCD
};

// samples s of a group of n with s % 8 == x
__host__ __device__ __forceinline__ unsigned int xcd_samples(int n, unsigned int x) {
    return (unsigned int)((n + 7 - (int)x) / 8);
}

// (group, sample, tile) of stage block b; false for the padding blocks of an
// XCD whose share is below per_xcd
__device__ __forceinline__ bool batch_stage_decode(const BatchParams& B, unsigned int b, int& gi,
                                                   unsigned int& sample, unsigned int& tile) {
    const unsigned int x = b % 8u;
    unsigned int k = b / 8u;
#pragma unroll
    for (int i = 0; i < MAX_GROUPS; ++i) {
        if (i >= B.ngroups) break;
        const unsigned int cnt = xcd_samples(B.g[i].nsamples, x) * B.g[i].tiles;
        if (k < cnt) {
            gi = i;
            sample = (k / B.g[i].tiles) * 8u + x;
            tile = k % B.g[i].tiles;
            return true;
        }
        k -= cnt;
    }
    return false;
}

__device__ __forceinline__ int batch_group(const BatchParams& B, unsigned int b) {
    int gi = 0;
#pragma unroll
    for (int i = 1; i < MAX_GROUPS; ++i)
        if (i < B.ngroups && b >= B.g[i].first_block) gi = i;
    return gi;
}

template <class K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void batch_stage_kernel(BatchParams B) {
    int gi = 0;
    unsigned int sample = 0, tile = 0;
    if (!batch_stage_decode(B, blockIdx.x, gi, sample, tile)) return;
    const GroupDesc& G = B.g[gi];
    StageParams P;
    P.in = G.in + (long long)sample * G.stride;
    P.u0 = G.u0 + (long long)sample * G.stride;
    P.out = G.out + (long long)sample * G.stride;
    P.ctrl = G.ctrl + sample;
    P.ncells = G.ncells;
    P.n[0] = G.n[0];
    P.n[1] = G.n[1];
    P.n[2] = G.n[2];
    P.bc = B.bc;
    P.mode = B.mode;
    P.gamma = B.gamma;
    K::run_tile(P, tile);
}

// Element-wise out = L*s + U0 for every sample (forward euler's finish with
// the constraint check, or rk2's U1); blocks as in the stage launch but
// 256 cells each.
template <class EQ>
__global__ __launch_bounds__(256) void batch_ew_kernel(BatchParams B, int fe_finish) {
    const unsigned int b = blockIdx.x;
    const int gi = batch_group(B, b);
    const GroupDesc& G = B.g[gi];
    const unsigned int local = b - G.first_block;
    const unsigned int sample = local / G.tiles;
    const long long i = (long long)(local % G.tiles) * 256 + threadIdx.x;
    Control* ctrl = G.ctrl + sample;
    if (!ctrl->active) return;
    constexpr int NV = EQ::NV;
    const double s = step_dt(ctrl) / ctrl->cell[0];
    bool bad = false;
    if (i < G.ncells) {
        const long long o = (long long)sample * G.stride;
        double r[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double x = G.in[o + v * G.ncells + i] * s;
            x = x + G.u0[o + v * G.ncells + i];
            r[v] = x;
        }
        if (fe_finish) bad = !EQ::obeys(r, B.gamma);
#pragma unroll
        for (int v = 0; v < NV; ++v) G.out[o + v * G.ncells + i] = r[v];
    }
    if (fe_finish) {
        if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&ctrl->block_violation, 1);
    }
}

// One sample's end of step (same bookkeeping as finish_step_kernel).
// Returns the statistics save index this step reached, or -1.
__device__ __forceinline__ int finish_sample(Control* c, int stats_adj) {
    if (!c->active) return -1;
    const double dt = step_dt(c);
    c->dt = dt;
    c->scale = dt / c->cell[0];
    c->t = c->t + dt;
    c->steps = c->steps + 1;
    c->last_dt = dt;
    int hit = -1;
    if (c->block_violation) {
        c->violated = 1;
    } else {
        for (int i = 0; i < c->n_adj; ++i) {
            if (c->t >= (double)c->adj_saved[i] * c->adj_interval[i] + c->adj_start[i]) {
                if (i == stats_adj) hit = (int)c->adj_saved[i];
                c->adj_saved[i] = c->adj_saved[i] + 1;
                c->saves_pending = c->saves_pending + 1;
            }
        }
    }
    c->block_violation = 0;
    clear_wave_speeds(c);
    c->active = (!(c->t >= c->t_end) && !c->violated) ? 1 : 0;
    return hit;
}

struct StatsGroup {
    Control* ctrl;
    const double* state;  // sample-major states (the solution buffer)
    long long stride, ncells;
    int nsamples;
    int nsaves;
    int* list;            // (sample, save index, count after the update) triples
    int* list_len;
    long long* counts;    // per save index: samples folded in so far
    int* active_count;
    double* mean_c;       // [nsaves][nv][ncells]
    double* m2_c;
    double* mean_p;       // rho, u..., p
    double* m2_p;
    unsigned int blocks;       // stats blocks of this group
    unsigned int first_block;
};

struct StatsParams {
    StatsGroup g[MAX_GROUPS];
    int ngroups;
    int stats_adj;  // adjuster whose saves record statistics, -1 for none
    int save_offset;  // adjuster save number of statistics save 0: 1, or 0 with an initial save
    double gamma;
};

// Welford fold of the listed samples, one thread per cell, list order.
template <class EQ>
__global__ __launch_bounds__(256) void batch_stats_kernel(StatsParams S) {
    constexpr int NV = EQ::NV;
    const unsigned int b = blockIdx.x;
    int gi = 0;
#pragma unroll
    for (int i = 1; i < MAX_GROUPS; ++i)
        if (i < S.ngroups && b >= S.g[i].first_block) gi = i;
    const StatsGroup& G = S.g[gi];
    const int len = *G.list_len;
    if (len == 0) return;
    const long long i = (long long)(b - G.first_block) * 256 + threadIdx.x;
    if (i >= G.ncells) return;
    const long long nc = G.ncells;
    for (int j = 0; j < len; ++j) {
        const int s = G.list[3 * j], k = G.list[3 * j + 1], cnt = G.list[3 * j + 2];
        if (k >= G.nsaves || cnt <= 0) continue;
        const double n = (double)cnt;
        double q[NV], p[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = G.state[(long long)s * G.stride + v * nc + i];
        if constexpr (EQ::NV > 1) {
            // rho, u = m/rho, p = (gamma-1)(E - 0.5 m.m/rho) (Euler::computeExtra)
            const typename EQ::All a = EQ::make_all(q, S.gamma);
            p[0] = q[0];
#pragma unroll
            for (int v = 1; v < NV - 1; ++v) p[v] = a.u[v - 1];
            p[NV - 1] = a.p;
        } else {
            p[0] = q[0];
        }
        const long long base = (long long)k * NV * nc + i;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long long o = base + v * nc;
            double m = G.mean_c[o];
            const double d = q[v] - m;
            m = m + d / n;
            G.mean_c[o] = m;
            G.m2_c[o] = G.m2_c[o] + d * (q[v] - m);
            double mp = G.mean_p[o];
            const double dp = p[v] - mp;
            mp = mp + dp / n;
            G.mean_p[o] = mp;
            G.m2_p[o] = G.m2_p[o] + dp * (p[v] - mp);
        }
    }
}

}  // namespace avh
