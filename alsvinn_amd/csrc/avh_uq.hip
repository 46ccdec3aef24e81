// Batched Monte Carlo UQ on the hip platform: sample generation (the
// reference's STLMersenne + Uniform), batched solver groups, end-of-step
// bookkeeping and on-device Welford statistics (C ABI in alsvinn_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <new>
#include <random>
#include <vector>

#include "avh_internal.hpp"

#pragma clang fp contract(off)

using namespace avh;

namespace avh {

// One block per group: every sample's end of step, then a stable (sample
// order) compaction of the samples that reached a statistics save, each with
// its Welford count.
__global__ __launch_bounds__(1024) void batch_finish_kernel(StatsParams S) {
    __shared__ int s_wave[16];
    __shared__ int s_base;
    __shared__ int s_active;
    const StatsGroup& G = S.g[blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (threadIdx.x == 0) {
        s_base = 0;
        s_active = 0;
    }
    __syncthreads();
    for (int base = 0; base < G.nsamples; base += blockDim.x) {
        const int s = base + threadIdx.x;
        int hit = -1;
        if (s < G.nsamples) {
            hit = finish_sample(G.ctrl + s, S.stats_adj);
            if (hit >= 0) hit -= S.save_offset;
            if (G.ctrl[s].active) atomicAdd(&s_active, 1);
        }
        const bool flag = hit >= 0;
        const unsigned long long ball = __ballot(flag);
        const int before = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) s_wave[wid] = __popcll(ball);
        __syncthreads();
        int off = s_base;
        for (int w = 0; w < wid; ++w) off += s_wave[w];
        if (flag) {
            G.list[3 * (off + before) + 0] = s;
            G.list[3 * (off + before) + 1] = hit;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int w = 0; w < nw; ++w) tot += s_wave[w];
            s_base += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // Welford counts, in list order
        for (int j = 0; j < s_base; ++j) {
            const int k = G.list[3 * j + 1];
            const long long nk = (k < G.nsaves) ? ++G.counts[k] : 0;
            G.list[3 * j + 2] = (int)nk;
        }
        *G.list_len = s_base;
        *G.active_count = s_active;
    }
}


// Statistics of the initial state (writeInitialTimestep): every sample at
// save index 0, in sample order.
__global__ __launch_bounds__(256) void batch_initial_save_kernel(StatsParams S) {
    const StatsGroup& G = S.g[blockIdx.x];
    for (int s = threadIdx.x; s < G.nsamples; s += blockDim.x) {
        G.list[3 * s + 0] = s;
        G.list[3 * s + 1] = 0;
        G.list[3 * s + 2] = s + 1;
        if (S.stats_adj >= 0) G.ctrl[s].adj_saved[S.stats_adj] = 1;
    }
    if (threadIdx.x == 0) {
        *G.list_len = G.nsamples;
        if (G.nsaves > 0) G.counts[0] = G.nsamples;
    }
}

// 2D Kelvin-Helmholtz fill from per-sample band interfaces lo[x] < y < hi[x]
// at cell midpoints: rho = 2, ux = -0.5 inside, rho = 1, ux = 0.5 outside,
// uy = 0, p = 2.5; conserved with m = u*rho, E = p/(gamma-1) + 0.5*rho*(u.u)
// (alsfvm/include/alsfvm/equation/euler/Euler.hpp:236-245).
__global__ __launch_bounds__(256) void kh2d_fill_kernel(double* state, long long stride, int nx, int ny,
                                                        const double* iface, int nsamples, double lower_y,
                                                        double dy, double gamma) {
    const long long nc = (long long)nx * ny;
    const long long total = nc * nsamples;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(t / nc);
        const long long i = t % nc;
        const int x = (int)(i % nx), y = (int)(i / nx);
        const double ym = lower_y + dy * y + dy / 2.0;
        const double lo = iface[(long long)s * 2 * nx + x];
        const double hi = iface[(long long)s * 2 * nx + nx + x];
        const bool band = (ym > lo) && (ym < hi);
        const double rho = band ? 2.0 : 1.0;
        const double ux = band ? -0.5 : 0.5;
        const double uy = 0.0;
        const double p = 2.5;
        double dot = ux * ux;
        dot = dot + uy * uy;
        double* q = state + (long long)s * stride;
        q[i] = rho;
        q[nc + i] = ux * rho;
        q[2 * nc + i] = uy * rho;
        q[3 * nc + i] = p / (gamma - 1) + 0.5 * rho * dot;
    }
}


// Kinetic energy of one velocity component per sample,
//   E_c = (sum over cells of 0.5 * rho * u_c * u_c) * dV,  u_c = m_c / rho,
// a functional of the conserved state modelled on the Moment functional
// (alsfvm/src/functional/Moment.cpp:31-45).  One block per sample; each
// thread sums a strided subset in cell order, then a fixed tree, so the
// result does not depend on timing.
__global__ __launch_bounds__(256) void kinetic_energy_kernel(const double* state, long long stride, long long ncells,
                                                             int comp, double dv, double* out) {
    __shared__ double red[256];
    const double* q = state + (long long)blockIdx.x * stride;
    double acc = 0.0;
    for (long long i = threadIdx.x; i < ncells; i += 256) {
        const double rho = q[i];
        const double u = q[(long long)(1 + comp) * ncells + i] / rho;
        acc = acc + 0.5 * rho * (u * u);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0] * dv;
}

template <int P>
__device__ __forceinline__ double fast_power(double x) {
    if constexpr (P == 1) {
        return x;
    } else {
        return x * fast_power<P - 1>(x);
    }
}

__device__ __forceinline__ int bc_value_index(int x, int n, int bc) {
    // ValueAtBoundary: periodic wraps, neumann clamps to the interior
    // (alsfvm/include/alsfvm/boundary/ValueAtBoundary.hpp)
    if (bc == AVH_BC_PERIODIC) {
        while (x < 0) x += n;
        return x % n;
    }
    return x < 0 ? 0 : (x > n - 1 ? n - 1 : x);
}

// Structure function of one variable per sample (the structure_cube
// statistic, alsfvm/include/alsfvm/functional/structure_common.hpp:36-154):
// for every cell and lag h, |u - u'|^p summed over the cells u' on the
// boundary of the square of half-width h around it (2 points in 1D, 8h in
// 2D), divided by the number of cells.  Block = 256 consecutive cells of one
// sample at one lag; partial sums in fixed order.
template <int P>
__global__ __launch_bounds__(256) void structure_partial_kernel(const double* state, long long stride, int var, int nx,
                                                                int ny, int dims, int bc, double p, double* partial) {
    __shared__ double red[256];
    const long long ncells = (long long)nx * ny;
    const int h = blockIdx.y + 1;
    const double* u = state + (long long)blockIdx.z * stride + (long long)var * ncells;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (c < ncells) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        const double uc = u[c];
        auto term = [&](int x, int y) {
            const double d = fabs(uc - u[(long long)bc_value_index(y, ny, bc) * nx + bc_value_index(x, nx, bc)]);
            if constexpr (P > 0) {
                acc = acc + fast_power<P>(d);
            } else {
                acc = acc + pow(d, p);
            }
        };
        if (dims == 1) {
            term(i - h, 0);
            term(i + h, 0);
        } else {
            // d = 0: columns x = i -/+ h, y in [j-h+1, j+h-1]
            for (int side = 0; side < 2; ++side) {
                const int x = side == 0 ? i - h : i + h;
                for (int y = j - h + 1; y < j + h; ++y) term(x, y);
            }
            // d = 1: rows y = j -/+ h, x in [i-h, i+h]
            for (int side = 0; side < 2; ++side) {
                const int y = side == 0 ? j - h : j + h;
                for (int x = i - h; x < i + h + 1; ++x) term(x, y);
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// out[s][0] = 0, out[s][h] = (sum of the partials of (s, h) in block order) / cells
__global__ __launch_bounds__(256) void structure_finish_kernel(const double* partial, int nchunks, int nh, int nsamples,
                                                               double ncells, double* out) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nsamples * nh) return;
    const int s = t / nh, h = t % nh;
    if (h == 0) {
        out[t] = 0.0;
        return;
    }
    const double* q = partial + ((long long)s * (nh - 1) + (h - 1)) * nchunks;
    double acc = 0.0;
    for (int k = 0; k < nchunks; ++k) acc = acc + q[k];
    out[t] = acc / ncells;
}

// End time of every sample: samples with t < t_end (and no constraint
// failure) run again.
__global__ __launch_bounds__(256) void set_end_time_kernel(Control* ctrl, int nsamples, double t_end, int* active_count) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < nsamples; s += 256) {
        Control* c = ctrl + s;
        c->t_end = t_end;
        c->active = (!(c->t >= t_end) && !c->violated) ? 1 : 0;
        if (c->active) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *active_count = cnt;
}

}  // namespace avh

namespace {

struct GroupHost {
    int n[3];
    int nsamples;
    long long ncells;
    double *A = nullptr, *B = nullptr, *C = nullptr;
    Control* ctrl = nullptr;
    double *mean_c = nullptr, *m2_c = nullptr, *mean_p = nullptr, *m2_p = nullptr;
    int* list = nullptr;
    int* list_len = nullptr;
    long long* counts = nullptr;
    int* active_count = nullptr;
};

template <class EQ>
void launch_batch_ew(const BatchParams& B, unsigned int nblocks, int fe_finish, hipStream_t st) {
    hipLaunchKernelGGL(batch_ew_kernel<EQ>, dim3(nblocks), dim3(256), 0, st, B, fe_finish);
}
template <class EQ>
void launch_batch_stats(const StatsParams& S, unsigned int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(batch_stats_kernel<EQ>, dim3(nblocks), dim3(256), 0, st, S);
}
typedef void (*BatchEw)(const BatchParams&, unsigned int, int, hipStream_t);
typedef void (*BatchStats)(const StatsParams&, unsigned int, hipStream_t);

BatchEw batch_ew_for(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return &launch_batch_ew<BurgersEq>;
    case AVH_EQ_EULER1: return &launch_batch_ew<EulerEq<1>>;
    case AVH_EQ_EULER2: return &launch_batch_ew<EulerEq<2>>;
    case AVH_EQ_EULER3: return &launch_batch_ew<EulerEq<3>>;
    }
    return nullptr;
}
BatchStats batch_stats_for(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return &launch_batch_stats<BurgersEq>;
    case AVH_EQ_EULER1: return &launch_batch_stats<EulerEq<1>>;
    case AVH_EQ_EULER2: return &launch_batch_stats<EulerEq<2>>;
    case AVH_EQ_EULER3: return &launch_batch_stats<EulerEq<3>>;
    }
    return nullptr;
}

}  // namespace

struct avh_batch {
    avh_ctx* ctx = nullptr;
    avh_config cfg{};
    int dim = 1, nv = 0;
    int ngroups = 0;
    GroupHost g[MAX_GROUPS];
    int nsaves = 0;         // statistics saves allocated
    int stats_adj = -1;     // adjuster index recording statistics
    int stats_initial = 0;  // record the initial state as save 0
    int initial_done = 0;
    Control proto{};        // per-sample initial Control
    hipStream_t stream = nullptr;
    BatchLauncher stage = nullptr;
    BatchEw ew = nullptr;
    BatchStats stats = nullptr;
    int graph_steps = 8;  // steps per captured hipGraph (0: ordinary launches)
    StepGraph graph;
    int shape = -1;       // tile shape of the stage kernel; -1: autotune on first use
    // groups the stage and element-wise launches cover: a group whose
    // samples have all stopped is dropped until the next avh_batch_advance
    bool live[MAX_GROUPS] = {true, true, true, true};
};

namespace {

void free_batch(avh_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->graph.reset();
    for (int i = 0; i < b->ngroups; ++i) {
        GroupHost& G = b->g[i];
        (void)hipFree(G.A);
        (void)hipFree(G.B);
        (void)hipFree(G.C);
        (void)hipFree(G.ctrl);
        (void)hipFree(G.mean_c);
        (void)hipFree(G.m2_c);
        (void)hipFree(G.mean_p);
        (void)hipFree(G.m2_p);
        (void)hipFree(G.list);
        (void)hipFree(G.list_len);
        (void)hipFree(G.counts);
        (void)hipFree(G.active_count);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Launch parameters over the live groups (slot j of P is group index[j]);
// nblocks = the launch's block count (8 * per_xcd for a stage launch, the
// element-wise blocks of every sample otherwise), 0 if no group is live.
BatchParams batch_params(const avh_batch* b, int mode, bool ew_blocks, unsigned int* nblocks,
                         int* index = nullptr) {
    BatchParams P;
    std::memset(&P, 0, sizeof P);
    P.bc = b->cfg.boundary;
    P.mode = mode;
    P.gamma = b->cfg.gamma;
    unsigned int first = 0;
    unsigned int per_xcd[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int j = 0;
    for (int i = 0; i < b->ngroups; ++i) {
        if (!b->live[i]) continue;
        const GroupHost& G = b->g[i];
        GroupDesc& D = P.g[j];
        if (index) index[j] = i;
        ++j;
        D.ctrl = G.ctrl;
        D.stride = (long long)b->nv * G.ncells;
        D.ncells = G.ncells;
        for (int d = 0; d < 3; ++d) D.n[d] = G.n[d];
        D.tiles = ew_blocks ? (unsigned int)((G.ncells + 255) / 256) : stage_tiles(G.n, b->dim, b->shape);
        D.first_block = first;
        D.nsamples = G.nsamples;
        first += D.tiles * (unsigned int)G.nsamples;
        for (unsigned int x = 0; x < 8; ++x) per_xcd[x] += xcd_samples(G.nsamples, x) * D.tiles;
    }
    P.ngroups = j;
    P.per_xcd = 0;
    for (unsigned int x = 0; x < 8; ++x) P.per_xcd = std::max(P.per_xcd, per_xcd[x]);
    *nblocks = ew_blocks ? first : 8u * P.per_xcd;
    return P;
}

StatsParams stats_params(const avh_batch* b, unsigned int* nblocks) {
    StatsParams S;
    std::memset(&S, 0, sizeof S);
    S.ngroups = b->ngroups;
    S.stats_adj = b->stats_adj;
    S.save_offset = b->stats_initial ? 0 : 1;
    S.gamma = b->cfg.gamma;
    unsigned int first = 0;
    for (int i = 0; i < b->ngroups; ++i) {
        const GroupHost& G = b->g[i];
        StatsGroup& T = S.g[i];
        T.ctrl = G.ctrl;
        T.state = G.A;
        T.stride = (long long)b->nv * G.ncells;
        T.ncells = G.ncells;
        T.nsamples = G.nsamples;
        T.nsaves = b->nsaves;
        T.list = G.list;
        T.list_len = G.list_len;
        T.counts = G.counts;
        T.active_count = G.active_count;
        T.mean_c = G.mean_c;
        T.m2_c = G.m2_c;
        T.mean_p = G.mean_p;
        T.m2_p = G.m2_p;
        T.blocks = (unsigned int)((G.ncells + 255) / 256);
        T.first_block = first;
        first += T.blocks;
    }
    *nblocks = first;
    return S;
}

TuneKey batch_tune_key(const avh_batch* b) {
    long long total = 0;
    for (int i = 0; i < b->ngroups; ++i) total += b->g[i].nsamples;
    return TuneKey(1, b->cfg.equation * 16 + b->cfg.reconstruction, b->cfg.flux, b->dim, b->g[0].n[0],
                   b->g[0].n[1], total);
}

// Times every tile shape with flux-only launches over all samples (they write
// the L0 buffers and add to the per-sample wave-speed maxima; the control
// blocks are restored afterwards) and keeps the fastest.
void set_all_live(avh_batch* b);

avh_status batch_autotune(avh_batch* b, int* best, double* ms, int* nshapes) {
    const int ns = num_shapes(b->dim);
    set_all_live(b);  // time the launches over every group
    AVH_HIP(hipStreamSynchronize(b->stream));
    std::vector<std::vector<Control>> saved(b->ngroups);
    for (int i = 0; i < b->ngroups; ++i) {
        saved[i].resize(b->g[i].nsamples);
        AVH_HIP(hipMemcpy(saved[i].data(), b->g[i].ctrl, sizeof(Control) * b->g[i].nsamples,
                          hipMemcpyDeviceToHost));
    }
    const int keep = b->shape;
    int bst = 0;
    double tmin = 0;
    avh_status r = AVH_OK;
    for (int k = 0; k < ns && r == AVH_OK; ++k) {
        double t = 0;
        if (ns > 1) {
            b->shape = k;
            unsigned int nb = 0;
            BatchParams P = batch_params(b, MODE_FLUX_ONLY, false, &nb);
            for (int i = 0; i < b->ngroups; ++i) {
                P.g[i].in = b->g[i].A;
                P.g[i].u0 = b->g[i].A;
                P.g[i].out = b->g[i].B;
            }
            r = time_launches(b->stream, 3, [&]() { return b->stage(P, nb, k, b->stream); }, &t);
        }
        if (ms) ms[k] = t;
        if (k == 0 || t < tmin) {
            tmin = t;
            bst = k;
        }
    }
    b->shape = keep;
    AVH_HIP(hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->ngroups; ++i)
        AVH_HIP(hipMemcpy(b->g[i].ctrl, saved[i].data(), sizeof(Control) * b->g[i].nsamples,
                          hipMemcpyHostToDevice));
    if (r) return r;
    if (b->shape != bst) b->graph.reset();
    b->shape = bst;
    tune_store(batch_tune_key(b), bst);
    if (best) *best = bst;
    if (nshapes) *nshapes = ns;
    return AVH_OK;
}

avh_status batch_ensure_shape(avh_batch* b) {
    if (b->shape >= 0) return AVH_OK;
    int s;
    if (tune_lookup(batch_tune_key(b), &s)) {
        b->shape = s;
        return AVH_OK;
    }
    return batch_autotune(b, nullptr, nullptr, nullptr);
}

// Launches per batch step: the stages (plus forward euler's / rk2's
// element-wise pass), the end-of-step kernel and, with statistics, the
// statistics kernel.
int batch_launches_per_step(const avh_batch* b) {
    const int stages = b->cfg.integrator == AVH_INT_FORWARDEULER ? 2 : 3;
    return stages + 1 + (b->nsaves > 0 ? 1 : 0);
}

void set_all_live(avh_batch* b) {
    bool changed = false;
    for (int i = 0; i < b->ngroups; ++i) {
        changed = changed || !b->live[i];
        b->live[i] = true;
    }
    if (changed) b->graph.reset();
}

avh_status enqueue_batch_step(avh_batch* b, EventMarks* tm = nullptr) {
    avh_status r;
    if ((r = mark(tm, b->stream))) return r;
    unsigned int nb = 0, nbe = 0;
    int index[MAX_GROUPS] = {0, 0, 0, 0};
    auto set_ptrs = [&](BatchParams& P, int which_in, int which_u0, int which_out) {
        for (int j = 0; j < P.ngroups; ++j) {
            const GroupHost& G = b->g[index[j]];
            double* bufs[3] = {G.A, G.B, G.C};
            P.g[j].in = bufs[which_in];
            P.g[j].u0 = bufs[which_u0];
            P.g[j].out = bufs[which_out];
        }
    };
    // a launch with no live group is left out (every sample has stopped;
    // the end-of-step kernel then leaves them unchanged)
    auto stage = [&](BatchParams& P) -> avh_status {
        if (nb) AVH_HIP(b->stage(P, nb, b->shape, b->stream));
        return mark(tm, b->stream);
    };
    auto ew = [&](BatchParams& E, int fe_finish) -> avh_status {
        if (nbe) {
            b->ew(E, nbe, fe_finish, b->stream);
            AVH_HIP(hipGetLastError());
        }
        return mark(tm, b->stream);
    };
    enum { A = 0, B = 1, C = 2 };
    BatchParams P = batch_params(b, MODE_L, false, &nb, index);
    set_ptrs(P, A, A, B);
    if ((r = stage(P))) return r;
    if (b->cfg.integrator == AVH_INT_FORWARDEULER) {
        BatchParams E = batch_params(b, MODE_L, true, &nbe, index);
        set_ptrs(E, B, A, A);
        if ((r = ew(E, 1))) return r;
    } else if (b->cfg.integrator == AVH_INT_RUNGEKUTTA2) {
        BatchParams E = batch_params(b, MODE_L, true, &nbe, index);
        set_ptrs(E, B, A, C);
        if ((r = ew(E, 0))) return r;
        P = batch_params(b, MODE_RK2_2, false, &nb, index);
        set_ptrs(P, C, A, A);
        if ((r = stage(P))) return r;
    } else {
        P = batch_params(b, MODE_RK3_2, false, &nb, index);
        set_ptrs(P, B, A, C);
        if ((r = stage(P))) return r;
        P = batch_params(b, MODE_RK3_3, false, &nb, index);
        set_ptrs(P, C, A, A);
        if ((r = stage(P))) return r;
    }
    unsigned int nbs = 0;
    StatsParams S = stats_params(b, &nbs);
    int maxs = 64;
    for (int i = 0; i < b->ngroups; ++i) maxs = std::max(maxs, std::min(1024, (b->g[i].nsamples + 63) / 64 * 64));
    hipLaunchKernelGGL(batch_finish_kernel, dim3(b->ngroups), dim3(maxs), 0, b->stream, S);
    AVH_HIP(hipGetLastError());
    if ((r = mark(tm, b->stream))) return r;
    if (b->nsaves > 0) {
        b->stats(S, nbs, b->stream);
        AVH_HIP(hipGetLastError());
        if ((r = mark(tm, b->stream))) return r;
    }
    return AVH_OK;
}

}  // namespace

extern "C" {

avh_status avh_batch_create(avh_ctx* ctx, const avh_config* cfg, int ngroups, const avh_group* groups,
                            double t_end, int stats_saves, int stats_write_initial, avh_batch** out) {
    if (!ctx || !cfg || !groups || !out) return fail(AVH_EINVAL, "null argument");
    *out = nullptr;
    if (ngroups < 1 || ngroups > MAX_GROUPS) return fail(AVH_EINVAL, "1 to 4 resolution groups");
    const int nv = nvars_of(cfg->equation);
    const int ng = ghosts_of(cfg->reconstruction);
    if (!nv || !ng) return fail(AVH_EINVAL, "unknown equation or reconstruction");
    if (cfg->integrator < AVH_INT_FORWARDEULER || cfg->integrator > AVH_INT_RUNGEKUTTA3)
        return fail(AVH_EINVAL, "unknown integrator");
    if (cfg->boundary != AVH_BC_PERIODIC && cfg->boundary != AVH_BC_NEUMANN)
        return fail(AVH_EINVAL, "unknown boundary");
    if (stats_saves < 0) return fail(AVH_EINVAL, "stats_saves must be >= 0");
    int dim = 0;
    for (int i = 0; i < ngroups; ++i) {
        const avh_group& G = groups[i];
        if (G.nsamples < 1) return fail(AVH_EINVAL, "every group needs at least one sample");
        for (int d = 0; d < 3; ++d)
            if (G.n[d] < 1) return fail(AVH_EINVAL, "grid sizes must be >= 1");
        const int gd = G.n[2] > 1 ? 3 : (G.n[1] > 1 ? 2 : 1);
        if (dim && gd != dim) return fail(AVH_EINVAL, "all groups need the same number of active axes");
        dim = gd;
        for (int d = 0; d < dim; ++d)
            if (G.n[d] < ng) return fail(AVH_EINVAL, "too few cells for the ghost width");
    }
    const Launchers L = launchers_for(cfg->equation, cfg->reconstruction, cfg->flux, dim);
    if (!L.batch)
        return fail(AVH_EINVAL,
                    "unsupported combination for batched runs (1D or 2D grids; burgers: godunov; euler<n>: hll, hll3)");
    AVH_HIP(hipSetDevice(ctx->device));
    auto b = new (std::nothrow) avh_batch;
    if (!b) return fail(AVH_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    b->cfg = *cfg;
    b->dim = dim;
    b->nv = nv;
    b->ngroups = ngroups;
    b->stage = L.batch;
    b->ew = batch_ew_for(cfg->equation);
    b->stats = batch_stats_for(cfg->equation);
    b->stats_initial = stats_write_initial ? 1 : 0;
    b->nsaves = stats_saves > 0 ? stats_saves + b->stats_initial : 0;
    Control& h = b->proto;
    std::memset(&h, 0, sizeof h);
    h.cfl = cfg->cfl;
    h.t_end = t_end;
    if (stats_saves > 0) {
        // FixedIntervalStatistics: interval endTime/numberOfSaves, no start
        // offset (alsuq/src/config/Setup.cpp:270-285,
        // alsuq/src/stats/FixedIntervalStatistics.cpp:33-41)
        b->stats_adj = 0;
        h.n_adj = 1;
        h.adj_interval[0] = t_end / stats_saves;
        h.adj_start[0] = 0;
        h.adj_saved[0] = stats_write_initial ? 0 : 1;
    }
    h.active = (0.0 < t_end) ? 1 : 0;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->stream = nullptr;
        free_batch(b);
        return fail(AVH_EHIP, "hipStreamCreate failed");
    }
    for (int i = 0; i < ngroups; ++i) {
        GroupHost& G = b->g[i];
        for (int d = 0; d < 3; ++d) G.n[d] = groups[i].n[d];
        G.nsamples = groups[i].nsamples;
        G.ncells = (long long)G.n[0] * G.n[1] * G.n[2];
        const size_t state = (size_t)nv * G.ncells * G.nsamples * sizeof(double);
        const size_t stat = (size_t)std::max(1, b->nsaves) * nv * G.ncells * sizeof(double);
        bool ok = hipMalloc(&G.A, state) == hipSuccess && hipMalloc(&G.B, state) == hipSuccess &&
                  hipMalloc(&G.C, state) == hipSuccess &&
                  hipMalloc(&G.ctrl, sizeof(Control) * G.nsamples) == hipSuccess &&
                  hipMalloc(&G.mean_c, stat) == hipSuccess && hipMalloc(&G.m2_c, stat) == hipSuccess &&
                  hipMalloc(&G.mean_p, stat) == hipSuccess && hipMalloc(&G.m2_p, stat) == hipSuccess &&
                  hipMalloc(&G.list, sizeof(int) * 3 * G.nsamples) == hipSuccess &&
                  hipMalloc(&G.list_len, sizeof(int)) == hipSuccess &&
                  hipMalloc(&G.counts, sizeof(long long) * std::max(1, b->nsaves)) == hipSuccess &&
                  hipMalloc(&G.active_count, sizeof(int)) == hipSuccess;
        if (!ok) {
            b->ngroups = i + 1;
            free_batch(b);
            return fail(AVH_ENOMEM, "device allocation failed");
        }
        ok = hipMemset(G.A, 0, state) == hipSuccess && hipMemset(G.B, 0, state) == hipSuccess &&
             hipMemset(G.C, 0, state) == hipSuccess && hipMemset(G.mean_c, 0, stat) == hipSuccess &&
             hipMemset(G.m2_c, 0, stat) == hipSuccess && hipMemset(G.mean_p, 0, stat) == hipSuccess &&
             hipMemset(G.m2_p, 0, stat) == hipSuccess && hipMemset(G.list_len, 0, sizeof(int)) == hipSuccess &&
             hipMemset(G.counts, 0, sizeof(long long) * std::max(1, b->nsaves)) == hipSuccess;
        Control c = h;
        for (int d = 0; d < 3; ++d) c.cell[d] = (cfg->upper[d] - cfg->lower[d]) / double(G.n[d]);
        if (G.n[1] == 1) c.cell[1] = 1;
        if (G.n[2] == 1) c.cell[2] = 1;
        std::vector<Control> cs(G.nsamples, c);
        const int act = c.active ? G.nsamples : 0;
        ok = ok && hipMemcpy(G.ctrl, cs.data(), sizeof(Control) * G.nsamples, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(G.active_count, &act, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) {
            b->ngroups = i + 1;
            free_batch(b);
            return fail(AVH_EHIP, "device initialisation failed");
        }
    }
    *out = b;
    return AVH_OK;
}

avh_status avh_batch_destroy(avh_batch* b) {
    free_batch(b);
    return AVH_OK;
}

avh_status avh_batch_add_fixed_interval_adjuster(avh_batch* b, double interval, double start, int write_initial) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    if (!(interval > 0)) return fail(AVH_EINVAL, "interval must be positive");
    if (b->proto.n_adj >= MAX_ADJ) return fail(AVH_EINVAL, "at most 4 adjusters");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    const int i = b->proto.n_adj++;
    b->proto.adj_interval[i] = interval;
    b->proto.adj_start[i] = start;
    b->proto.adj_saved[i] = write_initial ? 0 : 1;
    // the Runner's callWriters() at t = 0 (alsuq/src/run/Runner.cpp:68)
    if (0.0 >= (double)b->proto.adj_saved[i] * interval + start) b->proto.adj_saved[i] += 1;
    for (int k = 0; k < b->ngroups; ++k) {
        GroupHost& G = b->g[k];
        std::vector<Control> cs(G.nsamples);
        AVH_HIP(hipMemcpy(cs.data(), G.ctrl, sizeof(Control) * G.nsamples, hipMemcpyDeviceToHost));
        for (auto& c : cs) {
            c.n_adj = b->proto.n_adj;
            c.adj_interval[i] = interval;
            c.adj_start[i] = start;
            c.adj_saved[i] = b->proto.adj_saved[i];
        }
        AVH_HIP(hipMemcpy(G.ctrl, cs.data(), sizeof(Control) * G.nsamples, hipMemcpyHostToDevice));
    }
    return AVH_OK;
}

static avh_status check_gs(const avh_batch* b, int group, int sample) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    if (group < 0 || group >= b->ngroups) return fail(AVH_EINVAL, "no such group");
    if (sample < 0 || sample >= b->g[group].nsamples) return fail(AVH_EINVAL, "no such sample");
    return AVH_OK;
}

avh_status avh_batch_set_state(avh_batch* b, int group, int sample, const double* host_interior) {
    avh_status r = check_gs(b, group, sample);
    if (r) return r;
    if (!host_interior) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    const GroupHost& G = b->g[group];
    const size_t n = (size_t)b->nv * G.ncells;
    AVH_HIP(hipMemcpy(G.A + (size_t)sample * n, host_interior, n * sizeof(double), hipMemcpyHostToDevice));
    return AVH_OK;
}

avh_status avh_batch_get_state(avh_batch* b, int group, int sample, double* host_interior) {
    avh_status r = check_gs(b, group, sample);
    if (r) return r;
    if (!host_interior) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    const GroupHost& G = b->g[group];
    const size_t n = (size_t)b->nv * G.ncells;
    AVH_HIP(hipMemcpy(host_interior, G.A + (size_t)sample * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return AVH_OK;
}

avh_status avh_batch_set_kh2d(avh_batch* b, int group, const double* host_interfaces) {
    avh_status r = check_gs(b, group, 0);
    if (r) return r;
    if (!host_interfaces) return fail(AVH_EINVAL, "null argument");
    if (b->cfg.equation != AVH_EQ_EULER2 || b->dim != 2) return fail(AVH_EINVAL, "KH fill needs euler2 on 2D grids");
    AVH_HIP(hipSetDevice(b->ctx->device));
    const GroupHost& G = b->g[group];
    const size_t bytes = (size_t)G.nsamples * 2 * G.n[0] * sizeof(double);
    double* d = nullptr;
    AVH_HIP(hipMalloc(&d, bytes));
    hipError_t e = hipMemcpy(d, host_interfaces, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const double dy = (b->cfg.upper[1] - b->cfg.lower[1]) / double(G.n[1]);
        hipLaunchKernelGGL(kh2d_fill_kernel, dim3(2048), dim3(256), 0, b->stream, G.A, (long long)b->nv * G.ncells,
                           G.n[0], G.n[1], d, G.nsamples, b->cfg.lower[1], dy, b->cfg.gamma);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("KH fill: ") + hipGetErrorString(e));
    return AVH_OK;
}

avh_status avh_batch_advance(avh_batch* b, int64_t max_steps, int chunk, int64_t* steps_done, int64_t* active_left) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    if (chunk < 1) chunk = 1;
    AVH_HIP(hipSetDevice(b->ctx->device));
    if (b->stats_initial && !b->initial_done && b->nsaves > 0) {
        unsigned int nbs = 0;
        StatsParams S = stats_params(b, &nbs);
        hipLaunchKernelGGL(batch_initial_save_kernel, dim3(b->ngroups), dim3(256), 0, b->stream, S);
        AVH_HIP(hipGetLastError());
        b->stats(S, nbs, b->stream);
        AVH_HIP(hipGetLastError());
        b->initial_done = 1;
    }
    avh_status rs = batch_ensure_shape(b);
    if (rs) return rs;
    // every group takes part until a chunk shows that all its samples have
    // stopped (the end-of-step kernel counts the running samples of every
    // group, live or not, at every step)
    set_all_live(b);
    int64_t done = 0, left = 0;
    while (done < max_steps) {
        const int64_t k = std::min<int64_t>(chunk, max_steps - done);
        avh_status r = enqueue_steps_graphed(b->graph, b->graph_steps, b->stream, k,
                                             [b]() { return enqueue_batch_step(b); });
        if (r) return r;
        done += k;
        AVH_HIP(hipStreamSynchronize(b->stream));
        left = 0;
        for (int i = 0; i < b->ngroups; ++i) {
            int a = 0;
            AVH_HIP(hipMemcpy(&a, b->g[i].active_count, sizeof(int), hipMemcpyDeviceToHost));
            left += a;
            if (a == 0 && b->live[i]) {
                b->live[i] = false;  // drop the finished group from the launches
                b->graph.reset();
            }
        }
        if (left == 0) break;
    }
    if (steps_done) *steps_done = done;
    if (active_left) *active_left = left;
    return AVH_OK;
}

avh_status avh_batch_set_graph_steps(avh_batch* b, int steps) {
    if (!b || steps < 0 || steps > 4096) return fail(AVH_EINVAL, "bad argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    b->graph.reset();
    b->graph_steps = steps;
    return AVH_OK;
}

avh_status avh_batch_violations(avh_batch* b, int group, int64_t* count, int64_t* first_sample) {
    avh_status r = check_gs(b, group, 0);
    if (r) return r;
    if (!count || !first_sample) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    const GroupHost& G = b->g[group];
    std::vector<Control> cs(G.nsamples);
    AVH_HIP(hipMemcpy(cs.data(), G.ctrl, sizeof(Control) * G.nsamples, hipMemcpyDeviceToHost));
    *count = 0;
    *first_sample = -1;
    for (int s = 0; s < G.nsamples; ++s)
        if (cs[s].violated) {
            if (*first_sample < 0) *first_sample = s;
            *count += 1;
        }
    return AVH_OK;
}

avh_status avh_batch_set_end_time(avh_batch* b, double t_end) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    for (int i = 0; i < b->ngroups; ++i) {
        hipLaunchKernelGGL(set_end_time_kernel, dim3(1), dim3(256), 0, b->stream, b->g[i].ctrl, b->g[i].nsamples, t_end,
                           b->g[i].active_count);
        AVH_HIP(hipGetLastError());
    }
    b->proto.t_end = t_end;
    AVH_HIP(hipStreamSynchronize(b->stream));
    return AVH_OK;
}

avh_status avh_batch_kinetic_energy(avh_batch* b, int group, int component, double* out) {
    avh_status r = check_gs(b, group, 0);
    if (r) return r;
    if (!out) return fail(AVH_EINVAL, "null argument");
    if (b->cfg.equation == AVH_EQ_BURGERS || component < 0 || component >= b->nv - 2)
        return fail(AVH_EINVAL, "kinetic energy needs an euler equation and a velocity component < dimensions");
    AVH_HIP(hipSetDevice(b->ctx->device));
    const GroupHost& G = b->g[group];
    double dv = 1.0;
    for (int d = 0; d < 3; ++d)
        if (G.n[d] > 1) dv *= (b->cfg.upper[d] - b->cfg.lower[d]) / double(G.n[d]);
    double* d = nullptr;
    AVH_HIP(hipMalloc(&d, sizeof(double) * G.nsamples));
    hipLaunchKernelGGL(kinetic_energy_kernel, dim3(G.nsamples), dim3(256), 0, b->stream, G.A,
                       (long long)b->nv * G.ncells, G.ncells, component, dv, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(double) * G.nsamples, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("kinetic energy: ") + hipGetErrorString(e));
    return AVH_OK;
}

avh_status avh_batch_structure_cube(avh_batch* b, int group, int var, int number_of_h, double p, double* out) {
    avh_status r = check_gs(b, group, 0);
    if (r) return r;
    if (!out) return fail(AVH_EINVAL, "null argument");
    if (var < 0 || var >= b->nv) return fail(AVH_EINVAL, "no such variable");
    if (b->dim > 2) return fail(AVH_EINVAL, "structure functions of batches: 1D and 2D grids");
    const GroupHost& G = b->g[group];
    if (number_of_h < 2 || number_of_h > 65535) return fail(AVH_EINVAL, "number_of_h must be in [2, 65535]");
    if (!(p > 0)) return fail(AVH_EINVAL, "p must be positive");
    if (G.nsamples > 65535) return fail(AVH_EINVAL, "structure functions: at most 65535 samples per group");
    AVH_HIP(hipSetDevice(b->ctx->device));
    const unsigned int nchunks = (unsigned int)((G.ncells + 255) / 256);
    const size_t npart = (size_t)nchunks * (number_of_h - 1) * G.nsamples;
    double *part = nullptr, *d = nullptr;
    AVH_HIP(hipMalloc(&part, sizeof(double) * npart));
    if (hipMalloc(&d, sizeof(double) * G.nsamples * number_of_h) != hipSuccess) {
        (void)hipFree(part);
        return fail(AVH_ENOMEM, "device allocation failed");
    }
    const dim3 grid(nchunks, number_of_h - 1, G.nsamples);
    const long long stride = (long long)b->nv * G.ncells;
    const int pi = p == 1.0 ? 1 : p == 2.0 ? 2 : p == 3.0 ? 3 : p == 4.0 ? 4 : p == 5.0 ? 5 : 0;
    switch (pi) {
    case 1: hipLaunchKernelGGL(structure_partial_kernel<1>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    case 2: hipLaunchKernelGGL(structure_partial_kernel<2>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    case 3: hipLaunchKernelGGL(structure_partial_kernel<3>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    case 4: hipLaunchKernelGGL(structure_partial_kernel<4>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    case 5: hipLaunchKernelGGL(structure_partial_kernel<5>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    default: hipLaunchKernelGGL(structure_partial_kernel<0>, grid, dim3(256), 0, b->stream, G.A, stride, var, G.n[0], G.n[1], b->dim, b->cfg.boundary, p, part); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        const int nt = G.nsamples * number_of_h;
        hipLaunchKernelGGL(structure_finish_kernel, dim3((nt + 255) / 256), dim3(256), 0, b->stream, part, (int)nchunks,
                           number_of_h, G.nsamples, (double)G.ncells, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, sizeof(double) * G.nsamples * number_of_h, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(part);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("structure function: ") + hipGetErrorString(e));
    return AVH_OK;
}

avh_status avh_batch_set_tile_shape(avh_batch* b, int shape) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    if (shape < -1 || shape >= num_shapes(b->dim)) return fail(AVH_EINVAL, "no such tile shape");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    b->graph.reset();
    b->shape = shape;
    return AVH_OK;
}

avh_status avh_batch_autotune(avh_batch* b, int* best_shape, double* ms_per_shape, int* nshapes) {
    if (!b) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    return batch_autotune(b, best_shape, ms_per_shape, nshapes);
}

avh_status avh_batch_time_step_kernels(avh_batch* b, int64_t steps, double* ms, int* kernels) {
    if (!b || !ms || !kernels || steps < 1 || steps > 100000) return fail(AVH_EINVAL, "bad argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    const int nk = batch_launches_per_step(b);
    avh_status r = batch_ensure_shape(b);
    if (r) return r;
    EventMarks t;
    if ((r = t.create((size_t)steps * (nk + 1)))) return r;
    for (int64_t k = 0; k < steps; ++k)
        if ((r = enqueue_batch_step(b, &t))) return r;
    if ((r = t.average(steps, nk, ms))) return r;
    *kernels = nk;
    return AVH_OK;
}

avh_status avh_batch_get_progress(avh_batch* b, int group, int sample, avh_progress* p) {
    avh_status r = check_gs(b, group, sample);
    if (r) return r;
    if (!p) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    Control c;
    AVH_HIP(hipMemcpy(&c, b->g[group].ctrl + sample, sizeof c, hipMemcpyDeviceToHost));
    p->time = c.t;
    p->last_dt = c.last_dt;
    p->steps = c.steps;
    p->violated = c.violated;
    p->saves_pending = c.saves_pending;
    return AVH_OK;
}

avh_status avh_batch_get_statistics(avh_batch* b, int group, int save, int primitive, double* mean,
                                    double* variance, int64_t* count) {
    avh_status r = check_gs(b, group, 0);
    if (r) return r;
    if (!mean || !variance || !count) return fail(AVH_EINVAL, "null argument");
    if (save < 0 || save >= b->nsaves) return fail(AVH_EINVAL, "no such statistics save");
    AVH_HIP(hipSetDevice(b->ctx->device));
    AVH_HIP(hipStreamSynchronize(b->stream));
    const GroupHost& G = b->g[group];
    const size_t n = (size_t)b->nv * G.ncells;
    const size_t off = (size_t)save * n;
    long long cnt = 0;
    AVH_HIP(hipMemcpy(&cnt, G.counts + save, sizeof cnt, hipMemcpyDeviceToHost));
    AVH_HIP(hipMemcpy(mean, (primitive ? G.mean_p : G.mean_c) + off, n * sizeof(double), hipMemcpyDeviceToHost));
    AVH_HIP(hipMemcpy(variance, (primitive ? G.m2_p : G.m2_c) + off, n * sizeof(double), hipMemcpyDeviceToHost));
    // population variance M2 / M (the reference's M2/M - mean^2,
    // alsuq/src/stats/MeanVariance.cpp:55-67, StatisticsHelper.cpp:70-72)
    for (size_t i = 0; i < n; ++i) variance[i] = cnt > 0 ? variance[i] / (double)cnt : 0.0;
    *count = cnt;
    return AVH_OK;
}

avh_status avh_uq_generate(int nparams, const int* dims, const double* lower, const double* upper,
                           int64_t sample_start, int64_t nsamples, double* out) {
    if (nparams < 1 || !dims || !lower || !upper || !out || sample_start < 0 || nsamples < 0)
        return fail(AVH_EINVAL, "bad argument");
    // One default-seeded std::mt19937_64 shared by every parameter, with the
    // reference's fast-forward: component c of sample s of a parameter of
    // dimension D is preceded by s*D + c draws counted on the shared counter
    // (alsuq/src/generator/STLMersenne.cpp:25-85); uniform_real_distribution
    // <double>(0,1) scaled to [lower, upper) (alsuq/src/distribution/
    // Uniform.cpp:27-34); parameters in the caller's order, which is the
    // reference's parameter-name order.
    std::mt19937_64 engine;
    long counter = 0;
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    size_t o = 0;
    for (int64_t s = sample_start; s < sample_start + nsamples; ++s) {
        for (int p = 0; p < nparams; ++p) {
            const long dimension = dims[p];
            if (dimension < 1) return fail(AVH_EINVAL, "parameter dimensions must be >= 1");
            for (long c = 0; c < dimension; ++c) {
                if (counter / dimension < long(s)) {
                    const long add = long(s) * dimension - counter;
                    engine.discard(add);
                    counter += add;
                }
                if (counter % dimension < c) {
                    const long add = c - counter % dimension;
                    engine.discard(add);
                    counter += add;
                }
                counter++;
                const double x = dist(engine);
                out[o++] = x * (upper[p] - lower[p]) + lower[p];
            }
        }
    }
    return AVH_OK;
}

}  // extern "C"
