// C ABI of the hip platform (include/alsvinn_hip.h): solver objects, device
// buffers, the per-step kernel sequence of each integrator, the ghost-filled
// state export and the small element-wise kernels.
//
// Buffers (interior-only SoA, nvars * ncells doubles each):
//   A  the state U0 (always holds the current solution between steps)
//   B  L0 = unscaled net flux of stage 1
//   C  U1 (rungekutta2) or U2 (rungekutta3)
// Step sequences (reference: alsfvm/src/integrator/{ForwardEuler,RungeKutta2,
// RungeKutta3}.cpp, Simulator::incrementSolution Simulator.cpp:191-214):
//   forwardeuler: stage(L: A->B) ; ew(A = B*s + A, check, end step)
//   rungekutta2 : stage(L: A->B) ; ew(C = B*s + A) ; stage(RK2_2: C, A -> A)
//   rungekutta3 : stage(L: A->B) ; stage(RK3_2: B*s+A on the fly -> C) ;
//                 stage(RK3_3: C, A -> A)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/alsvinn_hip.h"
#include "avh_internal.hpp"
#include "avh_stage_inst.hpp"

#pragma clang fp contract(off)

using namespace avh;

namespace {

// End of a step (end_step) as its own launch, after forward euler's
// element-wise pass; the RK2 and RK3 final stages end the step themselves
// (StageParams::finish).
__global__ __launch_bounds__(64) void finish_step_kernel(Control* c) {
    if (threadIdx.x != 0 || !c->active) return;
    end_step(c, c->block_violation != 0);
}

enum EwMode { EW_FE_FINISH = 0, EW_RK2_U1 = 1 };

// out = L*s + U0 element-wise (Volume *= dt/dx; += U0).  EW_FE_FINISH writes
// in place into U0, checks the constraints and closes the step.
template <class EQ>
__global__ __launch_bounds__(256) void ew_kernel(const double* L, const double* u0, double* out,
                                                 Control* ctrl, long long nc, int mode, double gamma) {
    if (!ctrl->active) return;
    constexpr int NV = EQ::NV;
    const double s = step_dt(ctrl) / ctrl->cell[0];
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        double r[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double o = L[v * nc + i] * s;
            o = o + u0[v * nc + i];
            r[v] = o;
        }
        if (mode == EW_FE_FINISH) bad = bad || !EQ::obeys(r, gamma);
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v * nc + i] = r[v];
    }
    if (mode == EW_FE_FINISH) {
        if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&ctrl->block_violation, 1);
    }
}

// Ghosted export: every cell of the (n+2g)-sized volume takes the value of
// the interior cell the boundary map assigns to it.
__global__ __launch_bounds__(256) void ghost_fill_kernel(const double* in, double* out, int nv,
                                                         int n0, int n1, int n2, int g0, int g1,
                                                         int g2, int bc) {
    const long long t0 = n0 + 2 * g0, t1 = n1 + 2 * g1, t2 = n2 + 2 * g2;
    const long long total = t0 * t1 * t2;
    const long long nc = (long long)n0 * n1 * n2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % t0);
        const int y = (int)((i / t0) % t1);
        const int z = (int)(i / (t0 * t1));
        const int sx = bc_map(x - g0, n0, bc);
        const int sy = g1 ? bc_map(y - g1, n1, bc) : 0;
        const int sz = g2 ? bc_map(z - g2, n2, bc) : 0;
        const long long si = ((long long)sz * n1 + sy) * n0 + sx;
        for (int v = 0; v < nv; ++v) out[v * total + i] = in[v * nc + si];
    }
}

// Arithmetic self-test: the stage kernels' square root without range fix-ups
// (sqrt_core) and shared-reciprocal quotients (div_recip / div_shared)
// against the compiler's sqrt and divide, bit for bit, on operands drawn
// from a counter-based generator over the ranges the kernels' guards admit.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// a double in [2^elo, 2^ehi]: one draw in 64 is exactly one of the two
// bounds, the others have a uniformly random 52-bit mantissa and a binary
// exponent drawn uniformly from [elo, ehi - 1]
__device__ __forceinline__ double draw_double(unsigned long long r, int elo, int ehi) {
    if (((r >> 46) & 63ull) == 0) {
        const int e = ((r >> 45) & 1ull) ? ehi : elo;
        return __longlong_as_double((long long)((unsigned long long)(e + 1023) << 52));
    }
    const unsigned long long mant = r & ((1ull << 52) - 1);
    const int e = elo + (int)((r >> 52) % (unsigned long long)(ehi - elo));
    return __longlong_as_double((long long)(((unsigned long long)(e + 1023) << 52) | mant));
}

enum { ARITH_SQRT = 0, ARITH_DIV_RHO = 1, ARITH_DIV_WENO = 2, ARITH_DIV_ROE = 3, ARITH_FREXP = 4 };

// frexp exponent of a finite double from its bits: 0 for +-0; for normal x
// the biased exponent - 1022; for subnormal x = mant * 2^-1074 the position
// of mant's leading bit (x = f * 2^e with 0.5 <= f < 1)
__device__ __forceinline__ int frexp_exp_bits(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
    if (b == 0) return 0;
    const int be = (int)(b >> 52);
    if (be != 0) return be - 1022;
    return (64 - __clzll((long long)b)) - 1074;
}

__global__ __launch_bounds__(256) void arith_selftest_kernel(int kind, unsigned long long seed, long long n,
                                                             unsigned long long* bad, double* first) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const unsigned long long r0 = splitmix64(seed ^ (2ull * (unsigned long long)i));
        const unsigned long long r1 = splitmix64(seed ^ (2ull * (unsigned long long)i + 1ull));
        double a = 0, b = 0, got = 0, want = 0;
        if (kind == ARITH_SQRT) {
            // [2^-767, 2^1023] and DBL_MAX: the range SharedOps::sqrt_ admits
            a = ((r1 & 255ull) == 0) ? 0x1.fffffffffffffp+1023 : draw_double(r0, -767, 1023);
            got = sqrt_core(a);
            want = sqrt(a);
        } else if (kind == ARITH_DIV_RHO) {
            // denominators in [2^-100, 2^100], numerators zero or +-[2^-800, 2^600]:
            // the ranges of make_all_ops / sqrt_div_rho
            b = draw_double(r1, -100, 100);
            a = (r0 & 1023ull) == 0 ? 0.0 : draw_double(r0, -800, 600);
            if (r1 & (1ull << 63)) a = -a;
            got = div_shared(a, b, div_recip(b));
            want = a / b;
        } else if (kind == ARITH_DIV_WENO) {
            // weights a0, a1 in [2^-400, 2^53]: quotient a0 / (a0 + a1)
            const double a0 = draw_double(r0, -400, 53);
            const double a1 = draw_double(r1, -400, 53);
            a = a0;
            b = a0 + a1;
            got = div_shared(a, b, div_recip(b));
            want = a0 / (a0 + a1);
        } else if (kind == ARITH_FREXP) {
            // the exponent the range checks read (frexp_exp) on zeros,
            // subnormals and normals of either sign: 1 draw in 8 a signed
            // zero, 1 in 8 a subnormal, the rest normals over the whole range
            const unsigned long long sel = r1 & 7ull;
            if (sel == 0) {
                a = (r1 & 8ull) ? -0.0 : 0.0;
            } else if (sel == 1) {
                const unsigned long long mant = (r0 & ((1ull << 52) - 1)) >> ((r1 >> 4) % 52ull);
                a = __longlong_as_double((long long)(mant ? mant : 1ull));
            } else {
                a = draw_double(r0, -1022, 1023);
            }
            if (r1 & (1ull << 63)) a = -a;
            got = (double)frexp_exp(a);
            want = (double)frexp_exp_bits(a);
        } else {
            // Roe denominators sqrt(rhoL) + sqrt(rhoR) in [2^-50, 2^51],
            // numerators zero or +-[2^-800, 2^600]
            b = draw_double(r1, -50, 51);
            a = (r0 & 1023ull) == 0 ? 0.0 : draw_double(r0, -800, 600);
            if (r1 & (1ull << 62)) a = -a;
            got = div_shared(a, b, div_recip(b));
            want = a / b;
        }
        if (__double_as_longlong(got) != __double_as_longlong(want)) {
            if (atomicAdd(bad, 1ull) == 0ull) {
                first[0] = a;
                first[1] = b;
            }
        }
    }
}

typedef void (*EwLauncher)(const double*, const double*, double*, Control*, long long, int, double,
                           hipStream_t);

template <class EQ>
void launch_ew(const double* L, const double* u0, double* out, Control* c, long long nc, int mode,
               double gamma, hipStream_t st) {
    long long nb = (nc + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(ew_kernel<EQ>, dim3((unsigned)nb), dim3(256), 0, st, L, u0, out, c, nc, mode,
                       gamma);
}

// One sample folded into running mean / M2 arrays (Welford, n = samples
// including this one), for the conserved variables and for rho, u, p
// (MeanVariance, alsuq/src/stats/MeanVariance.cpp:30-67).
template <class EQ>
__global__ __launch_bounds__(256) void welford_kernel(const double* q, long long nc, double n, double gamma,
                                                      double* mean_c, double* m2_c, double* mean_p, double* m2_p) {
    constexpr int NV = EQ::NV;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nc) return;
    double u[NV], p[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = q[v * nc + i];
    if constexpr (EQ::NV > 1) {
        const typename EQ::All a = EQ::make_all(u, gamma);
        p[0] = u[0];
#pragma unroll
        for (int v = 1; v < NV - 1; ++v) p[v] = a.u[v - 1];
        p[NV - 1] = a.p;
    } else {
        p[0] = u[0];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const long long o = v * nc + i;
        double m = mean_c[o];
        const double d = u[v] - m;
        m = m + d / n;
        mean_c[o] = m;
        m2_c[o] = m2_c[o] + d * (u[v] - m);
        double mp = mean_p[o];
        const double dp = p[v] - mp;
        mp = mp + dp / n;
        mean_p[o] = mp;
        m2_p[o] = m2_p[o] + dp * (p[v] - mp);
    }
}

// Unfused reconstruction along one direction: the left- and right-face
// values of every interior cell (Reconstruction::performReconstruction,
// alsfvm/src/reconstruction/ReconstructionCPU.cpp:37-97, with the ghost
// cells taken through the boundary map).  One thread per cell; the stencil
// along `dir` is gathered into a register window centred on the cell.
template <int RECON, int NV>
__global__ __launch_bounds__(256) void reconstruct_kernel(const double* in, double* L, double* R, long long nc, int n0,
                                                          int n1, int n2, int dir, int bc) {
    constexpr int G = ReconTraits<RECON>::GHOST;
    constexpr int HW = G - 1 > 1 ? G - 1 : 1;
    constexpr int WZ = 2 * HW + 1;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nc) return;
    const int x = (int)(i % n0), y = (int)((i / n0) % n1), z = (int)(i / ((long long)n0 * n1));
    double u[WZ][NV];
#pragma unroll
    for (int k = 0; k < WZ; ++k) {
        int xx = x, yy = y, zz = z;
        if (dir == 0) xx = bc_map(x + k - HW, n0, bc);
        else if (dir == 1) yy = bc_map(y + k - HW, n1, bc);
        else zz = bc_map(z + k - HW, n2, bc);
        const long long gi = ((long long)zz * n1 + yy) * n0 + xx;
#pragma unroll
        for (int v = 0; v < NV; ++v) u[k][v] = in[v * nc + gi];
    }
    double l[NV], r[NV];
    recon_face_reg<RECON, NV, false, WZ>(u, HW, l);
    recon_face_reg<RECON, NV, true, WZ>(u, HW, r);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        L[v * nc + i] = l[v];
        R[v * nc + i] = r[v];
    }
}

template <int RECON, int NV>
void launch_reconstruct(const double* in, double* L, double* R, long long nc, const int* n, int dir, int bc,
                        hipStream_t st) {
    hipLaunchKernelGGL((reconstruct_kernel<RECON, NV>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, in, L,
                       R, nc, n[0], n[1], n[2], dir, bc);
}

template <int NV>
bool launch_reconstruct_nv(int recon, const double* in, double* L, double* R, long long nc, const int* n, int dir,
                           int bc, hipStream_t st) {
    switch (recon) {
    case RECON_NONE: launch_reconstruct<RECON_NONE, NV>(in, L, R, nc, n, dir, bc, st); return true;
    case RECON_WENO2: launch_reconstruct<RECON_WENO2, NV>(in, L, R, nc, n, dir, bc, st); return true;
    case RECON_ENO2: launch_reconstruct<RECON_ENO2, NV>(in, L, R, nc, n, dir, bc, st); return true;
    case RECON_ENO3: launch_reconstruct<RECON_ENO3, NV>(in, L, R, nc, n, dir, bc, st); return true;
    }
    return false;
}

typedef void (*WelfordLauncher)(const double*, long long, double, double, double*, double*, double*, double*,
                                hipStream_t);
template <class EQ>
void launch_welford(const double* q, long long nc, double n, double gamma, double* a, double* b, double* c, double* d,
                    hipStream_t st) {
    hipLaunchKernelGGL(welford_kernel<EQ>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, q, nc, n, gamma, a, b,
                       c, d);
}
WelfordLauncher welford_for(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return &launch_welford<BurgersEq>;
    case AVH_EQ_EULER1: return &launch_welford<EulerEq<1>>;
    case AVH_EQ_EULER2: return &launch_welford<EulerEq<2>>;
    case AVH_EQ_EULER3: return &launch_welford<EulerEq<3>>;
    }
    return nullptr;
}

EwLauncher ew_for(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return &launch_ew<BurgersEq>;
    case AVH_EQ_EULER1: return &launch_ew<EulerEq<1>>;
    case AVH_EQ_EULER2: return &launch_ew<EulerEq<2>>;
    case AVH_EQ_EULER3: return &launch_ew<EulerEq<3>>;
    }
    return nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// The time step that allocates nothing (avh_step_noalloc).  Its two stage
// buffers (L0 and U1/U2) and its control block live in `scratch`, a static
// device array that every GPU the library runs on holds until the process
// ends: 512 bytes for the control block, then two buffers of
// NOALLOC_MAX_VALUES doubles (64 MiB in all).  Host code never hands the
// array's name to a kernel; its device address is looked up once per device
// with hipGetSymbolAddress when a context is created, and kept per device.
constexpr long long NOALLOC_MAX_VALUES = 1ll << 22;  // nvars * cells of one stage buffer
constexpr long long NOALLOC_CTRL_DOUBLES = 64;
static_assert(sizeof(Control) <= NOALLOC_CTRL_DOUBLES * sizeof(double), "the control block fits its slot");
static __device__ __align__(256) double scratch[NOALLOC_CTRL_DOUBLES + 2 * NOALLOC_MAX_VALUES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device (`device`)
static avh_status noalloc_scratch(int device, double** out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(AVH_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(AVH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return AVH_OK;
}

// a fresh control block for one step: time 0, the CFL number and cell
// lengths, and dt_cap as a one-save fixed-interval adjuster (next save at
// 1 * dt_cap + 0, so the clip candidate is dt_cap exactly); every field
// written, since the block keeps the previous call's contents
static __global__ void noalloc_init_kernel(Control* c, double c0, double c1, double c2, double cfl, double cap) {
    if (threadIdx.x != 0) return;
    clear_wave_speeds(c);
    c->t = 0.0;
    c->dt = 0.0;
    c->scale = 0.0;
    c->last_dt = 0.0;
    c->t_end = __builtin_inf();
    c->cfl = cfl;
    c->cell[0] = c0;
    c->cell[1] = c1;
    c->cell[2] = c2;
    for (int i = 0; i < MAX_ADJ; ++i) {
        c->adj_interval[i] = 0.0;
        c->adj_start[i] = 0.0;
        c->adj_saved[i] = 0;
    }
    c->n_adj = 0;
    if (!__builtin_isinf(cap)) {
        c->n_adj = 1;
        c->adj_interval[0] = cap;
        c->adj_saved[0] = 1;
    }
    c->steps = 0;
    c->active = 1;
    c->violated = 0;
    c->arrive = 0ull;
    c->saves_pending = 0;
    c->pause_on_save = 0;
    c->block_violation = 0;
}

static __global__ void noalloc_result_kernel(const Control* c, double* dt, int* violated) {
    if (threadIdx.x != 0) return;
    if (dt) dt[0] = c->last_dt;
    if (violated) violated[0] = c->violated;
}

struct avh_solver {
    avh_ctx* ctx = nullptr;
    avh_config cfg{};
    int dim = 1, nv = 0, ng = 0;
    int g[3] = {0, 0, 0};
    long long ncells = 0;
    long long nghosted = 0;
    double *A = nullptr, *B = nullptr, *C = nullptr, *staging = nullptr;
    Control* ctrl = nullptr;
    Control h{};  // host mirror of ctrl, valid after pull()
    hipStream_t stream = nullptr;
    StageLauncher stage = nullptr;
    EwLauncher ew = nullptr;
    int graph_steps = 16;  // steps per captured hipGraph (0: ordinary launches)
    StepGraph graph;
    int shape = -1;        // tile shape of the stage kernel; -1: autotune on first use
};

static avh_status use_device(const avh_solver* s) {
    AVH_HIP(hipSetDevice(s->ctx->device));
    return AVH_OK;
}

static avh_status pull(avh_solver* s) {
    AVH_HIP(hipStreamSynchronize(s->stream));
    AVH_HIP(hipMemcpy(&s->h, s->ctrl, sizeof(Control), hipMemcpyDeviceToHost));
    return AVH_OK;
}

static avh_status push(avh_solver* s) {
    AVH_HIP(hipMemcpy(s->ctrl, &s->h, sizeof(Control), hipMemcpyHostToDevice));
    return AVH_OK;
}

static void recompute_active(Control& h) {
    h.active = (!(h.t >= h.t_end) && !h.violated && !(h.pause_on_save && h.saves_pending > 0)) ? 1 : 0;
}

static avh_status mark(avh_solver* s, EventMarks* t) { return avh::mark(t, s->stream); }

static avh_status enqueue_step(avh_solver* s, EventMarks* tm = nullptr) {
    StageParams P;
    P.ctrl = s->ctrl;
    P.ncells = s->ncells;
    for (int d = 0; d < 3; ++d) P.n[d] = s->cfg.n[d];
    P.bc = s->cfg.boundary;
    P.gamma = s->cfg.gamma;
    // stage 1: L0 = L(U0) into B, dt from the wave speeds
    P.in = s->A;
    P.u0 = s->A;
    P.out = s->B;
    P.mode = MODE_L;
    avh_status r;
    if ((r = mark(s, tm))) return r;
    AVH_HIP(s->stage(P, s->shape, s->stream));
    if ((r = mark(s, tm))) return r;
    if (s->cfg.integrator == AVH_INT_FORWARDEULER) {
        s->ew(s->B, s->A, s->A, s->ctrl, s->ncells, EW_FE_FINISH, s->cfg.gamma, s->stream);
        AVH_HIP(hipGetLastError());
        if ((r = mark(s, tm))) return r;
        hipLaunchKernelGGL(finish_step_kernel, dim3(1), dim3(64), 0, s->stream, s->ctrl);
        AVH_HIP(hipGetLastError());
        if ((r = mark(s, tm))) return r;
    } else if (s->cfg.integrator == AVH_INT_RUNGEKUTTA2) {
        s->ew(s->B, s->A, s->C, s->ctrl, s->ncells, EW_RK2_U1, s->cfg.gamma, s->stream);
        AVH_HIP(hipGetLastError());
        if ((r = mark(s, tm))) return r;
        P.in = s->C;
        P.u0 = s->A;
        P.out = s->A;
        P.mode = MODE_RK2_2;
        P.finish = 1;
        AVH_HIP(s->stage(P, s->shape, s->stream));
        if ((r = mark(s, tm))) return r;
    } else {
        P.in = s->B;
        P.u0 = s->A;
        P.out = s->C;
        P.mode = MODE_RK3_2;
        AVH_HIP(s->stage(P, s->shape, s->stream));
        if ((r = mark(s, tm))) return r;
        P.in = s->C;
        P.u0 = s->A;
        P.out = s->A;
        P.mode = MODE_RK3_3;
        P.finish = 1;
        AVH_HIP(s->stage(P, s->shape, s->stream));
        if ((r = mark(s, tm))) return r;
    }
    return AVH_OK;
}

static TuneKey tune_key(const avh_solver* s) {
    return TuneKey(0, s->cfg.equation * 16 + s->cfg.reconstruction, s->cfg.flux, s->dim, s->cfg.n[0], s->cfg.n[1],
                   s->cfg.n[2]);
}

// Times every tile shape on the current state and keeps the fastest: a
// flux-only launch (writing the L0 buffer, which stage 1 overwrites), and
// for RK3 also the second stage on its output (input formed on the fly from
// L0 and U0, writing the U2 buffer), whose extra reads weigh differently on
// different shapes.  The state is not written; the wave-speed maxima the
// launches add to are restored.
static avh_status autotune(avh_solver* s, int* best, double* ms, int* nshapes) {
    const int ns = num_shapes(s->dim);
    avh_status r = pull(s);
    if (r) return r;
    const Control saved = s->h;
    StageParams P;
    P.ctrl = s->ctrl;
    P.ncells = s->ncells;
    for (int d = 0; d < 3; ++d) P.n[d] = s->cfg.n[d];
    P.bc = s->cfg.boundary;
    P.gamma = s->cfg.gamma;
    P.in = s->A;
    P.u0 = s->A;
    P.out = s->B;
    P.mode = MODE_FLUX_ONLY;
    StageParams P2 = P;
    P2.in = s->B;
    P2.out = s->C;
    P2.mode = MODE_RK3_2;
    const bool rk3 = s->cfg.integrator == AVH_INT_RUNGEKUTTA3 && s->h.active;
    int b = 0;
    double tmin = 0;
    for (int k = 0; k < ns; ++k) {
        double t = 0;
        if (ns > 1) {
            r = time_launches(s->stream, 5,
                              [&]() {
                                  hipError_t e = s->stage(P, k, s->stream);
                                  if (e == hipSuccess && rk3) e = s->stage(P2, k, s->stream);
                                  return e;
                              },
                              &t);
            if (r) return r;
        }
        if (ms) ms[k] = t;
        if (k == 0 || t < tmin) {
            tmin = t;
            b = k;
        }
    }
    AVH_HIP(hipStreamSynchronize(s->stream));
    s->h = saved;
    if ((r = push(s))) return r;
    if (s->shape != b) s->graph.reset();
    s->shape = b;
    tune_store(tune_key(s), b);
    if (best) *best = b;
    if (nshapes) *nshapes = ns;
    return AVH_OK;
}

static avh_status ensure_shape(avh_solver* s) {
    if (s->shape >= 0) return AVH_OK;
    int b;
    if (tune_lookup(tune_key(s), &b)) {
        s->shape = b;
        return AVH_OK;
    }
    return autotune(s, nullptr, nullptr, nullptr);
}

static avh_status enqueue_steps(avh_solver* s, int64_t n) {
    avh_status r = ensure_shape(s);
    if (r) return r;
    return enqueue_steps_graphed(s->graph, s->graph_steps, s->stream, n, [s]() { return enqueue_step(s); });
}

extern "C" {

const char* avh_last_error(void) { return last_error_slot().c_str(); }

avh_status avh_ctx_create(int device, avh_ctx** ctx) {
    if (!ctx) return fail(AVH_EINVAL, "ctx is null");
    int count = 0;
    AVH_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(AVH_EINVAL, "no such device");
    AVH_HIP(hipSetDevice(device));
    // the no-allocation step's scratch address on this device, once
    double* sp = nullptr;
    avh_status r = noalloc_scratch(device, &sp);
    if (r) return r;
    auto c = new (std::nothrow) avh_ctx;
    if (!c) return fail(AVH_ENOMEM, "host allocation failed");
    c->device = device;
    *ctx = c;
    return AVH_OK;
}

avh_status avh_ctx_destroy(avh_ctx* ctx) {
    if (ctx && ctx->scratch) {
        (void)hipSetDevice(ctx->device);
        (void)hipFree(ctx->scratch);
    }
    delete ctx;
    return AVH_OK;
}

avh_status avh_solver_destroy(avh_solver* s) {
    if (!s) return AVH_OK;
    (void)hipSetDevice(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->graph.reset();
    (void)hipFree(s->A);
    (void)hipFree(s->B);
    (void)hipFree(s->C);
    (void)hipFree(s->staging);
    (void)hipFree(s->ctrl);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return AVH_OK;
}

avh_status avh_solver_create(avh_ctx* ctx, const avh_config* cfg, avh_solver** out) {
    if (!ctx || !cfg || !out) return fail(AVH_EINVAL, "null argument");
    *out = nullptr;
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    const int ng = ghosts_of(cfg->reconstruction);
    if (!ng) return fail(AVH_EINVAL, "unknown reconstruction");
    if (cfg->integrator < AVH_INT_FORWARDEULER || cfg->integrator > AVH_INT_RUNGEKUTTA3)
        return fail(AVH_EINVAL, "unknown integrator");
    if (cfg->boundary != AVH_BC_PERIODIC && cfg->boundary != AVH_BC_NEUMANN)
        return fail(AVH_EINVAL, "unknown boundary");
    for (int d = 0; d < 3; ++d)
        if (cfg->n[d] < 1) return fail(AVH_EINVAL, "grid sizes must be >= 1");
    const int dim = cfg->n[2] > 1 ? 3 : (cfg->n[1] > 1 ? 2 : 1);
    for (int d = 0; d < dim; ++d)
        if (cfg->n[d] < ng)
            return fail(AVH_EINVAL, "too few cells: every active axis needs at least as many cells as ghost cells");
    const StageLauncher st = launchers_for(cfg->equation, cfg->reconstruction, cfg->flux, dim).stage;
    if (!st)
        return fail(AVH_EINVAL,
                    "unsupported combination of equation, flux and grid dimension (burgers: godunov; "
                    "euler<n>: hll, hll3 with at most n active axes)");

    AVH_HIP(hipSetDevice(ctx->device));
    auto s = new (std::nothrow) avh_solver;
    if (!s) return fail(AVH_ENOMEM, "host allocation failed");
    s->ctx = ctx;
    s->cfg = *cfg;
    s->dim = dim;
    s->nv = nv;
    s->ng = ng;
    s->g[0] = ng;
    s->g[1] = dim > 1 ? ng : 0;
    s->g[2] = dim > 2 ? ng : 0;
    s->ncells = (long long)cfg->n[0] * cfg->n[1] * cfg->n[2];
    s->nghosted = (long long)(cfg->n[0] + 2 * s->g[0]) * (cfg->n[1] + 2 * s->g[1]) * (cfg->n[2] + 2 * s->g[2]);
    s->stage = st;
    s->ew = ew_for(cfg->equation);
    const size_t bytes = (size_t)nv * s->ncells * sizeof(double);
    if (hipMalloc(&s->A, bytes) != hipSuccess || hipMalloc(&s->B, bytes) != hipSuccess ||
        hipMalloc(&s->C, bytes) != hipSuccess ||
        hipMalloc(&s->staging, (size_t)nv * s->nghosted * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->ctrl, sizeof(Control)) != hipSuccess) {
        avh_solver_destroy(s);
        return fail(AVH_ENOMEM, "device allocation failed");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->stream = nullptr;
        avh_solver_destroy(s);
        return fail(AVH_EHIP, "hipStreamCreate failed");
    }
    if (hipMemset(s->A, 0, bytes) != hipSuccess || hipMemset(s->B, 0, bytes) != hipSuccess ||
        hipMemset(s->C, 0, bytes) != hipSuccess) {
        avh_solver_destroy(s);
        return fail(AVH_EHIP, "device initialisation failed");
    }
    // Grid::computeCellLengths: (top - origin)/N, 1 on single-cell axes
    // (alsfvm/src/grid/Grid.cpp:24-41)
    Control& h = s->h;
    std::memset(&h, 0, sizeof(Control));
    for (int d = 0; d < 3; ++d) h.cell[d] = (cfg->upper[d] - cfg->lower[d]) / double(cfg->n[d]);
    if (cfg->n[1] == 1) h.cell[1] = 1;
    if (cfg->n[2] == 1) h.cell[2] = 1;
    h.cfl = cfg->cfl;
    h.t_end = std::numeric_limits<double>::infinity();
    recompute_active(h);
    avh_status r = push(s);
    if (r != AVH_OK) {
        avh_solver_destroy(s);
        return r;
    }
    *out = s;
    return AVH_OK;
}

avh_status avh_ghost_cells(const avh_solver* s, int* ghosts) {
    if (!s || !ghosts) return fail(AVH_EINVAL, "null argument");
    *ghosts = s->ng;
    return AVH_OK;
}

avh_status avh_num_vars(const avh_solver* s, int* nvars) {
    if (!s || !nvars) return fail(AVH_EINVAL, "null argument");
    *nvars = s->nv;
    return AVH_OK;
}

avh_status avh_set_state(avh_solver* s, const double* host) {
    if (!s || !host) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    const int n0 = s->cfg.n[0], n1 = s->cfg.n[1], n2 = s->cfg.n[2];
    const long long t0 = n0 + 2 * s->g[0], t1 = n1 + 2 * s->g[1];
    std::vector<double> packed((size_t)s->nv * s->ncells);
    for (int v = 0; v < s->nv; ++v)
        for (int z = 0; z < n2; ++z)
            for (int y = 0; y < n1; ++y) {
                const double* src = host + (size_t)v * s->nghosted +
                                    ((size_t)(z + s->g[2]) * t1 + (y + s->g[1])) * t0 + s->g[0];
                double* dst = packed.data() + (size_t)v * s->ncells + ((size_t)z * n1 + y) * n0;
                std::memcpy(dst, src, (size_t)n0 * sizeof(double));
            }
    AVH_HIP(hipStreamSynchronize(s->stream));
    AVH_HIP(hipMemcpy(s->A, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice));
    return AVH_OK;
}

avh_status avh_set_state_device(avh_solver* s, const double* dev) {
    if (!s || !dev) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipMemcpyAsync(s->A, dev, (size_t)s->nv * s->ncells * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    AVH_HIP(hipStreamSynchronize(s->stream));
    return AVH_OK;
}

avh_status avh_get_state_device(avh_solver* s, double* dev) {
    if (!s || !dev) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipMemcpyAsync(dev, s->A, (size_t)s->nv * s->ncells * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    AVH_HIP(hipStreamSynchronize(s->stream));
    return AVH_OK;
}

avh_status avh_get_state(avh_solver* s, double* host) {
    if (!s || !host) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    long long nb = (s->nghosted + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(ghost_fill_kernel, dim3((unsigned)nb), dim3(256), 0, s->stream, s->A, s->staging, s->nv,
                       s->cfg.n[0], s->cfg.n[1], s->cfg.n[2], s->g[0], s->g[1], s->g[2], s->cfg.boundary);
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipMemcpyAsync(host, s->staging, (size_t)s->nv * s->nghosted * sizeof(double),
                           hipMemcpyDeviceToHost, s->stream));
    AVH_HIP(hipStreamSynchronize(s->stream));
    return AVH_OK;
}

avh_status avh_set_time(avh_solver* s, double time, int64_t steps) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    s->h.t = time;
    s->h.steps = steps;
    s->h.violated = 0;
    s->h.saves_pending = 0;
    recompute_active(s->h);
    return push(s);
}

avh_status avh_add_fixed_interval_adjuster(avh_solver* s, double interval, double start, int write_initial) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    if (!(interval > 0)) return fail(AVH_EINVAL, "interval must be positive");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    if (s->h.n_adj >= MAX_ADJ) return fail(AVH_EINVAL, "at most 4 fixed-interval adjusters");
    const int i = s->h.n_adj++;
    s->h.adj_interval[i] = interval;
    s->h.adj_start[i] = start;
    s->h.adj_saved[i] = write_initial ? 0 : 1;
    // the drivers' callWriters() at the current time (alsvinncli/src/main.cpp:210-213)
    if (s->h.t >= (double)s->h.adj_saved[i] * interval + start) {
        s->h.adj_saved[i] += 1;
        s->h.saves_pending += 1;
        recompute_active(s->h);
    }
    return push(s);
}

avh_status avh_set_end_time(avh_solver* s, double t_end) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    s->h.t_end = t_end;
    recompute_active(s->h);
    return push(s);
}

avh_status avh_enqueue_steps(avh_solver* s, int64_t steps) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    return enqueue_steps(s, steps);
}

avh_status avh_step_noalloc(avh_ctx* ctx, const avh_config* cfg, double* state, double dt_cap, double* dt_out,
                            int* violated_out, void* stream) {
    if (!ctx || !cfg || !state) return fail(AVH_EINVAL, "null argument");
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    const int ng = ghosts_of(cfg->reconstruction);
    if (!ng) return fail(AVH_EINVAL, "unknown reconstruction");
    if (cfg->integrator < AVH_INT_FORWARDEULER || cfg->integrator > AVH_INT_RUNGEKUTTA3)
        return fail(AVH_EINVAL, "unknown integrator");
    if (cfg->boundary != AVH_BC_PERIODIC && cfg->boundary != AVH_BC_NEUMANN)
        return fail(AVH_EINVAL, "unknown boundary");
    for (int d = 0; d < 3; ++d)
        if (cfg->n[d] < 1) return fail(AVH_EINVAL, "grid sizes must be >= 1");
    const int dim = cfg->n[2] > 1 ? 3 : (cfg->n[1] > 1 ? 2 : 1);
    for (int d = 0; d < dim; ++d)
        if (cfg->n[d] < ng)
            return fail(AVH_EINVAL, "too few cells: every active axis needs at least as many cells as ghost cells");
    const long long ncells = (long long)cfg->n[0] * cfg->n[1] * cfg->n[2];
    if ((long long)nv * ncells > NOALLOC_MAX_VALUES) {
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "state of %lld values (nvars * cells) exceeds the no-allocation scratch of %lld values",
                      (long long)nv * ncells, NOALLOC_MAX_VALUES);
        return fail(AVH_EINVAL, buf);
    }
    const StageLauncher stage = launchers_for(cfg->equation, cfg->reconstruction, cfg->flux, dim).stage;
    if (!stage)
        return fail(AVH_EINVAL, "unsupported combination of equation, flux and grid dimension");
    AVH_HIP(hipSetDevice(ctx->device));
    double* sp = nullptr;
    avh_status r = noalloc_scratch(ctx->device, &sp);
    if (r) return r;
    // tile shape: the autotuner's choice for this problem if a solver made
    // one, else a fixed default (timing would need events)
    int shape = dim == 1 ? 0 : (dim == 2 ? 3 : 1);
    int tuned;
    if (tune_lookup(TuneKey(0, cfg->equation * 16 + cfg->reconstruction, cfg->flux, dim, cfg->n[0], cfg->n[1],
                            cfg->n[2]),
                    &tuned))
        shape = tuned;
    Control* ctrl = (Control*)sp;
    double* B = sp + NOALLOC_CTRL_DOUBLES;
    double* C = B + NOALLOC_MAX_VALUES;
    const hipStream_t st = (hipStream_t)stream;
    double cell[3];
    for (int d = 0; d < 3; ++d) cell[d] = (cfg->upper[d] - cfg->lower[d]) / double(cfg->n[d]);
    if (cfg->n[1] == 1) cell[1] = 1;
    if (cfg->n[2] == 1) cell[2] = 1;
    hipLaunchKernelGGL(noalloc_init_kernel, dim3(1), dim3(64), 0, st, ctrl, cell[0], cell[1], cell[2], cfg->cfl,
                       dt_cap);
    AVH_HIP(hipGetLastError());
    const EwLauncher ew = ew_for(cfg->equation);
    StageParams P;
    P.ctrl = ctrl;
    P.ncells = ncells;
    for (int d = 0; d < 3; ++d) P.n[d] = cfg->n[d];
    P.bc = cfg->boundary;
    P.gamma = cfg->gamma;
    P.in = state;
    P.u0 = state;
    P.out = B;
    P.mode = MODE_L;
    AVH_HIP(stage(P, shape, st));
    if (cfg->integrator == AVH_INT_FORWARDEULER) {
        ew(B, state, state, ctrl, ncells, EW_FE_FINISH, cfg->gamma, st);
        AVH_HIP(hipGetLastError());
        hipLaunchKernelGGL(finish_step_kernel, dim3(1), dim3(64), 0, st, ctrl);
        AVH_HIP(hipGetLastError());
    } else if (cfg->integrator == AVH_INT_RUNGEKUTTA2) {
        ew(B, state, C, ctrl, ncells, EW_RK2_U1, cfg->gamma, st);
        AVH_HIP(hipGetLastError());
        P.in = C;
        P.out = state;
        P.mode = MODE_RK2_2;
        P.finish = 1;
        AVH_HIP(stage(P, shape, st));
    } else {
        P.in = B;
        P.out = C;
        P.mode = MODE_RK3_2;
        AVH_HIP(stage(P, shape, st));
        P.in = C;
        P.out = state;
        P.mode = MODE_RK3_3;
        P.finish = 1;
        AVH_HIP(stage(P, shape, st));
    }
    hipLaunchKernelGGL(noalloc_result_kernel, dim3(1), dim3(64), 0, st, ctrl, dt_out, violated_out);
    AVH_HIP(hipGetLastError());
    return AVH_OK;
}

avh_status avh_prepare_graph(avh_solver* s) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = ensure_shape(s))) return r;
    return ensure_step_graph(s->graph, s->graph_steps, s->stream, [s]() { return enqueue_step(s); });
}

avh_status avh_set_graph_steps(avh_solver* s, int steps) {
    if (!s || steps < 0 || steps > 4096) return fail(AVH_EINVAL, "bad argument");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(s->stream));
    s->graph.reset();
    s->graph_steps = steps;
    return AVH_OK;
}

avh_status avh_set_tile_shape(avh_solver* s, int shape) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    if (shape < -1 || shape >= num_shapes(s->dim)) return fail(AVH_EINVAL, "no such tile shape");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(s->stream));
    s->graph.reset();
    s->shape = shape;
    return AVH_OK;
}

avh_status avh_autotune(avh_solver* s, int* best_shape, double* ms_per_shape, int* nshapes) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    return autotune(s, best_shape, ms_per_shape, nshapes);
}

avh_status avh_synchronize(avh_solver* s) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(s->stream));
    return AVH_OK;
}

void* avh_stream(avh_solver* s) { return s ? (void*)s->stream : nullptr; }

avh_status avh_time_step_kernels(avh_solver* s, int64_t steps, double* ms, int* kernels) {
    if (!s || !ms || !kernels || steps < 1 || steps > 100000) return fail(AVH_EINVAL, "bad argument");
    avh_status r = use_device(s);
    if (r) return r;
    // launches per step: the stages, the element-wise pass of forward euler
    // / rk2, and forward euler's end-of-step kernel (the rk final stages end
    // the step themselves)
    const int nk = 3;
    if ((r = ensure_shape(s))) return r;
    EventMarks t;
    if ((r = t.create((size_t)steps * (nk + 1)))) return r;
    for (int64_t k = 0; k < steps; ++k)
        if ((r = enqueue_step(s, &t))) return r;
    if ((r = t.average(steps, nk, ms))) return r;
    *kernels = nk;
    return AVH_OK;
}

static void fill_progress(const avh_solver* s, avh_progress* p) {
    if (!p) return;
    p->time = s->h.t;
    p->last_dt = s->h.last_dt;
    p->steps = s->h.steps;
    p->violated = s->h.violated;
    p->saves_pending = s->h.saves_pending;
}

avh_status avh_get_progress(avh_solver* s, avh_progress* p) {
    if (!s || !p) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    fill_progress(s, p);
    return AVH_OK;
}

avh_status avh_ack_saves(avh_solver* s) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    s->h.saves_pending = 0;
    recompute_active(s->h);
    return push(s);
}

static avh_status constraint_error(const avh_solver* s) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "Simulation state does not obey constraints!  At time %.17g, number of timesteps performed %lld",
                  s->h.t, (long long)s->h.steps);
    return fail(AVH_ECONSTRAINT, buf);
}

avh_status avh_step(avh_solver* s, double* dt_out) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    if (s->h.violated) return constraint_error(s);
    // performStep runs regardless of the end time; saves stay counted
    const double t_end = s->h.t_end;
    const int pause = s->h.pause_on_save;
    s->h.t_end = std::numeric_limits<double>::infinity();
    s->h.pause_on_save = 0;
    s->h.active = 1;
    if ((r = push(s))) return r;
    if ((r = ensure_shape(s))) return r;
    if ((r = enqueue_step(s))) return r;
    if ((r = pull(s))) return r;
    s->h.t_end = t_end;
    s->h.pause_on_save = pause;
    recompute_active(s->h);
    if ((r = push(s))) return r;
    if (dt_out) *dt_out = s->h.last_dt;
    if (s->h.violated) return constraint_error(s);
    return AVH_OK;
}

avh_status avh_advance(avh_solver* s, double t_end, int64_t max_steps, int chunk, int pause_on_save,
                       avh_progress* progress) {
    if (!s) return fail(AVH_EINVAL, "null argument");
    if (chunk < 1) chunk = 1;
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = pull(s))) return r;
    if (s->h.violated) {
        fill_progress(s, progress);
        return constraint_error(s);
    }
    s->h.t_end = t_end;
    s->h.pause_on_save = pause_on_save ? 1 : 0;
    recompute_active(s->h);
    if ((r = push(s))) return r;
    int64_t done = 0;
    while (done < max_steps && s->h.active) {
        const int64_t k = (max_steps - done) < chunk ? (max_steps - done) : chunk;
        const int64_t before = s->h.steps;
        if ((r = enqueue_steps(s, k))) return r;
        if ((r = pull(s))) return r;
        done += s->h.steps - before;
        if (s->h.steps - before < k) break;  // the device loop ended inside the chunk
    }
    fill_progress(s, progress);
    if (s->h.violated) return constraint_error(s);
    return AVH_OK;
}

}  // extern "C"

// Device-side meanvar statistics for callers that run one solver per sample
// (the reference's Runner, alsuq/src/run/Runner.cpp:39-90): every save keeps
// a sample count and mean / M2 arrays of the conserved and the primitive
// variables.
struct avh_stats {
    avh_ctx* ctx = nullptr;
    int equation = 0, nv = 0, nsaves = 0;
    long long ncells = 0;
    double gamma = 1.4;
    std::vector<long long> count;
    double* mem = nullptr;  // [4][nsaves][nv][ncells]: mean_c, m2_c, mean_p, m2_p
    hipStream_t stream = nullptr;
};

extern "C" {

avh_status avh_stats_create(avh_ctx* ctx, int equation, const int n[3], double gamma, int nsaves, avh_stats** out) {
    if (!ctx || !n || !out) return fail(AVH_EINVAL, "null argument");
    *out = nullptr;
    const int nv = nvars_of(equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    if (nsaves < 1 || n[0] < 1 || n[1] < 1 || n[2] < 1) return fail(AVH_EINVAL, "bad sizes");
    AVH_HIP(hipSetDevice(ctx->device));
    auto t = new (std::nothrow) avh_stats;
    if (!t) return fail(AVH_ENOMEM, "host allocation failed");
    t->ctx = ctx;
    t->equation = equation;
    t->nv = nv;
    t->nsaves = nsaves;
    t->ncells = (long long)n[0] * n[1] * n[2];
    t->gamma = gamma;
    t->count.assign(nsaves, 0);
    const size_t bytes = sizeof(double) * 4 * (size_t)nsaves * nv * t->ncells;
    if (hipMalloc(&t->mem, bytes) != hipSuccess) {
        delete t;
        return fail(AVH_ENOMEM, "device allocation failed");
    }
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemset(t->mem, 0, bytes) != hipSuccess) {
        (void)hipFree(t->mem);
        if (t->stream) (void)hipStreamDestroy(t->stream);
        delete t;
        return fail(AVH_EHIP, "device initialisation failed");
    }
    *out = t;
    return AVH_OK;
}

avh_status avh_stats_destroy(avh_stats* t) {
    if (!t) return AVH_OK;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->stream);
    (void)hipFree(t->mem);
    (void)hipStreamDestroy(t->stream);
    delete t;
    return AVH_OK;
}

avh_status avh_stats_reset(avh_stats* t) {
    if (!t) return fail(AVH_EINVAL, "null argument");
    AVH_HIP(hipSetDevice(t->ctx->device));
    AVH_HIP(hipStreamSynchronize(t->stream));
    AVH_HIP(hipMemset(t->mem, 0, sizeof(double) * 4 * (size_t)t->nsaves * t->nv * t->ncells));
    std::fill(t->count.begin(), t->count.end(), 0);
    return AVH_OK;
}

avh_status avh_stats_accumulate(avh_stats* t, avh_solver* s, int save) {
    if (!t || !s) return fail(AVH_EINVAL, "null argument");
    if (save < 0 || save >= t->nsaves) return fail(AVH_EINVAL, "no such statistics save");
    if (s->cfg.equation != t->equation || s->ncells != t->ncells) return fail(AVH_EINVAL, "solver does not match");
    AVH_HIP(hipSetDevice(t->ctx->device));
    // the solver's state is final once its stream is idle
    AVH_HIP(hipStreamSynchronize(s->stream));
    const long long n = ++t->count[save];
    const size_t blk = (size_t)t->nsaves * t->nv * t->ncells;
    const size_t off = (size_t)save * t->nv * t->ncells;
    welford_for(t->equation)(s->A, t->ncells, (double)n, t->gamma, t->mem + off, t->mem + blk + off,
                             t->mem + 2 * blk + off, t->mem + 3 * blk + off, t->stream);
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipStreamSynchronize(t->stream));
    return AVH_OK;
}

avh_status avh_stats_read(avh_stats* t, int save, int primitive, double* mean, double* variance, int64_t* count) {
    if (!t || !mean || !variance || !count) return fail(AVH_EINVAL, "null argument");
    if (save < 0 || save >= t->nsaves) return fail(AVH_EINVAL, "no such statistics save");
    AVH_HIP(hipSetDevice(t->ctx->device));
    AVH_HIP(hipStreamSynchronize(t->stream));
    const size_t blk = (size_t)t->nsaves * t->nv * t->ncells;
    const size_t n = (size_t)t->nv * t->ncells;
    const size_t off = (size_t)save * n + (primitive ? 2 * blk : 0);
    AVH_HIP(hipMemcpy(mean, t->mem + off, n * sizeof(double), hipMemcpyDeviceToHost));
    AVH_HIP(hipMemcpy(variance, t->mem + blk + off, n * sizeof(double), hipMemcpyDeviceToHost));
    const long long c = t->count[save];
    for (size_t i = 0; i < n; ++i) variance[i] = c > 0 ? variance[i] / (double)c : 0.0;
    *count = c;
    return AVH_OK;
}

avh_status avh_arith_selftest(avh_ctx* ctx, int kind, uint64_t seed, int64_t n, int64_t* mismatches,
                              double* first_bad) {
    if (!ctx || !mismatches) return fail(AVH_EINVAL, "null argument");
    if (kind < ARITH_SQRT || kind > ARITH_FREXP) return fail(AVH_EINVAL, "unknown self-test kind");
    if (n < 0) return fail(AVH_EINVAL, "negative operand count");
    AVH_HIP(hipSetDevice(ctx->device));
    void* mem = nullptr;
    if (hipMalloc(&mem, 64) != hipSuccess) return fail(AVH_ENOMEM, "device allocation failed");
    unsigned long long* bad = (unsigned long long*)mem;
    double* first = (double*)((char*)mem + 16);
    hipError_t e = hipMemset(mem, 0, 64);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(arith_selftest_kernel, dim3(4096), dim3(256), 0, 0, kind, (unsigned long long)seed,
                           (long long)n, bad, first);
        e = hipGetLastError();
    }
    unsigned long long hbad = 0;
    double hfirst[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(&hbad, bad, sizeof hbad, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hfirst, first, sizeof hfirst, hipMemcpyDeviceToHost);
    (void)hipFree(mem);
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("arith self-test: ") + hipGetErrorString(e));
    *mismatches = (int64_t)hbad;
    if (first_bad) {
        first_bad[0] = hfirst[0];
        first_bad[1] = hfirst[1];
    }
    return AVH_OK;
}

avh_status avh_reconstruct(avh_solver* s, int direction, double* host_left, double* host_right) {
    if (!s || !host_left || !host_right) return fail(AVH_EINVAL, "null argument");
    if (direction < 0 || direction >= s->dim) return fail(AVH_EINVAL, "direction must be an active axis");
    avh_status r = use_device(s);
    if (r) return r;
    AVH_HIP(hipStreamSynchronize(s->stream));
    const size_t bytes = (size_t)s->nv * s->ncells * sizeof(double);
    // the B and C buffers are rewritten by every step; use them as scratch
    bool ok = false;
    switch (s->nv) {
    case 1: ok = launch_reconstruct_nv<1>(s->cfg.reconstruction, s->A, s->B, s->C, s->ncells, s->cfg.n, direction, s->cfg.boundary, s->stream); break;
    case 3: ok = launch_reconstruct_nv<3>(s->cfg.reconstruction, s->A, s->B, s->C, s->ncells, s->cfg.n, direction, s->cfg.boundary, s->stream); break;
    case 4: ok = launch_reconstruct_nv<4>(s->cfg.reconstruction, s->A, s->B, s->C, s->ncells, s->cfg.n, direction, s->cfg.boundary, s->stream); break;
    case 5: ok = launch_reconstruct_nv<5>(s->cfg.reconstruction, s->A, s->B, s->C, s->ncells, s->cfg.n, direction, s->cfg.boundary, s->stream); break;
    }
    if (!ok) return fail(AVH_EINVAL, "unsupported reconstruction");
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipMemcpyAsync(host_left, s->B, bytes, hipMemcpyDeviceToHost, s->stream));
    AVH_HIP(hipMemcpyAsync(host_right, s->C, bytes, hipMemcpyDeviceToHost, s->stream));
    AVH_HIP(hipStreamSynchronize(s->stream));
    return AVH_OK;
}

avh_status avh_compute_flux(avh_solver* s, double* host_out, double ws[3]) {
    if (!s || !host_out || !ws) return fail(AVH_EINVAL, "null argument");
    avh_status r = use_device(s);
    if (r) return r;
    if ((r = ensure_shape(s))) return r;
    if ((r = pull(s))) return r;
    StageParams P;
    P.ctrl = s->ctrl;
    P.ncells = s->ncells;
    for (int d = 0; d < 3; ++d) P.n[d] = s->cfg.n[d];
    P.bc = s->cfg.boundary;
    P.gamma = s->cfg.gamma;
    P.in = s->A;
    P.u0 = s->A;
    P.out = s->B;
    P.mode = MODE_FLUX_ONLY;
    AVH_HIP(s->stage(P, s->shape, s->stream));
    AVH_HIP(hipMemcpyAsync(host_out, s->B, (size_t)s->nv * s->ncells * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    if ((r = pull(s))) return r;
    for (int d = 0; d < 3; ++d) ws[d] = wave_speed(&s->h, d);
    clear_wave_speeds(&s->h);
    return push(s);
}

}  // extern "C"
