// The reference's per-call plug-in interfaces on device Volumes
// (include/alsvinn_hip.h, avh_numerical_flux ... avh_substep).
//
// A reference Volume keeps one array per conserved variable with its ghost
// cells, x fastest (alsfvm/src/volume/Volume.cpp:44-60).  The fused solver
// path keeps interior-only buffers and resolves ghosts through the boundary
// map; these entry points instead read and write the reference's layout, so
// that the reference's own Simulator, ConservedSystem and Integrator can
// drive the hip platform call by call:
//   numerical flux  the stage kernel in flux-only mode over a box of the
//                   interior, halo from the caller's ghost cells;
//   reconstruction  one thread per cell, a register window along the
//                   direction;
//   boundary        ghost cells from the interior by the boundary map;
//   constraints     every cell, ghosts included;
//   substep         numerical flux over the interior, dt from the wave
//                   speeds, then the integrator's Volume operations on
//                   every cell.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>

#include "../../include/alsvinn_hip.h"
#include "avh_internal.hpp"
#include "avh_stage3d.hpp"
#include "avh_stage_inst.hpp"

#pragma clang fp contract(off)

using namespace avh;

namespace {

struct Scratch {
    Control ctrl;  // wave-speed maxima
    int flag;      // constraint violation
};

// cells per axis of a volume, ghosts included
struct VolShape {
    int tot[3];
    long long cells;
};

VolShape shape_of(const avh_volume* v) {
    VolShape s;
    s.cells = 1;
    for (int d = 0; d < 3; ++d) {
        s.tot[d] = v->n[d] + 2 * v->ghosts[d];
        s.cells *= s.tot[d];
    }
    return s;
}

int dims_of(const int* n) { return n[2] > 1 ? 3 : (n[1] > 1 ? 2 : 1); }

// a volume with nv device arrays, n cells and at least min_ghost ghost cells
// on active axes (none on inactive ones)
avh_status check_volume(const avh_volume* v, int nv, int min_ghost, const char* what) {
    if (!v) return fail(AVH_EINVAL, std::string(what) + " is null");
    if (v->nvars != nv) return fail(AVH_EINVAL, std::string(what) + ": wrong number of variables");
    for (int i = 0; i < nv; ++i)
        if (!v->var[i]) return fail(AVH_EINVAL, std::string(what) + ": null variable array");
    const int dim = dims_of(v->n);
    for (int d = 0; d < 3; ++d) {
        if (v->n[d] < 1) return fail(AVH_EINVAL, std::string(what) + ": cell counts must be >= 1");
        if (d < dim) {
            if (v->ghosts[d] < min_ghost)
                return fail(AVH_EINVAL, std::string(what) + ": fewer ghost cells than the stencil needs");
            if (v->n[d] < v->ghosts[d])
                return fail(AVH_EINVAL, std::string(what) + ": fewer interior cells than ghost cells");
        } else if (v->ghosts[d] != 0) {
            return fail(AVH_EINVAL, std::string(what) + ": ghost cells on an inactive axis");
        }
    }
    return AVH_OK;
}

bool same_shape(const avh_volume* a, const avh_volume* b) {
    for (int d = 0; d < 3; ++d)
        if (a->n[d] != b->n[d] || a->ghosts[d] != b->ghosts[d]) return false;
    return true;
}

avh_status ctx_scratch(avh_ctx* ctx, Scratch** out) {
    if (!ctx->scratch) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(Scratch)) != hipSuccess) return fail(AVH_ENOMEM, "device allocation failed");
        ctx->scratch = p;
    }
    *out = (Scratch*)ctx->scratch;
    return AVH_OK;
}

// wave-speed maxima and flag to zero, by a kernel so that a captured call
// records no memset
__global__ void clear_scratch_kernel(Scratch* s) {
    if (threadIdx.x == 0) {
        clear_wave_speeds(&s->ctrl);
        s->flag = 0;
    }
}

// ---------------------------------------------------------------------------
// reconstruction

struct ReconVolParams {
    const double* in[5];
    double* left[5];
    double* right[5];
    int tot[3];
    int ghost[3];
    int org[3];  // first cell of the range, interior coordinates
    int ext[3];  // cells of the range per axis
    int dir;
    int indicator;
};

template <int RECON, int NV>
__global__ __launch_bounds__(256) void volume_reconstruct_kernel(ReconVolParams R) {
    constexpr int G = ReconTraits<RECON>::GHOST;
    constexpr int HW = G - 1 > 1 ? G - 1 : 1;
    constexpr int WZ = 2 * HW + 1;
    const long long total = (long long)R.ext[0] * R.ext[1] * R.ext[2];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        const int x = (int)(i % R.ext[0]) + R.org[0];
        const int y = (int)((i / R.ext[0]) % R.ext[1]) + R.org[1];
        const int z = (int)(i / ((long long)R.ext[0] * R.ext[1])) + R.org[2];
        const long long sx = 1, sy = R.tot[0], sz = (long long)R.tot[0] * R.tot[1];
        const long long c = (z + R.ghost[2]) * sz + (y + R.ghost[1]) * sy + (x + R.ghost[0]);
        const long long step = R.dir == 0 ? sx : (R.dir == 1 ? sy : sz);
        if (RECON == RECON_NONE) {
            // a copy of the cell itself (the range is the whole volume)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double q = R.in[v][c];
                R.left[v][c] = q;
                R.right[v][c] = q;
            }
            continue;
        }
        double w[WZ][NV];
#pragma unroll
        for (int k = 0; k < WZ; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) w[k][v] = R.in[v][c + (k - HW) * step];
        // ENO picks its stencil on the indicator: put it in slot 0
        const bool swap = (RECON == RECON_ENO2 || RECON == RECON_ENO3) && R.indicator != 0;
        if (swap) {
#pragma unroll
            for (int k = 0; k < WZ; ++k)
#pragma unroll
                for (int v = 1; v < NV; ++v)
                    if (v == R.indicator) {
                        const double t = w[k][0];
                        w[k][0] = w[k][v];
                        w[k][v] = t;
                    }
        }
        double l[NV], r[NV];
        recon_face_reg<RECON, NV, false, WZ>(w, HW, l);
        recon_face_reg<RECON, NV, true, WZ>(w, HW, r);
        if (swap) {
#pragma unroll
            for (int v = 1; v < NV; ++v)
                if (v == R.indicator) {
                    double t = l[0];
                    l[0] = l[v];
                    l[v] = t;
                    t = r[0];
                    r[0] = r[v];
                    r[v] = t;
                }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R.left[v][c] = l[v];
            R.right[v][c] = r[v];
        }
    }
}

template <int NV>
bool launch_volume_reconstruct(int recon, const ReconVolParams& R, unsigned nb, hipStream_t st) {
    switch (recon) {
    case RECON_NONE: hipLaunchKernelGGL((volume_reconstruct_kernel<RECON_NONE, NV>), dim3(nb), dim3(256), 0, st, R); return true;
    case RECON_WENO2: hipLaunchKernelGGL((volume_reconstruct_kernel<RECON_WENO2, NV>), dim3(nb), dim3(256), 0, st, R); return true;
    case RECON_ENO2: hipLaunchKernelGGL((volume_reconstruct_kernel<RECON_ENO2, NV>), dim3(nb), dim3(256), 0, st, R); return true;
    case RECON_ENO3: hipLaunchKernelGGL((volume_reconstruct_kernel<RECON_ENO3, NV>), dim3(nb), dim3(256), 0, st, R); return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// boundary, constraints, Volume operations

struct VolPtrs {
    double* v[5];
    int nv;
    int n[3];
    int g[3];
    int tot[3];
};

VolPtrs ptrs_of(const avh_volume* a) {
    VolPtrs p;
    for (int i = 0; i < 5; ++i) p.v[i] = i < a->nvars ? a->var[i] : nullptr;
    p.nv = a->nvars;
    for (int d = 0; d < 3; ++d) {
        p.n[d] = a->n[d];
        p.g[d] = a->ghosts[d];
        p.tot[d] = a->n[d] + 2 * a->ghosts[d];
    }
    return p;
}

// Every ghost cell takes the value of the interior cell the boundary map
// assigns to it; interior cells are only read, so the order does not matter.
__global__ __launch_bounds__(256) void volume_boundary_kernel(VolPtrs V, int bc) {
    const long long total = (long long)V.tot[0] * V.tot[1] * V.tot[2];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        const int x = (int)(i % V.tot[0]) - V.g[0];
        const int y = (int)((i / V.tot[0]) % V.tot[1]) - V.g[1];
        const int z = (int)(i / ((long long)V.tot[0] * V.tot[1])) - V.g[2];
        const bool ghost = x < 0 || x >= V.n[0] || y < 0 || y >= V.n[1] || z < 0 || z >= V.n[2];
        if (!ghost) continue;
        const int sx = bc_map(x, V.n[0], bc);
        const int sy = V.g[1] ? bc_map(y, V.n[1], bc) : y;
        const int sz = V.g[2] ? bc_map(z, V.n[2], bc) : z;
        const long long si = ((long long)(sz + V.g[2]) * V.tot[1] + (sy + V.g[1])) * V.tot[0] + (sx + V.g[0]);
        for (int v = 0; v < V.nv; ++v) V.v[v][i] = V.v[v][si];
    }
}

template <class EQ>
__global__ __launch_bounds__(256) void volume_constraints_kernel(VolPtrs V, double gamma, int* flag) {
    constexpr int NV = EQ::NV;
    const long long total = (long long)V.tot[0] * V.tot[1] * V.tot[2];
    bool bad = false;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        double q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = V.v[v][i];
        bad = bad || !EQ::obeys(q, gamma);
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flag, 1);
}

// The substep's Volume operations after the system evaluation, on every
// cell (RungeKutta3.cpp:49-86, RungeKutta2.cpp:37-67, ForwardEuler.cpp:32-51):
//   out = out * s + U_substep, then
//   rk2 substep 1: (out + U0) * 0.5
//   rk3 substep 1: (out * (1/3) + U0) * (3/4)
//   rk3 substep 2: (out * 2 + U0) * (1/3)
enum { COMBINE_PLAIN = 0, COMBINE_RK2_1 = 1, COMBINE_RK3_1 = 2, COMBINE_RK3_2 = 3 };

struct CombineParams {
    double* out[5];
    const double* us[5];
    const double* u0[5];
    int nv;
    long long cells;
    int mode;
    double s;
};

__global__ __launch_bounds__(256) void volume_combine_kernel(CombineParams C) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < C.cells; i += (long long)gridDim.x * 256) {
        for (int v = 0; v < C.nv; ++v) {
            double o = C.out[v][i] * C.s;
            o = o + C.us[v][i];
            if (C.mode == COMBINE_RK2_1) {
                o = o + C.u0[v][i];
                o = o * 0.5;
            } else if (C.mode == COMBINE_RK3_1) {
                o = o * (1. / 3.);
                o = o + C.u0[v][i];
                o = o * (3. / 4.);
            } else if (C.mode == COMBINE_RK3_2) {
                o = o * 2;
                o = o + C.u0[v][i];
                o = o * (1. / 3.);
            }
            C.out[v][i] = o;
        }
    }
}

unsigned grid_blocks(long long cells) {
    long long nb = (cells + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 8192) nb = 8192;
    return (unsigned)nb;
}

// flux over the box [start, n + end) into out; ws_host[3] the maxima
avh_status volume_flux(avh_ctx* ctx, const avh_config* cfg, const avh_volume* in, avh_volume* out, const int* start,
                       const int* end, double* ws_host, hipStream_t st) {
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    const int G = ghosts_of(cfg->reconstruction);
    if (!G) return fail(AVH_EINVAL, "unknown reconstruction");
    avh_status r;
    if ((r = check_volume(in, nv, G, "input volume"))) return r;
    if ((r = check_volume(out, nv, 0, "output volume"))) return r;
    for (int d = 0; d < 3; ++d)
        if (out->n[d] != in->n[d]) return fail(AVH_EINVAL, "input and output volumes differ in size");
    const int dim = dims_of(in->n);
    const Launchers L = launchers_for(cfg->equation, cfg->reconstruction, cfg->flux, dim);
    if (!L.volume)
        return fail(AVH_EINVAL, "unsupported combination of equation, flux and grid dimension");
    StageParams P;
    VolumeParams V;
    std::memset(&P, 0, sizeof P);
    std::memset(&V, 0, sizeof V);
    bool empty = false;
    for (int d = 0; d < 3; ++d) {
        const int s0 = d < dim ? start[d] : 0, e0 = d < dim ? end[d] : 0;
        if (d >= dim && (start[d] != 0 || end[d] != 0)) return fail(AVH_EINVAL, "start/end on an inactive axis");
        if (s0 < 0 || e0 > 0) return fail(AVH_EINVAL, "start must be >= 0 and end <= 0");
        const int ext = in->n[d] + e0 - s0;
        if (ext <= 0) empty = true;
        P.n[d] = ext;
        V.org[d] = s0;
        V.tot[d] = in->n[d] + 2 * in->ghosts[d];
        V.ghost[d] = in->ghosts[d];
        if (out->ghosts[d] != in->ghosts[d]) return fail(AVH_EINVAL, "input and output volumes differ in ghost cells");
    }
    for (int v = 0; v < nv; ++v) {
        V.in[v] = in->var[v];
        V.out[v] = out->var[v];
    }
    Scratch* sc = nullptr;
    if ((r = ctx_scratch(ctx, &sc))) return r;
    if (empty) {
        for (int d = 0; d < 3; ++d) ws_host[d] = 0.0;
        return AVH_OK;
    }
    P.ctrl = &sc->ctrl;
    P.mode = MODE_FLUX_ONLY;
    P.gamma = cfg->gamma;
    P.ncells = 0;
    hipLaunchKernelGGL(clear_scratch_kernel, dim3(1), dim3(64), 0, st, sc);
    AVH_HIP(hipGetLastError());
    AVH_HIP(L.volume(P, V, st));
    unsigned long long bits[3];
    AVH_HIP(hipMemcpyAsync(bits, sc->ctrl.ws_bits, sizeof bits, hipMemcpyDeviceToHost, st));
    AVH_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < 3; ++d) std::memcpy(&ws_host[d], &bits[d], sizeof(double));
    return AVH_OK;
}

}  // namespace

extern "C" {

avh_status avh_numerical_flux(avh_ctx* ctx, const avh_config* cfg, const avh_volume* in, avh_volume* out,
                              const int start[3], const int end[3], double ws[3], void* stream) {
    if (!ctx || !cfg || !start || !end || !ws) return fail(AVH_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(ctx->lock);
    AVH_HIP(hipSetDevice(ctx->device));
    return volume_flux(ctx, cfg, in, out, start, end, ws, (hipStream_t)stream);
}

avh_status avh_reconstruction(avh_ctx* ctx, const avh_config* cfg, const avh_volume* in, int direction,
                              int indicator_variable, avh_volume* left, avh_volume* right, const int start[3],
                              const int end[3], void* stream) {
    if (!ctx || !cfg || !start || !end) return fail(AVH_EINVAL, "null argument");
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    const int G = ghosts_of(cfg->reconstruction);
    if (!G) return fail(AVH_EINVAL, "unknown reconstruction");
    avh_status r;
    if ((r = check_volume(in, nv, G, "input volume"))) return r;
    if ((r = check_volume(left, nv, 0, "left volume"))) return r;
    if ((r = check_volume(right, nv, 0, "right volume"))) return r;
    if (!same_shape(in, left) || !same_shape(in, right)) return fail(AVH_EINVAL, "volumes differ in shape");
    const int dim = dims_of(in->n);
    if (direction < 0 || direction >= dim) return fail(AVH_EINVAL, "direction must be an active axis");
    if (indicator_variable < 0 || indicator_variable >= nv) return fail(AVH_EINVAL, "no such indicator variable");
    ReconVolParams R;
    std::memset(&R, 0, sizeof R);
    long long total = 1;
    for (int d = 0; d < 3; ++d) {
        if (d >= dim && (start[d] != 0 || end[d] != 0)) return fail(AVH_EINVAL, "start/end on an inactive axis");
        if (d < dim && (start[d] < 0 || end[d] > 0)) return fail(AVH_EINVAL, "start must be >= 0 and end <= 0");
        R.tot[d] = in->n[d] + 2 * in->ghosts[d];
        R.ghost[d] = in->ghosts[d];
        if (cfg->reconstruction == AVH_RECON_NONE) {
            // NoReconstruction copies every cell of the volume, ghost cells
            // included, whatever the range (NoReconstruction.cpp:46-60)
            R.org[d] = -in->ghosts[d];
            R.ext[d] = R.tot[d];
        } else {
            const int widen = d == direction ? 1 : 0;
            R.org[d] = start[d] - widen;
            R.ext[d] = in->n[d] + end[d] + widen - R.org[d];
            if (R.ext[d] < 0) R.ext[d] = 0;
        }
        total *= R.ext[d];
    }
    for (int v = 0; v < nv; ++v) {
        R.in[v] = in->var[v];
        R.left[v] = left->var[v];
        R.right[v] = right->var[v];
    }
    R.dir = direction;
    R.indicator = indicator_variable;
    std::lock_guard<std::mutex> g(ctx->lock);
    AVH_HIP(hipSetDevice(ctx->device));
    if (total == 0) return AVH_OK;
    const hipStream_t st = (hipStream_t)stream;
    bool ok = false;
    switch (nv) {
    case 1: ok = launch_volume_reconstruct<1>(cfg->reconstruction, R, grid_blocks(total), st); break;
    case 3: ok = launch_volume_reconstruct<3>(cfg->reconstruction, R, grid_blocks(total), st); break;
    case 4: ok = launch_volume_reconstruct<4>(cfg->reconstruction, R, grid_blocks(total), st); break;
    case 5: ok = launch_volume_reconstruct<5>(cfg->reconstruction, R, grid_blocks(total), st); break;
    }
    if (!ok) return fail(AVH_EINVAL, "unsupported reconstruction");
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipStreamSynchronize(st));
    return AVH_OK;
}

avh_status avh_apply_boundary(avh_ctx* ctx, int boundary, avh_volume* volume, void* stream) {
    if (!ctx || !volume) return fail(AVH_EINVAL, "null argument");
    if (boundary != AVH_BC_PERIODIC && boundary != AVH_BC_NEUMANN) return fail(AVH_EINVAL, "unknown boundary");
    if (volume->nvars < 1 || volume->nvars > 5) return fail(AVH_EINVAL, "1 to 5 variables");
    avh_status r;
    if ((r = check_volume(volume, volume->nvars, 0, "volume"))) return r;
    std::lock_guard<std::mutex> g(ctx->lock);
    AVH_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = (hipStream_t)stream;
    const VolShape s = shape_of(volume);
    hipLaunchKernelGGL(volume_boundary_kernel, dim3(grid_blocks(s.cells)), dim3(256), 0, st, ptrs_of(volume),
                       boundary);
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipStreamSynchronize(st));
    return AVH_OK;
}

avh_status avh_obeys_constraints(avh_ctx* ctx, const avh_config* cfg, const avh_volume* volume, int* obeys,
                                 void* stream) {
    if (!ctx || !cfg || !volume || !obeys) return fail(AVH_EINVAL, "null argument");
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    avh_status r;
    if ((r = check_volume(volume, nv, 0, "volume"))) return r;
    std::lock_guard<std::mutex> g(ctx->lock);
    AVH_HIP(hipSetDevice(ctx->device));
    Scratch* sc = nullptr;
    if ((r = ctx_scratch(ctx, &sc))) return r;
    const hipStream_t st = (hipStream_t)stream;
    const VolShape s = shape_of(volume);
    const VolPtrs V = ptrs_of(volume);
    hipLaunchKernelGGL(clear_scratch_kernel, dim3(1), dim3(64), 0, st, sc);
    const unsigned nb = grid_blocks(s.cells);
    switch (cfg->equation) {
    case AVH_EQ_BURGERS: hipLaunchKernelGGL(volume_constraints_kernel<BurgersEq>, dim3(nb), dim3(256), 0, st, V, cfg->gamma, &sc->flag); break;
    case AVH_EQ_EULER1: hipLaunchKernelGGL(volume_constraints_kernel<EulerEq<1>>, dim3(nb), dim3(256), 0, st, V, cfg->gamma, &sc->flag); break;
    case AVH_EQ_EULER2: hipLaunchKernelGGL(volume_constraints_kernel<EulerEq<2>>, dim3(nb), dim3(256), 0, st, V, cfg->gamma, &sc->flag); break;
    case AVH_EQ_EULER3: hipLaunchKernelGGL(volume_constraints_kernel<EulerEq<3>>, dim3(nb), dim3(256), 0, st, V, cfg->gamma, &sc->flag); break;
    }
    AVH_HIP(hipGetLastError());
    int flag = 0;
    AVH_HIP(hipMemcpyAsync(&flag, &sc->flag, sizeof flag, hipMemcpyDeviceToHost, st));
    AVH_HIP(hipStreamSynchronize(st));
    *obeys = flag ? 0 : 1;
    return AVH_OK;
}

avh_status avh_substep(avh_ctx* ctx, const avh_config* cfg, const avh_volume* const* inputs, int substep,
                       avh_volume* out, double dt, double dt_cap, double* dt_out, void* stream) {
    if (!ctx || !cfg || !inputs || !out || !dt_out) return fail(AVH_EINVAL, "null argument");
    const int nsub = cfg->integrator == AVH_INT_FORWARDEULER ? 1 : (cfg->integrator == AVH_INT_RUNGEKUTTA2 ? 2 : 3);
    if (cfg->integrator < AVH_INT_FORWARDEULER || cfg->integrator > AVH_INT_RUNGEKUTTA3)
        return fail(AVH_EINVAL, "unknown integrator");
    if (substep < 0 || substep >= nsub) return fail(AVH_EINVAL, "no such substep for this integrator");
    const int nv = nvars_of(cfg->equation);
    if (!nv) return fail(AVH_EINVAL, "unknown equation");
    avh_status r;
    for (int k = 0; k <= substep; ++k) {
        if ((r = check_volume(inputs[k], nv, 0, "input volume"))) return r;
        if (!same_shape(inputs[k], out)) return fail(AVH_EINVAL, "volumes differ in shape");
    }
    std::lock_guard<std::mutex> g(ctx->lock);
    AVH_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = (hipStream_t)stream;
    // ConservedSystem: the interior plus the side strips equal one call over
    // the whole interior (start = end = 0)
    const int zero[3] = {0, 0, 0};
    double ws[3];
    if ((r = volume_flux(ctx, cfg, inputs[substep], out, zero, zero, ws, st))) return r;
    // Grid::computeCellLengths (alsfvm/src/grid/Grid.cpp:24-41): 1 on
    // single-cell axes
    double cell[3];
    for (int d = 0; d < 3; ++d) cell[d] = (cfg->upper[d] - cfg->lower[d]) / double(out->n[d]);
    if (out->n[1] == 1) cell[1] = 1;
    if (out->n[2] == 1) cell[2] = 1;
    if (substep == 0) {
        double total = 0;
        for (int d = 0; d < 3; ++d) {
            if (cell[d] == 0) continue;
            total += ws[d] / cell[d];
        }
        dt = cfg->cfl / total;
        dt = (dt_cap < dt) ? dt_cap : dt;
    }
    CombineParams C;
    std::memset(&C, 0, sizeof C);
    for (int v = 0; v < nv; ++v) {
        C.out[v] = out->var[v];
        C.us[v] = inputs[substep]->var[v];
        C.u0[v] = inputs[0]->var[v];
    }
    C.nv = nv;
    C.cells = shape_of(out).cells;
    C.s = dt / cell[0];
    C.mode = COMBINE_PLAIN;
    if (cfg->integrator == AVH_INT_RUNGEKUTTA2 && substep == 1) C.mode = COMBINE_RK2_1;
    if (cfg->integrator == AVH_INT_RUNGEKUTTA3 && substep == 1) C.mode = COMBINE_RK3_1;
    if (cfg->integrator == AVH_INT_RUNGEKUTTA3 && substep == 2) C.mode = COMBINE_RK3_2;
    hipLaunchKernelGGL(volume_combine_kernel, dim3(grid_blocks(C.cells)), dim3(256), 0, st, C);
    AVH_HIP(hipGetLastError());
    AVH_HIP(hipStreamSynchronize(st));
    *dt_out = dt;
    return AVH_OK;
}

}  // extern "C"
