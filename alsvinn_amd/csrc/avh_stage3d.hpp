// Fused FP64 stage kernel for 3D grids: a TX x TY column tile marches along z.
//
// Per z-plane k of its chunk [z0, z1) a workgroup
//   * writes plane k (its own columns from registers, the x/y halo ring from
//     HBM through the boundary map) into an LDS plane buffer,
//   * evaluates the x- and y-interface fluxes of the plane from LDS (one work
//     item per interface, as in the 2D kernel) and accumulates them per cell,
//   * evaluates the z-interface flux F(k+1/2) of each column in registers from
//     a sliding window of planes [k+1-HW, k+1+HW] and the right face of
//     plane k kept from the previous iteration; F(k-1/2) is likewise kept,
//   * adds ((-F_z(k+1/2)) + F_z(k-1/2)) after the x and y terms -- the same
//     x, y, z accumulation order as NumericalFluxCPU::computeFlux
//     (alsfvm/src/numflux/NumericalFluxCPU.cpp:170-221) -- and applies the RK
//     stage combination.
// Each plane is read once per workgroup plus its x/y halo ring; no z halo is
// ever reloaded, and the tile has no z-direction quantisation loss.
#pragma once
#include "avh_stage.hpp"

#pragma clang fp contract(off)

namespace avh {

// ENO / WENO2 / none on a register window: u[i][v], stencil centre `c`
// (compile-time), unit stride.  Selects replace ENO's data-dependent index.
// OPS as in recon_face_ops (the WENO2 weights may share a reciprocal).
template <int RECON, int NV, bool RIGHT, int WZ, class OPS = PlainOps>
__device__ __forceinline__ void recon_face_reg(const double (&u)[WZ][NV], const int c, double* out,
                                               bool& ok) {
    if (RECON == RECON_NONE) {
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = u[c][v];
    } else if (RECON == RECON_WENO2) {
        const double eps = 1e-8;
        const double w0 = u[c][0];
        const double dp = u[c + 1][0] - w0;
        const double b0 = dp * dp;
        const double dm = w0 - u[c - 1][0];
        const double b1 = dm * dm;
        const double d0 = 3 * (eps + b0) * (eps + b0);
        const double d1 = 3 * (eps + b1) * (eps + b1);
        if (RIGHT) {
            const double a0 = 2 / d0;
            const double a1 = 1 / d1;
            double w0r, w1r;
            weno2_weights<OPS>(a0, a1, w0r, w1r, ok);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                out[v] = 0.5 * (w0r * u[c + 1][v] + (3 * w1r + w0r) * u[c][v] - w1r * u[c - 1][v]);
        } else {
            const double a0 = 1 / d0;
            const double a1 = 2 / d1;
            double w0l, w1l;
            weno2_weights<OPS>(a0, a1, w0l, w1l, ok);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                out[v] = 0.5 * (w1l * u[c - 1][v] + (3 * w0l + w1l) * u[c][v] - w0l * u[c + 1][v]);
        }
    } else {
        constexpr int ORDER = RECON == RECON_ENO2 ? 2 : 3;
        int shift = 0;
        {
            const double ddr = u[c + 1][0] - u[c][0];
            const double ddl = u[c][0] - u[c - 1][0];
            if (fabs(ddl) < fabs(ddr)) shift++;
        }
        if (ORDER == 3) {
            const double d0p = u[c + 2 < WZ ? c + 2 : WZ - 1][0] - u[c + 1][0];
            const double d00 = u[c + 1][0] - u[c][0];
            const double d0m = u[c][0] - u[c - 1][0];
            const double ddr = d0p - d00;
            const double ddl = d00 - d0m;
            if (fabs(ddl) < fabs(ddr)) shift++;
        }
        const int r = RIGHT ? shift + 1 : shift;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < ORDER; ++j) {
                // value at offset j - shift, shift in [0, ORDER-1]
                double val = u[c + j][v];
#pragma unroll
                for (int sft = 1; sft < ORDER; ++sft) {
                    const int idx = c + j - sft;
                    val = (shift == sft) ? u[idx < 0 ? 0 : idx][v] : val;
                }
                acc += eno_coef<ORDER>(r, j) * val;
            }
            out[v] = acc;
        }
    }
}

template <int RECON, int NV, bool RIGHT, int WZ>
__device__ __forceinline__ void recon_face_reg(const double (&u)[WZ][NV], const int c, double* out) {
    bool ok = true;
    recon_face_reg<RECON, NV, RIGHT, WZ, PlainOps>(u, c, out, ok);
}

// Both faces of the window's cell c.  WENO2 under SharedOps divides once per
// denominator: the left face takes a0 = 1/d0, a1 = 2/d1 and the right face
// a0 = 2/d0, a1 = 1/d1, and 2/d is formed as 2 * (1/d).  Doubling is exact
// and commutes with rounding while 2/d stays normal; the weight checks
// (every a0, a1 >= 2^-400, so 1/d0, 1/d1 >= 2^-400 and d0, d1 <= 2^400) hold
// only there, so a pair that passes them has the bits of the two divisions.
template <int RECON, int NV, int WZ, class OPS>
__device__ __forceinline__ void recon_faces_reg(const double (&u)[WZ][NV], const int c, double* outl,
                                                double* outr, bool& ok) {
    if constexpr (RECON == RECON_WENO2 && OPS::SHARED) {
        const double eps = 1e-8;
        const double w0 = u[c][0];
        const double dp = u[c + 1][0] - w0;
        const double b0 = dp * dp;
        const double dm = w0 - u[c - 1][0];
        const double b1 = dm * dm;
        const double d0 = 3 * (eps + b0) * (eps + b0);
        const double d1 = 3 * (eps + b1) * (eps + b1);
        const double r0 = 1 / d0;
        const double r1 = 1 / d1;
        double w0l, w1l, w0r, w1r;
        weno2_weights<OPS>(r0, 2 * r1, w0l, w1l, ok);
        weno2_weights<OPS>(2 * r0, r1, w0r, w1r, ok);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            outl[v] = 0.5 * (w1l * u[c - 1][v] + (3 * w0l + w1l) * u[c][v] - w0l * u[c + 1][v]);
            outr[v] = 0.5 * (w0r * u[c + 1][v] + (3 * w1r + w0r) * u[c][v] - w1r * u[c - 1][v]);
        }
    } else {
        recon_face_reg<RECON, NV, false, WZ, OPS>(u, c, outl, ok);
        recon_face_reg<RECON, NV, true, WZ, OPS>(u, c, outr, ok);
    }
}

// TWO_BARRIERS: the plane tile is double-buffered and the x- and y-interface
// fluxes have LDS buffers of their own, so a plane takes two barriers (plane
// published; fluxes published) instead of five, and the x, y and z flux
// evaluations of a plane run in one barrier-free stretch.
// MODE_CT >= 0 compiles the kernel for that StageMode alone (P.mode must
// equal it): the other modes' branches and registers drop out.
template <class EQ, int RECON, class FLUX, int TX, int TY, int LZ, int NTH, bool TWO_BARRIERS = false, int MODE_CT = -1>
struct MarchKernel3 {
    static constexpr int NV = EQ::NV;
    static constexpr int NC = TX * TY;  // columns: one per thread, threads past them work on the planes only
    static constexpr int NT = NTH;
    static_assert(NT >= NC && NT % 64 == 0, "one thread per column, whole wavefronts");
    static constexpr int G = ReconTraits<RECON>::GHOST;
    static constexpr int HW = G - 1 > 1 ? G - 1 : 1;  // z-window half width
    static constexpr int WZ = 2 * HW + 1;              // planes k+1-HW .. k+1+HW
    static constexpr int LXP = TX + 2 * G;
    static constexpr int LYP = TY + 2 * G;
    static constexpr int LP = LXP * LYP;  // plane cells incl. halo
    static constexpr int NIX = (TX + 1) * TY;
    static constexpr int NIY = TX * (TY + 1);
    static constexpr int NIMAX = NIX > NIY ? NIX : NIY;
    static constexpr int NHALO = LP - TX * TY;

    __host__ __device__ static unsigned int tiles(const int* n) {
        return (unsigned int)(((n[0] + TX - 1) / TX) * ((n[1] + TY - 1) / TY) * ((n[2] + LZ - 1) / LZ));
    }

    // Cell indices.  The kernels compiled for the first two RK3 stages index
    // with 32-bit values (their launcher takes them only for grids of fewer
    // than 2^29 cells, so byte offsets stay below 2^32): each load is then a
    // 32-bit offset from a variable's base held in scalar registers, with no
    // 64-bit address arithmetic per lane.  (The final stage, which also
    // stores, runs out of scalar registers that way and is slower.)
    static constexpr bool OFF32 = MODE_CT == MODE_L || MODE_CT == MODE_RK3_2;
    typedef typename std::conditional<OFF32, unsigned int, long long>::type Idx;
    __device__ __forceinline__ static Idx gidx(int x, int y, int z, int n0, int n1) {
        return ((Idx)z * n1 + y) * n0 + x;
    }
    __device__ __forceinline__ static double ld(const double* base, long long ncells, int v, Idx gi) {
        return soa_ld<OFF32>(base, ncells, v, gi);
    }
    __device__ __forceinline__ static void st(double* base, long long ncells, int v, Idx gi, double x) {
        soa_st<OFF32>(base, ncells, v, gi, x);
    }

    // one input value at a (boundary-mapped) global cell; in the on-the-fly
    // stage U0's values are also written to `keep` (stride NC) when given
    __device__ __forceinline__ static void fetch(const StageParams& P, bool otf, double s, Idx gi, double* q,
                                                 double* keep = nullptr) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double val = ld(P.in, P.ncells, v, gi);
            if (otf) {
                const double u = ld(P.u0, P.ncells, v, gi);
                if (keep) keep[v * NC] = u;
                val = val * s + u;
            }
            q[v] = val;
        }
    }

    // The kernel compiled for the on-the-fly stage (U1 = L0*s + U0 formed as
    // the tile loads) reads the column's U0 with each window plane; its RK
    // combination needs the same U0 again RING planes later, by which time
    // the line has left L2.  A thread-private LDS ring of RING planes keeps
    // it, where LDS still allows two workgroups per CU (the register budget
    // allows no more).
    static constexpr int RING = 2 + HW;
    static constexpr int LDS_BASE =
        8 * ((TWO_BARRIERS ? 2 : 1) * NV * LP + (TWO_BARRIERS ? 2 : 1) * NV * NIMAX + 3 * (NT / 64));
    static constexpr bool U0RING = MODE_CT == MODE_RK3_2 && LDS_BASE + 8 * RING * NV * NC <= 80 * 1024;

    // interface between plane cells ia and ia + SD under OPS
    template <class OPS, int DIR>
    __device__ __forceinline__ static double plane_interface(const double* sp, int ia, int SD, double gamma,
                                                             double* F, bool& ok) {
        double qa[NV], qb[NV];
        recon_face_ops<RECON, NV, true, OPS>(sp, LP, ia, SD, qa, ok);
        recon_face_ops<RECON, NV, false, OPS>(sp, LP, ia + SD, SD, qb, ok);
        const typename EQ::All aL = EQ::template make_all_ops<OPS>(qa, gamma, ok);
        const typename EQ::All aR = EQ::template make_all_ops<OPS>(qb, gamma, ok);
        return FLUX::template flux_ops<OPS, DIR, EQ>(aL, aR, gamma, F, ok);
    }

    template <int DIR>
    __device__ __forceinline__ static bool plane_item(int it, int ox, int oy, const int* n, int& ia) {
        int cx, cy;
        if (DIR == 0) {
            cx = it % (TX + 1) - 1;
            cy = it / (TX + 1);
        } else {
            cx = it % TX;
            cy = it / TX - 1;
        }
        ia = (cy + G) * LXP + (cx + G);
        return (ox + cx < n[0]) && (oy + cy < n[1]);
    }

    // A FAST flux evaluates each interface with SharedOps and redoes those
    // whose operands leave the guarded range with PlainOps (as the 2D stage
    // kernel's flux_pass); otherwise PlainOps throughout.
    template <int DIR>
    __device__ __forceinline__ static void plane_pass(const double* sp, double* sf, int ox, int oy, const int* n,
                                                      double gamma, double& ws) {
        constexpr int NI = DIR == 0 ? NIX : NIY;
        constexpr int SD = DIR == 0 ? 1 : LXP;
        static_assert((NI + NT - 1) / NT <= 32, "one redo bit per turn");
        typedef typename OpsFor<FLUX::FAST>::type Ops;
        unsigned int redo = 0;
        int m = 0;
        for (int it = threadIdx.x; it < NI; it += NT, ++m) {
            int ia;
            if (!plane_item<DIR>(it, ox, oy, n, ia)) continue;
            double F[NV];
            bool ok = true;
            const double w = plane_interface<Ops, DIR>(sp, ia, SD, gamma, F, ok);
            if (ok) ws = ws_max(ws, w);
            else redo |= 1u << m;
#pragma unroll
            for (int v = 0; v < NV; ++v) sf[v * NIMAX + it] = F[v];
        }
        if (FLUX::FAST && redo != 0) {
            m = 0;
            for (int it = threadIdx.x; it < NI; it += NT, ++m) {
                int ia;
                if (!((redo >> m) & 1u) || !plane_item<DIR>(it, ox, oy, n, ia)) continue;
                double F[NV];
                bool ok = true;
                ws = ws_max(ws, plane_interface<PlainOps, DIR>(sp, ia, SD, gamma, F, ok));
#pragma unroll
                for (int v = 0; v < NV; ++v) sf[v * NIMAX + it] = F[v];
            }
        }
    }

    __device__ static void run(const StageParams& P) {
        constexpr int NBUF = TWO_BARRIERS ? 2 : 1;
        __shared__ double sp_buf[NBUF][NV * LP];
        __shared__ double sf[NV * NIMAX];
        __shared__ double sf_y[TWO_BARRIERS ? NV * NIMAX : 1];
        __shared__ double s_ws[3][NT / 64];
        __shared__ double su0[U0RING ? RING * NV * NC : 1];
        double* const sfx = sf;
        double* const sfy = TWO_BARRIERS ? sf_y : sf;

        Control* ctrl = P.ctrl;
        const int mode = MODE_CT >= 0 ? MODE_CT : P.mode;
        if (!ctrl->active && mode != MODE_FLUX_ONLY) return;

        const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
        const int ntx = (n0 + TX - 1) / TX, nty = (n1 + TY - 1) / TY;
        const unsigned int tile = tile_id_xcd(blockIdx.x, gridDim.x);
        const int tx = tile % ntx;
        const int ty = (tile / ntx) % nty;
        const int tz = tile / (ntx * nty);
        const int ox = tx * TX, oy = ty * TY, z0 = tz * LZ;
        const int z1 = (z0 + LZ < n2) ? z0 + LZ : n2;
        const bool otf = (mode == MODE_RK3_2);
        const double s = (mode == MODE_FLUX_ONLY || mode == MODE_L) ? 0.0 : step_dt(ctrl) / ctrl->cell[0];

        // this thread's column.  A column past the grid edge (partial tile)
        // sits where the tile's x/y halo lies, so it loads the boundary image
        // of that position, exactly as the halo ring does.
        const bool has_col = (int)threadIdx.x < NC;
        const int cx = threadIdx.x % TX, cy = has_col ? (int)threadIdx.x / TX : 0;
        const bool col_ok = has_col && (ox + cx < n0) && (oy + cy < n1);
        const int gxc = bc_map(ox + cx < n0 + G ? ox + cx : n0 + G - 1, n0, P.bc);
        const int gyc = bc_map(oy + cy < n1 + G ? oy + cy : n1 + G - 1, n1, P.bc);

        double ws[3] = {0.0, 0.0, 0.0};
        double w[WZ][NV];
        double rprev[NV], fprev[NV];
#pragma unroll
        for (int i = 0; i < WZ; ++i)
#pragma unroll
            for (int v = 0; v < NV; ++v) w[i][v] = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) rprev[v] = fprev[v] = 0.0;

        // prologue: window centred at z0-1 -> R(z0-1); then centred at z0 ->
        // L(z0), R(z0), F_z(z0-1/2); then centred at z0+1 for the loop
        // ring slot of the column's plane p (z0 <= p): U0 at (p - z0) % RING
        const bool ring = U0RING && col_ok;
        double* const su0c = su0 + (U0RING ? threadIdx.x : 0);
        auto slot = [&](int p) { return su0c + ((p - z0) % RING) * NV * NC; };
        if (has_col) {
#pragma unroll
            for (int i = 0; i < WZ; ++i) {
                const int gz = bc_map(z0 - 1 - HW + i, n2, P.bc);
                fetch(P, otf, s, gidx(gxc, gyc, gz, n0, n1), w[i], (ring && i > HW) ? slot(z0 - 1 - HW + i) : nullptr);
            }
            double rm1[NV];
            recon_face_reg<RECON, NV, true, WZ>(w, HW, rm1);
#pragma unroll
            for (int i = 0; i < WZ - 1; ++i)
#pragma unroll
                for (int v = 0; v < NV; ++v) w[i][v] = w[i + 1][v];
            fetch(P, otf, s, gidx(gxc, gyc, bc_map(z0 + HW, n2, P.bc), n0, n1), w[WZ - 1],
                  ring ? slot(z0 + HW) : nullptr);
            if (col_ok) {
                double l0[NV];
                recon_face_reg<RECON, NV, false, WZ>(w, HW, l0);
                recon_face_reg<RECON, NV, true, WZ>(w, HW, rprev);
                const typename EQ::All aL = EQ::make_all(rm1, P.gamma);
                const typename EQ::All aR = EQ::make_all(l0, P.gamma);
                const double wz = FLUX::template flux<2, EQ>(aL, aR, P.gamma, fprev);
                ws[2] = ws_max(ws[2], wz);
            }
#pragma unroll
            for (int i = 0; i < WZ - 1; ++i)
#pragma unroll
                for (int v = 0; v < NV; ++v) w[i][v] = w[i + 1][v];
            fetch(P, otf, s, gidx(gxc, gyc, bc_map(z0 + 1 + HW, n2, P.bc), n0, n1), w[WZ - 1],
                  ring ? slot(z0 + 1 + HW) : nullptr);
        }

        bool bad = false;
        const bool final_stage = (mode == MODE_RK2_2 || mode == MODE_RK3_3);
        // This thread's cells of the x/y halo ring (HPT per thread at most):
        // their LDS slots and boundary-mapped positions in a plane do not
        // change with k.  The ring rows above and below the tile come first,
        // then the side columns.
        constexpr int HPT = (NHALO + NT - 1) / NT;
        int halo_lds[HPT];
        Idx halo_xy[HPT];
        bool has_halo[HPT];
#pragma unroll
        for (int j = 0; j < HPT; ++j) {
            const int h = threadIdx.x + j * NT;
            has_halo[j] = h < NHALO;
            halo_lds[j] = 0;
            halo_xy[j] = 0;
            if (has_halo[j]) {
                int lx, ly;
                if (h < G * LXP) {
                    ly = h / LXP;
                    lx = h % LXP;
                } else if (h < 2 * G * LXP) {
                    ly = TY + G + (h - G * LXP) / LXP;
                    lx = (h - G * LXP) % LXP;
                } else {
                    const int r = h - 2 * G * LXP;  // 2G side cells per interior row
                    ly = G + r / (2 * G);
                    const int c2 = r % (2 * G);
                    lx = c2 < G ? c2 : TX + c2;
                }
                int gx = ox + lx - G, gy = oy + ly - G;
                gx = gx < n0 + G ? gx : n0 + G - 1;
                gy = gy < n1 + G ? gy : n1 + G - 1;
                gx = bc_map(gx, n0, P.bc);
                gy = bc_map(gy, n1, P.bc);
                halo_lds[j] = ly * LXP + lx;
                halo_xy[j] = (Idx)gy * n0 + gx;
            }
        }
        const Idx plane = (Idx)n0 * n1;
        // the ring of the next plane is loaded while the current plane is
        // computed: raw input (and U0 for the on-the-fly stage) in registers
        double pin[HPT][NV], pu0[HPT][NV];
#pragma unroll
        for (int j = 0; j < HPT; ++j)
#pragma unroll
            for (int v = 0; v < NV; ++v) pin[j][v] = pu0[j][v] = 0.0;
#define AVH_PREFETCH_RING(KP)                                                   \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) {                           \
        if (has_halo[j]) {                                                      \
            const Idx gi = (Idx)(KP) * plane + halo_xy[j];                      \
            _Pragma("unroll") for (int v = 0; v < NV; ++v) {                    \
                pin[j][v] = ld(P.in, P.ncells, v, gi);                            \
                if (otf) pu0[j][v] = ld(P.u0, P.ncells, v, gi);                   \
            }                                                                   \
        }                                                                       \
    }
        AVH_PREFETCH_RING(z0)

        // RK combination ((L*s + U) * ca + U0) * cb (the factor 1 of the RK2
        // stage is exact); U0 of the column's next plane is loaded one plane
        // ahead, with the window's next plane
        const bool combine = !(mode == MODE_FLUX_ONLY || mode == MODE_L);
        const double ca = mode == MODE_RK3_2 ? 1. / 3. : (mode == MODE_RK3_3 ? 2.0 : 1.0);
        const double cb = mode == MODE_RK2_2 ? 0.5 : (mode == MODE_RK3_2 ? 3. / 4. : 1. / 3.);
        double u0n[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) u0n[v] = 0.0;
        if (combine && col_ok && !ring) {
            const Idx gi = gidx(ox + cx, oy + cy, z0, n0, n1);
#pragma unroll
            for (int v = 0; v < NV; ++v) u0n[v] = ld(P.u0, P.ncells, v, gi);
        }

        for (int k = z0; k < z1; ++k) {
            double* const sp = sp_buf[TWO_BARRIERS ? ((k - z0) & 1) : 0];
            // plane k into LDS: own column from the window (w[HW-1] is plane k),
            // the halo ring from the prefetched registers
            if (has_col) {
#pragma unroll
                for (int v = 0; v < NV; ++v) sp[v * LP + (cy + G) * LXP + (cx + G)] = w[HW - 1][v];
            }
#pragma unroll
            for (int j = 0; j < HPT; ++j) {
                if (has_halo[j]) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        double val = pin[j][v];
                        if (otf) val = val * s + pu0[j][v];
                        sp[v * LP + halo_lds[j]] = val;
                    }
                }
            }
            __syncthreads();
            if (k + 1 < z1) {
                AVH_PREFETCH_RING(k + 1)
            }

            double acc[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = 0.0;
            // z: F(k+1/2) from R(k) (kept) and L(k+1) (window centred at k+1)
            double fz[NV], rnext[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) fz[v] = rnext[v] = 0.0;
            // (only for the tile's columns inside the grid: a column past the
            // grid edge supplies the plane's halo but has no cells to update)
            auto z_flux = [&]() {
                if (col_ok) {
                    typedef typename OpsFor<FLUX::FAST>::type Ops;
                    double lz[NV];
                    bool ok = true;
                    recon_faces_reg<RECON, NV, WZ, Ops>(w, HW, lz, rnext, ok);
                    typename EQ::All aL = EQ::template make_all_ops<Ops>(rprev, P.gamma, ok);
                    typename EQ::All aR = EQ::template make_all_ops<Ops>(lz, P.gamma, ok);
                    double wz = FLUX::template flux_ops<Ops, 2, EQ>(aL, aR, P.gamma, fz, ok);
                    if (FLUX::FAST && !ok) {
                        // operands outside the guarded range: the plain path
                        recon_face_reg<RECON, NV, false, WZ>(w, HW, lz);
                        recon_face_reg<RECON, NV, true, WZ>(w, HW, rnext);
                        aL = EQ::make_all(rprev, P.gamma);
                        aR = EQ::make_all(lz, P.gamma);
                        bool ok2 = true;
                        wz = FLUX::template flux_ops<PlainOps, 2, EQ>(aL, aR, P.gamma, fz, ok2);
                    }
                    ws[2] = ws_max(ws[2], wz);
                }
            };
            auto add_x = [&]() {
                if (has_col) {
                    const int ip = cy * (TX + 1) + cx + 1;
#pragma unroll
                    for (int v = 0; v < NV; ++v)
                        acc[v] = acc[v] + ((-sfx[v * NIMAX + ip]) + sfx[v * NIMAX + ip - 1]);
                }
            };
            auto add_y = [&]() {
                if (has_col) {
                    const int ip = (cy + 1) * TX + cx;
#pragma unroll
                    for (int v = 0; v < NV; ++v)
                        acc[v] = acc[v] + ((-sfy[v * NIMAX + ip]) + sfy[v * NIMAX + ip - TX]);
                }
            };
            if constexpr (TWO_BARRIERS) {
                // every flux of the plane, then one barrier, then the sums in
                // the reference's x, y, z order
                plane_pass<0>(sp, sfx, ox, oy, P.n, P.gamma, ws[0]);
                plane_pass<1>(sp, sfy, ox, oy, P.n, P.gamma, ws[1]);
                z_flux();
                __syncthreads();
                add_x();
                add_y();
            } else {
                plane_pass<0>(sp, sfx, ox, oy, P.n, P.gamma, ws[0]);
                __syncthreads();
                add_x();
                __syncthreads();
                plane_pass<1>(sp, sfy, ox, oy, P.n, P.gamma, ws[1]);
                __syncthreads();
                add_y();
                z_flux();
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = acc[v] + ((-fz[v]) + fprev[v]);

            if (col_ok) {
                const Idx gi = gidx(ox + cx, oy + cy, k, n0, n1);
                if (ring) {
                    const double* const u0k = slot(k);
#pragma unroll
                    for (int v = 0; v < NV; ++v) u0n[v] = u0k[v * NC];
                }
                double r[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const double l = acc[v];
                    if (!combine) {
                        r[v] = l;
                    } else {
                        double o = l * s;
                        o = o + w[HW - 1][v];
                        o = o * ca;
                        o = o + u0n[v];
                        o = o * cb;
                        r[v] = o;
                    }
                }
                if (final_stage) bad = bad || !EQ::obeys_fast(r, P.gamma);
#pragma unroll
                for (int v = 0; v < NV; ++v) st(P.out, P.ncells, v, gi, r[v]);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                fprev[v] = fz[v];
                rprev[v] = rnext[v];
            }
            // advance the window to centre k+2
#pragma unroll
            for (int i = 0; i < WZ - 1; ++i)
#pragma unroll
                for (int v = 0; v < NV; ++v) w[i][v] = w[i + 1][v];
            // the window's next plane and the next write-back's U0, issued
            // together so that their latencies overlap
            if (has_col && k + 1 < z1) {
                const Idx gn = gidx(gxc, gyc, bc_map(k + 2 + HW, n2, P.bc), n0, n1);
                const Idx gw = gidx(ox + cx, oy + cy, k + 1, n0, n1);
                double cin[NV], cu0[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    cin[v] = ld(P.in, P.ncells, v, gn);
                    cu0[v] = 0.0;
                }
                if (otf) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) cu0[v] = ld(P.u0, P.ncells, v, gn);
                }
                if (combine && col_ok && !ring) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) u0n[v] = ld(P.u0, P.ncells, v, gw);
                }
                if (ring && k + 2 + HW < z1) {
                    double* const kp = slot(k + 2 + HW);
#pragma unroll
                    for (int v = 0; v < NV; ++v) kp[v * NC] = cu0[v];
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) w[WZ - 1][v] = otf ? cin[v] * s + cu0[v] : cin[v];
            }
            // one plane buffer: every thread must be done reading it (and the
            // flux buffer) before the next plane is written
            if constexpr (!TWO_BARRIERS) __syncthreads();
        }

#undef AVH_PREFETCH_RING
        if (mode == MODE_FLUX_ONLY || mode == MODE_L) {
            const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double wv = wave_max(ws[d]);
                if (lane == 0) s_ws[d][wid] = wv;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                double wv = 0.0;
#pragma unroll
                for (int i = 0; i < NT / 64; ++i) wv = ws_max(wv, s_ws[threadIdx.x][i]);
                if (wv > 0.0)
                    atomicMax(&ctrl->ws_bits[threadIdx.x], (unsigned long long)__double_as_longlong(wv));
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
__global__ __launch_bounds__(256) void march_kernel3(StageParams P) {
    K::run(P);
}

}  // namespace avh
