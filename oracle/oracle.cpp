/*
 * CPU ORACLE -- test infrastructure, not product code.
 *
 * A C++17/OpenMP restatement of the alsvinn/alsvinn v0.5.4 CPU path for one
 * finite-volume time step, written from the equations and following the
 * reference's passes and operation order so that its results are the
 * reference CPU results:
 *
 *   NumericalFluxCPU::computeFlux  = makeZero, then per direction
 *       reconstruction into left/right volumes, flux into a temp volume with
 *       a wave-speed max, combine        (alsfvm/src/numflux/NumericalFluxCPU.cpp:32-221)
 *   ConservedSystem::operator()   = interior call + 2*dim side strips
 *                                    (alsfvm/src/simulator/ConservedSystem.cpp:31-103)
 *   ForwardEuler / RungeKutta2 / RungeKutta3 substeps as Volume *=, +=
 *                                    (alsfvm/src/integrator/{ForwardEuler,RungeKutta2,RungeKutta3}.cpp)
 *   Integrator::computeTimestep     (alsfvm/src/integrator/Integrator.cpp:20-66)
 *   FixedIntervalWriter adjuster    (alsfvm/src/io/FixedIntervalWriter.cpp:33-53)
 *   BoundaryCPU periodic / neumann  (alsfvm/src/boundary/BoundaryCPU.cpp:30-121)
 *   CPUCellComputer::obeysConstraints (alsfvm/src/equation/CPUCellComputer.cpp:79-125)
 *   Simulator step driver           (alsfvm/src/simulator/Simulator.cpp:77-214)
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker and as the timed CPU baseline; the product
 * path never calls it.  Build with -O3 -fopenmp -ffp-contract=off (see
 * oracle/Makefile): no FMA contraction, no fast-math, IEEE divide and sqrt.
 *
 * Pins (tests/test_oracle_*.py): reference known answers (WENO2 closed-form
 * weights, HLL consistency, point fluxes, net flux of constant states,
 * ghost index maps, grid midpoints), convergence rates, RK polynomials,
 * conservation under periodic BCs, Burgers characteristics before the
 * shock, exact Riemann solution for Sod, and interior+strips == whole domain.
 *
 * Observables (tests/test_oracle_observables.py): oracle_structure_cube
 * follows computeStructureCubeCPU's loops and per-term division and is
 * pinned by the reference's StructureTestBoundary cases (constant, linear
 * 1D on N = 16, linear 2D on 16 x 16, p = 1..8, periodic and neumann) at the
 * reference's own 4-ulp bar, plus a vectorised brute force;
 * oracle_kinetic_energy (a new functional) by closed forms (uniform flow,
 * sin^2 under the periodic midpoint rule).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <random>
#include <functional>
#include <type_traits>

#include "../include/alsvinn_hip.h"

#if defined(__FP_FAST_FMA) && !defined(ORACLE_ALLOW_FMA)
/* not an error by itself; -ffp-contract=off is what matters */
#endif

namespace oracle {

using real = double;

/* ------------------------------------------------------------------ */
/* Volume: SoA, one array per variable, x fastest, ghosts on active axes
 * (alsfvm/src/volume/Volume.cpp:44-60, alsfvm/include/alsfvm/memory/View.hpp:116-118) */
struct Volume {
    int nv = 0;
    int nx = 1, ny = 1, nz = 1; /* total sizes incl. ghosts */
    int gx = 0, gy = 0, gz = 0; /* ghost counts */
    std::vector<real> d;

    Volume() = default;
    Volume(int nv_, int ix, int iy, int iz, int gx_, int gy_, int gz_)
        : nv(nv_), nx(ix + 2 * gx_), ny(iy + 2 * gy_), nz(iz + 2 * gz_), gx(gx_), gy(gy_), gz(gz_),
          d((size_t)nv_ * (ix + 2 * gx_) * (iy + 2 * gy_) * (iz + 2 * gz_), 0.0) {}
    size_t cells() const { return (size_t)nx * ny * nz; }
    real* var(int v) { return d.data() + (size_t)v * cells(); }
    const real* var(int v) const { return d.data() + (size_t)v * cells(); }
    size_t index(int x, int y, int z) const { return (size_t)z * nx * ny + (size_t)y * nx + x; }
};

/* Volume::operator*=, += element-wise over every cell (incl. ghosts),
 * alsfvm/src/volume/Volume.cpp:178-196 -> alsfvm/src/memory/HostMemory.cpp:97-201 */
static void vol_mul(Volume& a, real s) {
    const size_t n = a.d.size();
    real* p = a.d.data();
#pragma omp parallel for simd
    for (size_t i = 0; i < n; ++i) p[i] *= s;
}
static void vol_add(Volume& a, const Volume& b) {
    const size_t n = a.d.size();
    real* p = a.d.data();
    const real* q = b.d.data();
#pragma omp parallel for simd
    for (size_t i = 0; i < n; ++i) p[i] += q[i];
}

/* ------------------------------------------------------------------ */
/* Equations.  A state is NV doubles: burgers {u}; euler {rho, m[nsd], E}. */

struct Burgers {
    static constexpr int NV = 1;
    static constexpr int NSD = 3; /* the flux ignores the direction: any axis */
    real gamma = 0; /* unused */
    struct All {
        real u;
    };
    /* makeAllVariables: alsfvm/include/alsfvm/equation/burgers/Burgers.hpp:220-224 */
    All make_all(const real* q) const { return All{q[0]}; }
    /* point flux u*u/2: Burgers.hpp:158-164 */
    template <int DIR>
    void point_flux(const All& a, real* F) const { F[0] = a.u * a.u / 2; }
    /* getWeight = u: Burgers.hpp:231-233 */
    static real weight(const real* q) { return q[0]; }
    /* CPUCellComputer::obeysConstraints: conserved finite, then u<INF && u==u */
    bool obeys(const real* q) const {
        if (std::isnan(q[0]) || std::isinf(q[0])) return false;
        return q[0] < INFINITY && (q[0] == q[0]);
    }
};

template <int NSD_>
struct Euler {
    static constexpr int NSD = NSD_;
    static constexpr int NV = NSD_ + 2;
    real gamma = 1.4;
    struct All {
        real rho, m[NSD_], E, p, u[NSD_];
    };
    /* vec dot = x*ox + y*oy (+ z*oz), left to right: alsutils/include/alsutils/vec{1,2,3}.hpp */
    static real dot(const real* m) {
        real s = m[0] * m[0];
        for (int c = 1; c < NSD_; ++c) s = s + m[c] * m[c];
        return s;
    }
    /* computeExtra: ie = E - 0.5*m.m/rho; u = m/rho; p = (gamma-1)*ie
     * alsfvm/include/alsfvm/equation/euler/Euler.hpp:204-211 */
    All make_all(const real* q) const {
        All a;
        a.rho = q[0];
        for (int c = 0; c < NSD_; ++c) a.m[c] = q[1 + c];
        a.E = q[NSD_ + 1];
        const real ie = a.E - 0.5 * dot(a.m) / a.rho;
        for (int c = 0; c < NSD_; ++c) a.u[c] = a.m[c] / a.rho;
        a.p = (gamma - 1) * ie;
        return a;
    }
    /* computePointFlux: Euler.hpp:172-181 */
    template <int DIR>
    void point_flux(const All& a, real* F) const {
        F[0] = a.m[DIR];
        for (int c = 0; c < NSD_; ++c) F[1 + c] = a.m[c] * a.u[DIR];
        F[1 + DIR] += a.p;
        F[NSD_ + 1] = (a.E + a.p) * a.u[DIR];
    }
    static real weight(const real* q) { return q[0]; }
    /* CPUCellComputer::obeysConstraints (alsfvm/src/equation/CPUCellComputer.cpp:79-125)
     * + Euler::obeysConstraints (Euler.hpp:269-277) */
    bool obeys(const real* q) const {
        for (int i = 0; i < NV; ++i)
            if (std::isnan(q[i]) || std::isinf(q[i])) return false;
        All a = make_all(q);
        if (std::isnan(a.p) || std::isinf(a.p)) return false;
        for (int c = 0; c < NSD_; ++c)
            if (std::isnan(a.u[c]) || std::isinf(a.u[c])) return false;
        return std::isfinite(a.rho) && (a.rho == a.rho) && (a.rho > 0) && (a.p > 0);
    }
    /* computeConserved: m = rho*u, E = p/(gamma-1) + 0.5*rho*u.u (Euler.hpp:236-245) */
    void from_primitive(const real* prim, real* q) const {
        const real rho = prim[0];
        real u[NSD_];
        for (int c = 0; c < NSD_; ++c) u[c] = prim[1 + c];
        const real p = prim[NSD_ + 1];
        q[0] = rho;
        for (int c = 0; c < NSD_; ++c) q[1 + c] = u[c] * rho;
        q[NSD_ + 1] = p / (gamma - 1) + 0.5 * rho * dot(u);
    }
};

/* ------------------------------------------------------------------ */
/* Numerical fluxes.  Each returns the interface wave speed. */

/* burgers::Godunov: alsfvm/include/alsfvm/numflux/burgers/Godunov.hpp:40-60 */
struct Godunov {
    template <int DIR>
    static real flux(const Burgers& eq, const Burgers::All& L, const Burgers::All& R, real* F) {
        real fl[1], fr[1];
        eq.point_flux<DIR>(Burgers::All{std::fmax(L.u, real(0.0))}, fl);
        eq.point_flux<DIR>(Burgers::All{std::fmin(R.u, real(0.0))}, fr);
        F[0] = std::fmax(fl[0], fr[0]);
        return std::fmax(std::fabs(L.u), std::fabs(R.u));
    }
};

/* euler::HLL: alsfvm/include/alsfvm/numflux/euler/HLL.hpp */
struct HLL {
    template <int DIR, int NSD>
    static real flux(const Euler<NSD>& eq, const typename Euler<NSD>::All& L,
                     const typename Euler<NSD>::All& R, real* F) {
        constexpr int NV = NSD + 2;
        const real g = eq.gamma;
        const real wl = std::sqrt(L.rho);
        const real wr = std::sqrt(R.rho);
        const real rho = (L.rho + R.rho) / 2;
        (void)rho;
        real u[NSD];
        for (int c = 0; c < NSD; ++c) u[c] = (L.u[c] * wl + R.u[c] * wr) / (wl + wr);
        const real p = (L.p * wl + R.p * wr) / (wl + wr);
        const real cs = std::sqrt(g * p / rho);
        const real sl = std::fmin(L.u[DIR] - std::sqrt(g * L.p / L.rho), u[DIR] - cs);
        const real sr = std::fmax(R.u[DIR] + std::sqrt(g * R.p / R.rho), u[DIR] + cs);
        if (sl == 0) {
            for (int i = 0; i < NV; ++i) F[i] = 0;
        } else if (sl > 0) {
            eq.template point_flux<DIR>(L, F);
        } else if (sr < 0) {
            eq.template point_flux<DIR>(R, F);
        } else {
            real fl[NV], fr[NV];
            eq.template point_flux<DIR>(L, fl);
            eq.template point_flux<DIR>(R, fr);
            real ul[NV], ur[NV];
            ul[0] = L.rho; ur[0] = R.rho;
            for (int c = 0; c < NSD; ++c) { ul[1 + c] = L.m[c]; ur[1 + c] = R.m[c]; }
            ul[NSD + 1] = L.E; ur[NSD + 1] = R.E;
            const real srsl = sr * sl;
            const real den = sr - sl;
            for (int i = 0; i < NV; ++i)
                F[i] = (sr * fl[i] - sl * fr[i] + srsl * (ur[i] - ul[i])) / den;
        }
        return std::fmax(std::fabs(sl), std::fabs(sr));
    }
};

/* euler::HLL3: alsfvm/include/alsfvm/numflux/euler/HLL3.hpp:52-148 */
struct HLL3 {
    template <int DIR, int NSD>
    static real flux(const Euler<NSD>& eq, const typename Euler<NSD>::All& L,
                     const typename Euler<NSD>::All& R, real* F) {
        constexpr int NV = NSD + 2;
        const real g = eq.gamma;
        /* computeHLLSpeeds */
        const real wl = std::sqrt(L.rho);
        const real wr = std::sqrt(R.rho);
        const real rho = (L.rho + R.rho) / 2;
        real u[NSD];
        for (int c = 0; c < NSD; ++c) u[c] = (L.u[c] * wl + R.u[c] * wr) / (wl + wr);
        const real p = (wl * L.p + wr * R.p) / (wl + wr);
        const real cfl = std::sqrt(g * L.p / L.rho);
        const real cfr = std::sqrt(g * R.p / R.rho);
        const real correct = 0.5 * std::fmax(real(0), L.u[DIR] - R.u[DIR]);
        const real cs = std::sqrt(g * p / rho);
        const real sl = std::fmin(L.u[DIR] + correct - cfl, u[DIR] - cs);
        const real sr = std::fmax(R.u[DIR] - correct + cfr, u[DIR] + cs);

        real fl[NV], fr[NV];
        eq.template point_flux<DIR>(L, fl);
        eq.template point_flux<DIR>(R, fr);
        const real pl = L.p, pr = R.p;
        const real udl = L.u[DIR] - sl;
        const real udr = R.u[DIR] - sr;
        const real aa = udr * R.rho - udl * L.rho;
        const real sm = (R.m[DIR] * udr - L.m[DIR] * udl + pr - pl) / aa;
        real us[NSD];
        for (int c = 0; c < NSD; ++c)
            us[c] = (fr[1 + c] - fl[1 + c] - R.m[c] * sr + L.m[c] * sl) / aa;
        us[DIR] = sm;

        if (sl == 0) {
            for (int i = 0; i < NV; ++i) F[i] = 0;
        } else if (sl > 0) {
            for (int i = 0; i < NV; ++i) F[i] = fl[i];
        } else if (sr < 0) {
            for (int i = 0; i < NV; ++i) F[i] = fr[i];
        } else if (sm >= 0) {
            real mid[NV];
            mid[0] = L.rho * udl / (sm - sl);
            for (int c = 0; c < NSD; ++c) mid[1 + c] = us[c] * mid[0];
            const real pm = pl + L.rho * (L.u[DIR] - sm) * udl;
            mid[NSD + 1] = (udl * L.E + pl * L.u[DIR] - pm * sm) / (sm - sl);
            real uc[NV];
            uc[0] = L.rho;
            for (int c = 0; c < NSD; ++c) uc[1 + c] = L.m[c];
            uc[NSD + 1] = L.E;
            for (int i = 0; i < NV; ++i) F[i] = fl[i] + sl * (mid[i] - uc[i]);
        } else {
            real mid[NV];
            mid[0] = R.rho * udr / (sm - sr);
            for (int c = 0; c < NSD; ++c) mid[1 + c] = us[c] * mid[0];
            const real pm = pr + R.rho * (R.u[DIR] - sm) * udr;
            mid[NSD + 1] = (udr * R.E + pr * R.u[DIR] - pm * sm) / (sm - sr);
            real uc[NV];
            uc[0] = R.rho;
            for (int c = 0; c < NSD; ++c) uc[1 + c] = R.m[c];
            uc[NSD + 1] = R.E;
            for (int i = 0; i < NV; ++i) F[i] = fr[i] + sr * (mid[i] - uc[i]);
        }
        return std::fmax(std::fabs(sl), std::fabs(sr));
    }
};

/* ------------------------------------------------------------------ */
/* Reconstructions.  Interface of the reference Reconstruction class:
 * performReconstruction(in, direction, indicatorVariable, left, right, start, end)
 * (alsfvm/include/alsfvm/reconstruction/Reconstruction.hpp:46-58). */

struct ivec3 {
    int x = 0, y = 0, z = 0;
    int& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    int operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct Reconstruction {
    virtual ~Reconstruction() = default;
    virtual int ghosts() const = 0;
    virtual void reconstruct(const Volume& in, int dir, int ind, Volume& L, Volume& R,
                             ivec3 start, ivec3 end) = 0;
};

/* NoReconstruction: left = right = cell value over every cell
 * (alsfvm/src/reconstruction/NoReconstruction.cpp) */
struct NoRecon : Reconstruction {
    int ghosts() const override { return 1; }
    void reconstruct(const Volume& in, int, int, Volume& L, Volume& R, ivec3, ivec3) override {
        for (int v = 0; v < in.nv; ++v) {
            const real* pi = in.var(v);
            real* pl = L.var(v);
            real* pr = R.var(v);
            for (size_t i = 0; i < in.cells(); ++i) { pl[i] = pi[i]; pr[i] = pi[i]; }
        }
    }
};

/* WENO2 via ReconstructionCPU: alsfvm/include/alsfvm/reconstruction/WENO2.hpp:31-95,
 * alsfvm/src/reconstruction/ReconstructionCPU.cpp:37-97, epsilon 1e-8
 * (alsfvm/include/alsfvm/reconstruction/WENOCoefficients.hpp:21) */
struct Weno2 : Reconstruction {
    int dim;
    explicit Weno2(int d) : dim(d) {}
    int ghosts() const override { return 2; }
    void reconstruct(const Volume& in, int dir, int /*ind*/, Volume& L, Volume& R, ivec3 start,
                     ivec3 end) override {
        const real eps = 1e-8;
        const int dx = dir == 0, dy = dir == 1, dz = dir == 2;
        const int nx = in.nx, ny = in.ny, nz = in.nz;
        const int ngx = 2, ngy = (dim > 1) * 2, ngz = (dim > 2) * 2;
        const int sx = (ngx - dx) + start.x, sy = (ngy - dy) + start.y, sz = (ngz - dz) + start.z;
        const int ex = nx - (ngx - dx) + end.x, ey = ny - (ngy - dy) + end.y,
                  ez = nz - (ngz - dz) + end.z;
        const int nv = in.nv;
        const real* w = in.var(0); /* getWeight: rho (euler) or u (burgers) */
        for (int z = sz; z < ez; z++) {
#pragma omp parallel for
            for (int y = sy; y < ey; y++) {
                for (int x = sx; x < ex; x++) {
                    const size_t io = in.index(x, y, z);
                    const size_t ir = in.index(x + dx, y + dy, z + dz);
                    const size_t il = in.index(x - dx, y - dy, z - dz);
                    const real i0 = w[io];
                    const real d0 = w[ir] - i0;
                    const real b0 = d0 * d0;
                    const real d1 = i0 - w[il];
                    const real b1 = d1 * d1;
                    const real a0L = 1 / (3 * (eps + b0) * (eps + b0));
                    const real a1L = 2 / (3 * (eps + b1) * (eps + b1));
                    const real w0L = a0L / (a0L + a1L);
                    const real w1L = a1L / (a0L + a1L);
                    const real a0R = 2 / (3 * (eps + b0) * (eps + b0));
                    const real a1R = 1 / (3 * (eps + b1) * (eps + b1));
                    const real w0R = a0R / (a0R + a1R);
                    const real w1R = a1R / (a0R + a1R);
                    for (int v = 0; v < nv; ++v) {
                        const real* u = in.var(v);
                        L.var(v)[io] = 0.5 * (w1L * u[il] + (3 * w0L + w1L) * u[io] - w0L * u[ir]);
                        R.var(v)[io] = 0.5 * (w0R * u[ir] + (3 * w1R + w0R) * u[io] - w1R * u[il]);
                    }
                }
            }
        }
    }
};

/* ENOCPU<order>: alsfvm/src/reconstruction/ENOCPU.cpp:41-267, coefficient
 * tables alsfvm/src/reconstruction/ENOCoefficients.cpp:27-40 */
template <int ORDER>
struct Eno : Reconstruction {
    std::vector<std::vector<real>> dd; /* ORDER-1 divided-difference arrays */
    int tnx = 0, tny = 0, tnz = 0;
    static real coef(int r, int j) {
        if (ORDER == 2) {
            static const real c[3][2] = {{real(3.0 / 2.0), real(-1.0 / 2.0)},
                                         {real(1.0 / 2.0), real(1.0 / 2.0)},
                                         {real(-1.0 / 2.0), real(3.0 / 2.0)}};
            return c[r][j];
        } else {
            static const real c[4][3] = {{real(11.0 / 6.0), real(-7.0 / 6.0), real(1.0 / 3.0)},
                                         {real(1.0 / 3.0), real(5.0 / 6.0), real(-1.0 / 6.0)},
                                         {real(-1.0 / 6.0), real(5.0 / 6.0), real(1.0 / 3.0)},
                                         {real(1.0 / 3.0), real(-7.0 / 6.0), real(11.0 / 6.0)}};
            return c[r][j];
        }
    }
    int ghosts() const override { return ORDER; }
    void make_arrays(int nx, int ny, int nz) {
        tnx = nx; tny = ny; tnz = nz;
        dd.assign(ORDER - 1, std::vector<real>((size_t)nx * ny * nz, 0.0));
    }
    /* computeDividedDifferences (ENOCPU.cpp:205-262) */
    void divided(const real* pin, int nx, int ny, int nz, ivec3 dv, int level, real* pout,
                 ivec3 start, ivec3 end) {
        const int sx = (dv.x == 0) * ORDER + dv.x * level + start.x;
        const int sy = (ny > 1) * (dv.y == 0) * ORDER + dv.y * level + start.y;
        const int sz = (nz > 1) * (dv.z == 0) * ORDER + dv.z * level + start.z;
        const int ex = nx - ((dv.x == 0) * (ORDER - 1) + dv.x * (level - 1)) + end.x;
        const int ey = ny - ((ny > 1) * (dv.y == 0) * (ORDER - 1) + dv.y * (level - 1)) + end.y;
        const int ez = nz - ((nz > 1) * (dv.z == 0) * (ORDER - 1) + dv.z * (level - 1)) + end.z;
        for (int z = sz; z < ez; z++)
            for (int y = sy; y < ey; y++)
                for (int x = sx; x < ex; x++) {
                    const size_t ir = (size_t)z * nx * ny + (size_t)y * nx + x;
                    const size_t il = (size_t)(z - dv.z) * nx * ny + (size_t)(y - dv.y) * nx + (x - dv.x);
                    pout[il] = pin[ir] - pin[il];
                }
    }
    void reconstruct(const Volume& in, int dir, int ind, Volume& L, Volume& R, ivec3 start,
                     ivec3 end) override {
        const int nx = in.nx, ny = in.ny, nz = in.nz;
        if ((int)dd.empty() || tnx != nx || tny != ny || tnz != nz) make_arrays(nx, ny, nz);
        ivec3 dv;
        dv.x = dir == 0; dv.y = dir == 1; dv.z = dir == 2;
        divided(in.var(ind), nx, ny, nz, dv, 1, dd[0].data(), start, end);
        for (int i = 1; i < ORDER - 1; i++)
            divided(dd[i - 1].data(), nx, ny, nz, dv, i + 1, dd[i].data(), start, end);
        const int sx = (ORDER - dv.x) + start.x;
        const int sy = (ny > 1) * (ORDER - dv.y) + start.y;
        const int sz = (nz > 1) * (ORDER - dv.z) + start.z;
        const int ex = nx - (ORDER - dv.x) + end.x;
        const int ey = ny - (ny > 1) * (ORDER - dv.y) + end.y;
        const int ez = nz - (nz > 1) * (ORDER - dv.z) + end.z;
        for (int v = 0; v < in.nv; v++) {
            const real* pin = in.var(v);
            real* pl = L.var(v);
            real* pr = R.var(v);
            for (int z = sz; z < ez; z++) {
#pragma omp parallel for
                for (int y = sy; y < ey; y++) {
                    for (int x = sx; x < ex; x++) {
                        const size_t ir = (size_t)z * nx * ny + (size_t)y * nx + x;
                        const size_t il = (size_t)(z - dv.z) * nx * ny + (size_t)(y - dv.y) * nx + (x - dv.x);
                        int shift = 0;
                        for (int level = 0; level < ORDER - 1; level++) {
                            const real ddr = dd[level][ir];
                            const real ddl = dd[level][il];
                            if (std::fabs(ddl) < std::fabs(ddr)) shift++;
                        }
                        real lv = 0.0, rv = 0.0;
                        for (int j = 0; j < ORDER; j++) {
                            const size_t idx = (size_t)(z - (shift - j) * dv.z) * nx * ny +
                                               (size_t)(y - (shift - j) * dv.y) * nx +
                                               (x - (shift - j) * dv.x);
                            const real value = pin[idx];
                            lv += coef(shift, j) * value;
                            rv += coef(shift + 1, j) * value;
                        }
                        pl[ir] = lv;
                        pr[ir] = rv;
                    }
                }
            }
        }
    }
};

/* ------------------------------------------------------------------ */
/* NumericalFluxCPU<Flux, Equation, dimension> */

struct NumericalFlux {
    virtual ~NumericalFlux() = default;
    virtual int ghosts() const = 0;
    virtual void compute(const Volume& U, real ws[3], Volume& out, ivec3 start, ivec3 end) = 0;
};

template <class Flux, class Eq, int DIM>
struct NumericalFluxCPU : NumericalFlux {
    Eq eq;
    std::unique_ptr<Reconstruction> recon;
    Volume left, right, temp;

    NumericalFluxCPU(const Eq& e, std::unique_ptr<Reconstruction> r) : eq(e), recon(std::move(r)) {}
    /* max(stencil.size()-1, reconstruction ghosts) (NumericalFluxCPU.cpp:226-231) */
    int ghosts() const override { return std::max(1, recon->ghosts()); }

    void ensure(const Volume& U) {
        if (left.nx != U.nx || left.ny != U.ny || left.nz != U.nz || left.nv != U.nv) {
            left = Volume(U.nv, U.nx - 2 * U.gx, U.ny - 2 * U.gy, U.nz - 2 * U.gz, U.gx, U.gy, U.gz);
            right = left;
            temp = left;
        }
    }

    /* makeZero (NumericalFluxCPU.cpp:121-146) */
    void make_zero(Volume& out, ivec3 start, ivec3 end) {
        const int nx = out.nx, ny = out.ny, nz = out.nz;
        const int ngx = out.gx, ngy = out.gy, ngz = out.gz;
        for (int z = ngz + start.z; z < nz - ngz + end.z; ++z) {
#pragma omp parallel for
            for (int y = ngy + start.y; y < ny - ngy + end.y; ++y)
                for (int x = ngx + start.x; x < nx - ngx + end.x; ++x)
                    for (int v = 0; v < out.nv; ++v) out.var(v)[out.index(x, y, z)] = 0;
        }
    }

    /* computeNetFlux<Flux, Equation, direction> (NumericalFluxCPU.cpp:32-118) */
    template <int DIR>
    void net_flux(Volume& out, real& waveSpeed, ivec3 start, ivec3 end) {
        constexpr int NV = Eq::NV;
        const int xd = DIR == 0, yd = DIR == 1, zd = DIR == 2;
        const int nx = out.nx, ny = out.ny, nz = out.nz;
        const int ngx = out.gx, ngy = out.gy, ngz = out.gz;
        real wsc = 0;
        for (int z = ngz - zd + start.z; z < nz - ngz + end.z; ++z) {
#pragma omp parallel for reduction(max : wsc)
            for (int y = ngy - yd + start.y; y < ny - ngy + end.y; ++y) {
                for (int x = ngx - xd + start.x; x < nx - ngx + end.x; ++x) {
                    const size_t i0 = out.index(x, y, z);
                    const size_t i1 = out.index(x + xd, y + yd, z + zd);
                    real qr[NV], ql[NV];
                    for (int v = 0; v < NV; ++v) {
                        qr[v] = left.var(v)[i1];  /* value at j+1/2 from the right cell */
                        ql[v] = right.var(v)[i0]; /* value at j+1/2 from the left cell  */
                    }
                    const auto aR = eq.make_all(qr);
                    const auto aL = eq.make_all(ql);
                    real F[NV];
                    const real wl = Flux::template flux<DIR>(eq, aL, aR, F);
                    for (int v = 0; v < NV; ++v) temp.var(v)[i0] = (-1.0) * F[v];
                    wsc = std::max(wsc, wl);
                }
            }
        }
        waveSpeed = wsc;
        for (int z = ngz - zd + start.z; z < nz - ngz + end.z - zd; ++z) {
#pragma omp parallel for
            for (int y = ngy - yd + start.y; y < ny - ngy + end.y - yd; ++y) {
                for (int x = ngx - xd + start.x; x < nx - ngx + end.x - xd; ++x) {
                    const size_t ir = out.index(x + xd, y + yd, z + zd);
                    const size_t im = out.index(x, y, z);
                    for (int v = 0; v < NV; ++v) {
                        const real fmr = temp.var(v)[ir];
                        const real flm = (-1.0) * temp.var(v)[im];
                        out.var(v)[ir] += fmr + flm;
                    }
                }
            }
        }
    }

    /* NumericalFluxCPU::computeFlux (NumericalFluxCPU.cpp:170-221) */
    void compute(const Volume& U, real ws[3], Volume& out, ivec3 start, ivec3 end) override {
        ensure(U);
        make_zero(out, start, end);
        recon->reconstruct(U, 0, 0, left, right, start, end);
        net_flux<0>(out, ws[0], start, end);
        if (DIM > 1) {
            recon->reconstruct(U, 1, 0, left, right, start, end);
            net_flux<DIM >= 2 ? 1 : 0>(out, ws[1], start, end);
        }
        if (DIM > 2) {
            recon->reconstruct(U, 2, 0, left, right, start, end);
            net_flux<DIM >= 3 ? 2 : 0>(out, ws[2], start, end);
        }
    }
};

/* ------------------------------------------------------------------ */
/* ConservedSystem::operator() (alsfvm/src/simulator/ConservedSystem.cpp:31-103) */
struct System {
    std::unique_ptr<NumericalFlux> flux;
    int dim = 1;
    bool whole_domain = false; /* test switch: one whole-domain call instead of interior+strips */

    void operator()(const Volume& U, real ws[3], Volume& out) {
        if (whole_domain) {
            flux->compute(U, ws, out, ivec3{}, ivec3{});
            return;
        }
        ivec3 g;
        g.x = out.gx; g.y = out.gy; g.z = out.gz;
        ivec3 size;
        size.x = out.nx; size.y = out.ny; size.z = out.nz;
        real sides[3] = {0, 0, 0};
        ivec3 ng;
        ng.x = -g.x; ng.y = -g.y; ng.z = -g.z;
        flux->compute(U, ws, out, g, ng);
        for (int d = 0; d < dim; ++d) {
            ivec3 sb, eb;
            eb[d] = -size[d] + 3 * g[d];
            real wb[3] = {0, 0, 0};
            flux->compute(U, wb, out, sb, eb);
            for (int k = 0; k < 3; ++k) sides[k] = std::max(sides[k], wb[k]);
            ivec3 st, et;
            st[d] = size[d] - 3 * g[d];
            real wt[3] = {0, 0, 0};
            flux->compute(U, wt, out, st, et);
            for (int k = 0; k < 3; ++k) sides[k] = std::max(sides[k], wt[k]);
        }
        for (int k = 0; k < 3; ++k) ws[k] = std::max(ws[k], sides[k]);
    }
};

/* ------------------------------------------------------------------ */
/* Boundary (BoundaryCPU + Periodic / Neumann) */
static void apply_boundary(Volume& vol, int dim, int ngc, int bc) {
    const int nx = vol.nx, ny = vol.ny, nz = vol.nz;
    const int inx = nx - 2 * ngc, iny = ny - 2 * ngc, inz = nz - 2 * ngc; /* Periodic's nx,ny,nz */
    for (int v = 0; v < vol.nv; v++) {
        real* p = vol.var(v);
        auto at = [&](int x, int y, int z) -> real& { return p[(size_t)z * nx * ny + (size_t)y * nx + x]; };
        for (int d = 0; d < dim; d++) {
            for (int i = 0; i < 2; i++) {
                const int zDir = d == 2, yDir = d == 1, xDir = d == 0;
                const int zs = zDir ? (i == 0 ? ngc : nz - ngc - 1) : 0;
                const int ze = zDir ? zs + 1 : nz;
                const int ys = yDir ? (i == 0 ? ngc : ny - ngc - 1) : 0;
                const int ye = yDir ? ys + 1 : ny;
                const int xs = xDir ? (i == 0 ? ngc : nx - ngc - 1) : 0;
                const int xe = xDir ? xs + 1 : nx;
                const int sign = (i == 1) ? -1 : 1;
                for (int z = zs; z < ze; z++)
                    for (int y = ys; y < ye; y++)
                        for (int x = xs; x < xe; x++)
                            for (int k = 1; k <= ngc; k++) {
                                if (bc == AVH_BC_PERIODIC) {
                                    at(x - sign * k * xDir, y - sign * k * yDir, z - sign * k * zDir) =
                                        at(x + sign * (-k + inx) * xDir, y + sign * (-k + iny) * yDir,
                                           z + sign * (-k + inz) * zDir);
                                } else {
                                    at(x - sign * k * xDir, y - sign * k * yDir, z - sign * k * zDir) =
                                        at(x + sign * (k - 1) * xDir, y + sign * (k - 1) * yDir,
                                           z + sign * (k - 1) * zDir);
                                }
                            }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Timestep adjusters: FixedIntervalWriter (alsfvm/src/io/FixedIntervalWriter.cpp) */
struct FixedInterval {
    real interval, start;
    size_t numberSaved;
    real adjust(real dt, real t) const {
        if (numberSaved > 0) {
            const real next = numberSaved * interval + start;
            return std::min(dt, next - t);
        }
        return dt;
    }
    bool write(real t) {
        if (t >= numberSaved * interval + start) {
            numberSaved++;
            return true;
        }
        return false;
    }
};

/* ------------------------------------------------------------------ */
/* {ForwardEuler,RungeKutta2,RungeKutta3}::performSubstep
 * (alsfvm/src/integrator/ForwardEuler.cpp:32-51, RungeKutta2.cpp:37-67,
 *  RungeKutta3.cpp:49-86): system, dt at substep 0, then output *= dt/dx.x,
 * += U_substep and the stage combination, as whole-volume vector ops. */
template <class Sys, class Dt>
static real integrator_substep(int integrator, std::vector<Volume>& vols, int s, real dt, real cellx,
                               Sys&& system, Dt&& compute_dt) {
    Volume& out = vols[s + 1];
    real ws[3] = {0, 0, 0};
    system(vols[s], ws, out);
    if (integrator == AVH_INT_FORWARDEULER || s == 0) dt = compute_dt(ws);
    const real scale = dt / cellx;
    vol_mul(out, scale);
    vol_add(out, vols[s]);
    if (integrator == AVH_INT_RUNGEKUTTA2) {
        if (s == 1) {
            vol_add(out, vols[0]);
            vol_mul(out, 0.5);
        }
    } else if (integrator == AVH_INT_RUNGEKUTTA3) {
        if (s == 1) {
            vol_mul(out, 1. / 3.);
            vol_add(out, vols[0]);
            vol_mul(out, 3. / 4.);
        } else if (s == 2) {
            vol_mul(out, 2);
            vol_add(out, vols[0]);
            vol_mul(out, 1. / 3.);
        }
    }
    return dt;
}

/* ------------------------------------------------------------------ */
/* Simulator with Integrator (FE / RK2 / RK3) */

struct Simulator {
    avh_config cfg;
    int dim = 1, ng = 0, nv = 0;
    real cell[3] = {1, 1, 1};
    System system;
    std::vector<Volume> vols;
    std::vector<FixedInterval> adjusters;
    real time = 0;
    long steps = 0;
    real last_dt = 0;
    int substeps = 1;
    std::function<bool(const real*)> obeys_cell;
    long saves = 0;

    /* Integrator::computeTimestep + adjustTimestep (Integrator.cpp:20-66) */
    real compute_timestep(const real ws[3]) const {
        real total = 0;
        for (int d = 0; d < 3; ++d) {
            const real w = ws[d];
            if (cell[d] == 0) continue;
            total += w / cell[d];
        }
        real dt = cfg.cfl / total;
        for (const auto& a : adjusters) dt = a.adjust(dt, time);
        return dt;
    }

    real substep(int s, real dt) {
        return integrator_substep(
            cfg.integrator, vols, s, dt, cell[0], [this](const Volume& U, real* ws, Volume& out) { system(U, ws, out); },
            [this](const real* ws) { return compute_timestep(ws); });
    }

    bool check_constraints() const {
        const Volume& v = vols[0];
        bool ok = true;
        std::vector<real> q(nv);
        for (size_t i = 0; i < v.cells(); ++i) {
            for (int k = 0; k < nv; ++k) q[k] = v.var(k)[i];
            if (!obeys_cell(q.data())) ok = false;
        }
        return ok;
    }

    /* Simulator::performStep = incrementSolution; checkConstraints; callWriters */
    int step() {
        real dt = 0;
        for (int s = 0; s < substeps; ++s) {
            dt = substep(s, dt);
            apply_boundary(vols[s + 1], dim, ng, cfg.boundary);
        }
        std::swap(vols[0], vols[substeps]);
        time += dt;
        steps++;
        last_dt = dt;
        const bool ok = check_constraints();
        for (auto& a : adjusters)
            if (a.write(time)) saves++;
        return ok ? AVH_OK : AVH_ECONSTRAINT;
    }
};

template <class Eq>
static std::function<bool(const real*)> make_obeys(const Eq& eq) {
    return [eq](const real* q) { return eq.obeys(q); };
}

static std::unique_ptr<Reconstruction> make_recon(int which, int dim) {
    switch (which) {
    case AVH_RECON_NONE: return std::unique_ptr<Reconstruction>(new NoRecon());
    case AVH_RECON_WENO2: return std::unique_ptr<Reconstruction>(new Weno2(dim));
    case AVH_RECON_ENO2: return std::unique_ptr<Reconstruction>(new Eno<2>());
    case AVH_RECON_ENO3: return std::unique_ptr<Reconstruction>(new Eno<3>());
    }
    throw std::runtime_error("unknown reconstruction");
}

template <class Flux, class Eq>
static std::unique_ptr<NumericalFlux> make_flux_dim(const Eq& eq, int dim, std::unique_ptr<Reconstruction> r) {
    if (dim == 1) return std::unique_ptr<NumericalFlux>(new NumericalFluxCPU<Flux, Eq, 1>(eq, std::move(r)));
    if constexpr (Eq::NSD >= 2) {
        if (dim == 2) return std::unique_ptr<NumericalFlux>(new NumericalFluxCPU<Flux, Eq, 2>(eq, std::move(r)));
    }
    if constexpr (Eq::NSD >= 3) {
        if (dim == 3) return std::unique_ptr<NumericalFlux>(new NumericalFluxCPU<Flux, Eq, 3>(eq, std::move(r)));
    }
    throw std::runtime_error("grid has more active axes than the equation");
}

static int active_dim(const avh_config& c) {
    if (c.n[2] > 1) return 3;
    if (c.n[1] > 1) return 2;
    return 1;
}

static int num_vars(int equation) {
    switch (equation) {
    case AVH_EQ_BURGERS: return 1;
    case AVH_EQ_EULER1: return 3;
    case AVH_EQ_EULER2: return 4;
    case AVH_EQ_EULER3: return 5;
    }
    return 0;
}

/* Builds the flux/reconstruction pair the reference factories would build
 * (NumericalFluxFactory.cpp:43-200, ReconstructionFactory.cpp:41-355). */
static std::unique_ptr<NumericalFlux> make_numflux(const avh_config& c, std::function<bool(const real*)>* obeys) {
    const int dim = active_dim(c);
    auto r = make_recon(c.reconstruction, dim);
    if (c.equation == AVH_EQ_BURGERS) {
        if (c.flux != AVH_FLUX_GODUNOV) throw std::runtime_error("burgers supports flux godunov");
        Burgers eq;
        if (obeys) *obeys = make_obeys(eq);
        return make_flux_dim<Godunov>(eq, dim, std::move(r));
    }
    if (c.flux != AVH_FLUX_HLL && c.flux != AVH_FLUX_HLL3) throw std::runtime_error("euler supports hll, hll3");
    const bool h3 = c.flux == AVH_FLUX_HLL3;
#define ORACLE_EULER(NSD)                                                                  \
    {                                                                                      \
        Euler<NSD> eq;                                                                     \
        eq.gamma = c.gamma;                                                                \
        if (obeys) *obeys = make_obeys(eq);                                                \
        return h3 ? make_flux_dim<HLL3>(eq, dim, std::move(r)) : make_flux_dim<HLL>(eq, dim, std::move(r)); \
    }
    if (c.equation == AVH_EQ_EULER1) ORACLE_EULER(1)
    if (c.equation == AVH_EQ_EULER2) ORACLE_EULER(2)
    if (c.equation == AVH_EQ_EULER3) ORACLE_EULER(3)
#undef ORACLE_EULER
    throw std::runtime_error("unknown equation");
}

static void validate(const avh_config& c) {
    if (num_vars(c.equation) == 0) throw std::runtime_error("unknown equation");
    for (int d = 0; d < 3; ++d)
        if (c.n[d] < 1) throw std::runtime_error("n must be >= 1");
    const int dim = active_dim(c);
    if (c.equation >= AVH_EQ_EULER1 && c.equation <= AVH_EQ_EULER3 && dim > c.equation)
        throw std::runtime_error("grid has more active axes than the euler equation");
}

static Simulator* make_sim(const avh_config& c) {
    validate(c);
    auto s = new Simulator();
    s->cfg = c;
    s->dim = active_dim(c);
    s->nv = num_vars(c.equation);
    s->system.flux = make_numflux(c, &s->obeys_cell);
    s->system.dim = s->dim;
    s->ng = s->system.flux->ghosts();
    /* Grid cell lengths: (top-origin)/N, 1 on axes with one cell (Grid.cpp:24-41) */
    for (int d = 0; d < 3; ++d) s->cell[d] = (c.upper[d] - c.lower[d]) / real(c.n[d]);
    if (c.n[1] == 1) s->cell[1] = 1;
    if (c.n[2] == 1) s->cell[2] = 1;
    s->substeps = c.integrator == AVH_INT_FORWARDEULER ? 1 : (c.integrator == AVH_INT_RUNGEKUTTA2 ? 2 : 3);
    const int gy = s->dim > 1 ? s->ng : 0, gz = s->dim > 2 ? s->ng : 0;
    for (int i = 0; i <= s->substeps; ++i) s->vols.emplace_back(s->nv, c.n[0], c.n[1], c.n[2], s->ng, gy, gz);
    return s;
}

} // namespace oracle

/* ================================================================== */
/* extern "C" surface for the Python tests (ctypes).                   */

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void* oracle_sim_create(const avh_config* cfg) {
    try {
        return make_sim(*cfg);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_sim_destroy(void* h) { delete static_cast<Simulator*>(h); }

int oracle_sim_ghosts(void* h) { return static_cast<Simulator*>(h)->ng; }
int oracle_sim_nvars(void* h) { return static_cast<Simulator*>(h)->nv; }
size_t oracle_sim_cells(void* h) { return static_cast<Simulator*>(h)->vols[0].cells(); }

void oracle_sim_set_whole_domain(void* h, int on) { static_cast<Simulator*>(h)->system.whole_domain = on != 0; }

/* Simulator::setSimulationState: copy, then boundary (Simulator.cpp:100-103) */
int oracle_sim_set_state(void* h, const double* ghosted) {
    auto s = static_cast<Simulator*>(h);
    std::memcpy(s->vols[0].d.data(), ghosted, s->vols[0].d.size() * sizeof(double));
    apply_boundary(s->vols[0], s->dim, s->ng, s->cfg.boundary);
    return 0;
}

int oracle_sim_get_state(void* h, double* ghosted) {
    auto s = static_cast<Simulator*>(h);
    std::memcpy(ghosted, s->vols[0].d.data(), s->vols[0].d.size() * sizeof(double));
    return 0;
}

/* Registers a FixedIntervalWriter and performs the drivers' callWriters()
 * at the current time for it (alsvinncli/src/main.cpp:210-213,
 * alsuq/src/run/Runner.cpp:68): with writeInitialTimestep the initial state
 * is a save, numberSaved becomes 1 before the first step. */
int oracle_sim_add_fixed_interval(void* h, double interval, double start, int write_initial) {
    auto s = static_cast<Simulator*>(h);
    s->adjusters.push_back(FixedInterval{interval, start, (size_t)(write_initial ? 0 : 1)});
    if (s->adjusters.back().write(s->time)) s->saves++;
    return 0;
}

void oracle_sim_set_time(void* h, double t, long steps) {
    auto s = static_cast<Simulator*>(h);
    s->time = t;
    s->steps = steps;
}
double oracle_sim_time(void* h) { return static_cast<Simulator*>(h)->time; }
long oracle_sim_steps(void* h) { return static_cast<Simulator*>(h)->steps; }
double oracle_sim_last_dt(void* h) { return static_cast<Simulator*>(h)->last_dt; }
long oracle_sim_saves(void* h) { return static_cast<Simulator*>(h)->saves; }

int oracle_sim_step(void* h) {
    try {
        return static_cast<Simulator*>(h)->step();
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* while (!atEnd()) performStep(), at most max_steps (alsvinncli/src/main.cpp:210-232) */
int oracle_sim_run(void* h, double t_end, long max_steps) {
    auto s = static_cast<Simulator*>(h);
    long n = 0;
    while (!(s->time >= t_end) && n < max_steps) {
        int rc = oracle_sim_step(h);
        ++n;
        if (rc != 0) return rc;
    }
    return 0;
}

/* NumericalFlux::computeFlux over the whole domain (start = end = 0) on a
 * caller-provided ghosted state; `out` is caller-initialised like the
 * reference tests do (CUDANumericalFluxTest.cpp:240-246). */
int oracle_compute_flux(const avh_config* cfg, const double* U, double* out, double* ws) {
    try {
        std::unique_ptr<Simulator> s(make_sim(*cfg));
        std::memcpy(s->vols[0].d.data(), U, s->vols[0].d.size() * sizeof(double));
        std::memcpy(s->vols[1].d.data(), out, s->vols[1].d.size() * sizeof(double));
        ws[0] = ws[1] = ws[2] = 0;
        s->system.flux->compute(s->vols[0], ws, s->vols[1], ivec3{}, ivec3{});
        std::memcpy(out, s->vols[1].d.data(), s->vols[1].d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* NumericalFlux::computeFlux(U, ws, true, out, start, end) on a
 * caller-provided ghosted state and output (NumericalFluxCPU.cpp:170-221). */
int oracle_compute_flux_box(const avh_config* cfg, const double* U, double* out, double* ws, const int* start,
                            const int* end) {
    try {
        std::unique_ptr<Simulator> s(make_sim(*cfg));
        std::memcpy(s->vols[0].d.data(), U, s->vols[0].d.size() * sizeof(double));
        std::memcpy(s->vols[1].d.data(), out, s->vols[1].d.size() * sizeof(double));
        ws[0] = ws[1] = ws[2] = 0;
        s->system.flux->compute(s->vols[0], ws, s->vols[1], ivec3{start[0], start[1], start[2]},
                                ivec3{end[0], end[1], end[2]});
        std::memcpy(out, s->vols[1].d.data(), s->vols[1].d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* ConservedSystem::operator() on a ghosted state (interior + strips). */
int oracle_system(const avh_config* cfg, const double* U, double* out, double* ws, int whole_domain) {
    try {
        std::unique_ptr<Simulator> s(make_sim(*cfg));
        s->system.whole_domain = whole_domain != 0;
        std::memcpy(s->vols[0].d.data(), U, s->vols[0].d.size() * sizeof(double));
        std::memcpy(s->vols[1].d.data(), out, s->vols[1].d.size() * sizeof(double));
        ws[0] = ws[1] = ws[2] = 0;
        s->system(s->vols[0], ws, s->vols[1]);
        std::memcpy(out, s->vols[1].d.data(), s->vols[1].d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* Reconstruction::performReconstruction(in, dir, 0, left, right) over the
 * whole domain. */
int oracle_reconstruct(const avh_config* cfg, const double* U, int dir, double* left, double* right) {
    try {
        validate(*cfg);
        const int dim = active_dim(*cfg);
        auto r = make_recon(cfg->reconstruction, dim);
        const int g = std::max(1, r->ghosts());
        const int nv = num_vars(cfg->equation);
        Volume in(nv, cfg->n[0], cfg->n[1], cfg->n[2], g, dim > 1 ? g : 0, dim > 2 ? g : 0);
        std::memcpy(in.d.data(), U, in.d.size() * sizeof(double));
        Volume L = in, R = in;
        std::memcpy(L.d.data(), left, L.d.size() * sizeof(double));
        std::memcpy(R.d.data(), right, R.d.size() * sizeof(double));
        r->reconstruct(in, dir, 0, L, R, ivec3{}, ivec3{});
        std::memcpy(left, L.d.data(), L.d.size() * sizeof(double));
        std::memcpy(right, R.d.data(), R.d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* Reconstruction::performReconstruction(in, dir, indicator, left, right,
 * start, end) on caller-provided ghosted volumes. */
int oracle_reconstruct_box(const avh_config* cfg, const double* U, int dir, int indicator, double* left,
                           double* right, const int* start, const int* end) {
    try {
        validate(*cfg);
        const int dim = active_dim(*cfg);
        auto r = make_recon(cfg->reconstruction, dim);
        const int g = std::max(1, r->ghosts());
        const int nv = num_vars(cfg->equation);
        if (indicator < 0 || indicator >= nv) throw std::runtime_error("no such indicator variable");
        Volume in(nv, cfg->n[0], cfg->n[1], cfg->n[2], g, dim > 1 ? g : 0, dim > 2 ? g : 0);
        std::memcpy(in.d.data(), U, in.d.size() * sizeof(double));
        Volume L = in, R = in;
        std::memcpy(L.d.data(), left, L.d.size() * sizeof(double));
        std::memcpy(R.d.data(), right, R.d.size() * sizeof(double));
        r->reconstruct(in, dir, indicator, L, R, ivec3{start[0], start[1], start[2]}, ivec3{end[0], end[1], end[2]});
        std::memcpy(left, L.d.data(), L.d.size() * sizeof(double));
        std::memcpy(right, R.d.data(), R.d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* Interface flux F(left, right) in direction dir between two conserved
 * states, as computeFluxForStencil does after fetchAllVariables. */
double oracle_interface_flux(int equation, int flux, int dir, double gamma, const double* ql,
                             const double* qr, double* F) {
#define ORACLE_DIRS(FLUX, EQ)                                                         \
    {                                                                                 \
        auto aL = eq.make_all(ql);                                                    \
        auto aR = eq.make_all(qr);                                                    \
        if (dir == 0) return FLUX::flux<0>(eq, aL, aR, F);                            \
        if (dir == 1) return FLUX::flux<EQ::NSD >= 2 ? 1 : 0>(eq, aL, aR, F);         \
        return FLUX::flux<EQ::NSD >= 3 ? 2 : 0>(eq, aL, aR, F);                       \
    }
    if (equation == AVH_EQ_BURGERS) {
        Burgers eq;
        auto aL = eq.make_all(ql);
        auto aR = eq.make_all(qr);
        return Godunov::flux<0>(eq, aL, aR, F);
    }
#define ORACLE_E(NSD)                                                                 \
    if (equation == NSD) {                                                            \
        Euler<NSD> eq;                                                                \
        eq.gamma = gamma;                                                             \
        using E = Euler<NSD>;                                                         \
        if (flux == AVH_FLUX_HLL3) ORACLE_DIRS(HLL3, E) else ORACLE_DIRS(HLL, E)      \
    }
    ORACLE_E(1)
    ORACLE_E(2)
    ORACLE_E(3)
#undef ORACLE_E
#undef ORACLE_DIRS
    return -1;
}

/* Point flux of a conserved state (Euler/Burgers computePointFlux). */
int oracle_point_flux(int equation, int dir, double gamma, const double* q, double* F) {
    if (equation == AVH_EQ_BURGERS) {
        Burgers eq;
        eq.point_flux<0>(eq.make_all(q), F);
        return 0;
    }
#define ORACLE_P(NSD)                                                                  \
    if (equation == NSD) {                                                             \
        Euler<NSD> eq;                                                                 \
        eq.gamma = gamma;                                                              \
        auto a = eq.make_all(q);                                                       \
        if (dir == 0) eq.point_flux<0>(a, F);                                          \
        else if (dir == 1) eq.point_flux<NSD >= 2 ? 1 : 0>(a, F);                      \
        else eq.point_flux<NSD >= 3 ? 2 : 0>(a, F);                                    \
        return 0;                                                                      \
    }
    ORACLE_P(1)
    ORACLE_P(2)
    ORACLE_P(3)
#undef ORACLE_P
    return AVH_EINVAL;
}

/* BoundaryCPU::applyBoundaryConditions on a ghosted state. */
int oracle_apply_boundary(const avh_config* cfg, int ghosts, double* U) {
    try {
        validate(*cfg);
        const int dim = active_dim(*cfg);
        const int nv = num_vars(cfg->equation);
        Volume v(nv, cfg->n[0], cfg->n[1], cfg->n[2], ghosts, dim > 1 ? ghosts : 0, dim > 2 ? ghosts : 0);
        std::memcpy(v.d.data(), U, v.d.size() * sizeof(double));
        apply_boundary(v, dim, ghosts, cfg->boundary);
        std::memcpy(U, v.d.data(), v.d.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AVH_EINVAL;
    }
}

/* CellComputer::obeysConstraints over a whole ghosted state; 1 = obeys. */
int oracle_obeys(void* h, const double* ghosted) {
    auto s = static_cast<Simulator*>(h);
    std::vector<real> q(s->nv);
    const size_t n = s->vols[0].cells();
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < s->nv; ++k) q[k] = ghosted[(size_t)k * n + i];
        if (!s->obeys_cell(q.data())) return 0;
    }
    return 1;
}

/* Primitive -> conserved over ncells (CellComputer::computeFromPrimitive,
 * Euler::computeConserved / Burgers identity). SoA in and out. */
int oracle_from_primitive(int equation, double gamma, const double* prim, double* cons, size_t ncells) {
    const int nv = num_vars(equation);
    if (nv == 0) return AVH_EINVAL;
    for (size_t i = 0; i < ncells; ++i) {
        real p[5], q[5];
        for (int k = 0; k < nv; ++k) p[k] = prim[(size_t)k * ncells + i];
        if (equation == AVH_EQ_BURGERS) q[0] = p[0];
        else if (equation == AVH_EQ_EULER1) { Euler<1> e; e.gamma = gamma; e.from_primitive(p, q); }
        else if (equation == AVH_EQ_EULER2) { Euler<2> e; e.gamma = gamma; e.from_primitive(p, q); }
        else { Euler<3> e; e.gamma = gamma; e.from_primitive(p, q); }
        for (int k = 0; k < nv; ++k) cons[(size_t)k * ncells + i] = q[k];
    }
    return 0;
}

/* Point flux of an AllVariables given field by field: a = {rho, m[nsd],
 * E, p, u[nsd]} (EulerEquationTest builds such states directly,
 * test/library_tests/src/fvm/equation/EulerEquationTest.cpp:27). */
int oracle_point_flux_all(int nsd, int dir, const double* a, double* F) {
#define ORACLE_PA(NSD)                                                               \
    if (nsd == NSD) {                                                                \
        Euler<NSD> eq;                                                               \
        Euler<NSD>::All v;                                                           \
        v.rho = a[0];                                                                \
        for (int c = 0; c < NSD; ++c) v.m[c] = a[1 + c];                             \
        v.E = a[1 + NSD];                                                            \
        v.p = a[2 + NSD];                                                            \
        for (int c = 0; c < NSD; ++c) v.u[c] = a[3 + NSD + c];                       \
        if (dir == 0) eq.point_flux<0>(v, F);                                        \
        else if (dir == 1) eq.point_flux<NSD >= 2 ? 1 : 0>(v, F);                    \
        else eq.point_flux<NSD >= 3 ? 2 : 0>(v, F);                                  \
        return 0;                                                                    \
    }
    ORACLE_PA(1)
    ORACLE_PA(2)
    ORACLE_PA(3)
#undef ORACLE_PA
    return AVH_EINVAL;
}

/* Numerical flux between two AllVariables given field by field (HLLTest
 * passes AllVariables(1, {2,3,4}, 5, 6, {7,8,9}) straight to the flux,
 * test/library_tests/src/fvm/flux/HLLTest.cpp:24-71). */
double oracle_flux_all(int nsd, int flux, int dir, double gamma, const double* aL, const double* aR, double* F) {
#define ORACLE_FA(NSD)                                                                 \
    if (nsd == NSD) {                                                                  \
        Euler<NSD> eq;                                                                 \
        eq.gamma = gamma;                                                              \
        Euler<NSD>::All L, R;                                                          \
        const double* src[2] = {aL, aR};                                               \
        Euler<NSD>::All* dst[2] = {&L, &R};                                            \
        for (int k = 0; k < 2; ++k) {                                                  \
            const double* a = src[k];                                                  \
            dst[k]->rho = a[0];                                                        \
            for (int c = 0; c < NSD; ++c) dst[k]->m[c] = a[1 + c];                     \
            dst[k]->E = a[1 + NSD];                                                    \
            dst[k]->p = a[2 + NSD];                                                    \
            for (int c = 0; c < NSD; ++c) dst[k]->u[c] = a[3 + NSD + c];               \
        }                                                                              \
        if (flux == AVH_FLUX_HLL3) {                                                   \
            if (dir == 0) return HLL3::flux<0>(eq, L, R, F);                           \
            if (dir == 1) return HLL3::flux<NSD >= 2 ? 1 : 0>(eq, L, R, F);            \
            return HLL3::flux<NSD >= 3 ? 2 : 0>(eq, L, R, F);                          \
        }                                                                              \
        if (dir == 0) return HLL::flux<0>(eq, L, R, F);                                \
        if (dir == 1) return HLL::flux<NSD >= 2 ? 1 : 0>(eq, L, R, F);                 \
        return HLL::flux<NSD >= 3 ? 2 : 0>(eq, L, R, F);                               \
    }
    ORACLE_FA(1)
    ORACLE_FA(2)
    ORACLE_FA(3)
#undef ORACLE_FA
    return -1;
}

/* IntegratorConvergenceTest's ODESystem (test/library_tests/src/fvm/
 * integrator/IntegratorConvergenceTest.cpp:28-43): output = h*u, wave
 * speeds (1,0,0), cell sizes (1,1,1), cfl 1, so each substep integrates
 * du/dt = u with step h through the integrator's own vector arithmetic.
 * Runs `steps` steps from u0 and returns u. */
double oracle_integrate_ode(int integrator, double h, long steps, double u0) {
    const int substeps = integrator == AVH_INT_FORWARDEULER ? 1 : (integrator == AVH_INT_RUNGEKUTTA2 ? 2 : 3);
    std::vector<Volume> vols;
    for (int i = 0; i <= substeps; ++i) {
        vols.emplace_back(1, 1, 1, 1, 0, 0, 0);
        vols.back().d[0] = u0;
    }
    auto sys = [h](const Volume& U, real* ws, Volume& out) {
        out.d[0] = h * U.d[0];
        ws[0] = 1;
        ws[1] = 0;
        ws[2] = 0;
    };
    auto cdt = [](const real* ws) {
        real total = 0;
        const real cell[3] = {1, 1, 1};
        for (int d = 0; d < 3; ++d) total += ws[d] / cell[d];
        return real(1) / total;
    };
    for (long k = 0; k < steps; ++k) {
        real dt = 1;
        for (int s = 0; s < substeps; ++s) dt = integrator_substep(integrator, vols, s, dt, 1.0, sys, cdt);
        std::swap(vols[0], vols[substeps]);
    }
    return vols[0].d[0];
}

/* std::mt19937_64 draws (alsuq/src/generator/STLMersenne.cpp:38-85 uses the
 * default-seeded engine).  Writes n raw 64-bit outputs after discarding
 * `skip` values. */
void oracle_mt19937_64(uint64_t seed, uint64_t skip, uint64_t n, uint64_t* out) {
    std::mt19937_64 g(seed);
    g.discard(skip);
    for (uint64_t i = 0; i < n; ++i) out[i] = g();
}

/* std::uniform_real_distribution<double>(0,1) draws on a fresh engine per
 * value, advanced by `skip + i` (the STLMersenne discard pattern). */
void oracle_uniform01(uint64_t seed, uint64_t skip, uint64_t n, double* out) {
    std::mt19937_64 g(seed);
    g.discard(skip);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    for (uint64_t i = 0; i < n; ++i) out[i] = dist(g);
}

/* Kinetic energy of velocity component `comp` of one interior SoA state
 * (nv variables of ncells each): (sum in cell order of 0.5*rho*u*u) * dv,
 * u = m_comp / rho.  A new functional; the Moment functional
 * (alsfvm/src/functional/Moment.cpp:31-45) is its model. */
double oracle_kinetic_energy(const double* q, long ncells, int comp, double dv) {
    double acc = 0;
    for (long i = 0; i < ncells; ++i) {
        const double rho = q[i];
        const double u = q[(long)(1 + comp) * ncells + i] / rho;
        acc += 0.5 * rho * (u * u);
    }
    return acc * dv;
}

static int value_at_boundary(int x, int n, int bc) {
    /* ValueAtBoundary<PERIODIC/NEUMANN> (alsfvm/include/alsfvm/boundary/ValueAtBoundary.hpp) */
    if (bc == 0) {
        while (x < 0) x += n;
        return x % n;
    }
    return std::max(0, std::min(x, n - 1));
}

static double structure_power(double x, double p) {
    /* FastPower<k> for p = 1..5 (x * (x * ...)), PowPower otherwise
     * (alsfvm/include/alsfvm/functional/structure_common.hpp:156-180) */
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * (x * x);
    if (p == 4.0) return x * (x * (x * x));
    if (p == 5.0) return x * (x * (x * (x * x)));
    return std::pow(x, p);
}

/* computeStructureCubeCPU for one variable of an interior field u[ny][nx]
 * (dims 1 or 2): for k, j, i in order, for h = 1..number_of_h-1, the
 * forEachPointInComputeStructureCube ring in its loop order, each term
 * divided by the cell count as it is added
 * (alsfvm/include/alsfvm/functional/structure_common.hpp:36-154). */
void oracle_structure_cube(const double* u, int nx, int ny, int dims, int bc, int number_of_h, double p,
                           double* out) {
    const int nz = 1;
    for (int h = 0; h < number_of_h; ++h) out[h] = 0;
    const double ncell = double(nx * ny * nz);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const double uc = u[(long)j * nx + i];
            for (int h = 1; h < number_of_h; ++h) {
                for (int d = 0; d < dims; ++d) {
                    for (int side = 0; side < 2; ++side) {
                        const bool yDir = d == 1, xDir = d == 0;
                        const int yStart = yDir ? (side == 0 ? j - h : j + h) : (dims == 2 ? j - h + 1 : 0);
                        const int yEnd = yDir ? yStart + 1 : (dims == 2 ? j + h : 1);
                        const int xStart = xDir ? (side == 0 ? i - h : i + h) : i - h;
                        const int xEnd = xDir ? xStart + 1 : i + h + 1;
                        for (int y = yStart; y < yEnd; ++y)
                            for (int x = xStart; x < xEnd; ++x) {
                                const double v = u[(long)value_at_boundary(y, ny, bc) * nx +
                                                   value_at_boundary(x, nx, bc)];
                                out[h] += structure_power(std::fabs(uc - v), p) / ncell;
                            }
                    }
                }
            }
        }
    }
}

} /* extern "C" */
