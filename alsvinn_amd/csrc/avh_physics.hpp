// Device-side physics of the FP64 finite-volume stage: equation state
// algebra, reconstructions and numerical fluxes.  Every expression keeps the
// operation order of the reference CPU path so that results are bit-equal:
// no FMA contraction (also forced by -ffp-contract=off), IEEE divide and
// sqrt, the same association of sums and products.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace avh {

// Correctly rounded FP64 square root.  For x in [2^-767, DBL_MAX] this is
// the operation sequence the compiler emits for sqrt() on gfx950 -- v_rsq_f64,
// then g = x*r, h = r/2 and two Goldschmidt / Newton corrections by fused
// multiply-adds -- without the steps that are the identity on that range
// (scaling tiny arguments by 2^256 and the result back by 2^-128, and the
// zero / infinity fix-up), so the result has the same bits as sqrt(x).  Any
// other argument (tiny, zero, negative, infinite, NaN) takes sqrt() itself.
__device__ __forceinline__ double sqrt_rn(double x) {
    if (x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023) {
        const double r = __builtin_amdgcn_rsq(x);
        double g = x * r;
        double h = r * 0.5;
        const double d0 = fma(-h, g, 0.5);
        g = fma(g, d0, g);
        h = fma(h, d0, h);
        const double d1 = fma(-g, g, x);
        g = fma(d1, h, g);
        const double d2 = fma(-g, g, x);
        return fma(d2, h, g);
    }
    return sqrt(x);
}

// sqrt_rn or sqrt(): the same bits either way; the 3D kernel, whose code is
// large, keeps the compiler's single expansion.
template <bool FAST>
__device__ __forceinline__ double sqrt_sel(double x) {
    if constexpr (FAST) {
        return sqrt_rn(x);
    } else {
        return sqrt(x);
    }
}

// ---------------------------------------------------------------------------
// Equations.  A conserved state is NV doubles: burgers {u};
// euler<nsd> {rho, m[0..nsd-1], E}.

struct BurgersEq {
    static constexpr int NV = 1;
    static constexpr int NSD = 3;  // the flux is the same along every axis
    struct All {
        double u;
    };
    // makeAllVariables (alsfvm/include/alsfvm/equation/burgers/Burgers.hpp:220-224)
    __device__ __forceinline__ static All make_all(const double* q, double) { return All{q[0]}; }
    // computePointFlux: u*u/2 (Burgers.hpp:158-164)
    __device__ __forceinline__ static double point_flux(double u) { return u * u / 2; }
    // CPUCellComputer::obeysConstraints + Burgers::obeysConstraints
    // (alsfvm/src/equation/CPUCellComputer.cpp:79-125, Burgers.hpp:212-217)
    __device__ __forceinline__ static bool obeys(const double* q, double) {
        const double u = q[0];
        return !__builtin_isnan(u) && !__builtin_isinf(u);
    }
};

template <int NSD_>
struct EulerEq {
    static constexpr int NSD = NSD_;
    static constexpr int NV = NSD_ + 2;
    struct All {
        double rho, m[NSD_], E, p, u[NSD_];
    };
    // vec::dot, left to right (alsutils/include/alsutils/vec{1,2,3}.hpp)
    __device__ __forceinline__ static double dot(const double* m) {
        double s = m[0] * m[0];
#pragma unroll
        for (int c = 1; c < NSD_; ++c) s = s + m[c] * m[c];
        return s;
    }
    // computeExtra: ie = E - 0.5*m.m/rho, u = m/rho, p = (gamma-1)*ie
    // (alsfvm/include/alsfvm/equation/euler/Euler.hpp:204-211)
    __device__ __forceinline__ static All make_all(const double* q, double gamma) {
        All a;
        a.rho = q[0];
#pragma unroll
        for (int c = 0; c < NSD_; ++c) a.m[c] = q[1 + c];
        a.E = q[NSD_ + 1];
        const double ie = a.E - 0.5 * dot(a.m) / a.rho;
#pragma unroll
        for (int c = 0; c < NSD_; ++c) a.u[c] = a.m[c] / a.rho;
        a.p = (gamma - 1) * ie;
        return a;
    }
    // computePointFlux along DIR (Euler.hpp:172-181)
    template <int DIR>
    __device__ __forceinline__ static void point_flux(const All& a, double* F) {
        F[0] = a.m[DIR];
#pragma unroll
        for (int c = 0; c < NSD_; ++c) F[1 + c] = a.m[c] * a.u[DIR];
        F[1 + DIR] = F[1 + DIR] + a.p;
        F[NSD_ + 1] = (a.E + a.p) * a.u[DIR];
    }
    // conserved and extra variables finite, rho > 0, p > 0
    // (alsfvm/src/equation/CPUCellComputer.cpp:79-125, Euler.hpp:269-277)
    __device__ __forceinline__ static bool obeys(const double* q, double gamma) {
        bool ok = true;
#pragma unroll
        for (int i = 0; i < NV; ++i) ok = ok && !__builtin_isnan(q[i]) && !__builtin_isinf(q[i]);
        const All a = make_all(q, gamma);
        ok = ok && !__builtin_isnan(a.p) && !__builtin_isinf(a.p);
#pragma unroll
        for (int c = 0; c < NSD_; ++c) ok = ok && !__builtin_isnan(a.u[c]) && !__builtin_isinf(a.u[c]);
        return ok && (a.rho > 0) && (a.p > 0);
    }
};

// ---------------------------------------------------------------------------
// Numerical fluxes F(left, right) along DIR; each returns the interface wave
// speed used for the CFL condition.

// burgers::Godunov (alsfvm/include/alsfvm/numflux/burgers/Godunov.hpp:40-60):
// F = max(f(max(uL,0)), f(min(uR,0))), wave speed max(|uL|, |uR|).
struct GodunovFlux {
    typedef GodunovFlux Plain;
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const BurgersEq::All& L, const BurgersEq::All& R,
                                                  double, double* F) {
        const double fl = BurgersEq::point_flux(fmax(L.u, 0.0));
        const double fr = BurgersEq::point_flux(fmin(R.u, 0.0));
        F[0] = fmax(fl, fr);
        return fmax(fabs(L.u), fabs(R.u));
    }
};

// euler::HLL (alsfvm/include/alsfvm/numflux/euler/HLL.hpp).  The reference's
// if/else chain is evaluated branch-free: every candidate is computed and the
// chain's answer is selected, which gives the same value as the branch taken.
template <bool FAST_SQRT>
struct HLLFluxT {
    typedef HLLFluxT<false> Plain;
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const typename EQ::All& L, const typename EQ::All& R,
                                                  double g, double* F) {
        constexpr int NSD = EQ::NSD;
        constexpr int NV = EQ::NV;
        const double wl = sqrt_sel<FAST_SQRT>(L.rho);
        const double wr = sqrt_sel<FAST_SQRT>(R.rho);
        const double rho = (L.rho + R.rho) / 2;
        const double ud = (L.u[DIR] * wl + R.u[DIR] * wr) / (wl + wr);
        const double p = (L.p * wl + R.p * wr) / (wl + wr);
        const double cs = sqrt_sel<FAST_SQRT>(g * p / rho);
        const double sl = fmin(L.u[DIR] - sqrt_sel<FAST_SQRT>(g * L.p / L.rho), ud - cs);
        const double sr = fmax(R.u[DIR] + sqrt_sel<FAST_SQRT>(g * R.p / R.rho), ud + cs);
        double fl[NV], fr[NV], ul[NV], ur[NV];
        EQ::template point_flux<DIR>(L, fl);
        EQ::template point_flux<DIR>(R, fr);
        ul[0] = L.rho;
        ur[0] = R.rho;
#pragma unroll
        for (int c = 0; c < NSD; ++c) {
            ul[1 + c] = L.m[c];
            ur[1 + c] = R.m[c];
        }
        ul[NSD + 1] = L.E;
        ur[NSD + 1] = R.E;
        const double srsl = sr * sl;
        const double den = sr - sl;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double mid = (sr * fl[i] - sl * fr[i] + srsl * (ur[i] - ul[i])) / den;
            double v = mid;
            v = (sr < 0) ? fr[i] : v;
            v = (sl > 0) ? fl[i] : v;
            v = (sl == 0) ? 0.0 : v;
            F[i] = v;
        }
        return fmax(fabs(sl), fabs(sr));
    }
};

// euler::HLL3 (alsfvm/include/alsfvm/numflux/euler/HLL3.hpp:52-148).  The two
// star-state branches (sm >= 0 / sm < 0) share one formula on the side's
// variables, so the side is selected first and the formula evaluated once.
template <bool FAST_SQRT>
struct HLL3FluxT {
    typedef HLL3FluxT<false> Plain;
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const typename EQ::All& L, const typename EQ::All& R,
                                                  double g, double* F) {
        constexpr int NSD = EQ::NSD;
        constexpr int NV = EQ::NV;
        // computeHLLSpeeds
        const double wl = sqrt_sel<FAST_SQRT>(L.rho);
        const double wr = sqrt_sel<FAST_SQRT>(R.rho);
        const double rho = (L.rho + R.rho) / 2;
        const double ud = (L.u[DIR] * wl + R.u[DIR] * wr) / (wl + wr);
        const double p = (wl * L.p + wr * R.p) / (wl + wr);
        const double cfl = sqrt_sel<FAST_SQRT>(g * L.p / L.rho);
        const double cfr = sqrt_sel<FAST_SQRT>(g * R.p / R.rho);
        const double correct = 0.5 * fmax(0.0, L.u[DIR] - R.u[DIR]);
        const double cs = sqrt_sel<FAST_SQRT>(g * p / rho);
        const double sl = fmin(L.u[DIR] + correct - cfl, ud - cs);
        const double sr = fmax(R.u[DIR] - correct + cfr, ud + cs);

        double fl[NV], fr[NV];
        EQ::template point_flux<DIR>(L, fl);
        EQ::template point_flux<DIR>(R, fr);
        const double pl = L.p, pr = R.p;
        const double udl = L.u[DIR] - sl;
        const double udr = R.u[DIR] - sr;
        const double aa = udr * R.rho - udl * L.rho;
        const double sm = (R.m[DIR] * udr - L.m[DIR] * udl + pr - pl) / aa;

        const bool left = sm >= 0;
        const double rk = left ? L.rho : R.rho;
        const double udk = left ? udl : udr;
        const double sk = left ? sl : sr;
        const double pk = left ? pl : pr;
        const double uk = left ? L.u[DIR] : R.u[DIR];
        const double ek = left ? L.E : R.E;

        double mid[NV], uc[NV], fk[NV];
        mid[0] = rk * udk / (sm - sk);
#pragma unroll
        for (int c = 0; c < NSD; ++c) {
            // us = (fR.m - fL.m - sR*mR + sL*mL)/aa with us[DIR] = sm
            const double us =
                (c == DIR) ? sm : (fr[1 + c] - fl[1 + c] - R.m[c] * sr + L.m[c] * sl) / aa;
            mid[1 + c] = us * mid[0];
            uc[1 + c] = left ? L.m[c] : R.m[c];
        }
        const double pm = pk + rk * (uk - sm) * udk;
        mid[NSD + 1] = (udk * ek + pk * uk - pm * sm) / (sm - sk);
        uc[0] = rk;
        uc[NSD + 1] = ek;
#pragma unroll
        for (int i = 0; i < NV; ++i) fk[i] = left ? fl[i] : fr[i];
        // subsonic interfaces (sL < 0 < sR) take the star branch; when the
        // whole wavefront is subsonic the other three branches are skipped
        const bool star = !(sl == 0) && !(sl > 0) && !(sr < 0);
        if (__all(star)) {
#pragma unroll
            for (int i = 0; i < NV; ++i) F[i] = fk[i] + sk * (mid[i] - uc[i]);
        } else {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double v = fk[i] + sk * (mid[i] - uc[i]);
                v = (sr < 0) ? fr[i] : v;
                v = (sl > 0) ? fl[i] : v;
                v = (sl == 0) ? 0.0 : v;
                F[i] = v;
            }
        }
        return fmax(fabs(sl), fabs(sr));
    }
};

typedef HLLFluxT<true> HLLFlux;
typedef HLL3FluxT<true> HLL3Flux;

// ---------------------------------------------------------------------------
// Reconstructions on an LDS tile.  `s` points at variable 0 of the tile,
// `vs` is the per-variable stride, `i` the cell, `sd` the stride along the
// reconstruction direction.  The indicator is variable 0 (rho or u).

enum { RECON_NONE = 0, RECON_WENO2 = 1, RECON_ENO2 = 2, RECON_ENO3 = 3 };

template <int RECON>
struct ReconTraits;
template <>
struct ReconTraits<RECON_NONE> {
    static constexpr int GHOST = 1;
};
template <>
struct ReconTraits<RECON_WENO2> {
    static constexpr int GHOST = 2;
};
template <>
struct ReconTraits<RECON_ENO2> {
    static constexpr int GHOST = 2;
};
template <>
struct ReconTraits<RECON_ENO3> {
    static constexpr int GHOST = 3;
};

// WENO2 (alsfvm/include/alsfvm/reconstruction/WENO2.hpp:31-95; epsilon 1e-8,
// WENOCoefficients.hpp:21).  RIGHT selects the value at the cell's right face
// (the reference's `right` volume), otherwise the left face.
template <int NV, bool RIGHT>
__device__ __forceinline__ void weno2_face(const double* s, int vs, int i, int sd, double* out) {
    const double eps = 1e-8;
    const double w0 = s[i];
    const double dp = s[i + sd] - w0;
    const double b0 = dp * dp;
    const double dm = w0 - s[i - sd];
    const double b1 = dm * dm;
    const double d0 = 3 * (eps + b0) * (eps + b0);
    const double d1 = 3 * (eps + b1) * (eps + b1);
    if (RIGHT) {
        const double a0 = 2 / d0;
        const double a1 = 1 / d1;
        const double w0r = a0 / (a0 + a1);
        const double w1r = a1 / (a0 + a1);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* u = s + v * vs;
            out[v] = 0.5 * (w0r * u[i + sd] + (3 * w1r + w0r) * u[i] - w1r * u[i - sd]);
        }
    } else {
        const double a0 = 1 / d0;
        const double a1 = 2 / d1;
        const double w0l = a0 / (a0 + a1);
        const double w1l = a1 / (a0 + a1);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* u = s + v * vs;
            out[v] = 0.5 * (w1l * u[i - sd] + (3 * w0l + w1l) * u[i] - w0l * u[i + sd]);
        }
    }
}

// ENO coefficient tables (alsfvm/src/reconstruction/ENOCoefficients.cpp:27-40)
template <int ORDER>
__device__ __forceinline__ double eno_coef(int r, int j);
template <>
__device__ __forceinline__ double eno_coef<2>(int r, int j) {
    const double c00 = 3.0 / 2.0, c01 = -1.0 / 2.0, c10 = 1.0 / 2.0, c11 = 1.0 / 2.0,
                 c20 = -1.0 / 2.0, c21 = 3.0 / 2.0;
    const double a = (r == 0) ? c00 : ((r == 1) ? c10 : c20);
    const double b = (r == 0) ? c01 : ((r == 1) ? c11 : c21);
    return j == 0 ? a : b;
}
template <>
__device__ __forceinline__ double eno_coef<3>(int r, int j) {
    const double t[4][3] = {{11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0},
                            {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0},
                            {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
                            {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0}};
    double v = t[0][j];
    v = (r == 1) ? t[1][j] : v;
    v = (r == 2) ? t[2][j] : v;
    v = (r == 3) ? t[3][j] : v;
    return v;
}

// ENO<order> (alsfvm/src/reconstruction/ENOCPU.cpp:41-187): the stencil shift
// grows by one at each divided-difference level where |dd[x-1]| < |dd[x]|;
// divided differences are plain differences of the indicator,
// dd0[y] = w[y+1]-w[y], dd1[y] = dd0[y+1]-dd0[y] (ENOCPU.cpp:205-262).
// left face uses coefficients[shift], right face coefficients[shift+1].
template <int ORDER, int NV, bool RIGHT>
__device__ __forceinline__ void eno_face(const double* s, int vs, int i, int sd, double* out) {
    int shift = 0;
    {
        const double ddr = s[i + sd] - s[i];
        const double ddl = s[i] - s[i - sd];
        if (fabs(ddl) < fabs(ddr)) shift++;
    }
    if (ORDER == 3) {
        const double d0p = s[i + 2 * sd] - s[i + sd];
        const double d00 = s[i + sd] - s[i];
        const double d0m = s[i] - s[i - sd];
        const double ddr = d0p - d00;
        const double ddl = d00 - d0m;
        if (fabs(ddl) < fabs(ddr)) shift++;
    }
    const int r = RIGHT ? shift + 1 : shift;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double* u = s + v * vs;
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < ORDER; ++j) acc += eno_coef<ORDER>(r, j) * u[i + (j - shift) * sd];
        out[v] = acc;
    }
}

template <int RECON, int NV, bool RIGHT>
__device__ __forceinline__ void recon_face(const double* s, int vs, int i, int sd, double* out) {
    if (RECON == RECON_NONE) {
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = s[v * vs + i];
    } else if (RECON == RECON_WENO2) {
        weno2_face<NV, RIGHT>(s, vs, i, sd, out);
    } else if (RECON == RECON_ENO2) {
        eno_face<2, NV, RIGHT>(s, vs, i, sd, out);
    } else {
        eno_face<3, NV, RIGHT>(s, vs, i, sd, out);
    }
}

}  // namespace avh
