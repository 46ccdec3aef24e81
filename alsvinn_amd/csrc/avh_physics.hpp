// Device-side physics of the FP64 finite-volume stage: equation state
// algebra, reconstructions and numerical fluxes.  Every expression keeps the
// operation order of the reference CPU path so that results are bit-equal:
// no FMA contraction (also forced by -ffp-contract=off), IEEE divide and
// sqrt, the same association of sums and products.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace avh {

// Correctly rounded FP64 square root for x in [2^-767, DBL_MAX]: the
// operation sequence the compiler emits for sqrt() on gfx950 -- v_rsq_f64,
// then g = x*r, h = r/2 and two Goldschmidt / Newton corrections by fused
// multiply-adds -- without the steps that are the identity on that range
// (scaling tiny arguments by 2^256 and the result back by 2^-128, and the
// zero / infinity fix-up), so the result has the same bits as sqrt(x) there.
__device__ __forceinline__ double sqrt_core(double x) {
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

// Quotients that share a denominator.  The compiler divides n/d in FP64 on
// gfx950 as
//   d' = div_scale(d), n' = div_scale(n), r = rcp(d'),
//   r = fma(r, fma(-d', r, 1), r) twice, q0 = n' * r,
//   q = div_fmas(fma(-d', q0, n'), r, q0), result div_fixup(q, d, n).
// When d > 0, d and 1/d are normal, n is zero or has |n| >= 2^-968 and the
// quotient's binary exponent lies in [-1000, 760], the scalings are the
// identity (no VCC scale), div_fmas is a plain fma and div_fixup returns q
// -- or, for n = +-0, the zero with the sign of n.  The quotient is then a
// function of the refined reciprocal r alone, so r is computed once per
// denominator (div_recip) and each numerator costs one multiply and two
// fmas (div_shared).  The residual is formed as d*q0 - n and subtracted, the
// exact negation of the compiler's n - d*q0 added: nonzero numerators give
// the same bits, and n = +-0 gives +-0 as div_fixup does.
__device__ __forceinline__ double div_recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    return r;
}
__device__ __forceinline__ double div_shared(double n, double d, double r) {
    const double q0 = n * r;
    return fma(-fma(d, q0, -n), r, q0);
}
// lo <= n <= hi (false for zero, negative values, inf and NaN)
__device__ __forceinline__ bool positive_within(double n, double lo, double hi) {
    return (n >= lo) & (n <= hi);
}
// Binary exponent e of x = f * 2^e, 0.5 <= |f| < 1 (v_frexp_exp_i32_f64):
// 0 for +-0, the true exponent (<= -1022) for subnormals.  It is also 0 for
// inf and NaN, so a check built on it must exclude those by another test.
// `x is zero or |x| >= 2^(k-1)` is `frexp_exp(x) >= k`.
__device__ __forceinline__ int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }

// Arithmetic of an interface's reconstruction and flux.
//   PlainOps   the compiler's sqrt and divide everywhere.
//   SharedOps  sqrt_core, and shared reciprocals for the divisions by rho,
//              by sqrt(rhoL) + sqrt(rhoR) and by the WENO weight sum; every
//              operand is checked against the range where this gives the
//              bits of PlainOps, and a failed check clears `ok`, after which
//              the caller recomputes the interface with PlainOps.  The
//              checks feed a flag instead of branching so that the interface
//              stays one straight-line block for the scheduler.
// The checks combine with non-short-circuit '&' so that they stay lane
// masks and never turn into branches.
struct PlainOps {
    static constexpr bool SHARED = false;
    __device__ __forceinline__ static double sqrt_(double x, bool&) { return sqrt(x); }
    // sqrt of an argument already known to lie in [2^-767, DBL_MAX] under
    // SharedOps (a density that passed make_all_ops' check, or a quotient
    // of checked operands)
    __device__ __forceinline__ static double sqrt_in_range(double x) { return sqrt(x); }
};
struct SharedOps {
    static constexpr bool SHARED = true;
    __device__ __forceinline__ static double sqrt_(double x, bool& ok) {
        ok = ok & (x >= 0x1p-767) & (x <= 0x1.fffffffffffffp+1023);
        return sqrt_core(x);
    }
    __device__ __forceinline__ static double sqrt_in_range(double x) { return sqrt_core(x); }
};
template <bool SHARED>
struct OpsFor {
    typedef PlainOps type;
};
template <>
struct OpsFor<true> {
    typedef SharedOps type;
};

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
    template <class OPS>
    __device__ __forceinline__ static All make_all_ops(const double* q, double g, bool&) { return make_all(q, g); }
    // computePointFlux: u*u/2 (Burgers.hpp:158-164)
    __device__ __forceinline__ static double point_flux(double u) { return u * u / 2; }
    // CPUCellComputer::obeysConstraints + Burgers::obeysConstraints
    // (alsfvm/src/equation/CPUCellComputer.cpp:79-125, Burgers.hpp:212-217)
    __device__ __forceinline__ static bool obeys(const double* q, double) {
        const double u = q[0];
        return !__builtin_isnan(u) && !__builtin_isinf(u);
    }
    __device__ __forceinline__ static bool obeys_fast(const double* q, double g) { return obeys(q, g); }
};

template <int NSD_>
struct EulerEq {
    static constexpr int NSD = NSD_;
    static constexpr int NV = NSD_ + 2;
    struct All {
        double rho, m[NSD_], E, p, u[NSD_];
        double rinv;  // refined reciprocal of rho (div_recip), 0 when not formed
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
        a.rinv = 0.0;
        return a;
    }
    // make_all under OPS.  SharedOps forms one reciprocal of rho for the
    // divisions by rho.  It checks that rho is in [2^-100, 2^100], that
    // m.m <= 2^580 (so every |m[c]| <= 2^290 (1 + 2^-54); false for inf and
    // NaN) and, by binary exponent, that every momentum is zero or
    // |m[c]| >= 2^-433.  Then the numerators m[c] are zero or within
    // [2^-433, 2^291] and 0.5*m.m is zero or within [2^-867, 2^579] (the
    // computed sum of squares is at least its largest term, 2^-866 or more,
    // and nothing underflows), so every quotient lies within [2^-967, 2^679].
    template <class OPS>
    __device__ __forceinline__ static All make_all_ops(const double* q, double gamma, bool& ok) {
        if constexpr (!OPS::SHARED) {
            return make_all(q, gamma);
        } else {
            All a;
            a.rho = q[0];
#pragma unroll
            for (int c = 0; c < NSD_; ++c) a.m[c] = q[1 + c];
            a.E = q[NSD_ + 1];
            const double mm = dot(a.m);
            int emin = frexp_exp(a.m[0]);
#pragma unroll
            for (int c = 1; c < NSD_; ++c) emin = min(emin, frexp_exp(a.m[c]));
            ok = ok & positive_within(a.rho, 0x1p-100, 0x1p100) & (mm <= 0x1p580) & (emin >= -432);
            const double r = div_recip(a.rho);
            const double ie = a.E - div_shared(0.5 * mm, a.rho, r);
#pragma unroll
            for (int c = 0; c < NSD_; ++c) a.u[c] = div_shared(a.m[c], a.rho, r);
            a.rinv = r;
            a.p = (gamma - 1) * ie;
            return a;
        }
    }
    // sqrt(g*p / rho) for a state a from make_all_ops<OPS> (the sound speed;
    // n = g*p).  SharedOps checks that n lies in [2^-600, 2^600], so that the
    // quotient lies in [2^-700, 2^700], inside sqrt_core's range, and that
    // gamma lies in [2^-20, 2^20] (div_roe's bounds on p rest on it).
    template <class OPS>
    __device__ __forceinline__ static double sqrt_div_rho(double n, double g, const All& a, bool& ok) {
        if constexpr (!OPS::SHARED) {
            return sqrt(n / a.rho);
        } else {
            ok = ok & positive_within(n, 0x1p-600, 0x1p600) & positive_within(g, 0x1p-20, 0x1p20);
            return sqrt_core(div_shared(n, a.rho, a.rinv));
        }
    }
    // (n1, n2) / (wl + wr) with wl, wr the square roots of the two sides'
    // densities: n1 = uL wl + uR wr, n2 = wl pL + wr pR.  Under SharedOps the
    // densities passed make_all_ops' check, so wl + wr is in [2^-50, 2^51],
    // and the interface's other checks bound the numerators: |u| <= 2^391
    // (make_all_ops) and, with gamma in [2^-20, 2^20], p in [2^-620, 2^620]
    // on both sides (sqrt_div_rho), so |n1| <= 2^442 and n2, a sum of two
    // positive terms, lies in [2^-671, 2^671].  Only n1 can be small: it is checked, by
    // binary exponent, to be zero or at least 2^-800 in magnitude (it is
    // finite when the other checks pass).
    template <class OPS>
    __device__ __forceinline__ static void div_roe(double n1, double n2, double wl, double wr, double& q1,
                                                   double& q2, bool& ok) {
        if constexpr (!OPS::SHARED) {
            q1 = n1 / (wl + wr);
            q2 = n2 / (wl + wr);
        } else {
            ok = ok & (frexp_exp(n1) >= -799);
            const double den = wl + wr;
            const double r = div_recip(den);
            q1 = div_shared(n1, den, r);
            q2 = div_shared(n2, den, r);
        }
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
    // obeys() with computeExtra's quotients through one reciprocal of rho
    // (make_all_ops<SharedOps>) where the state lies in that range, which
    // gives the same extra variables and so the same answer; any other
    // state takes obeys() itself.
    __device__ __forceinline__ static bool obeys_fast(const double* q, double gamma) {
        bool in_range = true;
        const All a = make_all_ops<SharedOps>(q, gamma, in_range);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < NV; ++i) ok = ok & !__builtin_isnan(q[i]) & !__builtin_isinf(q[i]);
        ok = ok & !__builtin_isnan(a.p) & !__builtin_isinf(a.p);
#pragma unroll
        for (int c = 0; c < NSD_; ++c) ok = ok & !__builtin_isnan(a.u[c]) & !__builtin_isinf(a.u[c]);
        ok = ok & (a.rho > 0) & (a.p > 0);
        if (!in_range) ok = obeys(q, gamma);
        return ok;
    }
};

// ---------------------------------------------------------------------------
// Numerical fluxes F(left, right) along DIR; each returns the interface wave
// speed used for the CFL condition.

// burgers::Godunov (alsfvm/include/alsfvm/numflux/burgers/Godunov.hpp:40-60):
// F = max(f(max(uL,0)), f(min(uR,0))), wave speed max(|uL|, |uR|).
struct GodunovFlux {
    typedef GodunovFlux Plain;
    static constexpr bool FAST = false;
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const BurgersEq::All& L, const BurgersEq::All& R,
                                                  double, double* F) {
        const double fl = BurgersEq::point_flux(fmax(L.u, 0.0));
        const double fr = BurgersEq::point_flux(fmin(R.u, 0.0));
        F[0] = fmax(fl, fr);
        return fmax(fabs(L.u), fabs(R.u));
    }
    template <class OPS, int DIR, class EQ>
    __device__ __forceinline__ static double flux_ops(const BurgersEq::All& L, const BurgersEq::All& R, double g,
                                                      double* F, bool&) {
        return flux<DIR, EQ>(L, R, g, F);
    }
};

// euler::HLL (alsfvm/include/alsfvm/numflux/euler/HLL.hpp).  The reference's
// if/else chain is evaluated branch-free: every candidate is computed and the
// chain's answer is selected, which gives the same value as the branch taken.
template <bool FAST_>
struct HLLFluxT {
    typedef HLLFluxT<false> Plain;
    static constexpr bool FAST = FAST_;  // the stage kernel runs flux_ops<SharedOps> first
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const typename EQ::All& L, const typename EQ::All& R,
                                                  double g, double* F) {
        bool ok = true;
        return flux_ops<PlainOps, DIR, EQ>(L, R, g, F, ok);
    }
    template <class OPS, int DIR, class EQ>
    __device__ __forceinline__ static double flux_ops(const typename EQ::All& L, const typename EQ::All& R,
                                                      double g, double* F, bool& ok) {
        constexpr int NSD = EQ::NSD;
        constexpr int NV = EQ::NV;
        const double wl = OPS::sqrt_in_range(L.rho);
        const double wr = OPS::sqrt_in_range(R.rho);
        const double rho = (L.rho + R.rho) / 2;
        double ud, p;
        EQ::template div_roe<OPS>(L.u[DIR] * wl + R.u[DIR] * wr, L.p * wl + R.p * wr, wl, wr, ud, p, ok);
        const double cs = OPS::sqrt_(g * p / rho, ok);
        const double cfl = EQ::template sqrt_div_rho<OPS>(g * L.p, g, L, ok);
        const double cfr = EQ::template sqrt_div_rho<OPS>(g * R.p, g, R, ok);
        const double sl = fmin(L.u[DIR] - cfl, ud - cs);
        const double sr = fmax(R.u[DIR] + cfr, ud + cs);
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
template <bool FAST_>
struct HLL3FluxT {
    typedef HLL3FluxT<false> Plain;
    static constexpr bool FAST = FAST_;  // the stage kernel runs flux_ops<SharedOps> first
    // The star state of side k (k = L when s_m >= 0, else R): its speed sk,
    // the star conserved state `mid`, the side's conserved state `uc` and
    // its point flux `fk`, formed from side k's values by point_flux's
    // operations (the same bits as selecting between both point fluxes).
    // The transverse momentum fluxes in the star velocities are m[c] * u[DIR].
    template <int DIR, class EQ>
    __device__ __forceinline__ static void star_side(const typename EQ::All& L, const typename EQ::All& R, double sl,
                                                     double sr, double& sk, double* mid, double* uc, double* fk) {
        constexpr int NSD = EQ::NSD;
        const double pl = L.p, pr = R.p;
        const double udl = L.u[DIR] - sl;
        const double udr = R.u[DIR] - sr;
        const double aa = udr * R.rho - udl * L.rho;
        const double sm = (R.m[DIR] * udr - L.m[DIR] * udl + pr - pl) / aa;

        const bool left = sm >= 0;
        const double rk = left ? L.rho : R.rho;
        const double udk = left ? udl : udr;
        sk = left ? sl : sr;
        const double pk = left ? pl : pr;
        const double uk = left ? L.u[DIR] : R.u[DIR];
        const double ek = left ? L.E : R.E;

        mid[0] = rk * udk / (sm - sk);
#pragma unroll
        for (int c = 0; c < NSD; ++c) {
            // us = (fR.m - fL.m - sR*mR + sL*mL)/aa with us[DIR] = sm
            const double us = (c == DIR) ? sm
                                         : (R.m[c] * R.u[DIR] - L.m[c] * L.u[DIR] - R.m[c] * sr + L.m[c] * sl) / aa;
            mid[1 + c] = us * mid[0];
            uc[1 + c] = left ? L.m[c] : R.m[c];
        }
        const double pm = pk + rk * (uk - sm) * udk;
        mid[NSD + 1] = (udk * ek + pk * uk - pm * sm) / (sm - sk);
        uc[0] = rk;
        uc[NSD + 1] = ek;
        fk[0] = uc[1 + DIR];
#pragma unroll
        for (int c = 0; c < NSD; ++c) fk[1 + c] = uc[1 + c] * uk;
        fk[1 + DIR] = fk[1 + DIR] + pk;
        fk[NSD + 1] = (ek + pk) * uk;
    }
    template <int DIR, class EQ>
    __device__ __forceinline__ static double flux(const typename EQ::All& L, const typename EQ::All& R,
                                                  double g, double* F) {
        bool ok = true;
        return flux_ops<PlainOps, DIR, EQ>(L, R, g, F, ok);
    }
    template <class OPS, int DIR, class EQ>
    __device__ __forceinline__ static double flux_ops(const typename EQ::All& L, const typename EQ::All& R,
                                                      double g, double* F, bool& ok) {
        constexpr int NSD = EQ::NSD;
        constexpr int NV = EQ::NV;
        // computeHLLSpeeds
        const double wl = OPS::sqrt_in_range(L.rho);
        const double wr = OPS::sqrt_in_range(R.rho);
        const double rho = (L.rho + R.rho) / 2;
        double ud, p;
        EQ::template div_roe<OPS>(L.u[DIR] * wl + R.u[DIR] * wr, wl * L.p + wr * R.p, wl, wr, ud, p, ok);
        const double cfl = EQ::template sqrt_div_rho<OPS>(g * L.p, g, L, ok);
        const double cfr = EQ::template sqrt_div_rho<OPS>(g * R.p, g, R, ok);
        const double correct = 0.5 * fmax(0.0, L.u[DIR] - R.u[DIR]);
        const double cs = OPS::sqrt_(g * p / rho, ok);
        const double sl = fmin(L.u[DIR] + correct - cfl, ud - cs);
        const double sr = fmax(R.u[DIR] - correct + cfr, ud + cs);

        // subsonic interfaces (sL < 0 < sR) take the star branch.  Under
        // SharedOps a wavefront that is subsonic throughout skips the other
        // three branches and never forms both point fluxes (fewer values
        // live across the star formulas); otherwise both point fluxes are
        // formed first, as the general branch selects between them.
        const bool star = !(sl == 0) && !(sl > 0) && !(sr < 0);
        double sk, mid[NV], uc[NV], fk[NV];
        star_side<DIR, EQ>(L, R, sl, sr, sk, mid, uc, fk);
        if (__all(star)) {
#pragma unroll
            for (int i = 0; i < NV; ++i) F[i] = fk[i] + sk * (mid[i] - uc[i]);
        } else {
            double fl[NV], fr[NV];
            EQ::template point_flux<DIR>(L, fl);
            EQ::template point_flux<DIR>(R, fr);
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
// Under SharedOps the two weights share the reciprocal of a0 + a1, checked
// to need a0, a1 >= 2^-400: both are at most 2/(3*eps^2) < 2^53, so the
// numerators, the denominator and the quotients (within [2^-454, 1]) are in
// the range where div_shared equals the division.
template <class OPS>
__device__ __forceinline__ void weno2_weights(double a0, double a1, double& w0, double& w1, bool& ok) {
    if constexpr (OPS::SHARED) {
        ok = ok & (a0 >= 0x1p-400) & (a1 >= 0x1p-400);
        const double den = a0 + a1;
        const double r = div_recip(den);
        w0 = div_shared(a0, den, r);
        w1 = div_shared(a1, den, r);
    } else {
        w0 = a0 / (a0 + a1);
        w1 = a1 / (a0 + a1);
    }
}

template <int NV, bool RIGHT, class OPS>
__device__ __forceinline__ void weno2_face(const double* s, int vs, int i, int sd, double* out, bool& ok) {
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
        double w0r, w1r;
        weno2_weights<OPS>(a0, a1, w0r, w1r, ok);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* u = s + v * vs;
            out[v] = 0.5 * (w0r * u[i + sd] + (3 * w1r + w0r) * u[i] - w1r * u[i - sd]);
        }
    } else {
        const double a0 = 1 / d0;
        const double a1 = 2 / d1;
        double w0l, w1l;
        weno2_weights<OPS>(a0, a1, w0l, w1l, ok);
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

template <int RECON, int NV, bool RIGHT, class OPS>
__device__ __forceinline__ void recon_face_ops(const double* s, int vs, int i, int sd, double* out, bool& ok) {
    if (RECON == RECON_NONE) {
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = s[v * vs + i];
    } else if (RECON == RECON_WENO2) {
        weno2_face<NV, RIGHT, OPS>(s, vs, i, sd, out, ok);
    } else if (RECON == RECON_ENO2) {
        eno_face<2, NV, RIGHT>(s, vs, i, sd, out);
    } else {
        eno_face<3, NV, RIGHT>(s, vs, i, sd, out);
    }
}

template <int RECON, int NV, bool RIGHT>
__device__ __forceinline__ void recon_face(const double* s, int vs, int i, int sd, double* out) {
    bool ok = true;
    recon_face_ops<RECON, NV, RIGHT, PlainOps>(s, vs, i, sd, out, ok);
}

}  // namespace avh
