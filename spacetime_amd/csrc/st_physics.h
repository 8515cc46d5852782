// Device-side physics of the Kerr–Schild null-geodesic tracer, templated on
// the arithmetic type (float for the bulk path, double for promoted rays).
//
// Reference algorithm (SpaceTime.jl, src/metal.jl): ray generation :855-949,
// KS photon init :951-995, Bardeen launch test :1008-1094, step loop
// :1201-1260, RK4 :1382-1401, NaN bail :1414-1421, gas march :1440-1513,
// thin disc :1515-1571, fate + sky :1629-1704, helpers :314-724.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

// ---------------------------------------------------------------------------
// Arithmetic traits.  FP32 uses the hardware approximations directly
// (v_rsq_f32 / v_rcp_f32 / v_sqrt_f32, about 1 ulp); every argument they see
// on the hot path is a normal number well inside float range.  FP64 refines
// the hardware estimates by one Newton step (rsq, rcp, sqrt) and uses the
// correctly rounded library forms elsewhere (psqrt, exp, ...).
// ---------------------------------------------------------------------------
template <typename T> struct Ar;
template <> struct Ar<float> {
    static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float psqrt(float x) { return sqrtf(x); }
    static __device__ __forceinline__ float exp(float x) { return expf(x); }
    static __device__ __forceinline__ float log(float x) { return logf(x); }
    static __device__ __forceinline__ float acos(float x) { return acosf(x); }
    static __device__ __forceinline__ float atan2(float y, float x) { return atan2f(y, x); }
    static __device__ __forceinline__ float sin(float x) { return sinf(x); }
    static __device__ __forceinline__ float cos(float x) { return cosf(x); }
    static __device__ __forceinline__ float floor(float x) { return floorf(x); }
    static __device__ __forceinline__ float fabs(float x) { return fabsf(x); }
    static __device__ __forceinline__ float fmod(float x, float y) { return fmodf(x, y); }
};
template <> struct Ar<double> {
    // The hot-path forms start from the hardware estimates (v_rsq_f64,
    // v_rcp_f64: about 2^-29 relative) and take one Newton step, which
    // squares the error to FP64 round-off; the IEEE-exact sequences cost
    // several times as many instructions.
    static __device__ __forceinline__ double rsq(double x) {
        const double y = __builtin_amdgcn_rsq(x);
        const double e = __builtin_fma(-x * y, y, 1.0);
        return __builtin_fma(0.5 * y, e, y);
    }
    static __device__ __forceinline__ double rcp(double x) {
        const double r = __builtin_amdgcn_rcp(x);
        return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    }
    static __device__ __forceinline__ double sqrt(double x) {
        const double y = rsq(x);
        const double s = x * y;
        const double r = __builtin_fma(__builtin_fma(-s, s, x), 0.5 * y, s);
        return x > 0.0 ? r : x;   // rsq(0) = inf
    }
    static __device__ __forceinline__ double psqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double acos(double x) { return ::acos(x); }
    static __device__ __forceinline__ double atan2(double y, double x) { return ::atan2(y, x); }
    static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
    static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
    static __device__ __forceinline__ double floor(double x) { return ::floor(x); }
    static __device__ __forceinline__ double fabs(double x) { return ::fabs(x); }
    static __device__ __forceinline__ double fmod(double x, double y) { return ::fmod(x, y); }
};

// Transcendentals of the FP32 gas sampler, evaluated per density sample:
// the hardware exp2/log2 (about 1 ulp on the reduced argument) and an atan2
// built on an odd polynomial for atan on [0, 1] (coefficients fitted for
// this file by weighted least squares on Chebyshev nodes; max error
// 1.5e-7 rad in FP32).  The FP64 forms are the library functions.
template <typename T> struct GasMath;
template <> struct GasMath<float> {
    static __device__ __forceinline__ float exp(float x) {
        return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
    }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
    static __device__ __forceinline__ float atan2(float y, float x) {
        const float ax = fabsf(x), ay = fabsf(y);
        const float hi = fmaxf(ax, ay), lo = fminf(ax, ay);
        const float t = hi > 0.0f ? lo * __builtin_amdgcn_rcpf(hi) : 0.0f;
        const float u = t * t;
        float p = -0.00405456f;
        p = __builtin_fmaf(p, u, 0.02186294f);
        p = __builtin_fmaf(p, u, -0.05591232f);
        p = __builtin_fmaf(p, u, 0.09642198f);
        p = __builtin_fmaf(p, u, -0.13908631f);
        p = __builtin_fmaf(p, u, 0.19946566f);
        p = __builtin_fmaf(p, u, -0.33329861f);
        p = __builtin_fmaf(p, u, 0.99999934f);
        float r = p * t;
        if (ay > ax) r = 1.5707963267948966f - r;
        if (x < 0.0f) r = 3.141592653589793f - r;
        return y < 0.0f ? -r : r;
    }
};
template <> struct GasMath<double> {
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double atan2(double y, double x) { return ::atan2(y, x); }
};

// Division and square root of the per-sample gas shading: in FP32 the
// hardware reciprocal and square root (about 1 ulp) in place of the
// correctly rounded sequences; in FP64 the exact operations.
template <typename T> struct Fast;
template <> struct Fast<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};
template <> struct Fast<double> {
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
};

template <typename T> __device__ __forceinline__ T mn(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T mx(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T clampv(T x, T lo, T hi) { return mn(mx(x, lo), hi); }
// a*b + c, fused.
__device__ __forceinline__ float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fm(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Julia `mod(x, y)` for y > 0.  fmod(x, y) is x itself when |x| < y, so
// that common case skips the library call.
template <typename T> __device__ __forceinline__ T jmod(T x, T y) {
    T m = Ar<T>::fabs(x) < y ? x : Ar<T>::fmod(x, y);
    if (m < T(0)) m += y;
    if (m >= y) m -= y;
    return m;
}

// _sim_hash (src/disc_sim.jl:83-91): integer hash -> [0, 1) in 2^-24 steps.
__device__ __forceinline__ float sim_hash(int x, int y, int z) {
    uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^ (uint32_t)z * 0xcb1ab31fu;
    h ^= h >> 13;
    h *= 0x9e3779b9u;
    h ^= h >> 16;
    return (float)(h & 0x00ffffffu) * 5.9604644775390625e-8f;
}

struct State6 {
    template <typename T> struct S { T x, y, z, px, py, pz; };
};
template <typename T> using S6 = State6::S<T>;

// ---------------------------------------------------------------------------
// Geodesic RHS.  Schwarzschild: ks_rhs_mtl (src/metal.jl:324-340).  Kerr:
// kerr_rhs_mtl (src/metal.jl:466-508), the same algebra regrouped for the
// fewest VALU operations: one reciprocal serves both 1/Sigma and 1/(r^2+a^2),
// r and 1/r come from one reciprocal square root, df/dr and df/dz are
// written as multiples of f, and f l factors out of all of dp.  `rk2`
// returns the KS radius squared at the evaluation point (the kill test and
// step rule reuse it).
// ---------------------------------------------------------------------------
template <typename T, bool KERR>
__device__ __forceinline__ S6<T> rhs(const S6<T>& s, T pt, T M, T a, T& rk2) {
    using A = Ar<T>;
    S6<T> o;
    if constexpr (!KERR) {
        const T r2 = s.x * s.x + s.y * s.y + s.z * s.z;
        const T ir = A::rsq(r2);
        const T f = T(2) * M * ir;
        const T kap = (s.x * s.px + s.y * s.py + s.z * s.pz) * ir;
        const T ell = kap - pt;
        const T c1 = f * ell * ir;
        const T c2 = c1 * (T(0.5) * ell + kap) * ir;
        o.x = s.px - c1 * s.x;
        o.y = s.py - c1 * s.y;
        o.z = s.pz - c1 * s.z;
        o.px = c1 * s.px - c2 * s.x;
        o.py = c1 * s.py - c2 * s.y;
        o.pz = c1 * s.pz - c2 * s.z;
        rk2 = r2;
    } else {
        // r from the quartic r^4 - (rho^2 - a^2) r^2 - a^2 z^2 = 0 as
        // r^2 = w/2 + sqrt((w/2)^2 + a^2 z^2), w = rho^2 - a^2.
        const T a2 = a * a;
        const T z2 = s.z * s.z;
        const T a2z2 = a2 * z2;
        const T hw = fm(T(0.5), fm(s.x, s.x, fm(s.y, s.y, z2)), T(-0.5) * a2);
        const T r2 = mx(hw + A::sqrt(fm(hw, hw, a2z2)), T(1e-12));
        const T ir = A::rsq(r2);
        const T r = r2 * ir;
        const T Sg = fm(r2, r2, a2z2);                // Sigma = r^4 + a^2 z^2
        const T R2A = r2 + a2;
        const T q = A::rcp(Sg * R2A);
        const T invS = R2A * q;                       // 1/Sigma
        const T iRA = Sg * q;                         // 1/(r^2 + a^2)
        // l = ((r x + a y), (r y - a x)) / (r^2 + a^2), z / r
        const T rxay = fm(r, s.x, a * s.y);
        const T ryax = fm(r, s.y, -a * s.x);
        const T Lz = s.z * ir;
        const T N = fm(s.px, rxay, s.py * ryax);      // (r^2 + a^2) (l_x p_x + l_y p_y)
        const T zpir = Lz * s.pz;
        const T ell = fm(N, iRA, zpir) - pt;
        const T g = r2 * r * invS;                    // r^3 / Sigma, dr/dx = g x, dr/dy = g y
        const T fl = (T(2) * M) * g * ell;            // f l
        // dx = p - f l L
        const T c = fl * iRA;
        o.x = fm(-c, rxay, s.px);
        o.y = fm(-c, ryax, s.py);
        o.z = fm(-fl, Lz, s.pz);
        // dp = f l [ B grad r + d(l.p)/dx|_r ] - f l l a^2 z / Sigma zhat, with
        // B = l/(2r) (3 - 4 r^4/Sigma) + d(l.p)/dr   (so f l B = l^2/2 df/dr + f l dl.p/dr)
        const T u = fm(s.x, s.px, s.y * s.py);
        const T dLdr = fm(-zpir, ir, fm(T(-2) * r * N, iRA, u) * iRA);
        const T kh = fm(T(-2), (r2 * r2) * invS, T(1.5));
        const T Bg = fm(ell * ir, kh, dLdr) * g;
        const T ex = fm(r, s.px, -a * s.py) * iRA;
        const T ey = fm(r, s.py, a * s.px) * iRA;
        o.px = fl * fm(Bg, s.x, ex);
        o.py = fl * fm(Bg, s.y, ey);
        const T zt = fm(a2, fm(Bg, ir * ir, -ell * invS), Bg);
        o.pz = fl * fm(s.z, zt, s.pz * ir);
        rk2 = r2;
    }
    return o;
}

template <typename T>
__device__ __forceinline__ S6<T> axpy(const S6<T>& s, T c, const S6<T>& k) {
    return S6<T>{s.x + c * k.x, s.y + c * k.y, s.z + c * k.z, s.px + c * k.px, s.py + c * k.py, s.pz + c * k.pz};
}

// KS radius squared from the implicit quartic (src/metal.jl:1218-1220).
template <typename T, bool KERR>
__device__ __forceinline__ T ks_r2(T x, T y, T z, T a) {
    const T r2 = x * x + y * y + z * z;
    if constexpr (!KERR) return r2;
    const T a2 = a * a;
    const T w = r2 - a2;
    return T(0.5) * (w + Ar<T>::sqrt(w * w + T(4) * a2 * z * z));
}

// H = 1/2 p g^-1 p (src/gr.jl:196-203) in the kernel's chart.
template <typename T, bool KERR>
__device__ __forceinline__ T hamiltonian(const S6<T>& s, T pt, T M, T a) {
    using A = Ar<T>;
    T f, Lx, Ly, Lz;
    if constexpr (!KERR) {
        const T r = A::psqrt(s.x * s.x + s.y * s.y + s.z * s.z);
        f = T(2) * M / r;
        Lx = s.x / r; Ly = s.y / r; Lz = s.z / r;
    } else {
        const T a2 = a * a;
        const T r2 = mx(ks_r2<T, true>(s.x, s.y, s.z, a), T(1e-30));
        const T r = A::psqrt(r2);
        f = T(2) * M * r2 * r / (r2 * r2 + a2 * s.z * s.z);
        const T R2A = r2 + a2;
        Lx = (r * s.x + a * s.y) / R2A; Ly = (r * s.y - a * s.x) / R2A; Lz = s.z / r;
    }
    const T ell = -pt + Lx * s.px + Ly * s.py + Lz * s.pz;
    return T(0.5) * (-pt * pt + s.px * s.px + s.py * s.py + s.pz * s.pz - f * ell * ell);
}

// ISCO of the prograde orbit + plunge constants (src/metal.jl:371-380).
template <typename T>
__device__ __forceinline__ void isco(T M, T a, T& r_isco, T& E_isco, T& L_isco) {
    using A = Ar<T>;
    auto cb = [](T v) { return A::exp(A::log(mx(v, T(1e-12))) / T(3)); };
    const T q = a / M;
    const T z1 = T(1) + cb(T(1) - q * q) * (cb(T(1) + q) + cb(T(1) - q));
    const T z2 = A::psqrt(T(3) * q * q + z1 * z1);
    r_isco = M * (T(3) + z2 - A::psqrt(mx((T(3) - z1) * (T(3) + z1 + T(2) * z2), T(0))));
    E_isco = A::psqrt(T(1) - T(2) * M / (T(3) * r_isco));
    L_isco = (T(2) * M / (T(3) * T(1.7320508075688772))) * (T(1) + T(2) * A::psqrt(T(3) * r_isco / M - T(2)));
}

// Emitter shift -p.u/E (src/metal.jl:396-428).  FAST: FP32 divisions and
// square roots through the hardware reciprocal / square root (the gas
// sampler, once per density sample); the thin disc keeps the exact forms.
template <typename T, bool FAST = false>
__device__ __forceinline__ T emitter_opz(T sx, T sy, T psx, T psy, T E, T M, T a, T r_isco, T E_isco, T L_isco) {
    using A = Ar<T>;
    auto dv = [](T x, T y) { return FAST ? Fast<T>::div(x, y) : x / y; };
    auto sq = [](T x) { return FAST ? Fast<T>::sqrt(x) : A::psqrt(x); };
    const T a2 = a * a;
    const T s2 = sx * sx + sy * sy;
    const T r = sq(mx(s2 - a2, T(1e-6)));
    const T sqM = sq(M), sqr = sq(r), r32 = r * sqr;
    const T Dl = r * r - T(2) * M * r + a2;
    T ut, uphi, ur = T(0);
    if (r >= r_isco) {
        const T den = r32 - T(3) * M * sqr + T(2) * a * sqM;
        ut = dv(r32 + a * sqM, sq(r32) * sq(mx(den, T(1e-2))));
        uphi = dv(sqM, r32 + a * sqM) * ut;
    } else {
        const T inv_D = dv(T(1), mx(Dl, T(1e-3)));
        const T ir = dv(T(1), r);
        ut = ((r * r + a2 + T(2) * M * a2 * ir) * E_isco - T(2) * M * a * L_isco * ir) * inv_D;
        uphi = (T(2) * M * a * E_isco * ir + (T(1) - T(2) * M * ir) * L_isco) * inv_D;
        const T t1 = E_isco * (r * r + a2) - a * L_isco;
        const T t2 = L_isco - a * E_isco;
        const T Rr = t1 * t1 - Dl * (r * r + t2 * t2);
        ur = -dv(sq(mx(Rr, T(0))), r * r);
        ut += T(2) * M * r * inv_D * ur;
        uphi += a * inv_D * ur;
    }
    const T inv_s2 = dv(T(1), mx(s2, T(1e-6)));
    const T cphi = (r * sx + a * sy) * inv_s2;
    const T sphi = (r * sy - a * sx) * inv_s2;
    const T ux = ur * cphi - sy * uphi;
    const T uy = ur * sphi + sx * uphi;
    const T Es = A::fabs(E) > T(1e-6) ? E : T(1e-6);
    return clampv(ut - dv(psx * ux + psy * uy, Es), T(0.1), T(20));
}

}  // namespace st
