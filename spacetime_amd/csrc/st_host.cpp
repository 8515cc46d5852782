// Host-side helpers of libspacetime_hip (FP64, no device work): camera
// parameter blocks with the Kerr–Schild tetrad, the blackbody LUT, and the
// seeded synthetic sky and gas grids used as inputs.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "spacetime_hip.h"

namespace {

using V3 = std::array<double, 3>;
using V4 = std::array<double, 4>;

V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 normalized(const V3& a) {
    const double n = std::sqrt(dot3(a, a));
    return {a[0] / n, a[1] / n, a[2] / n};
}

// Kerr–Schild metric data at a point: g = eta + f l (x) l with covariant
// l = (1, lvec).
struct KSPoint {
    double f;
    V3 l;
};

KSPoint ks_point(const V3& p, double M, double a) {
    KSPoint k;
    if (a == 0.0) {
        const double r = std::sqrt(dot3(p, p));
        k.f = 2.0 * M / r;
        k.l = {p[0] / r, p[1] / r, p[2] / r};
        return k;
    }
    const double a2 = a * a;
    const double w = dot3(p, p) - a2;
    const double r2 = 0.5 * (w + std::sqrt(w * w + 4.0 * a2 * p[2] * p[2]));
    const double r = std::sqrt(std::max(r2, 1e-12));
    const double den = r2 + a2;
    k.l = {(r * p[0] + a * p[1]) / den, (r * p[1] - a * p[0]) / den, p[2] / r};
    k.f = 2.0 * M * r2 * r / (r2 * r2 + a2 * p[2] * p[2]);
    return k;
}

double gdot(const KSPoint& k, const V4& A, const V4& B) {
    const double lA = A[0] + k.l[0] * A[1] + k.l[1] * A[2] + k.l[2] * A[3];
    const double lB = B[0] + k.l[0] * B[1] + k.l[1] * B[2] + k.l[2] * B[3];
    return -A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + A[3] * B[3] + k.f * lA * lB;
}

V4 scale4(const V4& a, double s) { return {a[0] * s, a[1] * s, a[2] * s, a[3] * s}; }
V4 add4(const V4& a, const V4& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]}; }

// Observer 4-velocity, ks_camera_tetrad (src/preview.jl:99-163): static for
// r >= 2.5M and a radial faller dropped from 2.5M inside (a = 0); static,
// blended to the ZAMO on f in [0.72, 0.88] (a != 0).
V4 observer(const V3& p, double M, double a, const KSPoint& k) {
    if (a == 0.0) {
        const double r = std::sqrt(dot3(p, p));
        const double E = std::sqrt(std::max(1.0 - 2.0 / 2.5, 1.0 - k.f));
        const double vr = -std::sqrt(std::max(E * E - (1.0 - k.f), 0.0));
        const double w = (1.0 - E * (E - vr)) / (E - vr);
        const double lu = E + w;
        const double c = (w - k.f * lu) / r;
        return {E + k.f * lu, c * p[0], c * p[1], c * p[2]};
    }
    const double a2 = a * a;
    const double w = dot3(p, p) - a2;
    const double r2 = 0.5 * (w + std::sqrt(w * w + 4.0 * a2 * p[2] * p[2]));
    const double r = std::sqrt(std::max(r2, 1e-12));
    const double Sg = r2 * r2 + a2 * p[2] * p[2];
    const double Dl = std::max(r2 - 2.0 * M * r + a2, 1e-6);
    const double Ap = 2.0 * M * r / Dl;   // dA/dr with t_BL = t_KS - A(r)
    const V3 grad_r = {p[0] * r2 * r / Sg, p[1] * r2 * r / Sg, p[2] * r * (r2 + a2) / Sg};
    const double luc = 1.0 + Ap * dot3(k.l, grad_r);
    V4 uz = {1.0 + k.f * luc, Ap * grad_r[0] - k.f * luc * k.l[0], Ap * grad_r[1] - k.f * luc * k.l[1],
             Ap * grad_r[2] - k.f * luc * k.l[2]};
    V4 u_zamo = scale4(uz, 1.0 / std::sqrt(-gdot(k, uz, uz)));
    if (k.f >= 0.88) return u_zamo;
    const V4 u_stat = {1.0 / std::sqrt(1.0 - k.f), 0.0, 0.0, 0.0};
    double wb = std::min(std::max((k.f - 0.72) / 0.16, 0.0), 1.0);
    wb = wb * wb * (3.0 - 2.0 * wb);
    if (wb == 0.0) return u_stat;
    const V4 um = add4(scale4(u_stat, 1.0 - wb), scale4(u_zamo, wb));
    return scale4(um, 1.0 / std::sqrt(-gdot(k, um, um)));
}

void tetrad(const V3& pos, const V3& fwd, const V3& right, const V3& up, double M, double a, const V3& beta,
            V4& u, V4& Ef, V4& Er, V4& Eu) {
    const KSPoint k = ks_point(pos, M, a);
    u = observer(pos, M, a, k);
    // Gram–Schmidt under g with forward first (src/preview.jl:167-178)
    auto project_out = [&](V4 e, const V4& b, double sign) { return add4(e, scale4(b, -sign * gdot(k, e, b))); };
    Ef = {0.0, fwd[0], fwd[1], fwd[2]};
    Ef = project_out(Ef, u, -1.0);
    Ef = scale4(Ef, 1.0 / std::sqrt(gdot(k, Ef, Ef)));
    Er = {0.0, right[0], right[1], right[2]};
    Er = project_out(Er, u, -1.0);
    Er = project_out(Er, Ef, 1.0);
    Er = scale4(Er, 1.0 / std::sqrt(gdot(k, Er, Er)));
    Eu = {0.0, up[0], up[1], up[2]};
    Eu = project_out(Eu, u, -1.0);
    const double cf = gdot(k, Eu, Ef), cr = gdot(k, Eu, Er);
    Eu = add4(Eu, add4(scale4(Ef, -cf), scale4(Er, -cr)));
    Eu = scale4(Eu, 1.0 / std::sqrt(gdot(k, Eu, Eu)));
    // Lorentz boost by the camera-frame 3-velocity (src/preview.jl:184-196)
    const double b2raw = dot3(beta, beta);
    if (b2raw > 1e-12) {
        const double b2 = std::min(b2raw, 0.9801);
        const double sc = std::sqrt(b2 / b2raw);
        const V3 b = {beta[0] * sc, beta[1] * sc, beta[2] * sc};
        const double g = 1.0 / std::sqrt(1.0 - b2);
        const V4 bE = add4(add4(scale4(Ef, b[0]), scale4(Er, b[1])), scale4(Eu, b[2]));
        const V4 t = add4(scale4(bE, (g - 1.0) / b2), scale4(u, g));
        const V4 ub = scale4(add4(u, bE), g);
        Ef = add4(Ef, scale4(t, b[0]));
        Er = add4(Er, scale4(t, b[1]));
        Eu = add4(Eu, scale4(t, b[2]));
        u = ub;
    }
}

void pack(const V3& pos, double fov, const V4& u, const V4& Ef, const V4& Er, const V4& Eu, double fisheye_deg,
          double focus_dist, int equirect, double out[49]) {
    std::fill(out, out + 49, 0.0);
    out[0] = pos[0]; out[1] = pos[1]; out[2] = pos[2];
    out[3] = fov;
    for (int i = 0; i < 4; ++i) {
        out[4 + i] = Ef[i]; out[8 + i] = Er[i]; out[12 + i] = Eu[i]; out[16 + i] = u[i];
    }
    out[20] = equirect ? 2.0 : (fisheye_deg > 0.0 ? 1.0 : 0.0);
    out[21] = std::max(fisheye_deg, 0.0) * (M_PI / 180.0);
    out[24] = focus_dist;
}

// SplitMix64 / xoshiro256** (Blackman & Vigna).
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {  // Box–Muller
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Colour of a blackbody, CIE-Gaussian XYZ -> linear sRGB, max-normalised
// (src/blackbody.jl:15-35).
void bb_rgb(double T, double out[3]) {
    struct G { double w, mu, s1, s2; };
    static const G xg[3] = {{1.056, 599.8, 37.9, 31.0}, {0.362, 442.0, 16.0, 26.7}, {-0.065, 501.1, 20.4, 26.2}};
    static const G yg[2] = {{0.821, 568.8, 46.9, 40.5}, {0.286, 530.9, 16.3, 31.1}};
    static const G zg[2] = {{1.217, 437.0, 11.8, 36.0}, {0.681, 459.0, 26.0, 13.8}};
    auto lobe = [](const G& g, double lam) {
        const double s = lam < g.mu ? g.s1 : g.s2;
        const double t = (lam - g.mu) / s;
        return g.w * std::exp(-0.5 * t * t);
    };
    double X = 0.0, Y = 0.0, Z = 0.0;
    for (int i = 0; i <= 80; ++i) {
        const double lam = 380.0 + 5.0 * i;
        const double lm = lam * 1e-9;
        const double B = 1.0 / (lm * lm * lm * lm * lm * (std::exp(0.014388 / (lm * T)) - 1.0));
        X += B * (lobe(xg[0], lam) + lobe(xg[1], lam) + lobe(xg[2], lam));
        Y += B * (lobe(yg[0], lam) + lobe(yg[1], lam));
        Z += B * (lobe(zg[0], lam) + lobe(zg[1], lam));
    }
    const double r = std::max(3.2406 * X - 1.5372 * Y - 0.4986 * Z, 0.0);
    const double g = std::max(-0.9689 * X + 1.8758 * Y + 0.0415 * Z, 0.0);
    const double b = std::max(0.0557 * X - 0.2040 * Y + 1.0570 * Z, 0.0);
    const double m = std::max({r, g, b, 1e-10});
    out[0] = r / m; out[1] = g / m; out[2] = b / m;
}

thread_local std::string g_host_err;

}  // namespace

namespace st {
int fail(int code, const std::string& msg);  // records the message st_last_error() returns
}  // namespace st

extern "C" {

int st_camera_params_basis(const double pos[3], const double fwd[3], const double right[3], const double up[3],
                           double fov, const double beta[3], double M, double a, double out[49]) {
    if (!pos || !fwd || !right || !up || !out) return st::fail(ST_ERR_INVALID_ARG, "st_camera_params_basis: null pointer");
    if (!(M > 0.0) || std::fabs(a) > M) return st::fail(ST_ERR_INVALID_ARG, "st_camera_params_basis: need M > 0 and |a| <= M");
    const V3 p = {pos[0], pos[1], pos[2]};
    if (dot3(p, p) <= 0.0) return st::fail(ST_ERR_INVALID_ARG, "st_camera_params_basis: camera at the origin");
    V4 u, Ef, Er, Eu;
    const V3 b = beta ? V3{beta[0], beta[1], beta[2]} : V3{0.0, 0.0, 0.0};
    tetrad(p, {fwd[0], fwd[1], fwd[2]}, {right[0], right[1], right[2]}, {up[0], up[1], up[2]}, M, a, b, u, Ef, Er, Eu);
    pack(p, fov, u, Ef, Er, Eu, 0.0, 1.0, 0, out);
    return ST_OK;
}

int st_camera_params(const double pos[3], const double target[3], const double up[3], double fov,
                     const double velocity[3], double M, double a, double fisheye_deg, double focus_dist,
                     int equirect, double out[49]) {
    if (!pos || !target || !up || !out) return st::fail(ST_ERR_INVALID_ARG, "st_camera_params: null pointer");
    if (!(M > 0.0) || std::fabs(a) > M) return st::fail(ST_ERR_INVALID_ARG, "st_camera_params: need M > 0 and |a| <= M");
    const V3 p = {pos[0], pos[1], pos[2]};
    const V3 fwd = normalized({target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]});
    const V3 right = normalized(cross(fwd, {up[0], up[1], up[2]}));
    const V3 upl = cross(right, fwd);
    V3 beta = {0.0, 0.0, 0.0};
    if (velocity) {
        const V3 v = {velocity[0], velocity[1], velocity[2]};
        beta = {dot3(v, fwd), dot3(v, right), dot3(v, upl)};
    }
    V4 u, Ef, Er, Eu;
    tetrad(p, fwd, right, upl, M, a, beta, u, Ef, Er, Eu);
    pack(p, fov, u, Ef, Er, Eu, fisheye_deg, focus_dist, equirect, out);
    return ST_OK;
}

int st_blackbody_lut(double wb, double tmin, double tmax, int n, float* out) {
    if (!out || n < 1 || !(tmax > tmin) || !(wb > 0.0))
        return st::fail(ST_ERR_INVALID_ARG, "st_blackbody_lut: need out, n >= 1, tmax > tmin, wb > 0");
    double ref[3];
    bb_rgb(wb, ref);
    for (int i = 0; i < n; ++i) {
        const double T = n > 1 ? tmin + (tmax - tmin) * i / (double)(n - 1) : tmin;
        double c[3];
        bb_rgb(std::max(T, 1.0), c);
        for (int ch = 0; ch < 3; ++ch) out[3 * i + ch] = (float)(c[ch] / ref[ch]);
    }
    return ST_OK;
}

int st_make_sky(uint64_t seed, int w, int h, int nstars, double sigma, float* out) {
    if (!out || w < 2 || h < 2 || nstars < 0) return st::fail(ST_ERR_INVALID_ARG, "st_make_sky: need out, w, h >= 2, nstars >= 0");
    // smooth gradient in theta: warmer near the +z pole, bluer near -z
    for (int y = 0; y < h; ++y) {
        const double th = M_PI * y / (double)(h - 1);
        const double c = std::cos(th);
        const float r = (float)(0.020 + 0.012 * (0.5 + 0.5 * c));
        const float g = (float)(0.022 + 0.008 * (0.5 + 0.5 * c));
        const float b = (float)(0.026 + 0.014 * (0.5 - 0.5 * c));
        for (int x = 0; x < w; ++x) {
            float* p = out + ((size_t)y * w + x) * 3;
            p[0] = r; p[1] = g; p[2] = b;
        }
    }
    // stars: uniform on the sphere, log-normal brightness, bilinear splat in
    // the sampler's own texel mapping
    Rng rng(seed);
    for (int k = 0; k < nstars; ++k) {
        const double z = 2.0 * rng.uniform() - 1.0;
        const double ph = 2.0 * M_PI * rng.uniform();
        const double bright = std::exp(sigma * rng.normal());
        const double tint = rng.uniform();
        const double th = std::acos(z);
        const double xf = ph / (2.0 * M_PI) * (w - 1), yf = th / M_PI * (h - 1);
        const int x0 = std::min((int)std::floor(xf), w - 2), y0 = std::min((int)std::floor(yf), h - 2);
        const double fx = xf - x0, fy = yf - y0;
        const double col[3] = {bright * (0.85 + 0.3 * tint), bright, bright * (1.15 - 0.3 * tint)};
        const double wts[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
        const int xs[4] = {x0, x0 + 1, x0, x0 + 1}, ys[4] = {y0, y0, y0 + 1, y0 + 1};
        for (int t = 0; t < 4; ++t) {
            float* p = out + ((size_t)ys[t] * w + xs[t]) * 3;
            for (int ch = 0; ch < 3; ++ch) p[ch] += (float)(0.25 * col[ch] * wts[t]);
        }
    }
    return ST_OK;
}

int st_make_volume(uint64_t seed, int nr, int nphi, int nz, double sigma, float* out) {
    if (!out || nr < 2 || nphi < 1 || nz < 2) return st::fail(ST_ERR_INVALID_ARG, "st_make_volume: need out, nr >= 2, nphi >= 1, nz >= 2");
    // Gaussian random field as a sum of random plane waves in (log s, phi, z)
    // index space, periodic in phi; exponentiated to a unit-mean log-normal.
    Rng rng(seed);
    const int nwave = 48;
    struct Wv { double kr, kp, kz, ph, amp; };
    std::vector<Wv> wv(nwave);
    double norm = 0.0;
    for (int i = 0; i < nwave; ++i) {
        const double scale = std::pow(2.0, rng.uniform() * 5.0);   // 1..32 cycles
        wv[i].kr = 2.0 * M_PI * scale * (rng.uniform() - 0.5) / nr * 4.0;
        wv[i].kp = 2.0 * M_PI * std::floor(scale * rng.uniform() * 4.0 + 1.0) / nphi;
        wv[i].kz = 2.0 * M_PI * scale * (rng.uniform() - 0.5) / nz;
        wv[i].ph = 2.0 * M_PI * rng.uniform();
        wv[i].amp = 1.0 / scale;   // red spectrum
        norm += 0.5 * wv[i].amp * wv[i].amp;
    }
    const double inv = 1.0 / std::sqrt(norm);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < nphi; ++j)
            for (int i = 0; i < nr; ++i) {
                double g = 0.0;
                for (const Wv& w : wv) g += w.amp * std::cos(w.kr * i + w.kp * j + w.kz * k + w.ph);
                g *= inv;
                out[((size_t)k * nphi + j) * nr + i] = (float)std::exp(sigma * g - 0.5 * sigma * sigma);
            }
    return ST_OK;
}

}  // extern "C"
