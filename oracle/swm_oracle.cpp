// swm_oracle.cpp -- CPU restatement of ShallowWaters.jl's time-stepping path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle of the GPU library: only
// tests/, the smoke check and the cpu_baseline leg of bench.py load it, and only to check
// or time against it.  Nothing in the product path calls it.
//
// It restates the reference operator by operator: one loop nest per reference function,
// the same array shapes (src/preallocate.jl), the same 1-based column-major indexing and the
// same floating-point evaluation order (Julia evaluates a+b+c left to right, x^2 as x*x,
// and contracts nothing into FMAs; build with -ffp-contract=off, no fast-math).  Each
// function cites the reference lines it follows.
//
// Number formats: double, float, and binary16 / bfloat16 emulated on float with one rounding
// after each operation (float carries >= 2p+2 significand bits for both, so + - * / sqrt
// are correctly rounded; Julia's Float16 arithmetic is likewise Float32-then-round).
//
// Parity pins (tests/test_oracle.py, tests/test_setup.py): the exact-zero known answer of
// test/runtests.jl:4-9, closed forms of every stencil, rest-state tendencies equal to the
// wind, uniform zonal flow in Coriolis balance, exact mass conservation of continuity,
// x-shift equivariance away from the periodic seam (where the reference's AHβ!/AHγ! wrap
// breaks it), the PV advection schemes' exact conservation identities on non-divergent
// fluxes (enstrophy for Sadourny's, enstrophy and energy for Arakawa-Hsu's), exact tracer
// shifts and bilinear reproduction of linear fields, and IEEE
// binary16 / bfloat16 rounding.  Julia is not installed here, so the reference itself
// cannot be run; see the tests for what each pin covers.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/swm.h"

// ------------------------------------------------------------------ number formats

static inline float f32_round_odd(double x) {
    float f = (float)x;
    if (!std::isfinite(f) || (double)f == x) return f;
    uint32_t b;
    std::memcpy(&b, &f, 4);
    if (std::fabs((double)f) > std::fabs(x)) b -= 1u;
    b |= 1u;
    std::memcpy(&f, &b, 4);
    return f;
}

static inline float round_binary16(float x) {
    if (!std::isfinite(x)) return x;
    float ax = std::fabs(x);
    if (ax >= 65520.0f) return std::copysign(INFINITY, x);
    float q;
    if (ax < 6.103515625e-05f) {
        q = 5.9604644775390625e-08f;        // 2^-24, subnormal quantum
    } else {
        int e;
        std::frexp(ax, &e);
        q = std::ldexp(1.0f, e - 11);       // 10 fraction bits
    }
    float r = std::nearbyint(ax / q) * q;   // ties to even (default rounding mode)
    return std::copysign(r, x);
}

static inline float round_bfloat16(float x) {
    if (std::isnan(x)) return x;
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    b &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}

template <int KIND>  // 0: binary16, 1: bfloat16
struct Low {
    float v;
    static float rnd(float x) { return KIND == 0 ? round_binary16(x) : round_bfloat16(x); }
    Low() : v(0.0f) {}
    explicit Low(double x) : v(rnd(f32_round_odd(x))) {}
    explicit Low(float x) : v(rnd(x)) {}
    explicit Low(int x) : v(rnd((float)x)) {}
    template <int K2>
    explicit Low(Low<K2> o) : v(rnd(o.v)) {}
    static Low raw(float x) { Low r; r.v = x; return r; }
    explicit operator double() const { return (double)v; }
    explicit operator float() const { return v; }
    friend Low operator+(Low a, Low b) { return raw(rnd(a.v + b.v)); }
    friend Low operator-(Low a, Low b) { return raw(rnd(a.v - b.v)); }
    friend Low operator*(Low a, Low b) { return raw(rnd(a.v * b.v)); }
    friend Low operator/(Low a, Low b) { return raw(rnd(a.v / b.v)); }
    friend Low operator-(Low a) { return raw(-a.v); }
    Low &operator+=(Low b) { v = rnd(v + b.v); return *this; }
    friend bool operator<(Low a, Low b) { return a.v < b.v; }
    friend bool operator>(Low a, Low b) { return a.v > b.v; }
    friend bool operator==(Low a, Low b) { return a.v == b.v; }
};
typedef Low<0> f16;
typedef Low<1> bf16;

// Binary16 / bfloat16 values with float arithmetic (swm_config.t_arith32, a form of the
// library, not of the reference): a value entering the format (a conversion from Tprog or
// Float64, a T constant) is rounded to it; + - * / and sqrt are float's, rounded once to
// float and not back to the format.
// the planted mistake MUT_WIDE_ROUNDS (see the MUT_* list): Wide's arithmetic rounds to the
// format, which makes it Low's
static bool g_wide_rounds = false;
template <int KIND>
struct Wide {
    float v;
    static float rnd(float x) { return Low<KIND>::rnd(x); }
    static float op(float x) { return g_wide_rounds ? rnd(x) : x; }
    Wide() : v(0.0f) {}
    explicit Wide(double x) : v(rnd(f32_round_odd(x))) {}
    explicit Wide(float x) : v(rnd(x)) {}
    explicit Wide(int x) : v(rnd((float)x)) {}
    static Wide raw(float x) { Wide r; r.v = x; return r; }
    explicit operator double() const { return (double)v; }
    explicit operator float() const { return v; }
    friend Wide operator+(Wide a, Wide b) { return raw(op(a.v + b.v)); }
    friend Wide operator-(Wide a, Wide b) { return raw(op(a.v - b.v)); }
    friend Wide operator*(Wide a, Wide b) { return raw(op(a.v * b.v)); }
    friend Wide operator/(Wide a, Wide b) { return raw(op(a.v / b.v)); }
    friend Wide operator-(Wide a) { return raw(-a.v); }
    Wide &operator+=(Wide b) { v = op(v + b.v); return *this; }
    friend bool operator<(Wide a, Wide b) { return a.v < b.v; }
    friend bool operator>(Wide a, Wide b) { return a.v > b.v; }
    friend bool operator==(Wide a, Wide b) { return a.v == b.v; }
};
typedef Wide<0> f16w;
typedef Wide<1> bf16w;

// A float64 value carrying one directional derivative (forward-mode differentiation).  The
// model instantiated on Dual returns its outputs together with the tangent-linear map applied
// to the input tangents, operation by operation in the model's own order.  Kinks follow fixed
// conventions: sqrt'(0) = 0, floor and the integer indices derived from it are constants, a
// clipped departure coordinate is a constant.
struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    explicit Dual(double x) : v(x), d(0.0) {}
    explicit Dual(float x) : v(x), d(0.0) {}
    explicit Dual(int x) : v(x), d(0.0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    friend Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
    friend Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
    friend Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
    friend Dual operator/(Dual a, Dual b) {
        const double q = a.v / b.v;
        return Dual(q, (a.d - q * b.d) / b.v);
    }
    friend Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
    Dual &operator+=(Dual b) { v += b.v; d += b.d; return *this; }
    friend bool operator<(Dual a, Dual b) { return a.v < b.v; }
    friend bool operator>(Dual a, Dual b) { return a.v > b.v; }
    friend bool operator==(Dual a, Dual b) { return a.v == b.v; }
};

template <class T> static inline T num_sqrt(T x) { return std::sqrt(x); }
template <> inline f16 num_sqrt<f16>(f16 x) { return f16::raw(f16::rnd(std::sqrt(x.v))); }
template <> inline bf16 num_sqrt<bf16>(bf16 x) { return bf16::raw(bf16::rnd(std::sqrt(x.v))); }
template <> inline f16w num_sqrt<f16w>(f16w x) { return f16w::raw(f16w::op(std::sqrt(x.v))); }
template <> inline bf16w num_sqrt<bf16w>(bf16w x) { return bf16w::raw(bf16w::op(std::sqrt(x.v))); }
template <> inline Dual num_sqrt<Dual>(Dual x) {
    const double s = std::sqrt(x.v);
    return Dual(s, s > 0.0 ? x.d / (2.0 * s) : 0.0);
}
template <class T> static inline double to_f64(T x) { return (double)x; }
template <class T> static inline bool is_nan(T x) { return std::isnan((double)x); }

// Julia's T(x): conversion between formats with a single rounding.
template <class To, class From> static inline To cvt(From x) {
    if constexpr (std::is_same_v<To, From>) return x;
    else if constexpr (std::is_same_v<To, double>) return (double)x;
    else if constexpr (std::is_same_v<To, float>) {
        if constexpr (std::is_same_v<From, double>) return (float)x;
        else return (float)x;               // from f16/bf16: exact widening
    } else {
        if constexpr (std::is_same_v<From, double>) return To(x);
        else if constexpr (std::is_same_v<From, float>) return To(x);
        else return To((float)x);            // f16 <-> bf16
    }
}

// host storage of a format (binary16 / bfloat16 are 16-bit patterns)
static inline float half_bits_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float r;
    if (e == 0) {
        r = std::ldexp((float)m, -24);
        return s ? -r : r;
    }
    uint32_t b;
    if (e == 31) b = s | 0x7f800000u | (m << 13);
    else b = s | ((e + 112) << 23) | (m << 13);
    std::memcpy(&r, &b, 4);
    return r;
}
static inline uint16_t float_to_half_bits(float x) {  // x already binary16-representable
    uint32_t b;
    std::memcpy(&b, &x, 4);
    uint16_t s = (uint16_t)((b >> 16) & 0x8000u);
    float ax = std::fabs(x);
    if (std::isnan(x)) return s | 0x7e00u;
    if (std::isinf(x)) return s | 0x7c00u;
    if (ax < 6.103515625e-05f) return s | (uint16_t)std::lrint(ax * 16777216.0f);
    uint32_t e = ((b >> 23) & 0xff) - 112, m = (b >> 13) & 0x3ff;
    return s | (uint16_t)(e << 10) | (uint16_t)m;
}
template <class T> static inline T load_elem(const void *p, size_t k) {
    if constexpr (std::is_same_v<T, double>) return ((const double *)p)[k];
    else if constexpr (std::is_same_v<T, float>) return ((const float *)p)[k];
    else if constexpr (std::is_same_v<T, f16> || std::is_same_v<T, f16w>)
        return T::raw(half_bits_to_float(((const uint16_t *)p)[k]));
    else {
        uint32_t b = (uint32_t)((const uint16_t *)p)[k] << 16;
        float f;
        std::memcpy(&f, &b, 4);
        return T::raw(f);
    }
}
template <class T> static inline void store_elem(void *p, size_t k, T x) {
    if constexpr (std::is_same_v<T, double>) ((double *)p)[k] = x;
    else if constexpr (std::is_same_v<T, float>) ((float *)p)[k] = x;
    else if constexpr (std::is_same_v<T, f16>) ((uint16_t *)p)[k] = float_to_half_bits(x.v);
    else if constexpr (std::is_same_v<T, f16w>) ((uint16_t *)p)[k] = float_to_half_bits(round_binary16(x.v));
    else {
        uint32_t b;
        const float r = std::is_same_v<T, bf16w> ? round_bfloat16(x.v) : x.v;
        std::memcpy(&b, &r, 4);
        ((uint16_t *)p)[k] = (uint16_t)(b >> 16);
    }
}

// ------------------------------------------------------------------ Julia-layout arrays

template <class T> struct Arr {
    int m = 0, n = 0;
    std::vector<T> a;
    Arr() {}
    Arr(int m_, int n_) : m(m_), n(n_), a((size_t)m_ * n_, T(0.0)) {}
    T &operator()(int i, int j) { return a[(size_t)(i - 1) + (size_t)(j - 1) * m]; }
    const T &operator()(int i, int j) const { return a[(size_t)(i - 1) + (size_t)(j - 1) * m]; }
    size_t size() const { return a.size(); }
    T &end(int di, int j) { return (*this)(m - di, j); }
};

struct ShapeError {};
#define CHECK(cond) \
    do { if (!(cond)) throw ShapeError(); } while (0)

template <class T> static void load_arr(Arr<T> &A, const void *p) {
    for (size_t k = 0; k < A.size(); ++k) A.a[k] = load_elem<T>(p, k);
}
template <class T> static void store_arr(const Arr<T> &A, void *p) {
    for (size_t k = 0; k < A.size(); ++k) store_elem<T>(p, k, A.a[k]);
}
template <class D, class S> static void copyto(Arr<D> &d, const Arr<S> &s) {  // copyto!/convert
    CHECK(d.m == s.m && d.n == s.n);
    for (size_t k = 0; k < d.size(); ++k) d.a[k] = cvt<D>(s.a[k]);
}

// ------------------------------------------------------------------ gradients.jl

// ∂x!  gradients.jl:3-12
template <class T> static void dx(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m + 1 == u.m && o.n == u.n);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) o(i, j) = u(i + 1, j) - u(i, j);
}
// ∂y!  gradients.jl:16-25
template <class T> static void dy(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m == u.m && o.n + 1 == u.n);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) o(i, j) = u(i, j + 1) - u(i, j);
}
// ∇²!  gradients.jl:29-42 (low-precision resilient form)
template <class T> static void lap(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m + 2 == u.m && o.n + 2 == u.n);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) {
            T c = u(i + 1, j + 1);
            o(i, j) = ((u(i, j + 1) - c) + (u(i + 2, j + 1) - c)) + ((u(i + 1, j) - c) + (u(i + 1, j + 2) - c));
        }
}

// ------------------------------------------------------------------ interpolations.jl

// Ix!  interpolations.jl:3-14
template <class T> static void Ix(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m + 1 == u.m && o.n == u.n);
    const T half = T(0.5);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) o(i, j) = half * (u(i + 1, j) + u(i, j));
}
// Iy!  interpolations.jl:18-29
template <class T> static void Iy(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m == u.m && o.n + 1 == u.n);
    const T half = T(0.5);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) o(i, j) = half * (u(i, j + 1) + u(i, j));
}
// Ixy!  interpolations.jl:33-45
template <class T> static void Ixy(Arr<T> &o, const Arr<T> &u) {
    CHECK(o.m + 1 == u.m && o.n + 1 == u.n);
    const T q = T(0.25);
    for (int j = 1; j <= o.n; ++j)
        for (int i = 1; i <= o.m; ++i) o(i, j) = q * (u(i, j) + u(i + 1, j)) + q * (u(i, j + 1) + u(i + 1, j + 1));
}

// ------------------------------------------------------------------ the model

// ------------------------------------------------------------------ planted mistakes
// Test infrastructure of the pins: swmo_set_mutation(k) plants mistake k in this restatement
// so that a test can show its closed form fails under that mistake (0: none, the default).
enum {
    MUT_NONE = 0,
    MUT_DT_MINUS = 1,        // DT = (dudx - dvdy)^2; diffusion.jl:95 has +
    MUT_S12_NO_EP = 2,       // S12 from dLudy[i,j] instead of dLudy[i+ep,j]; diffusion.jl:148,197
    MUT_BU_NO_CD = 3,        // Bu without the drag coefficient cD; bottom_drag.jl:62
    MUT_SQRTKE_ROW = 4,      // sqrtKE from KEu[i+ep,j] instead of KEu[i+ep,j+1]; bottom_drag.jl:48
    MUT_ADD_BU_SHIFT = 5,    // u0 += Δt_diff·Bu[i-ep,j+1] instead of Bu[i+1-ep,j+1]; diffusion.jl:237
    MUT_RK_U1_FROM_U1 = 6,   // u1 = u1 + RKb·du instead of u + RKb·du; time_integration.jl:69
    MUT_SSPRK2_WEIGHTS = 7,  // u0 = b·u + a·u1 (weights swapped); time_integration.jl:108
    MUT_SSPRK3_KN = 8,       // kna and knb swapped at stage kn; time_integration.jl:129
    MUT_4SSPRK3_THIRDS = 9,  // u1 = 1/3·u + 2/3·u1; time_integration.jl:187
    MUT_RELAX_SIGN = 10,     // γ·(η - η_ref); continuity.jl:21
    MUT_FTIME_COS = 11,      // Ftime = cos(ω t); forcing.jl:216
    MUT_S11_ROW = 12,        // constant viscosity: S11 from dLudx[i,j+1]; diffusion.jl:188
    MUT_WIDE_ROUNDS = 13,    // Float16 / BFloat16 values with float arithmetic: every operation rounds to the format
};
static int g_mutation = MUT_NONE;

template <class T, class Tp, class Tc> struct Model {
    const swm_config &c;
    int nx, ny, nux, nuy, nvx, nvy, ep;
    bool periodic;
    // static fields (Forcing{T}, Grid.f_q; Grid.f_u, f_v for the linear dynamics)
    Arr<T> H, Fx, Fy, f_q, eta_ref, Feta, f_u, f_v;
    // PrognosticVars{Tprog}
    Arr<Tp> u, v, eta, sst;
    // RungeKuttaVars{Tprog}, TendencyVars{Tprog}   preallocate.jl:2-73
    Arr<Tp> u0, u1, v0, v1, eta0, eta1, du, dv, deta;
    // PrognosticVarsRHS{T}
    Arr<T> ur, vr, etar, sstr;
    // VolumeFluxVars{T}   preallocate.jl:77-116
    Arr<T> h, h_u, U, h_v, V, dUdx, dVdy;
    // VorticityVars{T}    preallocate.jl:119-169
    Arr<T> h_q, q, qhu, qhv, dudx, dudy, dvdx, dvdy;
    // BernoulliVars{T}    preallocate.jl:172-211
    Arr<T> u2, v2, KEu, KEv, p, dpdx, dpdy;
    // BottomdragVars{T}   preallocate.jl:214-250
    Arr<T> sqrtKE, sqrtKE_u, sqrtKE_v, Bu, Bv;
    // ArakawaHsuVars{T}   preallocate.jl:253-288
    Arr<T> qa, qb, qg, qd;
    // VorticityVars{T} of the Sadourny scheme  preallocate.jl:141-145
    Arr<T> q_u, V_u, q_v, U_v;
    // VorticityVars{T} of the linear dynamics (preallocate.jl:150-151) and the array g*η that
    // rhs_linear! allocates
    Arr<T> u_v, v_u, geta;
    // LaplaceVars{T}      preallocate.jl:291-329
    Arr<T> Lu, Lv, dLudx, dLudy, dLvdx, dLvdy;
    // SmagorinskyVars{T}  preallocate.jl:332-385
    Arr<T> DT, DS, nuSmag, DS_q, nuSmag_q, S12, S21, S11, S22, LLu1, LLu2, LLv1, LLv2;
    // SemiLagrangeVars{T} preallocate.jl:388-436
    Arr<T> xd, yd, um, vm, u_T, um_T, v_T, vm_T, uinterp, vinterp, ssti;
    // intermediate states of one step (set by swmo_tape): the RK stage inputs, u0 before and
    // after the dissipation, and the tracer's mid-point velocities after the step
    struct Tape {
        std::vector<Arr<Tp>> s;
        std::vector<Arr<Tp>> p;          // 4SSPRK3: u1, v1 of every stage after ghost_points_uv!
        Arr<Tp> au, av, ae, du, dv;
        Arr<T> um, vm;
    } *tape = nullptr;

    // Low-precision options (not in the reference; defined by this library): u, v carried as
    // s·u, s·v with s = 2^scale_log2, the factor 1/s absorbed where the equations are not
    // linear in u, v (PV weights, Bernoulli potential, divergence, drag speed, Smagorinsky
    // coefficient, tracer departure); with compensated, the RK increments are summed apart
    // and added to the state with Kahan's compensated summation (a persistent compensation
    // per variable).
    double sc = 1.0, sci = 1.0;
    bool comp = false;
    Arr<Tp> cu, cv, ce;                 // Kahan compensations of u, v, η

    explicit Model(const swm_config &cfg) : c(cfg) {
        sc = std::ldexp(1.0, c.scale_log2);
        sci = std::ldexp(1.0, -c.scale_log2);
        comp = c.compensated != 0;
        nx = c.nx;
        ny = c.ny;
        periodic = c.bc == SWM_PERIODIC;
        ep = periodic ? 1 : 0;
        nux = periodic ? nx : nx - 1;
        nuy = ny;
        nvx = nx;
        nvy = ny - 1;
        H = Arr<T>(nx + 2, ny + 2);
        Fx = Arr<T>(nux, nuy);
        Fy = f_v = Arr<T>(nvx, nvy);
        f_u = Arr<T>(nux, nuy);
        f_q = Arr<T>(nx + 1, ny + 1);
        eta_ref = Arr<T>(nx, ny);
        Feta = Arr<T>(nx, ny);
        u = u0 = u1 = du = cu = Arr<Tp>(nux + 4, nuy + 4);
        v = v0 = v1 = dv = cv = Arr<Tp>(nvx + 4, nvy + 4);
        eta = eta0 = eta1 = deta = ce = Arr<Tp>(nx + 2, ny + 2);
        sst = Arr<Tp>(nx + 2, ny);
        ur = u2 = um = Arr<T>(nux + 4, nuy + 4);
        vr = v2 = vm = Arr<T>(nvx + 4, nvy + 4);
        etar = h = p = sqrtKE = DT = DS = nuSmag = geta = Arr<T>(nx + 2, ny + 2);
        u_v = Arr<T>(nux + 3, nuy + 3);
        v_u = Arr<T>(nvx + 3, nvy + 3);
        sstr = ssti = Arr<T>(nx + 2, ny);
        h_u = U = dpdx = sqrtKE_u = Bu = Arr<T>(nx + 1, ny + 2);
        h_v = V = dpdy = sqrtKE_v = Bv = Arr<T>(nx + 2, ny + 1);
        dUdx = Arr<T>(nx, ny + 2);
        dVdy = Arr<T>(nx + 2, ny);
        h_q = q = nuSmag_q = S12 = S21 = Arr<T>(nx + 1, ny + 1);
        qhu = Arr<T>(nvx, nvy);
        qhv = Arr<T>(nux, nuy);
        dudx = KEu = u_T = um_T = Arr<T>(nux + 3, nuy + 4);
        dudy = Arr<T>(nux + 4, nuy + 3);
        dvdx = DS_q = Arr<T>(nvx + 3, nvy + 4);
        dvdy = KEv = v_T = vm_T = Arr<T>(nvx + 4, nvy + 3);
        qa = qd = xd = yd = uinterp = vinterp = Arr<T>(nx, ny);
        qb = qg = Arr<T>(nx + 1, ny);
        q_u = V_u = Arr<T>(nx + 1, ny);
        q_v = U_v = Arr<T>(nx, ny + 1);
        Lu = Arr<T>(nux + 2, nuy + 2);
        Lv = Arr<T>(nvx + 2, nvy + 2);
        dLudx = S11 = Arr<T>(nux + 1, nuy + 2);
        dLudy = Arr<T>(nux + 2, nuy + 1);
        dLvdx = Arr<T>(nvx + 1, nvy + 2);
        dLvdy = S22 = Arr<T>(nvx + 2, nvy + 1);
        LLu1 = Arr<T>(nux, nuy + 2);
        LLu2 = Arr<T>(nx + 1, ny);
        LLv1 = Arr<T>(nx, ny + 1);
        LLv2 = Arr<T>(nvx + 2, nvy);
    }

    // Base.convert of time_integration.jl:330-347: alias if the formats agree, else copy.
    const Arr<T> &rhs_view(const Arr<Tp> &x, Arr<T> &buf) {
        if constexpr (std::is_same_v<T, Tp>) return x;
        else {
            copyto(buf, x);
            return buf;
        }
    }

    // ------------------------------------------------------------ ghost_points.jl
    static Tp comm(Tp x) { return cvt<Tp>(cvt<Tc>(x)); }  // T(Tcomm(x)), ghost_points.jl:65

    void ghost_u_nonperiodic(Arr<Tp> &a) {                // ghost_points.jl:47-57
        const Tp oma = Tp(c.one_minus_alpha);
        for (int i = 1; i <= a.m; ++i) {
            a(i, 1) = oma * a(i, 4);
            a(i, 2) = oma * a(i, 3);
            a(i, a.n - 1) = oma * a(i, a.n - 2);
            a(i, a.n) = oma * a(i, a.n - 3);
        }
    }
    void ghost_u_periodic(Arr<Tp> &a) {                   // ghost_points.jl:60-73
        for (int j = 1; j <= a.n; ++j) {
            a(1, j) = comm(a(a.m - 3, j));
            a(2, j) = comm(a(a.m - 2, j));
            a(a.m - 1, j) = comm(a(3, j));
            a(a.m, j) = comm(a(4, j));
        }
    }
    void ghost_v_nonperiodic(Arr<Tp> &a) {                // ghost_points.jl:76-86
        const Tp oma = Tp(c.one_minus_alpha);
        for (int j = 1; j <= a.n; ++j) {
            a(1, j) = oma * a(4, j);
            a(2, j) = oma * a(3, j);
            a(a.m - 1, j) = oma * a(a.m - 2, j);
            a(a.m, j) = oma * a(a.m - 3, j);
        }
    }
    void ghost_v_periodic(Arr<Tp> &a) {                   // ghost_points.jl:89-99
        for (int j = 1; j <= a.n; ++j) {
            a(1, j) = comm(a(a.m - 3, j));
            a(2, j) = comm(a(a.m - 2, j));
            a(a.m - 1, j) = comm(a(3, j));
            a(a.m, j) = comm(a(4, j));
        }
    }
    void ghost_eta_nonperiodic(Arr<Tp> &a) {              // ghost_points.jl:102-116
        for (int j = 1; j <= a.n; ++j) {
            a(1, j) = a(2, j);
            a(a.m, j) = a(a.m - 1, j);
        }
        for (int i = 1; i <= a.m; ++i) {
            a(i, 1) = a(i, 2);
            a(i, a.n) = a(i, a.n - 1);
        }
    }
    void ghost_eta_periodic(Arr<Tp> &a) {                 // ghost_points.jl:119-133
        for (int j = 1; j <= a.n; ++j) {
            a(1, j) = comm(a(a.m - 1, j));
            a(a.m, j) = comm(a(2, j));
        }
        for (int i = 1; i <= a.m; ++i) {
            a(i, 1) = a(i, 2);
            a(i, a.n) = a(i, a.n - 1);
        }
    }
    template <class S> void ghost_sst_any(Arr<S> &a) {    // ghost_points.jl:136-179, halosstx=1, halossty=0
        const int hx = 1;
        for (int j = 1; j <= a.n; ++j)
            for (int i = 1; i <= hx; ++i) {
                if (periodic) {
                    a(i, j) = cvt<S>(cvt<Tc>(a(a.m - 2 * hx + i, j)));
                    a(a.m - hx + i, j) = cvt<S>(cvt<Tc>(a(hx + i, j)));
                } else {
                    a(i, j) = a(hx + 1, j);
                    a(a.m - i + 1, j) = a(a.m - hx, j);
                }
            }
    }
    void ghost_points(Arr<Tp> &uu, Arr<Tp> &vv, Arr<Tp> &ee) {  // ghost_points.jl:182-199
        if (periodic) { ghost_u_periodic(uu); ghost_v_periodic(vv); ghost_eta_periodic(ee); }
        else { ghost_u_nonperiodic(uu); ghost_v_nonperiodic(vv); ghost_eta_nonperiodic(ee); }
    }
    void ghost_points_uv(Arr<Tp> &uu, Arr<Tp> &vv) {           // ghost_points.jl:202-216
        if (periodic) { ghost_u_periodic(uu); ghost_v_periodic(vv); }
        else { ghost_u_nonperiodic(uu); ghost_v_nonperiodic(vv); }
    }
    void ghost_points_eta(Arr<Tp> &ee) {                       // ghost_points.jl:219-229
        if (periodic) ghost_eta_periodic(ee);
        else ghost_eta_nonperiodic(ee);
    }

    // ------------------------------------------------------------ rhs.jl
    void thickness(const Arr<T> &e) {                          // rhs.jl:131-139
        CHECK(h.m == e.m && h.n == e.n && H.m == e.m && H.n == e.n);
        for (size_t k = 0; k < h.size(); ++k) h.a[k] = e.a[k] + H.a[k];
    }
    void speed(const Arr<T> &uu, const Arr<T> &vv) {           // rhs.jl:142-159
        for (size_t k = 0; k < uu.size(); ++k) u2.a[k] = uu.a[k] * uu.a[k];
        for (size_t k = 0; k < vv.size(); ++k) v2.a[k] = vv.a[k] * vv.a[k];
    }
    void Uflux(const Arr<T> &uu) {                             // rhs.jl:277-291
        CHECK(U.m + 2 + ep == uu.m && U.n + 2 == uu.n);
        for (int j = 1; j <= U.n; ++j)
            for (int i = 1; i <= U.m; ++i) U(i, j) = uu(1 + ep + i, 1 + j) * h_u(i, j);
    }
    void Vflux(const Arr<T> &vv) {                             // rhs.jl:294-304
        CHECK(V.m + 2 == vv.m && V.n + 2 == vv.n);
        for (int j = 1; j <= V.n; ++j)
            for (int i = 1; i <= V.m; ++i) V(i, j) = vv(i + 1, j + 1) * h_v(i, j);
    }
    void UVfluxes(const Arr<T> &uu, const Arr<T> &vv, const Arr<T> &e) {  // rhs.jl:307-324
        thickness(e);
        Ix(h_u, h);
        Iy(h_v, h);
        Uflux(uu);
        Vflux(vv);
    }
    void PV() {                                                // PV_advection.jl:2-18
        CHECK(q.m + 2 == dvdx.m && q.n + 2 == dvdx.n && q.m + 2 + ep == dudy.m);
        for (int j = 1; j <= q.n; ++j)
            for (int i = 1; i <= q.m; ++i)
                q(i, j) = (f_q(i, j) + dvdx(i + 1, j + 1) - dudy(i + 1 + ep, j + 1)) / h_q(i, j);
    }
    void AHa() {                                               // PV_advection.jl:88-99
        const T k = T(1.0 / 12.0) * T(sci);
        for (int j = 1; j <= qa.n; ++j)
            for (int i = 1; i <= qa.m; ++i) qa(i, j) = k * (q(i, j) + q(i, j + 1) + q(i + 1, j + 1));
    }
    void AHb() {                                               // PV_advection.jl:103-115
        const T k = T(1.0 / 12.0) * T(sci);
        for (int j = 1; j <= qb.n; ++j) {
            qb(1, j) = k * (q(1, j) + q(1, j + 1) + q(q.m, j + 1));
            for (int i = 2; i <= qb.m; ++i) qb(i, j) = k * (q(i, j) + q(i, j + 1) + q(i - 1, j + 1));
        }
    }
    void AHg() {                                               // PV_advection.jl:119-131
        const T k = T(1.0 / 12.0) * T(sci);
        for (int j = 1; j <= qg.n; ++j) {
            qg(1, j) = k * (q(1, j) + q(1, j + 1) + q(q.m, j));
            for (int i = 2; i <= qg.m; ++i) qg(i, j) = k * (q(i, j) + q(i - 1, j) + q(i, j + 1));
        }
    }
    void AHd() {                                               // PV_advection.jl:135-146
        const T k = T(1.0 / 12.0) * T(sci);
        for (int j = 1; j <= qd.n; ++j)
            for (int i = 1; i <= qd.m; ++i) qd(i, j) = k * (q(i, j) + q(i, j + 1) + q(i + 1, j));
    }
    void advection_coriolis(const Arr<T> &uu, const Arr<T> &vv) {  // rhs.jl:89-128
        Ixy(h_q, h);
        dx(dvdx, vv);
        dy(dudy, uu);
        speed(uu, vv);
        Ix(KEu, u2);
        Iy(KEv, v2);
        PV();
        PV_combinations();
    }
    void PV_combinations() {                                   // rhs.jl:117-128
        if (c.adv_scheme == SWM_ADV_SADOURNY) {                // Iy!(q_u,q), Ix!(q_v,q)
            const T half = T(0.5) * T(sci);                    // (1/s: q and the fluxes both carry s)
            for (int j = 1; j <= q_u.n; ++j)
                for (int i = 1; i <= q_u.m; ++i) q_u(i, j) = half * (q(i, j + 1) + q(i, j));
            for (int j = 1; j <= q_v.n; ++j)
                for (int i = 1; i <= q_v.m; ++i) q_v(i, j) = half * (q(i + 1, j) + q(i, j));
            return;
        }
        AHa();
        AHb();
        AHg();
        AHd();
    }
    void PV_advection() {                                      // PV_advection.jl:21-27
        if (c.adv_scheme == SWM_ADV_SADOURNY) PV_Sadourny();
        else PV_ArakawaHsu();
    }
    void PV_Sadourny() {                                       // PV_advection.jl:31-59
        Ixy(V_u, V);
        Ixy(U_v, U);
        CHECK(qhv.m + 2 - ep == q_u.m && qhv.n == q_u.n && qhu.m == q_v.m && qhu.n + 2 == q_v.n);
        for (int j = 1; j <= qhv.n; ++j)
            for (int i = 1; i <= qhv.m; ++i) qhv(i, j) = q_u(i + 1 - ep, j) * V_u(i + 1 - ep, j);
        for (int j = 1; j <= qhu.n; ++j)
            for (int i = 1; i <= qhu.m; ++i) qhu(i, j) = q_v(i, j + 1) * U_v(i, j + 1);
    }
    void PV_ArakawaHsu() {                                     // PV_advection.jl:63-84
        for (int j = 1; j <= qhv.n; ++j)
            for (int i = 1; i <= qhv.m; ++i)
                qhv(i, j) = qa(1 - ep + i, j) * V(2 - ep + i, j + 1) + qb(1 - ep + i, j) * V(1 - ep + i, j + 1) +
                            qg(1 - ep + i, j) * V(1 - ep + i, j) + qd(1 - ep + i, j) * V(2 - ep + i, j);
        for (int j = 1; j <= qhu.n; ++j)
            for (int i = 1; i <= qhu.m; ++i)
                qhu(i, j) = qa(i, j) * U(i, j + 1) + qb(i + 1, j) * U(i + 1, j + 1) + qg(i + 1, j + 1) * U(i + 1, j + 2) +
                            qd(i, j + 1) * U(i, j + 2);
    }
    void bernoulli(const Arr<T> &e) {                          // rhs.jl:162-180
        CHECK(p.m + ep == KEu.m && p.n + 2 == KEu.n && p.m + 2 == KEv.m && p.n == KEv.n);
        const T half = T(0.5) * T(sci), g = T(c.g) * T(sc);
        for (int j = 1; j <= p.n; ++j)
            for (int i = 1; i <= p.m; ++i) p(i, j) = half * (KEu(i + ep, j + 1) + KEv(i + 1, j)) + g * e(i, j);
    }
    T Ftime(int64_t t, double w) const {                  // forcing.jl:215-217
        return cvt<T>(g_mutation == MUT_FTIME_COS ? std::cos(w * (double)t) : std::sin(w * (double)t));
    }
    void momentum_u(int64_t t) {                               // rhs.jl:218-245
        const int m = du.m - 4, n = du.n - 4;
        CHECK(m == qhv.m && n == qhv.n && m + 2 - ep == dpdx.m && n + 2 == dpdx.n && m == Fx.m && n == Fx.n);
        const T Fxt = c.seasonal_wind_x ? Ftime(t, c.omegaFx) : T(1.0);
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                du(i + 2, j + 2) = (cvt<Tp>(qhv(i, j)) - cvt<Tp>(dpdx(i + 1 - ep, j + 1))) + cvt<Tp>(Fxt * Fx(i, j));
    }
    void momentum_v(int64_t t) {                               // rhs.jl:248-274
        const int m = dv.m - 4, n = dv.n - 4;
        CHECK(m == qhu.m && n == qhu.n && m + 2 == dpdy.m && n + 2 == dpdy.n && m == Fy.m && n == Fy.n);
        const T Fyt = c.seasonal_wind_y ? Ftime(t, c.omegaFy) : T(1.0);
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                dv(i + 2, j + 2) = -(cvt<Tp>(qhu(i, j)) + cvt<Tp>(dpdy(i + 1, j + 1))) + cvt<Tp>(Fyt * Fy(i, j));
    }
    // rhs_linear!  rhs.jl:55-86: f·v and f·u on the u- and v-grids (fv!, fu!  rhs.jl:183-215)
    // and the gradient of g·η alone; U and V are not formed (continuity! under the classic RK
    // scheme then reads the zero fluxes it was allocated with).
    void rhs_linear(const Arr<T> &uu, const Arr<T> &vv, const Arr<T> &e, int64_t t) {
        const T g = T(c.g) * T(sc);
        for (size_t k = 0; k < geta.size(); ++k) geta.a[k] = g * e.a[k];
        dx(dpdx, geta);
        dy(dpdy, geta);
        Ixy(v_u, vv);
        Ixy(u_v, uu);
        CHECK(v_u.m == qhv.m + 4 - ep && v_u.n == qhv.n + 2 && u_v.m == qhu.m + 2 + ep && u_v.n == qhu.n + 4);
        for (int j = 1; j <= qhv.n; ++j)
            for (int i = 1; i <= qhv.m; ++i) qhv(i, j) = f_u(i, j) * v_u(i + 2 - ep, j + 1);
        for (int j = 1; j <= qhu.n; ++j)
            for (int i = 1; i <= qhu.m; ++i) qhu(i, j) = f_v(i, j) * u_v(i + 1 + ep, j + 2);
        momentum_u(t);
        momentum_v(t);
    }
    void rhs(const Arr<T> &uu, const Arr<T> &vv, const Arr<T> &e, int64_t t) {  // rhs.jl:2-51
        if (c.dynamics == SWM_LINEAR) return rhs_linear(uu, vv, e, t);
        UVfluxes(uu, vv, e);
        if (c.nstep_advcor == 0) advection_coriolis(uu, vv);
        PV_advection();
        bernoulli(e);
        dx(dpdx, p);
        dy(dpdy, p);
        momentum_u(t);
        momentum_v(t);
    }

    // ------------------------------------------------------------ continuity.jl
    void continuity(const Arr<T> &uu, const Arr<T> &vv, const Arr<T> &e, int64_t t) {  // continuity.jl:72-98
        if (c.time_scheme != SWM_RK) UVfluxes(uu, vv, e);
        dx(dUdx, U);
        dy(dVdy, V);
        const int m = deta.m - 2, n = deta.n - 2;
        CHECK(m == dUdx.m && n + 2 == dUdx.n && m + 2 == dVdy.m && n == dVdy.n);
        if (c.continuity == SWM_CONT_RELAX) {                  // continuity.jl:3-24
            const T gam = T(c.gamma);
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i)
                    deta(i + 1, j + 1) = -(cvt<Tp>(dUdx(i, j + 1)) + cvt<Tp>(dVdy(i + 1, j))) * Tp(sci) +
                                         cvt<Tp>(gam * (g_mutation == MUT_RELAX_SIGN ? e(i + 1, j + 1) - eta_ref(i, j)
                                                                                     : eta_ref(i, j) - e(i + 1, j + 1)));
        } else if (c.continuity == SWM_CONT_FORCING) {         // continuity.jl:27-49
            const T Fet = Ftime(t, c.omegaFeta);
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i)
                    deta(i + 1, j + 1) = -(cvt<Tp>(dUdx(i, j + 1)) + cvt<Tp>(dVdy(i + 1, j))) * Tp(sci) + cvt<Tp>(Fet * Feta(i, j));
        } else {                                               // continuity.jl:52-68
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) deta(i + 1, j + 1) = -(cvt<Tp>(dUdx(i, j + 1)) + cvt<Tp>(dVdy(i + 1, j))) * Tp(sci);
        }
    }

    // ------------------------------------------------------------ bottom_drag.jl
    void bottom_drag(const Arr<T> &uu, const Arr<T> &vv, const Arr<T> &e) {  // bottom_drag.jl:2-17
        if (c.bottom_drag == SWM_DRAG_QUADRATIC) {             // bottom_drag.jl:20-76
            const T cD = T(c.cD);
            thickness(e);
            Ix(h_u, h);
            Iy(h_v, h);
            speed(uu, vv);
            Ix(KEu, u2);
            Iy(KEv, v2);
            CHECK(sqrtKE.m + ep == KEu.m && sqrtKE.n + 2 == KEu.n);
            for (int j = 1; j <= sqrtKE.n; ++j)
                for (int i = 1; i <= sqrtKE.m; ++i)
                    sqrtKE(i, j) = num_sqrt(KEu(i + ep, g_mutation == MUT_SQRTKE_ROW ? j : j + 1) + KEv(i + 1, j)) * T(sci);
            Ix(sqrtKE_u, sqrtKE);
            Iy(sqrtKE_v, sqrtKE);
            for (int j = 1; j <= Bu.n; ++j)
                for (int i = 1; i <= Bu.m; ++i)
                    Bu(i, j) = (g_mutation == MUT_BU_NO_CD ? T(1.0) : cD) * sqrtKE_u(i, j) * uu(i + 1 + ep, j + 1) / h_u(i, j);
            for (int j = 1; j <= Bv.n; ++j)
                for (int i = 1; i <= Bv.m; ++i) Bv(i, j) = cD * sqrtKE_v(i, j) * vv(i + 1, j + 1) / h_v(i, j);
        } else if (c.bottom_drag == SWM_DRAG_LINEAR) {         // bottom_drag.jl:80-107
            const T rD = T(c.rD);
            for (int j = 1; j <= Bu.n; ++j)
                for (int i = 1; i <= Bu.m; ++i) Bu(i, j) = rD * uu(i + 1 + ep, j + 1);
            for (int j = 1; j <= Bv.n; ++j)
                for (int i = 1; i <= Bv.m; ++i) Bv(i, j) = rD * vv(i + 1, j + 1);
        } else {                                               // bottom_drag.jl:110-119
            for (auto &x : Bu.a) x = T(0.0);
            for (auto &x : Bv.a) x = T(0.0);
        }
    }

    // ------------------------------------------------------------ diffusion.jl
    void stress_tensor(const Arr<T> &uu, const Arr<T> &vv) {   // diffusion.jl:114-122
        lap(Lu, uu);
        lap(Lv, vv);
        dx(dLudx, Lu);
        dy(dLudy, Lu);
        dx(dLvdx, Lv);
        dy(dLvdy, Lv);
    }
    void smagorinsky_coeff() {                                 // diffusion.jl:68-111
        const T cS = T(c.cSmag);
        CHECK(DS_q.m + ep == dudy.m && DS_q.n == dudy.n && DS_q.m == dvdx.m && DS_q.n == dvdx.n);
        for (int j = 1; j <= DS_q.n; ++j)
            for (int i = 1; i <= DS_q.m; ++i) {
                T s = dudy(i + ep, j) + dvdx(i, j);
                DS_q(i, j) = s * s;
            }
        Ixy(DS, DS_q);
        CHECK(DT.m + ep == dudx.m && DT.n + 2 == dudx.n && DT.m + 2 == dvdy.m && DT.n == dvdy.n);
        for (int j = 1; j <= DT.n; ++j)
            for (int i = 1; i <= DT.m; ++i) {
                T s = g_mutation == MUT_DT_MINUS ? dudx(i + ep, j + 1) - dvdy(i + 1, j) : dudx(i + ep, j + 1) + dvdy(i + 1, j);
                DT(i, j) = s * s;
            }
        for (int j = 1; j <= nuSmag.n; ++j)
            for (int i = 1; i <= nuSmag.m; ++i) nuSmag(i, j) = cS * (num_sqrt(DS(i, j) + DT(i, j)) * T(sci));
        Ixy(nuSmag_q, nuSmag);
    }
    void viscous_tensor_smagorinsky() {                        // diffusion.jl:126-171
        CHECK(S11.m + 2 - ep == nuSmag.m && S11.n == nuSmag.n);
        for (int j = 1; j <= S11.n; ++j)
            for (int i = 1; i <= S11.m; ++i) S11(i, j) = nuSmag(i + 1 - ep, j) * dLudx(i, j);
        for (int j = 1; j <= S12.n; ++j)
            for (int i = 1; i <= S12.m; ++i) S12(i, j) = nuSmag_q(i, j) * dLudy(i + (g_mutation == MUT_S12_NO_EP ? 0 : ep), j);
        for (int j = 1; j <= S21.n; ++j)
            for (int i = 1; i <= S21.m; ++i) S21(i, j) = nuSmag_q(i, j) * dLvdx(i, j);
        for (int j = 1; j <= S22.n; ++j)
            for (int i = 1; i <= S22.m; ++i) S22(i, j) = nuSmag(i, j + 1) * dLvdy(i, j);
    }
    void viscous_tensor_constant() {                           // diffusion.jl:175-218
        const T nu = T(c.nuB);
        for (int j = 1; j <= S11.n; ++j)
            for (int i = 1; i <= S11.m; ++i) S11(i, j) = nu * dLudx(i, g_mutation == MUT_S11_ROW && j < S11.n ? j + 1 : j);
        for (int j = 1; j <= S12.n; ++j)
            for (int i = 1; i <= S12.m; ++i) S12(i, j) = nu * dLudy(i + (g_mutation == MUT_S12_NO_EP ? 0 : ep), j);
        for (int j = 1; j <= S21.n; ++j)
            for (int i = 1; i <= S21.m; ++i) S21(i, j) = nu * dLvdx(i, j);
        for (int j = 1; j <= S22.n; ++j)
            for (int i = 1; i <= S22.m; ++i) S22(i, j) = nu * dLvdy(i, j);
    }
    void diffusion(const Arr<T> &uu, const Arr<T> &vv) {      // diffusion.jl:2-64
        if (c.diffusion == SWM_DIFF_SMAGORINSKY) {
            dx(dudx, uu);
            dy(dvdy, vv);
            dx(dvdx, vv);
            dy(dudy, uu);
            stress_tensor(uu, vv);
            smagorinsky_coeff();
            viscous_tensor_smagorinsky();
        } else {
            stress_tensor(uu, vv);
            viscous_tensor_constant();
        }
        dx(LLu1, S11);
        dy(LLu2, S12);
        dx(LLv1, S21);
        dy(LLv2, S22);
    }
    void add_drag_diff_tendencies(Arr<Tp> &uu, Arr<Tp> &vv) { // diffusion.jl:221-251
        const Tp dtd = Tp(c.Dt_diff);
        int m = uu.m - 4, n = uu.n - 4;
        CHECK(m + 2 - ep == Bu.m && n + 2 == Bu.n && m == LLu1.m && n + 2 == LLu1.n && m + 2 - ep == LLu2.m && n == LLu2.n);
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                uu(i + 2, j + 2) += dtd * (cvt<Tp>(Bu(i + 1 - ep - (g_mutation == MUT_ADD_BU_SHIFT && i > 1 ? 1 : 0), j + 1)) +
                                           cvt<Tp>(LLu1(i, j + 1)) + cvt<Tp>(LLu2(i + 1 - ep, j)));
        m = vv.m - 4;
        n = vv.n - 4;
        CHECK(m + 2 == Bv.m && n + 2 == Bv.n && m == LLv1.m && n + 2 == LLv1.n && m + 2 == LLv2.m && n == LLv2.n);
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                vv(i + 2, j + 2) += dtd * (cvt<Tp>(Bv(i + 1, j + 1)) + cvt<Tp>(LLv1(i, j + 1)) + cvt<Tp>(LLv2(i + 1, j)));
    }

    // ------------------------------------------------------------ tracer_advection.jl
    // backtraj!  tracer_advection.jl:96-123
    void backtraj(Arr<T> &rd, T dt, const Arr<T> &uv) {
        int is, js;
        if (rd.m == uv.m && rd.n == uv.n) { is = 0; js = 0; }
        else if (rd.m + 4 == uv.m && rd.n + 2 == uv.n) { is = 2; js = 1; }
        else if (rd.m + 2 + ep == uv.m && rd.n + 4 == uv.n) { is = 1 + ep; js = 2; }
        else throw ShapeError();
        for (int j = 1; j <= rd.n; ++j)
            for (int i = 1; i <= rd.m; ++i) rd(i, j) = -dt * uv(i + is, j + js);
    }
    // clip / wrap  tracer_advection.jl:211-239 ; returns the absolute index, writes the coordinate
    static int clip(int64_t xyi, int ij, T xy, int mn, int h1, T *x0) {
        int64_t s = xyi + ij + h1;
        if (s < 1 + h1) { *x0 = T(0.0); return 1 + h1; }
        if (s > mn - 1) { *x0 = T(1.0); return mn - 1; }
        *x0 = xy - cvt<T>((double)xyi);
        return (int)s;
    }
    static int wrap(int64_t xyi, int ij, T xy, int mn, int h1, int h2, T *x0) {
        int64_t s = xyi + ij + h1;
        *x0 = xy - cvt<T>((double)xyi);
        if (s < 1 + h1) return (int)(s + mn - h1 - h2);
        if (s > mn - 1) return (int)(s - mn + h1 + h2);
        return (int)s;
    }
    static T bilin(T f00, T f10, T f01, T f11, T x, T y) {     // tracer_advection.jl:243-246
        const T one = T(1.0);
        return f00 * (one - x) * (one - y) + f10 * x * (one - y) + f01 * (one - x) * y + f11 * x * y;
    }
    static int64_t ifloor(T x) {                                // Int(floor(x)); NaN throws InexactError
        double d = std::floor(to_f64(x));
        if (!std::isfinite(d)) throw ShapeError();
        return (int64_t)d;
    }
    void interp_uv(Arr<T> &uvi, const Arr<T> &uv, const Arr<T> &xx, const Arr<T> &yy) {  // :130-173
        const int m = uvi.m, n = uvi.n, muv = uv.m, nuv = uv.n;
        int xh1, xh2, yh1, yh2;
        if (m + 2 + ep == muv && n + 4 == nuv) { xh1 = 1 + ep; xh2 = 1; yh1 = 2; yh2 = 2; }
        else if (m + 4 == muv && n + 2 == nuv) { xh1 = 2; xh2 = 2; yh1 = 1; yh2 = 1; }
        else throw ShapeError();
        (void)yh2;
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                int64_t xi = ifloor(xx(i, j)), yi = ifloor(yy(i, j));
                T x0, y0;
                int k = periodic ? wrap(xi, i, xx(i, j), muv, xh1, xh2, &x0) : clip(xi, i, xx(i, j), muv, xh1, &x0);
                int l = clip(yi, j, yy(i, j), nuv, yh1, &y0);
                if (k < 1 || k + 1 > muv || l < 1 || l + 1 > nuv) throw ShapeError();  // Julia BoundsError
                uvi(i, j) = bilin(uv(k, l), uv(k + 1, l), uv(k, l + 1), uv(k + 1, l + 1), x0, y0);
            }
    }
    void departure(const Arr<T> &uu, const Arr<T> &vv) {       // tracer_advection.jl:52-85
        Ix(u_T, uu);
        for (auto &x : u_T.a) x = x * T(sci);
        Ix(um_T, um);
        Iy(v_T, vv);
        for (auto &x : v_T.a) x = x * T(sci);
        Iy(vm_T, vm);
        backtraj(xd, T(c.half_dtadvu), u_T);
        backtraj(yd, T(c.half_dtadvv), v_T);
        interp_uv(uinterp, um_T, xd, yd);
        interp_uv(vinterp, vm_T, xd, yd);
        for (auto &x : uinterp.a) x = x * T(sci);
        for (auto &x : vinterp.a) x = x * T(sci);
        backtraj(xd, T(c.dtadvu), uinterp);
        backtraj(yd, T(c.dtadvv), vinterp);
    }
    void adv_sst(const Arr<T> &s) {                            // tracer_advection.jl:179-206
        const int m = ssti.m, n = ssti.n, hx = 1, hy = 0;
        for (int j = 1; j <= n - 2 * hy; ++j)
            for (int i = 1; i <= m - 2 * hx; ++i) {
                int64_t xi = ifloor(xd(i, j)), yi = ifloor(yd(i, j));
                T x0, y0;
                int k = periodic ? wrap(xi, i, xd(i, j), m, hx, hx, &x0) : clip(xi, i, xd(i, j), m, hx, &x0);
                int l = clip(yi, j, yd(i, j), n, hy, &y0);
                if (k < 1 || k + 1 > m || l < 1 || l + 1 > n) throw ShapeError();
                ssti(i + hx, j + hy) = bilin(s(k, l), s(k + 1, l), s(k, l + 1), s(k + 1, l + 1), x0, y0);
            }
    }
    void tracer_consumption() {                                // tracer_advection.jl:262-275
        const T j0 = T(c.jSST), smin = T(c.SSTmin);
        for (int j = 1; j <= ssti.n; ++j)
            for (int i = 2; i <= ssti.m - 1; ++i) ssti(i, j) += j0 * (smin - ssti(i, j));
    }
    void tracer(int64_t i, const Arr<T> &uu, const Arr<T> &vv) {  // tracer_advection.jl:1-44
        if (!c.tracer_advection) return;
        if ((i + c.nadvstep_half) % c.nadvstep == 0) {
            copyto(um, uu);
            copyto(vm, vv);
        }
        if (i % c.nadvstep == 0) {
            const Arr<T> &sr = rhs_view(sst, sstr);
            departure(uu, vv);
            adv_sst(sr);
            if (c.tracer_consumption) tracer_consumption();
            ghost_sst_any(ssti);
            copyto(sst, ssti);
        }
    }

    // ------------------------------------------------------------ time_integration.jl
    // axb!, caxb!, cxab!, cxayb!, dxaybzc!  time_integration.jl:253-326 (whole arrays)
    static void axb(Arr<Tp> &a, Tp x, const Arr<Tp> &b) {
        for (size_t k = 0; k < a.size(); ++k) a.a[k] += x * b.a[k];
    }
    static void caxb(Arr<Tp> &cc, const Arr<Tp> &a, Tp x, const Arr<Tp> &b) {
        for (size_t k = 0; k < a.size(); ++k) cc.a[k] = a.a[k] + x * b.a[k];
    }
    static void cxab(Arr<Tp> &cc, double x, const Arr<Tp> &a, const Arr<Tp> &b) {
        const Tp xT = cvt<Tp>(x);
        for (size_t k = 0; k < a.size(); ++k) cc.a[k] = xT * (a.a[k] + b.a[k]);
    }
    static void cxayb(Arr<Tp> &cc, double x, const Arr<Tp> &a, double y, const Arr<Tp> &b) {
        const Tp xT = cvt<Tp>(x), yT = cvt<Tp>(y);
        for (size_t k = 0; k < a.size(); ++k) cc.a[k] = xT * a.a[k] + yT * b.a[k];
    }
    static void dxaybzc(Arr<Tp> &d, double x, const Arr<Tp> &a, double y, const Arr<Tp> &b, double z, const Arr<Tp> &cc) {
        const Tp xT = cvt<Tp>(x), yT = cvt<Tp>(y), zT = cvt<Tp>(z);
        for (size_t k = 0; k < a.size(); ++k) d.a[k] = xT * a.a[k] + yT * b.a[k] + zT * cc.a[k];
    }

    // d = x + d with Kahan's compensated summation: y = d - c, t = x + y, c = (t - x) - y
    static void kahan(Arr<Tp> &d, const Arr<Tp> &x, Arr<Tp> &cc) {
        for (size_t k = 0; k < d.size(); ++k) {
            const Tp y = d.a[k] - cc.a[k];
            const Tp t = x.a[k] + y;
            cc.a[k] = (t - x.a[k]) - y;
            d.a[k] = t;
        }
    }

    bool has_nan() const {                                    // feedback.jl:61-72
        for (auto &x : u.a) if (is_nan(x)) return true;
        for (auto &x : v.a) if (is_nan(x)) return true;
        for (auto &x : eta.a) if (is_nan(x)) return true;
        for (auto &x : sst.a) if (is_nan(x)) return true;
        return false;
    }

    // Initial thicknesses and PV terms, time_integration.jl:20-36.  thickness! is called
    // there with the Tprog array η: Julia promotes η[i]+H[i] to the wider format and
    // rounds the sum to T on the store.
    void init() {
        using W = std::conditional_t<std::is_same_v<T, Dual>, Dual,
                                     std::conditional_t<std::is_same_v<T, double> || std::is_same_v<Tp, double>, double, float>>;
        for (size_t k = 0; k < h.size(); ++k) h.a[k] = cvt<T>(cvt<W>(eta.a[k]) + cvt<W>(H.a[k]));
        Ix(h_u, h);
        Iy(h_v, h);
        Ixy(h_q, h);
        const Arr<T> &ur_ = rhs_view(u, ur);
        const Arr<T> &vr_ = rhs_view(v, vr);
        advection_coriolis(ur_, vr_);
        PV_advection();
        copyto(u0, u);
        copyto(v0, v);
        copyto(eta0, eta);
    }

    // One pass of the loop body of time_integration.jl:45-243.  stop_stage>0 ends the step
    // after that many stages of the scheme (for stage-level parity).  Returns true when NaNs
    // were detected (the caller breaks).
    bool step(int64_t i, int64_t t, int stop_stage) {
        ghost_points(u, v, eta);                               // :48
        copyto(u1, u);
        copyto(v1, v);
        copyto(eta1, eta);
        int stages_run = 0;
        auto stop = [&]() { return stop_stage > 0 && stages_run >= stop_stage; };

        if (c.time_scheme == SWM_RK) {                         // :53-78
            if (comp) {                                        // u0 holds the increment Σ a_k du_k
                for (auto &x : u0.a) x = Tp(0.0);
                for (auto &x : v0.a) x = Tp(0.0);
                for (auto &x : eta0.a) x = Tp(0.0);
            }
            for (int rki = 1; rki <= c.RKo; ++rki) {
                if (rki > 1) ghost_points(u1, v1, eta1);
                if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
                const Arr<T> &a = rhs_view(u1, ur), &b = rhs_view(v1, vr), &e = rhs_view(eta1, etar);
                rhs(a, b, e, t);
                continuity(a, b, e, t);
                if (rki < c.RKo) {
                    const bool m6 = g_mutation == MUT_RK_U1_FROM_U1;
                    caxb(u1, m6 ? u1 : u, Tp(c.RKbDt[rki - 1]), du);
                    caxb(v1, m6 ? v1 : v, Tp(c.RKbDt[rki - 1]), dv);
                    caxb(eta1, m6 ? eta1 : eta, Tp(c.RKbDt[rki - 1]), deta);
                }
                axb(u0, Tp(c.RKaDt[rki - 1]), du);
                axb(v0, Tp(c.RKaDt[rki - 1]), dv);
                axb(eta0, Tp(c.RKaDt[rki - 1]), deta);
                if (comp && rki == c.RKo) {                    // u0 = u + Σ, compensated
                    kahan(u0, u, cu);
                    kahan(v0, v, cv);
                    kahan(eta0, eta, ce);
                }
                ++stages_run;
                if (stop()) return false;
            }
        } else if (c.time_scheme == SWM_SSPRK2) {              // :80-110
            const Tp dts = Tp(c.Dt_Ds);
            for (int rki = 1; rki <= c.RKs; ++rki) {
                if (rki > 1) ghost_points_eta(eta1);
                if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
                const Arr<T> &e = rhs_view(eta1, etar);
                rhs(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                axb(u1, dts, du);
                axb(v1, dts, dv);
                ghost_points_uv(u1, v1);
                continuity(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                axb(eta1, dts, deta);
                ++stages_run;
                if (stop() && rki < c.RKs) return false;
            }
            if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
            double a = 1.0 / c.RKs, b = (double)(c.RKs - 1) / c.RKs;
            if (g_mutation == MUT_SSPRK2_WEIGHTS) std::swap(a, b);
            cxayb(u0, a, u, b, u1);
            cxayb(v0, a, v, b, v1);
            cxayb(eta0, a, eta, b, eta1);
        } else if (c.time_scheme == SWM_SSPRK3) {              // :112-156
            const Tp dtn = Tp(c.ssp_Dt_Dn), dtnc = Tp(c.ssp_Dt_Dnc);
            for (int rki = 2; rki <= c.ssp_s + 1; ++rki) {
                if (rki > 2) ghost_points_eta(eta1);
                if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
                const Arr<T> &e = rhs_view(eta1, etar);
                rhs(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                const bool m8 = g_mutation == MUT_SSPRK3_KN;
                const double kna = m8 ? c.ssp_knb : c.ssp_kna, knb = m8 ? c.ssp_kna : c.ssp_knb;
                if (rki == c.ssp_kn) {
                    dxaybzc(u1, kna, u1, knb, u0, to_f64(dtnc), du);
                    dxaybzc(v1, kna, v1, knb, v0, to_f64(dtnc), dv);
                } else {
                    axb(u1, dtn, du);
                    axb(v1, dtn, dv);
                }
                ghost_points_uv(u1, v1);
                continuity(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                if (rki == c.ssp_kn) dxaybzc(eta1, kna, eta1, knb, eta0, to_f64(dtnc), deta);
                else axb(eta1, dtn, deta);
                if (rki == c.ssp_mn || rki == c.ssp_s + 1) {
                    copyto(u0, u1);
                    copyto(v0, v1);
                    ghost_points_eta(eta1);
                    copyto(eta0, eta1);
                }
                ++stages_run;
                if (stop()) return false;
            }
            if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
        } else {                                               // 4SSPRK3 :158-195
            const Tp dt = Tp(c.Dt_D);
            for (int rki = 1; rki <= 4; ++rki) {
                if (rki > 1) ghost_points(u1, v1, eta1);
                if (tape) { tape->s.push_back(u1); tape->s.push_back(v1); tape->s.push_back(eta1); }
                const Arr<T> &e = rhs_view(eta1, etar);
                rhs(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                caxb(u0, u1, dt, du);
                caxb(v0, v1, dt, dv);
                cxab(u1, 0.5, u1, u0);
                cxab(v1, 0.5, v1, v0);
                ghost_points_uv(u1, v1);
                if (tape) { tape->p.push_back(u1); tape->p.push_back(v1); }
                continuity(rhs_view(u1, ur), rhs_view(v1, vr), e, t);
                caxb(eta0, eta1, dt, deta);
                cxab(eta1, 0.5, eta1, eta0);
                if (rki == 3) {
                    const bool m9 = g_mutation == MUT_4SSPRK3_THIRDS;
                    const double x = m9 ? 1.0 / 3.0 : 2.0 / 3.0, y = m9 ? 2.0 / 3.0 : 1.0 / 3.0;
                    cxayb(u1, x, u, y, u1);
                    cxayb(v1, x, v, y, v1);
                    cxayb(eta1, x, eta, y, eta1);
                } else if (rki == 4) {
                    copyto(u0, u1);
                    copyto(v0, v1);
                    copyto(eta0, eta1);
                }
                ++stages_run;
                if (stop()) return false;
            }
        }

        if (stop_stage > 0) return false;                      // all stages run: state before :198
        ghost_points(u0, v0, eta0);                            // :198
        if (tape) { tape->au = u0; tape->av = v0; tape->ae = eta0; }
        const Arr<T> &a0 = rhs_view(u0, ur), &b0 = rhs_view(v0, vr), &e0 = rhs_view(eta0, etar);
        if (c.dynamics == SWM_NONLINEAR && c.nstep_advcor > 0 && i % c.nstep_advcor == 0) {   // :208-211
            UVfluxes(a0, b0, e0);
            advection_coriolis(a0, b0);
        }
        if (i % c.nstep_diff == 0) {                           // :216-221
            bottom_drag(a0, b0, e0);
            diffusion(a0, b0);
            add_drag_diff_tendencies(u0, v0);
            ghost_points_uv(u0, v0);
        }
        // :226-228 tracer with the post-dissipation velocities
        if (tape) { tape->du = u0; tape->dv = v0; }
        tracer(i, rhs_view(u0, ur), rhs_view(v0, vr));
        if (tape) { tape->um = um; tape->vm = vm; }
        if (c.nout > 0 && i % c.nout == 0 && has_nan()) return true;  // feedback.jl:136-148
        copyto(u, u0);                                         // :240-242
        copyto(v, v0);
        copyto(eta, eta0);
        return false;
    }
};

// ------------------------------------------------------------------ C entry points

struct OracleStatic {
    const void *H, *Fx, *Fy, *f_q, *eta_ref, *Feta, *f_u, *f_v;
};

template <class T, class Tp, class Tc>
static int run_typed(const swm_config *c, const OracleStatic *s, void *u, void *v, void *eta, void *sst,
                     int64_t nsteps, int64_t i0, int64_t t0, int32_t stop_stage, void **rk_out,
                     int64_t *steps_done, int64_t *nan_step) {
    Model<T, Tp, Tc> M(*c);
    load_arr(M.H, s->H);
    load_arr(M.Fx, s->Fx);
    load_arr(M.Fy, s->Fy);
    load_arr(M.f_q, s->f_q);
    if (s->eta_ref) load_arr(M.eta_ref, s->eta_ref);
    if (s->Feta) load_arr(M.Feta, s->Feta);
    if (c->dynamics == SWM_LINEAR) {
        if (!s->f_u || !s->f_v) throw ShapeError();
        load_arr(M.f_u, s->f_u);
        load_arr(M.f_v, s->f_v);
    }
    load_arr(M.u, u);
    load_arr(M.v, v);
    load_arr(M.eta, eta);
    load_arr(M.sst, sst);
    if (c->scale_log2 != 0) {                   // u, v and their forcing carried as s·u, s·v
        for (auto *A : {&M.Fx, &M.Fy, &M.f_q})
            for (auto &x : A->a) x = x * T(M.sc);
        for (auto *A : {&M.u, &M.v})
            for (auto &x : A->a) x = x * Tp(M.sc);
    }
    M.init();
    *steps_done = 0;
    *nan_step = 0;
    int64_t t = t0;
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t i = i0 + k;
        const bool last = k == nsteps - 1;
        const int ss = last ? stop_stage : 0;
        if (M.step(i, t, ss)) {
            *nan_step = i;
            break;
        }
        *steps_done = k + 1;
        if (ss > 0) break;
        t += c->dtint;
    }
    if (c->scale_log2 != 0)
        for (auto *A : {&M.u, &M.v, &M.u0, &M.v0, &M.u1, &M.v1, &M.du, &M.dv})
            for (auto &x : A->a) x = x * Tp(M.sci);
    store_arr(M.u, u);
    store_arr(M.v, v);
    store_arr(M.eta, eta);
    store_arr(M.sst, sst);
    if (rk_out) {
        Arr<Tp> *arrs[9] = {&M.u0, &M.v0, &M.eta0, &M.u1, &M.v1, &M.eta1, &M.du, &M.dv, &M.deta};
        for (int k = 0; k < 9; ++k)
            if (rk_out[k]) store_arr(*arrs[k], rk_out[k]);
    }
    return 0;
}

// One operator of the step on a state as given (Tprog storage, Julia layout, in place):
//   op 0  ghost_points! of (u, v, η) and of sst (what uploading a state applies,
//         time_integration.jl:48 and ghost_points.jl:136-199)
//   op 1  bottom_drag!, diffusion!, add_drag_diff_tendencies!, ghost_points_uv! of
//         (u, v, η) as u0, v0, η0 (time_integration.jl:216-221); aux[0..8] (nullable)
//         receive Bu, Bv, LLu1, LLu2, LLv1, LLv2, DT, DS, νSmag as float64
//   op 2  tracer! at step i with (u, v) as the post-dissipation velocities
//         (tracer_advection.jl:1-44); aux[0], aux[1] = um, vm (values of format T in
//         float64, in place)
template <class T, class Tp, class Tc>
static int op_typed(const swm_config *c, const OracleStatic *s, int op, void *u, void *v, void *eta, void *sst, int64_t i,
                    double **aux) {
    Model<T, Tp, Tc> M(*c);
    load_arr(M.H, s->H);
    load_arr(M.Fx, s->Fx);
    load_arr(M.Fy, s->Fy);
    load_arr(M.f_q, s->f_q);
    load_arr(M.u0, u);
    load_arr(M.v0, v);
    load_arr(M.eta0, eta);
    load_arr(M.sst, sst);
    const double sc = M.sc, sci = M.sci;
    for (auto *A : {&M.u0, &M.v0})
        for (auto &x : A->a) x = x * Tp(sc);
    auto get_aux = [&](Arr<T> &A, const double *p) {
        for (size_t k = 0; k < A.size(); ++k) A.a[k] = cvt<T>(p[k]) * T(sc);
    };
    auto put_aux = [&](const Arr<T> &A, double *p, double f) {
        if (p) for (size_t k = 0; k < A.size(); ++k) p[k] = to_f64(A.a[k]) * f;
    };
    if (op == 0) {
        M.ghost_points(M.u0, M.v0, M.eta0);
        M.ghost_sst_any(M.sst);
    } else if (op == 1) {
        const Arr<T> &a0 = M.rhs_view(M.u0, M.ur), &b0 = M.rhs_view(M.v0, M.vr), &e0 = M.rhs_view(M.eta0, M.etar);
        M.bottom_drag(a0, b0, e0);
        M.diffusion(a0, b0);
        M.add_drag_diff_tendencies(M.u0, M.v0);
        M.ghost_points_uv(M.u0, M.v0);
        if (aux) {
            const Arr<T> *A[9] = {&M.Bu, &M.Bv, &M.LLu1, &M.LLu2, &M.LLv1, &M.LLv2, &M.DT, &M.DS, &M.nuSmag};
            for (int k = 0; k < 9; ++k) put_aux(*A[k], aux[k], k < 6 ? sci : 1.0);
        }
    } else if (op == 2) {
        if (!aux || !aux[0] || !aux[1]) return -1;
        get_aux(M.um, aux[0]);
        get_aux(M.vm, aux[1]);
        M.tracer(i, M.rhs_view(M.u0, M.ur), M.rhs_view(M.v0, M.vr));
        put_aux(M.um, aux[0], sci);
        put_aux(M.vm, aux[1], sci);
    } else {
        return -1;
    }
    for (auto *A : {&M.u0, &M.v0})
        for (auto &x : A->a) x = x * Tp(sci);
    store_arr(M.u0, u);
    store_arr(M.v0, v);
    store_arr(M.eta0, eta);
    store_arr(M.sst, sst);
    return 0;
}

// Calls fn(T(), Tp(), Tc()) with the number types the config names: T = Tprog in {Float64,
// Float32} with any Tcomm, or T in {Float16, BFloat16} over Tprog = Float32 with Tcomm in
// {T, Float32, Float64} (the formats the library runs); -1 for any other combination.
template <class Fn> static int with_types(const swm_config *c, Fn &&fn) {
    auto comm = [&](auto t, auto tp) -> int {
        switch (c->Tcomm) {
        case SWM_F64: return fn(t, tp, double());
        case SWM_F32: return fn(t, tp, float());
        case SWM_F16: return fn(t, tp, f16());
        case SWM_BF16: return fn(t, tp, bf16());
        }
        return -1;
    };
    auto low = [&](auto t) -> int {                 // T narrower than Tprog = Float32
        using TT = decltype(t);
        if (c->Tcomm == c->T) return fn(t, float(), TT());
        if (c->Tcomm == SWM_F32) return fn(t, float(), float());
        if (c->Tcomm == SWM_F64) return fn(t, float(), double());
        return -1;
    };
    if (c->T == c->Tprog && c->T == SWM_F64) return comm(double(), double());
    if (c->T == c->Tprog && c->T == SWM_F32) return comm(float(), float());
    if (c->Tprog == SWM_F32 && c->T == SWM_F16) return c->t_arith32 ? low(f16w()) : low(f16());
    if (c->Tprog == SWM_F32 && c->T == SWM_BF16) return c->t_arith32 ? low(bf16w()) : low(bf16());
    return -1;
}

template <class T>
static int tape_typed(const swm_config *c, const double *const *st, int64_t i, int64_t t, const double *const *x,
                      double *const *out) {
    Model<T, T, T> M(*c);
    typename Model<T, T, T>::Tape tp;
    auto in = [](Arr<T> &A, const double *p) {
        if (p) for (size_t k = 0; k < A.size(); ++k) A.a[k] = (T)p[k];
    };
    in(M.H, st[0]); in(M.Fx, st[1]); in(M.Fy, st[2]); in(M.f_q, st[3]); in(M.eta_ref, st[4]); in(M.Feta, st[5]);
    if (c->dynamics == SWM_LINEAR) { in(M.f_u, st[6]); in(M.f_v, st[7]); }
    in(M.u, x[0]); in(M.v, x[1]); in(M.eta, x[2]); in(M.sst, x[3]); in(M.um, x[4]); in(M.vm, x[5]);
    M.init();
    M.tape = &tp;
    M.step(i, t, 0);
    int k = 0;
    auto put = [&](const Arr<T> &A) {
        if (out[k]) for (size_t q = 0; q < A.size(); ++q) out[k][q] = (double)A.a[q];
        ++k;
    };
    for (auto &A : tp.s) put(A);
    for (auto &A : tp.p) put(A);
    put(tp.au); put(tp.av); put(tp.ae);
    put(tp.du); put(tp.dv);
    put(tp.um); put(tp.vm);
    put(M.u); put(M.v); put(M.eta); put(M.sst);
    return 0;
}

extern "C" {

// Run nsteps steps of the time loop from Prog (u, v, eta, sst; format Tprog, Julia layout,
// updated in place).  rk_out (nullable, 9 nullable pointers) receives u0 v0 eta0 u1 v1 eta1
// du dv deta of the last step.  Returns 0, or -1 on a shape/option error or an index
// the reference would throw on (NaN in the tracer, departure outside the arrays).
int swmo_run(const swm_config *c, const void *H, const void *Fx, const void *Fy, const void *f_q,
             const void *eta_ref, const void *Feta, const void *f_u, const void *f_v, void *u, void *v, void *eta,
             void *sst, int64_t nsteps, int64_t i0, int64_t t0, int32_t stop_stage, void **rk_out, int64_t *steps_done,
             int64_t *nan_step) {
    OracleStatic s{H, Fx, Fy, f_q, eta_ref, Feta, f_u, f_v};
    try {
        return with_types(c, [&](auto t, auto tp, auto tc) {
            return run_typed<decltype(t), decltype(tp), decltype(tc)>(c, &s, u, v, eta, sst, nsteps, i0, t0, stop_stage,
                                                                      rk_out, steps_done, nan_step);
        });
    } catch (ShapeError &) {
        return -1;
    }
}

// tracer! alone (tracer_advection.jl:1-44) at step i for T = Tprog = double: u, v are the
// post-dissipation velocities, um/vm the mid-point snapshot (updated in place when due),
// sst the tracer (updated in place).  Closed-form pins of departure points and adv_sst!.
int swmo_tracer(const swm_config *c, int64_t i, const double *u, const double *v, double *um, double *vm, double *sst) {
    try {
        Model<double, double, double> M(*c);
        std::memcpy(M.um.a.data(), um, sizeof(double) * M.um.size());
        std::memcpy(M.vm.a.data(), vm, sizeof(double) * M.vm.size());
        std::memcpy(M.sst.a.data(), sst, sizeof(double) * M.sst.size());
        Arr<double> uu(M.u.m, M.u.n), vv(M.v.m, M.v.n);
        std::memcpy(uu.a.data(), u, sizeof(double) * uu.size());
        std::memcpy(vv.a.data(), v, sizeof(double) * vv.size());
        M.tracer(i, uu, vv);
        std::memcpy(um, M.um.a.data(), sizeof(double) * M.um.size());
        std::memcpy(vm, M.vm.a.data(), sizeof(double) * M.vm.size());
        std::memcpy(sst, M.sst.a.data(), sizeof(double) * M.sst.size());
    } catch (ShapeError &) {
        return -1;
    }
    return 0;
}

// The PV advection terms alone for T = Float64 (PV_advection.jl:22-146 with the q
// interpolations of rhs.jl:117-128): from q at the corners (nx+1, ny+1) and the mass fluxes U
// (nx+1, ny+2), V (nx+2, ny+1) given as they are, qhv (nux, nuy) and qhu (nvx, nvy) of the
// configured scheme.  Pins the schemes' conservation properties on fluxes chosen freely.
int swmo_pv_advection(const swm_config *c, const double *q, const double *U, const double *V, double *qhv,
                      double *qhu) {
    try {
        Model<double, double, double> M(*c);
        std::memcpy(M.q.a.data(), q, sizeof(double) * M.q.size());
        std::memcpy(M.U.a.data(), U, sizeof(double) * M.U.size());
        std::memcpy(M.V.a.data(), V, sizeof(double) * M.V.size());
        M.PV_combinations();
        M.PV_advection();
        std::memcpy(qhv, M.qhv.a.data(), sizeof(double) * M.qhv.size());
        std::memcpy(qhu, M.qhu.a.data(), sizeof(double) * M.qhu.size());
    } catch (ShapeError &) {
        return -1;
    }
    return 0;
}

// ------------------------------------------------------------------ tangent-linear model
// The model on Dual (T = Tprog = Tcomm = Float64 plus one directional derivative).  Arrays
// are float64 in Julia layout; x / dx are input values / tangents, y / dy receive output
// values / tangents (dx, dy entries may be NULL for zero / unwanted tangents).
//   op 0  nsteps steps of the time loop (time_integration.jl:45-243) from (u, v, η, sst)
//         after the ghost copies swm_upload_prog applies; x, y = (u, v, η, sst)
//   op 1  rhs! and continuity! of a classic RK stage (rhs.jl:2-51, continuity.jl:72-98) at
//         the stage input as given (no ghost copies); x = (u, v, η), y = (du, dv, dη) with the
//         shapes of u, v, η
//   op 2  bottom_drag!, diffusion!, add_drag_diff_tendencies! and ghost_points_uv!
//         (time_integration.jl:216-221) of (u0, v0, η0) as given; x = (u0, v0, η0), y = (u0, v0)
//   op 3  tracer! at step i0 (tracer_advection.jl:1-44); x = (u, v, um, vm, sst), y = (um, vm, sst)
// dFx (nullable): a tangent of the zonal wind stress Fx (nux, nuy), for any op.
static void load_dual(Arr<Dual> &A, const double *v, const double *d) {
    for (size_t k = 0; k < A.size(); ++k) A.a[k] = Dual(v[k], d ? d[k] : 0.0);
}
static void store_dual(const Arr<Dual> &A, double *v, double *d) {
    for (size_t k = 0; k < A.size(); ++k) {
        if (v) v[k] = A.a[k].v;
        if (d) d[k] = A.a[k].d;
    }
}

int swmo_tlm(const swm_config *c, const double *H, const double *Fx, const double *Fy, const double *f_q,
             const double *eta_ref, const double *Feta, const double *f_u, const double *f_v, int32_t op,
             int64_t nsteps, int64_t i0, int64_t t0,
             const double *const *x, const double *const *dx, double *const *y, double *const *dy, const double *dFx) {
    if (c->scale_log2 != 0 || c->compensated) return -1;   // the tangent-linear model is of the plain step
    auto dxk = [&](int k) { return dx ? dx[k] : nullptr; };
    auto dyk = [&](int k) { return dy ? dy[k] : nullptr; };
    try {
        Model<Dual, Dual, Dual> M(*c);
        load_dual(M.H, H, nullptr);
        load_dual(M.Fx, Fx, dFx);
        load_dual(M.Fy, Fy, nullptr);
        load_dual(M.f_q, f_q, nullptr);
        if (eta_ref) load_dual(M.eta_ref, eta_ref, nullptr);
        if (Feta) load_dual(M.Feta, Feta, nullptr);
        if (c->dynamics == SWM_LINEAR) {
            if (!f_u || !f_v) return -1;
            load_dual(M.f_u, f_u, nullptr);
            load_dual(M.f_v, f_v, nullptr);
        }
        if (op == 0) {
            load_dual(M.u, x[0], dxk(0));
            load_dual(M.v, x[1], dxk(1));
            load_dual(M.eta, x[2], dxk(2));
            load_dual(M.sst, x[3], dxk(3));
            M.ghost_points(M.u, M.v, M.eta);
            M.ghost_sst_any(M.sst);
            M.init();
            int64_t t = t0;
            for (int64_t k = 0; k < nsteps; ++k, t += c->dtint) M.step(i0 + k, t, 0);
            store_dual(M.u, y[0], dyk(0));
            store_dual(M.v, y[1], dyk(1));
            store_dual(M.eta, y[2], dyk(2));
            store_dual(M.sst, y[3], dyk(3));
        } else if (op == 1) {
            Arr<Dual> uu(M.u.m, M.u.n), vv(M.v.m, M.v.n), ee(M.eta.m, M.eta.n);
            load_dual(uu, x[0], dxk(0));
            load_dual(vv, x[1], dxk(1));
            load_dual(ee, x[2], dxk(2));
            M.rhs(uu, vv, ee, t0);
            M.continuity(uu, vv, ee, t0);
            store_dual(M.du, y[0], dyk(0));
            store_dual(M.dv, y[1], dyk(1));
            store_dual(M.deta, y[2], dyk(2));
        } else if (op == 2) {
            load_dual(M.u0, x[0], dxk(0));
            load_dual(M.v0, x[1], dxk(1));
            load_dual(M.eta0, x[2], dxk(2));
            M.bottom_drag(M.u0, M.v0, M.eta0);
            M.diffusion(M.u0, M.v0);
            M.add_drag_diff_tendencies(M.u0, M.v0);
            M.ghost_points_uv(M.u0, M.v0);
            store_dual(M.u0, y[0], dyk(0));
            store_dual(M.v0, y[1], dyk(1));
        } else if (op == 3) {
            Arr<Dual> uu(M.u.m, M.u.n), vv(M.v.m, M.v.n);
            load_dual(uu, x[0], dxk(0));
            load_dual(vv, x[1], dxk(1));
            load_dual(M.um, x[2], dxk(2));
            load_dual(M.vm, x[3], dxk(3));
            load_dual(M.sst, x[4], dxk(4));
            M.tracer(i0, uu, vv);
            store_dual(M.um, y[0], dyk(0));
            store_dual(M.vm, y[1], dyk(1));
            store_dual(M.sst, y[2], dyk(2));
        } else {
            return -1;
        }
    } catch (ShapeError &) {
        return -1;
    }
    return 0;
}

// One step at step i, time t from the state (u, v, η, sst, um, vm) as given (Prog with its
// ghost copies; um, vm the tracer's mid-point velocities), T = Tprog = Tcomm = Float64 or
// Float32 (c->T; arrays are float64 holding values of that format), returning its
// intermediate states: out[3k..3k+2] = the input (u1, v1, η1) of RK stage k+1 (k < RKo) --
// under SSPRK2 (SSPRK3) of stage k+1 for k < RKs (s), then u1, v1, η1 after the last stage;
// under 4SSPRK3 of stage k+1 for k < 4, followed by u1, v1 of each of the 4 stages after its
// ghost_points_uv! (:178) -- then u0, v0, η0 after ghost_points! (:198), then u0, v0 after
// the dissipation, then um, vm after the step, then the new u, v, η, sst.  Every scheme.
int swmo_tape(const swm_config *c, const double *H, const double *Fx, const double *Fy, const double *f_q,
              const double *eta_ref, const double *Feta, const double *f_u, const double *f_v, int64_t i, int64_t t,
              const double *u, const double *v, const double *eta, const double *sst, const double *um,
              const double *vm, double *const *out) {
    if (c->T != c->Tprog || c->Tcomm != c->Tprog ||
        c->scale_log2 != 0 || c->compensated)
        return -1;
    if (c->dynamics == SWM_LINEAR && (!f_u || !f_v)) return -1;
    const double *st[8] = {H, Fx, Fy, f_q, eta_ref, Feta, f_u, f_v};
    const double *x[6] = {u, v, eta, sst, um, vm};
    try {
        if (c->T == SWM_F64) return tape_typed<double>(c, st, i, t, x, out);
        if (c->T == SWM_F32) return tape_typed<float>(c, st, i, t, x, out);
    } catch (ShapeError &) {
        return -1;
    }
    return -1;
}

// Single-operator entry points on float64 arrays (closed-form pins of the stencils).
// op: 0 ∂x, 1 ∂y, 2 ∇², 3 Ix, 4 Iy, 5 Ixy.  `in` is (m_in, n_in); `out` gets the
// reference's output shape.
int swmo_stencil(int op, const double *in, int m_in, int n_in, double *out) {
    try {
        Arr<double> a(m_in, n_in), o;
        std::memcpy(a.a.data(), in, sizeof(double) * a.size());
        switch (op) {
        case 0: o = Arr<double>(m_in - 1, n_in); dx(o, a); break;
        case 1: o = Arr<double>(m_in, n_in - 1); dy(o, a); break;
        case 2: o = Arr<double>(m_in - 2, n_in - 2); lap(o, a); break;
        case 3: o = Arr<double>(m_in - 1, n_in); Ix(o, a); break;
        case 4: o = Arr<double>(m_in, n_in - 1); Iy(o, a); break;
        case 5: o = Arr<double>(m_in - 1, n_in - 1); Ixy(o, a); break;
        default: return -1;
        }
        std::memcpy(out, o.a.data(), sizeof(double) * o.size());
    } catch (ShapeError &) {
        return -1;
    }
    return 0;
}

// One operator of the step (see op_typed) in the config's number formats.  Returns 0, or
// -1 on a shape/option error or an index the reference would throw on.
int swmo_op(const swm_config *c, const void *H, const void *Fx, const void *Fy, const void *f_q, int32_t op, void *u,
            void *v, void *eta, void *sst, int64_t i, double **aux) {
    OracleStatic s{H, Fx, Fy, f_q, nullptr, nullptr, nullptr, nullptr};
    try {
        return with_types(c, [&](auto t, auto tp, auto tc) {
            return op_typed<decltype(t), decltype(tp), decltype(tc)>(c, &s, op, u, v, eta, sst, i, aux);
        });
    } catch (ShapeError &) {
        return -1;
    }
}

// Plant mistake k (the MUT_* list) in the restatement, 0 to remove it; returns the previous.
int swmo_set_mutation(int32_t k) {
    const int prev = g_mutation;
    g_mutation = k;
    g_wide_rounds = k == MUT_WIDE_ROUNDS;
    return prev;
}

// Format conversions of the oracle (Julia's T(x)), for checking the emulated formats.
// fmt: SWM_F16 or SWM_BF16; in float64, out float64 values of the rounded numbers.
int swmo_round(int fmt, const double *in, double *out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        if (fmt == SWM_F16) out[k] = (double)f16(in[k]).v;
        else if (fmt == SWM_BF16) out[k] = (double)bf16(in[k]).v;
        else if (fmt == SWM_F32) out[k] = (double)(float)in[k];
        else out[k] = in[k];
    }
    return 0;
}

// binary op in an emulated format: op 0 +, 1 -, 2 *, 3 /, 4 sqrt(a)
int swmo_arith(int fmt, int op, const double *a, const double *b, double *out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        if (fmt == SWM_F16) {
            f16 x = f16(a[k]), y = f16(b[k]), r;
            r = op == 0 ? x + y : op == 1 ? x - y : op == 2 ? x * y : op == 3 ? x / y : num_sqrt(x);
            out[k] = r.v;
        } else {
            bf16 x = bf16(a[k]), y = bf16(b[k]), r;
            r = op == 0 ? x + y : op == 1 ? x - y : op == 2 ? x * y : op == 3 ? x / y : num_sqrt(x);
            out[k] = r.v;
        }
    }
    return 0;
}

}  // extern "C"
