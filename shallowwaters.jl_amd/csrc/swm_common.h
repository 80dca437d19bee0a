// swm_common.h -- number formats, device layout, field access, cross-lane moves, ghost
// copies and wave tiling shared by the gfx950 kernels of the ShallowWaters.jl time step.
//
// Device layout.  Every field, whatever its staggering, lives in one padded 2-D layout
// indexed by T-cell coordinates (x, y): element (x, y) sits at element offset
// (x + OX) + (y + OY)*P of its allocation.  A field's staggering is implicit:
//   T-point  (η, h, H, p, ...)  centre of cell (x, y)
//   U-point  (u, U, du, Fx ...) right face of cell (x, y)
//   V-point  (v, V, dv, Fy ...) top face of cell (x, y)
//   Q-point  (q, f_q, ...)      top-right corner of cell (x, y)
// so every reference stencil becomes a fixed offset pattern in {-1,0,1}^2 and the
// reference's ep-dependent index shifts disappear into the mapping Julia index <-> (x, y):
//   u[ii,jj] -> (ii-2-ep, jj-2)   v[ii,jj] -> (ii-2, jj-2)   η[ii,jj] -> (ii-1, jj-1)
//   q[ii,jj] -> (ii-1, jj-1)      sst[ii,jj] -> (ii-1, jj)
// The padding (x from -64, y from -4) keeps every lane's access in bounds; padding is zero
// and never written.  All fields share the element pitch P; their byte pitch follows their
// storage format.
//
// Number formats (src/default_parameters.jl:3-6).  T is the format of the right-hand side,
// Tprog that of the prognostic and Runge-Kutta variables, Tcomm that of ghost copies.  A
// kernel is instantiated for a bundle Fmt<R, Q, ST>: R the arithmetic type of T, Q that of
// Tprog, ST the storage of T-format fields (static fields, the stage input u1, the tracer
// snapshots).  Binary16 and bfloat16 arithmetic is carried in float registers and rounded
// after every operation (lo<K>): float has at least 2p+2 significand bits for both, so each
// + - * / sqrt is the correctly rounded result of the narrow format, as in Julia.
//
// Kernel shape.  One wavefront owns a strip of 64 consecutive x (lane = x - xs) and a
// segment of rows, and marches upward in y keeping the few rows each stencil needs in
// registers.  Neighbours in x come from the adjacent lanes (cross-lane moves, no LDS, no
// barriers); every stencil of the step has an x-radius of at most 2, so lanes 2..61 of a
// strip produce output and strips advance by 60 columns.  Each output is written once,
// together with the ghost copies the reference's ghost_points! makes from it.
//
// Arithmetic follows the reference term by term in the same order (compiled with
// -ffp-contract=off; IEEE division and sqrt), so results are bitwise those of the CPU
// restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace swm {

constexpr int OX = 64;       // left padding in x (elements)
constexpr int OY = 4;        // bottom padding in y (rows)
constexpr int WAVE = 64;
constexpr int HALO = 2;      // x-radius of every fused stencil
constexpr int WOUT = WAVE - 2 * HALO;   // output columns per strip
constexpr int WAVES_PER_BLOCK = 4;
constexpr int NXCD = 8;      // accelerator complex dies of an MI355X, each with its own L2

// ------------------------------------------------------------------ number formats

enum { FMT_F64 = 0, FMT_F32 = 1, FMT_F16 = 2, FMT_BF16 = 3 };   // the SWM_F64 ... codes
// the library's bundles of Float16 / BFloat16 values with float arithmetic (swm_config.t_arith32)
enum { FMT_F16A = 4, FMT_BF16A = 5 };

struct h16 { uint16_t b; };  // storage of an IEEE binary16 value
struct b16 { uint16_t b; };  // storage of a bfloat16 value

__device__ __forceinline__ float round_f16(float x) { return (float)(_Float16)x; }
// bfloat16 rounding (RNE) of x, as a float: converting the pair (0, x) packs bf16(x) into the
// high half over a zero low half, which is the bit pattern of the float bf16(x) itself, so
// the conversion needs no shift back
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float round_bf16(float x) {
    return __builtin_bit_cast(float, __builtin_convertvector((f32x2){0.0f, x}, bf16x2));
}

// float64 -> float32 rounding to odd: a second rounding to binary16 / bfloat16 is then
// the correctly rounded result from float64
__device__ __forceinline__ float round_odd_f32(double x) {
    float f = (float)x;
    if ((double)f != x && __builtin_isfinite(f)) {
        uint32_t b = __builtin_bit_cast(uint32_t, f);
        if (__builtin_fabs((double)f) > __builtin_fabs(x)) b -= 1u;
        f = __builtin_bit_cast(float, b | 1u);
    }
    return f;
}

// Arithmetic in binary16 (K = FMT_F16) or bfloat16 (K = FMT_BF16): a float register holding
// a value of the format, rounded to it after every operation.  Binary16 + - * use the
// native half-precision instructions (correctly rounded: the same bits as the float
// operation rounded once, since float carries more than 2p+2 significand bits); the
// compiler then drops the conversions between consecutive half operations.  Division and
// sqrt go through float and one rounding.
template <int K> struct lo {
    float v;
    __device__ __forceinline__ static float rnd(float x) { return K == FMT_F16 ? round_f16(x) : round_bf16(x); }
    __device__ __forceinline__ static lo raw(float x) { lo r; r.v = x; return r; }
    __device__ __forceinline__ static _Float16 h(float x) { return (_Float16)x; }
    lo() = default;
    __device__ __forceinline__ lo(int i) : v(rnd((float)i)) {}
    // exact constants only (0.5, 0.25, 1.0 ...); inexact ones come pre-rounded from the host
    __device__ __forceinline__ explicit lo(double d) : v(rnd((float)d)) {}
    __device__ __forceinline__ friend lo operator+(lo a, lo b) {
        if constexpr (K == FMT_F16) return raw((float)(h(a.v) + h(b.v)));
        else return raw(rnd(a.v + b.v));
    }
    __device__ __forceinline__ friend lo operator-(lo a, lo b) {
        if constexpr (K == FMT_F16) return raw((float)(h(a.v) - h(b.v)));
        else return raw(rnd(a.v - b.v));
    }
    __device__ __forceinline__ friend lo operator*(lo a, lo b) {
        if constexpr (K == FMT_F16) return raw((float)(h(a.v) * h(b.v)));
        else return raw(rnd(a.v * b.v));
    }
    __device__ __forceinline__ friend lo operator/(lo a, lo b) { return raw(rnd(a.v / b.v)); }
    __device__ __forceinline__ friend lo operator-(lo a) { return raw(-a.v); }
    __device__ __forceinline__ lo &operator+=(lo b) { *this = *this + b; return *this; }
    __device__ __forceinline__ friend bool operator<(lo a, lo b) { return a.v < b.v; }
    __device__ __forceinline__ friend bool operator>(lo a, lo b) { return a.v > b.v; }
};

// Binary16 / bfloat16 values with float arithmetic (K = FMT_F16 / FMT_BF16; a form of this
// library, not of the reference): every value that enters T (the stage input converted from
// Tprog, the static fields, the T constants) is rounded to the format, the arithmetic on
// them is float's, each operation rounded once to float, and its results are not rounded
// back to the format.
template <int K> struct ar {
    float v;
    __device__ __forceinline__ static float rnd(float x) { return K == FMT_F16 ? round_f16(x) : round_bf16(x); }
    __device__ __forceinline__ static ar raw(float x) { ar r; r.v = x; return r; }
    ar() = default;
    __device__ __forceinline__ ar(int i) : v(rnd((float)i)) {}
    __device__ __forceinline__ explicit ar(double d) : v(rnd((float)d)) {}
    __device__ __forceinline__ friend ar operator+(ar a, ar b) { return raw(a.v + b.v); }
    __device__ __forceinline__ friend ar operator-(ar a, ar b) { return raw(a.v - b.v); }
    __device__ __forceinline__ friend ar operator*(ar a, ar b) { return raw(a.v * b.v); }
    __device__ __forceinline__ friend ar operator/(ar a, ar b) { return raw(a.v / b.v); }
    __device__ __forceinline__ friend ar operator-(ar a) { return raw(-a.v); }
    __device__ __forceinline__ ar &operator+=(ar b) { v = v + b.v; return *this; }
    __device__ __forceinline__ friend bool operator<(ar a, ar b) { return a.v < b.v; }
    __device__ __forceinline__ friend bool operator>(ar a, ar b) { return a.v > b.v; }
};
template <class R> struct is_ar : std::false_type {};
template <int K> struct is_ar<ar<K>> : std::true_type {};

// sqrt, floor and the finiteness test of every arithmetic type (IEEE sqrt, correctly rounded)
__device__ __forceinline__ float nsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double nsqrt(double x) { return ::sqrt(x); }
template <int K> __device__ __forceinline__ lo<K> nsqrt(lo<K> a) { return lo<K>::raw(lo<K>::rnd(sqrtf(a.v))); }
template <int K> __device__ __forceinline__ ar<K> nsqrt(ar<K> a) { return ar<K>::raw(sqrtf(a.v)); }
template <int K> __device__ __forceinline__ ar<K> nfloor(ar<K> a) { return ar<K>::raw(floorf(a.v)); }
template <int K> __device__ __forceinline__ bool nfinite(ar<K> a) { return __builtin_isfinite(a.v); }
template <int K> __device__ __forceinline__ float wide(ar<K> x) { return x.v; }
__device__ __forceinline__ float nfloor(float x) { return floorf(x); }
__device__ __forceinline__ double nfloor(double x) { return ::floor(x); }
template <int K> __device__ __forceinline__ lo<K> nfloor(lo<K> a) { return lo<K>::raw(floorf(a.v)); }
__device__ __forceinline__ bool nfinite(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool nfinite(double x) { return __builtin_isfinite(x); }
template <int K> __device__ __forceinline__ bool nfinite(lo<K> a) { return __builtin_isfinite(a.v); }

// the arithmetic type's value as float / double (exact)
__device__ __forceinline__ double wide(double x) { return x; }
__device__ __forceinline__ float wide(float x) { return x; }
template <int K> __device__ __forceinline__ float wide(lo<K> x) { return x.v; }

// T(x) for an integer-valued x of the arithmetic type's range (|x| < 2^24)
template <class R> __device__ __forceinline__ R from_int(long i) {
    if constexpr (std::is_same_v<R, double>) return (double)i;
    else if constexpr (std::is_same_v<R, float>) return (float)i;
    else return R::raw(R::rnd((float)i));
}

// A bundle of formats: R arithmetic of T, Q arithmetic and storage of Tprog, ST storage of
// the T-format fields.
template <class R_, class Q_, class ST_> struct Fmt {
    using R = R_;
    using Q = Q_;
    using ST = ST_;
    static constexpr bool mixed = !std::is_same_v<R_, Q_>;
};
using FmtF64 = Fmt<double, double, double>;
using FmtF32 = Fmt<float, float, float>;
using FmtF16 = Fmt<lo<FMT_F16>, float, h16>;     // T = Float16, Tprog = Float32
using FmtBF16 = Fmt<lo<FMT_BF16>, float, b16>;   // T = BFloat16, Tprog = Float32
using FmtF16A = Fmt<ar<FMT_F16>, float, h16>;    // Float16 values, float arithmetic, Tprog = Float32
using FmtBF16A = Fmt<ar<FMT_BF16>, float, b16>;  // BFloat16 values, float arithmetic, Tprog = Float32
// storage of q at the periodic wrap column (a T intermediate): T's storage, float where T's
// arithmetic is float's
template <class F> using QES = std::conditional_t<is_ar<typename F::R>::value, float, typename F::ST>;

// Tprog(x) of a T value (exact widening) and T(x) of a Tprog value (one rounding)
template <class Q, class R> __device__ __forceinline__ Q toQ(R r) {
    if constexpr (std::is_same_v<Q, R>) return r;
    else return (Q)wide(r);
}
template <class R, class Q> __device__ __forceinline__ R toR(Q q) {
    if constexpr (std::is_same_v<Q, R>) return q;
    else if constexpr (std::is_same_v<R, double> || std::is_same_v<R, float>) return (R)q;
    else if constexpr (std::is_same_v<Q, double>) return R::raw(R::rnd(round_odd_f32(q)));
    else return R::raw(R::rnd(q));
}

// Tcomm(x) of a value held in format `own` (ghost_points.jl:65): a no-op unless Tcomm is
// narrower; the caller's format then holds it exactly.
__device__ __forceinline__ double comm_round(double x, int cm) {
    if (cm == FMT_F32) return (double)(float)x;
    if (cm == FMT_F16) return (double)round_f16(round_odd_f32(x));
    if (cm == FMT_BF16) return (double)round_bf16(round_odd_f32(x));
    return x;
}
__device__ __forceinline__ float comm_round(float x, int cm) {
    if (cm == FMT_F16) return round_f16(x);
    if (cm == FMT_BF16) return round_bf16(x);
    return x;
}
template <int K> __device__ __forceinline__ lo<K> comm_round(lo<K> x, int) { return x; }   // Tcomm is T or Tprog
// Tcomm is T or wider: T(Tcomm(x)) of a T value is x (as the reference's copy between equal formats)
template <int K> __device__ __forceinline__ ar<K> comm_round(ar<K> x, int) { return x; }

// ------------------------------------------------------------------ field access

// All fields of a context live in one device arena; a kernel addresses them through one
// buffer descriptor.  A field is named by the byte offset of its allocation in the arena
// (its origin), element (x, y) at origin + colb(x) + rowb(y): the column part is a
// loop-invariant VGPR, origin + row a scalar, so marching a row costs one scalar add.
struct ArenaRef {
    const void *base;
    uint32_t bytes;
};

struct Arena {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ explicit Arena(ArenaRef a)
        : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.base), 0, (int)a.bytes, 0x00020000)) {}
};

// Storage formats: the register value type V, loads and stores through the arena.
template <class S> struct Sto;
template <> struct Sto<double> {
    using V = double;
    __device__ __forceinline__ static V ld(const Arena &a, uint32_t vo, uint32_t so) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(a.rs, vo, so, 0));
    }
    __device__ __forceinline__ static void st(const Arena &a, V v, uint32_t vo, uint32_t so) {
        using u2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(a.rs, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), a.rs, vo, so, 0);
    }
};
template <> struct Sto<float> {
    using V = float;
    __device__ __forceinline__ static V ld(const Arena &a, uint32_t vo, uint32_t so) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(a.rs, vo, so, 0));
    }
    __device__ __forceinline__ static void st(const Arena &a, V v, uint32_t vo, uint32_t so) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), a.rs, vo, so, 0);
    }
};
template <> struct Sto<h16> {
    using V = float;                     // a binary16 value, widened
    __device__ __forceinline__ static V ld(const Arena &a, uint32_t vo, uint32_t so) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(a.rs, vo, so, 0));
    }
    __device__ __forceinline__ static void st(const Arena &a, V v, uint32_t vo, uint32_t so) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, (_Float16)v), a.rs, vo, so, 0);
    }
};
template <> struct Sto<b16> {
    using V = float;                     // a bfloat16 value, widened
    __device__ __forceinline__ static V ld(const Arena &a, uint32_t vo, uint32_t so) {
        const uint32_t b = (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(a.rs, vo, so, 0);
        return __builtin_bit_cast(float, b << 16);
    }
    __device__ __forceinline__ static void st(const Arena &a, V v, uint32_t vo, uint32_t so) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(__builtin_bit_cast(uint32_t, v) >> 16), a.rs, vo, so, 0);
    }
};

// value of an arithmetic type to store in format S (rounding a wider value to S)
template <class S, class X> __device__ __forceinline__ typename Sto<S>::V enc(X x) {
    if constexpr (std::is_same_v<S, double>) return (double)wide(x);
    else if constexpr (std::is_same_v<S, float>) return (float)wide(x);
    else if constexpr (std::is_same_v<X, lo<FMT_F16>> || std::is_same_v<X, lo<FMT_BF16>>) return x.v;
    else if constexpr (std::is_same_v<S, h16>) return round_f16(wide(x));
    else return round_bf16(wide(x));
}
// arithmetic value R of a stored S value (exact for T-format storage, one rounding from Tprog)
template <class R, class S> __device__ __forceinline__ R dec(typename Sto<S>::V v) {
    if constexpr (std::is_same_v<S, h16> || std::is_same_v<S, b16>) {
        if constexpr (std::is_same_v<R, float> || std::is_same_v<R, double>) return (R)v;
        else return R::raw(v);
    } else return toR<R>(v);
}

// a field of the arena in storage format S: its origin and the arena it lives in
template <class S> struct Fld {
    using V = typename Sto<S>::V;
    const Arena *a;
    uint32_t o;
    __device__ __forceinline__ V ld(uint32_t vo, uint32_t so) const { return Sto<S>::ld(*a, vo, o + so); }
    template <class R> __device__ __forceinline__ R ldr(uint32_t vo, uint32_t so) const { return dec<R, S>(ld(vo, so)); }
    template <class X> __device__ __forceinline__ void st(X x, uint32_t vo, uint32_t so) const {
        Sto<S>::st(*a, enc<S>(x), vo, o + so);
    }
};

// byte offsets of column x and row y in the padded layout of a field stored as S
template <class S> __device__ __forceinline__ uint32_t colb(int x) { return (uint32_t)(x + OX) * (uint32_t)sizeof(S); }
template <class S> __device__ __forceinline__ uint32_t rowb(int y, int P) {
    return (uint32_t)(y + OY) * (uint32_t)P * (uint32_t)sizeof(S);
}

// ------------------------------------------------------------------ cross-lane moves
// value of lane+1 (x+1) and lane-1 (x-1); the edge lanes receive junk that only feeds
// the two halo lanes on each side.
// Whole-wave DPP shifts (a modifier on a VALU move, no LDS round trip):
// wave_shl:1 (0x130) gives lane i the value of lane i+1, wave_shr:1 (0x138) that of lane i-1.
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double dpp_f(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo_ = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi_ = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi_ << 32) | (uint32_t)lo_);
}
__device__ __forceinline__ float sh_xp(float v) { return dpp_f<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ float sh_xm(float v) { return dpp_f<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double sh_xp(double v) { return dpp_f<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double sh_xm(double v) { return dpp_f<DPP_WAVE_SHR1>(v); }
// binary16 values move as their 16-bit pattern (no conversion around the lane shift)
template <int CTRL> __device__ __forceinline__ float dpp_h(float v) {
    const int b = (int)__builtin_bit_cast(uint16_t, (_Float16)v);
    const int r = __builtin_amdgcn_update_dpp(0, b, CTRL, 0xF, 0xF, true);
    return (float)__builtin_bit_cast(_Float16, (uint16_t)r);
}
template <int K> __device__ __forceinline__ lo<K> sh_xp(lo<K> v) {
    if constexpr (K == FMT_F16) return lo<K>::raw(dpp_h<DPP_WAVE_SHL1>(v.v));
    else return lo<K>::raw(sh_xp(v.v));
}
template <int K> __device__ __forceinline__ lo<K> sh_xm(lo<K> v) {
    if constexpr (K == FMT_F16) return lo<K>::raw(dpp_h<DPP_WAVE_SHR1>(v.v));
    else return lo<K>::raw(sh_xm(v.v));
}
template <int K> __device__ __forceinline__ ar<K> sh_xp(ar<K> v) { return ar<K>::raw(sh_xp(v.v)); }
template <int K> __device__ __forceinline__ ar<K> sh_xm(ar<K> v) { return ar<K>::raw(sh_xm(v.v)); }

// ------------------------------------------------------------------ static fields

struct StaticRefs {
    uint32_t H;                        // T-point, 2-D
    uint32_t Fx, Fy, fq;               // U-, V-, Q-point; 1-D per row when x-invariant
    int Fx_2d, Fy_2d, fq_2d;           // 1: 2-D layout, 0: one value per row
    uint32_t eta_ref, Feta;            // T-point interior (continuity relax / forcing)
    uint32_t fu, fv;                   // f on the u- and v-grid (linear dynamics), layout of Fx, Fy
    int fu_2d, fv_2d;
};

// PV advection / Coriolis form of the right-hand side (rhs.jl:2-16, PV_advection.jl:21-27);
// ADV_KEEP (with ADV_AH or ADV_SAD): q and the kinetic part of the Bernoulli potential are
// read from the last update instead of formed (nstep_advcor > 0, rhs.jl:36-38)
enum { ADV_AH = 0, ADV_SAD = 1, ADV_LIN = 2, ADV_KEEP = 4 };
// a flag beside the right-hand side form of a fused pair: the first stage's output is also
// stored (the adjoint's taped step keeps every stage input)
enum { ADV_TAPE = 8 };

// byte offset of (x, y) in a static field that may be stored 1-D (one value per row)
template <class S> __device__ __forceinline__ uint32_t scolb(int x, int is2d) { return is2d ? colb<S>(x) : 0u; }
template <class S> __device__ __forceinline__ uint32_t srowb(int y, int P, int is2d) {
    return is2d ? rowb<S>(y, P) : (uint32_t)(y + OY) * (uint32_t)sizeof(S);
}

// continuity variant (continuity.jl:91-97), a runtime switch uniform over the launch
enum { CONT_ITSELF = 0, CONT_RELAX = 1, CONT_FORCING = 2 };

// ------------------------------------------------------------------ ghost copies on store
// ghost_points.jl:47-133: each interior value is written together with the halo cells the
// reference copies from it.  The common case (no ghost cell) costs one test of a flag
// precomputed per lane (x-copies) and a uniform row test (y-copies).  Values are in the
// format of the field the reference updates (Tprog, arithmetic type Q); the store rounds
// them to the field's storage S.  Periodic copies pass through Tcomm (cm).

// lanes whose u, v or η value is the source of an x-direction ghost copy
struct GhostLanes {
    bool u, v, e;
};
template <bool PER> __device__ __forceinline__ GhostLanes ghost_lanes(int x, int nx) {
    GhostLanes g;
    g.u = PER && (x == 0 || x == 1 || x == nx - 2 || x == nx - 1);
    g.v = x == 1 || x == 2 || x == nx - 1 || x == nx;
    g.e = x == 1 || x == nx;
    return g;
}

template <class S, bool PER, class Q>
__device__ __forceinline__ void ghost_u(const Fld<S> &b, int x, int y, Q w, int nx, int ny, int P, Q oma, int cm) {
    if (PER) {                                   // ghost_points_u_periodic!  :60-73
        const Q gw = comm_round(w, cm);
        const uint32_t so = rowb<S>(y, P);
        if (x == nx - 2) b.st(gw, colb<S>(-2), so);
        if (x == nx - 1) b.st(gw, colb<S>(-1), so);
        if (x == 0) b.st(gw, colb<S>(nx), so);
        if (x == 1) b.st(gw, colb<S>(nx + 1), so);
    } else {                                     // ghost_points_u_nonperiodic!  :47-57
        const Q gw = oma * w;
        const uint32_t vo = colb<S>(x);
        if (y == 1) b.st(gw, vo, rowb<S>(0, P));
        if (y == 2) b.st(gw, vo, rowb<S>(-1, P));
        if (y == ny) b.st(gw, vo, rowb<S>(ny + 1, P));
        if (y == ny - 1) b.st(gw, vo, rowb<S>(ny + 2, P));
    }
}

template <class S, bool PER, class Q>
__device__ __forceinline__ void ghost_v(const Fld<S> &b, int x, int y, Q w, int nx, int P, Q oma, int cm) {
    const uint32_t so = rowb<S>(y, P);
    if (PER) {                                   // ghost_points_v_periodic!  :89-99
        const Q gw = comm_round(w, cm);
        if (x == nx - 1) b.st(gw, colb<S>(-1), so);
        if (x == nx) b.st(gw, colb<S>(0), so);
        if (x == 1) b.st(gw, colb<S>(nx + 1), so);
        if (x == 2) b.st(gw, colb<S>(nx + 2), so);
    } else {                                     // ghost_points_v_nonperiodic!  :76-86
        const Q gw = oma * w;
        if (x == 1) b.st(gw, colb<S>(0), so);
        if (x == 2) b.st(gw, colb<S>(-1), so);
        if (x == nx) b.st(gw, colb<S>(nx + 1), so);
        if (x == nx - 1) b.st(gw, colb<S>(nx + 2), so);
    }
}

template <class S, bool PER, class Q>
__device__ __forceinline__ void ghost_e(const Fld<S> &b, int x, int y, Q w, int nx, int ny, int P, int cm) {
    // ghost_points_η_(non)periodic!  :102-133: x-copies first (through Tcomm when
    // periodic), then y-copies of whole rows
    const bool gy0 = y == 1, gy1 = y == ny;
    const uint32_t s0 = rowb<S>(0, P), s1 = rowb<S>(ny + 1, P), so = rowb<S>(y, P);
    const uint32_t vo = colb<S>(x);
    if (gy0) b.st(w, vo, s0);
    if (gy1) b.st(w, vo, s1);
    const Q gw = PER ? comm_round(w, cm) : w;
    if (x == (PER ? nx : 1)) {
        const uint32_t vg = colb<S>(0);
        b.st(gw, vg, so);
        if (gy0) b.st(gw, vg, s0);
        if (gy1) b.st(gw, vg, s1);
    }
    if (x == (PER ? 1 : nx)) {
        const uint32_t vg = colb<S>(nx + 1);
        b.st(gw, vg, so);
        if (gy0) b.st(gw, vg, s0);
        if (gy1) b.st(gw, vg, s1);
    }
}

// store of an interior value plus its ghost copies
template <class S, bool PER, class Q>
__device__ __forceinline__ void put_u(const Fld<S> &b, int x, uint32_t vo, int y, uint32_t so, Q w, bool glane, int nx,
                                      int ny, int P, Q oma, int cm) {
    b.st(w, vo, so);
    if (PER ? glane : (y <= 2 || y >= ny - 1)) ghost_u<S, PER>(b, x, y, w, nx, ny, P, oma, cm);
}
template <class S, bool PER, class Q>
__device__ __forceinline__ void put_v(const Fld<S> &b, int x, uint32_t vo, int y, uint32_t so, Q w, bool glane, int nx,
                                      int P, Q oma, int cm) {
    b.st(w, vo, so);
    if (glane) ghost_v<S, PER>(b, x, y, w, nx, P, oma, cm);
}
template <class S, bool PER, class Q>
__device__ __forceinline__ void put_e(const Fld<S> &b, int x, uint32_t vo, int y, uint32_t so, Q w, bool glane, int nx,
                                      int ny, int P, int cm) {
    b.st(w, vo, so);
    if (glane || y == 1 || y == ny) ghost_e<S, PER>(b, x, y, w, nx, ny, P, cm);
}

// ------------------------------------------------------------------ wave -> (strip, segment)

struct WaveTile {
    int lane, x, ya, yb;
    bool live;
};

// The launch's wave index.  Blocks are dispatched round-robin over the XCDs (block b to
// XCD b % NXCD).  The grid is a multiple of NXCD blocks; renumbering gives each XCD a
// contiguous run of tiles, so the halo rows two segments share sit in the same L2.  The
// index is uniform over the wave: readfirstlane makes it (and the rows) scalar.
__device__ __forceinline__ int wave_index() {
    const int per = (int)gridDim.x / NXCD;
    const int b = ((int)blockIdx.x % NXCD) * per + (int)blockIdx.x / NXCD;
    return b * WAVES_PER_BLOCK + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
}

// Tile w of a field: strips of WAVE lanes with an x-halo of H lanes on each side (outputs in
// lanes H..WAVE-H-1) times segments of L rows.
template <int H = HALO>
__device__ __forceinline__ WaveTile tile_of(int w, int nstrips, int nseg, int L, int ny) {
    WaveTile t;
    t.lane = (int)(threadIdx.x % WAVE);
    t.live = w >= 0 && w < nstrips * nseg;
    const int s = w % nstrips, g = w / nstrips;
    t.x = -H + (WAVE - 2 * H) * s + t.lane;
    t.ya = 1 + g * L;
    t.yb = min(t.ya + L, ny + 1);
    return t;
}

// A batched launch over the members of an ensemble: start[m] is the first tile of member
// m, start[nmem] the total.  Sets the member and the member's own tile index of tile w;
// false past the last tile.  Uniform over the wave (scalar loads).
__device__ __forceinline__ bool batch_locate(const int *__restrict__ start, int nmem, int w, int &m, int &wl) {
    if (w >= start[nmem]) return false;
    int lo = 0, hi = nmem;                 // start[lo] <= w < start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= w) lo = mid;
        else hi = mid;
    }
    m = lo;
    wl = w - start[lo];
    return true;
}

}  // namespace swm
