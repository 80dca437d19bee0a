// k_misc.h -- tracer advection, ghost points of an uploaded state, snapshot copies, NaN
// check and FP64 reductions (gfx950).
#pragma once

#include "swm_common.h"

namespace swm {

// element access through a plain pointer to a field stored as S (element (x, y) of a field
// whose (0,0) element p points at is p[x + y*P])
template <class S> __device__ __forceinline__ typename Sto<S>::V ldp(const S *p, long i) {
    if constexpr (std::is_same_v<S, h16>) return (float)__builtin_bit_cast(_Float16, p[i].b);
    else if constexpr (std::is_same_v<S, b16>) return __builtin_bit_cast(float, (uint32_t)p[i].b << 16);
    else return p[i];
}
template <class S> __device__ __forceinline__ void stp(S *p, long i, typename Sto<S>::V v) {
    if constexpr (std::is_same_v<S, h16>) p[i].b = __builtin_bit_cast(uint16_t, (_Float16)v);
    else if constexpr (std::is_same_v<S, b16>) p[i].b = (uint16_t)(__builtin_bit_cast(uint32_t, v) >> 16);
    else p[i] = v;
}

// ------------------------------------------------------------------ tracer (K3)
//
// tracer!  tracer_advection.jl:1-44 when i % nadvstep == 0: departure! (:52-85, with
// backtraj! :96-123 and interp_uv! :130-173 gathering the mid-point velocities um, vm
// interpolated to the T-grid on the fly), adv_sst! (:179-206), ghost_points_sst! and the
// copy back to sst.  One thread per T-cell.  idx(k,l) maps the reference's absolute
// indices of u_T/v_T/sst back to (x,y).  Arithmetic in T: u, v (= T(u0) after
// dissipation, time_integration.jl:226-228) and sst (Tprog) are rounded to T on load; um,
// vm are T snapshots; the result is written to sst in Tprog (exact).
template <class F> struct TracerArgs {
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    const Q *u, *v;                   // post-dissipation velocities (Prog of the next step)
    const ST *um, *vm;                // mid-point snapshots
    const Q *sst;                     // tracer before the advection
    Q *ssti;                          // after it (the other sst slot)
    R hdu, hdv, du_, dv_;             // ½dtadvu, ½dtadvv, dtadvu, dtadvv
    R jSST, SSTmin, sinv;
    int nx, ny, P, ep, consume, cm;
    int *bad;                         // bit 0: departure point NaN, bit 1: index outside the field
};

// the advection of T-cell (x, y), 1 <= x <= nx, 1 <= y <= ny
template <class F, bool PER>
__device__ __forceinline__ void tracer_cell(const TracerArgs<F> &A, int x, int y) {
    using R = typename F::R;
    using Q = typename F::Q;
    using ST = typename F::ST;
    const Q *u = A.u, *v = A.v, *sst = A.sst;
    const ST *um = A.um, *vm = A.vm;
    Q *ssti = A.ssti;
    const R hdu = A.hdu, hdv = A.hdv, du_ = A.du_, dv_ = A.dv_, jSST = A.jSST, SSTmin = A.SSTmin, sinv = A.sinv;
    const int nx = A.nx, ny = A.ny, P = A.P, ep = A.ep, consume = A.consume, cm = A.cm;
    int *bad = A.bad;
    const R half = R(0.5), one = R(1.0), zero = R(0.0);
    auto U = [&](int cx, int cy) { return toR<R>(u[cx + cy * P]); };
    auto V = [&](int cx, int cy) { return toR<R>(v[cx + cy * P]); };
    // Ix!(u_T,u): u_T(x,y) = ½(u(x) + u(x-1)); Iy!(v_T,v): v_T(x,y) = ½(v(y) + v(y-1))
    // (velocities carried scaled by s: the departure distances use u/s, v/s)
    const R uT = half * (U(x, y) + U(x - 1, y)) * sinv;
    const R vT = half * (V(x, y) + V(x, y - 1)) * sinv;
    const R xd0 = -hdu * uT, yd0 = -hdv * vT;
    // interp_uv! on um_T (x-index k <-> T-column k-1-ep, y-index l <-> row l-2) and
    // vm_T (k <-> column k-2, l <-> row l-1)
    auto interp = [&](const ST *f, bool is_u, R xx, R yy) -> R {
        const int muv = is_u ? nx + 2 + ep : nx + 4, nuv = is_u ? ny + 4 : ny + 2;
        const int xh1 = is_u ? 1 + ep : 2, xh2 = is_u ? 1 : 2, yh1 = is_u ? 2 : 1;
        const R fx = nfloor(xx), fy = nfloor(yy);
        if (!nfinite(fx) || !nfinite(fy)) { atomicOr(bad, 1); return zero; }
        const long xi = (long)wide(fx), yi = (long)wide(fy);
        long k = xi + x + xh1, l = yi + y + yh1;
        R x0 = xx - from_int<R>(xi), y0 = yy - from_int<R>(yi);
        if (PER) {
            if (k < 1 + xh1) k = k + muv - xh1 - xh2;
            else if (k > muv - 1) k = k - muv + xh1 + xh2;
        } else {
            if (k < 1 + xh1) { k = 1 + xh1; x0 = zero; }
            else if (k > muv - 1) { k = muv - 1; x0 = one; }
        }
        if (l < 1 + yh1) { l = 1 + yh1; y0 = zero; }
        else if (l > nuv - 1) { l = nuv - 1; y0 = one; }
        if (k < 1 || k + 1 > muv) { atomicOr(bad, 2); return zero; }
        auto F_ = [&](int cx, int cy) { return dec<R, ST>(ldp(f, cx + (long)cy * P)); };
        // value of f_T at reference index (kk, ll)
        auto fT = [&](long kk, long ll) -> R {
            if (is_u) { const int cx = (int)kk - 1 - ep, cy = (int)ll - 2; return half * (F_(cx, cy) + F_(cx - 1, cy)); }
            const int cx = (int)kk - 2, cy = (int)ll - 1;
            return half * (F_(cx, cy) + F_(cx, cy - 1));
        };
        const R f00 = fT(k, l), f10 = fT(k + 1, l), f01 = fT(k, l + 1), f11 = fT(k + 1, l + 1);
        return f00 * (one - x0) * (one - y0) + f10 * x0 * (one - y0) + f01 * (one - x0) * y0 + f11 * x0 * y0;
    };
    const R ui = interp(um, true, xd0, yd0) * sinv;
    const R vi = interp(vm, false, xd0, yd0) * sinv;
    const R xd = -du_ * ui, yd = -dv_ * vi;
    // adv_sst!  (sst index k <-> column k-1, l <-> row l; m = nx+2, n = ny, halo (1,0))
    const R fx = nfloor(xd), fy = nfloor(yd);
    if (!nfinite(fx) || !nfinite(fy)) { atomicOr(bad, 1); return; }
    const long xi = (long)wide(fx), yi = (long)wide(fy);
    const int m = nx + 2, n = ny;
    long k = xi + x + 1, l = yi + y;
    R x0 = xd - from_int<R>(xi), y0 = yd - from_int<R>(yi);
    if (PER) {
        if (k < 2) k = k + m - 2;
        else if (k > m - 1) k = k - m + 2;
    } else {
        if (k < 2) { k = 2; x0 = zero; }
        else if (k > m - 1) { k = m - 1; x0 = one; }
    }
    if (l < 1) { l = 1; y0 = zero; }
    else if (l > n - 1) { l = n - 1; y0 = one; }
    if (k < 1 || k + 1 > m) { atomicOr(bad, 2); return; }
    auto S = [&](long kk, long ll) -> R { return toR<R>(sst[(int)kk - 1 + (int)ll * P]); };
    R s = S(k, l) * (one - x0) * (one - y0) + S(k + 1, l) * x0 * (one - y0) + S(k, l + 1) * (one - x0) * y0 +
          S(k + 1, l + 1) * x0 * y0;
    if (consume) s += jSST * (SSTmin - s);       // tracer_consumption!  :262-275
    // ghost_points_sst!  ghost_points.jl:136-179 (halosstx=1, halossty=0) on the T array
    ssti[x + y * P] = toQ<Q>(s);
    const R gs = PER ? comm_round(s, cm) : s;
    if (x == (PER ? nx : 1)) ssti[y * P] = toQ<Q>(gs);
    if (x == (PER ? 1 : nx)) ssti[nx + 1 + y * P] = toQ<Q>(gs);
}

template <class F, bool PER>
__global__ void tracer_kernel(TracerArgs<F> A) {
    const int x = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x > A.nx) return;
    tracer_cell<F, PER>(A, x, y);
}

// ------------------------------------------------------------------ small kernels

// x ← x·f for a power of two f, n elements of storage S (velocities and their forcing
// carried scaled)
template <class S> __global__ void scale_kernel(S *p, size_t n, float f) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (std::is_same_v<S, double>) p[i] = p[i] * (double)f;
        else if constexpr (std::is_same_v<S, float>) p[i] = p[i] * f;
        else stp(p, (long)i, ldp(p, (long)i) * f);
    }
}

// ghost_points! of the uploaded state (ghost_points.jl:182-242), in the reference order:
// phase 0 writes the x-copies, phase 1 the y-copies (which read the x-copies for η).
// Fields in Tprog; periodic copies pass through Tcomm (cm).
template <class Q, bool PER>
__global__ void ghost_kernel(Q *u, Q *v, Q *e, Q *sst, int nx, int ny, int P, Q oma, int cm, int phase) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    auto c = [&](Q x) { return comm_round(x, cm); };
    if (phase == 0) {
        // rows of u (periodic: x-copies), v (x-copies), η (x-copies), sst (x-copies)
        const int y = k - 1;                     // y = -1 .. ny+2
        if (y < -1 || y > ny + 2) return;
        if (PER) {
            u[-2 + y * P] = c(u[nx - 2 + y * P]);
            u[-1 + y * P] = c(u[nx - 1 + y * P]);
            u[nx + y * P] = c(u[y * P]);
            u[nx + 1 + y * P] = c(u[1 + y * P]);
            if (y <= ny + 1) {
                v[-1 + y * P] = c(v[nx - 1 + y * P]);
                v[y * P] = c(v[nx + y * P]);
                v[nx + 1 + y * P] = c(v[1 + y * P]);
                v[nx + 2 + y * P] = c(v[2 + y * P]);
            }
            if (y >= 0 && y <= ny + 1) {
                e[y * P] = c(e[nx + y * P]);
                e[nx + 1 + y * P] = c(e[1 + y * P]);
            }
            if (y >= 1 && y <= ny) {
                sst[y * P] = c(sst[nx + y * P]);
                sst[nx + 1 + y * P] = c(sst[1 + y * P]);
            }
        } else {
            if (y <= ny + 1) {
                v[y * P] = oma * v[1 + y * P];
                v[-1 + y * P] = oma * v[2 + y * P];
                v[nx + 1 + y * P] = oma * v[nx + y * P];
                v[nx + 2 + y * P] = oma * v[nx - 1 + y * P];
            }
            if (y >= 0 && y <= ny + 1) {
                e[y * P] = e[1 + y * P];
                e[nx + 1 + y * P] = e[nx + y * P];
            }
            if (y >= 1 && y <= ny) {
                sst[y * P] = sst[1 + y * P];
                sst[nx + 1 + y * P] = sst[nx + y * P];
            }
        }
    } else {
        const int x = k - 2;                     // x = -2 .. nx+2
        if (x < -2 || x > nx + 2) return;
        const int xu_lo = PER ? -2 : -1, xu_hi = PER ? nx + 1 : nx + 1;
        if (!PER && x >= xu_lo && x <= xu_hi) {
            u[x] = oma * u[x + P];
            u[x - P] = oma * u[x + 2 * P];
            u[x + (ny + 1) * P] = oma * u[x + ny * P];
            u[x + (ny + 2) * P] = oma * u[x + (ny - 1) * P];
        }
        if (x >= 0 && x <= nx + 1) {
            e[x] = e[x + P];
            e[x + (ny + 1) * P] = e[x + ny * P];
        }
    }
}

// T(u0) into a T-format field: the tracer mid-point snapshot um ← u0rhs
// (tracer_advection.jl:12-17); dst and src point at the first element of the allocations
template <class F>
__global__ void snapshot_kernel(typename F::ST *__restrict__ dst, const typename F::Q *__restrict__ src, size_t n) {
    using R = typename F::R;
    using ST = typename F::ST;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) stp(dst, (long)i, enc<ST>(toR<R>(src[i])));
}

// any NaN in the listed fields (nan_detection!, feedback.jl:61-72); f[k] points at (0,0)
// of field k with rows y0[k]..y1[k] and columns x0[k]..x1[k]
struct NanArgs {
    const void *f[4];
    int x0[4], x1[4], y0[4], y1[4];
    int P, nf;
};
template <class Q>
__global__ void nan_kernel(NanArgs A, int *flag) {
    const int y = (int)blockIdx.y + A.y0[blockIdx.z];
    const int k = blockIdx.z;
    if (y > A.y1[k]) return;
    const Q *f = (const Q *)A.f[k];
    bool bad = false;
    for (int x = A.x0[k] + (int)threadIdx.x; x <= A.x1[k]; x += blockDim.x) bad |= isnan(f[x + y * A.P]);
    if (__any(bad) && (threadIdx.x % WAVE) == 0) atomicOr(flag, 1);
}

// FP64 mass and energy over the interior (swm_reduce): block partial sums in double.
template <class Q, class ST>
__global__ void reduce_kernel(const Q *u, const Q *v, const Q *e, const ST *H, int nx, int ny, int P, int ep, double g,
                              double sinv, double *part) {
    __shared__ double sm[2][256];
    double mass = 0.0, en = 0.0;
    const int y = 1 + blockIdx.x;
    for (int x = 1 + (int)threadIdx.x; x <= nx; x += blockDim.x) {
        const double ee = (double)e[x + y * P];
        const double h = ee + (double)ldp(H, x + (long)y * P);
        mass += ee;
        en += 0.5 * g * ee * ee;
        const double uu = (double)u[x + y * P] * sinv, vv = (double)v[x + y * P] * sinv;
        if (ep || x <= nx - 1) en += 0.5 * h * uu * uu;
        if (y <= ny - 1) en += 0.5 * h * vv * vv;
    }
    sm[0][threadIdx.x] = mass;
    sm[1][threadIdx.x] = en;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0][0];
        part[2 * blockIdx.x + 1] = sm[1][0];
    }
}

}  // namespace swm
