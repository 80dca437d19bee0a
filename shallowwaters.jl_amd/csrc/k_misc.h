// k_misc.h -- tracer advection, ghost points of an uploaded state, copies, NaN check and
// FP64 reductions (gfx950).
#pragma once

#include "swm_common.h"

namespace swm {

// ------------------------------------------------------------------ tracer (K3)
//
// tracer!  tracer_advection.jl:1-44 when i % nadvstep == 0: departure! (:52-85, with
// backtraj! :96-123 and interp_uv! :130-173 gathering the mid-point velocities um, vm
// interpolated to the T-grid on the fly), adv_sst! (:179-206), ghost_points_sst! and the
// copy back to sst.  One thread per T-cell.  idx(k,l) maps the reference's absolute
// indices of u_T/v_T/sst back to (x,y).
template <class R, bool PER>
__global__ void tracer_kernel(const R *u, const R *v, const R *um, const R *vm, const R *sst, R *ssti,
                              R hdu, R hdv, R du_, R dv_, int nx, int ny, int P, int ep, int consume,
                              R jSST, R SSTmin, int *bad) {
    const int x = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = 1 + (int)blockIdx.y;
    if (x > nx) return;
    const R half = R(0.5), one = R(1.0);
    // Ix!(u_T,u): u_T(x,y) = ½(u(x) + u(x-1)); Iy!(v_T,v): v_T(x,y) = ½(v(y) + v(y-1))
    const R uT = half * (u[x + y * P] + u[x - 1 + y * P]);
    const R vT = half * (v[x + y * P] + v[x + (y - 1) * P]);
    const R xd0 = -hdu * uT, yd0 = -hdv * vT;
    // interp_uv! on um_T (x-index k <-> T-column k-1-ep, y-index l <-> row l-2) and
    // vm_T (k <-> column k-2, l <-> row l-1)
    auto interp = [&](const R *f, bool is_u, R xx, R yy) -> R {
        const int muv = is_u ? nx + 2 + ep : nx + 4, nuv = is_u ? ny + 4 : ny + 2;
        const int xh1 = is_u ? 1 + ep : 2, xh2 = is_u ? 1 : 2, yh1 = is_u ? 2 : 1;
        const R fx = floor(xx), fy = floor(yy);
        if (!isfinite(fx) || !isfinite(fy)) { atomicOr(bad, 1); return R(0); }
        const long xi = (long)fx, yi = (long)fy;
        long k = xi + x + xh1, l = yi + y + yh1;
        R x0 = xx - R((double)xi), y0 = yy - R((double)yi);
        if (PER) {
            if (k < 1 + xh1) k = k + muv - xh1 - xh2;
            else if (k > muv - 1) k = k - muv + xh1 + xh2;
        } else {
            if (k < 1 + xh1) { k = 1 + xh1; x0 = R(0); }
            else if (k > muv - 1) { k = muv - 1; x0 = one; }
        }
        if (l < 1 + yh1) { l = 1 + yh1; y0 = R(0); }
        else if (l > nuv - 1) { l = nuv - 1; y0 = one; }
        if (k < 1 || k + 1 > muv) { atomicOr(bad, 2); return R(0); }
        // value of f_T at reference index (kk, ll)
        auto fT = [&](long kk, long ll) -> R {
            if (is_u) { const int cx = (int)kk - 1 - ep, cy = (int)ll - 2; return half * (f[cx + cy * P] + f[cx - 1 + cy * P]); }
            const int cx = (int)kk - 2, cy = (int)ll - 1;
            return half * (f[cx + cy * P] + f[cx + (cy - 1) * P]);
        };
        const R f00 = fT(k, l), f10 = fT(k + 1, l), f01 = fT(k, l + 1), f11 = fT(k + 1, l + 1);
        return f00 * (one - x0) * (one - y0) + f10 * x0 * (one - y0) + f01 * (one - x0) * y0 + f11 * x0 * y0;
    };
    const R ui = interp(um, true, xd0, yd0);
    const R vi = interp(vm, false, xd0, yd0);
    const R xd = -du_ * ui, yd = -dv_ * vi;
    // adv_sst!  (sst index k <-> column k-1, l <-> row l; m = nx+2, n = ny, halo (1,0))
    const R fx = floor(xd), fy = floor(yd);
    if (!isfinite(fx) || !isfinite(fy)) { atomicOr(bad, 1); return; }
    const long xi = (long)fx, yi = (long)fy;
    const int m = nx + 2, n = ny;
    long k = xi + x + 1, l = yi + y;
    R x0 = xd - R((double)xi), y0 = yd - R((double)yi);
    if (PER) {
        if (k < 2) k = k + m - 2;
        else if (k > m - 1) k = k - m + 2;
    } else {
        if (k < 2) { k = 2; x0 = R(0); }
        else if (k > m - 1) { k = m - 1; x0 = one; }
    }
    if (l < 1) { l = 1; y0 = R(0); }
    else if (l > n - 1) { l = n - 1; y0 = one; }
    if (k < 1 || k + 1 > m) { atomicOr(bad, 2); return; }
    auto S = [&](long kk, long ll) -> R { return sst[(int)kk - 1 + (int)ll * P]; };
    R s = S(k, l) * (one - x0) * (one - y0) + S(k + 1, l) * x0 * (one - y0) + S(k, l + 1) * (one - x0) * y0 +
          S(k + 1, l + 1) * x0 * y0;
    if (consume) s += jSST * (SSTmin - s);       // tracer_consumption!  :262-275
    // ghost_points_sst!  ghost_points.jl:136-179 (halosstx=1, halossty=0)
    ssti[x + y * P] = s;
    if (x == (PER ? nx : 1)) ssti[y * P] = s;
    if (x == (PER ? 1 : nx)) ssti[nx + 1 + y * P] = s;
}

// ------------------------------------------------------------------ small kernels

// ghost_points! of the uploaded state (ghost_points.jl:182-242), in the reference order:
// phase 0 writes the x-copies, phase 1 the y-copies (which read the x-copies for η).
template <class R, bool PER>
__global__ void ghost_kernel(R *u, R *v, R *e, R *sst, int nx, int ny, int P, R oma, int phase) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (phase == 0) {
        // rows of u (periodic: x-copies), v (x-copies), η (x-copies), sst (x-copies)
        const int y = k - 1;                     // y = -1 .. ny+2
        if (y < -1 || y > ny + 2) return;
        if (PER) {
            u[-2 + y * P] = u[nx - 2 + y * P];
            u[-1 + y * P] = u[nx - 1 + y * P];
            u[nx + y * P] = u[y * P];
            u[nx + 1 + y * P] = u[1 + y * P];
            if (y <= ny + 1) {
                v[-1 + y * P] = v[nx - 1 + y * P];
                v[y * P] = v[nx + y * P];
                v[nx + 1 + y * P] = v[1 + y * P];
                v[nx + 2 + y * P] = v[2 + y * P];
            }
            if (y >= 0 && y <= ny + 1) {
                e[y * P] = e[nx + y * P];
                e[nx + 1 + y * P] = e[1 + y * P];
            }
            if (y >= 1 && y <= ny) {
                sst[y * P] = sst[nx + y * P];
                sst[nx + 1 + y * P] = sst[1 + y * P];
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

// copy of a padded field (tracer mid-point snapshot um ← u0, tracer_advection.jl:12-17)
template <class R>
__global__ void copy_kernel(R *__restrict__ dst, const R *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// any NaN in the listed fields (nan_detection!, feedback.jl:61-72); f[k] points at (0,0)
// of field k with rows y0[k]..y1[k] and columns x0[k]..x1[k]
struct NanArgs {
    const void *f[4];
    int x0[4], x1[4], y0[4], y1[4];
    int P, nf;
};
template <class R>
__global__ void nan_kernel(NanArgs A, int *flag) {
    const int y = (int)blockIdx.y + A.y0[blockIdx.z];
    const int k = blockIdx.z;
    if (y > A.y1[k]) return;
    const R *f = (const R *)A.f[k];
    bool bad = false;
    for (int x = A.x0[k] + (int)threadIdx.x; x <= A.x1[k]; x += blockDim.x) bad |= isnan(f[x + y * A.P]);
    if (__any(bad) && (threadIdx.x % WAVE) == 0) atomicOr(flag, 1);
}

// FP64 mass and energy over the interior (swm_reduce): block partial sums in double.
template <class R>
__global__ void reduce_kernel(const R *u, const R *v, const R *e, const R *H, int nx, int ny, int P, int ep, double g,
                              double *part) {
    __shared__ double sm[2][256];
    double mass = 0.0, en = 0.0;
    const int y = 1 + blockIdx.x;
    for (int x = 1 + (int)threadIdx.x; x <= nx; x += blockDim.x) {
        const double ee = (double)e[x + y * P];
        const double h = ee + (double)H[x + y * P];
        mass += ee;
        en += 0.5 * g * ee * ee;
        const double uu = (double)u[x + y * P], vv = (double)v[x + y * P];
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
