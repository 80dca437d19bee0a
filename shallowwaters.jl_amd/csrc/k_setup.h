// k_setup.h -- the static fields of Forcing{T} and Grid evaluated on the device (gfx950).
//
// forcing.jl:11-217 and grid.jl:91-108 evaluate every field in Float64 from the grid
// coordinates and round it to T.  The kernels here do the same: the coordinates are formed
// from the cell indices with the reference's operations (Δ·i - Δ/2 ...), every formula is
// evaluated in double in the reference's order, and the value is rounded once to T on the
// store.  The transcendental functions are the device library's double-precision ones; the
// host's may differ from them in the last bit of a double, which changes a value of T only
// when it lies within that bit of a rounding boundary of T.
#pragma once

#include "swm_common.h"

namespace swm {

// topographies (forcing.jl:39-47) and zonal winds (forcing.jl:15-37)
enum { TOPO_FLAT = 0, TOPO_RIDGE = 1, TOPO_RIDGES = 2, TOPO_SEAMOUNT = 3 };
enum { WIND_CHANNEL = 0, WIND_SHEAR = 1, WIND_DOUBLE_GYRE = 2, WIND_CONSTANT = 3, WIND_NONE = 4 };

// which field a launch fills
enum { SET_H = 0, SET_FX = 1, SET_FY = 2, SET_FQ = 3, SET_FU = 4, SET_FV = 5, SET_ETA_REF = 6, SET_FETA = 7 };

struct SetupArgs {
    int field, topo, wind_x, wind_y;
    double Lx, Ly, D, H, th, w, Fx0, Fy0, rho, f0, beta, eta_refh, eta_refw, A, phik, wk, R, phi;
    int nx, ny, ep;
    // destination: element (i, j) of the reference's array (0-based) at p + i*sx + j*sy
    void *p;
    long sx, sy;
    int m, n;          // the reference's array shape
};

constexpr double SETUP_PI = 3.141592653589793;   // Float64(π), Julia's π in a Float64 expression

__device__ __forceinline__ double setup_value(const SetupArgs &a, int i, int j) {
    const double D = a.D, Lx = a.Lx, Ly = a.Ly;
    switch (a.field) {
    case SET_H: {                       // x_T_halo, y_T_halo: Δ·(i-1) - Δ/2 for i = 1 .. nx+2 (grid.jl:59-60)
        const double xx = D * (double)(i) - D / 2, yy = D * (double)(j) - D / 2;
        const double w2 = 2 * (a.w * a.w);
        if (a.topo == TOPO_FLAT) return a.H;                                                 // forcing.jl:191-195
        if (a.topo == TOPO_SEAMOUNT) {                                                       // forcing.jl:141-151
            const double dx = xx - Lx / 2, dy = yy - Ly / 2;
            const double bx = exp(-(dx * dx) / w2), by = exp(-(dy * dy) / w2);
            return a.H - a.th * bx * by;
        }
        if (a.topo == TOPO_RIDGE) {                                                          // forcing.jl:156-168
            const double dx = xx - Lx / 2;
            return a.H - a.th * exp(-(dx * dx) / w2);
        }
        const double d1 = xx - 0.99 * Lx / 4, d2 = xx - 1.01 * Lx / 2, d3 = xx - 0.99 * 3 * Lx / 4, d4 = xx - Lx;  // :171-188
        const double b0 = exp(-(xx * xx) / w2), b1 = exp(-(d1 * d1) / w2), b2 = exp(-(d2 * d2) / w2);
        const double b3 = exp(-(d3 * d3) / w2), b4 = exp(-(d4 * d4) / w2);
        return a.H - a.th * b0 - a.th * b1 - a.th * b2 - a.th * b3 - a.th * b4;
    }
    case SET_FX: {                      // y_u = Δ·j - Δ/2, j = 1 .. ny (grid.jl:41-43)
        const double c = D * a.Fx0 / a.rho / a.H;
        const double yy = D * (double)(j + 1) - D / 2;
        const double s = yy / Ly - 1.0 / 2;
        switch (a.wind_x) {
        case WIND_CHANNEL: { const double k = cos(SETUP_PI * s); return c * (k * k); }        // forcing.jl:57-68
        case WIND_SHEAR: return c * tanh(2 * SETUP_PI * s);                                   // :72-83
        case WIND_DOUBLE_GYRE: return c * (cos(2 * SETUP_PI * s) + 2 * sin(SETUP_PI * s));  // :88-98
        case WIND_CONSTANT: return c;                                                         // :101-111
        default: return 0.0;                                                                  // :114-123
        }
    }
    case SET_FY:                        // only constant and none are valid (the others build Fy on the u-grid)
        return a.wind_y == WIND_CONSTANT ? D * a.Fy0 / a.rho / a.H : 0.0;
    case SET_FQ: {                      // y_q = Δ·k - Δ, k = 1 .. ny+1 (grid.jl:99-108)
        const double yy = D * (double)(j + 1) - D;
        return D * (a.f0 + a.beta * (yy - Ly / 2));
    }
    case SET_FU: {                      // y_u (grid.jl:93)
        const double yy = D * (double)(j + 1) - D / 2;
        return D * (a.f0 + a.beta * (yy - Ly / 2));
    }
    case SET_FV: {                      // y_v = Δ·j, j = 1 .. ny-1 (grid.jl:94)
        const double yy = D * (double)(j + 1);
        return D * (a.f0 + a.beta * (yy - Ly / 2));
    }
    case SET_ETA_REF: {                 // forcing.jl:127-136 on the T points (x_T, y_T)
        const double yy = D * (double)(j + 1) - D / 2;
        return -(a.eta_refh / 2) * tanh(2 * SETUP_PI * (Ly / (4 * a.eta_refw)) * (yy / Ly - 1.0 / 2));
    }
    default: {                          // SET_FETA, forcing.jl:198-212
        const double yy = D * (double)(j + 1) - D / 2;
        const double mphi = 2 * SETUP_PI * a.R / 360.;
        const double y0 = Ly / 2 - (a.phi - a.phik) * mphi;
        const double d = yy - y0;
        return a.A * D * exp(-(d * d) / (2 * (a.wk * a.wk)));
    }
    }
}

// one thread per element (i, j) of the reference's array; ST the storage of T
template <class ST>
__global__ void setup_kernel(SetupArgs a) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (i >= a.m || j >= a.n) return;
    const double v = setup_value(a, i, j);
    ST *q = (ST *)a.p + (long)i * a.sx + (long)j * a.sy;
    if constexpr (std::is_same_v<ST, double>) *q = v;
    else if constexpr (std::is_same_v<ST, float>) *q = (float)v;
    else if constexpr (std::is_same_v<ST, h16>) q->b = __builtin_bit_cast(uint16_t, (_Float16)round_f16(round_odd_f32(v)));
    else q->b = (uint16_t)(__builtin_bit_cast(uint32_t, round_bf16(round_odd_f32(v))) >> 16);
}

}  // namespace swm
