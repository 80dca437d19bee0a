// adj_emu.cpp -- host build of the adjoint kernels (shallowwaters.jl_amd/csrc/adj_core.h).
//
// TEST INFRASTRUCTURE ONLY.  The kernels' march bodies are compiled here against an
// emulated wavefront: 64 lanes held in arrays, lane shifts done by copying, every load
// bounds-checked against the field it names.  Tests compare its results with the
// transposed Jacobian of the oracle's tangent-linear model on small grids, and the GPU
// build's results with it.  With nan_edges set, the lanes past a wave's edge read NaN
// instead of 0, so an output that depends on a lane outside its wave shows up as NaN.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../shallowwaters.jl_amd/csrc/adj_core.h"

using namespace swm::adj;

namespace {

// ------------------------------------------------------------------ emulated lanes
template <class T> struct Lanes {
    T l[AWAVE];
    Lanes() {}
    Lanes(T s) {
        for (auto &e : l) e = s;
    }
    T &operator[](int i) { return l[i]; }
    const T &operator[](int i) const { return l[i]; }
#define LANES_BIN(OP)                                                             \
    friend Lanes operator OP(const Lanes &a, const Lanes &b) {                    \
        Lanes r;                                                                  \
        for (int i = 0; i < AWAVE; ++i) r.l[i] = a.l[i] OP b.l[i];                \
        return r;                                                                 \
    }
    LANES_BIN(+) LANES_BIN(-) LANES_BIN(*) LANES_BIN(/)
#undef LANES_BIN
    friend Lanes operator-(const Lanes &a) {
        Lanes r;
        for (int i = 0; i < AWAVE; ++i) r.l[i] = -a.l[i];
        return r;
    }
#define LANES_CMP(OP)                                                             \
    friend Lanes<bool> operator OP(const Lanes &a, const Lanes &b) {              \
        Lanes<bool> r;                                                            \
        for (int i = 0; i < AWAVE; ++i) r.l[i] = a.l[i] OP b.l[i];                \
        return r;                                                                 \
    }
    LANES_CMP(<) LANES_CMP(>) LANES_CMP(<=) LANES_CMP(>=) LANES_CMP(==) LANES_CMP(!=)
#undef LANES_CMP
};
inline Lanes<bool> operator&(const Lanes<bool> &a, const Lanes<bool> &b) {
    Lanes<bool> r;
    for (int i = 0; i < AWAVE; ++i) r.l[i] = a.l[i] && b.l[i];
    return r;
}
inline Lanes<bool> operator|(const Lanes<bool> &a, const Lanes<bool> &b) {
    Lanes<bool> r;
    for (int i = 0; i < AWAVE; ++i) r.l[i] = a.l[i] || b.l[i];
    return r;
}
inline Lanes<bool> operator&(const Lanes<bool> &a, bool b) { return a & Lanes<bool>(b); }
inline Lanes<bool> operator!(const Lanes<bool> &a) {
    Lanes<bool> r;
    for (int i = 0; i < AWAVE; ++i) r.l[i] = !a.l[i];
    return r;
}

struct Fault {};
static char g_err[512];

struct Arena {
    std::vector<char> mem;
    int P = 0, rows = 0, es = 8;
    std::vector<uint32_t> origins;      // 2-D fields
    std::vector<uint32_t> origins1;     // 1-D fields
    size_t field_bytes() const { return (size_t)P * rows * es; }
    bool is2d(uint32_t o) const {
        for (uint32_t f : origins) if (f == o) return true;
        return false;
    }
    void check(uint32_t o, int x, int y, const char *what) const {
        bool ok = false;
        for (uint32_t f : origins)
            if (f == o) ok = x + AOX >= 0 && x + AOX < P && y + AOY >= 0 && y + AOY < rows;
        for (uint32_t f : origins1)
            if (f == o) ok = x == 0 && y + AOY >= 0 && y + AOY < rows;
        if (!ok) {
            snprintf(g_err, sizeof g_err, "%s out of bounds: field origin %u, x=%d y=%d", what, o, x, y);
            throw Fault();
        }
    }
    uint32_t add(bool two_d) {
        const size_t at = (mem.size() + 255) & ~(size_t)255;
        const size_t n = two_d ? field_bytes() : (size_t)rows * es;
        mem.resize(at + n, 0);
        (two_d ? origins : origins1).push_back((uint32_t)at);
        return (uint32_t)at;
    }
};

static bool g_nan_edges = false;

template <class R_> struct EmuWave {
    using R = R_;
    using V = Lanes<R>;
    using M = Lanes<bool>;
    using I = Lanes<int>;
    Arena *a;
    I x;
    M own;
    EmuWave(Arena *a_, int x0, int H) : a(a_) {
        for (int l = 0; l < AWAVE; ++l) {
            x[l] = x0 + l;
            own[l] = l >= H && l < AWAVE - H;
        }
    }
    R *addr(uint32_t o, int xc, int y) const {
        a->check(o, xc, y, "load/store");
        return (R *)(a->mem.data() + o + ((size_t)(xc + AOX) + (size_t)(y + AOY) * a->P) * sizeof(R));
    }
    R *addr1(uint32_t o, int y) const {
        a->check(o, 0, y, "row load/store");
        return (R *)(a->mem.data() + o + (size_t)(y + AOY) * sizeof(R));
    }
    V ld(uint32_t o, int y) const {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = *addr(o, x[l], y);
        return r;
    }
    V ld_at(uint32_t o, const I &xc, int y, const M &m) const {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = m[l] ? *addr(o, xc[l], y) : R(0);
        return r;
    }
    V ld_row1(uint32_t o, int y) const { return V(*addr1(o, y)); }
    V ld_s(uint32_t o, int y, int is2d) const { return is2d ? ld(o, y) : ld_row1(o, y); }
    void st(uint32_t o, int y, const V &v, const M &m) const {
        for (int l = 0; l < AWAVE; ++l)
            if (m[l]) *addr(o, x[l], y) = v[l];
    }
    void st_row1(uint32_t o, int y, const V &v, const M &m) const {
        for (int l = 0; l < AWAVE; ++l)
            if (m[l]) *addr1(o, y) = v[l];
    }
    static R edge() { return g_nan_edges ? R(NAN) : R(0); }
    static V xp(const V &v) {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = l + 1 < AWAVE ? v[l + 1] : edge();
        return r;
    }
    static V xm(const V &v) {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = l > 0 ? v[l - 1] : edge();
        return r;
    }
    static V sel(const M &m, const V &p, const V &q) {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = m[l] ? p[l] : q[l];
        return r;
    }
    static I seli(const M &m, int p, const I &q) {
        I r;
        for (int l = 0; l < AWAVE; ++l) r[l] = m[l] ? p : q[l];
        return r;
    }
    static I seli(const M &m, int p, int q) { return seli(m, p, I(q)); }
    static V vsqrt(const V &v) {
        V r;
        for (int l = 0; l < AWAVE; ++l) r[l] = std::sqrt(v[l]);
        return r;
    }
};

template <class R> struct Cells {
    Arena *a;
    R &at(uint32_t o, int x, int y) const {
        a->check(o, x, y, "cell");
        return *(R *)(a->mem.data() + o + ((size_t)(x + AOX) + (size_t)(y + AOY) * a->P) * sizeof(R));
    }
    int &ati(uint32_t o, int x, int y) const {
        a->check(o, x, y, "cell");
        return *(int *)(a->mem.data() + o + ((size_t)(x + AOX) + (size_t)(y + AOY) * a->P) * sizeof(R));
    }
    R &row1(uint32_t o, int y) const {
        a->check(o, 0, y, "row cell");
        return *(R *)(a->mem.data() + o + (size_t)(y + AOY) * sizeof(R));
    }
};

// ------------------------------------------------------------------ launcher
template <class R> struct EmuK {
    Arena *a;
    int L;
    int dmax_x = 0, dmax_y = 0;

    template <int H, class Fn> void march(int nx, int ny, int y0, int y1, Fn fn) {
        const int nstr = nstrips_for<H>(nx + 2);
        const int nseg = (y1 - y0 + L - 1) / L;
        for (int w = 0; w < nstr * nseg; ++w) {
            const Tile t = tile_of<H>(w, nstr, nseg, L, y0, y1);
            EmuWave<R> W(a, t.x_lane0, H);
            fn(W, t);
        }
    }
    template <bool PER, bool NEXT, int NX, bool PRE, int SCH> void stage_t(const StageAdjArgs<R> &S) {
        march<RT_HALO>(S.nx, S.ny, ADJ_Y0, adj_y1(S.ny), [&](EmuWave<R> &W, const Tile &t) {
            stage_adjoint<EmuWave<R>, PER, NEXT, NX, PRE, SCH>(W, S, t.ya, t.yb);
        });
    }
    template <bool PER, int SCH> void stage_p(const StageAdjArgs<R> &S, int next, int nx, int pre) {
        if (!next) stage_t<PER, false, 0, false, SCH>(S);
        else if (!pre) stage_t<PER, true, 0, false, SCH>(S);
        else if (nx == 0) stage_t<PER, true, 0, true, SCH>(S);
        else if (nx == 1) stage_t<PER, true, 1, true, SCH>(S);
        else stage_t<PER, true, 2, true, SCH>(S);
    }
    template <bool PER> void stage_s(const StageAdjArgs<R> &S, int next, int nx, int pre, int sch) {
        if (sch == ADJ_CONT) stage_t<PER, false, 0, false, ADJ_CONT>(S);
        else if (sch == ADJ_LIN) stage_p<PER, ADJ_LIN>(S, next, nx, pre);
        else if (sch == ADJ_SAD) stage_p<PER, ADJ_SAD>(S, next, nx, pre);
        else if (sch == (ADJ_AH | ADJ_KEEP)) stage_p<PER, ADJ_AH | ADJ_KEEP>(S, next, nx, pre);
        else if (sch == (ADJ_SAD | ADJ_KEEP)) stage_p<PER, ADJ_SAD | ADJ_KEEP>(S, next, nx, pre);
        else stage_p<PER, ADJ_AH>(S, next, nx, pre);
    }
    int stage(const StageAdjArgs<R> &S, int per, int next, int nx, int pre, int sch) {
        if (per) stage_s<true>(S, next, nx, pre, sch);
        else stage_s<false>(S, next, nx, pre, sch);
        return 0;
    }
    int qend(const StageAdjArgs<R> &S) {
        Cells<R> c{a};
        for (int y = 0; y <= S.ny; ++y) qend_row<R>(c, S, y);
        return 0;
    }
    int advcor_adj(const AdvcorAdjArgs &X) {
        Cells<R> c{a};
        for (int y = -1; y <= X.ny + 2; ++y)
            for (int x = -2; x <= X.nx + 2; ++x) advcor_adjoint_cell<R>(c, X, x, y);
        return 0;
    }
    int lqend_keep(const StageAdjArgs<R> &S) {
        Cells<R> c{a};
        for (int y = 0; y <= S.ny; ++y) lqend_keep_row<R>(c, S, y);
        return 0;
    }
    int lincomb(uint32_t d, uint32_t x, R ax, uint32_t y, R ay, uint32_t z, R az) {
        R *b = (R *)a->mem.data();
        const size_t n = (size_t)a->P * a->rows;
        for (size_t k = 0; k < n; ++k)
            lincomb_elem<R>(b + d / sizeof(R), b + x / sizeof(R), ax, y == NOF ? nullptr : b + y / sizeof(R), ay,
                            z == NOF ? nullptr : b + z / sizeof(R), az, (long)k);
        return 0;
    }
    int zero(uint32_t o) {
        R *b = (R *)(a->mem.data() + o);
        for (size_t k = 0; k < (size_t)a->P * a->rows; ++k) b[k] = R(0);
        return 0;
    }
    int qend_adj(const StageAdjArgs<R> &S, const UVE &mu) {
        Cells<R> c{a};
        for (int yy = 0; yy <= S.ny + 1; ++yy) qend_adjoint_row<R>(c, S, mu, yy);
        return 0;
    }
    template <bool PER, int DRAG, bool SMAG> void diss_t(const DissAdjArgs<R> &D) {
        march<DT_HALO>(D.nx, D.ny, ADJ_Y0, adj_y1(D.ny), [&](EmuWave<R> &W, const Tile &t) {
            diss_adjoint<EmuWave<R>, PER, DRAG, SMAG>(W, D, t.ya, t.yb);
        });
    }
    template <bool PER> void diss_p(const DissAdjArgs<R> &D, int drag, int smag) {
        if (drag == 0) { if (smag) diss_t<PER, 0, true>(D); else diss_t<PER, 0, false>(D); }
        else if (drag == 1) { if (smag) diss_t<PER, 1, true>(D); else diss_t<PER, 1, false>(D); }
        else { if (smag) diss_t<PER, 2, true>(D); else diss_t<PER, 2, false>(D); }
    }
    int diss(const DissAdjArgs<R> &D, int per, int drag, int smag) {
        if (per) diss_p<true>(D, drag, smag);
        else diss_p<false>(D, drag, smag);
        return 0;
    }
    template <bool PER> void tracer_t(const TracerAdjArgs<R> &T) {
        Cells<R> c{a};
        int dx = 0, dy = 0;
        for (int y = 1; y <= T.ny; ++y)
            for (int x = 1; x <= T.nx; ++x) {
                const int d = tracer_adj_source<R, PER>(c, T, x, y);
                dx = std::max(dx, d & 0xFFFF);
                dy = std::max(dy, d >> 16);
            }
        dmax_x = dx;
        dmax_y = dy;
        for (int Y = -1; Y <= T.ny + 3; ++Y)
            for (int X = -2; X <= T.nx + 3; ++X) tracer_adj_gather_uv<R, PER>(c, T, X, Y, dx, dy);
        for (int Y = -1; Y <= T.ny + 2; ++Y)
            for (int X = -2; X <= T.nx + 2; ++X) tracer_adj_gather<R, PER>(c, T, X, Y);
    }
    int tracer(const TracerAdjArgs<R> &T) {
        if (T.ep) tracer_t<true>(T);
        else tracer_t<false>(T);
        return 0;
    }
    int snapshot(uint32_t lu, uint32_t lum, uint32_t lv, uint32_t lvm) {
        const size_t n = (size_t)a->P * a->rows;
        R *b = (R *)a->mem.data();
        for (size_t k = 0; k < n; ++k) {
            snapshot_adjoint_elem<R>(b + lu / sizeof(R), b + lum / sizeof(R), (long)k);
            snapshot_adjoint_elem<R>(b + lv / sizeof(R), b + lvm / sizeof(R), (long)k);
        }
        return 0;
    }
};

}  // namespace

// ------------------------------------------------------------------ C entry points

extern "C" {

// Scalars of a configuration, as the GPU driver derives them from swm_config.
struct EmuCfg {
    int32_t nx, ny, per, RKo, cont, drag, smag, consume, tracer, nadvstep, nadvstep_half, nstep_diff;
    int32_t L, nan_edges, fmt, sad, lin, ssp, s4, s3, kn, mn;
    double a[4], b[4];
    double cs, sa, sb, cs2;
    double g, gamma, k12, oma, cD, rD, cS, nuB, dtd, hdu, hdv, du, dv, jSST;
};

const char *emu_last_error() { return g_err; }

}  // extern "C"

namespace {

// Julia arrays (float64, column-major) <-> padded fields
struct Shape {
    int m, n, x0, y0;
};
template <class R> void put(Arena &A, uint32_t o, const double *h, Shape s) {
    Cells<R> c{&A};
    for (int j = 0; j < s.n; ++j)
        for (int i = 0; i < s.m; ++i) c.at(o, s.x0 + i, s.y0 + j) = (R)h[i + (size_t)j * s.m];
}
template <class R> void get(Arena &A, uint32_t o, double *h, Shape s) {
    Cells<R> c{&A};
    for (int j = 0; j < s.n; ++j)
        for (int i = 0; i < s.m; ++i) h[i + (size_t)j * s.m] = (double)c.at(o, s.x0 + i, s.y0 + j);
}

struct Shapes {
    Shape u, v, e, sst, fq, fu, fv;
};
Shapes shapes(const EmuCfg &c) {
    const int ep = c.per ? 1 : 0, nux = c.per ? c.nx : c.nx - 1;
    Shapes s;
    s.u = {nux + 4, c.ny + 4, -1 - ep, -1};
    s.v = {c.nx + 4, c.ny - 1 + 4, -1, -1};
    s.e = {c.nx + 2, c.ny + 2, 0, 0};
    s.sst = {c.nx + 2, c.ny, 0, 1};
    s.fq = {c.nx + 1, c.ny + 1, 0, 0};
    s.fu = {nux, c.ny, 1 - ep, 1};
    s.fv = {c.nx, c.ny - 1, 1, 1};
    return s;
}

template <class R> AdjConsts<R> consts(const EmuCfg &c) {
    AdjConsts<R> k{};
    for (int j = 0; j < 4; ++j) { k.a[j] = (R)c.a[j]; k.b[j] = (R)c.b[j]; }
    k.g = (R)c.g; k.gamma = (R)c.gamma; k.k12 = (R)c.k12; k.oma = (R)c.oma;
    k.cD = (R)c.cD; k.rD = (R)c.rD; k.cS = (R)c.cS; k.nuB = (R)c.nuB; k.dtd = (R)c.dtd;
    k.hdu = (R)c.hdu; k.hdv = (R)c.hdv; k.du = (R)c.du; k.dv = (R)c.dv; k.jSST = (R)c.jSST;
    k.RKo = c.RKo; k.cont = c.cont; k.drag = c.drag; k.smag = c.smag; k.per = c.per; k.consume = c.consume;
    k.tracer = c.tracer; k.nadvstep = c.nadvstep; k.nadvstep_half = c.nadvstep_half; k.nstep_diff = c.nstep_diff;
    k.nx = c.nx; k.ny = c.ny; k.ep = c.per ? 1 : 0;
    k.sad = c.sad;
    k.lin = c.lin;
    k.ssp = c.ssp;
    k.s4 = c.s4;
    k.s3 = c.s3; k.kn = c.kn; k.mn = c.mn; k.cs2 = (R)c.cs2;
    k.cs = (R)c.cs; k.sa = (R)c.sa; k.sb = (R)c.sb;
    return k;
}

// op 1: μ = (∂R/∂s)ᵀ λd.   in: su sv se H fq gu gv ge (+ f_u f_v with the linear dynamics);
//       out: μu μv μe
// op 2: λ before the dissipation.  in: u0 v0 e0 H lu lv le;  out: λu λv λe
// op 3: tracer at step i (advection and snapshot as due).  in: u v um vm sst lum lvm lsst;
//       out: λu λv λum λvm λsst
// op 4: reverse of step i.  in: s_1..s_RKo (3 each; SSPRK2: w_1..w_RKs+1; SSPRK3: w_1..w_s+1; 4SSPRK3: w_1..w_4 and
//       the 4 stages' (u', v')), pre (3), post (2), um vm sst, H fq,
//       L (u v e sst um vm) (+ f_u f_v with the linear dynamics);  out: L (u v e sst um vm)
// op 5: fold of the ghost copies.  in: Lu Lv Le Lsst;  out: gu gv ge gsst
template <class R> int run_op(const EmuCfg &c, int op, int64_t i, const double *const *in, double *const *out) {
    Arena A;
    A.P = ((c.nx + 2 * AOX + 63) / 64) * 64;
    A.rows = c.ny + 2 * AOY + 4;
    A.es = sizeof(R);
    const Shapes sh = shapes(c);
    AdjConsts<R> k = consts<R>(c);
    k.P = A.P;
    EmuK<R> K{&A, c.L};
    auto uve = [&]() { UVE f; f.u = A.add(true); f.v = A.add(true); f.e = A.add(true); return f; };
    AdjFields F{};
    const int ns = c.ssp ? c.ssp + 1 : c.s4 ? 4 : c.s3 ? c.s3 + 1 : c.RKo;    // taped stage inputs
    for (int j = 0; j < ns; ++j) F.s[j] = uve();
    for (int j = 0; j < ns; ++j) {                       // SSPRK2, SSPRK3: u'_k = u_{k+1}; 4SSPRK3: fields of their own
        if (c.s4) F.p[j] = UVE{A.add(true), A.add(true), 0};
        else if (j + 1 < ns) F.p[j] = UVE{F.s[j + 1].u, F.s[j + 1].v, 0};
    }
    F.pre = uve();
    F.post_u = A.add(true); F.post_v = A.add(true); F.um = A.add(true); F.vm = A.add(true); F.sst = A.add(true);
    F.L = uve(); F.Z = uve(); F.Ma = uve(); F.Mb = uve(); F.A = uve();
    F.Lsst = A.add(true); F.Lsst2 = A.add(true); F.Lum = A.add(true); F.Lvm = A.add(true);
    F.H = A.add(true); F.fq = A.add(true); F.fq_2d = 1;
    F.fu = A.add(true); F.fv = A.add(true); F.fu_2d = F.fv_2d = 1;
    F.qend = A.add(false); F.lqend = A.add(false);
    F.tr0 = A.add(true); F.tr1 = A.add(true);
    F.tg0 = A.add(true); F.tg1 = A.add(true);
    F.zero = A.add(true);
    try {
        if (op == 1) {
            put<R>(A, F.s[0].u, in[0], sh.u); put<R>(A, F.s[0].v, in[1], sh.v); put<R>(A, F.s[0].e, in[2], sh.e);
            put<R>(A, F.H, in[3], sh.e); put<R>(A, F.fq, in[4], sh.fq);
            put<R>(A, F.L.u, in[5], sh.u); put<R>(A, F.L.v, in[6], sh.v); put<R>(A, F.L.e, in[7], sh.e);
            if (c.lin) { put<R>(A, F.fu, in[8], sh.fu); put<R>(A, F.fv, in[9], sh.fv); }
            StageAdjArgs<R> S{};
            S.s = F.s[0]; S.H = F.H; S.fq = F.fq; S.fq_2d = 1; S.qend = F.qend; S.lqend = F.lqend;
            S.fu = F.fu; S.fv = F.fv; S.fu_2d = S.fv_2d = 1;
            S.l0 = F.L; S.mu_in = F.Ma; S.mu_x[0] = S.mu_x[1] = F.A; S.mu_out = F.Mb; S.a_out = F.A;
            S.a = R(1); S.b = R(0);
            S.g = k.g; S.gamma = k.gamma; S.k12 = k.k12; S.oma = k.oma; S.cont = c.cont;
            S.nx = c.nx; S.ny = c.ny; S.P = A.P;
            const bool wrap = c.per && !c.sad && !c.lin;
            if (wrap) K.qend(S);
            K.stage(S, c.per, 0, 0, 0, c.lin ? ADJ_LIN : c.sad ? ADJ_SAD : ADJ_AH);
            if (wrap) K.qend_adj(S, S.mu_out);
            get<R>(A, F.Mb.u, out[0], sh.u); get<R>(A, F.Mb.v, out[1], sh.v); get<R>(A, F.Mb.e, out[2], sh.e);
        } else if (op == 2) {
            put<R>(A, F.pre.u, in[0], sh.u); put<R>(A, F.pre.v, in[1], sh.v); put<R>(A, F.pre.e, in[2], sh.e);
            put<R>(A, F.H, in[3], sh.e);
            put<R>(A, F.L.u, in[4], sh.u); put<R>(A, F.L.v, in[5], sh.v); put<R>(A, F.L.e, in[6], sh.e);
            DissAdjArgs<R> D{};
            D.s = F.pre; D.H = F.H; D.lin = F.L; D.out = F.Z;
            D.cD = k.cD; D.rD = k.rD; D.cS = k.cS; D.nuB = k.nuB; D.dtd = k.dtd; D.oma = k.oma;
            D.nx = c.nx; D.ny = c.ny; D.P = A.P;
            K.diss(D, c.per, c.drag, c.smag);
            get<R>(A, F.Z.u, out[0], sh.u); get<R>(A, F.Z.v, out[1], sh.v); get<R>(A, F.Z.e, out[2], sh.e);
        } else if (op == 3 || op == 4) {
            int p = 0;
            if (op == 4) {
                for (int j = 0; j < ns; ++j) {
                    put<R>(A, F.s[j].u, in[p++], sh.u); put<R>(A, F.s[j].v, in[p++], sh.v); put<R>(A, F.s[j].e, in[p++], sh.e);
                }
                for (int j = 0; c.s4 && j < 4; ++j) { put<R>(A, F.p[j].u, in[p++], sh.u); put<R>(A, F.p[j].v, in[p++], sh.v); }
                put<R>(A, F.pre.u, in[p++], sh.u); put<R>(A, F.pre.v, in[p++], sh.v); put<R>(A, F.pre.e, in[p++], sh.e);
            }
            put<R>(A, F.post_u, in[p++], sh.u); put<R>(A, F.post_v, in[p++], sh.v);
            put<R>(A, F.um, in[p++], sh.u); put<R>(A, F.vm, in[p++], sh.v); put<R>(A, F.sst, in[p++], sh.sst);
            if (op == 4) { put<R>(A, F.H, in[p++], sh.e); put<R>(A, F.fq, in[p++], sh.fq); }
            if (op == 4) { put<R>(A, F.L.u, in[p++], sh.u); put<R>(A, F.L.v, in[p++], sh.v); put<R>(A, F.L.e, in[p++], sh.e); }
            if (op == 4) put<R>(A, F.Lsst, in[p++], sh.sst);
            put<R>(A, F.Lum, in[p++], sh.u); put<R>(A, F.Lvm, in[p++], sh.v);
            if (op == 3) put<R>(A, F.Lsst, in[p++], sh.sst);
            if (op == 4 && c.lin) { put<R>(A, F.fu, in[p++], sh.fu); put<R>(A, F.fv, in[p++], sh.fv); }
            if (op == 3) {
                reverse_tracer<R>(K, F, k, i);
                get<R>(A, F.L.u, out[0], sh.u); get<R>(A, F.L.v, out[1], sh.v);
                get<R>(A, F.Lum, out[2], sh.u); get<R>(A, F.Lvm, out[3], sh.v); get<R>(A, F.Lsst, out[4], sh.sst);
            } else {
                reverse_step<R>(K, F, k, i);
                get<R>(A, F.L.u, out[0], sh.u); get<R>(A, F.L.v, out[1], sh.v); get<R>(A, F.L.e, out[2], sh.e);
                get<R>(A, F.Lsst, out[3], sh.sst); get<R>(A, F.Lum, out[4], sh.u); get<R>(A, F.Lvm, out[5], sh.v);
            }
        } else if (op == 5) {
            put<R>(A, F.L.u, in[0], sh.u); put<R>(A, F.L.v, in[1], sh.v); put<R>(A, F.L.e, in[2], sh.e);
            put<R>(A, F.Lsst, in[3], sh.sst);
            FoldOutArgs O{F.L, F.Z, F.Lsst, F.Lsst2, c.nx, c.ny, c.per};
            Cells<R> cc{&A};
            for (int y = -1; y <= c.ny + 2; ++y)
                for (int x = -2; x <= c.nx + 2; ++x) fold_out_cell<R>(cc, O, k.oma, x, y);
            get<R>(A, F.Z.u, out[0], sh.u); get<R>(A, F.Z.v, out[1], sh.v); get<R>(A, F.Z.e, out[2], sh.e);
            get<R>(A, F.Lsst2, out[3], sh.sst);
        } else {
            snprintf(g_err, sizeof g_err, "unknown op %d", op);
            return -1;
        }
    } catch (Fault &) {
        return -2;
    }
    return 0;
}

}  // namespace

extern "C" int emu_run(const EmuCfg *c, int32_t op, int64_t i, const double *const *in, double *const *out) {
    g_nan_edges = c->nan_edges != 0;
    g_err[0] = 0;
    if (c->fmt == 8) return run_op<double>(*c, op, i, in, out);
    return run_op<float>(*c, op, i, in, out);
}
