// k_persist.h -- one persistent launch that advances every member of an ensemble through
// many steps (gfx950).
//
// The launch holds as many blocks as the device keeps resident at once (a cooperative
// launch).  For each step it runs the step's plan -- the same items an ensemble step
// launches one kernel each for: the two fused RK stage pairs (and, periodic, the strips at
// the seam) and the dissipation, each over the tiles of every member -- then the members'
// tracer work and NaN checks due at that step (time_integration.jl:226-237), with a grid-wide
// barrier wherever a later phase reads what an earlier one wrote.  Waves take the tiles of a
// phase in turn (tile w, w + waves, ...), so any member sizes share the launch.  The step's
// plans alternate with the parity of the step (slot roles repeat every two steps); the
// events of a member follow its own schedule, computed here from its step counters.
#pragma once

#include "k_dissipation.h"
#include "k_misc.h"
#include "k_stage.h"

namespace swm {

// an item of a step's plan (a batched kernel of the graph path): kind 1 / 4 a fused stage
// pair (all strips / the strips clear of a periodic seam), 5 the seam strips of a periodic
// pair, 2 the dissipation; the members' arguments and the tile table (first tile of each
// member) at byte offsets in the plan.  sync: a grid barrier follows the item.
struct PItem {
    int kind, ma, mb, sync;
    uint32_t args, tab;
    int tiles;
};

constexpr int PERSIST_ITEMS = 8;
constexpr int SNAP_TILE = 512;        // elements per wave and tile of a snapshot copy

// a member's tracer and NaN-check work (host-filled)
template <class F> struct PMember {
    using Q = typename F::Q;
    using ST = typename F::ST;
    TracerArgs<F> tr;                 // u, v, sst, ssti set per step from the fields below
    const Q *tu[2], *tv[2];           // the velocities the tracer reads (the step's dissipation output), parity 0 / 1
    const Q *su[2], *sv[2];           // the same fields from their first element (snapshot copies)
    Q *sst[2];                        // the two tracer slots, element (0,0)
    ST *um, *vm;                      // snapshot targets from their first element
    long long elems;                  // elements of an allocation
    NanArgs nan[2];                   // Prog at the start of a step of parity 0 / 1; f[3] set per step
    int tracer, nadvstep, nout;
    int r_adv, r_half, r_nan;         // (i0 % nadvstep), ((i0 + nadvstep_half) % nadvstep), (i0 % nout)
    int rsst0;                        // tracer slot at i0
    int *nanflag;                     // set when this member's check finds NaNs
};

struct PCtl {
    unsigned long long count;         // arrivals of XCD leaders at grid barriers
    unsigned gen;                     // barriers completed
    int abort;                        // a barrier timed out: every block leaves
    int stop;                         // first step (offset from i0) whose NaN check found NaNs
    unsigned started;                 // blocks that have counted themselves into their XCD
    int pad[26];
    struct Group {                    // the blocks on one XCD, on a 128-byte line of their own
        unsigned long long count;     // their arrivals
        unsigned gen;                 // barriers their blocks may leave
        unsigned pop;                 // blocks of the launch on this XCD
        int pad[28];
    } grp[NXCD];
};

template <class F> struct PArgs {
    const char *plan[2];              // the plans of steps of parity 0 / 1 (device)
    const PItem *item;                // the items of steps of parity p: item[p * PERSIST_ITEMS + j] (device)
    int nitems[2];
    const PMember<F> *mem;
    int nm;
    int nsteps;
    PCtl *ctl;
    long long timeout;                // wall-clock ticks a barrier waits before it gives up
    int diss;                         // dissipation form: 2 * drag (quadratic 0, linear 1, none 2) + (constant diffusion)
    int ntrace;                       // barriers whose exit time block 0 records (0: none)
    long long *trace;                 // [0]: the start, [k]: exit of barrier k (wall-clock ticks)
};

// Grid-wide barrier.  Every block arrives before any leaves; stores before it are visible
// to every block after it.  The L2 caches of the XCDs are not coherent with each other, so
// the release writes an XCD's L2 back and the acquire invalidates it -- operations on the
// whole cache, which cost too much to issue once per block.  Blocks are grouped by the XCD
// they run on (hardware register XCC_ID): the last block of an XCD to arrive (its blocks'
// stores are then all in that L2) writes the L2 back and arrives for the XCD at the grid
// counter; the last XCD releases the grid; each XCD's leader invalidates its L2 and passes
// the release on to its XCD's blocks, which invalidate their CU's L1.  No address sees more
// than an XCD's worth of atomics and polls.  The launch is cooperative, so all blocks are
// resident and the waits end; a wait longer than `timeout` ticks of the wall clock sets
// ctl->abort, and then every block leaves at its next barrier.  Returns false once aborted
// (uniform over the block).
__device__ __forceinline__ bool wait_gen(PCtl *ctl, const unsigned *gen, unsigned epoch, long long timeout) {
    const long long t0 = (long long)wall_clock64();
    while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        if (__hip_atomic_load(&ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if ((long long)wall_clock64() - t0 > timeout) {
            atomicExch(&ctl->abort, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

struct GridSync {
    PCtl *ctl;
    unsigned epoch = 0, g = 0, pop = 0, nxcd = 0;
    long long timeout;
    long long *trace;
    int ntrace;

    // count this block into its XCD, then wait until every block has (the populations are
    // then final); false if aborted
    __device__ bool start() {
        __shared__ int ok;
        if (threadIdx.x == 0) {
            g = (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | 20) % NXCD;    // hwreg(HW_REG_XCC_ID, 0, 4)
            __hip_atomic_fetch_add(&ctl->grp[g].pop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&ctl->started, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            bool good = wait_gen(ctl, &ctl->started, gridDim.x, timeout);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            pop = __hip_atomic_load(&ctl->grp[g].pop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int k = 0; k < NXCD; ++k)
                nxcd += __hip_atomic_load(&ctl->grp[k].pop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0;
            ok = good;
            if (blockIdx.x == 0 && ntrace > 0) trace[0] = (long long)wall_clock64();
        }
        __syncthreads();
        return ok;
    }

    __device__ bool sync() {
        __shared__ int ok;
        __syncthreads();                 // the block's stores have reached its XCD's L2
        ++epoch;
        if (threadIdx.x == 0) {
            PCtl::Group &G = ctl->grp[g];
            const unsigned long long a = __hip_atomic_fetch_add(&G.count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a == (unsigned long long)pop * epoch - 1) {          // the XCD's last block
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");    // its L2 written back
                const unsigned long long b = __hip_atomic_fetch_add(&ctl->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (b == (unsigned long long)nxcd * epoch - 1) __hip_atomic_store(&ctl->gen, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else (void)wait_gen(ctl, &ctl->gen, epoch, timeout);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");    // its L2 (and this CU's L1) invalidated
                __hip_atomic_store(&G.gen, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                (void)wait_gen(ctl, &G.gen, epoch, timeout);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");    // this CU's L1 (and the L2) invalidated
            }
            ok = __hip_atomic_load(&ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
            if (blockIdx.x == 0 && (int)epoch < ntrace) trace[epoch] = (long long)wall_clock64();
        }
        __syncthreads();
        return ok;
    }
};

// the first tile of a phase wave w takes when the phase's tiles so far number `base`
__device__ __forceinline__ int first_tile(int w, int base, int nw) {
    const int r = (w - base) % nw;
    return r < 0 ? r + nw : r;
}

// tracer slot of a member at step offset s: flipped by every advection in [i0, i0 + s)
template <class F> __device__ __forceinline__ int rsst_at(const PMember<F> &M, int s) {
    if (!M.tracer) return M.rsst0;
    const int a0 = (M.nadvstep - M.r_adv) % M.nadvstep;      // offset of the first advection
    const int n = s > a0 ? (s - 1 - a0) / M.nadvstep + 1 : 0;
    return M.rsst0 ^ (n & 1);
}

// The persistent ensemble kernel for fused RK4 pairs with T = Tprog, the continuity
// itself and Arakawa-Hsu advection, any drag and diffusion.
template <class F, bool PER>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PER ? 3 : 4)))
ens_persist_kernel(PArgs<F> A) {
    using Q = typename F::Q;
    using ST = typename F::ST;
    using R = typename F::R;
    const int nw = (int)gridDim.x * WAVES_PER_BLOCK, w0 = wave_index();
    const int xcd = (int)blockIdx.x % NXCD, nlw = nw / NXCD;                 // blocks go to the XCDs round-robin
    const int lw = ((int)blockIdx.x / NXCD) * WAVES_PER_BLOCK + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int lane = (int)(threadIdx.x % WAVE);
    GridSync gs{A.ctl};
    gs.timeout = A.timeout;
    gs.trace = A.trace;
    gs.ntrace = A.ntrace;
    if (!gs.start()) return;
    for (int s = 0; s < A.nsteps; ++s) {
        const int p = s & 1;
        const char *pl = A.plan[p];
        for (int j = 0; j < A.nitems[p]; ++j) {
            const PItem it = A.item[p * PERSIST_ITEMS + j];
            const int *tab = (const int *)(pl + it.tab);
            // the phase's tiles split evenly over the XCDs, each XCD's share contiguous (the
            // halo rows of neighbouring segments then sit in one L2), its waves taking them in turn
            const int share = (it.tiles + NXCD - 1) / NXCD, x0 = xcd * share, x1 = min(x0 + share, it.tiles);
            for (int w = x0 + lw; w < x1; w += nlw) {
                int m, wl;
                batch_locate(tab, A.nm, w, m, wl);
                if (it.kind == 2) {
                    const DissArgs<F> a = ((const DissArgs<F> *)(pl + it.args))[m];
                    switch (A.diss) {                // drag (quadratic, linear, none) x (Smagorinsky, constant)
                    case 0: dissipation_body<F, PER, 0, true>(a, wl); break;
                    case 1: dissipation_body<F, PER, 0, false>(a, wl); break;
                    case 2: dissipation_body<F, PER, 1, true>(a, wl); break;
                    case 3: dissipation_body<F, PER, 1, false>(a, wl); break;
                    case 4: dissipation_body<F, PER, 2, true>(a, wl); break;
                    default: dissipation_body<F, PER, 2, false>(a, wl); break;
                    }
                } else {
                    const StageArgs<F> a = ((const StageArgs<F> *)(pl + it.args))[m];
                    if (PER && it.kind == 5) {
                        // (the bodies reload Prog for stage B: one register allocation serves
                        // every body here, and the extra rows would spill)
                        if (it.ma == M_RK_FIRST) fused_seam_body<F, M_RK_FIRST, M_RK_MID, false, false>(a, wl);
                        else fused_seam_body<F, M_RK_MID, M_RK_LAST, false, false>(a, wl);
                    } else if (it.ma == M_RK_FIRST) {
                        fused_body<F, M_RK_FIRST, M_RK_MID, false, PER, ADV_AH, false>(a, wl);
                    } else {
                        fused_body<F, M_RK_MID, M_RK_LAST, false, PER, ADV_AH, false>(a, wl);
                    }
                }
            }
            if (it.sync && !gs.sync()) return;
        }
        // the members' events of this step, in the reference's order: mid-point snapshot,
        // advection (tracer_advection.jl:12-42), NaN check (feedback.jl:136-148)
        bool any_snap = false, any_adv = false, both = false, any_nan = false;
        for (int m = 0; m < A.nm; ++m) {
            const PMember<F> &M = A.mem[m];
            const bool sn = M.tracer && (M.r_half + s) % M.nadvstep == 0;
            const bool ad = M.tracer && (M.r_adv + s) % M.nadvstep == 0;
            any_snap |= sn;
            any_adv |= ad;
            both |= sn && ad;
            any_nan |= M.nout > 0 && (M.r_nan + s) % M.nout == 0;
        }
        if (any_snap) {                      // um, vm <- T(u), T(v) of the step's output
            int base = 0;
            for (int m = 0; m < A.nm; ++m) {
                const PMember<F> &M = A.mem[m];
                if (!(M.tracer && (M.r_half + s) % M.nadvstep == 0)) continue;
                const int per_f = (int)((M.elems + SNAP_TILE - 1) / SNAP_TILE);
                for (int t = first_tile(w0, base, nw); t < 2 * per_f; t += nw) {
                    const bool isv = t >= per_f;
                    const Q *src = isv ? M.sv[p] : M.su[p];
                    ST *dst = isv ? M.vm : M.um;
                    const long long e0 = (long long)(isv ? t - per_f : t) * SNAP_TILE;
                    for (long long e = e0 + lane; e < e0 + SNAP_TILE && e < M.elems; e += WAVE)
                        stp(dst, (long)e, enc<ST>(toR<R>(src[e])));
                }
                base += 2 * per_f;
            }
            if (both && !gs.sync()) return;
        }
        if (any_adv) {                       // sst[dst] from sst[src], one row segment per tile
            int base = 0;
            for (int m = 0; m < A.nm; ++m) {
                const PMember<F> &M = A.mem[m];
                if (!(M.tracer && (M.r_adv + s) % M.nadvstep == 0)) continue;
                const int sx = (M.tr.nx + WAVE - 1) / WAVE, tiles = sx * M.tr.ny;
                TracerArgs<F> ta = M.tr;
                const int src = rsst_at(M, s);
                ta.u = M.tu[p];
                ta.v = M.tv[p];
                ta.sst = M.sst[src];
                ta.ssti = M.sst[1 - src];
                for (int t = first_tile(w0, base, nw); t < tiles; t += nw) {
                    const int y = 1 + t / sx, x = 1 + (t % sx) * WAVE + lane;
                    if (x <= ta.nx) tracer_cell<F, PER>(ta, x, y);
                }
                base += tiles;
            }
            if (any_nan && !gs.sync()) return;
        }
        if (any_nan) {                       // Prog of the step and the tracer after it
            int base = 0;
            for (int m = 0; m < A.nm; ++m) {
                const PMember<F> &M = A.mem[m];
                if (!(M.nout > 0 && (M.r_nan + s) % M.nout == 0)) continue;
                const NanArgs &N = M.nan[p];
                int nrow = 0;                // tiles per field: rows of the tallest
                for (int k = 0; k < 4; ++k) nrow = max(nrow, N.y1[k] - N.y0[k] + 1);
                const Q *f3 = M.sst[rsst_at(M, s + 1)];
                for (int t = first_tile(w0, base, nw); t < 4 * nrow; t += nw) {
                    const int k = t / nrow, y = N.y0[k] + t % nrow;
                    if (y > N.y1[k]) continue;
                    const Q *f = k == 3 ? f3 : (const Q *)N.f[k];
                    bool bad = false;
                    for (int x = N.x0[k] + lane; x <= N.x1[k]; x += WAVE) bad |= isnan(f[x + (long)y * N.P]);
                    if (__any(bad) && lane == 0) {
                        atomicOr(M.nanflag, 1);
                        atomicMin(&A.ctl->stop, s);
                    }
                }
                base += 4 * nrow;
            }
            if (!gs.sync()) return;
            if (__hip_atomic_load(&A.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= s) return;
        }
    }
}

}  // namespace swm
