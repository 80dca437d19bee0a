"""Benchmark: RK4 cell-updates/s of the ShallowWaters time step at 4096x4096 on one MI355X.

Workload: the nonperiodic double gyre at 4096x4096 (Lx = Ly = 3840 km, seamount, double-gyre
wind Fx0 = 0.12 Pa, free slip, Smagorinsky biharmonic diffusion, quadratic drag, tracer on,
RK4, Float32, dt = 13 s), started from a seeded smoothed η (σ = 0.1 m) with geostrophic
u, v.  A step is one full pass of the time loop (4 RK stages, ghost points, dissipation,
tracer when due) over the whole grid; the state stays resident in HBM.  The timed steps are
placed so that they contain a tracer advection (every 288 steps at this size).

Prints one JSON line.  `roofline` is for the kernel with the largest share of the step (a
fused pair of RK stages), timed with HIP events on the library's stream around each of its
launches inside the steps after the timed window (`kernel_ms_in_step`; `kernel_ms` holds
each kernel repeated on its own): `achieved` is the bytes that kernel must move (each field
it reads once, each output written once) over its average duration; `step_roofline` gives the whole step, also against the per-stage
synchronisation bytes of the reference's loop.  `cpu_baseline` times the CPU restatement
(oracle, one thread, the -O3 vectorised build) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX = 4096


def workload_parameter(swm, T="Float32", nx=NX, Tprog=None, **kw):
    return swm.workloads.double_gyre(nx=nx, T=T, Tprog=Tprog, **kw)


def bytes_per_cell_stage(es: int, est: int = None) -> dict:
    """Algorithmic HBM bytes per cell: each kernel reads every field it needs once and
    writes each output once.  es: bytes of a Tprog value (u, v, η, u0), est: of a T value
    (H, and the stage input u1 between RK stages when T differs from Tprog)."""
    est = es if est is None else est
    return {
        "stage1": 6 * es + 4 * est,        # read u,v,η (Tprog), H; write u1 (T), u0 (Tprog)
        "stage_mid": 9 * es + 7 * est,     # read u1 (T), H, u, u0; write u1 (T), u0
        "stage4": 6 * es + 4 * est,        # read u1 (T), H, u0; write u0
        "dissipation": 5 * es + est,       # read u0, v0, η0, H; write u, v
    }


def fused_bytes_per_cell(es: int, est: int = None) -> dict:
    """A fused pair of RK stages: the algorithmic bytes of the two stages it performs (each
    stage's sync-point minimum, as for the separate kernels) and the minimum the fused
    kernel itself moves (the intermediate stage stays in registers)."""
    b = bytes_per_cell_stage(es, est)
    est = es if est is None else est
    return {"fused12": {"alg": b["stage1"] + b["stage_mid"], "min": 6 * es + 4 * est},     # read u,v,η,H; write u1,u0
            "fused34": {"alg": b["stage_mid"] + b["stage4"], "min": 9 * es + 4 * est}}     # read u1,H,u,u0; write u0


# kernels of one step: unfused RK4 (periodic) and with fused stage pairs (nonperiodic)
STEP_KERNELS = {"unfused": ("stage1", "stage_mid", "stage_mid", "stage4", "dissipation"),
                "fused": ("fused12", "fused34", "dissipation")}


def timed_window(G, steps, warmup):
    """First warm-up step and first timed step: the timed steps contain the tracer's
    advection step (i = nadvstep) when the run has a tracer."""
    i1 = max(1 + warmup, G.nadvstep - steps // 2) if G.nadvstep > 0 else 1 + warmup
    return i1 - warmup, i1


def run_form(swm, T, Tprog, nx, steps, warmup, kernel_reps, **kw):
    """Time `steps` steps of the workload with right-hand side format T and state format
    Tprog (state resident in HBM), then the four step kernels with HIP events."""
    S = swm.make_setup(workload_parameter(swm, T=T, nx=nx, Tprog=Tprog, **kw))
    G = S.grid
    prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
    i0, i1 = timed_window(G, steps, warmup)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        done, nan = sim.step(warmup, i0, (i0 - 1) * G.dtint)
        assert nan == 0
        t0 = time.perf_counter()
        done, nan = sim.step(steps, i1, (i1 - 1) * G.dtint)   # returns after a stream sync
        elapsed = time.perf_counter() - t0
        assert (done, nan) == (steps, 0), (done, nan)
        mass, energy = sim.reduce()
        kms = {name: sim.time_kernel(k, kernel_reps)
               for name, k in (("stage_mid", 0), ("stage1", 1), ("stage4", 2), ("dissipation", 3))}
        if S.parameters.bc == "nonperiodic":      # the step runs stages 1+2 and 3+4 as fused pairs
            kms.update({name: sim.time_kernel(k, kernel_reps) for name, k in (("fused12", 5), ("fused34", 6))})
        # the same kernels timed inside the steps that follow the timed window (events around
        # each launch): what the step's kernels take in the order the step runs them
        i2 = i1 + steps
        kms["in_step"] = {k: ms for k, (ms, n) in sim.time_steps(kernel_reps, i2, (i2 - 1) * G.dtint).items()}
    return S, elapsed, kms, mass, energy


def run_ensemble(swm, nmem: int, steps: int, seq_steps: int, seed: int = 4):
    """The batched ensemble: nmem members with grid sizes drawn from {128, 256, 512, 1024}^2
    (L_ratio = 1), Fx0 ~ U(0.08, 0.16) Pa, H ~ U(400, 600) m, seeded smoothed η noise
    (σ = 0.05 m), Float32, RK4, the other parameters at their defaults.  Times `steps`
    steps of the tuned ensemble from the initial state as batched launches and graph
    replays (the default); the same steps in the persistent launch (every step, tracer and
    NaN-check steps included, in one cooperative kernel, its phases timed from inside); and
    `seq_steps` steps of each member on its own, one after the other."""
    import numpy as np
    setups = swm.workloads.ensemble_members(nmem, seed)
    sizes = np.array([S.grid.nx for S, _ in setups])
    cells = sum(S.grid.nx * S.grid.ny for S, _ in setups)
    sims = [swm.Simulation(S) for S, _ in setups]
    try:
        for sim, (_, p) in zip(sims, setups):
            sim.upload(p)
        # unbatched: each member alone
        t0 = time.perf_counter()
        for sim in sims:
            sim.step(seq_steps, 1, 0)
        seq = time.perf_counter() - t0
        for sim, (_, p) in zip(sims, setups):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            tuned = ens.tune(reps=3)
            t0 = time.perf_counter()
            done, nan, graph = ens.step(steps, 1)
            el = time.perf_counter() - t0
            assert done == steps and not any(nan), (done, nan)
            for sim, (_, p) in zip(sims, setups):
                sim.upload(p)
            ens.persistent = "timed"
            t0 = time.perf_counter()
            done, nan, _ = ens.step(steps, 1)
            el_p = time.perf_counter() - t0
            assert done == steps and not any(nan), (done, nan)
            persistent, why = ens.persistent_status()
            phases = ens.phase_ms()
    finally:
        for sim in sims:
            sim.close()
    value = cells * steps / el
    return {"members": nmem, "sizes": {str(int(n)): int((sizes == n).sum()) for n in (128, 256, 512, 1024)},
            "cells_per_step": cells, "steps": steps, "value": value, "unit": "cell-updates/s",
            "ms_per_step": el / steps * 1e3, "graph_steps": graph, "rows_per_wave_tuned": list(tuned),
            "persistent_launch": {"value": cells * steps / el_p, "ms_per_step": el_p / steps * 1e3,
                                  "steps_in_launch": persistent, "not_used_because": why or None,
                                  "phase_ms": phases},
            "unbatched_value": cells * seq_steps / seq, "speedup_vs_unbatched": value / (cells * seq_steps / seq)}


def adjoint_bytes_per_cell(es: int) -> dict:
    """Algorithmic HBM bytes per cell of the RK4 reverse step's kernels (each field read or
    written once): the last RK stage's adjoint reads the stage input, H and λ0 and writes
    μ; the middle ones also read μ of the later stage; the first reads the μ of all three
    later stages and writes the step input's adjoint; the dissipation adjoint reads the
    state before the dissipation, H and λ after it, and writes λ before it."""
    return {"stage_adj_last": 10 * es, "stage_adj_next": 13 * es, "stage_adj_mid": 13 * es, "stage_adj_first": 19 * es,
            "diss_adj": 10 * es}


def run_adjoint(swm, T, nx, nsteps, snapshots, reps):
    """The reverse mode on the workload with the tracer advecting (Uadv = 20 m/s: every 3
    steps): the step's kernels timed with HIP events, then a forward run of `nsteps` steps
    recorded with `snapshots` revolve checkpoints and its pullback for the loss of the final
    state (kinetic energy plus the η misfit at a seeded 1% of the cells against a perturbed
    twin run), giving the gradients with respect to the initial state and the wind stress
    field Fx; device time of the sweep.  Then the same with every state kept in HBM (one
    snapshot per step: the sweep recomputes only its taped steps).  Each forward run is timed
    on its second call, when the snapshot block exists."""
    import numpy as np
    S = swm.make_setup(workload_parameter(swm, T=T, nx=nx, Uadv=20.0))
    G = S.grid
    prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
    es = 4 if T == "Float32" else 8
    obs, mask = swm.workloads.observation_twin(S, prog, nsteps)
    def gradient(sim, nsnap):
        sim.upload(prog)
        sim.adjoint_forward(nsteps, 1, 0, nsnap)           # the snapshot block, allocated once
        sim.upload(prog)
        t0 = time.perf_counter()
        sim.adjoint_forward(nsteps, 1, 0, nsnap)
        fwd = time.perf_counter() - t0
        loss, ybar = swm.workloads.final_state_loss(S, sim.download(), obs, mask)
        t0 = time.perf_counter()
        g, nf = sim.pullback(ybar)
        gFx = sim.pullback_Fx()
        wall = time.perf_counter() - t0
        finite = all(bool(np.isfinite(a).all()) for a in (g.u, g.v, g.η, g.sst, gFx))
        return {"steps": nsteps, "snapshots": nsnap, "forward_steps_recomputed": nf,
                "forward_ms_per_step": fwd / nsteps * 1e3, "reverse_sweep_ms_per_step": sim.sweep_ms / nsteps,
                "gradient_ms_per_step": fwd / nsteps * 1e3 + sim.sweep_ms / nsteps,
                "gradient_over_forward": (fwd / nsteps * 1e3 + sim.sweep_ms / nsteps) / t["forward_step"],
                "pullback_wall_s_incl_host_transfers": wall, "gradient_finite": finite,
                "loss": loss, "gradients": ["eta0", "u0", "v0", "sst0", "Fx"], "tracer_every": G.nadvstep}

    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.adjoint_track_Fx(True)
        names = ("forward_step", "taped_step", "reverse_step", "stage_adjoint_mid", "dissipation_adjoint")
        t = {n: sim.adjoint_time(w, reps) for w, n in enumerate(names)}
        rev = gradient(sim, snapshots)
        rev_all = gradient(sim, nsteps - 1)                # every state kept in HBM: nothing recomputed
    cells = G.nx * G.ny
    b = adjoint_bytes_per_cell(es)
    step_bytes = sum(b.values()) * cells
    return {"kernel_ms": t, "reverse_over_forward": t["reverse_step"] / t["forward_step"],
            "taped_plus_reverse_over_forward": (t["taped_step"] + t["reverse_step"]) / t["forward_step"],
            "reverse_step_alg_bytes_per_cell": sum(b.values()),
            "reverse_step_GBs": step_bytes / (t["reverse_step"] * 1e-3) / 1e9,
            "stage_adjoint_mid_GBs": b["stage_adj_mid"] * cells / (t["stage_adjoint_mid"] * 1e-3) / 1e9,
            "dissipation_adjoint_GBs": b["diss_adj"] * cells / (t["dissipation_adjoint"] * 1e-3) / 1e9,
            "revolve": rev, "all_states_kept": rev_all}


def precision_report(swm, nx, days_list):
    """Errors of the Float32 and the Float16 / BFloat16 right-hand-side forms against the
    Float64 run of the same seeded state after each of `days_list` model days: relative L2
    of η, u, v over the interior, relative difference of total energy, and mass.  One JSON
    line per form and day, printed as each form finishes; a form whose run stops (NaN, a
    departure point off the grid) prints its error instead.

    Drift: the closed basin conserves Σ η exactly in exact arithmetic, so mass_drift_m, the
    change of the mean surface height since the initial state (Σ η / N, metres), is the
    rounding's doing; energy_change is the relative change of the total energy since the
    initial state (the flow is forced and damped, so energy_rel_f64 compares it with FP64's)."""
    import numpy as np
    # (T, Tprog, options): the scale of u, v and compensated RK increments are the
    # library's low-precision options.  Float16 intermediates of this state overflow for
    # s >= 8 (see scale_probe); BFloat16 has the Float32 range.
    forms = [("Float64", "Float64", {}), ("Float32", "Float32", {}), ("Float32", "Float32", {"compensated": True}),
             ("BFloat16", "Float32", {}), ("BFloat16", "Float32", {"scale": 64.0, "compensated": True}),
             ("BFloat16", "Float32", {"arithmetic": "Float32"}),
             ("Float16", "Float32", {}), ("Float16", "Float32", {"scale": 4.0, "compensated": True}),
             ("Float16", "Float32", {"arithmetic": "Float32"})]
    P0 = workload_parameter(swm, T="Float64", nx=nx)
    G0 = swm.make_setup(P0).grid
    marks = [int(np.ceil(d * 86400 / G0.dtint)) for d in days_list]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    ref = None
    for T, Tp, opt in forms:
        S = swm.make_setup(workload_parameter(swm, T=T, nx=nx, Tprog=Tp, **opt))
        prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
        snaps = {}
        t0 = time.perf_counter()
        try:
            with swm.Simulation(S) as sim:
                sim.upload(prog)
                snaps[0] = (None, None, None) + tuple(sim.reduce())
                done = 0
                for m in marks:
                    while done < m:
                        n = min(2000, m - done)
                        d, nan = sim.step(n, 1 + done, done * S.grid.dtint)
                        if nan or d != n:
                            raise RuntimeError(f"NaN after {done + d} steps")
                        done += n
                        print(f"# {T}/{Tp} {opt}: {done} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                              flush=True)
                    g = sim.download()
                    mass, energy = sim.reduce()
                    u, v, e, _ = swm.remove_halo(g.u, g.v, g.η, g.sst, S)
                    snaps[m] = (swm.to_float64(u, Tp), swm.to_float64(v, Tp), swm.to_float64(e, Tp), mass, energy)
        except RuntimeError as err:                     # includes the library's SwmError
            print(json.dumps({"T": T, "Tprog": Tp, "options": opt, "error": str(err)}), flush=True)
            continue
        if T == "Float64":
            ref = snaps
            continue
        N = G0.nx * G0.ny
        m0, e0 = snaps[0][3], snaps[0][4]
        for d, m in zip(days_list, marks):
            u, v, e, mass, energy = snaps[m]
            ur, vr, er, mr, enr = ref[m]
            print(json.dumps({"T": T, "Tprog": Tp, "options": opt, "nx": nx, "days": d, "steps": m,
                              "rel_l2_eta": rel(e, er), "rel_l2_u": rel(u, ur), "rel_l2_v": rel(v, vr),
                              "rel_energy": (energy - enr) / enr, "mass": mass, "mass_f64": mr,
                              "mass_drift_m": (mass - m0) / N, "mass_drift_f64_m": (mr - ref[0][3]) / N,
                              "energy_change": (energy - e0) / e0, "energy_change_f64": (enr - ref[0][4]) / ref[0][4]}),
                  flush=True)


def cpu_protocol(swm, orc, config2_cpu_steps: int = 100):
    """The CPU-beside-GPU table: the oracle (one thread) and the GPU on the same setups, one
    JSON line per row.  4096^2 workload: 5 timed steps after one warm-up step, Float32 and
    Float64; the README's 100x100 double gyre over 100 days (30,639 steps from rest, no
    tracer); the 2048x1024 periodic channel (tracer every 300 steps, seeded state),
    `config2_cpu_steps` steps on the CPU and 1000 on the GPU."""
    def row(name, S, prog, steps, cpu_steps, warm=1):
        cfg = swm.make_config(S)
        G = S.grid
        out = {"row": name, "nx": G.nx, "ny": G.ny, "T": S.parameters.T, "steps_gpu": steps, "steps_cpu": cpu_steps}
        if cpu_steps > 0:
            if warm:
                orc.run(S, cfg, prog, warm)
            t = time.perf_counter()
            orc.run(S, cfg, prog, cpu_steps)
            dt = time.perf_counter() - t
            out.update(cpu_cell_updates_per_s=G.nx * G.ny * cpu_steps / dt, cpu_seconds=dt, cpu_cores=1)
        with swm.Simulation(S) as sim:
            sim.upload(prog)
            sim.step(min(warm * 10, steps), 1, 0)
            sim.upload(prog)
            t = time.perf_counter()
            d, nan = sim.step(steps, 1, 0)
            dt = time.perf_counter() - t
        out.update(gpu_cell_updates_per_s=G.nx * G.ny * steps / dt, gpu_seconds=dt, gpu_steps_done=d, nan=nan)
        print(json.dumps(out), flush=True)

    for T in ("Float32", "Float64"):
        S = swm.make_setup(workload_parameter(swm, T=T))
        row(f"double_gyre_4096x4096 {T}", S, swm.seeded_initial_conditions(S, seed=2024, σ=0.1), 200, 5)
    S1 = swm.make_setup(swm.workloads.readme_double_gyre())
    row("readme_double_gyre_100x100 100 days", S1, swm.initial_conditions("Float32", S1), S1.grid.nt, S1.grid.nt)
    S2 = swm.make_setup(channel_2048(swm))
    row("channel_2048x1024 periodic", S2, swm.seeded_initial_conditions(S2, seed=2024, σ=0.1), 1000,
        config2_cpu_steps)


def channel_2048(swm, T="Float32", topography="flat"):
    """The 2048x1024 periodic channel (flat bottom, channel wind, tracer every 300 steps)."""
    return swm.workloads.channel(T, topography)


def channel_run(swm, T, steps, warmup, kernel_reps, topography="flat"):
    """The periodic channel alone on the GPU: `steps` timed steps from the seeded state and the
    step kernels timed with HIP events (the step runs its RK stages as fused pairs wrapping
    around the seam; with a topography whose H differs across the seam, the seam strips run
    on their own kernel beside them).  One JSON line."""
    S = swm.make_setup(channel_2048(swm, T, topography))
    G = S.grid
    with swm.Simulation(S) as sim:
        sim.upload(swm.seeded_initial_conditions(S, seed=2024, σ=0.1))
        sim.step(warmup, 1, 0)
        t = time.perf_counter()
        d, nan = sim.step(steps, 1 + warmup, warmup * G.dtint)
        dt = time.perf_counter() - t
        kms = {name: sim.time_kernel(k, kernel_reps) for name, k in (("fused12", 5), ("fused34", 6), ("dissipation", 3))}
    print(json.dumps({"workload": f"channel_{G.nx}x{G.ny}_periodic_{topography}", "T": T, "steps": steps, "steps_done": d,
                      "nan": nan, "ms_per_step": dt / steps * 1e3, "cell_updates_per_s": G.nx * G.ny * steps / dt,
                      "kernel_ms": kms}), flush=True)


def scale_probe(swm, nx):
    """How far the 16-bit right-hand sides run at `nx`² with velocities carried scaled by s,
    with and without compensated increments: the step count reached before the state turns
    NaN or a tracer departure point leaves the grid (one model day at most)."""
    day = int(round(86400 / swm.make_setup(workload_parameter(swm, nx=nx)).grid.dtint))
    for T, scales in (("Float16", (1.0, 2.0, 4.0, 8.0, 16.0, 64.0)), ("BFloat16", (1.0, 64.0))):
        for sc in scales:
            for comp in (False, True):
                S = swm.make_setup(workload_parameter(swm, T=T, nx=nx, Tprog="Float32", scale=sc, compensated=comp))
                prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
                done, status = 0, "ok"
                with swm.Simulation(S) as sim:
                    sim.upload(prog)
                    try:
                        for chunk in (1, 9, 90, 900, day - 1000):
                            d, nan = sim.step(chunk, 1 + done, done * S.grid.dtint)
                            done += d
                            if nan or d != chunk:
                                status = f"NaN by step {done}"
                                break
                    except RuntimeError as err:             # includes the library's SwmError
                        status = f"stopped after step {done}: {err}"
                print(json.dumps({"T": T, "Tprog": "Float32", "scale": sc, "compensated": comp, "steps_run": done,
                                  "status": status}), flush=True)


def cpu_baseline(swm, orc, sample_steps: int, nx: int):
    """The oracle on one host thread, both builds (oracle/Makefile): the -O3 vectorised one
    (x86-64-v3) is the reported baseline, the -O2 one a labelled second row.  Same
    arithmetic (no FMA contraction in either), so the same results."""
    P = workload_parameter(swm, nx=nx)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
    cfg = swm.make_config(S)
    host = f"{platform.processor() or platform.machine()}, nproc={os.cpu_count()}"
    rows = {}
    for variant, flags in (("o3", "g++ -O3 -march=x86-64-v3"), ("o2", "g++ -O2")):
        orc.use(variant)
        orc.run(S, cfg, prog, 1)                     # warm-up
        t = time.perf_counter()
        orc.run(S, cfg, prog, sample_steps)
        dt = time.perf_counter() - t
        rows[variant] = {"value": S.grid.nx * S.grid.ny * sample_steps / dt, "unit": "cell-updates/s", "cores": 1,
                         "kind": "port", "seconds": dt,
                         "sample": f"{sample_steps} steps of the same workload at {S.grid.nx}x{S.grid.ny}, "
                                   f"oracle/swm_oracle.cpp ({flags}, -ffp-contract=off, one thread) on {host}"}
    orc.use("o2")
    out = dict(rows["o3"])
    out["o2_build"] = rows["o2"]
    return out


def measured_traffic(workload: str, dtype: str, kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs kept
    in profiles/ (KiB per dispatch, each scaled by what the counter reads for a plain copy
    of known size); None when no measurement of this workload and kernel is recorded."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}_{dtype}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = json.load(f)
    k = m.get("kernels", {}).get(kernel)
    if k is None:
        return None
    return 1024.0 * (k["fetch_size_kib"] * m["fetch_scale"] + k["write_size_kib"] * m["write_scale"])


def copy_probe(mib: int, reps: int = 20):
    """A plain copy of `mib` MiB (read mib, write mib), timed with events: the calibration
    kernel for the HBM traffic counters and the practical copy bandwidth of the device."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "ms": ms,
                      "GBs": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=NX)
    ap.add_argument("--dtype", default="Float32", choices=["Float32", "Float64"])
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for cpu_baseline (0 skips)")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--form-steps", type=int, default=100,
                    help="steps timed for each reduced-precision right-hand-side form (0 skips them)")
    ap.add_argument("--precision-days", type=str, default="",
                    help="comma-separated model days: only run the Float32 / Float16 / BFloat16 forms against "
                         "Float64 and print their errors at those days, then exit")
    ap.add_argument("--cpu-protocol", action="store_true",
                    help="only print the CPU-beside-GPU rows (oracle on one thread and the GPU on the 4096^2 "
                         "workload in Float32 and Float64, the README's 100x100 100-day run, the 2048x1024 "
                         "periodic channel), then exit")
    ap.add_argument("--channel", action="store_true",
                    help="only time the 2048x1024 periodic channel (--steps, --warmup, --dtype) and its kernels")
    ap.add_argument("--topography", default="flat", choices=["flat", "ridge", "ridges", "seamount"],
                    help="topography of --channel")
    ap.add_argument("--scale-probe", action="store_true",
                    help="only report how far the Float16 / BFloat16 right-hand sides run with velocities scaled "
                         "by s = 1 .. 64, then exit")
    ap.add_argument("--ensemble-steps", type=int, default=2000,
                    help="steps timed for the 64-member batched ensemble (0 skips it)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only time the 64-member ensemble (--ensemble-steps on both paths), print it and exit")
    ap.add_argument("--ensemble-seq-steps", type=int, default=20,
                    help="steps of each ensemble member run on its own, for the unbatched comparison")
    ap.add_argument("--adjoint-nx", type=str, default="1024,4096",
                    help="grid sizes (comma-separated) of the reverse-mode measurement: the adjoint/forward "
                         "ratio is quoted at 1024^2 (200 steps, revolve), 4096^2 is the forward workload")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the revolve pullback measurement (0 skips the reverse-mode line)")
    ap.add_argument("--adjoint-snapshots", type=int, default=20)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only time a plain device-to-device copy of this many MiB (the reference kernel of "
                         "known traffic that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the time step is "
              f"benchmarked on a single MI355X; multi-GPU sharding is a design note, not a run).", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_oracle, load_package
    swm = load_package()
    if args.precision_days:
        precision_report(swm, args.nx, [float(d) for d in args.precision_days.split(",")])
        return
    if args.scale_probe:
        scale_probe(swm, args.nx)
        return
    if args.cpu_protocol:
        cpu_protocol(swm, load_oracle())
        return
    if args.channel:
        channel_run(swm, args.dtype, args.steps, args.warmup, args.kernel_reps, args.topography)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": run_ensemble(swm, 64, args.ensemble_steps, args.ensemble_seq_steps)}), flush=True)
        return
    es = 4 if args.dtype == "Float32" else 8
    # kernel times with HIP events on the library stream
    S, elapsed, kms, mass, energy = run_form(swm, args.dtype, args.dtype, args.nx, args.steps, args.warmup,
                                             args.kernel_reps)
    G = S.grid
    cells = G.nx * G.ny
    value = cells * args.steps / elapsed
    bpc = bytes_per_cell_stage(es)
    fb = fused_bytes_per_cell(es)
    alg = {**bpc, **{k: v["alg"] for k, v in fb.items()}}                 # bytes of the stages a kernel performs
    moved = {**bpc, **{k: v["min"] for k, v in fb.items()}}               # minimum bytes a kernel itself moves
    step = STEP_KERNELS["fused" if "fused12" in kms else "unfused"]
    kin = kms.pop("in_step")                                           # durations inside the step
    kdur = {k: kin.get(k, kms[k]) for k in step}
    dom = max(set(step), key=lambda k: kdur[k] * step.count(k))        # largest share of the step
    traffic = measured_traffic(f"double_gyre_{G.nx}x{G.ny}", "f32" if es == 4 else "f64", dom)
    achieved = moved[dom] * cells / (kdur[dom] * 1e-3) / 1e9
    step_bytes = sum(moved[k] for k in step) * cells
    ref_bytes = sum(bpc[k] for k in STEP_KERNELS["unfused"]) * cells   # per-stage sync points (232 B at FP32)
    out = {
        "metric": "RK4 cell-updates/s at 4096x4096 (nonperiodic double gyre)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if es == 4 else "f64",
        "data": "synthetic: seeded smoothed η (σ=0.1 m, seed 2024) with geostrophic u,v; seeded sinusoid plus uniform noise sst",
        "config": {"workload": f"double_gyre_{G.nx}x{G.ny}", "nx": G.nx, "ny": G.ny, "Lx_m": G.Lx, "bc": "nonperiodic",
                   "topography": "seamount", "wind": "double_gyre Fx0=0.12", "diffusion": "Smagorinsky",
                   "bottom_drag": "quadratic", "alpha": 0.0, "time_scheme": "RK4", "dt_s": G.dtint,
                   "tracer_every": G.nadvstep, "timed_steps_from": timed_window(G, args.steps, args.warmup)[1],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "bytes_per_cell": moved[dom], "kernel_ms": kdur[dom], "kernel_timing": "in step (HIP events)",
                     "traffic_GBs": traffic / (kdur[dom] * 1e-3) / 1e9 if traffic else None,
                     "stages_sync_bytes_per_cell": alg[dom],
                     "stages_sync_equivalent_GBs": alg[dom] * cells / (kdur[dom] * 1e-3) / 1e9},
        "step_roofline": {"kernels": list(step), "alg_bytes_per_cell_step": step_bytes / cells,
                          "achieved_GBs": step_bytes / (elapsed / args.steps) / 1e9,
                          "frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
                          "per_stage_sync_bytes_per_cell_step": ref_bytes / cells,
                          "per_stage_sync_equivalent_GBs": ref_bytes / (elapsed / args.steps) / 1e9},
        "kernel_ms": kms,
        "kernel_ms_in_step": kin,
        "diagnostics": {"mass": mass, "energy": energy},
    }
    if args.form_steps > 0 and args.dtype == "Float32":
        # the reduced-precision right-hand-side forms: Tprog = Float32 state, T = BFloat16 / Float16
        out["forms"] = {}
        for name, T, opt in (("BFloat16", "BFloat16", {}), ("Float16", "Float16", {}),
                             ("BFloat16+scale+compensated", "BFloat16", {"scale": 64.0, "compensated": True}),
                             ("Float16+scale+compensated", "Float16", {"scale": 4.0, "compensated": True}),
                             # BFloat16 / Float16 values with Float32 arithmetic
                             ("BFloat16-values+Float32-arithmetic", "BFloat16", {"arithmetic": "Float32"}),
                             ("Float16-values+Float32-arithmetic", "Float16", {"arithmetic": "Float32"})):
            Sf, el, km, _, _ = run_form(swm, T, "Float32", args.nx, args.form_steps, args.warmup, args.kernel_reps, **opt)
            b = bytes_per_cell_stage(4, 2)
            fstep = STEP_KERNELS["fused" if "fused12" in km else "unfused"]
            out["forms"][name] = {"value": cells * args.form_steps / el, "ms_per_step": el / args.form_steps * 1e3,
                               "speedup_vs_f32": (elapsed / args.steps) / (el / args.form_steps),
                               "kernel_ms": km, "bytes_per_cell_step": sum(
                                   {**b, **{k: v["min"] for k, v in fused_bytes_per_cell(4, 2).items()}}[k] for k in fstep)}
    if args.ensemble_steps > 0 and args.dtype == "Float32":
        out["ensemble"] = run_ensemble(swm, 64, args.ensemble_steps, args.ensemble_seq_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = {f"{n}x{n}": run_adjoint(swm, args.dtype, n, args.adjoint_steps, args.adjoint_snapshots,
                                                  max(5, args.kernel_reps // 2))
                          for n in (int(a) for a in args.adjoint_nx.split(","))}
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(swm, load_oracle(), args.cpu_steps, args.nx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
