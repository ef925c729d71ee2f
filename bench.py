#!/usr/bin/env python3
"""Headline benchmark: the L=18 FP64 Chebyshev propagator on one MI355X.

Workload (the single-large-state configuration, whose 64 realisations run
twelve at a time): 2^18 complex FP64 amplitudes per realisation, all-to-all
secular dipolar XXZ couplings (153 pairs) from random 2D positions at unit
density (min spacing 0.5, (1 - 3cos^2 theta)/r^3) plus Gaussian Hz disorder
(sigma 0.5), real-time steps of dt = 0.05 at epsilon = 1e-12,
autonormalisation off. One bench "step" is one Chebyshev time step of each of
the `--set` realisations (default 12), stepped together: every order is one
launch sequence for the set, each realisation with its own bound and order.

Prints ONE JSON line. `value` is algorithmic GB/s of the whole timed region
(inputs resident in HBM, 88 B per amplitude per general order); `roofline`
prices the dominant launch (a general Chebyshev order over the set) against
the 8 TB/s HBM peak; `single_state` gives one realisation alone; the CPU
baseline runs the same per-order work through the reference's own kernels
(oracle/_ref) or the oracle restatement, on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·ψ GB/s vs HBM peak at L=18; s to ±1e-3 echo error bar, CV vs plain MC"
HBM_PEAK_GBS = 8000.0
# HBM bytes per general order over the set at L=18, keyed by (order kernels,
# states in the set, states per launch): FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (factors from the in-place scale kernel of known size), summed over the
# set's launches and averaged over the benchmark's own steps;
# profiles/k_dense_L18_order.md, profiles/k_wht_L18_8states_order.md and
# profiles/k_wht_L18_12states_order.md.
WHT = "k_wht_cols<8> + k_wht_rows_bwd<K_ORDERN>"
MEASURED_TRAFFIC = {("k_dense<K_ORDERN, 18, 10>", 1, 1): 42417 * 1024,
                    (WHT, 8, 8): 384722 * 1024,
                    (WHT, 8, 4): 385260 * 1024,
                    (WHT, 12, 12): 576850 * 1024,
                    (WHT, 12, 6): 577387 * 1024}
# where each figure can be reproduced, keyed by (order kernels, states in the
# set): the kernel-trace summary whose per-dispatch means add up to
# `us_per_launch`, and the PMC runs behind `traffic` (scaled FETCH_SIZE +
# WRITE_SIZE, not measured by this run)
TRACE_PROFILE = {("k_dense<K_ORDERN, 18, 10>", 1): "profiles/k_dense_L18_order.md",
                 (WHT, 8): "profiles/k_wht_L18_8states_order.md",
                 (WHT, 12): "profiles/k_wht_L18_12states_order.md"}
TRAFFIC_PROFILE = TRACE_PROFILE
BYTES_PER_AMP_ORDER = 88          # read T_{n-1}, T_{n-2}, diag; write T_n; read+write psi
BYTES_FIRST, BYTES_SECOND = 40, 72  # order 1 (psi -> T_1), order 2 (fused psi update)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


SEED = 20240726


def synthetic_system(L: int, seed: int = SEED, index: int = 0, sigma: float = 0.5, min_sep: float = 0.5):
    """Realisation (seed, index) of the dipolar ensemble, drawn by libspx in
    the reference's order (secular (1 - 3cos^2 theta)/r^3 couplings at unit
    density, min spacing 0.5, Gaussian Hz of spread sigma); a normalised
    Gaussian initial state from numpy."""
    from spinsim_amd import spx
    d = spx.dipolar_realisation(seed, index, L, 1.0, min_sep, True, sigma)
    fields = np.zeros((L, 3))
    fields[:, 2] = d.hz
    psi = np.random.default_rng([seed, index]).standard_normal(2 << L)
    psi /= np.linalg.norm(psi)
    return fields, d.lo, d.hi, d.jx, d.jy, d.jz, psi


def step_bytes(order: int, n: int) -> int:
    if order <= 0:
        return 0
    if order == 1:
        return (BYTES_FIRST + 48) * n
    return (BYTES_FIRST + BYTES_SECOND + BYTES_PER_AMP_ORDER * (order - 2)) * n


def cpu_baseline(L, fields, lo, hi, jx, jy, jz, order_bytes, target_s=15.0):
    """The same per-order work on the host: negate + accumulateTwiceNormalized
    + addScaled (src/core/Propagator.cpp:58-72), one independent state per
    thread, like LocalRunner's realisation pool. Bounded sample."""
    from oracle import oracle as O
    threads = max(1, len(os.sched_getaffinity(0)))
    f = np.ascontiguousarray(fields)
    probe_orders = 2
    if O.ref_available():
        R = O.ref()
        kind = "reference"

        def run(orders):
            return R.ref_time_orders(L, O._d(f), lo.size, O._u32(lo), O._u32(hi), O._d(jx), O._d(jy),
                                     O._d(jz), orders, threads)
    else:
        kind = "port"
        h = O.Ham(L, fields, O.Couplings(lo, hi, jx, jy, jz))

        def run(orders):
            return O.lib().so_time_orders(h.h, orders, threads)
    t = run(probe_orders)
    per = t / probe_orders
    orders = max(2, int(target_s / max(per, 1e-6)))
    wall = run(orders)
    total_orders = orders * threads
    gbs = total_orders * order_bytes / wall / 1e9
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), model)
    except OSError:
        pass
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"{orders} Chebyshev orders x {threads} threads (one L={L} state each), "
                      f"{wall:.1f} s wall; {wall / orders * 1e3:.1f} ms per order per thread",
            "ms_per_order_per_thread": round(wall / orders * 1e3, 3),
            "cpu_model": model, "host_cpus_visible": len(os.sched_getaffinity(0))}


# ------------------------------------------------ control-variate leg

def control_variate_leg(spx, ctx):
    """Seconds to a +-1e-3 error bar on every point of the mean XY-16 echo at
    L = 16: plain batched Monte-Carlo against the control variate (the
    surrogate on 1e6 random geometries plus the exact correction), each arm
    run until its own standard error meets the bar; the one-shot cost with
    the surrogate's training counted is reported beside it
    (spinsim_amd/ensembles.py)."""
    from spinsim_amd import ensembles as E
    return E.control_variate(ctx, log=lambda m: log(f"[bench] {m}")).line


def config0_leg(spx, ctx, count=20):
    """The plain CPMG echo configuration on the device: L = 10, pi/2_x then
    8 x [10 x 0.05, ideal pi_y, 10 x 0.05, measure], corr:x and norm,
    `count` realisations in one batch (the LDS-resident launch)."""
    from spinsim_amd import ensembles as E
    b = spx.Batch(ctx, E.cpmg_program(10, 8), [(spx.OBS_CORR_X, 0), (spx.OBS_NORM, 0)], epsilon=1e-7,
                  reference_axis=0)
    b.add_dipolar(E.SEED, 0, count, 10, **E.DRAW)
    b.run()   # warm: module load and first-launch costs
    ctx.synchronize()
    t0 = time.perf_counter()
    b.run()
    dt = time.perf_counter() - t0
    return {"workload": f"L=10 CPMG, 8 echoes, {count} realisations", "s": round(dt, 4),
            "realisations_per_s": round(count / dt, 1)}


def cpu_realisation_rates(threads, target_s=8.0):
    """Realisations per second of the configurations' exact runs on the
    host's cores: the oracle's runRealization (the reference's algorithm and
    flags), one realisation per thread at a time as LocalRunner runs them.
    Bounded samples; best and median of repeats where a sample is short."""
    from oracle import oracle as O
    out = {}

    def timed(cfg, n):
        t0 = time.perf_counter()
        O.run_ensemble(cfg, n, threads=threads)
        return time.perf_counter() - t0

    # plain CPMG echo, L = 10, 20 realisations
    c0 = O.dipolar_config(10, O.cpmg_program(10, 8))
    reps = [timed(c0, 20)]
    while sum(reps) < target_s and len(reps) < 9:
        reps.append(timed(c0, 20))
    out["cpmg_L10"] = {"value": round(20 / min(reps), 2), "median": round(20 / float(np.median(reps)), 2),
                       "unit": "realisations/s", "sample": f"20 realisations x {len(reps)} repeats, best and median"}
    # uneven L in 8..12, XY-8 x 16: one realisation per thread at each size
    tot, n = 0.0, 0
    for L in range(8, 13):
        tot += timed(O.dipolar_config(L, O.xy8_program(L)), threads)
        n += threads
    out["xy8_L8_12"] = {"value": round(n / tot, 3), "unit": "realisations/s",
                        "sample": f"{threads} realisations at each L in 8..12 ({n}), {tot:.1f} s"}
    # L = 16 XY-16 with finite pulses: one realisation per thread
    t4 = timed(O.dipolar_config(16, O.xy16_program(16)), threads)
    out["xy16_L16"] = {"value": round(threads / t4, 3), "unit": "realisations/s",
                       "sample": f"{threads} realisations, {t4:.1f} s"}
    return out


def xy8_program(spx, blocks=16, omega=50.0, free_steps=10):
    """XY-8 (X Y X Y Y X Y X) with finite pulses: each pulse an evolve of
    tau_p = pi/omega under hx or hy = omega on every spin, preceded by free
    evolution free_steps x 0.05; a measure after each block."""
    P = spx.Program()
    tp = math.pi / omega
    P.measure()
    P.repeat(blocks)
    for ax in "XYXYYXYX":
        P.evolve(0.05, free_steps)
        hx, hy = (omega, 0.0) if ax == "X" else (0.0, omega)
        P.evolve_uniform(tp, 1, hx=hx, hy=hy, hz=0.0)
    P.measure()
    P.end()
    return P


def batched_leg(spx, ctx, count=4096, seed=20240726, sizes=None):
    """One launch of `count` realisations at uneven L in {8..12} (L from a
    hash of the index, no draws taken from the realisation stream), XY-8
    repeated 16 times, corr:x and norm per block: realisations per second.
    `sizes` (a list of L) replaces the hash: realisation i gets
    sizes[i % len(sizes)]."""
    obs = [(spx.OBS_CORR_X, 0), (spx.OBS_NORM, 0)]
    b = spx.Batch(ctx, xy8_program(spx), obs, epsilon=1e-7, reference_axis=0)
    g = np.random.default_rng(seed)
    Ls = []
    for i in range(count):
        L = 8 + ((i * 2654435761) >> 16) % 5 if sizes is None else sizes[i % len(sizes)]
        d = spx.dipolar_realisation(seed, i, L)
        f = np.zeros((L, 3))
        f[:, 2] = d.hz
        psi = g.standard_normal(2 << L)
        psi /= np.linalg.norm(psi)
        b.add(L, f, d.lo, d.hi, d.jx, d.jy, d.jz, psi)
        Ls.append(L)
    ctx.synchronize()
    t0 = time.perf_counter()
    _, v = b.run()
    dt = time.perf_counter() - t0
    if not np.isfinite(v).all():
        raise SystemExit("non-finite batched ensemble output")
    spread = "L uniform in 8..12" if sizes is None else "L in " + ",".join(map(str, sizes))
    return {"workload": f"{count} realisations, {spread}, XY-8 (finite pulses, Omega=50, tau=0.5) x 16, "
                        "corr:x + norm per block, one persistent launch for every size",
            "realisations": count, "s": round(dt, 4), "realisations_per_s": round(count / dt, 1),
            "mean_L": round(float(np.mean(Ls)), 2)}


def single_state_leg(spx, ctx, L, system, args):
    """The first realisation alone (one state per launch, the schedule tuned
    at width 1): algorithmic GB/s of its steps and its general order."""
    fields, lo, hi, jx, jy, jz, psi0 = system
    ham = spx.Hamiltonian(ctx, L)
    ham.set_fields(fields)
    ham.set_couplings(lo, hi, jx, jy, jz)
    coeffs, eff = spx.cheb_build(spx.REAL, ham.spectral_bound(), args.dt, args.epsilon)
    if not args.no_tune:
        ham.autotune(eff, nvec=1)
    order = coeffs.size - 1
    n = 1 << L
    psi = spx.StateVector(ctx, L, psi0)
    prop = spx.Propagator(ctx, L, args.epsilon)
    prop.steps(ham, spx.REAL, psi, coeffs, eff, args.warmup)
    ctx.record(5)
    prop.steps(ham, spx.REAL, psi, coeffs, eff, args.steps)
    ctx.record(6)
    reps = max(20, args.steps // 4)
    short = coeffs[: max(3, order // 2) + 1]
    prop.steps(ham, spx.REAL, psi, coeffs, eff, reps)
    ctx.record(7)
    prop.steps(ham, spx.REAL, psi, short, eff, reps)
    ctx.record(8)
    ctx.synchronize()
    ms = ctx.elapsed_ms(5, 6)
    per_order_ms = (ctx.elapsed_ms(6, 7) - ctx.elapsed_ms(7, 8)) / (reps * (order - (short.size - 1)))
    return {"value": round(step_bytes(order, n) * args.steps / (ms * 1e-3) / 1e9, 2), "unit": "GB/s",
            "order": int(order), "kernel": ham.kernel_name(),
            "us_per_order": round(per_order_ms * 1e3, 3),
            "order_GBs": round(BYTES_PER_AMP_ORDER * n / (per_order_ms * 1e-3) / 1e9, 2)}


def copy_probe(spx, L, reps=50):
    """In-place scale (read + write 16 B per amplitude): the plain streaming
    kernel that calibrates the HBM counters and the achievable bandwidth."""
    ctx = spx.Context(0)
    st = spx.StateVector(ctx, L, np.ones(2 << L) / math.sqrt(float(2 << L)))
    out = {"probe": "k_scale", "nspins": L}
    for _ in range(5):
        st.scale(1.0)
    ctx.record(0)
    for _ in range(reps):
        st.scale(1.0)
    ctx.record(1)
    ctx.synchronize()
    us = ctx.elapsed_ms(0, 1) * 1e3 / reps
    nbytes = 32 * (1 << L)
    out.update({"bytes_per_launch": nbytes, "us_per_launch": round(us, 3),
                "GB/s": round(nbytes / (us * 1e-6) / 1e9, 1)})
    return out


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nspins", type=int, default=18)
    ap.add_argument("--epsilon", type=float, default=1e-12)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cv", action="store_true", help="skip the control-variate leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch schedule")
    ap.add_argument("--set", type=int, default=12, metavar="S",
                    help="realisations stepped together, one launch per order (1..16; 1 = a single state)")
    ap.add_argument("--graphs", action="store_true", help="replay each step as a captured HIP graph")
    ap.add_argument("--streams", type=int, default=2, choices=(1, 2),
                    help="2: a set of 4..16 runs as two halves on two streams (the default); 1: one stream")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="L",
                    help="only time the in-place scale kernel on 2^L amplitudes")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the 4096-realisation uneven-L batch (its JSON object alone)")
    ap.add_argument("--batched-count", type=int, default=4096, metavar="R",
                    help="with --batched-only: realisations in the batch")
    ap.add_argument("--batched-sizes", default="", metavar="L[,L...]",
                    help="with --batched-only: cycle through these sizes instead of L = 8..12 by hash")
    return ap


def main():
    args = build_parser().parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(realisations would shard across GPUs with a host-side index-ordered fold)",
              file=sys.stderr)
        sys.exit(2)

    from spinsim_amd import spx
    if args.copy_probe:
        print(json.dumps(copy_probe(spx, args.copy_probe)), flush=True)
        return
    if args.batched_only:
        sizes = [int(x) for x in args.batched_sizes.split(",")] if args.batched_sizes else None
        if sizes is not None and not all(1 <= x <= 12 for x in sizes):
            raise SystemExit("--batched-sizes takes L in 1..12 (the resident runner's range)")
        print(json.dumps(batched_leg(spx, spx.Context(0), count=args.batched_count, sizes=sizes)), flush=True)
        return
    L = args.nspins
    n = 1 << L
    S = args.set
    if not 1 <= S <= 16:
        raise SystemExit("--set takes 1..16 realisations per launch")
    systems = [synthetic_system(L, index=i) for i in range(S)]
    fields, lo, hi, jx, jy, jz, _ = systems[0]
    log(f"[bench] L={L} pairs={lo.size} realisations={S} building device objects")
    ctx = spx.Context(0)
    hams, states = [], []
    for f, l, h_, x, y, z, p0 in systems:
        ham = spx.Hamiltonian(ctx, L)
        ham.set_fields(f)
        ham.set_couplings(l, h_, x, y, z)
        hams.append(ham)
        states.append(spx.StateVector(ctx, L, p0))
    built = [spx.cheb_build(spx.REAL, h.spectral_bound(), args.dt, args.epsilon) for h in hams]
    coeffs = [c for c, _ in built]
    effs = [e for _, e in built]
    orders = [c.size - 1 for c in coeffs]
    # states per launch: with two streams a set of 4..8 runs as two halves
    W = (S + 1) // 2 if args.streams == 2 and S >= 4 else S
    tune_report = None
    if not args.no_tune:
        # a set evolve settles each Hamiltonian's path at the set's
        # one-stream width S before it narrows to halves of W, so the
        # schedule is tuned at width S; the other realisations adopt it from
        # the tuner's cache (same shape and width)
        tune_report = hams[0].autotune(effs[0], nvec=S)
        spx.autotune_mode(1)
        log(f"[bench] autotune (width {S}): {tune_report[:300]}")
    prop = spx.Propagator(ctx, L, args.epsilon)
    prop.set_graphs(args.graphs)
    prop.set_streams(args.streams)
    log(f"[bench] bounds={[round(h.spectral_bound(), 3) for h in hams]} orders={orders}")

    prop.steps_set(hams, spx.REAL, states, coeffs, effs, args.warmup)
    ctx.synchronize()
    kernels = [h.kernel_name() for h in hams]
    log(f"[bench] warmup {args.warmup} steps done; kernels {kernels}")

    ctx.record(0)
    prop.steps_set(hams, spx.REAL, states, coeffs, effs, args.steps)
    ctx.record(1)
    ctx.synchronize()
    ms = ctx.elapsed_ms(0, 1)
    log(f"[bench] timed {args.steps} steps of {S} realisations: {ms:.2f} ms")
    total_bytes = sum(step_bytes(o, n) for o in orders) * args.steps
    value = total_bytes / (ms * 1e-3) / 1e9

    # Dominant launch: a general order over the S states (T_n = 2 Hhat T_{n-1}
    # - T_{n-2}; psi += c T_n). Its time is the difference between two runs of
    # the same launch stream whose expansions are cut to common lengths m and
    # s < m, so every launch carries all the states of its set (or half) and
    # the first- and second-order launches cancel out.
    reps = max(20, args.steps // 4)
    m = min(orders)
    sh = max(3, m // 2)

    def order_us(streams):
        """Time of one general order over the set, by difference of two runs
        cut to m and sh < m orders, with the set's launches on `streams`."""
        prop.set_streams(streams)
        prop.steps_set(hams, spx.REAL, states, coeffs, effs, 2)   # untimed: settles the launch width
        ctx.record(2)
        prop.steps_set(hams, spx.REAL, states, [c[: m + 1] for c in coeffs], effs, reps)
        ctx.record(3)
        prop.steps_set(hams, spx.REAL, states, [c[: sh + 1] for c in coeffs], effs, reps)
        ctx.record(4)
        ctx.synchronize()
        prop.set_streams(args.streams)
        return (ctx.elapsed_ms(2, 3) - ctx.elapsed_ms(3, 4)) / (reps * (m - sh)) * 1e3

    kernel_bytes = BYTES_PER_AMP_ORDER * n * S
    kernel_name = hams[0].kernel_name()
    # roofline: every launch of the order on one stream, the figure that a
    # kernel trace's per-dispatch durations add up to (the committed profile
    # reproduces it); with two streams the halves' dispatches overlap, and
    # that live figure is reported beside it
    per_order_us = order_us(1)
    achieved = kernel_bytes / (per_order_us * 1e-6) / 1e9
    log(f"[bench] general order over {S} states, one stream: {per_order_us:.2f} us -> {achieved:.1f} GB/s")
    two = None
    if W < S:
        two_us = order_us(2)
        two_gbs = kernel_bytes / (two_us * 1e-6) / 1e9
        two = {"us_per_launch": round(two_us, 3), "achieved": round(two_gbs, 2),
               "frac": round(two_gbs / HBM_PEAK_GBS, 4),
               "arrangement": f"two halves of {W} states, one launch sequence each, on two streams "
                              "(the timed steps' arrangement); events on the propagator's stream"}
        log(f"[bench] the same order as two halves on two streams: {two_us:.2f} us -> {two_gbs:.1f} GB/s")
    spx.autotune_mode(0)
    # the achievable streaming rate beside the datasheet peak: the in-place
    # scale kernel over 2^24 amplitudes (256 MiB each way, past every cache)
    copy = copy_probe(spx, 24)
    log(f"[bench] copy probe: {copy}")

    single = None
    if S > 1:
        single = single_state_leg(spx, ctx, L, systems[0], args)
        log(f"[bench] single state: {single}")

    for st in states:
        if not np.isfinite(st.download()).all():
            raise SystemExit("non-finite state after the timed run")

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"L={L} dipolar XXZ real-time Chebyshev steps, dt={args.dt}, "
                               f"eps={args.epsilon:g}, autonormalize off; {S} realisations (seed {SEED}, "
                               f"index 0..{S - 1}) stepped together, " + (
                                   f"one launch per order for each half of {W} on its own stream" if W < S
                                   else "one launch per order for the set"),
                   "nspins": L, "pairs": int(lo.size), "realisations": S, "orders": [int(o) for o in orders],
                   "spectral_bounds": [round(h.spectral_bound(), 4) for h in hams],
                   "schedule": tune_report.split(";")[0] if tune_report else "default",
                   "kernels": sorted(set(kernels))},
        "hpsi_per_s": round(args.steps * sum(orders) / (ms * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((kernel_name, S, S)),
                     "traffic_from": TRAFFIC_PROFILE.get((kernel_name, S)),
                     "kernel": kernel_name + f" (general Chebyshev order, {S} states per launch, one stream)",
                     "bytes_per_launch": kernel_bytes, "us_per_launch": round(per_order_us, 3),
                     "profile": TRACE_PROFILE.get((kernel_name, S)),
                     "copy_GBs": copy["GB/s"], "frac_of_copy": round(achieved / copy["GB/s"], 4),
                     "two_stream": two},
    }
    if single:
        line["single_state"] = single
    if not args.no_cv:
        log("[bench] control variate at L=16 ...")
        line["control_variate"] = control_variate_leg(spx, ctx)
        log("[bench] batched ensemble leg ...")
        line["batched_ensemble"] = batched_leg(spx, ctx)
        line["cpmg_L10"] = config0_leg(spx, ctx)
        cv = line["control_variate"]
        line["realisations_per_s"] = {
            "cpmg_L10": line["cpmg_L10"]["realisations_per_s"],
            "xy8_L8_12": line["batched_ensemble"]["realisations_per_s"],
            "xy16_L16": round(cv["plain"]["exact_runs"] / cv["plain"]["s"], 1)}
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(L, fields, lo, hi, jx, jy, jz, BYTES_PER_AMP_ORDER * n,
                                            args.cpu_seconds)
        line["cpu_baseline"]["realisations_per_s"] = cpu_realisation_rates(line["cpu_baseline"]["cores"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
