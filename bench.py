#!/usr/bin/env python3
"""Benchmark of the GLAC hot path on one MI355X.

Workload: a 32^4 periodic lattice at beta = 6.1 (the lattice of the headline
metric).  A *step* is one RK3 Wilson-flow step (eps = 0.01) with FP32 links
(two stored rows per link by default, ``--precision fp32`` for three), FP64
exponential coefficients and FP64 reductions, followed by the clover
measurement (plaquette, E, Q, Q(t)) -- one iteration of GLAC's
System::flowConfiguration loop (GLAC/system.cpp:494-508).  ``value`` is flow
steps per second with the field resident in HBM.

Also reported, each from its own timed region:
 - Metropolis sweeps/s (FP32 links, 10 hits, eps = 0.24, reunitarised every
   sweep) and FP64 flow steps/s;
 - ``roofline``: the flow-stage kernel's algorithmic bytes per launch over its
   mean launch time (HIP events on the library's stream), against 8 TB/s, with
   the HBM traffic per launch measured by rocprofv3 and the fraction of the
   launch the SIMDs spend issuing its VALU instructions (rocprofv3 SQ
   counters per wavefront, profiles/counters_<lattice>.json).  With the
   measurement fused into stage 0, that is RK3 stages 1 and 2;
   ``stage0_measure`` reports the fused stage 0 the same way;
 - ``indep_configs_per_s``: effectively independent configurations per second,
   configs/s / (2 tau_int(Q^2)), for the learned proposal with FP32 links,
   GLAC's RST proposal with the same FP32 links (the network's gain at equal
   precision) and GLAC's plain RST proposal with FP64 links
   (``--indep-seconds`` each);
 - ``cpu_baseline``: the FP64 oracle (its site loops on the host's OpenMP
   threads) doing flow steps, each with one measurement, on the same 32^4
   lattice.

Run ``python bench.py [--steps K] [--warmup W]``; one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "indep. configs/s (tau_int Q^2) + flow steps/s at 32^4, FP32/learned vs FP64 plain"


def link_bytes(fp32, rows=3):
    """Bytes of one stored link: 6 reals per row (3 complex)."""
    return 6 * rows * (4 if fp32 else 8)


def alg_bytes(fp32):
    return 32 if fp32 else 64       # 8 reals per su(3) accumulator entry


def flow_stage_bytes_per_launch(V, fp32, rows=3, stages=(0, 1, 2)):
    """Each RK3 stage launch reads 4V links and writes 4V links, and moves the
    su(3) accumulator: 4V entries written in stage 0, 4V read and written in
    stage 1, 4V read in stage 2 (24 s_L + 16 s_Z per site per step).  Mean
    over `stages`."""
    acc = {0: 4, 1: 8, 2: 4}
    per_site = [8 * link_bytes(fp32, rows) + acc[k] * alg_bytes(fp32) for k in stages]
    return sum(per_site) * V / len(per_site)


def metropolis_bytes_per_launch(V, fp32, rows=3):
    """A sub-sweep (one mu, one parity) reads all 4V links and writes V/2."""
    return 4.5 * V * link_bytes(fp32, rows)


SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4  # MI355X: 256 CUs x 4 SIMDs; a wave64 VALU op holds a SIMD 4 cycles


def profiled(n, kernels):
    """Mean over `kernels` of the committed rocprofv3 figures per launch
    (profiles/counters_<lattice>.json: SQ counters per wavefront, HBM bytes
    from FETCH_SIZE / WRITE_SIZE), or None when the lattice or a kernel was
    not profiled."""
    path = os.path.join(ROOT, "profiles", "counters_%s.json" % "x".join(map(str, n)))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)["kernels"]
    rows = [table.get(k) for k in kernels]
    if not all(rows):
        return None
    return {key: sum(r[key] for r in rows) / len(rows) for key in ("waves", "valu_per_wave", "traffic_bytes")}


def valu_issue_frac(prof, launch_ms):
    """Fraction of the launch time the SIMDs spend issuing the kernel's VALU
    instructions: waves x VALU instructions per wave x 4 cycles over
    1024 SIMDs x 2.4 GHz x the launch time measured in this run."""
    if prof is None:
        return None
    return prof["waves"] * prof["valu_per_wave"] * VALU_CYCLES / (SIMDS * CLOCK_HZ * launch_ms * 1e-3)


SHARED_THERM = 6000  # RST sweeps from the hot start before the variants of independent_configs split
VARIANTS = {"fp32c_learned": ("FP32_COMPRESSED", "learned"), "fp32c_rst": ("FP32_COMPRESSED", "rst"),
            "fp64_plain": ("FP64", "rst")}


def ratio_with_error(a, b):
    """indep. configs/s of a over b, with the error propagated from the two
    tau_int(Q^2) estimates (the configs/s rates are measured, not sampled)."""
    ratio = a["indep_configs_per_s"] / b["indep_configs_per_s"]
    rel = np.hypot(a["tau_int_q2_err"] / a["tau_int_q2"], b["tau_int_q2_err"] / b["tau_int_q2"])
    return ratio, ratio * rel


def independent_configs(G, n, beta, seconds, variants=tuple(VARIANTS)):
    """Effectively independent configurations per second,
    configs/s / (2 tau_int(Q^2)), for the learned proposal with compressed
    FP32 links, GLAC's fixed-eps RST proposal with the same FP32 links (the
    network's gain at equal precision) and GLAC's RST proposal with FP64
    links (the plain FP64 path).

    Shared start: Haar hot start and SHARED_THERM RST sweeps in compressed
    FP32 (about ten tau_int(Q^2) of the RST chain, so every variant starts
    with its topology settled).  Each variant continues from that field with
    its own streams: the learned variant collects 20 sweeps of (staple, link)
    samples and trains its network (3000 Adam steps, not counted in
    configs/s); then 300 sweeps of its own update and configurations spaced by
    NCor = 50 sweeps, each flowed 50 RK3 steps of 0.02 (t = 1) and measured,
    for about `seconds` of wall time.  tau_int(Q^2) by Wolff's Gamma method,
    in units of the spacing."""
    from glac_amd.system import EnsembleConfig, run_ensemble
    from glac_amd.proposal import train_from_chain
    out = {}
    with G.Lattice(n, G.FP32_COMPRESSED) as L:
        L.hot_start(31)
        L.metropolis(SHARED_THERM, beta, 0.24, 10, seed=3)
        start = L.download()
    for name in variants:
        prec, proposal = getattr(G, VARIANTS[name][0]), VARIANTS[name][1]
        seed = 9 + list(VARIANTS).index(name)
        L = G.Lattice(n, prec)
        L.upload(start)
        off = SHARED_THERM
        train_s = 0.0
        if proposal == "learned":
            t0 = time.perf_counter()
            train_from_chain(L, beta, seed=seed, sweep_offset=off, collect_sweeps=20)
            L.synchronize()
            train_s = time.perf_counter() - t0
            off += 20
        cfg = EnsembleConfig(dims=n, beta=beta, NTherm=300, NCf=1, NCor=50, NFlows=50, flowEpsilon=0.02, seed=seed,
                             proposal=proposal)
        pilot = run_ensemble(L, cfg, thermalise=True, sweep_offset=off)
        off += cfg.NTherm + cfg.NCor
        cfg.NCf = max(40, int(seconds / (pilot.update_seconds + pilot.flow_seconds)))
        r = run_ensemble(L, cfg, thermalise=False, sweep_offset=off,
                         progress=lambda i, q: log(f"[bench] {name} config {i}") if i % 100 == 0 else None)
        L.close()
        ti = r.tau_int_q2()
        rate = r.configs_per_second / (2.0 * ti["tau_int"])
        out[name] = {"indep_configs_per_s": rate, "configs_per_s": r.configs_per_second,
                     "tau_int_q2": ti["tau_int"], "tau_int_q2_err": ti["dtau"], "configs": len(r.q),
                     "q2_mean": float(np.mean(r.q ** 2)), "q2_err": ti["err"], "plaq_mean": float(r.plaq.mean()),
                     "acceptance": r.acceptance, "train_seconds": train_s}
        log(f"[bench] {name}: {r.configs_per_second:.2f} configs/s, tau_int(Q^2) {ti['tau_int']:.2f} "
            f"+- {ti['dtau']:.2f}, {rate:.3f} indep. configs/s, <Q^2> {np.mean(r.q ** 2):.2f}")
    if "fp32c_learned" in out and "fp64_plain" in out:
        r, e = ratio_with_error(out["fp32c_learned"], out["fp64_plain"])
        out["ratio_learned_fp32_over_plain_fp64"], out["ratio_err"] = r, e
    if "fp32c_learned" in out and "fp32c_rst" in out:
        r, e = ratio_with_error(out["fp32c_learned"], out["fp32c_rst"])
        out["same_precision"] = {"ratio_learned_over_rst_fp32c": r, "ratio_err": e,
                                 "tau_int_q2_ratio_rst_over_learned":
                                     out["fp32c_rst"]["tau_int_q2"] / out["fp32c_learned"]["tau_int_q2"]}
    out["protocol"] = (f"shared start: hot start + {SHARED_THERM} RST sweeps (FP32); per variant 300 sweeps, then "
                       "NCor 50 sweeps (10 hits, eps 0.24 fallback/RST), flow 50 x 0.02, Q at t = 1; "
                       "tau_int(Q^2) by Wolff's Gamma method (S = 1.5), errors from its estimate")
    # the same measurement run for several minutes per variant (committed), for comparison
    # with this run's shorter, noisier estimate
    ref = os.path.join(ROOT, "profiles", "indep_configs_%s.json" % "x".join(map(str, n)))
    if os.path.exists(ref):
        with open(ref) as f:
            r = json.load(f)
        out["long_run"] = {"ratio_learned_fp32_over_plain_fp64": r["ratio_learned_fp32_over_plain_fp64"],
                           "ratio_err": r["ratio_err"], "seconds_per_variant": r.get("seconds_per_variant"),
                           "same_precision": r.get("same_precision"),
                           "source": "profiles/" + os.path.basename(ref)}
        for k in VARIANTS:
            if k in r:
                out["long_run"][k + "_indep_configs_per_s"] = r[k]["indep_configs_per_s"]
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def learned_sweep_rate(G, n, beta, therm, sweeps, train_steps=3000):
    """Learned-proposal sweeps per second and the sub-sweep kernel's time,
    compressed FP32 links: a hot start thermalised by `therm` RST sweeps, the
    network trained on 20 more (`train_steps` Adam steps), two untimed
    learned sweeps, then `sweeps` timed ones."""
    from glac_amd.proposal import train_from_chain
    with G.Lattice(n, G.FP32_COMPRESSED) as L:
        L.hot_start(20261202)
        L.metropolis(therm, beta, 0.24, 10, seed=7)
        train_from_chain(L, beta, seed=7, sweep_offset=therm, collect_sweeps=20, steps=train_steps)
        L.metropolis_learned(2, beta, 0.24, 10, seed=8)
        L.synchronize()
        L.timing(True)
        L.timing_reset()
        t0 = time.perf_counter()
        acc = L.metropolis_learned(sweeps, beta, 0.24, 10, seed=8, sweep_offset=2)
        L.synchronize()
        el = time.perf_counter() - t0
        launches, ms = L.timing_read(G.K_METROPOLIS_LEARNED)
        L.timing(False)
    out = {"lattice": list(n), "sweeps": sweeps, "learned_sweeps_per_s": sweeps / el,
           "subsweep_kernel_ms": ms / launches, "acceptance": float(acc.mean())}
    log(f"[bench] learned: {out['learned_sweeps_per_s']:.1f} sweeps/s, sub-sweep {out['subsweep_kernel_ms']:.4f} ms, "
        f"acceptance {out['acceptance']:.4f}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--dims", default="32,32,32,32")
    ap.add_argument("--sweeps", type=int, default=100, help="timed Metropolis sweeps")
    ap.add_argument("--therm", type=int, default=1000,
                    help="thermalising Metropolis sweeps from the hot start before anything is timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4, help="flow steps of the CPU baseline sample")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--precision", choices=["fp32", "fp32c"], default="fp32c",
                    help="FP32 mode of the timed flow and sweeps: fp32 stores 3 rows per link, fp32c two")
    ap.add_argument("--indep-seconds", type=float, default=30.0,
                    help="wall seconds of ensemble generation per variant for the independent-configs/s "
                         "measurement (learned FP32, RST FP32, RST FP64); 0 skips it")
    ap.add_argument("--only-indep", action="store_true",
                    help="run only the independent-configs/s measurement and print its JSON")
    ap.add_argument("--indep-variants", default=",".join(VARIANTS),
                    help="comma-separated variants of the independent-configs/s measurement (%s)" % ", ".join(VARIANTS))
    ap.add_argument("--only-learned", action="store_true",
                    help="time only the learned-proposal sweep (compressed FP32 links): thermalise --therm RST "
                         "sweeps, train on 20 more, then --sweeps learned sweeps; print its JSON")
    ap.add_argument("--train-steps", type=int, default=3000, help="Adam steps of --only-learned's training")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the library runs a lattice on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the lattice by 4D domain decomposition is not implemented).",
              file=sys.stderr)
        sys.exit(2)

    import glac_amd as G

    n = tuple(int(v) for v in args.dims.split(","))
    V = int(np.prod(n))
    beta, eps_flow = 6.1, 0.01

    if args.only_learned:
        print(json.dumps(learned_sweep_rate(G, n, beta, args.therm, args.sweeps, args.train_steps)), flush=True)
        return

    if args.only_indep:
        variants = tuple(v for v in args.indep_variants.split(",") if v)
        unknown = [v for v in variants if v not in VARIANTS]
        if unknown:
            ap.error(f"unknown variants {unknown}")
        print(json.dumps(independent_configs(G, n, beta, args.indep_seconds, variants)), flush=True)
        return

    # ---------------------------------------------------------------- FP32 flow
    rows = 2 if args.precision == "fp32c" else 3
    L = G.Lattice(n, G.FP32_COMPRESSED if rows == 2 else G.FP32_LINKS)
    L.hot_start(20261202)
    log(f"[bench] lattice {n}, FP32 links ({rows} stored rows), {L.device_bytes / 2**20:.0f} MiB on device")
    # Metropolis: thermalising sweeps, then the timed sweeps
    L.metropolis(args.therm, beta, 0.24, 10, reunit_every=1, seed=7, sweep_offset=0)
    L.synchronize()
    L.timing(True)
    L.timing_reset()
    t0 = time.perf_counter()
    acc = L.metropolis(args.sweeps, beta, 0.24, 10, reunit_every=1, seed=7, sweep_offset=args.therm)
    L.synchronize()
    t_sweeps = time.perf_counter() - t0
    m_launches, m_ms = L.timing_read(G.K_METROPOLIS)
    L.timing(False)
    sweeps_per_s = args.sweeps / t_sweeps
    log(f"[bench] metropolis {args.sweeps} sweeps: {sweeps_per_s:.1f} sweeps/s, "
        f"acceptance {acc.mean():.4f}")

    L.copy(G.FLOW, G.MAIN)
    L.flow(args.warmup, eps_flow, field=G.FLOW, measure=True)
    L.synchronize()
    t0 = time.perf_counter()
    plaq, e_raw, q, qt = L.flow(args.steps, eps_flow, field=G.FLOW, measure=True)
    L.synchronize()
    elapsed = time.perf_counter() - t0
    # Per-kernel times: the same flow again from the same field, with a HIP
    # event pair around every launch (kept out of the run above, whose wall
    # time is the step rate).
    L.copy(G.FLOW, G.MAIN)
    L.flow(args.warmup, eps_flow, field=G.FLOW, measure=True)
    L.synchronize()
    L.timing(True)
    L.timing_reset()
    t1 = time.perf_counter()
    L.flow(args.steps, eps_flow, field=G.FLOW, measure=True)
    L.synchronize()
    elapsed_events = time.perf_counter() - t1
    f_launches, f_ms = L.timing_read(G.K_FLOW_STAGE)
    fm_launches, fm_ms = L.timing_read(G.K_FLOW_STAGE_MEASURE)
    meas_launches, meas_ms = L.timing_read(G.K_MEASURE)
    L.timing(False)
    U_therm = None if args.no_fp64 else L.download(G.MAIN)
    L.close()
    steps_per_s = args.steps / elapsed
    log(f"[bench] flow {args.precision}: {args.steps} steps in {elapsed:.3f} s = {steps_per_s:.1f} steps/s "
        f"({args.steps / elapsed_events:.1f} with per-kernel events); "
        f"stage kernel {f_ms / max(f_launches, 1):.4f} ms x {f_launches}, "
        f"stage 0 + clover {fm_ms / max(fm_launches, 1):.4f} ms x {fm_launches}, "
        f"clover kernel {meas_ms / max(meas_launches, 1):.4f} ms x {meas_launches}")

    # With the measurement fused into stage 0, the plain stage kernel is timed
    # on stages 1 and 2 only.
    plain_stages = (1, 2) if fm_launches else (0, 1, 2)
    stage_bytes = flow_stage_bytes_per_launch(V, True, rows, plain_stages)
    stage_ms = f_ms / max(f_launches, 1)
    achieved = stage_bytes / (stage_ms * 1e-3) / 1e9
    stage0_bytes = flow_stage_bytes_per_launch(V, True, rows, (0,))
    metro_bytes = metropolis_bytes_per_launch(V, True, rows)
    metro_ms = m_ms / max(m_launches, 1)
    metro_gbs = metro_bytes / (metro_ms * 1e-3) / 1e9
    prof_stage = profiled(n, [f"flow_stage_kernel<float, {rows}, {k}, false>" for k in plain_stages])
    prof_stage0 = profiled(n, [f"flow_stage_kernel<float, {rows}, 0, true>"])
    prof_metro = profiled(n, [f"metropolis_kernel<float, {rows}>"])

    # ---------------------------------------------------------------- FP64 flow
    fp64 = None
    if not args.no_fp64:
        L = G.Lattice(n, G.FP64)
        L.upload(U_therm, field=G.FLOW)       # the same thermalised configuration
        del U_therm
        k64 = max(args.steps // 3, 10)
        L.flow(min(args.warmup, 10), eps_flow, field=G.FLOW)
        L.synchronize()
        L.timing(True)
        L.timing_reset()
        t0 = time.perf_counter()
        L.flow(k64, eps_flow, field=G.FLOW)
        L.synchronize()
        el64 = time.perf_counter() - t0
        n64, ms64 = L.timing_read(G.K_FLOW_STAGE)
        L.close()
        ms_stage64 = ms64 / max(n64, 1)
        fp64 = {
            "flow_steps_per_s": k64 / el64,
            "stage_kernel_ms": ms_stage64,
            "stage_kernel_gbs": flow_stage_bytes_per_launch(V, False) / (ms_stage64 * 1e-3) / 1e9,
        }
        log(f"[bench] flow fp64: {fp64['flow_steps_per_s']:.1f} steps/s, stage kernel {ms_stage64:.4f} ms")

    # ---------------------------------------------------------------- indep. configs/s
    indep = None
    if args.indep_seconds > 0:
        indep = independent_configs(G, n, beta, args.indep_seconds)

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        U = O.hot_start_haar(n, 1)
        cpu_steps = args.cpu_steps
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            O.flow_step(U, n, eps_flow)
            O.measure(U, n)
        cpu_s = time.perf_counter() - t0
        threads = O.num_threads()
        cpu = {"value": cpu_steps / cpu_s, "unit": "flow steps/s", "cores": threads, "kind": "port",
               "sample": f"{cpu_steps} RK3 steps, each followed by a clover measurement, of the FP64 oracle "
                         f"on the full {'x'.join(map(str, n))} lattice, site loops on {threads} host threads "
                         f"(OpenMP), {cpu_s:.1f} s"}
        log(f"[bench] cpu baseline: {cpu_s / cpu_steps:.2f} s per step on {threads} threads")

    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "flow steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 links (2 stored rows, third rebuilt)" if rows == 2 else "fp32 links")
                 + ", fp64 exp coefficients and sums",
        "data": f"synthetic: Haar hot start + {args.therm} thermalising Metropolis sweeps at beta 6.1, "
                f"seed 20261202; flow timed from flow time {args.warmup * eps_flow:g}",
        "config": {"workload": "RK3 Wilson flow step + clover measurement", "lattice": list(n),
                   "beta": beta, "flow_eps": eps_flow, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": prof_stage["traffic_bytes"] if prof_stage else None,
                     "kernel": f"flow_stage_kernel<float, {rows}, {{{','.join(map(str, plain_stages))}}}, false>",
                     "alg_bytes_per_launch": stage_bytes, "mean_launch_ms": stage_ms,
                     "valu_issue_frac": valu_issue_frac(prof_stage, stage_ms)},
        "metropolis": {"sweeps_per_s": sweeps_per_s, "acceptance": float(acc.mean()),
                       "subsweep_kernel_ms": metro_ms, "subsweep_kernel_gbs": metro_gbs,
                       "hbm_frac": metro_gbs / HBM_PEAK_GBS, "valu_issue_frac": valu_issue_frac(prof_metro, metro_ms)},
        "measure_kernel_ms": meas_ms / max(meas_launches, 1),
        "flow_steps_per_s_with_kernel_events": args.steps / elapsed_events,
        "stage0_measure": ({"kernel": f"flow_stage_kernel<float, {rows}, 0, true>", "mean_launch_ms": fm_ms / fm_launches,
                            "alg_bytes_per_launch": stage0_bytes,
                            "achieved_gbs": stage0_bytes / (fm_ms / fm_launches * 1e-3) / 1e9,
                            "traffic": prof_stage0["traffic_bytes"] if prof_stage0 else None,
                            "valu_issue_frac": valu_issue_frac(prof_stage0, fm_ms / fm_launches)}
                           if fm_launches else None),
        "fp64_plain": fp64,
        "indep_configs_per_s": indep,
        "cpu_baseline": cpu,
        "final_obs": {"plaq": float(plaq[-1]), "E": float(-e_raw[-1] / 64), "Q": float(q[-1])},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
