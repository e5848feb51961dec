"""Throughput of the Monte Carlo link-drop path on one MI355X.

A step is one fused launch (``hpx_run_batch``) of 4096 independent drops of the
4x4 correlated 23-tap TDL, 1024-subcarrier 16-QAM OFDM link with MMSE
detection on ideal CSI; drop i of the campaign runs at SNR point i % 11 of
0..30 dB.  Inputs (link tables, per-drop noise powers) are resident in HBM
before timing; every drop draws its channel, bits and noise on the device.

Prints one JSON line (rank 0, N = 1).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "MC drops/s (4x4 1024-SC OFDM, MMSE) at BER parity; learned vs LMMSE est."
FP32_VECTOR_PEAK_TFLOPS = 157.3  # published MI355X FP32 peak (packed FP32 math)
FP32_UNPACKED_PEAK_TFLOPS = 78.6  # one FP32 FMA per lane and cycle: the kernels are built without packed FP32
HBM_PEAK_GBPS = 8000.0  # published MI355X HBM3E bandwidth
# HBM bytes of one fused 4x4 drop, from rocprofv3 FETCH_SIZE / WRITE_SIZE over 45 056-drop launches
# (FETCH x2.0 and WRITE x1.0, the factors a 256 MiB device copy reads in the same process):
# (1990 x 2 + 1408) KiB / 45056; the algorithmic part is ~12 B (noise power, error count) plus the
# link tables every drop reads from L2
HBM_BYTES_PER_DROP = (1990.0 * 2.0 + 1408.0) * 1024.0 / 45056.0
DROPS_PER_STEP = 11 * 4096  # one step: the campaign's 4096 drops at each of its 11 SNR points, one launch
SEED = 20261202


def algorithmic_flops_per_drop(cfg) -> float:
    """Per-drop work of the reference algorithm (LinkConfig.algorithmic_flops)."""

    return cfg.algorithmic_flops


def _oracle_drops(args):
    """Worker: oracle drops for budget_s seconds on one core (reference-faithful unless faithful is
    False: closed-form CSI and vectorised taps); returns the count, or with keep > 0 the count and
    (errors, bits, equalized) of the first `keep` drops, which it runs in any case."""

    cfg_name, first, stride, budget_s = args[:4]
    faithful = args[4] if len(args) > 4 else True
    keep = args[5] if len(args) > 5 else 0
    from threadpoolctl import threadpool_limits

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from hermespy_amd.configs import CONFIGS
    from oracle import drop as odrop  # the oracle, timed as the CPU baseline

    cfg = CONFIGS[cfg_name]()
    link = odrop.Link(cfg.num_rx, cfg.num_tx, cfg.num_subcarriers, cfg.words, cfg.prefix_ratio, cfg.modulation_order,
                      cfg.bandwidth, cfg.delays, cfg.power_profile, cfg.rice_factors, cfg.doppler, cfg.num_sinusoids,
                      cfg.corr_rx, cfg.corr_tx, cfg.detector, cfg.dc_suppression, cfg.gain, cfg.device_power)
    snr = np.asarray(cfg.snr_db)
    done = 0
    kept = []
    with threadpool_limits(1):
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s or done == 0 or done < keep:
            d = first + done * stride
            r = odrop.run_drop(link, SEED, d, float(snr[d % snr.size]), faithful=faithful)
            if done < keep:
                kept.append((r.errors, r.bits, r.equalized))
            done += 1
    return (done, kept) if keep else done


def cpu_baseline(cfg, budget_s: float = 15.0, all_core_budget_s: float = 10.0, workers: int = 16,
                 vectorised_budget_s: float = 5.0, gpu_drops=None, check_drops: int = 11) -> dict:
    """Reference-faithful float64 oracle (dense impulse response, per-path loop, pinv/solve): one host core,
    then one process per core for the all-core aggregate (as the reference runs one actor per CPU), and
    the oracle's vectorised mode (closed-form CSI, no dense impulse response) on one core for fairness.

    gpu_drops(k) -> (errors, equalized) of the fused path's drops 0 .. k-1 (the benchmark's drops):
    the first `check_drops` oracle drops of the single-core leg (one per SNR point) check them, giving
    the run's error against the float64 oracle (``fp64_check``)."""

    from concurrent.futures import ProcessPoolExecutor

    name = "mimo4x4-1024"
    keep = check_drops if gpu_drops is not None else 0
    t0 = time.perf_counter()
    res = _oracle_drops((name, 0, 1, budget_s, True, keep))
    dt = time.perf_counter() - t0
    done, kept = res if keep else (res, [])
    out = {"value": done / dt, "unit": "drops/s", "cores": 1, "kind": "port",
           "sample": f"{done} drops of {cfg.name}, reference-faithful float64 oracle (dense CIR, per-path loop), "
                     f"{dt:.1f} s on 1 host core of {os.cpu_count()}"}
    if keep:
        g_err, g_eq = gpu_drops(keep)
        rel = [float(np.linalg.norm(g_eq[i] - eq) / np.linalg.norm(eq)) for i, (_, _, eq) in enumerate(kept)]
        d_err = [int(g_err[i]) - int(e) for i, (e, _, _) in enumerate(kept)]
        bits = sum(b for _, b, _ in kept)
        out["fp64_check"] = {"drops": keep, "snr_db": [float(x) for x in np.asarray(cfg.snr_db)[np.arange(keep) % len(cfg.snr_db)]],
                             "max_rel_err_equalized": max(rel), "median_rel_err_equalized": float(np.median(rel)),
                             "delta_errors": d_err, "bits": int(bits),
                             "ber_gpu": float(sum(int(x) for x in g_err[:keep])) / bits,
                             "ber_fp64_oracle": float(sum(e for e, _, _ in kept)) / bits,
                             "oracle": "reference-faithful float64 (dense impulse response, pinv / stacked solve)"}
    workers = max(1, min(workers, os.cpu_count() or 1))
    if all_core_budget_s > 0 and workers > 1:
        t0 = time.perf_counter()
        with ProcessPoolExecutor(workers) as pool:
            counts = list(pool.map(_oracle_drops, [(name, w, workers, all_core_budget_s) for w in range(workers)]))
        dt = time.perf_counter() - t0
        out["all_cores"] = {"value": sum(counts) / dt, "cores": workers, "drops": sum(counts), "seconds": dt}
    if vectorised_budget_s > 0:
        t0 = time.perf_counter()
        done = _oracle_drops((name, 0, 1, vectorised_budget_s, False))
        dt = time.perf_counter() - t0
        out["vectorised"] = {"value": done / dt, "cores": 1, "drops": done, "seconds": dt,
                             "mode": "closed-form CSI, vectorised taps, float64"}
    return out


def plain_gpu_path(cfg, budget_s: float = 5.0) -> dict:
    """The plain single-drop GPU path: stage entry points, one drop at a time."""

    import torch

    from hermespy_amd.simulation import SingleDropLink

    plain = SingleDropLink(cfg, SEED)
    snr = np.asarray(cfg.snr_db)
    acc = torch.zeros(1, dtype=torch.int64, device=plain.device)
    plain.run_drop(0, float(snr[0]), acc)  # warm-up
    torch.cuda.synchronize()
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        plain.run_drop(done, float(snr[done % snr.size]), acc)
        done += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "drops/s", "drops": done}


def estimator_leg(drops: int = 4096, reps: int = 10) -> dict:
    """Pilot-based estimation on the 2x2 512-subcarrier comb-pilot link: NMSE and BER of the LMMSE
    estimator and of the learned estimator against the ideal CSI, and the throughput of each
    (pilot LS + interpolation, drops/s, HIP events)."""

    import torch

    from hermespy_amd import _lib
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.learned import WEIGHTS_PATH, ConvNetEstimator
    from hermespy_amd.lmmse import LMMSEChannelEstimation
    from hermespy_amd.simulation import BatchedLink

    cfg = mimo2x2_512_pilots()
    dev = torch.device("cuda")
    snr = np.asarray(cfg.snr_db)
    sidx = np.arange(drops) % snr.size
    npw = torch.as_tensor(cfg.noise_power(snr[sidx]), dtype=torch.float64, device=dev)
    link = BatchedLink(cfg)
    e = cfg.words * cfg.num_subcarriers
    err_ideal = torch.empty(drops, dtype=torch.int32, device=dev)
    grid = torch.empty((drops, cfg.num_rx, cfg.words, cfg.num_subcarriers), dtype=torch.complex64, device=dev)
    csi = torch.empty((drops, cfg.words, cfg.num_subcarriers), dtype=torch.complex64, device=dev)
    spatial = torch.empty((drops, cfg.num_rx, cfg.num_tx), dtype=torch.complex64, device=dev)
    link.run(SEED, 0, npw, err_ideal, doppler=cfg.draw_doppler(0, drops), grid=grid, csi=csi, spatial=spatial)
    lm = LMMSEChannelEstimation(cfg)
    rows = lm.rows_by_set(sidx)
    out = torch.empty((cfg.num_tx, drops, cfg.num_rx, e), dtype=torch.complex64, device=dev)
    lib, st = _lib.load(), _lib.stream_handle()

    def timed(fn) -> float:
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return drops * reps / (a.elapsed_time(b) / 1e3)

    def quality(est) -> dict:
        err = torch.zeros(drops, dtype=torch.float64, device=dev)
        pw = torch.zeros(drops, dtype=torch.float64, device=dev)
        _lib.check(lib.hpx_estimate_error(_lib.ptr(est), _lib.ptr(csi), _lib.ptr(spatial), drops, cfg.num_rx, cfg.num_tx,
                                          e, _lib.ptr(err), _lib.ptr(pw), st))
        bit_err = torch.empty(drops, dtype=torch.int32, device=dev)
        _lib.check(lib.hpx_detect_estimated(_lib.ptr(grid), _lib.ptr(est), drops, cfg.num_rx, cfg.num_tx, cfg.words,
                                            cfg.num_subcarriers, _lib.ptr(link.elements), _lib.ptr(link.word_data),
                                            cfg.num_data, cfg.modulation_order, _lib.HPX_DETECT_MMSE, _lib.ptr(npw), SEED,
                                            0, _lib.ptr(bit_err), None, st))
        en, pn, bn = err.cpu().numpy(), pw.cpu().numpy(), bit_err.cpu().numpy()
        nmse = [float(10 * np.log10(en[sidx == k].sum() / pn[sidx == k].sum())) for k in range(snr.size)]
        ber = [float(bn[sidx == k].sum() / ((sidx == k).sum() * cfg.bits_per_drop)) for k in range(snr.size)]
        return {"nmse_db": nmse, "ber": ber}

    res = {"workload": cfg.name, "drops": drops, "snr_db": snr.tolist()}
    ideal = err_ideal.cpu().numpy()
    res["ideal_csi_ber"] = [float(ideal[sidx == k].sum() / ((sidx == k).sum() * cfg.bits_per_drop))
                            for k in range(snr.size)]
    ls = lm.pilot_ls(grid, drops)
    res["lmmse"] = {"drops_per_s": timed(lambda: lm.apply(lm.pilot_ls(grid, drops), sidx, out, rows)),
                    "dtype": "f32", **quality(lm.apply(ls, sidx, out, rows))}
    if WEIGHTS_PATH.exists():
        net = ConvNetEstimator.load(cfg)
        res["learned"] = {"drops_per_s": timed(lambda: net.estimate_grid(grid, lm.pilot_index, lm.pilot_values, npw, out)),
                          "dtype": "bf16 weights and activations, f32 accumulate", "params": net.num_params,
                          **quality(net.estimate_grid(grid, lm.pilot_index, lm.pilot_values, npw, out))}
        res["throughput_ratio"] = res["learned"]["drops_per_s"] / res["lmmse"]["drops_per_s"]
        gap = np.asarray(res["learned"]["nmse_db"]) - np.asarray(res["lmmse"]["nmse_db"])
        res["nmse_gap_db"] = gap.tolist()
        res["max_nmse_gap_db"] = float(gap.max())
    else:
        res["learned"] = None
    return res


def uneven_leg(drops: int = 8192, reps: int = 5) -> dict:
    """The mixed population (antennas 1/2/4/8, N 128..2048, 5..23 paths, Doppler 10..900 Hz, QPSK to
    64-QAM, 11 SNR points) as one graph-captured submission after per-class autotuning."""

    import torch

    from hermespy_amd.uneven import UnevenBatch, draw_population, tune

    pop = draw_population(drops)
    ub = UnevenBatch(pop)
    tuning = tune(ub, reps=2)
    ub.capture()
    ub.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        ub.replay()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    bits = float(ub.bits.sum())
    flops = float(sum(c.algorithmic_flops for c in ub.links))
    achieved = flops / (ms / 1e3) / 1e12
    return {"workload": "uneven mix (8x8 at 2048 subcarriers in lean classes)", "drops": drops,
            "lean_drops": ub.num_lean, "ms_per_launch": ms,
            "drops_per_s": drops / (ms / 1e3), "bits_per_s": bits / (ms / 1e3), "classes": ub.num_classes,
            "launches": int(sum(len(lane) for lane in ub.lanes)), "streams": len(ub.lanes),
            "stream_cost_ms": [1e3 * x for x in ub.stream_cost],
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_VECTOR_PEAK_TFLOPS, "flops_per_launch": flops,
                         "note": "the population's algorithmic flops over the graph's span"},
            "class_choice": tuning["choice"], "class_choice_fields": ["workgroup_size", "fft_values"]}


def campaign_leg(num_samples: int = 2048) -> dict:
    """The 4x4 link's BER campaign over the 11 SNR sections through BatchedSimulation.simulate:
    section queue, indexed launches of up to 16 384 drops, device section statistics and the
    per-launch read-back of the stopping test, without stopping (default evaluator) and with
    the reference's rule at tolerance 1e-3, confidence 0.05, 1024 samples between tests."""

    import time

    import torch

    from hermespy_amd.configs import mimo4x4_1024
    from hermespy_amd.modem import BitErrorEvaluator
    from hermespy_amd.simulation import BatchedSimulation

    sim = BatchedSimulation(mimo4x4_1024(), seed=SEED)
    sim.simulate(64)  # warm-up
    out = {"workload": "mimo4x4-1024sc-16qam-mmse, 11 SNR sections", "num_samples": num_samples}
    for name, ev in (("full", None), ("early_stopping", BitErrorEvaluator(confidence=0.05, tolerance=1e-3,
                                                                         min_num_samples=1024))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sim.simulate(num_samples, ev)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        drops = int(res.evaluation_results[0].num_drops().sum())
        out[name] = {"drops": drops, "seconds": wall, "drops_per_s": drops / wall,
                     "drops_per_section": res.evaluation_results[0].num_drops().ravel().tolist(),
                     "ber": [float(x) for x in res.evaluation_results[0].to_array().ravel()]}
    return out


def massive_leg(drops: int = 2048, chunk: int = 1024) -> dict:
    """64x8 massive-MIMO uplink drops (4096-point OFDM, 3276 active subcarriers, 256-QAM, LMMSE)."""

    import torch

    from hermespy_amd.massive import MassiveBatch, MassiveConfig

    cfg = MassiveConfig()
    mb = MassiveBatch(cfg, chunk=chunk, keep_grid=False)
    snr = np.asarray(cfg.snr_db)
    err = torch.zeros(drops, dtype=torch.int32, device="cuda")
    npw = torch.as_tensor(cfg.noise_power(snr[np.arange(drops) % snr.size]), dtype=torch.float64, device="cuda")
    mb.run(SEED, 0, npw[:chunk], err[:chunk])
    torch.cuda.synchronize()
    err.zero_()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for s0 in range(0, drops, chunk):
        mb.run(SEED, s0, npw[s0:s0 + chunk], err[s0:s0 + chunk])
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    e = err.cpu().numpy()
    idx = np.arange(drops) % snr.size
    ber = [float(e[idx == k].sum() / ((idx == k).sum() * cfg.bits_per_drop)) for k in range(snr.size)]
    rate = drops / (ms / 1e3)
    mfma = cfg.mfma_flops_per_drop * rate / 1e12
    hbm = cfg.hbm_bytes_per_drop * rate / 1e9
    return {"workload": cfg.name, "drops": drops, "drops_per_launch": chunk, "ms": ms, "drops_per_s": rate,
            "bits_per_s": drops * cfg.bits_per_drop / (ms / 1e3), "snr_db": snr.tolist(), "ber": ber,
            "roofline": {"bound": "mfma", "achieved": mfma, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": mfma / FP32_VECTOR_PEAK_TFLOPS, "mfma_flops_per_drop": cfg.mfma_flops_per_drop,
                         "note": "float32 MFMA of the receiver over the whole pipeline's time (a lower bound on "
                                 "the receive kernel's own rate)",
                         "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBPS,
                                 "bytes_per_drop": cfg.hbm_bytes_per_drop},
                         "reference_equivalent": {"flops_per_drop": cfg.reference_flops_per_drop,
                                                  "tflops": cfg.reference_flops_per_drop * rate / 1e12}}}


def small_link_leg(drops_per_point: int = 1000, reps: int = 3) -> dict:
    """The small single-antenna link (64 subcarriers, QPSK, ZF, SNR 0..20 dB in 5 dB steps, 1000
    drops each, one launch): the float64 path and the float32 fused path, and the float32
    path's error against the float64 one on the same drops."""

    import torch

    from hermespy_amd.configs import siso_64
    from hermespy_amd.simulation import BatchedLink

    cfg = siso_64()
    snr = np.array([0.0, 5.0, 10.0, 15.0, 20.0])
    n = drops_per_point * snr.size
    idx = np.arange(n) // drops_per_point
    bl = BatchedLink(cfg)
    npw = torch.as_tensor(cfg.noise_power(snr[idx]), dtype=torch.float64, device="cuda")
    per_drop = cfg.words * cfg.num_subcarriers
    e64 = torch.zeros(n, dtype=torch.int32, device="cuda")
    e32 = torch.zeros(n, dtype=torch.int32, device="cuda")
    q64 = torch.empty((n, per_drop), dtype=torch.complex128, device="cuda")
    q32 = torch.empty((n, 1, per_drop), dtype=torch.complex64, device="cuda")
    bl.run_f64(SEED, 0, npw, e64, q64)
    bl.run(SEED, 0, npw, e32, q32)
    torch.cuda.synchronize()
    err = np.linalg.norm((q32[:, 0].to(torch.complex128) - q64).cpu().numpy(), axis=1)
    rel = err / np.linalg.norm(q64.cpu().numpy(), axis=1)
    res = {"workload": f"{cfg.name}, 5 SNR points x {drops_per_point} drops", "drops": n,
           "f32_rel_error_vs_f64": {"max": float(rel.max()), "median": float(np.median(rel))},
           "error_counts_equal": int((e64 == e32).sum().item())}
    for name, launch in (("f64", lambda: bl.run_f64(SEED, 0, npw, e64)), ("f32", lambda: bl.run(SEED, 0, npw, e32))):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            launch()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        e = (e64 if name == "f64" else e32).cpu().numpy()
        res[name] = {"ms": ms, "drops_per_s": n / (ms / 1e3),
                     "ber": [float(e[idx == k].sum() / (drops_per_point * cfg.bits_per_drop)) for k in range(snr.size)]}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--drops", type=int, default=DROPS_PER_STEP)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain single-drop GPU path")
    ap.add_argument("--no-tune", action="store_true", help="use the default workgroup size instead of the autotuner's")
    ap.add_argument("--no-estimators", action="store_true", help="skip the pilot-estimation leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the heterogeneous-batch leg")
    ap.add_argument("--no-massive", action="store_true", help="skip the massive-MIMO leg")
    ap.add_argument("--no-small", action="store_true", help="skip the float64 / float32 small-link leg")
    ap.add_argument("--no-campaign", action="store_true", help="skip the early-stopping campaign leg")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (drops shard across GPUs by "
              f"Philox counter range; this benchmark runs on one GPU)", file=sys.stderr)
        return 2

    import torch

    from hermespy_amd.configs import mimo4x4_1024
    from hermespy_amd.simulation import BatchedLink

    from hermespy_amd.autotune import tuned

    cfg = tuned(mimo4x4_1024(), verbose=True) if not args.no_tune else mimo4x4_1024()
    snr = np.asarray(cfg.snr_db)
    link = BatchedLink(cfg)
    dev = link.device
    n = args.drops
    total = args.warmup + args.steps
    first = [s * n for s in range(total)]
    npw = [torch.as_tensor(cfg.noise_power(snr[(f + np.arange(n)) % snr.size]), dtype=torch.float64, device=dev)
           for f in first]
    errors = torch.zeros((total, n), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    if args.copy_mib:
        src = torch.ones(args.copy_mib * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for s in range(args.warmup):
        link.run(SEED, first[s], npw[s], errors[s])
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for s in range(args.warmup, total):
        link.run(SEED, first[s], npw[s], errors[s])
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)  # HIP events on the launch stream
    ms_per_step = ms / args.steps
    drops_per_s = n * args.steps / (ms / 1e3)
    flops = algorithmic_flops_per_drop(cfg) * n
    achieved = flops / (ms_per_step / 1e3) / 1e12

    # BER of the timed campaign (sanity: finite and decreasing with SNR)
    err = errors[args.warmup:].cpu().numpy().reshape(-1)
    idx = (np.arange(args.warmup * n, total * n)) % snr.size
    ber = np.bincount(idx, weights=err, minlength=snr.size) / (np.bincount(idx, minlength=snr.size) * cfg.bits_per_drop)

    out = {
        "metric": METRIC,
        "value": drops_per_s,
        "unit": "drops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox channel realisations, bits and noise drawn on the device)",
        "config": {"workload": cfg.name, "drops_per_step": n, "snr_db": snr.tolist(),
                   "bits_per_drop": cfg.bits_per_drop, "detector": cfg.detector, "csi": "ideal",
                   "workgroup_size": cfg.workgroup_size or "default"},
        "bits_per_s": drops_per_s * cfg.bits_per_drop,
        "wall_s_timed": wall,
        "ber_by_snr": [float(b) for b in ber],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_VECTOR_PEAK_TFLOPS, "traffic": HBM_BYTES_PER_DROP * n,
                     "kernel": "batch_kernel<4,4,P,T>", "flops_per_drop": algorithmic_flops_per_drop(cfg),
                     "peak_unpacked_fp32": FP32_UNPACKED_PEAK_TFLOPS,
                     "frac_unpacked_fp32": achieved / FP32_UNPACKED_PEAK_TFLOPS},
    }
    if not args.no_plain:
        plain = plain_gpu_path(cfg)
        out["plain_gpu_path"] = plain
        out["speedup_vs_plain_gpu_path"] = drops_per_s / plain["value"]
    if not args.no_estimators:
        out["estimators"] = estimator_leg()
    if not args.no_uneven:
        out["uneven_batch"] = uneven_leg()
    if not args.no_massive:
        out["massive_mimo"] = massive_leg()
    if not args.no_small:
        out["small_link"] = small_link_leg()
    if not args.no_campaign:
        out["campaign"] = campaign_leg()
    if not args.no_cpu_baseline:
        def gpu_drops(k: int):
            """The benchmark's drops 0 .. k-1 again, with their equalised symbols exported."""

            e = torch.empty(k, dtype=torch.int32, device=dev)
            q = torch.empty((k, cfg.num_tx, cfg.num_data), dtype=torch.complex64, device=dev)
            link.run(SEED, 0, npw[0][:k].contiguous(), e, q)
            torch.cuda.synchronize()
            assert np.array_equal(e.cpu().numpy(), errors[0, :k].cpu().numpy())  # the timed launches' drops
            return e.cpu().numpy(), q.cpu().numpy().astype(np.complex128)

        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget, gpu_drops=gpu_drops)
        if "fp64_check" in out["cpu_baseline"]:
            chk = out["cpu_baseline"]["fp64_check"]
            out["error_vs_fp64_oracle"] = {k: chk[k] for k in ("drops", "max_rel_err_equalized", "delta_errors",
                                                              "ber_gpu", "ber_fp64_oracle")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
