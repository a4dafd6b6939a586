"""Benchmark: plain periodic PM on one MI355X at N = 2^24 particles, 256^3 grid.

One step = one KDK leapfrog step of the whole scene (deposit -> FFT Poisson +
spectral gradient -> fused CIC gather + kick + drift), inputs resident in HBM.
Prints one JSON line (rank 0) with value = particle-steps/s over exactly
--steps timed steps, the roofline of the dominant kernel (HIP events on the
launch stream inside the timed region), per-stage HBM rates, the force
rel-RMS against the FP64 Ewald oracle, and the CPU baseline (the PyTorch-CPU
FP32 restatement of the same step timed on this host).

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

METRIC = "particle-steps/s at N=2^24: surrogate vs plain PM at stated force rel-RMS vs FP64"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(name, N, G):
    """Compulsory HBM bytes per launch of each per-step kernel: every input
    read once and every output written once (FP32 particles and grids,
    complex64 half spectra of G*G*(G/2+1) points, the force grid counted at
    its 12 payload bytes per cell)."""
    G3 = G ** 3
    spec = 8 * G * G * (G // 2 + 1)
    return {
        "dep_tiles": 12 * N + 4 * G3,            # read positions, write the density grid
        "fft_rows_r2c": 4 * G3 + spec,           # density -> x half spectrum
        "fft_cols_fwd": 2 * spec,                # y transform in place
        "pm_zpass": 3 * spec,                    # z fwd + Green + z inv -> A (in place), B
        "pm_yinv": 5 * spec,                     # read A, B; write three y-inverse spectra
        "pm_xc2r": 3 * spec + 12 * G3,           # three x C2R rows -> force grid
        "gather_kdk": 48 * N + 12 * G3,          # r/w positions + velocities, read the force grid
    }.get(name)


def compulsory_bytes(stage, N, G):
    """Compulsory HBM bytes per step of each stage (every input read once and
    every output written once): deposit 12N + 4G^3; FFT Poisson + gradient
    R2C 4G^3 + S plus three C2R S + 4G^3 (S = 8 G^2 (G/2 + 1), the half
    spectrum); gather + KDK 48N + 12G^3."""
    G3 = G ** 3
    spec = 8 * G * G * (G // 2 + 1)
    return {"deposit": 12 * N + 4 * G3, "fft_poisson_grad": 16 * G3 + 4 * spec,
            "gather_kdk": 48 * N + 12 * G3}.get(stage)


STAGES = {  # stage -> kernels (their design bytes are the per-pass sums of algorithmic_bytes)
    "deposit": ("dep_bound", "dep_plan", "dep_tiles"),
    "fft_poisson_grad": ("fft_rows_r2c", "fft_cols_fwd", "pm_zpass", "pm_yinv", "pm_xc2r"),
    "gather_kdk": ("gather_kdk",),
    "resort_gate": ("pm_sort",),
}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_bench(a, N, G, pos_np, vel_np, dev, oracle, pm_value, pm_rel):
    """Train the PM correction network on the GPU (own conv kernels, Adam)
    on scenes of the same mixture (other seeds) against FP64 Ewald forces,
    state its force error on the workload's own targets, and time `a.steps`
    surrogate KDK steps of the workload (probe residual check included)."""
    import numpy as np
    import torch
    from augur_amd import _lib, ops, scenes
    from augur_amd.surrogate import PMSurrogate, rel_rms

    sur = PMSurrogate(G, 1.0, 1.0, dev, mode="correction")
    t0 = time.perf_counter()
    tuner = sur.net.autotune(1, G)   # conv tiles of this grid (forward and backward calls)
    t_tune = time.perf_counter() - t0
    torch.cuda.empty_cache()
    print(f"[bench] conv tiles autotuned in {t_tune:.2f}s: "
          f"{ {k: v['best'] for k, v in tuner.report().items()} }", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    train = []
    for k in range(a.sur_scenes):
        p, _ = scenes.cfg2(seed=1000 + k, n=N)
        pos = torch.from_numpy(p).to(dev)
        rng = np.random.default_rng(2000 + k)
        tgt = torch.from_numpy(np.sort(rng.choice(N, a.sur_targets, replace=False))).to(dev)
        train.append((pos, tgt, ops.ewald_f64(pos, 1.0 / N, tgt, L=1.0, eps=1.0 / G)))
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = [sur.train_step(*train[i % len(train)], mass_value=1.0 / N, lr=a.sur_lr) for i in range(a.sur_train_steps)]
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del train
    print(f"[bench] surrogate trained: {a.sur_train_steps} steps on {a.sur_scenes} scenes in {t_train:.1f}s "
          f"(data {t_data:.1f}s), loss {losses[0]:.3e} -> {losses[-1]:.3e}", file=sys.stderr, flush=True)

    pos0 = torch.from_numpy(pos_np).to(dev)
    rel_sur = None
    if oracle is not None:
        # the force of the timed path: a zero-length surrogate step (4-channel packed input, the output
        # layer adding into the PM planes, the fused gather), read back in the original order
        tg, a_ex = oracle
        st0 = sur.load(pos0, torch.zeros_like(pos0), mass_value=1.0 / N)
        sur.kdk(st0, 0.0, 1)
        a_sur = sur.pm.unload(st0)[2]
        rel_sur = rel_rms(a_sur[tg], a_ex)
        del st0, a_sur
        torch.cuda.empty_cache()
        print(f"[bench] surrogate force rel-RMS vs FP64 Ewald: {rel_sur:.4e} (plain PM {pm_rel:.4e})",
              file=sys.stderr, flush=True)

    st = sur.load(pos0, torch.from_numpy(vel_np).to(dev), mass_value=1.0 / N)
    probe = torch.arange(0, N, max(1, N // 4096), device=dev)
    sur.kdk(st, a.dt, max(1, a.warmup), probe=probe)
    torch.cuda.synchronize()
    # per-kernel HIP-event breakdown on its own run of the same steps, then the
    # timed run with the events of the residual conv layer only
    _lib.timing_enable(True)
    sur.kdk(st, a.dt, a.steps, probe=probe)
    torch.cuda.synchronize()
    timing = _lib.timing_collect()
    _lib.timing_enable(True)
    _lib.timing_filter("conv_fwd_res")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gate = sur.kdk(st, a.dt, a.steps, probe=probe)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    timing.update(_lib.timing_collect())
    _lib.timing_filter(None)
    _lib.timing_enable(False)
    value = N * a.steps / sec
    fell_back, residuals = gate.fell_back, gate.residuals
    res = max(residuals)
    at_2e2 = next((k for k, r in enumerate(residuals) if r > 2e-2), None)
    print(f"[bench] surrogate {a.steps} steps in {sec * 1e3:.2f} ms -> {value:.4e} particle-steps/s "
          f"(threshold {sur.check_threshold:.3e}, fell back: {fell_back}, max probe residual {res:.3e}; "
          f"a 2e-2 bound would fall back at step {at_2e2})", file=sys.stderr, flush=True)
    reuse = correction_reuse(a, sur, N, G, pos0, torch.from_numpy(vel_np).to(dev), probe, oracle, pm_value)
    conv = {}
    # algorithmic flop per voxel (2 * 27 * cin * cout of the layer; the kernels pad the 4 input
    # channels of the first layer to 16 and the 3 outputs of the last to 16, which is not counted)
    cin0 = sur.net.cin
    flops = {"conv_fwd_res": 2 * 27 * 32 * 32, "conv_fwd_gelu": 2 * 27 * cin0 * 32, "conv_fwd_out": 2 * 27 * 32 * 3}
    for name, fpv in flops.items():
        if name in timing:
            ms, cnt = timing[name]
            tf = fpv * G ** 3 / (ms / cnt / 1e3) / 1e12
            conv[name] = {"us_avg": ms / cnt * 1e3, "flop_per_voxel": fpv, "tflops": tf,
                          "frac": tf / MFMA_BF16_PEAK_TFLOPS}
    res_k = conv.get("conv_fwd_res")
    return {
        "value": value, "unit": "particle-steps/s", "ms_per_step": sec * 1e3 / a.steps,
        "speed_ratio_vs_plain_pm": value / pm_value,
        "force_rel_rms_vs_fp64": rel_sur,
        "rel_rms_ratio_vs_plain_pm": (rel_sur / pm_rel) if (rel_sur is not None and pm_rel) else None,
        "fell_back": bool(fell_back), "probe_residual": res, "probe_residuals": residuals,
        "fallback_check": {"threshold": sur.check_threshold, "probes": int(probe.numel()),
                           "rule": "twice the correction's residual against PM on its training targets "
                                   f"({sur.train_residual:.3e}), at least 2e-2",
                           "per_step": True, "on_device": True,
                           "at_2e-2_falls_back_at_step": at_2e2},
        "model": "PM + LocalNet3D correction (6 convs 3^3, width 32, BF16/FP32-acc), input [density contrast, "
                 "PM force], 3-channel correction",
        "train": {"scenes": a.sur_scenes, "steps": a.sur_train_steps, "lr": a.sur_lr, "targets": a.sur_targets,
                  "seconds": t_train, "loss_first": losses[0], "loss_last": losses[-1],
                  "data": "clumpy scenes of the workload's mixture, seeds 1000+, FP64 Ewald eps=1/G"},
        "roofline": {"kernel": "conv_fwd_res", "bound": "mfma",
                     "achieved": res_k["tflops"] if res_k else None, "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": res_k["frac"] if res_k else None, "traffic": None},
        "conv_kernels": conv,
        "conv_tiles": {"autotune_seconds": t_tune, "chosen": {k: v["best"] for k, v in tuner.report().items()}},
        "kernels": {k: {"us_avg": v[0] / v[1] * 1e3, "launches": v[1]} for k, v in timing.items()},
        "correction_reuse": reuse,
    }


def correction_reuse(a, sur, N, G, pos0, vel0, probe, oracle, pm_value, ks=(1, 2, 4, 8), steps=8, targets=4096):
    """The correction net run every k-th step only (PMSurrogate.kdk(reuse=k)):
    speed over `steps` steps of the workload for each k, and the force error at
    the stalest step (the correction made at step 0, used k - 1 steps later):
    PM(x_j) + correction(x_0) against FP64 Ewald at x_j for j = k - 1, on the
    first `targets` of the workload's targets, x_j from plain PM steps."""
    import torch
    from augur_amd import ops
    from augur_amd.surrogate import rel_rms

    out = []
    for k in ks:
        st = sur.load(pos0, vel0, mass_value=1.0 / N)
        sur.kdk(st, a.dt, steps, probe=probe, reuse=k)   # warm-up call of the same shape
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gate = sur.kdk(st, a.dt, steps, probe=probe, reuse=k)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        value = N * steps / sec
        row = {"k": k, "steps": steps, "ms_per_step": sec * 1e3 / steps, "value": value,
               "speed_ratio_vs_plain_pm": value / pm_value, "fell_back": gate.fell_back,
               "max_probe_residual": gate.residual}
        del st
        out.append(row)
    if oracle is not None:
        tg = oracle[0][:targets].contiguous()
        rho0, f0 = sur.pm.fields(pos0, mass_value=1.0 / N)
        corr = torch.empty(3, G, G, G, dtype=torch.float32, device=pos0.device)
        sur.net.forward_into(sur.packed_features(rho0, f0, (1.0 / N) * N / sur.L ** 3), corr, sur.force_scale,
                             accumulate=False)
        del rho0, f0
        for row in out:
            lag = row["k"] - 1
            st = sur.pm.load(pos0, vel0, mass_value=1.0 / N)
            if lag > 0:
                sur.pm.kdk(st, a.dt, lag)
            pj = sur.pm.unload(st)[0].contiguous()
            del st
            _, fj = sur.pm.fields(pj, mass_value=1.0 / N)
            a_ex = ops.ewald_f64(pj, 1.0 / N, tg, L=1.0, eps=1.0 / G)
            pt = pj[tg].contiguous()
            row["force_rel_rms_at_lag"] = {"lag_steps": lag, "targets": int(tg.numel()),
                                           "surrogate": rel_rms(sur.gather(fj + corr, pt), a_ex),
                                           "plain_pm": rel_rms(sur.gather(fj, pt), a_ex)}
            del fj, pj
            torch.cuda.empty_cache()
        del corr
    for row in out:
        e = row.get("force_rel_rms_at_lag", {})
        print(f"[bench] correction every {row['k']} steps: {row['ms_per_step']:.3f} ms/step "
              f"({row['speed_ratio_vs_plain_pm']:.3f}x plain PM), force rel-RMS "
              f"{e.get('surrogate', float('nan')):.4e} after {e.get('lag_steps', '-')} steps "
              f"(plain PM {e.get('plain_pm', float('nan')):.4e})", file=sys.stderr, flush=True)
    return out


def measured_traffic(kernel, N, G):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by what the
    same counters read for a copy of known size), or None when absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"pm_traffic_n{N}_g{G}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) >= 8 and w[0] == "kernel" and w[1] == kernel and w[6] == "traffic_bytes":
                return float(w[7])
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or None


def cpu_kdk_steps(pos, vel, G, dt, steps):
    """Median seconds per KDK step of the PyTorch-CPU FP32 port (after one warm-up step)."""
    import torch
    from oracle import torch_cpu as tc
    n = pos.shape[0]
    p, v = torch.from_numpy(pos), torch.from_numpy(vel)
    m = torch.full((n,), 1.0 / n)
    a = tc.pm_acc(p, m, G)
    p, v, a = tc.kdk_step(p, v, a, m, G, dt)
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        p, v, a = tc.kdk_step(p, v, a, m, G, dt)
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], sum(ts)


def cpu_baseline(pos, vel, G, dt, steps=1, extra=True):
    """The PyTorch-CPU FP32 port timed on this host (all cores): the workload's
    KDK step, and (extra) the Plummer cold collapse over all its 100 steps and
    the 2D reference-shape pm_acc (G=128, N=1e6).  Median per step after a
    warm-up, the timing convention of scripts/scatter_bh_1b.py:296-300."""
    import numpy as np
    import torch
    from augur_amd import scenes
    from oracle import torch_cpu as tc
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    n = pos.shape[0]
    sec, _ = cpu_kdk_steps(pos, vel, G, dt, steps)
    out = {"value": n / sec, "unit": "particle-steps/s", "cores": torch.get_num_threads(), "kind": "port",
           "cpu_model": cpu_model(),
           "sample": f"{steps} KDK step(s) (after a warm-up step) of the same N={n}, {G}^3 workload with the "
                     f"PyTorch-CPU FP32 restatement (oracle/torch_cpu.py), {sec:.2f} s/step"}
    if extra:
        p1, v1 = scenes.cfg1()
        s1, tot = cpu_kdk_steps(p1, v1, 64, 1e-3, 100)
        out["plummer_cold_collapse"] = {"N": len(p1), "grid": 64, "steps": 100, "s_per_step_median": s1,
                                        "seconds_total": tot, "value": len(p1) / s1, "unit": "particle-steps/s"}
        rng = np.random.default_rng(4738)
        n2 = 1000000
        p2 = torch.from_numpy(rng.normal(0.0, 8.0, (n2, 2)).astype(np.float32))
        m2 = torch.ones(n2)
        tc.pm_acc_2d(p2, m2, 128, 64.0)
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            tc.pm_acc_2d(p2, m2, 128, 64.0)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        out["pm_acc_2d"] = {"N": n2, "grid": 128, "extent": 64.0, "ms_median": ts[2] * 1e3,
                            "note": "the reference's 2D plain-PM force (scripts/pm_baseline_force_error.py:68-76) "
                                    "on a Gaussian scene; 1.5 ms on the reference's GPU"}
    return out


def large_surrogate_train(a, N, G, NC, dev):
    """The correction net for the large workload, trained at its own size: one
    scene of the same mixture (another seed) drawn on the GPU, FP64 Ewald
    forces on seeded targets, its PM grids made once (the scene is then
    dropped), conv tiles autotuned at G^3, then Adam steps on those grids."""
    import numpy as np
    import torch
    from augur_amd import ops, scenes
    from augur_amd.surrogate import PMSurrogate

    t0 = time.perf_counter()
    sur = PMSurrogate(G, 1.0, 1.0, dev, mode="correction")
    pos, vel = scenes.clumpy_device(N, dev, seed=1000, n_clumps=NC)
    del vel
    tg = torch.from_numpy(scenes.targets(N, a.targets, seed=1000)).to(dev)
    a_true = ops.ewald_f64(pos, 1.0 / N, tg, L=1.0, eps=1.0 / G)
    pt = pos[tg].contiguous()
    rho, force = sur.pm.fields(pos, mass_value=1.0 / N)
    del pos
    sur.pm._scratch = None
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    t_data = time.perf_counter() - t0
    sur.calibrate(rho, force)
    t0 = time.perf_counter()
    sur.net.autotune(1, G)
    torch.cuda.empty_cache()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = [sur.train_step_grids(rho, force, pt, a_true, lr=a.sur_lr) for _ in range(a.large_sur_steps)]
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del rho, force, pt, a_true
    torch.cuda.empty_cache()
    print(f"[bench] large surrogate: data {t_data:.1f}s, tiles {t_tune:.1f}s, {a.large_sur_steps} steps in "
          f"{t_train:.1f}s, loss {losses[0]:.3e} -> {losses[-1]:.3e}", file=sys.stderr, flush=True)
    info = {"scene": f"same mixture, seed 1000, N={N}", "targets": a.targets, "steps": a.large_sur_steps,
            "lr": a.sur_lr, "seconds_data": t_data, "seconds_tune": t_tune, "seconds_train": t_train,
            "loss_first": losses[0], "loss_last": losses[-1]}
    return sur, info


def large_bench(a):
    """The single-GPU large-scale workload: N = 2^30 (50% uniform + 50% in 4096
    Gaussian clumps, sigma 0.02, drawn on the GPU), 512^3 grid, KDK steps with
    the FP64 momentum and kinetic energy of every step (fused into the gather);
    force rel-RMS of the initial PM forces vs FP64 Ewald on seeded targets;
    peak device memory.  Then (unless --no-surrogate) the same number of
    surrogate KDK steps with a correction net trained at this size, with its
    force rel-RMS on the same targets."""
    import numpy as np
    import torch
    from augur_amd import _lib, ops, scenes
    from augur_amd.pm import PeriodicPM
    from augur_amd.surrogate import rel_rms

    dev = torch.device("cuda:0")
    N, G, NC = 1 << 30, 512, 4096
    sur = sur_info = None
    if not a.no_surrogate:
        sur, sur_info = large_surrogate_train(a, N, G, NC, dev)
        sur_info["peak_device_bytes_training"] = torch.cuda.max_memory_allocated()
        torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    pos, vel = scenes.clumpy_device(N, dev, seed=4738, n_clumps=NC)
    torch.cuda.synchronize()
    print(f"[bench] large scene N={N} drawn on the GPU in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    t0 = time.perf_counter()
    st = pm.load(pos, vel, mass_value=1.0 / N)
    del pos, vel
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    print(f"[bench] loaded (sort + initial PM forces) in {time.perf_counter() - t0:.2f}s", file=sys.stderr,
          flush=True)
    rel = None
    if not a.no_oracle:
        # targets are sorted slots (the state's order); all N sources
        tg = torch.from_numpy(scenes.targets(N, a.targets)).to(dev)
        t0 = time.perf_counter()
        a_ex = ops.ewald_f64(st.pos, 1.0 / N, tg, L=1.0, eps=1.0 / G)
        a_pm = st.acc[tg].double()
        rel = float((((a_pm - a_ex) ** 2).sum(-1).mean() / (a_ex ** 2).sum(-1).mean()).sqrt())
        print(f"[bench] PM force rel-RMS vs FP64 Ewald on {a.targets} targets: {rel:.4e} "
              f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
        if sur is not None:
            f, _ = sur.field(st.pos, mass_value=1.0 / N)
            sur.pm._scratch = None
            rel_sur = rel_rms(sur.gather(f, st.pos[tg].contiguous()), a_ex)
            del f
            torch.cuda.empty_cache()
            sur_info["force_rel_rms_vs_fp64"] = rel_sur
            sur_info["rel_rms_ratio_vs_plain_pm"] = rel_sur / rel
            print(f"[bench] surrogate force rel-RMS vs FP64 Ewald: {rel_sur:.4e} (plain PM {rel:.4e})",
                  file=sys.stderr, flush=True)
    m = 1.0 / N
    pm.kdk(st, a.dt, a.warmup) if a.warmup > 0 else None
    # per-kernel HIP-event breakdown on its own run, then the timed run with
    # the events of the dominant kernel only
    _lib.timing_enable(True)
    pm.kdk(st, a.dt, a.steps, diag=True)
    torch.cuda.synchronize()
    timing = _lib.timing_collect()
    dname = max(timing.items(), key=lambda kv: kv[1][0])[0]
    em0 = ops.energy_momentum_f64(st.vel, None, m)[:4]
    torch.cuda.synchronize()
    _lib.timing_enable(True)
    _lib.timing_filter(dname)
    t0 = time.perf_counter()
    # the FP64 momentum / kinetic sums of every step are fused into the gather
    diag = pm.kdk(st, a.dt, a.steps, diag=True)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    timing.update(_lib.timing_collect())
    _lib.timing_filter(None)
    _lib.timing_enable(False)
    value = N * a.steps / sec
    em = torch.cat([em0[None], diag]).cpu().numpy()
    print(f"[bench] {a.steps} steps (FP64 sums fused) in {sec * 1e3:.1f} ms -> {value:.4e} particle-steps/s",
          file=sys.stderr, flush=True)
    if sur is not None:
        # surrogate KDK steps of the same state (probe residual check on the first step of each call)
        probe = torch.arange(0, N, N // 4096, device=dev)
        sur.kdk(st, a.dt, 1, probe=probe)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fell_back, res = sur.kdk(st, a.dt, a.steps, probe=probe)
        torch.cuda.synchronize()
        sec_s = time.perf_counter() - t1
        sur_info.update({"value": N * a.steps / sec_s, "unit": "particle-steps/s", "ms_per_step": sec_s * 1e3 / a.steps,
                         "speed_ratio_vs_plain_pm": (N * a.steps / sec_s) / value, "fell_back": bool(fell_back),
                         "probe_residual": res})
        print(f"[bench] surrogate {a.steps} steps in {sec_s * 1e3:.1f} ms -> {N * a.steps / sec_s:.4e} "
              f"particle-steps/s (fell back: {fell_back})", file=sys.stderr, flush=True)
    kernels = {k: {"ms_total": v[0], "launches": v[1], "us_avg": v[0] / v[1] * 1e3} for k, v in timing.items()}
    dom = max(timing.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcnt) = dom
    ab = algorithmic_bytes(dname, N, G)
    achieved = ab / (dms / dcnt / 1e3) / 1e9 if ab else None
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": sec * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain periodic PM, N={N} (clumpy: 50% uniform + {NC} Gaussian clumps sigma 0.02, "
                               f"drawn on the GPU), {G}^3 grid, KDK dt={a.dt}, Maxwellian sigma_v=0.01, FP64 "
                               "momentum and kinetic energy of every step (fused into the gather, timed)",
                   "N": N, "grid": G, "dt": a.dt, "path": "plain_pm"},
        "force_rel_rms_vs_fp64": rel, "oracle": f"FP64 Ewald (periodic, eps=1/{G}) on {a.targets} targets",
        "momentum_drift": (np.abs(em[-1, :3] - em[0, :3])).tolist(), "kinetic_energy_per_step": em[:, 3].tolist(),
        "peak_device_bytes": torch.cuda.max_memory_allocated(),
        "roofline": {"kernel": dname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "avg_launch_us": dms / dcnt * 1e3, "algorithmic_bytes": ab},
        "kernels": kernels, "cpu_baseline": None, "surrogate": sur_info,
    }
    print(json.dumps(out), flush=True)


def cfg1_bench(a):
    """The Plummer cold collapse: N = 65,536 equal masses (m = 1/N) in a
    Plummer sphere of scale 0.05 centred in the unit periodic box, at rest,
    64^3 grid, 100 KDK steps of dt = 1e-3.  Force rel-RMS of the initial PM
    forces against FP64 sums with softening 1/64 on 4096 seeded targets: the
    isolated direct sum (the reference's exact_acc) and the periodic Ewald sum
    (the same boundary as the PM solve).  The 100 steps are timed as one
    eager call and as a HIP-graph replay of the same call (the state restored
    before each, outside the timed region)."""
    import numpy as np
    import torch
    from augur_amd import _lib, ops, scenes
    from augur_amd.pm import PeriodicPM

    dev = torch.device("cuda:0")
    N, G, dt, steps, eps = 65536, 64, 1e-3, 100, 1.0 / 64
    pos_np, vel_np = scenes.cfg1()
    pos, vel = torch.from_numpy(pos_np).to(dev), torch.from_numpy(vel_np).to(dev)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    tg = torch.from_numpy(scenes.targets(N, 4096)).to(dev)
    a_pm = pm.accel(pos, mass_value=1.0 / N)[tg].double()
    m = torch.full((N,), 1.0 / N, dtype=torch.float64, device=dev)
    a_dir = ops.direct_sum_f64(pos, eps, m, tg)
    a_ew = ops.ewald_f64(pos, 1.0 / N, tg, L=1.0, eps=eps)
    rr = lambda x, y: float((((x - y) ** 2).sum(-1).mean() / (y ** 2).sum(-1).mean()).sqrt())
    rel_dir, rel_ew, rel_ew_dir = rr(a_pm, a_dir), rr(a_pm, a_ew), rr(a_ew, a_dir)
    print(f"[bench] cfg1 PM force rel-RMS vs FP64 direct (isolated, eps 1/64) {rel_dir:.4e}, vs FP64 Ewald "
          f"(periodic, eps 1/64) {rel_ew:.4e}; Ewald vs isolated {rel_ew_dir:.4e}", file=sys.stderr, flush=True)
    # warm-up on a scratch state (kernel loads), then the timed calls
    w = pm.load(pos, vel, mass_value=1.0 / N)
    pm.kdk(w, dt, max(1, a.warmup))
    del w
    st = pm.load(pos, vel, mass_value=1.0 / N)
    saved = [t.clone() for t in (st.pos, st.vel, st.acc, st.ids, st.ws)]
    em0 = ops.energy_momentum_f64(st.vel, None, 1.0 / N)[:4]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pm.kdk(st, dt, steps)
    torch.cuda.synchronize()
    sec_eager = time.perf_counter() - t0
    p_eager = st.pos.clone()
    _lib.timing_enable(True)
    for t, sv in zip((st.pos, st.vel, st.acc, st.ids, st.ws), saved):
        t.copy_(sv)
    pm.kdk(st, dt, steps)
    torch.cuda.synchronize()
    timing = _lib.timing_collect()
    _lib.timing_enable(False)
    for t, sv in zip((st.pos, st.vel, st.acc, st.ids, st.ws), saved):
        t.copy_(sv)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        pm.kdk(st, dt, steps)
    secs = []
    for _ in range(3):
        for t, sv in zip((st.pos, st.vel, st.acc, st.ids, st.ws), saved):
            t.copy_(sv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        graph.replay()
        torch.cuda.synchronize()
        secs.append(time.perf_counter() - t0)
    sec = sorted(secs)[1]
    same = bool(torch.equal(st.pos, p_eager))
    em1 = ops.energy_momentum_f64(st.vel, None, 1.0 / N)[:4]
    p, v, _ = (t.cpu().double().numpy() for t in pm.unload(st))
    r = np.linalg.norm(p - 0.5, axis=1)
    value = N * steps / sec
    print(f"[bench] cfg1 {steps} steps: graph replay {sec * 1e3:.2f} ms ({value:.4e} particle-steps/s), eager "
          f"{sec_eager * 1e3:.2f} ms; replay equals eager: {same}", file=sys.stderr, flush=True)
    kernels = {k: {"ms_total": vv[0], "launches": vv[1], "us_avg": vv[0] / vv[1] * 1e3} for k, vv in timing.items()}
    cpu = None
    if not a.no_cpu_baseline:
        import torch as _t
        _t.set_num_threads(os.cpu_count() or 1)
        s1, tot = cpu_kdk_steps(pos_np, vel_np, G, dt, steps)
        cpu = {"value": N / s1, "unit": "particle-steps/s", "cores": _t.get_num_threads(), "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"all {steps} KDK steps of the same workload with the PyTorch-CPU FP32 restatement "
                         f"(oracle/torch_cpu.py): {tot:.2f} s, median {s1 * 1e3:.1f} ms/step"}
        print(f"[bench] CPU baseline {cpu['value']:.3e} p-steps/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)
    out = {
        "metric": "particle-steps/s, plain PM cold collapse", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": sec * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Plummer cold collapse, N={N} (a=0.05, centred, v=0, m=1/N), {G}^3 periodic grid, "
                               f"{steps} KDK steps dt={dt}", "N": N, "grid": G, "dt": dt, "path": "plain_pm"},
        "force_rel_rms_vs_fp64": rel_dir,
        "oracle": "FP64 isolated direct sum (eps=1/64) on 4096 seeded targets, initial scene",
        "force_rel_rms_vs_fp64_ewald": rel_ew, "ewald_vs_isolated_rel_rms": rel_ew_dir,
        "eager_ms_per_step": sec_eager * 1e3 / steps, "graph_replay_equals_eager": same,
        "momentum_drift": (em1[:3] - em0[:3]).abs().cpu().tolist(),
        "lagrangian_radii_final": {q: float(np.quantile(r, q)) for q in (0.1, 0.5, 0.9)},
        "roofline": None, "roofline_note": "16.6 MB of compulsory bytes per step sit in the 256 MB Infinity "
                                           "Cache; the step is launch-bound (about 30 launches), not HBM-bound",
        "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def train_bench(a):
    """Batched surrogate training: 64 scenes per batch with particle counts
    drawn log-uniformly in [2^14, 2^20] (Plummer with a in [0.02, 0.1], or the
    clumpy mix), 128^3 grids, FP64 Ewald targets on 8192 particles per scene,
    the gradient step captured in a HIP graph and replayed, Adam lr 3e-4 on
    FP32 masters.  Reports training iterations/s and the network's MFMA rate."""
    import numpy as np
    import torch
    from augur_amd import _lib, ops, scenes
    from augur_amd.surrogate import PMSurrogate, BatchGraph, ragged_batch, batch_grads, _batch_grids

    dev = torch.device("cuda:0")
    B, G, T = a.train_scenes, 128, a.train_targets
    rng = np.random.default_rng(4738)
    sizes = np.exp(rng.uniform(np.log(2 ** 14), np.log(2 ** 20), B)).astype(np.int64)
    sur = PMSurrogate(G, 1.0, 1.0, dev, mode="correction")
    batches = []
    t0 = time.perf_counter()
    for k in range(2):
        sc = []
        for i, n in enumerate(sizes):
            r = np.random.default_rng(1000 * k + i)
            p = (scenes.plummer(int(n), float(r.uniform(0.02, 0.1)), r) if i % 2 == 0
                 else scenes.clumpy(int(n), r, n_clumps=int(r.integers(4, 64)), sigma=float(r.uniform(0.01, 0.05))))
            sc.append(scenes.to_box_f32(p))
        pos, starts, mass = ragged_batch(sc, dev)
        tp, at = [], []
        for i, p in enumerate(sc):
            tg = torch.from_numpy(scenes.targets(len(p), T, seed=1000 * k + i)).to(dev)
            pb = pos[int(starts[i]): int(starts[i + 1])]
            at.append(ops.ewald_f64(pb, 1.0 / len(p), tg, L=1.0, eps=1.0 / G))
            tp.append(pb[tg])
        tpos, a_true = torch.stack(tp).contiguous(), torch.stack(at).contiguous()
        inv_norm = (1.0 / (a_true ** 2).sum()).reshape(1)
        dens = torch.ones(B, dtype=torch.float32, device=dev)
        batches.append((pos, starts, mass, tpos, a_true, inv_norm, dens))
    torch.cuda.synchronize()
    total = int(sizes.sum())
    print(f"[bench] {B} scenes x 2 batches ({total} particles per batch, {T} FP64 targets per scene) in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    rho, force = _batch_grids(sur, *batches[0][:3])
    sur.calibrate(rho[0], force[0])
    del rho, force
    t0 = time.perf_counter()
    tuner = sur.net.autotune(B, G)   # conv tiles of the batch (forward and backward calls)
    t_tune = time.perf_counter() - t0
    torch.cuda.empty_cache()
    print(f"[bench] conv tiles autotuned in {t_tune:.2f}s", file=sys.stderr, flush=True)
    # one eager step with the kernel timers on (they are not recorded inside a replay)
    _lib.timing_enable(True)
    loss0, grads = batch_grads(sur, *batches[0])
    sur.net.adam_step(grads, lr=3e-4, loss=loss0)
    torch.cuda.synchronize()
    timing = _lib.timing_collect()
    _lib.timing_enable(False)
    del grads
    torch.cuda.empty_cache()
    graph = BatchGraph(sur, *batches[0])
    for k in range(a.warmup):
        graph.step(*batches[k % 2], lr=3e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = []
    for k in range(a.steps):
        losses.append(graph.step(*batches[k % 2], lr=3e-4).clone())
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    it_s = a.steps / sec
    stepping = batched_stepping(a, sur, batches[0][:3], G, total, dev)
    # network flop per voxel: forward 2*27*(4*32 + 4*32*32 + 32*3); backward data and weight
    # gradients twice that (the first layer's data gradient is not needed)
    fwd = 2 * 27 * (4 * 32 + 4 * 32 * 32 + 32 * 3)
    flop = 3 * fwd * B * G ** 3
    conv_ms = sum(v[0] for k, v in timing.items() if k.startswith("conv") or k == "gelu_bwd")
    out = {
        "metric": "batched surrogate training iterations/s (64 ragged scenes, 128^3)", "value": it_s,
        "unit": "iterations/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": sec * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (FP32 accumulate, FP32 masters)", "data": "synthetic",
        "config": {"workload": f"{B} scenes per batch, N log-uniform in [2^14, 2^20] (Plummer a~U[0.02,0.1] / "
                               f"clumpy), {G}^3 grids, {T} FP64 Ewald targets per scene, gradient step in a "
                               "HIP graph, Adam lr 3e-4", "scenes": B, "grid": G, "particles_per_batch": total},
        "scene_steps_per_s": it_s * B, "particles_per_s": it_s * total,
        "network_tflops": flop / (sec / a.steps) / 1e12, "mfma_frac": flop / (sec / a.steps) / 1e12 / 2500.0,
        "eager_step_kernels_ms": {k: v[0] for k, v in timing.items()}, "eager_step_conv_ms": conv_ms,
        "conv_tiles": {"autotune_seconds": t_tune, "chosen": {k: v["best"] for k, v in tuner.report().items()}},
        "loss_first": float(loss0), "loss_last": float(losses[-1]),
        "batched_stepping": stepping,
        "peak_device_bytes": torch.cuda.max_memory_allocated(),
    }
    print(json.dumps(out), flush=True)


def batched_stepping(a, sur, batch, G, total, dev):
    """The batch's scenes stepped together: `a.steps` KDK steps of all of them
    (ragged deposit, batched FFT solve, optionally the trained net at batch B,
    one persistent gather/kick/drift launch per step) captured in one HIP
    graph, replayed once to warm up and once timed; plain PM and surrogate."""
    import torch
    from augur_amd import _lib
    from augur_amd.batch import RaggedStepper
    pos, starts, mass = batch
    out = {}
    for name, s in (("plain_pm", None), ("surrogate", sur)):
        st = RaggedStepper(G, device=dev, surrogate=s)
        _, P, M = st.sort(pos, starts, mass)   # each scene in cell order (set-up, not timed)
        V = torch.zeros_like(P)
        mass = M
        A = st.accel(P, starts, mass)
        graph = st.capture(P, V, A, starts, mass, a.dt, a.steps)
        graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        graph.replay()
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        # per-kernel breakdown from one eager call of the same steps
        _lib.timing_enable(True)
        st.kdk(P, V, A, starts, mass, a.dt, a.steps)
        torch.cuda.synchronize()
        tm = _lib.timing_collect()
        _lib.timing_enable(False)
        out[name] = {"particle_steps_per_s": total * a.steps / sec, "ms_per_step": sec * 1e3 / a.steps,
                     "steps": a.steps, "dt": a.dt,
                     "step_ragged_us_avg": tm["step_ragged"][0] / tm["step_ragged"][1] * 1e3
                     if "step_ragged" in tm else None,
                     "eager_kernels_ms": {k: v[0] for k, v in tm.items()}}
        print(f"[bench] batched {name}: {a.steps} steps of {total} particles in {sec * 1e3:.1f} ms -> "
              f"{out[name]['particle_steps_per_s']:.3e} particle-steps/s", file=sys.stderr, flush=True)
        del graph, P, V, A
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["default", "large", "train", "cfg1"], default="default",
                    help="large: N=2^30 on a 512^3 grid (4096 clumps), PM steps with FP64 reductions; "
                         "train: batched surrogate training on 64 ragged scenes at 128^3")
    ap.add_argument("--train-scenes", type=int, default=64)
    ap.add_argument("--train-targets", type=int, default=8192)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--targets", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-scenes", type=int, default=3, help="training scenes of the surrogate")
    ap.add_argument("--sur-train-steps", type=int, default=120)
    ap.add_argument("--sur-lr", type=float, default=1e-3)
    ap.add_argument("--sur-targets", type=int, default=4096, help="FP64 target particles per training scene")
    ap.add_argument("--large-sur-steps", type=int, default=40, help="surrogate training steps of the large workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, run one 256 MiB device-to-device copy (the byte-counter "
                         "calibration kernel of a rocprofv3 --pmc run)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "train":
        train_bench(a)
        return
    if a.workload == "cfg1":
        cfg1_bench(a)
        return
    if a.workload == "large":
        if a.steps == 10 and "--steps" not in sys.argv:
            a.steps = 5
        if a.targets == 16384 and "--targets" not in sys.argv:
            a.targets = 4096
        large_bench(a)
        return

    import numpy as np
    import torch
    from augur_amd import _lib, ops, scenes
    from augur_amd.pm import PeriodicPM

    dev = torch.device("cuda:0")
    N, G = a.n, a.grid
    t0 = time.perf_counter()
    pos_np, vel_np = scenes.cfg2(n=N)
    print(f"[bench] scene N={N} G={G} built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    st = pm.load(torch.from_numpy(pos_np).to(dev), torch.from_numpy(vel_np).to(dev), mass_value=1.0 / N)
    torch.cuda.synchronize()

    # force error of the plain PM path vs the FP64 periodic (Ewald) oracle,
    # on seeded targets of the initial scene
    rel = None
    if not a.no_oracle:
        tg = torch.from_numpy(scenes.targets(N, a.targets)).to(dev)
        pos0 = torch.from_numpy(pos_np).to(dev)
        a_pm = pm.accel(pos0, mass_value=1.0 / N)[tg].double()
        a_ex = ops.ewald_f64(pos0, 1.0 / N, tg, L=1.0, eps=1.0 / G)
        rel = float((((a_pm - a_ex) ** 2).sum(-1).mean() / (a_ex ** 2).sum(-1).mean()).sqrt())
        del pos0, a_pm
        print(f"[bench] PM force rel-RMS vs FP64 Ewald on {a.targets} targets: {rel:.4e}", file=sys.stderr,
              flush=True)

    pm.kdk(st, a.dt, a.warmup) if a.warmup > 0 else None
    torch.cuda.synchronize()
    # per-stage HIP-event breakdown on its own run of the same steps (an event
    # pair around every stage adds launch gaps), then the timed run with the
    # events of the dominant kernel only
    _lib.timing_enable(True)
    pm.kdk(st, a.dt, a.steps)
    torch.cuda.synchronize()
    timing = _lib.timing_collect()
    dom = max(timing.items(), key=lambda kv: kv[1][0])
    dname = dom[0]
    _lib.timing_enable(True)
    _lib.timing_filter(dname)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pm.kdk(st, a.dt, a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dtime = _lib.timing_collect()
    _lib.timing_filter(None)
    _lib.timing_enable(False)
    sec = t1 - t0
    value = N * a.steps / sec
    print(f"[bench] {a.steps} steps in {sec * 1e3:.2f} ms -> {value:.4e} particle-steps/s", file=sys.stderr,
          flush=True)

    if a.calibrate_copy:
        src = torch.ones(1 << 26, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 268435456 bytes read, 268435456 bytes written", file=sys.stderr, flush=True)
        del src, dst

    # dominant kernel (its launches in the timed run) and per-stage rates
    dms, dcnt = dtime[dname]
    avg_s = dms / dcnt / 1e3
    ab = algorithmic_bytes(dname, N, G)
    achieved = ab / avg_s / 1e9 if ab else None
    traffic = measured_traffic(dname, N, G)
    roofline = {"kernel": dname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "avg_launch_us": avg_s * 1e6, "algorithmic_bytes": ab}
    stages = {}
    for sname, kern in STAGES.items():
        ms = sum(timing.get(k, (0.0, 0))[0] for k in kern) / a.steps
        b = compulsory_bytes(sname, N, G)
        design = sum(algorithmic_bytes(k, N, G) or 0 for k in kern)
        stages[sname] = {"ms_per_step": ms, "algorithmic_bytes": b,
                         "GB_s": b / (ms / 1e3) / 1e9 if ms > 0 and b else None,
                         "frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS if ms > 0 and b else None,
                         "design_bytes": design or None}
    kernels = {k: {"ms_total": v[0], "launches": v[1], "us_avg": v[0] / v[1] * 1e3} for k, v in timing.items()}

    sur = None
    if not a.no_surrogate:
        sur = surrogate_bench(a, N, G, pos_np, vel_np, dev, None if a.no_oracle else (tg, a_ex), value, rel)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(pos_np, vel_np, G, a.dt)
        print(f"[bench] CPU baseline {cpu['value']:.3e} p-steps/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": sec * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain periodic PM, N={N} (clumpy: 50% uniform + 64 Gaussian clumps sigma 0.02), "
                               f"{G}^3 grid, KDK dt={a.dt}, Maxwellian sigma_v=0.01",
                   "N": N, "grid": G, "dt": a.dt, "path": "plain_pm"},
        "force_rel_rms_vs_fp64": rel, "oracle": f"FP64 Ewald (periodic, eps=1/{G})",
        "surrogate": sur,
        "roofline": roofline, "stages": stages, "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
