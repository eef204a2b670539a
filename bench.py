"""Headline benchmark: NLL+gradient evaluations per second of the unbinned likelihood.

Workload (one GPU): Lambda_c+ -> p K- pi+ with six resonances in all three pairings
(22 partial waves, 34 free real couplings, masses and widths fixed), 1M data events
accept-reject sampled from the model and 10M flat phase-space MC events, all resident in
HBM. One step = one FCN.nll_grad(x) call: the data log-sum and the MC normalisation with
their gradients. Couplings are seeded |g| ~ U(0.5, 2), arg g ~ U(-pi, pi).

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits: this build runs
on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tf-pwa_amd"))
sys.path.insert(0, ROOT)

CONFIG = os.path.join(ROOT, "tf-pwa_amd", "tf_pwa_amd", "configs", "lc_pkpi_6res.yml")
METRIC = "NLL+grad evals/s @1M data+10M MC, |dNLL|<1e-3 vs FP64; toy-fit p50/p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_couplings(vm, seed):
    rng = np.random.default_rng(seed)
    vals = {}
    for n in vm.trainable_vars:
        if n.endswith("r") and n[:-1] + "i" in vm.trainable_vars:
            vals[n] = rng.uniform(0.5, 2.0)
            vals[n[:-1] + "i"] = rng.uniform(-np.pi, np.pi)
    return vals


def make_workload(n_data, n_mc, seed=20261203):
    from tf_pwa_amd import ConfigLoader
    from tf_pwa_amd.phasespace import generate_phsp

    conf = ConfigLoader(CONFIG, seed=seed)
    conf.set_params(random_couplings(conf.vm, seed))
    amp = conf.get_amplitude()
    order = conf.dat_order
    masses = [conf.particles[n].mass for n in order]
    top = conf.particles[conf.top_name].mass
    rng = np.random.default_rng(seed + 1)
    t = time.time()
    mc_p4 = dict(zip(order, generate_phsp(top, masses, n_mc, rng)))
    mc = conf.data_from_p4(mc_p4)
    del mc_p4
    f_mc = amp(mc)
    fmax = 1.1 * float(f_mc.max())
    log("mc: {} events in {:.1f}s, <f>/fmax = {:.3f}".format(n_mc, time.time() - t,
                                                            f_mc.mean() / fmax))
    t = time.time()
    keep = {n: [] for n in order}
    got = 0
    chunk = 4_000_000
    while got < n_data:
        pool = dict(zip(order, generate_phsp(top, masses, chunk, rng)))
        ev = conf.data_from_p4(pool)
        f = amp(ev)
        del ev
        sel = f > rng.uniform(0.0, fmax, f.size)
        for n in order:
            keep[n].append(pool[n][sel])
        got += int(sel.sum())
    data_p4 = {n: np.concatenate(v)[:n_data] for n, v in keep.items()}
    data = conf.data_from_p4(data_p4)
    log("data: {} events in {:.1f}s".format(n_data, time.time() - t))
    return conf, data, mc, data_p4, mc


def cpu_baseline(conf, data_p4, n_data_full, n_mc_full, seconds=15.0):
    """The oracle (torch CPU float64 eager + autograd, the analogue of tf-pwa's TF eager CPU
    path) on a bounded sample, scaled per event to the full workload."""
    import torch

    from oracle import tfpwa_oracle as O

    threads = torch.get_num_threads()
    om = O.Model(CONFIG)
    om.set_params(conf.vm.get_all_dic())
    nd, nm = 4000, 40000
    d = {k: v[:nd] for k, v in data_p4.items()}
    mc = O.gen_mc(om, nm, 5)
    ad, am = O.cal_angle(om, d), O.cal_angle(om, mc)
    w, wm = O.fcn_weights(nd, None, nm)
    O.nll_grad(om, ad, am, w, wm)
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds or reps < 1:
        O.nll_grad(om, ad, am, w, wm)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    scale = (n_data_full + n_mc_full) / float(nd + nm)
    return {"value": 1.0 / (dt * scale), "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "{} data + {} MC events (oracle NLL+autograd gradient, {:.3f}s per "
                      "eval over {} evals), scaled by events to {} + {}".format(
                          nd, nm, dt, reps, n_data_full, n_mc_full)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-data", type=int, default=1_000_000)
    ap.add_argument("--n-mc", type=int, default=10_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (the counter calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (the likelihood is sharded by events across "
              "GPUs by design, but this build measures the single-GPU kernel)", file=sys.stderr)
        sys.exit(2)

    from tf_pwa_amd import FCN, _lib

    conf, data, mc, data_p4, _ = make_workload(args.n_data, args.n_mc)
    fcn = FCN(conf.get_model(), data, mc)
    x0 = np.array(conf.vm.get_all_val())
    for _ in range(args.warmup):
        fcn.nll_grad(x0)
    L = _lib.lib()
    L.tpa_timing_reset()
    L.tpa_timing_enable(1)
    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only the stream-sync helper here
        sync = lambda: None
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nll, g = fcn.nll_grad(x0)
    sync()
    elapsed = time.perf_counter() - t0
    L.tpa_timing_enable(0)
    import ctypes

    ms = ctypes.c_double(0)
    nl = ctypes.c_int64(0)
    L.tpa_timing_read(ctypes.byref(ms), ctypes.byref(nl))
    amp = conf.get_amplitude()
    bytes_per_event = amp.P * amp.lam_pad * 16 + 8
    launches_per_step = 2
    bytes_per_step = (args.n_data + args.n_mc) * bytes_per_event
    kernel_s = ms.value / 1e3
    achieved = bytes_per_step * args.steps / kernel_s / 1e9 if kernel_s > 0 else 0.0
    res = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Lc->pKpi 6 resonances, 22 partial waves, FP64 path",
                   "n_data": args.n_data, "n_mc": args.n_mc, "n_params": len(x0),
                   "path": "fp64"},
        "nll": nll,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_eval_fold", "launches": int(nl.value),
                     "kernel_ms_avg": ms.value / max(1, nl.value),
                     "bytes_per_event": bytes_per_event},
    }
    if args.calib_copy:
        import torch

        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        res["calib_copy_bytes"] = src.numel() * 8
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(conf, data_p4, args.n_data, args.n_mc)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
