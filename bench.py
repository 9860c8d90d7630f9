#!/usr/bin/env python3
"""Throughput of the MI355X fastDETECT2 optical transport.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (default): 100,000 pencil-beam x-ray histories at 25 keV on a
150-um CsI layer with 2.55-um columns (~2.0e5 deposition events, ~1.1e8
optical photons, synthetic events), the plain Monte Carlo transport.  A step
is one pass of the transport over that batch: device work-queue build plus
the persistent transport kernel, events already resident in HBM.  The JSON
line reports optical photons/s over K timed steps, the FP32-VALU roofline of
the transport kernel (algorithmic FLOPs from hybridmantis_amd/flops.py with
per-photon primitive counts measured by the oracle on a slice of the same
workload) and a CPU baseline: the project's restatement of the reference C
path (oracle/, the C path's own arithmetic and random streams) run on a
bounded sample in one process per host core, built -O2 and -O0.

Two more legs go into the same line: "batched" (4 geometries x 15-90 keV,
50,000 histories each, autotuned launch shape, HIP-graph replays) and
"surrogate" (1e6 histories at 50 keV: surrogate + MC fallback against the
full plain MC -- histories/s of both, PRF L1, detected-fraction and Swank
errors).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "optical photons/s (plain MC); histories/s surrogate vs MC at <=2% PRF error"
FP32_PEAK_TFLOPS = 157.3


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """CPU cores this process may use; the one-GPU box allots 16 per GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(p, ev, workers: int, photons_per_worker: int):
    """The project's restatement of the reference C path (oracle/fdo.c in
    the C path's own float/double arithmetic on its per-event RANECU
    streams: main/hybridMANTIS_c_ver1_0.c driver + the C branch of
    main/kernel_cuda_c_ver1_0.cu; bit-identical to the compiled reference on
    the golden fixtures) on the first events of the workload, one process
    per core over contiguous slices, ~photons_per_worker photons each.
    Timed built -O2 (the value) and -O0 (how the reference builds it,
    main/compile_ver1_0.sh:9-11), each with and without the driver's two
    per-event clears of an unused 501 x 501 u64 image
    (main/hybridMANTIS_c_ver1_0.c:205-207, 248-250)."""
    import numpy as np

    from oracle import pyoracle as po

    cum = np.cumsum(ev["N"].astype(np.int64))
    n_ev = int(np.searchsorted(cum, photons_per_worker * workers)) + 1
    sample = ev[:n_ev]
    photons = int(sample["N"].sum())
    po.build()
    rates = {}
    for level, path in (("O2", po.LIB_PATH), ("O0", po.LIB_O0_PATH)):
        for rezero in (False, True):
            os.environ["FDO_LIBRARY"] = path  # the spawned workers load this build
            try:
                _, secs = po.run_parallel(p, sample, p["num_hist"], workers=workers,
                                          rng=po.RNG_RANECU, seed=123456789,
                                          math=po.MATH_FAITHFUL, rezero=rezero)
            finally:
                os.environ.pop("FDO_LIBRARY", None)
            rates[(level, rezero)] = (photons / max(secs), max(secs))
    v, wall = rates[("O2", False)]
    return {"value": v, "unit": "optical photons/s", "cores": workers, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "value_with_image_rezero": rates[("O2", True)][0],
            "value_O0": rates[("O0", False)][0],
            "value_O0_with_image_rezero": rates[("O0", True)][0],
            "sample": f"first {n_ev} events ({photons} photons) of the workload; the "
                      "reference C path restated (oracle faithful arithmetic, per-event "
                      f"RANECU), {workers} processes over contiguous slices; slowest worker "
                      f"{wall:.1f} s at -O2, {rates[('O0', False)][1]:.1f} s at -O0; value "
                      "= -O2 without the driver's per-event image clears"}


def flops_model(p, ev, n_photons: int = 60000):
    import numpy as np

    from hybridmantis_amd import flops
    from oracle import pyoracle as po

    cum = np.cumsum(ev["N"].astype(np.int64))
    n_ev = int(np.searchsorted(cum, n_photons)) + 1
    r = po.run(p, ev[:n_ev], p["num_hist"], math=po.MATH_DEVICE)
    ph = float(r.counters[0])
    return flops.flops_per_photon(r.work, ph), flops.per_photon(r.work, ph)


def traffic_from_profile(name: str):
    """HBM bytes per launch recorded by the rocprofv3 PMC runs in
    profiles/<name> (line 'traffic_bytes_per_launch: N')."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def batched_leg(steps: int) -> dict:
    """Config-2 batch: 4 geometries x 15-90 keV, 50,000 histories each, in
    one persistent launch replayed from a HIP graph, with the launch shape
    picked by the autotuner on every 10th event of the batch."""
    from hybridmantis_amd import events, fd2

    plist, ev = events.batched_workload()
    n_ph = int(ev["N"].sum())
    with fd2.Transport(plist) as tr:
        buf = tr.upload_events(ev)
        sample = ev[::10]
        sbuf = tr.upload_events(sample)
        best, results = tr.autotune(sbuf, len(sample), repeats=1)
        sbuf.free()
        ms = []
        with tr.graph(buf, len(ev)) as g:
            g.launch()
            tr.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                g.launch()
                tr.sync()
                ms.append(tr.last_kernel_ms())
            wall = time.perf_counter() - t0
        gen = sum(tr.counters(k)["generated"] for k in range(len(plist)))
        assert gen == n_ph * (steps + 1)
        buf.free()
    return {"workload": "4 geometries (H, R) = (100, 2.55), (150, 2.55), (300, 4.0), "
                        "(500, 5.0) um x uniform 15-90 keV, 50,000 histories each, "
                        "one HIP-graph launch per step",
            "events": int(len(ev)), "photons": n_ph, "steps": steps,
            "photons_per_s": n_ph * steps / wall, "ms_per_launch": sum(ms) / len(ms),
            "histories_per_s": 4 * 50000 * steps / wall,
            "tuning": best, "autotune_candidates": len(results)}


BF16_PEAK_TFLOPS = 2500.0
MLP_FWD_FLOP = 2.0 * (5 * 256 + 2 * 256 * 256 + 256 * 65)  # per row


def mfma_rates() -> dict:
    """The surrogate network's BF16 MFMA GEMMs against the dense BF16 peak:
    inference = the forward pass on a device-resident batch of 65,536 rows
    (the surrogate path's batch), training = Adam steps at batch 4096
    (forward, loss, backward, update; 3x the forward's FLOPs per row)."""
    import numpy as np

    from hybridmantis_amd import fd2, surrogate as S

    out = {"peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s"}
    net = S.load()
    try:
        ms = net.time_forward(65536, reps=50)
        t = MLP_FWD_FLOP * 65536 / (ms * 1e-3) / 1e12
        out["inference"] = {"batch": 65536, "ms_per_forward": ms, "achieved": t,
                            "frac": t / BF16_PEAK_TFLOPS}
    finally:
        net.close()
    rng = np.random.default_rng(3)
    x = rng.random((65536, 5)).astype(np.float32)
    y = rng.random((65536, 65)).astype(np.float32)
    y[:, 1:] /= y[:, 1:].sum(1, keepdims=True)
    with fd2.Mlp(seed=1) as tr:
        tr.set_data(x, y)
        tr.train(20, batch=4096, lr=1e-3)
        steps = 300
        _, ms = tr.train(steps, batch=4096, lr=1e-3)
    t = 3.0 * MLP_FWD_FLOP * 4096 * steps / (ms * 1e-3) / 1e12
    out["training"] = {"batch": 4096, "steps": steps, "ms": ms, "achieved": t,
                       "frac": t / BF16_PEAK_TFLOPS}
    return out


def surrogate_leg() -> dict:
    """Config 4: surrogate + MC fallback against the full plain MC on 1e6
    histories at 50 keV (H = 500 um, R = 5 um)."""
    from hybridmantis_amd import surrogate_eval

    r = surrogate_eval.evaluate(1.0)
    keep = ("histories", "events", "photons", "mc_histories_per_s", "surrogate_histories_per_s",
            "speedup", "prf_l1_vs_ref", "prf_l1_mc_vs_ref", "prf_l1_bar", "prf_l1_pass",
            "prf_l1_ref_runs", "prf_l1", "prf_l1_9um_vs_ref", "prf_l1_9um_mc_vs_ref",
            "detected_fraction_rel_err", "swank_rel_err", "mean_detected_rel_err")
    out = {k: r[k] for k in keep if k in r}
    rt = r["routing"]
    out["routing"] = {k: rt[k] for k in ("n_events", "n_mc", "n_surrogate", "n_audit",
                                         "ms_total", "ms_mc", "ms_surrogate")}
    out["workload"] = "mono50: 1e6 histories at 50 keV, H=500 um, R=5 um"
    out["prf_l1_note"] = ("L1 at the 1-um pitch against the mean PRF of 4 plain-MC runs "
                          "of the same events; the same-size plain MC scores prf_l1_mc_vs_ref")
    out["mfma"] = mfma_rates()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-workers", type=int, default=0)
    ap.add_argument("--cpu-photons", type=int, default=150000,
                    help="photons per CPU process in the baseline sample")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 4-geometry leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(histories shard across GPUs with one tally all-reduce; not built). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    from hybridmantis_amd import events

    p, ev = events.workload("mono25x50")
    n_photons = int(ev["N"].sum())

    # CPU legs first, before this process touches the GPU
    cpu = None
    if not args.no_cpu:
        workers = args.cpu_workers or host_cores()
        cpu = cpu_baseline(p, ev, workers, args.cpu_photons)
    fpp, per_photon = flops_model(p, ev)

    from hybridmantis_amd import fd2

    tr = fd2.Transport(p)
    buf = tr.device_buffer(ev.nbytes)
    buf.upload(ev)
    for _ in range(args.warmup):
        tr.submit_device(buf, len(ev))
        tr.sync()
    tr.reset()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr.submit_device(buf, len(ev))
        tr.sync()
        kernel_ms.append(tr.last_kernel_ms())
    wall = time.perf_counter() - t0
    counters = tr.counters()
    assert counters["generated"] == n_photons * args.steps
    buf.free()
    tr.close()
    batched = None if args.no_batched else batched_leg(max(1, min(args.steps, 3)))
    surr = None if args.no_surrogate else surrogate_leg()

    avg_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = fpp * n_photons / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": n_photons * args.steps / wall,
        "unit": "optical photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mono25x50: 100,000 histories at 25 keV, pencil beam, "
                               "H=150 um, R=2.55 um, 501x501 PRF at 1 um, "
                               f"{len(ev)} events, {n_photons} optical photons",
                   "histories": p["num_hist"], "events": int(len(ev)),
                   "photons": n_photons, "parallelism": "1 GPU"},
        "roofline": {"bound": "fp32-valu", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic_from_profile("transport_mono25x50_pmc.md"),
                     "kernel_ms": avg_ms, "flops_per_photon": fpp},
        "cpu_baseline": cpu,
        "histories_per_s": p["num_hist"] * args.steps / wall,
        "batched": batched,
        "surrogate": surr,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
