"""Throughput of the tomographic 3x2pt projection path on one MI355X.

Workload: the full real-space 3x2pt layout (10 Smail source bins z0=0.6, 5
Gaussian lens bins, 55 ss + 50 gs + 5 gg pairs, 26 theta bins 2.5'-900',
N_ell = 500, LMAX = 50000, NLA IA, linear bias 1.2-2.0), plain FP64, over a
batch of draws that each carry their own cosmology (Omega_m, sigma8, h drawn
uniformly) and hence their own chi(z), G(z) and 400x100 P(k,z) tables.  One
step = the whole path (lens efficiencies, Limber C_ell, Legendre transform,
data-vector assembly) for every draw of the batch, inputs resident in HBM.

Prints one JSON line (see README of the driver contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cocoa-cosmolike-core_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix rates coincide)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_draws(B, seed):
    from cosmolike_amd import synth
    draws = []
    for cos, *_ in synth.random_draws(B, seed=seed, ragged=False):
        draws.append(synth.full_3x2pt_problem(cos))
    return draws


def algorithmic_work(p):
    """Per-draw algorithmic work of each kernel (FLOP), as defined in DESIGN.md."""
    nss, ngs, ngg = p.ns * (p.ns + 1) // 2, p.ns * p.nl, p.nl
    rows = p.n_ell + 19
    return {
        # 20 FLOP per pair-node integrand term + 40 FLOP per P(k,z) lookup
        "k_limber_sx": 20.0 * rows * 64 * (nss + ngs) + 40.0 * rows * 64 * (1 + p.nl),
        "k_limber_gg": 20.0 * rows * 96 * ngg + 40.0 * rows * 96 * p.nl,
        "k_legendre": 2.0 * p.ntheta * rows * (2 * nss + ngs + ngg),
        "k_legendre_full": 2.0 * p.ntheta * (p.lmax - 1) * (2 * nss + ngs + ngg),
    }


def cpu_baseline(draws, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the FP64 CPU restatement, timed as the CPU reference path
    o = oracle.Oracle().load(draws[0])
    o.datavector()  # warm: Legendre kernels are cosmology independent
    n = 0
    t0 = time.perf_counter()
    while True:
        d = draws[n % len(draws)]
        o.set_cosmology(d.omega_m, d.chi_z, d.chi, d.G_z, d.G, d.log10k, d.pk_z, d.lnP)
        o.touch()
        o.datavector()
        n += 1
        el = time.perf_counter() - t0
        if (el > seconds and n >= 3) or n >= 2000:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n / el, "unit": "data vectors/s", "cores": threads, "kind": "port",
            "sample": f"{n} draws of the same 3x2pt layout (FP64 oracle, OpenMP), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mode", choices=["folded", "full"], default="folded")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vjp", type=int, default=1, help="also time the batched VJP")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the draws of a batch "
              "are independent, so N GPUs would each take a share of the batch with no exchange.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from cosmolike_amd import Cosmolike, LEGENDRE_FOLDED, LEGENDRE_FULL

    t0 = time.perf_counter()
    draws = make_draws(args.batch, args.seed)
    log(f"bench: built {len(draws)} draws in {time.perf_counter() - t0:.1f}s")
    mode = LEGENDRE_FULL if args.mode == "full" else LEGENDRE_FOLDED
    c = Cosmolike(device=0).load(draws[0], mode)
    c.batch_upload(draws)
    n = c.datavector_size()
    dv = torch.empty(args.batch * n, dtype=torch.float64, device="cuda:0")
    for _ in range(args.warmup):
        c.batch_forward(dv.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    log("bench: warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c.batch_forward(dv.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = args.batch * args.steps / el
    log(f"bench: {args.steps} steps in {el:.3f}s -> {value:.1f} dv/s")

    # per-kernel durations with HIP events on the library's stream
    c.enable_timing(True)
    acc = {}
    nt = max(3, min(args.steps, 10))
    for _ in range(nt):
        c.batch_forward(dv.data_ptr())
        c.synchronize()
        for k, v in c.kernel_times().items():
            acc[k] = acc.get(k, 0.0) + v / nt
    c.enable_timing(False)

    # VJP throughput: one VJP of every draw's data vector (includes its forward recompute)
    vjp_value = None
    if args.vjp:
        B = args.batch
        ns, nl = draws[0].ns, draws[0].nl
        nknz = draws[0].lnP.size
        dvbar = torch.randn(B * n, dtype=torch.float64, device="cuda:0")
        g = {k: torch.empty(B * m, dtype=torch.float64, device="cuda:0")
             for k, m in (("lnP", nknz), ("src", ns), ("lens", nl), ("b1", nl), ("A", 1), ("m", ns))}
        ptrs = [g[k].data_ptr() for k in ("lnP", "src", "lens", "b1", "A", "m")]
        for _ in range(2):
            c.batch_vjp(dvbar.data_ptr(), *ptrs)
        c.synchronize()
        t1 = time.perf_counter()
        nv = max(3, args.steps // 2)
        for _ in range(nv):
            c.batch_vjp(dvbar.data_ptr(), *ptrs)
        c.synchronize()
        vjp_value = B * nv / (time.perf_counter() - t1)
        log(f"bench: VJP {vjp_value:.1f} per s")
    work = algorithmic_work(draws[0])
    if args.mode == "full":
        acc["k_legendre_full"] = acc.pop("k_legendre")
    dom = max(acc, key=acc.get)
    flop = work.get(dom, 0.0) * args.batch
    achieved = flop / (acc[dom] * 1e-3) / 1e12 if acc[dom] > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom,
                "kernel_ms": acc[dom], "kernel_ms_all": acc}
    log(f"bench: kernel ms {acc}")

    # sanity: the last batch agrees with itself (finite values)
    out = dv.view(args.batch, n)
    assert torch.isfinite(out).all().item()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(draws, args.cpu_seconds)
        log(f"bench: cpu baseline {cpu}")

    rec = {
        "metric": "3x2pt data vectors/s + VJPs/s at <=1e-4 rel err vs FP64 plain path, 1 GPU",
        "value": value, "unit": "data vectors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "full 3x2pt real space, plain FP64 (10 src + 5 lens bins, 55 ss + 50 gs "
                               "+ 5 gg pairs, 26 theta 2.5'-900', N_ell=500, LMAX=50000), batch of "
                               f"{args.batch} draws with their own cosmology",
                   "batch": args.batch, "legendre": args.mode, "datavector_len": n},
        "vjp_value": vjp_value, "vjp_unit": "VJPs/s (full data vector -> lnP, dz_src, dz_lens, b1, A_IA, m)",
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
