"""Throughput of the tomographic 3x2pt projection path on one MI355X.

Workload: the full real-space 3x2pt layout (10 Smail source bins z0=0.6, 5
Gaussian lens bins, 55 ss + 50 gs + 5 gg pairs, 26 theta bins 2.5'-900',
N_ell = 500, LMAX = 50000, NLA IA, linear bias 1.2-2.0), over a batch of
draws that each carry their own cosmology (Omega_m, sigma8, h drawn
uniformly) and hence their own chi(z), G(z) and 400x100 P(k,z) tables.  One
step = the whole path (lens efficiencies, Limber C_ell, Legendre transform,
data-vector assembly) for every draw of the batch, inputs resident in HBM.
--precision picks the plain FP64 path or a mixed-precision one (FP32 Limber
with an FP32 or 3-term split-BF16 MFMA Legendre contraction); for the mixed
ones the per-probe error against the plain FP64 GPU path on the same batch is
reported next to the throughput.

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
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _build_one(item):
    from cosmolike_amd import synth
    kind, spec = item
    if kind == "full":
        return synth.full_3x2pt_problem(spec)
    if kind == "nl":
        return synth.nonlimber_problem(spec)
    cs, ns, nl, nt = spec
    return synth.ragged_3x2pt_problem(cs, ns, nl, nt)


def _build(items):
    """Synthetic draws, built by a pool of worker processes (forked before
    anything touches the GPU)."""
    workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or len(items) < 16:
        return [_build_one(x) for x in items]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_build_one, items, chunksize=4)


def make_draws(B, seed):
    from cosmolike_amd import synth
    return _build([("full", cos) for cos, *_ in synth.random_draws(B, seed=seed, ragged=False)])


def make_nonlimber_draws(B, seed):
    from cosmolike_amd import synth
    return _build([("nl", cos) for cos, *_ in synth.random_draws(B, seed=seed, ragged=False)])


def make_ragged_draws(n, seed):
    from cosmolike_amd import synth
    return _build([("ragged", tuple(x)) for x in synth.random_draws(n, seed=seed, ragged=True)])


def algorithmic_work(p):
    """Per-draw algorithmic work of each kernel (FLOP): 20 per pair-node Limber
    integrand evaluation plus 40 per P(k,z) lookup, 2 per multiply-add of the
    Legendre contraction."""
    nss, ngs, ngg = p.ns * (p.ns + 1) // 2, p.ns * p.nl, p.nl
    rows = p.n_ell + 19
    return {
        # 20 FLOP per pair-node integrand term + 40 FLOP per P(k,z) lookup
        "k_limber_sx": 20.0 * rows * 64 * (nss + ngs) + 40.0 * rows * 64 * (1 + p.nl),
        "k_limber_gg": 20.0 * rows * 96 * ngg + 40.0 * rows * 96 * p.nl,
        "k_legendre": 2.0 * p.ntheta * rows * (2 * nss + ngs + ngg),
        "k_legendre_full": 2.0 * p.ntheta * (p.lmax - 1) * (2 * nss + ngs + ngg),
    }


def kernel_work_and_peak(name, work, precision):
    """(algorithmic FLOP per draw, peak TFLOP/s) of a kernel by its name."""
    base = name[:-4] if name.endswith("_f32") else name
    if name == "k_legendre_mx":
        base = "k_legendre"
    flop = work.get(base, 0.0)
    if name.endswith("_f32"):
        return flop, FP32_PEAK_TFLOPS
    if name == "k_legendre_mx":
        return flop, (BF16_PEAK_TFLOPS if precision == "bf16x3" else FP32_PEAK_TFLOPS)
    return flop, FP64_PEAK_TFLOPS


def probe_errors(dv, ref, ns, nl, nt):
    """Max and RMS of |dv - ref| / max(|ref|, 1e-3 max|ref| of the row) per probe."""
    n_ss, n_gs = ns * (ns + 1) // 2, ns * nl
    out, o = {}, 0
    for name, rows in (("xi+", n_ss), ("xi-", n_ss), ("gammat", n_gs), ("w", nl)):
        g = dv[:, o:o + rows * nt].reshape(dv.shape[0], rows, nt)
        r = ref[:, o:o + rows * nt].reshape(dv.shape[0], rows, nt)
        floor = 1e-3 * np.abs(r).max(axis=2, keepdims=True)
        e = np.abs(g - r) / np.maximum(np.abs(r), floor)
        out[name] = {"max": float(e.max()), "rms": float(np.sqrt((e ** 2).mean()))}
        o += rows * nt
    return out


def measured_traffic(kernel, batch, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE scaled by a plain device copy), or None."""
    path = os.path.join(ROOT, "profiles", f"projection_3x2pt_b{batch}_{precision}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def time_pairs(c, dv, dvbar, p0, B, n, ckpt, steps, fwd_s):
    """Forward + VJP pairs over the uploaded batch (gradients of ln P, the
    photo-z shifts, b1, A_IA, m); ckpt: the forward keeps the Limber
    integrands for the VJP instead of the VJP recomputing them.  fwd_s: the
    forward's own time per step, subtracted for the VJP rate."""
    import torch
    ns, nl = p0.ns, p0.nl
    nknz = p0.lnP.size
    g = {k: torch.empty(B * m, dtype=torch.float64, device="cuda:0")
         for k, m in (("lnP", nknz), ("src", ns), ("lens", nl), ("b1", nl), ("A", 1), ("m", ns))}
    ptrs = [g[k].data_ptr() for k in ("lnP", "src", "lens", "b1", "A", "m")]
    c.set_vjp_checkpoint(ckpt)
    tiles = c.autotune()  # forward and adjoint stages tuned for pairs
    for _ in range(2):
        c.batch_forward(dv.data_ptr())
        c.batch_vjp(dvbar.data_ptr(), *ptrs)
    c.synchronize()
    nv = max(3, steps // 2)
    t1 = time.perf_counter()
    for _ in range(nv):
        c.batch_forward(dv.data_ptr())
        c.batch_vjp(dvbar.data_ptr(), *ptrs)
    c.synchronize()
    t_pair = (time.perf_counter() - t1) / nv
    c.enable_timing(True)
    ms = {}
    for _ in range(3):
        c.batch_forward(dv.data_ptr())
        c.batch_vjp(dvbar.data_ptr(), *ptrs)
        c.synchronize()
        for k, v in c.kernel_times().items():
            if k.endswith("_adj") or k in ("k_lnp_gather", "lens_limits"):
                ms[k] = ms.get(k, 0.0) + v / 3
    c.enable_timing(False)
    c.set_vjp_checkpoint(False)
    return {"fwd_plus_vjp_value": B / t_pair, "vjp_value": B / max(t_pair - fwd_s, 1e-9),
            "vjp_kernel_ms": ms, "tile_rows": tiles}


def speedups(value, vjp, mixed, plain):
    """Throughput of each batched leg over the plain FP64 single-draw path."""
    out = {"forward": value / plain}
    for k, v in (vjp or {}).items():
        out[f"forward_plus_vjp_{k}"] = v["fwd_plus_vjp_value"] / plain
    for k, v in (mixed or {}).items():
        out[f"{k}_forward"] = v["value"] / plain
        if "fwd_plus_vjp_value" in v:
            out[f"{k}_forward_plus_vjp"] = v["fwd_plus_vjp_value"] / plain
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


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
            "cpu_model": cpu_model(),
            "sample": f"{n} draws of the same 3x2pt layout (FP64 oracle, OpenMP), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--mode", choices=["folded", "full"], default="folded")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16x3"], default="fp64")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--vjp", type=int, default=1, help="also time the batched VJP")
    ap.add_argument("--mixed", type=int, default=1,
                    help="also time the FP32 and BF16X3 forward and forward + VJP (with --precision fp64)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="also copy 512 MiB device to device once (HBM-counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--ragged", type=int, default=512,
                    help="draws of the multi-layout leg (3-10 src x 2-6 lens bins, 20-30 theta bins; 0 = skip)")
    ap.add_argument("--nonlimber", type=int, default=128,
                    help="draws of the non-Limber w leg (adopt_limber_gg = 0; 0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the draws of a batch "
              "are independent, so N GPUs would each take a share of the batch with no exchange.",
              file=sys.stderr)
        sys.exit(2)

    # synthetic inputs first: the builder forks worker processes, which must
    # happen before the HIP runtime is loaded
    t0 = time.perf_counter()
    draws = make_draws(args.batch, args.seed)
    probs = make_ragged_draws(args.ragged, args.seed + 1) if args.ragged > 0 else []
    nlprobs = make_nonlimber_draws(args.nonlimber, args.seed + 2) if args.nonlimber > 0 else []
    log(f"bench: built {len(draws)} + {len(probs)} + {len(nlprobs)} draws in {time.perf_counter() - t0:.1f}s")

    import torch
    from cosmolike_amd import Cosmolike, LEGENDRE_FOLDED, LEGENDRE_FULL
    mode = LEGENDRE_FULL if args.mode == "full" else LEGENDRE_FOLDED
    c = Cosmolike(device=0).load(draws[0], mode)
    c.batch_upload(draws)
    n = c.datavector_size()
    dv = torch.empty(args.batch * n, dtype=torch.float64, device="cuda:0")
    tile_rows = c.autotune()
    log(f"bench: autotuned rows per tile (fwd ss/gs, fwd gg, adj ss/gs, adj gg) {tile_rows}, "
        f"gg tables in LDS (fwd, adj) {c.gg_lds}")
    accuracy = None
    if args.precision != "fp64":
        # error of the mixed path against the plain FP64 GPU path on this batch
        ref = torch.empty_like(dv)
        c.batch_forward(ref.data_ptr())
        c.set_precision(args.precision)
        c.batch_forward(dv.data_ptr())
        c.synchronize()
        torch.cuda.synchronize()
        p0 = draws[0]
        accuracy = probe_errors(dv.view(args.batch, n).cpu().numpy(), ref.view(args.batch, n).cpu().numpy(),
                                p0.ns, p0.nl, p0.ntheta)
        log(f"bench: {args.precision} vs FP64 path {accuracy}")
        del ref
    for _ in range(args.warmup):
        c.batch_forward(dv.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    log("bench: warmup done")
    if args.traffic_probe:
        src = torch.ones(1 << 26, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: traffic probe copied {src.numel() * 8} bytes device to device")
        del src, dst
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
            if not k.endswith("_adj") and k not in ("k_lnp_gather", "lens_limits"):
                acc[k] = acc.get(k, 0.0) + v / nt
    c.enable_timing(False)

    # forward + VJP: one autograd step per draw (forward, then the VJP of the
    # whole data vector against a cotangent), timed as pairs; the VJP rate is
    # the pair time minus the forward time.  The VJP reuses the forward's
    # device state (node weights, lens efficiencies, C_ell rows).  Two
    # adjoints: recomputing the Limber integrands (the default), or reading a
    # checkpoint the forward stored (clh_set_vjp_checkpoint).
    vjp = {}
    if args.vjp:
        B = args.batch
        dvbar = torch.randn(B * n, dtype=torch.float64, device="cuda:0")
        for name, ck in (("recompute", False), ("checkpoint", True)):
            vjp[name] = time_pairs(c, dv, dvbar, draws[0], B, n, ck, args.steps, el / args.steps)
            log(f"bench: forward+VJP ({name}) {vjp[name]}")
        c.set_tile_rows(tile_rows)

    # the split-precision forward and forward + VJP on the same batch, with
    # each probe's error against the plain FP64 GPU path
    mixed = {}
    if args.mixed and args.precision == "fp64":
        ref = dv.view(args.batch, n).cpu().numpy()
        p0 = draws[0]
        for prec in ("fp32", "bf16x3"):
            c.set_precision(prec)
            mt = c.autotune()
            mdv = torch.empty_like(dv)
            for _ in range(args.warmup):
                c.batch_forward(mdv.data_ptr())
            c.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                c.batch_forward(mdv.data_ptr())
            c.synchronize()
            elm = time.perf_counter() - t1
            acc_m = probe_errors(mdv.view(args.batch, n).cpu().numpy(), ref, p0.ns, p0.nl, p0.ntheta)
            rec_m = {"value": args.batch * args.steps / elm, "unit": "data vectors/s", "tile_rows": mt,
                     "accuracy_vs_fp64_path": acc_m,
                     "max_rel_err": max(v["max"] for v in acc_m.values())}
            if args.vjp:
                pr = time_pairs(c, mdv, dvbar, p0, args.batch, n, False, args.steps, elm / args.steps)
                rec_m["fwd_plus_vjp_value"] = pr["fwd_plus_vjp_value"]
                rec_m["vjp_value"] = pr["vjp_value"]
            mixed[prec] = rec_m
            log(f"bench: {prec} {rec_m}")
            del mdv
        c.set_precision("fp64")
        c.set_tile_rows(tile_rows)

    work = algorithmic_work(draws[0])
    dom = max(acc, key=acc.get)
    flop1, peak = kernel_work_and_peak(dom, work, args.precision)
    flop = flop1 * args.batch
    achieved = flop / (acc[dom] * 1e-3) / 1e12 if acc[dom] > 0 else 0.0
    # the Limber kernels compute on the FP64 (FP32) vector ALUs: their roofline
    # is the vector peak of that type (on MI355X equal to the dense MFMA one)
    roofline = {"bound": "valu" if "legendre" not in dom else "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak,
                "units": "vector ALU" if "legendre" not in dom else "MFMA",
                "traffic": measured_traffic(dom, args.batch, args.precision), "kernel": dom,
                "kernel_ms": acc[dom], "kernel_ms_all": acc,
                "vjp_kernel_ms": vjp["recompute"]["vjp_kernel_ms"] if vjp else None}
    log(f"bench: kernel ms {acc}")

    # sanity: the last batch agrees with itself (finite values)
    out = dv.view(args.batch, n)
    assert torch.isfinite(out).all().item()

    # multi-layout batch: every draw with its own cosmology and its own survey
    # layout, all in one captured graph of per-stage launches
    ragged = None
    if args.ragged > 0:
        t1 = time.perf_counter()
        cr = Cosmolike(device=0)
        cache, lays = {}, []
        for pr in probs:
            key = (pr.ns, pr.nl, pr.ntheta)
            if key not in cache:
                cr.load(pr, mode)
                cache[key] = cr.capture_layout()
            lays.append(cache[key])
        cr.set_precision(args.precision)
        cr.batch_upload_layouts(probs, lays)
        rtile = cr.autotune()
        ntot = cr.batch_offsets()[-1]
        log(f"bench: ragged batch of {len(probs)} draws, {len(cache)} layouts, {ntot} outputs, "
            f"set up in {time.perf_counter() - t1:.1f}s")
        rdv = torch.empty(ntot, dtype=torch.float64, device="cuda:0")
        cr.batch_use_graph(True)
        for _ in range(args.warmup):
            cr.batch_forward(rdv.data_ptr())
        cr.synchronize()
        torch.cuda.synchronize()
        nr = max(5, args.steps // 4)
        t1 = time.perf_counter()
        for _ in range(nr):
            cr.batch_forward(rdv.data_ptr())
        cr.synchronize()
        torch.cuda.synchronize()
        elr = time.perf_counter() - t1
        assert torch.isfinite(rdv).all().item()
        ragged = {"value": len(probs) * nr / elr, "unit": "data vectors/s", "draws": len(probs),
                  "layouts": len(cache), "ms_per_step": 1e3 * elr / nr, "graph": True, "tile_rows": rtile,
                  "precision": args.precision}
        log(f"bench: ragged {ragged['value']:.1f} dv/s")
        del cr, rdv

    # non-Limber w (the reference's default adopt_limber_gg = 0): FFTLog C_gg
    # below ell 150 for every lens bin of every draw, FP64, graph replay
    nonlimber = None
    if args.nonlimber > 0:
        t1 = time.perf_counter()
        cn = Cosmolike(device=0).load(nlprobs[0], mode)
        cn.batch_upload(nlprobs)
        ndv = torch.empty(len(nlprobs) * n, dtype=torch.float64, device="cuda:0")
        cn.batch_use_graph(True)
        for _ in range(2):
            cn.batch_forward(ndv.data_ptr())
        cn.synchronize()
        torch.cuda.synchronize()
        log(f"bench: non-Limber batch of {len(nlprobs)} draws set up in {time.perf_counter() - t1:.1f}s")
        nn = max(3, args.steps // 4)
        t1 = time.perf_counter()
        for _ in range(nn):
            cn.batch_forward(ndv.data_ptr())
        cn.synchronize()
        torch.cuda.synchronize()
        eln = time.perf_counter() - t1
        assert torch.isfinite(ndv).all().item()
        nonlimber = {"value": len(nlprobs) * nn / eln, "unit": "data vectors/s", "draws": len(nlprobs),
                     "ms_per_step": 1e3 * eln / nn, "precision": "fp64",
                     "w": "non-Limber C_gg below ell 150 (C_cl_tomo FFTLog, tolerance 0.01), Limber above"}
        log(f"bench: non-Limber {nonlimber['value']:.1f} dv/s")
        del cn, ndv

    # the plain FP64 drop-in path: one draw at a time through the setters and
    # clh_compute_3x2pt_real (host buffers in and out), as the reference's
    # interface calls it
    c.set_precision("fp64")
    cp = Cosmolike(device=0).load(draws[0], mode)
    cp.compute_3x2pt_real()
    npl, t1 = 0, time.perf_counter()
    while True:
        cp.set_draw(draws[npl % len(draws)])
        cp.compute_3x2pt_real()
        npl += 1
        if time.perf_counter() - t1 > 2.0 and npl >= 10:
            break
    plain = {"value": npl / (time.perf_counter() - t1), "unit": "data vectors/s",
             "path": "per draw: cosmology/nuisance setters + clh_compute_3x2pt_real, FP64, host buffers"}
    log(f"bench: plain FP64 single-draw path {plain['value']:.1f} dv/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(draws, args.cpu_seconds)
        log(f"bench: cpu baseline {cpu}")

    rec = {
        "metric": "3x2pt data vectors/s + VJPs/s at <=1e-4 rel err vs FP64 plain path, 1 GPU",
        "value": value, "unit": "data vectors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "f32+bf16x3"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"full 3x2pt real space, {args.precision} (10 src + 5 lens bins, 55 ss + 50 gs "
                               "+ 5 gg pairs, 26 theta 2.5'-900', N_ell=500, LMAX=50000), batch of "
                               f"{args.batch} draws with their own cosmology",
                   "batch": args.batch, "legendre": args.mode, "precision": args.precision,
                   "tile_rows": tile_rows,
                   "datavector_len": n},
        "accuracy_vs_fp64_path": accuracy,
        "vjp_value": vjp["recompute"]["vjp_value"] if vjp else None,
        "vjp_unit": "VJPs/s (full data vector -> lnP, dz_src, dz_lens, b1, A_IA, m; FP64 adjoint kernels "
                    "recomputing the Limber integrands, timed as forward+VJP pairs minus the forward)",
        "fwd_plus_vjp_value": vjp["recompute"]["fwd_plus_vjp_value"] if vjp else None,
        "vjp_legs": vjp,
        "mixed_precision": mixed,
        "plain_fp64_single_draw": plain,
        "ragged_batch": ragged,
        "nonlimber_batch": nonlimber,
        "speedup_vs_plain_path": speedups(value, vjp, mixed, plain["value"]),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
