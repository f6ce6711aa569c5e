"""Benchmark of the three-phonon scattering + kappa path on one MI355X.

Workload: the 16-atom FCC mass superlattice (4-atom cubic FCC cell x4 along z, masses
40/80 amu alternating every two atomic layers, Morse IFCs), 10x10x4 q-mesh, 7
temperatures 100-700 K batched in one pass, synthetic inputs generated in-process.

A step is one full scattering pass (tc3_scatter): Phi(q2,q3) build, |Psi|^2 for every
visited irreducible triplet and every (b1,b2,b3), smeared-delta weights for all
temperatures, linewidths and the symmetrised scattering matrix.  `value` is
(visited triplets x (3N)^3) / time per step for the mixed-precision (split-BF16x3)
contraction; the FP64 form, the kappa wall times of both, the kappa error of the
mixed mode against FP64, the kernel roofline and a CPU baseline (the C restatement
of the reference's per-element loop on a bounded sample) are reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x3|fp32|fp64]
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
PEAK_HBM_GBPS = 8000.0
WORKLOAD = "sl16"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_traffic(kernel, triplets):
    """HBM bytes of one launch of `kernel` over `triplets` triplets, from the committed
    counter summary (HBM bytes per triplet of two rocprofv3 --pmc runs, FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "scatter_sl16_bf16x3_counters.md")
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0] == kernel:
                return float(cells[3]) * triplets, os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def kernel_stats(eng):
    import ctypes
    out = {}
    for w, name in ((0, "phi_u_w"), (1, "psi_epilogue"), (2, "xi_symmetrise"), (4, "xi_assembled")):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        b = ctypes.c_int64()
        eng.lib.tc3_get_kernel_stats(eng.ctx, w, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b))
        out[name] = (ms.value, n.value, b.value)
    return out


def scatter_pass(eng, temps, prec, steps, warmup):
    for _ in range(warmup):
        eng.scattering_rates(temps, precision=prec)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.scattering_rates(temps, precision=prec)
    return (time.perf_counter() - t0) / steps


def kappa_wall(eng, temps, prec):
    """harmonic + scattering + RTA + iterative BTE for all temperatures (no file I/O).
    Returns the wall time, kappa (SMA, iterative) and the engine's own stage times."""
    t0 = time.perf_counter()
    eng.dispersions()
    eng.scattering_rates(temps, precision=prec)
    ks, ki, _ = eng.kappa(offdiag=False)
    wall = time.perf_counter() - t0
    tm = eng.timings()
    stages = {k: float(tm[k]) for k in ("harmonic", "scatter", "kappa", "bte")}
    stages["bte_iterations"] = [int(x) for x in eng.niter]
    return wall, ks, ki, stages


def cpu_baseline(indir, grid, temperature, budget_s=15.0):
    from oracle import run as R
    from oracle import ref_port
    pre = R.setup(indir, grid)
    disp = pre[5]
    rows = [(q, b) for q in range(1, pre[4].nirr) for b in range(disp.nb) if disp.om_irr[q, b] > 1e-9]
    # calibrate on one (row, triplet), then size the sample to the budget
    _, t1, nthr, _ = ref_port.run_rows(pre, temperature, rows[:1], max_trip=1)
    n = max(1, int(budget_s * nthr / max(t1, 1e-6)))
    pick = rows[:: max(1, len(rows) // n)][:n] if n < len(rows) else rows
    per = max(1, n // len(pick))
    gam, dt, nthr, nel = ref_port.run_rows(pre, temperature, pick, max_trip=per)
    return dict(value=nel / dt, unit="3ph matrix elements/s", cores=nthr, kind="port",
                sample=f"{len(pick)} rows (q_irr, b1) x first {per} visited triplets each, "
                       f"all (b2,b3), T={temperature:g} K, {dt:.1f} s on {nthr} threads; "
                       "oracle/scatter_ref.c restates the reference's zgeru/dot_product row loop")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="only the timed mixed-precision pass (for profiler runs); prints no JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device-to-device copy (the reference kernel for scaling the HBM "
                         "byte counters); prints no JSON")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {a.numel() * 4 >> 20} MiB")
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: the scattering engine has no multi-GPU path "
              f"(requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    from tdep_amd import inputs
    from tdep_amd.engine import Engine, KAPPA_AU_TO_SI
    cfg = inputs.CONFIGS[WORKLOAD]
    grid, temps = cfg["qgrid"], list(cfg["temperatures"])
    with tempfile.TemporaryDirectory() as d:
        inputs.make_config(WORKLOAD, d)
        log("[bench] setup", WORKLOAD, grid, temps)
        eng = Engine.from_directory(d, grid, device=0)
        eng.dispersions()
        nb = eng.nb
        elements = eng.ntriplets * nb ** 3
        log(f"[bench] nirr={eng.nirr} triplets={eng.ntriplets} nb={nb} elements/step={elements:.3e}")
        # mixed (timed) pass
        t_mx = scatter_pass(eng, temps, args.precision, args.steps, args.warmup)
        ks_mx = kernel_stats(eng)
        tk = eng.timings()
        log(f"[bench] {args.precision}: {t_mx * 1e3:.2f} ms/step  kernels {ks_mx}  probes {eng.tuning()}")
        if args.quick:
            return
        # FP64 GPU form of the same pass
        t_64 = scatter_pass(eng, temps, "fp64", max(1, args.steps // 2), 1)
        ks_64 = kernel_stats(eng)
        log(f"[bench] fp64: {t_64 * 1e3:.2f} ms/step  kernels {ks_64}")
        # kappa wall time, both modes; one untimed run each first (kernel code loading and
        # graph instantiation on first use are process start-up, not part of the solve)
        kappa_wall(eng, temps, args.precision)
        kw_mx, ksma_mx, kit_mx, st_mx = kappa_wall(eng, temps, args.precision)
        kappa_wall(eng, temps, "fp64")
        kw_64, ksma_64, kit_64, st_64 = kappa_wall(eng, temps, "fp64")
        err = float(np.max(np.abs(kit_mx - kit_64)) / np.max(np.abs(kit_64)))
        err_sma = float(np.max(np.abs(ksma_mx - ksma_64)) / np.max(np.abs(ksma_64)))
        log(f"[bench] kappa wall {args.precision} {kw_mx:.3f}s fp64 {kw_64:.3f}s rel err {err:.2e}")
        # roofline of the dominant kernel (algorithmic complex-GEMM FLOPs, 8 per complex MAC)
        # k_uw: W = E2^T U is a dense nb^4 complex GEMM per triplet; U = Phi E3 is formed only
        # on the nonzero (a1, a2, a3) IFC blocks, i.e. nb^4 complex MACs times their fraction
        na = nb // 3
        blk = len(set(zip(eng.fc3["a1"].tolist(), eng.fc3["a2"].tolist(), eng.fc3["a3"].tolist())))
        frac_blk = blk / float(na ** 3)
        fpt_uw = 8.0 * nb ** 4 * (1.0 + frac_blk)
        f_uw = fpt_uw * eng.ntriplets
        f_psi = 8.0 * nb ** 4 * eng.ntriplets       # one nb^4 complex GEMM per triplet
        ms_uw, n_uw, batch = ks_mx["phi_u_w"]
        ms_psi, n_psi, _ = ks_mx["psi_epilogue"]
        if ms_uw >= ms_psi:
            dom, fl, ms, n = "k_uw (Phi build + U=Phi.E3 + W=E2^T.U)", f_uw, ms_uw, n_uw
        else:
            dom, fl, ms, n = "k_psi (Psi=E1^T.W + smeared-delta epilogue)", f_psi, ms_psi, n_psi
        # kernel statistics cover the last tc3_scatter call, i.e. one step.  Algorithmic
        # bytes per launch: W (2 nb^3 reals per triplet, FP32 in the mixed modes) written by
        # k_uw and read once by k_psi, plus what k_psi stores: the compact per-triplet sums
        # (2 nb per (b1, T)) when the rows are assembled afterwards, else the
        # scattering-matrix entries (FP32 in the mixed modes); tables and eigenvectors are
        # negligible.  The roofline reported is the one whose minimum time (algorithmic
        # work at peak) is larger.
        avg_ms = ms / max(1, n)
        es = 8 if args.precision == "fp64" else 4
        w_bytes = 2 * nb ** 3 * es * batch
        assembled = ks_mx["xi_assembled"][1] > 0
        if assembled:
            out_bytes = len(temps) * eng.ntriplets * nb * 2 * nb * es / max(1, n_psi)
        else:
            out_bytes = len(temps) * eng.nrows * eng.nfull * nb * es / max(1, n_psi)
        by = w_bytes if dom.startswith("k_uw") else w_bytes + out_bytes
        fl_launch = fl / max(1, n)
        issue = 3.0 if args.precision == "bf16x3" else 1.0
        peak = PEAK_TFLOPS[args.precision]
        t_mfma = fl_launch * issue / (peak * 1e12)
        t_hbm = by / (PEAK_HBM_GBPS * 1e9)
        mfma_view = dict(achieved=fl_launch / (avg_ms * 1e-3) / 1e12, peak=peak, unit="TFLOP/s",
                         frac=fl_launch / (avg_ms * 1e-3) / 1e12 / peak,
                         flop_per_launch=fl_launch, issued_min_ms_at_peak=t_mfma * 1e3)
        if args.precision == "bf16x3":
            mfma_view["issue_frac"] = 3 * mfma_view["frac"]
        hbm_view = dict(achieved=by / (avg_ms * 1e-3) / 1e9, peak=PEAK_HBM_GBPS, unit="GB/s",
                        frac=by / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, bytes_per_launch=by,
                        min_ms_at_peak=t_hbm * 1e3)
        traffic, tsrc = (None, None)
        if args.precision == "bf16x3":
            traffic, tsrc = profile_traffic("k_uw" if dom.startswith("k_uw") else "k_psi", batch)
        main = hbm_view if t_hbm >= t_mfma else mfma_view
        roof = dict(bound="hbm" if t_hbm >= t_mfma else "mfma", achieved=main["achieved"], peak=main["peak"],
                    unit=main["unit"], frac=main["frac"], traffic=traffic, traffic_source=tsrc, kernel=dom,
                    launches_per_step=n, avg_launch_ms=avg_ms, triplets_per_launch=batch,
                    flop_per_triplet=fpt_uw if dom.startswith("k_uw") else 8.0 * nb ** 4,
                    ifc_block_fraction=frac_blk, rows_assembled=assembled, hbm=hbm_view, mfma=mfma_view)
        if traffic:
            roof["measured_hbm_GBps"] = traffic / (avg_ms * 1e-3) / 1e9
        res = {
            "metric": "3ph matrix elements/s + κ wall-time, mixed vs FP64 GPU @≤1e-4 rel κ err",
            "value": elements / t_mx,
            "unit": "3ph matrix elements/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": t_mx * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": {"bf16x3": "bf16x3", "fp32": "f32", "fp64": "f64"}[args.precision],
            "data": "synthetic (Morse-potential IFCs generated in-process)",
            "config": {"workload": "16-atom FCC mass superlattice (40/80 amu), 10x10x4 q-mesh, "
                                   "7 temperatures 100-700 K, three-phonon, adaptive Gaussian",
                       "qgrid": list(grid), "temperatures": temps, "bands": nb,
                       "irreducible_q": eng.nirr, "visited_triplets": eng.ntriplets,
                       "precision": args.precision},
            "fp64": {"value": elements / t_64, "ms_per_step": t_64 * 1e3},
            "speedup_vs_fp64": t_64 / t_mx,
            "kappa_wall_s": {"mixed": kw_mx, "fp64": kw_64, "speedup": kw_64 / kw_mx,
                             "stages_mixed": st_mx, "stages_fp64": st_64},
            "kappa_rel_err_vs_fp64": {"iterative": err, "sma": err_sma},
            "kappa_xx_W_mK_300K": {"mixed": float(kit_mx[2][0, 0] * KAPPA_AU_TO_SI),
                                   "fp64": float(kit_64[2][0, 0] * KAPPA_AU_TO_SI)},
            "roofline": roof,
            "autotune": {"chosen_batch": batch,
                         "probes_ms": [[b, ms_] for b, ms_ in eng.tuning()]},
        }
        if not args.no_cpu_baseline:
            try:
                res["cpu_baseline"] = cpu_baseline(d, grid, 300.0)
            except Exception as e:  # the baseline is reported, never required
                res["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
