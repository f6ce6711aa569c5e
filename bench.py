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

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}   # dense MFMA peaks of the arithmetic used
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


def kernel_stats(eng, temps, prec):
    """Per-kernel device times of one more pass with direct launches: the timed passes replay
    the contraction from a HIP graph, which records no events between its kernels (the kernels
    and their arguments are the same either way)."""
    import ctypes
    os.environ["TC3_NO_GRAPH"] = "1"
    try:
        eng.scattering_rates(temps, precision=prec)
    finally:
        del os.environ["TC3_NO_GRAPH"]
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
    # the scattering matrix as the BTE product reads it: elements per temperature, bytes each,
    # rows stored per orbit of the little group of their q-point or in full
    n, es, per_orbit = eng.xi_storage()
    stages["xi_elements_per_T"], stages["xi_entry_bytes"], stages["xi_per_orbit"] = n, es, per_orbit
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
    ap.add_argument("--kappa-walls", type=int, default=0, metavar="N",
                    help="only N kappa walls (harmonic step, scattering pass, iterative BTE) in the "
                         "given precision, after one untimed wall (for profiler runs); prints no JSON")
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
        if args.kappa_walls:
            kappa_wall(eng, temps, args.precision)
            for i in range(args.kappa_walls):
                kw, _, _, st = kappa_wall(eng, temps, args.precision)
                log(f"[bench] kappa wall {i} {args.precision}: {kw * 1e3:.2f} ms, stages {st}")
            return
        # mixed (timed) pass
        t_mx = scatter_pass(eng, temps, args.precision, args.steps, args.warmup)
        ks_mx = kernel_stats(eng, temps, args.precision)
        tk = eng.timings()
        log(f"[bench] {args.precision}: {t_mx * 1e3:.2f} ms/step  kernels {ks_mx}  probes {eng.tuning()}")
        if args.quick:
            return
        # FP64 GPU form of the same pass
        t_64 = scatter_pass(eng, temps, "fp64", max(1, args.steps // 2), 1)
        ks_64 = kernel_stats(eng, temps, "fp64")
        log(f"[bench] fp64: {t_64 * 1e3:.2f} ms/step  kernels {ks_64}")
        # kappa wall time, both modes; one untimed run each first (kernel code loading and
        # graph instantiation on first use are process start-up, not part of the solve)
        kappa_wall(eng, temps, args.precision)
        kw_mx, ksma_mx, kit_mx, st_mx = kappa_wall(eng, temps, args.precision)
        lw_mx = np.array([eng.linewidths(i)[0] for i in range(len(temps))])
        kappa_wall(eng, temps, "fp64")
        kw_64, ksma_64, kit_64, st_64 = kappa_wall(eng, temps, "fp64")
        lw_64 = np.array([eng.linewidths(i)[0] for i in range(len(temps))])
        err = float(np.max(np.abs(kit_mx - kit_64)) / np.max(np.abs(kit_64)))
        err_sma = float(np.max(np.abs(ksma_mx - ksma_64)) / np.max(np.abs(ksma_64)))
        big = lw_64 > 1e-3 * lw_64.max(axis=(1, 2), keepdims=True)
        lw_rel = np.abs(lw_mx[big] / lw_64[big] - 1)
        lw_err = {args.precision: dict(max=float(lw_rel.max()), median=float(np.median(lw_rel)))}
        # the other reduced-precision mode's linewidths (one untimed pass)
        other = "fp32" if args.precision == "bf16x3" else "bf16x3"
        if args.precision != "fp64":
            eng.scattering_rates(temps, precision=other)
            lw_o = np.array([eng.linewidths(i)[0] for i in range(len(temps))])
            r_o = np.abs(lw_o[big] / lw_64[big] - 1)
            lw_err[other] = dict(max=float(r_o.max()), median=float(np.median(r_o)))
        lw_err["modes"] = "linewidths above 1e-3 of the largest, all temperatures"
        # |Psi|^2 of 32 triplets spread over the set, formed by the contraction kernels in each
        # precision (tc3_get_psi2), against FP64: normwise per triplet, and per element above
        # 1e-3 of its triplet's largest; elements with a frequency below lo_freqtol left out
        psi_err = {}
        pt = np.linspace(0, eng.ntriplets - 1, 32).astype(np.int64)
        p64, pq = eng.psi2(pt, "fp64")
        om_i, _, _ = eng.harmonic()
        om_f, _, _ = eng.harmonic(full=True)
        ftol = 1e-9   # lo_freqtol, Hartree
        for prec in ("fp32", "bf16x3"):
            pm, _ = eng.psi2(pt, prec)
            nrm, el = [], []
            for i, (a1, a2, a3) in enumerate(pq):
                m = ((om_i[a1] >= ftol)[:, None, None] & (om_f[a2] >= ftol)[None, :, None]
                     & (om_f[a3] >= ftol)[None, None, :])
                x, y = pm[i][m], p64[i][m]
                nrm.append(np.linalg.norm(x - y) / np.linalg.norm(y))
                bg = y > 1e-3 * y.max()
                el.append(np.abs(x[bg] / y[bg] - 1))
            el = np.concatenate(el)
            psi_err[prec] = dict(norm_max=float(max(nrm)), norm_median=float(np.median(nrm)),
                                 element_max=float(el.max()), element_median=float(np.median(el)))
        psi_err["sample"] = "32 triplets spread over the set; element errors over elements above 1e-3 of the triplet's largest"
        del p64
        # end to end: a fresh context from the input files to outfile.thermal_conductivity and
        # the grid file (autotuner probes, graph instantiation and uploads included)
        e2e = {}
        for prec in (args.precision, "fp64"):
            t0 = time.perf_counter()
            e2 = Engine.from_directory(d, grid, device=0)
            e2.dispersions()
            e2.scattering_rates(temps, precision=prec)
            e2.kappa()
            with open(os.path.join(d, "outfile.thermal_conductivity"), "w") as fh:
                fh.write(e2.outfile_text(0))
            e2.write_grid_hdf5(os.path.join(d, "outfile.thermal_conductivity_grid.hdf5"), 0)
            e2e[prec if prec == "fp64" else "mixed"] = time.perf_counter() - t0
            del e2
        log(f"[bench] kappa wall {args.precision} {kw_mx:.3f}s fp64 {kw_64:.3f}s rel err {err:.2e}")
        # Roofline (MFMA-bound contraction): the |Psi|^2 chain is three complex contractions of
        # nb^4 MACs per visited triplet, 24 nb^4 real FLOP algorithmically (8 per complex MAC),
        # whatever the kernels issue: k_uw forms U = Phi E3 and W = E2^T U (16 nb^4), k_psi
        # Psi = E1^T W (8 nb^4) and the smeared-delta epilogue.  Peak: BF16 dense MFMA (the
        # split-BF16x3 products issue three BF16 MFMAs each, so 1/3 is the attainable ceiling).
        # The kernel statistics cover the last tc3_scatter call, i.e. one step.
        ms_uw, n_uw, batch = ks_mx["phi_u_w"]
        ms_psi, n_psi, _ = ks_mx["psi_epilogue"]
        ms_asm, n_asm, _ = ks_mx["xi_assembled"]
        peak = PEAK_TFLOPS[args.precision]
        per_trip = {"k_uw": 16.0 * nb ** 4, "k_psi": 8.0 * nb ** 4}
        kern = {}
        for name, ms, n in (("k_uw", ms_uw, n_uw), ("k_psi", ms_psi, n_psi)):
            avg = ms / max(1, n)
            fl = per_trip[name] * batch
            kern[name] = dict(avg_launch_ms=avg, launches=n, triplets_per_launch=batch, flop_per_launch=fl,
                              achieved_TFLOPs=fl / (avg * 1e-3) / 1e12, frac=fl / (avg * 1e-3) / 1e12 / peak)
        t_con = (ms_uw + ms_psi) * 1e-3
        fl_con = 24.0 * nb ** 4 * eng.ntriplets
        contraction = dict(flop_per_step=fl_con, kernel_ms_per_step=t_con * 1e3,
                           achieved_TFLOPs=fl_con / t_con / 1e12, frac=fl_con / t_con / 1e12 / peak,
                           peak_TFLOPs=peak, unit="TFLOP/s")
        if args.precision == "bf16x3":
            contraction["frac_of_split_bf16x3_ceiling"] = 3 * contraction["frac"]
        dom = "k_psi" if ms_psi >= ms_uw else "k_uw"
        kd = kern[dom]
        # measured HBM traffic per launch (FETCH_SIZE + WRITE_SIZE, rocprofv3 counters): the W
        # intermediate (2 nb^3 reals per triplet) is not algorithmic work; its round trip is
        # the waste against the output (the compact per-triplet sums)
        traffic, tsrc = (None, None)
        if args.precision == "bf16x3":
            traffic, tsrc = profile_traffic(dom, batch)
        es = 8 if args.precision == "fp64" else 4
        out_trip = len(temps) * nb * 2 * nb * es
        w_trip = 2 * nb ** 3 * es
        roof = dict(bound="mfma", achieved=kd["achieved_TFLOPs"], peak=peak, unit="TFLOP/s", frac=kd["frac"],
                    traffic=traffic, traffic_source=tsrc, kernel=dom, avg_launch_ms=kd["avg_launch_ms"],
                    launches_per_step=kd["launches"], triplets_per_launch=batch,
                    flop_per_triplet=per_trip[dom], kernels=kern, contraction=contraction,
                    hbm_waste={"w_bytes_per_triplet": w_trip, "w_round_trip_bytes_per_step": 2 * w_trip * eng.ntriplets,
                               "output_bytes_per_triplet": out_trip})
        if traffic:
            roof["traffic_per_triplet_vs_output"] = traffic / batch / out_trip
        # HBM-bound stages: rows of the scattering matrix assembled from the compact sums
        # (read once, rows written once) and the BTE (every iteration of every temperature
        # streams its scattering matrix and the folded-out F).  The rows are stored as the
        # product reads them: one value per orbit of the little group of their q-point when
        # nothing is added after the orbit average (xi_per_orbit), else nfull nb per row.
        xi_T = st_mx["xi_elements_per_T"] * st_mx["xi_entry_bytes"]
        xi_full_T = eng.nrows * eng.nfull * nb * es
        if n_asm:
            by = len(temps) * (eng.ntriplets * nb * 2 * nb * es + xi_T)
            roof["k_xi_asm"] = dict(bytes=by, ms=ms_asm, GBps=by / (ms_asm * 1e-3) / 1e9,
                                    frac=by / (ms_asm * 1e-3) / 1e9 / PEAK_HBM_GBPS)
        it_mx = st_mx["bte_iterations"]
        by_bte = sum(it_mx) * (xi_T + 3 * eng.nfull * nb * 8)
        roof["bte"] = dict(bytes=by_bte, ms=st_mx["bte"] * 1e3, GBps=by_bte / st_mx["bte"] / 1e9,
                           frac=by_bte / st_mx["bte"] / 1e9 / PEAK_HBM_GBPS,
                           temperature_iterations=int(sum(it_mx)),
                           xi_bytes_per_T=xi_T, xi_full_bytes_per_T=xi_full_T,
                           xi_per_orbit=st_mx["xi_per_orbit"],
                           note="whole BTE (orbit sums of F, Xi.F, update, kappa log, checks) against the "
                                "Xi bytes as stored + F bytes")
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
            "linewidth_rel_err_vs_fp64": lw_err,
            "psi2_rel_err_vs_fp64": psi_err,
            "end_to_end_s": e2e,
            "kappa_xx_W_mK_300K": {"mixed": float(kit_mx[2][0, 0] * KAPPA_AU_TO_SI),
                                   "fp64": float(kit_64[2][0, 0] * KAPPA_AU_TO_SI)},
            "roofline": roof,
            "autotune": {"chosen_batch": batch,
                         "probes_ms": [[b, ms_] for b, ms_ in eng.tuning()[:4]],
                         "whole_triplet_k_psi_probe_ms": [[b, ms_] for b, ms_ in eng.tuning()[4:5]]},
        }
        if not args.no_cpu_baseline:
            try:
                res["cpu_baseline"] = cpu_baseline(d, grid, 300.0)
            except Exception as e:  # the baseline is reported, never required
                res["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
