"""Benchmark of the MI355X pressure projection (one JSON line on stdout).

Workload (default): 32768 blocks of 24^3 cells in a 32x32x32 periodic block
layout (768^3 = 452,984,832 cells), Gaussian-random staggered velocity with a
k^-5/3 energy spectrum at rms 1e-3 (seed 4, generated on the GPU), SIP
pressure solve.  A step is one outer (defect-correction) iteration of the
projection: `ninner` SIP inner iterations over every cell plus the FP64
residual/RHS/pressure update.  `value` = interior cells x inner iterations /
wall time of the K timed steps (GCell-updates/s) for the FP32-refined form;
the all-FP64 form is timed the same way and reported beside it, and so is
the FP32-refined form with one copy of the SIP factors per block
(`fp32-unshared`; the library shares one copy between blocks whose factors
are equal bit for bit, every block here).

The line also carries:
  roofline     -- one SIP inner iteration (forward + backward sweep, timed
                  with HIP events on the library's stream) priced at the
                  fused algorithmic bytes per cell-update (45 B FP32: DP,
                  RHS, AP, seven factors, mask read, DP written), against
                  the 8 TB/s peak and a measured device copy (torch copy_
                  of 2 x 1.8 GB); `step` prices the whole timed step
                  (ninner x 45 + 49 B per cell) over ms_per_step;
                  `traffic` is the HBM bytes per inner iteration from the
                  committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, and
                  `per_kernel` gives each sweep at its own bytes (factors
                  shared between blocks come from L2, so HBM carries less
                  than the algorithmic bytes);
  autotune     -- the pipeline shape the library's autotuner picked for the
                  plane kernels, with its per-shape timings;
  cpu_baseline -- the oracle (C restatement of the reference CPU path) timed
                  on this host on a bounded sample of the same workload;
  tts          -- time to a 1e-10 relative residual (FP32-refined and FP64)
                  for the configuration named in `tts.config`;
  rk_stage     -- (with --stage) the momentum right-hand side (tstle4) and
                  the 2N-storage update (rkstep) on the same fields, host
                  wall time per call over synchronised repetitions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "mglet-base_amd")]

# algorithmic bytes per cell per launch (all-fluid blocks read no mask)
#   forward sweep: HILF, RHS, AP, LW, LS, LB, LPR reads, RES write   (8)
#   backward sweep: RES, UE, UN, UT, HILF reads, HILF write          (6)
ALG_BYTES = {"fp32": {"forward": 32, "backward": 24},
             "fp64": {"forward": 64, "backward": 48}}
# fused algorithmic bytes per cell-update of one inner iteration (DP, RHS,
# AP and the seven SIP factors read at 4 or 8 B, the int8 mask read, DP
# written) and of the FP64 outer update (DP, HILF, RHS, AP, mask read; DP,
# RHS written), whatever the kernels move on chip
FUSED_BYTES = {"fp32": 45, "fp64": 89}
OUTER_BYTES = 49
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_velocity(g, nb, n, seed, rms=1e-3, spectrum=True):
    """Staggered velocity in the reference layout as CUDA tensors."""
    import torch
    N = [nb[a] * n[a] for a in range(3)]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    comps = []
    for c in range(3):
        if spectrum:
            kx = torch.fft.fftfreq(N[0], d=1.0 / N[0], device="cuda")
            ky = torch.fft.fftfreq(N[1], d=1.0 / N[1], device="cuda")
            kz = torch.fft.rfftfreq(N[2], d=1.0 / N[2], device="cuda")
            k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
            amp = torch.where(k2 > 0, k2.clamp(min=1.0) ** (-11.0 / 12.0), torch.zeros_like(k2))
            del k2
            re = torch.randn(amp.shape, generator=gen, device="cuda", dtype=torch.float64)
            im = torch.randn(amp.shape, generator=gen, device="cuda", dtype=torch.float64)
            spec = torch.complex(re * amp, im * amp)
            del re, im, amp
            f = torch.fft.irfftn(spec, s=N)
            del spec
        else:
            f = torch.rand(N, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        comps.append(f)
    scale = rms / torch.sqrt(sum((f * f).mean() for f in comps) / 3)
    out = []
    for f in comps:
        f = f * scale
        blk = (f.reshape(nb[0], n[0], nb[1], n[1], nb[2], n[2])
               .permute(4, 2, 0, 1, 3, 5).reshape(-1, n[0], n[1], n[2]))
        full = torch.zeros((blk.shape[0], n[0] + 4, n[1] + 4, n[2] + 4), dtype=torch.float64,
                           device="cuda")
        full[:, 2:-2, 2:-2, 2:-2] = blk
        out.append(full.reshape(-1))
        del f, blk
    torch.cuda.empty_cache()
    return out


def run_steps(solver, fields0, steps, ninner):
    """mgpoisl with nouter = steps on fresh copies; returns (wall s, stats)."""
    import torch
    u, v, w = (f.clone() for f in fields0)
    p = torch.zeros_like(u)
    solver.set_iterations(ninner=ninner, nouter=steps, nouter_min=steps + 1, eps_rel=0.0,
                          check_every=max(1, steps))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solver.mgpoisl(u, v, w, p, 1e-3)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    del u, v, w, p
    return t1 - t0, st


# stored positions per launch move U, V, W, P, G in and UO, VO, WO out (tstle4);
# P, DP, RHS in and P, DP out (rkstep)
STAGE_BYTES = {"tstle4": 64, "rkstep": 40}


def rk_stage_timing(solver, fields0, reps=5):
    """Mean wall ms per call of tstle4 and rkstep on the workload's fields
    (viscosity 1e-2, pressure 0); each call synchronises before returning."""
    import torch
    u, v, w = fields0
    p = torch.zeros_like(u)
    g = torch.full_like(u, 1e-2)
    out = {}
    solver.tstle4(u, v, w, p, g, 1e-2)          # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        uo, vo, wo = solver.tstle4(u, v, w, p, g, 1e-2)
        del uo, vo, wo
    torch.cuda.synchronize()
    out["tstle4_ms"] = (time.perf_counter() - t0) / reps * 1e3
    del g
    q, dq, rhs = p, torch.zeros_like(u), torch.ones_like(u)
    solver.rkstep(q, dq, rhs, -5.0 / 9.0, 1e-3)
    t0 = time.perf_counter()
    for _ in range(reps):
        solver.rkstep(q, dq, rhs, -5.0 / 9.0, 1e-3)
    torch.cuda.synchronize()
    out["rkstep_ms"] = (time.perf_counter() - t0) / reps * 1e3
    n = u.numel()
    for k in ("tstle4", "rkstep"):
        out[f"{k}_gbs"] = round(STAGE_BYTES[k] * n / (out[f"{k}_ms"] * 1e-3) / 1e9, 1)
        out[f"{k}_ms"] = round(out[f"{k}_ms"], 3)
    out["stored_positions"] = n
    out["alg_bytes_per_position"] = STAGE_BYTES
    return out


def time_to_solution(solver, fields0, ninner, eps, cap):
    import torch
    u, v, w = (f.clone() for f in fields0)
    p = torch.zeros_like(u)
    solver.set_iterations(ninner=ninner, nouter=cap, nouter_min=0, eps_rel=eps, check_every=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solver.mgpoisl(u, v, w, p, 1e-3)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, st, p


def cpu_baseline(target_s=12.0):
    """Oracle SIP inner iterations (hyperplane order) on a 4x4x4-block
    sample of 24^3 blocks, all host threads; GCell-updates/s."""
    import numpy as np
    import oracle
    from mgp.grids import uniform_box
    g = uniform_box((4, 4, 4), (24, 24, 24))
    o = oracle.Oracle(g)
    st = o.setup()
    rng = np.random.default_rng(4)
    rhs = np.where(g.interior_mask(), rng.standard_normal(g.n3d), 0.0)
    threads = oracle.cpu_count()
    t1 = o.time_sip(st, rhs, 1, threads)
    niter = max(2, int(target_s / max(t1, 1e-3)))
    t = o.time_sip(st, rhs, niter, threads)
    # one thread, about 3 s
    s1 = o.time_sip(st, rhs, 1, 1)
    n1 = max(2, int(3.0 / max(s1, 1e-3)))
    t_one = o.time_sip(st, rhs, n1, 1)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh
                          if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=round(g.ncells * niter / t / 1e9, 4), unit="GCell-updates/s",
                cores=threads, kind="port", cpu_model=model,
                single_thread=round(g.ncells * n1 / t_one / 1e9, 4),
                sample=f"{g.ngrid} blocks of 24^3 ({g.ncells} cells), {niter} hyperplane-SIP "
                       f"inner iterations (laplacephi+sipiter1+conn+sipiter2+conn), "
                       f"OpenMP over blocks, {t:.1f} s; single_thread: {n1} iterations "
                       f"on 1 thread, {t_one:.1f} s")


def copy_bandwidth(nbytes=1 << 31, reps=5):
    """Device copy rate in GB/s (bytes read + written) of torch's copy_ on
    two float32 tensors of `nbytes` each: the practical streaming rate the
    plane kernels are compared with."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / (best * 1e-3) / 1e9


def pmc_summary(fetch_dir, write_dir, out, cells=452984832):
    """Bytes per launch from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, both in KiB per dispatch), scaled by what the counters read
    for the library's plain convert kernel k_convert<float> (reads `cells`
    doubles, writes `cells` floats).  Writes the JSON the roofline line
    reads; prints the table."""
    import csv
    import re

    def load(d, counter):
        rows = []
        for f in Path(d).glob("*counter_collection.csv"):
            rows += list(csv.DictReader(open(f)))
        per = {}
        for r in rows:
            if r["Counter_Name"] != counter:
                continue
            name = re.sub(r"\(.*", "", r["Kernel_Name"]).strip()
            per.setdefault(name, []).append(float(r["Counter_Value"]) * 1024.0)
        return {k: sum(v) / len(v) for k, v in per.items()}

    fe, wr = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    conv = "void mgp::k_convert<float>"
    fs = 8.0 * cells / fe[conv]
    ws = 4.0 * cells / wr[conv]
    per = {}
    for k in sorted(fe):
        if "mgp::" in k and k in wr:
            per[k] = {"fetch_bytes": fe[k] * fs, "write_bytes": wr[k] * ws,
                      "total_bytes": fe[k] * fs + wr[k] * ws}
    res = {"workload": f"32768 blocks of 24^3, {cells} cells", "fetch_scale": fs,
           "write_scale": ws, "per_launch": per}
    Path(out).write_text(json.dumps(res, indent=1))
    print(f"fetch scale x{fs:.3f}, write scale x{ws:.3f}")
    print("| kernel | fetch | write | total |\n|---|---|---|---|")
    for k, v in per.items():
        print(f"| `{k}` | {v['fetch_bytes'] / 1e9:.2f} | {v['write_bytes'] / 1e9:.2f} "
              f"| {v['total_bytes'] / 1e9:.2f} |")


def main():
    if len(sys.argv) == 5 and sys.argv[1] == "--pmc-summary":
        pmc_summary(*sys.argv[2:5])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed outer iterations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ninner", type=int, default=5)
    ap.add_argument("--blocks", type=int, nargs=3, default=[32, 32, 32])
    ap.add_argument("--n", type=int, default=24)
    ap.add_argument("--tts", choices=["none", "c2", "main"], default="c2",
                    help="time-to-1e-10 on the 8x8x8-block case (c2), the main case, or skip")
    ap.add_argument("--tts-cap", type=int, default=4000, help="max outer iterations")
    ap.add_argument("--tts-eps", type=float, default=1e-10,
                    help="relative residual the time-to-solution runs stop at")
    ap.add_argument("--tts-history", default=None,
                    help="write each form's max|RHS| per outer iteration of the tts solve "
                         "to this .npz file")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64,fp32-unshared",
                    help="fp32 / fp64: the FP32-refined and all-FP64 forms; "
                         "fp32-unshared: FP32-refined with one copy of the SIP factors "
                         "per block (MGP_SHARED_FACTORS=0), timed the same way")
    ap.add_argument("--stage", action="store_true",
                    help="also time the RK stage kernels (tstle4, rkstep) on the workload")
    ap.add_argument("--unshared", action="store_true",
                    help="keep one copy of the SIP factors per block (MGP_SHARED_FACTORS=0)")
    args = ap.parse_args()
    if args.unshared:
        os.environ["MGP_SHARED_FACTORS"] = "0"

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (every block of the projection "
              "lives on one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mgp
    from mgp.grids import uniform_box

    nb = tuple(args.blocks)
    n = (args.n,) * 3
    log(f"[bench] building topology {nb} blocks of {n}")
    g = uniform_box(nb, n)
    log(f"[bench] generating velocity ({g.ncells} cells)")
    fields0 = make_velocity(g, nb, n, seed=4)
    forms = args.forms.split(",")
    res = {}
    stage = None
    filled = False
    for form in forms:
        log(f"[bench] init {form}")
        prec = form.split("-")[0]
        if form.endswith("-unshared"):
            os.environ["MGP_SHARED_FACTORS"] = "0"
        s = mgp.PressureSolver(g, ninner=args.ninner, nouter=1, precision=prec, check_every=4)
        if form.endswith("-unshared") and not args.unshared:
            os.environ.pop("MGP_SHARED_FACTORS", None)
        if not filled:      # the caller's connect: ghosts of the input velocity
            s.connect(fields0, layers=2, corners=True)
            filled = True
        if args.warmup > 0:
            run_steps(s, fields0, args.warmup, args.ninner)
        log(f"[bench] timing {form}: {args.steps} outer x {args.ninner} inner")
        wall, st = run_steps(s, fields0, args.steps, args.ninner)
        kt = s.time_kernels(niter=3)
        shape, shape_ms = s.plan_shape()
        if args.stage and stage is None:
            log("[bench] timing the RK stage kernels")
            stage = rk_stage_timing(s, fields0)
            log(f"[bench] rk stage: {stage}")
        res[form] = dict(wall=wall, stats=st, kernels=kt, info=s.info(),
                         factor_classes=s.factor_classes(),
                         autotune={"shape": shape, "backward_shape": s.plan_shape_backward(),
                                   "ms_per_outer_by_shape": [round(x, 3) for x in shape_ms]})
        log(f"[bench] {form}: {wall:.3f} s, device loop {st['solve_ms']:.1f} ms, kernels {kt}")
        s.close()
        del s
        torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth()
    cells = g.ncells
    main_form = "fp32" if "fp32" in res else forms[0]
    mp = main_form.split("-")[0]     # its precision
    r = res[main_form]
    upd = cells * args.ninner * args.steps
    value = upd / r["wall"] / 1e9
    fw_ms = r["kernels"]["forward_ms"]
    bw_ms = r["kernels"]["backward_ms"]
    # per-launch HBM bytes of the sweeps from the committed PMC summary
    traffic = {}
    tfile = ROOT / "profiles" / "traffic_sip_32768x24.json"
    if tfile.exists() and nb == (32, 32, 32) and n == (24, 24, 24):
        per = json.loads(tfile.read_text())["per_launch"]
        tname = "float" if mp == "fp32" else "double"
        for kern, key in (("forward", f"void mgp::k_forward<{tname}, false, false, true, "),
                          ("backward", f"void mgp::k_backward<{tname}, false, false, ")):
            hit = [k for k in per if k.startswith(key)]
            if hit:
                traffic[kern] = per[hit[0]]["total_bytes"] / 1e9
    upd_bytes = FUSED_BYTES[mp]
    inner_ms = fw_ms + bw_ms
    achieved = upd_bytes * cells / (inner_ms * 1e-3) / 1e9
    step_bytes = args.ninner * upd_bytes + OUTER_BYTES
    step_ms = r["wall"] / args.steps * 1e3
    step_gbs = step_bytes * cells / (step_ms * 1e-3) / 1e9
    per_kernel = {}
    for kern, ms in (("forward", fw_ms), ("backward", bw_ms)):
        a = ALG_BYTES[mp][kern] * cells / (ms * 1e-3) / 1e9
        t = traffic.get(kern)
        per_kernel[kern] = {"ms": round(ms, 4), "alg_bytes_per_cell": ALG_BYTES[mp][kern],
                            "achieved": round(a, 1), "frac": round(a / HBM_PEAK_GBS, 4),
                            "traffic": round(t, 3) if t else None,
                            # HBM bytes actually moved over the launch time:
                            # factors shared between blocks come from L2
                            "hbm_gbs": round(t / (ms * 1e-3), 1) if t else None}
    line = {
        "metric": "GCell-updates/s + time-to-1e-10 residual, FP32-refined vs FP64, 1 MI355X",
        "value": round(value, 3),
        "unit": "GCell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner sweeps, f64 outer" if mp == "fp32" else "f64",
        "data": "synthetic (seeded k^-5/3 Gaussian velocity, rms 1e-3)",
        "config": {"workload": f"SIP projection, {g.ngrid} blocks of {n[0]}^3 "
                               f"({nb[0]}x{nb[1]}x{nb[2]} periodic), {cells} cells",
                   "ninner": args.ninner, "step": f"1 outer iteration = {args.ninner} SIP "
                                                   "inner iterations + FP64 update"},
        "forms": {f: {"gcell_updates_per_s": round(cells * args.ninner * args.steps
                                                   / res[f]["wall"] / 1e9, 3),
                      "ms_per_step": round(res[f]["wall"] / args.steps * 1e3, 3),
                      # the outer-iteration loop alone (device time of the
                      # solve, without the per-call divergence, correction
                      # and exchanges)
                      "loop_ms_per_step": round(res[f]["stats"]["solve_ms"] / args.steps, 3),
                      "kernel_ms": {k: round(v, 4) for k, v in res[f]["kernels"].items()},
                      "device_gb": round(res[f]["info"]["device_bytes"] / 1e9, 2),
                      "factor_classes": res[f]["factor_classes"][0],
                      "autotune": res[f]["autotune"]}
                  for f in res},
        # one SIP inner iteration (forward + backward sweep) priced at the
        # fused algorithmic bytes per cell-update: DP, RHS, AP, the seven
        # factors and the mask read, DP written (45 B FP32, 89 B FP64); the
        # RES exchange and halo traffic are overhead, not counted
        "roofline": {"bound": "hbm",
                     "kernel": f"SIP inner iteration: k_forward + k_backward ({main_form})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (round(sum(traffic.values()), 3)
                                 if len(traffic) == 2 else None),
                     "traffic_unit": "GB per inner iteration (forward + backward launch; "
                                     "rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
                     "alg_bytes_per_cell_update": upd_bytes,
                     "inner_iteration_ms": round(inner_ms, 4),
                     "copy_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4),
                     # the whole timed step: ninner inner iterations plus the
                     # FP64 outer update (49 B per cell) over ms_per_step
                     "step": {"bytes_per_cell": step_bytes, "achieved": round(step_gbs, 1),
                              "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
                     "per_kernel": per_kernel},
    }
    if stage is not None:
        line["rk_stage"] = stage
    if "fp32" in res and "fp64" in res:
        line["fp32_over_fp64_throughput"] = round(res["fp64"]["wall"] / res["fp32"]["wall"], 3)

    if args.tts != "none":
        tnb = (8, 8, 8) if args.tts == "c2" else nb
        tg = uniform_box(tnb, n) if args.tts == "c2" else g
        tf = make_velocity(tg, tnb, n, seed=4) if args.tts == "c2" else fields0
        tts = {"config": f"{tg.ngrid} blocks of {n[0]}^3 ({tnb[0]}x{tnb[1]}x{tnb[2]} periodic)",
               "eps_rel": args.tts_eps}
        ps, hist = {}, {}
        tforms = [f for f in forms if not f.endswith("-unshared")]
        for form in tforms:
            s = mgp.PressureSolver(tg, ninner=args.ninner, nouter=1, precision=form)
            if form == tforms[0] and args.tts == "c2":
                s.connect(tf, layers=2, corners=True)
            time_to_solution(s, tf, args.ninner, args.tts_eps, 2)        # warm
            t, st, p = time_to_solution(s, tf, args.ninner, args.tts_eps, args.tts_cap)
            ps[form] = p
            if args.tts_history:
                hist[form] = s.history()
            tts[form] = {"seconds": round(t, 4), "inner_iterations": st["ipc"],
                         "outer_iterations": st["nouter_done"],
                         "converged": st["converged"],
                         "rel_residual": st["maxrhs"] / st["maxrhs0"],
                         # pipeline shape the autotuner picked for this batch
                         "autotune_shape": s.plan_shape()[0],
                         "autotune_backward_shape": s.plan_shape_backward()}
            log(f"[bench] tts {form}: {t:.3f} s, {st['ipc']} inner, conv {st['converged']}")
            s.close()
        if "fp32" in ps and "fp64" in ps:
            a, b = ps["fp32"], ps["fp64"]
            a = a - a.mean()
            b = b - b.mean()
            tts["fp32_pressure_err_vs_fp64"] = float((a - b).abs().max() / b.abs().max())
            tts["speedup_fp32_over_fp64"] = round(tts["fp64"]["seconds"] / tts["fp32"]["seconds"], 3)
        line["tts"] = tts
        if args.tts_history:
            import numpy as np
            np.savez(args.tts_history, **hist)

    if not args.no_cpu:
        log("[bench] cpu baseline")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
