"""Benchmark of the spin-projection grid kernel on one MI355X.

Workload (the plain-FP64 GPU configuration): synthetic factored Hamiltonian
n=64 spatial orbitals (NSO=128), N=32 electrons, L=192, Lebedev degree 11
(50 points) x 16 gamma = 800 grid points, complex128, seed 1002.

One step = one projected energy + gradient evaluation (EandG, PHF.py:55-91)
on the device-resident objective: Z in; grid sums, the 1 x 1 EigenSolver and
LocalG on the device; E0 and the (NSO-N) x N gradient block out.  The
Hamiltonian is built on the device from its factors; grid and Hamiltonian
stay resident.  value = grid points / s over the timed steps (of the fastest
form within 1e-6 Eh of the plain FP64 form).  The same line carries the
PHFTools_F drop-in path (host arrays, full Gradmat/Rdmmat per call;
--no-dropin skips it) and the 64-instance batch, the batched form of the
target (--no-batch skips it).

roofline: the dominant kernel, the ERI GEMM launch on the largest spin block
(D): algorithmic FLOPs per launch 2 x 2G x nD^2 (nD = 2 n^2, one product)
over its average launch time from HIP events on the stream it runs on;
frac_executed counts the split products the reduced forms execute.  Every
launch of the GEMM stage is listed beside it.

cpu_baseline: the oracle's CPU restatement on a bounded sample of grid points,
scaled to grid points / s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense matrix-core peaks (AMD public figures; int8 in OP/s)
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15, "f16x2": 2.5e15, "i8x4": 5.0e15}
# real products per complex-free product of each form (the split forms execute several)
NPROD = {"fp64": 1, "fp32": 1, "bf16x3": 6, "f16x2": 3, "i8x4": 10}
# algorithmic ERI GEMM FLOPs per evaluation (n spatial orbitals, G grid points,
# one density per point, real A; the re and im rows of each density count as
# two GEMM rows): dense 2 (2G) NSO^4 = 4 (2n)^4 G; spin blocks D (2n^2 x 2n^2)
# and X1, X2 (n^2 x n^2): 4 (4 + 2) n^4 G; spin-free (2J - K) on one set of
# rows and K on three: 4 (1 + 3) n^4 G
STRUCT_FLOPS = {"dense": lambda n, G: 4.0 * (2 * n) ** 4 * G,
                "spin-blocks": lambda n, G: 24.0 * n ** 4 * G,
                "spin-free": lambda n, G: 16.0 * n ** 4 * G}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(n=64, nocc=32, ngrid=(50, 16), seed=1002):
    from phf_pcc_amd.hamiltonian import synthetic_problem
    from phf_pcc_amd.spin_proj import grid_args
    t0 = time.time()
    prob = synthetic_problem(n, nocc, seed=seed, build_dense=True, ngrid=ngrid)
    log(f"[bench] built synthetic Hamiltonian n={n} NSO={2*n} in {time.time()-t0:.1f}s")
    ga = grid_args(n, ngrid, 0)
    return prob, ga


def cpu_baseline(prob, ga, ref, nocc, G, seconds):
    """The oracle's C + OpenMP restatement of BuildHSG (oracle/buildhsg_cref.c)
    on a bounded sample of the same grid and Hamiltonian, OpenMP over points."""
    from oracle import cref
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    htwo = np.ascontiguousarray(prob.htwo, dtype=np.complex128)
    allpts = [(il, iy, 0) for il in range(int(ga["ngrid"][0])) for iy in range(int(ga["ngrid"][1]))]
    args = dict(ngrid=ga["ngrid"], nci=ga["nci"], cmplxconj=1, ncisp=ga["ncisp"], nocc=nocc, r1=ga["r1"],
                r2=ga["r2"], rpg=ga["rpg"], weightsp=ga["weightsp"], weightpg=ga["weightpg"], roota=ga["roota"],
                rootb=ga["rootb"], rooty=ga["rooty"], qj=0, sp=1)
    npts = min(threads, G)
    t0 = time.perf_counter()
    cref.buildhsg(prob.hone, htwo, ref, prob.z, points=allpts[:npts], **args)
    t1 = time.perf_counter() - t0
    if t1 < seconds / 2 and npts < G:
        npts = min(G, max(npts, int(npts * seconds / max(t1, 1e-3)) // threads * threads))
        t0 = time.perf_counter()
        cref.buildhsg(prob.hone, htwo, ref, prob.z, points=allpts[:npts], **args)
        t1 = time.perf_counter() - t0
    return {"value": npts / t1, "unit": "grid pts/s", "cores": threads, "kind": "port",
            "sample": f"C/OpenMP restatement of BuildHSG (oracle/buildhsg_cref.c), {npts} of {G} grid points of "
                      f"the same workload in {t1:.1f} s; evals/s = value / {G}"}


def make_batch_workload(ninst=64, seed=1004):
    from phf_pcc_amd.hamiltonian import batch_instances
    return batch_instances(ninst, seed)


def batch_forms(forms, steps, warmup, ninst=64, baselines=True):
    """The batch workload: every instance evaluated (E and dE/dZ*) through
    phf_batch_eval per step, replayed from its hipGraph after the first call;
    one batch per form (spin blocks), plus the plain dense FP64 batch as the
    baseline.  Returns the record the JSON lines carry."""
    from phf_pcc_amd.batch import ProjectedHFBatch
    t0 = time.time()
    inst, zs, ns, gs = make_batch_workload(ninst)
    G = int(gs.sum())
    ns_ = ns.astype(np.float64)
    sflops = {m: float(np.sum(f(ns_, gs))) for m, f in STRUCT_FLOPS.items()}
    log(f"[bench] batch of {len(inst)} instances, n in [{ns.min()}, {ns.max()}], {G} grid points, "
        f"{sflops['spin-free']:.3e} algorithmic ERI GEMM flops per step in the spin-free form, "
        f"{sflops['spin-blocks']:.3e} over the spin blocks (host set-up {time.time()-t0:.1f}s)")
    if not baselines:
        # profiling runs: only the listed forms, no comparison
        for form in forms:
            bt = ProjectedHFBatch(inst, storage=form)
            for _ in range(max(1, warmup + steps)):
                bt.energy_and_gradient(zs, form, strict=False)
            log(f"[bench] batch {form}: {max(1, warmup + steps)} evaluations (profiling run, no baselines)")
            bt.close()
        return None
    if "fp64" not in forms:
        forms = forms + ["fp64"]
    # first the plain dense FP64 batch (202 GB of A: it runs while device memory
    # is unfragmented), then FP64 over the spin blocks, then every form in the
    # spin-free form (the instances' Refs are spin-restricted)
    runs = [("fp64_dense", "fp64", "dense"), ("fp64_spin_blocks", "fp64", "spin-blocks")]
    runs += [(f, f, "spin-free") for f in forms]
    # one child process per run: a batch whose launch sequence was captured in
    # a hipGraph does not return all of its device memory to the process when
    # it is destroyed (measured: about two thirds stay allocated), so runs in
    # one process would accumulate it; each child returns it at exit
    import subprocess
    import tempfile
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for key, form, want in runs:
            out = os.path.join(tmp, key + ".npz")
            cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "batch64", "--ninst", str(ninst),
                   "--batch-run", f"{form}:{want}:{out}", "--steps", str(steps), "--warmup", str(warmup)]
            r = subprocess.run(cmd)
            if r.returncode != 0:
                raise RuntimeError(f"batch run {key} failed with status {r.returncode}")
            d = np.load(out)
            meta = json.loads(str(d["meta"]))
            goff = np.cumsum([0] + [int(x) for x in d["gsizes"]])
            gflat = d["G"]
            res[key] = dict(E=d["E"], G=[gflat[goff[i]:goff[i + 1]] for i in range(len(goff) - 1)],
                            status=d["status"], **meta)
            log(f"[bench] batch {key} ({res[key]['mode']}): build {res[key]['build_s']:.1f}s  step "
                f"{res[key]['ms_step']:.2f} ms  device {res[key]['dev_ms']:.2f} ms  (eager GEMM "
                f"{res[key]['eager_gemm_ms']:.2f} ms)  instances failing the Hermiticity test: "
                f"{int((res[key]['status'] != 0).sum())}")
    base = res["fp64_dense"]
    passing = base["status"] == 0
    for f, r in res.items():
        de = np.abs(r["E"] - base["E"])
        dg = np.array([np.abs(a - b).max() for a, b in zip(r["G"], base["G"])])
        r["dE"] = float(de.max())
        r["dG"] = float(dg.max())
        r["dE_passing"] = float(de[passing].max()) if passing.any() else None
        log(f"[bench] batch {f}: max|dE| vs plain fp64 over all {len(de)} instances = {r['dE']:.3e} (over the "
            f"{int(passing.sum())} passing the Hermiticity test {r['dE_passing']:.3e})  max|dG| = {r['dG']:.3e}  "
            f"speedup vs plain fp64 = {r['value'] / base['value']:.2f}x  vs fp64 in the same structure = "
            f"{r['value'] / res['fp64']['value']:.2f}x  vs fp64 spin blocks = "
            f"{r['value'] / res['fp64_spin_blocks']['value']:.2f}x")
    okf = [r for k, r in res.items() if r["dE"] <= 1e-6 and k in forms]
    best = max(okf, key=lambda r: r["value"])
    bprec = best["prec"]
    nprod = NPROD[bprec]
    flops_best = sflops[best["mode"]]
    t_gemm = best["eager_gemm_ms"] * 1e-3
    grouped = bprec != "fp64"
    roof = {"bound": "mfma", "achieved": flops_best / t_gemm / 1e12, "peak": PEAK[bprec] / 1e12,
            "unit": "TFLOP/s", "frac": flops_best / t_gemm / PEAK[bprec],
            "frac_executed": nprod * flops_best / t_gemm / PEAK[bprec], "traffic": None,
            "kernel": _grouped_kernel(bprec, best["tuning"]) if grouped else "per-instance ERI GEMM launches",
            "algorithmic_flops_per_launch": flops_best, "executed_flops_per_launch": nprod * flops_best,
            "products": nprod, "structure": best["mode"], "ms_per_launch": best["eager_gemm_ms"],
            "time_source": "HIP events around the grouped launch (all instances) in the eager first evaluation"}
    traffic = load_traffic(roof["kernel"], "batch64")
    if traffic is not None:
        roof["traffic"] = traffic["bytes_per_launch"]
        roof["traffic_source"] = traffic["source"]
    forms_out = {f: {"grid_pts_per_s": r["value"], "evals_per_s": len(inst) * 1e3 / r["ms_step"],
                     "ms_per_step": r["ms_step"], "device_ms": r["dev_ms"], "eager_eri_gemm_ms": r["eager_gemm_ms"],
                     "build_s": r["build_s"], "max_abs_dE_vs_plain_fp64": r["dE"],
                     "max_abs_dE_vs_plain_fp64_passing": r["dE_passing"], "max_abs_dG_vs_plain_fp64": r["dG"],
                     "structure": r["mode"], "speedup_vs_plain_fp64": r["value"] / base["value"],
                     "speedup_vs_fp64_same_structure": r["value"] / res["fp64"]["value"],
                     "speedup_vs_fp64_spin_blocks": r["value"] / res["fp64_spin_blocks"]["value"],
                     "failing_instances": int((r["status"] != 0).sum()), "tuning": r["tuning"] or None}
                 for f, r in res.items()}
    for f, r in res.items():   # the grouped launch of every reduced form against its own MFMA peak
        if f != "fp64" and not f.startswith("fp64") and r["eager_gemm_ms"] > 0:
            fl = sflops[r["mode"]]
            np_ = NPROD[f]
            forms_out[f]["grouped_gemm"] = {"kernel": _grouped_kernel(f, r["tuning"]),
                                            "ms_per_launch": r["eager_gemm_ms"],
                                            "achieved_tflops": fl / (r["eager_gemm_ms"] * 1e-3) / 1e12,
                                            "peak_tflops": PEAK[f] / 1e12,
                                            "frac": fl / (r["eager_gemm_ms"] * 1e-3) / PEAK[f],
                                            "frac_executed": np_ * fl / (r["eager_gemm_ms"] * 1e-3) / PEAK[f]}
    return {"value": best["value"], "ms_per_step": best["ms_step"], "prec": bprec, "structure": best["mode"],
            "grid_points": G, "instances": len(inst),
            "workload": f"batched: {len(inst)} synthetic instances, n in {{16..96}}, N=n/2, L=3n, grids "
                        f"Lebedev {{6..50}} x gamma {{4,8,16}} (seed 1004), {G} grid points per step, "
                        "E + adjoint gradient per instance, hipGraph replay",
            "form": f"{bprec} ({best['mode']})",
            "evals_per_s": len(inst) * 1e3 / best["ms_step"],
            "speedup_vs_plain_fp64_batch": best["value"] / base["value"],
            "speedup_vs_fp64_same_structure": best["value"] / res["fp64"]["value"],
            "speedup_vs_fp64_spin_blocks": best["value"] / res["fp64_spin_blocks"]["value"],
            "instances_failing_hermiticity_test": int((~passing).sum()),
            "note": "an instance that fails the Hermiticity test (PHFTools.f90:195-200; coarse grids with complex Z) "
                    "costs the same work; its E0 = Re H / |Re S| is formed all the same and is compared too",
            "forms": forms_out, "roofline": roof}


def batch_run_child(spec, steps, warmup, ninst):
    """One run of the batch workload in this process (spec = form:structure:out):
    build, one eager evaluation (its ERI GEMM time), warm-up, timed graph
    replays; E, G, status and the timings go to the .npz `out`."""
    from phf_pcc_amd.batch import ProjectedHFBatch
    form, want, out = spec.split(":", 2)
    inst, zs, ns, gs = make_batch_workload(ninst)
    G = int(gs.sum())
    t0 = time.time()
    os.environ["PHF_SPINFREE"] = "1" if want == "spin-free" else "0"   # read by the instances' contexts
    bt = ProjectedHFBatch(inst, storage=form, structure=want != "dense")
    t_build = time.time() - t0
    e, g, _ = bt.energy_and_gradient(zs, form, strict=False)
    eager_total, eager_gemm, _ = bt.stats()
    for _ in range(max(0, warmup - 1)):
        e, g, _ = bt.energy_and_gradient(zs, form, strict=False)
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        e, g, _ = bt.energy_and_gradient(zs, form, strict=False)
        dev.append(bt.stats()[0])
    dt = time.perf_counter() - t0
    nsf = bt.spinfree_instances()
    mode = want if want != "spin-free" else ("spin-free" if nsf == len(inst) else "spin-blocks")
    assert want != "spin-free" or nsf in (0, len(inst)), nsf   # one structure per run
    meta = dict(ms_step=dt / steps * 1e3, value=G * steps / dt, dev_ms=float(np.mean(dev)), eager_gemm_ms=eager_gemm,
                eager_total_ms=eager_total, build_s=t_build, tuning=bt.tuning_report().strip(), prec=form, mode=mode)
    np.savez(out, E=np.asarray(e), G=np.concatenate([np.asarray(x).ravel() for x in g]),
             gsizes=np.array([np.asarray(x).size for x in g]), status=bt.status.copy(), meta=json.dumps(meta))
    bt.close()


def run_batch(args):
    """The batch workload as its own JSON line."""
    if args.batch_run:
        return batch_run_child(args.batch_run, args.steps, args.warmup, args.ninst)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    rec = batch_forms(forms, args.steps, args.warmup, args.ninst, baselines=not args.no_baselines)
    if rec is None:
        return
    out = {"metric": "projected E+grad evals/s (grid pts/s) vs plain FP64 GPU at 1e-6 Eh; % MFMA peak",
           "value": rec["value"], "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": rec["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": _form_dtype(rec["prec"]), "data": "synthetic",
           "config": {k: rec[k] for k in ("workload", "form", "evals_per_s", "speedup_vs_plain_fp64_batch",
                                          "speedup_vs_fp64_same_structure", "speedup_vs_fp64_spin_blocks",
                                          "instances_failing_hermiticity_test", "note", "forms")},
           "roofline": rec["roofline"], "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def _grouped_kernel(prec, tuning):
    """Kernel of a batch's grouped launch from its tuning record (the chosen
    tile configuration)."""
    cfg = None
    for rec in _report(tuning):
        if rec.get("grouped") and "chosen" in rec:
            cfg = rec["chosen"]["cfg"]
    return _cfg_kernel(prec, cfg, grouped=True)


def _tname(prec):
    """Element type of the grouped tile kernel's instantiation for a form."""
    return {"fp32": "f32", "bf16x3": "bf16", "f16x2": "f16", "i8x4": "i8"}[prec]


def _form_dtype(prec):
    return {"fp64": "f64", "fp32": "f32 (FP64 accumulate)", "bf16x3": "bf16x3 split (FP64 accumulate)",
            "f16x2": "f16x2 scaled split (FP64 accumulate)",
            "i8x4": "int8x4 scaled split (exact int32, FP64 accumulate)"}[prec]


def run_large(args):
    """n=128 (NSO=256), N=64, L=384, 50 x 24 grid: every precision form of the
    device-resident E + gradient on one factored Hamiltonian; errors against
    the FP64 form per energy and gradient."""
    from phf_pcc_amd.hamiltonian import synthetic_problem
    from phf_pcc_amd.model import ProjectedHF
    from phf_pcc_amd.spin_proj import grid_args
    n, nocc, ngrid = 128, 64, (50, 24)
    t0 = time.time()
    prob = synthetic_problem(n, nocc, seed=1003, ngrid=ngrid)
    ga = grid_args(n, ngrid, 0)
    model = ProjectedHF(nocc, prob.ref, ga, factors=(prob.h_spatial, prob.b))
    model.ctx.set_profiling(True)
    G = ngrid[0] * ngrid[1]
    nso = 2 * n
    log(f"[bench] n={n} problem and device Hamiltonian in {time.time()-t0:.1f}s")
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    if "fp64" not in forms:
        forms = ["fp64"] + forms
    res = {}

    def set_mode(m):
        model.ctx.set_blocked(m != "dense")
        model.ctx.set_spinfree(m == "spin-free")

    set_mode("spin-free")
    mode = "spin-free" if model.ctx.spinfree() else ("spin-blocks" if model.ctx.blocked() else "dense")
    runs = [("fp64_dense", "fp64", "dense"), ("fp64_spin_blocks", "fp64", "spin-blocks")]
    runs += [(f, f, mode) for f in forms]
    for key, f, m in runs:
        set_mode(m)
        model.precision = f
        for _ in range(args.warmup):
            model.energy_and_gradient(prob.z)
        st, t0 = [], time.perf_counter()
        for _ in range(args.steps):
            e, g = model.energy_and_gradient(prob.z)
            st.append(model.ctx.stage_times())
        dt = time.perf_counter() - t0
        sm = np.mean(np.array(st), axis=0)
        res[key] = dict(E=e, G=g, ms=dt / args.steps * 1e3, value=G * args.steps / dt, gemm_ms=sm[2], dev_ms=sm[0],
                        prec=f, mode=m, stages=sm)
        log(f"[bench] n=128 {key} ({m}): step {res[key]['ms']:.2f} ms  ERI GEMM {sm[2]:.2f} ms  E={e:.12f}")
    set_mode("spin-free")
    base = res["fp64_dense"]
    out_forms = {}
    for f, r in res.items():
        flops = STRUCT_FLOPS[r["mode"]](n, G)
        nprod = NPROD[r["prec"]]
        out_forms[f] = {"structure": r["mode"], "grid_pts_per_s": r["value"], "ms_per_step": r["ms"],
                        "eri_gemm_ms": r["gemm_ms"], "device_ms": r["dev_ms"], "E": r["E"],
                        "abs_dE_vs_fp64": abs(r["E"] - base["E"]),
                        "max_abs_dG_vs_fp64": float(np.abs(r["G"] - base["G"]).max()),
                        "speedup_vs_plain_fp64": r["value"] / base["value"],
                        "eri_gemm_algorithmic_flops": flops,
                        "mfma_frac": flops / (r["gemm_ms"] * 1e-3) / PEAK[r["prec"]],
                        "mfma_frac_executed": nprod * flops / (r["gemm_ms"] * 1e-3) / PEAK[r["prec"]]}
        log(f"[bench] n=128 {f}: |dE| = {out_forms[f]['abs_dE_vs_fp64']:.3e}  speedup vs plain fp64 "
            f"{out_forms[f]['speedup_vs_plain_fp64']:.2f}x  MFMA {100 * out_forms[f]['mfma_frac']:.1f}% "
            f"({100 * out_forms[f]['mfma_frac_executed']:.1f}% executed)")
    ok = [f for f in forms if out_forms[f]["abs_dE_vs_fp64"] <= 1e-6]
    best = max(ok, key=lambda f: res[f]["value"])
    bprec = res[best]["prec"]
    kernels, unpack = gemm_launches(mode, res[best]["stages"], G, n, bprec, model.ctx.tuning_report())
    dom = max(kernels, key=lambda k: k["ms_per_launch"])
    traffic = load_traffic(dom["kernel"], f"{n}-{dom['id']}")
    nprod = NPROD[bprec]
    out = {"metric": "projected E+grad evals/s (grid pts/s) vs plain FP64 GPU at 1e-6 Eh; % MFMA peak",
           "value": res[best]["value"], "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res[best]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": _form_dtype(bprec), "data": "synthetic",
           "config": {"workload": f"projected E+grad, synthetic n={n} (NSO={nso}), N={nocc}, L={3 * n}, grid 50x24={G} "
                                  "(seed 1003), device-resident objective", "form": f"{best} ({mode})",
                      "forms": out_forms},
           "roofline": {"bound": "mfma", "achieved": dom["achieved_tflops"], "peak": PEAK[bprec] / 1e12,
                        "unit": "TFLOP/s", "frac": dom["frac"], "frac_executed": dom["frac_executed"],
                        "traffic": None, "kernel": dom["kernel"], "launch": dom["launch"], "structure": mode,
                        "ms_per_launch": dom["ms_per_launch"],
                        "algorithmic_flops_per_launch": dom["algorithmic_flops_per_launch"],
                        "executed_flops_per_launch": nprod * dom["algorithmic_flops_per_launch"],
                        "products": nprod, "kernels": kernels, "unpack": unpack},
           "cpu_baseline": None}
    if traffic is not None:
        out["roofline"]["traffic"] = traffic["bytes_per_launch"]
        out["roofline"]["traffic_source"] = traffic["source"]
    print(json.dumps(out), flush=True)


def run_adjoint(args):
    """n=96 (NSO=192), N=48, L=288, 50 x 16 grid: the differentiable form.
    Autograd gradient against central differences (h = 1e-5, 32 random
    directions) and against the drop-in LocalG gradient; then 50 L-BFGS
    iterations from the seeded Z, timed, in the chosen precision."""
    import torch
    from phf_pcc_amd.hamiltonian import synthetic_problem
    from phf_pcc_amd.model import ProjectedEnergy, ProjectedHF, lbfgs_optimize
    from phf_pcc_amd.phf import eigen_solver, vec_decomp
    from phf_pcc_amd.spin_proj import grid_args
    n, nocc, ngrid = 96, 48, (50, 16)
    t0 = time.time()
    prob = synthetic_problem(n, nocc, seed=1005, ngrid=ngrid)
    ga = grid_args(n, ngrid, 0)
    model = ProjectedHF(nocc, prob.ref, ga, factors=(prob.h_spatial, prob.b))
    G = ngrid[0] * ngrid[1]
    log(f"[bench] n={n} problem and device Hamiltonian in {time.time()-t0:.1f}s")
    # gradient checks in FP64
    model.precision = "fp64"
    zr = torch.tensor(prob.z.real.copy(), requires_grad=True)
    zi = torch.tensor(prob.z.imag.copy(), requires_grad=True)
    e = ProjectedEnergy.apply(zr, zi, model)
    e.backward()
    grad = np.concatenate([zr.grad.numpy().ravel(), zi.grad.numpy().ravel()])
    x0 = np.concatenate([prob.z.real.ravel(), prob.z.imag.ravel()])
    nv = prob.z.size
    f = lambda x: model.energy_and_gradient((x[:nv] + 1j * x[nv:]).reshape(prob.z.shape))[0]
    rng = np.random.default_rng(1005)
    h = 1e-5
    fd_err = []
    t0 = time.time()
    for _ in range(32):
        d = rng.normal(size=x0.shape)
        d /= np.linalg.norm(d)
        fd = (f(x0 + h * d) - f(x0 - h * d)) / (2 * h)
        fd_err.append(abs(fd - grad @ d) / (1 + abs(fd)))
    t_fd = time.time() - t0
    h_, s_, gm, rm = model.buildhsg(prob.z)
    ev, vec = eigen_solver(h_, s_)
    fsp, fpg, fk = vec_decomp(vec[:, 0], model.ncisp, model.ncipg, 1)
    g_loc = model.localg(gm, rm, s_, ev[0], fsp, fpg, fk)[nocc:, :nocc]
    g_ana = np.concatenate([2 * g_loc.real.ravel(), 2 * g_loc.imag.ravel()])
    ana_err = float(np.abs(g_ana - grad).max())
    log(f"[bench] adjoint: max rel FD error over 32 directions {max(fd_err):.2e} ({t_fd:.1f}s); "
        f"max |autograd - LocalG| {ana_err:.2e}")
    res = {}
    forms = [f_.strip() for f_ in args.forms.split(",") if f_.strip()]
    for prec in forms:
        model.precision = prec
        model.nevals = 0
        t0 = time.perf_counter()
        z, energies = lbfgs_optimize(model, prob.z, steps=50)
        dt = time.perf_counter() - t0
        model.precision = "fp64"
        e_final64 = model.energy_and_gradient(z)[0]
        res[prec] = {"lbfgs_iterations": 50, "evaluations": len(energies), "seconds": dt,
                     "evals_per_s": len(energies) / dt, "grid_pts_per_s": len(energies) * G / dt,
                     "E_start": energies[0], "E_final": energies[-1], "E_final_fp64_check": e_final64}
        log(f"[bench] L-BFGS {prec}: {len(energies)} evaluations in {dt:.2f}s ({res[prec]['evals_per_s']:.1f} evals/s), "
            f"E {energies[0]:.8f} -> {energies[-1]:.8f} (FP64 re-evaluation {e_final64:.8f})")
    best = max(res, key=lambda k: res[k]["grid_pts_per_s"])
    out = {"metric": "projected E+grad evals/s (grid pts/s) vs plain FP64 GPU at 1e-6 Eh; % MFMA peak",
           "value": res[best]["grid_pts_per_s"], "unit": "grid pts/s", "n_gpus": 1, "steps": res[best]["evaluations"],
           "warmup": 0, "ms_per_step": res[best]["seconds"] / res[best]["evaluations"] * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": _form_dtype(best),
           "data": "synthetic",
           "config": {"workload": f"differentiable form: E + adjoint dE/dZ, synthetic n={n} (NSO={2 * n}), N={nocc}, "
                                  f"L={3 * n}, grid 50x16={G} (seed 1005), 50 L-BFGS iterations (torch, strong Wolfe)",
                      "form": best, "fd_check": {"directions": 32, "h": h, "max_rel_error": float(max(fd_err))},
                      "max_abs_autograd_minus_localg": ana_err, "lbfgs": res},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def run_buildsc(args):
    """Projected SD-space overlap times vector (BuildSC, PHFTools.f90:464-537)
    for LinDep: NSO=16 (8 spatial orbitals), N=8, Lebedev 14 x 8 gamma = 112
    spin-grid points in a random MO basis (seed 1006), blocks of 280 packed
    SD vectors (LinDep.py's Lanczos block; dim = 849) resident on the device.
    One step = one block product.  value = SD vectors / s.  roofline: the
    stage-2 FP64 complex GEMM (zgemm_kernel<2, 1>: the second pair transform
    with the weighted grid sum in its contraction, the largest launch; 8 real
    FLOPs per complex multiply-add as launched), its duration from HIP events
    on the library's stream; all four GEMMs of a step are reported beside it.
    cpu_baseline: the oracle's restatement of BuildSC (dense
    per-point BuildSC_Vec + ao2mo4, numpy) on a bounded sample of grid points
    for one vector, scaled to vectors / s over the full grid."""
    import torch
    from oracle import buildsc_oracle as B
    from phf_pcc_amd.sc import ProjectedOverlap
    from phf_pcc_amd.hamiltonian import random_unitary
    from phf_pcc_amd.spin_proj import grid_args
    nao, nocc, ngrid, nb = 8, 8, (14, 8), args.sc_block
    nso = 2 * nao
    rng = np.random.default_rng(1006)
    u = random_unitary(nso, rng)
    ga = grid_args(nao, ngrid, 0)
    r1 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r1"], u)
    r2 = np.einsum('pa,gpq,qb->gab', u.conj(), ga["r2"], u)
    op = ProjectedOverlap(nso, nocc, ga["ngrid"], r1, r2, ga["rpg"], ga["weightsp"], ga["weightpg"])
    G = int(np.prod(ga["ngrid"]))
    dim = op.dim
    x = rng.normal(size=(nb, dim)) + 1j * rng.normal(size=(nb, dim))
    dx = torch.tensor(x, dtype=torch.complex128, device="cuda")
    dy = torch.empty_like(dx)
    from phf_pcc_amd import _lib
    lib = _lib.load()
    h = op.ctx.handle

    def step():
        _lib.check(lib.phf_sc_apply_packed_dev(h, nb, dx.data_ptr(), dy.data_ptr()), h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    got = dy.cpu().numpy()
    # parity of vector 0 against the oracle over the full grid
    t = B.buildsc(*B.vec2mat(x[0], nso, nocc), nocc, ga["ngrid"], 0, r1, r2, ga["rpg"], ga["weightsp"],
                  ga["weightpg"], np.ones(1, complex), np.ones(1, complex))
    ref0 = B.mat2vec(*t, nocc)
    parity = float(np.abs(got[0] - ref0).max() / np.abs(ref0).max())
    log(f"[bench] buildsc: {args.steps} blocks of {nb} in {dt:.3f}s; parity vs oracle (vector 0) {parity:.2e}")
    # stage times of one profiled step
    op.ctx.set_profiling(True)
    step()
    st = op.ctx.stage_times()
    op.ctx.set_profiling(False)
    V = nso - nocc
    p12, p34 = V * (V - 1) // 2, nocc * (nocc - 1) // 2      # pair spaces a<b, i<j
    flops = 8.0 * (p12 * (nb * p34) * p12 * G             # stage 1: per point P12 x (nb P34) x P12
                   + (p12 * nb) * p34 * (G * p34)         # stage 2: (P12 nb) x P34 x (G P34)
                   + (V * nocc * nb) * G * (V * nocc)     # P
                   + (V * nocc) * (V * nocc * nb) * G)    # rank term
    gemm_ms = float(st[1] + st[2] + st[3] + st[4])
    nlaunch = int(st[7])
    all_gemm = flops / (gemm_ms * 1e-3)
    # dominant launch: the stage-2 GEMM (zgemm_kernel<2, 1>, one launch per step)
    flops2 = 8.0 * (p12 * nb) * p34 * (G * p34)
    achieved = flops2 / (float(st[2]) * 1e-3)
    log(f"[bench] buildsc stage ms: all {st[0]:.3f}, stage1 {st[1]:.3f}, stage2 {st[2]:.3f}, P {st[3]:.3f}, "
        f"rank {st[4]:.3f}, point kernel {st[5]:.3f}, rest {st[6]:.3f}; {nlaunch} GEMM launches at "
        f"{all_gemm / 1e12:.2f} TFLOP/s; stage-2 GEMM {achieved / 1e12:.2f} TFLOP/s")
    cpu = None
    if not args.no_cpu:
        pts = [(il, iy, 0) for il in range(int(ga["ngrid"][0])) for iy in range(int(ga["ngrid"][1]))]
        c0, c1, c2 = B.vec2mat(x[0], nso, nocc)
        done, tc = 0, 0.0
        t0c = time.perf_counter()
        while done < len(pts) and tc < args.cpu_seconds:
            B.buildsc(c0, c1, c2, nocc, ga["ngrid"], 0, r1, r2, ga["rpg"], ga["weightsp"], ga["weightpg"],
                      np.ones(1, complex), np.ones(1, complex), points=pts[done:done + 4])
            done += 4
            tc = time.perf_counter() - t0c
        per_vec = tc / done * G
        cpu = {"value": 1.0 / per_vec, "unit": "SD vectors/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
               "sample": f"oracle BuildSC (dense per-point BuildSC_Vec + ao2mo4, numpy einsum) on {done} of {G} "
                         f"grid points for one vector in {tc:.1f} s, scaled to the full grid"}
        log(f"[bench] buildsc CPU oracle: {cpu['value']:.3f} vectors/s")
    out = {"metric": "projected SD-space overlap x vector (BuildSC for LinDep): SD vectors/s",
           "value": nb * args.steps / dt, "unit": "SD vectors/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": f"BuildSC packed: NSO={nso}, N={nocc}, SD dim={dim}, spin grid 14x8={G} points "
                                  f"(seed 1006), block of {nb} vectors per step, device-resident",
                      "parity_rel_err_vector0": parity,
                      "stage_ms": {"all": st[0], "stage1_gemm": st[1], "stage2_gemm": st[2], "p_gemm": st[3],
                                   "rank_gemm": st[4], "point_kernel": st[5], "rest": st[6]}},
           "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK["fp64"], "traffic": load_traffic("zgemm_kernel<2, 1>", "so16-buildsc"),
                        "kernel": "zgemm_kernel<2, 1> (stage-2 pair transform with the grid sum)",
                        "flops_per_launch": flops2, "launches_per_step": 1,
                        "all_gemms": {"launches_per_step": nlaunch, "flops_per_step": flops,
                                      "achieved_tflops": all_gemm / 1e12}},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def run_pcc(args):
    """Projected-CC grid consumer, every output of PbarHbarOlap
    (FPUCC.f90:14-272, TrunODE = 1) through phf_pcc_olap: synthetic n=16
    (NSO=32), N=16, L=48 Hamiltonian in its MO basis (seed 1007), random
    complex T1, T2 ~ N(0, 0.05^2) (seed 1008), spin grid Lebedev 50 x 16
    gamma = 800 points as rotations in the MO basis.  One step = one call
    (host arrays in and out, as the Fortran routine takes them); value = grid
    points per second of device time (HIP events from after the uploads to
    the last grid sum).  roofline: the particle-particle ladder GEMM
    (1/2 <ab||ef> tau_efij, V^2 x N^2 x V^2 per point, one batched launch per
    chunk of points; 8 real FLOPs per complex multiply-add) on HIP events
    around its launches.  cpu_baseline: the oracle's restatement
    (numpy) of EvalKernel on a bounded sample of grid points, scaled."""
    from oracle import pcc_oracle as P
    from phf_pcc_amd import _lib
    from phf_pcc_amd.hamiltonian import synthetic_problem
    from phf_pcc_amd.pcc import pbar_hbar_olap
    from phf_pcc_amd.spin_proj import grid_args
    n, N, ngrid = 16, 16, (50, 16)
    nso, V = 2 * n, 2 * n - N
    p = synthetic_problem(n, N, seed=1007, build_dense=True)
    rng = np.random.default_rng(1008)
    t1 = 0.05 * (rng.normal(size=(V, N)) + 1j * rng.normal(size=(V, N)))
    t2 = P.asymm2(0.05 * (rng.normal(size=(V, V, N, N)) + 1j * rng.normal(size=(V, V, N, N)))) / 4
    ga = grid_args(n, ngrid, 0)
    ref = p.ref.astype(complex)
    tomo = lambda r: np.einsum('pa,gpq,qb->gab', ref.conj(), r, ref, optimize=True)
    pargs = (p.hone, p.htwo, t1, t2, N, 0.0, 0, 1, ga["ngrid"], tomo(ga["r1"]), tomo(ga["r2"]),
             np.eye(nso, dtype=complex)[None], None, ga["roota"], ga["rootb"], ga["rooty"], ga["weightsp"],
             np.ones((1, 1, 1)), np.ones(1, complex), np.ones(1, complex), np.ones(1, complex))
    G = int(np.prod(ga["ngrid"]))
    ctx = _lib.Context(0)
    ctx.set_profiling(True)
    for _ in range(args.warmup):
        out = pbar_hbar_olap(*pargs, ctx=ctx)
    dev, wall, st = [], [], None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = pbar_hbar_olap(*pargs, ctx=ctx)
        wall.append(time.perf_counter() - t0)
        st = ctx.stage_times()
        dev.append(float(st[0]))
    dev_ms, wall_s = float(np.mean(dev)), float(np.mean(wall))
    log(f"[bench] pcc: {args.steps} calls, device {dev_ms:.2f} ms, wall {wall_s * 1e3:.2f} ms per call; stages "
        f"zeroth {st[1]:.2f} ms, residual {st[2]:.2f} ms (ladder {st[3]:.2f} ms over {int(st[5])} launches), "
        f"sums {st[4]:.2f} ms")
    # parity on a sub-grid: the first 2 x 2 points against the oracle
    sub = list(pargs)
    sub[8] = np.array([2, 2])
    sub[9], sub[10] = sub[9][:2], sub[10][:2]
    sub[16] = ga["weightsp"][:2, :2]
    got = pbar_hbar_olap(*sub, ctx=ctx)
    want = P.pbar_hbar_olap(*sub)
    parity = max(float(np.abs(np.asarray(x) - np.asarray(y)).max() / np.abs(y).max()) for x, y in zip(got, want))
    log(f"[bench] pcc parity on a 2 x 2 sub-grid vs the oracle: max rel err {parity:.2e}")
    nlad = max(1, int(st[5]))
    chunk_pts = G / nlad
    lad_flops = 8.0 * V ** 4 * N * N * chunk_pts
    lad_ms = float(st[3]) / nlad
    achieved = lad_flops / (lad_ms * 1e-3)
    cpu = None
    if not args.no_cpu:
        rots = [(ia, iy) for ia in range(ngrid[0]) for iy in range(ngrid[1])]
        done, tc = 0, 0.0
        t0c = time.perf_counter()
        while done < len(rots) and tc < args.cpu_seconds:
            ia, iy = rots[done]
            P.eval_kernel(p.hone, p.htwo, t1, t2, 0.0, pargs[9][ia] @ pargs[10][iy], N)
            done += 1
            tc = time.perf_counter() - t0c
        cpu = {"value": done / tc, "unit": "grid pts/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
               "sample": f"oracle EvalKernel (numpy einsum, BLAS-ordered) on {done} of {G} grid points in {tc:.1f} s"}
        log(f"[bench] pcc CPU oracle: {cpu['value']:.2f} grid pts/s")
    rec = {"metric": "projected-CC grid consumer (PbarHbarOlap, every output): grid pts/s",
           "value": G / (dev_ms * 1e-3), "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": f"PbarHbarOlap TrunODE=1: synthetic n={n} (NSO={nso}), N={N}, L=48 (seed 1007), "
                                  f"T1, T2 ~ N(0, 0.05^2) (seed 1008), spin grid 50x16={G} points, all six outputs",
                      "wall_ms_per_call_with_host_transfers": wall_s * 1e3,
                      "stage_ms": {"device": float(st[0]), "zeroth_order": float(st[1]), "residual": float(st[2]),
                                   "ladder_gemm": float(st[3]), "sd_sums": float(st[4])},
                      "ladder_launches": int(st[5]), "parity_rel_err_2x2": parity},
           "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK["fp64"], "traffic": load_traffic("zgemm_kernel<2, 2>", "so32-pcc"),
                        "kernel": "zgemm_kernel<2, 2> (particle-particle ladder, 1/2 <ab||ef> tau_efij)",
                        "flops_per_launch": lad_flops, "ms_per_launch": lad_ms,
                        "time_source": "HIP events around the ladder launches on the library's stream"},
           "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(rec), flush=True)


def cpu_blas_gemm(prob, G):
    """Strong CPU baseline for the dominant stage: Gamma for every grid point as
    one BLAS dgemm A[(i,k),(j,l)] P (numpy's BLAS on all host cores; real A,
    re/im density columns).  Times the GEMM only (>= 95% of the per-point work
    at NSO = 128), so it bounds the CPU throughput from above."""
    nso = prob.nso
    n2 = nso * nso
    a = np.ascontiguousarray(prob.htwo.real.transpose(0, 2, 1, 3).reshape(n2, n2))
    rng = np.random.default_rng(0)
    p = rng.normal(size=(n2, 2 * G))
    a @ p[:, :16]
    t0 = time.perf_counter()
    a @ p
    t1 = time.perf_counter() - t0
    return {"value": G / t1, "unit": "grid pts/s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "blas", "sample": f"numpy BLAS dgemm {n2} x {n2} x {2 * G} (all {G} grid points) in {t1:.2f} s; "
                                      "ERI GEMM only"}


def gemm_kernel_name(prec, report):
    """Kernel of the ERI GEMM of form `prec` on its largest block, from the
    tuner's choice: configuration 4 is the 256 x 256 int8 tile, 5 the 256 x 256
    round-to-nearest 16-bit tile, the others the 128-row fold-tile kernels."""
    names = {"fp64": "eri_gemm_f64_kernel", "fp32": "eri_gemm_h16_kernel<f32>", "bf16x3": "eri_gemm_h16_kernel<bf16>",
             "f16x2": "eri_gemm_h16_kernel<f16>", "i8x4": "eri_gemm_h16_kernel<i8>"}
    code = {"fp32": 1, "bf16x3": 2, "f16x2": 3, "i8x4": 4}.get(prec)
    best_cols, cfg = -1, None
    for line in (report or "").splitlines():
        try:
            rec = json.loads(line)
        except ValueError:
            continue
        if code is not None and rec.get("prec") == code and "cols" in rec and rec["cols"] > best_cols:
            best_cols, cfg = rec["cols"], rec["chosen"]["cfg"]
    if prec == "i8x4" and cfg == 4:
        return "eri_gemm_i8_kernel"
    if prec in ("f16x2", "bf16x3") and cfg == 5:
        return "eri_gemm_rn16_kernel<%s>" % ("f16" if prec == "f16x2" else "bf16")
    return names[prec]


def x_group_kernel_name(prec, report):
    """Kernel of the exchange blocks' grouped launch, from the tuner's choice:
    the 256 x 256 tile (round-to-nearest 16-bit, exact int8) or the blocks'
    own tile in its grouped form."""
    t = _tname(prec)
    for line in (report or "").splitlines():
        try:
            rec = json.loads(line)
        except ValueError:
            continue
        ch = rec.get("exchange_blocks", {}).get("chosen")
        if ch == "grouped_256x256":
            return "eri_gemm_i8_grouped_kernel" if prec == "i8x4" else f"eri_gemm_rn16_grouped_kernel<{t}>"
        if ch == "grouped":
            return f"eri_gemm_h16_grouped_kernel<{t}>"
    return None


def _cfg_kernel(prec, cfg, grouped=False):
    """Kernel of tile configuration cfg for a form (4: 256 x 256 int8 tile, 5:
    256 x 256 round-to-nearest 16-bit tile, others: the 128-row tiles)."""
    if prec == "fp64":
        return "eri_gemm_f64_kernel"
    t = _tname(prec)
    if cfg == 4:
        return "eri_gemm_i8_grouped_kernel" if grouped else "eri_gemm_i8_kernel"
    if cfg == 5:
        return f"eri_gemm_rn16_grouped_kernel<{t}>" if grouped else f"eri_gemm_rn16_kernel<{t}>"
    return f"eri_gemm_h16_grouped_kernel<{t}>" if grouped else f"eri_gemm_h16_kernel<{t}>"


def _report(report):
    for line in (report or "").splitlines():
        try:
            yield json.loads(line)
        except ValueError:
            continue


def gemm_launches(mode, st, G, n, prec, report):
    """Every launch of the ERI GEMM stage of one evaluation from the profiled
    stage times st[8:] (HIP events): name, kernel, ms, algorithmic FLOPs (one
    product), achieved TFLOP/s and fractions of the form's peak; plus, for the
    spin-free form, the unpack into Gamma + h with its algorithmic bytes."""
    nprod, peak = NPROD[prec], PEAK[prec]
    code = {"fp32": 1, "bf16x3": 2, "f16x2": 3, "i8x4": 4}.get(prec)

    def row(ident, name, kernel, ms, alg):
        return {"id": ident, "launch": name, "kernel": kernel, "ms_per_launch": ms,
                "algorithmic_flops_per_launch": alg,
                "achieved_tflops": alg / (ms * 1e-3) / 1e12, "frac": alg / (ms * 1e-3) / peak,
                "frac_executed": nprod * alg / (ms * 1e-3) / peak}

    def tuned_cfg(rows, cols):
        for rec in _report(report):
            if rec.get("prec") == code and rec.get("rows") == rows and rec.get("cols") == cols and "chosen" in rec:
                return rec["chosen"]["cfg"]
        return None

    def ms(q):
        return float(st[8 + q]) if len(st) > 8 + q and st[8 + q] > 0 else None

    n2 = n * n
    if mode == "spin-free":
        a0, a1 = 2.0 * (2 * G) * n2 * n2, 2.0 * (6 * G) * n2 * n2
        c1 = tuned_cfg(6 * G, n2)
        chosen, half, split = None, False, False
        for rec in _report(report):
            if rec.get("prec") == code and "spin_free_units" in rec:
                chosen = rec["spin_free_units"]["chosen"]
                half = rec["spin_free_units"].get("unpack") == "aa_bb"
        if ms(1) is None:   # one grouped launch of both products, or its two halves on two streams
            cfg = (4 if prec == "i8x4" else 5) if chosen == "grouped_256x256" else c1
            split = half and os.environ.get("PHF_SPLIT_RJK", "1").strip() != "0"
            name = ("2J-K + K (the grouped tiles as two concurrent launches: Gamma_aa/bb product sets on the "
                    "main stream, exchange sets on a side stream; span from the first's start to the second's "
                    "end)") if split else "2J-K + K (one grouped launch)"
            kernels = [row("spinfree", name, _cfg_kernel(prec, cfg, True), ms(0), a0 + a1)]
            kernels[0]["launches"] = 2 if split else 1
        else:
            kernels = [row("2J-K", "2J-K", _cfg_kernel(prec, tuned_cfg(2 * G, n2)), ms(0), a0),
                       row("K", "K", _cfg_kernel(prec, c1), ms(1), a1)]
        # read the product sets, write their Gamma elements: all four (2G rows x (4 + 4) n^2 doubles),
        # or only Gamma_aa, Gamma_bb when the grouped launch wrote the exchange elements itself
        ub = 8.0 * (2 * G) * ((2 + 2) if half else (4 + 4)) * n2
        unpack = {"launch": "unpack into Gamma + h" + (" (aa, bb)" if half else "")
                  + (", beside the exchange launch" if ms(1) is None and split else ""), "kernel": "rjk_unpack_kernel",
                  "ms_per_launch": ms(2), "algorithmic_bytes": ub,
                  "GBps": ub / (ms(2) * 1e-3) / 1e9 if ms(2) else None,
                  "frac_hbm": ub / (ms(2) * 1e-3) / 8e12 if ms(2) else None}
        return kernels, unpack
    if mode == "spin-blocks":
        kname = gemm_kernel_name(prec, report)
        kernels = [row("D", "D", kname, ms(0), 2.0 * (2 * G) * (2 * n2) ** 2)]
        ax = 2.0 * (2 * G) * n2 * n2
        if ms(2) is None:
            kernels.append(row("X1+X2", "X1 + X2 (one grouped launch)", x_group_kernel_name(prec, report), ms(1),
                               2 * ax))
        else:
            kernels += [row("X1", "X1", kname, ms(1), ax), row("X2", "X2", kname, ms(2), ax)]
        return kernels, None
    return [row("dense", "dense", gemm_kernel_name(prec, report), ms(0), 2.0 * (2 * G) * (4 * n2) ** 2)], None


def load_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 counter summary (profiles/traffic.json), if present.  The key
    names the kernel and the workload (n and, for the spin-block GEMMs, the
    block), e.g. 'eri_gemm_h16_kernel<f16>@n64-blocks'."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
    except ValueError:
        return None
    rec = data.get(f"{kernel}@n{n}")
    return None if rec is None else {"bytes_per_launch": rec["bytes_per_launch"], "source": rec["source"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nocc", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,f16x2,i8x4",
                    help="comma list of ERI precisions to time (fp64 always runs: it is the reference form)")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip timing the PHFTools_F drop-in (host arrays in, Hmat/Smat/Gradmat/Rdmmat out per call)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-instance batch block of the default line")
    ap.add_argument("--batch-forms", default="f16x2,i8x4,fp32",
                    help="precisions of the batch block (fp64 and the plain dense FP64 batch always run)")
    ap.add_argument("--no-autotune", action="store_true", help="use the tile cost model instead of timing candidates")
    ap.add_argument("--sc-block", type=int, default=280, help="vectors per block of the buildsc workload")
    ap.add_argument("--workload", default="single",
                    choices=["single", "batch64", "large", "adjoint", "buildsc", "pcc"],
                    help="single: one n=64 instance (default); batch64: 64 uneven instances in one batch; "
                         "large: n=128, 50x24 grid, every precision; adjoint: n=96 gradient checks + 50 L-BFGS; "
                         "buildsc: BuildSC blocks; pcc: the projected-CC grid consumer (NSO=32, 800 points)")
    ap.add_argument("--ninst", type=int, default=64, help="instances of the batch64 workload")
    ap.add_argument("--batch-run", default=None, metavar="FORM:STRUCTURE:OUT",
                    help="batch64: one run in this process, results to the .npz OUT (the workload's own children)")
    ap.add_argument("--no-baselines", action="store_true",
                    help="batch64: evaluate only the listed forms and print no line (for profiler runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "read_bytes": nbytes, "write_bytes": nbytes}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the grid-point sharding over GPUs "
              "is a design in DESIGN.md, not a measured path", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        return run_batch(args)
    if args.workload == "large":
        return run_large(args)
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "buildsc":
        return run_buildsc(args)
    if args.workload == "pcc":
        return run_pcc(args)
    from phf_pcc_amd.PHFTools_F import phftools
    from phf_pcc_amd.model import ProjectedHF
    from phf_pcc_amd.phf import eigen_solver, vec_decomp

    prob, ga = make_workload(args.n, args.nocc)
    nso, nocc = prob.nso, prob.nocc
    G = int(ga["ngrid"][0] * ga["ngrid"][1])
    ref = prob.ref.astype(complex)
    t0 = time.time()
    model = ProjectedHF(nocc, ref, ga, factors=(prob.h_spatial, prob.b))
    log(f"[bench] device Hamiltonian (factored builder) and grid in {time.time()-t0:.1f}s")
    ctx = model.ctx
    ctx.set_profiling(True)

    def step():
        return model.energy_and_gradient(prob.z)

    def run_form(prec, stepfn=step, tag=""):
        """Timed steps without HIP events (the objective replays its captured
        launch graph), then as many profiled steps (eager launches with HIP
        events) for the device, stage and per-launch times."""
        model.precision = prec
        phftools.precision = prec
        sctx = stepfn.ctx if hasattr(stepfn, "ctx") else ctx
        sctx.set_profiling(False)
        for i in range(args.warmup):
            e, _ = stepfn()
        t0 = time.perf_counter()
        for i in range(args.steps):
            e, g = stepfn()
        dt = time.perf_counter() - t0
        sctx.set_profiling(True)
        gemm_ms, dev_ms, stages = [], [], []
        for i in range(args.steps):
            e2, g2 = stepfn()
            st = sctx.stage_times()
            dev_ms.append(st[0])
            gemm_ms.append(st[2])
            stages.append(st)
        assert abs(e2 - e) <= 1e-12 * max(1.0, abs(e)), (e, e2)   # replayed and eager evaluations agree
        r = {"prec": prec, "E": e, "G": g, "ms_step": dt / args.steps * 1e3,
             "value": G * args.steps / dt, "gemm_ms": float(np.mean(gemm_ms)), "dev_ms": float(np.mean(dev_ms)),
             "stages": np.mean(np.array(stages), axis=0)}
        log(f"[bench] {tag}{prec}: {args.steps} steps {dt:.3f}s  E={e:.12f}  step {r['ms_step']:.3f} ms  device "
            f"{r['dev_ms']:.3f} ms  ERI GEMM {r['gemm_ms']:.3f} ms")
        return r

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    if args.no_autotune:
        ctx.autotune(False)
    if "fp64" not in forms:
        forms = ["fp64"] + forms
    # the plain FP64 GPU form: MakeGamma over the dense NSO^2 x NSO^2 A
    ctx.set_blocked(False)
    plain = run_form("fp64", tag="plain dense ")
    # FP64 through the spin blocks D, X1, X2 of A (spin-pure Ref)
    ctx.set_blocked(True)
    ctx.set_spinfree(False)
    blocks64 = run_form("fp64", tag="spin-block ") if ctx.blocked() else None
    # every form through the spin-free form (spin-restricted Ref), else the blocks
    ctx.set_spinfree(True)
    mode = "spin-free" if ctx.spinfree() else ("spin-blocks" if ctx.blocked() else "dense")
    res = {f: run_form(f, tag=mode + " ") for f in forms}
    base = plain
    if blocks64 is not None:
        blocks64["dE"] = abs(blocks64["E"] - base["E"])
        blocks64["dG"] = float(np.abs(blocks64["G"] - base["G"]).max())
    for f, r in res.items():
        r["dE"] = abs(r["E"] - base["E"])
        r["dG"] = float(np.abs(r["G"] - base["G"]).max())
        log(f"[bench] {mode} {f}: |dE| vs plain fp64 = {r['dE']:.3e}  max|dG| = {r['dG']:.3e}  speedup vs plain "
            f"fp64 = {r['value'] / base['value']:.2f}x  vs fp64 {mode} = {r['value'] / res['fp64']['value']:.2f}x"
            + (f"  vs fp64 spin blocks = {r['value'] / blocks64['value']:.2f}x" if blocks64 else ""))
    dropin = {}
    if not args.no_dropin:
        phftools.set_hamiltonian(prob.hone, prob.htwo)
        dctx = phftools.context()
        dctx.set_profiling(True)
        call = (prob.hone, prob.htwo, ref, prob.z, ga["ngrid"], ga["nci"], 1, ga["ncisp"], nocc, ga["r1"], ga["r2"],
                ga["rpg"], ga["weightsp"], ga["weightpg"], ga["roota"], ga["rootb"], ga["rooty"], 0, 1, 0)

        def dstep():
            h, s_, gm, rm = phftools.buildhsg(*call)
            ev, vec = eigen_solver(h, s_)
            fsp, fpg, fk = vec_decomp(vec[:, 0], int(ga["ncisp"]), 1, 1)
            g = phftools.localg(gm, rm, s_, ev[0], fsp, fpg, fk, 1)
            return ev[0].real, g[nocc:, :nocc]
        dstep.ctx = dctx
        for f in ("fp64", "f16x2"):
            r = run_form(f, dstep, "drop-in ")
            dropin[f] = {"grid_pts_per_s": r["value"], "ms_per_step": r["ms_step"], "device_ms": r["dev_ms"],
                         "abs_dE_vs_plain_fp64": abs(r["E"] - plain["E"]),
                         "max_abs_dG_vs_plain_fp64": float(np.abs(r["G"] - plain["G"]).max())}
        dropin["path"] = ("phftools.buildhsg with the dense host htwo (uploaded once, cached by content hash), "
                          "Hmat/Smat/Gradmat/Rdmmat copied back per call, host EigenSolver, phftools.localg")
        phftools.invalidate()
        dctx.close()
    for line in ctx.tuning_report().splitlines():
        log(f"[bench] tuning {line}")
    ok = [r for r in res.values() if r["dE"] <= 1e-6]
    best = max(ok, key=lambda r: r["value"])
    n_sp = nso // 2
    flops_dense = STRUCT_FLOPS["dense"](n_sp, G)
    flops = STRUCT_FLOPS[mode](n_sp, G)
    nprod = NPROD[best["prec"]]
    peak = PEAK[best["prec"]]
    # each launch of the best form's ERI GEMM, timed with HIP events on the
    # stream it runs on: algorithmic FLOPs per launch = 2 x (density rows) x
    # (A rows) x (contraction length), one product
    kernels, unpack = gemm_launches(mode, best["stages"], G, n_sp, best["prec"], ctx.tuning_report())
    dom = max(kernels, key=lambda k: k["ms_per_launch"])
    t_gemm = best["gemm_ms"] * 1e-3
    roof = {"bound": "mfma", "achieved": dom["achieved_tflops"], "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": dom["frac"], "frac_executed": dom["frac_executed"], "traffic": None,
            "kernel": dom["kernel"], "launch": dom["launch"], "structure": mode,
            "ms_per_launch": dom["ms_per_launch"],
            "algorithmic_flops_per_launch": dom["algorithmic_flops_per_launch"],
            "executed_flops_per_launch": nprod * dom["algorithmic_flops_per_launch"], "products": nprod,
            "time_source": ("HIP events around this launch on the stream it runs on, mean over the timed steps"
                            if dom.get("launches", 1) == 1 else
                            "HIP events: start recorded before the first launch (main stream), end after the "
                            "second (side stream), mean over the timed steps"),
            "kernels": kernels, "unpack": unpack,
            "gemm_stage": {"ms": best["gemm_ms"], "launches": len(kernels) + (1 if unpack else 0),
                           "algorithmic_flops": flops, "achieved_tflops": flops / t_gemm / 1e12,
                           "frac": flops / t_gemm / peak, "frac_executed": nprod * flops / t_gemm / peak,
                           "note": ("the stage includes the unpack of the spin-free products"
                                    + (" (run beside the exchange launch)" if dom.get("launches", 1) == 2 else ""))
                           if unpack else None},
            "plain_fp64_dense_frac": flops_dense / (plain["gemm_ms"] * 1e-3) / PEAK["fp64"]}
    value = best["value"]
    ms_step = best["ms_step"]

    # HBM-bound stages of the plain FP64 form: algorithmic bytes per grid point
    # rotation 32 NSO N (read C, write T); density 16 NSO N + 16 N^2 + 2 s NSO^2
    # (read T, M; write the re/im rho planes, s = 8)
    st64 = plain["stages"]
    b_rot = 32.0 * nso * nocc * G
    b_den = (16.0 * nso * nocc + 16.0 * nocc ** 2 + 2 * 8 * nso ** 2) * G
    t_rot, t_den = st64[4] * 1e-3, (st64[5] + st64[6]) * 1e-3
    stage_out = {"rotation": {"ms": st64[4], "algorithmic_bytes": b_rot, "GBps": b_rot / t_rot / 1e9,
                              "frac_hbm": b_rot / t_rot / 8e12},
                 "density": {"ms": st64[5] + st64[6], "algorithmic_bytes": b_den, "GBps": b_den / t_den / 1e9,
                             "frac_hbm": b_den / t_den / 8e12, "pack_ms": st64[6], "overlap_lu_solve_ms": st64[5]},
                 "epilogue_ms": st64[3], "pack_split_ms_f16x2": (res["f16x2"]["stages"][6] + res["f16x2"]["stages"][7] if "f16x2" in res else None)}
    log(f"[bench] stages (fp64): rotation {st64[4]:.3f} ms = {stage_out['rotation']['GBps']:.0f} GB/s, density "
        f"{st64[5] + st64[6]:.3f} ms = {stage_out['density']['GBps']:.0f} GB/s (algorithmic bytes)")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, ga, ref, nocc, G, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.3f} grid pts/s on {cpu['cores']} threads ({cpu['sample']})")
        cpu["blas_gemm_variant"] = cpu_blas_gemm(prob, G)
        log(f"[bench] cpu BLAS ERI GEMM variant {cpu['blas_gemm_variant']['value']:.1f} grid pts/s "
            f"({cpu['blas_gemm_variant']['sample']})")

    traffic = load_traffic(roof["kernel"], f"{args.n}-{dom['id']}" if mode != "dense" else args.n)
    if traffic is not None:
        roof["traffic"] = traffic["bytes_per_launch"]
        roof["traffic_source"] = traffic["source"]

    def form_out(r, fl, structure):
        nprod_ = NPROD[r["prec"]]
        return {"structure": structure, "grid_pts_per_s": r["value"], "evals_per_s": r["value"] / G,
                "ms_per_step": r["ms_step"], "eri_gemm_ms": r["gemm_ms"], "device_ms": r["dev_ms"],
                "abs_dE_vs_plain_fp64": r.get("dE", 0.0), "max_abs_dG_vs_plain_fp64": r.get("dG", 0.0),
                "speedup_vs_plain_fp64": r["value"] / base["value"],
                "eri_gemm_algorithmic_flops": fl,
                "eri_gemm_frac": fl / (r["gemm_ms"] * 1e-3) / PEAK[r["prec"]],
                "eri_gemm_frac_executed": nprod_ * fl / (r["gemm_ms"] * 1e-3) / PEAK[r["prec"]],
                "launch_ms": [float(x) for x in r["stages"][8:]]}
    forms_out = {f: form_out(r, flops, mode) for f, r in res.items()}
    plain_out = form_out(plain, flops_dense, "dense")
    if blocks64 is not None:
        forms_out["fp64_spin_blocks"] = form_out(blocks64, STRUCT_FLOPS["spin-blocks"](n_sp, G), "spin-blocks")
    batch = None
    if not args.no_batch:
        # the batch runs in a child process with a device of its own: its plain
        # dense FP64 form alone holds 202 GB of A
        ctx.close()
        import subprocess
        cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "batch64", "--forms", args.batch_forms,
               "--steps", str(args.steps), "--warmup", str(args.warmup)]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
        if r.returncode != 0:
            log(f"[bench] batch child process failed with status {r.returncode}")
            sys.exit(r.returncode)
        line = json.loads(r.stdout.strip().splitlines()[-1])
        batch = dict(line["config"], value=line["value"], unit=line["unit"], ms_per_step=line["ms_per_step"],
                     dtype=line["dtype"], roofline=line["roofline"], command=" ".join(cmd[1:]))
    out = {"metric": "projected E+grad evals/s (grid pts/s) vs plain FP64 GPU at 1e-6 Eh; % MFMA peak",
           "value": value, "unit": "grid pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": _form_dtype(best["prec"]),
           "data": "synthetic",
           "config": {"workload": f"projected E+grad (EandG), synthetic n={args.n} (NSO={nso}), N={nocc}, "
                                  f"L={3*args.n}, grid 50x16={G}, device-resident objective",
                      "form": f"{best['prec']} ({mode})", "evals_per_s": value / G,
                      "grid_points": G, "speedup_vs_plain_fp64": value / base["value"],
                      "speedup_vs_fp64_same_structure": value / res["fp64"]["value"],
                      "speedup_vs_fp64_spin_blocks": value / blocks64["value"] if blocks64 else None,
                      "abs_dE_vs_fp64": best["dE"],
                      "plain_fp64": plain_out, "forms": forms_out,
                      "path": "device-resident objective (phf_eval_objective: grid sums, the 1 x 1 EigenSolver and "
                              "LocalG on the device), factored Hamiltonian "
                              "built on the device",
                      "hbm_stages_fp64": stage_out,
                      "drop_in_phftools": dropin or None,
                      "batch64": batch},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
