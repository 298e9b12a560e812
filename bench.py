"""Benchmark: grid export of a seeded random CSG tree on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload export1024|export512] [--no-cpu]

A step is one plain exact export of the workload: every sample of the grid
interpreted on the GPU, then naive surface nets on the device.  The tree is
built and uploaded before timing; the grid, the mesh and all intermediates
stay in HBM.  One JSON line is printed (see the keys below).

Workloads (named by what they are):
  export1024  2048 random primitives (seed 1), 1024^3 cells over [-8,8]^3
              (1025^3 = 1.08e9 samples)
  export512   256 random primitives (seed 1), 512^3 cells over [-6,6]^3

cpu_baseline: the reference's own threaded export (its SDFOctree at 0.25
cells behind the surface-nets vertex and face thunks, built from the
reference sources into oracle/_ref) on this host, over a bounded number of
z-slabs; the full-grid time is extrapolated from them.  Without oracle/_ref
the oracle's interpreter fills the same slabs ("port").
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from tangerine_amd import _lib, scenes  # noqa: E402
from tangerine_amd.export import Mesher  # noqa: E402

PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector peak (packed issue)
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # same without packed FP32 (not used by these kernels)
PEAK_HBM_GBS = 8000.0

# HBM bytes per k_eval_blocks launch, measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE (separate runs) on the plain exact export at 1024^3 (launch
# 2 points per lane, 4 waves), each counter scaled by what it reads for a
# 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  The algorithmic
# bytes of that launch are 4.54e9 (2,146,689 tiles x 2048 B of samples plus
# 64 B of sign bits each); counters cannot be read inside a timed run.
MEASURED_TRAFFIC_BYTES = {"export1024": 4.671e9}

WORKLOADS = {
    "export1024": dict(seed=1, primitives=2048, lo=-8.0, hi=8.0, cells=1024),
    "export512": dict(seed=1, primitives=256, lo=-6.0, hi=6.0, cells=512),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(spec, lo, hi, cells, budget_s: float, threads: int) -> dict:
    """Reference threaded export on `threads` host threads over a few z-slabs."""
    origin, step, cc = scenes.grid_for(lo, hi, cells)
    try:
        from oracle.oracle import REF_SO, Oracle, Reference
    except Exception as exc:  # pragma: no cover
        return {"error": f"oracle unavailable: {exc}"}
    if REF_SO.exists():
        ref = Reference()
        h = ref.build(spec)
        t0 = time.perf_counter()
        octree = ref.octree(h, 0.25)
        t_oct = time.perf_counter() - t0
        k0 = cells // 2
        slabs, elapsed, evals = 0, 0.0, 0
        while elapsed < budget_s and k0 + slabs < cells:
            t0 = time.perf_counter()
            _, _, ev = ref.export_threaded(octree, origin, step, cc, k0 + slabs, k0 + slabs + 1, threads)
            elapsed += time.perf_counter() - t0
            evals += ev
            slabs += 1
        ref.lib.tgref_octree_free(octree)
        ref.free(h)
        rate = evals / elapsed
        return {
            "value": rate, "unit": "evals/s", "cores": threads, "kind": "reference",
            "sample": (f"reference threaded export (SDFOctree 0.25 + surface-nets thunks) over {slabs} of "
                       f"{cells} z-slabs at the grid centre, {evals} SDFOctree::Eval calls in {elapsed:.1f} s; "
                       f"octree build {t_oct:.1f} s"),
            "export_ms_extrapolated": 1000.0 * (t_oct + elapsed * cells / slabs),
            "hardware_concurrency": int(ref.lib.tgref_hardware_concurrency()),
        }
    orc = Oracle()
    h = orc.build(spec)
    words = orc.compile(h)
    k0 = cells // 2
    planes, elapsed = 0, 0.0
    while elapsed < budget_s and planes < cells:
        t0 = time.perf_counter()
        orc.grid_interp(words, origin, step, cc, k0 + planes, k0 + planes + 1, threads)
        elapsed += time.perf_counter() - t0
        planes += 1
    evals = planes * (cells + 1) ** 2
    return {
        "value": evals / elapsed, "unit": "evals/s", "cores": threads, "kind": "port",
        "sample": f"oracle interpreter over {planes} z-planes of the grid, {evals} evaluations in {elapsed:.1f} s",
        "export_ms_extrapolated": None,
    }


def surrogate_leg(mesher, prog, wl, origin, step, cells, args) -> dict:
    """Fit the brick-culling surrogate, time culled exports of the same grid,
    and check them against the plain export that ran last."""
    import hashlib
    from tangerine_amd.surrogate import Surrogate

    def digest(m):
        v, t = m.mesh()
        return hashlib.sha256(v.tobytes()).hexdigest(), hashlib.sha256(t.tobytes()).hexdigest(), v.shape, t.shape

    plain_digest = digest(mesher)
    active = mesher.active_bricks().astype(bool)
    pad = 0.25
    lo = (wl["lo"] - pad,) * 3
    sur = Surrogate(lo, (wl["hi"] - wl["lo"]) + 2 * pad, sigma=2.0, clamp=args.sur_clamp, seed=wl["seed"])
    t0 = time.perf_counter()
    fit = sur.fit(prog, samples=args.sur_samples, calib_samples=200_000, surface_sigma=0.05, epochs=args.sur_epochs,
                  fast=args.fast)
    fit_wall = time.perf_counter() - t0
    loss = sur.loss_log()
    log(f"surrogate fit: {fit_wall:.2f} s, margin {fit['margin']:.4f}, mean residual {fit['mean_residual']:.4f}, "
        f"loss {loss[0]:.4g} -> {loss[-1]:.4g}")

    def culled_once(surrogate, certify):
        return mesher.export_culled(prog, surrogate, origin, step, cells, fast=args.fast, prune=not args.no_prune,
                                    certify=certify)

    def timed(surrogate, certify, label):
        for i in range(args.warmup):
            culled_once(surrogate, certify)
        totals, evals_ms = [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            st = culled_once(surrogate, certify)
            totals.append(st["total_ms"])
            evals_ms.append(st["eval_ms"])
            log(f"{label} step {i}: select+prune {st['prune_ms']:.1f} eval {st['eval_ms']:.1f} mesh "
                f"{st['mesh_ms']:.1f} total {st['total_ms']:.1f} ms, kept {st['bricks_flagged']}/{st['bricks_total']} "
                f"bricks ({st['bricks_certified']} by the certificate)")
        _lib.check(_lib.load().tg_device_synchronize(), "sync")
        wall = time.perf_counter() - t0
        kept = mesher.brick_flags().astype(bool)
        res = {
            "export_ms": 1000.0 * wall / args.steps,
            "export_ms_device": float(np.mean(totals)),
            "eval_ms": float(np.mean(evals_ms)),
            "identical_mesh": bool(digest(mesher) == plain_digest),
            "missed_bricks": int(np.count_nonzero(active & ~kept)),
            "bricks_kept": int(st["bricks_flagged"]),
            "bricks_certified": int(st["bricks_certified"]),
            "exact_evals": int(st["exact_evals"]),
        }
        return res, st

    guess, _ = timed(sur, False, "surrogate alone")
    cert_only, _ = timed(None, True, "certificate alone")
    res, st = timed(sur, True, "surrogate + certificate")
    sur.close()
    return {
        "export_ms_culled": res["export_ms"],
        "export_ms_culled_device": res["export_ms_device"],
        "eval_ms_culled": res["eval_ms"],
        "identical_mesh": res["identical_mesh"],
        "missed_bricks": res["missed_bricks"],
        "bricks_kept": res["bricks_kept"],
        "bricks_certified": res["bricks_certified"],
        "bricks_total": int(st["bricks_total"]),
        "bricks_active": int(active.sum()),
        "cull_threshold": float(st["cull_threshold"]),
        "exact_evals_culled": res["exact_evals"],
        "surrogate_alone": guess,
        "certificate_alone": cert_only,
        "fit": {"wall_ms": 1000.0 * fit_wall, "train_ms": fit["train_ms"], "sample_ms": fit["sample_ms"],
                "samples": fit["samples"], "calib_samples": fit["calib_samples"], "steps": fit["steps"],
                "margin": fit["margin"], "mean_residual": fit["mean_residual"],
                "loss_first": float(loss[0]), "loss_last": float(loss[-1]), "clamp": args.sur_clamp,
                "domain": [lo[0], lo[0] + (wl["hi"] - wl["lo"]) + 2 * pad]},
    }


def batch_leg(args) -> dict:
    """64 seeded trees (primitive counts log-uniform in [8, 2048], resolutions
    from {64,128,192,256,384}^3 over [-6,6]^3) exported by one persistent
    launch; the captured evaluation launch is then replayed and timed."""
    from tangerine_amd.batch import Batch
    specs = scenes.batch_specs(4, 64)
    progs = [scenes.build_sdf(spec).program() for spec, _ in specs]
    grids = [scenes.grid_for(-6.0, 6.0, res) for _, res in specs]
    b = Batch()
    t0 = time.perf_counter()
    first = b.export(progs, grids, fast=args.fast)
    first_wall = 1000.0 * (time.perf_counter() - t0)
    tune = None
    if not args.no_autotune:
        tune = b.autotune()
        log(f"batch autotune: {tune['points_per_lane']} points/lane, {tune['waves']} waves, "
            f"{tune['workgroups_per_cu']} workgroups/CU, {tune['eval_ms']:.2f} ms")
    for _ in range(args.warmup):
        b.replay()
    ms = [b.replay() for _ in range(max(args.steps, 1))]
    st = b.stats()
    b.close()
    replay_ms = float(np.mean(ms))
    out = {
        "trees": len(progs), "primitives": [int(p.info["brushes"]) for p in progs],
        "resolutions": [int(r) for _, r in specs], "samples": int(st["samples"]), "blocks": int(st["blocks"]),
        "first_export_ms": first_wall, "prune_ms": first["prune_ms"], "eval_ms_first": first["eval_ms"],
        "mesh_ms": first["mesh_ms"], "eval_ms_replay": replay_ms,
        "evals_per_s": st["exact_evals"] / (replay_ms * 1e-3),
        "tflops": st["flops"] / (replay_ms * 1e-3) / 1e12,
        "autotune": None if tune is None else {k: tune[k] for k in ("points_per_lane", "waves", "workgroups_per_cu",
                                                                      "eval_ms")},
    }
    log(f"batch: {out['samples']} samples in {len(progs)} trees, eval {replay_ms:.2f} ms per replay, "
        f"{out['evals_per_s']:.3e} evals/s")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="export1024", choices=sorted(WORKLOADS))
    ap.add_argument("--fast", action="store_true", help="contracted FP32 numerics instead of bit parity")
    ap.add_argument("--no-prune", action="store_true", help="interpret the whole program in every tile")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate-culled export leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default block-kernel launch")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-tree batched export leg")
    ap.add_argument("--tuning", default="", metavar="P,W[,L]",
                    help="block-kernel launch (points per lane, waves per workgroup, shortest block program pruned "
                         "per tile) instead of autotuning")
    ap.add_argument("--sur-samples", type=int, default=2 * 1024 * 1024)
    ap.add_argument("--sur-epochs", type=int, default=16)
    ap.add_argument("--sur-clamp", type=float, default=0.5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("TG_CPU_THREADS", "16")))
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the device copy kernel over BYTES (calibrates the HBM byte counters) and exit")
    args = ap.parse_args(argv)
    if args.copy_probe:
        lib = _lib.load()
        ms = C.c_double()
        _lib.check(lib.tg_device_copy_probe(args.copy_probe, 3, C.byref(ms)), "tg_device_copy_probe")
        print(json.dumps({"copy_bytes": args.copy_probe, "ms_per_copy": ms.value,
                          "gb_per_s": 2 * args.copy_probe / (ms.value * 1e-3) / 1e9}), flush=True)
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; run with --gpus 1 (sharding across "
                                   "GPUs is not built)"}))
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "single-GPU benchmark; do not launch under torch.distributed"}))
        return 2

    wl = WORKLOADS[args.workload]
    lib = _lib.load()
    _lib.check(lib.tg_device_init(0), "tg_device_init")
    spec = scenes.random_csg_spec(wl["seed"], wl["primitives"])
    node = scenes.build_sdf(spec)
    prog = node.program()
    origin, step, cells = scenes.grid_for(wl["lo"], wl["hi"], wl["cells"])
    mesher = Mesher()
    prune = not args.no_prune
    log(f"workload {args.workload}: {prog.info}")

    def step_once():
        return mesher.export(prog, origin, step, cells, fast=args.fast, prune=prune)

    tuning = None
    if args.tuning:
        vals = [int(v) for v in args.tuning.split(",")] + [0]
        pts, waves, min_len = vals[0], vals[1], vals[2]
        mesher.set_tuning(pts, waves, min_len)
        tuning = {"points_per_lane": pts, "waves": waves, "min_prune_len": min_len, "source": "--tuning"}
    elif prune and not args.no_autotune:
        t0 = time.perf_counter()
        tuning = mesher.autotune(prog, origin, step, cells, fast=args.fast)
        tuning["wall_ms"] = 1000.0 * (time.perf_counter() - t0)
        log(f"autotune: {tuning['points_per_lane']} points/lane, {tuning['waves']} waves, min prune length "
            f"{tuning['min_prune_len']}, {tuning['eval_ms']:.2f} ms ({tuning['wall_ms']:.0f} ms to tune)")

    for i in range(args.warmup):
        st = step_once()
        log(f"warmup {i}: eval {st['eval_ms']:.1f} ms mesh {st['mesh_ms']:.1f} ms")
    _lib.check(lib.tg_device_synchronize(), "sync")
    evals_ms, mesh_ms, totals, block_ms = [], [], [], []
    flops = evals = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step_once()
        evals_ms.append(st["eval_ms"])
        block_ms.append(st["block_ms"])
        mesh_ms.append(st["mesh_ms"])
        totals.append(st["total_ms"])
        flops, evals = st["flops"], st["exact_evals"]
        log(f"step {i}: eval {st['eval_ms']:.1f} ms mesh {st['mesh_ms']:.1f} ms total {st['total_ms']:.1f} ms")
    _lib.check(lib.tg_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    nverts, ntris = mesher.counts()
    culled = None
    if not args.no_surrogate:
        culled = surrogate_leg(mesher, prog, wl, origin, step, cells, args)
    mesher.close()
    batch = None if args.no_batch else batch_leg(args)

    ms_per_step = 1000.0 * wall / args.steps
    samples = (wl["cells"] + 1) ** 3
    eval_ms = float(np.mean(evals_ms))
    # the dominant kernel alone: k_eval_blocks (pruned) or k_eval_tiles (unpruned)
    launch_ms = float(np.mean(block_ms)) if prune and np.mean(block_ms) > 0 else eval_ms
    achieved = flops / (launch_ms * 1e-3) / 1e12
    out = {
        "metric": "exact SDF evals/s vs CPU; 1024^3 export ms, plain vs surrogate-culled, same mesh",
        "value": evals / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random CSG tree, generated in-process)",
        "config": {
            "workload": f"{args.workload}: plain exact export (grid eval + naive surface nets)",
            "primitives": wl["primitives"], "seed": wl["seed"], "cells": wl["cells"],
            "bounds": [wl["lo"], wl["hi"]], "samples": samples, "numerics": "fast" if args.fast else "exact",
            "pruning": "per-brick (32^3 blocks, 8^3 tiles)" if prune else "none",
        },
        "export_ms_plain": float(np.mean(totals)),
        "eval_ms": eval_ms,
        "mesh_ms": float(np.mean(mesh_ms)),
        "samples_per_s": samples / (ms_per_step * 1e-3),
        "mesh": {"vertices": nverts, "triangles": ntris},
        "autotune": tuning,
        "surrogate_culled": culled,
        "batch64": batch,
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": MEASURED_TRAFFIC_BYTES.get(args.workload) if prune and not args.fast else None,
            "traffic_gbs": (MEASURED_TRAFFIC_BYTES[args.workload] / (launch_ms * 1e-3) / 1e9
                            if prune and not args.fast and args.workload in MEASURED_TRAFFIC_BYTES else None),
            "hbm_peak_gbs": PEAK_HBM_GBS,
            "frac_of_nonpacked_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS,
            "kernel": "k_eval_blocks" if prune else "k_eval_tiles", "flops_per_launch": flops, "launch_ms": launch_ms,
            "launch": {"points_per_lane": st["points_per_lane"], "waves": st["waves"],
                       "lds_bytes": st["block_lds_bytes"], "block_max_len": st["block_max_len"],
                       "block_max_binary": st["block_max_binary"]},
        },
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(spec, wl["lo"], wl["hi"], wl["cells"], args.cpu_seconds, args.cpu_threads)
        cb = out["cpu_baseline"]
        if cb.get("value"):
            out["gpu_over_cpu_evals"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
