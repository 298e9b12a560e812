"""Benchmark: CSG signed-distance evaluation on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload export1024|export512] [--no-cpu] [--legs ...]

A step is one plain exact export of the headline workload: every sample of the
grid interpreted on the GPU (per-brick pruned programs), then naive surface
nets on the device.  The tree is built and uploaded before timing; the grid,
the mesh and all intermediates stay in HBM.  One JSON line is printed.

Workloads (named by what they are):
  export1024  2048 random primitives (seed 1), 1024^3 cells over [-8,8]^3
              (1025^3 = 1.08e9 samples)
  export512   256 random primitives (seed 1), 512^3 cells over [-6,6]^3
              (513^3 = 1.35e8 samples)

`value` is grid samples resolved exactly per second of wall time per export
(one exact interpreter evaluation per sample).  Further legs, each an object
of the same line (--legs picks them; default all):
  fast        the headline export in fast numerics (fused multiply-adds,
              hardware square root): its FP32 roofline beside the exact one
  surrogate   surrogate fit, then culled exports (surrogate + certificate,
              certificate alone, surrogate alone) compared with the plain one
  batch64     64 seeded trees (8..2048 primitives, {64..384}^3 over [-6,6]^3):
              cold and warm whole-batch exports, and the captured persistent
              evaluation replayed
  trace1080p  sphere tracing of the 256-primitive tree (seed 1) at 1920x1080
              from a seeded eye on the sphere of radius 12: at most 256 steps,
              hit at d <= 1e-4, far 40, central-difference normals (h = 1e-3)
  export512   the 512^3 workload above (when the headline is export1024)

cpu_baseline (rank 0, these host cores): the reference's own threaded export
(its SDFOctree at 0.25 cells behind the surface-nets vertex and face thunks,
built from the reference sources into oracle/_ref) over z-slabs spread evenly
through the grid, extrapolated to the whole grid; for tracing, the
reference's SDFNode::RayMarch and SDFNode::Gradient over rows spread through
the image, on the same number of threads.  Without oracle/_ref the oracle's
interpreter fills the slabs ("port").
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
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
PEAK_BF16_DENSE_TFLOPS = 2500.0

# HBM bytes per launch of the block-evaluation kernel (k_eval_batch<P,true>,
# persistent), measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (separate runs) on the plain exact exports, each counter scaled by what it
# reads for a 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  At
# 1024^3 the algorithmic bytes are 4.54e9 (2,146,689 tiles x 2048 B of samples
# plus 64 B of sign bits each); counters cannot be read inside a timed run.
MEASURED_TRAFFIC_BYTES = {"export1024": 4.816e9, "export512": 0.603e9}

WORKLOADS = {
    "export1024": dict(seed=1, primitives=2048, lo=-8.0, hi=8.0, cells=1024),
    "export512": dict(seed=1, primitives=256, lo=-6.0, hi=6.0, cells=512),
}
LEGS = ("fast", "surrogate", "batch64", "trace1080p", "export512")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def roofline(flops: float, ms: float, kernel: str, **extra) -> dict:
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    out = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
           "frac_of_nonpacked_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS, "kernel": kernel,
           "flops_per_launch": flops, "launch_ms": ms}
    out.update(extra)
    return out


def block_kernel(st: dict, prune: bool) -> str:
    """The grid-evaluation kernel an export's stats describe."""
    if not prune:
        return "k_eval_tiles"
    if st["block_workgroups"]:
        return f"k_eval_batch<{st['points_per_lane']},true> (persistent block evaluation, one tree)"
    return f"k_eval_blocks<{st['points_per_lane']}>"


# ---- CPU baselines -------------------------------------------------------------

def spread_slabs(cells: int, count: int) -> list[int]:
    """`count` z-slab indices spread evenly through [0, cells)."""
    return sorted({min(cells - 1, int((i + 0.5) * cells / count)) for i in range(count)})


def cpu_export_baseline(spec, lo, hi, cells, budget_s: float, threads: int) -> dict:
    """Reference threaded export on `threads` host threads over z-slabs spread
    through the grid (one cube slab each), extrapolated to all `cells`."""
    origin, step, cc = scenes.grid_for(lo, hi, cells)
    samples = (cells + 1) ** 3
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
        done, elapsed, evals = [], 0.0, 0
        # slabs in an order that stays spread however early the budget ends
        order = []
        for count in (2, 4, 8, 16, 32, 64, 128, 256, 512, 1024):
            order += [k for k in spread_slabs(cells, count) if k not in order]
        for k in order:
            if elapsed >= budget_s:
                break
            t0 = time.perf_counter()
            _, _, ev = ref.export_threaded(octree, origin, step, cc, k, k + 1, threads)
            elapsed += time.perf_counter() - t0
            evals += ev
            done.append(k)
        ref.lib.tgref_octree_free(octree)
        ref.free(h)
        export_s = t_oct + elapsed * cells / len(done)
        return {
            "value": samples / export_s, "unit": "samples/s", "cores": threads, "kind": "reference",
            "sample": (f"reference threaded export (SDFOctree 0.25 + surface-nets thunks) over {len(done)} of "
                       f"{cells} cube z-slabs spread through the grid ({min(done)}..{max(done)}), {evals} "
                       f"SDFOctree::Eval calls in {elapsed:.1f} s; octree build {t_oct:.1f} s; whole export "
                       f"extrapolated"),
            "export_ms_extrapolated": 1000.0 * export_s,
            "evals_per_s": evals / elapsed,
            "hardware_concurrency": int(ref.lib.tgref_hardware_concurrency()),
        }
    orc = Oracle()
    h = orc.build(spec)
    words = orc.compile(h)
    planes, elapsed = 0, 0.0
    for k in spread_slabs(cells, 64):
        if elapsed >= budget_s:
            break
        t0 = time.perf_counter()
        orc.grid_interp(words, origin, step, cc, k, k + 1, threads)
        elapsed += time.perf_counter() - t0
        planes += 1
    rate = planes * (cells + 1) ** 2 / elapsed
    return {
        "value": rate, "unit": "samples/s", "cores": threads, "kind": "port",
        "sample": f"oracle interpreter over {planes} z-planes spread through the grid in {elapsed:.1f} s",
        "export_ms_extrapolated": 1000.0 * samples / rate,
    }


def cpu_trace_baseline(spec, cam, budget_s: float, threads: int) -> dict:
    """Reference SDFNode::RayMarch (256 steps, epsilon 1e-4, no far plane)
    and SDFNode::Gradient at the hits over image rows spread through the
    frame, chunks of rows on `threads` threads."""
    from tangerine_amd import trace
    try:
        from oracle.oracle import REF_SO, Reference
    except Exception as exc:  # pragma: no cover
        return {"error": f"oracle unavailable: {exc}"}
    if not REF_SO.exists():
        return {"error": "oracle/_ref is not built"}
    ref = Reference()
    h = ref.build(spec)
    dirs = trace.camera_directions(cam).reshape(int(cam.height), int(cam.width), 3)
    eye = np.array(cam.eye[:], dtype=np.float32)
    rows = []
    for count in (8, 16, 32, 64, 128, 256, 512, 1080):
        rows += [r for r in spread_slabs(int(cam.height), count) if r not in rows]

    def march(r):
        d = dirs[r]
        s = np.repeat(eye[None, :], d.shape[0], axis=0)
        hit, t, pos = ref.raymarch(h, s, d, 256, 1e-4)
        if hit.any():
            ref.gradient(h, pos[hit])
        return d.shape[0]

    rays, elapsed, used = 0, 0.0, 0
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        while used < len(rows) and elapsed < budget_s:
            chunk = rows[used:used + threads]
            used += len(chunk)
            t0 = time.perf_counter()
            rays += sum(ex.map(march, chunk))
            elapsed += time.perf_counter() - t0
    ref.free(h)
    return {
        "value": rays / elapsed, "unit": "rays/s", "cores": threads, "kind": "reference",
        "sample": (f"reference SDFNode::RayMarch + Gradient over {used} of {int(cam.height)} image rows spread "
                   f"through the frame ({rays} rays in {elapsed:.1f} s)"),
        "frame_ms_extrapolated": 1000.0 * int(cam.width) * int(cam.height) / (rays / elapsed),
    }


# ---- GPU legs -------------------------------------------------------------------

def export_leg(mesher, prog, origin, step, cells, args, fast: bool, tune: bool, label: str):
    """Timed plain exports (args.warmup + args.steps); returns (result, last stats, tuning)."""
    lib = _lib.load()
    tuning = None
    if args.tuning:
        vals = [int(v) for v in args.tuning.split(",")] + [0, 0]
        mesher.set_tuning(vals[0], vals[1], vals[3], vals[2])
        tuning = {"points_per_lane": vals[0], "waves": vals[1], "workgroups_per_cu": vals[2],
                  "min_prune_len": vals[3], "source": "--tuning"}
    elif tune and not args.no_prune:
        t0 = time.perf_counter()
        tuning = mesher.autotune(prog, origin, step, cells, fast=fast)
        tuning["wall_ms"] = 1000.0 * (time.perf_counter() - t0)
        log(f"{label} autotune: {tuning['points_per_lane']} points/lane, {tuning['waves']} waves, "
            f"{tuning['workgroups_per_cu']} workgroups/CU, {tuning['eval_ms']:.2f} ms")
    for i in range(args.warmup):
        mesher.export(prog, origin, step, cells, fast=fast, prune=not args.no_prune)
    _lib.check(lib.tg_device_synchronize(), "sync")
    block_ms, totals, evals_ms, mesh_ms = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = mesher.export(prog, origin, step, cells, fast=fast, prune=not args.no_prune)
        block_ms.append(st["block_ms"])
        totals.append(st["total_ms"])
        evals_ms.append(st["eval_ms"])
        mesh_ms.append(st["mesh_ms"])
    _lib.check(lib.tg_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    log(f"{label}: {1000.0 * wall / args.steps:.2f} ms per export")
    res = {"ms_per_step": 1000.0 * wall / args.steps, "export_ms_device": float(np.mean(totals)),
           "eval_ms": float(np.mean(evals_ms)), "mesh_ms": float(np.mean(mesh_ms)),
           "block_ms": float(np.mean(block_ms))}
    return res, st, tuning


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
    log(f"surrogate fit: {fit_wall:.2f} s, margin {fit['margin']:.4f}, loss {loss[0]:.4g} -> {loss[-1]:.4g}")

    def timed(surrogate, certify, label):
        for i in range(args.warmup):
            mesher.export_culled(prog, surrogate, origin, step, cells, fast=args.fast, certify=certify)
        totals, evals_ms = [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            st = mesher.export_culled(prog, surrogate, origin, step, cells, fast=args.fast, certify=certify)
            totals.append(st["total_ms"])
            evals_ms.append(st["eval_ms"])
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
            "tiles_evaluated": int(st["tiles_evaluated"]),
        }
        log(f"{label}: {res['export_ms']:.2f} ms, {res['bricks_kept']} bricks kept, identical {res['identical_mesh']}")
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
        "tiles_evaluated": res["tiles_evaluated"],
        "surrogate_alone": guess,
        "certificate_alone": cert_only,
        "fit": {"wall_ms": 1000.0 * fit_wall, "train_ms": fit["train_ms"], "sample_ms": fit["sample_ms"],
                "samples": fit["samples"], "calib_samples": fit["calib_samples"], "steps": fit["steps"],
                "margin": fit["margin"], "mean_residual": fit["mean_residual"],
                "loss_first": float(loss[0]), "loss_last": float(loss[-1]), "clamp": args.sur_clamp,
                "domain": [lo[0], lo[0] + (wl["hi"] - wl["lo"]) + 2 * pad],
                "mfma": train_mfma(fit)},
    }


def train_mfma(fit: dict) -> dict:
    """BF16 MFMA work of the fit's Adam steps over the replayed training
    graphs (gradient and Adam kernels): per 32-point group, the forward
    products 2(96*64 + 64*64)*32 and the backward dW2, dA1 and dW1 products
    2(64*64 + 64*64 + 64*96)*32; 16,384 points per step."""
    per_group = 2 * (96 * 64 + 64 * 64) * 32 + 2 * (64 * 64 + 64 * 64 + 64 * 96) * 32
    flops = float(per_group) * (16384 // 32) * fit["steps"]
    tflops = flops / (fit["train_ms"] * 1e-3) / 1e12 if fit["train_ms"] > 0 else 0.0
    return {"flops": flops, "tflops": tflops, "peak_bf16_dense_tflops": PEAK_BF16_DENSE_TFLOPS,
            "frac": tflops / PEAK_BF16_DENSE_TFLOPS}


def batch_leg(args) -> dict:
    """64 seeded trees exported together: a cold export (buffers allocated),
    warm exports (the whole pipeline again: pruning, evaluation, surface nets),
    then the captured persistent evaluation replayed."""
    from tangerine_amd.batch import Batch
    specs = scenes.batch_specs(4, 64)
    progs = [scenes.build_sdf(spec).program() for spec, _ in specs]
    grids = [scenes.grid_for(-6.0, 6.0, res) for _, res in specs]
    samples = sum((res + 1) ** 3 for _, res in specs)
    b = Batch()
    t0 = time.perf_counter()
    b.export(progs, grids, fast=args.fast)
    cold_ms = 1000.0 * (time.perf_counter() - t0)
    tune = None
    if not args.no_autotune:
        tune = b.autotune()
        log(f"batch autotune: {tune['points_per_lane']} points/lane, {tune['waves']} waves, "
            f"{tune['workgroups_per_cu']} workgroups/CU, {tune['eval_ms']:.2f} ms")
    warm = []
    for _ in range(max(1, min(args.steps, 5))):
        t0 = time.perf_counter()
        st_warm = b.export(progs, grids, fast=args.fast)
        warm.append(1000.0 * (time.perf_counter() - t0))
    for _ in range(args.warmup):
        b.replay()
    ms = [b.replay() for _ in range(max(args.steps, 1))]
    st = b.stats()
    b.close()
    replay_ms = float(np.mean(ms))
    out = {
        "trees": len(progs), "primitives": [int(p.info["brushes"]) for p in progs],
        "resolutions": [int(r) for _, r in specs], "samples": samples, "blocks": int(st["blocks"]),
        "export_ms_cold": cold_ms, "export_ms_warm": float(np.mean(warm)),
        "warm_split_ms": {"prune": st_warm["prune_ms"], "eval": st_warm["eval_ms"], "mesh": st_warm["mesh_ms"]},
        "eval_ms_replay": replay_ms,
        "samples_per_s_replay": samples / (replay_ms * 1e-3),
        "samples_per_s_warm_export": samples / (float(np.mean(warm)) * 1e-3),
        "roofline": roofline(st["flops"], replay_ms, "k_eval_batch (captured replay)"),
        "launch": {"points_per_lane": st["points_per_lane"], "waves": st["waves"], "workgroups": st["workgroups"],
                   "lds_bytes": st["lds_bytes"], "max_len": st["max_len"]},
        "autotune": None if tune is None else {k: tune[k] for k in ("points_per_lane", "waves", "workgroups_per_cu",
                                                                      "eval_ms")},
    }
    log(f"batch: {samples} samples in {len(progs)} trees, replay {replay_ms:.2f} ms, warm export "
        f"{out['export_ms_warm']:.1f} ms, cold {cold_ms:.1f} ms")
    return out


def trace_leg(args):
    """Sphere tracing at 1920x1080 (seeded eye on the sphere of radius 12)."""
    from tangerine_amd import trace
    spec = scenes.random_csg_spec(1, 256)
    prog = scenes.build_sdf(spec).program()
    cam = trace.look_at(trace.seeded_eye(3), (0, 0, 0), (0, 0, 1), 60.0, 1920, 1080)
    tr = trace.Tracer()
    for _ in range(args.warmup):
        tr.render(prog, cam, 256, 1e-4, 40.0, 1e-3, fast=args.fast)
    _lib.check(_lib.load().tg_device_synchronize(), "sync")
    trace_ms, totals = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = tr.render(prog, cam, 256, 1e-4, 40.0, 1e-3, fast=args.fast)
        trace_ms.append(st["trace_ms"])
        totals.append(st["trace_ms"] + st["prune_ms"])
    _lib.check(_lib.load().tg_device_synchronize(), "sync")
    wall = time.perf_counter() - t0
    tr.close()
    frame_ms = 1000.0 * wall / args.steps
    k_ms = float(np.mean(trace_ms))
    rays = int(st["rays"])
    out = {
        "workload": "256 random primitives (seed 1), 1920x1080 rays from a seeded eye at radius 12, <= 256 steps, "
                    "eps 1e-4, far 40, central-difference normals h = 1e-3",
        "rays": rays, "hits": int(st["hits"]), "evaluations": int(st["evaluations"]),
        "frame_ms": frame_ms, "frame_ms_device": float(np.mean(totals)), "trace_ms": k_ms,
        "prune_ms": float(st["prune_ms"]),
        "rays_per_s": rays / (frame_ms * 1e-3),
        "evals_per_s": st["evaluations"] / (frame_ms * 1e-3),
        "roofline": roofline(st["flops"], k_ms, "k_trace"),
    }
    log(f"trace: {frame_ms:.2f} ms per frame, k_trace {k_ms:.2f} ms, {out['rays_per_s']:.3e} rays/s")
    return out, spec, cam


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="export1024", choices=sorted(WORKLOADS))
    ap.add_argument("--legs", default=",".join(LEGS), help=f"comma list of extra legs among {', '.join(LEGS)} "
                                                          "('' for none)")
    ap.add_argument("--fast", action="store_true", help="fast numerics for every leg instead of bit parity")
    ap.add_argument("--no-prune", action="store_true", help="interpret the whole program in every tile")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tuning", default="", metavar="P,W[,G[,L]]",
                    help="block-evaluation launch (points per lane, waves per workgroup, persistent workgroups per "
                         "CU or 0 for a workgroup per block, shortest block program pruned per tile) instead of "
                         "autotuning")
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
    legs = [x for x in args.legs.split(",") if x]
    for x in legs:
        if x not in LEGS:
            ap.error(f"unknown leg {x}")

    wl = WORKLOADS[args.workload]
    lib = _lib.load()
    _lib.check(lib.tg_device_init(0), "tg_device_init")
    spec = scenes.random_csg_spec(wl["seed"], wl["primitives"])
    prog = scenes.build_sdf(spec).program()
    origin, step, cells = scenes.grid_for(wl["lo"], wl["hi"], wl["cells"])
    samples = (wl["cells"] + 1) ** 3
    log(f"workload {args.workload}: {prog.info}")

    mesher = Mesher()
    res, st, tuning = export_leg(mesher, prog, origin, step, cells, args, args.fast, not args.no_autotune,
                                 args.workload)
    nverts, ntris = mesher.counts()
    prune = not args.no_prune
    launch_ms = res["block_ms"] if prune and res["block_ms"] > 0 else res["eval_ms"]
    kernel = block_kernel(st, prune)
    traffic = MEASURED_TRAFFIC_BYTES.get(args.workload) if prune and not args.fast else None
    rl = roofline(st["flops"], launch_ms, kernel,
                  traffic=traffic,
                  traffic_gbs=(traffic / (launch_ms * 1e-3) / 1e9) if traffic else None,
                  hbm_peak_gbs=PEAK_HBM_GBS,
                  launch={"points_per_lane": st["points_per_lane"], "waves": st["waves"],
                          "workgroups": st["block_workgroups"],
                          "lds_bytes": st["block_lds_bytes"], "block_max_len": st["block_max_len"],
                          "block_max_binary": st["block_max_binary"]})
    out = {
        "metric": "exact SDF evals/s vs CPU; 1024^3 export ms, plain vs surrogate-culled, same mesh",
        "value": samples / (res["ms_per_step"] * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
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
        "value_note": "grid samples resolved per second (one exact interpreter evaluation each)",
        "export_ms_plain": res["export_ms_device"],
        "eval_ms": res["eval_ms"],
        "mesh_ms": res["mesh_ms"],
        "mesh": {"vertices": nverts, "triangles": ntris},
        "autotune": tuning,
        "roofline": rl,
    }

    if "fast" in legs and not args.fast:
        fres, fst, ftune = export_leg(mesher, prog, origin, step, cells, args, True, not args.no_autotune,
                                      f"{args.workload} fast")
        fms = fres["block_ms"] if prune and fres["block_ms"] > 0 else fres["eval_ms"]
        out["fast"] = {"ms_per_step": fres["ms_per_step"], "samples_per_s": samples / (fres["ms_per_step"] * 1e-3),
                       "autotune": ftune, "roofline": roofline(fst["flops"], fms, block_kernel(fst, prune) + " (fast numerics)")}
        # back to the exact export (and its launch shape) for the surrogate leg's comparison
        if tuning:
            mesher.set_tuning(tuning["points_per_lane"], tuning["waves"], tuning.get("min_prune_len", 0),
                              tuning.get("workgroups_per_cu", 0))
        mesher.export(prog, origin, step, cells, prune=prune)
    if "surrogate" in legs:
        out["surrogate_culled"] = surrogate_leg(mesher, prog, wl, origin, step, cells, args)
    mesher.close()
    if "batch64" in legs:
        out["batch64"] = batch_leg(args)
    trace_case = None
    if "trace1080p" in legs:
        out["trace1080p"], tspec, tcam = trace_leg(args)
        trace_case = (tspec, tcam)
    e512_spec = None
    if "export512" in legs and args.workload != "export512":
        w2 = WORKLOADS["export512"]
        e512_spec = scenes.random_csg_spec(w2["seed"], w2["primitives"])
        p2 = scenes.build_sdf(e512_spec).program()
        o2, s2, c2 = scenes.grid_for(w2["lo"], w2["hi"], w2["cells"])
        m2 = Mesher()
        r2, st2, t2 = export_leg(m2, p2, o2, s2, c2, args, args.fast, not args.no_autotune, "export512")
        m2.close()
        n2 = (w2["cells"] + 1) ** 3
        ms2 = r2["block_ms"] if prune and r2["block_ms"] > 0 else r2["eval_ms"]
        out["export512"] = {"workload": "256 random primitives (seed 1), 512^3 cells over [-6,6]^3",
                            "samples": n2, "ms_per_step": r2["ms_per_step"], "export_ms_device": r2["export_ms_device"],
                            "samples_per_s": n2 / (r2["ms_per_step"] * 1e-3), "autotune": t2,
                            "roofline": roofline(st2["flops"], ms2, block_kernel(st2, prune),
                                                 traffic=MEASURED_TRAFFIC_BYTES["export512"]
                                                 if prune and not args.fast else None)}

    if not args.no_cpu:
        log("cpu baselines ...")
        cb = cpu_export_baseline(spec, wl["lo"], wl["hi"], wl["cells"], args.cpu_seconds, args.cpu_threads)
        out["cpu_baseline"] = cb
        if cb.get("value"):
            out["gpu_over_cpu_samples_per_s"] = out["value"] / cb["value"]
        if e512_spec is not None:
            w2 = WORKLOADS["export512"]
            cb2 = cpu_export_baseline(e512_spec, w2["lo"], w2["hi"], w2["cells"], 0.6 * args.cpu_seconds,
                                      args.cpu_threads)
            out["export512"]["cpu_baseline"] = cb2
            if cb2.get("value"):
                out["export512"]["gpu_over_cpu_samples_per_s"] = out["export512"]["samples_per_s"] / cb2["value"]
        if trace_case is not None:
            cbt = cpu_trace_baseline(trace_case[0], trace_case[1], 0.6 * args.cpu_seconds, args.cpu_threads)
            out["trace1080p"]["cpu_baseline"] = cbt
            if cbt.get("value"):
                out["trace1080p"]["gpu_over_cpu_rays_per_s"] = out["trace1080p"]["rays_per_s"] / cbt["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
