"""Writes the golden fixtures under tests/golden/ from the reference's own
sources (oracle/_ref/libtgref.so, built by `make -C oracle` from
/root/reference).  Run in the build container only:

    python tests/golden/make_golden.py                   # every fixture
    python tests/golden/make_golden.py --only-big-tree   # points_rand2048_s1.npz only

Fixtures are data only: inputs (specs, points, grids) and the reference's
outputs for them.  Every array is saved with numpy (no pickles).
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Reference  # noqa: E402
from tangerine_amd import scenes  # noqa: E402

OUT = Path(__file__).resolve().parent


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def canonical_faces(faces: np.ndarray) -> np.ndarray:
    """Faces as a set: each triangle rotated so its smallest id is first, rows sorted."""
    f = np.asarray(faces, dtype=np.uint64).reshape(-1, 3)
    if f.size == 0:
        return f
    k = np.argmin(f, axis=1)
    rolled = np.stack([f[np.arange(len(f)), (k + i) % 3] for i in range(3)], axis=1)
    order = np.lexsort((rolled[:, 2], rolled[:, 1], rolled[:, 0]))
    return rolled[order]


def grid_points(origin, step, cells) -> np.ndarray:
    nx, ny, nz = (int(c) + 1 for c in cells)
    ii = np.arange(nx, dtype=np.float32)
    jj = np.arange(ny, dtype=np.float32)
    kk = np.arange(nz, dtype=np.float32)
    x = np.float32(origin[0]) + ii * np.float32(step[0])
    y = np.float32(origin[1]) + jj * np.float32(step[1])
    z = np.float32(origin[2]) + kk * np.float32(step[2])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X, Y, Z], axis=-1).reshape(-1, 3)


def spec_array(spec) -> np.ndarray:
    return np.array(json.dumps(spec))


def points_fixture(ref, name, spec, ext, rng) -> dict:
    """Point queries of one tree: tree walk (SDFNode::Eval) and interpreter
    values at uniform points, gradients, and ray marches."""
    h = ref.build(spec)
    words = ref.compile(h)
    pts = rng.uniform(-ext, ext, size=(4096, 3)).astype(np.float32)
    tree_d = ref.eval(h, pts)
    interp_d = ref.interp_eval(h, pts)
    gpts = pts[:512]
    grad = ref.gradient(h, gpts)
    starts = rng.uniform(-2 * ext, 2 * ext, size=(512, 3)).astype(np.float32)
    dirs = rng.normal(size=(512, 3)).astype(np.float32)
    hit, travel, pos = ref.raymarch(h, starts, dirs, 100, 1e-3)
    np.savez_compressed(
        OUT / f"points_{name}.npz", spec=spec_array(spec), words=words, points=pts, tree_eval=tree_d,
        interp_eval=interp_d, bounds=ref.bounds(h), leaf_count=np.int64(ref.fn("leaf_count")(h)),
        stack_size=np.int64(ref.fn("stack_size")(h)), grad_points=gpts, gradient=grad,
        ray_start=starts, ray_dir=dirs, ray_hit=hit, ray_travel=travel, ray_pos=pos)
    out = {"words": int(words.size), "stack": int(ref.fn("stack_size")(h))}
    ref.free(h)
    return out


def big_tree_fixture(ref, summary) -> None:
    """The 2048-primitive tree of the 1024^3 export (seed 1) over [-8,8]^3:
    tree-walk values at uniform points and at points near its surface (ray
    hits), with its own generator so the other fixtures stay as they are."""
    summary["rand2048_s1"] = points_fixture(ref, "rand2048_s1", scenes.random_csg_spec(1, 2048), 8.0,
                                            np.random.default_rng(20261203))


def main() -> int:
    ref = Reference()
    summary = json.loads((OUT / "summary.json").read_text()) if (OUT / "summary.json").exists() else {}
    if "--only-big-tree" in sys.argv:
        big_tree_fixture(ref, summary)
        (OUT / "summary.json").write_text(json.dumps(summary, indent=1) + "\n")
        return 0
    rng = np.random.default_rng(20240601)

    # ---- point queries for three trees --------------------------------------
    trees = {
        "config1": (scenes.config1_spec(), 2.5),
        "rand16_s5": (scenes.random_csg_spec(5, 16), 6.0),
        "rand256_s1": (scenes.random_csg_spec(1, 256), 6.0),
    }
    for name, (spec, ext) in trees.items():
        summary[name] = points_fixture(ref, name, spec, ext, rng)

    # ---- Clip outputs (program words of the clipped tree) --------------------
    h = ref.build(scenes.config1_spec())
    clips = []
    for c, r in [((0, 0, 0), 0.1), ((0.9, 0, 0), 0.2), ((1.3, 0, 0), 0.05), ((0, 0, 1.2), 0.3),
                 ((3, 3, 3), 0.5), ((0, 0, 0), 3.0), ((-0.9, 0.1, 0.0), 0.15)]:
        ch = ref.fn("clip")(h, *[float(v) for v in c], float(r))
        w = ref.compile(ch) if ch else np.zeros(0, dtype=np.uint32)
        if ch:
            ref.free(ch)
        clips.append((np.array(c, dtype=np.float32), np.float32(r), w))
    np.savez_compressed(OUT / "clip_config1.npz", centers=np.stack([c[0] for c in clips]),
                        radii=np.array([c[1] for c in clips]),
                        lengths=np.array([c[2].size for c in clips]),
                        words=np.concatenate([c[2] for c in clips]))

    # ---- config 1 grid: interpreter samples and serial surface nets ---------
    origin, step, cells = scenes.grid_for(-2.0, 2.0, 128)
    pts = grid_points(origin, step, cells)
    interp = ref.interp_eval(h, pts)
    tree = ref.eval(h, pts)
    v, f = ref.surface_nets(interp, origin, step, cells)
    cf = canonical_faces(f)
    vt, ft = ref.surface_nets(tree, origin, step, cells)
    summary["config1_grid"] = {"vertices": int(len(v)), "faces": int(len(f))}
    np.savez_compressed(
        OUT / "grid_config1.npz", origin=np.array(origin, dtype=np.float32), step=np.array(step, dtype=np.float32),
        cells=np.array(cells, dtype=np.int64),
        interp_sha=np.array(sha(interp)), tree_sha=np.array(sha(tree)),
        interp_vs_tree_max=np.float32(np.max(np.abs(interp - tree))),
        interp_vs_tree_flips=np.int64(np.count_nonzero((interp >= 0) != (tree >= 0))),
        mesh_vertices=np.int64(len(v)), mesh_faces=np.int64(len(f)),
        mesh_vertex_sha=np.array(sha(v)), mesh_face_sha=np.array(sha(cf)),
        tree_mesh_vertices=np.int64(len(vt)), tree_mesh_faces=np.int64(len(ft)),
        tree_mesh_vertex_sha=np.array(sha(vt)), tree_mesh_face_sha=np.array(sha(canonical_faces(ft))))

    # ---- small grids with full meshes (surface nets on identical values) ----
    for name, spec, lo, hi, n in [("config1_c24", scenes.config1_spec(), -2.0, 2.0, 24),
                                   ("rand16_c40", scenes.random_csg_spec(5, 16), -6.0, 6.0, 40)]:
        hh = ref.build(spec)
        origin, step, cells = scenes.grid_for(lo, hi, n)
        pts = grid_points(origin, step, cells)
        vals = ref.interp_eval(hh, pts)
        v, f = ref.surface_nets(vals, origin, step, cells)
        np.savez_compressed(OUT / f"mesh_{name}.npz", spec=spec_array(spec),
                            origin=np.array(origin, dtype=np.float32), step=np.array(step, dtype=np.float32),
                            cells=np.array(cells, dtype=np.int64), values=vals, vertices=v,
                            faces=canonical_faces(f))
        summary[name] = {"vertices": int(len(v)), "faces": int(len(f))}
        ref.free(hh)
    ref.free(h)

    # ---- point clouds (PointCloudExportThread steps over the whole tree) ----
    for name, spec, lo, hi, step, iters in [("config1", scenes.config1_spec(), (-2.0, -1.5, -1.5), (2.0, 1.5, 1.5),
                                             0.05, (0, 2)),
                                            ("rand16_s5", scenes.random_csg_spec(5, 16), (-6.0,) * 3, (6.0,) * 3,
                                             0.15, (0, 1))]:
        hh = ref.build(spec)
        clouds = {f"points_it{it}": ref.point_cloud(hh, lo, hi, step, it) for it in iters}
        np.savez_compressed(OUT / f"cloud_{name}.npz", spec=spec_array(spec), lo=np.array(lo, dtype=np.float32),
                            hi=np.array(hi, dtype=np.float32), step=np.float32(step),
                            iterations=np.array(iters, dtype=np.int64), **clouds)
        summary[f"cloud_{name}"] = {k: int(len(v)) for k, v in clouds.items()}
        ref.free(hh)

    # ---- MagicaVoxel: VoxExport steps over the reference tree, and its
    # VoxWriter over a scattered multi-model cell list ----------------------
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        spec = scenes.config1_spec()
        hh = ref.build(spec)
        path = Path(tmp) / "a.vox"
        n = ref.vox_export(hh, 16.0, 3, path)
        data = np.frombuffer(path.read_bytes(), dtype=np.uint8)
        np.savez_compressed(OUT / "vox_config1_g16.npz", spec=spec_array(spec), bounds=ref.bounds(hh),
                            grid_size=np.float32(16.0), color_index=np.int64(3), voxels=np.int64(n), vox_bytes=data)
        summary["vox_config1_g16"] = {"voxels": n, "bytes": int(data.size)}
        ref.free(hh)
        vrng = np.random.default_rng(77)
        size = np.array([130, 20, 300], dtype=np.uint32)
        cells = vrng.choice(int(np.prod(size)), size=1500, replace=False).astype(np.uint32)
        ref.vox_write(size, cells, 200, path)
        data = np.frombuffer(path.read_bytes(), dtype=np.uint8)
        np.savez_compressed(OUT / "vox_write_multi.npz", size=size, cells=cells, color_index=np.int64(200), vox_bytes=data)
        summary["vox_write_multi"] = {"cells": int(cells.size), "bytes": int(data.size)}

    big_tree_fixture(ref, summary)
    (OUT / "summary.json").write_text(json.dumps(summary, indent=1) + "\n")
    print(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
