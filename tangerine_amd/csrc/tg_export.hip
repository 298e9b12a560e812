// Grid export on the device: the MeshExportThread shape (evaluate the tree on
// a regular grid, then naive surface nets; tangerine/export.cpp:320-381) with
// every stage on one HIP stream and no host round trip except the two
// element counts the output allocation needs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"
#include "tg_mesh.h"
#include "tg_types.h"

using namespace tg;

#define TG_CHECK_HIP(expr, what)                         \
	do                                                   \
	{                                                    \
		hipError_t _e = (expr);                          \
		if (_e != hipSuccess) return hip_fail(_e, what); \
	} while (0)

struct tg_mesher
{
	hipStream_t stream = nullptr;
	hipEvent_t ev[4] = {};
	DevBuf pool, signs, active, word_count, word_base, quad_count, quad_base, vertices, triangles, scan_tmp, dense;
	DevBuf counters;
	CellPrograms blocks;
	MeshArgs args{};
	uint64_t nverts = 0, ntris = 0;
	bool have_grid = false;
	tg_export_stats stats{};
	int points_per_lane = 8;      // preferred points per lane (tuned)
	size_t lds_budget = 80 * 1024;  // LDS per workgroup the block kernel may use
	int last_points_per_lane = 0;
};

namespace {

constexpr uint64_t kMaxAxis = 1u << 14;  // cells per axis (sample indices fit 32 bits per axis)

int validate_grid(const tg_grid* g)
{
	if (!g) return fail(TG_ERR_INVALID, "null grid");
	if (g->sx == 0 || g->sy == 0 || g->sz == 0) return fail(TG_ERR_INVALID, "grid needs at least one cell per axis");
	if (g->sx >= kMaxAxis || g->sy >= kMaxAxis || g->sz >= kMaxAxis)
		return fail(TG_ERR_LIMIT, "grid axis exceeds 16383 cells");
	if (!(g->dx > 0.0f && g->dy > 0.0f && g->dz > 0.0f)) return fail(TG_ERR_INVALID, "grid step must be positive");
	const uint64_t words = g->sy * g->sz * ((g->sx + 63) / 64);
	if (words >= (1ull << 31)) return fail(TG_ERR_LIMIT, "grid has too many cube rows");
	// Vertex ids and quad counts are 32-bit: at most 3 quads per cube.
	if (double(g->sx) * double(g->sy) * double(g->sz) * 3.0 >= 4294967295.0)
		return fail(TG_ERR_LIMIT, "grid has too many cells for 32-bit vertex ids");
	return TG_OK;
}

// Fills the layout part of the mesher arguments and sizes the sample buffers.
int layout(tg_mesher* m, const tg_grid* g)
{
	MeshArgs& a = m->args;
	a = MeshArgs{};
	a.gx = g->x; a.gy = g->y; a.gz = g->z;
	a.dx = g->dx; a.dy = g->dy; a.dz = g->dz;
	a.sx = uint32_t(g->sx); a.sy = uint32_t(g->sy); a.sz = uint32_t(g->sz);
	a.nx = a.sx + 1; a.ny = a.sy + 1; a.nz = a.sz + 1;
	a.tx = (a.nx + 7) / 8; a.ty = (a.ny + 7) / 8; a.tz = (a.nz + 7) / 8;
	a.cbx = (a.sx + 7) / 8; a.cby = (a.sy + 7) / 8; a.cbz = (a.sz + 7) / 8;
	a.row_bytes = ((a.tx + 7) / 8) * 8;
	a.aw = (a.sx + 63) / 64;
	const uint64_t tiles = uint64_t(a.tx) * a.ty * a.tz;
	int rc = m->pool.ensure(tiles * kTileSamples * sizeof(float));
	if (!rc) rc = m->signs.ensure(uint64_t(a.ny) * a.nz * a.row_bytes);
	if (rc) return rc;
	a.pool = m->pool.as<float>();
	a.signs = m->signs.as<uint8_t>();
	a.tile_slot = nullptr;
	a.brick_flags = nullptr;
	return TG_OK;
}

// Surface nets over the samples already in pool/signs.
int surface_nets(tg_mesher* m)
{
	MeshArgs& a = m->args;
	const uint64_t words = uint64_t(a.sy) * a.sz * a.aw;
	int rc = m->active.ensure(words * 8);
	if (!rc) rc = m->word_count.ensure(words * 4);
	if (!rc) rc = m->word_base.ensure(words * 4);
	if (!rc) rc = m->quad_count.ensure(words * 4);
	if (!rc) rc = m->quad_base.ensure(words * 4);
	if (rc) return rc;
	a.active = m->active.as<uint64_t>();
	a.word_count = m->word_count.as<uint32_t>();
	a.word_base = m->word_base.as<uint32_t>();
	a.quad_count = m->quad_count.as<uint32_t>();
	a.quad_base = m->quad_base.as<uint32_t>();
	size_t tmp = 0;
	TG_CHECK_HIP(exclusive_scan_u32(nullptr, nullptr, words, nullptr, tmp, m->stream), "scan sizing");
	rc = m->scan_tmp.ensure(tmp);
	if (rc) return rc;

	TG_CHECK_HIP(mesh_classify(a, m->stream), "classify launch");
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_u32(a.word_count, a.word_base, words, m->scan_tmp.ptr, tmp, m->stream), "vertex scan");
	uint32_t tail[2] = { 0, 0 };
	TG_CHECK_HIP(hipMemcpyAsync(&tail[0], a.word_base + words - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipMemcpyAsync(&tail[1], a.word_count + words - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "vertex count");
	m->nverts = uint64_t(tail[0]) + tail[1];
	rc = m->vertices.ensure(std::max<uint64_t>(m->nverts, 1) * 12);
	if (rc) return rc;
	a.vertices = m->vertices.as<float>();
	TG_CHECK_HIP(mesh_vertices(a, m->stream), "vertex launch");
	TG_CHECK_HIP(mesh_quad_count(a, m->stream), "quad count launch");
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_u32(a.quad_count, a.quad_base, words, m->scan_tmp.ptr, tmp, m->stream), "quad scan");
	TG_CHECK_HIP(hipMemcpyAsync(&tail[0], a.quad_base + words - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipMemcpyAsync(&tail[1], a.quad_count + words - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "quad count");
	const uint64_t quads = uint64_t(tail[0]) + tail[1];
	m->ntris = quads * 2;
	rc = m->triangles.ensure(std::max<uint64_t>(quads, 1) * 24);
	if (rc) return rc;
	a.triangles = m->triangles.as<uint32_t>();
	TG_CHECK_HIP(mesh_quad_write(a, m->stream), "quad write launch");
	return TG_OK;
}

constexpr size_t kMaxLds = 160 * 1024;

// Two-level pruned evaluation of every tile (k_prune_cells over 32^3-sample
// blocks, then k_eval_blocks).  Returns 1 if it ran, 0 if the program is not
// prunable (the caller then interprets the whole program per tile), <0 on
// error.
int eval_pruned(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags, const uint64_t* tile_mask,
	const uint32_t* block_list, uint32_t n_listed, uint64_t* evals, double* flops)
{
	MeshArgs& ma = m->args;
	const uint32_t counts[3] = { (ma.tx + kBlockTiles - 1) / kBlockTiles, (ma.ty + kBlockTiles - 1) / kBlockTiles,
	                             (ma.tz + kBlockTiles - 1) / kBlockTiles };
	// Block (32 samples per axis) centred at g + (32 b + 15.5) d: a cell grid with
	// origin g - d/2 and cell 32 d; the samples lie within 15.5 |d| of the centre.
	const float origin[3] = { g->x - 0.5f * g->dx, g->y - 0.5f * g->dy, g->z - 0.5f * g->dz };
	const float cell[3] = { float(kBlockSamples) * g->dx, float(kBlockSamples) * g->dy, float(kBlockSamples) * g->dz };
	const float half = 0.5f * float(kBlockSamples - 1);
	const float radius = half * std::sqrt(g->dx * g->dx + g->dy * g->dy + g->dz * g->dz);
	int rc = prune_cells(m->stream, p, origin, cell, radius, counts, (flags & TG_FLAG_FAST) != 0, m->blocks);
	if (rc) return rc;
	if (!m->blocks.valid) return 0;
	rc = m->counters.ensure(64);
	if (rc) return rc;
	unsigned long long* work = m->counters.as<unsigned long long>();
	TG_CHECK_HIP(hipMemsetAsync(m->counters.ptr, 0, 64, m->stream), "counter reset");

	BlockEvalArgs ba{};
	ba.prog = static_cast<const Instr*>(p->device.ptr);
	ba.pool = m->blocks.pool.as<const Entry>();
	ba.desc = m->blocks.desc.as<const BlockDesc>();
	ba.block_list = block_list;
	ba.tile_mask = tile_mask;
	ba.nblocks = block_list ? n_listed : uint32_t(m->blocks.ncells);
	ba.gx = g->x; ba.gy = g->y; ba.gz = g->z;
	ba.dx = g->dx; ba.dy = g->dy; ba.dz = g->dz;
	ba.nx = ma.nx; ba.ny = ma.ny; ba.nz = ma.nz;
	ba.tx = ma.tx; ba.ty = ma.ty; ba.tz = ma.tz;
	ba.bx = counts[0]; ba.by = counts[1]; ba.bz = counts[2];
	ba.tile_slot = ma.tile_slot;
	ba.out_pool = m->pool.as<float>();
	ba.signs = m->signs.as<uint8_t>();
	ba.row_bytes = ma.row_bytes;
	ba.slots = std::max<uint32_t>(1, m->blocks.max_stack > 0 ? m->blocks.max_stack - 1 : 1);
	ba.prune_tiles = 1;
	ba.work = work;
	ba.whole = p->entries.as<const Entry>();
	ba.n_whole = uint32_t(p->decoded.entries.size());
	ba.waves = 4;
	int pts = m->points_per_lane;
	while (pts > 2 && block_lds_bytes(ba.slots, pts, ba.waves) > m->lds_budget) pts >>= 1;
	if (block_lds_bytes(ba.slots, pts, ba.waves) > kMaxLds) return 0;
	m->last_points_per_lane = pts;
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_eval_blocks(ba, pts, m->stream)
	                                      : exact::launch_eval_blocks(ba, pts, m->stream);
	if (e != hipSuccess) return hip_fail(e, "block evaluation launch");
	unsigned long long wk[2] = { 0, 0 };
	TG_CHECK_HIP(hipMemcpyAsync(wk, work, 16, hipMemcpyDeviceToHost, m->stream), "work read");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "block evaluation");
	*evals = wk[0];
	*flops = double(wk[1]);
	return 1;
}

}  // namespace

extern "C" int tg_mesher_create(tg_mesher** out)
{
	if (!out) return fail(TG_ERR_INVALID, "tg_mesher_create: null argument");
	int rc = ensure_device();
	if (rc) return rc;
	auto* m = new tg_mesher;
	hipError_t e = hipStreamCreate(&m->stream);
	for (int i = 0; e == hipSuccess && i < 4; ++i) e = hipEventCreate(&m->ev[i]);
	if (e != hipSuccess)
	{
		delete m;
		return hip_fail(e, "mesher stream/events");
	}
	*out = m;
	return TG_OK;
}

extern "C" void tg_mesher_destroy(tg_mesher* m)
{
	if (!m) return;
	if (m->stream) hipStreamSynchronize(m->stream);
	for (DevBuf* b : { &m->pool, &m->signs, &m->active, &m->word_count, &m->word_base, &m->quad_count, &m->quad_base,
	                   &m->vertices, &m->triangles, &m->scan_tmp, &m->dense, &m->counters })
		b->release();
	m->blocks.release();
	for (hipEvent_t e : m->ev)
		if (e) hipEventDestroy(e);
	if (m->stream) hipStreamDestroy(m->stream);
	delete m;
}

extern "C" int tg_mesher_export(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags)
{
	if (!m || !p) return fail(TG_ERR_INVALID, "tg_mesher_export: null argument");
	int rc = validate_grid(g);
	if (rc) return rc;
	m->have_grid = false;
	rc = layout(m, g);
	if (rc) return rc;
	MeshArgs& ma = m->args;
	GridArgs ga{};
	ga.prog = static_cast<const Instr*>(p->device.ptr);
	ga.gx = g->x; ga.gy = g->y; ga.gz = g->z;
	ga.dx = g->dx; ga.dy = g->dy; ga.dz = g->dz;
	ga.nx = ma.nx; ga.ny = ma.ny; ga.nz = ma.nz;
	ga.tx = ma.tx; ga.ty = ma.ty; ga.tz = ma.tz;
	ga.ntiles = ma.tx * ma.ty * ma.tz;
	ga.tile_list = nullptr;
	ga.tile_slot = nullptr;
	ga.pool = m->pool.as<float>();
	ga.signs = m->signs.as<uint8_t>();
	ga.row_bytes = ma.row_bytes;
	ga.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);

	TG_CHECK_HIP(hipEventRecord(m->ev[0], m->stream), "event");
	uint64_t evals = uint64_t(ga.ntiles) * kTileSamples;
	double flops = double(evals) * p->decoded.flops;
	int pruned = 0;
	if (!(flags & TG_FLAG_NO_PRUNE))
	{
		pruned = eval_pruned(m, p, g, flags, nullptr, nullptr, 0, &evals, &flops);
		if (pruned < 0) return pruned;
	}
	if (!pruned)
	{
		evals = uint64_t(ga.ntiles) * kTileSamples;
		flops = double(evals) * p->decoded.flops;
		hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_tiles(ga, 8, m->stream)
		                                      : exact::launch_tiles(ga, 8, m->stream);
		if (e != hipSuccess) return hip_fail(e, "tile evaluation launch");
	}
	TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
	rc = surface_nets(m);
	if (rc) return rc;
	TG_CHECK_HIP(hipEventRecord(m->ev[2], m->stream), "event");
	TG_CHECK_HIP(hipEventSynchronize(m->ev[2]), "export");
	float t_eval = 0, t_mesh = 0, t_all = 0;
	hipEventElapsedTime(&t_eval, m->ev[0], m->ev[1]);
	hipEventElapsedTime(&t_mesh, m->ev[1], m->ev[2]);
	hipEventElapsedTime(&t_all, m->ev[0], m->ev[2]);
	tg_export_stats& s = m->stats;
	s = tg_export_stats{};
	s.eval_ms = t_eval;
	s.mesh_ms = t_mesh;
	s.total_ms = t_all;
	s.samples = uint64_t(ma.nx) * ma.ny * ma.nz;
	s.tiles_evaluated = evals / kTileSamples;
	s.exact_evals = evals;
	s.flops = flops;
	m->have_grid = true;
	return TG_OK;
}

extern "C" int tg_mesher_mesh_values(tg_mesher* m, const float* values, const tg_grid* g)
{
	if (!m || !values) return fail(TG_ERR_INVALID, "tg_mesher_mesh_values: null argument");
	int rc = validate_grid(g);
	if (rc) return rc;
	m->have_grid = false;
	rc = layout(m, g);
	if (rc) return rc;
	MeshArgs& a = m->args;
	const uint64_t n = uint64_t(a.nx) * a.ny * a.nz;
	rc = m->dense.ensure(n * 4);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemcpyAsync(m->dense.ptr, values, n * 4, hipMemcpyHostToDevice, m->stream), "values upload");
	TG_CHECK_HIP(mesh_dense_to_pool(a, m->dense.as<float>(), m->stream), "values to tiles");
	TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
	rc = surface_nets(m);
	if (rc) return rc;
	TG_CHECK_HIP(hipEventRecord(m->ev[2], m->stream), "event");
	TG_CHECK_HIP(hipEventSynchronize(m->ev[2]), "mesh");
	float t_mesh = 0;
	hipEventElapsedTime(&t_mesh, m->ev[1], m->ev[2]);
	m->stats = tg_export_stats{};
	m->stats.mesh_ms = t_mesh;
	m->stats.total_ms = t_mesh;
	m->stats.samples = n;
	m->have_grid = true;
	return TG_OK;
}

extern "C" int tg_mesher_counts(const tg_mesher* m, uint64_t* v, uint64_t* t)
{
	if (!m || !v || !t) return fail(TG_ERR_INVALID, "tg_mesher_counts: null argument");
	*v = m->nverts;
	*t = m->ntris;
	return TG_OK;
}

extern "C" int tg_mesher_copy_mesh(const tg_mesher* m, float* v, uint32_t* t)
{
	if (!m) return fail(TG_ERR_INVALID, "tg_mesher_copy_mesh: null argument");
	if (m->nverts && v)
		TG_CHECK_HIP(hipMemcpy(v, m->vertices.ptr, m->nverts * 12, hipMemcpyDeviceToHost), "vertex download");
	if (m->ntris && t)
		TG_CHECK_HIP(hipMemcpy(t, m->triangles.ptr, m->ntris * 12, hipMemcpyDeviceToHost), "triangle download");
	return TG_OK;
}

extern "C" int tg_mesher_copy_grid(const tg_mesher* m, float* values)
{
	if (!m || !values) return fail(TG_ERR_INVALID, "tg_mesher_copy_grid: null argument");
	if (!m->have_grid) return fail(TG_ERR_INVALID, "tg_mesher_copy_grid: no grid evaluated yet");
	tg_mesher* mm = const_cast<tg_mesher*>(m);
	const MeshArgs& a = mm->args;
	const uint64_t plane = uint64_t(a.nx) * a.ny;
	// Convert in slabs of z-planes to bound the staging buffer.
	const uint32_t slab = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(a.nz, (256ull << 20) / (plane * 4))));
	int rc = mm->dense.ensure(plane * slab * 4);
	if (rc) return rc;
	for (uint32_t k0 = 0; k0 < a.nz; k0 += slab)
	{
		const uint32_t nk = std::min(slab, a.nz - k0);
		TG_CHECK_HIP(mesh_pool_to_dense(a, mm->dense.as<float>(), k0, nk, mm->stream), "tiles to dense");
		TG_CHECK_HIP(hipMemcpyAsync(values + plane * k0, mm->dense.ptr, plane * nk * 4, hipMemcpyDeviceToHost, mm->stream),
			"grid download");
		TG_CHECK_HIP(hipStreamSynchronize(mm->stream), "grid download");
	}
	return TG_OK;
}

extern "C" int tg_mesher_stats(const tg_mesher* m, tg_export_stats* out)
{
	if (!m || !out) return fail(TG_ERR_INVALID, "tg_mesher_stats: null argument");
	*out = m->stats;
	return TG_OK;
}
