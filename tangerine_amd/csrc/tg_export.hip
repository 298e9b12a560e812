// Grid export on the device: the MeshExportThread shape (evaluate the tree on
// a regular grid, then naive surface nets; tangerine/export.cpp:320-381) with
// every stage on one HIP stream and no host round trip except the two
// element counts the output allocation needs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <random>
#include <vector>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"
#include "tg_mesh.h"
#include "tg_mlp.h"
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
	hipEvent_t ev[6] = {};
	DevBuf pool, signs, active, word_count, word_base, quad_count, quad_base, vertices, triangles, scan_tmp, dense, cubes;
	DevBuf normals, exchange;
	DevBuf cloud_flags, cloud_rank, vox_list;
	uint64_t nvox = 0;            // voxels of the last MagicaVoxel export
	bool have_normals = false;
	DevBuf counters;
	DevBuf brick_flags, brick_active, tile_mask, block_list, tile_list;
	uint64_t nbricks = 0;         // bricks of the last culled export (0: none)
	CellPrograms blocks;
	MeshArgs args{};
	uint64_t nverts = 0, ntris = 0;
	// room for the next mesh, from the last one (0: none yet): with it the
	// surface-net stages run without waiting for their counts
	uint64_t spec_verts = 0, spec_quads = 0;
	uint32_t* tail = nullptr;     // pinned: vertex and quad totals read back by surface nets, then
	                              // (from word 4) the block evaluation's two work counters
	uint32_t* tail_dev = nullptr; // the same words as the kernels address them
	bool have_grid = false;
	tg_export_stats stats{};
	int points_per_lane = 4;      // block kernel: points per lane (2, 4, 8; tg_mesher_autotune)
	int waves = 2;                // block kernel: waves per workgroup (1, 2, 4)
	int workgroups_per_cu = 8;    // 0: a workgroup per block (LDS tables); else persistent, global tables
	// launch shape of culled exports (tg_mesher_set_culled_tuning): their
	// near-surface tiles run deeper programs, which want more spill room per
	// wave than the plain export's tuning may leave
	int culled_points_per_lane = 4, culled_waves = 4, culled_workgroups_per_cu = 6;
	int cus = 0;                  // compute units of the device
	DevBuf persist_args, persist_head, persist_tables;  // the persistent launch's argument copy, queue head, tables
	DevBuf order_keys, order_values, order, order_tmp;   // its queue: blocks by program cost
	int last_workgroups = 0;
	uint32_t min_prune_len = 0;   // block kernel: shorter block programs skip per-tile pruning
	int last_points_per_lane = 0;
	int last_waves = 0;
	uint32_t last_max_len = 0, last_max_binary = 0, last_lds = 0;
	float last_block_ms = 0.0f;   // the last k_eval_blocks launch (HIP events)
	// Owned by a tg_batch (tg_batch_mesher): its buffers are captured in the
	// batch's graph, so calls that would lay them out again, and destroy, are
	// refused; the batch destroys it.
	bool batch_owned = false;
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
	a.aw = (a.sx + 63) / 64;
	a.sx_div = mesh_div_magic(a.sx);
	a.sy_div = mesh_div_magic(a.sy);
	a.aw_div = mesh_div_magic(a.aw);
	const uint64_t tiles = uint64_t(a.tx) * a.ty * a.tz;
	int rc = m->pool.ensure(tiles * kTileSamples * sizeof(float));
	if (!rc) rc = m->signs.ensure(tiles * 64);
	if (rc) return rc;
	a.pool = m->pool.as<float>();
	a.signs = m->signs.as<uint8_t>();
	a.tile_slot = nullptr;
	a.brick_flags = nullptr;
	return TG_OK;
}

// Surface nets over the samples already in pool/signs: classify cube rows,
// scan the per-word counts into vertex ids, list the active cubes, then one
// thread per active cube places its vertex and counts and (after a scan of the
// counts) writes its quads.  Three stages on the mesher's stream; each of the
// first two ends with an asynchronous copy of two counts into the pinned
// m->tail, which the next stage reads after the caller has synchronised the
// stream (the batch runs each stage over all its trees before waiting).
int surface_nets_classify(tg_mesher* m, bool copy_counts = true)
{
	MeshArgs& a = m->args;
	const uint64_t words = uint64_t(a.sy) * a.sz * a.aw;
	int rc = m->active.ensure(words * 8);
	if (!rc) rc = m->word_count.ensure(words * 4);
	if (!rc) rc = m->word_base.ensure(words * 4);
	if (rc) return rc;
	a.active = m->active.as<uint64_t>();
	a.word_count = m->word_count.as<uint32_t>();
	a.word_base = m->word_base.as<uint32_t>();
	size_t tmp = 0;
	TG_CHECK_HIP(exclusive_scan_u32(nullptr, nullptr, words, nullptr, tmp, m->stream), "scan sizing");
	rc = m->scan_tmp.ensure(tmp);
	if (rc) return rc;
	TG_CHECK_HIP(mesh_classify(a, m->stream), "classify launch");
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_u32(a.word_count, a.word_base, words, m->scan_tmp.ptr, tmp, m->stream), "vertex scan");
	if (copy_counts)
	{
		TG_CHECK_HIP(hipMemcpyAsync(&m->tail[0], a.word_base + words - 1, 4, hipMemcpyDeviceToHost, m->stream),
			"count copy");
		TG_CHECK_HIP(hipMemcpyAsync(&m->tail[1], a.word_count + words - 1, 4, hipMemcpyDeviceToHost, m->stream),
			"count copy");
	}
	return TG_OK;
}

// Room for a mesh after one of n vertices (or quads): a quarter more.
uint64_t room_for(uint64_t n) { return n + n / 4 + 4096; }

int surface_nets_vertices(tg_mesher* m)
{
	MeshArgs& a = m->args;
	m->nverts = uint64_t(m->tail[0]) + m->tail[1];
	const uint64_t nv = m->nverts;
	if (nv >= (1ull << 31)) return fail(TG_ERR_LIMIT, "surface nets: more than 2^31 vertices");
	// sized for the room the next mesh will run in (size_next_mesh), so
	// that it allocates nothing
	const uint64_t room = room_for(nv);
	int rc = m->vertices.ensure(room * 12);
	if (!rc) rc = m->cubes.ensure(room * 4);
	if (!rc) rc = m->quad_count.ensure(room * 4);
	if (!rc) rc = m->quad_base.ensure(room * 4);
	if (rc) return rc;
	a.vertices = m->vertices.as<float>();
	a.cubes = m->cubes.as<uint32_t>();
	a.cube_quads = m->quad_count.as<uint32_t>();
	a.cube_qbase = m->quad_base.as<uint32_t>();
	a.cap_verts = nv;
	m->tail[2] = m->tail[3] = 0;
	if (!nv) return TG_OK;
	TG_CHECK_HIP(mesh_list_cubes(a, m->stream), "cube list launch");
	TG_CHECK_HIP(mesh_vertices(a, nv, m->stream), "vertex launch");  // also each cube's quad flags
	size_t tmp = 0;
	TG_CHECK_HIP(exclusive_scan_quads(nullptr, nullptr, room, nullptr, tmp, m->stream), "scan sizing");
	rc = m->scan_tmp.ensure(tmp);
	if (rc) return rc;
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_quads(a.cube_quads, a.cube_qbase, nv, m->scan_tmp.ptr, tmp, m->stream), "quad scan");
	TG_CHECK_HIP(hipMemcpyAsync(&m->tail[2], a.cube_qbase + nv - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipMemcpyAsync(&m->tail[3], a.cube_quads + nv - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	return TG_OK;
}

// Room for the next mesh of this mesher: a quarter more than this one.
void size_next_mesh(tg_mesher* m)
{
	m->spec_verts = room_for(m->nverts);
	m->spec_quads = room_for(m->ntris / 2);
}

// Quads of the vertex stage's flags (m->tail[2..3]: the last cube's scanned
// base and its flags).
uint64_t quads_from_tail(const tg_mesher* m)
{
	return uint64_t(m->tail[2]) + uint64_t(__builtin_popcount(m->tail[3] & 7u));
}

int surface_nets_quads(tg_mesher* m)
{
	MeshArgs& a = m->args;
	const uint64_t quads = quads_from_tail(m);
	m->ntris = quads * 2;
	int rc = m->triangles.ensure(room_for(quads) * 24);
	if (rc) return rc;
	a.triangles = m->triangles.as<uint32_t>();
	a.cap_quads = quads;
	if (quads) TG_CHECK_HIP(mesh_quad_write(a, m->nverts, m->stream), "quad write launch");
	return TG_OK;
}

// All three stages at once, in the room m->spec_verts / m->spec_quads that an
// earlier mesh left: the kernels take the vertex count from the device and
// write within the room, and the host reads the counts once, at the end
// (surface_nets_room_check, after the stream is synchronised).
int surface_nets_in_room(tg_mesher* m)
{
	MeshArgs& a = m->args;
	const uint64_t cap_v = m->spec_verts, cap_q = m->spec_quads;
	// every buffer sized before the first launch (growing one frees it)
	size_t tmp = 0;
	TG_CHECK_HIP(exclusive_scan_quads(nullptr, nullptr, cap_v, nullptr, tmp, m->stream), "scan sizing");
	int rc = m->scan_tmp.ensure(tmp);
	if (!rc) rc = m->vertices.ensure(cap_v * 12);
	if (!rc) rc = m->cubes.ensure(cap_v * 4);
	if (!rc) rc = m->quad_count.ensure(cap_v * 4);
	if (!rc) rc = m->quad_base.ensure(cap_v * 4);
	if (!rc) rc = m->triangles.ensure(cap_q * 24);
	if (rc) return rc;
	rc = surface_nets_classify(m, false);
	if (rc) return rc;
	a.vertices = m->vertices.as<float>();
	a.cubes = m->cubes.as<uint32_t>();
	a.cube_quads = m->quad_count.as<uint32_t>();
	a.cube_qbase = m->quad_base.as<uint32_t>();
	a.triangles = m->triangles.as<uint32_t>();
	a.cap_verts = cap_v;
	a.cap_quads = cap_q;
	a.tail = m->tail_dev;  // the kernels write the totals: no copies to issue
	TG_CHECK_HIP(mesh_list_cubes(a, m->stream), "cube list launch");
	TG_CHECK_HIP(mesh_vertices(a, cap_v, m->stream), "vertex launch");
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_quads(a.cube_quads, a.cube_qbase, cap_v, m->scan_tmp.ptr, tmp, m->stream), "quad scan");
	TG_CHECK_HIP(mesh_quad_write(a, cap_v, m->stream), "quad write launch");
	a.tail = nullptr;
	return TG_OK;
}

// After surface_nets_in_room and a wait for the stream: whether the mesh fit
// the room (else it is incomplete and must be made the waiting way); sets the
// counts when it did.
bool surface_nets_room_check(tg_mesher* m)
{
	const uint64_t nv = uint64_t(m->tail[0]) + m->tail[1];
	if (nv > m->spec_verts) return false;
	// flags past the active cubes are zero, so the scan's total is the mesh's
	const uint64_t quads = quads_from_tail(m);
	if (quads > m->spec_quads) return false;
	m->nverts = nv;
	m->ntris = quads * 2;
	// a much smaller mesh than the room: shrink it for the next one
	if (m->nverts * 2 < m->spec_verts) size_next_mesh(m);
	return true;
}

// The waiting way: each stage after the previous stage's counts.
int surface_nets_waiting(tg_mesher* m)
{
	int rc = surface_nets_classify(m);
	if (rc) return rc;
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "vertex count");
	rc = surface_nets_vertices(m);
	if (rc) return rc;
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "quad count");
	rc = surface_nets_quads(m);
	if (rc) return rc;
	size_next_mesh(m);
	return TG_OK;
}


int surface_nets(tg_mesher* m)
{
	if (m->spec_verts > 0)
	{
		int rc = surface_nets_in_room(m);
		if (rc) return rc;
		TG_CHECK_HIP(hipStreamSynchronize(m->stream), "surface nets");
		if (surface_nets_room_check(m)) return TG_OK;
	}
	return surface_nets_waiting(m);
}

// Dynamic LDS a workgroup may ask for: the CU's 160 KiB less room for the
// kernels' static LDS (the persistent batch kernel's queue item), so that
// the persistent launches that split the CU's LDS evenly never overshoot.
constexpr size_t kMaxLds = 160 * 1024 - 256;

// Wave per block of 4^3 sample tiles, lane = tile (bit t = x + 4 y + 16 z).
// A kept cube brick b (cubes [8b, 8b + 8) per axis, corner samples
// [8b, 8b + 8]) reads sample tiles b and b + 1 on each axis, so tile i is
// needed when one of the bricks i - 1 and i (per axis) is kept: one ballot is
// the block's mask, stored whole (no atomics).  The lane whose tile index is
// a brick index counts that brick when it is kept (counter[0]).
__global__ void __launch_bounds__(256) k_tiles_from_bricks(const uint8_t* flags, uint32_t cbx, uint32_t cby,
	uint32_t cbz, uint32_t tx, uint32_t ty, uint32_t tz, uint32_t bx, uint32_t by, uint32_t bz,
	unsigned long long* mask, unsigned long long* counter)
{
	const uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
	const uint32_t lane = threadIdx.x & 63;
	if (w >= uint64_t(bx) * by * bz) return;  // whole waves
	const uint32_t bxi = uint32_t(w % bx), byi = uint32_t((w / bx) % by), bzi = uint32_t(w / (uint64_t(bx) * by));
	const uint32_t ix = bxi * 4 + (lane & 3), iy = byi * 4 + ((lane >> 2) & 3), iz = bzi * 4 + (lane >> 4);
	bool need = false, own = false;
	if (ix < tx && iy < ty && iz < tz)
		for (uint32_t dz = 0; dz < 2; ++dz)
			for (uint32_t dy = 0; dy < 2; ++dy)
				for (uint32_t dx = 0; dx < 2; ++dx)
				{
					if (ix < dx || iy < dy || iz < dz) continue;
					const uint32_t x = ix - dx, y = iy - dy, z = iz - dz;
					if (x >= cbx || y >= cby || z >= cbz) continue;
					const bool f = flags[x + uint64_t(cbx) * (y + uint64_t(cby) * z)] != 0;
					need = need || f;
					if (dx + dy + dz == 0) own = f;
				}
	const uint64_t m = __ballot(need), o = __ballot(own);
	if (lane == 0)
	{
		mask[w] = m;
		if (o) atomicAdd(counter, (unsigned long long)__popcll(o));
	}
}

// Thread per block: list the blocks with at least one tile to evaluate.
__global__ void __launch_bounds__(256) k_list_blocks(const unsigned long long* mask, uint32_t nblocks, uint32_t* list,
	unsigned long long* counter)
{
	const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nblocks || mask[b] == 0) return;
	list[atomicAdd(counter, 1ull)] = b;
}

// Thread per tile: list the tiles whose mask bit is set.
__global__ void __launch_bounds__(256) k_list_tiles(const unsigned long long* mask, uint32_t tx, uint32_t ty,
	uint32_t tz, uint32_t bx, uint32_t by, uint32_t* list, unsigned long long* counter)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= uint64_t(tx) * ty * tz) return;
	const uint32_t ix = uint32_t(t % tx), iy = uint32_t((t / tx) % ty), iz = uint32_t(t / (uint64_t(tx) * ty));
	const uint32_t blk = (ix >> 2) + bx * ((iy >> 2) + by * (iz >> 2));
	const uint32_t bit = (ix & 3) + 4 * (iy & 3) + 16 * (iz & 3);
	if ((mask[blk] >> bit) & 1ull) list[atomicAdd(counter, 1ull)] = uint32_t(t);
}

// Thread per active-cube word: mark the bricks holding an active cube.
__global__ void __launch_bounds__(256) k_active_bricks(const uint64_t* active, uint32_t aw, uint32_t sy, uint32_t sz,
	uint32_t cbx, uint32_t cby, uint8_t* out)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= uint64_t(sy) * sz * aw) return;
	const uint64_t act = active[t];
	if (!act) return;
	const uint32_t w = uint32_t(t % aw);
	const uint64_t row = t / aw;
	const uint32_t j = uint32_t(row % sy), k = uint32_t(row / sy);
	for (int b = 0; b < 8; ++b)
		if ((act >> (8 * b)) & 0xffull) out[(w * 8 + b) + uint64_t(cbx) * ((j >> 3) + uint64_t(cby) * (k >> 3))] = 1;
}

unsigned grid_for(uint64_t n) { return unsigned((n + 255) / 256); }

// Level-1 pruning: the program pruned for every 32^3-sample block (cell grid
// with origin g - d/2, cell 32 d; the samples lie within 15.5 |d| of a block
// centre).  m->blocks.valid is false when the program is not prunable.
// prune_blocks_start / prune_blocks_finish are its two halves (see
// prune_cells_start).
struct BlockGrid
{
	uint32_t counts[3];
	float origin[3], cell[3], radius;
};

BlockGrid block_grid(const tg_mesher* m, const tg_grid* g)
{
	const MeshArgs& ma = m->args;
	BlockGrid b{};
	b.counts[0] = (ma.tx + kBlockTiles - 1) / kBlockTiles;
	b.counts[1] = (ma.ty + kBlockTiles - 1) / kBlockTiles;
	b.counts[2] = (ma.tz + kBlockTiles - 1) / kBlockTiles;
	b.origin[0] = g->x - 0.5f * g->dx; b.origin[1] = g->y - 0.5f * g->dy; b.origin[2] = g->z - 0.5f * g->dz;
	b.cell[0] = float(kBlockSamples) * g->dx; b.cell[1] = float(kBlockSamples) * g->dy;
	b.cell[2] = float(kBlockSamples) * g->dz;
	const float half = 0.5f * float(kBlockSamples - 1);
	b.radius = half * std::sqrt(g->dx * g->dx + g->dy * g->dy + g->dz * g->dz);
	return b;
}

int prune_blocks_start(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags)
{
	const BlockGrid b = block_grid(m, g);
	return prune_cells_start(m->stream, p, b.origin, b.cell, b.radius, b.counts, (flags & TG_FLAG_FAST) != 0, m->blocks);
}

int prune_blocks_finish(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags)
{
	const BlockGrid b = block_grid(m, g);
	return prune_cells_finish(m->stream, p, b.origin, b.cell, b.counts, (flags & TG_FLAG_FAST) != 0, m->blocks);
}

int prune_blocks(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags)
{
	int rc = prune_blocks_start(m, p, g, flags);
	if (rc) return rc;
	return prune_blocks_finish(m, p, g, flags);
}

// Largest |p| over the grid's samples (the corners of its box).
float grid_reach(const tg_grid* g)
{
	const float o[3] = { g->x, g->y, g->z }, d[3] = { g->dx, g->dy, g->dz };
	const uint64_t n[3] = { g->sx, g->sy, g->sz };
	double s = 0.0;
	for (int i = 0; i < 3; ++i)
	{
		const double a = std::fabs(double(o[i])), b = std::fabs(double(o[i]) + double(n[i]) * double(d[i]));
		s += std::max(a, b) * std::max(a, b);
	}
	return float(std::sqrt(s));
}

// Block-kernel arguments of one grid (after prune_blocks).
BlockEvalArgs block_args(tg_mesher* m, const tg_program* p, const tg_grid* g, const uint64_t* tile_mask,
	const uint32_t* block_list, uint32_t n_listed, unsigned long long* work)
{
	const MeshArgs& ma = m->args;
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
	ba.bx = (ma.tx + kBlockTiles - 1) / kBlockTiles;
	ba.by = (ma.ty + kBlockTiles - 1) / kBlockTiles;
	ba.bz = (ma.tz + kBlockTiles - 1) / kBlockTiles;
	ba.tile_slot = ma.tile_slot;
	ba.out_pool = m->pool.as<float>();
	ba.signs = m->signs.as<uint8_t>();
	ba.max_depth = std::max<uint32_t>(1, m->blocks.max_stack);
	ba.slots = block_spill_slots(ba.max_depth, m->points_per_lane);
	ba.prune_tiles = 1;
	ba.work = work;
	ba.whole = p->entries.as<const Entry>();
	ba.n_whole = uint32_t(p->decoded.entries.size());
	ba.n_whole_binary = p->decoded.binary_entries;
	ba.max_binary = m->blocks.max_binary;
	ba.max_len = m->blocks.max_len;
	ba.waves = m->waves;
	ba.min_prune_len = m->min_prune_len;
	ba.slack = rounding_slack(p->decoded, grid_reach(g));
	return ba;
}

// The block evaluation's work counters, copied back into the pinned words 4..7
// of m->tail.
unsigned long long* work_readback(tg_mesher* m) { return reinterpret_cast<unsigned long long*>(m->tail + 4); }

// The persistent launch's queue: the n blocks (all, or those of `list`) by
// their program's cost, most expensive first, into m->order.
int block_order(tg_mesher* m, const BlockDesc* desc, const uint32_t* list, uint32_t n)
{
	size_t tmp = 0;
	TG_CHECK_HIP(blocks_by_cost(desc, list, n, nullptr, nullptr, nullptr, nullptr, tmp, m->stream), "sort sizing");
	int rc = m->order_keys.ensure(std::max<size_t>(n, 1) * 8);
	if (!rc) rc = m->order_values.ensure(std::max<size_t>(n, 1) * 4);
	if (!rc) rc = m->order.ensure(std::max<size_t>(n, 1) * 4);
	if (!rc) rc = m->order_tmp.ensure(std::max<size_t>(tmp, 1));
	if (rc) return rc;
	tmp = m->order_tmp.bytes;
	TG_CHECK_HIP(blocks_by_cost(desc, list, n, m->order_keys.as<uint32_t>(), m->order_values.as<uint32_t>(),
		m->order.as<uint32_t>(), m->order_tmp.ptr, tmp, m->stream), "block order");
	return TG_OK;
}

// Pruned evaluation (k_eval_blocks) of the listed blocks' masked tiles, or of
// every tile.  Needs m->blocks.valid.  `sorted`: m->order already holds the
// queue of every block (block_order, enqueued with the pruning).
// Asynchronous: once the caller has synchronised the stream, eval_blocks_read
// gives the work counters and the launch's duration.
int eval_blocks(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags, const uint64_t* tile_mask,
	const uint32_t* block_list, uint32_t n_listed, bool sorted = false)
{
	int rc = m->counters.ensure(64);
	if (rc) return rc;
	unsigned long long* work = m->counters.as<unsigned long long>();
	TG_CHECK_HIP(hipMemsetAsync(m->counters.ptr, 0, 64, m->stream), "counter reset");
	BlockEvalArgs ba = block_args(m, p, g, tile_mask, block_list, n_listed, work);
	// culled exports (a block list) have their own launch shape
	const bool culled = block_list != nullptr;
	int pts = culled ? m->culled_points_per_lane : m->points_per_lane;
	const int shape_waves = culled ? m->culled_waves : m->waves;
	const int shape_wgs = culled ? m->culled_workgroups_per_cu : m->workgroups_per_cu;
	ba.waves = shape_waves;
	if (shape_wgs > 0)
	{
		// Persistent launch: workgroups take blocks from a queue; each keeps its
		// block's decision tables and tile programs in its own slice of global
		// scratch, LDS holds only tile descriptors and the spill region.
		const size_t per_wg = kMaxLds / size_t(shape_wgs);
		const int waves = shape_waves;
		while (pts > 2 && block_lds_bytes_global(block_spill_slots(ba.max_depth, pts), pts, waves) > per_wg) pts >>= 1;
		if (block_lds_bytes_global(block_spill_slots(ba.max_depth, pts), pts, waves) > per_wg)
			return fail(TG_ERR_LIMIT, "block evaluation: spill region too large for the workgroups per CU");
		ba.slots = block_spill_slots(ba.max_depth, pts, (per_wg - kTileInfoBytes) / size_t(waves));
		ba.waves = waves;
		BatchArgs q{};
		q.ntrees = 1;
		q.items = nullptr;  // item k: block k of the one tree
		q.total = ba.nblocks;
		q.slots = ba.slots;
		q.max_binary = ba.max_binary;
		q.max_len = ba.max_len;
		q.max_depth = ba.max_depth;
		q.table_words = block_global_words(ba.max_binary, ba.max_len);
		const uint32_t wgs = uint32_t(std::max(1, m->cus * shape_wgs));
		rc = m->persist_args.ensure(sizeof(BlockEvalArgs));
		if (!rc) rc = m->persist_head.ensure(256);
		if (!rc) rc = m->persist_tables.ensure(size_t(wgs) * q.table_words * 4);
		if (rc) return rc;
		q.trees = m->persist_args.as<const BlockEvalArgs>();
		q.head = m->persist_head.as<unsigned int>();
		q.tables = m->persist_tables.as<uint32_t>();
		// the queue takes the most expensive block programs first, so that its
		// tail holds short ones
		if (!sorted)
		{
			rc = block_order(m, ba.desc, ba.block_list, ba.nblocks);
			if (rc) return rc;
		}
		ba.block_list = m->order.as<const uint32_t>();
		TG_CHECK_HIP(hipMemcpyAsync(m->persist_args.ptr, &ba, sizeof(ba), hipMemcpyHostToDevice, m->stream), "upload");
		m->last_points_per_lane = pts;
		m->last_waves = waves;
		m->last_workgroups = int(wgs);
		m->last_max_len = ba.max_len;
		m->last_max_binary = ba.max_binary;
		m->last_lds = uint32_t(block_lds_bytes_global(ba.slots, pts, waves));
		const bool fast = (flags & TG_FLAG_FAST) != 0;
		hipError_t e = fast ? fast::launch_batch_reset(q.head, nullptr, 0, m->stream)
		                    : exact::launch_batch_reset(q.head, nullptr, 0, m->stream);
		if (e != hipSuccess) return hip_fail(e, "queue reset");
		TG_CHECK_HIP(hipEventRecord(m->ev[4], m->stream), "event");
		e = fast ? fast::launch_batch(q, pts, uint32_t(waves), wgs, m->stream)
		         : exact::launch_batch(q, pts, uint32_t(waves), wgs, m->stream);
		if (e != hipSuccess) return hip_fail(e, "block evaluation launch");
		TG_CHECK_HIP(hipEventRecord(m->ev[5], m->stream), "event");
		TG_CHECK_HIP(hipMemcpyAsync(work_readback(m), work, 16, hipMemcpyDeviceToHost, m->stream), "work read");
		return TG_OK;
	}
	// Fit the LDS: fewer points per lane first, then fewer waves.
	auto lds_for = [&](int p, int w) {
		return block_lds_bytes(block_spill_slots(ba.max_depth, p), p, w, ba.max_binary, ba.max_len);
	};
	while (pts > 2 && lds_for(pts, ba.waves) > kMaxLds) pts >>= 1;
	while (ba.waves > 1 && lds_for(pts, ba.waves) > kMaxLds) ba.waves >>= 1;
	ba.slots = block_spill_slots(ba.max_depth, pts);
	if (block_lds_bytes(ba.slots, pts, ba.waves, ba.max_binary, ba.max_len) > kMaxLds)
	{
		// Block programs too long for per-tile lists in LDS: interpret them whole.
		ba.prune_tiles = 0;
		ba.max_binary = ba.max_len = 0;
	}
	m->last_points_per_lane = pts;
	m->last_waves = ba.waves;
	m->last_workgroups = 0;
	m->last_max_len = ba.max_len;
	m->last_max_binary = ba.max_binary;
	m->last_lds = uint32_t(block_lds_bytes(ba.slots, pts, ba.waves, ba.max_binary, ba.max_len));
	TG_CHECK_HIP(hipEventRecord(m->ev[4], m->stream), "event");
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_eval_blocks(ba, pts, m->stream)
	                                      : exact::launch_eval_blocks(ba, pts, m->stream);
	if (e != hipSuccess) return hip_fail(e, "block evaluation launch");
	TG_CHECK_HIP(hipEventRecord(m->ev[5], m->stream), "event");
	TG_CHECK_HIP(hipMemcpyAsync(work_readback(m), work, 16, hipMemcpyDeviceToHost, m->stream), "work read");
	return TG_OK;
}

// Results of the last eval_blocks, after the stream has been synchronised.
int eval_blocks_read(tg_mesher* m, uint64_t* evals, double* flops)
{
	float kms = 0;
	TG_CHECK_HIP(hipEventElapsedTime(&kms, m->ev[4], m->ev[5]), "timing");
	m->last_block_ms = kms;
	*evals = work_readback(m)[0];
	*flops = double(work_readback(m)[1]);
	return TG_OK;
}

}  // namespace

void destroy_mesher(tg_mesher* m);

extern "C" int tg_mesher_create(tg_mesher** out)
{
	if (!out) return fail(TG_ERR_INVALID, "tg_mesher_create: null argument");
	int rc = ensure_device();
	if (rc) return rc;
	auto* m = new tg_mesher;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e == hipSuccess) e = hipDeviceGetAttribute(&m->cus, hipDeviceAttributeMultiprocessorCount, dev);
	if (e == hipSuccess) e = hipStreamCreate(&m->stream);
	for (int i = 0; e == hipSuccess && i < 6; ++i) e = hipEventCreate(&m->ev[i]);
	if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->tail), 32, hipHostMallocDefault);
	if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&m->tail_dev), m->tail, 0);
	if (e != hipSuccess)
	{
		m->tail = nullptr;
		destroy_mesher(m);
		return hip_fail(e, "mesher stream/events");
	}
	*out = m;
	return TG_OK;
}

namespace {

// Refuses calls that re-lay out a batch-owned mesher's buffers.
int refuse_batch_owned(const tg_mesher* m, const char* what)
{
	if (m && m->batch_owned)
		return fail(TG_ERR_INVALID, std::string(what) + ": the mesher belongs to a tg_batch (read its results only)");
	return TG_OK;
}

}  // namespace

void destroy_mesher(tg_mesher* m)
{
	if (m->stream) (void)hipStreamSynchronize(m->stream);
	if (m->tail) (void)hipHostFree(m->tail);
	for (DevBuf* b : { &m->pool, &m->signs, &m->active, &m->word_count, &m->word_base, &m->quad_count, &m->quad_base,
	                   &m->vertices, &m->triangles, &m->scan_tmp, &m->dense, &m->counters, &m->brick_flags,
	                   &m->brick_active, &m->tile_mask, &m->block_list, &m->tile_list, &m->cubes, &m->normals,
	                   &m->exchange, &m->persist_args, &m->persist_head, &m->persist_tables, &m->order_keys,
	                   &m->order_values, &m->order, &m->order_tmp })
		b->release();
	m->blocks.release();
	for (hipEvent_t e : m->ev)
		if (e) (void)hipEventDestroy(e);
	if (m->stream) (void)hipStreamDestroy(m->stream);
	delete m;
}

extern "C" void tg_mesher_destroy(tg_mesher* m)
{
	if (!m || refuse_batch_owned(m, "tg_mesher_destroy")) return;
	destroy_mesher(m);
}

namespace {


// Brick selection of a culled export: brick flags from the surrogate (or none)
// plus, with the certificate, every brick whose exact centre value allows a
// sign change; then per-block tile masks and the list of blocks (or, for the
// unpruned evaluator, tiles) to evaluate.
struct Selection
{
	uint32_t listed = 0;
	uint64_t flagged = 0;      // kept bricks
	uint64_t certified = 0;    // bricks the certificate added
	float threshold = 0.0f;
	double cert_flops = 0.0;
};

int select_tiles(tg_mesher* m, const tg_program* p, const tg_grid* g, const tg_surrogate* s, bool certify, bool tiles,
	uint32_t flags, Selection& sel)
{
	MeshArgs& ma = m->args;
	const uint32_t bx = (ma.tx + kBlockTiles - 1) / kBlockTiles, by = (ma.ty + kBlockTiles - 1) / kBlockTiles,
	               bz = (ma.tz + kBlockTiles - 1) / kBlockTiles;
	const uint64_t nblocks = uint64_t(bx) * by * bz;
	const uint64_t ntiles = uint64_t(ma.tx) * ma.ty * ma.tz;
	const uint64_t nbricks = uint64_t(ma.cbx) * ma.cby * ma.cbz;
	int rc = m->brick_flags.ensure(nbricks);
	if (!rc) rc = m->tile_mask.ensure(nblocks * 8);
	if (!rc) rc = tiles ? m->tile_list.ensure(ntiles * 4) : m->block_list.ensure(nblocks * 4);
	if (!rc) rc = m->counters.ensure(64);
	if (rc) return rc;
	unsigned long long* ctr = m->counters.as<unsigned long long>();
	TG_CHECK_HIP(hipMemsetAsync(m->counters.ptr, 0, 64, m->stream), "counter reset");
	TG_CHECK_HIP(hipMemsetAsync(m->tile_mask.ptr, 0, nblocks * 8, m->stream), "mask reset");
	const float radius = 4.0f * std::sqrt(g->dx * g->dx + g->dy * g->dy + g->dz * g->dz);
	sel.threshold = 0.0f;
	if (!s)
		TG_CHECK_HIP(hipMemsetAsync(m->brick_flags.ptr, 0, nbricks, m->stream), "flag reset");
	else
	{
		mlp::PredictArgs pa{};
		pa.net = surrogate_net(s);
		pa.xyz = nullptr;
		pa.count = nbricks;
		pa.out = nullptr;
		pa.gx = g->x; pa.gy = g->y; pa.gz = g->z;
		pa.dx = g->dx; pa.dy = g->dy; pa.dz = g->dz;
		pa.cbx = ma.cbx; pa.cby = ma.cby; pa.cbz = ma.cbz;
		pa.radius = radius;
		pa.threshold = surrogate_threshold(s, radius, certify);
		pa.flags = m->brick_flags.as<uint8_t>();
		sel.threshold = pa.threshold;
		if (!(pa.threshold < pa.net.clamp))
		{
			// The clamped fit cannot certify a brick this large: keep them all.
			TG_CHECK_HIP(hipMemsetAsync(m->brick_flags.ptr, 1, nbricks, m->stream), "flag fill");
		}
		else
			TG_CHECK_HIP(mlp::launch_predict(pa, m->stream), "brick prediction launch");
	}
	if (certify)
	{
		const bool blocks = m->blocks.valid;
		CertifyArgs ca{};
		ca.prog = static_cast<const Instr*>(p->device.ptr);
		ca.pool = blocks ? m->blocks.pool.as<const Entry>() : nullptr;
		ca.desc = blocks ? m->blocks.desc.as<const BlockDesc>() : nullptr;
		ca.whole = p->entries.as<const Entry>();
		ca.n_whole = uint32_t(p->decoded.entries.size());
		ca.whole_flops = uint32_t(p->decoded.flops);
		ca.gx = g->x; ca.gy = g->y; ca.gz = g->z;
		ca.dx = g->dx; ca.dy = g->dy; ca.dz = g->dz;
		ca.bx = bx; ca.by = by; ca.bz = bz;
		ca.cbx = ma.cbx; ca.cby = ma.cby; ca.cbz = ma.cbz;
		ca.radius = radius * (1.0f + 1.0e-5f) + rounding_slack(p->decoded, grid_reach(g));
		ca.flags = m->brick_flags.as<uint8_t>();
		ca.counters = ctr + 2;
		ca.slots = std::max<uint32_t>(1, std::max(blocks ? m->blocks.max_stack : 0u, p->decoded.max_stack) - 1);
		hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_certify(ca, m->stream) : exact::launch_certify(ca, m->stream);
		if (e != hipSuccess) return hip_fail(e, "brick certificate launch");
	}
	hipLaunchKernelGGL(k_tiles_from_bricks, dim3(grid_for(uint64_t(nblocks) * 64)), dim3(256), 0, m->stream,
		m->brick_flags.as<const uint8_t>(), ma.cbx, ma.cby, ma.cbz, ma.tx, ma.ty, ma.tz, bx, by, bz,
		m->tile_mask.as<unsigned long long>(), ctr);
	TG_CHECK_HIP(hipGetLastError(), "tile selection launch");
	if (tiles)
		hipLaunchKernelGGL(k_list_tiles, dim3(grid_for(ntiles)), dim3(256), 0, m->stream,
			m->tile_mask.as<const unsigned long long>(), ma.tx, ma.ty, ma.tz, bx, by, m->tile_list.as<uint32_t>(), ctr + 1);
	else
		hipLaunchKernelGGL(k_list_blocks, dim3(grid_for(nblocks)), dim3(256), 0, m->stream,
			m->tile_mask.as<const unsigned long long>(), uint32_t(nblocks), m->block_list.as<uint32_t>(), ctr + 1);
	TG_CHECK_HIP(hipGetLastError(), "list launch");
	unsigned long long c[4] = { 0, 0, 0, 0 };
	TG_CHECK_HIP(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, m->stream), "count read");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "tile selection");
	sel.flagged = c[0];
	sel.listed = uint32_t(c[1]);
	sel.certified = c[2];
	sel.cert_flops = double(c[3]);
	m->nbricks = nbricks;
	ma.brick_flags = m->brick_flags.as<const uint8_t>();
	return TG_OK;
}

int run_export(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags, bool culled,
	const tg_surrogate* sur)
{
	int rc = validate_grid(g);
	if (rc) return rc;
	const bool certify = culled && (!sur || (flags & TG_FLAG_CERTIFY));
	if (culled)
	{
		if (!p->decoded.prunable())
			return fail(TG_ERR_PROGRAM, "culled export needs a prunable program (1-Lipschitz: no ellipsoid, no "
			                            "stretching scale; at most 128 nested blend / flate / field-scale operators)");
		if (sur && !(surrogate_threshold(sur, 0.0f) >= 0.0f))
			return fail(TG_ERR_INVALID, "culled export: surrogate is not calibrated");
	}
	m->have_grid = false;
	m->have_normals = false;
	m->nbricks = 0;
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
	ga.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);

	TG_CHECK_HIP(hipEventRecord(m->ev[0], m->stream), "event");
	m->blocks.valid = false;
	bool sorted = false;
	if (!(flags & TG_FLAG_NO_PRUNE) && p->decoded.prunable())
	{
		rc = prune_blocks_start(m, p, g, flags);
		if (rc) return rc;
		// The persistent launch's queue of every block, enqueued behind the
		// pruning so that the host issues it while the pruning runs; a rerun
		// of the emission changes the programs, and the queue is sorted again.
		if (!culled && m->workgroups_per_cu > 0 && m->blocks.state.pending)
		{
			rc = block_order(m, m->blocks.desc.as<const BlockDesc>(), nullptr, uint32_t(m->blocks.state.ncells));
			if (rc) return rc;
			sorted = true;
		}
		rc = prune_blocks_finish(m, p, g, flags);
		if (rc) return rc;
		sorted = sorted && !m->blocks.reran;
	}
	const bool blocks = m->blocks.valid;
	Selection sel;
	if (culled)
	{
		rc = select_tiles(m, p, g, sur, certify, !blocks, flags, sel);
		if (rc) return rc;
	}
	TG_CHECK_HIP(hipEventRecord(m->ev[3], m->stream), "event");
	uint64_t evals = 0;
	double flops = 0.0;
	if (blocks)
	{
		rc = culled ? eval_blocks(m, p, g, flags, m->tile_mask.as<const uint64_t>(), m->block_list.as<const uint32_t>(),
		                          sel.listed)
		            : eval_blocks(m, p, g, flags, nullptr, nullptr, 0, sorted);
		if (rc) return rc;
	}
	else
	{
		if (culled)
		{
			ga.tile_list = m->tile_list.as<const uint32_t>();
			ga.ntiles = sel.listed;
		}
		evals = uint64_t(ga.ntiles) * kTileSamples;
		// work over grid samples (all of them, or the listed tiles' lanes)
		flops = double(culled ? evals : uint64_t(ma.nx) * ma.ny * ma.nz) * p->decoded.flops;
		hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_tiles(ga, 8, m->stream)
		                                      : exact::launch_tiles(ga, 8, m->stream);
		if (e != hipSuccess) return hip_fail(e, "tile evaluation launch");
	}
	TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
	rc = surface_nets(m);
	if (rc) return rc;
	TG_CHECK_HIP(hipEventRecord(m->ev[2], m->stream), "event");
	TG_CHECK_HIP(hipEventSynchronize(m->ev[2]), "export");
	if (blocks)
	{
		rc = eval_blocks_read(m, &evals, &flops);
		if (rc) return rc;
	}
	float t_pre = 0, t_eval = 0, t_mesh = 0, t_all = 0;
	TG_CHECK_HIP(hipEventElapsedTime(&t_pre, m->ev[0], m->ev[3]), "timing");
	TG_CHECK_HIP(hipEventElapsedTime(&t_eval, m->ev[3], m->ev[1]), "timing");
	TG_CHECK_HIP(hipEventElapsedTime(&t_mesh, m->ev[1], m->ev[2]), "timing");
	TG_CHECK_HIP(hipEventElapsedTime(&t_all, m->ev[0], m->ev[2]), "timing");
	tg_export_stats& s = m->stats;
	s = tg_export_stats{};
	s.eval_ms = t_pre + t_eval;
	s.mesh_ms = t_mesh;
	s.total_ms = t_all;
	s.samples = uint64_t(ma.nx) * ma.ny * ma.nz;
	s.tiles_evaluated = evals / kTileSamples;
	s.exact_evals = evals;
	s.flops = flops;
	s.prune_ms = t_pre;
	s.block_ms = blocks ? m->last_block_ms : 0.0;
	s.points_per_lane = blocks ? uint32_t(m->last_points_per_lane) : 0u;
	s.waves = blocks ? uint32_t(m->last_waves) : 0u;
	s.block_max_len = blocks ? m->last_max_len : 0u;
	s.block_max_binary = blocks ? m->last_max_binary : 0u;
	s.block_lds_bytes = blocks ? m->last_lds : 0u;
	s.block_workgroups = blocks ? uint32_t(m->last_workgroups) : 0u;
	if (culled)
	{
		s.bricks_total = m->nbricks;
		s.bricks_flagged = sel.flagged;
		s.bricks_certified = sel.certified;
		s.cull_threshold = sel.threshold;
		s.flops += sel.cert_flops;
	}
	m->have_grid = !culled;
	return TG_OK;
}

}  // namespace

const float* tg::mesher_vertices(const tg_mesher* m) { return m->vertices.as<const float>(); }

extern "C" int tg_mesher_export(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags)
{
	if (!m || !p) return fail(TG_ERR_INVALID, "tg_mesher_export: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_export")) return rc;
	return run_export(m, p, g, flags, false, nullptr);
}

extern "C" int tg_mesher_export_culled(tg_mesher* m, const tg_program* p, const tg_grid* g, const tg_surrogate* s,
	uint32_t flags)
{
	if (!m || !p) return fail(TG_ERR_INVALID, "tg_mesher_export_culled: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_export_culled")) return rc;
	return run_export(m, p, g, flags, true, s);
}

extern "C" int tg_mesher_brick_flags(const tg_mesher* m, uint8_t* out, uint64_t capacity, uint64_t* count)
{
	if (!m || !count) return fail(TG_ERR_INVALID, "tg_mesher_brick_flags: null argument");
	*count = m->nbricks;
	if (out && capacity && m->nbricks)
		TG_CHECK_HIP(hipMemcpy(out, m->brick_flags.ptr, std::min(capacity, m->nbricks), hipMemcpyDeviceToHost),
			"flag download");
	return TG_OK;
}

extern "C" int tg_mesher_active_bricks(tg_mesher* m, uint8_t* out, uint64_t capacity, uint64_t* count)
{
	if (!m || !count) return fail(TG_ERR_INVALID, "tg_mesher_active_bricks: null argument");
	const MeshArgs& a = m->args;
	if (!a.active || a.sx == 0) return fail(TG_ERR_INVALID, "tg_mesher_active_bricks: nothing meshed yet");
	const uint64_t n = uint64_t(a.cbx) * a.cby * a.cbz;
	*count = n;
	if (!out || !capacity) return TG_OK;
	int rc = m->brick_active.ensure(n);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemsetAsync(m->brick_active.ptr, 0, n, m->stream), "reset");
	const uint64_t words = uint64_t(a.sy) * a.sz * a.aw;
	hipLaunchKernelGGL(k_active_bricks, dim3(grid_for(words)), dim3(256), 0, m->stream, a.active, a.aw, a.sy, a.sz,
		a.cbx, a.cby, m->brick_active.as<uint8_t>());
	TG_CHECK_HIP(hipGetLastError(), "active brick launch");
	TG_CHECK_HIP(hipMemcpyAsync(out, m->brick_active.ptr, std::min(capacity, n), hipMemcpyDeviceToHost, m->stream),
		"download");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "active bricks");
	return TG_OK;
}

extern "C" int tg_mesher_set_tuning(tg_mesher* m, uint32_t points_per_lane, uint32_t waves,
	uint32_t workgroups_per_cu, uint32_t min_prune_len)
{
	if (!m || !(points_per_lane == 2 || points_per_lane == 4 || points_per_lane == 8) ||
		!(waves == 1 || waves == 2 || waves == 4) || workgroups_per_cu > 16)
		return fail(TG_ERR_INVALID,
			"tg_mesher_set_tuning: points per lane in {2,4,8}, waves in {1,2,4}, workgroups per CU 0..16");
	if (int rc = refuse_batch_owned(m, "tg_mesher_set_tuning")) return rc;
	m->points_per_lane = int(points_per_lane);
	m->waves = int(waves);
	m->workgroups_per_cu = int(workgroups_per_cu);
	m->min_prune_len = min_prune_len;
	return TG_OK;
}

extern "C" int tg_mesher_set_emission(tg_mesher* m, uint32_t mode)
{
	if (!m) return fail(TG_ERR_INVALID, "tg_mesher_set_emission: null mesher");
	if (mode > 2) return fail(TG_ERR_INVALID, "tg_mesher_set_emission: mode 0 (automatic), 1 (walk) or 2 (per block)");
	if (int rc = refuse_batch_owned(m, "tg_mesher_set_emission")) return rc;
	m->blocks.emission = int(mode);
	return TG_OK;
}

extern "C" int tg_mesher_set_culled_tuning(tg_mesher* m, uint32_t points_per_lane, uint32_t waves,
	uint32_t workgroups_per_cu)
{
	if (!m || !(points_per_lane == 2 || points_per_lane == 4 || points_per_lane == 8) ||
		!(waves == 1 || waves == 2 || waves == 4) || workgroups_per_cu > 16)
		return fail(TG_ERR_INVALID,
			"tg_mesher_set_culled_tuning: points per lane in {2,4,8}, waves in {1,2,4}, workgroups per CU 0..16");
	if (int rc = refuse_batch_owned(m, "tg_mesher_set_culled_tuning")) return rc;
	m->culled_points_per_lane = int(points_per_lane);
	m->culled_waves = int(waves);
	m->culled_workgroups_per_cu = int(workgroups_per_cu);
	return TG_OK;
}

extern "C" int tg_mesher_autotune(tg_mesher* m, const tg_program* p, const tg_grid* g, uint32_t flags,
	tg_tune_result* out)
{
	if (!m || !p) return fail(TG_ERR_INVALID, "tg_mesher_autotune: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_autotune")) return rc;
	int rc = validate_grid(g);
	if (rc) return rc;
	if ((flags & TG_FLAG_NO_PRUNE) || !p->decoded.prunable())
		return fail(TG_ERR_INVALID, "tg_mesher_autotune: tunes the pruned block evaluator only");
	rc = layout(m, g);
	if (rc) return rc;
	m->have_grid = false;
	rc = prune_blocks(m, p, g, flags);
	if (rc) return rc;
	if (!m->blocks.valid) return fail(TG_ERR_INVALID, "tg_mesher_autotune: program not prunable here");
	tg_tune_result res{};
	// (waves, workgroups per CU): a workgroup per block (0) with LDS tables,
	// or persistent workgroups with global tables
	const int pts_c[3] = { 2, 4, 8 };
	const int shape_c[8][2] = { { 1, 0 }, { 2, 0 }, { 4, 0 }, { 2, 6 }, { 2, 8 }, { 2, 12 }, { 4, 4 }, { 4, 6 } };
	double best = 1e300;
	for (int pi = 0; pi < 3; ++pi)
		for (int wi = 0; wi < 8; ++wi)
		{
			m->points_per_lane = pts_c[pi];
			m->waves = shape_c[wi][0];
			m->workgroups_per_cu = shape_c[wi][1];
			double ms = 0.0;
			for (int rep = 0; rep < 3; ++rep)
			{
				TG_CHECK_HIP(hipEventRecord(m->ev[0], m->stream), "event");
				rc = eval_blocks(m, p, g, flags, nullptr, nullptr, 0);
				if (rc) return rc;
				TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
				TG_CHECK_HIP(hipEventSynchronize(m->ev[1]), "tuning");
				float t = 0;
				TG_CHECK_HIP(hipEventElapsedTime(&t, m->ev[0], m->ev[1]), "timing");
				if (rep > 0) ms += t;  // first run warms caches
			}
			ms *= 0.5;
			const int k = pi * 8 + wi;
			res.candidate_ms[k] = ms;
			// a candidate the LDS could not hold ran as a smaller one: record what ran
			res.candidate_points[k] = uint32_t(m->last_points_per_lane);
			res.candidate_waves[k] = uint32_t(m->last_waves);
			res.candidate_workgroups_per_cu[k] = uint32_t(m->workgroups_per_cu);
			if (ms < best)
			{
				best = ms;
				res.points_per_lane = uint32_t(m->last_points_per_lane);
				res.waves = uint32_t(m->last_waves);
				res.workgroups_per_cu = uint32_t(m->workgroups_per_cu);
			}
		}
	m->points_per_lane = int(res.points_per_lane);
	m->waves = int(res.waves);
	m->workgroups_per_cu = int(res.workgroups_per_cu);
	// then the shortest block program worth pruning per tile, at that launch
	const uint32_t len_c[3] = { 0, 16, 48 };
	res.min_prune_len = 0;
	for (int li = 0; li < 3; ++li)
	{
		m->min_prune_len = len_c[li];
		double ms = 0.0;
		for (int rep = 0; rep < 3; ++rep)
		{
			rc = eval_blocks(m, p, g, flags, nullptr, nullptr, 0);
			if (rc) return rc;
			TG_CHECK_HIP(hipStreamSynchronize(m->stream), "tuning");
			uint64_t evals = 0;
			double flops = 0.0;
			rc = eval_blocks_read(m, &evals, &flops);
			if (rc) return rc;
			if (rep > 0) ms += m->last_block_ms;
		}
		ms *= 0.5;
		res.candidate_len_ms[li] = ms;
		if (li == 0 || ms < best)
		{
			best = ms;
			res.min_prune_len = len_c[li];
		}
	}
	m->min_prune_len = res.min_prune_len;
	res.candidates = 24;
	res.eval_ms = best;
	if (out) *out = res;
	return TG_OK;
}

extern "C" int tg_mesher_mesh_values(tg_mesher* m, const float* values, const tg_grid* g)
{
	if (!m || !values) return fail(TG_ERR_INVALID, "tg_mesher_mesh_values: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_mesh_values")) return rc;
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
	TG_CHECK_HIP(hipEventElapsedTime(&t_mesh, m->ev[1], m->ev[2]), "timing");
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
	if (!m->have_grid)
		return fail(TG_ERR_INVALID, "tg_mesher_copy_grid: no fully evaluated grid (none yet, or the last export was culled)");
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

// ---- STL / PLY writers (tangerine/export.cpp:111-179, 181-317) -----------------

namespace {

// Triangle centroids ((a + b) + c) / 3, the glm order of export.cpp:134.
__global__ void __launch_bounds__(256) k_centroids(const float* v, const uint32_t* tri, uint64_t n, float* out)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const uint32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
	for (int k = 0; k < 3; ++k) out[3 * t + k] = ((v[3 * a + k] + v[3 * b + k]) + v[3 * c + k]) / 3.0f;
}

// Normalised tetrahedral gradients (SDFNode::Gradient, h = 1e-4) of the
// program at `n` device points; result copied to the host.
int gradients_at(tg_mesher* m, const tg_program* p, const float* dev_xyz, uint64_t n, uint32_t flags,
	std::vector<float>& out)
{
	out.assign(size_t(n) * 3, 0.0f);
	if (n == 0) return TG_OK;
	DevBuf dout;
	struct Guard { DevBuf* b; ~Guard() { b->release(); } } guard{ &dout };
	int rc = dout.ensure(n * 12);
	if (rc) return rc;
	GradArgs ga{};
	ga.prog = static_cast<const Instr*>(p->device.ptr);
	ga.xyz = dev_xyz;
	ga.out = dout.as<float>();
	ga.count = n;
	ga.h = 1.0e-4f;
	ga.tetra = 1;
	ga.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_gradient(ga, m->stream) : exact::launch_gradient(ga, m->stream);
	if (e != hipSuccess) return hip_fail(e, "gradient launch");
	TG_CHECK_HIP(hipMemcpyAsync(out.data(), dout.ptr, n * 12, hipMemcpyDeviceToHost, m->stream), "normal download");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "normals");
	return TG_OK;
}

struct File
{
	std::FILE* f = nullptr;
	~File()
	{
		if (f) std::fclose(f);
	}
};

}  // namespace

extern "C" int tg_mesher_write_stl(tg_mesher* m, const tg_program* p, const char* path, float scale, uint32_t flags)
{
	if (!m || !p || !path) return fail(TG_ERR_INVALID, "tg_mesher_write_stl: null argument");
	const uint64_t nt = m->ntris, nv = m->nverts;
	if (nt >= (1ull << 32)) return fail(TG_ERR_LIMIT, "tg_mesher_write_stl: more than 2^32-1 triangles");
	std::vector<float> verts(size_t(nv) * 3), normals;
	std::vector<uint32_t> tris(size_t(nt) * 3);
	int rc = tg_mesher_copy_mesh(m, verts.data(), tris.data());
	if (rc) return rc;
	if (nt)
	{
		DevBuf cen;
		struct Guard { DevBuf* b; ~Guard() { b->release(); } } guard{ &cen };
		rc = cen.ensure(nt * 12);
		if (rc) return rc;
		hipLaunchKernelGGL(k_centroids, dim3(grid_for(nt)), dim3(256), 0, m->stream, m->vertices.as<const float>(),
			m->triangles.as<const uint32_t>(), nt, cen.as<float>());
		TG_CHECK_HIP(hipGetLastError(), "centroid launch");
		rc = gradients_at(m, p, cen.as<const float>(), nt, flags, normals);
		if (rc) return rc;
	}
	for (float& v : verts) v *= scale;
	File out;
	out.f = std::fopen(path, "wb");
	if (!out.f) return fail(TG_ERR_INVALID, std::string("tg_mesher_write_stl: cannot open ") + path);
	char header[80] = {};
	const char* text = "STL generated by TangerineCSG";
	std::memcpy(header, text, std::strlen(text));
	const uint32_t count = uint32_t(nt);
	bool ok = std::fwrite(header, 1, 80, out.f) == 80 && std::fwrite(&count, 4, 1, out.f) == 1;
	std::vector<unsigned char> buf;
	buf.reserve(50 * 65536);
	for (uint64_t t = 0; ok && t < nt; ++t)
	{
		unsigned char rec[50];
		std::memcpy(rec, &normals[3 * t], 12);
		for (int k = 0; k < 3; ++k) std::memcpy(rec + 12 + 12 * k, &verts[3 * size_t(tris[3 * t + k])], 12);
		rec[48] = rec[49] = 0;
		buf.insert(buf.end(), rec, rec + 50);
		if (buf.size() >= 50 * 65536 || t + 1 == nt)
		{
			ok = std::fwrite(buf.data(), 1, buf.size(), out.f) == buf.size();
			buf.clear();
		}
	}
	// The reference pads (84 + 100 t) % 4 bytes, which is always zero.
	if (!ok) return fail(TG_ERR_INVALID, std::string("tg_mesher_write_stl: write failed: ") + path);
	return TG_OK;
}

extern "C" int tg_mesher_write_ply(tg_mesher* m, const tg_program* p, const char* path, float scale, uint32_t flags)
{
	if (!m || !p || !path) return fail(TG_ERR_INVALID, "tg_mesher_write_ply: null argument");
	const uint64_t nt = m->ntris, nv = m->nverts;
	std::vector<float> verts(size_t(nv) * 3), normals;
	std::vector<uint32_t> tris(size_t(nt) * 3);
	int rc = tg_mesher_copy_mesh(m, verts.data(), tris.data());
	if (rc) return rc;
	rc = gradients_at(m, p, m->vertices.as<const float>(), nv, flags, normals);
	if (rc) return rc;
	std::string header = "ply\nformat binary_little_endian 1.0\ncomment Created by Tangerine\nelement vertex " +
	                     std::to_string(nv) +
	                     "\nproperty float x\nproperty float y\nproperty float z\nproperty float nx\nproperty float "
	                     "ny\nproperty float nz\n";
	if (nt > 0) header += "element face " + std::to_string(nt) + "\nproperty list uchar uint vertex_indices\n";
	header += "end_header\n";
	File out;
	out.f = std::fopen(path, "wb");
	if (!out.f) return fail(TG_ERR_INVALID, std::string("tg_mesher_write_ply: cannot open ") + path);
	bool ok = std::fwrite(header.data(), 1, header.size(), out.f) == header.size();
	std::vector<unsigned char> buf;
	buf.reserve(24 * 65536);
	for (uint64_t v = 0; ok && v < nv; ++v)
	{
		float rec[6] = { verts[3 * v] * scale, verts[3 * v + 1] * scale, verts[3 * v + 2] * scale, normals[3 * v],
		                 normals[3 * v + 1], normals[3 * v + 2] };
		const unsigned char* r = reinterpret_cast<const unsigned char*>(rec);
		buf.insert(buf.end(), r, r + 24);
		if (buf.size() >= 24 * 65536 || v + 1 == nv)
		{
			ok = std::fwrite(buf.data(), 1, buf.size(), out.f) == buf.size();
			buf.clear();
		}
	}
	for (uint64_t t = 0; ok && t < nt; ++t)
	{
		unsigned char rec[13];
		rec[0] = 3;
		std::memcpy(rec + 1, &tris[3 * t], 12);
		buf.insert(buf.end(), rec, rec + 13);
		if (buf.size() >= 13 * 65536 || t + 1 == nt)
		{
			ok = std::fwrite(buf.data(), 1, buf.size(), out.f) == buf.size();
			buf.clear();
		}
	}
	if (!ok) return fail(TG_ERR_INVALID, std::string("tg_mesher_write_ply: write failed: ") + path);
	return TG_OK;
}

// ---- Sodapop meshing job (tangerine/sodapop.cpp:156-177, 337-401, 562-897) ---

namespace {

// The grid of NaiveSurfaceNetsScratch: the bounding cube of the model's
// bounds (AABB::BoundingCube, aabb.cpp:122-139), floor(density) samples per
// unit and at least 8 per axis, origin moved down two steps, three more cells.
int sodapop_grid(const float b[6], float density, tg_grid* g)
{
	float mn[3], mx[3], ext[3];
	for (int i = 0; i < 3; ++i)
	{
		mn[i] = b[i];
		mx[i] = b[3 + i];
		if (!std::isfinite(mn[i]) || !std::isfinite(mx[i]) || !(mx[i] > mn[i]))
			return fail(TG_ERR_INVALID, "sodapop export: bounds must be finite and well formed");
		ext[i] = mx[i] - mn[i];
	}
	const float longest = std::max(std::max(ext[0], ext[1]), ext[2]);
	for (int i = 0; i < 3; ++i)
	{
		const float pad = (longest - ext[i]) * 0.5f;
		mn[i] = mn[i] - pad;
		mx[i] = mx[i] + pad;
		ext[i] = mx[i] - mn[i];
	}
	const float d = std::floor(density);
	float o[3], step[3];
	uint64_t cells[3];
	for (int i = 0; i < 3; ++i)
	{
		const float spu = std::max(ext[i] * d, 8.0f);
		cells[i] = uint64_t(std::ceil(spu));
		step[i] = ext[i] / float(cells[i]);
		o[i] = mn[i] - step[i] * 2.0f;
		cells[i] += 3;
	}
	*g = tg_grid{ o[0], o[1], o[2], step[0], step[1], step[2], cells[0], cells[1], cells[2] };
	return TG_OK;
}

// ApplyVertexSequence (sodapop.cpp:337-394) with VertexSequence::Shuffle:
// the first half of the vertices swaps with a shuffled (mt19937 seeded 0) list
// of the second half; `exchange` maps every old index to its new one (an
// involution).
void vertex_sequence(uint64_t n, std::vector<uint32_t>& exchange)
{
	exchange.resize(n);
	for (uint64_t i = 0; i < n; ++i) exchange[i] = uint32_t(i);
	const size_t half = n / 2, mirror = (n % 2 == 0) ? half : half + 1;
	std::vector<size_t> seq;
	seq.reserve(mirror);
	for (size_t k = 0; k < half; ++k) seq.push_back(mirror + k);
	std::mt19937 gen;
	gen.seed(0);
	std::shuffle(seq.begin(), seq.end(), gen);
	if (mirror != half) seq.push_back(half);
	for (size_t t = 0; t < seq.size(); ++t)
	{
		exchange[t] = uint32_t(seq[t]);
		exchange[seq[t]] = uint32_t(t);
	}
}

__global__ void __launch_bounds__(256) k_permute_vertices(const float* src, float* dst, const uint32_t* exchange,
	uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t from = exchange[i];  // an involution: new i holds old exchange[i]
	dst[3 * i] = src[3 * size_t(from)];
	dst[3 * i + 1] = src[3 * size_t(from) + 1];
	dst[3 * i + 2] = src[3 * size_t(from) + 2];
}

__global__ void __launch_bounds__(256) k_remap_indices(uint32_t* tri, const uint32_t* exchange, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) tri[i] = exchange[tri[i]];
}

}  // namespace

extern "C" int tg_mesher_export_sodapop(tg_mesher* m, const tg_program* p, const float bounds[6], float density,
	int shuffle, uint32_t flags)
{
	if (!m || !p || !bounds) return fail(TG_ERR_INVALID, "tg_mesher_export_sodapop: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_export_sodapop")) return rc;
	if (!(density >= 1.0f)) return fail(TG_ERR_INVALID, "tg_mesher_export_sodapop: density must be >= 1");
	tg_grid g{};
	int rc = sodapop_grid(bounds, density, &g);
	if (rc) return rc;
	m->have_normals = false;
	rc = run_export(m, p, &g, flags, false, nullptr);
	if (rc) return rc;
	const uint64_t nv = m->nverts, ni = m->ntris * 3;
	if (shuffle && nv > 1)
	{
		std::vector<uint32_t> ex;
		vertex_sequence(nv, ex);
		rc = m->exchange.ensure(nv * 4);
		if (!rc) rc = m->normals.ensure(nv * 12);  // staging for the permuted positions
		if (rc) return rc;
		TG_CHECK_HIP(hipMemcpyAsync(m->exchange.ptr, ex.data(), nv * 4, hipMemcpyHostToDevice, m->stream), "upload");
		hipLaunchKernelGGL(k_permute_vertices, dim3(grid_for(nv)), dim3(256), 0, m->stream,
			m->vertices.as<const float>(), m->normals.as<float>(), m->exchange.as<const uint32_t>(), nv);
		TG_CHECK_HIP(hipGetLastError(), "permute launch");
		TG_CHECK_HIP(hipMemcpyAsync(m->vertices.ptr, m->normals.ptr, nv * 12, hipMemcpyDeviceToDevice, m->stream),
			"permute copy");
		if (ni)
		{
			hipLaunchKernelGGL(k_remap_indices, dim3(grid_for(ni)), dim3(256), 0, m->stream, m->triangles.as<uint32_t>(),
				m->exchange.as<const uint32_t>(), ni);
			TG_CHECK_HIP(hipGetLastError(), "remap launch");
		}
		TG_CHECK_HIP(hipStreamSynchronize(m->stream), "vertex sequence");
	}
	// normals: SDFOctree::Gradient at every (final) vertex (USE_GRADIENT_NORMALS)
	std::vector<float> normals;
	rc = gradients_at(m, p, m->vertices.as<const float>(), nv, flags, normals);
	if (rc) return rc;
	rc = m->normals.ensure(std::max<uint64_t>(nv, 1) * 12);
	if (rc) return rc;
	if (nv) TG_CHECK_HIP(hipMemcpy(m->normals.ptr, normals.data(), nv * 12, hipMemcpyHostToDevice), "normals");
	m->have_normals = true;
	return TG_OK;
}

extern "C" int tg_mesher_copy_normals(const tg_mesher* m, float* normals)
{
	if (!m || !normals) return fail(TG_ERR_INVALID, "tg_mesher_copy_normals: null argument");
	if (!m->have_normals) return fail(TG_ERR_INVALID, "tg_mesher_copy_normals: no normals (run a sodapop export)");
	if (m->nverts) TG_CHECK_HIP(hipMemcpy(normals, m->normals.ptr, m->nverts * 12, hipMemcpyDeviceToHost), "normals");
	return TG_OK;
}

// ---- point-cloud export (tangerine/export.cpp:384-480) --------------------------

extern "C" int tg_mesher_export_points(tg_mesher* m, const tg_program* p, const float lo[3], const float hi[3],
	const float step[3], int refine_iterations, uint32_t flags)
{
	if (!m || !p || !lo || !hi || !step) return fail(TG_ERR_INVALID, "tg_mesher_export_points: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_export_points")) return rc;
	if (refine_iterations < 0) return fail(TG_ERR_INVALID, "tg_mesher_export_points: negative refine iterations");
	CloudArgs a{};
	uint64_t n[3];
	float half[3];
	for (int i = 0; i < 3; ++i)
	{
		if (!std::isfinite(lo[i]) || !std::isfinite(hi[i]) || !(step[i] > 0.0f) || !std::isfinite(step[i]))
			return fail(TG_ERR_INVALID, "tg_mesher_export_points: bounds and step must be finite, step positive");
		// ivec3(ceil((Stop - Start) / Step)); no cells on an axis with Stop <= Start
		const float c = std::ceil((hi[i] - lo[i]) / step[i]);
		if (!(c < 2147483647.0f)) return fail(TG_ERR_LIMIT, "tg_mesher_export_points: too many cells per axis");
		n[i] = c > 0.0f ? uint64_t(c) : 0;
		half[i] = step[i] / 2.0f;
	}
	const uint64_t cells = n[0] * n[1] * n[2];
	if (cells >= (1ull << 31)) return fail(TG_ERR_LIMIT, "tg_mesher_export_points: more than 2^31-1 cells");
	a.prog = static_cast<const Instr*>(p->device.ptr);
	a.sx = lo[0]; a.sy = lo[1]; a.sz = lo[2];
	a.dx = step[0]; a.dy = step[1]; a.dz = step[2];
	a.hx = half[0]; a.hy = half[1]; a.hz = half[2];
	a.ix = uint32_t(n[0]); a.iy = uint32_t(n[1]); a.iz = uint32_t(n[2]);
	a.cells = cells;
	a.diagonal = std::sqrt((half[0] * half[0] + half[1] * half[1]) + half[2] * half[2]);  // glm length
	a.iterations = refine_iterations;
	a.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	m->have_grid = false;
	m->have_normals = false;
	m->nbricks = 0;
	m->nverts = m->ntris = 0;
	m->stats = tg_export_stats{};
	m->stats.samples = cells;
	const bool fast = (flags & TG_FLAG_FAST) != 0;
	int rc = m->cloud_flags.ensure(std::max<uint64_t>(cells, 1) * 4);
	if (!rc) rc = m->cloud_rank.ensure(std::max<uint64_t>(cells, 1) * 4);
	size_t tmp = 0;
	if (rc) return rc;
	TG_CHECK_HIP(exclusive_scan_u32(nullptr, nullptr, cells, nullptr, tmp, m->stream), "scan sizing");
	rc = m->scan_tmp.ensure(tmp);
	if (rc) return rc;
	a.flags = m->cloud_flags.as<uint32_t>();
	a.rank = m->cloud_rank.as<const uint32_t>();
	TG_CHECK_HIP(hipEventRecord(m->ev[0], m->stream), "event");
	uint64_t np = 0;
	if (cells)
	{
		hipError_t e = fast ? fast::launch_cloud(a, 0, m->stream) : exact::launch_cloud(a, 0, m->stream);
		if (e != hipSuccess) return hip_fail(e, "point select launch");
		tmp = m->scan_tmp.bytes;
		TG_CHECK_HIP(exclusive_scan_u32(a.flags, m->cloud_rank.as<uint32_t>(), cells, m->scan_tmp.ptr, tmp, m->stream),
			"point scan");
		uint32_t tail[2] = { 0, 0 };
		TG_CHECK_HIP(hipMemcpyAsync(&tail[0], a.rank + cells - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
		TG_CHECK_HIP(hipMemcpyAsync(&tail[1], a.flags + cells - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
		TG_CHECK_HIP(hipStreamSynchronize(m->stream), "point count");
		np = uint64_t(tail[0]) + tail[1];
	}
	rc = m->vertices.ensure(std::max<uint64_t>(np, 1) * 12);
	if (rc) return rc;
	a.points = m->vertices.as<float>();
	a.npoints = np;
	if (np)
	{
		hipError_t e = fast ? fast::launch_cloud(a, 1, m->stream) : exact::launch_cloud(a, 1, m->stream);
		if (e == hipSuccess && refine_iterations > 0)
			e = fast ? fast::launch_cloud(a, 2, m->stream) : exact::launch_cloud(a, 2, m->stream);
		if (e != hipSuccess) return hip_fail(e, "point refine launch");
	}
	TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
	TG_CHECK_HIP(hipEventSynchronize(m->ev[1]), "point cloud");
	float ms = 0.0f;
	TG_CHECK_HIP(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]), "event time");
	m->nverts = np;
	m->stats.eval_ms = ms;
	m->stats.total_ms = ms;
	m->stats.exact_evals = cells + np * uint64_t(refine_iterations) * 5;
	return TG_OK;
}

// ---- MagicaVoxel export (tangerine/magica.cpp:27-70) -------------------------

namespace {

// The .vox file that VoxExport's vox::VoxWriter (third_party/voxwriter)
// writes for voxels added in raster order: models ("cubes") of at most
// min(size, 126) voxels per axis, created in the order their first voxel
// comes; a scene graph of a root transform, a group and one transform and
// shape per model (its "_t" translation from the writer's centring formula,
// min-cube tracking included); no palette.  Little-endian chunks.
struct VoxFile
{
	std::vector<unsigned char> b;
	void u32(uint32_t v) { for (int k = 0; k < 4; ++k) b.push_back(uint8_t(v >> (8 * k))); }
	void i32(int32_t v) { u32(uint32_t(v)); }
	void id(const char* s) { b.insert(b.end(), s, s + 4); }
	void str(const std::string& s) { u32(uint32_t(s.size())); b.insert(b.end(), s.begin(), s.end()); }
	void chunk(const char* name, uint32_t content) { id(name); u32(content); u32(0); }
};

int write_vox(const char* path, const uint32_t size[3], const std::vector<uint32_t>& cells, uint8_t color)
{
	int32_t per[3];
	for (int a = 0; a < 3; ++a) per[a] = std::min<int32_t>(std::max<int32_t>(int32_t(size[a]), 0), 126);
	struct Cube
	{
		int32_t o[3];
		std::vector<uint8_t> xyzi;
	};
	std::vector<Cube> cubes;
	std::map<std::array<int32_t, 3>, size_t> index;
	int32_t min_cube[3] = { int32_t(1e7), int32_t(1e7), int32_t(1e7) };
	double lower[3] = { 1e7, 1e7, 1e7 }, upper[3] = { 0.0, 0.0, 0.0 };
	const uint64_t slice = uint64_t(size[0]) * size[1];
	for (const uint32_t c : cells)
	{
		const int32_t v[3] = { int32_t(c % size[0]), int32_t((c % slice) / size[0]), int32_t(c / slice) };
		const std::array<int32_t, 3> o = { int32_t(std::floor(double(v[0]) / double(per[0]))),
			                               int32_t(std::floor(double(v[1]) / double(per[1]))),
			                               int32_t(std::floor(double(v[2]) / double(per[2]))) };
		// VoxWriter::AddVoxel tracks y and z against the x minimum
		min_cube[0] = std::min(min_cube[0], o[0]);
		min_cube[1] = std::min(min_cube[0], o[1]);
		min_cube[2] = std::min(min_cube[0], o[2]);
		auto it = index.find(o);
		if (it == index.end())
		{
			it = index.emplace(o, cubes.size()).first;
			cubes.push_back(Cube{ { o[0], o[1], o[2] }, {} });
		}
		Cube& cube = cubes[it->second];
		for (int a = 0; a < 3; ++a)
		{
			lower[a] = std::min(lower[a], double(v[a]));
			upper[a] = std::max(upper[a], double(v[a]));
		}
		for (int a = 0; a < 3; ++a) cube.xyzi.push_back(uint8_t(v[a] % per[a]));
		cube.xyzi.push_back(color);
	}
	VoxFile f;
	f.id("VOX ");
	f.i32(150);
	f.id("MAIN");
	f.u32(0);
	f.u32(0);  // children bytes, patched below
	const size_t header = f.b.size();
	for (const Cube& c : cubes)
	{
		f.chunk("SIZE", 12);
		for (int a = 0; a < 3; ++a) f.i32(per[a] + 1);
		f.chunk("XYZI", uint32_t(4 + c.xyzi.size()));
		f.u32(uint32_t(c.xyzi.size() / 4));
		f.b.insert(f.b.end(), c.xyzi.begin(), c.xyzi.end());
	}
	const int32_t count = int32_t(cubes.size());
	// root transform (node 0) -> group (node 1)
	f.chunk("nTRN", 28);
	f.i32(0); f.u32(0); f.i32(1); f.i32(-1); f.i32(-1); f.i32(1); f.u32(0);
	f.chunk("nGRP", uint32_t(4 * (2 + count) + 4));
	f.i32(1); f.u32(0); f.i32(count);
	for (int32_t i = 0; i < count; ++i) f.i32(2 + 2 * i);
	for (int32_t i = 0; i < count; ++i)
	{
		const Cube& c = cubes[size_t(i)];
		int32_t t[3];
		for (int a = 0; a < 2; ++a)
		{
			const float cube_centre = (float(c.o[a] - min_cube[a]) + 0.5f) * float(per[a]);
			t[a] = int32_t(std::floor(double(cube_centre) - lower[a] - (upper[a] - lower[a]) * 0.5));
		}
		t[2] = int32_t(std::floor((float(c.o[2] - min_cube[2]) + 0.5f) * float(per[2])));
		const std::string value = std::to_string(t[0]) + " " + std::to_string(t[1]) + " " + std::to_string(t[2]);
		f.chunk("nTRN", uint32_t(20 + 4 + 4 + (4 + 2) + (4 + value.size())));
		f.i32(2 + 2 * i); f.u32(0); f.i32(3 + 2 * i); f.i32(-1); f.i32(0); f.i32(1);
		f.u32(1); f.str("_t"); f.str(value);
		f.chunk("nSHP", 20);
		f.i32(3 + 2 * i); f.u32(0); f.i32(1); f.i32(i); f.u32(0);
	}
	const uint32_t children = uint32_t(f.b.size() - header);
	for (int k = 0; k < 4; ++k) f.b[header - 4 + size_t(k)] = uint8_t(children >> (8 * k));
	File out;
	out.f = std::fopen(path, "wb");
	if (!out.f) return fail(TG_ERR_INVALID, std::string("tg_mesher_export_vox: cannot open ") + path);
	if (std::fwrite(f.b.data(), 1, f.b.size(), out.f) != f.b.size())
		return fail(TG_ERR_INVALID, std::string("tg_mesher_export_vox: write failed: ") + path);
	return TG_OK;
}

}  // namespace

extern "C" int tg_mesher_export_vox(tg_mesher* m, const tg_program* p, const float bounds[6], float grid_size,
	int color_index, const char* path, uint32_t flags, uint64_t* voxel_count)
{
	if (!m || !p || !bounds) return fail(TG_ERR_INVALID, "tg_mesher_export_vox: null argument");
	if (int rc = refuse_batch_owned(m, "tg_mesher_export_vox")) return rc;
	VoxArgs a{};
	for (int i = 0; i < 3; ++i)
	{
		a.lo[i] = bounds[i];
		a.hi[i] = bounds[3 + i];
		if (!std::isfinite(a.lo[i]) || !std::isfinite(a.hi[i]))
			return fail(TG_ERR_INVALID, "tg_mesher_export_vox: bounds must be finite");
		// ivec3(ceil(max - min) * grid_size)
		const float s = std::ceil(a.hi[i] - a.lo[i]) * grid_size;
		if (!(s >= 1.0f)) return fail(TG_ERR_INVALID, "tg_mesher_export_vox: empty voxel grid");
		if (!(s < 65536.0f)) return fail(TG_ERR_LIMIT, "tg_mesher_export_vox: more than 65535 voxels per axis");
		a.size[i] = uint32_t(s);
		a.fs[i] = float(a.size[i]);
	}
	a.cells = uint64_t(a.size[0]) * a.size[1] * a.size[2];
	if (a.cells >= (1ull << 31)) return fail(TG_ERR_LIMIT, "tg_mesher_export_vox: more than 2^31 - 1 cells");
	{
		// glm::distance(min, mix(min, max, 0.5 / size))
		float d2[3];
		for (int i = 0; i < 3; ++i)
		{
			const float t = 0.5f / a.fs[i];
			d2[i] = (a.lo[i] * (1.0f - t) + a.hi[i] * t) - a.lo[i];
		}
		a.radius = std::sqrt((d2[0] * d2[0] + d2[1] * d2[1]) + d2[2] * d2[2]);
	}
	a.prog = static_cast<const Instr*>(p->device.ptr);
	a.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	m->have_grid = false;
	m->have_normals = false;
	m->nbricks = 0;
	m->nverts = m->ntris = 0;
	m->nvox = 0;
	m->stats = tg_export_stats{};
	m->stats.samples = a.cells;
	const bool fast = (flags & TG_FLAG_FAST) != 0;
	int rc = m->cloud_flags.ensure(a.cells * 4);
	if (!rc) rc = m->cloud_rank.ensure(a.cells * 4);
	if (rc) return rc;
	size_t tmp = 0;
	TG_CHECK_HIP(exclusive_scan_u32(nullptr, nullptr, a.cells, nullptr, tmp, m->stream), "scan sizing");
	rc = m->scan_tmp.ensure(tmp);
	if (rc) return rc;
	a.flags = m->cloud_flags.as<uint32_t>();
	a.rank = m->cloud_rank.as<const uint32_t>();
	TG_CHECK_HIP(hipEventRecord(m->ev[0], m->stream), "event");
	hipError_t e = fast ? fast::launch_vox(a, 0, m->stream) : exact::launch_vox(a, 0, m->stream);
	if (e != hipSuccess) return hip_fail(e, "voxel select launch");
	tmp = m->scan_tmp.bytes;
	TG_CHECK_HIP(exclusive_scan_u32(a.flags, m->cloud_rank.as<uint32_t>(), a.cells, m->scan_tmp.ptr, tmp, m->stream),
		"voxel scan");
	uint32_t tail[2] = { 0, 0 };
	TG_CHECK_HIP(hipMemcpyAsync(&tail[0], a.rank + a.cells - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipMemcpyAsync(&tail[1], a.flags + a.cells - 1, 4, hipMemcpyDeviceToHost, m->stream), "count copy");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "voxel count");
	const uint64_t nv = uint64_t(tail[0]) + tail[1];
	rc = m->vox_list.ensure(std::max<uint64_t>(nv, 1) * 4);
	if (rc) return rc;
	a.list = m->vox_list.as<uint32_t>();
	if (nv)
	{
		e = fast ? fast::launch_vox(a, 1, m->stream) : exact::launch_vox(a, 1, m->stream);
		if (e != hipSuccess) return hip_fail(e, "voxel list launch");
	}
	TG_CHECK_HIP(hipEventRecord(m->ev[1], m->stream), "event");
	std::vector<uint32_t> cells(nv);
	if (nv) TG_CHECK_HIP(hipMemcpyAsync(cells.data(), a.list, nv * 4, hipMemcpyDeviceToHost, m->stream), "voxel download");
	TG_CHECK_HIP(hipStreamSynchronize(m->stream), "voxels");
	float ms = 0.0f;
	TG_CHECK_HIP(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]), "event time");
	m->stats.eval_ms = ms;
	m->stats.total_ms = ms;
	m->stats.exact_evals = a.cells;
	m->nvox = nv;
	if (voxel_count) *voxel_count = nv;
	if (path) return write_vox(path, a.size, cells, uint8_t(std::abs(color_index) % 255 + 1));
	return TG_OK;
}

extern "C" int tg_vox_write(const char* path, const uint32_t size[3], const uint32_t* cells, uint64_t count,
	int color_index)
{
	if (!path || !size || (count && !cells)) return fail(TG_ERR_INVALID, "tg_vox_write: null argument");
	const uint64_t total = uint64_t(size[0]) * size[1] * size[2];
	if (size[0] == 0 || size[1] == 0 || size[2] == 0) return fail(TG_ERR_INVALID, "tg_vox_write: empty grid");
	for (uint64_t i = 0; i < count; ++i)
		if (cells[i] >= total) return fail(TG_ERR_INVALID, "tg_vox_write: cell index outside the grid");
	std::vector<uint32_t> list(cells, cells + count);
	return write_vox(path, size, list, uint8_t(std::abs(color_index) % 255 + 1));
}

extern "C" int tg_mesher_copy_vox_cells(const tg_mesher* m, uint32_t* cells)
{
	if (!m || !cells) return fail(TG_ERR_INVALID, "tg_mesher_copy_vox_cells: null argument");
	if (m->nvox) TG_CHECK_HIP(hipMemcpy(cells, m->vox_list.ptr, m->nvox * 4, hipMemcpyDeviceToHost), "voxel download");
	return TG_OK;
}

extern "C" int tg_mesher_grid_get(const tg_mesher* m, tg_grid* out)
{
	if (!m || !out) return fail(TG_ERR_INVALID, "tg_mesher_grid_get: null argument");
	const MeshArgs& a = m->args;
	*out = tg_grid{ a.gx, a.gy, a.gz, a.dx, a.dy, a.dz, a.sx, a.sy, a.sz };
	return TG_OK;
}

// ---- batched export: many trees, one persistent launch -----------------------
//
// Every tree keeps its own mesher (grid layout, pruned block programs, sample
// tiles, mesh).  The evaluation of all trees' blocks is one persistent launch
// of k_eval_batch over a device queue of (tree, block) items, most expensive
// block program first (so that the tail of the queue holds short items); the
// launch and the reset of the queue head and work counters are captured in a
// hipGraph, replayed by tg_batch_export and tg_batch_replay.  Each workgroup
// keeps its block's decision tables and tile programs in its own slice of
// global scratch (LDS holds only the tile descriptors and the spill region),
// so programs of any length run at the same occupancy.

struct tg_batch
{
	hipStream_t stream = nullptr;
	hipEvent_t ev[2] = {};
	std::vector<tg_mesher*> meshers;
	std::vector<const tg_program*> programs;
	std::vector<tg_grid> grids;
	DevBuf trees, items, head, work;
	DevBuf tables;                // per-workgroup decision tables and tile pools
	PruneLaunches prune;          // every tree's level pruning and emission as shared launches
	DevBuf queue_sort;            // the queue's keys, values, tree offsets and sort storage
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	BatchArgs args = {};
	bool fast = false;
	int points_per_lane = 4;      // launch shape (tg_batch_set_tuning / autotune)
	int waves = 4;
	int workgroups_per_cu = 4;
	int cus = 0;
	int last_points_per_lane = 0, last_waves = 0, last_workgroups = 0;
	tg_batch_stats stats{};
};

namespace {

void drop_graph(tg_batch* b)
{
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
}

// Launch shape for (points, waves, workgroups per CU): the spill region gets
// all the LDS the workgroups per CU leave beside the tile descriptors; each
// workgroup gets its slice of the global tables.
int shape_batch(tg_batch* b, int points, int waves, int wgs_per_cu)
{
	BatchArgs& ba = b->args;
	int pts = points;
	const size_t per_wg = kMaxLds / size_t(std::max(1, wgs_per_cu));
	while (pts > 2 && block_lds_bytes_global(block_spill_slots(ba.max_depth, pts), pts, waves) > per_wg) pts >>= 1;
	if (block_lds_bytes_global(block_spill_slots(ba.max_depth, pts), pts, waves) > per_wg)
		return fail(TG_ERR_LIMIT, "batch: spill region too large for the workgroups per CU");
	ba.slots = block_spill_slots(ba.max_depth, pts, (per_wg - kTileInfoBytes) / size_t(waves));
	ba.pool_words = 0;
	ba.table_words = block_global_words(ba.max_binary, ba.max_len);
	const uint32_t wgs = uint32_t(std::max(1, b->cus * wgs_per_cu));
	int rc = b->tables.ensure(size_t(wgs) * ba.table_words * 4);
	if (rc) return rc;
	ba.tables = b->tables.as<uint32_t>();
	b->last_points_per_lane = pts;
	b->last_waves = waves;
	b->last_workgroups = int(wgs);
	return TG_OK;
}

// Captures the queue reset and the persistent launch with the current tuning.
int capture_batch(tg_batch* b)
{
	drop_graph(b);
	int rc = shape_batch(b, b->points_per_lane, b->waves, b->workgroups_per_cu);
	if (rc) return rc;
	const uint32_t nwork = uint32_t(2 * b->meshers.size());
	TG_CHECK_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "capture begin");
	hipError_t e = b->fast ? fast::launch_batch_reset(b->args.head, b->work.as<unsigned long long>(), nwork, b->stream)
	                       : exact::launch_batch_reset(b->args.head, b->work.as<unsigned long long>(), nwork, b->stream);
	if (e == hipSuccess)
		e = b->fast ? fast::launch_batch(b->args, b->last_points_per_lane, uint32_t(b->last_waves),
		                                 uint32_t(b->last_workgroups), b->stream)
		            : exact::launch_batch(b->args, b->last_points_per_lane, uint32_t(b->last_waves),
		                                  uint32_t(b->last_workgroups), b->stream);
	hipGraph_t g = nullptr;
	const hipError_t ee = hipStreamEndCapture(b->stream, &g);
	if (e != hipSuccess)
	{
		if (g) (void)hipGraphDestroy(g);
		return hip_fail(e, "batch capture");
	}
	if (ee != hipSuccess) return hip_fail(ee, "batch capture end");
	b->graph = g;
	TG_CHECK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "batch graph instantiate");
	return TG_OK;
}

int replay_batch(tg_batch* b, float* ms)
{
	TG_CHECK_HIP(hipEventRecord(b->ev[0], b->stream), "event");
	TG_CHECK_HIP(hipGraphLaunch(b->exec, b->stream), "batch graph launch");
	TG_CHECK_HIP(hipEventRecord(b->ev[1], b->stream), "event");
	TG_CHECK_HIP(hipEventSynchronize(b->ev[1]), "batch evaluation");
	TG_CHECK_HIP(hipEventElapsedTime(ms, b->ev[0], b->ev[1]), "timing");
	return TG_OK;
}

}  // namespace

extern "C" int tg_batch_create(tg_batch** out)
{
	if (!out) return fail(TG_ERR_INVALID, "tg_batch_create: null argument");
	int rc = ensure_device();
	if (rc) return rc;
	auto* b = new tg_batch;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
	if (e == hipSuccess) e = hipStreamCreate(&b->stream);
	for (int i = 0; e == hipSuccess && i < 2; ++i) e = hipEventCreate(&b->ev[i]);
	if (e != hipSuccess)
	{
		tg_batch_destroy(b);
		return hip_fail(e, "batch stream/events");
	}
	*out = b;
	return TG_OK;
}

extern "C" void tg_batch_destroy(tg_batch* b)
{
	if (!b) return;
	if (b->stream) (void)hipStreamSynchronize(b->stream);
	drop_graph(b);
	for (tg_mesher* m : b->meshers) destroy_mesher(m);
	for (DevBuf* d : { &b->trees, &b->items, &b->head, &b->work, &b->tables, &b->prune.dev, &b->queue_sort })
		d->release();
	for (hipEvent_t e : b->ev)
		if (e) (void)hipEventDestroy(e);
	if (b->stream) (void)hipStreamDestroy(b->stream);
	delete b;
}

extern "C" int tg_batch_set_tuning(tg_batch* b, uint32_t points_per_lane, uint32_t waves, uint32_t workgroups_per_cu)
{
	if (!b || !(points_per_lane == 2 || points_per_lane == 4 || points_per_lane == 8) ||
		!(waves == 1 || waves == 2 || waves == 4) || workgroups_per_cu < 1 || workgroups_per_cu > 16)
		return fail(TG_ERR_INVALID, "tg_batch_set_tuning: points per lane {2,4,8}, waves {1,2,4}, workgroups/CU 1..16");
	b->points_per_lane = int(points_per_lane);
	b->waves = int(waves);
	b->workgroups_per_cu = int(workgroups_per_cu);
	if (b->exec) return capture_batch(b);
	return TG_OK;
}

extern "C" int tg_batch_export(tg_batch* b, const tg_program* const* programs, const tg_grid* grids, uint32_t count,
	uint32_t flags)
{
	if (!b || (count && (!programs || !grids))) return fail(TG_ERR_INVALID, "tg_batch_export: null argument");
	if (count == 0 || count > 4096) return fail(TG_ERR_INVALID, "tg_batch_export: 1..4096 trees");
	if (flags & TG_FLAG_NO_PRUNE) return fail(TG_ERR_INVALID, "tg_batch_export: the batch evaluates pruned blocks");
	for (uint32_t i = 0; i < count; ++i)
	{
		if (!programs[i]) return fail(TG_ERR_INVALID, "tg_batch_export: null program");
		if (!programs[i]->decoded.prunable())
			return fail(TG_ERR_PROGRAM, "tg_batch_export: every tree must be prunable (1-Lipschitz, at most 128 "
			                            "nested blend / flate / field-scale operators)");
		int rc = validate_grid(&grids[i]);
		if (rc) return rc;
	}
	drop_graph(b);
	b->fast = (flags & TG_FLAG_FAST) != 0;
	while (b->meshers.size() < count)
	{
		tg_mesher* m = nullptr;
		int rc = tg_mesher_create(&m);
		if (rc) return rc;
		m->batch_owned = true;
		b->meshers.push_back(m);
	}
	while (b->meshers.size() > count)
	{
		destroy_mesher(b->meshers.back());
		b->meshers.pop_back();
	}
	b->programs.assign(programs, programs + count);
	b->grids.assign(grids, grids + count);
	auto t0 = std::chrono::steady_clock::now();
	int rc = b->work.ensure(size_t(count) * 16);
	if (!rc) rc = b->head.ensure(256);
	if (!rc) rc = b->trees.ensure(size_t(count) * sizeof(BlockEvalArgs));
	if (rc) return rc;
	// 1. layout and level-1 pruning per tree
	std::vector<BlockEvalArgs> args(count);
	BatchArgs ba{};
	uint64_t samples = 0;
	// every tree's pruning on the batch's stream: each height of every tree's
	// level pass in one launch, every tree's emission in one launch, one wait
	const bool fast = (flags & TG_FLAG_FAST) != 0;
	std::vector<CellGrid> cgs(count);
	std::vector<CellPrograms*> outs(count);
	for (uint32_t i = 0; i < count; ++i)
	{
		tg_mesher* m = b->meshers[i];
		m->have_grid = false;
		m->nbricks = 0;
		rc = layout(m, &grids[i]);
		if (rc) return rc;
		const BlockGrid bg = block_grid(m, &grids[i]);
		CellGrid& cg = cgs[i];
		for (int k = 0; k < 3; ++k)
		{
			cg.origin[k] = bg.origin[k];
			cg.cell[k] = bg.cell[k];
			cg.counts[k] = bg.counts[k];
		}
		cg.radius = bg.radius;
		outs[i] = &m->blocks;
	}
	rc = prune_cells_many(b->stream, programs, cgs.data(), count, fast, outs.data(), b->prune);
	if (rc) return rc;
	for (uint32_t i = 0; i < count; ++i)
	{
		tg_mesher* m = b->meshers[i];
		rc = prune_cells_finish(b->stream, programs[i], cgs[i].origin, cgs[i].cell, cgs[i].counts, fast, m->blocks);
		if (rc) return rc;
		if (!m->blocks.valid) return fail(TG_ERR_LIMIT, "tg_batch_export: a tree could not be pruned");
		args[i] = block_args(m, programs[i], &grids[i], nullptr, nullptr, 0, b->work.as<unsigned long long>() + 2 * i);
		ba.max_depth = std::max(ba.max_depth, args[i].max_depth);
		ba.max_binary = std::max(ba.max_binary, args[i].max_binary);
		ba.max_len = std::max(ba.max_len, args[i].max_len);
		samples += uint64_t(m->args.nx) * m->args.ny * m->args.nz;
	}
	// 2. the queue, sorted on the device: every tree's blocks, the most
	// expensive block program first
	std::vector<uint32_t> first(count + 1, 0);
	for (uint32_t i = 0; i < count; ++i)
	{
		// queue items pack tree << 20 | block (count <= 4096 trees)
		if (args[i].nblocks >= (1u << 20)) return fail(TG_ERR_LIMIT, "tg_batch_export: a tree has 2^20 blocks or more");
		first[i + 1] = first[i] + args[i].nblocks;
	}
	const uint32_t total = first[count];
	if (total >= (1u << 30)) return fail(TG_ERR_LIMIT, "tg_batch_export: too many blocks");
	size_t sort_tmp = 0;
	TG_CHECK_HIP(batch_queue_by_cost(nullptr, nullptr, count, total, nullptr, nullptr, nullptr, sort_tmp, b->stream),
		"queue sort sizing");
	rc = b->items.ensure(std::max<size_t>(2 * size_t(total), 2) * 4);
	if (!rc) rc = b->queue_sort.ensure(std::max<size_t>(4 * size_t(total), 1) * 4 + 256 + std::max<size_t>(sort_tmp, 256) +
	                                  (size_t(count) + 1) * 4);
	if (rc) return rc;
	uint32_t* sort_buf = b->queue_sort.as<uint32_t>();
	uint32_t* first_dev = sort_buf + 4 * size_t(total);
	void* sort_temp = reinterpret_cast<unsigned char*>(b->queue_sort.ptr) +
	                  ((4 * size_t(total) + count + 1) * 4 + 255) / 256 * 256;
	TG_CHECK_HIP(hipMemcpyAsync(b->trees.ptr, args.data(), size_t(count) * sizeof(BlockEvalArgs), hipMemcpyHostToDevice,
		b->stream), "upload");
	TG_CHECK_HIP(hipMemcpyAsync(first_dev, first.data(), (size_t(count) + 1) * 4, hipMemcpyHostToDevice, b->stream),
		"upload");
	TG_CHECK_HIP(batch_queue_by_cost(b->trees.as<const BlockEvalArgs>(), first_dev, count, total, sort_buf,
		b->items.as<uint32_t>(), sort_temp, sort_tmp, b->stream), "queue sort");
	TG_CHECK_HIP(hipStreamSynchronize(b->stream), "queue sort");
	ba.trees = b->trees.as<const BlockEvalArgs>();
	ba.items = b->items.as<const uint32_t>();
	ba.ntrees = count;
	ba.total = total;
	ba.head = b->head.as<unsigned int>();
	b->args = ba;
	auto t1 = std::chrono::steady_clock::now();
	// 3. one persistent launch (captured), then surface nets per tree
	rc = capture_batch(b);
	if (rc) return rc;
	float eval_ms = 0;
	rc = replay_batch(b, &eval_ms);
	if (rc) return rc;
	std::vector<unsigned long long> wk(2 * size_t(count));
	TG_CHECK_HIP(hipMemcpy(wk.data(), b->work.ptr, wk.size() * 8, hipMemcpyDeviceToHost), "work read");
	auto t2 = std::chrono::steady_clock::now();
	double mesh_ms = 0;
	uint64_t evals = 0;
	double flops = 0;
	// surface nets on the trees' streams: in the room an earlier export left
	// (all stages of every such tree issued before any wait), and the others,
	// with those that outgrew their room, the waiting way, each stage over
	// every tree before waiting
	std::vector<uint8_t> done(count, 0);
	for (uint32_t i = 0; i < count; ++i)
		if (b->meshers[i]->spec_verts > 0)
		{
			rc = surface_nets_in_room(b->meshers[i]);
			if (rc) return rc;
			done[i] = 1;
		}
	for (uint32_t i = 0; i < count; ++i)
		if (done[i])
		{
			TG_CHECK_HIP(hipStreamSynchronize(b->meshers[i]->stream), "surface nets");
			done[i] = surface_nets_room_check(b->meshers[i]) ? 1 : 0;
		}
	for (uint32_t i = 0; i < count; ++i)
	{
		if (done[i]) continue;
		rc = surface_nets_classify(b->meshers[i]);
		if (rc) return rc;
	}
	for (uint32_t i = 0; i < count; ++i)
	{
		if (done[i]) continue;
		TG_CHECK_HIP(hipStreamSynchronize(b->meshers[i]->stream), "vertex count");
		rc = surface_nets_vertices(b->meshers[i]);
		if (rc) return rc;
	}
	for (uint32_t i = 0; i < count; ++i)
	{
		if (done[i]) continue;
		TG_CHECK_HIP(hipStreamSynchronize(b->meshers[i]->stream), "quad count");
		rc = surface_nets_quads(b->meshers[i]);
		if (rc) return rc;
		size_next_mesh(b->meshers[i]);
	}
	for (uint32_t i = 0; i < count; ++i)
	{
		tg_mesher* m = b->meshers[i];
		TG_CHECK_HIP(hipStreamSynchronize(m->stream), "batch mesh");
		m->have_grid = true;
		tg_export_stats& s = m->stats;
		s = tg_export_stats{};
		s.samples = uint64_t(m->args.nx) * m->args.ny * m->args.nz;
		s.exact_evals = wk[2 * i];
		s.tiles_evaluated = wk[2 * i] / kTileSamples;
		s.flops = double(wk[2 * i + 1]);
		evals += wk[2 * i];
		flops += double(wk[2 * i + 1]);
	}
	auto t3 = std::chrono::steady_clock::now();
	mesh_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
	tg_batch_stats& st = b->stats;
	st = tg_batch_stats{};
	st.prune_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	st.eval_ms = eval_ms;
	st.mesh_ms = mesh_ms;
	st.total_ms = std::chrono::duration<double, std::milli>(t3 - t0).count();
	st.trees = count;
	st.blocks = b->args.total;
	st.samples = samples;
	st.exact_evals = evals;
	st.flops = flops;
	st.points_per_lane = uint32_t(b->last_points_per_lane);
	st.waves = uint32_t(b->last_waves);
	st.workgroups = uint32_t(b->last_workgroups);
	const BatchArgs& q = b->args;
	st.lds_bytes = uint32_t(block_lds_bytes_global(q.slots, std::max(2, b->last_points_per_lane),
		std::max(1, b->last_waves)));
	st.max_len = q.max_len;
	st.max_binary = q.max_binary;
	st.max_depth = q.max_depth;
	return TG_OK;
}

extern "C" int tg_batch_replay(tg_batch* b, double* eval_ms)
{
	if (!b || !eval_ms) return fail(TG_ERR_INVALID, "tg_batch_replay: null argument");
	if (!b->exec) return fail(TG_ERR_INVALID, "tg_batch_replay: no batch exported yet");
	float ms = 0;
	int rc = replay_batch(b, &ms);
	if (rc) return rc;
	*eval_ms = ms;
	return TG_OK;
}

extern "C" int tg_batch_autotune(tg_batch* b, tg_batch_tune_result* out)
{
	if (!b) return fail(TG_ERR_INVALID, "tg_batch_autotune: null argument");
	if (!b->exec) return fail(TG_ERR_INVALID, "tg_batch_autotune: export a batch first");
	// points per lane x (waves, workgroups per CU): 18 shapes
	const int pts_c[3] = { 2, 4, 8 };
	const int shape_c[6][2] = { { 4, 2 }, { 4, 3 }, { 4, 4 }, { 4, 6 }, { 2, 6 }, { 2, 8 } };
	tg_batch_tune_result r{};
	double best = 1e300;
	uint32_t k = 0;
	for (int pi = 0; pi < 3; ++pi)
		for (int si = 0; si < 6; ++si)
		{
			b->points_per_lane = pts_c[pi];
			b->waves = shape_c[si][0];
			b->workgroups_per_cu = shape_c[si][1];
			int rc = capture_batch(b);
			if (rc) return rc;
			float t = 0, sum = 0;
			for (int rep = 0; rep < 3; ++rep)
			{
				rc = replay_batch(b, &t);
				if (rc) return rc;
				if (rep > 0) sum += t;
			}
			const double ms = 0.5 * double(sum);
			r.candidate_points[k] = uint32_t(b->last_points_per_lane);
			r.candidate_waves[k] = uint32_t(b->last_waves);
			r.candidate_workgroups_per_cu[k] = uint32_t(shape_c[si][1]);
			r.candidate_ms[k] = ms;
			if (ms < best)
			{
				best = ms;
				r.points_per_lane = uint32_t(b->last_points_per_lane);
				r.waves = uint32_t(b->last_waves);
				r.workgroups_per_cu = uint32_t(shape_c[si][1]);
			}
			++k;
		}
	r.candidates = k;
	r.eval_ms = best;
	b->points_per_lane = int(r.points_per_lane);
	b->waves = int(r.waves);
	b->workgroups_per_cu = int(r.workgroups_per_cu);
	int rc = capture_batch(b);
	if (rc) return rc;
	if (out) *out = r;
	return TG_OK;
}

extern "C" tg_mesher* tg_batch_mesher(tg_batch* b, uint32_t index)
{
	if (!b || index >= b->meshers.size())
	{
		fail(TG_ERR_INVALID, "tg_batch_mesher: index out of range");
		return nullptr;
	}
	return b->meshers[index];
}

extern "C" int tg_batch_stats_get(const tg_batch* b, tg_batch_stats* out)
{
	if (!b || !out) return fail(TG_ERR_INVALID, "tg_batch_stats_get: null argument");
	*out = b->stats;
	return TG_OK;
}
