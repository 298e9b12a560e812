// Kernel argument blocks and host launchers for the evaluation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tg_types.h"

namespace tg {

// Dynamic LDS bytes per wave for the interpreter spill region.
__host__ __device__ inline size_t spill_bytes_per_wave(uint32_t slots, int points_per_lane)
{
	return size_t(slots) * size_t(points_per_lane) * 64 * sizeof(float);
}

struct PointArgs
{
	const Instr* prog;
	const float* xyz;   // 3 floats per point
	float* out;
	uint64_t count;
	uint32_t slots;     // LDS spill slots per wave (max_stack - 1, at least 1)
};

struct GridArgs
{
	const Instr* prog;
	float gx, gy, gz;        // grid origin
	float dx, dy, dz;        // sample step
	uint32_t nx, ny, nz;     // samples per axis
	uint32_t tx, ty, tz;     // tiles per axis
	uint32_t ntiles;         // tiles to evaluate (list length or tx*ty*tz)
	const uint32_t* tile_list;  // tile ids to evaluate, or null for all
	const uint32_t* tile_slot;  // pool slot per tile id, or null for identity
	float* pool;             // 512 floats per slot, [z][y][x] inside a tile
	uint8_t* signs;          // sign bits: 64 bytes per tile id, byte (z & 7) * 8 + (y & 7), x bits LSB first
	uint32_t slots;
};

// ---- pruned programs ----------------------------------------------------------
//
// A program is also viewed as a list of entries in postfix order, each naming
// a record and its children, so that it can be pruned per cell without
// touching the records: entry i of a pruned program says "execute record
// `rec`" (or negate the top of stack), and where its operands' roots are.
struct Entry
{
	uint32_t rec;   // record index (bits 0..21) | transform kind << 22 | opcode << 24 (6 bits; OP_NEGATE: no
	                // record) | kEntryPlusZero
	uint32_t l;     // binary: position of the left operand's root; unary: operand root
	uint32_t r;     // binary: position of the right operand's root
	uint32_t ord;   // binary: ordinal among the program's binary entries
};

constexpr uint32_t kRecMask = 0x003fffffu;
// Entry flag: after this entry, top += +0.0.  Pruning sets it on the kept
// operand's root when a BlendInter or BlendDiff drops one operand: the
// reference adds the blend's zero term in double (+0.0), which turns a kept
// -0.0 into +0.0 (BlendUnion subtracts it, which keeps -0.0).
constexpr uint32_t kEntryPlusZero = 1u << 31;
__host__ __device__ inline uint32_t entry_op(uint32_t rec) { return (rec >> 24) & 0x3fu; }
constexpr int kBlockTiles = 4;                 // tiles per block edge
constexpr int kTilesPerBlock = 64;             // 4^3 tiles = one lane each
constexpr int kBlockSamples = kTile * kBlockTiles;

struct BlockDesc
{
	uint32_t offset;     // first entry in the block-program pool
	uint32_t length;     // entries
	uint32_t max_stack;  // interpreter stack high-water mark
	uint32_t flops;      // cost-table work of one evaluation
};

// Level 1: prune the whole program for every cell of a regular cell grid
// (cell (i,j,k) centred at origin + (i+.5, j+.5, k+.5) * cell, valid within
// `radius` of its centre).  Pruned programs are packed into `pool`; a cell
// whose program does not fit (or whose pruning walk overflows) is marked to
// run the whole program.  counters: [0] pool entries asked for, [1] stack
// high-water mark over cells, [2] / [3] most binary entries / entries of a
// packed program, [4] cells left with the whole program.
struct PruneArgs
{
	const Instr* prog;
	const Entry* entries;   // whole program as entries
	uint32_t n_entries;
	uint32_t n_binary;
	float ox, oy, oz;       // cell grid origin
	float cx, cy, cz;       // cell size per axis
	float radius;           // validity radius around each cell centre
	uint32_t nx, ny, nz;    // cells per axis
	uint32_t slots;         // LDS spill slots (whole-program max stack - 1)
	uint32_t whole_flops;   // cost of one whole-program evaluation
	Entry* pool;
	uint64_t capacity;      // pool entries
	BlockDesc* desc;        // one per cell
	unsigned long long* counters;
	uint64_t* scratch;      // per wave: 2 n_binary + n_entries words (decision masks, liveness)
	uint32_t cells_per_wave;  // 1..64 (lanes beyond carry no cell)
	int masks_ready;          // decision masks and liveness already in scratch (k_prune_level): skip both passes
	const uint32_t* counts;   // with masks_ready: program length per lane, [group][64]
	uint32_t splits;          // with masks_ready: waves per cell group (1, 2 or 4), each emitting 64 / splits columns
	int columns;              // with masks_ready: a wave per cell, a lane per entry (k_prune_columns)
};

// Level-parallel decision pass: one launch per entry height (leaves 0, a
// parent one above its highest child); wave = (entry of the level, cell
// group of k_prune_cells), lane = cell.  Every entry's value at every cell
// centre is kept in `values`, so a parent reads its operands instead of
// re-running the program; binary entries record the same decision masks as
// k_prune_cells' own pass.
struct LevelArgs
{
	const Instr* prog;
	const Entry* entries;
	const uint32_t* list;     // entry positions of this height
	uint32_t count;           // entries of this height
	uint32_t ngroups;         // cell groups (waves of k_prune_cells)
	uint32_t groups_per_wave; // consecutive cell groups one wave runs (the entry's record loaded once)
	uint32_t cells_per_wave;
	float ox, oy, oz, cx, cy, cz, radius;
	uint32_t nx, ny, nz;
	float* values;            // [entry][group][64 lanes]
	uint64_t* scratch;        // decision masks and liveness, as PruneArgs::scratch
	uint32_t n_binary, n_entries;
	int liveness;             // 0: values and decisions (heights upward); 1: liveness (heights downward)
	uint32_t* counts;         // liveness: entries each lane's program will hold, [group][64] (zeroed)
};

// Level 2 + evaluation: one workgroup per block.
struct BlockEvalArgs
{
	const Instr* prog;
	const Entry* pool;
	const BlockDesc* desc;
	const uint32_t* block_list;  // blocks to run (null: all)
	const uint64_t* tile_mask;   // per block: tiles to evaluate (null: all)
	uint32_t nblocks;            // launched blocks
	float gx, gy, gz, dx, dy, dz;
	uint32_t nx, ny, nz;
	uint32_t tx, ty, tz;
	uint32_t bx, by, bz;
	const uint32_t* tile_slot;
	float* out_pool;
	uint8_t* signs;              // 64 bytes per tile id (see GridArgs)
	uint32_t slots;              // LDS spill slots per wave, each P floats per lane (block_spill_slots)
	uint32_t max_depth;          // deepest block program's stack (host: sizes `slots`)
	int prune_tiles;             // 0: evaluate block programs unpruned per tile
	uint32_t min_prune_len;      // blocks with shorter programs skip the per-tile pruning
	unsigned long long* work;    // [0] lanes interpreted (whole tiles), [1] cost units over the grid's samples
	const Entry* whole;          // whole program, for blocks whose pruning overflowed
	uint32_t n_whole;
	uint32_t n_whole_binary;
	uint32_t max_binary;         // LDS tables: binary entries / entries of any block program
	uint32_t max_len;
	int waves;                   // waves per workgroup
	float slack;                 // rounding slack of the per-tile pruning decisions (rounding_slack)
};

// Persistent batched evaluation of many trees' blocks (one queue item = one
// block of one tree).
struct BatchArgs
{
	const BlockEvalArgs* trees;  // per tree (device array)
	const uint32_t* items;       // [2 * total]: tree, block of each queue item, in queue order (null: item k is
	                             // block k of tree 0)
	uint32_t ntrees;
	uint32_t total;              // queue items
	unsigned int* head;          // queue head (reset to 0 before each launch)
	uint32_t slots, max_binary, max_len;  // layout: the largest over the trees
	uint32_t max_depth;                   // deepest block program over the trees (host: sizes `slots`)
	uint32_t pool_words;                  // tile pool in LDS (block_pool_words; 0: the waves' lists only)
	uint32_t* tables;                     // non-null: per-workgroup decision tables and tile pool in global
	                                      // memory, table_words each (block_global_words)
	uint64_t table_words;
};

// Exact brick certificate: wave per block (4^3 cube bricks of 8^3 cubes,
// lane = brick), the block's pruned program evaluated at every brick centre c.
// A brick holding a sign change has |d(c)| <= radius (1-Lipschitz field) plus
// the rounding slack; such bricks get flags[b] = 1 (never cleared).
struct CertifyArgs
{
	const Instr* prog;
	const Entry* pool;
	const BlockDesc* desc;       // null: every block runs the whole program
	const Entry* whole;
	uint32_t n_whole;
	uint32_t whole_flops;
	float gx, gy, gz, dx, dy, dz;
	uint32_t bx, by, bz;         // blocks per axis
	uint32_t cbx, cby, cbz;      // bricks per axis
	float radius;                // brick half-diagonal (1 + 1e-5) + rounding slack (rounding_slack)
	uint8_t* flags;
	unsigned long long* counters;  // [0] bricks newly flagged, [1] cost units executed
	uint32_t slots;
};

// ---- sphere tracing -------------------------------------------------------------

constexpr int kTraceLevels = 6;

// A ray still marching at the end of a pass: what the next pass resumes from
// (its direction and start follow from `ray`).
struct TraceRay
{
	uint32_t ray, steps;
	float t, px, py, pz;
};

struct TraceLevel
{
	const BlockDesc* desc;   // null: level absent
	const Entry* pool;       // this level's pruned programs
	uint32_t n;              // cells per axis (cubic cells over the level's cube)
	float cell;              // cell edge
	float o[3];              // the cube's lowest corner
	float inv;               // RN(1 / cell)
};

struct TraceArgs
{
	const Instr* prog;
	const Entry* whole;
	uint32_t n_whole;
	uint32_t whole_flops;
	TraceLevel levels[kTraceLevels];  // coarse to fine
	int nlevels;

	// rays: camera (width * height pixels) or arrays (nrays)
	int camera;
	float eye[3], fwd[3], right[3], up[3];
	float kx, ky;
	uint32_t width, height;
	const float* starts;
	const float* dirs;
	uint32_t nrays;
	int max_steps;
	float eps, far_, h;               // h <= 0: no normals
	float* t_out;
	float* n_out;                     // 3 per ray, or null
	float* p_out;                     // 3 per ray, or null
	uint32_t* steps_out;              // or null
	uint8_t* hit_out;                 // or null
	unsigned long long* work;         // [0] evaluations, [1] cost units, [2] hits
	uint32_t slots;

	// Passes: a pass marches each of its rays at most `budget` steps (<= 0:
	// to max_steps) and appends the rays still marching to q_out with their
	// cell keys (trace_key) in q_keys (q_out null: every ray ends in this
	// pass).  The first pass takes its rays from the camera tiles or the
	// arrays; a resumed pass takes q_in[q_order[i]] for i < q_count,
	// q_per_wave (1..64) to a wave.
	int budget;
	const TraceRay* q_in;
	const uint32_t* q_order;
	uint32_t q_count;
	uint32_t q_per_wave;
	TraceRay* q_out;
	uint32_t* q_keys;
	uint32_t* q_out_count;
};

// Point-cloud export: cells of a regular point grid kept near the surface,
// then refined along the gradient.
struct CloudArgs
{
	const Instr* prog;
	float sx, sy, sz;          // start (model minimum)
	float dx, dy, dz;          // step
	float hx, hy, hz;          // half step
	uint32_t ix, iy, iz;       // cells per axis
	uint64_t cells;
	float diagonal;            // length of the half step
	uint32_t* flags;           // per cell: kept
	const uint32_t* rank;      // exclusive scan of flags
	float* points;             // 3 per kept cell
	uint64_t npoints;
	int iterations;            // refinement steps
	uint32_t slots;
};

// VoxExport's voxel selection: cell (x, y, z) of a size[0] x size[1] x size[2]
// grid over [lo, hi] is kept when |d| <= radius at its centre.
struct VoxArgs
{
	const Instr* prog;
	float lo[3], hi[3];
	float fs[3];               // float(size)
	uint32_t size[3];
	uint64_t cells;
	float radius;
	uint32_t* flags;           // per cell: kept
	const uint32_t* rank;      // exclusive scan of flags
	uint32_t* list;            // kept cell indices, raster order
	uint32_t slots;
};

// LuaPivotTowards over a batch of queries (one thread each).
struct PivotArgs
{
	const Instr* prog;
	const float* pivot;        // 3 per query
	const float* heading;      // 3 per query
	const float* down;         // 3 per query
	uint64_t count;
	float max_dist, margin, max_angle;  // max_angle in radians
	int max_iterations;        // loop turns before status 3
	float* tail_out;           // 3 per query
	float* heading_out;        // 3 per query
	int32_t* status_out;       // 0 ok, 2 heading antiparallel to down, 3 iteration cap
	uint32_t slots;
};

// Central-difference or tetrahedral gradients at points.
struct GradArgs
{
	const Instr* prog;
	const float* xyz;
	float* out;
	uint64_t count;
	float h;
	int tetra;                        // 1: SDFNode::Gradient's tetrahedral rule
	uint32_t slots;
};

// Launchers, one set per numerics build (exact = reference bit parity,
// fast = contracted FP32).  `p` is points per lane for the tile kernel (4 or 8).
namespace exact {
hipError_t launch_points(const PointArgs& a, hipStream_t s);
hipError_t launch_prune_level(const LevelArgs& a, hipStream_t s);
// many programs' passes in one launch: job k owns waves [wave_base[k], wave_base[k + 1]) of `waves`
hipError_t launch_prune_level_multi(const LevelArgs* jobs, const uint32_t* wave_base, uint32_t njobs, uint32_t waves,
	hipStream_t s);
// many programs' emissions in one launch: job k owns blocks [block_base[k], ...) of `blocks`
hipError_t launch_prune_cells_multi(const PruneArgs* jobs, const uint32_t* block_base, uint32_t njobs, uint32_t blocks,
	uint32_t max_slots, hipStream_t s);
// the same a wave per cell (PruneArgs::columns): job k owns cells [cell_base[k], ...) of `cells`
hipError_t launch_prune_columns_multi(const PruneArgs* jobs, const uint32_t* cell_base, uint32_t njobs, uint32_t cells,
	uint32_t max_entries, hipStream_t s);
hipError_t launch_vox(const VoxArgs& a, int stage, hipStream_t s);  // 0 select, 1 list
hipError_t launch_pivot(const PivotArgs& a, hipStream_t s);
hipError_t launch_cloud(const CloudArgs& a, int stage, hipStream_t s);  // 0 select, 1 points, 2 refine
hipError_t launch_batch(const BatchArgs& b, int p, uint32_t waves, uint32_t workgroups, hipStream_t s);
hipError_t launch_batch_reset(unsigned int* head, unsigned long long* work, uint32_t nwork, hipStream_t s);
hipError_t launch_certify(const CertifyArgs& a, hipStream_t s);
hipError_t launch_tiles(const GridArgs& a, int p, hipStream_t s);
hipError_t launch_prune_cells(const PruneArgs& a, hipStream_t s);
hipError_t launch_eval_blocks(const BlockEvalArgs& a, int p, hipStream_t s);
hipError_t launch_trace(const TraceArgs& a, hipStream_t s);
// resumed pass (a.q_in set): a wave per a.q_per_wave of the a.q_count rays
hipError_t launch_trace_resume(const TraceArgs& a, hipStream_t s);
hipError_t launch_gradient(const GradArgs& a, hipStream_t s);
// all 2^32 float inputs of the exact build's square root against sqrtf
hipError_t launch_sqrt_check(uint32_t scramble, unsigned long long* mismatches, hipStream_t s);
hipError_t launch_quotient_check(float b, unsigned long long* counts, hipStream_t s);
}
namespace fast {
hipError_t launch_points(const PointArgs& a, hipStream_t s);
hipError_t launch_prune_level(const LevelArgs& a, hipStream_t s);
// many programs' passes in one launch: job k owns waves [wave_base[k], wave_base[k + 1]) of `waves`
hipError_t launch_prune_level_multi(const LevelArgs* jobs, const uint32_t* wave_base, uint32_t njobs, uint32_t waves,
	hipStream_t s);
// many programs' emissions in one launch: job k owns blocks [block_base[k], ...) of `blocks`
hipError_t launch_prune_cells_multi(const PruneArgs* jobs, const uint32_t* block_base, uint32_t njobs, uint32_t blocks,
	uint32_t max_slots, hipStream_t s);
// the same a wave per cell (PruneArgs::columns): job k owns cells [cell_base[k], ...) of `cells`
hipError_t launch_prune_columns_multi(const PruneArgs* jobs, const uint32_t* cell_base, uint32_t njobs, uint32_t cells,
	uint32_t max_entries, hipStream_t s);
hipError_t launch_vox(const VoxArgs& a, int stage, hipStream_t s);  // 0 select, 1 list
hipError_t launch_pivot(const PivotArgs& a, hipStream_t s);
hipError_t launch_cloud(const CloudArgs& a, int stage, hipStream_t s);  // 0 select, 1 points, 2 refine
hipError_t launch_batch(const BatchArgs& b, int p, uint32_t waves, uint32_t workgroups, hipStream_t s);
hipError_t launch_batch_reset(unsigned int* head, unsigned long long* work, uint32_t nwork, hipStream_t s);
hipError_t launch_certify(const CertifyArgs& a, hipStream_t s);
hipError_t launch_tiles(const GridArgs& a, int p, hipStream_t s);
hipError_t launch_prune_cells(const PruneArgs& a, hipStream_t s);
hipError_t launch_eval_blocks(const BlockEvalArgs& a, int p, hipStream_t s);
hipError_t launch_trace(const TraceArgs& a, hipStream_t s);
// resumed pass (a.q_in set): a wave per a.q_per_wave of the a.q_count rays
hipError_t launch_trace_resume(const TraceArgs& a, hipStream_t s);
hipError_t launch_gradient(const GradArgs& a, hipStream_t s);
}

// LDS layout of the block kernel (bytes), shared by launcher and kernel:
// decision masks (2 x max_binary) and liveness (max_len) of the block program,
// then per wave a tile program (max_len words) and the spill region.
// LDS spill slots of the block kernel, each P floats per lane.  A tile whose
// program needs more than `slots` slots is interpreted one point per lane,
// seeing the region as slots * P one-float slots, which must hold the deepest
// block program (the wave-0 decision pass runs one point per lane too).
// Beyond that minimum the region takes up to `budget_bytes` per wave (by
// default 3 KiB: enough for the short programs of most tiles, small enough
// for high occupancy).
__host__ __device__ inline uint32_t block_spill_slots(uint32_t max_depth, int p, size_t budget_bytes = 3072)
{
	const uint32_t need = max_depth > 1 ? max_depth - 1 : 1;  // one-float slots
	const uint32_t floor_slots = (need + uint32_t(p) - 1) / uint32_t(p);
	const uint32_t budget = uint32_t(budget_bytes / (size_t(p) * 64u * 4u));
	const uint32_t want = budget < need ? budget : need;
	return floor_slots > want ? floor_slots : (want > 0 ? want : 1u);
}
// Block kernel LDS: decision masks and liveness of the block program, one
// uint4 per tile (its program's place in the tile pool, length, stack depth,
// cost), then every wave's program list (together: the tile pool), then
// every wave's spill region.
constexpr size_t kTileInfoBytes = 64 * 16;
__host__ __device__ inline size_t block_table_bytes(uint32_t max_binary, uint32_t max_len)
{
	return size_t(max_binary) * 16 + ((size_t(max_len) * 8 + 15) & ~size_t(15)) + kTileInfoBytes;
}
__host__ __device__ inline size_t block_list_bytes(uint32_t max_len)
{
	return (size_t(max_len) * 4 + 15) & ~size_t(15);
}
// Tile pool words: at least every wave's list (a wave emits a tile program
// there when the block's 64 do not fit the pool together), more when the
// launch asks for it; a multiple of 4 words.
__host__ __device__ inline uint32_t block_pool_words(uint32_t pool_words, int waves, uint32_t max_len)
{
	const uint32_t lists = uint32_t(waves) * uint32_t(block_list_bytes(max_len) / 4);
	return ((pool_words > lists ? pool_words : lists) + 3u) & ~3u;
}
__host__ __device__ inline size_t block_lds_bytes(uint32_t slots, int p, int waves, uint32_t max_binary,
	uint32_t max_len, uint32_t pool_words = 0)
{
	return block_table_bytes(max_binary, max_len) + size_t(block_pool_words(pool_words, waves, max_len)) * 4 +
	       size_t(waves) * spill_bytes_per_wave(slots, p);
}
// Blocks with their decision tables and tile pool in global memory: words of
// one workgroup's slice (masks, liveness, then 64 tile programs of at most
// max_len entries each), and the LDS that remains (tile descriptors, spill).
__host__ __device__ inline uint64_t block_global_words(uint32_t max_binary, uint32_t max_len)
{
	return 4 * uint64_t(max_binary) + 2 * uint64_t(max_len) + uint64_t(kTilesPerBlock) * max_len + 64;
}
__host__ __device__ inline size_t block_lds_bytes_global(uint32_t slots, int p, int waves)
{
	return kTileInfoBytes + size_t(waves) * spill_bytes_per_wave(slots, p);
}
// Cell pruning: per-lane position stacks and the spill region (the decision
// masks and liveness live in global scratch, prune_scratch_words per wave).
__host__ __device__ inline size_t prune_lds_bytes(uint32_t slots)
{
	return size_t(kMaxStack) * 64 * 4 + spill_bytes_per_wave(slots, 1);
}
__host__ __device__ inline size_t prune_scratch_words(uint32_t n_binary, uint32_t n_entries)
{
	return 2 * size_t(n_binary) + n_entries;
}

}  // namespace tg
