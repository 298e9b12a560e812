// Device surface-nets stages (see tg_mesh.hip for the data layout).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

struct BlockDesc;

struct MeshArgs
{
	// grid (same meaning as isosurface::regular_grid_t)
	float gx, gy, gz, dx, dy, dz;
	uint32_t sx, sy, sz;          // cubes per axis
	uint32_t nx, ny, nz;          // samples per axis (= s + 1)
	uint32_t tx, ty, tz;          // sample tiles per axis
	uint64_t sx_div, sy_div, aw_div;  // division magics of sx, sy, aw (mesh_div_magic)
	uint32_t cbx, cby, cbz;       // cube bricks per axis (8^3 cubes each)
	// inputs
	const float* pool;
	const uint32_t* tile_slot;    // null: identity
	const uint8_t* signs;         // 64 bytes per tile id (tg_mesh.hip)
	const uint8_t* brick_flags;   // null: every cube brick may hold active cubes
	// intermediates
	uint32_t aw;                  // 64-bit words per cube row
	uint64_t* active;
	uint32_t* word_count;
	uint32_t* word_base;
	uint32_t* cubes;              // active cubes in vertex-id order: i + sx (j + sy k)
	uint32_t* cube_quads;         // quad flags per active cube (tg_mesh.hip: quad_flags)
	uint32_t* cube_qbase;         // exclusive scan of the quad counts
	uint64_t cap_verts;           // room for active cubes (cubes, vertices, quad flags and bases)
	uint64_t cap_quads;           // room for quads (triangle pairs)
	uint32_t* tail;               // non-null: pinned host words the kernels write the totals to
	                              // (k_vertices: the last word's vertex base and count,
	                              // k_quad_write: the last cube's quad base and flags)
	// outputs
	float* vertices;
	uint32_t* triangles;
};

// n / d for every 32-bit n is floor(M n / 2^64) with M = ceil(2^64 / d) (Lemire,
// Kaser and Kurz, "Faster remainder by direct computation", Software:
// Practice and Experience 49(6), 2019); M = 0 stands for d = 1.
inline uint64_t mesh_div_magic(uint32_t d) { return d <= 1 ? 0ull : ~0ull / d + 1ull; }

hipError_t mesh_classify(const MeshArgs& a, hipStream_t s);
// Thread per active cube (after the vertex-id scan): cube list, vertices,
// quad flags and, after the scan of their counts, triangles.
hipError_t mesh_list_cubes(const MeshArgs& a, hipStream_t s);
hipError_t mesh_vertices(const MeshArgs& a, uint64_t nverts, hipStream_t s);
hipError_t mesh_quad_write(const MeshArgs& a, uint64_t nverts, hipStream_t s);
hipError_t mesh_pool_to_dense(const MeshArgs& a, float* dense, uint32_t k0, uint32_t nk, hipStream_t s);
hipError_t mesh_dense_to_pool(const MeshArgs& a, const float* dense, hipStream_t s);
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, size_t& temp_bytes, hipStream_t s);
// Exclusive scan of the quad counts (bits 0..2) of n quad-flag words.
hipError_t exclusive_scan_quads(const uint32_t* flags, uint32_t* out, uint64_t n, void* temp, size_t& temp_bytes,
	hipStream_t s);
// Blocks (all n, or the n of `list`) ordered by their pruned program's cost,
// most expensive first, into `order`; keys needs 2n words, values n.  With
// temp = null only sizes temp_bytes.
// The batch queue: every tree's blocks (tree t owns items [first[t],
// first[t + 1])), most expensive block program first (stable), written to
// `items` as (tree, block) word pairs; sort_buf holds 4 total words.  With
// temp null, sets temp_bytes only.
hipError_t batch_queue_by_cost(const BlockEvalArgs* trees, const uint32_t* first, uint32_t ntrees, uint32_t total,
	uint32_t* sort_buf, uint32_t* items, void* temp, size_t& temp_bytes, hipStream_t s);
hipError_t blocks_by_cost(const BlockDesc* desc, const uint32_t* list, uint32_t n, uint32_t* keys,
	uint32_t* values, uint32_t* order, void* temp, size_t& temp_bytes, hipStream_t s);

}  // namespace tg
