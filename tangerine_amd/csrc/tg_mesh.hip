// Device naive surface nets over a sampled distance grid.
//
// Follows the serial reference isosurface::surface_nets
// (third_party/naive-surface-nets/src/surface_nets.cpp:30-580):
//   * a cube (i,j,k), 0 <= i < sx, is active iff one of its 12 edges has
//     corner signs that differ under `s >= 0` (:728-803);
//   * its vertex is the mean of the edge crossings p1 + t (p2 - p1),
//     t = (0 - s1) / (s2 - s1), accumulated in edge order 0..11 and mapped to
//     world space by min + (max - min) * c / s (:258-296), all in float
//     without contraction (this file is built with -ffp-contract=off);
//   * vertex ids follow raster order of the cube index i + j sx + k sx sy,
//     which is the serial reference's numbering (:91-93, :298-300);
//   * each active cube with i, j, k > 0 emits, for the directed edges (0,4),
//     (3,0), (0,1), a quad over the neighbour cubes {0,1,2}, {0,5,4},
//     {2,3,4} when all three are active, ordered {0,1,2} if the edge's second
//     value is greater than its first and {2,1,0} otherwise, split into the
//     triangles (v0,v1,v2), (v0,v2,v3) (:384-556).  The reference emits faces
//     in hash-map order; here they come in raster order of the cube, then edge
//     order, then triangle order.
//
// Data layout (built by the evaluation kernels):
//   pool  : 8^3-sample tiles, 512 floats each, [z][y][x] inside a tile; the
//           slot of tile t is tile_slot[t] (identity when null).
//   signs : one bit per sample (1 = s >= 0), 64 bytes per tile id (not slot):
//           byte (k & 7) * 8 + (j & 7) of tile (i >> 3, j >> 3, k >> 3) holds
//           x = 8 (i >> 3) .. + 7, LSB first (one wave's ballot per z-slice).
//   active: one bit per cube; cube row (j, k) is `aw` 64-bit words at
//           (k * sy + j) * aw.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "tg_kernels.h"
#include "tg_mesh.h"
#include "tg_types.h"

namespace tg {

namespace {

__device__ __forceinline__ float pool_at(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k)
{
	const uint32_t tile = (i >> 3) + a.tx * ((j >> 3) + a.ty * (k >> 3));
	const uint32_t slot = a.tile_slot ? a.tile_slot[tile] : tile;
	return a.pool[size_t(slot) * kTileSamples + ((k & 7) * 64 + (j & 7) * 8 + (i & 7))];
}

// n / d from d's magic (mesh_div_magic: M = ceil(2^64 / d), 0 for d = 1).
__device__ __forceinline__ uint32_t div_magic(uint32_t n, uint64_t m)
{
	if (m == 0) return n;
	const uint64_t lo = uint64_t(uint32_t(m)) * n;
	const uint64_t hi = (m >> 32) * n + (lo >> 32);
	return uint32_t(hi >> 32);
}

__device__ __forceinline__ bool cube_active(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k)
{
	const uint64_t word = a.active[(size_t(k) * a.sy + j) * a.aw + (i >> 6)];
	return (word >> (i & 63)) & 1ull;
}

__device__ __forceinline__ uint32_t vertex_id(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k)
{
	const size_t wi = (size_t(k) * a.sy + j) * a.aw + (i >> 6);
	const uint64_t below = a.active[wi] & ((1ull << (i & 63)) - 1ull);
	return a.word_base[wi] + uint32_t(__popcll(below));
}

// Workgroup per (segment of 8 active words = 64 cube columns, 8 cube rows in
// y, 8 in z).  Task (tile lt, slice kk) reads the 8-byte z-slices it needs
// from the tile-major sign array (8 consecutive tasks read one tile's 64
// bytes) and forms the active bits of its 8 x 8 cubes byte-parallel in a
// 64-bit word; the bytes are transposed through LDS into raster-order active
// words (a cube is active when its 8 corner signs differ).
constexpr uint32_t kSegWords = 8;  // active words per segment (64 sample tiles)

__device__ __forceinline__ uint64_t sign_slice(const MeshArgs& a, uint32_t tix, uint32_t tiy, uint32_t tiz, uint32_t z)
{
	if (tix >= a.tx || tiy >= a.ty || tiz >= a.tz) return 0ull;
	const size_t tile = tix + size_t(a.tx) * (tiy + size_t(a.ty) * tiz);
	return reinterpret_cast<const uint64_t*>(a.signs)[tile * 8 + z];
}

// Active bits of the 8 x 8 cubes of tile column tix at cube slice kk: byte jj
// holds cubes x = 8 tix .. 8 tix + 7 of row (8 cy + jj, 8 cz + kk).
__device__ __forceinline__ void corner_masks(const MeshArgs& a, uint32_t tix, uint32_t cy, uint32_t cz, uint32_t kk,
	uint64_t& all, uint64_t& any)
{
	const uint64_t s = sign_slice(a, tix, cy, cz, kk);
	const uint64_t sz = kk < 7 ? sign_slice(a, tix, cy, cz, kk + 1) : sign_slice(a, tix, cy, cz + 1, 0);
	const uint64_t y0 = sign_slice(a, tix, cy + 1, cz, kk) & 0xffull;
	const uint64_t y0z = (kk < 7 ? sign_slice(a, tix, cy + 1, cz, kk + 1) : sign_slice(a, tix, cy + 1, cz + 1, 0)) & 0xffull;
	const uint64_t s_up = (s >> 8) | (y0 << 56), sz_up = (sz >> 8) | (y0z << 56);  // rows jj + 1
	all = s & s_up & sz & sz_up;
	any = s | s_up | sz | sz_up;
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b % 8).  This
// numbering gives each XCD a contiguous range of logical workgroups, so that
// the sign words that neighbouring workgroups share are read through one L2.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t n)
{
	const uint32_t per = n / 8, rem = n % 8, x = b % 8, k = b / 8;
	return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

__global__ void __launch_bounds__(256) k_classify(MeshArgs a, uint32_t nseg)
{
	constexpr uint32_t kTiles = kSegWords * 8 + 1;  // + the next tile, for the x + 1 corner
	__shared__ uint64_t lds[8 * 8 * kSegWords];      // [kk][jj][8 tiles per word] bytes
	__shared__ uint64_t all_s[kTiles * 8], any_s[kTiles * 8];
	const uint32_t blk = xcd_contiguous(blockIdx.x, gridDim.x);
	const uint32_t seg = blk % nseg;
	const uint32_t cy = (blk / nseg) % a.cby, cz = blk / (nseg * a.cby);
	const uint32_t t0 = seg * kSegWords * 8;
	// (launched with 256 threads: fixed trip counts, so that every thread's
	// sign loads are issued before the first wait)
#pragma unroll
	for (uint32_t q = threadIdx.x; q < kTiles * 8; q += 256)
		corner_masks(a, t0 + (q >> 3), cy, cz, q & 7, all_s[q], any_s[q]);
	__syncthreads();
	uint8_t* bytes = reinterpret_cast<uint8_t*>(lds);
#pragma unroll
	for (uint32_t q = threadIdx.x; q < 512; q += 256)
	{
		const uint32_t kk = q & 7, lt = q >> 3;
		const uint64_t all0 = all_s[q], any0 = any_s[q], all1 = all_s[q + 8], any1 = any_s[q + 8];
		// corner x = i + 1: shift each byte down by one, bit 7 from the next tile's bit 0
		const uint64_t lo7 = 0x7f7f7f7f7f7f7f7full, b0 = 0x0101010101010101ull;
		const uint64_t all_hi = ((all0 >> 1) & lo7) | ((all1 & b0) << 7);
		const uint64_t any_hi = ((any0 >> 1) & lo7) | ((any1 & b0) << 7);
		const uint64_t act = ~(all0 & all_hi) & (any0 | any_hi);
#pragma unroll
		for (uint32_t jj = 0; jj < 8; ++jj) bytes[(kk * 8 + jj) * 64 + lt] = uint8_t(act >> (8 * jj));
	}
	__syncthreads();
#pragma unroll
	for (uint32_t o = threadIdx.x; o < 512; o += 256)
	{
		const uint32_t ww = o & 7, jj = (o >> 3) & 7, kk = o >> 6;
		const uint32_t w = seg * kSegWords + ww, j = cy * 8 + jj, k = cz * 8 + kk;
		if (w >= a.aw || j >= a.sy || k >= a.sz) continue;
		uint64_t act = lds[(kk * 8 + jj) * 8 + ww];
		const uint64_t first = uint64_t(w) * 64;
		if (first + 64 > a.sx)
		{
			const uint32_t keep = (a.sx > first) ? uint32_t(a.sx - first) : 0u;
			act &= (keep >= 64) ? ~0ull : ((1ull << keep) - 1ull);
		}
		if (a.brick_flags)
		{
			// Only cube bricks (8^3 cubes) flagged for exact evaluation may hold
			// active cubes; the flags hold one byte per brick.
			uint64_t m = 0;
#pragma unroll
			for (int b = 0; b < 8; ++b)
			{
				const uint32_t bi = w * 8 + b;
				if (bi < a.cbx && a.brick_flags[bi + a.cbx * (cy + a.cby * cz)]) m |= 0xffull << (8 * b);
			}
			act &= m;
		}
		const size_t t = (size_t(k) * a.sy + j) * a.aw + w;
		a.active[t] = act;
		a.word_count[t] = uint32_t(__popcll(act));
	}
}

// Vertex of active cube (i, j, k): the mean of the edge crossings, mapped to
// world space.
// Corner samples of cube (i, j, k) in the reference cube numbering.
__device__ __forceinline__ void cube_corners(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k, float (&s)[8])
{
	s[0] = pool_at(a, i, j, k);
	s[1] = pool_at(a, i + 1, j, k);
	s[2] = pool_at(a, i + 1, j + 1, k);
	s[3] = pool_at(a, i, j + 1, k);
	s[4] = pool_at(a, i, j, k + 1);
	s[5] = pool_at(a, i + 1, j, k + 1);
	s[6] = pool_at(a, i + 1, j + 1, k + 1);
	s[7] = pool_at(a, i, j + 1, k + 1);
}

__device__ __forceinline__ void place_vertex(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k, const float (&s)[8],
	float* v)
{
	const float maxx = a.gx + float(a.sx) * a.dx;
	const float maxy = a.gy + float(a.sy) * a.dy;
	const float maxz = a.gz + float(a.sz) * a.dz;
	const float fsx = float(a.sx), fsy = float(a.sy), fsz = float(a.sz);
	// corner offsets (x, y, z) and the 12 edges of the reference cube numbering
	constexpr uint8_t cx[8] = { 0, 1, 1, 0, 0, 1, 1, 0 };
	constexpr uint8_t cy[8] = { 0, 0, 1, 1, 0, 0, 1, 1 };
	constexpr uint8_t cz[8] = { 0, 0, 0, 0, 1, 1, 1, 1 };
	constexpr uint8_t e0[12] = { 0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3 };
	constexpr uint8_t e1[12] = { 1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7 };
	const float fi = float(i), fj = float(j), fk = float(k);
	float sx_ = 0.0f, sy_ = 0.0f, sz_ = 0.0f;
	int n = 0;
	// Every edge is computed and a non-crossing one adds +0.0 (no branches:
	// a wave's cubes cross most edges between them).  The sums start at +0.0
	// and every term is >= 0 (coordinates of the cube's corners, t in
	// [0, 1]), so adding +0.0 leaves them bit for bit as skipping would.
#pragma unroll
	for (int e = 0; e < 12; ++e)
	{
		const float v1 = s[e0[e]], v2 = s[e1[e]];
		const bool cross = (v1 >= 0.0f) != (v2 >= 0.0f);
		const float p1x = fi + float(cx[e0[e]]), p1y = fj + float(cy[e0[e]]), p1z = fk + float(cz[e0[e]]);
		const float p2x = fi + float(cx[e1[e]]), p2y = fj + float(cy[e1[e]]), p2z = fk + float(cz[e1[e]]);
		const float tt = (0.0f - v1) / (v2 - v1);
		sx_ = sx_ + (cross ? p1x + tt * (p2x - p1x) : 0.0f);
		sy_ = sy_ + (cross ? p1y + tt * (p2y - p1y) : 0.0f);
		sz_ = sz_ + (cross ? p1z + tt * (p2z - p1z) : 0.0f);
		n += cross ? 1 : 0;
	}
	const float fn = float(n);
	const float gcx = sx_ / fn, gcy = sy_ / fn, gcz = sz_ / fn;
	v[0] = a.gx + (maxx - a.gx) * (gcx - 0.0f) / (fsx - 0.0f);
	v[1] = a.gy + (maxy - a.gy) * (gcy - 0.0f) / (fsy - 0.0f);
	v[2] = a.gz + (maxz - a.gz) * (gcz - 0.0f) / (fsz - 0.0f);
}

// Thread per four cube-row words (one 32-byte load when whole): each word's
// active cubes into the cube list at their vertex ids.
constexpr uint32_t kListWords = 4;

__global__ void __launch_bounds__(256) k_list_cubes(MeshArgs a)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint64_t total = uint64_t(a.sy) * a.sz * a.aw;
	const uint64_t w0 = t * kListWords;
	if (w0 >= total) return;
	uint64_t act[kListWords] = { 0, 0, 0, 0 };
	if (w0 + kListWords <= total)
	{
		const ulonglong2* src = reinterpret_cast<const ulonglong2*>(a.active + w0);
		const ulonglong2 lo = src[0], hi = src[1];
		act[0] = lo.x; act[1] = lo.y; act[2] = hi.x; act[3] = hi.y;
	}
	else
		for (uint32_t q = 0; q < kListWords && w0 + q < total; ++q) act[q] = a.active[w0 + q];
#pragma unroll
	for (uint32_t q = 0; q < kListWords; ++q)
	{
		uint64_t bits = act[q];
		if (!bits) continue;
		const uint32_t wi = uint32_t(w0 + q);  // words < 2^31
		const uint32_t row = div_magic(wi, a.aw_div);
		const uint32_t w = wi - row * a.aw;
		const uint32_t base_index = w * 64 + row * a.sx;  // i + sx (j + sy k) for i = 64 w
		uint32_t id = a.word_base[wi];
		while (bits)
		{
			const int b = __ffsll((long long)bits) - 1;
			bits &= bits - 1;
			if (id < a.cap_verts) a.cubes[id] = base_index + uint32_t(b);
			++id;
		}
	}
}

__device__ __forceinline__ void cube_coords(const MeshArgs& a, uint32_t c, uint32_t& i, uint32_t& j, uint32_t& k)
{
	const uint32_t r = div_magic(c, a.sx_div);
	i = c - r * a.sx;
	k = div_magic(r, a.sy_div);
	j = r - k * a.sy;
}

// The three quads of a cube, one per directed edge (0,4), (3,0), (0,1), over
// its neighbour cubes 0..5 (surface_nets.cpp:415-423): cube c is
// (i - (c == 0 || c == 1 || c == 5), j - (1 <= c && c <= 3), k - (c >= 3)).
constexpr uint8_t kQuadNbr[3][3] = { { 0, 1, 2 }, { 0, 5, 4 }, { 2, 3, 4 } };

// Quads of one active cube as flags: bit q (q < 3) when quad q exists (its
// three neighbour cubes are active), bit 3 + q when the quad's edge has its
// second value greater than its first (vertex order {0,1,2} rather than
// {2,1,0}).  s: the cube's corner samples.
__device__ __forceinline__ uint32_t quad_flags(const MeshArgs& a, uint32_t i, uint32_t j, uint32_t k, const float (&s)[8])
{
	if (i == 0 || j == 0 || k == 0) return 0;
	const uint32_t nx[6] = { i - 1, i - 1, i, i, i, i - 1 };
	const uint32_t ny[6] = { j, j - 1, j - 1, j - 1, j, j };
	const uint32_t nz[6] = { k, k, k, k - 1, k - 1, k - 1 };
	// edge (first, second) values: (0,4) = samples (i,j,k), (i,j,k+1);
	// (3,0) = (i,j+1,k), (i,j,k); (0,1) = (i,j,k), (i+1,j,k)
	const bool fwd[3] = { s[4] > s[0], s[0] > s[3], s[1] > s[0] };
	uint32_t flags = 0;
#pragma unroll
	for (int q = 0; q < 3; ++q)
	{
		bool all = true;
#pragma unroll
		for (int t = 0; t < 3; ++t) all = all && cube_active(a, nx[kQuadNbr[q][t]], ny[kQuadNbr[q][t]], nz[kQuadNbr[q][t]]);
		if (all) flags |= (1u << q) | (fwd[q] ? 8u << q : 0u);
	}
	return flags;
}

// Active cubes (vertices) of the classified grid: the scanned word counts' total.
__device__ __forceinline__ uint64_t active_cubes(const MeshArgs& a)
{
	const uint64_t last = uint64_t(a.sy) * a.sz * a.aw - 1;
	return uint64_t(a.word_base[last]) + a.word_count[last];
}

// Thread per active cube (launched for n >= the active cubes, within the
// room cap_verts): vertex placement and the cube's quad flags (their quad
// counts are scanned before k_quad_write); threads past the active cubes
// write zero flags.
__global__ void __launch_bounds__(256) k_vertices(MeshArgs a, uint64_t n)
{
	const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (id >= n) return;
	if (id == 0 && a.tail)
	{
		const uint64_t last = uint64_t(a.sy) * a.sz * a.aw - 1;
		a.tail[0] = a.word_base[last];
		a.tail[1] = a.word_count[last];
	}
	if (id >= active_cubes(a))
	{
		a.cube_quads[id] = 0;
		return;
	}
	uint32_t i, j, k;
	cube_coords(a, a.cubes[id], i, j, k);
	float s[8];
	cube_corners(a, i, j, k, s);
	place_vertex(a, i, j, k, s, a.vertices + size_t(id) * 3);
	a.cube_quads[id] = quad_flags(a, i, j, k, s);
}

// Thread per active cube (vertex `id`: active cubes are listed in vertex
// order; launched like k_vertices): its quads at the scanned base, each as
// the two triangles (v0, v1, v2), (v0, v2, v3), when they fit the room
// cap_quads.
__global__ void __launch_bounds__(256) k_quad_write(MeshArgs a, uint64_t n)
{
	const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (id >= n) return;
	if (id == 0 && a.tail)
	{
		a.tail[2] = a.cube_qbase[n - 1];
		a.tail[3] = a.cube_quads[n - 1];
	}
	const uint32_t flags = a.cube_quads[id];
	if ((flags & 7u) == 0) return;
	const uint32_t base = a.cube_qbase[id];
	if (uint64_t(base) + uint32_t(__popc(flags & 7u)) > a.cap_quads) return;  // no room: the caller redoes the mesh
	uint32_t i, j, k;
	cube_coords(a, a.cubes[id], i, j, k);
	uint32_t* out = a.triangles + size_t(base) * 6;
#pragma unroll
	for (int q = 0; q < 3; ++q)
	{
		if (!((flags >> q) & 1u)) continue;
		uint32_t n[3];
#pragma unroll
		for (int t = 0; t < 3; ++t)
		{
			const int c = kQuadNbr[q][t];
			const uint32_t dx = (c == 0 || c == 1 || c == 5) ? 1u : 0u;
			const uint32_t dy = (1 <= c && c <= 3) ? 1u : 0u;
			const uint32_t dz = c >= 3 ? 1u : 0u;
			n[t] = vertex_id(a, i - dx, j - dy, k - dz);
		}
		const bool fwd = (flags >> (3 + q)) & 1u;
		const uint32_t v0 = uint32_t(id);
		const uint32_t v1 = fwd ? n[0] : n[2], v2 = n[1], v3 = fwd ? n[2] : n[0];
		// 24 bytes at an 8-byte aligned place (quad bases are whole quads), as
		// three 8-byte stores
		uint2* tri = reinterpret_cast<uint2*>(out);
		tri[0] = make_uint2(v0, v1);
		tri[1] = make_uint2(v2, v0);
		tri[2] = make_uint2(v2, v3);
		out += 6;
	}
}

// Pool tiles -> dense x-fastest grid (for copying samples back to the host).
__global__ void __launch_bounds__(256) k_pool_to_dense(MeshArgs a, float* dense, uint32_t k0, uint32_t nk)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint64_t plane = uint64_t(a.nx) * a.ny;
	if (t >= plane * nk) return;
	const uint32_t i = uint32_t(t % a.nx);
	const uint32_t j = uint32_t((t / a.nx) % a.ny);
	const uint32_t k = k0 + uint32_t(t / plane);
	dense[t] = pool_at(a, i, j, k);
}

// Dense x-fastest grid -> pool tiles and sign rows (host-supplied samples).
__global__ void __launch_bounds__(256) k_dense_to_pool(MeshArgs a, const float* dense, float* pool)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint64_t plane = uint64_t(a.nx) * a.ny;
	if (t >= plane * a.nz) return;
	const uint32_t i = uint32_t(t % a.nx);
	const uint32_t j = uint32_t((t / a.nx) % a.ny);
	const uint32_t k = uint32_t(t / plane);
	const uint32_t tile = (i >> 3) + a.tx * ((j >> 3) + a.ty * (k >> 3));
	const float v = dense[t];
	pool[size_t(tile) * kTileSamples + ((k & 7) * 64 + (j & 7) * 8 + (i & 7))] = v;
}

__global__ void __launch_bounds__(256) k_dense_signs(MeshArgs a, const float* dense, uint8_t* signs)
{
	// Thread per sign byte: tile t / 64, row (j & 7) = t & 7, (k & 7) = (t / 8) & 7.
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= uint64_t(a.tx) * a.ty * a.tz * 64) return;
	const uint64_t tile = t / 64;
	const uint32_t tix = uint32_t(tile % a.tx), tiy = uint32_t((tile / a.tx) % a.ty);
	const uint32_t tiz = uint32_t(tile / (uint64_t(a.tx) * a.ty));
	const uint32_t j = tiy * 8 + uint32_t(t & 7), k = tiz * 8 + uint32_t((t >> 3) & 7);
	uint8_t byte = 0;
	if (j < a.ny && k < a.nz)
		for (int x = 0; x < 8; ++x)
		{
			const uint32_t i = tix * 8 + x;
			if (i < a.nx && dense[(size_t(k) * a.ny + j) * a.nx + i] >= 0.0f) byte |= uint8_t(1u << x);
		}
	signs[t] = byte;
}

inline unsigned blocks_for(uint64_t n, unsigned threads) { return unsigned((n + threads - 1) / threads); }

}  // namespace

hipError_t mesh_classify(const MeshArgs& a, hipStream_t s)
{
	const uint32_t nseg = (a.aw + kSegWords - 1) / kSegWords;
	const uint64_t blocks = uint64_t(nseg) * a.cby * a.cbz;
	if (blocks == 0) return hipSuccess;
	if (blocks >= (1ull << 31)) return hipErrorInvalidValue;
	hipLaunchKernelGGL(k_classify, dim3(unsigned(blocks)), dim3(256), 0, s, a, nseg);
	return hipGetLastError();
}

hipError_t mesh_list_cubes(const MeshArgs& a, hipStream_t s)
{
	const uint64_t n = uint64_t(a.sy) * a.sz * a.aw;
	hipLaunchKernelGGL(k_list_cubes, dim3(blocks_for((n + kListWords - 1) / kListWords, 256)), dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t mesh_vertices(const MeshArgs& a, uint64_t nverts, hipStream_t s)
{
	if (nverts == 0) return hipSuccess;
	hipLaunchKernelGGL(k_vertices, dim3(blocks_for(nverts, 256)), dim3(256), 0, s, a, nverts);
	return hipGetLastError();
}

hipError_t mesh_quad_write(const MeshArgs& a, uint64_t nverts, hipStream_t s)
{
	if (nverts == 0) return hipSuccess;
	hipLaunchKernelGGL(k_quad_write, dim3(blocks_for(nverts, 256)), dim3(256), 0, s, a, nverts);
	return hipGetLastError();
}

hipError_t mesh_pool_to_dense(const MeshArgs& a, float* dense, uint32_t k0, uint32_t nk, hipStream_t s)
{
	const uint64_t n = uint64_t(a.nx) * a.ny * nk;
	hipLaunchKernelGGL(k_pool_to_dense, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, dense, k0, nk);
	return hipGetLastError();
}

hipError_t mesh_dense_to_pool(const MeshArgs& a, const float* dense, hipStream_t s)
{
	const uint64_t n = uint64_t(a.nx) * a.ny * a.nz;
	hipLaunchKernelGGL(k_dense_to_pool, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, dense, const_cast<float*>(a.pool));
	const uint64_t m = uint64_t(a.tx) * a.ty * a.tz * 64;
	hipLaunchKernelGGL(k_dense_signs, dim3(blocks_for(m, 256)), dim3(256), 0, s, a, dense, const_cast<uint8_t*>(a.signs));
	return hipGetLastError();
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, size_t& temp_bytes, hipStream_t s)
{
	return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, int(n), s);
}

namespace {
struct QuadCount
{
	__host__ __device__ uint32_t operator()(uint32_t flags) const { return uint32_t(__builtin_popcount(flags & 7u)); }
};
}  // namespace

hipError_t exclusive_scan_quads(const uint32_t* flags, uint32_t* out, uint64_t n, void* temp, size_t& temp_bytes,
	hipStream_t s)
{
	hipcub::TransformInputIterator<uint32_t, QuadCount, const uint32_t*> counts(flags, QuadCount{});
	return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, out, int(n), s);
}

namespace {

// Thread per block (or per listed block): its program's cost (keys) and its
// index (values) for the sort below.
// Sort key of a block program's cost: its position on a log scale with eight
// steps per octave (8 bits), so that one radix pass orders the queue; costs
// within an eighth of an octave keep their block order (the sort is stable).
__device__ __forceinline__ uint32_t cost_key(uint32_t flops)
{
	const uint32_t c = flops + 1u - (flops == 0xffffffffu ? 1u : 0u);
	const uint32_t e = 31u - uint32_t(__builtin_clz(c));
	const uint32_t frac = e >= 3u ? (c >> (e - 3u)) & 7u : (c << (3u - e)) & 7u;
	return e * 8u + frac;
}
constexpr int kCostKeyBits = 8;

__global__ void __launch_bounds__(256) k_block_costs(const BlockDesc* desc, const uint32_t* list, uint32_t n,
	uint32_t* keys, uint32_t* values)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t b = list ? list[i] : i;
	keys[i] = cost_key(desc[b].flops);
	values[i] = b;
}

}  // namespace

namespace {

// Thread per item of a batch queue: tree t owns items [first[t], first[t + 1]);
// key = its block program's cost, value = tree << 20 | block.
__global__ void __launch_bounds__(256) k_batch_keys(const BlockEvalArgs* trees, const uint32_t* first, uint32_t ntrees,
	uint32_t total, uint32_t* keys, uint32_t* values)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= total) return;
	uint32_t lo = 0, hi = ntrees;
	while (hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) / 2;
		if (first[mid] <= i)
			lo = mid;
		else
			hi = mid;
	}
	const uint32_t b = i - first[lo];
	keys[i] = cost_key(trees[lo].desc[b].flops);
	values[i] = (lo << 20) | b;
}

// The sorted values as (tree, block) word pairs.
__global__ void __launch_bounds__(256) k_batch_items(const uint32_t* values, uint32_t total, uint32_t* items)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= total) return;
	items[2 * size_t(i)] = values[i] >> 20;
	items[2 * size_t(i) + 1] = values[i] & 0xfffffu;
}

}  // namespace

hipError_t batch_queue_by_cost(const BlockEvalArgs* trees, const uint32_t* first, uint32_t ntrees, uint32_t total,
	uint32_t* sort_buf, uint32_t* items, void* temp, size_t& temp_bytes, hipStream_t s)
{
	uint32_t* keys = sort_buf;
	uint32_t* keys_out = sort_buf + total;
	uint32_t* values = sort_buf + 2 * size_t(total);
	uint32_t* values_out = sort_buf + 3 * size_t(total);
	if (!temp)
		return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys, keys_out, values, values_out,
			int(total), 0, kCostKeyBits, s);
	if (total == 0) return hipSuccess;
	hipLaunchKernelGGL(k_batch_keys, dim3((total + 255) / 256), dim3(256), 0, s, trees, first, ntrees, total, keys,
		values);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys, keys_out, values, values_out, int(total),
		0, kCostKeyBits, s);
	if (e != hipSuccess) return e;
	hipLaunchKernelGGL(k_batch_items, dim3((total + 255) / 256), dim3(256), 0, s, values_out, total, items);
	return hipGetLastError();
}

hipError_t blocks_by_cost(const BlockDesc* desc, const uint32_t* list, uint32_t n, uint32_t* keys,
	uint32_t* values, uint32_t* order, void* temp, size_t& temp_bytes, hipStream_t s)
{
	uint32_t* keys_out = keys + n;
	if (temp)
	{
		if (n == 0) return hipSuccess;
		hipLaunchKernelGGL(k_block_costs, dim3((n + 255) / 256), dim3(256), 0, s, desc, list, n, keys, values);
		const hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys, keys_out, values, order, int(n), 0,
		kCostKeyBits, s);
}

}  // namespace tg
