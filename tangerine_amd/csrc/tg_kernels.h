// Kernel argument blocks and host launchers for the evaluation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tg_types.h"

namespace tg {

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
	uint8_t* signs;          // sign bit rows: byte (k*ny + j)*row_bytes + tile_x
	uint32_t row_bytes;
	uint32_t slots;
};

// Launchers, one set per numerics build (exact = reference bit parity,
// fast = contracted FP32).  `p` is points per lane for the tile kernel (4 or 8).
namespace exact {
hipError_t launch_points(const PointArgs& a, hipStream_t s);
hipError_t launch_tiles(const GridArgs& a, int p, hipStream_t s);
}
namespace fast {
hipError_t launch_points(const PointArgs& a, hipStream_t s);
hipError_t launch_tiles(const GridArgs& a, int p, hipStream_t s);
}

// Dynamic LDS bytes per wave for the interpreter spill region.
inline size_t spill_bytes_per_wave(uint32_t slots, int points_per_lane)
{
	return size_t(slots) * size_t(points_per_lane) * 64 * sizeof(float);
}

}  // namespace tg
