// Internal pieces shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "tg_program.h"

namespace tg {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int ensure_device();

// Grow-only device allocation.
struct DevBuf
{
	void* ptr = nullptr;
	size_t bytes = 0;
	int ensure(size_t need);
	void release();
	template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace tg

struct tg_program
{
	uint64_t uid = 0;    // unique per uploaded program (caches of derived device data key on it)
	tg::DecodedProgram decoded;
	tg::DevBuf device;   // Instr records
	tg::DevBuf entries;  // whole program as pruning entries
};

namespace tg {
int upload_program(const DecodedProgram& d, tg_program** out);

// Progress of a cell pruning between prune_cells_start and prune_cells_finish.
struct PruneState
{
	bool pending = false;        // a pass was issued; its counters are not read yet
	bool level_parallel = false;
	uint64_t ncells = 0, waves = 0, capacity = 0;
	uint32_t cells_per_wave = 0;
	float decide_radius = 0.0f;
	size_t heights = 0;
	std::vector<uint32_t> level_first;  // first entry position of every height
	std::vector<uint32_t> level_order;  // entry positions by height (the upload's source, kept until finish)
};

// Level-1 pruned programs of a cell grid (BlockDesc per cell + entry pool).
struct CellPrograms
{
	DevBuf pool, desc, counters, scratch;
	DevBuf values, levels, lens;  // level-parallel passes: entry values per cell, entries by height, program lengths
	uint64_t ncells = 0;
	uint64_t entries = 0;     // pool entries in use
	uint32_t max_stack = 0;   // over all cells
	uint32_t max_binary = 0;  // most binary entries of a packed cell program
	uint32_t max_len = 0;     // most entries of a packed cell program
	uint64_t whole_cells = 0; // cells left with the whole program (pool overflow)
	bool valid = false;       // false: the program is not prunable here
	bool reran = false;       // the last pruning emitted again after a pool overflow (desc changed)
	int emission = 0;         // after level-parallel passes: 0 automatic, 1 sequential walk, 2 a wave per cell
	uint64_t levels_uid = 0;  // program whose entries by height `levels` holds (0: none)
	size_t levels_bytes = 0;  // ... in the allocation of this size (a grown buffer is a new one)
	PruneState state;
	unsigned long long* host = nullptr;  // pinned copy of the pass counters
	void release()
	{
		if (host) (void)hipHostFree(host);
		host = nullptr;
		pool.release();
		desc.release();
		counters.release();
		scratch.release();
		values.release();
		levels.release();
		levels_uid = 0;
		lens.release();
	}
};

int prune_cells(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3], float radius,
	const uint32_t counts[3], bool fast, CellPrograms& out);

// A cell grid to prune a program over.
struct CellGrid
{
	float origin[3], cell[3], radius;
	uint32_t counts[3];
};

// Job tables of prune_cells_many (kept between calls: the upload reads `host`).
struct PruneLaunches
{
	DevBuf dev;
	std::vector<uint8_t> host;
};

// prune_cells_start for n programs on one stream: every program's height-h
// pass in one launch per height, then every emission in one launch (instead
// of a launch per program and height).  Each is then finished with
// prune_cells_finish on the same stream.
int prune_cells_many(hipStream_t stream, const tg_program* const* progs, const CellGrid* grids, uint32_t n, bool fast,
	CellPrograms* const* outs, PruneLaunches& work);
// The same in two halves: start issues the passes on `stream` without waiting;
// finish waits for the stream, reruns the emission if the pool overflowed and
// fills the counts (same arguments).  Between them the caller may issue other
// work (other trees' pruning on other streams).
int prune_cells_start(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	float radius, const uint32_t counts[3], bool fast, CellPrograms& out);
int prune_cells_finish(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	const uint32_t counts[3], bool fast, CellPrograms& out);
}

struct tg_mesher;
struct tg_surrogate;
#include "tg_mlp.h"

namespace tg {
// Device vertices of the mesher's last export (3 floats each).
const float* mesher_vertices(const tg_mesher* m);
mlp::Net surrogate_net(const tg_surrogate* s);
// Cull threshold for bricks of half-diagonal r (< 0: not calibrated); the
// certified threshold may be lower (skipped bricks are checked exactly).
float surrogate_threshold(const tg_surrogate* s, float r, bool certified = false);
}
