// Host driver of the level-1 cell pruning shared by grid export and tracing.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"

#define TG_CHECK_HIP(expr, what)                         \
	do                                                   \
	{                                                    \
		hipError_t _e = (expr);                          \
		if (_e != hipSuccess) return hip_fail(_e, what); \
	} while (0)

namespace tg {

namespace {

constexpr uint32_t kMaxLevels = 64;                 // more heights: the in-kernel decision pass
constexpr uint64_t kMaxValueBytes = 2ull << 30;    // entry values kept for the level-parallel pass

// Entry positions ordered by height (leaves 0, a parent one above its highest
// child) and the first position of every height.  Returns false when the
// program has more than kMaxLevels heights.
bool entry_levels(const DecodedProgram& d, std::vector<uint32_t>& order, std::vector<uint32_t>& first)
{
	const size_t n = d.entries.size();
	std::vector<uint32_t> h(n, 0);
	uint32_t top = 0;
	for (size_t e = 0; e < n; ++e)
	{
		const HostEntry& x = d.entries[e];
		const uint32_t op = entry_op(x.rec);
		if (op - OP_SPHERE <= OP_PLANE - OP_SPHERE)
			h[e] = 0;
		else if (op == OP_FLATE || op == OP_SCALE_FIELD)
			h[e] = h[x.l] + 1;
		else
			h[e] = std::max(h[x.l], h[x.r]) + 1;
		top = std::max(top, h[e]);
		if (top >= kMaxLevels) return false;
	}
	first.assign(top + 2, 0);
	for (size_t e = 0; e < n; ++e) ++first[h[e] + 1];
	for (uint32_t k = 1; k < first.size(); ++k) first[k] += first[k - 1];
	order.assign(n, 0);
	std::vector<uint32_t> at(first.begin(), first.end() - 1);
	for (size_t e = 0; e < n; ++e) order[at[h[e]]++] = uint32_t(e);
	return true;
}

}  // namespace

// Grids of at most this many cells (twice as many for programs of 1024
// entries or more) emit a wave per cell (PruneArgs::columns).
constexpr uint64_t kColumnCells = 2048;

namespace {

// One emission pass (and, on the first, the level-parallel decision passes)
// with the pool capacity out.state.capacity; the five counters are copied to
// out.host (pinned) on `stream`.
// Emission arguments of a prepared pruning (pool capacity: state.capacity).
PruneArgs prune_args(const tg_program* p, const float origin[3], const float cell[3], const uint32_t counts[3],
	const CellPrograms& out)
{
	const DecodedProgram& d = p->decoded;
	const PruneState& ps = out.state;
	PruneArgs pa{};
	pa.prog = static_cast<const Instr*>(p->device.ptr);
	pa.entries = p->entries.as<const Entry>();
	pa.n_entries = uint32_t(d.entries.size());
	pa.n_binary = d.binary_entries;
	pa.ox = origin[0]; pa.oy = origin[1]; pa.oz = origin[2];
	pa.cx = cell[0]; pa.cy = cell[1]; pa.cz = cell[2];
	pa.radius = ps.decide_radius;
	pa.nx = counts[0]; pa.ny = counts[1]; pa.nz = counts[2];
	pa.slots = std::max<uint32_t>(1, d.max_stack - 1);
	pa.whole_flops = uint32_t(d.flops);
	pa.pool = out.pool.as<Entry>();
	pa.capacity = ps.capacity;
	pa.desc = out.desc.as<BlockDesc>();
	pa.counters = out.counters.as<unsigned long long>();
	pa.scratch = out.scratch.as<uint64_t>();
	pa.cells_per_wave = ps.cells_per_wave;
	// masks, liveness and lengths stay in scratch for the rerun after a pool overflow
	pa.masks_ready = ps.level_parallel ? 1 : 0;
	pa.counts = out.lens.as<const uint32_t>();
	// emission only: four waves per cell group while that leaves under ~8
	// waves per SIMD
	pa.splits = (ps.level_parallel && ps.waves * 4 <= 8192) ? 4u : 1u;
	// A wave per cell (a lane per entry) scans the whole program for every
	// cell: it wins on grids of few cells, whose programs are long (the walk's
	// time is the longest walk), and loses on many cells with short programs
	// (1024^3 blocks: 32,768 cells, 2.3 ms against the walk's 1.0 ms).
	const bool few = ps.ncells <= kColumnCells || (ps.ncells <= 2 * kColumnCells && d.entries.size() >= 1024);
	pa.columns = (ps.level_parallel && (out.emission == 2 || (out.emission == 0 && few))) ? 1 : 0;
	return pa;
}

// Level-parallel pass arguments (without the height's list) of a prepared pruning.
LevelArgs level_args(const PruneArgs& pa, const CellPrograms& out)
{
	const PruneState& ps = out.state;
	LevelArgs la{};
	la.prog = pa.prog;
	la.entries = pa.entries;
	la.ngroups = uint32_t(ps.waves);
	la.cells_per_wave = ps.cells_per_wave;
	la.ox = pa.ox; la.oy = pa.oy; la.oz = pa.oz;
	la.cx = pa.cx; la.cy = pa.cy; la.cz = pa.cz;
	la.radius = ps.decide_radius;
	la.nx = pa.nx; la.ny = pa.ny; la.nz = pa.nz;
	la.values = out.values.as<float>();
	la.scratch = pa.scratch;
	la.n_binary = pa.n_binary;
	la.n_entries = pa.n_entries;
	la.counts = out.lens.as<uint32_t>();
	return la;
}

// Steps of the level-parallel pass: heights 1 .. H - 1 upward, then H - 1 ..
// 1 downward (the parents count their leaf children), or, for a leaf root
// (H = 1), the root's own liveness step.
size_t level_steps(const PruneState& ps)
{
	return ps.heights >= 2 ? 2 * ps.heights - 1 : 2 * ps.heights;
}

// Height step `step` (1 .. level_steps - 1) of the level-parallel pass: values
// and decisions upward from height 1 (leaf operands are evaluated in place),
// then liveness and lengths downward.
void level_step(LevelArgs& la, const CellPrograms& out, size_t step)
{
	const PruneState& ps = out.state;
	const size_t heights = ps.heights;
	const size_t h = step < heights ? step : 2 * heights - 1 - step;
	la.liveness = step < heights ? 0 : 1;
	la.list = out.levels.as<const uint32_t>() + ps.level_first[h];
	la.count = ps.level_first[h + 1] - ps.level_first[h];
	// a few cell groups per wave while enough waves remain to fill the GPU
	la.groups_per_wave = 1;
	while (la.groups_per_wave < 8 && uint64_t(la.count) * ps.waves / (la.groups_per_wave * 2) >= 8192)
		la.groups_per_wave *= 2;
}

// One emission pass (and, on the first, the level-parallel decision passes)
// with the pool capacity out.state.capacity; the five counters are copied to
// out.host (pinned) on `stream`.
int prune_pass(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	const uint32_t counts[3], bool fast, CellPrograms& out, bool first)
{
	PruneState& ps = out.state;
	int rc = out.pool.ensure(ps.capacity * sizeof(Entry));
	if (rc) return rc;
	hipError_t e = hipMemsetAsync(out.counters.ptr, 0, 64, stream);
	if (e != hipSuccess) return hip_fail(e, "prune counters");
	const PruneArgs pa = prune_args(p, origin, cell, counts, out);
	const uint64_t waves = ps.waves;
	if (first && ps.level_parallel)
	{
		LevelArgs la = level_args(pa, out);
		e = hipMemsetAsync(out.lens.ptr, 0, waves * 64 * 4, stream);
		if (e != hipSuccess) return hip_fail(e, "length reset");
		for (size_t step = 1; step < level_steps(ps); ++step)
		{
			level_step(la, out, step);
			e = fast ? fast::launch_prune_level(la, stream) : exact::launch_prune_level(la, stream);
			if (e != hipSuccess) return hip_fail(e, "level pruning launch");
		}
	}
	e = fast ? fast::launch_prune_cells(pa, stream) : exact::launch_prune_cells(pa, stream);
	if (e != hipSuccess) return hip_fail(e, "cell pruning launch");
	e = hipMemcpyAsync(out.host, out.counters.ptr, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream);
	if (e != hipSuccess) return hip_fail(e, "cell pruning counters");
	return TG_OK;
}

}  // namespace

namespace {

// Everything of prune_cells_start before the passes: sizes, buffers and the
// upload of the entries by height.  *ready is false when the program is not
// pruned here.
int prune_prepare(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	float radius, const uint32_t counts[3], CellPrograms& out, bool* ready)
{
	const DecodedProgram& d = p->decoded;
	*ready = false;
	out.valid = false;
	out.state.pending = false;
	if (!d.prunable() || d.entries.empty() || d.entries.size() >= 65535) return TG_OK;

	const uint64_t ncells = uint64_t(counts[0]) * counts[1] * counts[2];
	if (ncells == 0 || ncells >= (1ull << 31)) return TG_OK;
	if (!out.host)
	{
		hipError_t e = hipHostMalloc(&out.host, 64, hipHostMallocDefault);
		if (e != hipSuccess)
		{
			out.host = nullptr;
			return hip_fail(e, "pinned pruning counters");
		}
	}
	PruneState& ps = out.state;
	ps.ncells = ncells;
	// decision radius: the cell radius with the Lipschitz tolerance and the
	// rounding slack for points of the cells' box
	float pmax = 0.0f;
	for (int i = 0; i < 3; ++i)
	{
		const float a = std::fabs(origin[i]), b = std::fabs(origin[i] + float(counts[i]) * cell[i]);
		pmax += std::max(a, b) * std::max(a, b);
	}
	ps.decide_radius = radius * (1.0f + 1.0e-5f) + rounding_slack(d, std::sqrt(pmax));
	// Fewer cells per wave only when there are very few cells (every wave
	// gathers the whole program, so smaller groups cost L2 traffic).
	uint32_t cpw = 64;
	while (cpw > 16 && (ncells + cpw - 1) / cpw < 256) cpw >>= 1;
	ps.cells_per_wave = cpw;
	ps.waves = (ncells + cpw - 1) / cpw;
	int rc = out.desc.ensure(ncells * sizeof(BlockDesc));
	if (!rc) rc = out.counters.ensure(64);
	if (!rc) rc = out.scratch.ensure(ps.waves * prune_scratch_words(d.binary_entries, uint32_t(d.entries.size())) * 8);
	if (rc) return rc;
	// Level-parallel decision pass when the heights are few and the entry
	// values fit: it replaces the per-wave walk over the whole program.
	std::vector<uint32_t>& level_order = ps.level_order;
	ps.level_parallel = entry_levels(d, level_order, ps.level_first) &&
	                    uint64_t(d.entries.size()) * ps.waves * 64 * 4 <= kMaxValueBytes;
	if (ps.level_parallel)
	{
		ps.heights = ps.level_first.size() - 1;
		rc = out.values.ensure(uint64_t(d.entries.size()) * ps.waves * 64 * 4);
		if (!rc) rc = out.levels.ensure(level_order.size() * 4);
		if (!rc) rc = out.lens.ensure(ps.waves * 64 * 4);
		if (rc) return rc;
		// the entries by height depend on the program alone: uploaded once per
		// program (and buffer)
		if (out.levels_uid != p->uid || out.levels_bytes != out.levels.bytes)
		{
			hipError_t e = hipMemcpyAsync(out.levels.ptr, level_order.data(), level_order.size() * 4,
				hipMemcpyHostToDevice, stream);
			if (e != hipSuccess) return hip_fail(e, "level upload");
			out.levels_uid = p->uid;
			out.levels_bytes = out.levels.bytes;
		}
	}
	// First guess: a quarter of the whole program per cell, at most 1 GiB; grown
	// to the exact total and rerun if any cell did not fit (prune_cells_finish).
	ps.capacity = std::max<uint64_t>(out.pool.bytes / sizeof(Entry),
		std::min<uint64_t>(ncells * std::max<size_t>(d.entries.size() / 4, 16), (1ull << 30) / sizeof(Entry)));
	*ready = true;
	return TG_OK;
}

}  // namespace

int prune_cells_start(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	float radius, const uint32_t counts[3], bool fast, CellPrograms& out)
{
	bool ready = false;
	int rc = prune_prepare(stream, p, origin, cell, radius, counts, out, &ready);
	if (rc || !ready) return rc;
	rc = prune_pass(stream, p, origin, cell, counts, fast, out, true);
	if (rc) return rc;
	out.state.pending = true;
	return TG_OK;
}

int prune_cells_many(hipStream_t stream, const tg_program* const* progs, const CellGrid* grids, uint32_t n, bool fast,
	CellPrograms* const* outs, PruneLaunches& work)
{
	std::vector<uint8_t> ready(n, 0);
	std::vector<PruneArgs> pas(n);
	size_t steps = 0;
	for (uint32_t i = 0; i < n; ++i)
	{
		bool r = false;
		int rc = prune_prepare(stream, progs[i], grids[i].origin, grids[i].cell, grids[i].radius, grids[i].counts,
			*outs[i], &r);
		if (rc) return rc;
		if (!r) continue;
		CellPrograms& out = *outs[i];
		ready[i] = 1;
		rc = out.pool.ensure(out.state.capacity * sizeof(Entry));
		if (rc) return rc;
		TG_CHECK_HIP(hipMemsetAsync(out.counters.ptr, 0, 64, stream), "prune counters");
		if (out.state.level_parallel)
		{
			TG_CHECK_HIP(hipMemsetAsync(out.lens.ptr, 0, out.state.waves * 64 * 4, stream), "length reset");
			steps = std::max(steps, level_steps(out.state));
		}
		pas[i] = prune_args(progs[i], grids[i].origin, grids[i].cell, grids[i].counts, out);
	}
	// host blob: per height step, its level jobs then their wave bases; then
	// the emission jobs and their block bases
	std::vector<uint8_t>& blob = work.host;
	blob.clear();
	auto put = [&](const void* src, size_t bytes) {
		const size_t at = blob.size();
		blob.resize(at + ((bytes + 15) & ~size_t(15)));
		std::memcpy(blob.data() + at, src, bytes);
		return at;
	};
	struct Launch
	{
		size_t jobs, bases;
		uint32_t njobs, units;
	};
	std::vector<Launch> level_launches;
	for (size_t step = 1; step < steps; ++step)
	{
		std::vector<LevelArgs> jobs;
		std::vector<uint32_t> base;
		uint32_t waves = 0;
		for (uint32_t i = 0; i < n; ++i)
		{
			const CellPrograms& out = *outs[i];
			if (!ready[i] || !out.state.level_parallel || step >= level_steps(out.state)) continue;
			LevelArgs la = level_args(pas[i], out);
			level_step(la, out, step);
			const uint32_t per_entry = (la.ngroups + la.groups_per_wave - 1) / la.groups_per_wave;
			base.push_back(waves);
			waves += la.count * per_entry;
			jobs.push_back(la);
		}
		if (jobs.empty()) continue;
		const size_t j = put(jobs.data(), jobs.size() * sizeof(LevelArgs));
		const size_t b = put(base.data(), base.size() * 4);
		level_launches.push_back(Launch{ j, b, uint32_t(jobs.size()), waves });
	}
	// emission jobs: the sequential walk's (a wave per group part) and the
	// column emission's (a wave per cell), one launch each
	std::vector<PruneArgs> ejobs, cjobs;
	std::vector<uint32_t> ebase, cbase;
	uint32_t blocks = 0, max_slots = 1, cells = 0, max_entries = 1;
	for (uint32_t i = 0; i < n; ++i)
	{
		if (!ready[i]) continue;
		const PruneArgs& pa = pas[i];
		const uint32_t ncells = pa.nx * pa.ny * pa.nz;
		if (pa.masks_ready && pa.columns)
		{
			cbase.push_back(cells);
			cells += ncells;
			max_entries = std::max(max_entries, pa.n_entries);
			cjobs.push_back(pa);
			continue;
		}
		const uint32_t splits = pa.masks_ready ? pa.splits : 1u;
		ebase.push_back(blocks);
		blocks += (ncells + pa.cells_per_wave - 1) / pa.cells_per_wave * splits;
		max_slots = std::max(max_slots, pa.slots);
		ejobs.push_back(pa);
	}
	cbase.push_back(cells);
	size_t ej = 0, eb = 0, cj = 0, cb = 0;
	if (!ejobs.empty())
	{
		ej = put(ejobs.data(), ejobs.size() * sizeof(PruneArgs));
		eb = put(ebase.data(), ebase.size() * 4);
	}
	if (!cjobs.empty())
	{
		cj = put(cjobs.data(), cjobs.size() * sizeof(PruneArgs));
		cb = put(cbase.data(), cbase.size() * 4);
	}
	if (!blob.empty())
	{
		int rc = work.dev.ensure(blob.size());
		if (rc) return rc;
		TG_CHECK_HIP(hipMemcpyAsync(work.dev.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice, stream),
			"pruning job upload");
	}
	const uint8_t* dev = work.dev.as<const uint8_t>();
	for (const Launch& l : level_launches)
	{
		const LevelArgs* jobs = reinterpret_cast<const LevelArgs*>(dev + l.jobs);
		const uint32_t* base = reinterpret_cast<const uint32_t*>(dev + l.bases);
		const hipError_t e = fast ? fast::launch_prune_level_multi(jobs, base, l.njobs, l.units, stream)
		                          : exact::launch_prune_level_multi(jobs, base, l.njobs, l.units, stream);
		if (e != hipSuccess) return hip_fail(e, "level pruning launch");
	}
	if (!ejobs.empty())
	{
		const PruneArgs* jobs = reinterpret_cast<const PruneArgs*>(dev + ej);
		const uint32_t* base = reinterpret_cast<const uint32_t*>(dev + eb);
		const hipError_t e =
			fast ? fast::launch_prune_cells_multi(jobs, base, uint32_t(ejobs.size()), blocks, max_slots, stream)
			     : exact::launch_prune_cells_multi(jobs, base, uint32_t(ejobs.size()), blocks, max_slots, stream);
		if (e != hipSuccess) return hip_fail(e, "cell pruning launch");
	}
	if (!cjobs.empty())
	{
		const PruneArgs* jobs = reinterpret_cast<const PruneArgs*>(dev + cj);
		const uint32_t* base = reinterpret_cast<const uint32_t*>(dev + cb);
		const hipError_t e =
			fast ? fast::launch_prune_columns_multi(jobs, base, uint32_t(cjobs.size()), cells, max_entries, stream)
			     : exact::launch_prune_columns_multi(jobs, base, uint32_t(cjobs.size()), cells, max_entries, stream);
		if (e != hipSuccess) return hip_fail(e, "cell pruning launch");
	}
	for (uint32_t i = 0; i < n; ++i)
	{
		if (!ready[i]) continue;
		CellPrograms& out = *outs[i];
		TG_CHECK_HIP(hipMemcpyAsync(out.host, out.counters.ptr, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
			stream), "cell pruning counters");
		out.state.pending = true;
	}
	return TG_OK;
}

int prune_cells_finish(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3],
	const uint32_t counts[3], bool fast, CellPrograms& out)
{
	PruneState& ps = out.state;
	if (!ps.pending) return TG_OK;
	ps.pending = false;
	out.reran = false;
	hipError_t e = hipStreamSynchronize(stream);
	if (e != hipSuccess) return hip_fail(e, "cell pruning");
	const unsigned long long* c = out.host;
	if (c[0] > ps.capacity && c[0] * sizeof(Entry) <= (size_t(8) << 30))
	{
		// some cells did not fit: rerun the emission with the exact total
		out.reran = true;
		ps.capacity = c[0];
		int rc = prune_pass(stream, p, origin, cell, counts, fast, out, false);
		if (rc) return rc;
		e = hipStreamSynchronize(stream);
		if (e != hipSuccess) return hip_fail(e, "cell pruning");
	}
	out.max_stack = uint32_t(c[1]);
	out.max_binary = uint32_t(c[2]);
	out.max_len = uint32_t(c[3]);
	out.whole_cells = c[4];
	out.entries = c[0];
	out.ncells = ps.ncells;
	out.valid = true;
	return TG_OK;
}

int prune_cells(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3], float radius,
	const uint32_t counts[3], bool fast, CellPrograms& out)
{
	int rc = prune_cells_start(stream, p, origin, cell, radius, counts, fast, out);
	if (rc) return rc;
	return prune_cells_finish(stream, p, origin, cell, counts, fast, out);
}

}  // namespace tg
