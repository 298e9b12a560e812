// Host driver of the level-1 cell pruning shared by grid export and tracing.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"

namespace tg {

int prune_cells(hipStream_t stream, const tg_program* p, const float origin[3], const float cell[3], float radius,
	const uint32_t counts[3], bool fast, CellPrograms& out)
{
	const DecodedProgram& d = p->decoded;
	out.valid = false;
	if (!d.lipschitz || d.entries.empty() || d.entries.size() >= 65535) return TG_OK;
	const uint32_t whole_slots = std::max<uint32_t>(1, d.max_stack - 1);

	const uint64_t ncells = uint64_t(counts[0]) * counts[1] * counts[2];
	if (ncells == 0 || ncells >= (1ull << 31)) return TG_OK;
	// Fewer cells per wave only when there are very few cells (every wave
	// gathers the whole program, so smaller groups cost L2 traffic).
	uint32_t cpw = 64;
	while (cpw > 16 && (ncells + cpw - 1) / cpw < 256) cpw >>= 1;
	const uint64_t waves = (ncells + cpw - 1) / cpw;
	int rc = out.desc.ensure(ncells * sizeof(BlockDesc));
	if (!rc) rc = out.counters.ensure(64);
	if (!rc) rc = out.scratch.ensure(waves * prune_scratch_words(d.binary_entries, uint32_t(d.entries.size())) * 8);
	if (rc) return rc;
	// First guess: a quarter of the whole program per cell, at most 1 GiB; grown
	// to the exact total and rerun if any cell did not fit.
	uint64_t capacity = std::max<uint64_t>(out.pool.bytes / sizeof(Entry),
		std::min<uint64_t>(ncells * std::max<size_t>(d.entries.size() / 4, 16), (1ull << 30) / sizeof(Entry)));
	for (int attempt = 0; attempt < 2; ++attempt)
	{
		rc = out.pool.ensure(capacity * sizeof(Entry));
		if (rc) return rc;
		hipError_t e = hipMemsetAsync(out.counters.ptr, 0, 64, stream);
		if (e != hipSuccess) return hip_fail(e, "prune counters");
		PruneArgs pa{};
		pa.prog = static_cast<const Instr*>(p->device.ptr);
		pa.entries = p->entries.as<const Entry>();
		pa.n_entries = uint32_t(d.entries.size());
		pa.n_binary = d.binary_entries;
		pa.ox = origin[0]; pa.oy = origin[1]; pa.oz = origin[2];
		pa.cx = cell[0]; pa.cy = cell[1]; pa.cz = cell[2];
		pa.radius = radius;
		pa.nx = counts[0]; pa.ny = counts[1]; pa.nz = counts[2];
		pa.slots = whole_slots;
		pa.whole_flops = uint32_t(d.flops);
		pa.pool = out.pool.as<Entry>();
		pa.capacity = capacity;
		pa.desc = out.desc.as<BlockDesc>();
		pa.counters = out.counters.as<unsigned long long>();
		pa.scratch = out.scratch.as<uint64_t>();
		pa.cells_per_wave = cpw;
		e = fast ? fast::launch_prune_cells(pa, stream) : exact::launch_prune_cells(pa, stream);
		if (e != hipSuccess) return hip_fail(e, "cell pruning launch");
		unsigned long long c[5] = { 0, 0, 0, 0, 0 };
		e = hipMemcpyAsync(c, out.counters.ptr, sizeof(c), hipMemcpyDeviceToHost, stream);
		if (e == hipSuccess) e = hipStreamSynchronize(stream);
		if (e != hipSuccess) return hip_fail(e, "cell pruning");
		out.max_stack = uint32_t(c[1]);
		out.max_binary = uint32_t(c[2]);
		out.max_len = uint32_t(c[3]);
		out.whole_cells = c[4];
		out.entries = c[0];
		if (c[0] <= capacity || c[0] * sizeof(Entry) > (size_t(8) << 30)) break;
		capacity = c[0];
	}
	out.ncells = ncells;
	out.valid = true;
	return TG_OK;
}

}  // namespace tg
