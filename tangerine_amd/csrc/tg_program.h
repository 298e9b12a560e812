// Reference ProgramBuffer words -> device instruction records.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "tg_types.h"

namespace tg {

// Entry view of a program (see Entry in tg_kernels.h): rec | kind << 22 | op << 24, l, r, ord.
struct HostEntry
{
	uint32_t rec, l, r, ord;
};

struct DecodedProgram
{
	std::vector<Instr> instrs;   // ends with OP_STOP
	std::vector<HostEntry> entries;  // every record but Stop, in order
	uint32_t binary_entries = 0;
	uint32_t max_stack = 0;      // interpreter stack high-water mark
	uint32_t brushes = 0;
	uint32_t set_ops = 0;
	bool lipschitz = true;       // false if any ellipsoid (not a distance bound)
	// Most blend / flate / field-scale operators on one root-to-brush path:
	// each rounds its value once, and the pruning decisions' value-proportional
	// margin (1e-5 (|a| + |b|), about 167 roundings of 2^-24) covers at most
	// kMaxRoundingDepth of them.
	uint32_t rounding_depth = 0;
	bool prunable() const;       // lipschitz, rounding_depth <= kMaxRoundingDepth
	double flops = 0.0;          // algorithmic FP32 work per evaluation (cost table)
	// Largest world-space size of a brush: its field scale times (|its frame
	// translation| + the sum of its parameter magnitudes); bounds the operand
	// magnitudes a brush's arithmetic sees besides the point itself.
	float leaf_extent = 0.0f;
	// A box holding every brush's surface (each brush as a ball around its
	// frame's origin, radius = field scale x the sum of its parameter
	// magnitudes), widened by the largest blend threshold and flate radius;
	// scene_finite is false with a plane.  Only a placement hint (the tracer's
	// cell domain), never a bound the evaluation relies on.
	float scene_lo[3] = { 0, 0, 0 }, scene_hi[3] = { 0, 0, 0 };
	bool scene_finite = false;
};

// Validates stack discipline and opcodes the way SDFInterpreter::Eval would
// (tangerine/sdf_evaluator.cpp:1397-1600) and re-encodes the program.
// Returns false with a message on malformed input.
bool decode_words(const uint32_t* words, size_t count, DecodedProgram& out, std::string& error);

// Absolute rounding slack of the pruning decisions and the brick certificate
// for points p with |p| <= pmax (see the definition in tg_program.cpp).
float rounding_slack(const DecodedProgram& d, float pmax);

// Algorithmic cost of one evaluation of a record (adds, muls, min/max and
// selects count 1, an FMA 2, sqrt and divide 1 each).  Fixed table.
double instr_flops(const Instr& in);

}  // namespace tg
