// Shared host/device definitions for the MI355X CSG distance path.
//
// Opcode numbering and word layout follow the reference ProgramBuffer
// (tangerine/sdf_evaluator.h:83-107): programs arrive as 32-bit words and are
// re-encoded here into fixed 64-byte instruction records so that one
// instruction is a single wave-uniform 16-dword scalar load.
#pragma once

#include <stdint.h>

namespace tg {

enum Opcode : uint32_t
{
	OP_STOP = 0,
	OP_SPHERE,
	OP_ELLIPSOID,
	OP_BOX,
	OP_TORUS,
	OP_CYLINDER,
	OP_CONE,
	OP_CONINDER,
	OP_PLANE,
	OP_UNION,
	OP_INTER,
	OP_DIFF,
	OP_BLEND_UNION,
	OP_BLEND_INTER,
	OP_BLEND_DIFF,
	OP_FLATE,
	OP_OFFSET,
	OP_MATRIX,
	OP_SCALE_FIELD,
	// Device-only opcodes (never appear in reference words).
	OP_NEGATE = 32,  // top = -top   (emitted by brick pruning of Diff)
	OP_COUNT
};

// Transform kinds folded into a brush record (bits 8..9 of word 0).
enum XformKind : uint32_t
{
	XF_NONE = 0,
	XF_OFFSET = 1,  // P = p + (m[3], m[7], m[11])
	XF_MATRIX = 2,  // P = rows of m (3x4), glm order ((c0 x + c1 y) + (c2 z + c3))
};

// One instruction = 16 dwords.
//   brush:  w[0] = opcode | kind << 8;  w[1..12] = 3x4 matrix rows; w[13..15] = params
//   set op: w[0] = opcode | kSwapOperands (if set);  w[1] = start of the operand
//           evaluated first, w[2] = start of the operand evaluated second;
//           f[3] = threshold; f[4] = 0.25 / threshold (blend only);
//           w[5..6] = RN(0.25 / threshold), w[7..8] = 4 threshold as doubles,
//           low word first (blend only)
//   flate / scale field: f[1] = radius / scale
//   negate / stop: w[0] only
struct Instr
{
	uint32_t w[16];
};

static_assert(sizeof(Instr) == 64, "instruction record must be 64 bytes");

// Set-operator flag (bit 30 of a record's word 0 and of an entry word): the
// right operand was evaluated first, so the left operand is on top of the
// stack.  Programs deeper than the device stack are reordered so that every
// operator evaluates its deeper operand first (Sethi-Ullman order); the
// operator still combines (left, right) as the source program does.
constexpr uint32_t kSwapOperands = 1u << 30;

inline uint32_t instr_op(const Instr& I) { return I.w[0] & 0xffu; }
inline uint32_t instr_kind(const Instr& I) { return (I.w[0] >> 8) & 3u; }

// Regular grid, same fields and meaning as isosurface::regular_grid_t
// (third_party/naive-surface-nets/include/regular_grid.h): origin, cell size
// and cell counts; samples are (s+1) per axis at origin + float(i) * d.
struct Grid
{
	float x, y, z;
	float dx, dy, dz;
	uint64_t sx, sy, sz;
};

constexpr int kTile = 8;          // samples per tile edge (8^3 = 512 samples / wave)
constexpr int kTileSamples = 512;
constexpr int kMaxStack = 16;     // interpreter stack slots per point (after operand reordering)

}  // namespace tg
