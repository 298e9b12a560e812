#include "tg_program.h"

#include <cstring>

namespace tg {

namespace {

float as_float(uint32_t u)
{
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

uint32_t as_word(float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}

int brush_param_count(uint32_t op)
{
	switch (op)
	{
	case OP_SPHERE: return 1;
	case OP_TORUS: case OP_CYLINDER: case OP_CONE: return 2;
	case OP_ELLIPSOID: case OP_BOX: case OP_CONINDER: case OP_PLANE: return 3;
	default: return -1;
	}
}

bool is_brush(uint32_t op) { return op >= OP_SPHERE && op <= OP_PLANE; }
bool is_binary(uint32_t op) { return op >= OP_UNION && op <= OP_BLEND_DIFF; }
bool is_blend(uint32_t op) { return op >= OP_BLEND_UNION && op <= OP_BLEND_DIFF; }

}  // namespace

double instr_flops(const Instr& in)
{
	const uint32_t op = instr_op(in);
	double xf = 0.0;
	if (is_brush(op))
	{
		const uint32_t k = instr_kind(in);
		xf = (k == XF_MATRIX) ? 18.0 : (k == XF_OFFSET) ? 3.0 : 0.0;
	}
	switch (op)
	{
	case OP_SPHERE: return xf + 7.0;
	case OP_ELLIPSOID: return xf + 25.0;
	case OP_BOX: return xf + 16.0;
	case OP_TORUS: return xf + 10.0;
	case OP_CYLINDER: return xf + 15.0;
	case OP_PLANE: return xf + 5.0;
	case OP_CONE: return xf + 43.0;
	case OP_CONINDER: return xf + 38.0;
	case OP_UNION: case OP_INTER: case OP_DIFF: return 1.0;
	case OP_BLEND_UNION: case OP_BLEND_INTER: case OP_BLEND_DIFF: return 8.0;
	case OP_FLATE: case OP_SCALE_FIELD: case OP_NEGATE: return 1.0;
	default: return 0.0;
	}
}

bool decode_words(const uint32_t* w, size_t n, DecodedProgram& out, std::string& err)
{
	out = DecodedProgram();
	// Pending point transform: the last Offset/Matrix before a brush applies to it
	// (the interpreter overwrites Point from EvalPoint and resets it after each brush).
	uint32_t pending_kind = XF_NONE;
	float pending[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
	std::vector<uint32_t> starts;  // subtree start record per stack entry
	size_t pc = 0;
	bool stopped = false;
	while (pc < n)
	{
		const uint32_t op = w[pc++];
		if (op == OP_STOP)
		{
			if (pc != n)
			{
				err = "Stop is not the last word of the program";
				return false;
			}
			if (starts.size() != 1)
			{
				err = "program leaves " + std::to_string(starts.size()) + " values on the stack (expected 1)";
				return false;
			}
			Instr s{};
			s.w[0] = OP_STOP;
			out.instrs.push_back(s);
			stopped = true;
			break;
		}
		if (is_brush(op))
		{
			const int np = brush_param_count(op);
			if (pc + np > n)
			{
				err = "truncated brush parameters";
				return false;
			}
			Instr b{};
			b.w[0] = op | (pending_kind << 8);
			for (int i = 0; i < 12; ++i) b.w[1 + i] = as_word(pending[i]);
			for (int i = 0; i < np; ++i) b.w[13 + i] = w[pc + i];
			pc += np;
			starts.push_back(uint32_t(out.instrs.size()));
			out.instrs.push_back(b);
			out.brushes++;
			if (op == OP_ELLIPSOID) out.lipschitz = false;
			out.max_stack = std::max<uint32_t>(out.max_stack, uint32_t(starts.size()));
			pending_kind = XF_NONE;
			const float ident[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
			std::memcpy(pending, ident, sizeof(pending));
			continue;
		}
		if (is_binary(op))
		{
			if (starts.size() < 2)
			{
				err = "set operator with fewer than two operands on the stack";
				return false;
			}
			Instr s{};
			s.w[0] = op;
			s.w[2] = starts.back();
			starts.pop_back();
			s.w[1] = starts.back();
			if (is_blend(op))
			{
				if (pc >= n)
				{
					err = "truncated blend threshold";
					return false;
				}
				const float t = as_float(w[pc++]);
				s.w[3] = as_word(t);
				s.w[4] = as_word(0.25f / t);
			}
			out.instrs.push_back(s);
			out.set_ops++;
			continue;
		}
		if (op == OP_FLATE || op == OP_SCALE_FIELD)
		{
			if (starts.empty())
			{
				err = "unary field operator on an empty stack";
				return false;
			}
			if (pc >= n)
			{
				err = "truncated operand";
				return false;
			}
			Instr u{};
			u.w[0] = op;
			u.w[1] = w[pc++];
			out.instrs.push_back(u);
			continue;
		}
		if (op == OP_OFFSET)
		{
			if (pc + 3 > n)
			{
				err = "truncated offset";
				return false;
			}
			const float ident[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
			std::memcpy(pending, ident, sizeof(pending));
			pending[3] = as_float(w[pc]);
			pending[7] = as_float(w[pc + 1]);
			pending[11] = as_float(w[pc + 2]);
			pending_kind = XF_OFFSET;
			pc += 3;
			continue;
		}
		if (op == OP_MATRIX)
		{
			if (pc + 16 > n)
			{
				err = "truncated matrix";
				return false;
			}
			// words are column-major c*4 + r; records hold rows (r*4 + c), rows 0..2.
			for (int r = 0; r < 3; ++r)
				for (int c = 0; c < 4; ++c) pending[r * 4 + c] = as_float(w[pc + c * 4 + r]);
			pending_kind = XF_MATRIX;
			pc += 16;
			continue;
		}
		err = "unknown opcode " + std::to_string(op) + " at word " + std::to_string(pc - 1);
		return false;
	}
	if (!stopped)
	{
		err = "program terminated without a Stop word";
		return false;
	}
	if (out.max_stack > uint32_t(kMaxStack))
	{
		err = "program needs " + std::to_string(out.max_stack) + " stack slots; the device interpreter holds " +
		      std::to_string(kMaxStack);
		return false;
	}
	for (const Instr& in : out.instrs) out.flops += instr_flops(in);
	return true;
}

}  // namespace tg
