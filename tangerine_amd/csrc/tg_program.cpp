#include "tg_program.h"

#include <algorithm>
#include <cmath>
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

// Largest singular value of the 3x3 part of a record's point transform
// (power iteration on M^T M, double precision).
double spectral_norm3(const float* rows)
{
	double a[3][3] = {};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			for (int k = 0; k < 3; ++k) a[i][j] += double(rows[k * 4 + i]) * double(rows[k * 4 + j]);
	double v[3] = { 1.0, 0.7548776662, 0.5698402910 };
	double lambda = 0.0;
	for (int it = 0; it < 64; ++it)
	{
		double w[3] = { 0, 0, 0 };
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j) w[i] += a[i][j] * v[j];
		const double n = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
		if (n == 0.0) return 0.0;
		for (int i = 0; i < 3; ++i) v[i] = w[i] / n;
		lambda = n;
	}
	return std::sqrt(lambda);
}

// Reorders the records (Stop last) so that every set operator evaluates the
// operand that needs the deeper stack first (Sethi-Ullman order), flagging the
// operators whose right operand now runs first (kSwapOperands).  The stack
// then holds at most the tree's Strahler number of values: a chain of any
// length needs 2, and 17 needs a balanced tree of 65,536 brushes.  Values do
// not change: a brush record carries its own point transform, so a subtree
// evaluates to the same bits wherever it runs, and a flagged operator still
// combines (left, right).  Returns the stack high-water mark.
uint32_t reorder_operands(std::vector<Instr>& in)
{
	const size_t n = in.size() - 1;
	std::vector<uint32_t> need(n), first(n), second(n);
	std::vector<uint8_t> swap(n, 0);
	for (size_t i = 0; i < n; ++i)
	{
		const uint32_t op = instr_op(in[i]);
		if (is_brush(op))
			need[i] = 1;
		else if (is_binary(op))
		{
			// roots: the left operand ends right before the right one starts
			const uint32_t a = in[i].w[2] - 1, b = uint32_t(i) - 1;
			swap[i] = need[b] > need[a];
			first[i] = swap[i] ? b : a;
			second[i] = swap[i] ? a : b;
			need[i] = std::max(std::max(need[a], need[b]), std::min(need[a], need[b]) + 1);
		}
		else
		{
			first[i] = uint32_t(i) - 1;  // flate / field scale
			need[i] = need[i - 1];
		}
	}
	// post-order walk from the root (the last record before Stop)
	std::vector<Instr> out;
	out.reserve(in.size());
	std::vector<uint32_t> start(n);  // new index of each subtree's first record
	struct Frame
	{
		uint32_t node, state;
	};
	std::vector<Frame> todo;
	todo.push_back(Frame{ uint32_t(n - 1), 0 });
	while (!todo.empty())
	{
		const uint32_t i = todo.back().node, state = todo.back().state;
		const uint32_t op = instr_op(in[i]);
		if (state == 0) start[i] = uint32_t(out.size());
		const uint32_t arity = is_brush(op) ? 0u : is_binary(op) ? 2u : 1u;
		if (state < arity)
		{
			todo.back().state = state + 1;
			todo.push_back(Frame{ state == 0 ? first[i] : second[i], 0 });
			continue;
		}
		Instr r = in[i];
		if (arity == 2)
		{
			r.w[0] = (r.w[0] & ~kSwapOperands) | (swap[i] ? kSwapOperands : 0u);
			r.w[1] = start[first[i]];
			r.w[2] = start[second[i]];
		}
		out.push_back(r);
		todo.pop_back();
	}
	out.push_back(in[n]);
	in.swap(out);
	return need[n - 1];
}

}  // namespace

constexpr uint32_t kMaxRoundingDepth = 128;

bool DecodedProgram::prunable() const { return lipschitz && rounding_depth <= kMaxRoundingDepth; }

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

// Pruning drops an operand of a set operator when, at a cell centre c, the
// computed operand values a, b satisfy (for union) b - a >= 2 (r' + 1e-5
// (|a| + |b|)), r' = r (1 + 1e-5) + slack, r the cell radius.  Both fields
// are 1-Lipschitz (the decoder checks the transform norm to 1e-5), so their
// exact values move by at most r (1 + 1e-5) inside the cell; the computed
// values differ from the exact ones by the rounding error E(p) of the
// interpreter, which holds at c and at every sample p.  Dropping is exact when
// 2 slack >= E_a(c) + E_a(p) + E_b(c) + E_b(p).  Per brush, the float
// arithmetic (a 3x4 transform in four roundings per component, then at most
// a few dozen operations of the primitive, the divides and square roots
// included) errs by at most K u (|p| + leaf extent) in world units, u =
// 2^-24, with K = 64 taken generously; set operators propagate the larger
// operand error (each is 1-Lipschitz in either operand; min, max and negation
// round nothing, a blend adds u |value|), and flate / field scale add
// u |value|, which the value-proportional 1e-5 term covers for up to 160 such
// roundings on a path (programs with more than 128 are not pruned).  Hence slack = 2 K u (pmax + extent),
// with pmax bounding |p| over the domain; it never goes below 2e-4.  The
// oracle's float64 interpreter measures the per-brush constant on random and
// adversarial trees (tests/test_oracle.py).
float rounding_slack(const DecodedProgram& d, float pmax)
{
	const double u = std::ldexp(1.0, -24), K = 64.0;
	const double s = 2.0 * K * u * (double(std::fabs(pmax)) + double(d.leaf_extent));
	return float(std::max(s, 2.0e-4));
}

bool decode_words(const uint32_t* w, size_t n, DecodedProgram& out, std::string& err)
{
	out = DecodedProgram();
	// Pending point transform: the last Offset/Matrix before a brush applies to it
	// (the interpreter overwrites Point from EvalPoint and resets it after each brush).
	uint32_t pending_kind = XF_NONE;
	float pending[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
	std::vector<uint32_t> starts;  // subtree start record per stack entry
	double last_brush_norm = 1.0;  // transform norm of the brush just emitted
	double last_brush_extent = 0.0;  // its extent in its own frame (scaled by a following ScaleField)
	double last_centre[3] = { 0, 0, 0 };  // world position of its frame's origin
	double last_radius = 0.0;        // its parameter magnitudes in its own frame
	double widen = 0.0;              // largest blend threshold / flate radius
	bool any_plane = false, any_brush = false;
	double box_lo[3] = { 1e300, 1e300, 1e300 }, box_hi[3] = { -1e300, -1e300, -1e300 };
	auto close_brush = [&](double scale) {
		const double r = last_radius * scale;
		for (int k = 0; k < 3; ++k)
		{
			box_lo[k] = std::min(box_lo[k], last_centre[k] - r);
			box_hi[k] = std::max(box_hi[k], last_centre[k] + r);
		}
		any_brush = true;
	};
	bool brush_open = false;  // a brush whose field scale may still follow
	size_t pc = 0;
	bool stopped = false;
	while (pc < n)
	{
		const uint32_t op = w[pc++];
		if (op != OP_SCALE_FIELD)
		{
			if (last_brush_norm > 1.0 + 1e-5) out.lipschitz = false;
			last_brush_norm = 1.0;
			if (brush_open) close_brush(1.0);
			brush_open = false;
		}
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
			// Pruning relies on every brush being a 1-Lipschitz field: exact
			// distance functions seen through a rigid (rotation + uniform scale,
			// undone by ScaleField) transform.
			if (op == OP_ELLIPSOID) out.lipschitz = false;
			if (op == OP_PLANE)
			{
				const double nx = as_float(b.w[13]), ny = as_float(b.w[14]), nz = as_float(b.w[15]);
				if (std::sqrt(nx * nx + ny * ny + nz * nz) > 1.0 + 1e-5) out.lipschitz = false;
			}
			// checked at the next record, which may be the ScaleField undoing a scale
			last_brush_norm = (pending_kind == XF_MATRIX) ? spectral_norm3(pending) : 1.0;
			{
				double mag = std::sqrt(double(pending[3]) * pending[3] + double(pending[7]) * pending[7] +
				                       double(pending[11]) * pending[11]);
				for (int i = 0; i < np; ++i) mag += std::fabs(double(as_float(w[pc - np + i])));
				if (op == OP_CONE) mag += std::fabs(double(as_float(w[pc - np])) * as_float(w[pc - np + 1]));
				last_brush_extent = mag;
				out.leaf_extent = std::max(out.leaf_extent, float(mag));
				// frame origin in world space: q = A p + b vanishes at p = -A^-1 b
				double radius = 0.0;
				for (int i = 0; i < np; ++i) radius += std::fabs(double(as_float(w[pc - np + i])));
				if (op == OP_CONE) radius += std::fabs(double(as_float(w[pc - np])) * as_float(w[pc - np + 1]));
				last_radius = radius;
				if (op == OP_PLANE) any_plane = true;
				const double b[3] = { pending[3], pending[7], pending[11] };
				if (pending_kind == XF_MATRIX)
				{
					const double A[3][3] = { { pending[0], pending[1], pending[2] }, { pending[4], pending[5], pending[6] },
					                         { pending[8], pending[9], pending[10] } };
					const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
					                   A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
					                   A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
					if (std::fabs(det) > 1e-30)
					{
						// Cramer's rule for A c = -b
						for (int k = 0; k < 3; ++k)
						{
							double M[3][3];
							for (int r = 0; r < 3; ++r)
								for (int c = 0; c < 3; ++c) M[r][c] = (c == k) ? -b[r] : A[r][c];
							last_centre[k] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
							                  M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
							                  M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / det;
						}
					}
					else
						any_plane = true;  // degenerate frame: no finite placement
				}
				else
					for (int k = 0; k < 3; ++k) last_centre[k] = -b[k];
				brush_open = true;
			}
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
				widen = std::max(widen, std::fabs(double(t)));
				s.w[3] = as_word(t);
				s.w[4] = as_word(0.25f / t);
				// exact build: RN(0.25 / t) and 4 t in double, low word first
				const double yq = 0.25 / double(t), t4 = 4.0 * double(t);
				uint64_t b[2];
				std::memcpy(&b[0], &yq, 8);
				std::memcpy(&b[1], &t4, 8);
				s.w[5] = uint32_t(b[0]);
				s.w[6] = uint32_t(b[0] >> 32);
				s.w[7] = uint32_t(b[1]);
				s.w[8] = uint32_t(b[1] >> 32);
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
			if (op == OP_SCALE_FIELD)
			{
				// field scale s after a transform of norm n: Lipschitz n * |s|
				const double ls = last_brush_norm * std::fabs(double(as_float(u.w[1])));
				if (ls > 1.0 + 1e-5) out.lipschitz = false;
				out.leaf_extent = std::max(out.leaf_extent,
				                           float(last_brush_extent * std::fabs(double(as_float(u.w[1])))));
				if (brush_open) close_brush(std::fabs(double(as_float(u.w[1]))));
				brush_open = false;
				last_brush_norm = 1.0;
			}
			else
				widen = std::max(widen, 2.0 * std::fabs(double(as_float(u.w[1]))));
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
	out.scene_finite = any_brush && !any_plane;
	if (out.scene_finite)
		for (int k = 0; k < 3; ++k)
		{
			out.scene_lo[k] = float(box_lo[k] - widen);
			out.scene_hi[k] = float(box_hi[k] + widen);
		}
	if (out.max_stack > uint32_t(kMaxStack)) out.max_stack = reorder_operands(out.instrs);
	if (out.max_stack > uint32_t(kMaxStack))
	{
		err = "program needs " + std::to_string(out.max_stack) + " stack slots in any operand order; the device "
		      "interpreter holds " + std::to_string(kMaxStack);
		return false;
	}
	// Record cost in bits 16..23 of word 0 (device-side work accounting), and
	// the entry view with operand roots for per-cell pruning.
	uint32_t ord = 0;
	std::vector<uint32_t> rdepth(out.instrs.size(), 0);
	for (size_t i = 0; i < out.instrs.size(); ++i)
	{
		Instr& in = out.instrs[i];
		const double c = instr_flops(in);
		out.flops += c;
		in.w[0] = (in.w[0] & (0xffffu | kSwapOperands)) | (uint32_t(c) << 16);
		const uint32_t op = instr_op(in);
		if (op == OP_STOP) break;
		HostEntry e{ uint32_t(i) | (instr_kind(in) << 22) | (op << 24) | (in.w[0] & kSwapOperands), 0, 0, 0 };
		if (is_binary(op))
		{
			// the operand evaluated first ends right before the second one starts
			const uint32_t below = in.w[2] - 1, top = uint32_t(i) - 1;
			const bool swapped = (in.w[0] & kSwapOperands) != 0;
			e.l = swapped ? top : below;
			e.r = swapped ? below : top;
			e.ord = ord++;
			rdepth[i] = std::max(rdepth[e.l], rdepth[e.r]) + (is_blend(op) ? 1u : 0u);
		}
		else if (op == OP_FLATE || op == OP_SCALE_FIELD)
		{
			e.l = uint32_t(i) - 1;
			rdepth[i] = rdepth[e.l] + 1;
		}
		out.rounding_depth = std::max(out.rounding_depth, rdepth[i]);
		out.entries.push_back(e);
	}
	out.binary_entries = ord;
	if (out.instrs.size() >= (1u << 22))
	{
		err = "program exceeds 2^22 records";
		return false;
	}
	return true;
}

}  // namespace tg
