// Host CSG tree for the device path.  Semantics follow the reference node
// family; the comments cite the file:line each rule comes from.
//
// Matrix words: the reference emits inverse(T * S * R) computed in float with
// glm 0.9.9.8 (tangerine/sdf_evaluator.cpp:409-429, transform.cpp:49-55).  To
// hand the GPU the same bits the reference interpreter sees, the matrix below
// is built with glm's published operation order: mat3_cast of the quaternion
// (gtc/quaternion.inl:41-72), translate (ext/matrix_transform.inl), scale_slow
// = m * diag(s) (ext/matrix_transform.inl:89-96), mat4 products summed left to
// right (detail/type_mat4x4.inl:630-648) and the cofactor inverse
// (detail/func_matrix.inl:294-351).

#include "tg_tree.h"
#include "tg_types.h"

#include <cmath>
#include <cstring>
#include <limits>

namespace tg {

namespace {

// Column-major 4x4: col[c][r].
struct Mat4
{
	float col[4][4];
};

Mat4 identity4()
{
	Mat4 m{};
	for (int i = 0; i < 4; ++i) m.col[i][i] = 1.0f;
	return m;
}

Mat4 mul4(const Mat4& a, const Mat4& b)
{
	Mat4 r{};
	for (int c = 0; c < 4; ++c)
	{
		for (int row = 0; row < 4; ++row)
		{
			float acc = a.col[0][row] * b.col[c][0];
			acc = acc + a.col[1][row] * b.col[c][1];
			acc = acc + a.col[2][row] * b.col[c][2];
			acc = acc + a.col[3][row] * b.col[c][3];
			r.col[c][row] = acc;
		}
	}
	return r;
}

Mat4 rotation4(Quat q)
{
	const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
	const float xz = q.x * q.z, xy = q.x * q.y, yz = q.y * q.z;
	const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
	Mat4 m = identity4();
	m.col[0][0] = 1.0f - 2.0f * (yy + zz);
	m.col[0][1] = 2.0f * (xy + wz);
	m.col[0][2] = 2.0f * (xz - wy);
	m.col[1][0] = 2.0f * (xy - wz);
	m.col[1][1] = 1.0f - 2.0f * (xx + zz);
	m.col[1][2] = 2.0f * (yz + wx);
	m.col[2][0] = 2.0f * (xz + wy);
	m.col[2][1] = 2.0f * (yz - wx);
	m.col[2][2] = 1.0f - 2.0f * (xx + yy);
	return m;
}

Mat4 inverse4(const Mat4& M)
{
	// 2x2 minors of the lower rows, named by the column pair and row pair they span.
	auto e = [&](int c, int r) { return M.col[c][r]; };
	const float k00 = e(2,2) * e(3,3) - e(3,2) * e(2,3);
	const float k02 = e(1,2) * e(3,3) - e(3,2) * e(1,3);
	const float k03 = e(1,2) * e(2,3) - e(2,2) * e(1,3);
	const float k04 = e(2,1) * e(3,3) - e(3,1) * e(2,3);
	const float k06 = e(1,1) * e(3,3) - e(3,1) * e(1,3);
	const float k07 = e(1,1) * e(2,3) - e(2,1) * e(1,3);
	const float k08 = e(2,1) * e(3,2) - e(3,1) * e(2,2);
	const float k10 = e(1,1) * e(3,2) - e(3,1) * e(1,2);
	const float k11 = e(1,1) * e(2,2) - e(2,1) * e(1,2);
	const float k12 = e(2,0) * e(3,3) - e(3,0) * e(2,3);
	const float k14 = e(1,0) * e(3,3) - e(3,0) * e(1,3);
	const float k15 = e(1,0) * e(2,3) - e(2,0) * e(1,3);
	const float k16 = e(2,0) * e(3,2) - e(3,0) * e(2,2);
	const float k18 = e(1,0) * e(3,2) - e(3,0) * e(1,2);
	const float k19 = e(1,0) * e(2,2) - e(2,0) * e(1,2);
	const float k20 = e(2,0) * e(3,1) - e(3,0) * e(2,1);
	const float k22 = e(1,0) * e(3,1) - e(3,0) * e(1,1);
	const float k23 = e(1,0) * e(2,1) - e(2,0) * e(1,1);

	const float F[6][4] = {
		{ k00, k00, k02, k03 }, { k04, k04, k06, k07 }, { k08, k08, k10, k11 },
		{ k12, k12, k14, k15 }, { k16, k16, k18, k19 }, { k20, k20, k22, k23 } };
	// V[r][i]: row r of the first two columns arranged as (m1r, m0r, m0r, m0r).
	float V[4][4];
	for (int r = 0; r < 4; ++r)
	{
		V[r][0] = e(1, r);
		V[r][1] = V[r][2] = V[r][3] = e(0, r);
	}
	const float sign_a[4] = { 1, -1, 1, -1 };
	const float sign_b[4] = { -1, 1, -1, 1 };
	Mat4 inv{};
	for (int i = 0; i < 4; ++i)
	{
		inv.col[0][i] = ((V[1][i] * F[0][i] - V[2][i] * F[1][i]) + V[3][i] * F[2][i]) * sign_a[i];
		inv.col[1][i] = ((V[0][i] * F[0][i] - V[2][i] * F[3][i]) + V[3][i] * F[4][i]) * sign_b[i];
		inv.col[2][i] = ((V[0][i] * F[1][i] - V[1][i] * F[3][i]) + V[3][i] * F[5][i]) * sign_a[i];
		inv.col[3][i] = ((V[0][i] * F[2][i] - V[1][i] * F[4][i]) + V[2][i] * F[5][i]) * sign_b[i];
	}
	const float p0 = M.col[0][0] * inv.col[0][0];
	const float p1 = M.col[0][1] * inv.col[1][0];
	const float p2 = M.col[0][2] * inv.col[2][0];
	const float p3 = M.col[0][3] * inv.col[3][0];
	const float det = (p0 + p1) + (p2 + p3);
	const float rdet = 1.0f / det;
	for (int c = 0; c < 4; ++c)
		for (int r = 0; r < 4; ++r) inv.col[c][r] = inv.col[c][r] * rdet;
	return inv;
}

Mat4 transform_matrix(const Transform3& t)
{
	Mat4 rot = rotation4(t.rotation);
	// translate(identity, t): column 3 = ((c0 tx + c1 ty) + c2 tz) + c3
	Mat4 tr = identity4();
	const Mat4 id = identity4();
	for (int r = 0; r < 4; ++r)
	{
		float acc = id.col[0][r] * t.translation.x;
		acc = acc + id.col[1][r] * t.translation.y;
		acc = acc + id.col[2][r] * t.translation.z;
		tr.col[3][r] = acc + id.col[3][r];
	}
	Mat4 sc = identity4();
	sc.col[0][0] = sc.col[1][1] = sc.col[2][2] = t.scale;
	return mul4(mul4(tr, sc), rot);
}

Vec3 cross(Vec3 a, Vec3 b)
{
	return { a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y };
}

// q * v in glm's order (detail/type_quat.inl:343-350).
Vec3 rotate(Quat q, Vec3 v)
{
	const Vec3 u{ q.x, q.y, q.z };
	const Vec3 uv = cross(u, v);
	const Vec3 uuv = cross(u, uv);
	const Vec3 s{ uv.x * q.w + uuv.x, uv.y * q.w + uuv.y, uv.z * q.w + uuv.z };
	return { v.x + s.x * 2.0f, v.y + s.y * 2.0f, v.z + s.z * 2.0f };
}

Quat qmul(Quat p, Quat q)
{
	return { p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z,
	         p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
	         p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z,
	         p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x };
}

float vmin(float a, float b) { return (b < a) ? b : a; }
float vmax(float a, float b) { return (a < b) ? b : a; }
Vec3 vmin3(Vec3 a, Vec3 b) { return { vmin(a.x, b.x), vmin(a.y, b.y), vmin(a.z, b.z) }; }
Vec3 vmax3(Vec3 a, Vec3 b) { return { vmax(a.x, b.x), vmax(a.y, b.y), vmax(a.z, b.z) }; }

bool is_identity(Quat q) { return q.w == 1.0f && q.x == 0.0f && q.y == 0.0f && q.z == 0.0f; }
bool is_blend(uint32_t op) { return op == OP_BLEND_UNION || op == OP_BLEND_INTER || op == OP_BLEND_DIFF; }
bool is_diff(uint32_t op) { return op == OP_DIFF || op == OP_BLEND_DIFF; }
bool is_union(uint32_t op) { return op == OP_UNION || op == OP_BLEND_UNION; }
bool is_inter(uint32_t op) { return op == OP_INTER || op == OP_BLEND_INTER; }

int brush_params(uint32_t op)
{
	return op == OP_SPHERE ? 1 : (op == OP_TORUS || op == OP_CYLINDER || op == OP_CONE) ? 2 : 3;
}

NodeRef brush(uint32_t op, float a, float b, float c, Vec3 half)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::Kind::Brush;
	n->opcode = op;
	n->params[0] = a; n->params[1] = b; n->params[2] = c;
	n->local_bounds = { { -half.x, -half.y, -half.z }, half };
	n->stack_size = 1;
	n->leaves = 1;
	return n;
}

void push_float(std::vector<uint32_t>& w, float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	w.push_back(u);
}

Vec3 apply(const Transform3& t, Vec3 p)
{
	Vec3 r = rotate(t.rotation, { p.x * t.scale, p.y * t.scale, p.z * t.scale });
	return { r.x + t.translation.x, r.y + t.translation.y, r.z + t.translation.z };
}

}  // namespace

NodeRef make_sphere(float r) { return brush(OP_SPHERE, r, 0, 0, { r, r, r }); }
NodeRef make_ellipsoid(float x, float y, float z) { return brush(OP_ELLIPSOID, x, y, z, { x, y, z }); }
NodeRef make_box(float x, float y, float z) { return brush(OP_BOX, x, y, z, { x, y, z }); }

NodeRef make_torus(float major, float minor)
{
	const float outer = major + minor;
	return brush(OP_TORUS, major, minor, 0, { outer, outer, minor });
}

NodeRef make_cylinder(float r, float h) { return brush(OP_CYLINDER, r, h, 0, { r, r, h }); }

NodeRef make_plane(float x, float y, float z)
{
	const float len = std::sqrt((x * x + y * y) + z * z);
	const float nx = x / len, ny = y / len, nz = z / len;
	const float inf = std::numeric_limits<float>::infinity();
	NodeRef n = brush(OP_PLANE, nx, ny, nz, { inf, inf, inf });
	// Half-infinite bounds only for axis normals (sdf_evaluator.cpp:1268-1292).
	Box3& b = n->local_bounds;
	if (nx == -1.0f) b.lo.x = 0.0f;
	else if (nx == 1.0f) b.hi.x = 0.0f;
	else if (ny == -1.0f) b.lo.y = 0.0f;
	else if (ny == 1.0f) b.hi.y = 0.0f;
	else if (nz == -1.0f) b.lo.z = 0.0f;
	else if (nz == 1.0f) b.hi.z = 0.0f;
	return n;
}

NodeRef make_cone(float r, float h)
{
	// Stored as {tangent = r/h, h}; bounds half-height h*.5 (sdf_evaluator.cpp:1296-1305).
	return brush(OP_CONE, r / h, h, 0, { r, r, float(double(h) * .5) });
}

NodeRef make_coninder(float lo, float hi, float height)
{
	const float half = float(double(height) * .5);
	const float m = std::fmax(lo, hi);
	return brush(OP_CONINDER, lo, hi, half, { m, m, half });
}

NodeRef make_set(uint32_t op, NodeRef l, NodeRef r, float t)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::Kind::Set;
	n->opcode = op;
	n->threshold = t;
	// Left-lean swap for commutative families (sdf_evaluator.cpp:759-771).
	if (!is_diff(op) && r->stack_size > l->stack_size) std::swap(l, r);
	n->stack_size = std::max(l->stack_size, r->stack_size + 1);
	n->leaves = l->leaves + r->leaves;
	n->lhs = std::move(l);
	n->rhs = std::move(r);
	return n;
}

NodeRef make_flate(NodeRef c, float radius)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::Kind::Flate;
	n->threshold = radius;
	n->stack_size = c->stack_size;
	n->leaves = c->leaves;
	n->lhs = std::move(c);
	return n;
}

NodeRef copy_tree(const NodeRef& n)
{
	switch (n->kind)
	{
	case Node::Kind::Brush: return std::make_shared<Node>(*n);
	case Node::Kind::Set: return make_set(n->opcode, copy_tree(n->lhs), copy_tree(n->rhs), n->threshold);
	default: return make_flate(copy_tree(n->lhs), n->threshold);
	}
}

// Transforms propagate to the brushes (sdf_evaluator.cpp:926-943; transform.cpp:29-46).
void move_tree(Node* n, Vec3 o)
{
	if (n->kind == Node::Kind::Brush)
	{
		Vec3& t = n->xform.translation;
		t = { t.x + o.x, t.y + o.y, t.z + o.z };
		return;
	}
	move_tree(n->lhs.get(), o);
	if (n->rhs) move_tree(n->rhs.get(), o);
}

void rotate_tree(Node* n, Quat q)
{
	if (n->kind == Node::Kind::Brush)
	{
		n->xform.translation = rotate(q, n->xform.translation);
		n->xform.rotation = qmul(q, n->xform.rotation);
		return;
	}
	rotate_tree(n->lhs.get(), q);
	if (n->rhs) rotate_tree(n->rhs.get(), q);
}

void scale_tree(Node* n, float s)
{
	if (n->kind == Node::Kind::Brush)
	{
		Vec3& t = n->xform.translation;
		t = { t.x * s, t.y * s, t.z * s };
		n->xform.scale = n->xform.scale * s;
		return;
	}
	if (n->kind == Node::Kind::Set)
	{
		n->threshold = n->threshold * s;  // set nodes scale their blend radius
		scale_tree(n->lhs.get(), s);
		scale_tree(n->rhs.get(), s);
		return;
	}
	scale_tree(n->lhs.get(), s);  // flate radius is not scaled (sdf_evaluator.cpp:1115-1118)
}

Quat axis_rotation(int axis, float degrees)
{
	// SDF::RotateX/Y/Z (sdf_evaluator.cpp:1180-1202): glm::radians in float, half angle, sinf/cosf.
	const float rad = degrees * static_cast<float>(0.01745329251994329576923690768489);
	const float half = float(double(rad) * .5);
	const float s = std::sin(half), c = std::cos(half);
	Quat q{ c, 0.0f, 0.0f, 0.0f };
	if (axis == 0) q.x = s;
	else if (axis == 1) q.y = s;
	else q.z = s;
	return q;
}

Box3 tree_bounds(const Node* n)
{
	if (n->kind == Node::Kind::Brush)
	{
		const Transform3& t = n->xform;
		const Box3& b = n->local_bounds;
		if (is_identity(t.rotation))
		{
			return { { b.lo.x * t.scale + t.translation.x, b.lo.y * t.scale + t.translation.y, b.lo.z * t.scale + t.translation.z },
			         { b.hi.x * t.scale + t.translation.x, b.hi.y * t.scale + t.translation.y, b.hi.z * t.scale + t.translation.z } };
		}
		// Transform the 8 corners (EvaluatorTransform::Apply, sdf_evaluator.cpp:365-397).
		const Vec3 A = b.lo, B = b.hi;
		const Vec3 others[7] = { B, { B.x, A.y, A.z }, { A.x, B.y, A.z }, { A.x, A.y, B.z },
		                         { A.x, B.y, B.z }, { B.x, A.y, B.z }, { B.x, B.y, A.z } };
		Box3 o;
		o.lo = o.hi = apply(t, A);
		for (const Vec3& p : others)
		{
			const Vec3 q = apply(t, p);
			o.lo = vmin3(o.lo, q);
			o.hi = vmax3(o.hi, q);
		}
		return o;
	}
	if (n->kind == Node::Kind::Flate)
	{
		Box3 b = tree_bounds(n->lhs.get());
		const float pad = n->threshold * 2.0f;
		b.hi = { b.hi.x + pad, b.hi.y + pad, b.hi.z + pad };
		b.lo = { b.lo.x - pad, b.lo.y - pad, b.lo.z - pad };
		return b;
	}
	const Box3 l = tree_bounds(n->lhs.get()), r = tree_bounds(n->rhs.get());
	Box3 c;
	if (is_union(n->opcode)) c = { vmin3(l.lo, r.lo), vmax3(l.hi, r.hi) };
	else if (is_diff(n->opcode)) c = l;
	else c = { vmax3(l.lo, r.lo), vmin3(l.hi, r.hi) };
	if (is_blend(n->opcode))
	{
		// Blend region padding (sdf_evaluator.cpp:878-886).
		const float t = n->threshold;
		Vec3 lo = vmax3(l.lo, r.lo), hi = vmin3(l.hi, r.hi);
		lo = { lo.x - t, lo.y - t, lo.z - t };
		hi = { hi.x + t, hi.y + t, hi.z + t };
		c.lo = vmin3(c.lo, lo);
		c.hi = vmax3(c.hi, hi);
	}
	return c;
}

bool tree_has_finite_bounds(const Node* n)
{
	if (n->kind == Node::Kind::Brush)
	{
		const Box3& b = n->local_bounds;
		return !(std::isinf(b.lo.x) || std::isinf(b.lo.y) || std::isinf(b.lo.z) ||
		         std::isinf(b.hi.x) || std::isinf(b.hi.y) || std::isinf(b.hi.z));
	}
	if (n->kind == Node::Kind::Set)
		return tree_has_finite_bounds(n->lhs.get()) || tree_has_finite_bounds(n->rhs.get());
	return tree_has_finite_bounds(n->lhs.get());
}

static void compile_rec(const Node* n, std::vector<uint32_t>& w)
{
	if (n->kind == Node::Kind::Brush)
	{
		const Transform3& t = n->xform;
		const bool rot = !is_identity(t.rotation);
		const bool scl = t.scale != 1.0f;
		const bool trn = !(t.translation.x == 0.0f && t.translation.y == 0.0f && t.translation.z == 0.0f);
		if (rot || scl)
		{
			const Mat4 inv = inverse4(transform_matrix(t));
			w.push_back(OP_MATRIX);
			for (int c = 0; c < 4; ++c)
				for (int r = 0; r < 4; ++r) push_float(w, inv.col[c][r]);
		}
		else if (trn)
		{
			w.push_back(OP_OFFSET);
			push_float(w, -t.translation.x);
			push_float(w, -t.translation.y);
			push_float(w, -t.translation.z);
		}
		w.push_back(n->opcode);
		for (int i = 0; i < brush_params(n->opcode); ++i) push_float(w, n->params[i]);
		if (scl)
		{
			w.push_back(OP_SCALE_FIELD);
			push_float(w, t.scale);
		}
		return;
	}
	if (n->kind == Node::Kind::Set)
	{
		compile_rec(n->lhs.get(), w);
		compile_rec(n->rhs.get(), w);
		w.push_back(n->opcode);
		if (is_blend(n->opcode)) push_float(w, n->threshold);
		return;
	}
	compile_rec(n->lhs.get(), w);
	w.push_back(OP_FLATE);
	push_float(w, n->threshold);
}

void compile_tree(const Node* n, std::vector<uint32_t>& words)
{
	words.clear();
	compile_rec(n, words);
	words.push_back(OP_STOP);
}

}  // namespace tg
