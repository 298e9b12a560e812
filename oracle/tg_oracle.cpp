// ORACLE -- test infrastructure, not product code.
//
// A plain C++ restatement of the reference's CPU CSG signed-distance path,
// used only by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg as the checker.  The product (tangerine_amd/) never links
// or calls it.
//
// What it restates, with the reference file:line each part follows:
//   distance math            tangerine/sdf_evaluator.cpp:165-295  (SDFMath::*)
//   scalar min/max           tangerine/glm_common.h:39-54         (fminf/fmaxf)
//   brush / set / flate Eval tangerine/sdf_evaluator.cpp:463-466, 774-780, 1059-1062
//   transform                tangerine/transform.cpp:29-67
//   quaternion algebra       glm 0.9.9.8 detail/type_quat.inl:16-23,282-292,343-350,
//                            ext/quaternion_common.inl:113-122, detail/func_geometric.inl:68-79
//   Compile (words)          tangerine/sdf_evaluator.cpp:41-69, 409-429, 495-507, 915-924, 1098-1103, 1376-1382
//   ToMatrix / inverse       tangerine/transform.cpp:49-55; glm 0.9.9.8 gtc/quaternion.inl:41-72,
//                            detail/func_matrix.inl:294-351 (cofactor inverse), detail/type_mat4x4.inl:555-572
//   interpreter              tangerine/sdf_evaluator.cpp:1384-1605
//   Clip                     tangerine/sdf_evaluator.cpp:468-478, 782-850, 1064-1075
//   Bounds                   tangerine/sdf_evaluator.cpp:330-397, 484-487, 857-890, 1082-1088, 1206-1310
//   Gradient (tetrahedral)   tangerine/sdf_evaluator.cpp:298-333
//   RayMarch                 tangerine/sdf_evaluator.cpp:336-353
//   serial surface nets      third_party/naive-surface-nets/src/surface_nets.cpp:30-580
//
// Functions with no reference counterpart and how they are pinned:
//   tgo_gradient_central     axis-aligned central differences; pinned by closed-form
//                            normals of spheres/boxes in tests/test_oracle.py.
//   tgo_raymarch_far         RayMarch with a far cutoff; pinned by ray-sphere closed form.
//
// Build flags: -O2 -ffp-contract=off (the reference Release build has no FMA
// contraction on x86-64 baseline), no fast-math.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tgo {

// ---------------------------------------------------------------------------
// Small vector algebra in the reference's evaluation order.

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct Q4 { float w, x, y, z; };  // glm::quat member order used by the ctor quat(w,x,y,z)

static inline float fmin_(float a, float b) { return std::fmin(a, b); }
static inline float fmax_(float a, float b) { return std::fmax(a, b); }
// glm scalar min/max used inside glm::clamp and vector max: (y < x) ? y : x etc.
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
static inline float gsign(float x) { return float((0.0f < x) - (x < 0.0f)); }

static inline float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
static inline float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float len2(V2 a) { return std::sqrt(dot2(a, a)); }
static inline float len3(V3 a) { return std::sqrt(dot3(a, a)); }
static inline V3 sub3(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
static inline V3 add3(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
static inline V3 mul3s(V3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
static inline V3 div3s(V3 a, float s) { return { a.x / s, a.y / s, a.z / s }; }

static inline V3 cross3(V3 a, V3 b)
{
	return { a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y };
}

// q * v (glm detail/type_quat.inl:343-350)
static inline V3 qrotate(Q4 q, V3 v)
{
	V3 u = { q.x, q.y, q.z };
	V3 uv = cross3(u, v);
	V3 uuv = cross3(u, uv);
	V3 t = add3(mul3s(uv, q.w), uuv);
	return add3(v, mul3s(t, 2.0f));
}

static inline Q4 qmul(Q4 p, Q4 q)
{
	return {
		p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z,
		p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
		p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z,
		p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x };
}

static inline Q4 qinverse(Q4 q)
{
	float d = (q.w * q.w + q.x * q.x) + (q.y * q.y + q.z * q.z);
	return { q.w / d, -q.x / d, -q.y / d, -q.z / d };
}

// Column-major 4x4, c[col][row].
struct M4 { float c[4][4]; };

static M4 mat_identity()
{
	M4 m;
	for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) m.c[i][j] = (i == j) ? 1.0f : 0.0f;
	return m;
}

static M4 mat_mul(const M4& a, const M4& b)
{
	// column j of result = a0*b[j][0] + a1*b[j][1] + a2*b[j][2] + a3*b[j][3],
	// summed left to right (glm detail/type_mat4x4.inl:630-648)
	M4 r;
	for (int j = 0; j < 4; ++j)
	{
		for (int row = 0; row < 4; ++row)
		{
			float m0 = a.c[0][row] * b.c[j][0];
			float m1 = a.c[1][row] * b.c[j][1];
			float m2 = a.c[2][row] * b.c[j][2];
			float m3 = a.c[3][row] * b.c[j][3];
			r.c[j][row] = ((m0 + m1) + m2) + m3;
		}
	}
	return r;
}

static M4 mat_from_quat(Q4 q)
{
	M4 r = mat_identity();
	float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
	float xz = q.x * q.z, xy = q.x * q.y, yz = q.y * q.z;
	float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
	r.c[0][0] = 1.0f - 2.0f * (yy + zz);
	r.c[0][1] = 2.0f * (xy + wz);
	r.c[0][2] = 2.0f * (xz - wy);
	r.c[1][0] = 2.0f * (xy - wz);
	r.c[1][1] = 1.0f - 2.0f * (xx + zz);
	r.c[1][2] = 2.0f * (yz + wx);
	r.c[2][0] = 2.0f * (xz + wy);
	r.c[2][1] = 2.0f * (yz - wx);
	r.c[2][2] = 1.0f - 2.0f * (xx + yy);
	return r;
}

// Cofactor inverse in the order of glm detail/func_matrix.inl:294-351.
static M4 mat_inverse(const M4& M)
{
	auto m = [&](int c, int r) { return M.c[c][r]; };
	float c00 = m(2,2) * m(3,3) - m(3,2) * m(2,3);
	float c02 = m(1,2) * m(3,3) - m(3,2) * m(1,3);
	float c03 = m(1,2) * m(2,3) - m(2,2) * m(1,3);
	float c04 = m(2,1) * m(3,3) - m(3,1) * m(2,3);
	float c06 = m(1,1) * m(3,3) - m(3,1) * m(1,3);
	float c07 = m(1,1) * m(2,3) - m(2,1) * m(1,3);
	float c08 = m(2,1) * m(3,2) - m(3,1) * m(2,2);
	float c10 = m(1,1) * m(3,2) - m(3,1) * m(1,2);
	float c11 = m(1,1) * m(2,2) - m(2,1) * m(1,2);
	float c12 = m(2,0) * m(3,3) - m(3,0) * m(2,3);
	float c14 = m(1,0) * m(3,3) - m(3,0) * m(1,3);
	float c15 = m(1,0) * m(2,3) - m(2,0) * m(1,3);
	float c16 = m(2,0) * m(3,2) - m(3,0) * m(2,2);
	float c18 = m(1,0) * m(3,2) - m(3,0) * m(1,2);
	float c19 = m(1,0) * m(2,2) - m(2,0) * m(1,2);
	float c20 = m(2,0) * m(3,1) - m(3,0) * m(2,1);
	float c22 = m(1,0) * m(3,1) - m(3,0) * m(1,1);
	float c23 = m(1,0) * m(2,1) - m(2,0) * m(1,1);
	float f0[4] = { c00, c00, c02, c03 };
	float f1[4] = { c04, c04, c06, c07 };
	float f2[4] = { c08, c08, c10, c11 };
	float f3[4] = { c12, c12, c14, c15 };
	float f4[4] = { c16, c16, c18, c19 };
	float f5[4] = { c20, c20, c22, c23 };
	float v0[4] = { m(1,0), m(0,0), m(0,0), m(0,0) };
	float v1[4] = { m(1,1), m(0,1), m(0,1), m(0,1) };
	float v2[4] = { m(1,2), m(0,2), m(0,2), m(0,2) };
	float v3[4] = { m(1,3), m(0,3), m(0,3), m(0,3) };
	float sa[4] = { 1, -1, 1, -1 };
	float sb[4] = { -1, 1, -1, 1 };
	M4 inv;
	for (int i = 0; i < 4; ++i)
	{
		inv.c[0][i] = ((v1[i] * f0[i] - v2[i] * f1[i]) + v3[i] * f2[i]) * sa[i];
		inv.c[1][i] = ((v0[i] * f0[i] - v2[i] * f3[i]) + v3[i] * f4[i]) * sb[i];
		inv.c[2][i] = ((v0[i] * f1[i] - v1[i] * f3[i]) + v3[i] * f5[i]) * sa[i];
		inv.c[3][i] = ((v0[i] * f2[i] - v1[i] * f4[i]) + v2[i] * f5[i]) * sb[i];
	}
	float row0[4] = { inv.c[0][0], inv.c[1][0], inv.c[2][0], inv.c[3][0] };
	float d0 = M.c[0][0] * row0[0], d1 = M.c[0][1] * row0[1], d2 = M.c[0][2] * row0[2], d3 = M.c[0][3] * row0[3];
	float det = (d0 + d1) + (d2 + d3);
	float one_over = 1.0f / det;
	for (int c = 0; c < 4; ++c) for (int r = 0; r < 4; ++r) inv.c[c][r] = inv.c[c][r] * one_over;
	return inv;
}

// ---------------------------------------------------------------------------
// Distance math (sdf_evaluator.cpp:165-295), including the double islands.

enum Op : uint32_t
{
	Stop = 0, Sphere, Ellipsoid, Box, Torus, Cylinder, Cone, Coninder, Plane,
	Union, Inter, Diff, BlendUnion, BlendInter, BlendDiff, Flate,
	Offset, Matrix, ScaleField
};

static float sd_sphere(V3 p, float r) { return len3(p) - r; }

static float sd_ellipsoid(V3 p, V3 r)
{
	float k0 = len3({ p.x / r.x, p.y / r.y, p.z / r.z });
	V3 rr = { r.x * r.x, r.y * r.y, r.z * r.z };
	float k1 = len3({ p.x / rr.x, p.y / rr.y, p.z / rr.z });
	return float(double(k0) * (double(k0) - 1.0) / double(k1));
}

static float sd_box(V3 p, V3 e)
{
	V3 a = { std::fabs(p.x) - e.x, std::fabs(p.y) - e.y, std::fabs(p.z) - e.z };
	V3 m = { gmax(a.x, 0.0f), gmax(a.y, 0.0f), gmax(a.z, 0.0f) };
	return len3(m) + fmin_(fmax_(fmax_(a.x, a.y), a.z), 0.0f);
}

static float sd_torus(V3 p, float R, float r)
{
	V2 q = { len2({ p.x, p.y }) - R, p.z };
	return len2(q) - r;
}

static float sd_cylinder(V3 p, float r, float h)
{
	V2 d = { std::fabs(len2({ p.x, p.y })) - r, std::fabs(p.z) - h };
	V2 m = { gmax(d.x, 0.0f), gmax(d.y, 0.0f) };
	return fmin_(fmax_(d.x, d.y), 0.0f) + len2(m);
}

static float sd_plane(V3 p, V3 n) { return dot3(p, n); }

static float sd_cone(V3 p, float tangent, float height)
{
	V2 q = { height * tangent, height * -1.0f };
	V2 w = { len2({ p.x, p.y }), float(double(height) * -0.5 + double(p.z)) };
	float t0 = gclamp(dot2(w, q) / dot2(q, q), 0.0f, 1.0f);
	V2 a = { w.x - q.x * t0, w.y - q.y * t0 };
	float t1 = gclamp(w.x / q.x, 0.0f, 1.0f);
	V2 b = { w.x - q.x * t1, w.y - q.y * 1.0f };
	float k = gsign(q.y);
	float d = fmin_(dot2(a, a), dot2(b, b));
	float s = fmax_(k * (w.x * q.y - w.y * q.x), k * (w.y - q.y));
	return std::sqrt(d) * gsign(s);
}

static float sd_coninder(V3 p, float rl, float rh, float h)
{
	V2 q = { len2({ p.x, p.y }), p.z };
	V2 k1 = { rh, h };
	V2 k2 = { rh - rl, float(2.0 * double(h)) };
	V2 ca = { q.x - fmin_(q.x, (q.y < 0.0f) ? rl : rh), std::fabs(q.y) - h };
	V2 kq = { k1.x - q.x, k1.y - q.y };
	float t = gclamp(dot2(kq, k2) / dot2(k2, k2), 0.0f, 1.0f);
	V2 cb = { (q.x - k1.x) + k2.x * t, (q.y - k1.y) + k2.y * t };
	float s = (cb.x < 0.0f && ca.y < 0.0f) ? -1.0f : 1.0f;
	return s * std::sqrt(fmin_(dot2(ca, ca), dot2(cb, cb)));
}

static float op_blend_union(float a, float b, float t)
{
	float h = fmax_(t - std::fabs(a - b), 0.0f);
	return float(double(fmin_(a, b)) - double(h * h) * 0.25 / double(t));
}

static float op_blend_inter(float a, float b, float t)
{
	float h = fmax_(t - std::fabs(a - b), 0.0f);
	return float(double(fmax_(a, b)) + double(h * h) * 0.25 / double(t));
}

static float op_blend_diff(float a, float b, float t)
{
	float h = fmax_(t - std::fabs(a + b), 0.0f);
	return float(double(fmax_(a, -b)) + double(h * h) * 0.25 / double(t));
}

static float eval_brush(uint32_t op, const float* prm, V3 p)
{
	switch (op)
	{
	case Sphere: return sd_sphere(p, prm[0]);
	case Ellipsoid: return sd_ellipsoid(p, { prm[0], prm[1], prm[2] });
	case Box: return sd_box(p, { prm[0], prm[1], prm[2] });
	case Torus: return sd_torus(p, prm[0], prm[1]);
	case Cylinder: return sd_cylinder(p, prm[0], prm[1]);
	case Cone: return sd_cone(p, prm[0], prm[1]);
	case Coninder: return sd_coninder(p, prm[0], prm[1], prm[2]);
	case Plane: return sd_plane(p, { prm[0], prm[1], prm[2] });
	default: return 0.0f;
	}
}

static float eval_set(uint32_t op, float a, float b, float t)
{
	switch (op)
	{
	case Union: return fmin_(a, b);
	case Inter: return fmax_(a, b);
	case Diff: return fmax_(a, -b);
	case BlendUnion: return op_blend_union(a, b, t);
	case BlendInter: return op_blend_inter(a, b, t);
	case BlendDiff: return op_blend_diff(a, b, t);
	default: return 0.0f;
	}
}

static int brush_param_count(uint32_t op)
{
	switch (op)
	{
	case Sphere: return 1;
	case Torus: case Cylinder: case Cone: return 2;
	default: return 3;
	}
}

// ---------------------------------------------------------------------------
// Tree.

struct AABBo { V3 mn, mx; };

struct Xform
{
	Q4 rot = { 1, 0, 0, 0 };
	V3 trans = { 0, 0, 0 };
	float scale = 1.0f;
};

struct Node;
using NodeP = std::shared_ptr<Node>;

struct Node
{
	enum Kind { BrushK, SetK, FlateK } kind;
	uint32_t op = 0;
	float prm[3] = { 0, 0, 0 };
	AABBo local_bounds{};
	Xform xf;
	NodeP lhs, rhs;
	float threshold = 0.0f;
	size_t stack = 1;
};

static NodeP make_brush(uint32_t op, float a, float b, float c, AABBo bounds)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::BrushK;
	n->op = op;
	n->prm[0] = a; n->prm[1] = b; n->prm[2] = c;
	n->local_bounds = bounds;
	n->stack = 1;
	return n;
}

static bool is_diff(uint32_t op) { return op == Diff || op == BlendDiff; }
static bool is_blend(uint32_t op) { return op >= BlendUnion && op <= BlendDiff; }
static int family(uint32_t op) { return (op == Union || op == BlendUnion) ? 0 : (op == Inter || op == BlendInter) ? 1 : 2; }

// SetNode ctor with left-lean swap (sdf_evaluator.cpp:737-772).
static NodeP make_set(uint32_t op, NodeP l, NodeP r, float t)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::SetK;
	n->op = op;
	n->threshold = t;
	if (!is_diff(op) && r->stack > l->stack) std::swap(l, r);
	n->lhs = l;
	n->rhs = r;
	n->stack = std::max(l->stack, r->stack + 1);
	return n;
}

static NodeP make_flate(NodeP c, float r)
{
	auto n = std::make_shared<Node>();
	n->kind = Node::FlateK;
	n->lhs = c;
	n->threshold = r;
	n->stack = c->stack;
	return n;
}

static NodeP copy_node(const NodeP& n)
{
	if (!n) return nullptr;
	switch (n->kind)
	{
	case Node::BrushK: return std::make_shared<Node>(*n);
	case Node::SetK: return make_set(n->op, copy_node(n->lhs), copy_node(n->rhs), n->threshold);
	default: return make_flate(copy_node(n->lhs), n->threshold);
	}
}

static void move_node(Node* n, V3 o)
{
	if (n->kind == Node::BrushK) n->xf.trans = add3(n->xf.trans, o);
	else { move_node(n->lhs.get(), o); if (n->rhs) move_node(n->rhs.get(), o); }
}

static void rotate_node(Node* n, Q4 q)
{
	if (n->kind == Node::BrushK)
	{
		n->xf.trans = qrotate(q, n->xf.trans);
		n->xf.rot = qmul(q, n->xf.rot);
	}
	else { rotate_node(n->lhs.get(), q); if (n->rhs) rotate_node(n->rhs.get(), q); }
}

static void scale_node(Node* n, float s)
{
	if (n->kind == Node::BrushK)
	{
		n->xf.trans = mul3s(n->xf.trans, s);
		n->xf.scale = n->xf.scale * s;
	}
	else if (n->kind == Node::SetK)
	{
		n->threshold = n->threshold * s;
		scale_node(n->lhs.get(), s);
		scale_node(n->rhs.get(), s);
	}
	else scale_node(n->lhs.get(), s);
}

static float eval_node(const Node* n, V3 p)
{
	switch (n->kind)
	{
	case Node::BrushK:
	{
		// ApplyInv: rotate(inverse(q), p - t) / s, then distance * s
		V3 local = div3s(qrotate(qinverse(n->xf.rot), sub3(p, n->xf.trans)), n->xf.scale);
		return eval_brush(n->op, n->prm, local) * n->xf.scale;
	}
	case Node::SetK:
		return eval_set(n->op, eval_node(n->lhs.get(), p), eval_node(n->rhs.get(), p), n->threshold);
	default:
		return eval_node(n->lhs.get(), p) - n->threshold;
	}
}

static int leaf_count(const Node* n)
{
	if (n->kind == Node::BrushK) return 1;
	if (n->kind == Node::SetK) return leaf_count(n->lhs.get()) + leaf_count(n->rhs.get());
	return leaf_count(n->lhs.get());
}

// Clip (sdf_evaluator.cpp:468-478, 782-850, 1064-1075).
static NodeP clip_node(const NodeP& n, V3 p, float r)
{
	if (eval_node(n.get(), p) > r) return nullptr;
	if (n->kind == Node::BrushK) return copy_node(n);
	if (n->kind == Node::FlateK)
	{
		// The reference wraps a null child here (sdf_evaluator.cpp:1066-1070); an
		// empty result is returned instead.
		NodeP c = clip_node(n->lhs, p, r + n->threshold);
		return c ? make_flate(c, n->threshold) : nullptr;
	}
	int fam = family(n->op);
	if (is_blend(n->op))
	{
		NodeP l = clip_node(n->lhs, p, r + n->threshold);
		NodeP rr = clip_node(n->rhs, p, r + n->threshold);
		if (l && rr) return make_set(n->op, l, rr, n->threshold);
		if (fam == 1) return nullptr;
	}
	NodeP l = clip_node(n->lhs, p, r);
	NodeP rr = clip_node(n->rhs, p, r);
	if (l && rr) return make_set(n->op, l, rr, n->threshold);
	if (fam == 0) return l ? l : rr;
	if (fam == 2) return l;
	return nullptr;
}

// Compile (sdf_evaluator.cpp:409-429, 495-507, 915-924, 1098-1103).
static M4 xform_matrix(const Xform& x)
{
	M4 rot = mat_from_quat(x.rot);
	M4 tr = mat_identity();
	tr.c[3][0] = x.trans.x; tr.c[3][1] = x.trans.y; tr.c[3][2] = x.trans.z;
	// translate(identity, t): column 3 = m0*t.x + m1*t.y + m2*t.z + m3 (glm ext/matrix_transform.inl)
	{
		M4 id = mat_identity();
		for (int r = 0; r < 4; ++r)
		{
			tr.c[3][r] = ((id.c[0][r] * x.trans.x + id.c[1][r] * x.trans.y) + id.c[2][r] * x.trans.z) + id.c[3][r];
		}
	}
	// scale_slow(m, v) = m * diag(v.x, v.y, v.z, 1)
	M4 sc = mat_identity();
	sc.c[0][0] = x.scale; sc.c[1][1] = x.scale; sc.c[2][2] = x.scale;
	M4 ts = mat_mul(tr, sc);
	return mat_mul(ts, rot);
}

static void push_f(std::vector<uint32_t>& w, float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	w.push_back(u);
}

static void compile_node(const Node* n, std::vector<uint32_t>& w)
{
	if (n->kind == Node::BrushK)
	{
		const Xform& x = n->xf;
		bool rot = !(x.rot.w == 1.0f && x.rot.x == 0.0f && x.rot.y == 0.0f && x.rot.z == 0.0f);
		bool scl = x.scale != 1.0f;
		bool trn = !(x.trans.x == 0.0f && x.trans.y == 0.0f && x.trans.z == 0.0f);
		if (rot || scl)
		{
			M4 inv = mat_inverse(xform_matrix(x));
			w.push_back(Matrix);
			for (int c = 0; c < 4; ++c) for (int r = 0; r < 4; ++r) push_f(w, inv.c[c][r]);
		}
		else if (trn)
		{
			w.push_back(Offset);
			push_f(w, -x.trans.x); push_f(w, -x.trans.y); push_f(w, -x.trans.z);
		}
		w.push_back(n->op);
		for (int i = 0; i < brush_param_count(n->op); ++i) push_f(w, n->prm[i]);
		if (scl)
		{
			w.push_back(ScaleField);
			push_f(w, x.scale);
		}
	}
	else if (n->kind == Node::SetK)
	{
		compile_node(n->lhs.get(), w);
		compile_node(n->rhs.get(), w);
		w.push_back(n->op);
		if (is_blend(n->op)) push_f(w, n->threshold);
	}
	else
	{
		compile_node(n->lhs.get(), w);
		w.push_back(Flate);
		push_f(w, n->threshold);
	}
}

// Bounds (sdf_evaluator.cpp:330-397 EvaluatorTransform::Apply, 857-890, 1082-1088).
static V3 xform_apply(const Xform& x, V3 p) { return add3(qrotate(x.rot, mul3s(p, x.scale)), x.trans); }

static AABBo bounds_node(const Node* n)
{
	if (n->kind == Node::BrushK)
	{
		const Xform& x = n->xf;
		AABBo b = n->local_bounds;
		bool ident = x.rot.w == 1.0f && x.rot.x == 0.0f && x.rot.y == 0.0f && x.rot.z == 0.0f;
		if (ident)
		{
			return { add3(mul3s(b.mn, x.scale), x.trans), add3(mul3s(b.mx, x.scale), x.trans) };
		}
		V3 A = b.mn, B = b.mx;
		V3 pts[7] = { B, { B.x, A.y, A.z }, { A.x, B.y, A.z }, { A.x, A.y, B.z },
			{ A.x, B.y, B.z }, { B.x, A.y, B.z }, { B.x, B.y, A.z } };
		AABBo o;
		o.mn = xform_apply(x, A);
		o.mx = o.mn;
		for (V3 q : pts)
		{
			V3 t = xform_apply(x, q);
			o.mn = { gmin(o.mn.x, t.x), gmin(o.mn.y, t.y), gmin(o.mn.z, t.z) };
			o.mx = { gmax(o.mx.x, t.x), gmax(o.mx.y, t.y), gmax(o.mx.z, t.z) };
		}
		return o;
	}
	if (n->kind == Node::FlateK)
	{
		AABBo b = bounds_node(n->lhs.get());
		float r2 = n->threshold * 2.0f;
		b.mx = add3(b.mx, { r2, r2, r2 });
		b.mn = sub3(b.mn, { r2, r2, r2 });
		return b;
	}
	AABBo l = bounds_node(n->lhs.get()), r = bounds_node(n->rhs.get());
	AABBo c;
	int fam = family(n->op);
	if (fam == 0)
	{
		c.mn = { gmin(l.mn.x, r.mn.x), gmin(l.mn.y, r.mn.y), gmin(l.mn.z, r.mn.z) };
		c.mx = { gmax(l.mx.x, r.mx.x), gmax(l.mx.y, r.mx.y), gmax(l.mx.z, r.mx.z) };
	}
	else if (fam == 2) c = l;
	else
	{
		c.mn = { gmax(l.mn.x, r.mn.x), gmax(l.mn.y, r.mn.y), gmax(l.mn.z, r.mn.z) };
		c.mx = { gmin(l.mx.x, r.mx.x), gmin(l.mx.y, r.mx.y), gmin(l.mx.z, r.mx.z) };
	}
	if (is_blend(n->op))
	{
		float t = n->threshold;
		V3 lmn = { gmax(l.mn.x, r.mn.x) - t, gmax(l.mn.y, r.mn.y) - t, gmax(l.mn.z, r.mn.z) - t };
		V3 lmx = { gmin(l.mx.x, r.mx.x) + t, gmin(l.mx.y, r.mx.y) + t, gmin(l.mx.z, r.mx.z) + t };
		c.mn = { gmin(c.mn.x, lmn.x), gmin(c.mn.y, lmn.y), gmin(c.mn.z, lmn.z) };
		c.mx = { gmax(c.mx.x, lmx.x), gmax(c.mx.y, lmx.y), gmax(c.mx.z, lmx.z) };
	}
	return c;
}

// ---------------------------------------------------------------------------
// Interpreter over ProgramBuffer words (sdf_evaluator.cpp:1384-1605).

static inline float wf(const uint32_t* w, size_t i)
{
	float f;
	std::memcpy(&f, &w[i], 4);
	return f;
}

static float interpret(const uint32_t* w, size_t n, V3 ep, bool* ok)
{
	// the reference sizes its stack from the tree (StackSize, no cap): 256
	// slots here, grown on the heap for deeper programs
	float small[256];
	std::vector<float> big;
	float* stack = small;
	size_t cap = 256;
	size_t sp = 0;
	size_t pc = 0;
	V3 p = ep;
	*ok = true;
	while (pc < n)
	{
		uint32_t op = w[pc++];
		switch (op)
		{
		case Stop:
			if (sp != 1) *ok = false;
			return sp > 0 ? stack[sp - 1] : 0.0f;
		case Sphere: case Ellipsoid: case Box: case Torus: case Cylinder: case Cone: case Coninder: case Plane:
		{
			int np = brush_param_count(op);
			float prm[3] = { 0, 0, 0 };
			for (int i = 0; i < np; ++i) prm[i] = wf(w, pc + i);
			pc += np;
			if (sp >= cap)
			{
				big.resize(2 * cap);
				if (stack == small) std::memcpy(big.data(), small, sizeof(small));
				stack = big.data();
				cap = big.size();
			}
			stack[sp++] = eval_brush(op, prm, p);
			p = ep;
			break;
		}
		case Union: case Inter: case Diff: case BlendUnion: case BlendInter: case BlendDiff:
		{
			if (sp < 2) { *ok = false; return 0.0f; }
			float r = stack[--sp];
			float l = stack[--sp];
			float t = 0.0f;
			if (is_blend(op)) t = wf(w, pc++);
			stack[sp++] = eval_set(op, l, r, t);
			break;
		}
		case Flate:
			if (sp < 1) { *ok = false; return 0.0f; }
			stack[sp - 1] -= wf(w, pc++);
			break;
		case Offset:
			p = { ep.x + wf(w, pc), ep.y + wf(w, pc + 1), ep.z + wf(w, pc + 2) };
			pc += 3;
			break;
		case Matrix:
		{
			// (M * vec4(p,1)).xyz, column-major, glm order ((c0x + c1y) + (c2z + c3))
			float m[16];
			for (int i = 0; i < 16; ++i) m[i] = wf(w, pc + i);
			pc += 16;
			V3 q;
			q.x = (m[0] * ep.x + m[4] * ep.y) + (m[8] * ep.z + m[12] * 1.0f);
			q.y = (m[1] * ep.x + m[5] * ep.y) + (m[9] * ep.z + m[13] * 1.0f);
			q.z = (m[2] * ep.x + m[6] * ep.y) + (m[10] * ep.z + m[14] * 1.0f);
			p = q;
			break;
		}
		case ScaleField:
			if (sp < 1) { *ok = false; return 0.0f; }
			stack[sp - 1] *= wf(w, pc++);
			break;
		default:
			*ok = false;
			return 0.0f;
		}
	}
	*ok = false;
	return 0.0f;
}

// ---------------------------------------------------------------------------
// Gradient and ray march (sdf_evaluator.cpp:298-353).

static V3 gradient_tetra(const Node* n, V3 p)
{
	const float h = 0.0001f;
	V2 o = { 1.0f * h, -1.0f * h };
	V3 kxyy = { o.x, o.y, o.y }, kyyx = { o.y, o.y, o.x }, kyxy = { o.y, o.x, o.y }, kxxx = { o.x, o.x, o.x };
	V3 g = add3(add3(add3(mul3s(kxyy, eval_node(n, add3(p, kxyy))), mul3s(kyyx, eval_node(n, add3(p, kyyx)))),
		mul3s(kyxy, eval_node(n, add3(p, kyxy)))), mul3s(kxxx, eval_node(n, add3(p, kxxx))));
	float l2 = dot3(g, g);
	if (l2 == 0.0f)
	{
		float d = eval_node(n, p);
		V3 v = { eval_node(n, add3(p, kxyy)) - d, eval_node(n, add3(p, kyxy)) - d, eval_node(n, add3(p, kyyx)) - d };
		float l = len3(v);
		return div3s(v, l);
	}
	return div3s(g, std::sqrt(l2));
}

}  // namespace tgo

// ---------------------------------------------------------------------------
// C ABI (ctypes-facing).  Handles are heap-allocated shared pointers.

using namespace tgo;

namespace {
struct H { NodeP n; };
H* wrap(NodeP n) { return n ? new H{ n } : nullptr; }
AABBo sym(V3 hi) { return { { -hi.x, -hi.y, -hi.z }, hi }; }
}

// ApplyVertexSequence with VertexSequence::Shuffle (tangerine/sodapop.cpp:
// 337-394): exchange[i] = the index vertex i moves to.  The generator is
// std::mt19937 seeded 0 and the shuffle is the C++ library's std::shuffle
// (libstdc++ here, as in the reference's Linux builds).
static void vertex_sequence(uint64_t n, uint32_t* exchange)
{
	for (uint64_t i = 0; i < n; ++i) exchange[i] = uint32_t(i);
	const size_t half = size_t(n) / 2;
	const size_t mirror = (n % 2) == 0 ? half : half + 1;
	std::vector<size_t> sequence;
	sequence.reserve(mirror);
	for (size_t k = 0; k < half; ++k) sequence.push_back(mirror + k);
	std::mt19937 generator;
	generator.seed(0);
	std::shuffle(sequence.begin(), sequence.end(), generator);
	if (mirror != half) sequence.push_back(half);
	for (size_t target = 0; target < sequence.size(); ++target)
	{
		exchange[target] = uint32_t(sequence[target]);
		exchange[sequence[target]] = uint32_t(target);
	}
}

extern "C" {

void* tgo_sphere(float r) { return wrap(make_brush(Sphere, r, 0, 0, sym({ r, r, r }))); }
void* tgo_ellipsoid(float x, float y, float z) { return wrap(make_brush(Ellipsoid, x, y, z, sym({ x, y, z }))); }
void* tgo_box(float x, float y, float z) { return wrap(make_brush(Box, x, y, z, sym({ x, y, z }))); }
void* tgo_torus(float R, float r)
{
	float o = R + r;
	return wrap(make_brush(Torus, R, r, 0, sym({ o, o, r })));
}
void* tgo_cylinder(float r, float h) { return wrap(make_brush(Cylinder, r, h, 0, sym({ r, r, h }))); }
void* tgo_plane(float x, float y, float z)
{
	float l = len3({ x, y, z });
	V3 n = div3s({ x, y, z }, l);
	const float inf = std::numeric_limits<float>::infinity();
	AABBo b = sym({ inf, inf, inf });
	if (n.x == -1.0f) b.mn.x = 0.0f;
	else if (n.x == 1.0f) b.mx.x = 0.0f;
	else if (n.y == -1.0f) b.mn.y = 0.0f;
	else if (n.y == 1.0f) b.mx.y = 0.0f;
	else if (n.z == -1.0f) b.mn.z = 0.0f;
	else if (n.z == 1.0f) b.mx.z = 0.0f;
	return wrap(make_brush(Plane, n.x, n.y, n.z, b));
}
void* tgo_cone(float r, float h)
{
	float t = r / h;
	return wrap(make_brush(Cone, t, h, 0, sym({ r, r, float(double(h) * .5) })));
}
void* tgo_coninder(float rl, float rh, float height)
{
	float half = float(double(height) * .5);
	float mr = fmax_(rl, rh);
	return wrap(make_brush(Coninder, rl, rh, half, sym({ mr, mr, half })));
}

void* tgo_op(int kind, float t, void* a, void* b)
{
	static const uint32_t ops[6] = { Union, Inter, Diff, BlendUnion, BlendInter, BlendDiff };
	if (kind < 0 || kind > 5) return nullptr;
	return wrap(make_set(ops[kind], ((H*)a)->n, ((H*)b)->n, t));
}

void* tgo_flate(void* a, float r) { return wrap(make_flate(((H*)a)->n, r)); }

void* tgo_move(void* a, float x, float y, float z)
{
	NodeP n = copy_node(((H*)a)->n);
	move_node(n.get(), { x, y, z });
	return wrap(n);
}

void* tgo_rotate(void* a, float x, float y, float z, float w)
{
	NodeP n = copy_node(((H*)a)->n);
	rotate_node(n.get(), { w, x, y, z });
	return wrap(n);
}

// SDF::RotateX/Y/Z (sdf_evaluator.cpp:1180-1202): half-angle from float radians.
void* tgo_rotate_axis(void* a, int axis, float degrees)
{
	float rad = degrees * float(0.01745329251994329576923690768489);
	float r = float(double(rad) * .5);
	float s = std::sin(r), c = std::cos(r);
	Q4 q = { c, 0, 0, 0 };
	if (axis == 0) q.x = s; else if (axis == 1) q.y = s; else q.z = s;
	NodeP n = copy_node(((H*)a)->n);
	rotate_node(n.get(), q);
	return wrap(n);
}

void* tgo_scale(void* a, float s)
{
	NodeP n = copy_node(((H*)a)->n);
	scale_node(n.get(), s);
	return wrap(n);
}

void tgo_free(void* a) { delete (H*)a; }
int tgo_leaf_count(void* a) { return leaf_count(((H*)a)->n.get()); }
uint64_t tgo_stack_size(void* a) { return ((H*)a)->n->stack; }

void tgo_bounds(void* a, float* out)
{
	AABBo b = bounds_node(((H*)a)->n.get());
	out[0] = b.mn.x; out[1] = b.mn.y; out[2] = b.mn.z;
	out[3] = b.mx.x; out[4] = b.mx.y; out[5] = b.mx.z;
}

void tgo_eval(void* a, const float* p, uint64_t n, float* out)
{
	const Node* nd = ((H*)a)->n.get();
	for (uint64_t i = 0; i < n; ++i) out[i] = eval_node(nd, { p[3 * i], p[3 * i + 1], p[3 * i + 2] });
}

uint64_t tgo_compile(void* a, uint32_t* words, uint64_t cap)
{
	std::vector<uint32_t> w;
	compile_node(((H*)a)->n.get(), w);
	w.push_back(Stop);
	for (size_t i = 0; i < w.size() && i < cap; ++i) words[i] = w[i];
	return w.size();
}

// Returns 0 on success, -1 if the program is malformed (stack discipline / opcode).
int tgo_interp_words(const uint32_t* words, uint64_t nwords, const float* p, uint64_t n, float* out)
{
	for (uint64_t i = 0; i < n; ++i)
	{
		bool ok;
		out[i] = interpret(words, nwords, { p[3 * i], p[3 * i + 1], p[3 * i + 2] }, &ok);
		if (!ok) return -1;
	}
	return 0;
}

// Multi-threaded dense-grid fill with the interpreter (CPU baseline "port" leg).
// Samples (count+1) per axis, x fastest, z-slabs [k0, k1).
int tgo_grid_interp(const uint32_t* words, uint64_t nwords, const float* origin, const float* step,
	const uint64_t* count, uint64_t k0, uint64_t k1, int threads, float* out)
{
	const uint64_t nx = count[0] + 1, ny = count[1] + 1;
	const uint64_t slice = nx * ny, total = slice * (k1 - k0);
	if (threads <= 0) threads = std::max(int(std::thread::hardware_concurrency()), 2);
	std::atomic<uint64_t> next(0);
	std::atomic<int> bad(0);
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; ++t)
	{
		pool.emplace_back([&]()
		{
			for (;;)
			{
				uint64_t base = next.fetch_add(4096);
				if (base >= total) break;
				uint64_t end = std::min(total, base + 4096);
				for (uint64_t s = base; s < end; ++s)
				{
					uint64_t k = s / slice + k0, j = (s % slice) / nx, i = s % nx;
					V3 p = { origin[0] + float(i) * step[0], origin[1] + float(j) * step[1], origin[2] + float(k) * step[2] };
					bool ok;
					out[s] = interpret(words, nwords, p, &ok);
					if (!ok) bad.store(1);
				}
			}
		});
	}
	for (auto& t : pool) t.join();
	return bad.load() ? -1 : threads;
}

void* tgo_clip(void* a, float x, float y, float z, float r)
{
	return wrap(clip_node(((H*)a)->n, { x, y, z }, r));
}

void tgo_gradient(void* a, const float* p, uint64_t n, float* out)
{
	const Node* nd = ((H*)a)->n.get();
	for (uint64_t i = 0; i < n; ++i)
	{
		V3 g = gradient_tetra(nd, { p[3 * i], p[3 * i + 1], p[3 * i + 2] });
		out[3 * i] = g.x; out[3 * i + 1] = g.y; out[3 * i + 2] = g.z;
	}
}

// Axis-aligned central differences on the interpreter:
// g = (f(p+h ex) - f(p-h ex), ..., ...) / |.|  (no reference counterpart).
int tgo_gradient_central_words(const uint32_t* words, uint64_t nwords, const float* p, uint64_t n, float h, float* out)
{
	for (uint64_t i = 0; i < n; ++i)
	{
		V3 c = { p[3 * i], p[3 * i + 1], p[3 * i + 2] };
		bool ok = true, k;
		float g[3];
		for (int a = 0; a < 3; ++a)
		{
			V3 lo = c, hi = c;
			if (a == 0) { hi.x = c.x + h; lo.x = c.x - h; }
			if (a == 1) { hi.y = c.y + h; lo.y = c.y - h; }
			if (a == 2) { hi.z = c.z + h; lo.z = c.z - h; }
			g[a] = interpret(words, nwords, hi, &k) - interpret(words, nwords, lo, &ok);
			ok = ok && k;
		}
		if (!ok) return -1;
		float l = std::sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
		out[3 * i] = g[0] / l; out[3 * i + 1] = g[1] / l; out[3 * i + 2] = g[2] / l;
	}
	return 0;
}

// SDFNode::Gradient's tetrahedral rule (sdf_evaluator.cpp:298-333) over the
// interpreter instead of the tree walk, at one point.
static bool gradient_tetra_words(const uint32_t* words, uint64_t nwords, V3 c, float h, V3* out)
{
	V2 o = { 1.0f * h, -1.0f * h };
	V3 kxyy = { o.x, o.y, o.y }, kyyx = { o.y, o.y, o.x }, kyxy = { o.y, o.x, o.y }, kxxx = { o.x, o.x, o.x };
	bool ok = true, k1, k2, k3, k4;
	float e1 = interpret(words, nwords, add3(c, kxyy), &k1);
	float e2 = interpret(words, nwords, add3(c, kyyx), &k2);
	float e3 = interpret(words, nwords, add3(c, kyxy), &k3);
	float e4 = interpret(words, nwords, add3(c, kxxx), &k4);
	ok = k1 && k2 && k3 && k4;
	V3 g = add3(add3(add3(mul3s(kxyy, e1), mul3s(kyyx, e2)), mul3s(kyxy, e3)), mul3s(kxxx, e4));
	float l2 = dot3(g, g);
	if (l2 == 0.0f)
	{
		float d0 = interpret(words, nwords, c, &k1);
		V3 v = { interpret(words, nwords, add3(c, kxyy), &k2) - d0, interpret(words, nwords, add3(c, kyxy), &k3) - d0,
			interpret(words, nwords, add3(c, kyyx), &k4) - d0 };
		ok = ok && k1 && k2 && k3 && k4;
		*out = mul3s(v, 1.0f / std::sqrt(dot3(v, v)));
	}
	else
	{
		*out = div3s(g, std::sqrt(l2));
	}
	return ok;
}

int tgo_gradient_tetra_words(const uint32_t* words, uint64_t nwords, const float* p, uint64_t n, float h, float* out)
{
	for (uint64_t i = 0; i < n; ++i)
	{
		V3 r;
		if (!gradient_tetra_words(words, nwords, { p[3 * i], p[3 * i + 1], p[3 * i + 2] }, h, &r)) return -1;
		out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
	}
	return 0;
}

// PointCloudExportThread (export.cpp:384-480) over the interpreter of the
// whole program, serially in raster order: cell centres
// (float(i) * Step + Start) + Half with |d| < length(Half), then
// `iterations` steps Cursor += Gradient * -d, clamp to the cell, keep if
// within length(Half) of the centre.  Writes up to `cap` points and returns
// how many there are, or -1 for a malformed program.
int64_t tgo_point_cloud(const uint32_t* words, uint64_t nwords, const float* lo, const float* hi, const float* step,
	int iterations, float* out, uint64_t cap)
{
	const V3 start = { lo[0], lo[1], lo[2] }, st = { step[0], step[1], step[2] };
	const V3 half = { st.x / 2.0f, st.y / 2.0f, st.z / 2.0f };
	const float diagonal = len3(half);
	int64_t it[3];
	for (int a = 0; a < 3; ++a)
	{
		const float c = std::ceil((hi[a] - lo[a]) / step[a]);
		it[a] = c > 0.0f ? int64_t(c) : 0;
	}
	const int64_t slice = it[0] * it[1], total = slice * it[2];
	std::vector<V3> verts;
	for (int64_t i = 0; i < total; ++i)
	{
		const float z = float(i / slice) * st.z + start.z;
		const float y = float((i % slice) / it[0]) * st.y + start.y;
		const float x = float(i % it[0]) * st.x + start.x;
		const V3 cursor = add3({ x, y, z }, half);
		bool ok;
		const float d = interpret(words, nwords, cursor, &ok);
		if (!ok) return -1;
		if (std::fabs(d) < diagonal) verts.push_back(cursor);
	}
	for (V3& v : verts)
	{
		const V3 low = { v.x - half.x, v.y - half.y, v.z - half.z }, high = add3(v, half);
		V3 c = v;
		for (int r = 0; r < iterations; ++r)
		{
			V3 dir;
			bool ok = gradient_tetra_words(words, nwords, c, 1.0e-4f, &dir);
			const float dist = float(double(interpret(words, nwords, c, &ok)) * -1.0);
			c = add3(c, mul3s(dir, dist));
		}
		c = { gclamp(c.x, low.x, high.x), gclamp(c.y, low.y, high.y), gclamp(c.z, low.z, high.z) };
		const V3 e = { v.x - c.x, v.y - c.y, v.z - c.z };
		if (len3(e) <= diagonal) v = c;
	}
	for (size_t i = 0; i < verts.size() && i < cap; ++i)
	{
		out[3 * i] = verts[i].x; out[3 * i + 1] = verts[i].y; out[3 * i + 2] = verts[i].z;
	}
	return int64_t(verts.size());
}

// SDFNode::RayMarch (sdf_evaluator.cpp:336-353) on the tree.
void tgo_raymarch(void* a, const float* start, const float* dir, uint64_t n, int iters, float eps,
	uint8_t* hit, float* travel, float* pos)
{
	const Node* nd = ((H*)a)->n.get();
	for (uint64_t i = 0; i < n; ++i)
	{
		V3 s = { start[3 * i], start[3 * i + 1], start[3 * i + 2] };
		V3 d = { dir[3 * i], dir[3 * i + 1], dir[3 * i + 2] };
		// glm::normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt
		d = mul3s(d, 1.0f / std::sqrt(dot3(d, d)));
		V3 ps = s;
		float t = 0.0f;
		bool h = false;
		for (int it = 0; it < iters; ++it)
		{
			float dist = eval_node(nd, ps);
			if (dist <= eps) { h = true; break; }
			t += dist;
			ps = add3(mul3s(d, t), s);
		}
		if (!h) t = std::numeric_limits<float>::infinity();
		hit[i] = h;
		travel[i] = t;
		pos[3 * i] = ps.x; pos[3 * i + 1] = ps.y; pos[3 * i + 2] = ps.z;
	}
}

// Sphere tracing over the interpreter with a far cutoff (no reference
// counterpart for the cutoff): the direction is used as given (callers pass
// unit vectors); hit when d <= eps, miss when t > far or steps run out.
// Returns per ray in `steps` the distance evaluations made, the hit test included.
int tgo_raymarch_far_words(const uint32_t* words, uint64_t nwords, const float* start, const float* dir, uint64_t n,
	int iters, float eps, float far_, uint8_t* hit, float* travel, int32_t* steps)
{
	for (uint64_t i = 0; i < n; ++i)
	{
		V3 s = { start[3 * i], start[3 * i + 1], start[3 * i + 2] };
		V3 d = { dir[3 * i], dir[3 * i + 1], dir[3 * i + 2] };
		V3 ps = s;
		float t = 0.0f;
		bool h = false;
		int evals = 0;  // distance evaluations along the ray, the hit test included
		for (int it = 0; it < iters; ++it)
		{
			bool ok;
			float dist = interpret(words, nwords, ps, &ok);
			if (!ok) return -1;
			++evals;
			if (dist <= eps) { h = true; break; }
			t += dist;
			if (t > far_) break;
			ps = add3(mul3s(d, t), s);
		}
		hit[i] = h;
		travel[i] = h ? t : std::numeric_limits<float>::infinity();
		steps[i] = evals;
	}
	return 0;
}

// LuaPivotTowards (lua_sdf.cpp:452-559) over the interpreter: the glm
// steps written out (radians in float, the 0.5 degree and 0.001 tests in
// double as the reference's double literals make them).  Returns 0, 2 when
// the heading is antiparallel to down (the reference draws random axes; no
// output), 3 when `max_iterations` turns of the loop did not end it, or -1
// for a malformed program.
int tgo_pivot_towards_words(const uint32_t* words, uint64_t nwords, float max_dist, float margin, float max_angle_deg,
	const float* pivot_in, const float* heading_in, const float* down_in, int max_iterations, float* tail_out,
	float* heading_out)
{
	const float deg = 0.01745329251994329576923690768489f;
	const float max_angle = max_angle_deg * deg;
	const V3 pivot = { pivot_in[0], pivot_in[1], pivot_in[2] };
	V3 heading = { heading_in[0], heading_in[1], heading_in[2] };
	V3 down = { down_in[0], down_in[1], down_in[2] };
	V3 tail, axis;
	bool ok = true;
	const float hl = len3(heading);
	if (hl == 0.0f || std::isnan(hl) || std::isinf(hl))
	{
		tail = pivot;
		heading = { 0.0f, 0.0f, 0.0f };
	}
	else
	{
		heading = div3s(heading, hl);
		const float dl = len3(down);
		if (dl == 0.0f || std::isnan(dl) || std::isinf(dl))
			down = { 0.0f, 0.0f, -1.0f };
		else
			down = div3s(down, dl);
		tail = add3(mul3s(heading, max_dist), pivot);
		axis = cross3(heading, down);
		if (len3(axis) == 0.0f)
		{
			if (double(len3(sub3(heading, down))) <= 0.001)
				heading = down;
			else
				return 2;
		}
		else
		{
			axis = mul3s(axis, 1.0f / std::sqrt(dot3(axis, axis)));
			float remaining = max_angle;
			const float fnord = max_dist * max_dist * 2.0f;
			int it = 0;
			for (;; ++it)
			{
				if (it >= max_iterations) return 3;
				float dist = interpret(words, nwords, tail, &ok);
				if (!ok) return -1;
				if (dist <= margin) break;
				dist = gmin(std::fabs(dist), max_dist);
				const float cosine = std::fabs((fnord - dist * dist) / fnord);
				const float maybe = std::acos(cosine);
				const V3 arm = sub3(tail, pivot);
				const float downward = std::fabs(std::acos(dot3(mul3s(arm, 1.0f / std::sqrt(dot3(arm, arm))), down)));
				const float angle = gmin(remaining, gmin(downward, maybe));
				remaining -= angle;
				const float s = std::sin(angle * 0.5f);
				const Q4 rot = { std::cos(angle * 0.5f), axis.x * s, axis.y * s, axis.z * s };
				tail = add3(qrotate(rot, sub3(tail, pivot)), pivot);
				if (double(angle) <= 0.5 * 0.01745329251994329576923690768489) break;
			}
			const V3 arm = sub3(tail, pivot);
			const float cosine = gmin(std::fabs(dot3(mul3s(arm, 1.0f / std::sqrt(dot3(arm, arm))), heading)), 1.0f);
			const float final_angle = gmin(max_angle, std::acos(cosine));
			const float s = std::sin(final_angle * 0.5f);
			const Q4 rot = { std::cos(final_angle * 0.5f), axis.x * s, axis.y * s, axis.z * s };
			heading = qrotate(rot, heading);
			tail = add3(mul3s(heading, max_dist), pivot);
		}
	}
	tail_out[0] = tail.x; tail_out[1] = tail.y; tail_out[2] = tail.z;
	heading_out[0] = heading.x; heading_out[1] = heading.y; heading_out[2] = heading.z;
	return 0;
}

// Picking (DeliverMouseButton with SDFModel::RayMarch, sdf_model.cpp:150-206,
// 259-264) over the interpreter: model m's program is words[offsets[m] ..
// offsets[m + 1]); transforms are 8 floats, rotation (x, y, z, w),
// translation, scale.  Local ray: rotate(inverse(q), start - t) / s and
// rotate(inverse(q), dir); RayMarch with 1000 steps, epsilon 0.001; nearest
// hit by local travel, the first on ties.  Returns the model or -1.
int tgo_pick_words(const uint32_t* words, const uint64_t* offsets, const float* xf, uint32_t count, const float* start,
	const float* dir, float* travel_out, float* position_out)
{
	float nearest = std::numeric_limits<float>::infinity();
	int match = -1;
	const V3 s0 = { start[0], start[1], start[2] }, d0 = { dir[0], dir[1], dir[2] };
	for (uint32_t m = 0; m < count; ++m)
	{
		const float* t = xf + 8 * m;
		const Q4 inv = qinverse(Q4{ t[3], t[0], t[1], t[2] });
		const V3 s = div3s(qrotate(inv, sub3(s0, V3{ t[4], t[5], t[6] })), t[7]);
		V3 d = qrotate(inv, d0);
		d = mul3s(d, 1.0f / std::sqrt(dot3(d, d)));
		const uint32_t* w = words + offsets[m];
		const uint64_t nw = offsets[m + 1] - offsets[m];
		V3 ps = s;
		float tr = 0.0f;
		bool h = false;
		for (int it = 0; it < 1000; ++it)
		{
			bool ok;
			const float dist = interpret(w, nw, ps, &ok);
			if (!ok) return -2;
			if (dist <= 0.001f) { h = true; break; }
			tr += dist;
			ps = add3(mul3s(d, tr), s);
		}
		if (h && tr < nearest)
		{
			nearest = tr;
			match = int(m);
			position_out[0] = ps.x; position_out[1] = ps.y; position_out[2] = ps.z;
		}
	}
	*travel_out = nearest;
	return match;
}

// VoxExport's voxel selection (magica.cpp:27-70) over the interpreter, in
// raster order: Size = ivec3(ceil(max - min) * grid_size), radius = the
// distance from min to mix(min, max, 0.5 / Size) (glm's mix, min (1 - t) +
// max t), cell (x, y, z) kept when
// |d(mix(min, max, (x + .5, y + .5, z + .5) / Size))| <= radius.  Writes the
// size and up to `cap` cell indices x + Size.x (y + Size.y z); returns the
// count, -1 for a malformed program, -2 for an empty grid.
int64_t tgo_vox_cells(const uint32_t* words, uint64_t nwords, const float* mn, const float* mx, float grid_size,
	int32_t* size_out, uint32_t* cells, uint64_t cap)
{
	int32_t size[3];
	for (int a = 0; a < 3; ++a) size[a] = int32_t(std::ceil(mx[a] - mn[a]) * grid_size);
	for (int a = 0; a < 3; ++a) size_out[a] = size[a];
	if (size[0] <= 0 || size[1] <= 0 || size[2] <= 0) return -2;
	const V3 lo = { mn[0], mn[1], mn[2] }, hi = { mx[0], mx[1], mx[2] };
	const V3 fs = { float(size[0]), float(size[1]), float(size[2]) };
	// glm::mix(a, b, t) = a * (1 - t) + b * t (glm detail/func_common.inl:87)
	auto mix = [&](V3 t) {
		return V3{ lo.x * (1.0f - t.x) + hi.x * t.x, lo.y * (1.0f - t.y) + hi.y * t.y, lo.z * (1.0f - t.z) + hi.z * t.z };
	};
	const V3 p0 = mix(V3{ 0.5f / fs.x, 0.5f / fs.y, 0.5f / fs.z });
	const float radius = len3(sub3(p0, lo));
	const int64_t slice = int64_t(size[0]) * size[1], total = slice * size[2];
	int64_t count = 0;
	for (int64_t i = 0; i < total; ++i)
	{
		const int64_t z = i / slice, y = (i % slice) / size[0], x = i % size[0];
		const V3 p = mix(V3{ float(double(x) + .5) / fs.x, float(double(y) + .5) / fs.y, float(double(z) + .5) / fs.z });
		bool ok;
		const float d = interpret(words, nwords, p, &ok);
		if (!ok) return -1;
		if (std::fabs(d) <= radius)
		{
			if (uint64_t(count) < cap) cells[count] = uint32_t(i);
			++count;
		}
	}
	return count;
}

// ---------------------------------------------------------------------------
// Serial naive surface nets (surface_nets.cpp:30-580) over sample values.
// Values: (sx+1)(sy+1)(sz+1), x fastest.  Vertex ids in raster order of the
// active cube index x + y*sx + z*sx*sy.  Faces are emitted in raster order of
// the active cube (the reference emits them in unordered_map order; callers
// compare faces as sets).

struct OMesh
{
	std::vector<float> v;
	std::vector<uint64_t> f;
};

void* tgo_surface_nets(const float* values, const float* origin, const float* step, const uint64_t* count)
{
	const uint64_t sx = count[0], sy = count[1], sz = count[2];
	const uint64_t nx = sx + 1, ny = sy + 1;
	const float gx = origin[0], gy = origin[1], gz = origin[2];
	const float dx = step[0], dy = step[1], dz = step[2];
	auto val = [&](uint64_t i, uint64_t j, uint64_t k) { return values[i + nx * (j + ny * k)]; };
	static const int corner[8][3] = { {0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1} };
	static const int edges[12][2] = { {0,1},{1,2},{2,3},{3,0},{4,5},{5,6},{6,7},{7,4},{0,4},{1,5},{2,6},{3,7} };
	const float maxx = gx + float(sx) * dx, maxy = gy + float(sy) * dy, maxz = gz + float(sz) * dz;

	OMesh* m = new OMesh;
	std::unordered_map<uint64_t, uint64_t> ids;
	std::vector<uint64_t> order;
	for (uint64_t k = 0; k < sz; ++k)
	for (uint64_t j = 0; j < sy; ++j)
	for (uint64_t i = 0; i < sx; ++i)
	{
		float s[8];
		for (int c = 0; c < 8; ++c) s[c] = val(i + corner[c][0], j + corner[c][1], k + corner[c][2]);
		bool bip[12];
		bool any = false;
		for (int e = 0; e < 12; ++e)
		{
			bip[e] = (s[edges[e][0]] >= 0.0f) != (s[edges[e][1]] >= 0.0f);
			any = any || bip[e];
		}
		if (!any) continue;
		float sumx = 0.f, sumy = 0.f, sumz = 0.f;
		int cnt = 0;
		for (int e = 0; e < 12; ++e)
		{
			if (!bip[e]) continue;
			int a = edges[e][0], b = edges[e][1];
			float p1x = float(i + corner[a][0]), p1y = float(j + corner[a][1]), p1z = float(k + corner[a][2]);
			float p2x = float(i + corner[b][0]), p2y = float(j + corner[b][1]), p2z = float(k + corner[b][2]);
			float t = (0.0f - s[a]) / (s[b] - s[a]);
			sumx = sumx + (p1x + t * (p2x - p1x));
			sumy = sumy + (p1y + t * (p2y - p1y));
			sumz = sumz + (p1z + t * (p2z - p1z));
			++cnt;
		}
		float fc = float(cnt);
		float cx = sumx / fc, cy = sumy / fc, cz = sumz / fc;
		float vx = gx + (maxx - gx) * (cx - 0.f) / (float(sx) - 0.f);
		float vy = gy + (maxy - gy) * (cy - 0.f) / (float(sy) - 0.f);
		float vz = gz + (maxz - gz) * (cz - 0.f) / (float(sz) - 0.f);
		uint64_t idx = i + j * sx + k * sx * sy;
		ids[idx] = m->v.size() / 3;
		order.push_back(idx);
		m->v.push_back(vx); m->v.push_back(vy); m->v.push_back(vz);
	}
	for (uint64_t idx : order)
	{
		uint64_t i = idx % sx, j = (idx / sx) % sy, k = idx / (sx * sy);
		if (i == 0 || j == 0 || k == 0) continue;
		uint64_t nb[6][3] = { {i-1,j,k},{i-1,j-1,k},{i,j-1,k},{i,j-1,k-1},{i,j,k-1},{i-1,j,k-1} };
		float e04[2] = { val(i, j, k), val(i, j, k + 1) };
		float e30[2] = { val(i, j + 1, k), val(i, j, k) };
		float e01[2] = { val(i, j, k), val(i + 1, j, k) };
		const float* es[3] = { e04, e30, e01 };
		static const int quad[3][3] = { {0,1,2},{0,5,4},{2,3,4} };
		for (int q = 0; q < 3; ++q)
		{
			uint64_t nid[3];
			bool ok = true;
			for (int t = 0; t < 3; ++t)
			{
				const uint64_t* g = nb[quad[q][t]];
				auto it = ids.find(g[0] + g[1] * sx + g[2] * sx * sy);
				if (it == ids.end()) { ok = false; break; }
				nid[t] = it->second;
			}
			if (!ok) continue;
			bool fwd = es[q][1] > es[q][0];
			uint64_t v0 = ids[idx];
			uint64_t v1 = fwd ? nid[0] : nid[2], v2 = nid[1], v3 = fwd ? nid[2] : nid[0];
			m->f.push_back(v0); m->f.push_back(v1); m->f.push_back(v2);
			m->f.push_back(v0); m->f.push_back(v2); m->f.push_back(v3);
		}
	}
	return m;
}

uint64_t tgo_mesh_vertex_count(void* m) { return ((OMesh*)m)->v.size() / 3; }
uint64_t tgo_mesh_face_count(void* m) { return ((OMesh*)m)->f.size() / 3; }
void tgo_mesh_copy(void* m, float* v, uint64_t* f)
{
	OMesh* o = (OMesh*)m;
	std::memcpy(v, o->v.data(), o->v.size() * 4);
	std::memcpy(f, o->f.data(), o->f.size() * 8);
}
void tgo_mesh_free(void* m) { delete (OMesh*)m; }


// mt19937 self-check value (the standard's 10000th output for the default seed)
// and the vertex exchange map of ApplyVertexSequence.
uint32_t tgo_mt19937_10000th(void)
{
	std::mt19937 g;
	uint32_t v = 0;
	for (int i = 0; i < 10000; ++i) v = g();
	return v;
}

void tgo_vertex_sequence(uint64_t n, uint32_t* exchange) { vertex_sequence(n, exchange); }

}  // extern "C"
