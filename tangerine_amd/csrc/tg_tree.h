// Host-side CSG tree: the SDFNode surface (tangerine/sdf_evaluator.h:171-264)
// restated for the device path.  Trees are immutable shared graphs; every
// transform copies first, as the Lua bindings do (tangerine/lua_sdf.cpp:56-170).
// Evaluation lives on the GPU; the host keeps only what the reference does on
// the host around evaluation: construction, bounds and compilation.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace tg {

struct Vec3 { float x, y, z; };
struct Quat { float w, x, y, z; };

struct Transform3
{
	Quat rotation{ 1.0f, 0.0f, 0.0f, 0.0f };
	Vec3 translation{ 0.0f, 0.0f, 0.0f };
	float scale = 1.0f;
};

struct Box3 { Vec3 lo, hi; };

struct Node;
using NodeRef = std::shared_ptr<Node>;

struct Node
{
	enum class Kind { Brush, Set, Flate };
	Kind kind = Kind::Brush;
	uint32_t opcode = 0;          // brush or set opcode (reference numbering)
	float params[3] = { 0, 0, 0 };
	Box3 local_bounds{};
	Transform3 xform;             // brushes only
	NodeRef lhs, rhs;             // set: both; flate: lhs
	float threshold = 0.0f;       // blend threshold or flate radius
	uint32_t stack_size = 1;
	uint32_t leaves = 1;
};

// Builders (SDF::* factories, tangerine/sdf_evaluator.cpp:1206-1372).
NodeRef make_sphere(float radius);
NodeRef make_ellipsoid(float rx, float ry, float rz);
NodeRef make_box(float ex, float ey, float ez);
NodeRef make_torus(float major, float minor);
NodeRef make_cylinder(float radius, float half_height);
NodeRef make_plane(float nx, float ny, float nz);
NodeRef make_cone(float radius, float height);
NodeRef make_coninder(float radius_lo, float radius_hi, float height);
NodeRef make_set(uint32_t opcode, NodeRef lhs, NodeRef rhs, float threshold);
NodeRef make_flate(NodeRef child, float radius);

NodeRef copy_tree(const NodeRef& n);
void move_tree(Node* n, Vec3 offset);
void rotate_tree(Node* n, Quat q);
void scale_tree(Node* n, float s);
Quat axis_rotation(int axis, float degrees);

Box3 tree_bounds(const Node* n);
bool tree_has_finite_bounds(const Node* n);

// Postfix words in reference ProgramBuffer layout, terminated by Stop.
void compile_tree(const Node* n, std::vector<uint32_t>& words);

}  // namespace tg
