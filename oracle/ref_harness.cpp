// Test-infrastructure only: a C-ABI harness around the reference's own
// CSG / octree / surface-nets sources, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/libtgref.so.  It is used to generate the
// golden vectors under tests/golden/ and, optionally, as the timed CPU
// baseline ("kind": "reference").  Nothing in the product links it.
//
// Every entry point forwards to the reference implementation unchanged:
//   builders      -> SDF::* factories            tangerine/sdf_evaluator.cpp:1169-1372
//   transforms    -> Lua copy-then-transform     tangerine/lua_sdf.cpp:56-170
//   eval          -> SDFNode::Eval               tangerine/sdf_evaluator.cpp:463,774,1059
//   interp_eval   -> SDFInterpreter::Eval        tangerine/sdf_evaluator.cpp:1376-1605
//   clip          -> SDFNode::Clip               tangerine/sdf_evaluator.cpp:468,782,1064
//   octree        -> SDFOctree                   tangerine/sdf_evaluator.cpp:1609-2005
//   gradient      -> SDFNode::Gradient           tangerine/sdf_evaluator.cpp:298-333
//   raymarch      -> SDFNode::RayMarch           tangerine/sdf_evaluator.cpp:336-353
//   surface nets  -> isosurface::surface_nets    third_party/naive-surface-nets/src/surface_nets.cpp:30-580
//   threaded export -> AsyncParallelSurfaceNets thunks (surface_nets.cpp:829-1122)
//                      driven like MeshExportThread (tangerine/export.cpp:320-381)
//                      on a pool of max(hw-1,2)+1 threads (tangerine/scheduler.cpp:141-148)
//   pool grid fill  -> Pool (tangerine/threadpool.cpp:27-40), loop shape of
//                      tangerine/export.cpp:401-427, SDFOctree::Eval(Exact)
//   point cloud     -> the steps of PointCloudExportThread (export.cpp:384-480;
//                      export.cpp itself needs SDL headers) over the whole tree
//   pivot_towards   -> the steps of LuaPivotTowards (lua_sdf.cpp:452-559; the
//                      Lua sources need Lua headers) over the reference tree
//   pick            -> DeliverMouseButton's nearest hit over SDFModel::RayMarch
//                      (sdf_model.cpp:150-206, 259-264) with Transform
//   vox export      -> the steps of VoxExport (magica.cpp:27-70) serially,
//                      written by third_party/voxwriter's vox::VoxWriter

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "sdf_evaluator.h"
#include "transform.h"
#include "surface_nets.h"
#include "threadpool.h"
#include "VoxWriter.h"

using namespace glm;

namespace {

struct Handle
{
	SDFNodeShared Node;
};

struct Octree
{
	SDFOctreeShared Tree;
};

struct Mesh
{
	isosurface::mesh M;
	uint64_t Evals = 0;
};

Handle* Wrap(SDFNodeShared Node)
{
	if (!Node)
	{
		return nullptr;
	}
	return new Handle{ Node };
}

isosurface::regular_grid_t MakeGrid(const float* Origin, const float* Step, const uint64_t* Count)
{
	isosurface::regular_grid_t Grid;
	Grid.x = Origin[0];
	Grid.y = Origin[1];
	Grid.z = Origin[2];
	Grid.dx = Step[0];
	Grid.dy = Step[1];
	Grid.dz = Step[2];
	Grid.sx = Count[0];
	Grid.sy = Count[1];
	Grid.sz = Count[2];
	return Grid;
}

}  // namespace

extern "C" {

void* tgref_sphere(float R) { return Wrap(SDF::Sphere(R)); }
void* tgref_ellipsoid(float X, float Y, float Z) { return Wrap(SDF::Ellipsoid(X, Y, Z)); }
void* tgref_box(float X, float Y, float Z) { return Wrap(SDF::Box(X, Y, Z)); }
void* tgref_torus(float Major, float Minor) { return Wrap(SDF::Torus(Major, Minor)); }
void* tgref_cylinder(float R, float H) { return Wrap(SDF::Cylinder(R, H)); }
void* tgref_plane(float X, float Y, float Z) { return Wrap(SDF::Plane(X, Y, Z)); }
void* tgref_cone(float R, float H) { return Wrap(SDF::Cone(R, H)); }
void* tgref_coninder(float L, float H, float Height) { return Wrap(SDF::Coninder(L, H, Height)); }

// Set operators take the operands as the Lua bindings do (no copy, lua_sdf.cpp:348-383).
void* tgref_op(int Kind, float Threshold, void* A, void* B)
{
	SDFNodeShared& L = ((Handle*)A)->Node;
	SDFNodeShared& R = ((Handle*)B)->Node;
	switch (Kind)
	{
	case 0: return Wrap(SDF::Union(L, R));
	case 1: return Wrap(SDF::Inter(L, R));
	case 2: return Wrap(SDF::Diff(L, R));
	case 3: return Wrap(SDF::BlendUnion(Threshold, L, R));
	case 4: return Wrap(SDF::BlendInter(Threshold, L, R));
	case 5: return Wrap(SDF::BlendDiff(Threshold, L, R));
	default: return nullptr;
	}
}

void* tgref_flate(void* A, float Radius)
{
	return Wrap(SDF::Flate(((Handle*)A)->Node, Radius));
}

// Transforms copy first, as every Lua transform binding does (lua_sdf.cpp:56-170).
void* tgref_move(void* A, float X, float Y, float Z)
{
	SDFNodeShared N = ((Handle*)A)->Node->Copy();
	N->Move(vec3(X, Y, Z));
	return Wrap(N);
}

void* tgref_rotate(void* A, float X, float Y, float Z, float W)
{
	SDFNodeShared N = ((Handle*)A)->Node->Copy();
	N->Rotate(quat(W, X, Y, Z));
	return Wrap(N);
}

void* tgref_rotate_axis(void* A, int Axis, float Degrees)
{
	SDFNodeShared N = ((Handle*)A)->Node->Copy();
	if (Axis == 0) SDF::RotateX(N, Degrees);
	else if (Axis == 1) SDF::RotateY(N, Degrees);
	else SDF::RotateZ(N, Degrees);
	return Wrap(N);
}

void* tgref_scale(void* A, float S)
{
	SDFNodeShared N = ((Handle*)A)->Node->Copy();
	N->Scale(S);
	return Wrap(N);
}

void tgref_free(void* A) { delete (Handle*)A; }

int tgref_leaf_count(void* A) { return ((Handle*)A)->Node->LeafCount(); }
uint64_t tgref_stack_size(void* A) { return ((Handle*)A)->Node->StackSize; }

void tgref_bounds(void* A, float* Out)
{
	AABB B = ((Handle*)A)->Node->Bounds();
	Out[0] = B.Min.x; Out[1] = B.Min.y; Out[2] = B.Min.z;
	Out[3] = B.Max.x; Out[4] = B.Max.y; Out[5] = B.Max.z;
}

void tgref_eval(void* A, const float* P, uint64_t N, float* Out)
{
	SDFNode* Node = ((Handle*)A)->Node.get();
	for (uint64_t i = 0; i < N; ++i)
	{
		Out[i] = Node->Eval(vec3(P[3 * i], P[3 * i + 1], P[3 * i + 2]));
	}
}

uint64_t tgref_compile(void* A, uint32_t* Words, uint64_t Capacity)
{
	SDFInterpreter Interp(((Handle*)A)->Node);
	uint64_t N = Interp.Program.Size();
	for (uint64_t i = 0; i < N && i < Capacity; ++i)
	{
		std::memcpy(&Words[i], &Interp.Program.Words[i], 4);
	}
	return N;
}

void tgref_interp_eval(void* A, const float* P, uint64_t N, float* Out)
{
	SDFInterpreter Interp(((Handle*)A)->Node);
	for (uint64_t i = 0; i < N; ++i)
	{
		Out[i] = Interp.Eval(vec3(P[3 * i], P[3 * i + 1], P[3 * i + 2]));
	}
}

void* tgref_clip(void* A, float X, float Y, float Z, float R)
{
	return Wrap(((Handle*)A)->Node->Clip(vec3(X, Y, Z), R));
}

void tgref_gradient(void* A, const float* P, uint64_t N, float* Out)
{
	SDFNode* Node = ((Handle*)A)->Node.get();
	for (uint64_t i = 0; i < N; ++i)
	{
		vec3 G = Node->Gradient(vec3(P[3 * i], P[3 * i + 1], P[3 * i + 2]));
		Out[3 * i] = G.x; Out[3 * i + 1] = G.y; Out[3 * i + 2] = G.z;
	}
}

void tgref_raymarch(void* A, const float* Start, const float* Dir, uint64_t N, int Iterations, float Epsilon,
	uint8_t* Hit, float* Travel, float* Position)
{
	SDFNode* Node = ((Handle*)A)->Node.get();
	for (uint64_t i = 0; i < N; ++i)
	{
		RayHit R = Node->RayMarch(
			vec3(Start[3 * i], Start[3 * i + 1], Start[3 * i + 2]),
			vec3(Dir[3 * i], Dir[3 * i + 1], Dir[3 * i + 2]), Iterations, Epsilon);
		Hit[i] = R.Hit ? 1 : 0;
		Travel[i] = R.Travel;
		Position[3 * i] = R.Position.x; Position[3 * i + 1] = R.Position.y; Position[3 * i + 2] = R.Position.z;
	}
}

void* tgref_octree_create(void* A, float TargetSize, int Coalesce, int MaxDepth, float Margin)
{
	SDFOctreeShared T = SDFOctree::Create(((Handle*)A)->Node, TargetSize, Coalesce != 0, MaxDepth, Margin);
	if (!T) return nullptr;
	return new Octree{ T };
}

void tgref_octree_free(void* O) { delete (Octree*)O; }

void tgref_octree_eval(void* O, const float* P, uint64_t N, int Exact, float* Out)
{
	SDFOctree* T = ((Octree*)O)->Tree.get();
	for (uint64_t i = 0; i < N; ++i)
	{
		Out[i] = T->Eval(vec3(P[3 * i], P[3 * i + 1], P[3 * i + 2]), Exact != 0);
	}
}

// Dense grid fill with SDFOctree::Eval(Exact) on the reference Pool: every
// Pool thread claims sample indices from one atomic counter, as in
// export.cpp:401-427.  Samples are (Count+1) per axis, x fastest.
// Only z-slabs [K0, K1) are filled (for bounded CPU timing).
int tgref_pool_grid_eval(void* O, const float* Origin, const float* Step, const uint64_t* Count,
	uint64_t K0, uint64_t K1, float* Out)
{
	SDFOctree* T = ((Octree*)O)->Tree.get();
	const uint64_t NX = Count[0] + 1, NY = Count[1] + 1;
	const uint64_t Slice = NX * NY;
	const uint64_t Total = Slice * (K1 - K0);
	std::atomic<uint64_t> Next(0);
	Pool([&]()
	{
		for (;;)
		{
			uint64_t n = Next.fetch_add(1);
			if (n >= Total) break;
			uint64_t k = n / Slice + K0;
			uint64_t j = (n % Slice) / NX;
			uint64_t i = n % NX;
			float X = Origin[0] + float(i) * Step[0];
			float Y = Origin[1] + float(j) * Step[1];
			float Z = Origin[2] + float(k) * Step[2];
			Out[n] = T->Eval(vec3(X, Y, Z), true);
		}
	});
	return std::max(int(std::thread::hardware_concurrency()), 2);
}

// PointCloudExportThread (tangerine/export.cpp:384-480) with the octree
// replaced by the whole tree: the distance from SDFInterpreter::Eval, the
// refinement direction from SDFNode::Gradient, as SDFOctree::Eval and
// SDFOctree::Gradient call them on a leaf.  Serial, so the points come in
// raster order.  Writes up to Capacity points; returns the count.
int64_t tgref_point_cloud(void* A, const float* Lo, const float* Hi, const float* StepIn, int RefineIterations,
	float* Out, uint64_t Capacity)
{
	SDFNodeShared Node = ((Handle*)A)->Node;
	SDFInterpreter Interp(Node);
	const vec3 Start(Lo[0], Lo[1], Lo[2]), Stop(Hi[0], Hi[1], Hi[2]), Step(StepIn[0], StepIn[1], StepIn[2]);
	const vec3 Half = Step / vec3(2.0);
	const float Diagonal = length(Half);
	const ivec3 Iterations = ivec3(ceil((Stop - Start) / Step));
	const int Slice = Iterations.x * Iterations.y;
	const int TotalCells = Iterations.x * Iterations.y * Iterations.z;
	std::vector<vec3> Vertices;
	for (int i = 0; i < TotalCells; ++i)
	{
		float Z = float(i / Slice) * Step.z + Start.z;
		float Y = float((i % Slice) / Iterations.x) * Step.y + Start.y;
		float X = float(i % Iterations.x) * Step.x + Start.x;
		vec3 Cursor = vec3(X, Y, Z) + Half;
		float Dist = Interp.Eval(Cursor);
		if (abs(Dist) < Diagonal)
		{
			Vertices.push_back(Cursor);
		}
	}
	for (vec3& Vertex : Vertices)
	{
		vec3 Low = Vertex - vec3(Half);
		vec3 High = Vertex + vec3(Half);
		vec3 Cursor = Vertex;
		for (int r = 0; r < RefineIterations; ++r)
		{
			vec3 RayDir = Node->Gradient(Cursor);
			float Dist = Interp.Eval(Cursor) * -1.0;
			Cursor += RayDir * Dist;
		}
		Cursor = clamp(Cursor, Low, High);
		if (distance(Cursor, Vertex) <= Diagonal)
		{
			Vertex = Cursor;
		}
	}
	for (uint64_t i = 0; i < Vertices.size() && i < Capacity; ++i)
	{
		Out[3 * i] = Vertices[i].x; Out[3 * i + 1] = Vertices[i].y; Out[3 * i + 2] = Vertices[i].z;
	}
	return int64_t(Vertices.size());
}

// LuaPivotTowards (tangerine/lua_sdf.cpp:452-559) over the reference tree:
// the same glm steps, Lua argument handling aside.  Returns 0, or 2 when the
// heading is antiparallel to down (the reference then draws random axes from
// the Lua environment's generator; nothing is written).
int tgref_pivot_towards(void* A, float MaxDist, float Margin, float MaxAngleDegrees, const float* PivotIn,
	const float* HeadingIn, const float* DownIn, float* TailOut, float* HeadingOut)
{
	SDFNode* Node = ((Handle*)A)->Node.get();
	const float MaxAngle = glm::radians(MaxAngleDegrees);
	glm::vec3 Pivot(PivotIn[0], PivotIn[1], PivotIn[2]);
	glm::vec3 Heading(HeadingIn[0], HeadingIn[1], HeadingIn[2]);
	glm::vec3 Down(DownIn[0], DownIn[1], DownIn[2]);
	glm::vec3 Tail;
	glm::vec3 Axis;
	{
		float HeadingLen = glm::length(Heading);
		if (HeadingLen == 0.0 || glm::isnan(HeadingLen) || glm::isinf(HeadingLen))
		{
			Tail = Pivot;
			Heading = glm::vec3(0.0, 0.0, 0.0);
			goto finish;
		}
		else
		{
			Heading /= HeadingLen;
		}
	}
	{
		float DownLen = glm::length(Down);
		if (DownLen == 0.0 || glm::isnan(DownLen) || glm::isinf(DownLen))
		{
			Down = glm::vec3(0, 0, -1);
		}
		else
		{
			Down /= DownLen;
		}
	}
	Tail = Heading * MaxDist + Pivot;
	Axis = glm::cross(Heading, Down);
	{
		float AxisLen = glm::length(Axis);
		if (AxisLen == 0.0)
		{
			if (glm::distance(Down, Heading) <= 0.001)
			{
				Heading = Down;
				goto finish;
			}
			return 2;
		}
	}
	Axis = glm::normalize(Axis);
	{
		float Dist;
		float RemainingMaxAngle = MaxAngle;
		const float Fnord = MaxDist * MaxDist * 2;
		while (true)
		{
			Dist = Node->Eval(Tail);
			if (Dist <= Margin)
			{
				break;
			}
			Dist = glm::min(glm::abs(Dist), MaxDist);
			float Cosine = glm::abs((Fnord - Dist * Dist) / Fnord);
			float MaybeAngle = glm::acos(Cosine);
			float MaxDownward = glm::abs(glm::acos(glm::dot(glm::normalize(Tail - Pivot), Down)));
			float Angle = glm::min(RemainingMaxAngle, glm::min(MaxDownward, MaybeAngle));
			RemainingMaxAngle -= Angle;
			glm::quat Rotation = glm::angleAxis(Angle, Axis);
			Tail = glm::rotate(Rotation, Tail - Pivot) + Pivot;
			if (Angle <= glm::radians(0.5))
			{
				break;
			}
		}
		const float Cosine = glm::min(glm::abs(glm::dot(glm::normalize(Tail - Pivot), Heading)), 1.0f);
		const float MaybeAngle = glm::acos(Cosine);
		const float FinalAngle = glm::min(MaxAngle, MaybeAngle);
		glm::quat Rotation = glm::angleAxis(FinalAngle, Axis);
		Heading = glm::rotate(Rotation, Heading);
		Tail = Heading * MaxDist + Pivot;
	}
finish:
	for (int i = 0; i < 3; ++i)
	{
		TailOut[i] = Tail[i];
		HeadingOut[i] = Heading[i];
	}
	return 0;
}

// Mouse picking (DeliverMouseButton, tangerine/sdf_model.cpp:150-206, with
// SDFModel::RayMarch, :259-264): every model's ray in its local frame
// (Transform::ApplyInv and the inverse rotation), RayMarch with 1000 steps
// and the default epsilon; the nearest hit by local travel wins (the first
// on ties).  Transforms: 8 floats each, rotation (x, y, z, w), translation,
// scale.  Returns the model index or -1; travel and the local position.
int tgref_pick(void* const* Models, const float* Transforms, uint32_t Count, const float* Start, const float* Dir,
	float* TravelOut, float* PositionOut)
{
	float Nearest = std::numeric_limits<float>::infinity();
	int Match = -1;
	const glm::vec3 RayStart(Start[0], Start[1], Start[2]), RayDir(Dir[0], Dir[1], Dir[2]);
	for (uint32_t m = 0; m < Count; ++m)
	{
		const float* T = Transforms + 8 * m;
		Transform LocalToWorld;
		LocalToWorld.Rotation = glm::quat(T[3], T[0], T[1], T[2]);
		LocalToWorld.Translation = glm::vec3(T[4], T[5], T[6]);
		LocalToWorld.Scalation = T[7];
		glm::vec3 RelativeOrigin = LocalToWorld.ApplyInv(RayStart);
		glm::vec3 RelativeRayDir = glm::rotate(glm::inverse(LocalToWorld.Rotation), RayDir);
		RayHit Query = ((Handle*)Models[m])->Node->RayMarch(RelativeOrigin, RelativeRayDir, 1000);
		if (Query.Hit && Query.Travel < Nearest)
		{
			Nearest = Query.Travel;
			Match = int(m);
			PositionOut[0] = Query.Position.x;
			PositionOut[1] = Query.Position.y;
			PositionOut[2] = Query.Position.z;
		}
	}
	*TravelOut = Nearest;
	return Match;
}

// VoxExport (tangerine/magica.cpp:27-70; magica.cpp needs the GUI's extern.h)
// over the reference tree, serially in raster order, written by the
// reference's own vox::VoxWriter.  Returns the voxel count, or -1 if the
// grid is empty.
int64_t tgref_vox_export(void* A, float GridSize, int ColorIndex, const char* Path)
{
	SDFNodeShared Evaluator = ((Handle*)A)->Node;
	AABB Bounds = Evaluator->Bounds();
	glm::ivec3 Size = glm::ivec3(glm::ceil(Bounds.Max - Bounds.Min) * GridSize);
	if (Size.x <= 0 || Size.y <= 0 || Size.z <= 0)
	{
		return -1;
	}
	float Radius;
	{
		glm::vec3 Alpha = glm::vec3(.5, .5, .5) / glm::vec3(Size);
		Radius = glm::distance(Bounds.Min, glm::mix(Bounds.Min, Bounds.Max, Alpha));
	}
	const int Slice = Size.x * Size.y;
	const int TotalCells = Size.x * Size.y * Size.z;
	vox::VoxWriter Writer(Size.x, Size.y, Size.z);
	int64_t Count = 0;
	for (int i = 0; i < TotalCells; ++i)
	{
		int z = i / Slice;
		int y = (i % Slice) / Size.x;
		int x = i % Size.x;
		glm::vec3 Alpha = glm::vec3(x + .5, y + .5, z + .5) / glm::vec3(Size);
		glm::vec3 Point = glm::mix(Bounds.Min, Bounds.Max, Alpha);
		float Dist = Evaluator->Eval(Point);
		if (std::abs(Dist) <= Radius)
		{
			Writer.AddVoxel(x, y, z, abs(ColorIndex) % 255 + 1);
			++Count;
		}
	}
	std::string PathStr(Path);
	Writer.SaveToFile(PathStr);
	return Count;
}

// vox::VoxWriter over a list of raster cell indices, added in the given order.
void tgref_vox_write(const uint32_t* Size, const uint32_t* Cells, uint64_t Count, int ColorIndex, const char* Path)
{
	const int32_t SX = int32_t(Size[0]), SY = int32_t(Size[1]), SZ = int32_t(Size[2]);
	vox::VoxWriter Writer(SX, SY, SZ);
	const uint64_t Slice = uint64_t(Size[0]) * Size[1];
	for (uint64_t i = 0; i < Count; ++i)
	{
		const uint64_t c = Cells[i];
		Writer.AddVoxel(int32_t(c % Size[0]), int32_t((c % Slice) / Size[0]), int32_t(c / Slice), abs(ColorIndex) % 255 + 1);
	}
	std::string PathStr(Path);
	Writer.SaveToFile(PathStr);
}

// Serial reference surface nets over precomputed sample values (identical-grid
// topology checks).  Values are (Count+1) per axis, x fastest.
void* tgref_surface_nets_values(const float* Values, const float* Origin, const float* Step, const uint64_t* Count)
{
	isosurface::regular_grid_t Grid = MakeGrid(Origin, Step, Count);
	const uint64_t NX = Count[0] + 1, NY = Count[1] + 1;
	auto Lookup = [&](float X, float Y, float Z) -> float
	{
		uint64_t i = (uint64_t)std::llround((X - Grid.x) / Grid.dx);
		uint64_t j = (uint64_t)std::llround((Y - Grid.y) / Grid.dy);
		uint64_t k = (uint64_t)std::llround((Z - Grid.z) / Grid.dz);
		return Values[i + NX * (j + NY * k)];
	};
	Mesh* M = new Mesh;
	std::atomic_bool Live(true);
	isosurface::surface_nets(Lookup, Grid, M->M, Live, 0.f);
	return M;
}

// Threaded export in the shape of MeshExportThread: octree (TargetSize .25)
// evaluation behind the surface-nets thunks, vertex loop over z-slabs
// [K0, K1) then face loop, both spread over Threads workers claiming work
// items from atomic counters.  Threads == 0 picks max(hw-1,2)+1.
void* tgref_export_threaded(void* O, const float* Origin, const float* Step, const uint64_t* Count,
	uint64_t K0, uint64_t K1, int Threads)
{
	SDFOctree* T = ((Octree*)O)->Tree.get();
	std::atomic<uint64_t> Evals(0);
	isosurface::AsyncParallelSurfaceNets Task;
	Task.Grid = MakeGrid(Origin, Step, Count);
	Task.ImplicitFunction = [&](float X, float Y, float Z) -> float
	{
		return T->Eval(vec3(X, Y, Z));
	};
	Task.Setup();
	if (Threads <= 0)
	{
		Threads = std::max(int(std::thread::hardware_concurrency()) - 1, 2) + 1;
	}
	std::vector<uint64_t> Slabs;
	for (uint64_t k = K0; k < K1 && k < Task.FirstLoopDomain.size(); ++k) Slabs.push_back(k);
	{
		std::atomic<uint64_t> Next(0);
		std::vector<std::thread> Workers;
		for (int t = 0; t < Threads; ++t)
		{
			Workers.emplace_back([&]()
			{
				for (;;)
				{
					uint64_t n = Next.fetch_add(1);
					if (n >= Slabs.size()) break;
					Task.FirstLoopThunk(Task, Slabs[n]);
				}
			});
		}
		for (auto& W : Workers) W.join();
	}
	std::vector<std::pair<const std::size_t, std::uint64_t>*> Items;
	for (auto& KV : Task.SecondLoopDomain) Items.push_back(&KV);
	{
		std::atomic<uint64_t> Next(0);
		std::vector<std::thread> Workers;
		for (int t = 0; t < Threads; ++t)
		{
			Workers.emplace_back([&]()
			{
				for (;;)
				{
					uint64_t n = Next.fetch_add(1);
					if (n >= Items.size()) break;
					Task.SecondLoopThunk(Task, *Items[n]);
				}
			});
		}
		for (auto& W : Workers) W.join();
	}
	Mesh* M = new Mesh;
	M->M = Task.OutputMesh;
	const uint64_t NX = Task.Grid.sx, NY = Task.Grid.sy;
	M->Evals = 8 * NX * NY * Slabs.size() + 6 * Items.size();
	return M;
}

int tgref_threads_default() { return std::max(int(std::thread::hardware_concurrency()) - 1, 2) + 1; }
int tgref_hardware_concurrency() { return int(std::thread::hardware_concurrency()); }

uint64_t tgref_mesh_vertex_count(void* M) { return ((Mesh*)M)->M.vertices_.size(); }
uint64_t tgref_mesh_face_count(void* M) { return ((Mesh*)M)->M.faces_.size(); }
uint64_t tgref_mesh_evals(void* M) { return ((Mesh*)M)->Evals; }

void tgref_mesh_copy(void* M, float* Vertices, uint64_t* Faces)
{
	Mesh* Ms = (Mesh*)M;
	for (size_t i = 0; i < Ms->M.vertices_.size(); ++i)
	{
		Vertices[3 * i] = Ms->M.vertices_[i].x;
		Vertices[3 * i + 1] = Ms->M.vertices_[i].y;
		Vertices[3 * i + 2] = Ms->M.vertices_[i].z;
	}
	for (size_t i = 0; i < Ms->M.faces_.size(); ++i)
	{
		Faces[3 * i] = Ms->M.faces_[i].v0;
		Faces[3 * i + 1] = Ms->M.faces_[i].v1;
		Faces[3 * i + 2] = Ms->M.faces_[i].v2;
	}
}

void tgref_mesh_free(void* M) { delete (Mesh*)M; }

}  // extern "C"
