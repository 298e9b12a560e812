/*
 * tangerine_gpu.h -- C ABI of the MI355X CSG signed-distance path.
 *
 * Drop-in for the evaluation path of Tangerine's SDFNode tree: trees are
 * built with the same factories and transforms the Lua API calls, flattened
 * to the reference bytecode, and evaluated on the GPU over point lists, dense
 * grids (with naive surface nets) and camera rays.
 *
 * Conventions
 *   - Every call returns TG_OK (0) or a negative TG_ERR_* code; on failure
 *     tg_last_error() returns a message for the calling thread.  Builders
 *     return NULL on failure.  The library never aborts the process.
 *   - Host buffers are caller-owned.  Opaque handles are released with the
 *     matching *_release / *_destroy call.
 *   - Trees (tg_node) are immutable and may be shared between threads; every
 *     transform returns a new tree, as the Lua bindings copy before they
 *     transform (tangerine/lua_sdf.cpp:56-170).
 *   - A tg_mesher or tg_tracer owns one HIP stream and its device buffers;
 *     calls on different handles may run concurrently, calls on one handle
 *     must be serialised by the caller.
 *   - Plain pointers and sizes only; no C++ or framework types cross the ABI.
 *
 * Reference interfaces replaced (file:line in Aeva/tangerine):
 *   builders        SDF::Sphere ... SDF::Flate       tangerine/sdf_evaluator.h:223-264,
 *                   stale C ABI Make*Brush/Make*Op   tangerine/c_sdf.cpp:118-178
 *   transforms      SDFNode::Move/Rotate/Scale       tangerine/sdf_evaluator.h:189-193,
 *                   MoveTree/RotateTree              tangerine/c_sdf.cpp:86-98
 *   compile         SDFNode::Compile/ProgramBuffer   tangerine/sdf_evaluator.h:110-163,185
 *   point eval      EvalTree / SDFNode::Eval         tangerine/c_sdf.cpp:30-34, sdf_evaluator.h:175
 *   gradient        SDFNode::Gradient                tangerine/sdf_evaluator.h:187
 *   ray march       RayMarchTree / SDFNode::RayMarch tangerine/c_sdf.cpp:56-67, sdf_evaluator.h:210
 *   grid export     MeshExportThread + par_surface_nets
 *                                                   tangerine/export.cpp:320-381,
 *                                                   third_party/naive-surface-nets/include/surface_nets.h:62-79
 */
#ifndef TANGERINE_GPU_H
#define TANGERINE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_ERR_INVALID (-1)  /* bad argument */
#define TG_ERR_DEVICE (-2)   /* HIP runtime failure or no device */
#define TG_ERR_PROGRAM (-3)  /* malformed program words */
#define TG_ERR_NOMEM (-4)    /* device allocation failed */
#define TG_ERR_LIMIT (-5)    /* size beyond a documented limit */

/* Evaluation flags. */
#define TG_FLAG_FAST 1u      /* contracted FP32 numerics; default is bit parity with SDFInterpreter::Eval */
#define TG_FLAG_NO_PRUNE 2u  /* interpret the whole program in every tile (no per-brick pruning) */
#define TG_FLAG_CERTIFY 4u   /* culled export: exact Lipschitz certificate for every skipped brick */

typedef struct tg_node tg_node;
typedef struct tg_program tg_program;
typedef struct tg_mesher tg_mesher;
typedef struct tg_tracer tg_tracer;
typedef struct tg_surrogate tg_surrogate;
typedef struct tg_batch tg_batch;

/* Pinhole camera: pixel (i, j) looks along
 *   normalize(forward + sx*tan_half_x*right + sy*tan_half_y*up),
 *   sx = (i + .5) * (2/width) - 1, sy = 1 - (j + .5) * (2/height). */
typedef struct tg_camera
{
	float eye[3];
	float forward[3];
	float right[3];
	float up[3];
	float tan_half_x, tan_half_y;
	uint32_t width, height;
} tg_camera;

typedef struct tg_trace_stats
{
	double trace_ms;           /* tracing kernel (HIP events) */
	double prune_ms;           /* cell-program pruning (HIP events) */
	uint64_t rays;
	uint64_t hits;
	uint64_t evaluations;      /* exact evaluations, march steps and normal taps */
	double flops;              /* algorithmic FP32 work executed (cost table) */
} tg_trace_stats;

/* Same fields and meaning as isosurface::regular_grid_t
 * (third_party/naive-surface-nets/include/regular_grid.h): origin, cell
 * size and cell counts; samples are (s+1) per axis at origin + i*d. */
typedef struct tg_grid
{
	float x, y, z;
	float dx, dy, dz;
	uint64_t sx, sy, sz;
} tg_grid;

typedef struct tg_program_info
{
	uint64_t instructions;     /* device records, including Stop */
	uint32_t max_stack;        /* device interpreter stack high-water mark: the reference's StackSize, or for
	                            * programs deeper than 16 slots the depth after the decoder runs each set
	                            * operator's deeper operand first (values unchanged) */
	uint32_t brushes;
	uint32_t set_ops;
	uint32_t lipschitz;        /* 1 if every primitive is a distance bound (no ellipsoid) */
	double flops_per_eval;     /* algorithmic FP32 work of one full-program evaluation */
	float leaf_extent;         /* largest world-space brush size (field scale x (|frame translation| + sum
	                            * |parameters|)); sets the rounding slack of pruning and certification */
	float reserved;
} tg_program_info;

typedef struct tg_export_stats
{
	double eval_ms;            /* grid evaluation kernels (HIP events) */
	double mesh_ms;            /* surface-nets kernels (HIP events) */
	double total_ms;           /* whole export on the device stream */
	uint64_t samples;          /* (sx+1)(sy+1)(sz+1) */
	uint64_t tiles_evaluated;  /* 8^3-sample tiles interpreted */
	uint64_t exact_evals;      /* lanes interpreted (whole 8^3 tiles, edge tiles' lanes past the grid included) */
	double flops;              /* algorithmic FP32 work (cost table) of the programs run at the grid's samples */
	uint64_t bricks_total;     /* culled export: 8^3-cube bricks of the grid (0 otherwise) */
	uint64_t bricks_flagged;   /* culled export: bricks the surrogate kept for exact evaluation */
	double prune_ms;           /* block pruning, and for a culled export brick selection (HIP events; inside eval_ms) */
	double cull_threshold;     /* culled export: bricks with |prediction at the centre| above this were skipped */
	uint64_t bricks_certified; /* culled export: bricks the exact certificate kept that the surrogate skipped */
	double block_ms;           /* the pruned block-evaluation kernel alone (HIP events; 0 if it did not run) */
	uint32_t points_per_lane;  /* its launch: points per lane and waves per workgroup */
	uint32_t waves;
	uint32_t block_max_len;    /* longest pruned block program (entries) and its binary entries */
	uint32_t block_max_binary;
	uint32_t block_lds_bytes;  /* LDS per workgroup of the block kernel */
	uint32_t block_workgroups; /* persistent workgroups of that launch (0: a workgroup per block) */
} tg_export_stats;

typedef struct tg_surrogate_stats
{
	double sample_ms;          /* coarse mesh, sample drawing and exact labels (host clock) */
	double train_ms;           /* replayed training graphs (HIP events) */
	double total_ms;           /* whole fit (host clock) */
	uint64_t samples;          /* training samples */
	uint64_t calib_samples;    /* held-out calibration samples */
	uint64_t surface_vertices; /* vertices of the coarse exact mesh the near-surface samples start from */
	uint32_t steps;            /* Adam steps */
	float final_loss;          /* mean squared error of the last batch */
	float margin;              /* max |prediction - clamped exact distance| over the calibration samples (reported) */
	float mean_residual;       /* mean of the same (the cull threshold's slack) */
} tg_surrogate_stats;

/* ---- errors / device ----------------------------------------------------- */
const char* tg_last_error(void);
int tg_device_init(int device);
int tg_device_synchronize(void);
/* Measurement utility (no reference counterpart): copies `bytes` between two
 * device allocations it makes itself with a plain streaming kernel, `repeat`
 * times; used to calibrate HBM byte counters against a known traffic. */
int tg_device_copy_probe(uint64_t bytes, int repeat, double* ms_per_copy);
/* Self-check (no reference counterpart): evaluates the exact build's square
 * root for all 2^32 float inputs against the compiler's correctly rounded
 * sqrtf and counts results that differ (NaNs compare equal).  scramble = 0
 * gives each wave 64 consecutive bit patterns, 1 a bijective mix of them. */
int tg_check_sqrt(uint32_t scramble, uint64_t* mismatches);
/* Self-check (no reference counterpart): for each divisor b, the exact
 * build's division by a record constant (RN(1/b) once, then one multiply and
 * two fused multiply-adds per point) against the correctly rounded a / b for
 * every float a that takes that path; counts differing results and the
 * values checked. */
int tg_check_quotient(const float* divisors, uint32_t n, uint64_t* mismatches, uint64_t* checked);

/* ---- tree builders (SDF::* factories; radii and half extents) -------------- */
tg_node* tg_sphere(float radius);                                        /* sdf_evaluator.cpp:1206 */
tg_node* tg_ellipsoid(float rx, float ry, float rz);                     /* :1216 */
tg_node* tg_box(float ex, float ey, float ez);                           /* :1227 */
tg_node* tg_torus(float major_radius, float minor_radius);               /* :1238 */
tg_node* tg_cylinder(float radius, float half_height);                   /* :1250 */
tg_node* tg_plane(float nx, float ny, float nz);                         /* :1260 */
tg_node* tg_cone(float radius, float height);                            /* :1296 */
tg_node* tg_coninder(float radius_low, float radius_high, float height); /* :1307 */
/* Set operators share their operands, as SDF::Union(LHS, RHS) does (:1320-1354). */
tg_node* tg_union(const tg_node* lhs, const tg_node* rhs);
tg_node* tg_inter(const tg_node* lhs, const tg_node* rhs);
tg_node* tg_diff(const tg_node* lhs, const tg_node* rhs);
tg_node* tg_blend_union(float threshold, const tg_node* lhs, const tg_node* rhs);
tg_node* tg_blend_inter(float threshold, const tg_node* lhs, const tg_node* rhs);
tg_node* tg_blend_diff(float threshold, const tg_node* lhs, const tg_node* rhs);
tg_node* tg_flate(const tg_node* node, float radius);                    /* :1356 */
/* Transforms return a transformed copy. */
tg_node* tg_move(const tg_node* node, float x, float y, float z);
tg_node* tg_rotate(const tg_node* node, float qx, float qy, float qz, float qw);
tg_node* tg_rotate_x(const tg_node* node, float degrees);                /* SDF::RotateX :1180 */
tg_node* tg_rotate_y(const tg_node* node, float degrees);
tg_node* tg_rotate_z(const tg_node* node, float degrees);
tg_node* tg_scale(const tg_node* node, float scale);
void tg_node_release(tg_node* node);                                     /* DiscardTree c_sdf.cpp:70 */
int tg_node_bounds(const tg_node* node, float min_max[6]);               /* SDFNode::Bounds */
int tg_node_has_finite_bounds(const tg_node* node);                      /* TreeHasFiniteBounds c_sdf.cpp:77 */
int tg_node_leaf_count(const tg_node* node);                             /* SDFNode::LeafCount */
int tg_node_stack_size(const tg_node* node);                             /* SDFNode::StackSize */
/* Reference ProgramBuffer words terminated by Stop (SDFInterpreter ctor,
 * sdf_evaluator.cpp:1376-1382).  Writes at most `capacity` words and sets
 * *count to the full length. */
int tg_node_compile(const tg_node* node, uint32_t* words, uint64_t capacity, uint64_t* count);

/* ---- device programs ------------------------------------------------------ */
int tg_program_from_node(const tg_node* node, tg_program** out);
/* Any stack depth (the reference sizes its stack from StackSize); TG_ERR_LIMIT
 * only for programs that need more than 16 slots in every operand order (a
 * balanced tree of 65,536 brushes or more). */
int tg_program_from_words(const uint32_t* words, uint64_t count, tg_program** out);
int tg_program_info_get(const tg_program* program, tg_program_info* out);
/* Host only: validates words as tg_program_from_words does and reports the
 * program's properties without touching a device. */
int tg_words_inspect(const uint32_t* words, uint64_t count, tg_program_info* out);
void tg_program_destroy(tg_program* program);

/* ---- point queries (EvalTree / SDFNode::Eval, batched) --------------------- */
/* xyz: 3 floats per point; host buffers. */
int tg_eval_points(const tg_program* program, const float* xyz, uint64_t count, float* out, uint32_t flags);
/* Same with device buffers on `stream` (a hipStream_t, or NULL for the null stream). */
int tg_eval_points_device(const tg_program* program, const float* xyz, uint64_t count, float* out,
                          uint32_t flags, void* stream);

/* Gradients at points.  mode 0: SDFNode::Gradient's tetrahedral rule with its
 * fallback (sdf_evaluator.cpp:298-333; pass h = 1e-4 for the reference step);
 * mode 1: axis central differences with step h.  out: 3 floats per point. */
int tg_gradient_points(const tg_program* program, const float* xyz, uint64_t count, float h, int mode,
                       float* out, uint32_t flags);

/* ---- sphere tracing (SDFNode::RayMarch, batched) ------------------------- */
/* Rays from host arrays (directions are normalized as glm::normalize does);
 * far = INFINITY reproduces RayMarch exactly.  hit: 1 byte per ray; travel is
 * INFINITY on a miss; position: 3 floats per ray (last position on a miss). */
int tg_raymarch(const tg_program* program, const float* starts, const float* dirs, uint64_t count,
                int max_steps, float epsilon, float far_distance, uint32_t flags,
                uint8_t* hit, float* travel, float* position);
/* tg_raymarch without any device allocation (no reference counterpart; for
 * callers that cannot allocate: tight per-frame query loops such as the Lua
 * ray_cast / magnet bindings, serving loops).  Same arguments and results as
 * tg_raymarch; rays are staged through a fixed device array of the library
 * (TG_RAYMARCH_STATIC_MAX rays, 41 bytes each, held by every device the
 * library runs on until the process ends).  count > TG_RAYMARCH_STATIC_MAX
 * returns TG_ERR_LIMIT.  Every call on one device shares that array: calls
 * that overlap on one device (from several threads) race and must be
 * serialised by the caller. */
#define TG_RAYMARCH_STATIC_MAX 65536u
int tg_raymarch_static(const tg_program* program, const float* starts, const float* dirs, uint64_t count,
                       int max_steps, float epsilon, float far_distance, uint32_t flags,
                       uint8_t* hit, float* travel, float* position);
/* ---- Lua queries (tangerine/lua_sdf.cpp:386-559, sdf_model.cpp:150-264) --- */
/* pivot_towards (LuaPivotTowards, lua_sdf.cpp:452-559) for `count` queries:
 * pivots, headings, downs: 3 floats each; the scalar arguments as in Lua
 * (max_angle in degrees).  Per query `status`: 0 done (tail and heading
 * written), 2 heading antiparallel to down (the reference then draws random
 * pivot axes from the Lua environment's generator: the caller does that),
 * 3 the loop ran `max_iterations` turns (the reference would keep turning).
 * Distances come from the program (interpreter), as eval does. */
int tg_pivot_towards(const tg_program* program, uint64_t count, const float* pivots, const float* headings,
                     const float* downs, float max_dist, float margin, float max_angle_degrees, int max_iterations,
                     float* tails, float* headings_out, int32_t* status, uint32_t flags);
/* Mouse picking (DeliverMouseButton with SDFModel::RayMarch, sdf_model.cpp:
 * 150-206, 259-264) for `ray_count` rays against `model_count` models: each
 * ray moved into every model's frame (transforms: 8 floats per model,
 * rotation quaternion x, y, z, w, translation, scale; Transform::ApplyInv),
 * marched 1000 steps with epsilon 0.001; the smallest local travel wins, the
 * earlier model on ties.  model_out: index or -1; travel_out: INFINITY on a
 * miss; position_out (3 floats per ray, may be NULL): the hit in the model's
 * local frame, as the reference reports it. */
int tg_pick(const tg_program* const* programs, const float* transforms, uint32_t model_count, const float* starts,
            const float* dirs, uint64_t ray_count, uint32_t flags, int32_t* model_out, float* travel_out,
            float* position_out);

/* Camera looking from eye at target; up is replaced when parallel to the view. */
int tg_camera_look_at(tg_camera* out, const float eye[3], const float target[3], const float up[3],
                      float fov_y_degrees, uint32_t width, uint32_t height);
int tg_tracer_create(tg_tracer** out);
void tg_tracer_destroy(tg_tracer* tracer);
/* Cube [lo, lo + size]^3 over which per-cell programs are built; positions
 * outside it run the whole program.  Default: the origin-centred cube holding
 * the eye. */
int tg_tracer_set_domain(tg_tracer* tracer, const float lo[3], float size);
/* Marching passes of a render: the first marches every pixel's ray at most
 * first_budget steps, 64 rays to a wave; each later pass the rays still
 * marching (sorted by the cell they are in) at most `budget` more, packed so
 * that about resume_waves waves run them, until max_steps.  first_budget <= 0:
 * one pass.  Results do not depend on these.  Default 48, 208, 65536. */
int tg_tracer_set_passes(tg_tracer* tracer, int first_budget, int budget, uint32_t resume_waves);
/* Renders width*height rays on the device; hit normals by central differences
 * with step normal_h (<= 0: none).  Results stay on the device until copied. */
int tg_tracer_render(tg_tracer* tracer, const tg_program* program, const tg_camera* camera, int max_steps,
                     float epsilon, float far_distance, float normal_h, uint32_t flags);
/* t: 1 float per pixel (INFINITY = miss); normals: 3 floats per pixel or NULL;
 * steps: distance evaluations along each pixel's ray (the final hit test
 * included), 1 uint32 per pixel, or NULL.  Row-major, row 0 at the top. */
int tg_tracer_copy(const tg_tracer* tracer, float* t, float* normals, uint32_t* steps);
int tg_tracer_stats(const tg_tracer* tracer, tg_trace_stats* out);

/* ---- grid export (MeshExportThread shape: evaluate, then naive surface nets) */
int tg_mesher_create(tg_mesher** out);
void tg_mesher_destroy(tg_mesher* mesher);
int tg_mesher_export(tg_mesher* mesher, const tg_program* program, const tg_grid* grid, uint32_t flags);
/* Mesh caller-supplied samples ((sx+1)(sy+1)(sz+1) floats, x fastest). */
int tg_mesher_mesh_values(tg_mesher* mesher, const float* values, const tg_grid* grid);
int tg_mesher_counts(const tg_mesher* mesher, uint64_t* vertices, uint64_t* triangles);
/* vertices: 3 floats each; triangles: 3 uint32 vertex ids each. */
int tg_mesher_copy_mesh(const tg_mesher* mesher, float* vertices, uint32_t* triangles);
/* Samples of the last export, x fastest. */
int tg_mesher_copy_grid(const tg_mesher* mesher, float* values);
int tg_mesher_stats(const tg_mesher* mesher, tg_export_stats* out);
/* Binary STL of the last export's mesh (WriteSTL, tangerine/export.cpp:111-179):
 * 80-byte header "STL generated by TangerineCSG", uint32 triangle count, per
 * triangle the normalised tetrahedral gradient (SDFNode::Gradient, h = 1e-4)
 * of `program` at the triangle's centroid, the three vertices times `scale`
 * and a zero uint16. */
int tg_mesher_write_stl(tg_mesher* mesher, const tg_program* program, const char* path, float scale,
                        uint32_t flags);
/* Binary little-endian PLY (PlyHeader / WritePLY, export.cpp:181-317): per
 * vertex x y z times `scale` and the gradient at the unscaled vertex, per face
 * uchar 3 and three 32-bit vertex ids.  No colour (no paint on this path). */
int tg_mesher_write_ply(tg_mesher* mesher, const tg_program* program, const char* path, float scale,
                        uint32_t flags);

/* Grid of the last export (origin, step, cells). */
int tg_mesher_grid_get(const tg_mesher* mesher, tg_grid* out);

/* ---- point-cloud export (PointCloudExportThread, export.cpp:384-480) ---------
 * Cells of the grid from `lo` with `step`, ceil((hi - lo) / step) per axis:
 * a cell's centre (float(i) * step + lo) + step / 2 is kept when the program's
 * |distance| there is below the half step's length; kept centres (raster
 * order; the threaded reference appends them in thread order) then take
 * `refine_iterations` steps p -= d(p) * gradient(p) (tetrahedral gradient,
 * h = 1e-4), are clamped to their cell and replace the centre when within the
 * half step's length of it.  The points become the mesher's vertices (no
 * triangles): tg_mesher_counts, tg_mesher_copy_mesh, and tg_mesher_write_ply
 * writes the reference's point-cloud PLY (normals = gradients, no faces).
 * Whole program, exact numerics unless TG_FLAG_FAST; at most 2^31 - 1 cells. */
int tg_mesher_export_points(tg_mesher* mesher, const tg_program* program, const float lo[3], const float hi[3],
                            const float step[3], int refine_iterations, uint32_t flags);

/* ---- MagicaVoxel export (VoxExport, magica.cpp:27-70) -------------------------
 * A grid of ivec3(ceil(max - min) * grid_size) voxels over `bounds` (min xyz,
 * max xyz: the model's SDFNode::Bounds); voxel (x, y, z) is kept when the
 * program's |distance| at mix(min, max, (x + .5, y + .5, z + .5) / size) is at
 * most the distance from min to mix(min, max, 0.5 / size).  When `path` is
 * not NULL the voxels are written as the .vox file the reference's
 * vox::VoxWriter writes for them added in raster order (colour index
 * |color_index| % 255 + 1, no palette); the reference adds them in thread
 * order.  voxel_count (may be NULL) receives the count;
 * tg_mesher_copy_vox_cells copies the kept cell indices x + size.x (y +
 * size.y z), raster order.  At most 2^31 - 1 cells. */
int tg_mesher_export_vox(tg_mesher* mesher, const tg_program* program, const float bounds[6], float grid_size,
                         int color_index, const char* path, uint32_t flags, uint64_t* voxel_count);
int tg_mesher_copy_vox_cells(const tg_mesher* mesher, uint32_t* cells);
/* The same .vox writer for a caller's cells (indices x + size.x (y + size.y
 * z), in the order vox::VoxWriter::AddVoxel would receive them; host only). */
int tg_vox_write(const char* path, const uint32_t size[3], const uint32_t* cells, uint64_t count, int color_index);

/* ---- Sodapop meshing job (MeshingJob::NaiveSurfaceNets, sodapop.cpp:562-897) --
 * The interactive mesher's grid (NaiveSurfaceNetsScratch, sodapop.cpp:156-177):
 * the bounding cube of `bounds` (min xyz, max xyz: the model's SDFNode::Bounds),
 * floor(density) samples per unit (at least 8 per axis; the reference passes
 * 20 + the model's density push), origin moved down two steps, three more
 * cells per axis; the samples are exact (the reference's inexact octree values
 * differ only away from the surface, where they are clamped to +-100 and do not
 * change the mesh).  With `shuffle`, vertices are reordered as
 * ApplyVertexSequence does for VertexSequence::Shuffle (sodapop.cpp:337-394:
 * std::mt19937 seeded 0, std::shuffle; libstdc++'s shuffle) and the triangle
 * indices remapped.  Normals: the tetrahedral gradient at every final vertex
 * (USE_GRADIENT_NORMALS, sodapop.cpp:40, 805-817), read with
 * tg_mesher_copy_normals.  Vertices before the shuffle follow raster order
 * (the reference's come in task order, which varies between runs). */
int tg_mesher_export_sodapop(tg_mesher* mesher, const tg_program* program, const float bounds[6], float density,
                             int shuffle, uint32_t flags);
int tg_mesher_copy_normals(const tg_mesher* mesher, float* normals);

/* ---- block-evaluator tuning (no reference counterpart) ----------------------
 * The pruned block evaluation has three launch parameters: points per lane
 * (2, 4 or 8 samples interpreted together by one lane: more work per fetched
 * instruction, more registers and LDS), waves per workgroup (1, 2, 4) and
 * workgroups per CU: 0 launches a workgroup per 32^3-sample block with its
 * decision tables and tile programs in LDS; 1..16 launches that many
 * persistent workgroups per CU, which take blocks from a queue and keep the
 * tables and tile programs in global scratch (LDS: the spill region only).
 * One work parameter: the shortest block program still pruned per tile.
 * tg_mesher_autotune prunes `program` for `grid`, times every launch on the
 * device, then the threshold at the fastest launch, and keeps the best for
 * later exports on this mesher.  Every setting writes the same samples. */
typedef struct tg_tune_result
{
	uint32_t points_per_lane;     /* chosen */
	uint32_t waves;
	uint32_t workgroups_per_cu;   /* chosen: 0 = a workgroup per block */
	uint32_t min_prune_len;       /* chosen: shorter block programs skip per-tile pruning */
	double eval_ms;               /* block kernel time of the chosen configuration */
	uint32_t candidates;
	uint32_t candidate_points[24]; /* configuration that actually ran (LDS may force fewer points per lane) */
	uint32_t candidate_waves[24];
	uint32_t candidate_workgroups_per_cu[24];
	double candidate_ms[24];
	double candidate_len_ms[3];   /* min_prune_len 0, 16, 48 at the chosen launch */
} tg_tune_result;
int tg_mesher_autotune(tg_mesher* mesher, const tg_program* program, const tg_grid* grid, uint32_t flags,
                       tg_tune_result* out);
int tg_mesher_set_tuning(tg_mesher* mesher, uint32_t points_per_lane, uint32_t waves, uint32_t workgroups_per_cu,
                         uint32_t min_prune_len);
/* Launch shape of culled exports (tg_mesher_export_culled), kept apart from
 * the plain export's: their near-surface tiles run deeper programs, which
 * want more spill room per wave (default 4 points per lane, 4 waves, 6
 * persistent workgroups per CU).  Every setting writes the same samples. */
int tg_mesher_set_culled_tuning(tg_mesher* mesher, uint32_t points_per_lane, uint32_t waves,
                                uint32_t workgroups_per_cu);
/* How the mesher's block pruning emits each block's program once its
 * decision and liveness passes ran: 0 automatic (the default: a wave per
 * block, with a lane per program entry, on grids of at most 2048 blocks, or
 * 4096 for programs of 1024 entries or more; otherwise the walk), 1 one wave
 * walking the entries of 16 blocks at a time, 2 a wave per block.  Every
 * setting emits the same programs, so exports give the same samples and
 * mesh. */
int tg_mesher_set_emission(tg_mesher* mesher, uint32_t mode);

/* ---- batched export (no reference counterpart; closest: one Sodapop
 * MeshingJob per model on Scheduler workers, sodapop.cpp:214-224,
 * scheduler.cpp:307-330) ------------------------------------------------------
 * Exports many trees, each on its own grid: every tree is pruned per block,
 * then all trees' blocks are evaluated by one persistent launch that takes
 * work items (one block of one tree) from a device queue, heaviest trees
 * first; the launch is captured in a hipGraph.  Each tree's samples and mesh
 * stay in its own mesher (tg_batch_mesher), so tg_mesher_counts /
 * tg_mesher_copy_mesh / tg_mesher_copy_grid read them.  Programs must be
 * 1-Lipschitz and stay alive while the batch is replayed. */
typedef struct tg_batch_stats
{
	double prune_ms;           /* layout and per-tree block pruning (host clock) */
	double eval_ms;            /* the persistent launch (graph replay, HIP events) */
	double mesh_ms;            /* surface nets of every tree (host clock) */
	double total_ms;
	uint64_t trees;
	uint64_t blocks;           /* queue items */
	uint64_t samples;
	uint64_t exact_evals;
	double flops;
	uint32_t points_per_lane, waves, workgroups;
	uint32_t lds_bytes;        /* LDS per workgroup of the short-program launch (the long one if none) */
	uint32_t max_len, max_binary, max_depth;  /* its layout: the largest over its queue */
} tg_batch_stats;
typedef struct tg_batch_tune_result
{
	uint32_t points_per_lane, waves, workgroups_per_cu;  /* chosen */
	double eval_ms;
	uint32_t candidates;
	uint32_t candidate_points[18], candidate_waves[18], candidate_workgroups_per_cu[18];
	double candidate_ms[18];
} tg_batch_tune_result;
int tg_batch_create(tg_batch** out);
void tg_batch_destroy(tg_batch* batch);
int tg_batch_export(tg_batch* batch, const tg_program* const* programs, const tg_grid* grids, uint32_t count,
                    uint32_t flags);
/* Replays the captured evaluation launch of the last export (same trees). */
int tg_batch_replay(tg_batch* batch, double* eval_ms);
/* Times points per lane {2,4,8} x waves {2,4} x persistent workgroups per CU
 * {1,2,4} for the short-program queue (blocks of at most 256 entries) on the
 * last exported batch, keeps the fastest, then picks the long-program queue's
 * shape among seven; eval_ms is the best whole-batch time seen. */
int tg_batch_autotune(tg_batch* batch, tg_batch_tune_result* out);
int tg_batch_set_tuning(tg_batch* batch, uint32_t points_per_lane, uint32_t waves, uint32_t workgroups_per_cu);
/* Tree `index`'s mesher, owned by the batch: read its results (counts, mesh,
 * grid, stats, bricks, STL/PLY writers); exports, tuning and destroy on it
 * return TG_ERR_INVALID (its buffers are captured in the batch's graph). */
tg_mesher* tg_batch_mesher(tg_batch* batch, uint32_t index);
int tg_batch_stats_get(const tg_batch* batch, tg_batch_stats* out);

/* ---- learned brick culling (no reference counterpart) ----------------------
 * A Fourier-feature MLP (3 -> 96 features, sigma-scaled Gaussian frequencies
 * -> 64 SiLU -> 64 SiLU -> 1; BF16 matrix-core operands, FP32 accumulation)
 * fitted to clamp(d, -clamp, clamp) over the cube [lo, lo + size]^3.  Its only
 * use is to choose which 8^3-cube bricks a culled export evaluates exactly.
 * A brick of half-diagonal r whose centre c holds surface has |d(c)| <= r,
 * so it is skipped only when |prediction(c)| exceeds the largest |prediction|
 * seen at held-out calibration points with |exact d| <= r, plus the mean
 * calibration residual (or, after tg_surrogate_set_margin, when it exceeds
 * r + margin).  With TG_FLAG_CERTIFY every skipped brick is checked exactly,
 * so the bar is the 90th percentile of those |prediction|s instead of the
 * largest (the certificate puts back what that misses).  Every sample and
 * vertex of the output still comes
 * from the exact evaluator; a brick holding surface that the surrogate
 * wrongly skipped would be missing from the mesh, which
 * tg_mesher_active_bricks lets a caller audit against a plain export. */
int tg_surrogate_create(tg_surrogate** out, const float lo[3], float size, float sigma, float clamp, uint64_t seed);
void tg_surrogate_destroy(tg_surrogate* surrogate);
/* Read (write = 0) or replace (write = 1) the FP32 parameters (10433 floats:
 * W1[64][96], b1[64], W2[64][64], b2[64], w3[64], b3) and the frequencies
 * (48 x 3 floats); either pointer may be NULL. */
int tg_surrogate_params(tg_surrogate* surrogate, float* params, float* fourier, int write);
int tg_surrogate_predict(const tg_surrogate* surrogate, const float* xyz, uint64_t count, float* out);
/* Gradient of the mean squared error against clamp(target) over count
 * points (count a positive multiple of 32) and the loss itself. */
int tg_surrogate_gradients(tg_surrogate* surrogate, const float* xyz, const float* target, uint64_t count,
                           float* grad, float* loss);
/* Draw `samples` training and `calib_samples` calibration points on the
 * device (half uniform, half near the surface of a 128^3 exact export of the
 * domain with N(0, surface_sigma^2) offsets), label them with `exact`, run
 * `epochs` passes of Adam (batches of 16384) and calibrate the margin. */
int tg_surrogate_fit(tg_surrogate* surrogate, const tg_program* exact, uint64_t samples, uint64_t calib_samples,
                     float surface_sigma, uint32_t epochs, uint32_t flags, tg_surrogate_stats* stats);
/* Replace the calibrated threshold by r + margin (margin >= 0). */
int tg_surrogate_set_margin(tg_surrogate* surrogate, float margin);
int tg_surrogate_get_stats(const tg_surrogate* surrogate, tg_surrogate_stats* out);
/* Mean training loss per Adam step of the last fit. */
int tg_surrogate_loss_log(const tg_surrogate* surrogate, float* out, uint64_t capacity, uint64_t* count);
/* Export that evaluates exactly only the tiles of kept bricks.  Needs a
 * 1-Lipschitz program.  With a calibrated surrogate, bricks are kept by its
 * prediction; with TG_FLAG_CERTIFY (or surrogate = NULL) the block's pruned
 * program is also evaluated exactly at every skipped brick's centre c, and a
 * brick with |d(c)| <= its half-diagonal (+ rounding slack) is kept, so no
 * brick holding a sign change can be skipped.  The samples of skipped tiles
 * are not computed, so tg_mesher_copy_grid refuses afterwards. */
int tg_mesher_export_culled(tg_mesher* mesher, const tg_program* program, const tg_grid* grid,
                            const tg_surrogate* surrogate, uint32_t flags);
/* One byte per 8^3-cube brick (x fastest): the kept bricks of the last culled
 * export, or the bricks holding an active cube of the last export. */
int tg_mesher_brick_flags(const tg_mesher* mesher, uint8_t* out, uint64_t capacity, uint64_t* count);
int tg_mesher_active_bricks(tg_mesher* mesher, uint8_t* out, uint64_t capacity, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* TANGERINE_GPU_H */
