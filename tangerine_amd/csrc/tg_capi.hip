#include <atomic>
// C ABI of the MI355X CSG distance path (declared in include/tangerine_gpu.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"
#include "tg_mesh.h"
#include "tg_program.h"
#include "tg_tree.h"

using namespace tg;

// ---------------------------------------------------------------------------
// errors

namespace tg {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
	g_last_error = msg;
	return code;
}

int hip_fail(hipError_t e, const char* what)
{
	return fail(TG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int DevBuf::ensure(size_t need)
{
	if (need <= bytes) return TG_OK;
	if (ptr) hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
	size_t grow = std::max(need, size_t(256));
	hipError_t e = hipMalloc(&ptr, grow);
	if (e != hipSuccess)
	{
		ptr = nullptr;
		hipGetLastError();
		return fail(TG_ERR_NOMEM, "hipMalloc of " + std::to_string(grow) + " bytes failed: " + hipGetErrorString(e));
	}
	bytes = grow;
	return TG_OK;
}

void DevBuf::release()
{
	if (ptr) hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}

int ensure_device()
{
	static int ready = -1;
	if (ready == 1) return TG_OK;
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n == 0)
	{
		hipGetLastError();
		return fail(TG_ERR_DEVICE, "no HIP device available");
	}
	ready = 1;
	return TG_OK;
}

}  // namespace tg

struct tg_node
{
	NodeRef ref;
};

#define TG_CHECK_HIP(expr, what)                 \
	do                                           \
	{                                            \
		hipError_t _e = (expr);                  \
		if (_e != hipSuccess) return hip_fail(_e, what); \
	} while (0)

extern "C" const char* tg_last_error(void) { return g_last_error.c_str(); }

extern "C" int tg_device_init(int device)
{
	int rc = ensure_device();
	if (rc) return rc;
	TG_CHECK_HIP(hipSetDevice(device), "hipSetDevice");
	return TG_OK;
}

extern "C" int tg_device_synchronize(void)
{
	TG_CHECK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
	return TG_OK;
}

// Plain streaming copy, 16 bytes per lane per iteration.
__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n)
{
	const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int tg_device_copy_probe(uint64_t bytes, int repeat, double* ms_per_copy)
{
	int rc = ensure_device();
	if (rc) return rc;
	if (bytes < 16 || repeat < 1) return fail(TG_ERR_INVALID, "tg_device_copy_probe: bytes >= 16, repeat >= 1");
	const uint64_t n = bytes / 16;
	DevBuf a, b;
	rc = a.ensure(n * 16);
	if (!rc) rc = b.ensure(n * 16);
	if (rc)
	{
		a.release();
		return rc;
	}
	hipEvent_t e0, e1;
	hipEventCreate(&e0);
	hipEventCreate(&e1);
	hipMemset(a.ptr, 1, n * 16);
	const unsigned blocks = unsigned(std::min<uint64_t>((n + 255) / 256, 256u * 8u * 8u));
	hipEventRecord(e0, nullptr);
	for (int r = 0; r < repeat; ++r)
		hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a.as<const uint4>(), b.as<uint4>(), n);
	hipEventRecord(e1, nullptr);
	hipError_t e = hipEventSynchronize(e1);
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	a.release();
	b.release();
	if (e != hipSuccess) return hip_fail(e, "copy probe");
	if (ms_per_copy) *ms_per_copy = double(ms) / repeat;
	return TG_OK;
}

// ---------------------------------------------------------------------------
// builders

static tg_node* wrap(NodeRef r) { return r ? new tg_node{ std::move(r) } : nullptr; }

static bool finite_all(std::initializer_list<float> v)
{
	for (float f : v)
		if (!std::isfinite(f)) return false;
	return true;
}

static tg_node* bad_arg(const char* what)
{
	g_last_error = what;
	return nullptr;
}

extern "C" tg_node* tg_sphere(float r)
{
	if (!finite_all({ r })) return bad_arg("tg_sphere: non-finite radius");
	return wrap(make_sphere(r));
}
extern "C" tg_node* tg_ellipsoid(float x, float y, float z)
{
	if (!finite_all({ x, y, z })) return bad_arg("tg_ellipsoid: non-finite radius");
	return wrap(make_ellipsoid(x, y, z));
}
extern "C" tg_node* tg_box(float x, float y, float z)
{
	if (!finite_all({ x, y, z })) return bad_arg("tg_box: non-finite extent");
	return wrap(make_box(x, y, z));
}
extern "C" tg_node* tg_torus(float major, float minor)
{
	if (!finite_all({ major, minor })) return bad_arg("tg_torus: non-finite radius");
	return wrap(make_torus(major, minor));
}
extern "C" tg_node* tg_cylinder(float r, float h)
{
	if (!finite_all({ r, h })) return bad_arg("tg_cylinder: non-finite size");
	return wrap(make_cylinder(r, h));
}
extern "C" tg_node* tg_plane(float x, float y, float z)
{
	if (!finite_all({ x, y, z }) || (x == 0.0f && y == 0.0f && z == 0.0f)) return bad_arg("tg_plane: degenerate normal");
	return wrap(make_plane(x, y, z));
}
extern "C" tg_node* tg_cone(float r, float h)
{
	if (!finite_all({ r, h }) || h == 0.0f) return bad_arg("tg_cone: degenerate size");
	return wrap(make_cone(r, h));
}
extern "C" tg_node* tg_coninder(float lo, float hi, float h)
{
	if (!finite_all({ lo, hi, h })) return bad_arg("tg_coninder: non-finite size");
	return wrap(make_coninder(lo, hi, h));
}

static tg_node* set_op(uint32_t op, float t, const tg_node* l, const tg_node* r)
{
	if (!l || !r) return bad_arg("set operator: null operand");
	if (!std::isfinite(t)) return bad_arg("set operator: non-finite threshold");
	return wrap(make_set(op, l->ref, r->ref, t));
}

extern "C" tg_node* tg_union(const tg_node* l, const tg_node* r) { return set_op(OP_UNION, 0.0f, l, r); }
extern "C" tg_node* tg_inter(const tg_node* l, const tg_node* r) { return set_op(OP_INTER, 0.0f, l, r); }
extern "C" tg_node* tg_diff(const tg_node* l, const tg_node* r) { return set_op(OP_DIFF, 0.0f, l, r); }
extern "C" tg_node* tg_blend_union(float t, const tg_node* l, const tg_node* r) { return set_op(OP_BLEND_UNION, t, l, r); }
extern "C" tg_node* tg_blend_inter(float t, const tg_node* l, const tg_node* r) { return set_op(OP_BLEND_INTER, t, l, r); }
extern "C" tg_node* tg_blend_diff(float t, const tg_node* l, const tg_node* r) { return set_op(OP_BLEND_DIFF, t, l, r); }

extern "C" tg_node* tg_flate(const tg_node* n, float radius)
{
	if (!n) return bad_arg("tg_flate: null node");
	return wrap(make_flate(n->ref, radius));
}

extern "C" tg_node* tg_move(const tg_node* n, float x, float y, float z)
{
	if (!n) return bad_arg("tg_move: null node");
	NodeRef c = copy_tree(n->ref);
	move_tree(c.get(), { x, y, z });
	return wrap(c);
}

extern "C" tg_node* tg_rotate(const tg_node* n, float qx, float qy, float qz, float qw)
{
	if (!n) return bad_arg("tg_rotate: null node");
	NodeRef c = copy_tree(n->ref);
	rotate_tree(c.get(), { qw, qx, qy, qz });
	return wrap(c);
}

static tg_node* rotate_axis(const tg_node* n, int axis, float degrees)
{
	if (!n) return bad_arg("tg_rotate_*: null node");
	NodeRef c = copy_tree(n->ref);
	rotate_tree(c.get(), axis_rotation(axis, degrees));
	return wrap(c);
}

extern "C" tg_node* tg_rotate_x(const tg_node* n, float d) { return rotate_axis(n, 0, d); }
extern "C" tg_node* tg_rotate_y(const tg_node* n, float d) { return rotate_axis(n, 1, d); }
extern "C" tg_node* tg_rotate_z(const tg_node* n, float d) { return rotate_axis(n, 2, d); }

extern "C" tg_node* tg_scale(const tg_node* n, float s)
{
	if (!n) return bad_arg("tg_scale: null node");
	NodeRef c = copy_tree(n->ref);
	scale_tree(c.get(), s);
	return wrap(c);
}

extern "C" void tg_node_release(tg_node* n) { delete n; }

extern "C" int tg_node_bounds(const tg_node* n, float* mm)
{
	if (!n || !mm) return fail(TG_ERR_INVALID, "tg_node_bounds: null argument");
	Box3 b = tree_bounds(n->ref.get());
	mm[0] = b.lo.x; mm[1] = b.lo.y; mm[2] = b.lo.z;
	mm[3] = b.hi.x; mm[4] = b.hi.y; mm[5] = b.hi.z;
	return TG_OK;
}

extern "C" int tg_node_has_finite_bounds(const tg_node* n) { return n ? int(tree_has_finite_bounds(n->ref.get())) : 0; }
extern "C" int tg_node_leaf_count(const tg_node* n) { return n ? int(n->ref->leaves) : 0; }
extern "C" int tg_node_stack_size(const tg_node* n) { return n ? int(n->ref->stack_size) : 0; }

extern "C" int tg_node_compile(const tg_node* n, uint32_t* words, uint64_t cap, uint64_t* count)
{
	if (!n || !count) return fail(TG_ERR_INVALID, "tg_node_compile: null argument");
	std::vector<uint32_t> w;
	compile_tree(n->ref.get(), w);
	*count = w.size();
	if (words)
	{
		const uint64_t k = std::min<uint64_t>(cap, w.size());
		std::memcpy(words, w.data(), k * 4);
	}
	return TG_OK;
}

// ---------------------------------------------------------------------------
// programs

int tg::upload_program(const DecodedProgram& d, tg_program** out)
{
	int rc = ensure_device();
	if (rc) return rc;
	static std::atomic<uint64_t> next_uid{ 1 };
	auto p = std::make_unique<tg_program>();
	p->uid = next_uid++;
	p->decoded = d;
	rc = p->device.ensure(d.instrs.size() * sizeof(Instr));
	if (!rc) rc = p->entries.ensure(std::max<size_t>(d.entries.size(), 1) * sizeof(HostEntry));
	if (rc)
	{
		p->device.release();
		return rc;
	}
	TG_CHECK_HIP(hipMemcpy(p->device.ptr, d.instrs.data(), d.instrs.size() * sizeof(Instr), hipMemcpyHostToDevice),
		"program upload");
	TG_CHECK_HIP(hipMemcpy(p->entries.ptr, d.entries.data(), d.entries.size() * sizeof(HostEntry), hipMemcpyHostToDevice),
		"program entries upload");
	*out = p.release();
	return TG_OK;
}

extern "C" int tg_program_from_words(const uint32_t* words, uint64_t count, tg_program** out)
{
	if (!words || !out || count == 0) return fail(TG_ERR_INVALID, "tg_program_from_words: null or empty argument");
	DecodedProgram d;
	std::string err;
	if (!decode_words(words, count, d, err))
	{
		return fail(err.find("stack slots") != std::string::npos ? TG_ERR_LIMIT : TG_ERR_PROGRAM,
			"tg_program_from_words: " + err);
	}
	return upload_program(d, out);
}

extern "C" int tg_program_from_node(const tg_node* n, tg_program** out)
{
	if (!n || !out) return fail(TG_ERR_INVALID, "tg_program_from_node: null argument");
	std::vector<uint32_t> w;
	compile_tree(n->ref.get(), w);
	return tg_program_from_words(w.data(), w.size(), out);
}

static void fill_info(const DecodedProgram& d, tg_program_info* out)
{
	out->instructions = d.instrs.size();
	out->max_stack = d.max_stack;
	out->brushes = d.brushes;
	out->set_ops = d.set_ops;
	out->lipschitz = d.lipschitz ? 1u : 0u;
	out->flops_per_eval = d.flops;
	out->leaf_extent = d.leaf_extent;
	out->reserved = 0.0f;
}

extern "C" int tg_program_info_get(const tg_program* p, tg_program_info* out)
{
	if (!p || !out) return fail(TG_ERR_INVALID, "tg_program_info_get: null argument");
	fill_info(p->decoded, out);
	return TG_OK;
}

extern "C" int tg_words_inspect(const uint32_t* words, uint64_t count, tg_program_info* out)
{
	if (!words || !out || count == 0) return fail(TG_ERR_INVALID, "tg_words_inspect: null or empty argument");
	DecodedProgram d;
	std::string err;
	if (!decode_words(words, count, d, err))
		return fail(err.find("stack slots") != std::string::npos ? TG_ERR_LIMIT : TG_ERR_PROGRAM, "tg_words_inspect: " + err);
	fill_info(d, out);
	return TG_OK;
}

extern "C" void tg_program_destroy(tg_program* p)
{
	if (!p) return;
	p->device.release();
	p->entries.release();
	delete p;
}

// ---------------------------------------------------------------------------
// point queries

extern "C" int tg_eval_points_device(const tg_program* p, const float* xyz, uint64_t n, float* out, uint32_t flags,
	void* stream)
{
	if (!p || (n && (!xyz || !out))) return fail(TG_ERR_INVALID, "tg_eval_points_device: null argument");
	if (n == 0) return TG_OK;
	PointArgs a{};
	a.prog = static_cast<const Instr*>(p->device.ptr);
	a.xyz = xyz;
	a.out = out;
	a.count = n;
	a.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	hipStream_t s = static_cast<hipStream_t>(stream);
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_points(a, s) : exact::launch_points(a, s);
	if (e != hipSuccess) return hip_fail(e, "eval points launch");
	return TG_OK;
}

extern "C" int tg_check_sqrt(uint32_t scramble, uint64_t* mismatches)
{
	if (!mismatches) return fail(TG_ERR_INVALID, "tg_check_sqrt: null argument");
	DevBuf d;
	int rc = d.ensure(8);
	if (rc) return rc;
	struct Guard { DevBuf& a; ~Guard() { a.release(); } } g{ d };
	TG_CHECK_HIP(hipMemset(d.ptr, 0, 8), "reset");
	TG_CHECK_HIP(exact::launch_sqrt_check(scramble, d.as<unsigned long long>(), nullptr), "sqrt check");
	unsigned long long n = 0;
	TG_CHECK_HIP(hipMemcpy(&n, d.ptr, 8, hipMemcpyDeviceToHost), "sqrt check");
	*mismatches = n;
	return TG_OK;
}

extern "C" int tg_check_quotient(const float* divisors, uint32_t n, uint64_t* mismatches, uint64_t* checked)
{
	if (!mismatches || !checked || (n && !divisors)) return fail(TG_ERR_INVALID, "tg_check_quotient: null argument");
	DevBuf d;
	int rc = d.ensure(16);
	if (rc) return rc;
	struct Guard { DevBuf& a; ~Guard() { a.release(); } } g{ d };
	TG_CHECK_HIP(hipMemset(d.ptr, 0, 16), "reset");
	for (uint32_t i = 0; i < n; ++i)
		TG_CHECK_HIP(exact::launch_quotient_check(divisors[i], d.as<unsigned long long>(), nullptr), "quotient check");
	unsigned long long c[2] = { 0, 0 };
	TG_CHECK_HIP(hipMemcpy(c, d.ptr, 16, hipMemcpyDeviceToHost), "quotient check");
	*mismatches = c[0];
	*checked = c[1];
	return TG_OK;
}

extern "C" int tg_eval_points(const tg_program* p, const float* xyz, uint64_t n, float* out, uint32_t flags)
{
	if (!p || (n && (!xyz || !out))) return fail(TG_ERR_INVALID, "tg_eval_points: null argument");
	if (n == 0) return TG_OK;
	DevBuf din, dout;
	int rc = din.ensure(n * 12);
	if (!rc) rc = dout.ensure(n * 4);
	if (rc) return rc;
	struct Guard { DevBuf& a; DevBuf& b; ~Guard() { a.release(); b.release(); } } g{ din, dout };
	TG_CHECK_HIP(hipMemcpy(din.ptr, xyz, n * 12, hipMemcpyHostToDevice), "points upload");
	rc = tg_eval_points_device(p, static_cast<const float*>(din.ptr), n, static_cast<float*>(dout.ptr), flags, nullptr);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemcpy(out, dout.ptr, n * 4, hipMemcpyDeviceToHost), "points download");
	return TG_OK;
}
