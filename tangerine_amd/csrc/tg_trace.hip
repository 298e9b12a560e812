// Sphere tracing and gradients on the device (SDFNode::RayMarch and
// SDFNode::Gradient batched; tangerine/sdf_evaluator.cpp:298-353).
//
// For tracing, pruned programs are built for a four-level hierarchy of cubic
// cells (8^3 .. 64^3 cells over a cube holding the scene bounds and the eye);
// each march step of a wave runs the program of the finest cell holding all
// of its live rays, which equals the whole program at every point of that cell.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"
#include "tg_tree.h"

using namespace tg;

#define TG_CHECK_HIP(expr, what)                         \
	do                                                   \
	{                                                    \
		hipError_t _e = (expr);                          \
		if (_e != hipSuccess) return hip_fail(_e, what); \
	} while (0)

// A model transform as passed to a kernel: rotation (x, y, z, w), translation, scale.
struct Transform8
{
	float x, y, z, w, tx, ty, tz, s;
};

// Marching passes of a camera render.  The first pass marches every pixel's
// ray at most kFirstBudget steps, 64 rays (an 8x8 tile) to a wave; nearly all
// rays end there.  Each later pass takes the rays still marching, sorted by
// the cell they are in, at most kPassBudget steps, packed so that about
// kResumeWaves waves run them (at most 64 rays, at least one to a wave): the
// long marches of a frame no longer share a wave, where each waited on the
// others' cell programs.  (1920x1080, 256 primitives: 10.6 ms in one pass,
// 9.0 ms as 48 + 208 steps with the rays left one to a wave.)
constexpr int kFirstBudget = 48;
constexpr int kPassBudget = 208;
constexpr uint32_t kResumeWaves = 65536;
constexpr int kKeyBits = 27;  // trace_key: level above an 18-bit Morton code

struct tg_tracer
{
	hipStream_t stream = nullptr;
	hipEvent_t ev[3] = {};
	DevBuf t, normals, steps, work;
	DevBuf queue[2], keys, order, qctl, sort_temp;  // rays carried between passes
	uint32_t* host_count = nullptr;                  // pinned
	int first_budget = kFirstBudget, budget = kPassBudget;
	uint32_t resume_waves = kResumeWaves;
	CellPrograms levels[kTraceLevels];
	float dom_lo[3] = { 0, 0, 0 };
	float dom_size = 0.0f;        // 0: derived from the camera
	uint32_t width = 0, height = 0;
	tg_trace_stats stats{};
};

namespace {

// Cell hierarchy: an outer level of 8^3 cells over `outer` (the cube holding
// the eye), then 8, 16, 32, 64 cells per axis over `inner` (the scene's cube,
// where rays spend their steps).  Each level is a cube (lowest corner, edge).
struct CubeDomain
{
	float lo[3];
	float size;
};

// Levels coarse to fine: doms[L] split into ns[L]^3 cells.
int build_levels(hipStream_t stream, const tg_program* p, const CubeDomain* doms, const uint32_t* ns, int count,
	bool fast, CellPrograms* levels, TraceArgs& ta)
{
	ta.nlevels = 0;
	for (int L = 0; L < count && L < kTraceLevels; ++L)
	{
		const CubeDomain& dom = doms[L];
		const uint32_t n = ns[L];
		const float cell = dom.size / float(n);
		const float c3[3] = { cell, cell, cell };
		const uint32_t counts[3] = { n, n, n };
		const float radius = 0.5f * cell * 1.7320508f;
		int rc = prune_cells(stream, p, dom.lo, c3, radius, counts, fast, levels[L]);
		if (rc) return rc;
		if (!levels[L].valid) return TG_OK;
		ta.levels[L] = TraceLevel{ levels[L].desc.as<const BlockDesc>(), levels[L].pool.as<const Entry>(), n, cell,
		                           { dom.lo[0], dom.lo[1], dom.lo[2] }, 1.0f / cell };
		ta.nlevels = L + 1;
	}
	return TG_OK;
}

__global__ void k_iota(uint32_t* v, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) v[i] = uint32_t(i);
}

// Queues, keys (unsorted, then sorted), the identity order (first n words of
// `order`; the sorted order follows it) and the sort's temporary storage for
// n rays.
int ensure_pass_buffers(tg_tracer* tr, uint64_t n)
{
	const bool grow = tr->order.bytes < 2 * n * 4;
	int rc = tr->queue[0].ensure(n * sizeof(TraceRay));
	if (!rc) rc = tr->queue[1].ensure(n * sizeof(TraceRay));
	if (!rc) rc = tr->keys.ensure(2 * n * 4);
	if (!rc) rc = tr->order.ensure(2 * n * 4);
	if (rc) return rc;
	size_t tb = 0;
	TG_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint32_t*)nullptr, (uint32_t*)nullptr,
		(const uint32_t*)nullptr, (uint32_t*)nullptr, int(n), 0, kKeyBits, tr->stream), "sort size");
	rc = tr->sort_temp.ensure(std::max<size_t>(tb, 256));
	if (rc) return rc;
	if (!tr->host_count)
	{
		hipError_t e = hipHostMalloc(&tr->host_count, 64, hipHostMallocDefault);
		if (e != hipSuccess)
		{
			tr->host_count = nullptr;
			return hip_fail(e, "pinned pass count");
		}
	}
	if (grow)
	{
		hipLaunchKernelGGL(k_iota, dim3(unsigned((n + 255) / 256)), dim3(256), 0, tr->stream, tr->order.as<uint32_t>(), n);
		TG_CHECK_HIP(hipGetLastError(), "order init");
	}
	return TG_OK;
}

void base_args(TraceArgs& ta, const tg_program* p)
{
	ta = TraceArgs{};
	ta.prog = static_cast<const Instr*>(p->device.ptr);
	ta.whole = p->entries.as<const Entry>();
	ta.n_whole = uint32_t(p->decoded.entries.size());
	ta.whole_flops = uint32_t(p->decoded.flops);
	ta.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
}

}  // namespace

extern "C" int tg_camera_look_at(tg_camera* out, const float eye[3], const float target[3], const float up[3],
	float fov_y_degrees, uint32_t width, uint32_t height)
{
	if (!out || !eye || !target || !up || width == 0 || height == 0 || !(fov_y_degrees > 0.0f && fov_y_degrees < 180.0f))
		return fail(TG_ERR_INVALID, "tg_camera_look_at: bad argument");
	auto norm = [](float v[3]) {
		const float l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
		if (!(l > 0.0f)) return false;
		v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
		return true;
	};
	auto cross = [](const float a[3], const float b[3], float o[3]) {
		o[0] = a[1] * b[2] - a[2] * b[1];
		o[1] = a[2] * b[0] - a[0] * b[2];
		o[2] = a[0] * b[1] - a[1] * b[0];
	};
	float f[3] = { target[0] - eye[0], target[1] - eye[1], target[2] - eye[2] };
	if (!norm(f)) return fail(TG_ERR_INVALID, "tg_camera_look_at: eye equals target");
	float u[3] = { up[0], up[1], up[2] };
	float r[3];
	cross(f, u, r);
	if (!norm(r))
	{
		// up parallel to the view direction: use the axis least aligned with it
		const float ax[3][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
		int best = 0;
		for (int i = 1; i < 3; ++i)
			if (std::fabs(f[i]) < std::fabs(f[best])) best = i;
		cross(f, ax[best], r);
		norm(r);
	}
	float tu[3];
	cross(r, f, tu);
	for (int i = 0; i < 3; ++i)
	{
		out->eye[i] = eye[i];
		out->forward[i] = f[i];
		out->right[i] = r[i];
		out->up[i] = tu[i];
	}
	const double half = double(fov_y_degrees) * 0.5 * 3.14159265358979323846 / 180.0;
	out->tan_half_y = float(std::tan(half));
	out->tan_half_x = float(std::tan(half) * double(width) / double(height));
	out->width = width;
	out->height = height;
	return TG_OK;
}

extern "C" int tg_tracer_create(tg_tracer** out)
{
	if (!out) return fail(TG_ERR_INVALID, "tg_tracer_create: null argument");
	int rc = ensure_device();
	if (rc) return rc;
	auto* tr = new tg_tracer;
	hipError_t e = hipStreamCreate(&tr->stream);
	for (int i = 0; e == hipSuccess && i < 3; ++i) e = hipEventCreate(&tr->ev[i]);
	if (e != hipSuccess)
	{
		delete tr;
		return hip_fail(e, "tracer stream/events");
	}
	*out = tr;
	return TG_OK;
}

extern "C" void tg_tracer_destroy(tg_tracer* tr)
{
	if (!tr) return;
	if (tr->stream) hipStreamSynchronize(tr->stream);
	for (DevBuf* b : { &tr->t, &tr->normals, &tr->steps, &tr->work, &tr->queue[0], &tr->queue[1], &tr->keys,
	                   &tr->order, &tr->qctl, &tr->sort_temp })
		b->release();
	if (tr->host_count) (void)hipHostFree(tr->host_count);
	for (CellPrograms& l : tr->levels) l.release();
	for (hipEvent_t e : tr->ev)
		if (e) hipEventDestroy(e);
	if (tr->stream) hipStreamDestroy(tr->stream);
	delete tr;
}

extern "C" int tg_tracer_set_passes(tg_tracer* tr, int first_budget, int budget, uint32_t resume_waves)
{
	if (!tr || (first_budget > 0 && (budget <= 0 || resume_waves == 0)))
		return fail(TG_ERR_INVALID, "tg_tracer_set_passes: bad argument");
	tr->first_budget = first_budget;
	tr->budget = budget;
	tr->resume_waves = resume_waves;
	return TG_OK;
}

extern "C" int tg_tracer_set_domain(tg_tracer* tr, const float lo[3], float size)
{
	if (!tr || !lo || !(size > 0.0f)) return fail(TG_ERR_INVALID, "tg_tracer_set_domain: bad argument");
	for (int i = 0; i < 3; ++i) tr->dom_lo[i] = lo[i];
	tr->dom_size = size;
	return TG_OK;
}

extern "C" int tg_tracer_render(tg_tracer* tr, const tg_program* p, const tg_camera* cam, int max_steps, float eps,
	float far_distance, float normal_h, uint32_t flags)
{
	if (!tr || !p || !cam) return fail(TG_ERR_INVALID, "tg_tracer_render: null argument");
	if (cam->width == 0 || cam->height == 0 || uint64_t(cam->width) * cam->height >= (1ull << 31))
		return fail(TG_ERR_LIMIT, "tg_tracer_render: image size out of range");
	const bool fast = (flags & TG_FLAG_FAST) != 0;
	const uint64_t n = uint64_t(cam->width) * cam->height;
	// passes: the first, then enough resumed ones to reach max_steps
	const bool passes = tr->first_budget > 0 && tr->first_budget < max_steps;
	const int npass = passes ? 1 + (max_steps - tr->first_budget + tr->budget - 1) / tr->budget : 1;
	int rc = tr->t.ensure(n * 4);
	if (!rc) rc = tr->normals.ensure(n * 12);
	if (!rc) rc = tr->steps.ensure(n * 4);
	if (!rc) rc = tr->work.ensure(64);
	if (!rc) rc = tr->qctl.ensure(size_t(npass) * 4);
	if (!rc && passes) rc = ensure_pass_buffers(tr, n);
	if (rc) return rc;
	TraceArgs ta;
	base_args(ta, p);
	TG_CHECK_HIP(hipEventRecord(tr->ev[0], tr->stream), "event");
	if (!(flags & TG_FLAG_NO_PRUNE))
	{
		// Outer level: the origin-centred cube that holds the eye.  Inner
		// levels: the caller's domain, or the cube around the program's scene
		// box (outer otherwise).
		CubeDomain outer{}, inner{};
		float reach = 1.0f;
		for (int i = 0; i < 3; ++i) reach = std::max(reach, std::fabs(cam->eye[i]) + 1.0f);
		outer.size = 2.0f * reach;
		for (int i = 0; i < 3; ++i) outer.lo[i] = -reach;
		inner = outer;
		// the cube every ray stays in until the far cutoff
		CubeDomain farthest{};
		const float far_reach = reach + far_distance + 1.0f;
		const bool far_level = std::isfinite(far_reach) && far_distance > 0.0f && far_reach < 1e30f;
		farthest.size = 2.0f * far_reach;
		for (int i = 0; i < 3; ++i) farthest.lo[i] = -far_reach;
		const DecodedProgram& d = p->decoded;
		if (tr->dom_size > 0.0f)
		{
			for (int i = 0; i < 3; ++i) inner.lo[i] = tr->dom_lo[i];
			inner.size = tr->dom_size;
		}
		else if (d.scene_finite)
		{
			float edge = 0.0f;
			for (int i = 0; i < 3; ++i) edge = std::max(edge, d.scene_hi[i] - d.scene_lo[i]);
			edge = edge * 1.01f + 1e-3f;
			if (edge < outer.size)
			{
				inner.size = edge;
				for (int i = 0; i < 3; ++i) inner.lo[i] = 0.5f * (d.scene_lo[i] + d.scene_hi[i]) - 0.5f * edge;
			}
		}
		// far cube (rays leaving the scene march on through few, small cell
		// programs), the eye's cube, then 16^3 and 64^3 cells over the scene.
		// (1080p, 256 primitives, pruning + marching: 8/16/32/64 levels over
		// the scene 11.2 ms, 16/64 10.7, 32/64 10.6, 64 alone 10.5, 48/96 12.3:
		// the waterfall over per-ray cells makes the middle levels cheap to drop.)
		CubeDomain doms[kTraceLevels];
		uint32_t ns[kTraceLevels];
		int count = 0;
		if (far_level)
		{
			doms[count] = farthest;
			ns[count++] = 8;
		}
		doms[count] = outer;
		ns[count++] = 8;
		for (uint32_t n : { 16u, 64u })
		{
			doms[count] = inner;
			ns[count++] = n;
		}
		rc = build_levels(tr->stream, p, doms, ns, count, fast, tr->levels, ta);
		if (rc) return rc;
	}
	TG_CHECK_HIP(hipEventRecord(tr->ev[1], tr->stream), "event");
	ta.camera = 1;
	for (int i = 0; i < 3; ++i)
	{
		ta.eye[i] = cam->eye[i];
		ta.fwd[i] = cam->forward[i];
		ta.right[i] = cam->right[i];
		ta.up[i] = cam->up[i];
	}
	ta.kx = cam->tan_half_x;
	ta.ky = cam->tan_half_y;
	ta.width = cam->width;
	ta.height = cam->height;
	ta.max_steps = max_steps;
	ta.eps = eps;
	ta.far_ = far_distance;
	ta.h = normal_h;
	ta.t_out = tr->t.as<float>();
	ta.n_out = tr->normals.as<float>();
	ta.steps_out = tr->steps.as<uint32_t>();
	ta.work = tr->work.as<unsigned long long>();
	TG_CHECK_HIP(hipMemsetAsync(tr->work.ptr, 0, 64, tr->stream), "work reset");
	// per pass: the count of rays it queued
	uint32_t* ctl = tr->qctl.as<uint32_t>();
	TG_CHECK_HIP(hipMemsetAsync(ctl, 0, size_t(npass) * 4, tr->stream), "pass counters");
	int taken = 0;
	for (int k = 0; k < npass; ++k)
	{
		const int left = max_steps - taken;
		ta.budget = passes ? std::min(left, k == 0 ? tr->first_budget : tr->budget) : max_steps;
		const bool last = k + 1 == npass;
		ta.q_out = last ? nullptr : tr->queue[k & 1].as<TraceRay>();
		ta.q_keys = last ? nullptr : tr->keys.as<uint32_t>();
		ta.q_out_count = last ? nullptr : ctl + k;
		hipError_t e;
		if (k == 0)
			e = fast ? fast::launch_trace(ta, tr->stream) : exact::launch_trace(ta, tr->stream);
		else
		{
			// the rays the previous pass left, in key order (one cell's rays together)
			TG_CHECK_HIP(hipMemcpyAsync(tr->host_count, ctl + k - 1, 4, hipMemcpyDeviceToHost, tr->stream), "pass count");
			TG_CHECK_HIP(hipStreamSynchronize(tr->stream), "pass count");
			const uint32_t m = *tr->host_count;
			if (m == 0) break;
			uint32_t* keys = tr->keys.as<uint32_t>();
			size_t tb = tr->sort_temp.bytes;
			TG_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tr->sort_temp.ptr, tb, keys, keys + n,
				tr->order.as<const uint32_t>(), tr->order.as<uint32_t>() + n, int(m), 0, kKeyBits, tr->stream),
				"pass sort");
			ta.q_in = tr->queue[(k - 1) & 1].as<const TraceRay>();
			ta.q_order = tr->order.as<const uint32_t>() + n;
			ta.q_count = m;
			ta.q_per_wave = std::max<uint32_t>(1, std::min<uint32_t>(64, (m + tr->resume_waves - 1) / tr->resume_waves));
			e = fast ? fast::launch_trace_resume(ta, tr->stream) : exact::launch_trace_resume(ta, tr->stream);
		}
		if (e != hipSuccess) return hip_fail(e, "trace launch");
		taken += ta.budget;
	}
	TG_CHECK_HIP(hipEventRecord(tr->ev[2], tr->stream), "event");
	unsigned long long wk[3] = { 0, 0, 0 };
	TG_CHECK_HIP(hipMemcpyAsync(wk, tr->work.ptr, 24, hipMemcpyDeviceToHost, tr->stream), "work read");
	TG_CHECK_HIP(hipStreamSynchronize(tr->stream), "trace");
	float t_prune = 0, t_trace = 0;
	hipEventElapsedTime(&t_prune, tr->ev[0], tr->ev[1]);
	hipEventElapsedTime(&t_trace, tr->ev[1], tr->ev[2]);
	tr->width = cam->width;
	tr->height = cam->height;
	tr->stats = tg_trace_stats{ t_trace, t_prune, n, wk[2], wk[0], double(wk[1]) };
	return TG_OK;
}

extern "C" int tg_tracer_copy(const tg_tracer* tr, float* t, float* normals, uint32_t* steps)
{
	if (!tr) return fail(TG_ERR_INVALID, "tg_tracer_copy: null argument");
	const uint64_t n = uint64_t(tr->width) * tr->height;
	if (n == 0) return fail(TG_ERR_INVALID, "tg_tracer_copy: nothing rendered yet");
	if (t) TG_CHECK_HIP(hipMemcpy(t, tr->t.ptr, n * 4, hipMemcpyDeviceToHost), "t download");
	if (normals) TG_CHECK_HIP(hipMemcpy(normals, tr->normals.ptr, n * 12, hipMemcpyDeviceToHost), "normal download");
	if (steps) TG_CHECK_HIP(hipMemcpy(steps, tr->steps.ptr, n * 4, hipMemcpyDeviceToHost), "steps download");
	return TG_OK;
}

extern "C" int tg_tracer_stats(const tg_tracer* tr, tg_trace_stats* out)
{
	if (!tr || !out) return fail(TG_ERR_INVALID, "tg_tracer_stats: null argument");
	*out = tr->stats;
	return TG_OK;
}

namespace {

// One batched march of `count` rays whose inputs and outputs are device
// arrays already holding / sized for them; results copied to the host arrays.
int raymarch_staged(const tg_program* p, const float* starts, const float* dirs, uint64_t count, int max_steps,
	float eps, float far_distance, uint32_t flags, uint8_t* hit, float* travel, float* position, float* ds, float* dd,
	float* dt, float* dp, uint8_t* dh)
{
	TG_CHECK_HIP(hipMemcpy(ds, starts, count * 12, hipMemcpyHostToDevice), "ray upload");
	TG_CHECK_HIP(hipMemcpy(dd, dirs, count * 12, hipMemcpyHostToDevice), "ray upload");
	TraceArgs ta;
	base_args(ta, p);
	ta.camera = 0;
	ta.starts = ds;
	ta.dirs = dd;
	ta.nrays = uint32_t(count);
	ta.max_steps = max_steps;
	ta.eps = eps;
	ta.far_ = far_distance;
	ta.h = 0.0f;
	ta.t_out = dt;
	ta.p_out = dp;
	ta.hit_out = dh;
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_trace(ta, nullptr) : exact::launch_trace(ta, nullptr);
	if (e != hipSuccess) return hip_fail(e, "raymarch launch");
	TG_CHECK_HIP(hipMemcpy(travel, dt, count * 4, hipMemcpyDeviceToHost), "travel download");
	if (position) TG_CHECK_HIP(hipMemcpy(position, dp, count * 12, hipMemcpyDeviceToHost), "position download");
	if (hit) TG_CHECK_HIP(hipMemcpy(hit, dh, count, hipMemcpyDeviceToHost), "hit download");
	return TG_OK;
}

constexpr uint64_t kStaticRays = TG_RAYMARCH_STATIC_MAX;
constexpr size_t kStaticBytesPerRay = 41;  // start 12, direction 12, travel 4, position 12, hit 1

}  // namespace

// Staging of tg_raymarch_static: one copy per device, for the whole process.
// In host code this name is only a handle for hipGetSymbolAddress; kernels
// and copies take the device address that call returns.
static __device__ __align__(256) unsigned char scratch[kStaticBytesPerRay * kStaticRays];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int static_scratch(void** out)
{
	int dev = 0;
	TG_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	if (dev < 0 || dev >= kMaxDevices) return fail(TG_ERR_LIMIT, "tg_raymarch_static: device index beyond 63");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_ptr[dev])
	{
		void* ptr = nullptr;
		TG_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		if (!ptr) return fail(TG_ERR_DEVICE, "tg_raymarch_static: no device address for the staging array");
		g_scratch_ptr[dev] = ptr;
	}
	*out = g_scratch_ptr[dev];
	return TG_OK;
}

}  // namespace

extern "C" int tg_raymarch(const tg_program* p, const float* starts, const float* dirs, uint64_t count, int max_steps,
	float eps, float far_distance, uint32_t flags, uint8_t* hit, float* travel, float* position)
{
	if (!p || (count && (!starts || !dirs || !travel))) return fail(TG_ERR_INVALID, "tg_raymarch: null argument");
	if (count == 0) return TG_OK;
	if (count >= (1ull << 31)) return fail(TG_ERR_LIMIT, "tg_raymarch: too many rays");
	int rc = ensure_device();
	if (rc) return rc;
	DevBuf ds, dd, dt, dp, dh;
	struct Guard { DevBuf* b[5]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &ds, &dd, &dt, &dp, &dh } };
	rc = ds.ensure(count * 12);
	if (!rc) rc = dd.ensure(count * 12);
	if (!rc) rc = dt.ensure(count * 4);
	if (!rc) rc = dp.ensure(count * 12);
	if (!rc) rc = dh.ensure(count);
	if (rc) return rc;
	return raymarch_staged(p, starts, dirs, count, max_steps, eps, far_distance, flags, hit, travel, position,
		ds.as<float>(), dd.as<float>(), dt.as<float>(), dp.as<float>(), dh.as<uint8_t>());
}

extern "C" int tg_raymarch_static(const tg_program* p, const float* starts, const float* dirs, uint64_t count,
	int max_steps, float eps, float far_distance, uint32_t flags, uint8_t* hit, float* travel, float* position)
{
	if (!p || (count && (!starts || !dirs || !travel))) return fail(TG_ERR_INVALID, "tg_raymarch_static: null argument");
	if (count == 0) return TG_OK;
	if (count > kStaticRays) return fail(TG_ERR_LIMIT, "tg_raymarch_static: more than TG_RAYMARCH_STATIC_MAX rays");
	int rc = ensure_device();
	if (rc) return rc;
	void* base = nullptr;
	rc = static_scratch(&base);
	if (rc) return rc;
	unsigned char* b = static_cast<unsigned char*>(base);
	float* ds = reinterpret_cast<float*>(b);
	float* dd = reinterpret_cast<float*>(b + 12 * kStaticRays);
	float* dt = reinterpret_cast<float*>(b + 24 * kStaticRays);
	float* dp = reinterpret_cast<float*>(b + 28 * kStaticRays);
	uint8_t* dh = b + 40 * kStaticRays;
	return raymarch_staged(p, starts, dirs, count, max_steps, eps, far_distance, flags, hit, travel, position, ds, dd,
		dt, dp, dh);
}

extern "C" int tg_gradient_points(const tg_program* p, const float* xyz, uint64_t count, float h, int mode, float* out,
	uint32_t flags)
{
	if (!p || (count && (!xyz || !out))) return fail(TG_ERR_INVALID, "tg_gradient_points: null argument");
	if (mode != 0 && mode != 1) return fail(TG_ERR_INVALID, "tg_gradient_points: mode must be 0 or 1");
	if (count == 0) return TG_OK;
	int rc = ensure_device();
	if (rc) return rc;
	DevBuf din, dout;
	struct Guard { DevBuf* b[2]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &din, &dout } };
	rc = din.ensure(count * 12);
	if (!rc) rc = dout.ensure(count * 12);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemcpy(din.ptr, xyz, count * 12, hipMemcpyHostToDevice), "points upload");
	GradArgs ga{};
	ga.prog = static_cast<const Instr*>(p->device.ptr);
	ga.xyz = din.as<const float>();
	ga.out = dout.as<float>();
	ga.count = count;
	ga.h = h;
	ga.tetra = mode == 0 ? 1 : 0;
	ga.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_gradient(ga, nullptr) : exact::launch_gradient(ga, nullptr);
	if (e != hipSuccess) return hip_fail(e, "gradient launch");
	TG_CHECK_HIP(hipMemcpy(out, dout.ptr, count * 12, hipMemcpyDeviceToHost), "gradient download");
	return TG_OK;
}

// ---- Lua queries: pivot_towards and mouse picking --------------------------------

extern "C" int tg_pivot_towards(const tg_program* p, uint64_t count, const float* pivots, const float* headings,
	const float* downs, float max_dist, float margin, float max_angle_degrees, int max_iterations, float* tails,
	float* headings_out, int32_t* status, uint32_t flags)
{
	if (!p || (count && (!pivots || !headings || !downs || !tails || !headings_out || !status)))
		return fail(TG_ERR_INVALID, "tg_pivot_towards: null argument");
	if (max_iterations < 1) return fail(TG_ERR_INVALID, "tg_pivot_towards: max_iterations must be positive");
	if (count == 0) return TG_OK;
	if (count >= (1ull << 31)) return fail(TG_ERR_LIMIT, "tg_pivot_towards: too many queries");
	int rc = ensure_device();
	if (rc) return rc;
	DevBuf din, dout, dst;
	struct Guard { DevBuf* b[3]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &din, &dout, &dst } };
	rc = din.ensure(count * 36);
	if (!rc) rc = dout.ensure(count * 24);
	if (!rc) rc = dst.ensure(count * 4);
	if (rc) return rc;
	float* in = din.as<float>();
	TG_CHECK_HIP(hipMemcpy(in, pivots, count * 12, hipMemcpyHostToDevice), "query upload");
	TG_CHECK_HIP(hipMemcpy(in + 3 * count, headings, count * 12, hipMemcpyHostToDevice), "query upload");
	TG_CHECK_HIP(hipMemcpy(in + 6 * count, downs, count * 12, hipMemcpyHostToDevice), "query upload");
	PivotArgs a{};
	a.prog = static_cast<const Instr*>(p->device.ptr);
	a.pivot = in;
	a.heading = in + 3 * count;
	a.down = in + 6 * count;
	a.count = count;
	a.max_dist = max_dist;
	a.margin = margin;
	a.max_angle = max_angle_degrees * 0.01745329251994329576923690768489f;  // glm::radians(float)
	a.max_iterations = max_iterations;
	a.tail_out = dout.as<float>();
	a.heading_out = dout.as<float>() + 3 * count;
	a.status_out = dst.as<int32_t>();
	a.slots = std::max<uint32_t>(1, p->decoded.max_stack - 1);
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_pivot(a, nullptr) : exact::launch_pivot(a, nullptr);
	if (e != hipSuccess) return hip_fail(e, "pivot launch");
	std::vector<float> out(size_t(count) * 6);
	TG_CHECK_HIP(hipMemcpy(out.data(), dout.ptr, count * 24, hipMemcpyDeviceToHost), "pivot download");
	TG_CHECK_HIP(hipMemcpy(status, dst.ptr, count * 4, hipMemcpyDeviceToHost), "pivot download");
	for (uint64_t i = 0; i < count; ++i)
	{
		if (status[i] != 0) continue;  // outputs left as the caller had them
		for (int k = 0; k < 3; ++k)
		{
			tails[3 * i + k] = out[3 * i + k];
			headings_out[3 * i + k] = out[3 * (count + i) + k];
		}
	}
	return TG_OK;
}

namespace {

// Every ray in model space: Transform::ApplyInv for the start,
// rotate(inverse(q), dir) for the direction (transform.cpp:64-67,
// sdf_model.cpp:259-264).  xf: rotation (x, y, z, w), translation, scale.
__global__ void __launch_bounds__(256) k_local_rays(const float* starts, const float* dirs, uint64_t n, Transform8 xf,
	float* out_starts, float* out_dirs)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	// inverse(q) = conjugate(q) / dot(q, q), dot = (w w + x x) + (y y + z z)
	const float qd = (xf.w * xf.w + xf.x * xf.x) + (xf.y * xf.y + xf.z * xf.z);
	const float w = xf.w / qd, qx = -xf.x / qd, qy = -xf.y / qd, qz = -xf.z / qd;
	auto rot = [&](float& x, float& y, float& z) {
		const float uvx = qy * z - y * qz, uvy = qz * x - z * qx, uvz = qx * y - x * qy;
		const float uuvx = qy * uvz - uvy * qz, uuvy = qz * uvx - uvz * qx, uuvz = qx * uvy - uvx * qy;
		x = x + (uvx * w + uuvx) * 2.0f;
		y = y + (uvy * w + uuvy) * 2.0f;
		z = z + (uvz * w + uuvz) * 2.0f;
	};
	float sx = starts[3 * i] - xf.tx, sy = starts[3 * i + 1] - xf.ty, sz = starts[3 * i + 2] - xf.tz;
	rot(sx, sy, sz);
	out_starts[3 * i] = sx / xf.s;
	out_starts[3 * i + 1] = sy / xf.s;
	out_starts[3 * i + 2] = sz / xf.s;
	float dx = dirs[3 * i], dy = dirs[3 * i + 1], dz = dirs[3 * i + 2];
	rot(dx, dy, dz);
	out_dirs[3 * i] = dx;
	out_dirs[3 * i + 1] = dy;
	out_dirs[3 * i + 2] = dz;
}

// Keeps, per ray, the hit with the smallest local travel (strictly smaller:
// the earlier model wins ties, as DeliverMouseButton's `Travel < Nearest`).
__global__ void __launch_bounds__(256) k_pick_reduce(const uint8_t* hit, const float* travel, const float* pos,
	uint64_t n, int32_t model, int32_t* best_model, float* best_travel, float* best_pos)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	if (hit[i] && travel[i] < best_travel[i])
	{
		best_travel[i] = travel[i];
		best_model[i] = model;
		best_pos[3 * i] = pos[3 * i];
		best_pos[3 * i + 1] = pos[3 * i + 1];
		best_pos[3 * i + 2] = pos[3 * i + 2];
	}
}

__global__ void __launch_bounds__(256) k_pick_init(uint64_t n, int32_t* best_model, float* best_travel, float* best_pos)
{
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	best_model[i] = -1;
	best_travel[i] = __builtin_huge_valf();
	best_pos[3 * i] = best_pos[3 * i + 1] = best_pos[3 * i + 2] = 0.0f;
}

}  // namespace

extern "C" int tg_pick(const tg_program* const* programs, const float* transforms, uint32_t model_count,
	const float* starts, const float* dirs, uint64_t ray_count, uint32_t flags, int32_t* model_out, float* travel_out,
	float* position_out)
{
	if ((model_count && (!programs || !transforms)) || (ray_count && (!starts || !dirs || !model_out || !travel_out)))
		return fail(TG_ERR_INVALID, "tg_pick: null argument");
	for (uint32_t m = 0; m < model_count; ++m)
		if (!programs[m]) return fail(TG_ERR_INVALID, "tg_pick: null program");
	if (ray_count == 0) return TG_OK;
	if (ray_count >= (1ull << 31)) return fail(TG_ERR_LIMIT, "tg_pick: too many rays");
	int rc = ensure_device();
	if (rc) return rc;
	DevBuf rays, local, res, best;
	struct Guard { DevBuf* b[4]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &rays, &local, &res, &best } };
	const uint64_t n = ray_count;
	rc = rays.ensure(n * 24);
	if (!rc) rc = local.ensure(n * 24);
	if (!rc) rc = res.ensure(n * 17);    // travel, position, hit
	if (!rc) rc = best.ensure(n * 20);   // model, travel, position
	if (rc) return rc;
	float* ray_s = rays.as<float>();
	float* ray_d = ray_s + 3 * n;
	float* loc_s = local.as<float>();
	float* loc_d = loc_s + 3 * n;
	float* r_t = res.as<float>();
	float* r_p = r_t + n;
	uint8_t* r_h = reinterpret_cast<uint8_t*>(r_p + 3 * n);
	int32_t* b_m = best.as<int32_t>();
	float* b_t = reinterpret_cast<float*>(b_m + n);
	float* b_p = b_t + n;
	TG_CHECK_HIP(hipMemcpy(ray_s, starts, n * 12, hipMemcpyHostToDevice), "ray upload");
	TG_CHECK_HIP(hipMemcpy(ray_d, dirs, n * 12, hipMemcpyHostToDevice), "ray upload");
	const unsigned blocks = unsigned((n + 255) / 256);
	hipLaunchKernelGGL(k_pick_init, dim3(blocks), dim3(256), 0, nullptr, n, b_m, b_t, b_p);
	TG_CHECK_HIP(hipGetLastError(), "pick init");
	for (uint32_t m = 0; m < model_count; ++m)
	{
		const float* t = transforms + 8 * size_t(m);
		if (!(t[7] != 0.0f)) return fail(TG_ERR_INVALID, "tg_pick: zero scale");
		const Transform8 xf{ t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7] };
		hipLaunchKernelGGL(k_local_rays, dim3(blocks), dim3(256), 0, nullptr, ray_s, ray_d, n, xf, loc_s, loc_d);
		TG_CHECK_HIP(hipGetLastError(), "local rays");
		TraceArgs ta;
		base_args(ta, programs[m]);
		ta.camera = 0;
		ta.starts = loc_s;
		ta.dirs = loc_d;
		ta.nrays = uint32_t(n);
		ta.max_steps = 1000;  // SDFModel::RayMarch: Evaluator->RayMarch(origin, dir, 1000), epsilon 0.001
		ta.eps = 0.001f;
		ta.far_ = __builtin_huge_valf();
		ta.h = 0.0f;
		ta.t_out = r_t;
		ta.p_out = r_p;
		ta.hit_out = r_h;
		hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_trace(ta, nullptr) : exact::launch_trace(ta, nullptr);
		if (e != hipSuccess) return hip_fail(e, "pick trace launch");
		hipLaunchKernelGGL(k_pick_reduce, dim3(blocks), dim3(256), 0, nullptr, r_h, r_t, r_p, n, int32_t(m), b_m, b_t, b_p);
		TG_CHECK_HIP(hipGetLastError(), "pick reduce");
	}
	TG_CHECK_HIP(hipMemcpy(model_out, b_m, n * 4, hipMemcpyDeviceToHost), "pick download");
	TG_CHECK_HIP(hipMemcpy(travel_out, b_t, n * 4, hipMemcpyDeviceToHost), "pick download");
	if (position_out) TG_CHECK_HIP(hipMemcpy(position_out, b_p, n * 12, hipMemcpyDeviceToHost), "pick download");
	return TG_OK;
}
