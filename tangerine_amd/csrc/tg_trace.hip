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

struct tg_tracer
{
	hipStream_t stream = nullptr;
	hipEvent_t ev[3] = {};
	DevBuf t, normals, steps, work;
	CellPrograms levels[kTraceLevels];
	float dom_lo[3] = { 0, 0, 0 };
	float dom_size = 0.0f;        // 0: derived from the camera
	uint32_t width = 0, height = 0;
	tg_trace_stats stats{};
};

namespace {

// Cell hierarchy over the cube [lo, lo + size]^3: levels of 8, 16, 32, 64 cells.
int build_levels(hipStream_t stream, const tg_program* p, const float lo[3], float size, bool fast,
	CellPrograms* levels, TraceArgs& ta)
{
	ta.nlevels = 0;
	for (int L = 0; L < kTraceLevels; ++L)
	{
		const uint32_t n = 8u << L;
		const float cell = size / float(n);
		const float c3[3] = { cell, cell, cell };
		const uint32_t counts[3] = { n, n, n };
		const float radius = 0.5f * cell * 1.7320508f;
		int rc = prune_cells(stream, p, lo, c3, radius, counts, fast, levels[L]);
		if (rc) return rc;
		if (!levels[L].valid) return TG_OK;
		ta.levels[L] = TraceLevel{ levels[L].desc.as<const BlockDesc>(), levels[L].pool.as<const Entry>(), n, cell };
		ta.nlevels = L + 1;
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
	for (DevBuf* b : { &tr->t, &tr->normals, &tr->steps, &tr->work }) b->release();
	for (CellPrograms& l : tr->levels) l.release();
	for (hipEvent_t e : tr->ev)
		if (e) hipEventDestroy(e);
	if (tr->stream) hipStreamDestroy(tr->stream);
	delete tr;
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
	int rc = tr->t.ensure(n * 4);
	if (!rc) rc = tr->normals.ensure(n * 12);
	if (!rc) rc = tr->steps.ensure(n * 4);
	if (!rc) rc = tr->work.ensure(64);
	if (rc) return rc;
	TraceArgs ta;
	base_args(ta, p);
	TG_CHECK_HIP(hipEventRecord(tr->ev[0], tr->stream), "event");
	if (!(flags & TG_FLAG_NO_PRUNE))
	{
		float lo[3], size;
		if (tr->dom_size > 0.0f)
		{
			for (int i = 0; i < 3; ++i) lo[i] = tr->dom_lo[i];
			size = tr->dom_size;
		}
		else
		{
			// Default domain: the origin-centred cube that holds the eye.
			float reach = 1.0f;
			for (int i = 0; i < 3; ++i) reach = std::max(reach, std::fabs(cam->eye[i]) + 1.0f);
			size = 2.0f * reach;
			for (int i = 0; i < 3; ++i) lo[i] = -reach;
		}
		rc = build_levels(tr->stream, p, lo, size, fast, tr->levels, ta);
		if (rc) return rc;
		for (int i = 0; i < 3; ++i) ta.dom[i] = lo[i];
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
	hipError_t e = fast ? fast::launch_trace(ta, tr->stream) : exact::launch_trace(ta, tr->stream);
	if (e != hipSuccess) return hip_fail(e, "trace launch");
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
	TG_CHECK_HIP(hipMemcpy(ds.ptr, starts, count * 12, hipMemcpyHostToDevice), "ray upload");
	TG_CHECK_HIP(hipMemcpy(dd.ptr, dirs, count * 12, hipMemcpyHostToDevice), "ray upload");
	TraceArgs ta;
	base_args(ta, p);
	ta.camera = 0;
	ta.starts = ds.as<const float>();
	ta.dirs = dd.as<const float>();
	ta.nrays = uint32_t(count);
	ta.max_steps = max_steps;
	ta.eps = eps;
	ta.far_ = far_distance;
	ta.h = 0.0f;
	ta.t_out = dt.as<float>();
	ta.p_out = dp.as<float>();
	ta.hit_out = dh.as<uint8_t>();
	hipError_t e = (flags & TG_FLAG_FAST) ? fast::launch_trace(ta, nullptr) : exact::launch_trace(ta, nullptr);
	if (e != hipSuccess) return hip_fail(e, "raymarch launch");
	TG_CHECK_HIP(hipMemcpy(travel, dt.ptr, count * 4, hipMemcpyDeviceToHost), "travel download");
	if (position) TG_CHECK_HIP(hipMemcpy(position, dp.ptr, count * 12, hipMemcpyDeviceToHost), "position download");
	if (hit) TG_CHECK_HIP(hipMemcpy(hit, dh.ptr, count, hipMemcpyDeviceToHost), "hit download");
	return TG_OK;
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
