// Learned brick-culling surrogate (no reference counterpart).
//
// fit():  samples are drawn on the device (half uniform in the domain, half
//         near the surface: vertices of a coarse exact export plus Gaussian
//         offsets), labelled by the exact evaluator, and fitted with Adam;
//         the training steps of one epoch are captured in a hipGraph and
//         replayed.  Separately drawn calibration samples (prediction and
//         exact distance) are kept on the host.
// threshold(r): a brick of half-diagonal r whose centre c holds surface has
//         |d(c)| <= r (the field is 1-Lipschitz), so the brick may be skipped
//         only if |prediction(c)| exceeds every prediction seen at calibration
//         points with |d| <= r.  The threshold is that maximum (at least r)
//         plus the mean calibration residual as slack; a margin set by the
//         caller replaces the calibration: threshold = r + margin.
// The culled export (tg_export.hip) evaluates exactly every tile that a
// flagged brick touches, so the surrogate only decides which work is skipped.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/tangerine_gpu.h"
#include "tg_capi_internal.h"
#include "tg_kernels.h"
#include "tg_mlp.h"

using namespace tg;

#define TG_CHECK_HIP(expr, what)                         \
	do                                                   \
	{                                                    \
		hipError_t _e = (expr);                          \
		if (_e != hipSuccess) return hip_fail(_e, what); \
	} while (0)

struct tg_surrogate
{
	hipStream_t stream = nullptr;
	hipEvent_t ev[4] = {};
	float lo[3] = { 0, 0, 0 };
	float size = 1.0f;
	float sigma = 2.0f;
	float clamp = 0.5f;
	uint64_t seed = 1;
	float margin = -1.0f;       // caller's fixed margin (< 0: use the calibration)
	bool calibrated = false;
	std::vector<float> calib_pred, calib_dist;  // held-out predictions and exact distances
	float mean_residual = 0.0f;
	// last threshold per mode (uncertified, certified) and the r it was for;
	// cleared whenever the calibration changes
	mutable float cache_r[2] = { -1.0f, -1.0f }, cache_t[2] = { 0.0f, 0.0f };
	DevBuf params, fourier, adam_m, adam_v, xyz, target, order, partial, loss_partial, counters, loss_log, tmp;
	tg_surrogate_stats stats{};
};

namespace tg {

mlp::Net surrogate_net(const tg_surrogate* s)
{
	mlp::Net n{};
	n.params = s->params.as<const float>();
	n.fourier = s->fourier.as<const float>();
	for (int i = 0; i < 3; ++i) n.lo[i] = s->lo[i];
	n.inv_half = 2.0f / s->size;
	n.clamp = s->clamp;
	return n;
}

float surrogate_threshold(const tg_surrogate* s, float r, bool certified)
{
	if (s->margin >= 0.0f) return r + s->margin;
	if (!s->calibrated) return -1.0f;
	const int mode = certified ? 1 : 0;
	if (s->cache_r[mode] == r) return s->cache_t[mode];
	// |prediction| at held-out points whose exact |d| <= r: a brick of that
	// half-diagonal holding surface looks like one of them.  Uncertified, the
	// largest is the bar; certified, every skipped brick is checked exactly
	// anyway, so the bar is their 90th percentile (fewer bricks kept for
	// nothing, the rest put back by the certificate).
	std::vector<float> near;
	float t = r;
	const size_t n = s->calib_pred.size();
	for (size_t i = 0; i < n; ++i)
		if (std::fabs(s->calib_dist[i]) <= r)
		{
			const float a = std::fabs(s->calib_pred[i]);
			if (certified)
				near.push_back(a);
			else
				t = std::max(t, a);
		}
	if (certified && !near.empty())
	{
		const size_t k = size_t(0.9 * double(near.size() - 1));
		std::nth_element(near.begin(), near.begin() + std::ptrdiff_t(k), near.end());
		t = std::max(t, near[k]);
	}
	s->cache_r[mode] = r;
	s->cache_t[mode] = t + s->mean_residual;
	return s->cache_t[mode];
}

}  // namespace tg

namespace {

// Counter-based generator: 32-bit hash of (seed, index).
__device__ __forceinline__ uint32_t hash32(uint64_t x)
{
	x ^= x >> 33;
	x *= 0xff51afd7ed558ccdull;
	x ^= x >> 33;
	x *= 0xc4ceb9fe1a85ec53ull;
	x ^= x >> 33;
	return uint32_t(x);
}

__device__ __forceinline__ float uniform01(uint64_t seed, uint64_t i)
{
	return (float(hash32(seed * 0x9e3779b97f4a7c15ull + i) >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// Samples [first, first + n): even i uniform in the domain, odd i a random
// surface vertex plus N(0, sigma^2) per axis (Box-Muller).
__global__ void k_make_samples(float* xyz, uint64_t first, uint64_t n, uint64_t seed, float lox, float loy, float loz,
	float size, const float* verts, uint64_t nverts, float sigma)
{
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const uint64_t i = first + t;
	float p[3];
	if ((i & 1) == 0 || nverts == 0)
	{
		p[0] = lox + size * uniform01(seed, 8 * i);
		p[1] = loy + size * uniform01(seed, 8 * i + 1);
		p[2] = loz + size * uniform01(seed, 8 * i + 2);
	}
	else
	{
		const uint64_t v = uint64_t(uniform01(seed, 8 * i + 3) * float(nverts)) % nverts;
		const float u1 = uniform01(seed, 8 * i + 4), u2 = uniform01(seed, 8 * i + 5), u3 = uniform01(seed, 8 * i + 6),
		            u4 = uniform01(seed, 8 * i + 7);
		const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
		const float g0 = r1 * cospif(2.0f * u2), g1 = r1 * sinpif(2.0f * u2), g2 = r2 * cospif(2.0f * u4);
		p[0] = verts[3 * v] + sigma * g0;
		p[1] = verts[3 * v + 1] + sigma * g1;
		p[2] = verts[3 * v + 2] + sigma * g2;
	}
	xyz[3 * t] = p[0];
	xyz[3 * t + 1] = p[1];
	xyz[3 * t + 2] = p[2];
}

void init_params(std::vector<float>& p, std::vector<float>& fourier, uint64_t seed, float sigma)
{
	std::mt19937_64 rng(seed);
	std::normal_distribution<float> normal(0.0f, 1.0f);
	fourier.resize(mlp::kFreq * 3);
	for (float& f : fourier) f = sigma * normal(rng);
	p.assign(mlp::kParams, 0.0f);
	auto glorot = [&](int off, int fan_out, int fan_in) {
		const float a = std::sqrt(6.0f / float(fan_in + fan_out));
		std::uniform_real_distribution<float> u(-a, a);
		for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = u(rng);
	};
	glorot(mlp::kW1, mlp::kHidden, mlp::kIn);
	glorot(mlp::kW2, mlp::kHidden, mlp::kHidden);
	glorot(mlp::kW3, 1, mlp::kHidden);
}

int upload_params(tg_surrogate* s, const std::vector<float>& p, const std::vector<float>& f)
{
	int rc = s->params.ensure(mlp::kParams * 4);
	if (!rc) rc = s->fourier.ensure(mlp::kFreq * 3 * 4);
	if (!rc) rc = s->adam_m.ensure(mlp::kParams * 4);
	if (!rc) rc = s->adam_v.ensure(mlp::kParams * 4);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemcpy(s->params.ptr, p.data(), mlp::kParams * 4, hipMemcpyHostToDevice), "params upload");
	TG_CHECK_HIP(hipMemcpy(s->fourier.ptr, f.data(), mlp::kFreq * 3 * 4, hipMemcpyHostToDevice), "fourier upload");
	TG_CHECK_HIP(hipMemset(s->adam_m.ptr, 0, mlp::kParams * 4), "adam reset");
	TG_CHECK_HIP(hipMemset(s->adam_v.ptr, 0, mlp::kParams * 4), "adam reset");
	return TG_OK;
}

}  // namespace

extern "C" int tg_surrogate_create(tg_surrogate** out, const float lo[3], float size, float sigma, float clamp,
	uint64_t seed)
{
	if (!out || !lo || !(size > 0.0f) || !(sigma > 0.0f) || !(clamp > 0.0f))
		return fail(TG_ERR_INVALID, "tg_surrogate_create: bad argument");
	int rc = ensure_device();
	if (rc) return rc;
	hipError_t pe = mlp::prepare();
	if (pe != hipSuccess) return hip_fail(pe, "surrogate kernel attributes");
	auto* s = new tg_surrogate;
	for (int i = 0; i < 3; ++i) s->lo[i] = lo[i];
	s->size = size;
	s->sigma = sigma;
	s->clamp = clamp;
	s->seed = seed;
	hipError_t e = hipStreamCreate(&s->stream);
	for (int i = 0; e == hipSuccess && i < 4; ++i) e = hipEventCreate(&s->ev[i]);
	if (e != hipSuccess)
	{
		delete s;
		return hip_fail(e, "surrogate stream/events");
	}
	std::vector<float> p, f;
	init_params(p, f, seed, sigma);
	rc = upload_params(s, p, f);
	if (rc)
	{
		tg_surrogate_destroy(s);
		return rc;
	}
	*out = s;
	return TG_OK;
}

extern "C" void tg_surrogate_destroy(tg_surrogate* s)
{
	if (!s) return;
	if (s->stream) hipStreamSynchronize(s->stream);
	for (DevBuf* b : { &s->params, &s->fourier, &s->adam_m, &s->adam_v, &s->xyz, &s->target, &s->order, &s->partial,
	                   &s->loss_partial, &s->counters, &s->loss_log, &s->tmp })
		b->release();
	for (hipEvent_t e : s->ev)
		if (e) hipEventDestroy(e);
	if (s->stream) hipStreamDestroy(s->stream);
	delete s;
}

extern "C" int tg_surrogate_params(tg_surrogate* s, float* params, float* fourier, int write)
{
	if (!s) return fail(TG_ERR_INVALID, "tg_surrogate_params: null argument");
	if (write)
	{
		if (params) TG_CHECK_HIP(hipMemcpy(s->params.ptr, params, mlp::kParams * 4, hipMemcpyHostToDevice), "params");
		if (fourier) TG_CHECK_HIP(hipMemcpy(s->fourier.ptr, fourier, mlp::kFreq * 3 * 4, hipMemcpyHostToDevice), "fourier");
	}
	else
	{
		if (params) TG_CHECK_HIP(hipMemcpy(params, s->params.ptr, mlp::kParams * 4, hipMemcpyDeviceToHost), "params");
		if (fourier) TG_CHECK_HIP(hipMemcpy(fourier, s->fourier.ptr, mlp::kFreq * 3 * 4, hipMemcpyDeviceToHost), "fourier");
	}
	return TG_OK;
}

extern "C" int tg_surrogate_predict(const tg_surrogate* s, const float* xyz, uint64_t n, float* out)
{
	if (!s || (n && (!xyz || !out))) return fail(TG_ERR_INVALID, "tg_surrogate_predict: null argument");
	if (n == 0) return TG_OK;
	DevBuf din, dout;
	struct Guard { DevBuf* b[2]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &din, &dout } };
	int rc = din.ensure(n * 12);
	if (!rc) rc = dout.ensure(n * 4);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemcpy(din.ptr, xyz, n * 12, hipMemcpyHostToDevice), "points upload");
	mlp::PredictArgs pa{};
	pa.net = surrogate_net(s);
	pa.xyz = din.as<const float>();
	pa.count = n;
	pa.out = dout.as<float>();
	TG_CHECK_HIP(mlp::launch_predict(pa, s->stream), "predict launch");
	TG_CHECK_HIP(hipMemcpyAsync(out, dout.ptr, n * 4, hipMemcpyDeviceToHost, s->stream), "prediction download");
	TG_CHECK_HIP(hipStreamSynchronize(s->stream), "predict");
	return TG_OK;
}

// Gradient of the mean squared error over n points (n a multiple of 32).
extern "C" int tg_surrogate_gradients(tg_surrogate* s, const float* xyz, const float* target, uint64_t n, float* grad,
	float* loss)
{
	if (!s || !xyz || !target || !grad) return fail(TG_ERR_INVALID, "tg_surrogate_gradients: null argument");
	if (n == 0 || n % mlp::kGroup || n >= (1ull << 31))
		return fail(TG_ERR_INVALID, "tg_surrogate_gradients: n must be a positive multiple of 32");
	DevBuf dx, dt, dord, dpart, dloss;
	struct Guard { DevBuf* b[5]; ~Guard() { for (DevBuf* x : b) x->release(); } } guard{ { &dx, &dt, &dord, &dpart, &dloss } };
	const uint32_t blocks = uint32_t(n / mlp::kGroup);
	int rc = dx.ensure(n * 12);
	if (!rc) rc = dt.ensure(n * 4);
	if (!rc) rc = dord.ensure(n * 4);
	if (!rc) rc = dpart.ensure(size_t(blocks) * mlp::kParams * 4);
	if (!rc) rc = dloss.ensure(blocks * 4);
	if (rc) return rc;
	std::vector<uint32_t> ord(n);
	std::iota(ord.begin(), ord.end(), 0u);
	TG_CHECK_HIP(hipMemcpy(dx.ptr, xyz, n * 12, hipMemcpyHostToDevice), "upload");
	TG_CHECK_HIP(hipMemcpy(dt.ptr, target, n * 4, hipMemcpyHostToDevice), "upload");
	TG_CHECK_HIP(hipMemcpy(dord.ptr, ord.data(), n * 4, hipMemcpyHostToDevice), "upload");
	mlp::TrainArgs ta{};
	ta.net = surrogate_net(s);
	ta.xyz = dx.as<const float>();
	ta.target = dt.as<const float>();
	ta.order = dord.as<const uint32_t>();
	ta.batch_offset = 0;
	ta.batch = uint32_t(n);
	ta.batches = 1;
	ta.groups_per_block = 1;
	ta.partial = dpart.as<float>();
	ta.loss_partial = dloss.as<float>();
	TG_CHECK_HIP(mlp::launch_train_grad(ta, blocks, s->stream), "gradient launch");
	std::vector<float> part(size_t(blocks) * mlp::kParams), lp(blocks);
	TG_CHECK_HIP(hipMemcpyAsync(part.data(), dpart.ptr, part.size() * 4, hipMemcpyDeviceToHost, s->stream), "download");
	TG_CHECK_HIP(hipMemcpyAsync(lp.data(), dloss.ptr, blocks * 4, hipMemcpyDeviceToHost, s->stream), "download");
	TG_CHECK_HIP(hipStreamSynchronize(s->stream), "gradients");
	for (int i = 0; i < mlp::kParams; ++i)
	{
		double acc = 0.0;
		for (uint32_t b = 0; b < blocks; ++b) acc += part[size_t(b) * mlp::kParams + i];
		grad[i] = float(acc);
	}
	if (loss)
	{
		double l = 0.0;
		for (float v : lp) l += v;
		*loss = float(l / double(n));
	}
	return TG_OK;
}

extern "C" int tg_surrogate_fit(tg_surrogate* s, const tg_program* exact, uint64_t samples, uint64_t calib_samples,
	float surface_sigma, uint32_t epochs, uint32_t flags, tg_surrogate_stats* stats)
{
	if (!s || !exact) return fail(TG_ERR_INVALID, "tg_surrogate_fit: null argument");
	// two waves per workgroup, a group of 32 points each: 256 workgroups
	const uint32_t batch = 16384, groups_per_block = 2;
	const uint32_t blocks = batch / (mlp::kGroup * groups_per_block);
	const uint64_t batches = samples / batch;
	if (batches == 0 || batches >= (1u << 20) || calib_samples == 0 || epochs == 0)
		return fail(TG_ERR_INVALID, "tg_surrogate_fit: need >= 16384 samples, calibration samples and epochs");
	const uint64_t n_train = batches * batch;
	const uint64_t n_all = n_train + calib_samples;
	auto t0 = std::chrono::steady_clock::now();

	// 1. surface vertices from a coarse exact export of the domain (128 cells per axis)
	tg_mesher* coarse = nullptr;
	int rc = tg_mesher_create(&coarse);
	if (rc) return rc;
	struct MesherGuard { tg_mesher* m; ~MesherGuard() { tg_mesher_destroy(m); } } mg{ coarse };
	tg_grid g{ s->lo[0], s->lo[1], s->lo[2], s->size / 128.0f, s->size / 128.0f, s->size / 128.0f, 128, 128, 128 };
	rc = tg_mesher_export(coarse, exact, &g, flags & TG_FLAG_FAST);
	if (rc) return rc;
	uint64_t nverts = 0, ntris = 0;
	tg_mesher_counts(coarse, &nverts, &ntris);
	const float* verts = mesher_vertices(coarse);

	// 2. samples and exact labels
	rc = s->xyz.ensure(n_all * 12);
	if (!rc) rc = s->target.ensure(n_all * 4);
	if (rc) return rc;
	hipLaunchKernelGGL(k_make_samples, dim3(unsigned((n_all + 255) / 256)), dim3(256), 0, s->stream, s->xyz.as<float>(),
		0ull, n_all, s->seed, s->lo[0], s->lo[1], s->lo[2], s->size, verts, nverts, surface_sigma);
	TG_CHECK_HIP(hipGetLastError(), "sample launch");
	rc = tg_eval_points_device(exact, s->xyz.as<const float>(), n_all, s->target.as<float>(), flags & TG_FLAG_FAST,
		s->stream);
	if (rc) return rc;
	TG_CHECK_HIP(hipStreamSynchronize(s->stream), "sampling");
	auto t1 = std::chrono::steady_clock::now();

	// 3. training: one permutation, batches in order, Adam; one epoch captured as a graph
	{
		std::vector<uint32_t> ord(n_train);
		std::iota(ord.begin(), ord.end(), 0u);
		std::mt19937_64 rng(s->seed ^ 0x5eedull);
		std::shuffle(ord.begin(), ord.end(), rng);
		rc = s->order.ensure(n_train * 4);
		if (rc) return rc;
		TG_CHECK_HIP(hipMemcpy(s->order.ptr, ord.data(), n_train * 4, hipMemcpyHostToDevice), "order upload");
	}
	const uint64_t total_steps = batches * epochs;
	rc = s->partial.ensure(size_t(blocks) * mlp::kParams * 4);
	if (!rc) rc = s->loss_partial.ensure(blocks * 4);
	if (!rc) rc = s->counters.ensure(64);
	if (!rc) rc = s->loss_log.ensure(total_steps * 4);
	if (rc) return rc;
	TG_CHECK_HIP(hipMemset(s->counters.ptr, 0, 64), "counter reset");
	TG_CHECK_HIP(hipMemset(s->adam_m.ptr, 0, mlp::kParams * 4), "adam reset");
	TG_CHECK_HIP(hipMemset(s->adam_v.ptr, 0, mlp::kParams * 4), "adam reset");
	uint64_t* step = s->counters.as<uint64_t>();
	uint64_t* batch_counter = step + 1;
	mlp::TrainArgs ta{};
	ta.net = surrogate_net(s);
	ta.xyz = s->xyz.as<const float>();
	ta.target = s->target.as<const float>();
	ta.order = s->order.as<const uint32_t>();
	ta.batch_counter = batch_counter;
	ta.batches = uint32_t(batches);
	ta.batch = batch;
	ta.groups_per_block = groups_per_block;
	ta.partial = s->partial.as<float>();
	ta.loss_partial = s->loss_partial.as<float>();
	mlp::AdamArgs aa{};
	aa.params = s->params.as<float>();
	aa.m = s->adam_m.as<float>();
	aa.v = s->adam_v.as<float>();
	aa.partial = s->partial.as<const float>();
	aa.loss_partial = s->loss_partial.as<const float>();
	aa.blocks = blocks;
	aa.step = step;
	aa.lr = 2e-3f;
	aa.beta1 = 0.9f;
	aa.beta2 = 0.999f;
	aa.eps = 1e-8f;
	aa.decay_steps = uint32_t(total_steps);
	aa.loss_log = s->loss_log.as<float>();
	aa.log_capacity = uint32_t(total_steps);
	aa.batch = batch;
	// capture one epoch
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	TG_CHECK_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "capture begin");
	hipError_t ce = hipSuccess;
	for (uint64_t b = 0; b < batches && ce == hipSuccess; ++b)
	{
		ce = mlp::launch_train_grad(ta, blocks, s->stream);
		if (ce == hipSuccess) ce = mlp::launch_adam(aa, s->stream);
		if (ce == hipSuccess) ce = mlp::launch_step_advance(step, batch_counter, s->stream);
	}
	hipError_t ee = hipStreamEndCapture(s->stream, &graph);
	if (ce != hipSuccess) return hip_fail(ce, "training capture");
	if (ee != hipSuccess) return hip_fail(ee, "capture end");
	TG_CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
	TG_CHECK_HIP(hipEventRecord(s->ev[0], s->stream), "event");
	for (uint32_t e = 0; e < epochs; ++e) TG_CHECK_HIP(hipGraphLaunch(exec, s->stream), "graph launch");
	TG_CHECK_HIP(hipEventRecord(s->ev[1], s->stream), "event");
	TG_CHECK_HIP(hipStreamSynchronize(s->stream), "training");
	hipGraphExecDestroy(exec);
	hipGraphDestroy(graph);
	float train_ms = 0;
	hipEventElapsedTime(&train_ms, s->ev[0], s->ev[1]);

	// 4. calibration on the held-out samples
	rc = s->tmp.ensure(calib_samples * 4);
	if (rc) return rc;
	mlp::PredictArgs pa{};
	pa.net = surrogate_net(s);
	pa.xyz = s->xyz.as<const float>() + n_train * 3;
	pa.count = calib_samples;
	pa.out = s->tmp.as<float>();
	TG_CHECK_HIP(mlp::launch_predict(pa, s->stream), "calibration predict");
	std::vector<float> pred(calib_samples), truth(calib_samples);
	TG_CHECK_HIP(hipMemcpyAsync(pred.data(), s->tmp.ptr, calib_samples * 4, hipMemcpyDeviceToHost, s->stream), "download");
	TG_CHECK_HIP(hipMemcpyAsync(truth.data(), s->target.as<float>() + n_train, calib_samples * 4, hipMemcpyDeviceToHost,
		s->stream), "download");
	float final_loss = 0.0f;
	TG_CHECK_HIP(hipMemcpyAsync(&final_loss, s->loss_log.as<float>() + total_steps - 1, 4, hipMemcpyDeviceToHost,
		s->stream), "download");
	TG_CHECK_HIP(hipStreamSynchronize(s->stream), "calibration");
	double max_res = 0.0, sum_res = 0.0;
	for (uint64_t i = 0; i < calib_samples; ++i)
	{
		const float t = std::fmin(std::fmax(truth[i], -s->clamp), s->clamp);
		const double r = std::fabs(double(pred[i]) - double(t));
		max_res = std::max(max_res, r);
		sum_res += r;
	}
	s->calib_pred = std::move(pred);
	s->calib_dist = std::move(truth);
	s->mean_residual = float(sum_res / double(calib_samples));
	s->calibrated = true;
	s->margin = -1.0f;
	s->cache_r[0] = s->cache_r[1] = -1.0f;
	auto t2 = std::chrono::steady_clock::now();
	tg_surrogate_stats st{};
	st.sample_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	st.train_ms = train_ms;
	st.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
	st.samples = n_train;
	st.steps = uint32_t(total_steps);
	st.final_loss = final_loss;
	st.calib_samples = calib_samples;
	st.margin = float(max_res);
	st.mean_residual = s->mean_residual;
	st.surface_vertices = nverts;
	s->stats = st;
	if (stats) *stats = st;
	return TG_OK;
}

extern "C" int tg_surrogate_set_margin(tg_surrogate* s, float margin)
{
	if (!s || !(margin >= 0.0f)) return fail(TG_ERR_INVALID, "tg_surrogate_set_margin: bad argument");
	s->margin = margin;
	return TG_OK;
}

extern "C" int tg_surrogate_get_stats(const tg_surrogate* s, tg_surrogate_stats* out)
{
	if (!s || !out) return fail(TG_ERR_INVALID, "tg_surrogate_get_stats: null argument");
	*out = s->stats;
	return TG_OK;
}

extern "C" int tg_surrogate_loss_log(const tg_surrogate* s, float* out, uint64_t capacity, uint64_t* count)
{
	if (!s || !count) return fail(TG_ERR_INVALID, "tg_surrogate_loss_log: null argument");
	*count = s->stats.steps;
	if (out && capacity)
	{
		const uint64_t k = std::min<uint64_t>(capacity, s->stats.steps);
		TG_CHECK_HIP(hipMemcpy(out, s->loss_log.ptr, k * 4, hipMemcpyDeviceToHost), "loss download");
	}
	return TG_OK;
}
