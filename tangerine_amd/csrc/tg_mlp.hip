// Fourier-feature MLP surrogate: forward (prediction, brick flagging) and
// training (forward + backward + Adam) kernels for gfx950.
//
// Matrix products use v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation).  Operands live in LDS with the reduction index contiguous:
// an A operand is a row-major [M][K] array, a B operand is an [N][K] array
// (each output column's K values contiguous), so that every lane fetches its
// eight K values with one 16-byte read.  For a 16x16 output tile the lane
// layout is
//   A[i][k]: lane = i + 16 * (k / 8), element k % 8
//   B[k][j]: lane = j + 16 * (k / 8), element k % 8
//   D[i][j]: lane = j + 16 * (i / 4), element i % 4
// A wave processes 32 points at a time; hidden activations are written to LDS
// both point-major ([point][unit], the next layer's B operand) and unit-major
// ([unit][point], operands of the weight-gradient products whose reduction
// runs over points).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "tg_mlp.h"

namespace tg {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f)
{
	// round to nearest even (finite inputs)
	uint32_t u = __float_as_uint(f);
	u += 0x7fffu + ((u >> 16) & 1u);
	return uint16_t(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ float silu_grad(float z)
{
	const float s = 1.0f / (1.0f + __expf(-z));
	return s * (1.0f + z * (1.0f - s));
}

// acc += A[16m.., k-block kb] * B[k-block kb, 16n..]
__device__ __forceinline__ f32x4 mma(const uint16_t* A, int lda, int m, const uint16_t* B, int ldb, int n, int kb,
	f32x4 acc, int lane)
{
	const int r = lane & 15, q = lane >> 4;
	const uint4 av = *reinterpret_cast<const uint4*>(A + (16 * m + r) * lda + 32 * kb + 8 * q);
	const uint4 bv = *reinterpret_cast<const uint4*>(B + (16 * n + r) * ldb + 32 * kb + 8 * q);
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc,
		0, 0, 0);
}

// Shared weights of one workgroup (BF16 operands, FP32 vectors).
struct Weights
{
	uint16_t w1[kHidden * kIn];       // [out][in]
	uint16_t w2[kHidden * kHidden];   // [out][in]
	uint16_t w2t[kHidden * kHidden];  // [in][out] (backward)
	float b1[kHidden], b2[kHidden], w3[kHidden];
	float b3;
	float fourier[kFreq * 3];
};

__device__ void load_weights(Weights& w, const Net& net, bool transposed, int tid, int nthreads)
{
	const float* p = net.params;
	for (int i = tid; i < kHidden * kIn; i += nthreads) w.w1[i] = to_bf16(p[kW1 + i]);
	for (int i = tid; i < kHidden * kHidden; i += nthreads)
	{
		const float v = p[kW2 + i];
		w.w2[i] = to_bf16(v);
		if (transposed)
		{
			const int o = i / kHidden, in = i % kHidden;
			w.w2t[in * kHidden + o] = to_bf16(v);
		}
	}
	for (int i = tid; i < kHidden; i += nthreads)
	{
		w.b1[i] = p[kB1 + i];
		w.b2[i] = p[kB2 + i];
		w.w3[i] = p[kW3 + i];
	}
	for (int i = tid; i < kFreq * 3; i += nthreads) w.fourier[i] = net.fourier[i];
	if (tid == 0) w.b3 = p[kB3];
}

// Per-wave activation buffers (BF16).
struct Acts
{
	uint16_t fpm[kGroup * kIn];       // [pt][feature]
	uint16_t fum[kIn * kGroup];       // [feature][pt]
	uint16_t a1pm[kGroup * kHidden];  // [pt][unit]
	uint16_t a1um[kHidden * kGroup];  // [unit][pt]
	uint16_t dz2pm[kGroup * kHidden];
	uint16_t dz2um[kHidden * kGroup];
	uint16_t dz1um[kHidden * kGroup];
};

// Fourier features of point `pt` (lanes 0..31: sin, lanes 32..63: cos of the
// same 48 phases), x' = (p - lo) * inv_half - 1, phase in revolutions.
__device__ __forceinline__ void features(const Net& net, const Weights& w, uint16_t* fpm, uint16_t* fum, float px,
	float py, float pz, int lane)
{
	const int pt = lane & 31;
	const bool cosine = lane >= 32;
	const float x = (px - net.lo[0]) * net.inv_half - 1.0f;
	const float y = (py - net.lo[1]) * net.inv_half - 1.0f;
	const float z = (pz - net.lo[2]) * net.inv_half - 1.0f;
	for (int f = 0; f < kFreq; ++f)
	{
		const float ph = (w.fourier[3 * f] * x + w.fourier[3 * f + 1] * y) + w.fourier[3 * f + 2] * z;
		const float v = cosine ? __builtin_amdgcn_cosf(ph) : __builtin_amdgcn_sinf(ph);
		const int k = (cosine ? kFreq : 0) + f;
		const uint16_t b = to_bf16(v);
		fpm[pt * kIn + k] = b;
		if (fum) fum[k * kGroup + pt] = b;
	}
}

// Forward pass of 32 points.  Returns z1, z2 (pre-activations, D layout:
// z[m][n][i] = unit 16m + 4q + i, point 16n + (lane & 15)) and y for the
// lane's two points (n = 0, 1).
struct Fwd
{
	float z1[4][2][4];
	float z2[4][2][4];
	float y[2];
};

__device__ __forceinline__ void forward(const Weights& w, const uint16_t* fpm, uint16_t* a1pm, uint16_t* a1um, Fwd& f,
	int lane)
{
	const int r = lane & 15, q = lane >> 4;
	__builtin_amdgcn_s_waitcnt(0);
	__builtin_amdgcn_wave_barrier();
#pragma unroll
	for (int m = 0; m < 4; ++m)
#pragma unroll
		for (int n = 0; n < 2; ++n)
		{
			f32x4 acc = { 0, 0, 0, 0 };
#pragma unroll
			for (int kb = 0; kb < kIn / 32; ++kb) acc = mma(w.w1, kIn, m, fpm, kIn, n, kb, acc, lane);
#pragma unroll
			for (int i = 0; i < 4; ++i) f.z1[m][n][i] = acc[i] + w.b1[16 * m + 4 * q + i];
		}
#pragma unroll
	for (int m = 0; m < 4; ++m)
#pragma unroll
		for (int n = 0; n < 2; ++n)
		{
			const int pt = 16 * n + r;
			uint16_t h[4];
#pragma unroll
			for (int i = 0; i < 4; ++i) h[i] = to_bf16(silu(f.z1[m][n][i]));
			const int u0 = 16 * m + 4 * q;
			uint2 packed = make_uint2(uint32_t(h[0]) | (uint32_t(h[1]) << 16), uint32_t(h[2]) | (uint32_t(h[3]) << 16));
			*reinterpret_cast<uint2*>(a1pm + pt * kHidden + u0) = packed;
			if (a1um)
#pragma unroll
				for (int i = 0; i < 4; ++i) a1um[(u0 + i) * kGroup + pt] = h[i];
		}
	__builtin_amdgcn_s_waitcnt(0);
	__builtin_amdgcn_wave_barrier();
	float part[2] = { 0.0f, 0.0f };
#pragma unroll
	for (int m = 0; m < 4; ++m)
#pragma unroll
		for (int n = 0; n < 2; ++n)
		{
			f32x4 acc = { 0, 0, 0, 0 };
#pragma unroll
			for (int kb = 0; kb < kHidden / 32; ++kb) acc = mma(w.w2, kHidden, m, a1pm, kHidden, n, kb, acc, lane);
#pragma unroll
			for (int i = 0; i < 4; ++i)
			{
				const int u = 16 * m + 4 * q + i;
				const float z = acc[i] + w.b2[u];
				f.z2[m][n][i] = z;
				part[n] += w.w3[u] * silu(z);
			}
		}
#pragma unroll
	for (int n = 0; n < 2; ++n)
	{
		float v = part[n];
		v += __shfl_xor(v, 16);
		v += __shfl_xor(v, 32);
		f.y[n] = v + w.b3;
	}
}

// ---- prediction / brick flags --------------------------------------------------

struct ActsFwd
{
	uint16_t fpm[kGroup * kIn];
	uint16_t a1pm[kGroup * kHidden];
};

struct PredictShared
{
	Weights w;
	ActsFwd a[4];
};

__global__ void __launch_bounds__(256) k_predict(PredictArgs p)
{
	extern __shared__ __align__(16) unsigned char raw[];
	PredictShared& sh = *reinterpret_cast<PredictShared*>(raw);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	load_weights(sh.w, p.net, false, threadIdx.x, blockDim.x);
	__syncthreads();
	ActsFwd& a = sh.a[wave];
	const uint64_t groups = (p.count + kGroup - 1) / kGroup;
	const uint64_t stride = uint64_t(gridDim.x) * (blockDim.x >> 6);
	for (uint64_t g = uint64_t(blockIdx.x) * (blockDim.x >> 6) + wave; g < groups; g += stride)
	{
		const uint64_t i = g * kGroup + (lane & 31);
		const uint64_t j = i < p.count ? i : p.count - 1;
		float px, py, pz;
		if (p.xyz)
		{
			px = p.xyz[3 * j]; py = p.xyz[3 * j + 1]; pz = p.xyz[3 * j + 2];
		}
		else
		{
			// centre of cube brick j: cells [8b, 8b + 8) -> sample span [8b, 8b + 8]
			const uint32_t bx = uint32_t(j % p.cbx), by = uint32_t((j / p.cbx) % p.cby), bz = uint32_t(j / (uint64_t(p.cbx) * p.cby));
			px = p.gx + (float(bx * 8) + 4.0f) * p.dx;
			py = p.gy + (float(by * 8) + 4.0f) * p.dy;
			pz = p.gz + (float(bz * 8) + 4.0f) * p.dz;
		}
		__builtin_amdgcn_wave_barrier();
		features(p.net, sh.w, a.fpm, nullptr, px, py, pz, lane);
		Fwd f;
		forward(sh.w, a.fpm, a.a1pm, nullptr, f, lane);
		if (lane < 32)
		{
			const int n = lane >> 4;  // lanes 0..15 hold points 0..15 (n=0), 16..31 hold 16..31 (n=1)
			const uint64_t k = g * kGroup + uint64_t(lane);
			const float y = n ? f.y[1] : f.y[0];
			if (k < p.count)
			{
				if (p.out) p.out[k] = y;
				if (p.flags)
				{
					// brick centre of this lane's point (lanes 0..31 own points 0..31)
					const float ux = fabsf((px - p.net.lo[0]) * p.net.inv_half - 1.0f);
					const float uy = fabsf((py - p.net.lo[1]) * p.net.inv_half - 1.0f);
					const float uz = fabsf((pz - p.net.lo[2]) * p.net.inv_half - 1.0f);
					const float reach = p.radius * p.net.inv_half;
					const bool outside = fmaxf(ux, fmaxf(uy, uz)) + reach > 1.0f;
					// keep unless the prediction certifies distance (NaN keeps)
					p.flags[k] = (outside || !(fabsf(y) > p.threshold)) ? 1 : 0;
				}
			}
		}
	}
}

// ---- training ----------------------------------------------------------------------

// Two waves per workgroup: the weights and the gradient accumulator are
// shared, each wave has its own activations (135 KiB of LDS, one workgroup
// per CU), so a 16384-point batch (512 groups of 32) fills 256 CUs with
// every wave taking one group.
constexpr int kTrainWaves = 2;

struct TrainShared
{
	Weights w;
	Acts a[kTrainWaves];
	float acc[kParams];
	float loss;
};

__device__ __forceinline__ void lds_add(float* p, float v) { atomicAdd(p, v); }

// Sum over the 16 lanes that share q (i.e. over the lane's points), result in
// every lane.
__device__ __forceinline__ float sum16(float v)
{
	v += __shfl_xor(v, 1);
	v += __shfl_xor(v, 2);
	v += __shfl_xor(v, 4);
	v += __shfl_xor(v, 8);
	return v;
}

__global__ void __launch_bounds__(64 * kTrainWaves) k_train_grad(TrainArgs t)
{
	extern __shared__ __align__(16) unsigned char raw[];
	TrainShared& sh = *reinterpret_cast<TrainShared*>(raw);
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = lane & 15, q = lane >> 4;
	load_weights(sh.w, t.net, true, tid, 64 * kTrainWaves);
	for (int i = tid; i < kParams; i += 64 * kTrainWaves) sh.acc[i] = 0.0f;
	if (tid == 0) sh.loss = 0.0f;
	__syncthreads();
	uint64_t base = t.batch_offset;
	if (t.batch_counter) base = (*t.batch_counter % t.batches) * uint64_t(t.batch);
	const float scale = 2.0f / float(t.batch);
	Acts& a = sh.a[wave];
	for (uint32_t gi = uint32_t(wave); gi < t.groups_per_block; gi += kTrainWaves)
	{
		const uint64_t g = uint64_t(blockIdx.x) * t.groups_per_block + gi;
		const uint64_t s0 = base + g * kGroup;
		const uint32_t idx = t.order[s0 + (lane & 31)];
		__builtin_amdgcn_wave_barrier();
		features(t.net, sh.w, a.fpm, a.fum, t.xyz[3 * idx], t.xyz[3 * idx + 1], t.xyz[3 * idx + 2], lane);
		Fwd f;
		forward(sh.w, a.fpm, a.a1pm, a.a1um, f, lane);
		// targets of the lane's two points, dL/dy of the mean squared error
		float dy[2];
#pragma unroll
		for (int n = 0; n < 2; ++n)
		{
			const uint32_t id = t.order[s0 + 16 * n + r];
			const float target = fminf(fmaxf(t.target[id], -t.net.clamp), t.net.clamp);
			const float e = f.y[n] - target;
			dy[n] = scale * e;
			if (q == 0) lds_add(&sh.loss, e * e);
		}
		// output layer
		float gb3 = 0.0f;
#pragma unroll
		for (int m = 0; m < 4; ++m)
#pragma unroll
			for (int i = 0; i < 4; ++i)
			{
				const int u = 16 * m + 4 * q + i;
				float gw3 = 0.0f, gb2 = 0.0f;
				uint16_t dz[2];
#pragma unroll
				for (int n = 0; n < 2; ++n)
				{
					const float z = f.z2[m][n][i];
					gw3 += dy[n] * silu(z);
					const float d = dy[n] * sh.w.w3[u] * silu_grad(z);
					gb2 += d;
					dz[n] = to_bf16(d);
					a.dz2um[u * kGroup + 16 * n + r] = dz[n];
					a.dz2pm[(16 * n + r) * kHidden + u] = dz[n];
				}
				gw3 = sum16(gw3);
				gb2 = sum16(gb2);
				if (r == 0)
				{
					lds_add(&sh.acc[kW3 + u], gw3);
					lds_add(&sh.acc[kB2 + u], gb2);
				}
			}
		gb3 = sum16(dy[0] + dy[1]);
		if (lane == 0) lds_add(&sh.acc[kB3], gb3);
		__builtin_amdgcn_s_waitcnt(0);
		__builtin_amdgcn_wave_barrier();
		// dW2[o][i] = sum_pt dz2[o][pt] a1[i][pt]
#pragma unroll
		for (int m = 0; m < 4; ++m)
#pragma unroll
			for (int n = 0; n < 4; ++n)
			{
				f32x4 acc = { 0, 0, 0, 0 };
				acc = mma(a.dz2um, kGroup, m, a.a1um, kGroup, n, 0, acc, lane);
#pragma unroll
				for (int i = 0; i < 4; ++i) lds_add(&sh.acc[kW2 + (16 * m + 4 * q + i) * kHidden + 16 * n + r], acc[i]);
			}
		// da1[u][pt] = sum_o W2[o][u] dz2[o][pt]; dz1 = da1 * silu'(z1)
#pragma unroll
		for (int m = 0; m < 4; ++m)
		{
			float gb1[4] = { 0, 0, 0, 0 };
#pragma unroll
			for (int n = 0; n < 2; ++n)
			{
				f32x4 acc = { 0, 0, 0, 0 };
#pragma unroll
				for (int kb = 0; kb < kHidden / 32; ++kb) acc = mma(sh.w.w2t, kHidden, m, a.dz2pm, kHidden, n, kb, acc, lane);
#pragma unroll
				for (int i = 0; i < 4; ++i)
				{
					const float d = acc[i] * silu_grad(f.z1[m][n][i]);
					gb1[i] += d;
					a.dz1um[(16 * m + 4 * q + i) * kGroup + 16 * n + r] = to_bf16(d);
				}
			}
#pragma unroll
			for (int i = 0; i < 4; ++i)
			{
				const float v = sum16(gb1[i]);
				if (r == 0) lds_add(&sh.acc[kB1 + 16 * m + 4 * q + i], v);
			}
		}
		__builtin_amdgcn_s_waitcnt(0);
		__builtin_amdgcn_wave_barrier();
		// dW1[o][k] = sum_pt dz1[o][pt] f[k][pt]
#pragma unroll
		for (int m = 0; m < 4; ++m)
#pragma unroll
			for (int n = 0; n < kIn / 16; ++n)
			{
				f32x4 acc = { 0, 0, 0, 0 };
				acc = mma(a.dz1um, kGroup, m, a.fum, kGroup, n, 0, acc, lane);
#pragma unroll
				for (int i = 0; i < 4; ++i) lds_add(&sh.acc[kW1 + (16 * m + 4 * q + i) * kIn + 16 * n + r], acc[i]);
			}
	}
	__syncthreads();
	float* out = t.partial + size_t(blockIdx.x) * kParams;
	for (int i = tid; i < kParams; i += 64 * kTrainWaves) out[i] = sh.acc[i];
	if (tid == 0 && t.loss_partial) t.loss_partial[blockIdx.x] = sh.loss;
}

// A workgroup per 64 parameters: wave w sums the block partials b = w mod 4
// for its 64 parameters (lane = parameter), the four sums are added in a
// fixed order (deterministic), then Adam with bias correction and cosine
// learning-rate decay to 5%.  Workgroup 0 also sums the loss partials.
constexpr int kAdamWaves = 4;

__global__ void __launch_bounds__(64 * kAdamWaves) k_adam(AdamArgs a)
{
	__shared__ float part[kAdamWaves][64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t i = blockIdx.x * 64 + uint32_t(lane);
	const uint64_t step = *a.step + 1;
	float acc[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
	if (i < uint32_t(kParams))
	{
		// four independent chains keep several loads in flight
		uint32_t b = uint32_t(wave), k = 0;
		for (; b < a.blocks; b += kAdamWaves, k = (k + 1) & 3u) acc[k] += a.partial[size_t(b) * kParams + i];
	}
	part[wave][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
	__syncthreads();
	if (wave == 0 && i < uint32_t(kParams))
	{
		const float g = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
		const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
		const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
		a.m[i] = m;
		a.v[i] = v;
		const float c1 = 1.0f - powf(a.beta1, float(step));
		const float c2 = 1.0f - powf(a.beta2, float(step));
		const float frac = a.decay_steps ? fminf(1.0f, float(step) / float(a.decay_steps)) : 0.0f;
		const float lr = a.lr * (0.05f + 0.95f * 0.5f * (1.0f + cospif(frac)));
		a.params[i] -= lr * (m / c1) / (sqrtf(v / c2) + a.eps);
	}
	if (blockIdx.x == 0 && a.loss_log && step - 1 < a.log_capacity)
	{
		__syncthreads();
		float l = 0.0f;
		for (uint32_t b = threadIdx.x; b < a.blocks; b += 64 * kAdamWaves) l += a.loss_partial[b];
		part[wave][lane] = l;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			float t = 0.0f;
			for (int w = 0; w < kAdamWaves; ++w)
				for (int j = 0; j < 64; ++j) t += part[w][j];
			a.loss_log[step - 1] = t / float(a.batch);
		}
	}
}

__global__ void k_step_advance(uint64_t* step, uint64_t* batch_counter)
{
	*step += 1;
	if (batch_counter) *batch_counter += 1;
}

}  // namespace

hipError_t prepare()
{
	hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_predict),
		hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(PredictShared)));
	if (e == hipSuccess)
		e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_train_grad), hipFuncAttributeMaxDynamicSharedMemorySize,
			int(sizeof(TrainShared)));
	return e;
}

hipError_t launch_predict(const PredictArgs& p, hipStream_t s)
{
	if (p.count == 0) return hipSuccess;
	const uint64_t groups = (p.count + kGroup - 1) / kGroup;
	const unsigned blocks = unsigned(std::min<uint64_t>((groups + 3) / 4, 4096));
	hipLaunchKernelGGL(k_predict, dim3(blocks), dim3(256), sizeof(PredictShared), s, p);
	return hipGetLastError();
}

hipError_t launch_train_grad(const TrainArgs& t, uint32_t blocks, hipStream_t s)
{
	hipLaunchKernelGGL(k_train_grad, dim3(blocks), dim3(64 * kTrainWaves), sizeof(TrainShared), s, t);
	return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t s)
{
	hipLaunchKernelGGL(k_adam, dim3((kParams + 63) / 64), dim3(64 * kAdamWaves), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_step_advance(uint64_t* step, uint64_t* batch_counter, hipStream_t s)
{
	hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, s, step, batch_counter);
	return hipGetLastError();
}

}  // namespace mlp
}  // namespace tg
