// Fourier-feature MLP surrogate of a signed distance field (no reference
// counterpart): 3 -> 2F Fourier features -> 64 SiLU -> 64 SiLU -> 1, BF16
// operands on MFMA with FP32 accumulation.  Used only to propose which bricks
// of a grid may hold the surface; every output sample and vertex still comes
// from the exact evaluator.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {
namespace mlp {

constexpr int kFreq = 48;                 // Fourier frequencies
constexpr int kIn = 2 * kFreq;            // 96 features (sin, cos)
constexpr int kHidden = 64;
constexpr int kGroup = 32;                // points per wave pass

// FP32 parameter vector layout (offsets in floats).
constexpr int kW1 = 0;                          // [64][96]  out x in
constexpr int kB1 = kW1 + kHidden * kIn;        // [64]
constexpr int kW2 = kB1 + kHidden;              // [64][64]  out x in
constexpr int kB2 = kW2 + kHidden * kHidden;    // [64]
constexpr int kW3 = kB2 + kHidden;              // [64]
constexpr int kB3 = kW3 + kHidden;              // [1]
constexpr int kParams = kB3 + 1;                // 10433

struct Net
{
	const float* params;      // FP32 master weights (kParams)
	const float* fourier;     // [kFreq][3] frequencies (input scaled to [-1, 1])
	float lo[3];              // domain origin
	float inv_half;           // 2 / domain size: x' = (p - lo) * inv_half - 1
	float clamp;              // training target: clamp(d, -clamp, clamp)
};

struct PredictArgs
{
	Net net;
	const float* xyz;         // points (or null: cube-brick centres of a grid)
	uint64_t count;
	float* out;               // predicted distance per point (or null)
	// brick flagging (xyz == null): cube bricks of 8^3 cells of a grid
	float gx, gy, gz, dx, dy, dz;
	uint32_t cbx, cby, cbz;   // cube bricks per axis
	float threshold;          // keep a brick unless |prediction at its centre| > threshold
	float radius;             // brick half-diagonal: bricks reaching outside the domain are flagged
	uint8_t* flags;
};

struct TrainArgs
{
	Net net;
	const float* xyz;         // training points
	const float* target;      // exact distances
	const uint32_t* order;    // sample permutation
	uint64_t batch_offset;    // first sample of the batch in `order` (or a device counter)
	const uint64_t* batch_counter;  // if set: batch index -> offset = (*counter % batches) * batch
	uint32_t batches;
	uint32_t batch;           // samples per batch (multiple of 32)
	uint32_t groups_per_block;
	float* partial;           // [blocks][kParams] gradients of the mean squared error
	float* loss_partial;      // [blocks] sum of squared errors
};

struct AdamArgs
{
	float* params;
	float* m;
	float* v;
	const float* partial;     // [blocks][kParams]
	const float* loss_partial;
	uint32_t blocks;
	uint64_t* step;           // device step counter (incremented by the step kernel)
	float lr, beta1, beta2, eps;
	uint32_t decay_steps;     // cosine decay of lr to 5% over this many steps
	float* loss_log;          // [steps] mean loss per step (or null)
	uint32_t log_capacity;
	uint32_t batch;
};

// Sets the kernels' LDS limits; call once before the first launch (and
// before any graph capture).
hipError_t prepare();
hipError_t launch_predict(const PredictArgs& a, hipStream_t s);
hipError_t launch_train_grad(const TrainArgs& a, uint32_t blocks, hipStream_t s);
hipError_t launch_adam(const AdamArgs& a, hipStream_t s);
hipError_t launch_step_advance(uint64_t* step, uint64_t* batch_counter, hipStream_t s);

}  // namespace mlp
}  // namespace tg
