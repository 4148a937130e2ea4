// Radiosity solver forms (radiosity.hip): many right-hand sides at once over
// the link structure of a DeviceScene.
#pragma once

#include <stdint.h>

#include "device.h"

namespace lgx {

enum RadForm {
	kSparseF64 = 0, // gather over the links, FP64 (the accuracy reference on the GPU)
	kSparseF32 = 1, // gather over the links, FP32
	kDenseF32 = 2,  // dense G = F diag(area) in FP32, plain FP32 FMAs
	kDenseBF16 = 3, // dense G in BF16, MFMA with FP32 accumulation
	kDenseF16 = 4,  // dense G in FP16 (power-of-two scaled), MFMA with FP32 accumulation
};

struct RadTimes {
	double build_ms = 0;      // dense matrix construction
	double series_ms = 0;     // the bounces (all right-hand sides)
	double per_bounce_ms = 0; // series_ms / bounces
	double refine_ms = 0;     // refinement sweeps
};

class RadSolver {
public:
	explicit RadSolver(DeviceScene& dev);
	~RadSolver();
	// E: N x R row-major emitted light; column c uses diffuse channel c % 3.
	void set_problem(const double* E, uint32_t R, const float* diffuse3, const float* area);
	void build_dense(int form);
	// total (N x R row-major) = sum_{k=0..bounces} K^k E; dense forms may add
	// `refine` refinement sweeps with the FP64 link residual.
	void solve(int form, int bounces, int refine, double* total);
	RadTimes times;
	double build_ms = 0;
	uint32_t N = 0, Np = 0, R = 0;

private:
	void release();
	uint32_t ksplit(uint32_t rows_per_block, uint32_t min_blocks) const;
	void apply_dense(int form, const float* xT, float* yT, float* part, void* xq, uint16_t* xqlo, bool split);
	void column_scales(int form, const float* xT);
	DeviceScene& dev;
	void* s_ = nullptr;
	float* g32_ = nullptr;
	uint16_t* g16_ = nullptr;   // 16-bit matrix (hi plane)
	uint16_t* g16lo_ = nullptr; // remainder plane for the compensated product
	int built_form_ = -1;
	uint32_t ncu_ = 256; // compute units of the device
	float gscale_ = 1.0f;
	double* E_ = nullptr;
	float* diff_ = nullptr;
	float* area_ = nullptr;
	float* cscale_ = nullptr;
	float* cunscale_ = nullptr;
};

} // namespace lgx
