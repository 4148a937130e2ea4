// Radiosity solver forms (radiosity.hip): many right-hand sides at once over
// the link structure of a DeviceScene, forward and reverse mode.
#pragma once

#include <stdint.h>

#include <vector>

#include "device.h"

namespace lgx {

enum RadForm {
	kSparseF64 = 0, // gather over the links, FP64 (the accuracy reference on the GPU)
	kSparseF32 = 1, // gather over the links, FP32
	kDenseF32 = 2,  // dense H = F in FP32, plain FP32 FMAs
	kDenseBF16 = 3, // H in BF16 (the tiles that hold links), MFMA with FP32 accumulation; outliers in FP32
	kDenseF16 = 4,  // H in FP16 (power-of-two scaled, the tiles that hold links), MFMA with FP32 accumulation;
	                // outliers in FP32
	kTileF32 = 5,   // H in FP32 as the same stored tiles, FP32 matrix-core MFMA (the fair FP32 form)
	kF16x3 = 6,     // FP16 hi + lo planes of H and the operand, three MFMA products per tile, every bounce
	kBF16x3 = 7,    // the same in BF16
};
constexpr int kNumForms = 8;

struct RadTimes {
	double build_ms = 0;      // dense matrix construction
	double series_ms = 0;     // the bounces (all right-hand sides)
	double per_bounce_ms = 0; // series_ms / bounces (reverse sweep / bounces for solve_grad)
	double refine_ms = 0;     // refinement sweeps (the reverse sweep for solve_grad)
};

class RadSolver {
public:
	explicit RadSolver(DeviceScene& dev);
	~RadSolver();
	// E: N x R row-major emitted light; column c uses diffuse channel c % 3.
	void set_problem(const double* E, uint32_t R, const float* diffuse3, const float* area);
	void build_dense(int form);
	// total (N x R row-major) = sum_{k=0..bounces} K^k E; dense forms may add
	// `refine` refinement sweeps with an FP64 residual.
	void solve(int form, int bounces, int refine, double* total);
	// Reverse mode: for the upstream gradient g = dL/dtotal (N x R), the
	// total again, lam = dL/dE and dalb = dL/d diffuse per column (the
	// albedo of column c's channel; sum the columns of a channel).
	void solve_grad(int form, int bounces, const double* g, double* total, double* lam, double* dalb);
	RadTimes times;
	double build_ms = 0;
	uint32_t N = 0, Np = 0, R = 0;
	uint64_t ntiles() const { return ntiles_; } // stored 256 x 32 tiles of the 16-bit forms
	uint64_t noutliers() const { return nout_; } // factors the 16-bit forms apply in FP32
	bool ntiles_built() const { return tiles_built_; }

private:
	void release();
	void build_tiles();
	void build_outliers();
	void ensure_work();
	void ensure_part();
	void ensure_perm_colors();
	uint32_t ksplit(uint32_t rows_per_block, uint32_t min_blocks) const;
	// chain: the next call is apply_dense(form, yT, .., split, pre), so the
	// reduction prepares its operand (BF16) or its column scales (FP16)
	void apply_dense(int form, const float* xT, float* yT, bool split, bool pre, bool chain = false);
	void apply_tiles_f32(const float* xT, float* yT, bool pre);
	void apply_op(int form, const double* x, double* y, bool pre);
	DeviceScene& dev;
	void* s_ = nullptr;
	float* g32_ = nullptr;
	uint16_t* g16_ = nullptr;   // 16-bit matrix (hi plane)
	uint16_t* g16lo_ = nullptr; // remainder plane for the compensated product
	float* g32t_ = nullptr;     // FP32 tile store of form kTileF32 (same tiles as the 16-bit forms)
	int32_t* tmap_ = nullptr;   // tile (256-row block, k stage) -> place in the 16-bit store, or -1
	uint32_t* tks_ = nullptr;   // k stage of each stored tile (row blocks in order, stages ascending)
	uint32_t* chunks_ = nullptr; // work segments {row block, first tile, end tile, slot}
	uint32_t* segoff_ = nullptr; // first segment of each persistent block (nblocks_ + 1 entries)
	uint32_t* nslot_ = nullptr;  // segments (partial-sum slots) per row block
	uint64_t ntiles_ = 0, nchunks_ = 0;
	uint32_t nblocks_ = 1;  // persistent blocks of the tile products (one per compute unit)
	uint32_t maxslot_ = 0;  // most segments of one row block
	uint32_t part_slots_ = 0; // slots part_ holds
	bool tiles_built_ = false;
	uint32_t* perm_ = nullptr; // position -> patch (Morton order, padding last)
	uint32_t* pinv_ = nullptr; // patch -> position
	int built_form_ = -1;
	uint32_t ncu_ = 256; // compute units of the device
	float gscale_ = 1.0f;
	double* E_ = nullptr;
	float* diff_ = nullptr;
	float* area_ = nullptr;
	// dense-product work buffers
	float* part_ = nullptr;
	float* xq_ = nullptr;
	uint16_t* xqlo_ = nullptr;
	float* xT_ = nullptr;
	float* yT_ = nullptr;
	float* rm32a_ = nullptr;
	float* rm32b_ = nullptr;
	float* rmx_ = nullptr;  // sparse FP32 forms: the operand (x diffuse) area prepared once per product
	double* rmx64_ = nullptr; // the same for the FP64 form
	unsigned* cmax_ = nullptr; // column maxima of the FP16 operand (its scales): two planes of kRPad
	int cmax_cur_ = 0;         // the plane of the current operand
	const float* chain_x_ = nullptr; // operand (or its scales) already prepared from this vector ..
	int chain_key_ = -1;             // .. for this storage / split / pre
	float* dperm_ = nullptr;   // diffuse at each Morton position (3 per position, padding zero)
	float* aperm_ = nullptr;   // area at each Morton position
	bool pcol_ok_ = false;     // dperm_ / aperm_ hold the current problem's colours
	// outlier split of the 16-bit forms: factors >= othr_ are applied in FP32
	// by the reduction (CSR by Morton position), the rest go to the tiles
	uint32_t* orow_ = nullptr;
	uint32_t* ocol_ = nullptr;
	float* oval_ = nullptr;
	uint64_t nout_ = 0;
	float othr_ = 0.0f;
	int bulk_exp_ = -1000; // largest factor exponent left in the tiles
	bool out_built_ = false;
};

} // namespace lgx
