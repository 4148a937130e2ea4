// Radiosity solver forms over the link structure built by rad_links().
//
// The reference's bounce loop (lighter.cpp:767-797, 1104-1123) computes the
// truncated series T = sum_{k=0..b} K^k E per colour channel, with
// K = F diag(area) diag(diffuse_c) and F the symmetric matrix of link form
// factors (zero for pairs without a link).  Here it is evaluated for many
// right-hand sides at once (columns of E; column c uses channel c % 3):
//
//  * sparse gathers over the symmetric CSR links, FP32 or FP64;
//  * dense H = F stored FP32 and multiplied with plain FP32 FMAs;
//  * H stored BF16 or FP16 (power-of-two scaled) as the 256 x 32 tiles that
//    hold links, the patches in Morton order, and multiplied with gfx950
//    MFMA instructions accumulating in FP32;
//  * the MFMA form followed by iterative refinement sweeps whose residual
//    r_k = K z_{k-1} - z_k uses the exact link operator in FP64.
//
// Layouts: dense FP32 H is row-major Np x Np (rows = receiving patch), the
// dense right-hand sides are column-major (R x Np, each column contiguous),
// the sparse ones row-major (Np x R).  Np is N rounded up to kPadM.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "device.h"
#include "geom.h"
#include "radiosity.h"

namespace lgx {

#define HIPCHECK(x)                                                                                          \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8))); // 16-bit MFMA operand fragment as loaded

constexpr int kB = 256;          // threads per block
constexpr int kTileM = 128;      // rows per block of the FP32 GEMM
constexpr int kPadM = 256;       // Np is a multiple of this (rows per block of the plain MFMA GEMM)
constexpr int kColTiles = 12;    // 16-wide column tiles: R padded to 192
constexpr int kRPad = 16 * kColTiles;

static inline unsigned gridn(uint64_t n, unsigned b = kB) { return (unsigned)((n + b - 1) / b); }

// 16-byte piece permutation inside a 16-row x 64-byte group of the 16-bit
// GEMM operands' LDS images: the rows one ds_read_b128 serves together land
// in distinct banks
__device__ __forceinline__ int piece_swz(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }

template <class T> static T* dalloc(size_t n)
{
	void* p = nullptr;
	HIPCHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
	return (T*)p;
}
template <class T> static void dfree(T*& p)
{
	if (p)
		(void)hipFree((void*)p);
	p = nullptr;
}

// A device array owned for one solve: freed on every exit, errors included.
template <class T> struct DevBuf {
	T* p = nullptr;
	DevBuf() = default;
	explicit DevBuf(size_t n) : p(dalloc<T>(n)) {}
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
	DevBuf& operator=(DevBuf&& o) noexcept
	{
		if (this != &o) {
			dfree(p);
			p = o.p;
			o.p = nullptr;
		}
		return *this;
	}
	~DevBuf() { dfree(p); }
	operator T*() const { return p; }
};

// ---------------------------------------------------------------------------
// conversions
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t to_bf16(float x)
{
	uint32_t u = __float_as_uint(x);
	if ((u & 0x7f800000u) == 0x7f800000u)
		return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u)); // inf / quiet NaN
	u += 0x7fffu + ((u >> 16) & 1u);                                 // round to nearest even
	return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t to_f16(float x)
{
	_Float16 h = (_Float16)x;
	uint16_t r;
	memcpy(&r, &h, 2);
	return r;
}

template <int P> struct Prec; // storage precision of the dense matrix / operand
template <> struct Prec<0> {   // BF16
	static __device__ __forceinline__ uint16_t cvt(float x) { return to_bf16(x); }
};
template <> struct Prec<1> { // FP16
	static __device__ __forceinline__ uint16_t cvt(float x) { return to_f16(x); }
};

template <int P>
__device__ __forceinline__ float from16(uint16_t v)
{
	if (P == 0)
		return __uint_as_float((uint32_t)v << 16);
	_Float16 h;
	memcpy(&h, &v, 2);
	return (float)h;
}

// ---------------------------------------------------------------------------
// dense FP32 matrix H[i][j] = f_ij (the symmetric link factors, zero off the
// links), row-major Np x Np, from the link CSR: one thread per row writes
// its links into the zeroed matrix
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_dense_build(uint32_t N, uint32_t Np, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                    float* __restrict__ G)
{
	const uint32_t i = blockIdx.x * kB + threadIdx.x;
	if (i >= N)
		return;
	float* row = G + (uint64_t)i * Np;
	for (uint32_t o = rowptr[i]; o < rowptr[i + 1]; ++o)
		row[col[o]] = val[o];
}

// ---------------------------------------------------------------------------
// 16-bit forms: H' = Pi H Pi^T for the patch order Pi (pinv[i] = position of
// patch i), stored as the 256-row x 32-column tiles that hold a link.
// Tile flags and values come from the symmetric CSR links; one wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_tile_flags(uint32_t N, uint32_t Np, const uint32_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ col,
                                                   const uint32_t* __restrict__ pinv, uint8_t* __restrict__ flags)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	const uint64_t rowbase = (uint64_t)(pinv[i] >> 8) * (Np / 32);
	for (uint32_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += 64)
		flags[rowbase + (pinv[col[k]] >> 5)] = 1; // every writer stores 1
}

// Values into the tile store (zeroed first), scaled by a power of two, and
// the remainder planes; factors >= othr (the outliers) stay out.  LDS-image layout of a tile: row r = i' & 255 in
// 1 KB group r >> 4 at 64 bytes per row; column kk = j' & 31 in 16-byte piece
// kk >> 3 at slot (kk >> 3) ^ piece_swz(r & 15), element kk & 7.
template <int P>
__global__ __launch_bounds__(kB) void k_tile_scatter(uint32_t N, uint32_t Np, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col,
                                                     const float* __restrict__ val,
                                                     const uint32_t* __restrict__ pinv,
                                                     const int32_t* __restrict__ tmap, float scale, float othr,
                                                     uint16_t* __restrict__ G, uint16_t* __restrict__ Glo)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	const uint32_t ip = pinv[i], r = ip & 255, rr = ip & 15;
	const uint64_t rowbase = (uint64_t)(ip >> 8) * (Np / 32);
	const uint64_t inrow = (uint64_t)(r >> 4) * 512 + rr * 32;
	for (uint32_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += 64) {
		const uint32_t jp = pinv[col[k]], kk = jp & 31;
		const int32_t tile = tmap[rowbase + (jp >> 5)];
		if (tile < 0 || val[k] >= othr)
			continue; // (not reached: the flags came from the same links); an outlier
		const uint64_t o = (uint64_t)tile * 8192 + inrow + (uint32_t)(((kk >> 3) ^ piece_swz(rr)) * 8) + (kk & 7);
		const float a = val[k] * scale;
		const uint16_t h = Prec<P>::cvt(a);
		G[o] = h;
		Glo[o] = Prec<P>::cvt(a - from16<P>(h));
	}
}

// ---------------------------------------------------------------------------
// Outlier split of the 16-bit forms.  A handful of near-coincident patch
// pairs at the edges where two surfaces meet have factors 10^5 - 10^6 times
// the median (1 / r^4 in the form factor), and they dominate both the
// radiosity total and the rounding error of a 16-bit product: rounded to
// FP16 they alone leave a relative error near 2^-12 in the total, however
// many links average the rest.  The factors in the top binades (at most
// kOutBinades below the largest and at most max(nnz / kOutCapDiv,
// kOutCapMin) of them, whole binades only) leave the tiles and are applied
// in FP32 by the product's reduction, from the FP32 operand:
// y = H_bulk16 x16 + H_out x.
// ---------------------------------------------------------------------------
constexpr int kHistBias = 160, kHistBins = 320; // bin = ilogb(f) + kHistBias (FP32 subnormals included)
constexpr int kOutBinades = 12;
constexpr uint64_t kOutCapDiv = 4096, kOutCapMin = 256;

__global__ __launch_bounds__(kB) void k_link_hist(uint64_t nnz, const float* __restrict__ val,
                                                  unsigned* __restrict__ hist)
{
	__shared__ unsigned h[kHistBins];
	for (int b = threadIdx.x; b < kHistBins; b += kB)
		h[b] = 0;
	__syncthreads();
	for (uint64_t k = (uint64_t)blockIdx.x * kB + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * kB) {
		const float v = val[k];
		if (v > 0.0f && v < INFINITY)
			atomicAdd(&h[ilogbf(v) + kHistBias], 1u);
	}
	__syncthreads();
	for (int b = threadIdx.x; b < kHistBins; b += kB)
		if (h[b])
			atomicAdd(hist + b, h[b]);
}

// outliers per Morton position (one wave per patch row; cnt starts at zero)
__global__ __launch_bounds__(kB) void k_outlier_count(uint32_t N, const uint32_t* __restrict__ rowptr,
                                                      const float* __restrict__ val, float thr,
                                                      const uint32_t* __restrict__ pinv, uint32_t* __restrict__ cnt)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	uint32_t n = 0;
	for (uint32_t k0 = rowptr[i]; k0 < rowptr[i + 1]; k0 += 64) {
		const uint32_t k = k0 + lane;
		n += (uint32_t)__popcll(__ballot(k < rowptr[i + 1] && val[k] >= thr));
	}
	if (lane == 0)
		cnt[pinv[i]] = n;
}

// the outliers of each row in link order: other position and factor
__global__ __launch_bounds__(kB) void k_outlier_fill(uint32_t N, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col, const float* __restrict__ val,
                                                     float thr, const uint32_t* __restrict__ pinv,
                                                     const uint32_t* __restrict__ orow, uint32_t* __restrict__ ocol,
                                                     float* __restrict__ oval)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	uint32_t o = orow[pinv[i]];
	for (uint32_t k0 = rowptr[i]; k0 < rowptr[i + 1]; k0 += 64) {
		const uint32_t k = k0 + lane;
		const bool hit = k < rowptr[i + 1] && val[k] >= thr;
		const uint64_t m = __ballot(hit);
		if (hit) {
			const uint32_t at = o + (uint32_t)__popcll(m & ((1ull << lane) - 1));
			ocol[at] = pinv[col[k]];
			oval[at] = val[k];
		}
		o += (uint32_t)__popcll(m);
	}
}

// ---------------------------------------------------------------------------
// Power-of-two column scale of the FP16 operand from its column maximum:
// each column's largest entry lands near 2^13, so neither growth of the
// iterates nor small values leave FP16's range.  BF16 has FP32's exponent
// range and keeps scale 1 (cmax null).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float col_scale_of(const unsigned* cmax, uint32_t c)
{
	if (!cmax)
		return 1.0f;
	const float m = __uint_as_float(cmax[c]);
	return (m > 0.0f && m < INFINITY) ? ldexpf(1.0f, 14 - ilogbf(m) - 1) : 1.0f;
}

// Diffuse and area of the patch at each Morton position (padding: zero).
__global__ __launch_bounds__(kB) void k_permute_colors(const float* __restrict__ diffuse,
                                                       const float* __restrict__ area,
                                                       const uint32_t* __restrict__ perm, uint32_t N, uint32_t Np,
                                                       float* __restrict__ dperm, float* __restrict__ aperm)
{
	const uint32_t p = blockIdx.x * kB + threadIdx.x;
	if (p >= Np)
		return;
	const uint32_t src = perm[p];
	const bool in = src < N;
	dperm[3 * p + 0] = in ? diffuse[3 * src + 0] : 0.0f;
	dperm[3 * p + 1] = in ? diffuse[3 * src + 1] : 0.0f;
	dperm[3 * p + 2] = in ? diffuse[3 * src + 2] : 0.0f;
	aperm[p] = in ? area[src] : 0.0f;
}

// ---------------------------------------------------------------------------
// Operand passes of the 16-bit tile products, one kernel for three jobs:
//   source   the column-major FP32 vector x (positions in Morton order), or
//            the sum of a product's partial-sum planes (slots 0 .. nslot of
//            the row block, in slot order) divided by the product's scales,
//            plus the outliers' products in FP32 (in CSR order);
//   y        that source written out (the product's result);
//   cmax_out column maxima of |w| by atomicMax (non-negative floats order
//            like their bits, so the result does not depend on the order of
//            the atomics; the buffer starts at zero);
//   xq / xlo w times the column scale from cmax_op, rounded to 16 bits, and
//            the remainder, as LDS images: stage j / 32, column group c / 16,
//            row c % 16, piece (j % 32) / 8 at slot piece ^ piece_swz(c % 16);
// where w = (x[c][j] * diffuse_j[c % 3]) * area_j for K = H diag(area)
// diag(diffuse_c) (pre), or x[c][j] for the bare H (the adjoint sweep).  The
// padding positions have zero diffuse and area (with pre off they hold zero
// in x); columns R .. 191 are zero.
// Block: one 16-column group x one 256-position row block; wave: two k
// stages; lane: column r = lane >> 2 of the group, piece pc = lane & 3, so a
// wave writes one contiguous 1 KB group per stage.  Fusing the reduction of
// one product with the operand of the next (the series z_b = K z_{b-1})
// saves a pass over z: the BF16 forms take the operand straight from the
// reduction, the FP16 forms the column maxima that set its scales.
// ---------------------------------------------------------------------------
struct OperandPass {
	const float* part = nullptr;      // partial-sum planes (null: read x)
	const uint32_t* nslot = nullptr;  // slots per row block
	const unsigned* cmax_in = nullptr; // the product's operand scales (FP16)
	float gscale = 1.0f;              // the product's matrix scale
	const float* x = nullptr;
	float* y = nullptr;
	const float* dperm = nullptr;
	const float* aperm = nullptr;
	int pre = 0;
	unsigned* cmax_out = nullptr;
	const unsigned* cmax_op = nullptr; // scales of the operand written here (FP16)
	uint16_t* xq = nullptr;
	uint16_t* xlo = nullptr;
	// reduction: + H_out w_in, the outliers (CSR by Morton position) times
	// the product's FP32 operand w_in (from its input vector xin)
	const uint32_t* orow = nullptr;
	const uint32_t* ocol = nullptr;
	const float* oval = nullptr;
	const float* xin = nullptr;
	uint32_t Np = 0, R = 0;
};

template <int P>
__global__ __launch_bounds__(kB) void k_operand(const OperandPass a)
{
	const uint32_t cg = blockIdx.x % (kRPad / 16), rb = blockIdx.x / (kRPad / 16);
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane >> 2, pc = lane & 3;
	const uint32_t c = cg * 16 + r;
	const bool live = c < a.R;
	const uint64_t cbase = (uint64_t)c * a.Np;
	float inv = 1.0f;
	uint32_t ns = 0;
	if (a.part) {
		ns = a.nslot[rb];
		inv = 1.0f / (col_scale_of(a.cmax_in, c) * a.gscale); // powers of two: exact
	}
	const float sc = a.xq ? col_scale_of(a.cmax_op, c) : 1.0f;
	float m = 0.0f;
#pragma unroll
	for (int k = 0; k < 2; ++k) {
		const uint32_t stage = rb * 8 + wave * 2 + k, j0 = stage * 32 + pc * 8;
		float v[8];
#pragma unroll
		for (int e = 0; e < 8; ++e)
			v[e] = 0.0f;
		if (a.part) {
			for (uint32_t q = 0; q < ns; ++q) {
				const float* src = a.part + (uint64_t)q * kRPad * a.Np + cbase + j0;
				const float4 u0 = *reinterpret_cast<const float4*>(src);
				const float4 u1 = *reinterpret_cast<const float4*>(src + 4);
				v[0] += u0.x, v[1] += u0.y, v[2] += u0.z, v[3] += u0.w;
				v[4] += u1.x, v[5] += u1.y, v[6] += u1.z, v[7] += u1.w;
			}
#pragma unroll
			for (int e = 0; e < 8; ++e)
				v[e] *= inv;
			if (a.orow && live && a.orow[j0] != a.orow[j0 + 8]) {
				for (int e = 0; e < 8; ++e)
					for (uint32_t k = a.orow[j0 + e]; k < a.orow[j0 + e + 1]; ++k) {
						const uint32_t q = a.ocol[k];
						float w = a.xin[cbase + q];
						if (a.pre)
							w = (w * a.dperm[3 * q + c % 3]) * a.aperm[q];
						v[e] += a.oval[k] * w;
					}
			}
		} else if (live) {
			const float4 u0 = *reinterpret_cast<const float4*>(a.x + cbase + j0);
			const float4 u1 = *reinterpret_cast<const float4*>(a.x + cbase + j0 + 4);
			v[0] = u0.x, v[1] = u0.y, v[2] = u0.z, v[3] = u0.w, v[4] = u1.x, v[5] = u1.y, v[6] = u1.z, v[7] = u1.w;
		}
		if (a.y && live) {
			*reinterpret_cast<float4*>(a.y + cbase + j0) = make_float4(v[0], v[1], v[2], v[3]);
			*reinterpret_cast<float4*>(a.y + cbase + j0 + 4) = make_float4(v[4], v[5], v[6], v[7]);
		}
		if (!a.xq && !a.cmax_out)
			continue;
#pragma unroll
		for (int e = 0; e < 8; ++e) {
			if (!live)
				v[e] = 0.0f;
			else if (a.pre)
				v[e] = (v[e] * a.dperm[3 * (j0 + e) + c % 3]) * a.aperm[j0 + e];
			m = fmaxf(m, fabsf(v[e]));
		}
		if (a.xq) {
			uint32_t h[4], l[4];
#pragma unroll
			for (int e = 0; e < 8; e += 2) {
				const float w0 = v[e] * sc, w1 = v[e + 1] * sc;
				const uint16_t h0 = Prec<P>::cvt(w0), h1 = Prec<P>::cvt(w1);
				h[e / 2] = h0 | (uint32_t)h1 << 16;
				l[e / 2] = Prec<P>::cvt(w0 - from16<P>(h0)) | (uint32_t)Prec<P>::cvt(w1 - from16<P>(h1)) << 16;
			}
			const uint64_t o = (((uint64_t)stage * (kRPad / 16) + cg) * 64 + r * 4 + (pc ^ piece_swz(r))) * 8;
			*reinterpret_cast<uint4*>(a.xq + o) = make_uint4(h[0], h[1], h[2], h[3]);
			if (a.xlo)
				*reinterpret_cast<uint4*>(a.xlo + o) = make_uint4(l[0], l[1], l[2], l[3]);
		}
	}
	if (a.cmax_out) {
		m = fmaxf(m, __shfl_xor(m, 1));
		m = fmaxf(m, __shfl_xor(m, 2));
		__shared__ float wm[4][16];
		if (pc == 0)
			wm[wave][r] = m;
		__syncthreads();
		if (threadIdx.x < 16) {
			const float b = fmaxf(fmaxf(wm[0][threadIdx.x], wm[1][threadIdx.x]), fmaxf(wm[2][threadIdx.x], wm[3][threadIdx.x]));
			if (b > 0.0f)
				atomicMax(a.cmax_out + cg * 16 + threadIdx.x, __float_as_uint(b));
		}
	}
}

// FP32 operand of the dense FP32 form, patch order: Xq[c][j] =
// (x[c][j] * diffuse[j][c % 3]) * area_j (pre) or x[c][j]
__global__ __launch_bounds__(kB) void k_prep_x32(const float* __restrict__ xT, const float* __restrict__ diffuse,
                                                 const float* __restrict__ area, uint32_t N, uint32_t Np, uint32_t R,
                                                 int pre, float* __restrict__ xq)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)kRPad * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), j = (uint32_t)(t % Np);
	float v = 0.0f;
	if (c < R && j < N) {
		v = xT[(uint64_t)c * Np + j];
		if (pre)
			v = (v * diffuse[3 * j + c % 3]) * area[j];
	}
	xq[t] = v;
}

// ---------------------------------------------------------------------------
// MFMA building blocks of the 16-bit tile products: v_mfma_f32_16x16x32
// (bf16 or f16) with FP32 accumulators; lane l supplies row (l & 15) of A and
// column (l & 15) of B with the same eight k indices 8 (l >> 4) + e, so the k
// order inside an instruction is immaterial; output element i of a 16 x 16
// tile sits at row 4 (l >> 4) + i, column l & 15.  k advances in stages of
// 32; operand stages go HBM -> LDS by global_load_lds_dwordx4, 1 KB (one
// 16-row x 64-byte group) per instruction, and the 16-byte pieces of a row
// are permuted by piece_swz(row) so that the rows one ds_read_b128 serves
// together fall in distinct banks.
// ---------------------------------------------------------------------------
constexpr int kKS = 32; // k per stage
constexpr int kMaxPD = 6; // deepest LDS ring

// The 16-bit matrix and operand are stored as LDS images (see k_tile_scatter
// / k_operand): 1 KB groups of 16 rows x 64 bytes.
constexpr uint64_t kTileElems = 512;                         // 16-bit elements per 1 KB group
constexpr uint64_t kGStageElems = 16 * kTileElems;           // a stored 256-row x 32 tile
constexpr uint64_t kXStageElems = (kRPad / 16) * kTileElems; // an operand stage (192 x 32)

template <int P>
__device__ __forceinline__ f32x4 mfma16(i16x8 a, i16x8 b, f32x4 c)
{
	if (P == 0)
		return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
		                                               0, 0, 0);
	else
		return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
		                                              0, 0, 0);
}


typedef __attribute__((address_space(3))) void lds_void;


// wait until at most N vector-memory operations of this wave are outstanding
template <int N>
__device__ __forceinline__ void wait_vm()
{
	static_assert(N >= 0 && N < 64, "vmcnt range");
	asm volatile("" ::: "memory");
	__builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
	asm volatile("" ::: "memory");
}

__device__ __forceinline__ void block_barrier()
{
	asm volatile("" ::: "memory");
	__builtin_amdgcn_s_barrier();
	asm volatile("" ::: "memory");
}

// 16 bytes per lane from global memory to LDS at dst + 16 * lane
__device__ __forceinline__ void dma16(const uint16_t* src, unsigned char* dst)
{
	__builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)dst, 16, 0, 0);
}

// ---------------------------------------------------------------------------
// The same product with the matrix fragments loaded straight into registers.
//
// Block: 4 waves over one 256-row block (64 rows = 4 row tiles per wave) x
// all 12 column tiles.  The matrix is per-wave data: each lane loads its A
// fragment (16 bytes per row tile and plane) from the tile store into
// registers D stages ahead; in the store's LDS-image layout lane l's piece
// (row l & 15, k piece l >> 4) lies at 16 (4 row + (piece ^ piece_swz(row)))
// bytes into the row tile's 1 KB group, so a wave's loads of one row tile
// cover one contiguous KB.  The operand stage (192 x 32 per plane, 12 KB) is
// shared by the block's waves: it goes HBM -> LDS by global_load_lds_dwordx4
// into a ring of D slots (one variable each, so that the compiler's
// wait-count tracking tells the slots apart), each wave copying 3 of the 12
// groups per plane.
// Stage s (ring slot s % D, unrolled so that every slot is a constant):
// wait until this wave's loads and copies of stage s have landed (the
// D - 1 later stages may still be in flight), barrier (every wave's copies
// are in), read the operand fragments, wait for those reads and barrier
// (every wave is done with the slot), copy stage s + D into the slot,
// multiply, load stage s + D's matrix fragments into the registers just
// used.  Past the end the last stage is fetched again (never used), which
// keeps every stage's wait count exact.  The matrix is read once per
// product, so its loads are non-temporal: they leave the L2 to the operand
// stages, which every row block of the same k stage reads again.
// ---------------------------------------------------------------------------
template <int NPL, int D>
struct TileLane {
	static constexpr int kXPlane = kRPad * kKS * 2; // operand plane per stage (12 KB)
	static constexpr int kOps = NPL * (3 + 4);      // loads and copies per wave per stage
	const uint16_t *G, *Glo, *xq, *xlo;
	const uint32_t* lst; // k stages of the run's tiles
	uint64_t first;      // the run's first tile
	uint32_t aofs;       // element offset of this lane's A piece in a 1 KB group
	int rofs;            // byte offset of this lane's B fragment in a 1 KB group
	int wave, lane;
	unsigned char* slot[kMaxPD];
};

__device__ __forceinline__ void lds_reads_done()
{
	asm volatile("" ::: "memory");
	__builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0), vmcnt and expcnt left alone
	asm volatile("" ::: "memory");
}

template <int NPL, int D>
__device__ __forceinline__ void tile_copy_operand(const TileLane<NPL, D>& L, uint32_t st, unsigned char* slot)
{
	const uint64_t ks = L.lst[st];
#pragma unroll
	for (int pl = 0; pl < NPL; ++pl) {
		const uint16_t* X = (pl ? L.xlo : L.xq) + ks * kXStageElems + 8 * L.lane;
#pragma unroll
		for (int j = 0; j < 3; ++j)
			dma16(X + (uint64_t)(3 * L.wave + j) * kTileElems,
			      slot + pl * TileLane<NPL, D>::kXPlane + (3 * L.wave + j) * 1024);
	}
}

template <int NPL, int D>
__device__ __forceinline__ void tile_load_matrix(const TileLane<NPL, D>& L, uint32_t st, i16x8 (&a)[NPL][4])
{
	const uint64_t base = (L.first + st) * kGStageElems + (uint64_t)(4 * L.wave) * kTileElems + L.aofs;
#pragma unroll
	for (int pl = 0; pl < NPL; ++pl) {
		const uint16_t* M = pl ? L.Glo : L.G;
#pragma unroll
		for (int t = 0; t < 4; ++t)
			a[pl][t] = __builtin_nontemporal_load(reinterpret_cast<const i16x8*>(M + base + t * kTileElems));
	}
}

template <int P, int NPL, int D, int B>
__device__ __forceinline__ void tile_step(int s, int nst, const TileLane<NPL, D>& L, i16x8 (&a)[D][NPL][4],
                                          f32x4 (&acc)[4][kColTiles])
{
	typedef TileLane<NPL, D> T;
	wait_vm<T::kOps * (D - 1)>();
	block_barrier();
	unsigned char* slot = L.slot[B];
	i16x8 bx[NPL][kColTiles];
#pragma unroll
	for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
		for (int c = 0; c < kColTiles; ++c)
			bx[pl][c] = *reinterpret_cast<const i16x8*>(slot + pl * T::kXPlane + c * 1024 + L.rofs);
	lds_reads_done();
	block_barrier();
	const int ahead = s + D < nst ? s + D : nst - 1;
	tile_copy_operand<NPL, D>(L, (uint32_t)ahead, slot);
#pragma unroll
	for (int c = 0; c < kColTiles; ++c) {
#pragma unroll
		for (int t = 0; t < 4; ++t)
			acc[t][c] = mfma16<P>(a[B][0][t], bx[0][c], acc[t][c]);
		if (NPL == 2) {
			// error-compensated: + G_hi X_lo + G_lo X_hi
#pragma unroll
			for (int t = 0; t < 4; ++t) {
				acc[t][c] = mfma16<P>(a[B][0][t], bx[NPL - 1][c], acc[t][c]);
				acc[t][c] = mfma16<P>(a[B][NPL - 1][t], bx[0][c], acc[t][c]);
			}
		}
	}
	tile_load_matrix<NPL, D>(L, (uint32_t)ahead, a[B]);
}

template <int P, int NPL, int D, int... Bs>
__device__ __forceinline__ void tile_steps(int s, int nst, const TileLane<NPL, D>& L, i16x8 (&a)[D][NPL][4],
                                           f32x4 (&acc)[4][kColTiles], std::integer_sequence<int, Bs...>)
{
	(tile_step<P, NPL, D, Bs>(s + Bs, nst, L, a, acc), ...);
}

template <int P, int NPL, int D, int... Bs>
__device__ __forceinline__ void tile_tail(int s, int nst, const TileLane<NPL, D>& L, i16x8 (&a)[D][NPL][4],
                                          f32x4 (&acc)[4][kColTiles], std::integer_sequence<int, Bs...>)
{
	((s + Bs < nst ? tile_step<P, NPL, D, Bs>(s + Bs, nst, L, a, acc) : void()), ...);
}

template <int NPL, int D, int... Bs>
__device__ __forceinline__ void tile_prologue(int nst, const TileLane<NPL, D>& L, i16x8 (&a)[D][NPL][4],
                                              std::integer_sequence<int, Bs...>)
{
	((tile_copy_operand<NPL, D>(L, (uint32_t)(Bs < nst ? Bs : nst - 1), L.slot[Bs]),
	  tile_load_matrix<NPL, D>(L, (uint32_t)(Bs < nst ? Bs : nst - 1), a[Bs])),
	 ...);
}

template <int P, int NPL, int D>
__global__ __launch_bounds__(kB) void k_gemm_tiles(const uint16_t* __restrict__ G, const uint16_t* __restrict__ Glo,
                                                   const uint16_t* __restrict__ xq, const uint16_t* __restrict__ xlo,
                                                   const uint4* __restrict__ chunks,
                                                   const uint32_t* __restrict__ seg_off,
                                                   const uint32_t* __restrict__ tks, uint32_t Np,
                                                   float* __restrict__ part)
{
	typedef TileLane<NPL, D> T;
	static_assert(D >= 2 && D <= kMaxPD, "ring depth");
	constexpr int kSlot = NPL * T::kXPlane;
	__shared__ __align__(1024) unsigned char lds0[kSlot], lds1[kSlot], lds2[D > 2 ? kSlot : 16],
	    lds3[D > 3 ? kSlot : 16], lds4[D > 4 ? kSlot : 16], lds5[D > 5 ? kSlot : 16];
	const int tid = threadIdx.x, lane = tid & 63;
	const int r = lane & 15, g = lane >> 4;
	T L;
	L.slot[0] = lds0;
	L.slot[1] = lds1;
	L.slot[2] = lds2;
	L.slot[3] = lds3;
	L.slot[4] = lds4;
	L.slot[5] = lds5;
	L.G = G;
	L.Glo = Glo;
	L.xq = xq;
	L.xlo = xlo;
	L.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	L.lane = lane;
	// XCD-aware order, as in k_gemm_mfma
	const uint32_t nb = gridDim.x, x = blockIdx.x % 8, per = nb / 8, rem = nb % 8;
	const uint32_t bid = x * per + (x < rem ? x : rem) + blockIdx.x / 8;
	L.aofs = 8u * (uint32_t)(4 * r + (g ^ piece_swz(r)));
	L.rofs = r * 64 + 16 * (g ^ piece_swz(r));
	for (uint32_t q = seg_off[bid]; q < seg_off[bid + 1]; ++q) {
		const uint4 ch = chunks[q];
		const uint32_t i0 = ch.x * 256 + L.wave * 64;
		const int nst = (int)(ch.z - ch.y);
		L.lst = tks + ch.y;
		L.first = ch.y;
		f32x4 acc[4][kColTiles];
#pragma unroll
		for (int t = 0; t < 4; ++t)
#pragma unroll
			for (int c = 0; c < kColTiles; ++c)
				acc[t][c] = f32x4{0, 0, 0, 0};
		if (nst > 0) {
			i16x8 a[D][NPL][4];
			tile_prologue<NPL, D>(nst, L, a, std::make_integer_sequence<int, D>{});
			int s = 0;
			for (; s + D <= nst; s += D)
				tile_steps<P, NPL, D>(s, nst, L, a, acc, std::make_integer_sequence<int, D>{});
			tile_tail<P, NPL, D>(s, nst, L, a, acc, std::make_integer_sequence<int, D - 1>{});
			wait_vm<0>();    // no copy lands after the segment (or the block) is done with the ring
			block_barrier(); // every wave has finished with the ring before the next segment refills it
		}
		float* out = part + (uint64_t)ch.w * kRPad * Np;
#pragma unroll
		for (int t = 0; t < 4; ++t)
#pragma unroll
			for (int c = 0; c < kColTiles; ++c) {
				const uint32_t col = c * 16 + r, i = i0 + t * 16 + 4 * g;
				*reinterpret_cast<float4*>(out + (uint64_t)col * Np + i) =
				    make_float4(acc[t][c][0], acc[t][c][1], acc[t][c][2], acc[t][c][3]);
			}
	}
}

constexpr int kTilesD1 = 4, kTilesD2 = 3; // ring depths of the plain and compensated products

constexpr uint32_t kMaxSlots = 16; // partial-sum planes allocated at least
constexpr uint64_t kMinShare = 16;  // fewest tiles per persistent block of the tile products

// ---------------------------------------------------------------------------
// Tile-sparse FP32 GEMM on the FP32 matrix cores: the stored tiles, work runs
// and slots of the 16-bit forms, values in FP32.  v_mfma_f32_16x16x4_f32:
// lane l supplies A[row l & 15][k] and B[k][column l & 15] for one k of the
// instruction's four; over the eight instructions of a 32-wide k stage, lane
// group g = l >> 4 takes k = 8g .. 8g + 7 in turn, for A and B alike, so the
// four k of every instruction match.  Output element i of a 16 x 16 tile
// sits at row 4 (l >> 4) + i, column l & 15, as for the 16-bit MFMA.
//
// LDS images (and the HBM layouts, so that every DMA copies one contiguous
// KB): 2 KB groups of 16 rows x 128 bytes, the eight 16-byte pieces of a row
// at slot piece ^ (row & 7), so that the rows one ds_read_b128 serves
// together fall in distinct banks.  A stored tile is 16 groups (256 rows x
// 32 k, 32 KB); an operand stage 12 groups (192 columns x 32 k, 24 KB).
// Block: 4 waves x 64 rows (4 row tiles) x all 192 columns; k stages
// through a two-buffer LDS ring (112 KB), the next stage's DMA in flight
// while the current one is multiplied (a stage is 384 MFMAs per wave, the
// kernel is bound by FP32 matrix throughput, not by its loads).
// ---------------------------------------------------------------------------
constexpr uint64_t kT32Group = 512;                      // floats per 2 KB group
constexpr uint64_t kT32Tile = 16 * kT32Group;            // a stored tile
constexpr uint64_t kT32XStage = (kRPad / 16) * kT32Group; // an operand stage
constexpr int kT32XBytes = (int)(kT32XStage * 4);          // 24 KB
constexpr int kT32Stage = kT32XBytes + (int)(kT32Tile * 4); // 56 KB

__device__ __forceinline__ int t32_slot(int piece, int row) { return piece ^ (row & 7); }

__device__ __forceinline__ void dma16f(const float* src, unsigned char* dst)
{
	__builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)dst, 16, 0, 0);
}

// stage st of the run (tile first + st, operand stage lst[st]) into buf
__device__ __forceinline__ void t32_issue(const float* G, const float* xq, uint64_t first, const uint32_t* lst,
                                          uint32_t st, unsigned char* buf, int wave, int lane)
{
	const uint64_t ks = lst[st];
	const float* xs = xq + ks * kT32XStage + 4 * lane;
	const float* gs = G + (first + st) * kT32Tile + 4 * lane;
#pragma unroll
	for (int j = 0; j < 6; ++j) // the operand's 24 KB: 6 per wave
		dma16f(xs + (uint64_t)(6 * wave + j) * 256, buf + (6 * wave + j) * 1024);
#pragma unroll
	for (int j = 0; j < 8; ++j) // the wave's own 64 rows: 4 groups of 2 KB
		dma16f(gs + (uint64_t)(8 * wave + j) * 256, buf + kT32XBytes + (8 * wave + j) * 1024);
}

__global__ __launch_bounds__(kB) void k_gemm_f32t(const float* __restrict__ G, const float* __restrict__ xq,
                                                  const uint4* __restrict__ chunks,
                                                  const uint32_t* __restrict__ seg_off,
                                                  const uint32_t* __restrict__ tks, uint32_t Np,
                                                  float* __restrict__ part)
{
	constexpr int RT = 4; // row tiles per wave
	__shared__ __align__(1024) unsigned char lds0[kT32Stage], lds1[kT32Stage];
	const int tid = threadIdx.x, lane = tid & 63;
	const int r = lane & 15, g = lane >> 4;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	// XCD-aware order, as in k_gemm_mfma
	const uint32_t nb = gridDim.x, x = blockIdx.x % 8, per = nb / 8, rem = nb % 8;
	const uint32_t bid = x * per + (x < rem ? x : rem) + blockIdx.x / 8;
	const int s0 = t32_slot(2 * g, r) * 16, s1 = t32_slot(2 * g + 1, r) * 16;
	for (uint32_t q = seg_off[bid]; q < seg_off[bid + 1]; ++q) {
	const uint4 ch = chunks[q];
	const uint32_t i0 = ch.x * 256 + wave * 16 * RT;
	const int nst = (int)(ch.z - ch.y);
	const uint32_t* lst = tks + ch.y;

	f32x4 acc[RT][kColTiles];
#pragma unroll
	for (int t = 0; t < RT; ++t)
#pragma unroll
		for (int c = 0; c < kColTiles; ++c)
			acc[t][c] = f32x4{0, 0, 0, 0};

	for (int st = 0; st < nst; ++st) {
		unsigned char* cur = (st & 1) ? lds1 : lds0;
		unsigned char* nxt = (st & 1) ? lds0 : lds1;
		if (st == 0)
			t32_issue(G, xq, ch.y, lst, 0, cur, wave, lane);
		wait_vm<0>();    // this wave's part of stage st has landed
		block_barrier(); // every wave's part is in; stage st-1's buffer is free
		float4 a[RT][2], b[kColTiles][2];
#pragma unroll
		for (int t = 0; t < RT; ++t) {
			const unsigned char* rowp = cur + kT32XBytes + (wave * RT + t) * 2048 + r * 128;
			a[t][0] = *reinterpret_cast<const float4*>(rowp + s0);
			a[t][1] = *reinterpret_cast<const float4*>(rowp + s1);
		}
#pragma unroll
		for (int c = 0; c < kColTiles; ++c) {
			const unsigned char* colp = cur + c * 2048 + r * 128;
			b[c][0] = *reinterpret_cast<const float4*>(colp + s0);
			b[c][1] = *reinterpret_cast<const float4*>(colp + s1);
		}
		// every fragment is read before the next DMA is issued (see gemm_step)
		if (st + 1 < nst)
			t32_issue(G, xq, ch.y, lst, st + 1, nxt, wave, lane);
#pragma unroll
		for (int e = 0; e < 8; ++e)
#pragma unroll
			for (int t = 0; t < RT; ++t)
#pragma unroll
				for (int c = 0; c < kColTiles; ++c) {
					const float av = e < 4 ? a[t][0][e] : a[t][1][e - 4];
					const float bv = e < 4 ? b[c][0][e] : b[c][1][e - 4];
					acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t][c], 0, 0, 0);
				}
	}
	block_barrier(); // every wave is done with both buffers before the next segment refills them
	float* out = part + (uint64_t)ch.w * kRPad * Np;
#pragma unroll
	for (int t = 0; t < RT; ++t)
#pragma unroll
		for (int c = 0; c < kColTiles; ++c) {
			const uint32_t col = c * 16 + r, i = i0 + t * 16 + 4 * g;
			*reinterpret_cast<float4*>(out + (uint64_t)col * Np + i) =
			    make_float4(acc[t][c][0], acc[t][c][1], acc[t][c][2], acc[t][c][3]);
		}
	}
}

// FP32 tile store values from the symmetric CSR links (zeroed first); one
// wave per row, as k_tile_scatter
__global__ __launch_bounds__(kB) void k_tile_scatter_f32(uint32_t N, uint32_t Np, const uint32_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ col,
                                                         const float* __restrict__ val,
                                                         const uint32_t* __restrict__ pinv,
                                                         const int32_t* __restrict__ tmap, float* __restrict__ G)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	const uint32_t ip = pinv[i], rr = ip & 15;
	const uint64_t rowbase = (uint64_t)(ip >> 8) * (Np / 32);
	const uint64_t inrow = (uint64_t)((ip & 255) >> 4) * kT32Group + rr * 32;
	for (uint32_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += 64) {
		const uint32_t jp = pinv[col[k]], kk = jp & 31;
		const int32_t tile = tmap[rowbase + (jp >> 5)];
		if (tile < 0)
			continue; // not reached: the flags came from the same links
		G[(uint64_t)tile * kT32Tile + inrow + t32_slot(kk >> 2, rr) * 4 + (kk & 3)] = val[k];
	}
}

// FP32 operand of the tile form at Morton positions j .. j + 3 of column c,
// one 16-byte piece per lane: (x[c][j] * diffuse_j[c % 3]) * area_j (pre) or
// x[c][j], in the LDS-image layout: stage j / 32, column group c / 16, row
// c % 16, piece (j % 32) / 4 at slot piece ^ (c % 8)
__global__ __launch_bounds__(kB) void k_prep_x32t(const float* __restrict__ xT, const float* __restrict__ dperm,
                                                  const float* __restrict__ aperm, uint32_t Np, uint32_t R, int pre,
                                                  float* __restrict__ xq)
{
	const uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	const uint32_t pieces = Np / 4;
	if (t >= (uint64_t)kRPad * pieces)
		return;
	const uint32_t c = (uint32_t)(t / pieces), j0 = (uint32_t)(t % pieces) * 4;
	float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	if (c < R) {
		v = *reinterpret_cast<const float4*>(xT + (uint64_t)c * Np + j0);
		if (pre) {
			v.x = (v.x * dperm[3 * (j0 + 0) + c % 3]) * aperm[j0 + 0];
			v.y = (v.y * dperm[3 * (j0 + 1) + c % 3]) * aperm[j0 + 1];
			v.z = (v.z * dperm[3 * (j0 + 2) + c % 3]) * aperm[j0 + 2];
			v.w = (v.w * dperm[3 * (j0 + 3) + c % 3]) * aperm[j0 + 3];
		}
	}
	const uint32_t kk = j0 & 31, cc = c & 15;
	*reinterpret_cast<float4*>(xq + ((uint64_t)(j0 >> 5) * (kRPad / 16) + (c >> 4)) * kT32Group + cc * 32 +
	                           t32_slot(kk >> 2, cc) * 4) = v;
}

// ---------------------------------------------------------------------------
// plain FP32 GEMM with the same blocking: 256 threads, each 8 rows x 12
// columns, k staged through LDS 16 at a time
// ---------------------------------------------------------------------------
constexpr int kK32 = 16;

__global__ __launch_bounds__(kB) void k_gemm_f32(const float* __restrict__ G, const float* __restrict__ xq,
                                                 uint32_t Np, uint32_t ksplit, float* __restrict__ part)
{
	__shared__ float as[kK32][kTileM + 4];
	__shared__ float bs[kK32][kRPad + 4];
	const int tid = threadIdx.x;
	const int tr = tid / 16, tc = tid % 16;
	const uint32_t i0 = blockIdx.x * kTileM;
	const uint32_t klen = Np / ksplit;
	const uint32_t kbeg = blockIdx.y * klen, kend = kbeg + klen;
	float acc[8][12];
#pragma unroll
	for (int a = 0; a < 8; ++a)
#pragma unroll
		for (int b = 0; b < 12; ++b)
			acc[a][b] = 0.0f;
	for (uint32_t kb = kbeg; kb < kend; kb += kK32) {
		__syncthreads();
		// A tile 128 x 16: 512 float4, two per thread
#pragma unroll
		for (int q = 0; q < 2; ++q) {
			int p = q * kB + tid, row = p / 4, kq = p % 4;
			float4 v = *reinterpret_cast<const float4*>(G + (uint64_t)(i0 + row) * Np + kb + 4 * kq);
			as[4 * kq + 0][row] = v.x;
			as[4 * kq + 1][row] = v.y;
			as[4 * kq + 2][row] = v.z;
			as[4 * kq + 3][row] = v.w;
		}
		// B tile 192 x 16: 768 float4, three per thread
#pragma unroll
		for (int q = 0; q < 3; ++q) {
			int p = q * kB + tid, c = p / 4, kq = p % 4;
			float4 v = *reinterpret_cast<const float4*>(xq + (uint64_t)c * Np + kb + 4 * kq);
			bs[4 * kq + 0][c] = v.x;
			bs[4 * kq + 1][c] = v.y;
			bs[4 * kq + 2][c] = v.z;
			bs[4 * kq + 3][c] = v.w;
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < kK32; ++k) {
			float a[8], b[12];
#pragma unroll
			for (int x = 0; x < 8; ++x)
				a[x] = as[k][tr * 8 + x];
#pragma unroll
			for (int y = 0; y < 12; ++y)
				b[y] = bs[k][tc + 16 * y];
#pragma unroll
			for (int x = 0; x < 8; ++x)
#pragma unroll
				for (int y = 0; y < 12; ++y)
					acc[x][y] += a[x] * b[y];
		}
	}
	float* out = part + (uint64_t)blockIdx.y * kRPad * Np;
#pragma unroll
	for (int y = 0; y < 12; ++y) {
		uint32_t c = tc + 16 * y;
#pragma unroll
		for (int x = 0; x < 8; x += 4)
			*reinterpret_cast<float4*>(out + (uint64_t)c * Np + i0 + tr * 8 + x) =
			    make_float4(acc[x][y], acc[x + 1][y], acc[x + 2][y], acc[x + 3][y]);
	}
}

// y[c][i] = (sum_q part[q][c][i]) / (s_c * gscale), q in slot order: ksplit
// slots, or nslot[i / 256] for the chunks of 256-row block i / 256
__global__ __launch_bounds__(kB) void k_reduce_split(const float* __restrict__ part, uint32_t ksplit,
                                                     const uint32_t* __restrict__ nslot, uint32_t Np, uint32_t R,
                                                     const unsigned* __restrict__ cmax, float gscale,
                                                     float* __restrict__ y)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)R * Np)
		return;
	const uint32_t c = (uint32_t)(t / Np), i = (uint32_t)(t % Np);
	const uint32_t ns = nslot ? nslot[i >> 8] : ksplit;
	float s = 0.0f;
	for (uint32_t q = 0; q < ns; ++q)
		s += part[(uint64_t)q * kRPad * Np + t];
	y[t] = s * (1.0f / (col_scale_of(cmax, c) * gscale)); // powers of two: exact
}

// ---------------------------------------------------------------------------
// sparse gather over the symmetric CSR links.  The operand is prepared once
// per product, xp[j][c] = (x[j][c] * diffuse[j][c % 3]) * area_j (PRE, the
// reference's operation order) or x[j][c]; then one wave per row, lanes
// over columns (up to 3 per lane for R <= 192), the row's links taken 64 at
// a time (one coalesced load of columns and factors, broadcast lane by lane):
// y[i][c] = sum_k xp[col_k][c] * f_k in link order.
// ---------------------------------------------------------------------------
template <class T, bool PRE>
__global__ __launch_bounds__(kB) void k_prep_rows(uint64_t n, uint32_t R, const T* __restrict__ x,
                                                  const float* __restrict__ diffuse, const float* __restrict__ area,
                                                  T* __restrict__ xp)
{
	const uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= n)
		return;
	if (PRE) {
		const uint32_t j = (uint32_t)(t / R), c = (uint32_t)(t % R);
		xp[t] = (x[t] * (T)diffuse[3 * j + c % 3]) * (T)area[j];
	} else {
		xp[t] = x[t];
	}
}

template <class T>
__global__ __launch_bounds__(kB) void k_gather(uint32_t N, uint32_t R, const uint32_t* __restrict__ rowptr,
                                               const uint32_t* __restrict__ col, const float* __restrict__ val,
                                               const T* __restrict__ xp, T* __restrict__ y)
{
	const uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	const uint32_t c0 = lane, c1 = lane + 64, c2 = lane + 128;
	const bool u0 = c0 < R, u1 = c1 < R, u2 = c2 < R;
	T a0 = 0, a1 = 0, a2 = 0;
	const uint32_t kb = rowptr[i], ke = rowptr[i + 1];
	for (uint32_t k0 = kb; k0 < ke; k0 += 64) {
		const uint32_t n = min(64u, ke - k0);
		const uint32_t mj = (uint32_t)lane < n ? col[k0 + lane] : 0u;
		const float mf = (uint32_t)lane < n ? val[k0 + lane] : 0.0f;
		for (uint32_t q = 0; q < n; ++q) {
			const uint32_t j = (uint32_t)__shfl((int)mj, (int)q);
			const T f = (T)__shfl(mf, (int)q);
			const T* row = xp + (uint64_t)j * R;
			if (u0)
				a0 += row[c0] * f;
			if (u1)
				a1 += row[c1] * f;
			if (u2)
				a2 += row[c2] * f;
		}
	}
	T* out = y + (uint64_t)i * R;
	if (u0)
		out[c0] = a0;
	if (u1)
		out[c1] = a1;
	if (u2)
		out[c2] = a2;
}

// layout changes between the row-major (sparse) and column-major (dense) sets
// (position p holds patch perm[p]; null perm / pinv: the identity)
template <class A, class B>
__global__ void k_to_colmajor(const A* __restrict__ x, uint32_t N, uint32_t Np, uint32_t R,
                              const uint32_t* __restrict__ perm, B* __restrict__ xT)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)R * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), p = (uint32_t)(t % Np);
	uint32_t j = perm ? perm[p] : p;
	xT[t] = j < N ? (B)x[(uint64_t)j * R + c] : (B)0;
}

template <class A, class B>
__global__ void k_to_rowmajor(const A* __restrict__ xT, uint32_t N, uint32_t Np, uint32_t R,
                              const uint32_t* __restrict__ pinv, B* __restrict__ x)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)N * R)
		return;
	uint32_t j = (uint32_t)(t / R), c = (uint32_t)(t % R);
	x[t] = (B)xT[(uint64_t)c * Np + (pinv ? pinv[j] : j)];
}

template <class A, class B> __global__ void k_copy_cast(uint64_t n, const A* __restrict__ x, B* __restrict__ y)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t < n)
		y[t] = (B)x[t];
}

template <class A, class B> __global__ void k_axpy(uint64_t n, const A* __restrict__ x, B* __restrict__ y)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t < n)
		y[t] += (B)x[t];
}

// one step of the adjoint sweep (row-major N x R, FP64), with u = H ybar_k:
// dalb += area * u * S_{k-1}, ybar_{k-1} = diffuse_c * area * u, lam += ybar_{k-1}
__global__ void k_adj_step(uint32_t N, uint32_t R, const double* __restrict__ u, const double* __restrict__ sprev,
                           const float* __restrict__ diffuse, const float* __restrict__ area,
                           double* __restrict__ dalb, double* __restrict__ ybar, double* __restrict__ lam)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)N * R)
		return;
	uint32_t i = (uint32_t)(t / R), c = (uint32_t)(t % R);
	double au = (double)area[i] * u[t];
	dalb[t] += au * sprev[t];
	double yb = (double)diffuse[3 * i + c % 3] * au;
	ybar[t] = yb;
	lam[t] += yb;
}

// r = a - b, formed in FP64
__global__ void k_resid64(uint64_t n, const float* __restrict__ a, const float* __restrict__ b,
                          double* __restrict__ r)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t < n)
		r[t] = (double)a[t] - (double)b[t];
}

// ===========================================================================
// host driver
// ===========================================================================
struct Ev {
	hipEvent_t a, b;
	hipStream_t s;
	explicit Ev(hipStream_t st) : s(st)
	{
		HIPCHECK(hipEventCreate(&a));
		HIPCHECK(hipEventCreate(&b));
		HIPCHECK(hipEventRecord(a, s));
	}
	double stop()
	{
		HIPCHECK(hipEventRecord(b, s));
		HIPCHECK(hipEventSynchronize(b));
		float ms = 0;
		HIPCHECK(hipEventElapsedTime(&ms, a, b));
		return ms;
	}
	~Ev()
	{
		(void)hipEventDestroy(a);
		(void)hipEventDestroy(b);
	}
};

RadSolver::RadSolver(DeviceScene& d) : dev(d)
{
	if (!d.rad.rowptr64)
		throw std::runtime_error("RadSolver: build the links first");
	if (!d.rad.rowptr)
		throw std::runtime_error("RadSolver: the solver forms take fewer than 2^32 link entries (" +
		                         std::to_string(d.rad.nnz) + " given)");
	N = (uint32_t)d.npatches;
	Np = ((N + kPadM - 1) / kPadM) * kPadM;
	if (Np < 512)
		Np = 512;
	s_ = d.stream();
	int dv = 0, ncu = 0;
	HIPCHECK(hipGetDevice(&dv));
	HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dv));
	ncu_ = ncu > 0 ? (uint32_t)ncu : 256;
}

RadSolver::~RadSolver() { release(); }

void RadSolver::release()
{
	chain_x_ = nullptr;
	dfree(g32_);
	dfree(g16_);
	dfree(g16lo_);
	dfree(g32t_);
	dfree(tmap_);
	dfree(tks_);
	dfree(chunks_);
	dfree(segoff_);
	dfree(nslot_);
	dfree(perm_);
	dfree(pinv_);
	dfree(E_);
	dfree(diff_);
	dfree(area_);
	dfree(part_);
	dfree(xq_);
	dfree(xqlo_);
	dfree(xT_);
	dfree(yT_);
	dfree(rm32a_);
	dfree(rm32b_);
	dfree(rmx_);
	dfree(rmx64_);
	dfree(cmax_);
	dfree(dperm_);
	dfree(aperm_);
	dfree(orow_);
	dfree(ocol_);
	dfree(oval_);
}

void RadSolver::set_problem(const double* E, uint32_t R_, const float* diffuse3, const float* area)
{
	chain_x_ = nullptr;
	if (R_ == 0 || R_ > (uint32_t)kRPad)
		throw std::runtime_error("RadSolver: 1..192 right-hand sides");
	if (R_ != R) {
		dfree(rm32a_);
		dfree(rm32b_);
		dfree(rmx_);
		dfree(rmx64_);
	}
	R = R_;
	hipStream_t s = (hipStream_t)s_;
	dfree(E_);
	dfree(diff_);
	dfree(area_);
	pcol_ok_ = false;
	E_ = dalloc<double>((size_t)N * R);
	diff_ = dalloc<float>(3 * (size_t)N);
	area_ = dalloc<float>((size_t)Np);
	HIPCHECK(hipMemcpyAsync(E_, E, 8 * (size_t)N * R, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(diff_, diffuse3, 12 * (size_t)N, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(area_, 0, 4 * (size_t)Np, s));
	HIPCHECK(hipMemcpyAsync(area_, area, 4 * (size_t)N, hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
}

// Spread the bits of a 10-bit integer to every third bit.
static inline uint32_t spread3(uint32_t v)
{
	v &= 0x3ffu;
	v = (v | (v << 16)) & 0x030000ffu;
	v = (v | (v << 8)) & 0x0300f00fu;
	v = (v | (v << 4)) & 0x030c30c3u;
	v = (v | (v << 2)) & 0x09249249u;
	return v;
}

// The 16-bit forms multiply H' = Pi H Pi^T, the patches taken along a Morton
// curve of their positions (ties in index order, the padding last), so that
// a 256-row block and a 32-column stage are each a compact patch of space.
// Only the tiles (256 rows x one 32-column k stage) that hold a link are
// stored: the k stages of the stored tiles (tks_, row blocks in order,
// stages ascending), for the build each tile's place in the store (tmap_,
// -1 for none), and the work runs over them (chunks_, nslot_).  Links reach some metres (f >= 0.001), so
// where patches are spread over space most tiles are empty; a small room
// whose patches all see each other stores nearly every tile.
void RadSolver::build_tiles()
{
	if (tiles_built_)
		return;
	hipStream_t s = (hipStream_t)s_;
	const uint32_t nrb = Np / 256, nks = Np / 32;
	const uint64_t n = (uint64_t)nrb * nks;
	{
		std::vector<float> pos(3 * (size_t)N);
		if (N)
			HIPCHECK(hipMemcpyAsync(pos.data(), dev.rad.pos, 12 * (size_t)N, hipMemcpyDeviceToHost, s));
		HIPCHECK(hipStreamSynchronize(s));
		float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (uint32_t i = 0; i < N; ++i)
			for (int a = 0; a < 3; ++a) {
				mn[a] = fminf(mn[a], pos[3 * i + a]);
				mx[a] = fmaxf(mx[a], pos[3 * i + a]);
			}
		const float ext = fmaxf(fmaxf(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
		const float q = ext > 0.0f ? 1023.0f / ext : 0.0f;
		std::vector<uint64_t> key(N);
		for (uint32_t i = 0; i < N; ++i) {
			uint32_t c = 0;
			for (int a = 0; a < 3; ++a) {
				float f = (pos[3 * i + a] - mn[a]) * q;
				uint32_t u = f > 0.0f ? (uint32_t)fminf(f, 1023.0f) : 0u; // NaN -> 0
				c |= spread3(u) << a;
			}
			key[i] = ((uint64_t)c << 32) | i;
		}
		std::sort(key.begin(), key.end());
		std::vector<uint32_t> perm(Np), pinv(Np);
		for (uint32_t p = 0; p < Np; ++p)
			perm[p] = p < N ? (uint32_t)key[p] : p;
		for (uint32_t p = 0; p < Np; ++p)
			pinv[perm[p]] = p;
		perm_ = dalloc<uint32_t>(Np);
		pinv_ = dalloc<uint32_t>(Np);
		HIPCHECK(hipMemcpyAsync(perm_, perm.data(), 4 * (size_t)Np, hipMemcpyHostToDevice, s));
		HIPCHECK(hipMemcpyAsync(pinv_, pinv.data(), 4 * (size_t)Np, hipMemcpyHostToDevice, s));
		HIPCHECK(hipStreamSynchronize(s));
	}
	std::vector<uint8_t> f(n);
	{
		DevBuf<uint8_t> df(n);
		HIPCHECK(hipMemsetAsync(df, 0, n, s));
		k_tile_flags<<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, pinv_, df);
		HIPCHECK(hipGetLastError());
		HIPCHECK(hipMemcpyAsync(f.data(), df, n, hipMemcpyDeviceToHost, s));
		HIPCHECK(hipStreamSynchronize(s));
	}
	std::vector<int32_t> map(n, -1);
	std::vector<uint32_t> off(nrb + 1, 0), ks;
	uint32_t maxcnt = 0;
	for (uint32_t rb = 0; rb < nrb; ++rb) {
		for (uint32_t w = 0; w < nks; ++w)
			if (f[(uint64_t)rb * nks + w]) {
				map[(uint64_t)rb * nks + w] = (int32_t)ks.size();
				ks.push_back(w);
			}
		off[rb + 1] = (uint32_t)ks.size();
		maxcnt = std::max(maxcnt, off[rb + 1] - off[rb]);
	}
	ntiles_ = ks.size();
	// Work: the stored tiles in row-block order, split into nblocks_ equal
	// contiguous shares, one per persistent block (the products hold one
	// block per compute unit); a share is cut where it crosses from one row
	// block to the next, and each piece (segment {row block, first tile, end
	// tile, slot}) writes its own partial-sum slot of its row block.  Every
	// block then multiplies the same number of tiles.
	(void)maxcnt;
	// at least kMinShare tiles per block: small stores use fewer blocks
	// rather than many one-tile pieces (every piece is a partial-sum slot)
	nblocks_ = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ncu_, (ntiles_ + kMinShare - 1) / kMinShare));
	std::vector<uint32_t> ch, segoff(nblocks_ + 1, 0), nsl(nrb, 0);
	for (uint32_t b = 0, rb = 0; b < nblocks_; ++b) {
		uint64_t t0 = ntiles_ * b / nblocks_;
		const uint64_t t1 = ntiles_ * (b + 1) / nblocks_;
		segoff[b] = (uint32_t)(ch.size() / 4);
		while (t0 < t1) {
			while (off[rb + 1] <= t0)
				++rb; // the row block holding tile t0 (empty ones skipped)
			const uint64_t e = std::min<uint64_t>(t1, off[rb + 1]);
			ch.insert(ch.end(), {rb, (uint32_t)t0, (uint32_t)e, nsl[rb]++});
			t0 = e;
		}
	}
	segoff[nblocks_] = (uint32_t)(ch.size() / 4);
	maxslot_ = nsl.empty() ? 0 : *std::max_element(nsl.begin(), nsl.end());
	ensure_part(); // the products write maxslot_ partial-sum planes
	nchunks_ = ch.size() / 4;
	tmap_ = dalloc<int32_t>(n);
	tks_ = dalloc<uint32_t>(ntiles_);
	chunks_ = dalloc<uint32_t>(ch.size());
	segoff_ = dalloc<uint32_t>(segoff.size());
	nslot_ = dalloc<uint32_t>(nrb);
	HIPCHECK(hipMemcpyAsync(segoff_, segoff.data(), 4 * segoff.size(), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(tmap_, map.data(), 4 * n, hipMemcpyHostToDevice, s));
	if (ntiles_)
		HIPCHECK(hipMemcpyAsync(tks_, ks.data(), 4 * ntiles_, hipMemcpyHostToDevice, s));
	if (nchunks_)
		HIPCHECK(hipMemcpyAsync(chunks_, ch.data(), 4 * ch.size(), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(nslot_, nsl.data(), 4 * (size_t)nrb, hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	tiles_built_ = true;
}

// Storage of a form's matrix (the compensated forms multiply the 16-bit
// stores with their remainder planes) and whether its product is the
// error-compensated one.
static int storage_of(int form) { return form == kF16x3 ? kDenseF16 : form == kBF16x3 ? kDenseBF16 : form; }
static bool split_of(int form) { return form == kF16x3 || form == kBF16x3; }

// The outlier threshold from the factors' exponent histogram (whole binades
// from the top, within kOutBinades of the largest factor and at most
// max(nnz / kOutCapDiv, kOutCapMin) factors), the outliers' CSR by Morton
// position, and the largest exponent left in the tiles.
void RadSolver::build_outliers()
{
	if (out_built_)
		return;
	hipStream_t s = (hipStream_t)s_;
	DevBuf<unsigned> dh(kHistBins);
	HIPCHECK(hipMemsetAsync(dh, 0, 4 * kHistBins, s));
	k_link_hist<<<1024, kB, 0, s>>>(dev.rad.nnz, dev.rad.val, dh);
	std::vector<unsigned> h(kHistBins);
	HIPCHECK(hipMemcpyAsync(h.data(), dh, 4 * kHistBins, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	int top = kHistBins - 1;
	while (top >= 0 && !h[top])
		--top;
	const uint64_t cap = std::max(dev.rad.nnz / kOutCapDiv, kOutCapMin);
	int lowest = top + 1; // lowest outlier bin
	uint64_t cum = 0;
	for (int b = top; b >= 0 && b > top - kOutBinades && cum + h[b] <= cap; --b) {
		cum += h[b];
		lowest = b;
	}
	othr_ = (top >= 0 && lowest <= top) ? ldexpf(1.0f, lowest - kHistBias) : INFINITY;
	int bulk = lowest - 1;
	while (bulk >= 0 && !h[bulk])
		--bulk;
	bulk_exp_ = bulk >= 0 ? bulk - kHistBias : -1000;
	nout_ = cum;
	dfree(orow_);
	dfree(ocol_);
	dfree(oval_);
	if (nout_) {
		DevBuf<uint32_t> cnt(Np);
		HIPCHECK(hipMemsetAsync(cnt, 0, 4 * (size_t)Np, s));
		k_outlier_count<<<gridn(N, kB / 64), kB, 0, s>>>(N, dev.rad.rowptr, dev.rad.val, othr_, pinv_, cnt);
		std::vector<uint32_t> c(Np), off(Np + 1, 0);
		HIPCHECK(hipMemcpyAsync(c.data(), cnt, 4 * (size_t)Np, hipMemcpyDeviceToHost, s));
		HIPCHECK(hipStreamSynchronize(s));
		for (uint32_t p = 0; p < Np; ++p)
			off[p + 1] = off[p] + c[p];
		if (off[Np] != nout_)
			throw std::runtime_error("RadSolver: outlier count differs from the histogram");
		orow_ = dalloc<uint32_t>(Np + 1);
		ocol_ = dalloc<uint32_t>(nout_);
		oval_ = dalloc<float>(nout_);
		HIPCHECK(hipMemcpyAsync(orow_, off.data(), 4 * ((size_t)Np + 1), hipMemcpyHostToDevice, s));
		k_outlier_fill<<<gridn(N, kB / 64), kB, 0, s>>>(N, dev.rad.rowptr, dev.rad.col, dev.rad.val, othr_, pinv_,
		                                                 orow_, ocol_, oval_);
		HIPCHECK(hipStreamSynchronize(s)); // off lives on this stack
	}
	out_built_ = true;
}

void RadSolver::build_dense(int form)
{
	hipStream_t s = (hipStream_t)s_;
	const int st = storage_of(form);
	Ev ev(s);
	if (st == kDenseF32) {
		if (!g32_)
			g32_ = dalloc<float>((uint64_t)Np * Np);
		HIPCHECK(hipMemsetAsync(g32_, 0, 4 * (uint64_t)Np * Np, s));
		if (N)
			k_dense_build<<<gridn(N), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val, g32_);
	} else if (st == kTileF32) {
		build_tiles();
		if (!g32t_)
			g32t_ = dalloc<float>(ntiles_ * kT32Tile);
		HIPCHECK(hipMemsetAsync(g32t_, 0, 4 * ntiles_ * kT32Tile, s));
		k_tile_scatter_f32<<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val, pinv_,
		                                                    tmap_, g32t_);
	} else {
		build_tiles();
		if (!g16_)
			g16_ = dalloc<uint16_t>(ntiles_ * kGStageElems);
		if (!g16lo_)
			g16lo_ = dalloc<uint16_t>(ntiles_ * kGStageElems);
		HIPCHECK(hipMemsetAsync(g16_, 0, 2 * ntiles_ * kGStageElems, s));
		HIPCHECK(hipMemsetAsync(g16lo_, 0, 2 * ntiles_ * kGStageElems, s));
		build_outliers();
		gscale_ = 1.0f;
		if (st == kDenseF16) {
			// scale so that the largest factor in the tiles sits near 2^13:
			// small factors stay in FP16's normal range
			gscale_ = bulk_exp_ > -1000 ? ldexpf(1.0f, 13 - bulk_exp_) : 1.0f;
			k_tile_scatter<1><<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val,
			                                                   pinv_, tmap_, gscale_, othr_, g16_, g16lo_);
		} else {
			k_tile_scatter<0><<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val,
			                                                   pinv_, tmap_, 1.0f, othr_, g16_, g16lo_);
		}
	}
	HIPCHECK(hipGetLastError());
	build_ms = ev.stop();
	built_form_ = st;
}

uint32_t RadSolver::ksplit(uint32_t rows_per_block, uint32_t min_blocks) const
{
	// split k (in whole stages) until every CU has a block
	uint32_t rows = Np / rows_per_block, s = 1;
	while (rows * s < min_blocks && (Np / (s * 2)) % 32 == 0 && s < 64)
		s *= 2;
	return s;
}

// Partial-sum planes (kRPad x Np floats each): as many as the dense FP32
// form's k split or the tile products' most segments of one row block.
void RadSolver::ensure_part()
{
	const uint32_t need = std::max({ksplit(kTileM, 2048), kMaxSlots, maxslot_});
	if (part_ && need <= part_slots_)
		return;
	dfree(part_);
	part_slots_ = 0;
	part_ = dalloc<float>((uint64_t)need * kRPad * Np);
	part_slots_ = need;
}

// work buffers of the products (kept across solves)
void RadSolver::ensure_work()
{
	const uint64_t n = (uint64_t)kRPad * Np;
	ensure_part();
	if (!xq_) {
		xq_ = dalloc<float>(n); // FP32 operand, or the 16-bit one in its first half
		xqlo_ = dalloc<uint16_t>(n);
		xT_ = dalloc<float>(n);
		yT_ = dalloc<float>(n);
		cmax_ = dalloc<unsigned>(2 * kRPad);
	}
	if (!rm32a_) {
		rm32a_ = dalloc<float>((uint64_t)Np * R);
		rm32b_ = dalloc<float>((uint64_t)Np * R);
		rmx_ = dalloc<float>((uint64_t)Np * R);
		rmx64_ = dalloc<double>((uint64_t)Np * R);
	}
}

// y = K x (pre) or H x over the FP32 tile store (Morton order in and out)
// Diffuse and area in Morton order for the tile forms' operand preparation
// (after set_problem and the tile build).
void RadSolver::ensure_perm_colors()
{
	if (pcol_ok_)
		return;
	if (!dperm_)
		dperm_ = dalloc<float>(3 * (size_t)Np);
	if (!aperm_)
		aperm_ = dalloc<float>((size_t)Np);
	k_permute_colors<<<gridn(Np), kB, 0, (hipStream_t)s_>>>(diff_, area_, perm_, N, Np, dperm_, aperm_);
	HIPCHECK(hipGetLastError());
	pcol_ok_ = true;
}

void RadSolver::apply_tiles_f32(const float* xT, float* yT, bool pre)
{
	hipStream_t s = (hipStream_t)s_;
	if (part_slots_ < maxslot_)
		throw std::runtime_error("RadSolver: partial-sum planes fewer than the tile segments need");
	ensure_perm_colors();
	k_prep_x32t<<<gridn((uint64_t)kRPad * (Np / 4)), kB, 0, s>>>(xT, dperm_, aperm_, Np, R, pre ? 1 : 0, xq_);
	if (nchunks_)
		k_gemm_f32t<<<nblocks_, kB, 0, s>>>(g32t_, xq_, reinterpret_cast<const uint4*>(chunks_), segoff_, tks_, Np,
		                                     part_);
	k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part_, 0, nslot_, Np, R, nullptr, 1.0f, yT);
	HIPCHECK(hipGetLastError());
}

// y (column-major, R x Np, FP32) = K x (pre) or H x for the dense forms; the
// tile forms take and give their vectors in Morton order (position p holds
// patch perm_[p]), the dense FP32 form in patch order.
// split: the error-compensated product of the 16-bit forms (hi + lo planes).
void RadSolver::apply_dense(int form, const float* xT, float* yT, bool split, bool pre, bool chain)
{
	hipStream_t s = (hipStream_t)s_;
	const int st = storage_of(form);
	split = split || split_of(form);
	uint64_t nq = (uint64_t)kRPad * Np;
	const int key = st * 4 + (split ? 2 : 0) + (pre ? 1 : 0);
	const bool ready = chain_x_ == xT && chain_key_ == key;
	chain_x_ = nullptr;
	if (st == kTileF32) {
		apply_tiles_f32(xT, yT, pre);
		return;
	}
	if (st == kDenseF32) {
		uint32_t ks = ksplit(kTileM, 2048);
		dim3 grid(Np / kTileM, ks);
		k_prep_x32<<<gridn(nq), kB, 0, s>>>(xT, diff_, area_, N, Np, R, pre ? 1 : 0, xq_);
		k_gemm_f32<<<grid, kB, 0, s>>>(g32_, xq_, Np, ks, part_);
		k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part_, ks, nullptr, Np, R, nullptr, 1.0f, yT);
	} else {
		const bool f16 = st == kDenseF16;
		uint16_t* q = reinterpret_cast<uint16_t*>(xq_);
		uint16_t* lo = split ? xqlo_ : nullptr;
		unsigned* cm = f16 ? cmax_ + cmax_cur_ * kRPad : nullptr;
		ensure_perm_colors();
		const unsigned og = (Np / 256) * (kRPad / 16); // k_operand blocks
		OperandPass op;
		op.x = xT;
		op.dperm = dperm_;
		op.aperm = aperm_;
		op.pre = pre ? 1 : 0;
		op.Np = Np;
		op.R = R;
		if (f16 && !ready) { // column maxima first: they set the operand's scales
			HIPCHECK(hipMemsetAsync(cm, 0, 4 * kRPad, s));
			op.cmax_out = cm;
			k_operand<1><<<og, kB, 0, s>>>(op);
			op.cmax_out = nullptr;
		}
		if (f16 || !ready) {
			op.cmax_op = cm;
			op.xq = q;
			op.xlo = lo;
			if (f16)
				k_operand<1><<<og, kB, 0, s>>>(op);
			else
				k_operand<0><<<og, kB, 0, s>>>(op);
		}
		if (part_slots_ < maxslot_)
			throw std::runtime_error("RadSolver: partial-sum planes fewer than the tile segments need");
		if (nchunks_) {
			const dim3 grid(nblocks_);
			const uint4* chs = reinterpret_cast<const uint4*>(chunks_);
			if (f16 && split)
				k_gemm_tiles<1, 2, kTilesD2><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, chs, segoff_, tks_, Np, part_);
			else if (f16)
				k_gemm_tiles<1, 1, kTilesD1><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, chs, segoff_, tks_, Np,
				                                                 part_);
			else if (split)
				k_gemm_tiles<0, 2, kTilesD2><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, chs, segoff_, tks_, Np, part_);
			else
				k_gemm_tiles<0, 1, kTilesD1><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, chs, segoff_, tks_, Np,
				                                                 part_);
		}
		// the reduction; chained, it also prepares the next product's operand
		// (BF16) or the column maxima that scale it (FP16)
		OperandPass rd;
		rd.part = part_;
		rd.nslot = nslot_;
		rd.cmax_in = cm;
		rd.gscale = gscale_;
		rd.y = yT;
		rd.dperm = dperm_;
		rd.aperm = aperm_;
		rd.pre = pre ? 1 : 0;
		rd.Np = Np;
		rd.R = R;
		if (nout_) {
			rd.orow = orow_;
			rd.ocol = ocol_;
			rd.oval = oval_;
			rd.xin = xT;
		}
		if (chain && f16) {
			cmax_cur_ ^= 1;
			rd.cmax_out = cmax_ + cmax_cur_ * kRPad;
			HIPCHECK(hipMemsetAsync(rd.cmax_out, 0, 4 * kRPad, s));
		} else if (chain) {
			rd.xq = q;
			rd.xlo = lo;
		}
		if (f16)
			k_operand<1><<<og, kB, 0, s>>>(rd);
		else
			k_operand<0><<<og, kB, 0, s>>>(rd);
		if (chain) {
			chain_x_ = yT;
			chain_key_ = key;
		}
	}
	HIPCHECK(hipGetLastError());
}

// y = K x (pre) or H x, row-major N x R FP64 in and out, in the given form
void RadSolver::apply_op(int form, const double* x, double* y, bool pre)
{
	hipStream_t s = (hipStream_t)s_;
	const uint64_t nR = (uint64_t)N * R, nRp = (uint64_t)R * Np;
	const unsigned sgrid = gridn(N, kB / 64);
	const uint32_t* rp = dev.rad.rowptr;
	const uint32_t* cl = dev.rad.col;
	const float* vl = dev.rad.val;
	if (form == kSparseF64) {
		if (pre)
			k_prep_rows<double, true><<<gridn(nR), kB, 0, s>>>(nR, R, x, diff_, area_, rmx64_);
		else
			k_prep_rows<double, false><<<gridn(nR), kB, 0, s>>>(nR, R, x, diff_, area_, rmx64_);
		k_gather<double><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, rmx64_, y);
	} else if (form == kSparseF32) {
		k_copy_cast<double, float><<<gridn(nR), kB, 0, s>>>(nR, x, rm32a_);
		if (pre)
			k_prep_rows<float, true><<<gridn(nR), kB, 0, s>>>(nR, R, rm32a_, diff_, area_, rmx_);
		else
			k_prep_rows<float, false><<<gridn(nR), kB, 0, s>>>(nR, R, rm32a_, diff_, area_, rmx_);
		k_gather<float><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, rmx_, rm32b_);
		k_copy_cast<float, double><<<gridn(nR), kB, 0, s>>>(nR, rm32b_, y);
	} else {
		const bool mo = storage_of(form) != kDenseF32; // Morton order
		k_to_colmajor<double, float><<<gridn(nRp), kB, 0, s>>>(x, N, Np, R, mo ? perm_ : nullptr, xT_);
		apply_dense(form, xT_, yT_, false, pre);
		k_to_rowmajor<float, double><<<gridn(nR), kB, 0, s>>>(yT_, N, Np, R, mo ? pinv_ : nullptr, y);
	}
	HIPCHECK(hipGetLastError());
}

void RadSolver::solve(int form, int bounces, int refine, double* total_out)
{
	chain_x_ = nullptr;
	if (form < 0 || form >= kNumForms)
		throw std::runtime_error("RadSolver: unknown form " + std::to_string(form));
	if (refine > 0 && form != kDenseBF16 && form != kDenseF16)
		throw std::runtime_error("RadSolver: refinement sweeps apply to the plain 16-bit forms (3, 4)");
	hipStream_t s = (hipStream_t)s_;
	uint64_t nR = (uint64_t)N * R, nRp = (uint64_t)R * Np;
	times = RadTimes{};
	ensure_work();
	// FP64 running total
	DevBuf<double> tot(nR);
	HIPCHECK(hipMemcpyAsync(tot, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	if (form == kSparseF64 || form == kSparseF32) {
		DevBuf<double> x(nR), y(nR);
		HIPCHECK(hipMemcpyAsync(x, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
		Ev ev(s);
		for (int b = 0; b < bounces; ++b) {
			apply_op(form, x, y, true);
			k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, y, tot);
			std::swap(x, y);
		}
		times.series_ms = ev.stop();
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
	} else {
		if (built_form_ != storage_of(form))
			build_dense(form);
		times.build_ms = build_ms;
		// z_k (column-major FP32), k = 0..b
		std::vector<DevBuf<float>> z;
		for (int b = 0; b <= bounces; ++b)
			z.emplace_back(nRp);
		// the tile forms keep z in Morton order
		const bool mo = storage_of(form) != kDenseF32;
		const uint32_t* po = mo ? perm_ : nullptr;
		const uint32_t* pi = mo ? pinv_ : nullptr;
		k_to_colmajor<double, float><<<gridn(nRp), kB, 0, s>>>(E_, N, Np, R, po, z[0]);
		{
			Ev ev(s);
			for (int b = 1; b <= bounces; ++b)
				apply_dense(form, z[b - 1], z[b], false, true, b < bounces);
			times.series_ms = ev.stop();
		}
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
		// Iterative refinement on the block system z_b = K z_{b-1}, z_0 = E.
		// One sweep: residuals rho_b = K z_{b-1} - z_b of the current iterate,
		// with K z_{b-1} from the error-compensated (hi + lo) MFMA product and
		// the difference taken in FP64; corrections d_b = Kdense d_{b-1} + rho_b
		// (d_0 = 0: z_0 = E is exact) with the plain 16-bit product; then
		// z += d.  z_{b-1} is corrected only after rho_b has used its old value.
		DevBuf<double> zr(nR);
		if (refine > 0) {
			DevBuf<double> rho(nRp);
			DevBuf<float> kz(nRp), d(nRp), dprev(nRp);
			Ev ev(s);
			for (int it = 0; it < refine; ++it) {
				HIPCHECK(hipMemsetAsync(dprev, 0, 4 * nRp, s));
				for (int b = 1; b <= bounces; ++b) {
					apply_dense(form, z[b - 1], kz, true, true);                          // K z_{b-1}
					k_resid64<<<gridn(nRp), kB, 0, s>>>(nRp, kz, z[b], rho);             // rho_b (FP64)
					apply_dense(form, dprev, d, false, true);                             // Kdense d_{b-1}
					k_axpy<double, float><<<gridn(nRp), kB, 0, s>>>(nRp, rho, d);        // d_b
					k_axpy<float, float><<<gridn(nRp), kB, 0, s>>>(nRp, dprev, z[b - 1]); // apply d_{b-1}
					std::swap(d, dprev);
				}
				k_axpy<float, float><<<gridn(nRp), kB, 0, s>>>(nRp, dprev, z[bounces]);
			}
			times.refine_ms = ev.stop();
		}
		HIPCHECK(hipGetLastError());
		// total = sum_k z_k in FP64 (z_0 = E exactly)
		for (int b = 1; b <= bounces; ++b) {
			k_to_rowmajor<float, double><<<gridn(nR), kB, 0, s>>>(z[b], N, Np, R, pi, zr);
			k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, zr, tot);
		}
	}
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(total_out, tot, 8 * nR, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
}

// Reverse mode of T = sum_{k=0..B} K^k E, K = H diag(area) diag(diffuse_c).
// Forward partial sums S_k = sum_{m<=k} K^m E are kept (T = S_B is the Horner
// recurrence S_k = E + K S_{k-1}); its adjoint, from ybar_B = g:
//   u_k = H ybar_k            (H symmetric, so K^T = diag(diffuse_c) diag(area) H)
//   dalb += area * u_k * S_{k-1}
//   ybar_{k-1} = diffuse_c * area * u_k
//   lam = sum_k ybar_k        (= dL/dE; also dL/d emission)
// Every product, forward and reverse, is the form's own (the compensated
// forms' three MFMA products included).
void RadSolver::solve_grad(int form, int bounces, const double* g, double* total_out, double* lam_out,
                           double* dalb_out)
{
	chain_x_ = nullptr;
	if (form < 0 || form >= kNumForms)
		throw std::runtime_error("RadSolver: unknown form " + std::to_string(form));
	hipStream_t s = (hipStream_t)s_;
	const uint64_t nR = (uint64_t)N * R;
	times = RadTimes{};
	ensure_work();
	if (form >= kDenseF32 && built_form_ != storage_of(form))
		build_dense(form);
	const int B = std::max(0, bounces);
	std::vector<DevBuf<double>> S;
	for (int k = 0; k < std::max(1, B); ++k)
		S.emplace_back(nR);
	DevBuf<double> z(nR), zn(nR), tot(nR), ybar(nR), u(nR), lam(nR), dalb(nR);
	Ev ev(s);
	// forward: S_0 = E, z_k = K z_{k-1}, S_k = S_{k-1} + z_k
	HIPCHECK(hipMemcpyAsync(S[0], E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	HIPCHECK(hipMemcpyAsync(z, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	HIPCHECK(hipMemcpyAsync(tot, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	for (int k = 1; k <= B; ++k) {
		apply_op(form, z, zn, true);
		k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, zn, tot);
		if (k < B)
			HIPCHECK(hipMemcpyAsync(S[k], tot, 8 * nR, hipMemcpyDeviceToDevice, s));
		std::swap(z, zn);
	}
	times.series_ms = ev.stop();
	// reverse sweep
	Ev ev2(s);
	HIPCHECK(hipMemcpyAsync(ybar, g, 8 * nR, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(lam, ybar, 8 * nR, hipMemcpyDeviceToDevice, s));
	HIPCHECK(hipMemsetAsync(dalb, 0, 8 * nR, s));
	for (int k = B; k >= 1; --k) {
		apply_op(form, ybar, u, false);
		k_adj_step<<<gridn(nR), kB, 0, s>>>(N, R, u, S[k - 1], diff_, area_, dalb, ybar, lam);
	}
	times.refine_ms = ev2.stop(); // the reverse sweep's time
	times.per_bounce_ms = times.refine_ms / std::max(1, B);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(total_out, tot, 8 * nR, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipMemcpyAsync(lam_out, lam, 8 * nR, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipMemcpyAsync(dalb_out, dalb, 8 * nR, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
}

} // namespace lgx
