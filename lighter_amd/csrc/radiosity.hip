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

// form factor of the pair, lower index as A (lighter.cpp:740-749)
__device__ __forceinline__ float link_factor(const float* pos, const float* nrm, uint32_t i, uint32_t j)
{
	uint32_t a = i < j ? i : j, b = i < j ? j : i;
	f3 d = ld3(pos + 3 * b) - ld3(pos + 3 * a);
	float da = dot(ld3(nrm + 3 * a), d);
	float db = dot(ld3(nrm + 3 * b), -d);
	float l2 = lensq(d);
	return da * db / (l2 * l2 * kPiF);
}

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
// links), row-major Np x Np, from the occlusion bitmask; one thread per
// 32-column word
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_dense_build(const uint32_t* __restrict__ bits, uint64_t words, uint32_t N,
                                                    uint32_t Np, const float* __restrict__ pos,
                                                    const float* __restrict__ nrm, float* __restrict__ G)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	uint64_t wpr = Np / 32; // words per padded row
	if (t >= (uint64_t)Np * wpr)
		return;
	uint32_t i = (uint32_t)(t / wpr), w = (uint32_t)(t % wpr);
	uint32_t m = (i < N && w < words) ? bits[(uint64_t)i * words + w] : 0u;
	float v[32];
#pragma unroll
	for (int b = 0; b < 32; ++b) {
		uint32_t j = w * 32 + b;
		v[b] = ((m >> b) & 1u) ? link_factor(pos, nrm, i, j) : 0.0f;
	}
	float4* o = reinterpret_cast<float4*>(G + (uint64_t)i * Np + (uint64_t)w * 32);
#pragma unroll
	for (int q = 0; q < 8; ++q)
		o[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
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
// the remainder planes.  LDS-image layout of a tile: row r = i' & 255 in
// 1 KB group r >> 4 at 64 bytes per row; column kk = j' & 31 in 16-byte piece
// kk >> 3 at slot (kk >> 3) ^ piece_swz(r & 15), element kk & 7.
template <int P>
__global__ __launch_bounds__(kB) void k_tile_scatter(uint32_t N, uint32_t Np, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col,
                                                     const float* __restrict__ val,
                                                     const uint32_t* __restrict__ pinv,
                                                     const int32_t* __restrict__ tmap, float scale,
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
		if (tile < 0)
			continue; // not reached: the flags came from the same links
		const uint64_t o = (uint64_t)tile * 8192 + inrow + (uint32_t)(((kk >> 3) ^ piece_swz(rr)) * 8) + (kk & 7);
		const float a = val[k] * scale;
		const uint16_t h = Prec<P>::cvt(a);
		G[o] = h;
		Glo[o] = Prec<P>::cvt(a - from16<P>(h));
	}
}

// largest link factor (for the FP16 scale); grid-stride over the stored
// links, one atomic per wave
__global__ __launch_bounds__(kB) void k_max_link(uint64_t nnz, const float* __restrict__ val, unsigned* out)
{
	float m = 0.0f;
	for (uint64_t k = (uint64_t)blockIdx.x * kB + threadIdx.x; k < nnz; k += (uint64_t)gridDim.x * kB)
		m = fmaxf(m, val[k]);
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		m = fmaxf(m, __shfl_xor(m, o));
	if ((threadIdx.x & 63) == 0)
		atomicMax(out, __float_as_uint(m)); // non-negative floats order like their bits
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

// Column maxima of the operand before scaling: block (x, c) takes positions
// x * 2048 .. of column c.  A maximum does not depend on the order of the
// atomics, so the result is deterministic; cmax starts at zero.
__global__ __launch_bounds__(kB) void k_col_max(const float* __restrict__ xT, const float* __restrict__ diffuse,
                                                const float* __restrict__ area, const uint32_t* __restrict__ perm,
                                                uint32_t N, uint32_t Np, int pre, unsigned* __restrict__ cmax)
{
	const uint32_t c = blockIdx.y, p0 = blockIdx.x * 2048, p1 = min(Np, p0 + 2048);
	float m = 0.0f;
	for (uint32_t p = p0 + threadIdx.x; p < p1; p += kB) {
		const uint32_t src = perm[p];
		if (src < N) {
			float v = xT[(uint64_t)c * Np + p];
			if (pre)
				v = (v * diffuse[3 * src + c % 3]) * area[src];
			m = fmaxf(m, fabsf(v));
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		m = fmaxf(m, __shfl_xor(m, o));
	if ((threadIdx.x & 63) == 0)
		atomicMax(cmax + c, __float_as_uint(m)); // non-negative floats order like their bits
}

// ---------------------------------------------------------------------------
// right-hand-side preparation at position j (patch perm[j], Morton order):
// Xq[c][j] = (x[c][j] * diffuse[perm[j]][c % 3]) * area[perm[j]] * s_c for
// K = H diag(area) diag(diffuse_c) (pre), or x[c][j] * s_c for the bare H
// (the adjoint sweep); s_c is the column scale
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kB) void k_prep_x16(const float* __restrict__ xT, const float* __restrict__ diffuse,
                                                 const float* __restrict__ area, const unsigned* __restrict__ cmax,
                                                 const uint32_t* __restrict__ perm, uint32_t N, uint32_t Np,
                                                 uint32_t R, int pre, uint16_t* __restrict__ xq,
                                                 uint16_t* __restrict__ xlo)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)kRPad * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), j = (uint32_t)(t % Np);
	const uint32_t src = perm[j];
	float v = 0.0f;
	if (c < R && src < N) {
		v = xT[(uint64_t)c * Np + j];
		if (pre)
			v = (v * diffuse[3 * src + c % 3]) * area[src];
		v *= col_scale_of(cmax, c);
	}
	uint16_t h = Prec<P>::cvt(v);
	// LDS-image layout: stage j / 32, column group c / 16, row c % 16, piece
	// (j % 32) / 8 at slot piece ^ piece_swz(c % 16), element j % 8
	const uint32_t kk = j & 31, r = c & 15;
	const uint64_t o = (((uint64_t)(j >> 5) * (kRPad / 16) + (c >> 4)) * 64 + r * 4 + ((kk >> 3) ^ piece_swz(r))) * 8 +
	                   (kk & 7);
	xq[o] = h;
	if (xlo)
		xlo[o] = Prec<P>::cvt(v - from16<P>(h));
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
// MFMA GEMM: part[s][c][i] = sum_{j in split s} G[i][j] * Xq[c][j]
//
// Block: 4 waves, each RT row tiles of 16 x all 12 column tiles (R padded to
// 192), v_mfma_f32_16x16x32 (bf16 or f16) with FP32 accumulators.  Lane l
// supplies row (l & 15) of A and column (l & 15) of B with the same eight k
// indices 8*(l>>4)+e, so the k order inside an instruction is immaterial;
// output element i of a 16x16 tile sits at row 4*(l>>4)+i, column l & 15.
//
// k advances in stages of 32.  A stage (the block's matrix rows x 32 and the
// right-hand-side chunk 192 x 32, per plane) goes HBM -> LDS by
// global_load_lds_dwordx4 (no register staging), PD stages in flight in a
// ring of PD LDS buffers.  LDS holds 16-row x 64-byte groups of 1 KB, one
// DMA instruction each; the 16-byte pieces of a row are permuted by
// piece_swz(row) so that the rows one ds_read_b128 serves together fall in
// distinct banks.  Each wave fetches its own matrix rows and a quarter of
// the right-hand-side chunk, waits on its own DMA count, then a barrier makes
// every wave's part visible.
// ---------------------------------------------------------------------------
constexpr int kKS = 32; // k per stage
constexpr int kMaxPD = 6; // deepest LDS ring

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

template <int NPL, int RT>
struct GemmShape {
	static constexpr int kRows = 64 * RT;                       // block rows
	static constexpr int kXBytes = kRPad * kKS * 2;             // right-hand-side plane per stage
	static constexpr int kGBytes = kRows * kKS * 2;             // matrix plane per stage
	static constexpr int kStage = NPL * (kXBytes + kGBytes);    // LDS bytes per stage
	static constexpr int kOps = NPL * (kColTiles / 4 + RT);     // DMA instructions per wave per stage
};

// Issue stage st (counted from the split's first stage) into LDS buffer
// base.  The 16-bit matrix and operand are stored as LDS images (see
// k_tile_scatter / k_prep_x16): every DMA instruction copies one contiguous
// KB, and a block's stage is one contiguous tile.
constexpr uint64_t kTileElems = 512;             // 16-bit elements per 1 KB group
constexpr uint64_t kGStageElems = 16 * kTileElems; // a 256-row stage of the matrix
constexpr uint64_t kXStageElems = (kRPad / 16) * kTileElems; // a stage of the operand

template <int NPL, int RT>
__device__ __forceinline__ void gemm_issue(const uint16_t* G, const uint16_t* Glo, const uint16_t* xq,
                                           const uint16_t* xlo, uint64_t xsrc, uint64_t gsrc, uint64_t gstride,
                                           const uint32_t* lst, uint32_t st, unsigned char* base, int wave)
{
	typedef GemmShape<NPL, RT> S;
	const uint64_t ks = lst[st]; // k stage of the block's st-th listed tile
#pragma unroll
	for (int pl = 0; pl < NPL; ++pl) {
		const uint16_t* X = pl ? xlo : xq;
		const uint16_t* M = pl ? Glo : G;
		unsigned char* xb = base + pl * S::kXBytes + wave * (kColTiles / 4) * 1024;
		unsigned char* gb = base + NPL * S::kXBytes + pl * S::kGBytes + wave * RT * 1024;
#pragma unroll
		for (int j = 0; j < kColTiles / 4; ++j)
			dma16(X + xsrc + ks * kXStageElems + j * kTileElems, xb + j * 1024);
#pragma unroll
		for (int t = 0; t < RT; ++t)
			dma16(M + gsrc + (uint64_t)st * gstride + t * kTileElems, gb + t * 1024);
	}
}

template <int NPL, int RT>
struct GemmLane {
	uint64_t xsrc, gsrc; // DMA source offsets (elements) of this lane at the split's first stage
	uint64_t gstride;    // matrix elements from one stage to the next
	const uint32_t* lst; // k stages of the split's tiles
	int wave, rofs;      // wave index; ds_read_b128 offset of the fragment in a 1 KB group
	unsigned char* buf[kMaxPD]; // the LDS stage ring
};

// one k stage (buffer B of PD): wait for it, free the previous buffer, issue
// the stage PD - 1 ahead into that buffer, multiply
template <int P, int NPL, int RT, int PD, int B>
__device__ __forceinline__ void gemm_step(int s, int nst, const GemmLane<NPL, RT>& L, const uint16_t* G,
                                          const uint16_t* Glo, const uint16_t* xq, const uint16_t* xlo,
                                          uint64_t Np, f32x4 (&acc)[RT][kColTiles])
{
	typedef GemmShape<NPL, RT> S;
	// stage s has landed once at most the later PD - 2 stages' DMAs are pending
	wait_vm<S::kOps * (PD - 2)>();
	block_barrier(); // every wave's part of stage s is in; stage s-1's buffer is free
	// all of this stage's fragments are read before the next DMA is issued:
	// the compiler cannot tell that DMA's buffer from this one and would
	// otherwise wait for every DMA in flight
	const unsigned char* base = L.buf[B];
	i16x8 a[NPL][RT], bx[NPL][kColTiles];
#pragma unroll
	for (int pl = 0; pl < NPL; ++pl) {
#pragma unroll
		for (int t = 0; t < RT; ++t)
			a[pl][t] = *reinterpret_cast<const i16x8*>(base + NPL * S::kXBytes + pl * S::kGBytes +
			                                           (L.wave * RT + t) * 1024 + L.rofs);
#pragma unroll
		for (int c = 0; c < kColTiles; ++c)
			bx[pl][c] = *reinterpret_cast<const i16x8*>(base + pl * S::kXBytes + c * 1024 + L.rofs);
	}
	// always PD - 1 stages in flight: past the end, the last stage is fetched
	// again into the freed buffer (never read), which keeps the wait counts
	// above exact
	constexpr int BN = (B + PD - 1) % PD;
	const int ahead = s + PD - 1 < nst ? s + PD - 1 : nst - 1;
	gemm_issue<NPL, RT>(G, Glo, xq, xlo, L.xsrc, L.gsrc, L.gstride, L.lst, ahead, L.buf[BN], L.wave);
#pragma unroll
	for (int c = 0; c < kColTiles; ++c) {
#pragma unroll
		for (int t = 0; t < RT; ++t)
			acc[t][c] = mfma16<P>(a[0][t], bx[0][c], acc[t][c]);
		if (NPL == 2) {
			// error-compensated: + G_hi X_lo + G_lo X_hi
#pragma unroll
			for (int t = 0; t < RT; ++t) {
				acc[t][c] = mfma16<P>(a[0][t], bx[NPL - 1][c], acc[t][c]);
				acc[t][c] = mfma16<P>(a[NPL - 1][t], bx[0][c], acc[t][c]);
			}
		}
	}
}

template <int P, int NPL, int RT, int PD, int... Bs>
__device__ __forceinline__ void gemm_steps(int s, const GemmLane<NPL, RT>& L, const uint16_t* G,
                                           const uint16_t* Glo, const uint16_t* xq, const uint16_t* xlo,
                                           uint64_t Np, f32x4 (&acc)[RT][kColTiles], int nst,
                                           std::integer_sequence<int, Bs...>)
{
	(gemm_step<P, NPL, RT, PD, Bs>(s + Bs, nst, L, G, Glo, xq, xlo, Np, acc), ...);
}

template <int NPL, int RT, int... Bs>
__device__ __forceinline__ void gemm_prologue(int nst, const GemmLane<NPL, RT>& L, const uint16_t* G,
                                              const uint16_t* Glo, const uint16_t* xq, const uint16_t* xlo,
                                              uint64_t Np, std::integer_sequence<int, Bs...>)
{
	// stage Bs into buffer Bs (the last stage again where there are fewer)
	(gemm_issue<NPL, RT>(G, Glo, xq, xlo, L.xsrc, L.gsrc, L.gstride, L.lst, Bs < nst ? Bs : nst - 1, L.buf[Bs],
	                     L.wave),
	 ...);
}

template <int P, int NPL, int RT, int PD, int... Bs>
__device__ __forceinline__ void gemm_tail(int s, int nst, const GemmLane<NPL, RT>& L, const uint16_t* G,
                                          const uint16_t* Glo, const uint16_t* xq, const uint16_t* xlo,
                                          uint64_t Np, f32x4 (&acc)[RT][kColTiles],
                                          std::integer_sequence<int, Bs...>)
{
	((s + Bs < nst ? gemm_step<P, NPL, RT, PD, Bs>(s + Bs, nst, L, G, Glo, xq, xlo, Np, acc) : void()), ...);
}

// y = G x over the stored tiles.  A chunk {rb, tbeg, tend, slot} is a run of
// 256-row block rb's stored tiles (tks[t]: the k stage of tile t); blocks
// 256 / kRows at a time take its rows and write their partial sums to the
// slot's plane of part.
template <int P, int NPL, int RT, int PD>
__global__ __launch_bounds__(kB) void k_gemm_mfma(const uint16_t* __restrict__ G, const uint16_t* __restrict__ Glo,
                                                  const uint16_t* __restrict__ xq, const uint16_t* __restrict__ xlo,
                                                  const uint4* __restrict__ chunks, const uint32_t* __restrict__ tks,
                                                  uint32_t Np, float* __restrict__ part)
{
	typedef GemmShape<NPL, RT> S;
	static_assert(PD >= 2, "pipeline depth");
	const int tid = threadIdx.x, lane = tid & 63;
	const int r = lane & 15, g = lane >> 4;
	// one variable per stage buffer: the compiler's wait-count tracking can
	// then tell the DMA into one buffer from the reads of another (with a
	// single array every LDS read would wait for every DMA in flight)
	static_assert(PD <= kMaxPD, "pipeline depth");
	__shared__ __align__(1024) unsigned char lds0[S::kStage], lds1[S::kStage], lds2[PD > 2 ? S::kStage : 16],
	    lds3[PD > 3 ? S::kStage : 16], lds4[PD > 4 ? S::kStage : 16], lds5[PD > 5 ? S::kStage : 16];
	GemmLane<NPL, RT> L;
	L.buf[0] = lds0;
	L.buf[1] = lds1;
	L.buf[2] = lds2;
	L.buf[3] = lds3;
	L.buf[4] = lds4;
	L.buf[5] = lds5;
	L.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	constexpr uint32_t kBPC = 256 / S::kRows; // blocks per chunk
	// XCD-aware order: block b runs on XCD b % 8, so XCD x is given the x-th
	// contiguous eighth of the work; neighbouring chunks (neighbouring
	// patches in Morton order, hence overlapping operand stages) then share
	// that XCD's L2
	const uint32_t nb = gridDim.x, x = blockIdx.x % 8, per = nb / 8, rem = nb % 8;
	const uint32_t bid = x * per + (x < rem ? x : rem) + blockIdx.x / 8;
	const uint4 ch = chunks[bid / kBPC];
	const uint32_t row0 = ch.x * 256 + (bid % kBPC) * S::kRows; // the block's first row
	const uint32_t i0 = row0 + L.wave * 16 * RT;                       // this wave's first row
	const uint32_t gbase = (row0 & 255) >> 4;
	const int nst = (int)(ch.z - ch.y);
	// DMA sources in the LDS-image layouts: lane l copies the l-th 16 bytes
	// of each 1 KB group.  Stored tile t (16 KB) holds group g at
	// (t * 16 + g) KB; operand stage ks, column group cg at (ks * 12 + cg) KB.
	L.lst = tks + ch.y;
	L.gstride = kGStageElems;
	L.gsrc = ((uint64_t)ch.y * 16 + gbase + (uint64_t)L.wave * RT) * kTileElems + 8 * lane;
	L.xsrc = ((uint64_t)(kColTiles / 4) * L.wave) * kTileElems + 8 * lane;
	L.rofs = r * 64 + 16 * (g ^ piece_swz(r));

	f32x4 acc[RT][kColTiles];
#pragma unroll
	for (int t = 0; t < RT; ++t)
#pragma unroll
		for (int c = 0; c < kColTiles; ++c)
			acc[t][c] = f32x4{0, 0, 0, 0};

	if (nst > 0) { // a split without tiles writes zeros
		// prologue: stages 0 .. PD-2 into buffers 0 .. PD-2
		gemm_prologue<NPL, RT>(nst, L, G, Glo, xq, xlo, Np, std::make_integer_sequence<int, PD - 1>{});
		int s = 0;
		for (; s + PD <= nst; s += PD)
			gemm_steps<P, NPL, RT, PD>(s, L, G, Glo, xq, xlo, Np, acc, nst, std::make_integer_sequence<int, PD>{});
		// fewer than PD stages left
		gemm_tail<P, NPL, RT, PD>(s, nst, L, G, Glo, xq, xlo, Np, acc, std::make_integer_sequence<int, PD - 1>{});
		wait_vm<0>(); // no DMA may land after the block has left its LDS
	}

	// part[slot][c][i]: column-major per slot
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

// the two instantiations: plain (64 rows per wave, 4 stages in flight) and
// error-compensated hi + lo (32 rows per wave, 3 stages)
constexpr int kPlainRT = 4, kPlainPD = 4, kSplitRT = 2, kSplitPD = 3;
constexpr uint32_t kMaxSlots = 16; // partial-sum slots per 256-row block

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
// sparse gather over the symmetric CSR links, one wave per row, lanes over
// columns: y[i][c] = sum_k ((x[j][c] * diffuse[j][c%3]) * area_j) * f_k
// (PRE, the reference's operation order) or sum_k x[j][c] * f_k
// ---------------------------------------------------------------------------
template <class T, bool PRE>
__global__ __launch_bounds__(kB) void k_spmm(uint32_t N, uint32_t R, const uint32_t* __restrict__ rowptr,
                                             const uint32_t* __restrict__ col, const float* __restrict__ val,
                                             const float* __restrict__ diffuse, const float* __restrict__ area,
                                             const T* __restrict__ x, T* __restrict__ y)
{
	uint32_t i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
	int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	for (uint32_t c0 = 0; c0 < R; c0 += 64) {
		uint32_t c = c0 + lane;
		T acc = 0;
		if (c < R) {
			for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
				uint32_t j = col[k];
				if (PRE)
					acc += ((x[(uint64_t)j * R + c] * (T)diffuse[3 * j + c % 3]) * (T)area[j]) * (T)val[k];
				else
					acc += x[(uint64_t)j * R + c] * (T)val[k];
			}
			y[(uint64_t)i * R + c] = acc;
		}
	}
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
	if (!d.rad.rowptr)
		throw std::runtime_error("RadSolver: build the links first");
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
	dfree(g32_);
	dfree(g16_);
	dfree(g16lo_);
	dfree(tmap_);
	dfree(tks_);
	dfree(chunks_);
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
	dfree(cmax_);
}

void RadSolver::set_problem(const double* E, uint32_t R_, const float* diffuse3, const float* area)
{
	if (R_ == 0 || R_ > (uint32_t)kRPad)
		throw std::runtime_error("RadSolver: 1..192 right-hand sides");
	if (R_ != R) {
		dfree(rm32a_);
		dfree(rm32b_);
	}
	R = R_;
	hipStream_t s = (hipStream_t)s_;
	dfree(E_);
	dfree(diff_);
	dfree(area_);
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
	uint8_t* df = dalloc<uint8_t>(n);
	HIPCHECK(hipMemsetAsync(df, 0, n, s));
	k_tile_flags<<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, pinv_, df);
	HIPCHECK(hipGetLastError());
	std::vector<uint8_t> f(n);
	HIPCHECK(hipMemcpyAsync(f.data(), df, n, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	dfree(df);
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
	// Work chunks: a row block's list in runs of about ntiles / (2 CUs)
	// tiles (at most kMaxSlots runs per block), so that the blocks with the
	// most links do not set the product's time; each run writes its own
	// partial-sum slot.
	const uint64_t run = std::max<uint64_t>(
	    {8, (ntiles_ + 2 * ncu_ - 1) / (2 * ncu_), ((uint64_t)maxcnt + kMaxSlots - 1) / kMaxSlots});
	std::vector<uint32_t> ch, nsl(nrb, 0);
	for (uint32_t rb = 0; rb < nrb; ++rb) {
		const uint64_t cnt = off[rb + 1] - off[rb];
		const uint32_t nch = (uint32_t)((cnt + run - 1) / run);
		for (uint32_t q = 0; q < nch; ++q) {
			ch.push_back(rb);
			ch.push_back(off[rb] + (uint32_t)(cnt * q / nch));
			ch.push_back(off[rb] + (uint32_t)(cnt * (q + 1) / nch));
			ch.push_back(q);
		}
		nsl[rb] = nch;
	}
	nchunks_ = ch.size() / 4;
	tmap_ = dalloc<int32_t>(n);
	tks_ = dalloc<uint32_t>(ntiles_);
	chunks_ = dalloc<uint32_t>(ch.size());
	nslot_ = dalloc<uint32_t>(nrb);
	HIPCHECK(hipMemcpyAsync(tmap_, map.data(), 4 * n, hipMemcpyHostToDevice, s));
	if (ntiles_)
		HIPCHECK(hipMemcpyAsync(tks_, ks.data(), 4 * ntiles_, hipMemcpyHostToDevice, s));
	if (nchunks_)
		HIPCHECK(hipMemcpyAsync(chunks_, ch.data(), 4 * ch.size(), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(nslot_, nsl.data(), 4 * (size_t)nrb, hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	tiles_built_ = true;
}

void RadSolver::build_dense(int form)
{
	hipStream_t s = (hipStream_t)s_;
	uint64_t threads = (uint64_t)Np * (Np / 32);
	Ev ev(s);
	if (form == kDenseF32) {
		if (!g32_)
			g32_ = dalloc<float>((uint64_t)Np * Np);
		k_dense_build<<<gridn(threads), kB, 0, s>>>(dev.rad.bits, dev.rad.words, N, Np, dev.rad.pos, dev.rad.nrm,
		                                            g32_);
	} else {
		build_tiles();
		if (!g16_)
			g16_ = dalloc<uint16_t>(ntiles_ * kGStageElems);
		if (!g16lo_)
			g16lo_ = dalloc<uint16_t>(ntiles_ * kGStageElems);
		HIPCHECK(hipMemsetAsync(g16_, 0, 2 * ntiles_ * kGStageElems, s));
		HIPCHECK(hipMemsetAsync(g16lo_, 0, 2 * ntiles_ * kGStageElems, s));
		gscale_ = 1.0f;
		if (form == kDenseF16) {
			// scale so that the largest factor sits near 2^13: small factors
			// stay in FP16's normal range
			unsigned* dm = dalloc<unsigned>(1);
			HIPCHECK(hipMemsetAsync(dm, 0, 4, s));
			k_max_link<<<2048, kB, 0, s>>>(dev.rad.nnz, dev.rad.val, dm);
			unsigned hm = 0;
			HIPCHECK(hipMemcpyAsync(&hm, dm, 4, hipMemcpyDeviceToHost, s));
			HIPCHECK(hipStreamSynchronize(s));
			dfree(dm);
			float mx;
			memcpy(&mx, &hm, 4);
			gscale_ = mx > 0 ? ldexpf(1.0f, 14 - ilogbf(mx) - 1) : 1.0f;
			k_tile_scatter<1><<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val,
			                                                   pinv_, tmap_, gscale_, g16_, g16lo_);
		} else {
			k_tile_scatter<0><<<gridn(N, kB / 64), kB, 0, s>>>(N, Np, dev.rad.rowptr, dev.rad.col, dev.rad.val,
			                                                   pinv_, tmap_, 1.0f, g16_, g16lo_);
		}
	}
	HIPCHECK(hipGetLastError());
	build_ms = ev.stop();
	built_form_ = form;
}

uint32_t RadSolver::ksplit(uint32_t rows_per_block, uint32_t min_blocks) const
{
	// split k (in whole stages) until every CU has a block
	uint32_t rows = Np / rows_per_block, s = 1;
	while (rows * s < min_blocks && (Np / (s * 2)) % 32 == 0 && s < 64)
		s *= 2;
	return s;
}

// work buffers of the dense products (kept across solves)
void RadSolver::ensure_work()
{
	const uint64_t n = (uint64_t)kRPad * Np;
	if (!part_) {
		const uint32_t kmax = std::max(ksplit(kTileM, 2048), kMaxSlots);
		part_ = dalloc<float>((uint64_t)kmax * n);
		xq_ = dalloc<float>(n); // FP32 operand, or the 16-bit one in its first half
		xqlo_ = dalloc<uint16_t>(n);
		xT_ = dalloc<float>(n);
		yT_ = dalloc<float>(n);
		cmax_ = dalloc<unsigned>(kRPad);
	}
	if (!rm32a_) {
		rm32a_ = dalloc<float>((uint64_t)Np * R);
		rm32b_ = dalloc<float>((uint64_t)Np * R);
	}
}

// y (column-major, R x Np, FP32) = K x (pre) or H x for the dense forms; the
// 16-bit forms take and give their vectors in Morton order (position p holds
// patch perm_[p]), the FP32 form in patch order.
// split: the error-compensated product of the 16-bit forms (hi + lo planes).
void RadSolver::apply_dense(int form, const float* xT, float* yT, bool split, bool pre)
{
	hipStream_t s = (hipStream_t)s_;
	uint64_t nq = (uint64_t)kRPad * Np;
	if (form == kDenseF32) {
		uint32_t ks = ksplit(kTileM, 2048);
		dim3 grid(Np / kTileM, ks);
		k_prep_x32<<<gridn(nq), kB, 0, s>>>(xT, diff_, area_, N, Np, R, pre ? 1 : 0, xq_);
		k_gemm_f32<<<grid, kB, 0, s>>>(g32_, xq_, Np, ks, part_);
		k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part_, ks, nullptr, Np, R, nullptr, 1.0f, yT);
	} else {
		uint16_t* q = reinterpret_cast<uint16_t*>(xq_);
		uint16_t* lo = split ? xqlo_ : nullptr;
		const unsigned* cm = nullptr;
		if (form == kDenseF16) {
			HIPCHECK(hipMemsetAsync(cmax_, 0, 4 * kRPad, s));
			k_col_max<<<dim3(gridn(Np, 2048), R), kB, 0, s>>>(xT, diff_, area_, perm_, N, Np, pre ? 1 : 0, cmax_);
			cm = cmax_;
			k_prep_x16<1><<<gridn(nq), kB, 0, s>>>(xT, diff_, area_, cm, perm_, N, Np, R, pre ? 1 : 0, q, lo);
		} else {
			k_prep_x16<0><<<gridn(nq), kB, 0, s>>>(xT, diff_, area_, cm, perm_, N, Np, R, pre ? 1 : 0, q, lo);
		}
		if (nchunks_) {
			const uint32_t bpc = 256 / (64 * (split ? kSplitRT : kPlainRT));
			const dim3 grid((unsigned)(nchunks_ * bpc));
			const uint4* chs = reinterpret_cast<const uint4*>(chunks_);
			if (form == kDenseF16 && split)
				k_gemm_mfma<1, 2, kSplitRT, kSplitPD><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, chs, tks_, Np, part_);
			else if (form == kDenseF16)
				k_gemm_mfma<1, 1, kPlainRT, kPlainPD><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, chs, tks_, Np,
				                                                          part_);
			else if (split)
				k_gemm_mfma<0, 2, kSplitRT, kSplitPD><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, chs, tks_, Np, part_);
			else
				k_gemm_mfma<0, 1, kPlainRT, kPlainPD><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, chs, tks_, Np,
				                                                          part_);
		}
		k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part_, 0, nslot_, Np, R, cm, gscale_, yT);
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
			k_spmm<double, true><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, x, y);
		else
			k_spmm<double, false><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, x, y);
	} else if (form == kSparseF32) {
		k_copy_cast<double, float><<<gridn(nR), kB, 0, s>>>(nR, x, rm32a_);
		if (pre)
			k_spmm<float, true><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, rm32a_, rm32b_);
		else
			k_spmm<float, false><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, rm32a_, rm32b_);
		k_copy_cast<float, double><<<gridn(nR), kB, 0, s>>>(nR, rm32b_, y);
	} else {
		const bool mo = form != kDenseF32; // Morton order
		k_to_colmajor<double, float><<<gridn(nRp), kB, 0, s>>>(x, N, Np, R, mo ? perm_ : nullptr, xT_);
		apply_dense(form, xT_, yT_, false, pre);
		k_to_rowmajor<float, double><<<gridn(nR), kB, 0, s>>>(yT_, N, Np, R, mo ? pinv_ : nullptr, y);
	}
	HIPCHECK(hipGetLastError());
}

void RadSolver::solve(int form, int bounces, int refine, double* total_out)
{
	hipStream_t s = (hipStream_t)s_;
	uint64_t nR = (uint64_t)N * R, nRp = (uint64_t)R * Np;
	times = RadTimes{};
	ensure_work();
	// FP64 running total
	double* tot = dalloc<double>(nR);
	HIPCHECK(hipMemcpyAsync(tot, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	if (form == kSparseF64 || form == kSparseF32) {
		double *x = dalloc<double>(nR), *y = dalloc<double>(nR);
		HIPCHECK(hipMemcpyAsync(x, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
		Ev ev(s);
		for (int b = 0; b < bounces; ++b) {
			apply_op(form, x, y, true);
			k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, y, tot);
			std::swap(x, y);
		}
		times.series_ms = ev.stop();
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
		dfree(x);
		dfree(y);
	} else {
		if (built_form_ != form)
			build_dense(form);
		times.build_ms = build_ms;
		// z_k (column-major FP32), k = 0..b
		std::vector<float*> z(bounces + 1);
		for (auto& p : z)
			p = dalloc<float>(nRp);
		// the 16-bit forms keep z in Morton order
		const uint32_t* po = form != kDenseF32 ? perm_ : nullptr;
		const uint32_t* pi = form != kDenseF32 ? pinv_ : nullptr;
		k_to_colmajor<double, float><<<gridn(nRp), kB, 0, s>>>(E_, N, Np, R, po, z[0]);
		{
			Ev ev(s);
			for (int b = 1; b <= bounces; ++b)
				apply_dense(form, z[b - 1], z[b], false, true);
			times.series_ms = ev.stop();
		}
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
		// Iterative refinement on the block system z_b = K z_{b-1}, z_0 = E.
		// One sweep: residuals rho_b = K z_{b-1} - z_b of the current iterate,
		// with K z_{b-1} from the error-compensated (hi + lo) MFMA product and
		// the difference taken in FP64; corrections d_b = Kdense d_{b-1} + rho_b
		// (d_0 = 0: z_0 = E is exact) with the plain 16-bit product; then
		// z += d.  z_{b-1} is corrected only after rho_b has used its old value.
		const bool half = form != kDenseF32;
		double* zr = dalloc<double>(nR);
		double* rho = refine > 0 ? dalloc<double>(nRp) : nullptr;
		float* kz = refine > 0 ? dalloc<float>(nRp) : nullptr;
		float* d = refine > 0 ? dalloc<float>(nRp) : nullptr;
		float* dprev = refine > 0 ? dalloc<float>(nRp) : nullptr;
		{
			Ev ev(s);
			for (int it = 0; it < refine; ++it) {
				HIPCHECK(hipMemsetAsync(dprev, 0, 4 * nRp, s));
				for (int b = 1; b <= bounces; ++b) {
					apply_dense(form, z[b - 1], kz, half, true);                          // K z_{b-1}
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
		for (auto& p : z)
			dfree(p);
		dfree(zr);
		dfree(kz);
		dfree(rho);
		dfree(d);
		dfree(dprev);
	}
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(total_out, tot, 8 * nR, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	dfree(tot);
}

// Reverse mode of T = sum_{k=0..B} K^k E, K = H diag(area) diag(diffuse_c).
// Forward partial sums S_k = sum_{m<=k} K^m E are kept (T = S_B is the Horner
// recurrence S_k = E + K S_{k-1}); its adjoint, from ybar_B = g:
//   u_k = H ybar_k            (H symmetric, so K^T = diag(diffuse_c) diag(area) H)
//   dalb += area * u_k * S_{k-1}
//   ybar_{k-1} = diffuse_c * area * u_k
//   lam = sum_k ybar_k        (= dL/dE; also dL/d emission)
void RadSolver::solve_grad(int form, int bounces, const double* g, double* total_out, double* lam_out,
                           double* dalb_out)
{
	hipStream_t s = (hipStream_t)s_;
	const uint64_t nR = (uint64_t)N * R;
	times = RadTimes{};
	ensure_work();
	if (form >= kDenseF32 && built_form_ != form)
		build_dense(form);
	const int B = std::max(0, bounces);
	std::vector<double*> S(std::max(1, B));
	for (auto& p : S)
		p = dalloc<double>(nR);
	double* z = dalloc<double>(nR);
	double* zn = dalloc<double>(nR);
	double* tot = dalloc<double>(nR);
	double* ybar = dalloc<double>(nR);
	double* u = dalloc<double>(nR);
	double* lam = dalloc<double>(nR);
	double* dalb = dalloc<double>(nR);
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
	for (auto& p : S)
		dfree(p);
	dfree(z);
	dfree(zn);
	dfree(tot);
	dfree(ybar);
	dfree(u);
	dfree(lam);
	dfree(dalb);
}

} // namespace lgx
