// Radiosity solver forms over the link structure built by rad_links().
//
// The reference's bounce loop (lighter.cpp:767-797, 1104-1123) computes the
// truncated series T = sum_{k=0..b} K^k E per colour channel, with
// K = F diag(area) diag(diffuse_c) and F the symmetric matrix of link form
// factors (zero for pairs without a link).  Here it is evaluated for many
// right-hand sides at once (columns of E; column c uses channel c % 3):
//
//  * sparse gathers over the symmetric CSR links, FP32 or FP64;
//  * dense G = F diag(area) stored FP32 and multiplied with plain FP32 FMAs;
//  * dense G stored BF16 or FP16 (power-of-two scaled) and multiplied with
//    gfx950 MFMA instructions accumulating in FP32;
//  * the MFMA form followed by iterative refinement sweeps whose residual
//    r_k = K z_{k-1} - z_k uses the exact link operator in FP64.
//
// Layouts: dense G is row-major Np x Np (rows = receiving patch), the dense
// right-hand sides are column-major (R x Np, each column contiguous), the
// sparse ones row-major (Np x R).  Np is N rounded up to kTileM.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <stdexcept>
#include <string>
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

constexpr int kB = 256;          // threads per block
constexpr int kTileM = 128;      // rows per block of the dense GEMMs
constexpr int kKStage = 64;      // k per LDS stage
constexpr int kColTiles = 12;    // 16-wide column tiles: R padded to 192
constexpr int kRPad = 16 * kColTiles;

static inline unsigned gridn(uint64_t n, unsigned b = kB) { return (unsigned)((n + b - 1) / b); }

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

// ---------------------------------------------------------------------------
// dense matrix G[i][j] = f_ij * area_j from the occlusion bitmask; one thread
// per 32-column word.  FP32 (P = -1) or 16-bit with a power-of-two scale.
// ---------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ float from16(uint16_t v)
{
	if (P == 0)
		return __uint_as_float((uint32_t)v << 16);
	_Float16 h;
	memcpy(&h, &v, 2);
	return (float)h;
}

template <int P>
__global__ __launch_bounds__(kB) void k_dense_build(const uint32_t* __restrict__ bits, uint64_t words, uint32_t N,
                                                    uint32_t Np, const float* __restrict__ pos,
                                                    const float* __restrict__ nrm, const float* __restrict__ area,
                                                    float scale, void* __restrict__ G, uint16_t* __restrict__ Glo)
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
		v[b] = ((m >> b) & 1u) ? link_factor(pos, nrm, i, j) * area[j] : 0.0f;
	}
	if (P < 0) {
		float4* o = reinterpret_cast<float4*>((float*)G + (uint64_t)i * Np + (uint64_t)w * 32);
#pragma unroll
		for (int q = 0; q < 8; ++q)
			o[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
	} else {
		constexpr int PP = P < 0 ? 0 : P;
		uint32_t pk[16], pl[16];
#pragma unroll
		for (int q = 0; q < 16; ++q) {
			float a = v[2 * q] * scale, b = v[2 * q + 1] * scale;
			uint16_t ha = Prec<PP>::cvt(a), hb = Prec<PP>::cvt(b);
			pk[q] = (uint32_t)ha | ((uint32_t)hb << 16);
			// remainder plane: what the 16-bit value left out
			pl[q] = (uint32_t)Prec<PP>::cvt(a - from16<PP>(ha)) | ((uint32_t)Prec<PP>::cvt(b - from16<PP>(hb)) << 16);
		}
		uint4* o = reinterpret_cast<uint4*>((uint16_t*)G + (uint64_t)i * Np + (uint64_t)w * 32);
#pragma unroll
		for (int q = 0; q < 4; ++q)
			o[q] = make_uint4(pk[4 * q], pk[4 * q + 1], pk[4 * q + 2], pk[4 * q + 3]);
		if (Glo) {
			uint4* ol = reinterpret_cast<uint4*>(Glo + (uint64_t)i * Np + (uint64_t)w * 32);
#pragma unroll
			for (int q = 0; q < 4; ++q)
				ol[q] = make_uint4(pl[4 * q], pl[4 * q + 1], pl[4 * q + 2], pl[4 * q + 3]);
		}
	}
}

// largest |G| entry (for the FP16 scale): max over links of f * area_j
__global__ void k_max_link(uint32_t N, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                           const float* __restrict__ val, const float* __restrict__ area, unsigned* out)
{
	uint32_t i = blockIdx.x * kB + threadIdx.x;
	if (i >= N)
		return;
	float m = 0.0f;
	for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
		m = fmaxf(m, val[k] * area[col[k]]);
	atomicMax(out, __float_as_uint(m)); // non-negative floats order like their bits
}

// ---------------------------------------------------------------------------
// right-hand-side preparation: Xq[c][j] = x[c][j] * diffuse[j][c % 3] * s_c
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kB) void k_prep_x16(const float* __restrict__ xT, const float* __restrict__ diffuse,
                                                 const float* __restrict__ cscale, uint32_t N, uint32_t Np,
                                                 uint32_t R, uint16_t* __restrict__ xq, uint16_t* __restrict__ xlo)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)kRPad * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), j = (uint32_t)(t % Np);
	float v = 0.0f;
	if (c < R && j < N)
		v = xT[(uint64_t)c * Np + j] * diffuse[3 * j + c % 3] * cscale[c];
	uint16_t h = Prec<P>::cvt(v);
	xq[t] = h;
	if (xlo)
		xlo[t] = Prec<P>::cvt(v - from16<P>(h));
}

__global__ __launch_bounds__(kB) void k_prep_x32(const float* __restrict__ xT, const float* __restrict__ diffuse,
                                                 uint32_t N, uint32_t Np, uint32_t R, float* __restrict__ xq)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)kRPad * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), j = (uint32_t)(t % Np);
	xq[t] = (c < R && j < N) ? xT[(uint64_t)c * Np + j] * diffuse[3 * j + c % 3] : 0.0f;
}

// ---------------------------------------------------------------------------
// MFMA GEMM: part[s][c][i] = sum_{j in split s} G[i][j] * Xq[c][j]
// Block: 128 rows x 192 columns, 4 waves of 32 rows; v_mfma_f32_16x16x32
// (bf16 or f16) with FP32 accumulators.  Lane l supplies row (l & 15) of A
// and column (l & 15) of B with the same eight k indices 8*(l>>4)+e, so the
// k order inside an instruction is immaterial; output element i of a 16x16
// tile sits at row 4*(l>>4)+i, column l & 15.
// The right-hand-side chunk (192 x 64) is staged through LDS per k stage.
// ---------------------------------------------------------------------------
constexpr int kLdsStride = kKStage + 8; // 16-bit elements per LDS column (pad)
constexpr int kXPieces = (kRPad * kKStage / 8) / kB; // 16-byte pieces per thread per stage

template <int P>
__device__ __forceinline__ f32x4 mfma16(uint4 a, uint4 b, f32x4 c)
{
	if (P == 0) {
		bf16x8 av, bv;
		memcpy(&av, &a, 16);
		memcpy(&bv, &b, 16);
		return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
	} else {
		f16x8 av, bv;
		memcpy(&av, &a, 16);
		memcpy(&bv, &b, 16);
		return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, c, 0, 0, 0);
	}
}

// SPLIT: the matrix and the right-hand side come as hi + lo planes and the
// kernel forms G_hi X_hi + G_hi X_lo + G_lo X_hi (the lo x lo term is below
// FP32 resolution), an error-compensated product.
// The next k stage's loads (matrix fragments to registers, right-hand-side
// chunk to registers then to the other LDS buffer) are issued before the
// current stage's MFMAs, so HBM traffic overlaps the math.
template <int P, bool SPLIT>
__global__ __launch_bounds__(kB) void k_gemm_mfma(const uint16_t* __restrict__ G, const uint16_t* __restrict__ Glo,
                                                  const uint16_t* __restrict__ xq, const uint16_t* __restrict__ xlo,
                                                  uint32_t Np, uint32_t ksplit, float* __restrict__ part)
{
	constexpr int NP = SPLIT ? 2 : 1; // planes
	__shared__ uint16_t xs[2][NP][kRPad * kLdsStride];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const uint32_t i0 = blockIdx.x * kTileM + wave * 32;
	const uint32_t klen = Np / ksplit;
	const uint32_t kbeg = blockIdx.y * klen, kend = kbeg + klen;
	const int r = lane & 15, g = lane >> 4;
	f32x4 acc[2][kColTiles];
#pragma unroll
	for (int a = 0; a < 2; ++a)
#pragma unroll
		for (int t = 0; t < kColTiles; ++t)
			acc[a][t] = f32x4{0, 0, 0, 0};
	const uint64_t row0 = (uint64_t)(i0 + r) * Np, row1 = (uint64_t)(i0 + 16 + r) * Np;
	// matrix fragments [plane][row tile][k half]; x chunk pieces [plane][piece]
	uint4 ga[NP][2][2], xr[NP][kXPieces];
	auto load_stage = [&](uint32_t kb) {
#pragma unroll
		for (int pl = 0; pl < NP; ++pl) {
			const uint16_t* M = pl ? Glo : G;
			const uint16_t* X = pl ? xlo : xq;
			ga[pl][0][0] = *reinterpret_cast<const uint4*>(M + row0 + kb + 8 * g);
			ga[pl][0][1] = *reinterpret_cast<const uint4*>(M + row0 + kb + 32 + 8 * g);
			ga[pl][1][0] = *reinterpret_cast<const uint4*>(M + row1 + kb + 8 * g);
			ga[pl][1][1] = *reinterpret_cast<const uint4*>(M + row1 + kb + 32 + 8 * g);
#pragma unroll
			for (int q = 0; q < kXPieces; ++q) {
				int piece = q * kB + tid, c = piece / (kKStage / 8), o = piece % (kKStage / 8);
				xr[pl][q] = *reinterpret_cast<const uint4*>(X + (uint64_t)c * Np + kb + o * 8);
			}
		}
	};
	auto store_x = [&](int buf) {
#pragma unroll
		for (int pl = 0; pl < NP; ++pl)
#pragma unroll
			for (int q = 0; q < kXPieces; ++q) {
				int piece = q * kB + tid, c = piece / (kKStage / 8), o = piece % (kKStage / 8);
				*reinterpret_cast<uint4*>(&xs[buf][pl][c * kLdsStride + o * 8]) = xr[pl][q];
			}
	};
	load_stage(kbeg);
	store_x(0);
	__syncthreads();
	int buf = 0;
	for (uint32_t kb = kbeg; kb < kend; kb += kKStage) {
		uint4 cur[NP][2][2];
#pragma unroll
		for (int pl = 0; pl < NP; ++pl)
#pragma unroll
			for (int a = 0; a < 2; ++a)
#pragma unroll
				for (int h = 0; h < 2; ++h)
					cur[pl][a][h] = ga[pl][a][h];
		const bool more = kb + kKStage < kend;
		if (more)
			load_stage(kb + kKStage);
#pragma unroll
		for (int h = 0; h < 2; ++h) {
#pragma unroll
			for (int t = 0; t < kColTiles; ++t) {
				const int off = (t * 16 + r) * kLdsStride + h * 32 + 8 * g;
				uint4 Bv = *reinterpret_cast<const uint4*>(&xs[buf][0][off]);
				acc[0][t] = mfma16<P>(cur[0][0][h], Bv, acc[0][t]);
				acc[1][t] = mfma16<P>(cur[0][1][h], Bv, acc[1][t]);
				if (SPLIT) {
					uint4 Bl = *reinterpret_cast<const uint4*>(&xs[buf][NP - 1][off]);
					acc[0][t] = mfma16<P>(cur[0][0][h], Bl, acc[0][t]);
					acc[1][t] = mfma16<P>(cur[0][1][h], Bl, acc[1][t]);
					acc[0][t] = mfma16<P>(cur[NP - 1][0][h], Bv, acc[0][t]);
					acc[1][t] = mfma16<P>(cur[NP - 1][1][h], Bv, acc[1][t]);
				}
			}
		}
		if (more)
			store_x(buf ^ 1);
		__syncthreads();
		buf ^= 1;
	}
	// part[s][c][i]: column-major per split
	float* out = part + (uint64_t)blockIdx.y * kRPad * Np;
#pragma unroll
	for (int a = 0; a < 2; ++a)
#pragma unroll
		for (int t = 0; t < kColTiles; ++t) {
			uint32_t c = t * 16 + r;
			uint32_t i = i0 + a * 16 + 4 * g;
			*reinterpret_cast<float4*>(out + (uint64_t)c * Np + i) =
			    make_float4(acc[a][t][0], acc[a][t][1], acc[a][t][2], acc[a][t][3]);
		}
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

// y[c][i] = (sum_s part[s][c][i]) * unscale_c, fixed order over s
__global__ __launch_bounds__(kB) void k_reduce_split(const float* __restrict__ part, uint32_t ksplit, uint32_t Np,
                                                     uint32_t R, const float* __restrict__ unscale,
                                                     float* __restrict__ y)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)R * Np)
		return;
	uint32_t c = (uint32_t)(t / Np);
	float s = 0.0f;
	for (uint32_t q = 0; q < ksplit; ++q)
		s += part[(uint64_t)q * kRPad * Np + t];
	y[t] = s * (unscale ? unscale[c] : 1.0f);
}

// ---------------------------------------------------------------------------
// sparse gather over the symmetric CSR links, one wave per row, lanes over
// columns: y[i][c] = sum_k ((x[j][c] * diffuse[j][c%3]) * area_j) * f_k
// ---------------------------------------------------------------------------
template <class T>
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
				acc += ((x[(uint64_t)j * R + c] * (T)diffuse[3 * j + c % 3]) * (T)area[j]) * (T)val[k];
			}
			y[(uint64_t)i * R + c] = acc;
		}
	}
}

// layout changes between the row-major (sparse) and column-major (dense) sets
template <class A, class B>
__global__ void k_to_colmajor(const A* __restrict__ x, uint32_t N, uint32_t Np, uint32_t R, B* __restrict__ xT)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)R * Np)
		return;
	uint32_t c = (uint32_t)(t / Np), j = (uint32_t)(t % Np);
	xT[t] = j < N ? (B)x[(uint64_t)j * R + c] : (B)0;
}

template <class A, class B>
__global__ void k_to_rowmajor(const A* __restrict__ xT, uint32_t N, uint32_t Np, uint32_t R, B* __restrict__ x)
{
	uint64_t t = (uint64_t)blockIdx.x * kB + threadIdx.x;
	if (t >= (uint64_t)N * R)
		return;
	uint32_t j = (uint32_t)(t / R), c = (uint32_t)(t % R);
	x[t] = (B)xT[(uint64_t)c * Np + j];
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
	Np = ((N + kTileM - 1) / kTileM) * kTileM;
	if (Np < 512)
		Np = 512;
	s_ = d.stream();
}

RadSolver::~RadSolver() { release(); }

void RadSolver::release()
{
	dfree(g32_);
	dfree(g16_);
	dfree(g16lo_);
	dfree(E_);
	dfree(diff_);
	dfree(area_);
}

void RadSolver::set_problem(const double* E, uint32_t R_, const float* diffuse3, const float* area)
{
	if (R_ == 0 || R_ > (uint32_t)kRPad)
		throw std::runtime_error("RadSolver: 1..192 right-hand sides");
	R = R_;
	built_form_ = -1; // G carries the areas: rebuild for a new problem
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

void RadSolver::build_dense(int form)
{
	hipStream_t s = (hipStream_t)s_;
	uint64_t threads = (uint64_t)Np * (Np / 32);
	Ev ev(s);
	if (form == kDenseF32) {
		if (!g32_)
			g32_ = dalloc<float>((uint64_t)Np * Np);
		k_dense_build<-1><<<gridn(threads), kB, 0, s>>>(dev.rad.bits, dev.rad.words, N, Np, dev.rad.pos,
		                                                dev.rad.nrm, area_, 1.0f, g32_, nullptr);
	} else {
		if (!g16_)
			g16_ = dalloc<uint16_t>((uint64_t)Np * Np);
		if (!g16lo_)
			g16lo_ = dalloc<uint16_t>((uint64_t)Np * Np);
		gscale_ = 1.0f;
		if (form == kDenseF16) {
			// scale so that the largest entry sits near 2^14: small factors
			// stay in FP16's normal range
			unsigned* dm = dalloc<unsigned>(1);
			HIPCHECK(hipMemsetAsync(dm, 0, 4, s));
			k_max_link<<<gridn(N), kB, 0, s>>>(N, dev.rad.rowptr, dev.rad.col, dev.rad.val, area_, dm);
			unsigned hm = 0;
			HIPCHECK(hipMemcpyAsync(&hm, dm, 4, hipMemcpyDeviceToHost, s));
			HIPCHECK(hipStreamSynchronize(s));
			dfree(dm);
			float mx;
			memcpy(&mx, &hm, 4);
			gscale_ = mx > 0 ? ldexpf(1.0f, 14 - ilogbf(mx) - 1) : 1.0f;
			k_dense_build<1><<<gridn(threads), kB, 0, s>>>(dev.rad.bits, dev.rad.words, N, Np, dev.rad.pos,
			                                               dev.rad.nrm, area_, gscale_, g16_, g16lo_);
		} else {
			k_dense_build<0><<<gridn(threads), kB, 0, s>>>(dev.rad.bits, dev.rad.words, N, Np, dev.rad.pos,
			                                               dev.rad.nrm, area_, 1.0f, g16_, g16lo_);
		}
	}
	HIPCHECK(hipGetLastError());
	build_ms = ev.stop();
	built_form_ = form;
}

uint32_t RadSolver::ksplit() const
{
	// enough blocks to fill the chip: split k until >= 2048 blocks
	uint32_t rows = Np / kTileM, s = 1;
	while (rows * s < 2048 && (Np / (s * 2)) % kKStage == 0 && s < 64)
		s *= 2;
	return s;
}

// y (column-major, R x Np, FP32) = K x for the dense forms.  split: the
// error-compensated product of the 16-bit forms (hi + lo planes); xqlo is
// the right-hand side's lo plane.
void RadSolver::apply_dense(int form, const float* xT, float* yT, float* part, void* xq, uint16_t* xqlo,
                            bool split)
{
	hipStream_t s = (hipStream_t)s_;
	uint32_t ks = ksplit();
	dim3 grid(Np / kTileM, ks);
	uint64_t nq = (uint64_t)kRPad * Np;
	if (form == kDenseF32) {
		k_prep_x32<<<gridn(nq), kB, 0, s>>>(xT, diff_, N, Np, R, (float*)xq);
		k_gemm_f32<<<grid, kB, 0, s>>>(g32_, (const float*)xq, Np, ks, part);
		k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part, ks, Np, R, nullptr, yT);
	} else {
		// per-column power-of-two scale keeps FP16 operands in range
		uint16_t* lo = split ? xqlo : nullptr;
		const uint16_t* q = (const uint16_t*)xq;
		if (form == kDenseF16) {
			k_prep_x16<1><<<gridn(nq), kB, 0, s>>>(xT, diff_, cscale_, N, Np, R, (uint16_t*)xq, lo);
			if (split)
				k_gemm_mfma<1, true><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, Np, ks, part);
			else
				k_gemm_mfma<1, false><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, Np, ks, part);
		} else {
			k_prep_x16<0><<<gridn(nq), kB, 0, s>>>(xT, diff_, cscale_, N, Np, R, (uint16_t*)xq, lo);
			if (split)
				k_gemm_mfma<0, true><<<grid, kB, 0, s>>>(g16_, g16lo_, q, lo, Np, ks, part);
			else
				k_gemm_mfma<0, false><<<grid, kB, 0, s>>>(g16_, nullptr, q, nullptr, Np, ks, part);
		}
		k_reduce_split<<<gridn((uint64_t)R * Np), kB, 0, s>>>(part, ks, Np, R, cunscale_, yT);
	}
	HIPCHECK(hipGetLastError());
}

// per-column scales for the 16-bit operand: FP16 needs the column maxima
void RadSolver::column_scales(int form, const float* xT)
{
	std::vector<float> hx((size_t)R * Np), sc(kRPad, 1.0f), us(kRPad, 1.0f);
	if (form == kDenseF16) {
		HIPCHECK(hipMemcpy(hx.data(), xT, 4 * (size_t)R * Np, hipMemcpyDeviceToHost));
		std::vector<float> hd(3 * (size_t)N);
		HIPCHECK(hipMemcpy(hd.data(), diff_, 12 * (size_t)N, hipMemcpyDeviceToHost));
		for (uint32_t c = 0; c < R; ++c) {
			float m = 0;
			for (uint32_t j = 0; j < N; ++j)
				m = fmaxf(m, fabsf(hx[(size_t)c * Np + j] * hd[3 * j + c % 3]));
			float e = m > 0 ? ldexpf(1.0f, 14 - ilogbf(m) - 1) : 1.0f;
			sc[c] = e;
			us[c] = 1.0f / (e * gscale_);
		}
	} else {
		for (uint32_t c = 0; c < R; ++c)
			us[c] = 1.0f / gscale_;
	}
	HIPCHECK(hipMemcpy(cscale_, sc.data(), 4 * kRPad, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(cunscale_, us.data(), 4 * kRPad, hipMemcpyHostToDevice));
}

void RadSolver::solve(int form, int bounces, int refine, double* total_out)
{
	hipStream_t s = (hipStream_t)s_;
	uint64_t nR = (uint64_t)N * R, nRp = (uint64_t)R * Np;
	times = RadTimes{};
	// FP64 running total and the exact sparse operator's workspace
	double* tot = dalloc<double>(nR);
	HIPCHECK(hipMemcpyAsync(tot, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
	const uint32_t* rp = dev.rad.rowptr;
	const uint32_t* cl = dev.rad.col;
	const float* vl = dev.rad.val;
	unsigned sgrid = gridn(N, kB / 64);
	if (form == kSparseF64 || form == kSparseF32) {
		Ev ev(s);
		if (form == kSparseF64) {
			double *x = dalloc<double>(nR), *y = dalloc<double>(nR);
			HIPCHECK(hipMemcpyAsync(x, E_, 8 * nR, hipMemcpyDeviceToDevice, s));
			for (int b = 0; b < bounces; ++b) {
				k_spmm<double><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, x, y);
				k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, y, tot);
				std::swap(x, y);
			}
			dfree(x);
			dfree(y);
		} else {
			float *x = dalloc<float>(nR), *y = dalloc<float>(nR);
			k_copy_cast<double, float><<<gridn(nR), kB, 0, s>>>(nR, E_, x);
			Ev ev2(s);
			for (int b = 0; b < bounces; ++b) {
				k_spmm<float><<<sgrid, kB, 0, s>>>(N, R, rp, cl, vl, diff_, area_, x, y);
				k_axpy<float, double><<<gridn(nR), kB, 0, s>>>(nR, y, tot);
				std::swap(x, y);
			}
			times.series_ms = ev2.stop();
			dfree(x);
			dfree(y);
		}
		double ms = ev.stop();
		if (form == kSparseF64)
			times.series_ms = ms;
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
	} else {
		if (built_form_ != form)
			build_dense(form);
		times.build_ms = build_ms;
		// z_k (column-major FP32), k = 0..b
		std::vector<float*> z(bounces + 1);
		for (auto& p : z)
			p = dalloc<float>(nRp);
		float* part = dalloc<float>((uint64_t)ksplit() * kRPad * Np);
		const bool half = form != kDenseF32;
		void* xq = half ? (void*)dalloc<uint16_t>((uint64_t)kRPad * Np) : (void*)dalloc<float>((uint64_t)kRPad * Np);
		uint16_t* xqlo = half ? dalloc<uint16_t>((uint64_t)kRPad * Np) : nullptr;
		cscale_ = dalloc<float>(kRPad);
		cunscale_ = dalloc<float>(kRPad);
		k_to_colmajor<double, float><<<gridn(nRp), kB, 0, s>>>(E_, N, Np, R, z[0]);
		column_scales(form, z[0]);
		{
			Ev ev(s);
			for (int b = 1; b <= bounces; ++b)
				apply_dense(form, z[b - 1], z[b], part, xq, nullptr, false);
			times.series_ms = ev.stop();
		}
		times.per_bounce_ms = times.series_ms / std::max(1, bounces);
		// Iterative refinement on the block system z_b = K z_{b-1}, z_0 = E.
		// One sweep: residuals rho_b = K z_{b-1} - z_b of the current iterate,
		// with K z_{b-1} from the error-compensated (hi + lo) MFMA product and
		// the difference taken in FP64; corrections d_b = Kdense d_{b-1} + rho_b
		// (d_0 = 0: z_0 = E is exact) with the plain 16-bit product; then
		// z += d.  z_{b-1} is corrected only after rho_b has used its old value.
		double* zr = dalloc<double>(nR);
		double* rho = dalloc<double>(nRp);
		float* kz = dalloc<float>(nRp);
		float* d = dalloc<float>(nRp);
		float* dprev = dalloc<float>(nRp);
		{
			Ev ev(s);
			for (int it = 0; it < refine; ++it) {
				HIPCHECK(hipMemsetAsync(dprev, 0, 4 * nRp, s));
				for (int b = 1; b <= bounces; ++b) {
					apply_dense(form, z[b - 1], kz, part, xq, xqlo, half);               // K z_{b-1}
					k_resid64<<<gridn(nRp), kB, 0, s>>>(nRp, kz, z[b], rho);             // rho_b (FP64)
					apply_dense(form, dprev, d, part, xq, nullptr, false);               // Kdense d_{b-1}
					k_axpy<double, float><<<gridn(nRp), kB, 0, s>>>(nRp, rho, d);        // d_b
					k_axpy<float, float><<<gridn(nRp), kB, 0, s>>>(nRp, dprev, z[b - 1]); // apply d_{b-1}
					std::swap(d, dprev);
				}
				k_axpy<float, float><<<gridn(nRp), kB, 0, s>>>(nRp, dprev, z[bounces]);
			}
			times.refine_ms = ev.stop();
		}
		dfree(xqlo);
		HIPCHECK(hipGetLastError());
		// total = sum_k z_k in FP64 (z_0 = E exactly)
		for (int b = 1; b <= bounces; ++b) {
			k_to_rowmajor<float, double><<<gridn(nR), kB, 0, s>>>(z[b], N, Np, R, zr);
			k_axpy<double, double><<<gridn(nR), kB, 0, s>>>(nR, zr, tot);
		}
		for (auto& p : z)
			dfree(p);
		dfree(part);
		if (form == kDenseF32)
			dfree(*(float**)&xq);
		else
			dfree(*(uint16_t**)&xq);
		dfree(cscale_);
		dfree(cunscale_);
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

} // namespace lgx
