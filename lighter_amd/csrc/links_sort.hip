// Column order within the rows of the radiosity link CSR: a segmented radix
// sort (rocPRIM) over the rows, so that every row lists its links by
// ascending patch index, as the all-pairs enumeration produces them.  The
// sort takes 32-bit sizes and offsets, so the rows go in runs of at most
// 2^31 entries, each with its offsets made relative to the run.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"

namespace lgx {

#define SORTCHECK(x)                                                                                     \
	do {                                                                                                 \
		hipError_t e_ = (x);                                                                             \
		if (e_ != hipSuccess)                                                                            \
			throw std::runtime_error(std::string("link column sort: ") + hipGetErrorString(e_) + ": " #x); \
	} while (0)

__global__ void k_rel_offsets(const uint64_t* __restrict__ rowptr, uint32_t r0, uint32_t n, uint64_t base,
                              uint32_t* __restrict__ out)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i <= n)
		out[i] = (uint32_t)(rowptr[r0 + i] - base);
}

void sort_row_columns(const uint32_t* in, uint32_t* out, uint64_t nnz, uint32_t nrows, const uint64_t* rowptr,
                      int bits, void* stream)
{
	hipStream_t s = (hipStream_t)stream;
	if (nnz == 0 || nrows == 0)
		return;
	std::vector<uint64_t> rp((size_t)nrows + 1);
	SORTCHECK(hipMemcpyAsync(rp.data(), rowptr, 8 * ((size_t)nrows + 1), hipMemcpyDeviceToHost, s));
	SORTCHECK(hipStreamSynchronize(s));
	constexpr uint64_t kRun = (uint64_t)1 << 31;
	uint32_t* rel = nullptr;
	void* tmp = nullptr;
	size_t tmp_bytes = 0;
	uint32_t rel_cap = 0;
	try {
		for (uint32_t r0 = 0; r0 < nrows;) {
			uint32_t r1 = r0 + 1;
			if (rp[r1] - rp[r0] > kRun)
				throw std::runtime_error("link column sort: a row holds more than 2^31 links");
			while (r1 < nrows && rp[r1 + 1] - rp[r0] <= kRun)
				++r1;
			const uint64_t base = rp[r0];
			const uint32_t size = (uint32_t)(rp[r1] - base), nseg = r1 - r0;
			if (nseg + 1 > rel_cap) {
				if (rel)
					SORTCHECK(hipFree(rel));
				rel = nullptr;
				rel_cap = nseg + 1;
				SORTCHECK(hipMalloc(&rel, 4 * (size_t)rel_cap));
			}
			k_rel_offsets<<<(nseg + 256) / 256, 256, 0, s>>>(rowptr, r0, nseg, base, rel);
			SORTCHECK(hipGetLastError());
			size_t bytes = 0;
			SORTCHECK(rocprim::segmented_radix_sort_keys(nullptr, bytes, in + base, out + base, size, nseg, rel,
			                                             rel + 1, 0, (unsigned)bits, s));
			if (bytes > tmp_bytes) {
				if (tmp)
					SORTCHECK(hipFree(tmp));
				tmp = nullptr;
				tmp_bytes = bytes;
				SORTCHECK(hipMalloc(&tmp, tmp_bytes));
			}
			SORTCHECK(rocprim::segmented_radix_sort_keys(tmp, bytes, in + base, out + base, size, nseg, rel, rel + 1,
			                                             0, (unsigned)bits, s));
			SORTCHECK(hipStreamSynchronize(s)); // rel and tmp are reused by the next run
			r0 = r1;
		}
	} catch (...) {
		(void)hipStreamSynchronize(s);
		if (rel)
			(void)hipFree(rel);
		if (tmp)
			(void)hipFree(tmp);
		throw;
	}
	if (rel)
		SORTCHECK(hipFree(rel));
	if (tmp)
		SORTCHECK(hipFree(tmp));
}

} // namespace lgx
