// Column order within the rows of the radiosity link CSR: a segmented radix
// sort (rocPRIM) over the rows, so that every row lists its links by
// ascending patch index, as the all-pairs enumeration produces them.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <stdexcept>
#include <string>

#include "device.h"

namespace lgx {

void sort_row_columns(const uint32_t* in, uint32_t* out, uint32_t nnz, uint32_t nrows, const uint32_t* rowptr,
                      int bits, void* stream)
{
	hipStream_t s = (hipStream_t)stream;
	size_t bytes = 0;
	hipError_t e = rocprim::segmented_radix_sort_keys(nullptr, bytes, in, out, nnz, nrows, rowptr, rowptr + 1, 0,
	                                                  (unsigned)bits, s);
	if (e != hipSuccess)
		throw std::runtime_error(std::string("link column sort: ") + hipGetErrorString(e));
	void* tmp = nullptr;
	e = hipMalloc(&tmp, bytes ? bytes : 1);
	if (e != hipSuccess)
		throw std::runtime_error(std::string("link column sort: ") + hipGetErrorString(e));
	e = rocprim::segmented_radix_sort_keys(tmp, bytes, in, out, nnz, nrows, rowptr, rowptr + 1, 0, (unsigned)bits, s);
	const hipError_t e2 = hipStreamSynchronize(s);
	(void)hipFree(tmp);
	if (e != hipSuccess || e2 != hipSuccess)
		throw std::runtime_error(std::string("link column sort: ") + hipGetErrorString(e != hipSuccess ? e : e2));
}

} // namespace lgx
