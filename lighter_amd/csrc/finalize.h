// Lightmap export on the GPU (lighter.cpp:849-958, lighter_math.cpp:357-424):
// the samples' colours scattered into each instance's atlas, three rounds of
// edge extrapolation into unset texels, the separable Gaussian blur (two
// convolve-and-transpose passes) and the optional 2x downsample, for every
// instance in one set of launches.  Each kernel evaluates exactly the host
// restatement's operations per texel, in its order, so the images are
// identical.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace lgx {

// One instance's image: atlas size, its samples (colour and texel index),
// and where the finished image goes (caller-owned, final_w x final_h x 3).
struct FinImage {
	uint32_t w, h;
	size_t nsamples;
	const float* rgb;     // 3 floats per sample
	const uint32_t* loc;  // texel index per sample (x + y w)
	float* out;           // final_w * final_h * 3 floats
	uint32_t final_w, final_h;
};

struct FinTimes {
	double pack_ms = 0, upload_ms = 0, kernels_ms = 0, download_ms = 0;
	size_t texels = 0;
};

// blur_size and ds2x as the config gives them; kernel: the normalised
// Gaussian taps (2 ceil(blur_size) + 1 of them) when blur_size != 0.
// d_rgb, when not null, holds every image's sample colours on the device,
// back to back in image order (3 floats per sample; then FinImage::rgb is
// not read).  nthreads host threads pack and unpack.  Throws
// std::runtime_error on HIP errors.
void finalize_gpu(const std::vector<FinImage>& imgs, const float* d_rgb, float blur_size,
                  const std::vector<float>& kernel, bool ds2x, int nthreads, FinTimes* times);

} // namespace lgx
