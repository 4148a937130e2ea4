// Sample placement on the GPU: OffsetSample and the correction in front of
// overlapping faces (lighter.cpp:436-466, lighter_math.cpp:991-1044) for
// every rasterised texel of the scene in one launch.  The kernel walks the
// host's own per-instance trees (BoxTree, SplitTree) and the instance tree,
// flattened as they are, in the host's visit orders, and evaluates each
// triangle with the same shared helpers (geom.h), so it places every sample
// exactly where the host restatement does.
#pragma once

#include <stddef.h>

#include <vector>

#include "host_geom.h"

namespace lgx {

// One instance's collision data as the placement reads it.
struct PlaceInstance {
	const TriSet* tris;
	const SplitTree* split;
	bool shadow;
};

// Host wall time of the steps around the kernel and the kernel itself (HIP
// events on the placement stream), milliseconds.
struct PlaceTimes {
	double flatten_ms = 0, upload_ms = 0, kernel_ms = 0, download_ms = 0;
	size_t samples = 0;
};

// One run of samples (one instance's, say): pos moved in place, nrm
// (normalised) and dist (the square root of the texel's world area), n each.
struct PlaceSamples {
	f3* pos;
	const f3* nrm;
	const float* dist;
	size_t n;
};

// Places every run's samples in one launch.  max_corr and corr_min_dot as
// the config gives them (corr_min_dot is the cosine of max_correct_angle).
// The trees are flattened and the runs packed and unpacked on nthreads host
// threads.  Throws std::runtime_error on HIP errors or when a tree is deeper
// than the kernel's stacks.
void place_samples_gpu(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree,
                       const std::vector<PlaceSamples>& runs, float max_corr, float corr_min_dot, int nthreads,
                       PlaceTimes* times);

} // namespace lgx
