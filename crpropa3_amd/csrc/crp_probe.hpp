// Measurement kernel behind crp_measure_gather.
#pragma once

#include <hip/hip_runtime.h>

namespace crp {

// Reads of W x 16 bytes from random 128-byte-aligned lines of a table, one line per lane and
// iteration: the access pattern of one cell-table field evaluation (W = 6) without the arithmetic.
template <int W>
__global__ void __launch_bounds__(256) k_probe_gather(const float4 *__restrict__ tab, unsigned nlines, int iters,
                                                      float *out) {
	const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
	unsigned h = tid * 2654435761u + 12345u;
	float acc = 0.f;
	for (int j = 0; j < iters; j++) {
		h ^= h >> 16; h *= 0x7feb352dU; h ^= h >> 15; h *= 0x846ca68bU; h ^= h >> 16;
		const float4 *p = tab + (size_t)(h % nlines) * 8; // line < nlines, W <= 8 float4 per line
		float4 v[W];
#pragma unroll
		for (int w = 0; w < W; w++)
			v[w] = p[w];
#pragma unroll
		for (int w = 0; w < W; w++)
			acc += (v[w].x + v[w].y) + (v[w].z + v[w].w);
		h += (unsigned)j;
	}
	out[tid] = acc;
}

} // namespace crp
