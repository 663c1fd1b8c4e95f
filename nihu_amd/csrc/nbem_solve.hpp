// nbem_solve.hpp -- matrix-vector product of the solve step.
//
// y = A x with A a column-major complex64 matrix (the assembled Burton-Miller
// or collocation matrix, rows x cols, leading dimension ld) and x, y complex128
// vectors, every product and sum in FP64: the Krylov solver of the serving
// loop (NiHu's application solves with GMRES:
// nihu/applications/helmholtz/helmholtz_3d_coll.cpp:99-108) then converges to
// the solution of the stored complex64 system without the matrix ever being
// widened.  Bound by reading A from HBM (8 bytes per entry).
//
// Pass 1: a workgroup of 256 lanes (one row per lane) runs over a chunk of
// columns; x_j is wave-uniform (scalar loads) and each column's 256 entries
// are read as contiguous 2 KB; every lane keeps its partial sum in
// registers and writes it to work[chunk][row].  Pass 2 sums the chunks of a
// row in order, so the result does not depend on scheduling.

#pragma once

#include "nbem_kernels.hpp"

namespace nbk {

struct GemvArgs {
	const float2 *A;
	long long rows, cols, ld;
	const double2 *x;
	double2 *y;
	double2 *work;      // [chunks][rows]
	int row_tiles, chunks;
	long long clen;     // columns per chunk
};

__global__ void __launch_bounds__(256) gemv_partial_kernel(GemvArgs a)
{
	const int tile = blockIdx.x % a.row_tiles;
	const int ch = blockIdx.x / a.row_tiles;
	const long long i = (long long)tile * 256 + threadIdx.x;
	const bool ok = i < a.rows;
	const long long ii = ok ? i : a.rows - 1;
	const long long j0 = (long long)ch * a.clen;
	const long long j1 = min(j0 + a.clen, a.cols);
	const auto x = uni((const double *)a.x);   // x[2 j], x[2 j + 1]: re, im of x_j
	double re = 0.0, im = 0.0;
	long long j = j0;
	// eight columns per step: eight independent loads in flight per lane
	for (; j + 8 <= j1; j += 8) {
		float2 v[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) v[u] = a.A[(j + u) * a.ld + ii];
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			const double xr = x[2 * (j + u)], xi = x[2 * (j + u) + 1];
			const double ar = v[u].x, ai = v[u].y;
			re = __fma_rn(ar, xr, __fma_rn(-ai, xi, re));
			im = __fma_rn(ar, xi, __fma_rn(ai, xr, im));
		}
	}
	for (; j < j1; ++j) {
		const float2 v = a.A[j * a.ld + ii];
		const double xr = x[2 * j], xi = x[2 * j + 1];
		const double ar = v.x, ai = v.y;
		re = __fma_rn(ar, xr, __fma_rn(-ai, xi, re));
		im = __fma_rn(ar, xi, __fma_rn(ai, xr, im));
	}
	if (ok) a.work[(long long)ch * a.rows + i] = make_double2(re, im);
}

__global__ void __launch_bounds__(256) gemv_reduce_kernel(GemvArgs a)
{
	const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
	if (i >= a.rows) return;
	double re = 0.0, im = 0.0;
	for (int c = 0; c < a.chunks; ++c) {
		const double2 v = a.work[(long long)c * a.rows + i];
		re += v.x;
		im += v.y;
	}
	a.y[i] = make_double2(re, im);
}

// launch shape: enough chunks for >= 4096 workgroups, chunks of a multiple of
// 8 columns and at least 64
inline void gemv_shape(long long rows, long long cols, int *row_tiles, int *chunks, long long *clen)
{
	const long long rt = (rows + 255) / 256;
	long long want = (4096 + rt - 1) / rt;
	long long cl = (cols + want - 1) / want;
	cl = ((cl + 7) / 8) * 8;
	if (cl < 64) cl = 64;
	*row_tiles = (int)rt;
	*clen = cl;
	*chunks = (int)((cols + cl - 1) / cl);
}

} // namespace nbk
