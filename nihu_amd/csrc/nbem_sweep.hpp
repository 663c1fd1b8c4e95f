// nbem_sweep.hpp -- Burton-Miller matrices for many wavenumbers on one mesh.
//
// Regular pairs: one persistent launch.  A workgroup of 256 lanes (one test
// row per lane) takes work items (row tile, trial-column tile, batch of KB
// wavenumbers) with a fixed stride until none are left.  For each pair the
// wavenumber-independent geometry of every quadrature point (r, r - d, the
// weight over r, 1/r^2 and the test-normal projection) is computed once and
// kept in registers while the KB wavenumbers of the batch reuse it; per
// wavenumber only the phase, the hardware sin/cos and the kernel algebra
// remain.  The FP64 near-singular and diagonal passes compute their
// wavenumber-independent (Laplace static) parts once per pair into a buffer
// and add the dynamic part per (pair, wavenumber).

#pragma once

#include "nbem_kernels.hpp"

namespace nbk {

struct SweepArgs {
	int n, row_tiles, col_tiles, tile_cols, n_k, n_items;
	int nr, nc;                 // rows / columns processed (n, or a leading block when tuning)
	const double *c;            // centres [3][n]
	const float *nu;            // unit normals [3][n]
	const double *thr;          // [NTHR][n]
	const QPoint *qt;           // quadrature table of the fixed degree [n][Q]
	const double *k_rev;        // [n_k] k / (2 pi), FP64
	const float *kf;            // [n_k] k
	const float *krevf;         // [n_k] k / (2 pi)
	const float2 *alpha;        // [n_k]
	float2 *A, *B;
	long long ld, kstride;
};

template <int KB, int Q>
__global__ void __launch_bounds__(256) sweep_regular_kernel(SweepArgs a)
{
	const auto sc = uni(a.c);
	const auto snu = uni(a.nu);
	const auto thr = uni(a.thr);
	const auto k_rev = uni(a.k_rev);
	const auto kf = uni(a.kf);
	const auto krevf = uni(a.krevf);
	const auto alpha = uni((const float *)a.alpha);
	for (int item = blockIdx.x; item < a.n_items; item += gridDim.x) {
		const int rt = item % a.row_tiles;
		const int rest = item / a.row_tiles;
		const int ct = rest % a.col_tiles;
		const int f0 = (rest / a.col_tiles) * KB;
		const int i = rt * 256 + threadIdx.x;
		const bool row_ok = i < a.nr;
		const int ii = row_ok ? i : a.nr - 1;
		const double cix = a.c[ii], ciy = a.c[a.n + ii], ciz = a.c[2 * a.n + ii];
		const float nxx = a.nu[ii], nxy = a.nu[a.n + ii], nxz = a.nu[2 * a.n + ii];
		row_ready(cix, ciy, ciz, nxx, nxy, nxz, ii);
		const int j0 = ct * a.tile_cols;
		const int j1 = min(j0 + a.tile_cols, a.nc);
		for (int j = j0; j < j1; ++j) {
			const double cjx = sc[j], cjy = sc[a.n + j], cjz = sc[2 * a.n + j];
			const double d2 = d2_pair(cix, ciy, ciz, cjx, cjy, cjz);
			const bool act = row_ok && !(d2 < thr[THR_NEAR * a.n + j]) && i != j;
			if (__ballot(act) == 0) continue;
			const float dcx = (float)(cjx - cix), dcy = (float)(cjy - ciy), dcz = (float)(cjz - ciz);
			const float dcx2 = 2.0f * dcx, dcy2 = 2.0f * dcy, dcz2 = 2.0f * dcz;
			const float d2f = (float)d2;
			double y = __builtin_amdgcn_rsq(d2);
			y = y * (1.5 - 0.5 * d2 * y * y);
			const double d = d2 * y;
			const float df = (float)d;
			const float nyx = snu[j], nyy = snu[a.n + j], nyz = snu[2 * a.n + j];
			const float av = dcx * nyx + dcy * nyy + dcz * nyz;
			const float nn = nxx * nyx + nxy * nyy + nxz * nyz;
			const float b0 = dcx * nxx + dcy * nxy + dcz * nxz;
			// wavenumber-independent geometry of the points: the phase
			// correction and, per kernel sum, the weight W that multiplies
			// e^{-i th} (g = w/r, G-type sums carry 1/r^2, Q = G (1 + i k r)
			// splits into a G sum and a k r G sum, b = r.n_x, w2 = b / r^2;
			// the E w2 sum of the hypersingular k^2 term has the weight
			// g w2 = w b / r^3 of the G b sum and reuses it)
			float gdel[Q], W[7][Q];
			const auto qf = uni((const float *)(a.qt + (size_t)j * Q));
#pragma unroll
			for (int q = 0; q < Q; ++q) {
				const float ox = qf[8 * q], oy = qf[8 * q + 1], oz = qf[8 * q + 2], w = qf[8 * q + 3], oo = qf[8 * q + 4];
				const float u = __builtin_fmaf(dcx2, ox, __builtin_fmaf(dcy2, oy, __builtin_fmaf(dcz2, oz, oo)));
				const float r2 = d2f + u;
				const float rinv = __builtin_amdgcn_rsqf(r2);
				const float r = r2 * rinv;
				gdel[q] = u * __builtin_amdgcn_rcpf(r + df);
				const float g = w * rinv;
				const float gi2 = g * (rinv * rinv);
				const float b = __builtin_fmaf(oz, nxz, __builtin_fmaf(oy, nxy, __builtin_fmaf(ox, nxx, b0)));
				const float w2 = b * (rinv * rinv);
				W[0][q] = g;              // SLP:  sum E
				W[1][q] = gi2;            // sum G
				W[2][q] = gi2 * r;        // sum r G
				W[3][q] = gi2 * b;        // sum G b
				W[4][q] = W[3][q] * r;    // sum r G b
				W[5][q] = gi2 * w2;       // sum G w2
				W[6][q] = W[5][q] * r;    // sum r G w2
			}
			for (int fb = 0; fb < KB; ++fb) {
				const int f = f0 + fb;
				if (f >= a.n_k) break;
				double td = d * k_rev[f];
				td -= rint(td);
				const float t0 = (float)td;
				const float k = kf[f], kr = krevf[f];
				// A[m] = sum_q W[m][q] (cos t_q, sin t_q); e^{-i th} = (cos, -sin)
				float2 A[7];
#pragma unroll
				for (int m = 0; m < 7; ++m) A[m] = zero2();
#pragma unroll
				for (int q = 0; q < Q; ++q) {
					const float t = __builtin_fmaf(kr, gdel[q], t0);
					const float c = __builtin_amdgcn_cosf(t), sn = __builtin_amdgcn_sinf(t);
#pragma unroll
					for (int m = 0; m < 7; ++m) {
						A[m].x = __builtin_fmaf(W[m][q], c, A[m].x);
						A[m].y = __builtin_fmaf(W[m][q], sn, A[m].y);
					}
				}
				// complex sums: S_m = (A_m.x, -A_m.y); sum Q = S1 + i k S2 etc.
				const float2 s0 = make_float2(A[0].x, -A[0].y);
				const float2 s1 = make_float2(A[1].x + k * A[2].y, -A[1].y + k * A[2].x);
				const float2 s2 = make_float2(A[3].x + k * A[4].y, -A[3].y + k * A[4].x);
				const float2 sqw = make_float2(A[5].x + k * A[6].y, -A[5].y + k * A[6].x);
				const float kk = k * k;
				const float2 s3 = make_float2(3.0f * sqw.x - kk * A[3].x, 3.0f * sqw.y + kk * A[3].y);
				if (act) {
					// SLP = s0, DLP = -a s1, DLPt = s2, HSP = nn s1 - a s3
					const float2 al = make_float2(alpha[2 * f], alpha[2 * f + 1]);
					const float hsr = nn * s1.x - av * s3.x, hsi = nn * s1.y - av * s3.y;
					const float ar = al.x * hsr - al.y * hsi - av * s1.x;
					const float ai = al.x * hsi + al.y * hsr - av * s1.y;
					const float br = s0.x + al.x * s2.x - al.y * s2.y;
					const float bi = s0.y + al.x * s2.y + al.y * s2.x;
					const long long idx = (long long)f * a.kstride + (long long)j * a.ld + i;
					a.A[idx] = make_float2(ar, ai);
					a.B[idx] = make_float2(br, bi);
				}
			}
		}
	}
}

// NiHu's distance-driven policy over many wavenumbers in one persistent
// launch: per pair the FP64 distance, the near-singular test and the
// per-kernel degrees of the distance tables are found once; each wavenumber
// of the work item's batch then reduces its own phase of k d in FP64 and
// evaluates the pair as the single-wavenumber pass does (the one-point rule
// for waves beyond the last breakpoint, else each distinct degree of the wave
// once for all kernel groups), so every wavenumber's matrices equal those of
// one Burton-Miller assembly at that wavenumber.
struct SweepNihuArgs {
	RegArgs r;                  // the mesh's tables, thresholds and records (its k, alpha, out unused)
	int n_k, n_items, col_tiles;
	const double *k_rev;        // [n_k] k / (2 pi), FP64
	const float *kf;            // [n_k] k
	const float *krevf;         // [n_k] k / (2 pi)
	const float2 *alpha;        // [n_k]
	float2 *A, *B;
	long long ld, kstride;
};

template <int KB>
__global__ void __launch_bounds__(256) sweep_nihu_kernel(SweepNihuArgs sa)
{
	const RegArgs &a = sa.r;
	const auto rec = uni(a.rec);
	const auto k_rev = uni(sa.k_rev);
	const auto kf = uni(sa.kf);
	const auto krevf = uni(sa.krevf);
	const auto alpha = uni((const float *)sa.alpha);
	for (int item = blockIdx.x; item < sa.n_items; item += gridDim.x) {
		const int rt = item % a.row_tiles;
		const int rest = item / a.row_tiles;
		const int ct = rest % sa.col_tiles;
		const int f0 = (rest / sa.col_tiles) * KB;
		const int i = rt * 256 + threadIdx.x;
		const bool row_ok = i < a.n_rows;
		const int ii = row_ok ? i : a.n_rows - 1;
		const double cix = a.tc[ii], ciy = a.tc[a.tn + ii], ciz = a.tc[2 * a.tn + ii];
		PairCtx p;
		p.nxx = a.tnu[ii]; p.nxy = a.tnu[a.tn + ii]; p.nxz = a.tnu[2 * a.tn + ii];
		row_ready(cix, ciy, ciz, p.nxx, p.nxy, p.nxz, ii);
		const int j0 = ct * a.tile_cols;
		const int j1 = min(j0 + a.tile_cols, a.n_cols);
		for (int j = j0; j < j1; ++j) {
			const double dxd = rec[j].c[0] - cix, dyd = rec[j].c[1] - ciy, dzd = rec[j].c[2] - ciz;
			const double d2 = __fma_rn(dzd, dzd, __fma_rn(dyd, dyd, dxd * dxd));
			const bool act = row_ok && !(d2 < rec[j].thr[THR_NEAR]) && i != j;
			if (__ballot(act) == 0) continue;
			p.dcx = (float)dxd; p.dcy = (float)dyd; p.dcz = (float)dzd;
			p.d2f = (float)d2;
			double y = __builtin_amdgcn_rsq(d2);
			y = y * (1.5 - 0.5 * d2 * y * y);
			const double d = d2 * y;
			const float rinv_d = (float)y;
			p.df = (float)d;
			const float nyx = rec[j].nu[0], nyy = rec[j].nu[1], nyz = rec[j].nu[2], w1 = rec[j].w1;
			const float av = p.dcx * nyx + p.dcy * nyy + p.dcz * nyz;
			const float nn = p.nxx * nyx + p.nxy * nyy + p.nxz * nyz;
			p.b0 = p.dcx * p.nxx + p.dcy * p.nxy + p.dcz * p.nxz;
			const bool allfar = __ballot(act && d2 < rec[j].thr[THR_22_3]) == 0;
			// degrees: the field complexity plus the distance-table degree; in
			// an all-far wave every kernel takes the field complexity alone
			const int fc = a.field_complexity;
			int dS = act ? fc : -1, dD = dS, dH = dS;
			if (!allfar) {
				double t[NTHR];
#pragma unroll
				for (int q = 0; q < NTHR; ++q) t[q] = rec[j].thr[q];
				dS = act ? fc + table_degree_rec<1>(d2, t) : -1;
				dD = act ? fc + table_degree_rec<2>(d2, t) : -1;
				dH = act ? fc + table_degree_rec<3>(d2, t) : -1;
			}
			const long long col = (long long)j * sa.ld;
			for (int fb = 0; fb < KB; ++fb) {
				const int f = f0 + fb;
				if (f >= sa.n_k) break;
				double tt = d * k_rev[f];
				tt -= rint(tt);
				p.t0 = (float)tt;
				const float k = kf[f];
				float2 s[6] = {zero2(), zero2(), zero2(), zero2(), zero2(), zero2()};
				if (allfar && fc <= 1) {
					if (act) accumulate_centroid<HELM, ALL4>(p, w1, rinv_d, k, s);
				} else if (allfar) {
					eval_degrees<HELM, ALL4>(a, j, dS, p, w1, rinv_d, k, krevf[f], s);
				} else {
					eval_groups<HELM, ALL4>(a, j, dS, dD, dH, p, w1, rinv_d, k, krevf[f], s);
				}
				if (act) {
					float2 K[4];
					kernel_values<HELM, ALL4>(s, av, nn, k, !allfar, K);
					const float2 al = make_float2(alpha[2 * f], alpha[2 * f + 1]);
					const float ar = al.x * K[3].x - al.y * K[3].y + K[1].x;
					const float ai = al.x * K[3].y + al.y * K[3].x + K[1].y;
					const float br = K[0].x + al.x * K[2].x - al.y * K[2].y;
					const float bi = K[0].y + al.x * K[2].y + al.y * K[2].x;
					const long long at = (long long)f * sa.kstride + col;
					*entry_ptr<float2>(sa.A, at, (unsigned)i * 8u) = make_float2(ar, ai);
					*entry_ptr<float2>(sa.B, at, (unsigned)i * 8u) = make_float2(br, bi);
				}
			}
		}
	}
}

// Wavenumber-independent parts of the FP64 pairs: near pairs (p < n_near)
// store the Laplace static integrals of SLP, DLP, DLPt, HSP; diagonal pairs
// (n_near <= p < n_near + n) store diag_static's four sums.
struct StaticArgs {
	int n, n_near;
	const int *pi, *pj;
	const double *c, *nu64, *X;
	const Rules *rules;
	double *st;                 // [(n_near + n) * 4]
};

__global__ void __launch_bounds__(64) fp64_static_kernel(StaticArgs a)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= a.n_near + a.n) return;
	const bool diag = p >= a.n_near;
	const int i = diag ? p - a.n_near : a.pi[p];
	const int j = diag ? i : a.pj[p];
	D3 V[3];
	for (int v = 0; v < 3; ++v)
		V[v] = {a.X[(3 * v) * a.n + j], a.X[(3 * v + 1) * a.n + j], a.X[(3 * v + 2) * a.n + j]};
	const D3 x0 = {a.c[i], a.c[a.n + i], a.c[2 * a.n + i]};
	double st[4];
	if (diag) {
		diag_static(V, x0, st);
	} else {
		const D3 nx = {a.nu64[i], a.nu64[a.n + i], a.nu64[2 * a.n + i]};
		near_static<ALL4>(V, x0, nx, *a.rules, st);
	}
	for (int t = 0; t < 4; ++t) a.st[(size_t)p * 4 + t] = st[t];
}

struct DynArgs {
	int n, n_near, n_k;
	int p0;                     // first pair of the launch (pairs p0 .. n_near + n - 1)
	const int *pi, *pj;
	const double *c, *nu64, *X, *J;
	const Rules *rules;
	const double *st;
	const double *k;            // [n_k]
	const double2 *alpha;       // [n_k]
	float2 *A, *B;
	long long ld, kstride;
};

// one lane per (FP64 pair, wavenumber): consecutive lanes take consecutive pairs
__global__ void __launch_bounds__(64) fp64_dynamic_kernel(DynArgs a)
{
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long np = (long long)a.n_near + a.n - a.p0;
	if (t >= np * a.n_k) return;
	const int f = (int)(t / np);
	const int p = a.p0 + (int)(t % np);
	const bool diag = p >= a.n_near;
	const int i = diag ? p - a.n_near : a.pi[p];
	const int j = diag ? i : a.pj[p];
	const Rules &R = *a.rules;
	D3 V[3];
	for (int v = 0; v < 3; ++v)
		V[v] = {a.X[(3 * v) * a.n + j], a.X[(3 * v + 1) * a.n + j], a.X[(3 * v + 2) * a.n + j]};
	const D3 x0 = {a.c[i], a.c[a.n + i], a.c[2 * a.n + i]};
	const double k = a.k[f];
	const double J = a.J[j];
	double st[4];
	for (int q = 0; q < 4; ++q) st[q] = a.st[(size_t)p * 4 + q];
	double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
	const double2 al = a.alpha[f];
	double ar, ai, br, bi;
	if (diag) {
		double2 dyn[2];
		diag_dynamic<SLP | HSP>(V, x0, J, k, R, dyn);
		diag_combine<HELM, SLP | HSP>(st, dyn, k, s);
	} else {
		const D3 nx = {a.nu64[i], a.nu64[a.n + i], a.nu64[2 * a.n + i]};
		const D3 ny = {a.nu64[j], a.nu64[a.n + j], a.nu64[2 * a.n + j]};
		for (int q = 0; q < 4; ++q) s[q].x = st[q];
		near_dynamic<ALL4>(V, x0, nx, ny, J, k, R, s);
	}
	ar = al.x * s[3].x - al.y * s[3].y + s[1].x;
	ai = al.x * s[3].y + al.y * s[3].x + s[1].y;
	br = s[0].x + al.x * s[2].x - al.y * s[2].y;
	bi = s[0].y + al.x * s[2].y + al.y * s[2].x;
	if (diag) {
		ar += -0.5;
		br += al.x / 2.0;
		bi += al.y / 2.0;
	}
	const long long idx = (long long)f * a.kstride + (long long)j * a.ld + i;
	a.A[idx] = make_float2((float)ar, (float)ai);
	a.B[idx] = make_float2((float)br, (float)bi);
}

// Near pairs over uniformly spaced wavenumbers k_f = k_0 + f dk: one lane per
// (near pair, block of KBLK consecutive wavenumbers); at each of the 42 points
// e^{-i k r} comes from sincos once per block and is rotated by e^{-i dk r}
// (one FP64 complex product) from one wavenumber to the next; the other
// factors use the actual k_f.  Same terms as fp64_dynamic_kernel otherwise.
constexpr int KBLK = 4;

__global__ void __launch_bounds__(64) fp64_dynamic_rec_kernel(DynArgs a, double dk)
{
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int nblk = (a.n_k + KBLK - 1) / KBLK;
	if (t >= (long long)a.n_near * nblk) return;
	const int fb = (int)(t / a.n_near);
	const int p = (int)(t % a.n_near);
	const int i = a.pi[p], j = a.pj[p];
	const Rules &R = *a.rules;
	D3 V[3];
	for (int v = 0; v < 3; ++v)
		V[v] = {a.X[(3 * v) * a.n + j], a.X[(3 * v + 1) * a.n + j], a.X[(3 * v + 2) * a.n + j]};
	const D3 x0 = {a.c[i], a.c[a.n + i], a.c[2 * a.n + i]};
	const D3 nx = {a.nu64[i], a.nu64[a.n + i], a.nu64[2 * a.n + i]};
	const D3 ny = {a.nu64[j], a.nu64[a.n + j], a.nu64[2 * a.n + j]};
	const double J = a.J[j];
	const int f0 = fb * KBLK;
	double kf[KBLK];
	double2 s[KBLK][4];
#pragma unroll
	for (int m = 0; m < KBLK; ++m) {
		kf[m] = a.k[min(f0 + m, a.n_k - 1)];
		for (int q = 0; q < 4; ++q) s[m][q] = make_double2(a.st[(size_t)p * 4 + q], 0.0);
	}
	const double nn = d3dot(nx, ny);
	for (int q = 0; q < 42; ++q) {
		const D3 y = tri_x(V, R.d14xi[q], R.d14eta[q]);
		const D3 rv = d3sub(y, x0);
		const double r2 = d3dot(rv, rv);
		const double rinv = rsqrt64(r2);
		const double r = r2 * rinv;
		const double g0 = rinv * (R.d14w[q] * J) * (1.0 / (4.0 * PI_D));
		const double rdny = d3dot(rv, ny) * rinv;
		const double rdnx = -d3dot(rv, nx) * rinv;
		double sn, cs, sd, cd;
		sincos64(kf[0] * r, &sn, &cs);
		sincos64(dk * r, &sd, &cd);
#pragma unroll
		for (int m = 0; m < KBLK; ++m) {
			dynamic_terms<ALL4>(cs, sn, kf[m] * r, g0, rinv, rdny, rdnx, nn, s[m]);
			// e^{-i (k + dk) r} = e^{-i k r} e^{-i dk r}
			const double c2 = cs * cd - sn * sd;
			sn = sn * cd + cs * sd;
			cs = c2;
		}
	}
#pragma unroll
	for (int m = 0; m < KBLK; ++m) {
		const int f = f0 + m;
		if (f >= a.n_k) break;
		const double2 al = a.alpha[f];
		const double ar = al.x * s[m][3].x - al.y * s[m][3].y + s[m][1].x;
		const double ai = al.x * s[m][3].y + al.y * s[m][3].x + s[m][1].y;
		const double br = s[m][0].x + al.x * s[m][2].x - al.y * s[m][2].y;
		const double bi = s[m][0].y + al.x * s[m][2].y + al.y * s[m][2].x;
		const long long idx = (long long)f * a.kstride + (long long)j * a.ld + i;
		a.A[idx] = make_float2((float)ar, (float)ai);
		a.B[idx] = make_float2((float)br, (float)bi);
	}
}

} // namespace nbk
