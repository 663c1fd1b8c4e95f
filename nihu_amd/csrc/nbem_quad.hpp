// nbem_quad.hpp -- quadrilateral elements (4-node bilinear, possibly warped)
// with constant fields in collocation: the regular FP32 pass and the FP64
// near-singular / singular pass (gfx950).
//
// Geometry (nihu/library/lib_shape.hpp:215-310, lib_domain.cpp:74-94):
//   x(xi, eta) = sum_a N_a X_a on [-1, 1]^2, N = (1 -+ xi)(1 -+ eta) / 4 with
//   corners (-1,-1), (1,-1), (1,1), (-1,1); n = dx/dxi x dx/deta varies over
//   the element; collocation point x(0, 0); R = mean edge length.
// Data in HBM per quadrilateral mesh (SoA, n elements): corners X[12][n]
// FP64, centres c[3][n] FP64, unit normals at the centre; quadrature tables
// per Gauss order N (N x N points, N = degree / 2 + 1):
//   QuadPt[n][N*N] = offset from the centre, w J / (4 pi), unit normal, |offset|^2
// (FP32).
//
// Regular pass: as for triangles, one lane per test row, a wave-uniform trial
// column whose record (thresholds, centre, normal, one-point weight) comes
// through the scalar cache; the regular degree is 1 (the jacobian order of
// quad_1) plus the distance-table degree (nihu/core/complexity_estimator.hpp:
// 48-62), a Gauss square rule; the normal varies over the element, so r.n_y
// and n_x.n_y are per point.  The one-point rule (every far pair) is the
// centre with weight 4 J(0, 0): no table load, FP64-reduced phase.
// FP64 pass:
//   diagonal (one lane per element): SLP / DLP / DLPt by the order-7 Duffy blind rule around the
//     centre (nihu/core/singular_accelerator.hpp:363-456,
//     nihu/core/duffy_quadrature.hpp); HSP by Guiggiani's method with orders
//     (5, 9) (nihu/library/guiggiani_1992.hpp:138-330, Laurent coefficients
//     of the Laplace HSP core nihu/library/laplace_kernel.hpp:266-295);
//   near (d / R < limit; 4 lanes per pair): Helmholtz SLP and Laplace SLP / DLP / DLPt by polar
//     integration around the image of x0 of K J - K_lin J_lin plus the planar
//     integral over the linearised element (orders 15 x 15; Laplace HSP
//     100 x 100) (nihu/library/nearly_singular_collocational.hpp:71-203);
//     Helmholtz DLP / DLPt / HSP by the Telles-transformed order-30 Gauss rule
//     (nihu/library/nearly_singular_collocational_telles.hpp:66-103,
//     nihu/library/telles_1987.hpp:30-96); the image of x0 by Newton-Raphson
//     (nihu/core/inverse_mapping.hpp:37-118).

#pragma once

#include "nbem_kernels.hpp"

namespace nbk {

constexpr int QN_MAX = 8;    // largest Gauss order per axis of the regular tables

struct QuadPt { float ox, oy, oz, w, nx, ny, nz, oo; };

struct QuadArgs {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	const int *rmap, *cmap;      // result row / column of each test / trial element (null: its own index)
	int tn, sn;
	const double *tc;            // test collocation points [3][tn]
	const float *tnu;            // test unit normals [3][tn]
	const TrialRec *rec;         // trial records [sn]: thresholds, centre, unit normal, 4 J / (4 pi)
	const QuadPt *qt[QN_MAX + 1];   // by Gauss order per axis
	int fixed_n;                 // > 0: fixed rule with fixed_n x fixed_n points
	float k, k_revf;
	double k_rev;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	float2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

struct QuadCtx {
	float dcx, dcy, dcz;         // c_j - x (FP32)
	float d2f, df, t0;           // |c_j - x|^2, |c_j - x|, reduced phase of k d (revolutions)
	float rinv_d;                // 1 / |c_j - x| from FP64
	float nxx, nxy, nxz;         // test unit normal
	float nyx, nyy, nyz;         // trial unit normal at the centre
	float w1;                    // weight of the one-point rule: 4 J(0, 0) / (4 pi)
	float ad;                    // e^{ki d} (complex wavenumber k + i ki; CK only)
};

// add one point's E = (ec, es), th = k r, 1/r, a = r.n_y, b = r.n_x, nn = n_x.n_y;
// CK (complex wavenumber k + i ki): E carries e^{ki r}, kih = ki r, and
// i (k + i ki) r = -kih + i th, (k + i ki)^2 = k^2 - ki^2 + 2 i k ki
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void quad_add(float ec, float es, float th, float rinv, float a, float b, float nn, float k,
	float2 (&K)[4], float kih = 0.0f, float ki = 0.0f)
{
	if (MASK & SLP) { K[0].x += ec; K[0].y += es; }
	if (MASK & (DLP | DLPT | HSP)) {
		const float ri2 = rinv * rinv;
		const float Gr = ec * ri2, Gi = es * ri2;
		float Qr = __builtin_fmaf(-th, Gi, Gr), Qi = __builtin_fmaf(th, Gr, Gi);
		if (CK) { Qr = __builtin_fmaf(-kih, Gr, Qr); Qi = __builtin_fmaf(-kih, Gi, Qi); }
		if (MASK & DLP) { K[1].x = __builtin_fmaf(-a, Qr, K[1].x); K[1].y = __builtin_fmaf(-a, Qi, K[1].y); }
		if (MASK & DLPT) { K[2].x = __builtin_fmaf(b, Qr, K[2].x); K[2].y = __builtin_fmaf(b, Qi, K[2].y); }
		if (MASK & HSP) {
			const float kk = FAM == HELM ? (CK ? __builtin_fmaf(k, k, -ki * ki) : k * k) : 0.0f;
			const float ab = a * b * ri2;
			float hr = __builtin_fmaf(3.0f, Qr, -kk * ec), hi = __builtin_fmaf(3.0f, Qi, -kk * es);
			if (CK) {
				const float kq = 2.0f * k * ki;
				hr = __builtin_fmaf(kq, es, hr);
				hi = __builtin_fmaf(-kq, ec, hi);
			}
			K[3].x += __builtin_fmaf(nn, Qr, -ab * hr);
			K[3].y += __builtin_fmaf(nn, Qi, -ab * hi);
		}
	}
}

// One-point Gauss rule (order 1): the point is the centre, so r = d with the
// FP64-reduced phase, and the normal is the centre's.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void quad_centre(const QuadCtx &p, float k, float2 (&K)[4], float ki = 0.0f)
{
	const float g = CK ? p.w1 * p.rinv_d * p.ad : p.w1 * p.rinv_d;
	float ec = g, es = 0.0f, th = 0.0f;
	if (FAM == HELM) {
		ec = g * __builtin_amdgcn_cosf(p.t0);
		es = -g * __builtin_amdgcn_sinf(p.t0);
		th = k * p.df;
	}
	float a = 0.0f, b = 0.0f, nn = 0.0f;
	if (MASK & (DLP | HSP)) a = p.dcx * p.nyx + p.dcy * p.nyy + p.dcz * p.nyz;
	if (MASK & (DLPT | HSP)) b = p.dcx * p.nxx + p.dcy * p.nxy + p.dcz * p.nxz;
	if (MASK & HSP) nn = p.nxx * p.nyx + p.nxy * p.nyy + p.nxz * p.nyz;
	quad_add<FAM, MASK, CK>(ec, es, th, p.rinv_d, a, b, nn, k, K, CK ? ki * p.df : 0.0f, ki);
}

// Kernel values times the weights at the points of one trial element (FP32):
//   SLP  E,  DLP -a Q,  DLPt b Q,  HSP  nn Q - a b (3 Q - k^2 E) / r^2
// with E = w e^{-ikr} / r, Q = E (1 + ikr) / r^2, a = r.n_y, b = r.n_x,
// nn = n_x.n_y at the point (the kernels of nihu/library/helmholtz_kernel.hpp:
// 169-228 and normal_derivative_kernel.hpp:114-257 with r = y - x).  The
// phase is t0 + (k / 2pi) u / (r + d) revolutions, u = r^2 - d^2 =
// 2 (c_j - x).off + |off|^2, as in the triangle pass.  CK: the weight takes
// e^{ki r} = e^{ki d} 2^{ki log2(e) (r - d)}.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void quad_points(const QuadPt *tab, int nq, const QuadCtx &p, float k_revf, float k,
	float2 (&K)[4], float ki = 0.0f)
{
	const float kil2e = ki * 1.4426950408889634f;
	const auto qf = uni((const float *)tab);
	const float dcx2 = 2.0f * p.dcx, dcy2 = 2.0f * p.dcy, dcz2 = 2.0f * p.dcz;
	for (int q = 0; q < nq; ++q) {
		const float ox = qf[8 * q], oy = qf[8 * q + 1], oz = qf[8 * q + 2], w = qf[8 * q + 3];
		const float nqx = qf[8 * q + 4], nqy = qf[8 * q + 5], nqz = qf[8 * q + 6], oo = qf[8 * q + 7];
		const float u = __builtin_fmaf(dcx2, ox, __builtin_fmaf(dcy2, oy, __builtin_fmaf(dcz2, oz, oo)));
		const float r2 = p.d2f + u;
		const float rinv = __builtin_amdgcn_rsqf(r2);
		float g = w * rinv;
		float ec, es, th = 0.0f, kih = 0.0f;
		if (FAM == HELM) {
			const float r = r2 * rinv;
			const float rd = u * __builtin_amdgcn_rcpf(r + p.df);
			const float t = __builtin_fmaf(k_revf, rd, p.t0);
			if (CK) {
				g *= p.ad * exp2f(kil2e * rd);
				kih = ki * r;
			}
			ec = g * __builtin_amdgcn_cosf(t);
			es = -g * __builtin_amdgcn_sinf(t);
			th = k * r;
		} else {
			ec = g;
			es = 0.0f;
		}
		const float rx = p.dcx + ox, ry = p.dcy + oy, rz = p.dcz + oz;
		float a = 0.0f, b = 0.0f, nn = 0.0f;
		if (MASK & (DLP | HSP)) a = rx * nqx + ry * nqy + rz * nqz;
		if (MASK & (DLPT | HSP)) b = rx * p.nxx + ry * p.nxy + rz * p.nxz;
		if (MASK & HSP) nn = p.nxx * nqx + p.nxy * nqy + p.nxz * nqz;
		quad_add<FAM, MASK, CK>(ec, es, th, rinv, a, b, nn, k, K, kih, ki);
	}
}

// Lanes at their own Gauss order n (n <= 0: idle), one distinct order at a
// time so that the table address stays wave-uniform.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void quad_orders(const QuadArgs &a, int j, int n, const QuadCtx &p, float2 (&K)[4])
{
	const float ki = (float)a.ki;
	if (__ballot(n == 1) != 0) {
		if (n == 1) quad_centre<FAM, MASK, CK>(p, a.k, K, ki);
		n = n == 1 ? 0 : n;
	}
	if (__ballot(n > 0) == 0) return;
	for (;;) {
		const unsigned long long m = __ballot(n > 0);
		if (m == 0) break;
		const int lane = __ffsll((long long)m) - 1;
		const int nsel = __builtin_amdgcn_readlane(n, lane);
		const unsigned long long msel = __ballot(n == nsel);
		if ((msel >> __lane_id()) & 1ull) {
			const int nq = nsel * nsel;
			quad_points<FAM, MASK, CK>(a.qt[nsel] + (size_t)j * nq, nq, p, a.k_revf, a.k, K, ki);
			n = 0;
		}
	}
}

// Kernel groups at different Gauss orders (per lane: SLP nS, DLP / DLPt nD,
// HSP nH; <= 0 idle): each distinct order present in the wave is evaluated
// once with every kernel of KM into temporary values, and each group whose
// order for this lane is the current one takes its kernel values by a select
// (every group is served once, so its value is the rule's sum from -0 in
// point order, as a separate evaluation per group gives it).
template <int FAM, int KM, bool CK = false>
__device__ __forceinline__ void quad_groups(const QuadArgs &a, int j, int nS, int nD, int nH, const QuadCtx &p,
	float2 (&K)[4])
{
	constexpr int NONE = 1 << 20;
	const float ki = (float)a.ki;
	int pS = (KM & SLP) && nS > 0 ? nS : NONE;
	int pD = (KM & (DLP | DLPT)) && nD > 0 ? nD : NONE;
	int pH = (KM & HSP) && nH > 0 ? nH : NONE;
	for (;;) {
		const int mine = min(pS, min(pD, pH));
		const unsigned long long m = __ballot(mine < NONE);
		if (m == 0) break;
		const int lane = __ffsll((long long)m) - 1;
		const int nsel = __builtin_amdgcn_readlane(mine, lane);
		const bool fS = pS == nsel, fD = pD == nsel, fH = pH == nsel;
		float2 t[4] = {zero2(), zero2(), zero2(), zero2()};
		if (fS || fD || fH) {
			if (nsel == 1) quad_centre<FAM, KM, CK>(p, a.k, t, ki);
			else quad_points<FAM, KM, CK>(a.qt[nsel] + (size_t)j * (nsel * nsel), nsel * nsel, p, a.k_revf, a.k, t,
				ki);
			if (fS) K[0] = t[0];
			if (fD) { K[1] = t[1]; K[2] = t[2]; }
			if (fH) K[3] = t[3];
		}
		pS = fS ? NONE : pS;
		pD = fD ? NONE : pD;
		pH = fH ? NONE : pH;
	}
}

// Gauss order per axis of a Gauss square rule of a degree
__device__ __forceinline__ int gauss_order(int degree) { return degree / 2 + 1; }

// MAP: stores through a.rmap / a.cmap (a block of a mesh with both element
// types); without it the plain j * ld + i
template <int FAM, int KM, int MODE, bool MAP, bool CK = false>
__global__ void __launch_bounds__(256) quad_regular_kernel(QuadArgs a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int tile_c = blockIdx.x / a.row_tiles;
	const int i = tile_r * 256 + threadIdx.x;
	const bool row_ok = i < a.n_rows;
	const int ii = row_ok ? i : a.n_rows - 1;
	const int orow = MAP && a.rmap ? a.rmap[ii] : i;
	const double cix = a.tc[ii], ciy = a.tc[a.tn + ii], ciz = a.tc[2 * a.tn + ii];
	QuadCtx p;
	p.nxx = p.nxy = p.nxz = 0.0f;
	if (KM & (DLPT | HSP)) { p.nxx = a.tnu[ii]; p.nxy = a.tnu[a.tn + ii]; p.nxz = a.tnu[2 * a.tn + ii]; }
	row_ready(cix, ciy, ciz, p.nxx, p.nxy, p.nxz, orow);
	const int j0 = tile_c * a.tile_cols;
	const int j1 = min(j0 + a.tile_cols, a.n_cols);
	const auto rec = uni(a.rec);
	// the trial record fields of the one-point path, loaded one column ahead
	struct Head { double cx, cy, cz, near, far; float nx, ny, nz, w1; };
	auto head = [&](int jj) {
		return Head{rec[jj].c[0], rec[jj].c[1], rec[jj].c[2], rec[jj].thr[THR_NEAR], rec[jj].thr[THR_22_3],
			rec[jj].nu[0], rec[jj].nu[1], rec[jj].nu[2], rec[jj].w1};
	};
	Head nxt = head(j0);
	for (int j = j0; j < j1; ++j) {
		const Head h = nxt;
		// c_j - c_i in FP64 once (d2 equals d2_pair(c_i, c_j) bit for bit)
		const double dxd = h.cx - cix, dyd = h.cy - ciy, dzd = h.cz - ciz;
		const double d2 = __fma_rn(dzd, dzd, __fma_rn(dyd, dyd, dxd * dxd));
		const bool act = row_ok && !(d2 < h.near) && !(a.same_mesh && i == (MAP && a.cmap ? a.cmap[j] : j));
		nxt = head(j + 1 < j1 ? j + 1 : j);
		if (__ballot(act) == 0) continue;
		p.dcx = (float)dxd; p.dcy = (float)dyd; p.dcz = (float)dzd;
		p.d2f = (float)d2;
		p.nyx = h.nx; p.nyy = h.ny; p.nyz = h.nz;
		p.w1 = h.w1;
		// d and 1/d in FP64 from the hardware reciprocal square root and one Newton step
		double y = __builtin_amdgcn_rsq(d2);
		y = y * (1.5 - 0.5 * d2 * y * y);
		const double d = d2 * y;
		p.rinv_d = (float)y;
		p.df = (float)d;
		if (FAM == HELM) {
			double tt = d * a.k_rev;
			tt -= rint(tt);
			p.t0 = (float)tt;
		} else {
			p.t0 = 0.0f;
		}
		p.ad = CK ? exp2f((float)(d * (a.ki * 1.4426950408889634))) : 1.0f;
		float2 K[4] = {zero2(), zero2(), zero2(), zero2()};
		if (a.fixed_n > 0) {
			quad_orders<FAM, KM, CK>(a, j, act ? a.fixed_n : 0, p, K);
		} else {
			// Gauss order per kernel group: degree 1 + table degree; beyond the
			// last breakpoint (22.3 R) every table gives 0, order 1 (one point),
			// and a wave of such pairs skips the selects
			if (__ballot(act && d2 < h.far) == 0) {
				// no pair of the wave inside 22.3 R: the one-point rule for all
				// kernels at once (inactive lanes store nothing)
				quad_centre<FAM, KM, CK>(p, a.k, K, (float)a.ki);
			} else {
				// the distance-table thresholds of the record
				double t[NTHR];
#pragma unroll
				for (int q = 0; q < NTHR; ++q) t[q] = rec[j].thr[q];
				const int nS = (KM & SLP) && act ? gauss_order(1 + table_degree_rec<1>(d2, t)) : 0;
				const int nD = (KM & (DLP | DLPT)) && act ? gauss_order(1 + table_degree_rec<2>(d2, t)) : 0;
				const int nH = (KM & HSP) && act ? gauss_order(1 + table_degree_rec<3>(d2, t)) : 0;
				quad_groups<FAM, KM, CK>(a, j, nS, nD, nH, p, K);
			}
		}
		if (act) {
			const long long col = (long long)(MAP && a.cmap ? a.cmap[j] : j) * a.ld;
			if (MODE == OUT_SINGLE) {
				const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
				const float re = a.alpha.x * K[kk].x - a.alpha.y * K[kk].y;
				const float im = a.alpha.x * K[kk].y + a.alpha.y * K[kk].x;
				store_entry(a.out0, col + orow, a.dtype, a.accumulate, re, im);
			} else {
				const float ar = a.alpha.x * K[3].x - a.alpha.y * K[3].y + K[1].x;
				const float ai = a.alpha.x * K[3].y + a.alpha.y * K[3].x + K[1].y;
				const float br = K[0].x + a.alpha.x * K[2].x - a.alpha.y * K[2].y;
				const float bi = K[0].y + a.alpha.x * K[2].y + a.alpha.y * K[2].x;
				store_pair(a.out0, a.out1, col, orow, a.dtype, a.accumulate, ar, ai, br, bi);
			}
		}
	}
}

// ---------------------------------------------------------------- FP64 pass

// A bilinear quadrilateral in power form: x(xi, eta) = c0 + xi A + eta B +
// xi eta C, the same map as sum_a N_a X_a with N = (1 -+ xi)(1 -+ eta) / 4
// (A = (-X0 + X1 + X2 - X3) / 4, B = (-X0 - X1 + X2 + X3) / 4,
// C = (X0 - X1 + X2 - X3) / 4, c0 the mean of the corners): a point costs 9
// FMAs and the tangents dx/dxi = A + eta C, dx/deta = B + xi C 6 more.
struct Quad { D3 c0, A, B, C; };

__device__ __forceinline__ Quad quad_from_corners(const D3 (&X)[4])
{
	Quad Q;
	Q.c0 = {(X[0].x + X[1].x + X[2].x + X[3].x) * 0.25, (X[0].y + X[1].y + X[2].y + X[3].y) * 0.25,
		(X[0].z + X[1].z + X[2].z + X[3].z) * 0.25};
	Q.A = {(-X[0].x + X[1].x + X[2].x - X[3].x) * 0.25, (-X[0].y + X[1].y + X[2].y - X[3].y) * 0.25,
		(-X[0].z + X[1].z + X[2].z - X[3].z) * 0.25};
	Q.B = {(-X[0].x - X[1].x + X[2].x + X[3].x) * 0.25, (-X[0].y - X[1].y + X[2].y + X[3].y) * 0.25,
		(-X[0].z - X[1].z + X[2].z + X[3].z) * 0.25};
	Q.C = {(X[0].x - X[1].x + X[2].x - X[3].x) * 0.25, (X[0].y - X[1].y + X[2].y - X[3].y) * 0.25,
		(X[0].z - X[1].z + X[2].z - X[3].z) * 0.25};
	return Q;
}

__device__ __forceinline__ D3 quad_x(const Quad &Q, double xi, double eta)
{
	const double xe = xi * eta;
	return {__fma_rn(xe, Q.C.x, __fma_rn(eta, Q.B.x, __fma_rn(xi, Q.A.x, Q.c0.x))),
		__fma_rn(xe, Q.C.y, __fma_rn(eta, Q.B.y, __fma_rn(xi, Q.A.y, Q.c0.y))),
		__fma_rn(xe, Q.C.z, __fma_rn(eta, Q.B.z, __fma_rn(xi, Q.A.z, Q.c0.z)))};
}

__device__ __forceinline__ void quad_dx(const Quad &Q, double xi, double eta, D3 &dxi, D3 &deta)
{
	dxi = {__fma_rn(eta, Q.C.x, Q.A.x), __fma_rn(eta, Q.C.y, Q.A.y), __fma_rn(eta, Q.C.z, Q.A.z)};
	deta = {__fma_rn(xi, Q.C.x, Q.B.x), __fma_rn(xi, Q.C.y, Q.B.y), __fma_rn(xi, Q.C.z, Q.B.z)};
}

// d2x / dxi deta (the pure second derivatives of the bilinear map vanish)
__device__ __forceinline__ D3 quad_dxe(const Quad &Q) { return Q.C; }

__device__ __forceinline__ D3 d3add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }

// Kernels of MASK (FAM = Helmholtz or Laplace) times wj at one point, FP64,
// added to s: SLP s[0], DLP s[1], DLPt s[2], HSP s[3] (1 / (4 pi) included).
// CK: Helmholtz with the complex wavenumber k + i ki.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void kernel_point(D3 x, D3 y, D3 nx, D3 ny, double k, double wj, double2 (&s)[4],
	double ki = 0.0)
{
	const D3 rv = d3sub(y, x);
	const double r2 = d3dot(rv, rv);
	const double rinv = rsqrt64(r2);
	if (FAM == HELM && CK) {
		double2 e, z;
		wave64(k, ki, r2 * rinv, &e, &z);
		double rdny = 0.0, rdnx = 0.0, nn = 0.0;
		if (MASK & (DLP | DLPT | HSP)) {
			rdny = d3dot(rv, ny) * rinv;
			rdnx = -d3dot(rv, nx) * rinv;
			if (MASK & HSP) nn = d3dot(nx, ny);
		}
		wave_terms_c<MASK>(e, z, rinv * wj * (1.0 / (4.0 * PI_D)), rinv, rdny, rdnx, nn, 0.0, s);
		return;
	}
	double cs = 1.0, sn = 0.0, kr = 0.0;
	if (FAM == HELM) {
		kr = k * (r2 * rinv);
		sincos64(kr, &sn, &cs);
	}
	const double g0 = rinv * wj * (1.0 / (4.0 * PI_D));
	if (MASK & SLP) { s[0].x += cs * g0; s[0].y += -sn * g0; }
	if (MASK & (DLP | DLPT | HSP)) {
		const double rdny = d3dot(rv, ny) * rinv;
		const double rdnx = -d3dot(rv, nx) * rinv;
		const double g1 = g0 * rinv;
		// e^{-ikr} (-1 - ikr)
		const double er = -cs - kr * sn, ei = sn - kr * cs;
		if (MASK & DLP) { s[1].x += er * g1 * rdny; s[1].y += ei * g1 * rdny; }
		if (MASK & DLPT) { s[2].x += er * g1 * rdnx; s[2].y += ei * g1 * rdnx; }
		if (MASK & HSP) {
			const double g3 = g1 * rinv;
			const double cc = rdnx * rdny, nn = d3dot(nx, ny);
			const double ar = (3.0 - kr * kr) * cc + nn, ai = kr * (3.0 * cc + nn);
			s[3].x += (cs * ar + sn * ai) * g3;
			s[3].y += (cs * ai - sn * ar) * g3;
		}
	}
}

// kernels at the point (xi, eta) of the element, times J(xi) w
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void quad_kernel_at(const Quad &Q, double xi, double eta, D3 x0, D3 nx0, double k,
	double w, double2 (&s)[4], double ki = 0.0)
{
	D3 a, b;
	quad_dx(Q, xi, eta, a, b);
	const D3 n = d3cross(a, b);
	const double n2 = d3dot(n, n);
	const double Jinv = rsqrt64(n2);
	kernel_point<FAM, MASK, CK>(x0, quad_x(Q, xi, eta), nx0, d3scale(n, Jinv), k, n2 * Jinv * w, s, ki);
}

// Singular SLP / DLP / DLPt: the order-7 Duffy rule around the centre
template <int FAM, int MASK, bool CK = false>
__device__ void quad_blind(const Quad &Q, D3 x0, D3 nx0, double k, const Rules &R, double2 (&s)[4], double ki = 0.0)
{
	for (int q = 0; q < 64; ++q)
		quad_kernel_at<FAM, MASK, CK>(Q, R.duffy[q][0], R.duffy[q][1], x0, nx0, k, R.duffy[q][2], s, ki);
}

// Guiggiani's finite-part integral of the HSP kernel at the centre, radial
// order 5 (3 points), angular order 9 (5 points); x0 = x(0, 0), n_x the unit
// normal there.  A complex wavenumber (CK) leaves the Laurent terms as they
// are: it enters the HSP kernel first at order 1 / r, bounded under rho.
template <int FAM, bool CK = false>
__device__ double2 quad_guiggiani(const Quad &Q, double k, const Rules &R, double ki = 0.0)
{
	D3 dxi, deta;
	quad_dx(Q, 0.0, 0.0, dxi, deta);
	const D3 dxe = quad_dxe(Q);
	const D3 nJ = d3cross(dxi, deta);
	const D3 n0 = d3scale(nJ, 1.0 / d3norm(nJ));
	const D3 x0 = quad_x(Q, 0.0, 0.0);
	// T maps xi to a plane in which the first-order distance is unit length
	const double la = d3norm(dxi), lb = d3norm(deta);
	const double cosg = d3dot(dxi, deta) / la / lb;
	const double u = lb / la;
	const double T00 = la, T01 = la * cosg * u, T11 = la * sqrt(1.0 - cosg * cosg) * u;
	const double detT = T00 * T11;
	const double Ti00 = T11 / detT, Ti01 = -T01 / detT, Ti11 = T00 / detT;
	const D3 J0 = d3scale(nJ, 1.0 / detT);
	const double cx[4] = {-1.0, 1.0, 1.0, -1.0}, cy[4] = {-1.0, -1.0, 1.0, 1.0};
	double th_lim[4], th0[4], dist[4];
	for (int n = 0; n < 4; ++n) {
		const int m = n == 3 ? 0 : n + 1;
		const double c1x = T00 * cx[n] + T01 * cy[n], c1y = T11 * cy[n];
		const double c2x = T00 * cx[m] + T01 * cy[m], c2y = T11 * cy[m];
		double lx = 1.0, ly = 0.0;
		const double ln = sqrt((c2x - c1x) * (c2x - c1x) + (c2y - c1y) * (c2y - c1y));
		if (ln > 1e-3) { lx = (c2x - c1x) / ln; ly = (c2y - c1y) / ln; }
		const double dl = c1x * lx + c1y * ly;
		const double d0x = c1x - lx * dl, d0y = c1y - ly * dl;
		th_lim[n] = atan2(c1y, c1x);
		th0[n] = atan2(d0y, d0x);
		dist[n] = sqrt(d0x * d0x + d0y * d0y);
	}
	double2 I = make_double2(0.0, 0.0);
	const double f4 = 1.0 / (4.0 * PI_D);
	for (int n = 0; n < 4; ++n) {
		const double t1 = th_lim[n];
		double t2 = th_lim[n == 3 ? 0 : n + 1];
		if (fabs(t2 - t1) < 1e-3) continue;
		if (fabs(t2 - t1) > PI_D) t2 += 2.0 * PI_D;
		for (int it = 0; it < 5; ++it) {
			const double th = ((1.0 - R.gl5x[it]) * t1 + (1.0 + R.gl5x[it]) * t2) / 2.0;
			const double wt = R.gl5w[it] * (t2 - t1) / 2.0;
			double st, ct;
			sincos64(th, &st, &ct);
			const double vx = Ti00 * ct + Ti01 * st, vy = Ti11 * st;
			const D3 r1 = d3add(d3scale(dxi, vx), d3scale(deta, vy));
			const D3 r2 = d3scale(dxe, vx * vy);
			const D3 J1 = d3scale(d3add(d3scale(d3cross(dxi, dxe), vx), d3scale(d3cross(dxe, deta), vy)), 1.0 / detT);
			// Laurent coefficients (constant field)
			const D3 g1 = d3scale(r1, d3dot(r2, J0) + d3dot(r1, J1));
			const double a0 = -d3dot(J0, n0);
			const double a1 = d3dot(d3sub(d3scale(g1, 3.0), J1), n0);
			const double Sm2 = -3.0 * d3dot(r1, r2);
			const double Fm1 = -(Sm2 * a0 + a1) * f4;
			const double Fm2 = -a0 * f4;
			const double rho_lim = dist[n] / cos(th - th0[n]);
			I.x += (Fm1 * log(rho_lim) - Fm2 / rho_lim) * wt;
			for (int ir = 0; ir < 3; ++ir) {
				const double rho = (1.0 + R.gl3x[ir]) * rho_lim / 2.0;
				const double wr = R.gl3w[ir] * rho_lim / 2.0;
				const double ex = rho * ct, ey = rho * st;
				const double xi = Ti00 * ex + Ti01 * ey, eta = Ti11 * ey;
				double2 F[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
				D3 a, b;
				quad_dx(Q, xi, eta, a, b);
				const D3 nv = d3cross(a, b);
				const double J = d3norm(nv);
				kernel_point<FAM, HSP, CK>(x0, quad_x(Q, xi, eta), n0, d3scale(nv, 1.0 / J), k, J / detT * rho, F, ki);
				const double sing = (Fm1 + Fm2 / rho) / rho;
				I.x += (F[3].x - sing) * (wt * wr);
				I.y += F[3].y * (wt * wr);
			}
		}
	}
	return I;
}

// Image (xi, eta) and normal offset zeta of x0: x(xi) + zeta n / sqrt|n| = x0
// by Newton-Raphson from 0 (at most max_iter steps -- NiHu's 100 -- with
// |residual| < tol).
__device__ bool quad_inverse(const Quad &Q, D3 x0, double tol, int max_iter, double &xi, double &eta, double &zeta)
{
	double r0 = 0.0, r1 = 0.0, r2 = 0.0;
	const D3 dxe = quad_dxe(Q);
	for (int it = 0; it < max_iter; ++it) {
		D3 a, b;
		quad_dx(Q, r0, r1, a, b);
		D3 n = d3cross(a, b);
		const double sj = sqrt(d3norm(n));
		n = d3scale(n, 1.0 / sj);
		const D3 f = d3sub(d3add(quad_x(Q, r0, r1), d3scale(n, r2)), x0);
		if (d3norm(f) < tol) { xi = r0; eta = r1; zeta = r2; return true; }
		const D3 c0 = d3add(a, d3scale(d3cross(a, dxe), r2 / sj));
		const D3 c1 = d3add(b, d3scale(d3cross(dxe, b), r2 / sj));
		// solve [c0 c1 n] d = f by Gaussian elimination with partial pivoting
		double M[3][4] = {{c0.x, c1.x, n.x, f.x}, {c0.y, c1.y, n.y, f.y}, {c0.z, c1.z, n.z, f.z}};
		for (int col = 0; col < 3; ++col) {
			int piv = col;
			for (int r = col + 1; r < 3; ++r)
				if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
			if (piv != col)
				for (int c = 0; c < 4; ++c) { const double tmp = M[piv][c]; M[piv][c] = M[col][c]; M[col][c] = tmp; }
			for (int r = col + 1; r < 3; ++r) {
				const double fct = M[r][col] / M[col][col];
				for (int c = col; c < 4; ++c) M[r][c] -= fct * M[col][c];
			}
		}
		const double d2 = M[2][3] / M[2][2];
		const double d1 = (M[1][3] - M[1][2] * d2) / M[1][1];
		const double d0 = (M[0][3] - M[0][1] * d1 - M[0][2] * d2) / M[0][0];
		r0 -= d0; r1 -= d1; r2 -= d2;
	}
	return false;
}

// Near-singular, polar: around the image xi0 of x0, K J - K_lin J_lin over
// the reference square (radial and angular Gauss orders NR, NT given as rule
// arrays), plus the planar integral over the linearised element.
// lane / nl: this lane takes every nl-th angular line, planar-part point and
// dynamic point of the pair (the caller sums the lanes).
template <int FAM, int MASK, bool CK = false>
__device__ bool quad_near_polar(const Quad &Q, D3 x0, D3 nx0, double k, int nr, const double *grx, const double *grw,
	int nt, const double *gtx, const double *gtw, const Rules &R, double2 (&s)[4], int lane, int nl, int max_iter,
	double ki = 0.0)
{
	double xi0, eta0, zeta;
	if (!quad_inverse(Q, x0, 1e-5, max_iter, xi0, eta0, zeta)) return false;
	// the tangent parallelogram at xi0: x_lin(xi) = xl0 + a (xi - xi0) + b (eta - eta0)
	D3 a, b;
	quad_dx(Q, xi0, eta0, a, b);
	const D3 xl0 = quad_x(Q, xi0, eta0);
	const D3 nlin = d3cross(a, b);
	const double Jl = d3norm(nlin);
	const D3 nul = d3scale(nlin, 1.0 / Jl);
	const double cx[4] = {-1.0, 1.0, 1.0, -1.0}, cy[4] = {-1.0, -1.0, 1.0, 1.0};
	D3 L[4];
	for (int c = 0; c < 4; ++c) L[c] = d3add(xl0, d3add(d3scale(a, cx[c] - xi0), d3scale(b, cy[c] - eta0)));
	double th_lim[4], th0[4], dist[4];
	for (int n = 0; n < 4; ++n) {
		const int m = n == 3 ? 0 : n + 1;
		double lx = 1.0, ly = 0.0;
		const double ln = sqrt((cx[m] - cx[n]) * (cx[m] - cx[n]) + (cy[m] - cy[n]) * (cy[m] - cy[n]));
		if (ln > 1e-3) { lx = (cx[m] - cx[n]) / ln; ly = (cy[m] - cy[n]) / ln; }
		const double d1x = cx[n] - xi0, d1y = cy[n] - eta0;
		const double dl = d1x * lx + d1y * ly;
		const double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		th_lim[n] = atan2(d1y, d1x);
		th0[n] = atan2(d0y, d0x);
		dist[n] = sqrt(d0x * d0x + d0y * d0y);
	}
	for (int n = 0; n < 4; ++n) {
		double t1 = th_lim[n], t2 = th_lim[n == 3 ? 0 : n + 1];
		if (fabs(t2 - t1) > PI_D) {
			if (t2 > t1) t1 += 2.0 * PI_D;
			else t2 += 2.0 * PI_D;
		}
		if (fabs(t2 - t1) < 1e-3) continue;
		for (int it = (lane + nt * nl - (nt * n) % nl) % nl; it < nt; it += nl) {
			const double th = ((1.0 - gtx[it]) * t1 + (1.0 + gtx[it]) * t2) / 2.0;
			const double wt = gtw[it] * (t2 - t1) / 2.0;
			double st, ct;
			sincos64(th, &st, &ct);
			const double rho_lim = dist[n] / cos(th - th0[n]);
			for (int ir = 0; ir < nr; ++ir) {
				const double rho = (1.0 + grx[ir]) * rho_lim / 2.0;
				const double w = wt * (grw[ir] * rho_lim / 2.0) * rho;
				const double dxi = rho * ct, deta = rho * st;
				quad_kernel_at<FAM, MASK, CK>(Q, dxi + xi0, deta + eta0, x0, nx0, k, w, s, ki);
				const D3 yl = d3add(xl0, d3add(d3scale(a, dxi), d3scale(b, deta)));
				kernel_point<FAM, MASK, CK>(x0, yl, nx0, nul, k, -Jl * w, s, ki);
			}
		}
	}
	// planar integral over the linearised element: Laplace static part and,
	// for Helmholtz, the 8 x 8-point Gauss rule of K_helm - K_lap
	double st[4] = {0, 0, 0, 0};
	near_static<MASK, 4>(L, x0, nx0, R, st, lane, nl);
	for (int t = 0; t < 4; ++t) s[t].x += st[t];
	if (FAM == HELM)
		for (int q = lane; q < 64; q += nl) {
			const int i = q >> 3, j = q & 7;
			const D3 yl = d3add(xl0, d3add(d3scale(a, R.gl8x[i] - xi0), d3scale(b, R.gl8x[j] - eta0)));
			dynamic_point<MASK, CK>(x0, yl, nx0, nul, k, R.gl8w[i] * R.gl8w[j] * Jl, s, ki);
		}
	return true;
}

// Telles' cubic map of one axis of the order-30 rule towards e (distance
// parameter rb): x -> ca x^3 + cb x^2 + cc x + cd, jacobian 3 ca x^2 + 2 cb x + cc
__device__ __forceinline__ void telles_coeffs(double e, double rb, double &ca, double &cb, double &cc, double &cd)
{
	const double f = 1.0 + 2.0 * rb;
	const double q = 1.0 / (2.0 * f) * ((e * (3.0 - 2.0 * rb) - 2.0 * e * e * e / f) * 1.0 / f - e);
	const double p = 1.0 / (3.0 * f * f) * (4.0 * rb * (1.0 - rb) + 3.0 * (1.0 - e * e));
	const double sq = sqrt(q * q + p * p * p);
	const double gb = cbrt(-q + sq) + cbrt(-q - sq) + e / f;
	const double Qd = 1.0 + 3.0 * gb * gb;
	ca = (1.0 - rb) / Qd;
	cb = -3.0 * (1.0 - rb) * gb / Qd;
	cc = (3.0 * gb * gb + rb) / Qd;
	cd = -cb;
}

template <int FAM, int MASK, bool CK = false>
__device__ bool quad_near_telles(const Quad &Q, D3 x0, D3 nx0, double k, const Rules &R, double2 (&s)[4], int lane,
	int nl, int max_iter, double ki = 0.0)
{
	double xi0, eta0, zeta;
	if (!quad_inverse(Q, x0, 1e-6, max_iter, xi0, eta0, zeta)) return false;
	double rb;
	if (zeta < 0.05) rb = 0.0;
	else if (zeta <= 1.3) rb = 0.85 + 0.24 * log(zeta);
	else if (zeta <= 3.618) rb = 0.893 + 0.0832 * log(zeta);
	else rb = 1.0;
	double a0, b0, c0, e0, a1, b1, c1, e1;
	telles_coeffs(xi0, rb, a0, b0, c0, e0);
	telles_coeffs(eta0, rb, a1, b1, c1, e1);
	for (int i = lane; i < 16; i += nl) {
		const double x = R.gl16x[i];
		const double xi = a0 * x * x * x + b0 * x * x + c0 * x + e0;
		const double jx = 3.0 * a0 * x * x + 2.0 * b0 * x + c0;
		for (int j = 0; j < 16; ++j) {
			const double y = R.gl16x[j];
			const double eta = a1 * y * y * y + b1 * y * y + c1 * y + e1;
			const double jy = 3.0 * a1 * y * y + 2.0 * b1 * y + c1;
			quad_kernel_at<FAM, MASK, CK>(Q, xi, eta, x0, nx0, k, R.gl16w[i] * R.gl16w[j] * jx * jy, s, ki);
		}
	}
	return true;
}

// ---------------------------------------------------------------- FP64 regular pass

// The FP64 precision policy: every regular pair in FP64 (NiHu's double
// kernels), one group of kernels at a time at its own Gauss order, the
// points from a per-element FP64 table (y, w J, unit normal).
struct Quad64Args {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	const int *rmap, *cmap;      // result row / column of each test / trial element (null: its own index)
	int tn, sn;
	const double *tc, *tnu;      // test collocation points, unit normals [3][tn]
	const double *sc;            // trial centres [3][sn]
	const double *thr;           // [NTHR][sn]
	const double *qt[QN_MAX + 1];   // per Gauss order N: [sn][N * N][8]
	int fixed_n;                 // > 0: fixed rule with fixed_n x fixed_n points
	double k;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	double2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

template <int FAM, int MASK, bool CK>
__device__ __forceinline__ void quad_eval64(const Quad64Args &a, int j, int n, D3 x0, D3 nx, double2 (&s)[4])
{
	if (n <= 0) return;
	const int nq = n * n;
	const double *t = a.qt[n] + (size_t)j * nq * 8;
	for (int q = 0; q < nq; ++q) {
		const D3 y = {t[8 * q], t[8 * q + 1], t[8 * q + 2]};
		const D3 ny = {t[8 * q + 4], t[8 * q + 5], t[8 * q + 6]};
		kernel_point<FAM, MASK, CK>(x0, y, nx, ny, a.k, t[8 * q + 3], s, a.ki);
	}
}

template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(256) quad_regular64_kernel(Quad64Args a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int tile_c = blockIdx.x / a.row_tiles;
	const int i = tile_r * 256 + threadIdx.x;
	if (i >= a.n_rows) return;
	const int orow = a.rmap ? a.rmap[i] : i;
	const D3 x0 = {a.tc[i], a.tc[a.tn + i], a.tc[2 * a.tn + i]};
	const D3 nx = {a.tnu[i], a.tnu[a.tn + i], a.tnu[2 * a.tn + i]};
	const int j0 = tile_c * a.tile_cols;
	const int j1 = min(j0 + a.tile_cols, a.n_cols);
	for (int j = j0; j < j1; ++j) {
		// classification exactly as quad_regular_kernel: d2 from c_j - c_i in FP64
		const double d2 = d2_pair(a.sc[j], a.sc[a.sn + j], a.sc[2 * a.sn + j], x0.x, x0.y, x0.z);
		const int cj = a.cmap ? a.cmap[j] : j;
		if (d2 < a.thr[THR_NEAR * a.sn + j] || (a.same_mesh && i == cj)) continue;
		const int fn = a.fixed_n;
		double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
		if (KM & SLP)
			quad_eval64<FAM, SLP, CK>(a, j, fn > 0 ? fn : gauss_order(1 + table_degree<1>(d2, a.thr, a.sn, j)), x0, nx,
				s);
		if (KM & (DLP | DLPT))
			quad_eval64<FAM, KM & (DLP | DLPT), CK>(a, j,
				fn > 0 ? fn : gauss_order(1 + table_degree<2>(d2, a.thr, a.sn, j)), x0, nx, s);
		if (KM & HSP)
			quad_eval64<FAM, HSP, CK>(a, j, fn > 0 ? fn : gauss_order(1 + table_degree<3>(d2, a.thr, a.sn, j)), x0, nx,
				s);
		const long long idx = (long long)cj * a.ld + orow;
		if (MODE == OUT_SINGLE) {
			const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
			const double re = a.alpha.x * s[kk].x - a.alpha.y * s[kk].y;
			const double im = a.alpha.x * s[kk].y + a.alpha.y * s[kk].x;
			store_entry64(a.out0, idx, a.dtype, a.accumulate, re, im);
		} else {
			const double ar = a.alpha.x * s[3].x - a.alpha.y * s[3].y + s[1].x;
			const double ai = a.alpha.x * s[3].y + a.alpha.y * s[3].x + s[1].y;
			const double br = s[0].x + a.alpha.x * s[2].x - a.alpha.y * s[2].y;
			const double bi = s[0].y + a.alpha.x * s[2].y + a.alpha.y * s[2].x;
			store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
			store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
		}
	}
}

// store one FP64 pair result (single operator or the Burton-Miller pair)
template <int KM, int MODE, bool DIAG>
__device__ __forceinline__ void quad_store64(const NearArgs &a, int i, int j, const double2 (&s)[4])
{
	const long long idx = (long long)(a.cmap ? a.cmap[j] : j) * a.ld + (a.rmap ? a.rmap[i] : i);
	if (MODE == OUT_SINGLE) {
		const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
		const double re = a.alpha.x * s[kk].x - a.alpha.y * s[kk].y;
		const double im = a.alpha.x * s[kk].y + a.alpha.y * s[kk].x;
		store_entry64(a.out0, idx, a.dtype, a.accumulate, re, im);
	} else {
		double ar = a.alpha.x * s[3].x - a.alpha.y * s[3].y + s[1].x;
		double ai = a.alpha.x * s[3].y + a.alpha.y * s[3].x + s[1].y;
		double br = s[0].x + a.alpha.x * s[2].x - a.alpha.y * s[2].y;
		double bi = s[0].y + a.alpha.x * s[2].y + a.alpha.y * s[2].x;
		if (DIAG) {
			ar += -0.5;
			br += a.alpha.x / 2.0;
			bi += a.alpha.y / 2.0;
		}
		store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
		store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
	}
}

__device__ __forceinline__ void quad_load(const NearArgs &a, int i, int j, Quad &Q, D3 &x0, D3 &nx)
{
	D3 X[4];
	for (int v = 0; v < 4; ++v)
		X[v] = {a.X[(3 * v) * a.sn + j], a.X[(3 * v + 1) * a.sn + j], a.X[(3 * v + 2) * a.sn + j]};
	Q = quad_from_corners(X);
	x0 = {a.tc[i], a.tc[a.tn + i], a.tc[2 * a.tn + i]};
	nx = {a.tnu[i], a.tnu[a.tn + i], a.tnu[2 * a.tn + i]};
}

// Diagonal pairs: one lane per element (Duffy rule, Guiggiani).
template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(64) quad_diag_kernel(NearArgs a)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= a.n_pairs) return;
	const Rules &R = *a.rules;
	Quad Q;
	D3 x0, nx;
	quad_load(a, p, p, Q, x0, nx);
	double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
	if (KM & (SLP | DLP | DLPT)) quad_blind<FAM, KM & (SLP | DLP | DLPT), CK>(Q, x0, nx, a.k, R, s, a.ki);
	if (KM & HSP) s[3] = quad_guiggiani<FAM, CK>(Q, a.k, R, a.ki);
	quad_store64<KM, MODE, true>(a, p, p, s);
}

// Near-singular pairs: QLPP lanes per pair share the angular lines, planar
// and Telles points (each lane repeats the pair's inverse mapping), then a
// butterfly sum over the lanes; the pair's first lane stores.
constexpr int QLPP = 4;

template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(64, 3) quad_near_kernel(NearArgs a)
{
	const int gid = blockIdx.x * blockDim.x + threadIdx.x;
	const int p = gid / QLPP, lane = gid % QLPP;
	// lanes past the last pair repeat it and do not store
	const bool valid = p < a.n_pairs;
	const int pp = valid ? p : a.n_pairs - 1;
	const int i = a.pi[pp], j = a.pj[pp];
	const Rules &R = *a.rules;
	Quad Q;
	D3 x0, nx;
	quad_load(a, i, j, Q, x0, nx);
	double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
	bool ok = true;
	if (FAM == HELM) {
		if (KM & SLP)
			ok = quad_near_polar<FAM, SLP, CK>(Q, x0, nx, a.k, 8, R.gl8x, R.gl8w, 8, R.gl8x, R.gl8w, R, s, lane, QLPP,
				a.inv_iter, a.ki);
		if (ok && (KM & (DLP | DLPT | HSP)))
			ok = quad_near_telles<FAM, KM & (DLP | DLPT | HSP), CK>(Q, x0, nx, a.k, R, s, lane, QLPP, a.inv_iter,
				a.ki);
	} else {
		if (KM & (SLP | DLP | DLPT))
			ok = quad_near_polar<FAM, KM & (SLP | DLP | DLPT)>(Q, x0, nx, 0.0, 8, R.gl8x, R.gl8w, 8, R.gl8x, R.gl8w, R, s,
				lane, QLPP, a.inv_iter);
		if (ok && (KM & HSP))
			ok = quad_near_polar<FAM, HSP>(Q, x0, nx, 0.0, 51, R.gl51x, R.gl51w, 51, R.gl51x, R.gl51w, R, s, lane, QLPP,
				a.inv_iter);
	}
	for (int off = QLPP / 2; off >= 1; off >>= 1)
		for (int t = 0; t < 4; ++t) {
			s[t].x += __shfl_xor(s[t].x, off, QLPP);
			s[t].y += __shfl_xor(s[t].y, off, QLPP);
		}
	if (!ok) {
		// the inverse mapping did not converge (NiHu throws): NaN marks the
		// entry and the pair is counted, so that the call reports an error
		const double nan = __builtin_nan("");
		for (int t = 0; t < 4; ++t) s[t] = make_double2(nan, nan);
		if (valid && lane == 0 && a.status) atomicAdd(a.status, 1);
	}
	if (valid && lane == 0) quad_store64<KM, MODE, false>(a, i, j, s);
}

} // namespace nbk
