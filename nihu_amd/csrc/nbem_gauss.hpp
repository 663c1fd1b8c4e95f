// nbem_gauss.hpp -- the quadrilateral gauss field in collocation (gfx950):
// four dofs per 4-node quadrilateral at the 2 x 2 Gauss points
// (nihu/library/quad_1_gauss_field.hpp, quad_1_gauss_shape_set.hpp; nodes
// (-,-), (+,-), (-,+), (+,+) at +-sqrt(3)/3, lib_shape.cpp:128-134), shape
// functions N_b = (1 + s_b sqrt3 xi)(1 + t_b sqrt3 eta) / 4, dof 4 e + b.
// Test side: the dirac of the field (rows 4 i + a: node a of element i) or
// the dirac of a constant field (rows = elements: field points).
//
// Each element pair is a block of (test nodes) x 4 entries.  Classification
// by element centres, as for constant fields (the block is regular, near or
// singular as a whole; helmholtz_nearly_singular_integrals.hpp:281-293):
//   regular (FP32): one Gauss square rule per kernel group of degree 2 (the
//     field's polynomial order 1 + jacobian order 1,
//     nihu/core/complexity_estimator.hpp:47-75) + the distance-table degree;
//     one lane per test node, the trial element wave-uniform as in
//     quad_regular_kernel; per trial point the kernels are evaluated once and
//     distributed over the four dofs with N_b(xi_q) (a per-order table shared
//     by all elements) -- double_integral.hpp:445-467;
//   near (FP64, d / R < 1.5 between centres): per test node the polar method
//     with F = K J N(xi) - K_lin J_lin N(xi0) and the planar integral times
//     N(xi0) (nearly_singular_collocational.hpp:86-197), or Telles with N(xi)
//     in the weights; QLPP lanes per (pair, node);
//   singular (FP64, the same element): per node xi_a the order-7 Duffy blind
//     rule around xi_a, and Guiggiani's method at xi_a for HSP with the
//     shape-function terms of the Laurent coefficients, a0 N_b(xi_a) and
//     a1 N_b(xi_a) + a0 grad N_b(xi_a).v (guiggiani_1992.hpp:180-330,
//     laplace_kernel.hpp:266-295); one lane per (element, node).

#pragma once

#include "nbem_quad.hpp"

namespace nbk {

constexpr int GDOF = 4;

// N_b at the points of the Gauss square rule of each order (rows as in the
// QuadPt tables: xi outer, eta inner)
struct GaussShapes {
	float n[QN_MAX + 1][QN_MAX * QN_MAX][GDOF];
	double n64[QN_MAX + 1][QN_MAX * QN_MAX][GDOF];   // the same in FP64 (FP64 precision policy)
};

struct GaussArgs {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	int nodes;                   // test nodes per test element: 4 (gauss field) or 1 (constant field)
	int tn, te;                  // rows (collocation points), test elements
	const double *tx;            // collocation points [3][tn]
	const float *tnu;            // unit normals there [3][tn]
	const double *tc;            // test element centres [3][te] (classification)
	const TrialRec *rec;         // trial records: thresholds, centre
	const QuadPt *qt[QN_MAX + 1];   // trial Gauss square tables by order
	const GaussShapes *gs;
	int fixed_n;                 // > 0: fixed rule with fixed_n x fixed_n points
	float k, k_revf;
	double k_rev;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	float2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

// One point's kernel values of MASK (quad_add into zero: E, -a Q, b Q,
// nn Q - a b (3 Q - k^2 E) / r^2 with E = (ec, es), Q = E (1 + i th) / r^2,
// and CK's attenuation and kappa^2).
template <int FAM, int MASK, bool CK>
__device__ __forceinline__ void gauss_vals(float ec, float es, float th, float rinv, float a, float b, float nn,
	float k, float kih, float ki, float2 (&V)[4])
{
	if (MASK & SLP) V[0] = make_float2(ec, es);
	if (MASK & (DLP | DLPT | HSP)) {
		const float ri2 = rinv * rinv;
		const float Gr = ec * ri2, Gi = es * ri2;
		float Qr = __builtin_fmaf(-th, Gi, Gr), Qi = __builtin_fmaf(th, Gr, Gi);
		if (CK) { Qr = __builtin_fmaf(-kih, Gr, Qr); Qi = __builtin_fmaf(-kih, Gi, Qi); }
		if (MASK & DLP) V[1] = make_float2(-a * Qr, -a * Qi);
		if (MASK & DLPT) V[2] = make_float2(b * Qr, b * Qi);
		if (MASK & HSP) {
			const float kk = FAM == HELM ? (CK ? __builtin_fmaf(k, k, -ki * ki) : k * k) : 0.0f;
			const float ab = a * b * ri2;
			float hr = __builtin_fmaf(3.0f, Qr, -kk * ec), hi = __builtin_fmaf(3.0f, Qi, -kk * es);
			if (CK) {
				const float kq = 2.0f * k * ki;
				hr = __builtin_fmaf(kq, es, hr);
				hi = __builtin_fmaf(-kq, ec, hi);
			}
			V[3] = make_float2(__builtin_fmaf(nn, Qr, -ab * hr), __builtin_fmaf(nn, Qi, -ab * hi));
		}
	}
}

// The kernels of MASK at the points of one trial element, each point's values
// distributed over the four dofs with the shape functions: T[b][kernel].  Point geometry and phase as in quad_points.
// O2: the order-2 rule, whose points are the field's nodes, so N_b(xi_q) is
// 1 at the node of dof b and 0 at the others (to rounding, 6e-17): point q
// goes to dof (0, 2, 1, 3)[q] (points xi outer, eta inner; dofs xi first).
template <int FAM, int MASK, bool CK, bool O2 = false>
__device__ __forceinline__ void gauss_points(const QuadPt *tab, const float *nsh, int nq, const QuadCtx &p,
	float k_revf, float k, float ki, float2 (&T)[GDOF][4])
{
	const auto qf = uni((const float *)tab);
	const auto nf = uni(nsh);
	const float kil2e = ki * 1.4426950408889634f;
	const float dcx2 = 2.0f * p.dcx, dcy2 = 2.0f * p.dcy, dcz2 = 2.0f * p.dcz;
	const int nqq = O2 ? 4 : nq;
#pragma unroll (O2 ? 4 : 2)
	for (int q = 0; q < nqq; ++q) {
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
		float2 V[4] = {zero2(), zero2(), zero2(), zero2()};
		gauss_vals<FAM, MASK, CK>(ec, es, th, rinv, a, b, nn, k, kih, ki, V);
		if (O2) {
			const int d = q == 1 ? 2 : q == 2 ? 1 : q;
#pragma unroll
			for (int t = 0; t < 4; ++t)
				if (MASK & (1 << t)) {
					T[d][t].x += V[t].x;
					if (FAM == HELM) T[d][t].y += V[t].y;
				}
		} else {
#pragma unroll
			for (int d = 0; d < GDOF; ++d) {
				const float N = nf[GDOF * q + d];
#pragma unroll
				for (int t = 0; t < 4; ++t)
					if (MASK & (1 << t)) {
						T[d][t].x = __builtin_fmaf(N, V[t].x, T[d][t].x);
						if (FAM == HELM) T[d][t].y = __builtin_fmaf(N, V[t].y, T[d][t].y);
					}
			}
		}
	}
}

// Kernel groups at their own Gauss orders (SLP nS, DLP / DLPt nD, HSP nH;
// <= 0 idle), each distinct order of the wave evaluated once, as quad_groups.
// An order above 2 that no lane of the wave takes for SLP (or for SLP and the
// double layers) is evaluated without those kernels: a lane's degrees grow
// from SLP to DLP to HSP at every distance, so the higher orders mostly serve
// the hypersingular group alone.
template <int FAM, int KM, bool CK>
__device__ __forceinline__ void gauss_groups(const GaussArgs &a, int j, int nS, int nD, int nH, const QuadCtx &p,
	float2 (&K)[GDOF][4])
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
		const bool needS = __ballot(fS) != 0, needD = __ballot(fD) != 0;
		if (fS || fD || fH) {
			float2 t[GDOF][4];
#pragma unroll
			for (int d = 0; d < GDOF; ++d) t[d][0] = t[d][1] = t[d][2] = t[d][3] = zero2();
			const int nq = nsel * nsel;
			const QuadPt *tab = a.qt[nsel] + (size_t)j * nq;
			const float *nsh = &a.gs->n[nsel][0][0];
			if (nsel == 2)
				gauss_points<FAM, KM, CK, true>(tab, nullptr, 4, p, a.k_revf, a.k, ki, t);
			else if (!needS && !needD)
				gauss_points<FAM, KM & HSP, CK>(tab, nsh, nq, p, a.k_revf, a.k, ki, t);
			else if (!needS)
				gauss_points<FAM, KM & (DLP | DLPT | HSP), CK>(tab, nsh, nq, p, a.k_revf, a.k, ki, t);
			else
				gauss_points<FAM, KM, CK>(tab, nsh, nq, p, a.k_revf, a.k, ki, t);
#pragma unroll
			for (int d = 0; d < GDOF; ++d) {
				if (fS) K[d][0] = t[d][0];
				if (fD) { K[d][1] = t[d][1]; K[d][2] = t[d][2]; }
				if (fH) K[d][3] = t[d][3];
			}
		}
		pS = fS ? NONE : pS;
		pD = fD ? NONE : pD;
		pH = fH ? NONE : pH;
	}
}

template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(256) gauss_regular_kernel(GaussArgs a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int tile_c = blockIdx.x / a.row_tiles;
	const int r = tile_r * 256 + threadIdx.x;
	const bool row_ok = r < a.n_rows;
	const int rr = row_ok ? r : a.n_rows - 1;
	const int ie = a.nodes == GDOF ? rr >> 2 : rr;
	const double xx = a.tx[rr], xy = a.tx[a.tn + rr], xz = a.tx[2 * a.tn + rr];
	const double cix = a.tc[ie], ciy = a.tc[a.te + ie], ciz = a.tc[2 * a.te + ie];
	QuadCtx p;
	p.nxx = p.nxy = p.nxz = 0.0f;
	if (KM & (DLPT | HSP)) { p.nxx = a.tnu[rr]; p.nxy = a.tnu[a.tn + rr]; p.nxz = a.tnu[2 * a.tn + rr]; }
	row_ready(cix, ciy, ciz, p.nxx, p.nxy, p.nxz, ie);
	row_ready(xx, xy, xz, 0.0f, 0.0f, 0.0f, rr);
	const int j0 = tile_c * a.tile_cols;
	const int j1 = min(j0 + a.tile_cols, a.n_cols);
	const auto rec = uni(a.rec);
	const float ki = (float)a.ki;
	for (int j = j0; j < j1; ++j) {
		const double hx = rec[j].c[0], hy = rec[j].c[1], hz = rec[j].c[2];
		// classification by element centres (d2 as the constant-field pass forms it)
		const double d2 = d2_pair(hx, hy, hz, cix, ciy, ciz);
		const bool act = row_ok && !(d2 < rec[j].thr[THR_NEAR]) && !(a.same_mesh && ie == j);
		if (__ballot(act) == 0) continue;
		// the kernels' geometry from the collocation point: c_j - x in FP64
		const double dxd = hx - xx, dyd = hy - xy, dzd = hz - xz;
		const double e2 = __fma_rn(dzd, dzd, __fma_rn(dyd, dyd, dxd * dxd));
		p.dcx = (float)dxd; p.dcy = (float)dyd; p.dcz = (float)dzd;
		p.d2f = (float)e2;
		double y = __builtin_amdgcn_rsq(e2);
		y = y * (1.5 - 0.5 * e2 * y * y);
		const double d = e2 * y;
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
		float2 K[GDOF][4];
#pragma unroll
		for (int dd = 0; dd < GDOF; ++dd) K[dd][0] = K[dd][1] = K[dd][2] = K[dd][3] = zero2();
		if (a.fixed_n > 0) {
			const int nq = a.fixed_n * a.fixed_n;
			if (act) {
				if (a.fixed_n == 2)
					gauss_points<FAM, KM, CK, true>(a.qt[2] + (size_t)j * 4, nullptr, 4, p, a.k_revf, a.k, ki, K);
				else
					gauss_points<FAM, KM, CK>(a.qt[a.fixed_n] + (size_t)j * nq, &a.gs->n[a.fixed_n][0][0], nq, p,
						a.k_revf, a.k, ki, K);
			}
		} else if (__ballot(act && d2 < rec[j].thr[THR_22_3]) == 0) {
			// beyond the last breakpoint every table degree is 0: order 2 for all kernels
			if (act) gauss_points<FAM, KM, CK, true>(a.qt[2] + (size_t)j * 4, nullptr, 4, p, a.k_revf, a.k, ki, K);
		} else {
			double t[NTHR];
#pragma unroll
			for (int q = 0; q < NTHR; ++q) t[q] = rec[j].thr[q];
			const int nS = (KM & SLP) && act ? gauss_order(2 + table_degree_rec<1>(d2, t)) : 0;
			const int nD = (KM & (DLP | DLPT)) && act ? gauss_order(2 + table_degree_rec<2>(d2, t)) : 0;
			const int nH = (KM & HSP) && act ? gauss_order(2 + table_degree_rec<3>(d2, t)) : 0;
			gauss_groups<FAM, KM, CK>(a, j, nS, nD, nH, p, K);
		}
		if (act) {
#pragma unroll
			for (int dd = 0; dd < GDOF; ++dd) {
				const long long col = (long long)(GDOF * j + dd) * a.ld;
				if (MODE == OUT_SINGLE) {
					const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
					const float re = a.alpha.x * K[dd][kk].x - a.alpha.y * K[dd][kk].y;
					const float im = a.alpha.x * K[dd][kk].y + a.alpha.y * K[dd][kk].x;
					store_entry(a.out0, col + r, a.dtype, a.accumulate, re, im);
				} else {
					const float ar = a.alpha.x * K[dd][3].x - a.alpha.y * K[dd][3].y + K[dd][1].x;
					const float ai = a.alpha.x * K[dd][3].y + a.alpha.y * K[dd][3].x + K[dd][1].y;
					const float br = K[dd][0].x + a.alpha.x * K[dd][2].x - a.alpha.y * K[dd][2].y;
					const float bi = K[dd][0].y + a.alpha.x * K[dd][2].y + a.alpha.y * K[dd][2].x;
					store_pair(a.out0, a.out1, col, r, a.dtype, a.accumulate, ar, ai, br, bi);
				}
			}
		}
	}
}

// ---------------------------------------------------------------- FP64 pass

struct GaussNearArgs {
	int n_pairs;                 // near: element pairs; diagonal: elements
	int nodes;                   // test nodes per test element (4 or 1)
	const int *pi, *pj;          // near pair list
	int tn;                      // rows
	const double *tx, *tnu;      // collocation points and unit normals [3][tn] (FP64)
	int sn;
	const double *X;             // trial corners [12][sn]
	double k, ki;
	double2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
	const Rules *rules;
	int *status;                 // pairs without an inverse image (null: none kept)
	int inv_iter;
};

// the shape functions at (xi, eta), and the nodes
__device__ __forceinline__ void gshape64(double xi, double eta, double (&N)[GDOF])
{
	const double s3 = 1.7320508075688772935274463415059;
	const double ax = s3 * xi, ay = s3 * eta;
	N[0] = (1.0 - ax) * (1.0 - ay) / 4.0;
	N[1] = (1.0 + ax) * (1.0 - ay) / 4.0;
	N[2] = (1.0 - ax) * (1.0 + ay) / 4.0;
	N[3] = (1.0 + ax) * (1.0 + ay) / 4.0;
}

__device__ __forceinline__ void gdshape64(double xi, double eta, double (&dx)[GDOF], double (&dy)[GDOF])
{
	const double s3 = 1.7320508075688772935274463415059;
	const double ax = s3 * xi, ay = s3 * eta;
	dx[0] = -s3 * (1.0 - ay) / 4.0; dy[0] = -(1.0 - ax) * s3 / 4.0;
	dx[1] = s3 * (1.0 - ay) / 4.0;  dy[1] = -(1.0 + ax) * s3 / 4.0;
	dx[2] = -s3 * (1.0 + ay) / 4.0; dy[2] = (1.0 - ax) * s3 / 4.0;
	dx[3] = s3 * (1.0 + ay) / 4.0;  dy[3] = (1.0 + ax) * s3 / 4.0;
}

__device__ __forceinline__ void gnode64(int a, double &xi, double &eta)
{
	const double g = 1.7320508075688772935274463415059 / 3.0;
	xi = (a & 1) ? g : -g;
	eta = (a & 2) ? g : -g;
}

// s[b][t] += N[b] v[t] for the kernels of MASK
template <int MASK>
__device__ __forceinline__ void gspread(const double (&N)[GDOF], const double2 (&v)[4], double2 (&s)[GDOF][4])
{
#pragma unroll
	for (int d = 0; d < GDOF; ++d)
#pragma unroll
		for (int t = 0; t < 4; ++t)
			if (MASK & (1 << t)) { s[d][t].x += N[d] * v[t].x; s[d][t].y += N[d] * v[t].y; }
}

// polar near-singular method with the shape functions (quad_near_polar)
template <int FAM, int MASK, bool CK>
__device__ bool gauss_near_polar(const Quad &Q, D3 x0, D3 nx0, double k, int nr, const double *grx, const double *grw,
	int nt, const double *gtx, const double *gtw, const Rules &R, double2 (&s)[GDOF][4], int lane, int nl,
	int max_iter, double ki)
{
	double xi0, eta0, zeta;
	if (!quad_inverse(Q, x0, 1e-5, max_iter, xi0, eta0, zeta)) return false;
	D3 a, b;
	quad_dx(Q, xi0, eta0, a, b);
	const D3 xl0 = quad_x(Q, xi0, eta0);
	const D3 nlin = d3cross(a, b);
	const double Jl = d3norm(nlin);
	const D3 nul = d3scale(nlin, 1.0 / Jl);
	double N0[GDOF];
	gshape64(xi0, eta0, N0);
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
				double2 f[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
				double2 fl[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
				quad_kernel_at<FAM, MASK, CK>(Q, dxi + xi0, deta + eta0, x0, nx0, k, w, f, ki);
				const D3 yl = d3add(xl0, d3add(d3scale(a, dxi), d3scale(b, deta)));
				kernel_point<FAM, MASK, CK>(x0, yl, nx0, nul, k, -Jl * w, fl, ki);
				double N[GDOF];
				gshape64(dxi + xi0, deta + eta0, N);
				gspread<MASK>(N, f, s);
				gspread<MASK>(N0, fl, s);
			}
		}
	}
	// planar integral over the linearised element, times N(xi0)
	double st[4] = {0, 0, 0, 0};
	near_static<MASK, 4>(L, x0, nx0, R, st, lane, nl);
	double2 pl[4];
	for (int t = 0; t < 4; ++t) pl[t] = make_double2(st[t], 0.0);
	if (FAM == HELM)
		for (int q = lane; q < 64; q += nl) {
			const int i = q >> 3, j = q & 7;
			const D3 yl = d3add(xl0, d3add(d3scale(a, R.gl8x[i] - xi0), d3scale(b, R.gl8x[j] - eta0)));
			dynamic_point<MASK, CK>(x0, yl, nx0, nul, k, R.gl8w[i] * R.gl8w[j] * Jl, pl, ki);
		}
	gspread<MASK>(N0, pl, s);
	return true;
}

// Telles-transformed order-30 rule with the shape functions (quad_near_telles)
template <int FAM, int MASK, bool CK>
__device__ bool gauss_near_telles(const Quad &Q, D3 x0, D3 nx0, double k, const Rules &R, double2 (&s)[GDOF][4],
	int lane, int nl, int max_iter, double ki)
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
			double2 f[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
			quad_kernel_at<FAM, MASK, CK>(Q, xi, eta, x0, nx0, k, R.gl16w[i] * R.gl16w[j] * jx * jy, f, ki);
			double N[GDOF];
			gshape64(xi, eta, N);
			gspread<MASK>(N, f, s);
		}
	}
	return true;
}

// Guiggiani's finite part at node xi0 with the shape functions
// (quad_guiggiani with the centre replaced by xi0: the sides seen from
// e0 = T xi0; per dof the Laurent coefficients a0 N_b(xi0) and
// a1 N_b(xi0) + a0 grad N_b(xi0).v, v the direction in the reference square).
template <int FAM, bool CK>
__device__ void gauss_guiggiani(const Quad &Q, double xi0x, double xi0y, double k, const Rules &R, double ki,
	double2 (&s)[GDOF][4])
{
	D3 dxi, deta;
	quad_dx(Q, xi0x, xi0y, dxi, deta);
	const D3 dxe = quad_dxe(Q);
	const D3 nJ = d3cross(dxi, deta);
	const D3 n0 = d3scale(nJ, 1.0 / d3norm(nJ));
	const D3 x0 = quad_x(Q, xi0x, xi0y);
	double N0[GDOF], dNx[GDOF], dNy[GDOF];
	gshape64(xi0x, xi0y, N0);
	gdshape64(xi0x, xi0y, dNx, dNy);
	const double la = d3norm(dxi), lb = d3norm(deta);
	const double cosg = d3dot(dxi, deta) / la / lb;
	const double u = lb / la;
	const double T00 = la, T01 = la * cosg * u, T11 = la * sqrt(1.0 - cosg * cosg) * u;
	const double detT = T00 * T11;
	const double Ti00 = T11 / detT, Ti01 = -T01 / detT, Ti11 = T00 / detT;
	const double e0x = T00 * xi0x + T01 * xi0y, e0y = T11 * xi0y;
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
		const double d1x = c1x - e0x, d1y = c1y - e0y;
		const double dl = d1x * lx + d1y * ly;
		const double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		th_lim[n] = atan2(d1y, d1x);
		th0[n] = atan2(d0y, d0x);
		dist[n] = sqrt(d0x * d0x + d0y * d0y);
	}
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
			const D3 g1 = d3scale(r1, d3dot(r2, J0) + d3dot(r1, J1));
			const double a0 = -d3dot(J0, n0);
			const double a1 = d3dot(d3sub(d3scale(g1, 3.0), J1), n0);
			const double Sm2 = -3.0 * d3dot(r1, r2);
			const double rho_lim = dist[n] / cos(th - th0[n]);
			const double lr = log(rho_lim);
			double Fm1[GDOF], Fm2[GDOF];
			for (int d = 0; d < GDOF; ++d) {
				const double A0 = a0 * N0[d];
				const double A1 = a1 * N0[d] + a0 * (dNx[d] * vx + dNy[d] * vy);
				Fm1[d] = -(Sm2 * A0 + A1) * f4;
				Fm2[d] = -A0 * f4;
				s[d][3].x += (Fm1[d] * lr - Fm2[d] / rho_lim) * wt;
			}
			for (int ir = 0; ir < 3; ++ir) {
				const double rho = (1.0 + R.gl3x[ir]) * rho_lim / 2.0;
				const double wr = R.gl3w[ir] * rho_lim / 2.0;
				const double ex = rho * ct, ey = rho * st;
				const double xi = Ti00 * ex + Ti01 * ey + xi0x, eta = Ti11 * ey + xi0y;
				double2 F[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
				D3 a, b;
				quad_dx(Q, xi, eta, a, b);
				const D3 nv = d3cross(a, b);
				const double J = d3norm(nv);
				kernel_point<FAM, HSP, CK>(x0, quad_x(Q, xi, eta), n0, d3scale(nv, 1.0 / J), k, J / detT * rho, F, ki);
				double N[GDOF];
				gshape64(xi, eta, N);
				for (int d = 0; d < GDOF; ++d) {
					const double sing = (Fm1[d] + Fm2[d] / rho) / rho;
					s[d][3].x += (F[3].x * N[d] - sing) * (wt * wr);
					s[d][3].y += F[3].y * N[d] * (wt * wr);
				}
			}
		}
	}
}

// The FP64 precision policy for the gauss field: every regular block in
// FP64, each kernel group at its Gauss order from the trial element's FP64
// table (y, w J, unit normal), each point's values spread over the four dofs
// with the FP64 shape-function table.
struct Gauss64Args {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	int nodes;                   // test nodes per test element (4 or 1)
	int tn, te;
	const double *tx, *tnu;      // collocation points, unit normals [3][tn]
	const double *tc;            // test element centres [3][te]
	const double *sc;            // trial centres [3][sn]
	int sn;
	const double *thr;           // [NTHR][sn]
	const double *qt[QN_MAX + 1];   // per Gauss order N: [sn][N * N][8]
	const GaussShapes *gs;
	int fixed_n;
	double k, ki;
	double2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

template <int FAM, int MASK, bool CK>
__device__ __forceinline__ void gauss_eval64(const Gauss64Args &a, int j, int n, D3 x0, D3 nx,
	double2 (&s)[GDOF][4])
{
	if (n <= 0) return;
	const int nq = n * n;
	const double *t = a.qt[n] + (size_t)j * nq * 8;
	for (int q = 0; q < nq; ++q) {
		const D3 y = {t[8 * q], t[8 * q + 1], t[8 * q + 2]};
		const D3 ny = {t[8 * q + 4], t[8 * q + 5], t[8 * q + 6]};
		double2 f[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
		kernel_point<FAM, MASK, CK>(x0, y, nx, ny, a.k, t[8 * q + 3], f, a.ki);
		double N[GDOF];
		for (int d = 0; d < GDOF; ++d) N[d] = a.gs->n64[n][q][d];
		gspread<MASK>(N, f, s);
	}
}

template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(256) gauss_regular64_kernel(Gauss64Args a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int tile_c = blockIdx.x / a.row_tiles;
	const int r = tile_r * 256 + threadIdx.x;
	if (r >= a.n_rows) return;
	const int ie = a.nodes == GDOF ? r >> 2 : r;
	const D3 x0 = {a.tx[r], a.tx[a.tn + r], a.tx[2 * a.tn + r]};
	const D3 nx = {a.tnu[r], a.tnu[a.tn + r], a.tnu[2 * a.tn + r]};
	const double cix = a.tc[ie], ciy = a.tc[a.te + ie], ciz = a.tc[2 * a.te + ie];
	const int j0 = tile_c * a.tile_cols;
	const int j1 = min(j0 + a.tile_cols, a.n_cols);
	for (int j = j0; j < j1; ++j) {
		// classification by element centres, as gauss_regular_kernel
		const double d2 = d2_pair(a.sc[j], a.sc[a.sn + j], a.sc[2 * a.sn + j], cix, ciy, ciz);
		if (d2 < a.thr[THR_NEAR * a.sn + j] || (a.same_mesh && ie == j)) continue;
		const int fn = a.fixed_n;
		double2 s[GDOF][4];
		for (int d = 0; d < GDOF; ++d)
			for (int t = 0; t < 4; ++t) s[d][t] = make_double2(0, 0);
		if (KM & SLP)
			gauss_eval64<FAM, SLP, CK>(a, j, fn > 0 ? fn : gauss_order(2 + table_degree<1>(d2, a.thr, a.sn, j)), x0, nx,
				s);
		if (KM & (DLP | DLPT))
			gauss_eval64<FAM, KM & (DLP | DLPT), CK>(a, j,
				fn > 0 ? fn : gauss_order(2 + table_degree<2>(d2, a.thr, a.sn, j)), x0, nx, s);
		if (KM & HSP)
			gauss_eval64<FAM, HSP, CK>(a, j, fn > 0 ? fn : gauss_order(2 + table_degree<3>(d2, a.thr, a.sn, j)), x0, nx,
				s);
		for (int d = 0; d < GDOF; ++d) {
			const long long idx = (long long)(GDOF * j + d) * a.ld + r;
			if (MODE == OUT_SINGLE) {
				const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
				const double re = a.alpha.x * s[d][kk].x - a.alpha.y * s[d][kk].y;
				const double im = a.alpha.x * s[d][kk].y + a.alpha.y * s[d][kk].x;
				store_entry64(a.out0, idx, a.dtype, a.accumulate, re, im);
			} else {
				const double ar = a.alpha.x * s[d][3].x - a.alpha.y * s[d][3].y + s[d][1].x;
				const double ai = a.alpha.x * s[d][3].y + a.alpha.y * s[d][3].x + s[d][1].y;
				const double br = s[d][0].x + a.alpha.x * s[d][2].x - a.alpha.y * s[d][2].y;
				const double bi = s[d][0].y + a.alpha.x * s[d][2].y + a.alpha.y * s[d][2].x;
				store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
				store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
			}
		}
	}
}

// store the (row, 4 dofs of element j) entries; DIAG: the identity of the
// Burton-Miller pair at the row's own dof
template <int KM, int MODE, bool DIAG>
__device__ __forceinline__ void gauss_store64(const GaussNearArgs &a, int row, int j, int node,
	const double2 (&s)[GDOF][4])
{
#pragma unroll
	for (int d = 0; d < GDOF; ++d) {
		const long long idx = (long long)(GDOF * j + d) * a.ld + row;
		if (MODE == OUT_SINGLE) {
			const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
			const double re = a.alpha.x * s[d][kk].x - a.alpha.y * s[d][kk].y;
			const double im = a.alpha.x * s[d][kk].y + a.alpha.y * s[d][kk].x;
			store_entry64(a.out0, idx, a.dtype, a.accumulate, re, im);
		} else {
			double ar = a.alpha.x * s[d][3].x - a.alpha.y * s[d][3].y + s[d][1].x;
			double ai = a.alpha.x * s[d][3].y + a.alpha.y * s[d][3].x + s[d][1].y;
			double br = s[d][0].x + a.alpha.x * s[d][2].x - a.alpha.y * s[d][2].y;
			double bi = s[d][0].y + a.alpha.x * s[d][2].y + a.alpha.y * s[d][2].x;
			if (DIAG && d == node) {
				ar += -0.5;
				br += a.alpha.x / 2.0;
				bi += a.alpha.y / 2.0;
			}
			store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
			store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
		}
	}
}

__device__ __forceinline__ void gauss_load(const GaussNearArgs &a, int row, int j, Quad &Q, D3 &x0, D3 &nx)
{
	D3 X[4];
	for (int v = 0; v < 4; ++v)
		X[v] = {a.X[(3 * v) * a.sn + j], a.X[(3 * v + 1) * a.sn + j], a.X[(3 * v + 2) * a.sn + j]};
	Q = quad_from_corners(X);
	x0 = {a.tx[row], a.tx[a.tn + row], a.tx[2 * a.tn + row]};
	nx = {a.tnu[row], a.tnu[a.tn + row], a.tnu[2 * a.tn + row]};
}

// Diagonal blocks: one lane per (element, node): Duffy around the node for
// SLP / DLP / DLPt, Guiggiani there for HSP.
template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(64) gauss_diag_kernel(GaussNearArgs a)
{
	const int gid = blockIdx.x * blockDim.x + threadIdx.x;
	const int e = gid >> 2, node = gid & 3;
	if (e >= a.n_pairs) return;
	const Rules &R = *a.rules;
	const int row = GDOF * e + node;
	Quad Q;
	D3 x0, nx;
	gauss_load(a, row, e, Q, x0, nx);
	double2 s[GDOF][4];
	for (int d = 0; d < GDOF; ++d)
		for (int t = 0; t < 4; ++t) s[d][t] = make_double2(0, 0);
	constexpr int WEAK = KM & (SLP | DLP | DLPT);
	if (WEAK)
		for (int q = 0; q < 64; ++q) {
			const double xi = R.gduffy[node][q][0], eta = R.gduffy[node][q][1];
			double2 f[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
			quad_kernel_at<FAM, WEAK, CK>(Q, xi, eta, x0, nx, a.k, R.gduffy[node][q][2], f, a.ki);
			double N[GDOF];
			gshape64(xi, eta, N);
			gspread<WEAK>(N, f, s);
		}
	if (KM & HSP) {
		double xi0, eta0;
		gnode64(node, xi0, eta0);
		gauss_guiggiani<FAM, CK>(Q, xi0, eta0, a.k, R, a.ki, s);
	}
	gauss_store64<KM, MODE, true>(a, row, e, node, s);
}

// Near-singular pairs: QLPP lanes per (pair, test node), then a butterfly sum
// over the lanes; the first lane stores the node's four dofs.
template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(64) gauss_near_kernel(GaussNearArgs a)
{
	const int gid = blockIdx.x * blockDim.x + threadIdx.x;
	const int lane = gid % QLPP;
	const int pn = gid / QLPP;
	const int p = pn / a.nodes, node = pn % a.nodes;
	// lanes past the last pair repeat it and do not store
	const bool valid = p < a.n_pairs;
	const int pp = valid ? p : a.n_pairs - 1;
	const int i = a.pi[pp], j = a.pj[pp];
	const int row = i * a.nodes + node;
	const Rules &R = *a.rules;
	Quad Q;
	D3 x0, nx;
	gauss_load(a, row, j, Q, x0, nx);
	double2 s[GDOF][4];
	for (int d = 0; d < GDOF; ++d)
		for (int t = 0; t < 4; ++t) s[d][t] = make_double2(0, 0);
	bool ok = true;
	if (FAM == HELM) {
		if (KM & SLP)
			ok = gauss_near_polar<FAM, SLP, CK>(Q, x0, nx, a.k, 8, R.gl8x, R.gl8w, 8, R.gl8x, R.gl8w, R, s, lane, QLPP,
				a.inv_iter, a.ki);
		if (ok && (KM & (DLP | DLPT | HSP)))
			ok = gauss_near_telles<FAM, KM & (DLP | DLPT | HSP), CK>(Q, x0, nx, a.k, R, s, lane, QLPP, a.inv_iter, a.ki);
	} else {
		if (KM & (SLP | DLP | DLPT))
			ok = gauss_near_polar<FAM, KM & (SLP | DLP | DLPT), false>(Q, x0, nx, 0.0, 8, R.gl8x, R.gl8w, 8, R.gl8x,
				R.gl8w, R, s, lane, QLPP, a.inv_iter, 0.0);
		if (ok && (KM & HSP))
			ok = gauss_near_polar<FAM, HSP, false>(Q, x0, nx, 0.0, 51, R.gl51x, R.gl51w, 51, R.gl51x, R.gl51w, R, s,
				lane, QLPP, a.inv_iter, 0.0);
	}
	for (int off = QLPP / 2; off >= 1; off >>= 1)
		for (int d = 0; d < GDOF; ++d)
			for (int t = 0; t < 4; ++t) {
				s[d][t].x += __shfl_xor(s[d][t].x, off, QLPP);
				s[d][t].y += __shfl_xor(s[d][t].y, off, QLPP);
			}
	if (!ok) {
		const double nan = __builtin_nan("");
		for (int d = 0; d < GDOF; ++d)
			for (int t = 0; t < 4; ++t) s[d][t] = make_double2(nan, nan);
		if (valid && lane == 0 && a.status) atomicAdd(a.status, 1);
	}
	if (valid && lane == 0) gauss_store64<KM, MODE, false>(a, row, j, node, s);
}

} // namespace nbk
