// nbem_galquad.hpp -- Galerkin formalism on quadrilaterals (gfx950).
//
// Fields on quad_1 elements: constant (one dof per element, ND = 1) or
// isoparametric bilinear (one dof per mesh node, ND = 4 shape functions).
// Entry (dof_a, dof_b) = sum over element pairs (e, f) of
//   int_e N_a(x) int_f K(x, y) N_b(y) dy dx
// (nihu/core/double_integral.hpp:181-206 regular, :214-248 singular).
//
// Regular pairs (FP32): the organisation of the triangle kernel -- a wave
// holds a tile of up to 64 test elements (one per lane), the trial element is
// wave-uniform (scalar loads) -- with the N x N Gauss square rule of the
// pair's degree on both elements (degree = field complexity + the kernel's
// distance interval, or fixed; N = degree / 2 + 1).  Quadrilaterals are not
// flat, so every point row of the per-element tables carries its own unit
// normal: row = o (point - centre), |o|^2, n, o.n, w J N_b (b = 0..3), w J.
//   r = y - x = (c_f - c_e) + o_q - a_p,  r^2 = d^2 + u with
//   u = 2 (c_f - c_e - a_p).o_q + |o_q|^2 + |a_p|^2 - 2 (c_f - c_e).a_p;
//   r.n_q = (c_f - c_e - a_p).n_q + o_q.n_q;  r.n_p = (c_f - c_e - a_p).n_p + o_q.n_p;
// the phase as in the collocation kernels (FP64 fract(k d / 2 pi) per pair,
// k (r - d) = k u / (r + d) per point).  Nodal fields keep the lane-private
// LDS column slots of the triangle kernel (4 corners x GALQ_CS slots per lane)
// and add each evicted column with one global atomic per row node.
// Singular pairs (FP64, same mesh): coincident, edge- and corner-adjacent
// element pairs, one wave per pair over the 2048 / 1536 / 1280 points of the
// quad x quad rules, the element corners permuted to the rules' canonical
// configuration (nihu/core/singular_accelerator.hpp:283-325).

#pragma once

#include "nbem_galerkin.hpp"
#include "nbem_quad.hpp"

namespace nbk {

constexpr int GALQ_CS = 6;      // column-node slots per lane (nodal fields)
constexpr int GALQ_NMAX = 8;    // Gauss points per axis (degree <= 14)
constexpr int GALQ_ROW = 16;    // floats per point row

struct GalQArgs {
	int n_tiles, chunk;          // test wave tiles; trial sequence positions per chunk
	int tn, sn;                  // element counts of test / trial mesh (SoA strides)
	int same_mesh;
	const int *t_elem;           // nodal: [n_tiles][64] test element of each lane (-1: idle)
	const uint2 *t_contrib;      // nodal: [n_tiles][128] contributors lane * 4 + corner of local row r (lane r % 64)
	const int *t_rows;           // nodal: [n_tiles][128] node of each local row (-1: unused)
	const int *t_quad;           // test element nodes [4][tn]
	const int *s_order;          // nodal: trial element at each sequence position
	const int *s_col;            // nodal: [sn][5] previous slot occupant of corners 0..3, packed slots
	const int *s_fin;            // nodal: [chunks][GALQ_CS] slot occupants at the end of a chunk
	const int *s_quad;           // trial element nodes [4][sn]
	const double *tc, *sc;       // centres x(0, 0) [3][n]
	const double *thr;           // trial distance thresholds [NTHR][sn]
	const float *tq[GALQ_NMAX + 1], *sq[GALQ_NMAX + 1];   // point rows by points per axis
	int fixed_n;                 // > 0: the fixed rule's points per axis
	int complexity;              // field complexity: 1 constant, 2 nodal
	float k, k_revf;             // wavenumber, k / (2 pi)
	double k_rev;
	float ki;                    // imaginary part of a complex wavenumber (kernels with CK)
	float2 alpha;                // alpha / (4 pi)
	void *out;
	long long ld;
	int dtype, accumulate;
};

// One pair over the diadic N x N rule: per test point p (lane's row) the
// trial sums S_b = sum_q K(x_p, y_q) w_q J_q N_b(q) (ND = 4) or w_q J_q
// (ND = 1), then blk[a][b] += w_p J_p N_a(p) S_b.  Kernel values without
// 1/(4 pi): SLP E, DLP -Q r.n_y, DLPt Q r.n_x, HSP n_x.n_y Q - (3Q - k^2 E r^2)
// r.n_x r.n_y / r^2 with E = e^{-i th}/r, G = E/r^2, Q = G (1 + i th),
// th = k r, th G = k e^{-i th}/r^2.  CK: complex wavenumber k + i kim, the
// attenuation e^{kim r} on c and sn, Q gains -kim r G, k^2 becomes (k + i kim)^2.
template <int FAM, int KM, int ND, bool CK = false>
__device__ __forceinline__ void galq_points(const float *tp, const float *sq_in, int nq, const GalPair &P,
	float k_revf, float k, float2 (&blk)[ND][ND], float kim)
{
	const auto sf = uni(sq_in);
	constexpr bool IM = FAM == HELM;
	for (int p = 0; p < nq; ++p) {
		const float4 t0 = *(const float4 *)(tp + GALQ_ROW * p);
		const float4 t1 = *(const float4 *)(tp + GALQ_ROW * p + 4);
		const float4 t2 = *(const float4 *)(tp + GALQ_ROW * p + 8);
		const float tw = tp[GALQ_ROW * p + 12];
		const float ax = t0.x, ay = t0.y, az = t0.z;
		// h = c_f - c_e - a_p, ex = 2 h
		const float hx = 0.5f * P.ex0 - ax, hy = 0.5f * P.ey0 - ay, hz = 0.5f * P.ez0 - az;
		const float ex = 2.0f * hx, ey = 2.0f * hy, ez = 2.0f * hz;
		const float vp = t0.w - __builtin_fmaf(P.ex0, ax, __builtin_fmaf(P.ey0, ay, P.ez0 * az));
		const float nxx = t1.x, nxy = t1.y, nxz = t1.z;
		float rnx0 = 0.0f;
		if (KM & (DLPT | HSP)) rnx0 = __builtin_fmaf(hx, nxx, __builtin_fmaf(hy, nxy, hz * nxz));
		float2 S[ND];
#pragma unroll
		for (int b = 0; b < ND; ++b) S[b] = zero2();
		for (int q = 0; q < nq; ++q) {
			const int o = GALQ_ROW * q;
			const float ox = sf[o], oy = sf[o + 1], oz = sf[o + 2];
			const float u = __builtin_fmaf(ex, ox, __builtin_fmaf(ey, oy, __builtin_fmaf(ez, oz, sf[o + 3] + vp)));
			const float r2 = P.d2f + u;
			const float rinv = __builtin_amdgcn_rsqf(r2);
			float c = 1.0f, sn = 0.0f;
			if (IM) {
				const float rd = u * __builtin_amdgcn_rcpf(__builtin_fmaf(r2, rinv, P.df));   // r - d
				const float tt = __builtin_fmaf(k_revf, rd, P.t0);
				c = __builtin_amdgcn_cosf(tt);
				sn = __builtin_amdgcn_sinf(tt);
				if (CK) {
					const float am = P.ad * __builtin_amdgcn_exp2f(kim * 1.4426950408889634f * rd);
					c *= am;
					sn *= am;
				}
			}
			float kr, ki = 0.0f;
			if (KM == SLP) {
				kr = rinv * c;
				if (IM) ki = -rinv * sn;
			} else {
				const float ri2 = rinv * rinv;
				const float g3 = rinv * ri2;
				const float Gr = g3 * c, Gi = IM ? -g3 * sn : 0.0f;
				const float kh = k * ri2;
				float Qr = IM ? __builtin_fmaf(kh, sn, Gr) : Gr;
				float Qi = IM ? __builtin_fmaf(kh, c, Gi) : 0.0f;
				if (CK) {
					const float kih = kim * ri2;
					Qr = __builtin_fmaf(-kih, c, Qr);
					Qi = __builtin_fmaf(kih, sn, Qi);
				}
				const float nyx = sf[o + 4], nyy = sf[o + 5], nyz = sf[o + 6];
				float rny = 0.0f;
				if (KM & (DLP | HSP)) rny = __builtin_fmaf(hx, nyx, __builtin_fmaf(hy, nyy, __builtin_fmaf(hz, nyz, sf[o + 7])));
				if (KM == DLP) {
					kr = -Qr * rny; ki = -Qi * rny;
				} else {
					const float rnx = rnx0 + __builtin_fmaf(ox, nxx, __builtin_fmaf(oy, nxy, oz * nxz));
					if (KM == DLPT) {
						kr = Qr * rnx; ki = Qi * rnx;
					} else {
						const float nn = __builtin_fmaf(nxx, nyx, __builtin_fmaf(nxy, nyy, nxz * nyz));
						const float cf = rnx * rny * ri2;
						float hr, hi;
						if (CK) {
							const float kr2 = (k * k - kim * kim) * rinv, ki2 = 2.0f * k * kim * rinv;
							hr = 3.0f * Qr - (kr2 * c + ki2 * sn);
							hi = 3.0f * Qi - (ki2 * c - kr2 * sn);
						} else {
							const float kk = k * k * rinv;
							hr = __builtin_fmaf(3.0f, Qr, -kk * c);
							hi = __builtin_fmaf(3.0f, Qi, kk * sn);
						}
						kr = __builtin_fmaf(nn, Qr, -hr * cf);
						ki = __builtin_fmaf(nn, Qi, -hi * cf);
					}
				}
			}
			if (ND == 1) {
				const float w = sf[o + 12];
				S[0].x = __builtin_fmaf(kr, w, S[0].x);
				if (IM || CK) S[0].y = __builtin_fmaf(ki, w, S[0].y);
			} else {
#pragma unroll
				for (int b = 0; b < ND; ++b) {
					const float w = sf[o + 8 + b];
					S[b].x = __builtin_fmaf(kr, w, S[b].x);
					if (IM || CK) S[b].y = __builtin_fmaf(ki, w, S[b].y);
				}
			}
		}
		// w J N_a of the test point (ND = 4), w J (ND = 1)
		const float w4[4] = {t2.x, t2.y, t2.z, t2.w};
		float wa[ND];
#pragma unroll
		for (int x = 0; x < ND; ++x) wa[x] = ND == 1 ? tw : w4[x];
#pragma unroll
		for (int x = 0; x < ND; ++x)
#pragma unroll
			for (int y = 0; y < ND; ++y) {
				blk[x][y].x = __builtin_fmaf(wa[x], S[y].x, blk[x][y].x);
				blk[x][y].y = __builtin_fmaf(wa[x], S[y].y, blk[x][y].y);
			}
	}
}

// per-lane accumulator of corner c in column slot s: acc[(s * 4 + c) * 64 + lane]
__device__ __forceinline__ float2 *galq_acc(float2 *acc, int slot, int corner, int lane)
{
	return acc + (slot * 4 + corner) * GAL_TILE + lane;
}

// add one column slot to the result and clear it: lane l sums its local rows
// l and l + 64 over their contributors, then every lane clears its own four
// entries
__device__ __forceinline__ void galq_flush(float2 *acc, int lane, int slot, int col, const int (&rnode)[2],
	const uint2 (&con)[2], const GalQArgs &a)
{
	float2 v[2];
#pragma unroll
	for (int r = 0; r < 2; ++r) {
		v[r] = make_float2(0.0f, 0.0f);
#pragma unroll
		for (int q = 0; q < 8; ++q) {
			const unsigned c = ((q < 4 ? con[r].x : con[r].y) >> (8 * (q & 3))) & 255u;
			if (c != 255u) {
				const float2 t = *galq_acc(acc, slot, c & 3u, c >> 2);
				v[r].x += t.x;
				v[r].y += t.y;
			}
		}
	}
#pragma unroll
	for (int c = 0; c < 4; ++c) *galq_acc(acc, slot, c, lane) = make_float2(0.0f, 0.0f);
#pragma unroll
	for (int r = 0; r < 2; ++r)
		if (rnode[r] >= 0 && (v[r].x != 0.0f || v[r].y != 0.0f)) {
			const float re = a.alpha.x * v[r].x - a.alpha.y * v[r].y;
			const float im = a.alpha.x * v[r].y + a.alpha.y * v[r].x;
			atomic_add_entry(a.out, (long long)col * a.ld + rnode[r], a.dtype, re, im);
		}
}

// Grid: (test tiles / 4) x chunks, 256 lanes (4 waves, one test tile each).
template <int FAM, int KM, int ND, bool CK = false>
__global__ void __launch_bounds__(256) galq_regular_kernel(GalQArgs a)
{
	__shared__ float2 acc_all[ND == 4 ? 4 * GALQ_CS * 4 * GAL_TILE : 1];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int wg_tiles = (a.n_tiles + 3) >> 2;
	const int tile = (int)(blockIdx.x % wg_tiles) * 4 + wave;
	const int chunk = (int)(blockIdx.x / wg_tiles);
	if (tile >= a.n_tiles) return;     // no block-wide barriers below
	float2 *acc = acc_all + (ND == 4 ? wave * GALQ_CS * 4 * GAL_TILE : 0);
	constexpr int ORDER = KM == SLP ? 1 : KM == HSP ? 3 : 2;

	int e;
	if (ND == 4) e = a.t_elem[tile * GAL_TILE + lane];
	else e = tile * GAL_TILE + lane < a.tn ? tile * GAL_TILE + lane : -1;
	const bool lane_ok = e >= 0;
	const int ee = lane_ok ? e : (ND == 4 ? a.t_elem[tile * GAL_TILE] : tile * GAL_TILE);
	const double cix = a.tc[ee], ciy = a.tc[a.tn + ee], ciz = a.tc[2 * a.tn + ee];
	int tv[4] = {-1, -1, -1, -1};
	if (a.same_mesh)
#pragma unroll
		for (int c = 0; c < 4; ++c) tv[c] = a.t_quad[c * a.tn + ee];
	int rnode[2] = {-1, -1};
	uint2 con[2] = {make_uint2(0xffffffffu, 0xffffffffu), make_uint2(0xffffffffu, 0xffffffffu)};
	if (ND == 4) {
#pragma unroll
		for (int r = 0; r < 2; ++r) {
			con[r] = a.t_contrib[(tile * 2 + r) * GAL_TILE + lane];
			rnode[r] = a.t_rows[(tile * 2 + r) * GAL_TILE + lane];
		}
#pragma unroll
		for (int s = 0; s < GALQ_CS; ++s)
#pragma unroll
			for (int c = 0; c < 4; ++c) *galq_acc(acc, s, c, lane) = make_float2(0.0f, 0.0f);
	}
	const int i0 = chunk * a.chunk;
	const int i1 = min(i0 + a.chunk, a.sn);
	const auto sc = uni(a.sc);
	const auto squad = uni(a.s_quad);
	const auto sord = uni(a.s_order);
	const auto scol = uni(a.s_col);
	GalPair P;
	P.nxx = P.nxy = P.nxz = 0.0f;
	P.nyx = P.nyy = P.nyz = 0.0f;
	P.dcny = P.dcnx = P.nn = 0.0f;
	P.ad = 1.0f;
	for (int i = i0; i < i1; ++i) {
		const int f = ND == 4 ? sord[i] : i;
		int sl[4] = {0, 0, 0, 0};
		if (ND == 4) {
			const int pk = scol[5 * i + 4];
#pragma unroll
			for (int b = 0; b < 4; ++b) {
				sl[b] = (pk >> (5 * b)) & 31;
				const int v = scol[5 * i + b];
				if (v >= 0) galq_flush(acc, lane, sl[b], v, rnode, con, a);
			}
		}
		const double cjx = sc[f], cjy = sc[a.sn + f], cjz = sc[2 * a.sn + f];
		const double d2 = d2_pair(cix, ciy, ciz, cjx, cjy, cjz);
		bool sing = false;
		if (a.same_mesh) {
#pragma unroll
			for (int c = 0; c < 4; ++c) {
				const int fv = squad[c * a.sn + f];
				sing = sing || tv[0] == fv || tv[1] == fv || tv[2] == fv || tv[3] == fv;
			}
		}
		const bool act = lane_ok && !sing;
		float2 blk[ND][ND];
#pragma unroll
		for (int x = 0; x < ND; ++x)
#pragma unroll
			for (int y = 0; y < ND; ++y) blk[x][y] = zero2();
		if (__ballot(act) != 0) {
			P.ex0 = (float)(2.0 * (cjx - cix)); P.ey0 = (float)(2.0 * (cjy - ciy)); P.ez0 = (float)(2.0 * (cjz - ciz));
			P.d2f = (float)d2;
			if (FAM == HELM) {
				double y = __builtin_amdgcn_rsq(d2);
				y = y * (1.5 - 0.5 * d2 * y * y);
				const double d = d2 * y;
				double t = d * a.k_rev;
				t -= rint(t);
				P.df = (float)d;
				P.t0 = (float)t;
				if (CK) P.ad = __builtin_amdgcn_exp2f((float)(d * (a.ki * 1.4426950408889634)));
			} else {
				P.df = 0.0f; P.t0 = 0.0f;
			}
			// points per axis per lane; lanes are served one distinct rule at a time
			int nsel = -1;
			if (act) nsel = a.fixed_n > 0 ? a.fixed_n : (a.complexity + table_degree<ORDER>(d2, a.thr, a.sn, f)) / 2 + 1;
			for (;;) {
				const unsigned long long m = __ballot(nsel > 0);
				if (m == 0) break;
				const int N = __builtin_amdgcn_readlane(nsel, __ffsll((long long)m) - 1);
				const unsigned long long msel = __ballot(nsel == N);
				if ((msel >> __lane_id()) & 1ull) {
					const int nq = N * N;
					galq_points<FAM, KM, ND, CK>(a.tq[N] + (size_t)ee * nq * GALQ_ROW,
						a.sq[N] + (size_t)f * nq * GALQ_ROW, nq, P, a.k_revf, a.k, blk, a.ki);
					nsel = -1;
				}
			}
		}
		if (ND == 1) {
			// singular entries: zero here, the FP64 pass adds them
			if (lane_ok && (act || !a.accumulate)) {
				const float re = a.alpha.x * blk[0][0].x - a.alpha.y * blk[0][0].y;
				const float im = a.alpha.x * blk[0][0].y + a.alpha.y * blk[0][0].x;
				store_entry(a.out, (long long)f * a.ld + e, a.dtype, a.accumulate, re, im);
			}
		} else if (act) {
#pragma unroll
			for (int x = 0; x < ND; ++x)
#pragma unroll
				for (int y = 0; y < ND; ++y) {
					float2 *p = galq_acc(acc, sl[y], x, lane);
					float2 v = *p;
					v.x += blk[x][y].x;
					v.y += blk[x][y].y;
					*p = v;
				}
		}
	}
	if (ND == 4) {
		const auto sfin = uni(a.s_fin);
		for (int s = 0; s < GALQ_CS; ++s) {
			const int v = sfin[chunk * GALQ_CS + s];
			if (v >= 0) galq_flush(acc, lane, s, v, rnode, con, a);
		}
	}
}

struct GalQSingArgs {
	int n_pairs;
	const int *pe, *pf, *pinfo;  // test / trial element, match | ind1 << 4 | ind2 << 8
	int n;                       // elements (SoA stride)
	const double *X;             // corners [12][n]
	const int *t_dof, *s_dof;    // row / column dof of each element corner [4][n] (nodes,
	                             // or the element at every corner for a constant field)
	const double *rule[3];       // canonical rules by match dimension: [points][5]
	int rule_n[3];
	double k;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	double2 alpha;
	void *out;
	long long ld;
	int dtype;
};

// One wave per singular pair.  The rule's corner c of the test element is the
// element's corner (ind1 + c) % 4; of the trial element (ind2 + 1 - c) mod 4
// for a common edge (it runs the other way) and (ind2 + c) % 4 for a common
// corner.  The corner coordinates are permuted accordingly (a rotation or
// reflection of the reference square, |J| = 1), so that the shape function
// N_c at a rule point belongs to the permuted corner c; a reflection turns
// the cross product of the tangents around, and the unit normal is taken
// with the element's own orientation.
template <int FAM, int KM, int ND, bool CK = false>
__global__ void __launch_bounds__(64) galq_singular_kernel(GalQSingArgs a)
{
	const int p = blockIdx.x, lane = threadIdx.x;
	if (p >= a.n_pairs) return;
	const int e = a.pe[p], f = a.pf[p], info = a.pinfo[p];
	const int match = info & 15, i1 = (info >> 4) & 15, i2 = (info >> 8) & 15;
	int ct[4], cs[4];
	for (int c = 0; c < 4; ++c) {
		ct[c] = match == 2 ? c : (i1 + c) & 3;
		cs[c] = match == 2 ? c : match == 1 ? (i2 + 5 - c) & 3 : (i2 + c) & 3;
	}
	const int n = a.n;
	D3 Vt[4], Vs[4];
	for (int c = 0; c < 4; ++c) {
		Vt[c] = {a.X[(3 * ct[c]) * n + e], a.X[(3 * ct[c] + 1) * n + e], a.X[(3 * ct[c] + 2) * n + e]};
		Vs[c] = {a.X[(3 * cs[c]) * n + f], a.X[(3 * cs[c] + 1) * n + f], a.X[(3 * cs[c] + 2) * n + f]};
	}
	const Quad Qt = quad_from_corners(Vt), Qs = quad_from_corners(Vs);
	// a reflection of the corners (the trial element of an edge match)
	// reverses dx/dxi x dx/deta: the element's own normal is the opposite
	const double ot = ((ct[1] - ct[0]) & 3) == 1 ? 1.0 : -1.0;
	const double os = ((cs[1] - cs[0]) & 3) == 1 ? 1.0 : -1.0;
	const double *R = a.rule[match];
	const int nr = a.rule_n[match];
	double acc[ND * ND * 2];
	for (int q = 0; q < ND * ND * 2; ++q) acc[q] = 0.0;
	for (int q = lane; q < nr; q += 64) {
		const double txi = R[5 * q], teta = R[5 * q + 1], sxi = R[5 * q + 2], seta = R[5 * q + 3], w = R[5 * q + 4];
		D3 a1, b1, a2, b2;
		quad_dx(Qt, txi, teta, a1, b1);
		quad_dx(Qs, sxi, seta, a2, b2);
		const D3 n1 = d3cross(a1, b1), n2 = d3cross(a2, b2);
		const double J1 = d3norm(n1), J2 = d3norm(n2);
		const D3 x = quad_x(Qt, txi, teta), y = quad_x(Qs, sxi, seta);
		const D3 u1 = d3scale(n1, ot / J1), u2 = d3scale(n2, os / J2);
		const double2 K = CK ? kernel_value64_c<KM>(x, y, u1, u2, a.k, a.ki) : kernel_value64<FAM, KM>(x, y, u1, u2, a.k);
		const double wt = J1 * w, ws = J2;
		if (ND == 1) {
			acc[0] += wt * K.x * ws;
			acc[1] += wt * K.y * ws;
		} else {
			const double Nt[4] = {0.25 * (1 - txi) * (1 - teta), 0.25 * (1 + txi) * (1 - teta),
				0.25 * (1 + txi) * (1 + teta), 0.25 * (1 - txi) * (1 + teta)};
			const double Ns[4] = {0.25 * (1 - sxi) * (1 - seta), 0.25 * (1 + sxi) * (1 - seta),
				0.25 * (1 + sxi) * (1 + seta), 0.25 * (1 - sxi) * (1 + seta)};
			for (int x4 = 0; x4 < 4; ++x4)
				for (int y4 = 0; y4 < 4; ++y4) {
					const double ta = Nt[x4] * wt, sb = Ns[y4] * ws;
					acc[2 * (4 * x4 + y4)] += ta * K.x * sb;
					acc[2 * (4 * x4 + y4) + 1] += ta * K.y * sb;
				}
		}
	}
	for (int q = 0; q < ND * ND * 2; ++q)
		for (int off = 32; off >= 1; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
	if (lane != 0) return;
	for (int x4 = 0; x4 < ND; ++x4)
		for (int y4 = 0; y4 < ND; ++y4) {
			const double vr = acc[2 * (ND * x4 + y4)], vi = acc[2 * (ND * x4 + y4) + 1];
			const double re = a.alpha.x * vr - a.alpha.y * vi, im = a.alpha.x * vi + a.alpha.y * vr;
			if (ND == 1) {
				store_entry64(a.out, (long long)f * a.ld + e, a.dtype, 1, re, im);
			} else {
				const int row = a.t_dof[ct[x4] * n + e], col = a.s_dof[cs[y4] * n + f];
				atomic_add_entry64(a.out, (long long)col * a.ld + row, a.dtype, re, im);
			}
		}
}

// identity (mass matrix) on quadrilaterals: the Gauss rule of degree
// shape orders + jacobian order 1 (nihu/core/single_integral.hpp:65-140):
// constant fields 4 J(0, 0) on the diagonal; nodal fields
// sum over the 2 x 2 points of N_a N_b J w per element, through the dof maps
// (a constant side names the element at every corner: the sums over its
// corners are the 2 x 2 rule of N_b J, NiHu's degree 2 for a constant test
// and a bilinear trial)
__global__ void galq_identity_kernel(void *out, long long ld, int n, const double *X, const int *t_map,
	const int *s_map, int nodal, int dtype, double ar, double ai)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n) return;
	D3 V[4];
	for (int c = 0; c < 4; ++c) V[c] = {X[(3 * c) * n + e], X[(3 * c + 1) * n + e], X[(3 * c + 2) * n + e]};
	const Quad Q = quad_from_corners(V);
	if (!nodal) {
		D3 a, b;
		quad_dx(Q, 0.0, 0.0, a, b);
		const double m = 4.0 * d3norm(d3cross(a, b));
		store_entry64(out, (long long)e * ld + e, dtype, 1, ar * m, ai * m);
		return;
	}
	const double g = 0.57735026918962576451;   // 1 / sqrt 3, weights 1
	double M[4][4] = {};
	for (int q = 0; q < 4; ++q) {
		const double xi = (q & 2) ? g : -g, eta = (q & 1) ? g : -g;
		D3 a, b;
		quad_dx(Q, xi, eta, a, b);
		const double J = d3norm(d3cross(a, b));
		const double N[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta), 0.25 * (1 + xi) * (1 + eta),
			0.25 * (1 - xi) * (1 + eta)};
		for (int x = 0; x < 4; ++x)
			for (int y = 0; y < 4; ++y) M[x][y] += N[x] * N[y] * J;
	}
	for (int x = 0; x < 4; ++x)
		for (int y = 0; y < 4; ++y)
			atomic_add_entry64(out, (long long)s_map[y * n + e] * ld + t_map[x * n + e], dtype, ar * M[x][y], ai * M[x][y]);
}

} // namespace nbk
