// nbem_galerkin.hpp -- Galerkin formalism on flat triangles (gfx950).
//
// Fields: constant (one dof per element, NDOF = 1) or isoparametric linear
// (one dof per mesh node, NDOF = 3 shape functions per element).
// Entry (dof_a, dof_b) = sum over element pairs (e, f) of
//   int_e N_a(x) int_f K(x, y) N_b(y) dy dx
// (nihu/core/double_integral.hpp:181-206 regular, :214-248 singular).
//
// Regular pairs (FP32): a wave takes a tile of up to 64 test elements (one
// per lane) and sweeps a chunk of the trial element sequence; the trial
// element is wave-uniform (scalar loads).  Each pair uses the diadic product
// of the same Dunavant rule on both elements, degree = field complexity +
// distance interval, or a fixed degree.
//   * constant fields: the pair's entry is one store, as in collocation;
//   * nodal fields: the 3x3 block of a pair belongs to nodes that other pairs
//     share.  Each lane keeps, in LDS private to it, its element's 3 corners x
//     GAL_CS column-node slots (plain read-add-write: shared accumulators with
//     LDS atomics serialise on bank and address conflicts); a host-built
//     schedule assigns the trial corners to slots (least recently used) and
//     names the node a slot held before.  That column is then reduced over the
//     lanes whose elements share a row node (a per-tile contributor table,
//     lane l owning local row l) and added to the result with one global
//     atomic per row.  Every column node of the chunk ends in one or a few
//     such flushes instead of one atomic per pair and entry.
// Singular pairs (FP64, same mesh only): coincident, edge- and vertex-adjacent
// element pairs, one wave per pair over the 512 / 1536 points of the
// Duffy-type rule, rotated to the shared corners as
// nihu/core/singular_accelerator.hpp:137-359 does.

#pragma once

#include "nbem_kernels.hpp"

namespace nbk {

constexpr int GAL_TILE = 64;    // test elements per wave tile
constexpr int GAL_RMAX = 64;    // distinct nodes per wave tile (nodal fields)
constexpr int GAL_CS = 8;       // column-node slots per lane (nodal fields)

struct GalArgs {
	int n_tiles, chunk;          // test wave tiles; trial sequence positions per chunk
	int tn, sn;                  // element counts of test / trial mesh (SoA strides)
	int same_mesh;
	const int *t_elem;           // nodal: [n_tiles][64] test element of each lane (-1: idle)
	const uint2 *t_contrib;      // nodal: [n_tiles][64] contributors lane * 4 + corner of each
	                             //        local row, 8 bits each, 0xff: none
	const int *t_rows;           // nodal: node of each local row (-1: unused)
	const int *t_tri;            // test element nodes [3][tn]
	const int *s_order;          // nodal: trial element at each sequence position
	const int *s_col;            // nodal: [sn][4] previous slot occupant of corners 0..2, packed slots
	const int *s_fin;            // nodal: [chunks][GAL_CS] slot occupants at the end of a chunk
	const int *s_tri;            // trial element nodes [3][sn]
	const double *tc, *sc;       // centres [3][n]
	const float *tnu, *snu;      // unit normals [3][n]
	const double *thr;           // trial distance thresholds [NTHR][sn]
	const QPoint *tq[NSLOT], *sq[NSLOT];
	int qn[NSLOT];
	const float2 *ref;           // [NSLOT][REF_MAXQ] reference points (xi, eta) of each degree's rule
	int fixed_degree;            // >= 0: fixed rule
	float k, k_revf;             // wavenumber, k / (2 pi)
	double k_rev;
	float ki;                    // imaginary part of a complex wavenumber (kernels with CK)
	float2 alpha;
	void *out;
	long long ld;
	int dtype, accumulate;
};

// test-side quadrature point: offset a from the element centre, |a|^2, w J
struct GTest { float ax, ay, az, aa, wj; };

struct GalPair {
	float ex0, ey0, ez0;         // 2 (c_f - c_e)
	float d2f, df, t0;           // |c_f - c_e|^2, |c_f - c_e|, fract(k |c_f - c_e| / 2 pi)
	float nyx, nyy, nyz;         // trial unit normal
	float nxx, nxy, nxz;         // test unit normal
	float dcny, dcnx, nn;        // (c_f - c_e).n_y, (c_f - c_e).n_x, n_x.n_y
	float ad;                    // complex wavenumber k + i ki: e^{ki d}
};

__device__ __forceinline__ void atomic_add_entry(void *out, long long idx, int dtype, float re, float im)
{
	switch (dtype) {
	case DT_C64: {
		float *p = (float *)out + 2 * idx;
		unsafeAtomicAdd(p, re);
		unsafeAtomicAdd(p + 1, im);
		break;
	}
	case DT_C128: {
		double *p = (double *)out + 2 * idx;
		unsafeAtomicAdd(p, (double)re);
		unsafeAtomicAdd(p + 1, (double)im);
		break;
	}
	case DT_F32: unsafeAtomicAdd((float *)out + idx, re); break;
	default: unsafeAtomicAdd((double *)out + idx, (double)re); break;
	}
}

__device__ __forceinline__ void atomic_add_entry64(void *out, long long idx, int dtype, double re, double im)
{
	switch (dtype) {
	case DT_C64: {
		float *p = (float *)out + 2 * idx;
		unsafeAtomicAdd(p, (float)re);
		unsafeAtomicAdd(p + 1, (float)im);
		break;
	}
	case DT_C128: {
		double *p = (double *)out + 2 * idx;
		unsafeAtomicAdd(p, re);
		unsafeAtomicAdd(p + 1, im);
		break;
	}
	case DT_F32: unsafeAtomicAdd((float *)out + idx, (float)re); break;
	default: unsafeAtomicAdd((double *)out + idx, re); break;
	}
}

// One pair over the diadic rule: for each test point p (tp(p), per lane) the
// sums S_b = sum_q K(x_p, y_q) w_q J_f N_b(q) / (4 pi) over the trial points
// (wave-uniform table rows), then block[a][b] += w_p J_e N_a(p) S_b.
// r = y_q - x_p = (c_f - c_e) + (o_q - a_p); with d = |c_f - c_e|,
//   r^2 = d^2 + u,  u = 2 (c_f - c_e - a_p).o_q + |o_q|^2 + |a_p|^2 - 2 (c_f - c_e).a_p
// and the phase k r = k d + k u / (r + d) as in the collocational kernel.
// r.n_y = (c_f - c_e - a_p).n_y (o_q lies in the trial plane) and
// r.n_x = (c_f - c_e + o_q).n_x (a_p lies in the test plane).
// Kernel values: SLP E, DLP -Q r.n_y, DLPt Q r.n_x,
// HSP n_x.n_y Q - (3Q - th^2 G) r.n_x r.n_y / r^2 with E = e^{-i th}/r,
// G = E/r^2, Q = G (1 + i th), th = k r; the 1/(4 pi) rides on the trial weights.
// G comes first and th G = k e^{-i th}/r^2 (no r or th needed); th^2 G = k^2 E.
// DLP's -r.n_y does not depend on q and multiplies the test weight instead.
// Nodal fields: the table rows carry w_q xi_q and w_q eta_q, so the sums are
// T = sum K w (N_0 + N_1 + N_2 = 1), S_1 and S_2, and blk[a][0] receives the
// T sums; the caller turns that column into N_0's, T - S_1 - S_2, after the
// pair.  ref: the rule's reference points (xi_p, eta_p) for the test weights.
// CK: complex wavenumber k + i kim: e = e^{kim r} (c - i sn) (the attenuation
// scales c and sn: every kernel is e times a polynomial in i(k + i kim) r),
// Q gains -kim r G and the (k r)^2 G term takes (k + i kim)^2.
template <int FAM, int KM, int NDOF, int NQ, class TP, bool CK = false>
__device__ __forceinline__ void gal_points(TP tp, const QPoint *sq_in, const float2 *ref, int nq_rt, const GalPair &P,
	float k_revf, float k, float2 (&blk)[NDOF][NDOF], float kim = 0.0f)
{
	const int nq = NQ > 0 ? NQ : nq_rt;
	const auto sf = uni((const float *)sq_in);
	const auto rf = uni((const float *)ref);
	constexpr bool IM = FAM == HELM;
	constexpr int UNROLL = NQ > 0 ? NQ : 1;
#pragma unroll UNROLL
	for (int p = 0; p < nq; ++p) {
		const GTest t = tp(p);
		const float ex = __builtin_fmaf(-2.0f, t.ax, P.ex0);
		const float ey = __builtin_fmaf(-2.0f, t.ay, P.ey0);
		const float ez = __builtin_fmaf(-2.0f, t.az, P.ez0);
		const float vp = t.aa - __builtin_fmaf(P.ex0, t.ax, __builtin_fmaf(P.ey0, t.ay, P.ez0 * t.az));
		float rny = 0.0f;
		if (KM & (DLP | HSP))
			rny = P.dcny - __builtin_fmaf(t.ax, P.nyx, __builtin_fmaf(t.ay, P.nyy, t.az * P.nyz));
		float2 S[NDOF];
#pragma unroll
		for (int b = 0; b < NDOF; ++b) S[b] = zero2();
#pragma unroll UNROLL
		for (int q = 0; q < nq; ++q) {
			const float ox = sf[8 * q], oy = sf[8 * q + 1], oz = sf[8 * q + 2], w = sf[8 * q + 3], oo = sf[8 * q + 4];
			const float u = __builtin_fmaf(ex, ox, __builtin_fmaf(ey, oy, __builtin_fmaf(ez, oz, oo + vp)));
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
				// Q = G (1 + i th): Re = Gr + k sn / r^2, Im = Gi + k c / r^2
				float Qr = IM ? __builtin_fmaf(kh, sn, Gr) : Gr;
				float Qi = IM ? __builtin_fmaf(kh, c, Gi) : 0.0f;
				if (CK) {
					const float kih = kim * ri2;
					Qr = __builtin_fmaf(-kih, c, Qr);
					Qi = __builtin_fmaf(kih, sn, Qi);
				}
				if (KM == DLP) {
					kr = Qr; ki = Qi;
				} else {
					const float rnx = P.dcnx + __builtin_fmaf(ox, P.nxx, __builtin_fmaf(oy, P.nxy, oz * P.nxz));
					if (KM == DLPT) {
						kr = Qr * rnx; ki = Qi * rnx;
					} else {
						const float cf = rnx * rny * ri2;
						// 3 Q - th^2 G with th^2 G = k^2 E = k^2 e^{-i th} / r
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
						kr = __builtin_fmaf(P.nn, Qr, -hr * cf);
						ki = __builtin_fmaf(P.nn, Qi, -hi * cf);
					}
				}
			}
			if (NDOF == 1) {
				S[0].x = __builtin_fmaf(kr, w, S[0].x);
				if (IM) S[0].y = __builtin_fmaf(ki, w, S[0].y);
			} else {
				const float w1 = sf[8 * q + 5], w2 = sf[8 * q + 6];
				S[0].x = __builtin_fmaf(kr, w, S[0].x);
				S[1].x = __builtin_fmaf(kr, w1, S[1].x);
				S[2].x = __builtin_fmaf(kr, w2, S[2].x);
				if (IM) {
					S[0].y = __builtin_fmaf(ki, w, S[0].y);
					S[1].y = __builtin_fmaf(ki, w1, S[1].y);
					S[2].y = __builtin_fmaf(ki, w2, S[2].y);
				}
			}
		}
		const float wt = KM == DLP ? -rny * t.wj : t.wj;
		if (NDOF == 1) {
			blk[0][0].x = __builtin_fmaf(wt, S[0].x, blk[0][0].x);
			if (IM) blk[0][0].y = __builtin_fmaf(wt, S[0].y, blk[0][0].y);
		} else {
			// the test element carries the same rule: N_a at the p-th reference point
			const float xi = rf[2 * p], eta = rf[2 * p + 1];
			const float wa[3] = {wt * (1.0f - xi - eta), wt * xi, wt * eta};
#pragma unroll
			for (int a = 0; a < 3; ++a)
#pragma unroll
				for (int b = 0; b < 3; ++b) {
					blk[a][b].x = __builtin_fmaf(wa[a], S[b].x, blk[a][b].x);
					if (IM) blk[a][b].y = __builtin_fmaf(wa[a], S[b].y, blk[a][b].y);
				}
		}
	}
}

// policy-driven degree per lane; lanes are served one distinct degree at a time
template <int FAM, int KM, int NDOF, bool CK = false>
__device__ __forceinline__ void gal_eval_dynamic(const GalArgs &a, int e, int f, double d2, bool act,
	const GalPair &P, float2 (&blk)[NDOF][NDOF])
{
	constexpr int ORDER = KM == SLP ? 1 : KM == HSP ? 3 : 2;
	int deg = a.fixed_degree >= 0 ? a.fixed_degree : (NDOF == 3 ? 1 : 0) + table_degree<ORDER>(d2, a.thr, a.sn, f);
	deg = act ? deg : -1;
	for (;;) {
		const unsigned long long m = __ballot(deg >= 0);
		if (m == 0) break;
		const int dsel = __builtin_amdgcn_readlane(deg, __ffsll((long long)m) - 1);
		// lanes picked by a ballot mask (not deg == dsel): dsel stays wave-uniform
		const unsigned long long msel = __ballot(deg == dsel);
		if ((msel >> __lane_id()) & 1ull) {
			const int nq = a.qn[dsel];
			const float *tq = (const float *)(a.tq[dsel] + (size_t)e * nq);
			auto tp = [&](int p) { return GTest{tq[8 * p], tq[8 * p + 1], tq[8 * p + 2], tq[8 * p + 4], tq[8 * p + 7]}; };
			gal_points<FAM, KM, NDOF, 0, decltype(tp), CK>(tp, a.sq[dsel] + (size_t)f * nq, a.ref + dsel * REF_MAXQ, nq,
				P, a.k_revf, a.k, blk, a.ki);
			deg = -1;
		}
	}
}

// per-lane accumulator of corner c in column slot s: acc[(s * 3 + c) * 64 + lane]
__device__ __forceinline__ float2 *gal_acc(float2 *acc, int slot, int corner, int lane)
{
	return acc + (slot * 3 + corner) * GAL_TILE + lane;
}

// add one column slot to the result and clear it: lane l sums local row l
// over its contributors, then every lane clears its own three entries
__device__ __forceinline__ void gal_flush(float2 *acc, int lane, int slot, int col, int rnode, uint2 con,
	const GalArgs &a)
{
	float2 v = make_float2(0.0f, 0.0f);
#pragma unroll
	for (int q = 0; q < 8; ++q) {
		const unsigned c = ((q < 4 ? con.x : con.y) >> (8 * (q & 3))) & 255u;
		if (c != 255u) {
			const float2 t = *gal_acc(acc, slot, c & 3u, c >> 2);
			v.x += t.x;
			v.y += t.y;
		}
	}
#pragma unroll
	for (int c = 0; c < 3; ++c) *gal_acc(acc, slot, c, lane) = make_float2(0.0f, 0.0f);
	if (rnode >= 0 && (v.x != 0.0f || v.y != 0.0f)) {
		const float re = a.alpha.x * v.x - a.alpha.y * v.y;
		const float im = a.alpha.x * v.y + a.alpha.y * v.x;
		atomic_add_entry(a.out, (long long)col * a.ld + rnode, a.dtype, re, im);
	}
}

// FIXQ > 0: fixed rule with FIXQ points per element (test points held in
// registers); 0: any policy.  Grid: (test tiles / 4) x chunks, 256 lanes.
template <int FAM, int KM, int NDOF, int FIXQ, bool CK = false>
__global__ void __launch_bounds__(256) galerkin_regular_kernel(GalArgs a)
{
	__shared__ float2 acc_all[NDOF == 3 ? 4 * GAL_CS * 3 * GAL_TILE : 1];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int wg_tiles = (a.n_tiles + 3) >> 2;
	const int tile = (int)(blockIdx.x % wg_tiles) * 4 + wave;
	const int chunk = (int)(blockIdx.x / wg_tiles);
	if (tile >= a.n_tiles) return;     // no block-wide barriers below
	float2 *acc = acc_all + (NDOF == 3 ? wave * GAL_CS * 3 * GAL_TILE : 0);

	int e;
	if (NDOF == 3) e = a.t_elem[tile * GAL_TILE + lane];
	else e = tile * GAL_TILE + lane < a.tn ? tile * GAL_TILE + lane : -1;
	const bool lane_ok = e >= 0;
	const int ee = lane_ok ? e : (NDOF == 3 ? a.t_elem[tile * GAL_TILE] : tile * GAL_TILE);
	const double cix = a.tc[ee], ciy = a.tc[a.tn + ee], ciz = a.tc[2 * a.tn + ee];
	GalPair P;
	P.nxx = P.nxy = P.nxz = 0.0f;
	if (KM & (DLPT | HSP)) { P.nxx = a.tnu[ee]; P.nxy = a.tnu[a.tn + ee]; P.nxz = a.tnu[2 * a.tn + ee]; }
	int tv0 = -1, tv1 = -1, tv2 = -1;
	if (a.same_mesh) { tv0 = a.t_tri[ee]; tv1 = a.t_tri[a.tn + ee]; tv2 = a.t_tri[2 * a.tn + ee]; }
	int rnode = -1;
	uint2 con = make_uint2(0xffffffffu, 0xffffffffu);
	if (NDOF == 3) {
		con = a.t_contrib[tile * GAL_TILE + lane];
		rnode = a.t_rows[tile * GAL_TILE + lane];
#pragma unroll
		for (int s = 0; s < GAL_CS; ++s)
#pragma unroll
			for (int c = 0; c < 3; ++c) *gal_acc(acc, s, c, lane) = make_float2(0.0f, 0.0f);
	}
	GTest tpr[FIXQ > 0 ? FIXQ : 1];
	if (FIXQ > 0) {
		const float *q = (const float *)(a.tq[a.fixed_degree] + (size_t)ee * FIXQ);
#pragma unroll
		for (int p = 0; p < FIXQ; ++p) tpr[p] = GTest{q[8 * p], q[8 * p + 1], q[8 * p + 2], q[8 * p + 4], q[8 * p + 7]};
	}
	const int i0 = chunk * a.chunk;
	const int i1 = min(i0 + a.chunk, a.sn);
	const auto sc = uni(a.sc);
	const auto snu = uni(a.snu);
	const auto stri = uni(a.s_tri);
	const auto sord = uni(a.s_order);
	const auto scol = uni(a.s_col);
	for (int i = i0; i < i1; ++i) {
		const int f = NDOF == 3 ? sord[i] : i;
		int sl0 = 0, sl1 = 0, sl2 = 0;
		if (NDOF == 3) {
			const int v0 = scol[4 * i], v1 = scol[4 * i + 1], v2 = scol[4 * i + 2], sl = scol[4 * i + 3];
			sl0 = sl & 31; sl1 = (sl >> 5) & 31; sl2 = (sl >> 10) & 31;
			if (v0 >= 0) gal_flush(acc, lane, sl0, v0, rnode, con, a);
			if (v1 >= 0) gal_flush(acc, lane, sl1, v1, rnode, con, a);
			if (v2 >= 0) gal_flush(acc, lane, sl2, v2, rnode, con, a);
		}
		const double cjx = sc[f], cjy = sc[a.sn + f], cjz = sc[2 * a.sn + f];
		const double d2 = d2_pair(cix, ciy, ciz, cjx, cjy, cjz);
		bool sing = false;
		if (a.same_mesh) {
			const int f0 = stri[f], f1 = stri[a.sn + f], f2 = stri[2 * a.sn + f];
			sing = tv0 == f0 || tv0 == f1 || tv0 == f2 || tv1 == f0 || tv1 == f1 || tv1 == f2 ||
				tv2 == f0 || tv2 == f1 || tv2 == f2;
		}
		const bool act = lane_ok && !sing;
		float2 blk[NDOF][NDOF];
#pragma unroll
		for (int x = 0; x < NDOF; ++x)
#pragma unroll
			for (int y = 0; y < NDOF; ++y) blk[x][y] = zero2();
		if (__ballot(act) != 0) {
			const float dcx = (float)(cjx - cix), dcy = (float)(cjy - ciy), dcz = (float)(cjz - ciz);
			P.ex0 = 2.0f * dcx; P.ey0 = 2.0f * dcy; P.ez0 = 2.0f * dcz;
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
			P.nyx = P.nyy = P.nyz = 0.0f;
			P.dcny = P.dcnx = P.nn = 0.0f;
			if (KM & (DLP | HSP)) {
				P.nyx = snu[f]; P.nyy = snu[a.sn + f]; P.nyz = snu[2 * a.sn + f];
				P.dcny = dcx * P.nyx + dcy * P.nyy + dcz * P.nyz;
			}
			if (KM & (DLPT | HSP)) P.dcnx = dcx * P.nxx + dcy * P.nxy + dcz * P.nxz;
			if (KM & HSP) P.nn = P.nxx * P.nyx + P.nxy * P.nyy + P.nxz * P.nyz;
			if (FIXQ > 0) {
				if (act) {
					auto tp = [&](int p) { return tpr[p]; };
					gal_points<FAM, KM, NDOF, FIXQ, decltype(tp), CK>(tp, a.sq[a.fixed_degree] + (size_t)f * FIXQ,
						a.ref + a.fixed_degree * REF_MAXQ, FIXQ, P, a.k_revf, a.k, blk, a.ki);
				}
			} else {
				gal_eval_dynamic<FAM, KM, NDOF, CK>(a, ee, f, d2, act, P, blk);
			}
		}
		if (NDOF == 1) {
			// singular entries: zero here, the FP64 pass adds them
			if (lane_ok && (act || !a.accumulate)) {
				const float re = a.alpha.x * blk[0][0].x - a.alpha.y * blk[0][0].y;
				const float im = a.alpha.x * blk[0][0].y + a.alpha.y * blk[0][0].x;
				store_entry(a.out, (long long)f * a.ld + e, a.dtype, a.accumulate, re, im);
			}
		} else if (act) {
			// column 0 holds the sums with the trial weight w (T); N_0's is T - S_1 - S_2
#pragma unroll
			for (int x = 0; x < 3; ++x) {
				blk[x][0].x -= blk[x][1].x + blk[x][2].x;
				blk[x][0].y -= blk[x][1].y + blk[x][2].y;
			}
			const int cl[3] = {sl0, sl1, sl2};
#pragma unroll
			for (int x = 0; x < 3; ++x)
#pragma unroll
				for (int y = 0; y < 3; ++y) {
					float2 *p = gal_acc(acc, cl[y], x, lane);
					float2 v = *p;
					v.x += blk[x][y].x;
					v.y += blk[x][y].y;
					*p = v;
				}
		}
	}
	if (NDOF == 3) {
		const auto sfin = uni(a.s_fin);
		for (int s = 0; s < GAL_CS; ++s) {
			const int v = sfin[chunk * GAL_CS + s];
			if (v >= 0) gal_flush(acc, lane, s, v, rnode, con, a);
		}
	}
}

// ---------------------------------------------------------------- FP64 pass

// full kernel value in FP64, r = y - x (nihu/library/helmholtz_kernel.hpp:169-228,
// laplace_kernel.hpp:133-171, normal_derivative_kernel.hpp:114-257)
template <int FAM, int KM>
__device__ __forceinline__ double2 kernel_value64(D3 x, D3 y, D3 nx, D3 ny, double k)
{
	const D3 rv = d3sub(y, x);
	const double r2 = d3dot(rv, rv);
	const double rinv = rsqrt64(r2);
	const double r = r2 * rinv;
	double sn = 0.0, cs = 1.0, kr = 0.0;
	if (FAM == HELM) {
		kr = k * r;
		sincos64(kr, &sn, &cs);
	}
	const double f = 1.0 / (4.0 * PI_D);
	// e^{-ikr} = cs - i sn
	const double g1 = f * rinv;
	if (KM == SLP) return make_double2(cs * g1, -sn * g1);
	const double rdny = d3dot(rv, ny) * rinv, rdnx = -d3dot(rv, nx) * rinv;
	const double g3 = g1 * rinv * rinv;
	if (KM == DLP || KM == DLPT) {
		// e (-1 - ikr) / (4 pi r^2) (r.n_y / r  or  -r.n_x / r)
		const double g = g3 * r * (KM == DLP ? rdny : rdnx);
		return make_double2((-cs - kr * sn) * g, (sn - kr * cs) * g);
	}
	// e / (4 pi r^3) [(3 + 3ikr - (kr)^2) rdnx rdny + (1 + ikr) n_x.n_y]
	const double cc = rdnx * rdny, nn = d3dot(nx, ny);
	const double ar = (3.0 - kr * kr) * cc + nn, ai = kr * (3.0 * cc + nn);
	return make_double2((cs * ar + sn * ai) * g3, (cs * ai - sn * ar) * g3);
}

// the same for a complex wavenumber k + i ki
template <int KM>
__device__ __forceinline__ double2 kernel_value64_c(D3 x, D3 y, D3 nx, D3 ny, double k, double ki)
{
	const D3 rv = d3sub(y, x);
	const double r2 = d3dot(rv, rv);
	const double rinv = rsqrt64(r2);
	double2 e, z;
	wave64(k, ki, r2 * rinv, &e, &z);
	double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
	wave_terms_c<KM>(e, z, (1.0 / (4.0 * PI_D)) * rinv, rinv, d3dot(rv, ny) * rinv, -d3dot(rv, nx) * rinv,
		(KM & HSP) ? d3dot(nx, ny) : 0.0, 0.0, s);
	return s[KM == SLP ? 0 : KM == DLP ? 1 : KM == DLPT ? 2 : 3];
}

struct GalSingArgs {
	int n_pairs;
	const int *pe, *pf, *pinfo;  // test / trial element, match | ind1 << 4 | ind2 << 8
	int n;                       // elements (SoA stride)
	const double *X;             // vertices [9][n]
	const double *J;             // Jacobians
	const double *nu;            // unit normals [3][n]
	const int *t_dof, *s_dof;    // row / column dof of each element corner [3][n] (nodes,
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

// One wave per singular pair.  The rule's corner k of the test element is the
// element's corner (ind1 + k) % 3; of the trial element (ind2 + 1 - k) mod 3
// for a common edge (it runs the other way) and (ind2 + k) % 3 for a common
// vertex.  The vertex coordinates are permuted accordingly, so that the
// shape function N_k at a rule point belongs to the permuted corner k.
template <int FAM, int KM, int NDOF, bool CK = false>
__global__ void __launch_bounds__(64) galerkin_singular_kernel(GalSingArgs a)
{
	const int p = blockIdx.x, lane = threadIdx.x;
	if (p >= a.n_pairs) return;
	const int e = a.pe[p], f = a.pf[p], info = a.pinfo[p];
	const int match = info & 15, i1 = (info >> 4) & 15, i2 = (info >> 8) & 15;
	int ct[3], cs[3];
	for (int c = 0; c < 3; ++c) {
		ct[c] = match == 2 ? c : (i1 + c) % 3;
		cs[c] = match == 2 ? c : match == 1 ? (i2 + 4 - c) % 3 : (i2 + c) % 3;
	}
	const int n = a.n;
	D3 Vt[3], Vs[3];
	for (int c = 0; c < 3; ++c) {
		Vt[c] = {a.X[(3 * ct[c]) * n + e], a.X[(3 * ct[c] + 1) * n + e], a.X[(3 * ct[c] + 2) * n + e]};
		Vs[c] = {a.X[(3 * cs[c]) * n + f], a.X[(3 * cs[c] + 1) * n + f], a.X[(3 * cs[c] + 2) * n + f]};
	}
	const D3 nx = {a.nu[e], a.nu[n + e], a.nu[2 * n + e]};
	const D3 ny = {a.nu[f], a.nu[n + f], a.nu[2 * n + f]};
	const double Je = a.J[e], Jf = a.J[f];
	const double *R = a.rule[match];
	const int nr = a.rule_n[match];
	double acc[NDOF * NDOF * 2];
	for (int q = 0; q < NDOF * NDOF * 2; ++q) acc[q] = 0.0;
	for (int q = lane; q < nr; q += 64) {
		const double txi = R[5 * q], teta = R[5 * q + 1], sxi = R[5 * q + 2], seta = R[5 * q + 3], w = R[5 * q + 4];
		const D3 x = tri_x(Vt, txi, teta), y = tri_x(Vs, sxi, seta);
		const double2 K = CK ? kernel_value64_c<KM>(x, y, nx, ny, a.k, a.ki) : kernel_value64<FAM, KM>(x, y, nx, ny, a.k);
		const double wt = Je * w, ws = Jf * 1.0;
		if (NDOF == 1) {
			acc[0] += wt * K.x * ws;
			acc[1] += wt * K.y * ws;
		} else {
			const double Nt[3] = {1.0 - txi - teta, txi, teta}, Ns[3] = {1.0 - sxi - seta, sxi, seta};
			for (int x3 = 0; x3 < 3; ++x3)
				for (int y3 = 0; y3 < 3; ++y3) {
					const double ta = Nt[x3] * wt, sb = Ns[y3] * ws;
					acc[2 * (3 * x3 + y3)] += ta * K.x * sb;
					acc[2 * (3 * x3 + y3) + 1] += ta * K.y * sb;
				}
		}
	}
	for (int q = 0; q < NDOF * NDOF * 2; ++q)
		for (int off = 32; off >= 1; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
	if (lane != 0) return;
	for (int x3 = 0; x3 < NDOF; ++x3)
		for (int y3 = 0; y3 < NDOF; ++y3) {
			const double vr = acc[2 * (NDOF * x3 + y3)], vi = acc[2 * (NDOF * x3 + y3) + 1];
			const double re = a.alpha.x * vr - a.alpha.y * vi, im = a.alpha.x * vi + a.alpha.y * vr;
			if (NDOF == 1) {
				store_entry64(a.out, (long long)f * a.ld + e, a.dtype, 1, re, im);
			} else {
				const int row = a.t_dof[ct[x3] * n + e], col = a.s_dof[cs[y3] * n + f];
				atomic_add_entry64(a.out, (long long)col * a.ld + row, a.dtype, re, im);
			}
		}
}

// identity (mass matrix): constant fields J/2 on the diagonal; nodal fields
// J/12 (a = b) and J/24 (a != b) per element, the values of the degree-2
// Dunavant rule of nihu/core/single_integral.hpp:65-140; through dof maps
// that name the element at every corner of a constant side, the sums are
// J/6, the integral of a linear shape function (exact, as NiHu's degree-1
// rule for a constant test and a linear trial)
__global__ void galerkin_identity_kernel(void *out, long long ld, int n, const double *J, const int *t_map,
	const int *s_map, int nodal, int dtype, double ar, double ai)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n) return;
	const double j = J[e];
	if (!nodal) {
		store_entry64(out, (long long)e * ld + e, dtype, 1, ar * (j * 0.5), ai * (j * 0.5));
		return;
	}
	for (int x = 0; x < 3; ++x)
		for (int y = 0; y < 3; ++y) {
			const double m = x == y ? j / 12.0 : j / 24.0;
			atomic_add_entry64(out, (long long)s_map[y * n + e] * ld + t_map[x * n + e], dtype, ar * m, ai * m);
		}
}

} // namespace nbk
