// nbem_kernels.hpp -- HIP kernels of the dense BEM assembly engine (gfx950).
//
// Data layout in HBM (per mesh, structure of arrays, n = number of triangles):
//   centres   c[3][n]  FP64        unit normals nu[3][n] FP32 and FP64
//   vertices  X[9][n]  FP64        Jacobians J[n] FP64
//   thresholds thr[NTHR][n] FP64: squared-distance limits per trial element
//              (slot 0 near-singular, slots 1.. the distance-table breakpoints)
//   quadrature tables per Dunavant degree: QPoint[n][Q] FP32 offsets of the
//              points from the element centre, weight*J/(4 pi), |offset|^2
// Result matrices are column-major: entry (i, j) at out[j * ld + i].
//
// Regular pairs: one lane per test row i, a workgroup of 256 lanes sweeps a
// tile of trial columns; the column index is wave-uniform so trial data come
// through scalar loads and a column's 64 results leave as one 512-byte store.
// Near-singular and diagonal pairs: one lane per pair in FP64.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbk {

enum { SLP = 1, DLP = 2, DLPT = 4, HSP = 8, ALL4 = 15 };
enum { LAP = 0, HELM = 1 };
enum { OUT_SINGLE = 0, OUT_BM = 1 };
enum { DT_C64 = 0, DT_C128 = 1, DT_F32 = 2, DT_F64 = 3 };

// squared-distance threshold slots; the breakpoint values are those of the
// accuracy-3 distance tables (nihu/library/distance_kernel_intervals.hpp:116-155)
enum { THR_NEAR = 0, THR_1_0, THR_1_1, THR_1_4, THR_1_5, THR_1_8, THR_1_9, THR_2_8, THR_3_6, THR_9_1,
	THR_15_8, THR_22_3, NTHR };

constexpr int NSLOT = 16;   // quadrature tables indexed by Dunavant degree 0..14 (+1 spare)
constexpr int REF_MAXQ = 48;   // points of the largest Dunavant rule used (degree 14: 42), rounded up

// offset from the element centre, w J / (4 pi), |offset|^2, the same weight
// times the reference coordinates xi and eta (the linear shape functions N_1,
// N_2 of the Galerkin pass), w J
struct QPoint { float ox, oy, oz, w, oo, wxi, weta, wj; };

// Per trial element, read as one 160-byte block through the scalar cache by
// the regular pass: distance thresholds, centre, unit normal and the weight
// 0.5 J / (4 pi) of the one-point (centroid) Dunavant rule of degrees 0 and 1.
struct TrialRec {
	double thr[NTHR];
	double c[3];
	float nu[3];
	float w1;
	float pad[6];
};
static_assert(sizeof(TrialRec) == 160, "TrialRec layout");

struct RegArgs {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	const int *rmap, *cmap;      // result row / column of each test / trial element (null: its own index)
	int tn, sn;                  // element counts of the test / trial mesh (SoA strides)
	const double *tc;            // test centres [3][tn]
	const float *tnu;            // test unit normals [3][tn]
	const double *sc;            // trial centres [3][sn]
	const float *snu;            // trial unit normals [3][sn]
	const double *thr;           // [NTHR][sn]
	const TrialRec *rec;         // [sn]
	const QPoint *qt[NSLOT];     // per degree
	int qn[NSLOT];
	int fixed_degree;            // >= 0: fixed rule
	int field_complexity;        // added to the table degree (0 for constant fields)
	float k, k_revf;             // wavenumber, k / (2 pi)
	double k_rev;
	float ki;                    // imaginary part of a complex wavenumber (kernels with CK)
	float2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

// ---------------------------------------------------------------- helpers

// Empty sum for FP32 accumulators: -0 is the additive identity (x + -0 = x for
// every x, while 0 + -0 = +0), so the compiler may drop the first addition of
// each sum (fma(a, b, -0) is a multiplication); the sums are otherwise the same.
__device__ __forceinline__ float2 zero2() { return make_float2(-0.0f, -0.0f); }

// View of read-only global memory in the constant address space: loads from a
// wave-uniform address through it become scalar loads (SGPRs, scalar cache).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *uni(const T *p)
{
	return (const __attribute__((address_space(4))) T *)(size_t)p;
}

// The row's data (loaded once, before the column loop) are consumed here, so
// that the wait for their loads happens once, before the loop.  Vector stores
// count in the same counter as vector loads: a first use of these registers
// inside the loop makes every iteration wait until the previous columns'
// stores have drained.
__device__ __forceinline__ void row_ready(double x, double y, double z, float a, float b, float c, int r)
{
	__asm__ volatile("" ::"v"(x), "v"(y), "v"(z), "v"(a), "v"(b), "v"(c), "v"(r));
}

__device__ __forceinline__ double d2_pair(double ax, double ay, double az, double bx, double by, double bz)
{
	double dx = ax - bx, dy = ay - by, dz = az - bz;
	return __fma_rn(dz, dz, __fma_rn(dy, dy, dx * dx));
}

// Entry (row, col) of a column-major result as base + col (wave-uniform, in
// elements of the dtype) + row (per lane): the stores then take a scalar base
// address and a 32-bit lane offset
__device__ __forceinline__ unsigned dtype_size(int dtype)
{
	return dtype == DT_C128 ? 16u : dtype == DT_F32 ? 4u : 8u;
}

// column base (wave-uniform) plus the lane's byte offset as a 32-bit value
// (rows x 16 bytes < 4 GiB), one offset for every dtype
template <class T>
__device__ __forceinline__ T *entry_ptr(void *out, long long col, unsigned row_bytes)
{
	return (T *)((char *)out + col * (long long)sizeof(T) + row_bytes);
}

__device__ __forceinline__ void store_entry(void *out, long long idx, int dtype, int accumulate, float re, float im)
{
	switch (dtype) {
	case DT_C64: {
		float2 *p = (float2 *)out + idx;
		if (accumulate) { float2 v = *p; v.x += re; v.y += im; *p = v; }
		else *p = make_float2(re, im);
		break;
	}
	case DT_C128: {
		double2 *p = (double2 *)out + idx;
		if (accumulate) { double2 v = *p; v.x += re; v.y += im; *p = v; }
		else *p = make_double2(re, im);
		break;
	}
	case DT_F32: {
		float *p = (float *)out + idx;
		if (accumulate) *p += re; else *p = re;
		break;
	}
	default: {
		double *p = (double *)out + idx;
		if (accumulate) *p += re; else *p = re;
		break;
	}
	}
}

// Both Burton-Miller entries of one pair (A to out0, B to out1) under one
// dtype branch, so that neither value is computed inside a store branch (the
// compiler would otherwise sink the B sums past the A store and keep the
// per-point terms alive, costing registers and occupancy).  With accumulate
// the A entry is written before the B entry is read, as two single stores would.
__device__ __forceinline__ void store_pair(void *out0, void *out1, long long col, int row, int dtype, int accumulate,
	float ar, float ai, float br, float bi)
{
	const unsigned rb = (unsigned)row * dtype_size(dtype);
	switch (dtype) {
	case DT_C64: {
		float2 *p = entry_ptr<float2>(out0, col, rb), *q = entry_ptr<float2>(out1, col, rb);
		if (accumulate) {
			const float2 u = *p;
			*p = make_float2(u.x + ar, u.y + ai);
			const float2 v = *q;
			*q = make_float2(v.x + br, v.y + bi);
		} else {
			*p = make_float2(ar, ai);
			*q = make_float2(br, bi);
		}
		break;
	}
	case DT_C128: {
		double2 *p = entry_ptr<double2>(out0, col, rb), *q = entry_ptr<double2>(out1, col, rb);
		if (accumulate) {
			const double2 u = *p;
			*p = make_double2(u.x + ar, u.y + ai);
			const double2 v = *q;
			*q = make_double2(v.x + br, v.y + bi);
		} else {
			*p = make_double2(ar, ai);
			*q = make_double2(br, bi);
		}
		break;
	}
	case DT_F32: {
		float *p = entry_ptr<float>(out0, col, rb), *q = entry_ptr<float>(out1, col, rb);
		if (accumulate) { *p += ar; *q += br; } else { *p = ar; *q = br; }
		break;
	}
	default: {
		double *p = entry_ptr<double>(out0, col, rb), *q = entry_ptr<double>(out1, col, rb);
		if (accumulate) { *p += ar; *q += br; } else { *p = ar; *q = br; }
		break;
	}
	}
}

__device__ __forceinline__ void store_entry64(void *out, long long idx, int dtype, int accumulate, double re, double im)
{
	switch (dtype) {
	case DT_C64: {
		float2 *p = (float2 *)out + idx;
		if (accumulate) { float2 v = *p; *p = make_float2((float)(v.x + re), (float)(v.y + im)); }
		else *p = make_float2((float)re, (float)im);
		break;
	}
	case DT_C128: {
		double2 *p = (double2 *)out + idx;
		if (accumulate) { double2 v = *p; v.x += re; v.y += im; *p = v; }
		else *p = make_double2(re, im);
		break;
	}
	case DT_F32: {
		float *p = (float *)out + idx;
		if (accumulate) *p = (float)(*p + re); else *p = (float)re;
		break;
	}
	default: {
		double *p = (double *)out + idx;
		if (accumulate) *p += re; else *p = re;
		break;
	}
	}
}

// Regular Dunavant degree of one pair from the squared-distance thresholds of
// the trial element: the first breakpoint x with d/R < x (nihu/tmp/interval.hpp:173-185).
template <int ORDER>
__device__ __forceinline__ int table_degree(double d2, const double *thr, int sn, int j)
{
	if (ORDER == 1) {          // inverse<1>: 1.1 -> 6, 1.8 -> 4, 9.1 -> 2, else 0
		if (d2 < thr[THR_1_1 * sn + j]) return 6;
		if (d2 < thr[THR_1_8 * sn + j]) return 4;
		if (d2 < thr[THR_9_1 * sn + j]) return 2;
		return 0;
	} else if (ORDER == 2) {   // inverse<2>: 1.1 -> 8, 1.5 -> 6, 2.8 -> 4, 15.8 -> 2, else 0
		if (d2 < thr[THR_1_1 * sn + j]) return 8;
		if (d2 < thr[THR_1_5 * sn + j]) return 6;
		if (d2 < thr[THR_2_8 * sn + j]) return 4;
		if (d2 < thr[THR_15_8 * sn + j]) return 2;
		return 0;
	} else {                   // inverse<3>: 1.0, 1.1 -> 9, 1.4 -> 8, 1.9 -> 6, 3.6 -> 4, 22.3 -> 2
		if (d2 < thr[THR_1_1 * sn + j]) return 9;
		if (d2 < thr[THR_1_4 * sn + j]) return 8;
		if (d2 < thr[THR_1_9 * sn + j]) return 6;
		if (d2 < thr[THR_3_6 * sn + j]) return 4;
		if (d2 < thr[THR_22_3 * sn + j]) return 2;
		return 0;
	}
}

struct PairCtx {
	float dcx, dcy, dcz;       // c_j - c_i (FP32)
	float d2f, df, t0;         // |c_j - c_i|^2, |c_j - c_i|, fract(k |c_j - c_i| / 2pi) in [-1/2, 1/2]
	float nxx, nxy, nxz;       // test unit normal
	float b0;                  // (c_j - c_i) . n_x
	float ad;                  // complex wavenumber k + i ki: e^{ki d} (else unused)
};

// Per-pair accumulators of the regular pass (complex):
//   s[0] = sum E                    SLP   = s[0]
//   s[1] = sum Q   (DLP group)      DLP   = -a s[1]
//   s[2] = sum Q b                  DLPt  = s[2]
//   s[3] = sum Q b/r^2              HSP   = nx.ny sum Q - a (3 s[3] - k^2 s[5])
//   s[4] = sum Q   (HSP group, when it is evaluated apart from DLP)
//   s[5] = sum G b
// with, per quadrature point y = c_j + off and r = y - x (|r| = r):
//   E = w e^{-i th} / r,  G = E / r^2,  Q = G (1 + i th),  th = k r,
//   a = r.n_y = (c_j - x).n_y (off lies in the trial plane), b = r.n_x,
// which restates SLP e/(4 pi r), DLP e(-1-ikr)/(4 pi r^2) (r.ny)/r,
// DLPt e(-1-ikr)/(4 pi r^2) (-r.nx)/r and
// HSP e/(4 pi r^3)[(3+3ikr-(kr)^2) rdnx rdny + (1+ikr) nx.ny]
// (nihu/library/helmholtz_kernel.hpp:169-228, normal_derivative_kernel.hpp:114-257);
// since th^2 G / r^2 = k^2 G, the (kr)^2 term is k^2 times a sum of G b;
// the weights w carry J w_q / (4 pi).
// r^2 = d^2 + u with u = 2 (c_j - x).off + |off|^2, and the phase is
// t0 + (k/2pi) u / (r + d) revolutions: t0 is the pair's FP64-reduced phase of
// k d and u / (r + d) = r - d keeps full FP32 relative accuracy; sin and cos
// are the hardware instructions, which take revolutions.
// Add one point to the sums from e^{-i th} = c - i sn, k, w, 1/r, r^2 and
// b = r.n_x.  With a derivative kernel in MASK, G = w/r^3 e^{-i th} comes
// first, E = G r^2 is folded into the SLP sum's FMA and th G = k (w/r^2) e^{-i th}
// needs neither r nor th.
// CK: complex wavenumber k + i ki.  The caller folds the attenuation e^{ki r}
// into w, and Q = G (1 + i(k + i ki) r) gains -ki r G = -ki h e:
// Re Q = Gr + k h sn - ki h c, Im Q = Gi + k h c + ki h sn.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void add_point(float c, float sn, float k, float w, float rinv, float r2, float b,
	float2 (&s)[6], float ki = 0.0f)
{
	// Laplace (c = 1, sn = 0): the imaginary sums stay zero
	constexpr bool IM = FAM == HELM;
	if (!(MASK & (DLP | DLPT | HSP))) {
		const float g = w * rinv;
		s[0].x = __builtin_fmaf(g, c, s[0].x);
		if (IM) s[0].y = __builtin_fmaf(-g, sn, s[0].y);
		return;
	}
	const float ri2 = rinv * rinv;
	const float h = w * ri2;
	const float g3 = h * rinv;
	const float Gr = g3 * c, Gi = -g3 * sn;
	// Q = G (1 + i th): Re = Gr - th Gi = Gr + k h sn, Im = Gi + th Gr = Gi + k h c
	const float kh = k * h;
	float Qr = IM ? __builtin_fmaf(kh, sn, Gr) : Gr;
	float Qi = IM ? __builtin_fmaf(kh, c, Gi) : 0.0f;
	if (CK) {
		const float kih = ki * h;
		Qr = __builtin_fmaf(-kih, c, Qr);
		Qi = __builtin_fmaf(kih, sn, Qi);
	}
	if (MASK & SLP) {
		s[0].x = __builtin_fmaf(Gr, r2, s[0].x);
		if (IM) s[0].y = __builtin_fmaf(Gi, r2, s[0].y);
	}
	if (MASK & DLP) { s[1].x += Qr; if (IM) s[1].y += Qi; }
	if ((MASK & HSP) && !(MASK & DLP)) { s[4].x += Qr; if (IM) s[4].y += Qi; }
	if (MASK & DLPT) {
		s[2].x = __builtin_fmaf(Qr, b, s[2].x);
		if (IM) s[2].y = __builtin_fmaf(Qi, b, s[2].y);
	}
	if (MASK & HSP) {
		const float w2 = b * ri2;
		s[3].x = __builtin_fmaf(Qr, w2, s[3].x);
		if (IM) {
			s[3].y = __builtin_fmaf(Qi, w2, s[3].y);
			s[5].x = __builtin_fmaf(Gr, b, s[5].x);
			s[5].y = __builtin_fmaf(Gi, b, s[5].y);
		}
	}
}

// NQ > 0: compile-time point count (fully unrolled).
template <int FAM, int MASK, int NQ, bool CK = false>
__device__ __forceinline__ void accumulate_points(const QPoint *qp_in, int nq_rt, const PairCtx &p,
	float k_revf, float k, float2 (&s)[6], float ki = 0.0f)
{
	const int nq = NQ > 0 ? NQ : nq_rt;
	const auto qf = uni((const float *)qp_in);
	const float dcx2 = 2.0f * p.dcx, dcy2 = 2.0f * p.dcy, dcz2 = 2.0f * p.dcz;
	constexpr int UNROLL = NQ > 0 ? NQ : 1;
#pragma unroll UNROLL
	for (int q = 0; q < nq; ++q) {
		const float ox = qf[8 * q], oy = qf[8 * q + 1], oz = qf[8 * q + 2], w = qf[8 * q + 3], oo = qf[8 * q + 4];
		const float u = __builtin_fmaf(dcx2, ox, __builtin_fmaf(dcy2, oy, __builtin_fmaf(dcz2, oz, oo)));
		const float r2 = p.d2f + u;
		const float rinv = __builtin_amdgcn_rsqf(r2);
		float c = 1.0f, sn = 0.0f, wa = w;
		if (FAM == HELM) {
			const float rd = u * __builtin_amdgcn_rcpf(__builtin_fmaf(r2, rinv, p.df));   // r - d
			const float t = __builtin_fmaf(k_revf, rd, p.t0);
			c = __builtin_amdgcn_cosf(t);
			sn = __builtin_amdgcn_sinf(t);
			// e^{ki r} = e^{ki d} 2^{ki log2(e) (r - d)}
			if (CK) wa = w * (p.ad * __builtin_amdgcn_exp2f(ki * 1.4426950408889634f * rd));
		}
		float b = 0.0f;
		if (MASK & (DLPT | HSP))
			b = __builtin_fmaf(oz, p.nxz, __builtin_fmaf(oy, p.nxy, __builtin_fmaf(ox, p.nxx, p.b0)));
		add_point<FAM, MASK, CK>(c, sn, k, wa, rinv, r2, b, s, ki);
	}
}

// One-point (centroid) rule: y = c_j, so r = d, the phase is the pair's
// FP64-reduced k d and b = (c_j - x).n_x; rinv_d = 1/d, w1 = 0.5 J / (4 pi).
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void accumulate_centroid(const PairCtx &p, float w1, float rinv_d, float k,
	float2 (&s)[6], float ki = 0.0f)
{
	float c = 1.0f, sn = 0.0f;
	if (FAM == HELM) {
		c = __builtin_amdgcn_cosf(p.t0);
		sn = __builtin_amdgcn_sinf(p.t0);
	}
	add_point<FAM, MASK, CK>(c, sn, k, CK ? w1 * p.ad : w1, rinv_d, p.d2f, p.b0, s, ki);
}

// Evaluate the kernels of MASK for the active lanes, each lane at its own
// quadrature degree (deg < 0: lane idle); lanes are served one distinct
// degree at a time so that the table pointer stays wave-uniform.  One-point
// rules (degrees 0 and 1) take an unrolled path.
// Degrees 0 and 1 are the one-point centroid rule (Dunavant; the host checks
// its tables for this) and take accumulate_centroid: no table load.
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void eval_degrees(const RegArgs &a, int j, int deg, const PairCtx &p, float w1,
	float rinv_d, float k, float k_revf, float2 (&s)[6], float ki = 0.0f)
{
	if (__ballot(deg >= 0 && deg <= 1) != 0) {
		if (deg >= 0 && deg <= 1) accumulate_centroid<FAM, MASK, CK>(p, w1, rinv_d, k, s, ki);
		deg = deg <= 1 ? -1 : deg;
	}
	// wave-uniform guard: waves without a multi-point lane skip the loop and
	// the accumulator moves it costs
	if (__ballot(deg >= 0) == 0) return;
	for (;;) {
		const unsigned long long m = __ballot(deg >= 0);
		if (m == 0) break;
		const int lane = __ffsll((long long)m) - 1;
		const int dsel = __builtin_amdgcn_readlane(deg, lane);
		// lanes picked by a ballot mask rather than by deg == dsel, so that the
		// compiler keeps dsel (and the table address) wave-uniform
		const unsigned long long msel = __ballot(deg == dsel);
		if ((msel >> __lane_id()) & 1ull) {
			const int nq = a.qn[dsel];
			accumulate_points<FAM, MASK, 0, CK>(a.qt[dsel] + (size_t)j * nq, nq, p, k_revf, k, s, ki);
			deg = -1;
		}
	}
}

// The kernel groups of one pair may need different degrees (SLP, DLP / DLPt,
// HSP each from its own distance table).  The distinct degrees present in the
// wave are taken one at a time (wave-uniform, so the table address stays
// uniform): the rule's points are evaluated once with every kernel of KM into
// temporary sums, and each group whose degree for this lane is the current
// one takes its sums by a select.  Every group of a lane is served exactly
// once, so its sums are the rule's sums from -0 in point order, as a
// separate evaluation per group would give them.  Degree < 0: group idle.
// s[4] receives the hypersingular group's sum of Q (the double layer's is s[1]).
template <int FAM, int KM, bool CK = false>
__device__ __forceinline__ void eval_groups(const RegArgs &a, int j, int dS, int dD, int dH, const PairCtx &p,
	float w1, float rinv_d, float k, float k_revf, float2 (&s)[6], float ki = 0.0f)
{
	constexpr int NONE = 1 << 20;
	int pS = (KM & SLP) && dS >= 0 ? dS : NONE;
	int pD = (KM & (DLP | DLPT)) && dD >= 0 ? dD : NONE;
	int pH = (KM & HSP) && dH >= 0 ? dH : NONE;
	for (;;) {
		const int mine = min(pS, min(pD, pH));
		const unsigned long long m = __ballot(mine < NONE);
		if (m == 0) break;
		const int lane = __ffsll((long long)m) - 1;
		const int dsel = __builtin_amdgcn_readlane(mine, lane);
		const bool fS = pS == dsel, fD = pD == dsel, fH = pH == dsel;
		float2 t[6] = {zero2(), zero2(), zero2(), zero2(), zero2(), zero2()};
		if (fS || fD || fH) {
			if (dsel <= 1) {
				accumulate_centroid<FAM, KM, CK>(p, w1, rinv_d, k, t, ki);
			} else {
				const int nq = a.qn[dsel];
				accumulate_points<FAM, KM, 0, CK>(a.qt[dsel] + (size_t)j * nq, nq, p, k_revf, k, t, ki);
			}
			// the Q sum serves both the double layer (s[1]) and the hypersingular (s[4]) groups
			const float2 tq = (KM & DLP) ? t[1] : t[4];
			if (fS) s[0] = t[0];
			if (fD) { s[1] = tq; s[2] = t[2]; }
			if (fH) { s[3] = t[3]; s[4] = tq; s[5] = t[5]; }
		}
		pS = fS ? NONE : pS;
		pD = fD ? NONE : pD;
		pH = fH ? NONE : pH;
	}
}

// Kernel values from the accumulators of add_point: SLP s[0], DLP -a s[1],
// DLPt s[2], HSP nn sum Q - a (3 s[3] - k^2 s[5]) with sum Q in s[4] when the
// hypersingular group had its own sums (hsp_apart), else s[1].
// CK: complex wavenumber k + i ki, the (k r)^2 G term takes (k + i ki)^2.
template <int FAM, int KM, bool CK = false>
__device__ __forceinline__ void kernel_values(const float2 (&s)[6], float av, float nn, float k, bool hsp_apart,
	float2 (&K)[4], float ki = 0.0f)
{
	K[0] = s[0];
	K[1] = make_float2(-av * s[1].x, -av * s[1].y);
	K[2] = s[2];
	const float2 sq = (hsp_apart || !(KM & DLP)) ? s[4] : s[1];
	float hr, hi;
	if (CK) {
		const float kr2 = k * k - ki * ki, ki2 = 2.0f * k * ki;
		hr = 3.0f * s[3].x - (kr2 * s[5].x - ki2 * s[5].y);
		hi = 3.0f * s[3].y - (kr2 * s[5].y + ki2 * s[5].x);
	} else {
		const float kk = FAM == HELM ? k * k : 0.0f;
		hr = __builtin_fmaf(3.0f, s[3].x, -kk * s[5].x);
		hi = __builtin_fmaf(3.0f, s[3].y, -kk * s[5].y);
	}
	K[3] = make_float2(nn * sq.x - av * hr, nn * sq.y - av * hi);
}

// table_degree from the thresholds of a trial record, as selects
template <int ORDER>
__device__ __forceinline__ int table_degree_rec(double d2, const double (&t)[NTHR])
{
	if (ORDER == 1)
		return d2 < t[THR_1_1] ? 6 : d2 < t[THR_1_8] ? 4 : d2 < t[THR_9_1] ? 2 : 0;
	if (ORDER == 2)
		return d2 < t[THR_1_1] ? 8 : d2 < t[THR_1_5] ? 6 : d2 < t[THR_2_8] ? 4 : d2 < t[THR_15_8] ? 2 : 0;
	return d2 < t[THR_1_1] ? 9 : d2 < t[THR_1_4] ? 8 : d2 < t[THR_1_9] ? 6 : d2 < t[THR_3_6] ? 4 :
		d2 < t[THR_22_3] ? 2 : 0;
}

// FIXQ > 0: fixed rule with FIXQ points (the caller guarantees a.fixed_degree
// has that many); 0: any policy.  MAP: stores go through a.rmap / a.cmap (a
// block of a mesh with both element types); without it the store address is
// the plain j * ld + i of the single-type meshes.
// CK: complex wavenumber k + i ki (Helmholtz): the attenuation e^{ki r} rides
// on the weights (e^{ki d} per pair from the FP64 distance, times
// 2^{ki log2(e) (r - d)} per point) and the kernels take i(k + i ki) r.
template <int FAM, int KM, int MODE, int FIXQ, bool MAP, bool CK = false>
__global__ void __launch_bounds__(256) regular_kernel(RegArgs a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int tile_c = blockIdx.x / a.row_tiles;
	const int i = tile_r * 256 + threadIdx.x;
	const bool row_ok = i < a.n_rows;
	const int ii = row_ok ? i : a.n_rows - 1;
	const int orow = MAP && a.rmap ? a.rmap[ii] : i;
	const double cix = a.tc[ii], ciy = a.tc[a.tn + ii], ciz = a.tc[2 * a.tn + ii];
	PairCtx p;
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
		// c_j - c_i in FP64 once: its squares equal those of c_i - c_j, so d2
		// is d2_pair(c_i, c_j) bit for bit (the classification stays exact)
		const double dxd = h.cx - cix, dyd = h.cy - ciy, dzd = h.cz - ciz;
		const double d2 = __fma_rn(dzd, dzd, __fma_rn(dyd, dyd, dxd * dxd));
		const bool act = row_ok && !(d2 < h.near) && !(a.same_mesh && i == (MAP && a.cmap ? a.cmap[j] : j));
		nxt = head(j + 1 < j1 ? j + 1 : j);
		if (__ballot(act) == 0) continue;
		const float dcx = (float)dxd, dcy = (float)dyd, dcz = (float)dzd;
		p.dcx = dcx; p.dcy = dcy; p.dcz = dcz;
		p.d2f = (float)d2;
		// d and 1/d in FP64 from the hardware reciprocal square root and one Newton step
		double y = __builtin_amdgcn_rsq(d2);
		y = y * (1.5 - 0.5 * d2 * y * y);
		const double d = d2 * y;
		const float rinv_d = (float)y;
		if (FAM == HELM) {
			double tt = d * a.k_rev;
			tt -= rint(tt);
			p.df = (float)d;
			p.t0 = (float)tt;
			if (CK) p.ad = __builtin_amdgcn_exp2f((float)(d * (a.ki * 1.4426950408889634)));
		} else {
			p.df = 0.0f; p.t0 = 0.0f;
		}
		float av = 0.0f, nn = 0.0f;
		if (KM & (DLP | HSP)) {
			const float nyx = h.nx, nyy = h.ny, nyz = h.nz;
			av = dcx * nyx + dcy * nyy + dcz * nyz;
			if (KM & HSP) nn = p.nxx * nyx + p.nxy * nyy + p.nxz * nyz;
		}
		if (KM & (DLPT | HSP)) p.b0 = dcx * p.nxx + dcy * p.nxy + dcz * p.nxz;
		float2 s[6] = {zero2(), zero2(), zero2(), zero2(), zero2(), zero2()};
		// hsp_apart: the hypersingular sum of Q went to s[4] (its own degree)
		bool hsp_apart = false;
		if (FIXQ > 0) {
			if (act)
				accumulate_points<FAM, KM, FIXQ, CK>(a.qt[a.fixed_degree] + (size_t)j * FIXQ, FIXQ, p, a.k_revf, a.k, s,
					a.ki);
		} else if (a.fixed_degree >= 0) {
			eval_degrees<FAM, KM, CK>(a, j, act ? a.fixed_degree : -1, p, h.w1, rinv_d, a.k, a.k_revf, s, a.ki);
		} else {
			// per-kernel degrees from the distance tables; where they coincide
			// (all far pairs: one point each) all kernels share one pass
			const float w1 = h.w1;
			const int fc = a.field_complexity;
			if (__ballot(act && d2 < h.far) == 0) {
				// beyond the last breakpoint (22.3 R, the largest of all tables)
				// every table degree is 0: the whole wave takes degree fc for all
				// kernels in one pass (degrees 0 and 1: the centroid rule)
				if (fc <= 1) {
					if (act) accumulate_centroid<FAM, KM, CK>(p, w1, rinv_d, a.k, s, a.ki);
				} else {
					eval_degrees<FAM, KM, CK>(a, j, act ? fc : -1, p, w1, rinv_d, a.k, a.k_revf, s, a.ki);
				}
			} else {
				// the distance-table thresholds of the record
				double t[NTHR];
#pragma unroll
				for (int q = 0; q < NTHR; ++q) t[q] = rec[j].thr[q];
				const int dS = (KM & SLP) && act ? fc + table_degree_rec<1>(d2, t) : -1;
				const int dD = (KM & (DLP | DLPT)) && act ? fc + table_degree_rec<2>(d2, t) : -1;
				const int dH = (KM & HSP) && act ? fc + table_degree_rec<3>(d2, t) : -1;
				eval_groups<FAM, KM, CK>(a, j, dS, dD, dH, p, w1, rinv_d, a.k, a.k_revf, s, a.ki);
				// the hypersingular sum of Q is in s[4]
				hsp_apart = true;
			}
		}
		if (act) {
			float2 K[4];
			kernel_values<FAM, KM, CK>(s, av, nn, a.k, hsp_apart, K, a.ki);
			const long long col = (long long)(MAP && a.cmap ? a.cmap[j] : j) * a.ld;
			if (MODE == OUT_SINGLE) {
				const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
				const float re = a.alpha.x * K[kk].x - a.alpha.y * K[kk].y;
				const float im = a.alpha.x * K[kk].y + a.alpha.y * K[kk].x;
				store_entry(a.out0, col + orow, a.dtype, a.accumulate, re, im);
			} else {
				// A = alpha D + H, B = G + alpha Ht
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

// line and triangle rules of the near-singular and singular formulas
struct Rules {
	double gl6x[6], gl6w[6];          // Gauss-Legendre order 10 (6 points)
	double gl21x[21], gl21w[21];      // Gauss-Legendre order 40 (21 points)
	double d7xi[13], d7eta[13], d7w[13];      // Dunavant degree 7
	double d14xi[42], d14eta[42], d14w[42];   // Dunavant degree 14
	// quadrilaterals: Gauss-Legendre orders 5, 9, 14 / 15, 30, 100 and the
	// order-7 Duffy blind rule around the centre of [-1, 1]^2 (xi, eta, w)
	double gl3x[3], gl3w[3], gl5x[5], gl5w[5], gl8x[8], gl8w[8], gl16x[16], gl16w[16], gl51x[51], gl51w[51];
	double duffy[64][3];
	// the same rule around each node of the quadrilateral gauss field
	// ((-,-), (+,-), (-,+), (+,+) at +-sqrt(3)/3)
	double gduffy[4][64][3];
	// reference points (xi, eta) of the Dunavant rule of each degree, FP32
	// (the test-side shape functions of the regular Galerkin pass)
	float2 ref[NSLOT][REF_MAXQ];
};

struct NearArgs {
	int n_pairs;
	const int *rmap, *cmap;      // result row / column of each test / trial element (null: its own index)
	const int *pi, *pj;          // pair list (near pass); diagonal pass uses (p, p)
	int tn, sn;
	const double *tc;            // test centres [3][tn]
	const double *tnu;           // test unit normals FP64 [3][tn]
	const double *X;             // trial vertices [9][sn]
	const double *sJ;            // trial Jacobians
	const double *snu;           // trial unit normals FP64 [3][sn]
	double k;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	double2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
	const Rules *rules;
	// quadrilateral near pass: count of pairs whose collocation point has no
	// image on the element (the inverse mapping did not converge); null: none
	// kept; and the Newton step limit of that mapping (NiHu: 100)
	int *status;
	int inv_iter;
};

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 d3sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 d3cross(D3 a, D3 b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double d3norm(D3 a) { return sqrt(d3dot(a, a)); }
__device__ __forceinline__ D3 d3scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }

__device__ __forceinline__ D3 tri_x(const D3 V[3], double xi, double eta)
{
	const double N0 = 1.0 - xi - eta;
	return {V[0].x * N0 + V[1].x * xi + V[2].x * eta, V[0].y * N0 + V[1].y * xi + V[2].y * eta,
		V[0].z * N0 + V[1].z * xi + V[2].z * eta};
}

constexpr double PI_D = 3.141592653589793238462643383279502884;

// 1/sqrt(x) in FP64: hardware estimate refined by two Newton steps
__device__ __forceinline__ double rsqrt64(double x)
{
	double y = __builtin_amdgcn_rsq(x);
	y = y * (1.5 - 0.5 * x * y * y);
	const double e = __fma_rn(-x * y, y, 1.0);      // 1 - x y^2
	return __fma_rn(0.5 * y, e, y);
}

// 1/x in FP64 (x finite and nonzero): hardware estimate refined by two Newton steps
__device__ __forceinline__ double rcp64(double x)
{
	double y = __builtin_amdgcn_rcp(x);
	double e = __fma_rn(-x, y, 1.0);
	y = __fma_rn(y, e, y);
	e = __fma_rn(-x, y, 1.0);
	return __fma_rn(y, e, y);
}

// sin and cos in FP64 for the arguments of the quadrature loops (angles and
// k r): r = x - n pi/2 with pi/2 split in three parts and FMAs (Cody and
// Waite's reduction, "Software Manual for the Elementary Functions", 1980;
// the neglected tail of pi/2 is 2e-49, so the reduction stays exact to
// rounding far beyond the |x| < 1e4 met here), then the Taylor series of sin
// (to r^15) and cos (to r^16) on |r| <= pi/4, whose truncation is below
// 5e-17; measured max abs error 1.5e-16 on |x| < 1e7.  NaN and infinities
// give NaN.
// Taylor coefficients of sin (-1/15! ... -1/3!) and cos (1/16! ... -1/2!), read
// through the scalar cache: held in SGPRs, each Horner step is one v_fma_f64
// (as VGPR-resident literals every step costs a copy plus a v_fmac_f64)
__constant__ double SINCOS_TAYLOR[15] = {
	-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0,
	-1.0 / 6.0,
	1.0 / 20922789888000.0, -1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0,
	-1.0 / 720.0, 1.0 / 24.0, -0.5};

__device__ __forceinline__ void sincos64(double x, double *s, double *c)
{
	const auto C = uni((const double *)SINCOS_TAYLOR);
	const double n = rint(x * 0.63661977236758138);         // 2 / pi
	double r = __fma_rn(-n, 1.5707963267948966, x);
	r = __fma_rn(-n, 6.123233995736766e-17, r);
	r = __fma_rn(-n, -1.4973849048591698e-33, r);
	const double r2 = r * r;
	double ps = C[0];
#pragma unroll
	for (int i = 1; i < 7; ++i) ps = __fma_rn(ps, r2, C[i]);
	const double sr = __fma_rn(ps * r2, r, r);
	double pc = C[7];
#pragma unroll
	for (int i = 8; i < 15; ++i) pc = __fma_rn(pc, r2, C[i]);
	const double cr = __fma_rn(pc, r2, 1.0);
	// x = n pi/2 + r: quadrant q = n mod 4 rotates (sin r, cos r)
	const int q = (int)__fma_rn(-4.0, floor(n * 0.25), n);
	const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
	*s = (q & 2) ? -ss : ss;
	*c = ((q + 1) & 2) ? -cc : cc;
}

// 1/21, 1/19, ..., 1/3: the series atanh(s) = s + s^3/3 + s^5/5 + ... (through
// the scalar cache, as for sincos64)
__constant__ double ATANH_SERIES[10] = {1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0,
	1.0 / 9.0, 1.0 / 7.0, 1.0 / 5.0, 1.0 / 3.0};

// natural logarithm in FP64: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log x = e log 2 + 2 atanh((m - 1) / (m + 1)), |s| <= 0.172, the series to
// s^21 (truncation below 3e-17), log 2 in two parts; measured max relative
// error 3.3e-16 on [1e-12, 1e12] (against long double).  0 gives -inf,
// negative values and NaN give NaN, +inf gives +inf.
__device__ __forceinline__ double log64(double x)
{
	const auto C = uni((const double *)ATANH_SERIES);
	int e;
	double m = frexp(x, &e);
	if (m < 0.70710678118654752) { m *= 2.0; e -= 1; }
	const double sv = (m - 1.0) * rcp64(m + 1.0);
	const double s2 = sv * sv;
	double p = C[0];
#pragma unroll
	for (int i = 1; i < 10; ++i) p = __fma_rn(p, s2, C[i]);
	const double at = __fma_rn(p * s2, sv, sv);
	const double r = __fma_rn((double)e, 0.6931471805599453, __fma_rn((double)e, 2.3190468138462996e-17, 2.0 * at));
	if (x > 0.0 && x < __builtin_inf()) return r;
	return x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan(""));
}

// K_helm - K_lap of the kernels of MASK from e^{-ikr} = cs - i sn, kr and the
// point's geometry (g0 = w J / (4 pi r), rinv = 1/r, rdny, rdnx, nn)
template <int MASK>
__device__ __forceinline__ void dynamic_terms(double cs, double sn, double kr, double g0, double rinv, double rdny,
	double rdnx, double nn, double2 (&s)[4])
{
	if (MASK & SLP) {
		s[0].x += (cs - 1.0) * g0;
		s[0].y += -sn * g0;
	}
	if (MASK & (DLP | DLPT | HSP)) {
		const double g1 = g0 * rinv;
		// e(-1 - ikr) - (-1)
		const double er = -cs - kr * sn + 1.0;
		const double ei = -kr * cs + sn;
		if (MASK & DLP) { s[1].x += er * g1 * rdny; s[1].y += ei * g1 * rdny; }
		if (MASK & DLPT) { s[2].x += er * g1 * rdnx; s[2].y += ei * g1 * rdnx; }
		if (MASK & HSP) {
			const double g3 = g1 * rinv;
			const double cc = rdnx * rdny;
			// e [(3 - kr^2 + 3ikr) cc + (1 + ikr) nn] - [3 cc + nn]
			const double ar = (3.0 - kr * kr) * cc + nn, ai = kr * (3.0 * cc + nn);
			s[3].x += ((cs * ar + sn * ai) - (3.0 * cc + nn)) * g3;
			s[3].y += (cs * ai - sn * ar) * g3;
		}
	}
}

// ---- complex wavenumbers k + i ki (FP64): e = e^{-i(k + i ki) r} =
// e^{ki r} (cos kr - i sin kr) and z = i (k + i ki) r = -ki r + i k r; the
// kernels are polynomials in z times e:
//   SLP e g0,  DLP e (-1 - z) g1 (r.n_y)/r,  DLPt the same with (-r.n_x)/r,
//   HSP e [(3 + 3z + z^2) rdnx rdny + (1 + z) n_x.n_y] g3
// (g0 = w J / (4 pi r), g1 = g0 / r, g3 = g1 / r^2; the real-k forms with z = i k r).
__device__ __forceinline__ double2 cmul64(double2 a, double2 b)
{
	return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void wave64(double k, double ki, double r, double2 *e, double2 *z)
{
	double sn, cs;
	sincos64(k * r, &sn, &cs);
	const double am = exp(ki * r);
	*e = make_double2(am * cs, -am * sn);
	*z = make_double2(-ki * r, k * r);
}

// K_helm - K_lap (minus = 1) or K_helm (minus = 0) of the kernels of MASK
template <int MASK>
__device__ __forceinline__ void wave_terms_c(double2 e, double2 z, double g0, double rinv, double rdny, double rdnx,
	double nn, double minus, double2 (&s)[4])
{
	if (MASK & SLP) { s[0].x += (e.x - minus) * g0; s[0].y += e.y * g0; }
	if (MASK & (DLP | DLPT | HSP)) {
		const double g1 = g0 * rinv;
		const double2 f = cmul64(e, make_double2(-1.0 - z.x, -z.y));
		const double fr = f.x + minus, fi = f.y;
		if (MASK & DLP) { s[1].x += fr * g1 * rdny; s[1].y += fi * g1 * rdny; }
		if (MASK & DLPT) { s[2].x += fr * g1 * rdnx; s[2].y += fi * g1 * rdnx; }
		if (MASK & HSP) {
			const double g3 = g1 * rinv, cc = rdnx * rdny;
			const double2 z2 = cmul64(z, z);
			const double2 h = cmul64(e, make_double2((3.0 + 3.0 * z.x + z2.x) * cc + (1.0 + z.x) * nn,
				(3.0 * z.y + z2.y) * cc + z.y * nn));
			s[3].x += (h.x - minus * (3.0 * cc + nn)) * g3;
			s[3].y += h.y * g3;
		}
	}
}

// K_helm - K_lap for the kernels of MASK at one point (FP64), r = y - x;
// CK: the complex wavenumber k + i ki
template <int MASK, bool CK = false>
__device__ __forceinline__ void dynamic_point(D3 x, D3 y, D3 nx, D3 ny, double k, double wj, double2 (&s)[4],
	double ki = 0.0)
{
	const D3 rv = d3sub(y, x);
	const double r2 = d3dot(rv, rv);
	const double rinv = rsqrt64(r2);
	const double r = r2 * rinv;
	const double g0 = rinv * wj * (1.0 / (4.0 * PI_D));
	double rdny = 0.0, rdnx = 0.0, nn = 0.0;
	if (MASK & (DLP | DLPT | HSP)) {
		rdny = d3dot(rv, ny) * rinv;
		rdnx = -d3dot(rv, nx) * rinv;
		if (MASK & HSP) nn = d3dot(nx, ny);
	}
	if (CK) {
		double2 e, z;
		wave64(k, ki, r, &e, &z);
		wave_terms_c<MASK>(e, z, g0, rinv, rdny, rdnx, nn, 1.0, s);
		return;
	}
	const double kr = k * r;
	double sn, cs;
	sincos64(kr, &sn, &cs);
	// e^{-ikr} = cs - i sn
	dynamic_terms<MASK>(cs, sn, kr, g0, rinv, rdny, rdnx, nn, s);
}

// Laplace static part of the near-singular integral over a flat triangle in
// its own plane (nihu/library/laplace_nearly_singular_integrals.hpp:62-389,
// nihu/library/plane_element_helper.hpp:76-128), all kernels of MASK.
// Along a ray at angle th from the projected point, the in-plane distance to
// side n is P = dist_n / cos(th - th0_n), evaluated here as
// dist_n / (cos th cos th0_n + sin th sin th0_n) with the sin and cos of th
// that DLPt and HSP need anyway.
// lane / nl: this lane evaluates every nl-th point (lane 0 of 1: all of them),
// for callers that split one pair's integral over several lanes.
template <int MASK, int NC = 3>
__device__ void near_static(const D3 V[NC], D3 x0g, D3 nxg, const Rules &R, double (&st)[4], int lane = 0, int nl = 1)
{
	const D3 v1 = d3sub(V[1], V[0]), v2 = d3sub(V[2], V[0]);
	const D3 T0 = d3scale(v1, 1.0 / d3norm(v1));
	D3 T2 = d3cross(v1, v2);
	T2 = d3scale(T2, 1.0 / d3norm(T2));
	D3 T1 = d3cross(T2, T0);
	T1 = d3scale(T1, 1.0 / d3norm(T1));
	D3 C[NC];
	for (int a = 0; a < NC; ++a) C[a] = {d3dot(T0, V[a]), d3dot(T1, V[a]), d3dot(T2, V[a])};
	const D3 x0 = {d3dot(T0, x0g), d3dot(T1, x0g), d3dot(T2, x0g)};
	const D3 nx = {d3dot(T0, nxg), d3dot(T1, nxg), d3dot(T2, nxg)};
	double th_lim[NC], c0[NC], s0[NC], dist[NC], dinv[NC];
	for (int n = 0; n < NC; ++n) {
		const int m = n == NC - 1 ? 0 : n + 1;
		double lx = C[m].x - C[n].x, ly = C[m].y - C[n].y;
		const double lninv = rsqrt64(lx * lx + ly * ly);
		lx *= lninv; ly *= lninv;
		const double d1x = C[n].x - x0.x, d1y = C[n].y - x0.y;
		const double dl = d1x * lx + d1y * ly;
		const double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		th_lim[n] = atan2(d1y, d1x);
		const double d02 = d0x * d0x + d0y * d0y;
		// direction of the foot of the perpendicular: (cos th0, sin th0);
		// th0 = atan2(0, 0) = 0 when the projected point lies on the side's line
		if (d02 > 0.0) {
			const double d0inv = rsqrt64(d02);
			dist[n] = d02 * d0inv;
			c0[n] = d0x * d0inv;
			s0[n] = d0y * d0inv;
		} else {
			dist[n] = 0.0;
			c0[n] = 1.0;
			s0[n] = 0.0;
		}
		dinv[n] = 1.0 / dist[n];
	}
	const double z = x0.z - C[0].z;
	const double az = fabs(z);
	const double z2 = z * z;
	const double sgz = (0.0 < z) - (z < 0.0);
	const double zinv = 1.0 / z, nzz = nx.z * z;
	double rs = 0.0, rd = 0.0, rt = 0.0, rh = 0.0;
	// one Gauss node of the SLP / DLP / DLPt rule: cd = cos(th - th0_n),
	// A = n_x cos th + n_y sin th (in-plane components of the test normal)
	auto node6 = [&](int n, double cd, double A, double w) {
		const double P = dist[n] * rcp64(cd);
		const double r2 = P * P + z2;
		const double rinv = rsqrt64(r2);
		const double r = r2 * rinv;
		if (MASK & SLP) rs += (r - az) * w;
		if (MASK & DLP) rd += (1.0 - az * rinv) * w;
		if (MASK & DLPT) {
			const double rdot = P * A - nx.z * z;
			rt += (-rdot * rinv + A * log64(P + r) - nx.z * sgz) * w;
		}
	};
	// one node of the HSP rule: -P^2 / r^2 / z * (-r.n / r) with r^2 = P^2 + z^2
	// and r.n = P A - n_z z, written with q = 1/P as
	// (A - n_z z q) / (z (1 + z^2 q^2)^{3/2}); a side whose line holds the
	// projected point (dist 0, P = 0) adds 0; in the plane: -n_z / P
	auto node21 = [&](int n, double cd, double A, double w) {
		const double Pinv = cd * dinv[n];
		double integrand;
		if (az > 1e-12) {
			const double rs3 = rsqrt64(__fma_rn(z2 * Pinv, Pinv, 1.0));
			integrand = dist[n] > 0.0 ? __fma_rn(-nzz, Pinv, A) * (rs3 * rs3 * rs3) * zinv : 0.0;
		} else {
			integrand = -nx.z * Pinv;
		}
		rh += integrand * w;
	};
	for (int n = 0; n < NC; ++n) {
		double t1 = th_lim[n], t2 = th_lim[n == NC - 1 ? 0 : n + 1];
		if (fabs(t2 - t1) > PI_D) {
			if (t2 > t1) t1 += 2.0 * PI_D;
			else t2 += 2.0 * PI_D;
		}
		const double jac = (t2 - t1) / 2.0, tm = (t1 + t2) / 2.0;
		// the Gauss nodes th = tm + jac x_g come in mirror pairs tm +- a (the
		// rules are symmetric, x_{N-1-g} = -x_g): with the sin and cos of tm
		// and of a, cos(th - th0_n) = Cm cos a -+ Sm sin a and
		// n_x cos th + n_y sin th = Am cos a +- Bm sin a
		double sm, cm;
		sincos64(tm, &sm, &cm);
		const double Cm = cm * c0[n] + sm * s0[n], Sm = sm * c0[n] - cm * s0[n];
		const double Am = nx.x * cm + nx.y * sm, Bm = nx.y * cm - nx.x * sm;
		if (MASK & (SLP | DLP | DLPT)) {
			for (int g = (lane + 3 * nl - (3 * n) % nl) % nl; g < 3; g += nl) {
				double sa, ca;
				sincos64(jac * R.gl6x[g], &sa, &ca);
				const double w = R.gl6w[g] * jac;
				node6(n, __fma_rn(-Sm, sa, Cm * ca), __fma_rn(Bm, sa, Am * ca), w);
				node6(n, __fma_rn(Sm, sa, Cm * ca), __fma_rn(-Bm, sa, Am * ca), w);
			}
		}
		if (MASK & HSP) {
			// 10 mirror pairs and the middle node x_10 = 0 (th = tm)
			for (int g = (lane + 11 * nl - (11 * n) % nl) % nl; g < 11; g += nl) {
				const double w = R.gl21w[g] * jac;
				if (g < 10) {
					double sa, ca;
					sincos64(jac * R.gl21x[g], &sa, &ca);
					node21(n, __fma_rn(-Sm, sa, Cm * ca), __fma_rn(Bm, sa, Am * ca), w);
					node21(n, __fma_rn(Sm, sa, Cm * ca), __fma_rn(-Bm, sa, Am * ca), w);
				} else {
					node21(n, Cm, Am, w);
				}
			}
		}
	}
	const double f = 1.0 / (4.0 * PI_D);
	st[0] = rs * f;
	st[1] = sgz * rd * f;
	st[2] = rt * f;
	st[3] = rh * f;
}

// Diagonal integrals at the centre of a flat triangle
// (nihu/library/helmholtz_singular_integrals.hpp:703-839,
//  nihu/library/laplace_singular_integrals.hpp:456-505, plane_element_helper.hpp:42-74);
// DLP and DLPt vanish (normal_derivative_singular_integrals.hpp:13-58).
// Static (wavenumber-independent) part: st[0] = sum r sin(a) ln(tan((a+t)/2)/tan(a/2))
// over all sides (SLP), st[1] = the same over sides with theta >= 1e-3 (IG0 of
// the Helmholtz HSP), st[2] = sum (cos(a+t) - cos a)/(r sin a) over sides with
// theta >= 1e-3 (IddG0), st[3] = the same over all sides (Laplace HSP).
__device__ void diag_static(const D3 V[3], D3 x0, double (&st)[4])
{
	double r[3], theta[3], alpha[3];
	D3 Rv[3], Cv[3];
	for (int a = 0; a < 3; ++a) {
		Rv[a] = d3sub(V[a], x0);
		r[a] = d3norm(Rv[a]);
		Rv[a] = d3scale(Rv[a], 1.0 / r[a]);
		Cv[a] = d3sub(V[a], V[a == 2 ? 0 : a + 1]);
		Cv[a] = d3scale(Cv[a], 1.0 / d3norm(Cv[a]));
	}
	for (int a = 0; a < 3; ++a) {
		const double cs = d3dot(Rv[a], Rv[a == 2 ? 0 : a + 1]);
		theta[a] = cs >= 1.0 ? 0.0 : acos(cs);
		alpha[a] = acos(d3dot(Rv[a], Cv[a]));
	}
	st[0] = st[1] = st[2] = st[3] = 0.0;
	for (int a = 0; a < 3; ++a) {
		const double sa = sin(alpha[a]);
		const double lg = r[a] * sa * log(tan((alpha[a] + theta[a]) / 2.0) / tan(alpha[a] / 2.0));
		const double dd = (cos(alpha[a] + theta[a]) - cos(alpha[a])) / (r[a] * sa);
		st[0] += lg;
		st[3] += dd;
		if (theta[a] < 1e-3) continue;
		st[1] += lg;
		st[2] += dd;
	}
}

// Wavenumber-dependent (13-point) part of the Helmholtz diagonal integrals:
// dyn[0] for SLP, dyn[1] for HSP, not yet divided by 4 pi.
template <int MASK, bool CK = false>
__device__ void diag_dynamic(const D3 V[3], D3 x0, double J, double k, const Rules &Rr, double2 (&dyn)[2],
	double ki = 0.0)
{
	double dr = 0.0, di = 0.0, ar = 0.0, ai = 0.0;
	for (int q = 0; q < 13; ++q) {
		const double rq = d3norm(d3sub(tri_x(V, Rr.d7xi[q], Rr.d7eta[q]), x0));
		if (CK) {
			// complex wavenumber kc = k + i ki, z = i kc rq, e = e^{-z}
			const double wJ = Rr.d7w[q] * J;
			double2 e, z;
			wave64(k, ki, rq, &e, &z);
			const double2 kc = make_double2(k, ki);
			if (MASK & SLP) {
				// -i kc e^{-i h} sinc(h), h = kc rq / 2: -(1 - e) / rq, or the series for |h| <= 1e-3
				double2 v;
				const double2 h = make_double2(k * rq / 2.0, ki * rq / 2.0);
				if (sqrt(h.x * h.x + h.y * h.y) > 1e-3) {
					v = make_double2(-(1.0 - e.x) / rq, e.y / rq);
				} else {
					// sinc(h) = 1 - h^2/6 (1 - h^2/20)
					const double2 h2 = cmul64(h, h);
					const double2 c6 = cmul64(make_double2(h2.x / 6.0, h2.y / 6.0),
						make_double2(1.0 - h2.x / 20.0, -h2.y / 20.0));
					const double2 sc = make_double2(1.0 - c6.x, -c6.y);
					// e^{-i h} = e^{ki rq / 2} (cos(k rq / 2) - i sin(k rq / 2))
					double2 eh, zh;
					wave64(k, ki, rq / 2.0, &eh, &zh);
					v = cmul64(cmul64(make_double2(ki, -k), eh), sc);
				}
				dr += v.x * wJ;
				di += v.y * wJ;
			}
			if (MASK & HSP) {
				double2 v;
				if (fabs(rq) > 1e-3) {
					// (e (1 + z) - 1 + z^2 / 2) / r^3
					const double2 z2 = cmul64(z, z);
					const double2 f = cmul64(e, make_double2(1.0 + z.x, z.y));
					const double r3 = rq * rq * rq;
					v = make_double2((f.x - 1.0 + z2.x / 2.0) / r3, (f.y + z2.y / 2.0) / r3);
				} else {
					// -i kc^3 (1/3 - z(1/8 - z(1/30 - z(1/144 - z(1/840 - z/5760)))))
					double2 pp = make_double2(1.0 / 840.0 - z.x / 5760.0, -z.y / 5760.0);
					const double c2[4] = {1.0 / 144.0, 1.0 / 30.0, 1.0 / 8.0, 1.0 / 3.0};
					for (int t = 0; t < 4; ++t) {
						const double2 zp = cmul64(z, pp);
						pp = make_double2(c2[t] - zp.x, -zp.y);
					}
					const double2 k3 = cmul64(cmul64(kc, kc), kc);
					v = cmul64(make_double2(k3.y, -k3.x), pp);
				}
				ar += v.x * wJ;
				ai += v.y * wJ;
			}
			continue;
		}
		if (MASK & SLP) {
			// -ik e^{-ik rq/2} sinc(k rq/2)
			const double h = k * rq / 2.0;
			double sh, ch;
			sincos64(h, &sh, &ch);
			const double sc = fabs(h) > 1e-3 ? sh / h : 1.0 - h * h / 6.0 * (1.0 - h * h / 20.0);
			dr += (-k * sh * sc) * Rr.d7w[q] * J;
			di += (-k * ch * sc) * Rr.d7w[q] * J;
		}
		if (MASK & HSP) {
			double vr, vi;
			if (fabs(rq) > 1e-3) {
				// (e^{-ikr}(1 + ikr) - 1 + (ikr)^2/2) / r^3
				const double kr = k * rq;
				double sn, cs;
				sincos64(kr, &sn, &cs);
				const double er = cs + kr * sn, ei = kr * cs - sn;
				vr = (er - 1.0 - kr * kr / 2.0) / (rq * rq * rq);
				vi = ei / (rq * rq * rq);
			} else {
				// -i k^3 (1/3 - ikr(1/8 - ikr(1/30 - ikr(1/144 - ikr(1/840 - ikr/5760)))))
				const double kr = k * rq;
				double pr = 1.0 / 840.0, pi = -kr / 5760.0;
				const double c2[4] = {1.0 / 144.0, 1.0 / 30.0, 1.0 / 8.0, 1.0 / 3.0};
				for (int t = 0; t < 4; ++t) {
					const double nr = c2[t] + kr * pi;
					const double ni = -kr * pr;
					pr = nr; pi = ni;
				}
				const double k3 = k * k * k;
				vr = k3 * pi;
				vi = -k3 * pr;
			}
			ar += vr * Rr.d7w[q] * J;
			ai += vi * Rr.d7w[q] * J;
		}
	}
	dyn[0] = make_double2(dr, di);
	dyn[1] = make_double2(ar, ai);
}

// combine the static and dynamic diagonal parts into kernel values
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void diag_combine(const double (&st)[4], const double2 (&dyn)[2], double k, double2 (&s)[4],
	double ki = 0.0)
{
	const double f = 1.0 / (4.0 * PI_D);
	if (MASK & SLP) s[0] = FAM == HELM ? make_double2((st[0] + dyn[0].x) * f, dyn[0].y * f) : make_double2(st[0] * f, 0.0);
	if (MASK & HSP) {
		if (CK)   // + kc^2 / 2 IG0 with kc = k + i ki
			s[3] = make_double2((st[2] + (k * k - ki * ki) / 2.0 * st[1] + dyn[1].x) * f,
				(k * ki * st[1] + dyn[1].y) * f);
		else
			s[3] = FAM == HELM ? make_double2((st[2] + k * k / 2.0 * st[1] + dyn[1].x) * f, dyn[1].y * f)
			                   : make_double2(st[3] * f, 0.0);
	}
}

// dynamic part of the Helmholtz near-singular integral: sum over the 42
// Dunavant-14 points of (K_helm - K_lap) J w
// (nihu/library/helmholtz_nearly_singular_integrals.hpp:48-219).
// The element is flat: r = (V0 - x) + xi (V1 - V0) + eta (V2 - V0), so r.n_y
// is the same at every point and r.n_x is linear in (xi, eta).
template <int MASK, bool CK = false>
__device__ __forceinline__ void near_dynamic(const D3 V[3], D3 x0, D3 nx, D3 ny, double J, double k, const Rules &R,
	double2 (&s)[4], double ki = 0.0)
{
	const D3 a0 = d3sub(V[0], x0), e1 = d3sub(V[1], V[0]), e2 = d3sub(V[2], V[0]);
	const double any = d3dot(a0, ny);
	const double anx = d3dot(a0, nx), e1nx = d3dot(e1, nx), e2nx = d3dot(e2, nx);
	const double nn = (MASK & HSP) ? d3dot(nx, ny) : 0.0;
	const double J4 = J * (1.0 / (4.0 * PI_D));
	for (int q = 0; q < 42; ++q) {
		const double xi = R.d14xi[q], eta = R.d14eta[q];
		const D3 rv = {__fma_rn(eta, e2.x, __fma_rn(xi, e1.x, a0.x)), __fma_rn(eta, e2.y, __fma_rn(xi, e1.y, a0.y)),
			__fma_rn(eta, e2.z, __fma_rn(xi, e1.z, a0.z))};
		const double r2 = d3dot(rv, rv);
		const double rinv = rsqrt64(r2);
		const double g0 = rinv * (R.d14w[q] * J4);
		double rdny = 0.0, rdnx = 0.0;
		if (MASK & (DLP | DLPT | HSP)) {
			rdny = any * rinv;
			rdnx = -__fma_rn(eta, e2nx, __fma_rn(xi, e1nx, anx)) * rinv;
		}
		if (CK) {
			double2 e, z;
			wave64(k, ki, r2 * rinv, &e, &z);
			wave_terms_c<MASK>(e, z, g0, rinv, rdny, rdnx, nn, 1.0, s);
		} else {
			const double kr = k * (r2 * rinv);
			double sn, cs;
			sincos64(kr, &sn, &cs);
			dynamic_terms<MASK>(cs, sn, kr, g0, rinv, rdny, rdnx, nn, s);
		}
	}
}

template <int FAM, int KM, int MODE, bool DIAG, bool CK = false>
__global__ void __launch_bounds__(64) fp64_pass_kernel(NearArgs a)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= a.n_pairs) return;
	const int i = DIAG ? p : a.pi[p];
	const int j = DIAG ? p : a.pj[p];
	const Rules &R = *a.rules;
	D3 V[3];
	for (int v = 0; v < 3; ++v)
		V[v] = {a.X[(3 * v) * a.sn + j], a.X[(3 * v + 1) * a.sn + j], a.X[(3 * v + 2) * a.sn + j]};
	const D3 x0 = {a.tc[i], a.tc[a.tn + i], a.tc[2 * a.tn + i]};
	const D3 nx = {a.tnu[i], a.tnu[a.tn + i], a.tnu[2 * a.tn + i]};
	const D3 ny = {a.snu[j], a.snu[a.sn + j], a.snu[2 * a.sn + j]};
	const double J = a.sJ[j];
	double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
	if (DIAG) {
		double st[4];
		double2 dyn[2] = {make_double2(0, 0), make_double2(0, 0)};
		diag_static(V, x0, st);
		if (FAM == HELM) diag_dynamic<KM & (SLP | HSP), CK>(V, x0, J, a.k, R, dyn, a.ki);
		diag_combine<FAM, KM & (SLP | HSP), CK>(st, dyn, a.k, s, a.ki);
	} else {
		double st[4] = {0, 0, 0, 0};
		near_static<KM>(V, x0, nx, R, st);
		for (int t = 0; t < 4; ++t) s[t].x = st[t];
		if (FAM == HELM) near_dynamic<KM, CK>(V, x0, nx, ny, J, a.k, R, s, a.ki);
	}
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
			ar += -0.5;                      // - I/2
			br += a.alpha.x / 2.0;           // + alpha/2 I
			bi += a.alpha.y / 2.0;
		}
		store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
		store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
	}
}

// ---------------------------------------------------------------- regular pairs in FP64

// The regular pass of the FP64 precision policy (every pair in FP64, the
// plain method of nihu/core/double_integral.hpp:543-589): the same pair
// classes and per-kernel degrees as regular_kernel, each quadrature point
// evaluated in FP64 from the point tables of nbg_quad_table64 (offset of the
// point from its element's centre, w J / (4 pi)).
struct Reg64Args {
	int n_rows, n_cols, same_mesh, row_tiles, tile_cols;
	const int *rmap, *cmap;      // result row / column of each test / trial element (null: its own index)
	int tn, sn;
	const double *tc, *tnu;      // test centres, unit normals [3][tn]
	const double *sc, *snu;      // trial centres, unit normals [3][sn]
	const double *thr;           // [NTHR][sn]
	const double *qt[NSLOT];     // per degree: [sn][Q][4]
	int qn[NSLOT];
	int fixed_degree;            // >= 0: fixed rule
	double k;
	double ki;                   // imaginary part of a complex wavenumber (kernels with CK)
	double2 alpha;
	void *out0, *out1;
	long long ld;
	int dtype, accumulate;
};

// the kernels of MASK at one point, rv = y - x (FP64): s += K w J
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void full_point64(D3 rv, D3 nx, D3 ny, double k, double w4, double2 (&s)[4],
	double ki = 0.0)
{
	const double r2 = d3dot(rv, rv);
	const double rinv = rsqrt64(r2);
	if (CK) {
		double2 e, z;
		wave64(k, ki, r2 * rinv, &e, &z);
		const double nn = (MASK & HSP) ? d3dot(nx, ny) : 0.0;
		wave_terms_c<MASK>(e, z, w4 * rinv, rinv, d3dot(rv, ny) * rinv, -d3dot(rv, nx) * rinv, nn, 0.0, s);
		return;
	}
	double sn = 0.0, cs = 1.0, kr = 0.0;
	if (FAM == HELM) {
		kr = k * (r2 * rinv);
		sincos64(kr, &sn, &cs);
	}
	// e^{-ikr} = cs - i sn; g0 = w J / (4 pi r)
	const double g0 = w4 * rinv;
	if (MASK & SLP) { s[0].x += cs * g0; s[0].y += -sn * g0; }
	if (MASK & (DLP | DLPT | HSP)) {
		const double g1 = g0 * rinv;
		// e (-1 - ikr)
		const double er = -cs - kr * sn, ei = -kr * cs + sn;
		const double rdny = d3dot(rv, ny) * rinv, rdnx = -d3dot(rv, nx) * rinv;
		if (MASK & DLP) { s[1].x += er * g1 * rdny; s[1].y += ei * g1 * rdny; }
		if (MASK & DLPT) { s[2].x += er * g1 * rdnx; s[2].y += ei * g1 * rdnx; }
		if (MASK & HSP) {
			const double g3 = g1 * rinv, cc = rdnx * rdny, nn = d3dot(nx, ny);
			// e [(3 - kr^2 + 3ikr) cc + (1 + ikr) nn]
			const double ar = (3.0 - kr * kr) * cc + nn, ai = kr * (3.0 * cc + nn);
			s[3].x += (cs * ar + sn * ai) * g3;
			s[3].y += (cs * ai - sn * ar) * g3;
		}
	}
}

// the kernels of MASK over the rule of each lane's degree (-1: none), one
// distinct degree of the wave at a time (wave-uniform table address)
template <int FAM, int MASK, bool CK = false>
__device__ __forceinline__ void eval64(const Reg64Args &a, int j, int deg, D3 dc, D3 nx, D3 ny, double2 (&s)[4])
{
	for (;;) {
		const unsigned long long m = __ballot(deg >= 0);
		if (m == 0) break;
		const int dsel = __builtin_amdgcn_readlane(deg, __ffsll((long long)m) - 1);
		const unsigned long long msel = __ballot(deg == dsel);
		if ((msel >> __lane_id()) & 1ull) {
			const int nq = a.qn[dsel];
			const double *q = a.qt[dsel] + (size_t)j * nq * 4;
			for (int t = 0; t < nq; ++t) {
				const D3 rv = {dc.x + q[4 * t], dc.y + q[4 * t + 1], dc.z + q[4 * t + 2]};
				full_point64<FAM, MASK, CK>(rv, nx, ny, a.k, q[4 * t + 3], s, a.ki);
			}
			deg = -1;
		}
	}
}

template <int FAM, int KM, int MODE, bool CK = false>
__global__ void __launch_bounds__(256) regular64_kernel(Reg64Args a)
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
		// classification exactly as regular_kernel: d2 from c_j - c_i in FP64
		const D3 dc = {a.sc[j] - x0.x, a.sc[a.sn + j] - x0.y, a.sc[2 * a.sn + j] - x0.z};
		const double d2 = __fma_rn(dc.z, dc.z, __fma_rn(dc.y, dc.y, dc.x * dc.x));
		const int cj = a.cmap ? a.cmap[j] : j;
		const bool act = !(d2 < a.thr[THR_NEAR * a.sn + j]) && !(a.same_mesh && i == cj);
		const D3 ny = {a.snu[j], a.snu[a.sn + j], a.snu[2 * a.sn + j]};
		const int fd = a.fixed_degree;
		double2 s[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
		if (KM & SLP)
			eval64<FAM, SLP, CK>(a, j, !act ? -1 : fd >= 0 ? fd : table_degree<1>(d2, a.thr, a.sn, j), dc, nx, ny, s);
		if (KM & (DLP | DLPT))
			eval64<FAM, KM & (DLP | DLPT), CK>(a, j, !act ? -1 : fd >= 0 ? fd : table_degree<2>(d2, a.thr, a.sn, j),
				dc, nx, ny, s);
		if (KM & HSP)
			eval64<FAM, HSP, CK>(a, j, !act ? -1 : fd >= 0 ? fd : table_degree<3>(d2, a.thr, a.sn, j), dc, nx, ny, s);
		if (!act) continue;
		const long long idx = (long long)cj * a.ld + orow;
		if (MODE == OUT_SINGLE) {
			const int kk = (KM & SLP) ? 0 : (KM & DLP) ? 1 : (KM & DLPT) ? 2 : 3;
			const double re = a.alpha.x * s[kk].x - a.alpha.y * s[kk].y;
			const double im = a.alpha.x * s[kk].y + a.alpha.y * s[kk].x;
			store_entry64(a.out0, idx, a.dtype, a.accumulate, re, im);
		} else {
			// A = alpha D + H, B = G + alpha Ht
			const double ar = a.alpha.x * s[3].x - a.alpha.y * s[3].y + s[1].x;
			const double ai = a.alpha.x * s[3].y + a.alpha.y * s[3].x + s[1].y;
			const double br = s[0].x + a.alpha.x * s[2].x - a.alpha.y * s[2].y;
			const double bi = s[0].y + a.alpha.x * s[2].y + a.alpha.y * s[2].x;
			store_entry64(a.out0, idx, a.dtype, a.accumulate, ar, ai);
			store_entry64(a.out1, idx, a.dtype, a.accumulate, br, bi);
		}
	}
}

// ---------------------------------------------------------------- near list, classes, identity

struct ListArgs {
	int n_rows, n_cols, same_mesh, row_tiles, chunk, n_chunks, tn, sn;
	const double *tc, *sc, *thr_near;
	int *counts;            // [n_rows][n_chunks]
	const int *offsets;     // exclusive scan of counts
	int *pi, *pj;
};

template <bool FILL>
__global__ void __launch_bounds__(256) near_list_kernel(ListArgs a)
{
	const int tile_r = blockIdx.x % a.row_tiles;
	const int ch = blockIdx.x / a.row_tiles;
	const int i = tile_r * 256 + threadIdx.x;
	if (i >= a.n_rows) return;
	const double cix = a.tc[i], ciy = a.tc[a.tn + i], ciz = a.tc[2 * a.tn + i];
	const int j0 = ch * a.chunk, j1 = min(j0 + a.chunk, a.n_cols);
	int cnt = 0;
	int base = FILL ? a.offsets[(long long)i * a.n_chunks + ch] : 0;
	for (int j = j0; j < j1; ++j) {
		const double d2 = d2_pair(cix, ciy, ciz, a.sc[j], a.sc[a.sn + j], a.sc[2 * a.sn + j]);
		if (d2 < a.thr_near[j] && !(a.same_mesh && i == j)) {
			if (FILL) { a.pi[base + cnt] = i; a.pj[base + cnt] = j; }
			++cnt;
		}
	}
	if (!FILL) a.counts[(long long)i * a.n_chunks + ch] = cnt;
}

struct ClassArgs {
	int n_rows, n_cols, same_mesh, tn, sn, order, fixed_degree, field_complexity;
	const double *tc, *sc, *thr;
	int8_t *cls;
};

__global__ void __launch_bounds__(256) classify_kernel(ClassArgs a)
{
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)a.n_rows * a.n_cols) return;
	const int i = (int)(t / a.n_cols), j = (int)(t % a.n_cols);
	const double d2 = d2_pair(a.tc[i], a.tc[a.tn + i], a.tc[2 * a.tn + i], a.sc[j], a.sc[a.sn + j], a.sc[2 * a.sn + j]);
	int c;
	if (a.same_mesh && i == j) c = -1;
	else if (d2 < a.thr[THR_NEAR * a.sn + j]) c = -2;
	else if (a.fixed_degree >= 0) c = a.fixed_degree;
	else if (a.order == 1) c = a.field_complexity + table_degree<1>(d2, a.thr, a.sn, j);
	else if (a.order == 2) c = a.field_complexity + table_degree<2>(d2, a.thr, a.sn, j);
	else c = a.field_complexity + table_degree<3>(d2, a.thr, a.sn, j);
	a.cls[t] = (int8_t)c;
}

// The FP64 routines of the near-singular and singular passes, evaluated
// elementwise (fn 0: sincos64 -> (sin, cos); 1: log64; 2: rsqrt64; 3: rcp64;
// one value each in y[2 i], y[2 i + 1] = 0 for the last three).
__global__ void math_kernel(int fn, const double *x, double *y, long long n)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double v = x[i];
	double a = 0.0, b = 0.0;
	switch (fn) {
	case 0: sincos64(v, &a, &b); break;
	case 1: a = log64(v); break;
	case 2: a = rsqrt64(v); break;
	default: a = rcp64(v); break;
	}
	y[2 * i] = a;
	y[2 * i + 1] = b;
}

__global__ void identity_kernel(void *out, long long ld, int n, int dtype, double are, double aim)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) store_entry64(out, (long long)i * ld + i, dtype, 1, are, aim);
}

// zero a column-major rows x cols block with leading dimension ld (a fill
// kernel rather than a memset so that callers may capture it in a graph):
// columns along the grid's y, 16-byte stores where a column allows them
__global__ void zero_kernel(void *out, long long ld, long long rows, long long cols, int elem_bytes)
{
	const long long col_bytes = rows * elem_bytes, ld_bytes = ld * elem_bytes;
	const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x, dt = (long long)gridDim.x * blockDim.x;
	for (long long c = blockIdx.y; c < cols; c += gridDim.y) {
		char *base = (char *)out + c * ld_bytes;
		if ((((size_t)base | (size_t)col_bytes) & 15) == 0) {
			uint4 *q = (uint4 *)base;
			for (long long w = t0; w < col_bytes / 16; w += dt) q[w] = make_uint4(0u, 0u, 0u, 0u);
		} else {
			// elements of 4, 8 or 16 bytes: every column starts on 4 bytes
			uint32_t *q = (uint32_t *)base;
			for (long long w = t0; w < col_bytes / 4; w += dt) q[w] = 0u;
		}
	}
}

} // namespace nbk
