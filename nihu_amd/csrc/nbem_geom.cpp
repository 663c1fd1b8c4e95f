// nbem_geom.cpp -- host-side FP64 geometry and quadrature tables of the engine.
//
// Compiled with -ffp-contract=off: element centres, sizes and the pair
// classification thresholds must round exactly as written, so that the GPU's
// comparisons  d2 < threshold  reproduce NiHu's  sqrt(d2) / R < x  tests.
//
// Geometry of a flat 3-node triangle (tria_1_elem):
//   x(xi, eta) = (1 - xi - eta) X0 + xi X1 + eta X2   (nihu/library/lib_shape.hpp:134-199)
//   centre      = x(1/3, 1/3)                          (nihu/core/element.hpp:276, lib_domain.cpp:70)
//   jacobian    = (X1 - X0) x (X2 - X0), J = |.|       (nihu/core/element.hpp:503-516)
//   R           = mean edge length over edges 01,12,20 (nihu/core/element.hpp:281-291)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

namespace {

struct P3 { double x, y, z; };

inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double nrm(P3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

inline P3 shape_x(const P3 X[3], double xi, double eta)
{
	double N0 = 1.0 - xi - eta, N1 = xi, N2 = eta;
	return {X[0].x * N0 + X[1].x * N1 + X[2].x * N2,
		X[0].y * N0 + X[1].y * N1 + X[2].y * N2,
		X[0].z * N0 + X[1].z * N1 + X[2].z * N2};
}

// Dunavant rules (weights sum to 1/2) as orbits of the symmetry group of the
// triangle: kind 1 = centre (a, a); kind 3 = (a, a) and its two images with b;
// kind 6 = all six arrangements of (a, b, c).  Constants of the published
// table, 15 significant digits (nihu/core/gaussian_quadrature.hpp:258-507).
struct Orbit { int kind; double a, b, c, w; };

const double T3 = 1. / 3.;
const double C15 = 0.333333333333333;

const std::vector<Orbit> &dunavant_orbits(int degree)
{
	static const std::vector<Orbit> table[15] = {
		/* 0 */ {{1, T3, 0, 0, 0.5}},
		/* 1 */ {{1, T3, 0, 0, 0.5}},
		/* 2 */ {{3, 1. / 6., 4. / 6., 0, 1. / 6.}},
		/* 3 */ {{1, T3, 0, 0, -0.281250000000000}, {3, 1. / 5., 3. / 5., 0, 0.260416666666667}},
		/* 4 */ {{3, 0.445948490915965, 0.108103018168070, 0, 0.111690794839006},
			{3, 0.091576213509771, 0.816847572980459, 0, 0.054975871827661}},
		/* 5 */ {{1, C15, 0, 0, 0.112500000000000},
			{3, 0.470142064105115, 0.059715871789770, 0, 0.066197076394253},
			{3, 0.101286507323456, 0.797426985353087, 0, 0.062969590272414}},
		/* 6 */ {{3, 0.249286745170910, 0.501426509658179, 0, 0.058393137863189},
			{3, 0.063089014491502, 0.873821971016996, 0, 0.025422453185104},
			{6, 0.310352451033784, 0.053145049844817, 0.636502499121399, 0.041425537809187}},
		/* 7 */ {{1, C15, 0, 0, -0.074785022233841},
			{3, 0.260345966079040, 0.479308067841920, 0, 0.087807628716604},
			{3, 0.065130102902216, 0.869739794195568, 0, 0.026673617804419},
			{6, 0.312865496004874, 0.048690315425316, 0.638444188569810, 0.038556880445128}},
		/* 8 */ {{1, C15, 0, 0, 0.072157803838894},
			{3, 0.459292588292723, 0.081414823414554, 0, 0.047545817133642},
			{3, 0.170569307751760, 0.658861384496480, 0, 0.051608685267359},
			{3, 0.050547228317031, 0.898905543365938, 0, 0.016229248811599},
			{6, 0.263112829634638, 0.008394777409958, 0.728492392955404, 0.013615157087218}},
		/* 9 */ {{1, C15, 0, 0, 0.048567898141400},
			{3, 0.489682519198738, 0.020634961602525, 0, 0.015667350113570},
			{3, 0.437089591492937, 0.125820817014127, 0, 0.038913770502387},
			{3, 0.188203535619033, 0.623592928761935, 0, 0.039823869463605},
			{3, 0.044729513394453, 0.910540973211095, 0, 0.012788837829349},
			{6, 0.221962989160766, 0.036838412054736, 0.741198598784498, 0.021641769688645}},
		/* 10 */ {{1, C15, 0, 0, 0.045408995191377},
			{3, 0.485577633383657, 0.028844733232685, 0, 0.018362978878233},
			{3, 0.109481575485037, 0.781036849029926, 0, 0.022660529717764},
			{6, 0.141707219414880, 0.307939838764121, 0.550352941820999, 0.036378958422710},
			{6, 0.025003534762686, 0.246672560639903, 0.728323904597411, 0.014163621265528},
			{6, 0.009540815400299, 0.066803251012200, 0.923655933587500, 0.004710833481867}},
		/* 11 = 12 */ {},
		/* 12 */ {{3, 0.488217389773805, 0.023565220452390, 0, 0.012865533220227},
			{3, 0.439724392294460, 0.120551215411079, 0, 0.021846272269019},
			{3, 0.271210385012116, 0.457579229975768, 0, 0.031429112108943},
			{3, 0.127576145541586, 0.744847708916828, 0, 0.017398056465355},
			{3, 0.021317350453210, 0.957365299093579, 0, 0.003083130525780},
			{6, 0.115343494534698, 0.275713269685514, 0.608943235779788, 0.020185778883191},
			{6, 0.022838332222257, 0.281325580989940, 0.695836086787803, 0.011178386601152},
			{6, 0.025734050548330, 0.116251915907597, 0.858014033544073, 0.008658115554329}},
		/* 13 */ {{1, C15, 0, 0, 0.026260461700401},
			{3, 0.495048184939705, 0.009903630120591, 0, 0.005640072604665},
			{3, 0.468716635109574, 0.062566729780852, 0, 0.015711759181227},
			{3, 0.414521336801277, 0.170957326397447, 0, 0.023536251252097},
			{3, 0.229399572042831, 0.541200855914337, 0, 0.023681793268178},
			{3, 0.114424495196330, 0.771151009607340, 0, 0.015583764522897},
			{3, 0.024811391363459, 0.950377217273082, 0, 0.003987885732537},
			{6, 0.094853828379579, 0.268794997058761, 0.636351174561660, 0.018424201364366},
			{6, 0.018100773278807, 0.291730066734288, 0.690169159986905, 0.008700731651911},
			{6, 0.022233076674090, 0.126357385491669, 0.851409537834241, 0.007760893419522}},
		/* 14 */ {{3, 0.488963910362179, 0.022072179275643, 0, 0.010941790684715},
			{3, 0.417644719340454, 0.164710561319092, 0, 0.016394176772063},
			{3, 0.273477528308839, 0.453044943382323, 0, 0.025887052253646},
			{3, 0.177205532412543, 0.645588935174913, 0, 0.021081294368497},
			{3, 0.061799883090873, 0.876400233818255, 0, 0.007216849834889},
			{3, 0.019390961248701, 0.961218077502598, 0, 0.002461701801200},
			{6, 0.057124757403648, 0.172266687821356, 0.770608554774996, 0.012332876606282},
			{6, 0.092916249356972, 0.336861459796345, 0.570222290846683, 0.019285755393531},
			{6, 0.014646950055654, 0.298372882136258, 0.686980167808088, 0.007218154056767},
			{6, 0.001268330932872, 0.118974497696957, 0.879757171370171, 0.002505114419250}},
	};
	return table[degree == 11 ? 12 : degree];
}

// Legendre P_n and derivative by the three-term recurrence
void legendre(int n, double x, double &p, double &dp)
{
	double p0 = 1.0, p1 = x;
	for (int k = 2; k <= n; ++k) {
		double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
		p0 = p1;
		p1 = pk;
	}
	p = n == 0 ? 1.0 : p1;
	dp = n == 0 ? 0.0 : n * (x * p1 - p0) / (x * x - 1.0);
}

} // namespace

extern "C" {

// Points (xi, eta, w) of the Dunavant rule of a degree; returns the count or -1.
int nbg_dunavant(int degree, double *xi, double *eta, double *w, int cap)
{
	if (degree < 0 || degree > 14) return -1;
	int n = 0;
	auto put = [&](double a, double b, double wt) {
		if (n < cap) { xi[n] = a; eta[n] = b; w[n] = wt; }
		++n;
	};
	for (auto const &o : dunavant_orbits(degree)) {
		if (o.kind == 1) put(o.a, o.a, o.w);
		else if (o.kind == 3) { put(o.a, o.b, o.w); put(o.a, o.a, o.w); put(o.b, o.a, o.w); }
		else {
			put(o.a, o.b, o.w); put(o.b, o.c, o.w); put(o.c, o.a, o.w);
			put(o.b, o.a, o.w); put(o.c, o.b, o.w); put(o.a, o.c, o.w);
		}
	}
	return n <= cap ? n : -1;
}

// n-point Gauss-Legendre rule on [-1, 1], ascending nodes (Newton on P_n).
void nbg_gauss_legendre(int n, double *x, double *w)
{
	const double pi = 3.141592653589793238462643383279502884;
	for (int i = 0; i < n; ++i) {
		double t = std::cos(pi * (i + 0.75) / (n + 0.5));
		for (int it = 0; it < 100; ++it) {
			double p, dp;
			legendre(n, t, p, dp);
			double d = p / dp;
			t -= d;
			if (std::abs(d) < 1e-17) break;
		}
		double p, dp;
		legendre(n, t, p, dp);
		x[n - 1 - i] = t;
		w[n - 1 - i] = 2.0 / ((1.0 - t * t) * dp * dp);
	}
}

// Element geometry.  Arrays are structure-of-arrays: c[3][n], nu[3][n],
// X[9][n] (X0x X0y X0z X1x ... X2z).  Returns 0, or 1 for a bad node index,
// 2 for a degenerate element.
int nbg_geometry(const double *xyz, int n_nodes, const int32_t *tri, int n, double *c, double *nu,
	double *J, double *R, double *diam, double *X)
{
	for (int e = 0; e < n; ++e) {
		P3 V[3];
		for (int a = 0; a < 3; ++a) {
			int v = tri[3 * e + a];
			if (v < 0 || v >= n_nodes) return 1;
			V[a] = {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
		}
		P3 cc = shape_x(V, 1. / 3., 1. / 3.);
		P3 e1 = sub(V[1], V[0]), e2 = sub(V[2], V[0]);
		P3 nv = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
		double j = nrm(nv);
		if (!(j > 0.0)) return 2;
		double l0 = nrm(sub(V[1], V[0])), l1 = nrm(sub(V[2], V[1])), l2 = nrm(sub(V[0], V[2]));
		double d = 0.0;
		d += l0; d += l1; d += l2;
		c[e] = cc.x; c[n + e] = cc.y; c[2 * n + e] = cc.z;
		nu[e] = nv.x / j; nu[n + e] = nv.y / j; nu[2 * n + e] = nv.z / j;
		J[e] = j;
		R[e] = d / 3.0;
		diam[e] = std::max(l0, std::max(l1, l2));
		for (int a = 0; a < 3; ++a) {
			X[(3 * a) * n + e] = V[a].x;
			X[(3 * a + 1) * n + e] = V[a].y;
			X[(3 * a + 2) * n + e] = V[a].z;
		}
	}
	return 0;
}

// Smallest d2 >= 0 for which  sqrt(d2) / scale < limit  is false, so that the
// test is exactly  d2 < threshold  (both sqrt and / are correctly rounded and
// monotone).  +inf when the limit is infinite.
double nbg_threshold(double scale, double limit)
{
	if (!(limit < 1e300)) return std::numeric_limits<double>::infinity();
	if (!(limit > 0.0)) return 0.0;
	auto pred = [&](double d2) { return std::sqrt(d2) / scale < limit; };
	double g = (limit * scale) * (limit * scale);
	const double inf = std::numeric_limits<double>::infinity();
	if (pred(g)) {
		while (pred(g)) g = std::nextafter(g, inf);
		return g;
	}
	for (;;) {
		double p = std::nextafter(g, 0.0);
		if (p == g || pred(p)) return g;
		g = p;
	}
}

// Quadrature point table of one Dunavant degree for all elements:
// out[(e * Q + q) * 8 + {0..7}] = (y_q - c_e) xyz, w_q * J_e / (4 pi), |y_q - c_e|^2,
// w_q * J_e / (4 pi) * xi_q, w_q * J_e / (4 pi) * eta_q, w_q * J_e (floats
// rounded from FP64).
int nbg_quad_table(int degree, int n, const double *c, const double *J, const double *X, float *out)
{
	double xi[64], eta[64], w[64];
	int Q = nbg_dunavant(degree, xi, eta, w, 64);
	if (Q < 0) return -1;
	const double four_pi = 4.0 * 3.141592653589793238462643383279502884;
	for (int e = 0; e < n; ++e) {
		P3 V[3];
		for (int a = 0; a < 3; ++a)
			V[a] = {X[(3 * a) * n + e], X[(3 * a + 1) * n + e], X[(3 * a + 2) * n + e]};
		P3 cc = {c[e], c[n + e], c[2 * n + e]};
		for (int q = 0; q < Q; ++q) {
			P3 o = sub(shape_x(V, xi[q], eta[q]), cc);
			float *p = out + ((size_t)e * Q + q) * 8;
			p[0] = (float)o.x; p[1] = (float)o.y; p[2] = (float)o.z;
			p[3] = (float)(w[q] * J[e] / four_pi);
			p[4] = (float)(o.x * o.x + o.y * o.y + o.z * o.z);
			p[5] = (float)(w[q] * J[e] / four_pi * xi[q]);
			p[6] = (float)(w[q] * J[e] / four_pi * eta[q]);
			p[7] = (float)(w[q] * J[e]);
		}
	}
	return Q;
}

// The same points in FP64 for the FP64 regular pass:
// out[(e * Q + q) * 4 + {0..3}] = (y_q - c_e) xyz, w_q * J_e / (4 pi).
int nbg_quad_table64(int degree, int n, const double *c, const double *J, const double *X, double *out)
{
	double xi[64], eta[64], w[64];
	int Q = nbg_dunavant(degree, xi, eta, w, 64);
	if (Q < 0) return -1;
	const double four_pi = 4.0 * 3.141592653589793238462643383279502884;
	for (int e = 0; e < n; ++e) {
		P3 V[3];
		for (int a = 0; a < 3; ++a)
			V[a] = {X[(3 * a) * n + e], X[(3 * a + 1) * n + e], X[(3 * a + 2) * n + e]};
		P3 cc = {c[e], c[n + e], c[2 * n + e]};
		for (int q = 0; q < Q; ++q) {
			P3 o = sub(shape_x(V, xi[q], eta[q]), cc);
			double *p = out + ((size_t)e * Q + q) * 4;
			p[0] = o.x; p[1] = o.y; p[2] = o.z;
			p[3] = w[q] * J[e] / four_pi;
		}
	}
	return Q;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Quadrilaterals (quad_1_elem): x(xi, eta) = sum_a N_a X_a on [-1, 1]^2 with
// N = (1 -+ xi)(1 -+ eta) / 4 for the corners (-1,-1), (1,-1), (1,1), (-1,1)
// (nihu/library/lib_shape.hpp:215-310, lib_domain.cpp:74-94); centre x(0, 0),
// R = mean edge length over edges 01 12 23 30 (nihu/core/element.hpp:281-291),
// diameter = largest corner distance.  Sums are accumulated from zero in
// corner order, as the oracle does, so that centres and R round identically.

namespace {
const double QCX[4] = {-1.0, 1.0, 1.0, -1.0}, QCY[4] = {-1.0, -1.0, 1.0, 1.0};

inline P3 quad_x(const P3 X[4], double xi, double eta)
{
	double N[4] = {(1.0 - xi) * (1.0 - eta) / 4.0, (1.0 + xi) * (1.0 - eta) / 4.0,
		(1.0 + xi) * (1.0 + eta) / 4.0, (1.0 - xi) * (1.0 + eta) / 4.0};
	P3 x = {0.0, 0.0, 0.0};
	for (int a = 0; a < 4; ++a) x = {x.x + X[a].x * N[a], x.y + X[a].y * N[a], x.z + X[a].z * N[a]};
	return x;
}

inline P3 quad_normal(const P3 X[4], double xi, double eta)
{
	double A[4] = {-(1.0 - eta) / 4.0, (1.0 - eta) / 4.0, (1.0 + eta) / 4.0, -(1.0 + eta) / 4.0};
	double B[4] = {-(1.0 - xi) / 4.0, -(1.0 + xi) / 4.0, (1.0 + xi) / 4.0, (1.0 - xi) / 4.0};
	P3 a = {0.0, 0.0, 0.0}, b = {0.0, 0.0, 0.0};
	for (int c = 0; c < 4; ++c) {
		a = {a.x + X[c].x * A[c], a.y + X[c].y * A[c], a.z + X[c].z * A[c]};
		b = {b.x + X[c].x * B[c], b.y + X[c].y * B[c], b.z + X[c].z * B[c]};
	}
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

inline void load_quad(const double *X, int n, int e, P3 V[4])
{
	for (int a = 0; a < 4; ++a) V[a] = {X[(3 * a) * n + e], X[(3 * a + 1) * n + e], X[(3 * a + 2) * n + e]};
}
} // namespace

extern "C" {

// Quadrilateral geometry: c[3][n], nu[3][n] (unit normal at the centre),
// J[n] (|normal| at the centre), R[n], diam[n], X[12][n].  Returns 0, 1 for a
// bad node index, 2 for a degenerate element (zero normal at a Gauss point of
// order 2 or at the centre).
int nbg_geometry_quad(const double *xyz, int n_nodes, const int32_t *quad, int n, double *c, double *nu,
	double *J, double *R, double *diam, double *X)
{
	for (int e = 0; e < n; ++e) {
		P3 V[4];
		for (int a = 0; a < 4; ++a) {
			int v = quad[4 * e + a];
			if (v < 0 || v >= n_nodes) return 1;
			V[a] = {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
		}
		P3 cc = quad_x(V, 0.0, 0.0);
		P3 nv = quad_normal(V, 0.0, 0.0);
		double j = nrm(nv);
		if (!(j > 0.0)) return 2;
		const double g = 0.577350269189625764509148780502;
		for (int q = 0; q < 4; ++q)
			if (!(nrm(quad_normal(V, (q & 1) ? g : -g, (q & 2) ? g : -g)) > 0.0)) return 2;
		double d = 0.0;
		for (int a = 0; a < 4; ++a) d += nrm(sub(V[(a + 1) % 4], V[a]));
		double dm = 0.0;
		for (int a = 0; a < 4; ++a)
			for (int b = a + 1; b < 4; ++b) dm = std::max(dm, nrm(sub(V[b], V[a])));
		c[e] = cc.x; c[n + e] = cc.y; c[2 * n + e] = cc.z;
		nu[e] = nv.x / j; nu[n + e] = nv.y / j; nu[2 * n + e] = nv.z / j;
		J[e] = j;
		R[e] = d / 4.0;
		diam[e] = dm;
		for (int a = 0; a < 4; ++a) {
			X[(3 * a) * n + e] = V[a].x;
			X[(3 * a + 1) * n + e] = V[a].y;
			X[(3 * a + 2) * n + e] = V[a].z;
		}
	}
	return 0;
}

// Gauss square rule with N x N points (xi outer, eta inner) for all
// quadrilaterals: out[(e * N*N + q) * 8 + {0..7}] = (y_q - c_e) xyz,
// w_q J(xi_q) / (4 pi), unit normal at y_q xyz, |y_q - c_e|^2 (floats).
int nbg_quad4_table(int N, int n, const double *c, const double *X, float *out)
{
	if (N < 1 || N > 64) return -1;
	std::vector<double> gx(N), gw(N);
	nbg_gauss_legendre(N, gx.data(), gw.data());
	const double four_pi = 4.0 * 3.141592653589793238462643383279502884;
	for (int e = 0; e < n; ++e) {
		P3 V[4];
		load_quad(X, n, e, V);
		P3 cc = {c[e], c[n + e], c[2 * n + e]};
		for (int i = 0; i < N; ++i)
			for (int j = 0; j < N; ++j) {
				P3 y = quad_x(V, gx[i], gx[j]);
				P3 nv = quad_normal(V, gx[i], gx[j]);
				double J = nrm(nv);
				P3 o = sub(y, cc);
				float *p = out + ((size_t)e * N * N + (size_t)i * N + j) * 8;
				p[0] = (float)o.x; p[1] = (float)o.y; p[2] = (float)o.z;
				p[3] = (float)(gw[i] * gw[j] * J / four_pi);
				p[4] = (float)(nv.x / J); p[5] = (float)(nv.y / J); p[6] = (float)(nv.z / J);
				p[7] = (float)(o.x * o.x + o.y * o.y + o.z * o.z);
			}
	}
	return N * N;
}

// Gauss square rule with N x N points (xi outer, eta inner) for the Galerkin
// passes on quadrilaterals: out[(e * N*N + q) * 16 + {0..15}] =
// o = y_q - c_e (xyz), |o|^2, unit normal n at y_q (xyz), o.n,
// w_q J(xi_q) N_b(xi_q) for the four corners b, w_q J(xi_q), 0, 0, 0 (floats).
int nbg_quad_gal_table(int N, int n, const double *c, const double *X, float *out)
{
	if (N < 1 || N > 64) return -1;
	std::vector<double> gx(N), gw(N);
	nbg_gauss_legendre(N, gx.data(), gw.data());
	for (int e = 0; e < n; ++e) {
		P3 V[4];
		load_quad(X, n, e, V);
		const P3 cc = {c[e], c[n + e], c[2 * n + e]};
		for (int i = 0; i < N; ++i)
			for (int j = 0; j < N; ++j) {
				const double xi = gx[i], eta = gx[j];
				const P3 y = quad_x(V, xi, eta);
				const P3 nv = quad_normal(V, xi, eta);
				const double J = nrm(nv);
				const P3 o = sub(y, cc);
				const P3 u = {nv.x / J, nv.y / J, nv.z / J};
				const double w = gw[i] * gw[j] * J;
				const double N4[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
					0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)};
				float *p = out + ((size_t)e * N * N + (size_t)i * N + j) * 16;
				p[0] = (float)o.x; p[1] = (float)o.y; p[2] = (float)o.z;
				p[3] = (float)(o.x * o.x + o.y * o.y + o.z * o.z);
				p[4] = (float)u.x; p[5] = (float)u.y; p[6] = (float)u.z;
				p[7] = (float)(o.x * u.x + o.y * u.y + o.z * u.z);
				for (int b = 0; b < 4; ++b) p[8 + b] = (float)(w * N4[b]);
				p[12] = (float)w;
				p[13] = p[14] = p[15] = 0.0f;
			}
	}
	return N * N;
}

// FP64 Gauss square rule with N x N points (xi outer, eta inner) for all
// quadrilaterals: out[(e * N*N + q) * 8 + {0..7}] = y_q xyz, w_q J(xi_q),
// unit normal at y_q xyz, 0 (the FP64 precision policy's regular pass).
int nbg_quad4_table64(int N, int n, const double *X, double *out)
{
	if (N < 1 || N > 64) return -1;
	std::vector<double> gx(N), gw(N);
	nbg_gauss_legendre(N, gx.data(), gw.data());
	for (int e = 0; e < n; ++e) {
		P3 V[4];
		load_quad(X, n, e, V);
		for (int i = 0; i < N; ++i)
			for (int j = 0; j < N; ++j) {
				const P3 y = quad_x(V, gx[i], gx[j]);
				const P3 nv = quad_normal(V, gx[i], gx[j]);
				const double J = nrm(nv);
				double *p = out + ((size_t)e * N * N + (size_t)i * N + j) * 8;
				p[0] = y.x; p[1] = y.y; p[2] = y.z;
				p[3] = gw[i] * gw[j] * J;
				p[4] = nv.x / J; p[5] = nv.y / J; p[6] = nv.z / J;
				p[7] = 0.0;
			}
	}
	return N * N;
}

// Nodes of the quadrilateral gauss field (nihu/library/lib_shape.cpp:128-134:
// (-,-), (+,-), (-,+), (+,+) at +-sqrt(3)/3): node a of element e is row
// 4 e + a of x[3][4n] (the collocation point x(xi_a)) and nu[3][4n] (the unit
// normal there).
int nbg_gauss_nodes(int n, const double *X, double *x, double *nu)
{
	const int m = 4 * n;
	const double g = std::sqrt(3.0) / 3.0;
	for (int e = 0; e < n; ++e) {
		P3 V[4];
		load_quad(X, n, e, V);
		for (int a = 0; a < 4; ++a) {
			const double xi = (a & 1) ? g : -g, eta = (a & 2) ? g : -g;
			const P3 y = quad_x(V, xi, eta);
			const P3 nv = quad_normal(V, xi, eta);
			const double j = nrm(nv);
			if (!(j > 0.0)) return 2;
			const int r = 4 * e + a;
			x[r] = y.x; x[m + r] = y.y; x[2 * m + r] = y.z;
			nu[r] = nv.x / j; nu[m + r] = nv.y / j; nu[2 * m + r] = nv.z / j;
		}
	}
	return 0;
}

// Shape functions of the gauss field, N_b = (1 + s_b sqrt3 xi)(1 + t_b sqrt3
// eta) / 4 (quad_1_gauss_shape_set.hpp), at the N x N Gauss square rule's
// points in the order of nbg_quad4_table: out[q * 4 + b].
int nbg_gauss_shapes(int N, double *out)
{
	if (N < 1 || N > 64) return -1;
	std::vector<double> gx(N), gw(N);
	nbg_gauss_legendre(N, gx.data(), gw.data());
	const double s3 = std::sqrt(3.0);
	for (int i = 0; i < N; ++i)
		for (int j = 0; j < N; ++j)
			for (int b = 0; b < 4; ++b) {
				const double sx = (b & 1) ? 1.0 : -1.0, sy = (b & 2) ? 1.0 : -1.0;
				out[(i * N + j) * 4 + b] = (1.0 + sx * s3 * gx[i]) * (1.0 + sy * s3 * gx[j]) / 4.0;
			}
	return N * N;
}

// Duffy-type blind rule around (x0, y0) of [-1, 1]^2 from the Gauss square
// rule with ng x ng points: four triangles (x0, corner s, corner s+1), each
// the image of the square under the bilinear map of the degenerate
// quadrilateral (x0, x0, corner s, corner s+1), weights times the Jacobian
// (nihu/core/duffy_quadrature.hpp, nihu/core/quadrature.hpp:105-127).
// out: 4 ng^2 rows (xi, eta, w).
int nbg_quad_duffy(int ng, double x0, double y0, double *out)
{
	std::vector<double> gx(ng), gw(ng);
	nbg_gauss_legendre(ng, gx.data(), gw.data());
	int k = 0;
	for (int s = 0; s < 4; ++s) {
		const double cx[4] = {x0, x0, QCX[s], QCX[(s + 1) % 4]};
		const double cy[4] = {y0, y0, QCY[s], QCY[(s + 1) % 4]};
		for (int i = 0; i < ng; ++i)
			for (int j = 0; j < ng; ++j) {
				const double u = gx[i], v = gx[j];
				const double N[4] = {(1.0 - u) * (1.0 - v) / 4.0, (1.0 + u) * (1.0 - v) / 4.0,
					(1.0 + u) * (1.0 + v) / 4.0, (1.0 - u) * (1.0 + v) / 4.0};
				const double Du[4] = {-(1.0 - v) / 4.0, (1.0 - v) / 4.0, (1.0 + v) / 4.0, -(1.0 + v) / 4.0};
				const double Dv[4] = {-(1.0 - u) / 4.0, -(1.0 + u) / 4.0, (1.0 + u) / 4.0, (1.0 - u) / 4.0};
				double px = 0, py = 0, a = 0, b = 0, cc = 0, d = 0;
				for (int t = 0; t < 4; ++t) {
					px += N[t] * cx[t]; py += N[t] * cy[t];
					a += Du[t] * cx[t]; b += Du[t] * cy[t];
					cc += Dv[t] * cx[t]; d += Dv[t] * cy[t];
				}
				out[3 * k] = px; out[3 * k + 1] = py; out[3 * k + 2] = gw[i] * gw[j] * (a * d - cc * b);
				++k;
			}
	}
	return k;
}

} // extern "C"
