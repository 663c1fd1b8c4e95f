// nbem_oracle.cpp -- FP64 CPU restatement of NiHu's dense BEM assembly for
// flat triangles (collocation with constant fields, Galerkin with constant
// and linear isoparametric fields) and bilinear quadrilaterals (collocation
// with constant fields and with the gauss field, Galerkin with constant and
// bilinear isoparametric fields), 3D Laplace and Helmholtz SLP / DLP / DLPt /
// HSP, real and complex wavenumbers.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker of the HIP engine in
// nihu_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load the library built from it, and only as the checker or as the
// timed CPU baseline -- never as the thing measured or shipped.
//
// Every routine cites the NiHu source it restates (paths relative to the
// NiHu repository root).  NiHu itself cannot be built here (it needs Eigen and
// Boost), so parity is pinned by:
//   * the golden element integrals of nihu/gtest (near-singular SLP, Laplace
//     and Helmholtz),
//   * the finite-difference kernel check of nihu/gtest,
//   * closed forms and brute-force quadrature (diagonal integrals, near
//     integrals, Gauss-Legendre nodes, Dunavant exactness),
//   * known-answer systems (transparent point sources, pulsating sphere),
//   * for the Galerkin part: exact integrals of polynomials over two
//     triangles (the Duffy-type rules), brute-force integration of coincident,
//     edge- and vertex-adjacent pairs under every labelling, the closed-form
//     solid angle, the analytic self-potential of a square (NiHu's
//     lenoir_salles_2012 test), Gauss's identity, the mass-matrix closed form.
//   * for quadrilaterals (collocation, constant fields): closed forms of the
//     planar SLP and hypersingular integrals over polygons (Duffy rule and
//     Guiggiani's method at NiHu's guiggiani_test shapes and points), an
//     independent Hadamard finite part on warped elements (integral outside a
//     ball minus 1/(2 eps), Richardson), brute-force subdivided quadrature of
//     the near-singular methods (polar with the linearised element, Telles),
//     the planar part equal to the sum over two golden-pinned triangles,
//     inverse mapping round trips, Gauss's identity and a transparent
//     point-source answer on a cubed sphere;
//   * for Galerkin on quadrilaterals: the closed-form double integral of 1/r
//     over the unit square, a semi-analytic value for squares sharing an
//     edge or a corner (closed-form inner integral, graded outer rule), every
//     cyclic relabelling of element pairs, brute-force tensor rules for
//     regular warped pairs, the mass matrix of a square, Gauss's identity and
//     the pulsating sphere.
// The node-overlap scan keeps NiHu's start-corner rule, including the
// labelling in which it misses an edge's orientation (documented in the
// Galerkin tests).
// Deliberate, documented departures from NiHu's arithmetic (all at the level
// of one rounding):
//   * the 3x3 plane rotation is applied as T^T (T is orthonormal) where NiHu
//     uses Eigen's partialPivLu().solve;
//   * Gauss-Legendre nodes come from Newton iteration on P_n where NiHu uses
//     Eigen's Golub-Welsch eigen-solve;
//   * squared distances for pair classification are fma(dz,dz,fma(dy,dy,dx*dx)),
//     the convention the GPU engine shares bit for bit.
// The file is compiled with -ffp-contract=off so that no other fusion happens.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::complex<double> cplx;
static const double PI = 3.141592653589793238462643383279502884;   // boost double_constants::pi
static const double TWO_PI = 6.283185307179586476925286766559005768;

struct V3 { double x, y, z; };
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalized(V3 a) { return a / norm(a); }

// sgn and sinc: nihu/util/math_functions.hpp:42-61
static inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }
static inline double sinc(double x)
{
	if (std::abs(x) > 1e-3)
		return std::sin(x) / x;
	return 1. - x * x / 6. * (1. - x * x / 20.);
}
// the same for a complex argument (complex wavenumbers)
static inline std::complex<double> sinc(std::complex<double> x)
{
	if (std::abs(x) > 1e-3)
		return std::sin(x) / x;
	return 1. - x * x / 6. * (1. - x * x / 20.);
}

// ---------------------------------------------------------------------------
// Quadrature
// ---------------------------------------------------------------------------
struct QP { double xi, eta, w; };

// Dunavant rules on the reference triangle (0,0),(1,0),(0,1), weights summing
// to 1/2; nihu/core/gaussian_quadrature.hpp:209-512 (point counts :209-225).
// Written out as explicit point lists in NiHu's order.
static std::vector<QP> dunavant(int degree)
{
	const double t3 = 1. / 3.;
	const double c = 0.333333333333333;
	switch (degree) {
	case 0: case 1:
		return {{t3, t3, 0.5}};
	case 2:
		return {{1. / 6., 4. / 6., 1. / 6.}, {1. / 6., 1. / 6., 1. / 6.}, {4. / 6., 1. / 6., 1. / 6.}};
	case 3:
		return {{t3, t3, -0.281250000000000}, {1. / 5., 3. / 5., 0.260416666666667},
			{1. / 5., 1. / 5., 0.260416666666667}, {3. / 5., 1. / 5., 0.260416666666667}};
	default: break;
	}
	// higher orders: orbits {a, b, w} (two-coordinate orbit a,b,a with b=1-2a
	// listed as in NiHu) and {a, b, c, w} (six permutations)
	std::vector<QP> q;
	auto s21 = [&](double a, double b, double w) {
		q.push_back({a, b, w}); q.push_back({a, a, w}); q.push_back({b, a, w});
	};
	auto s21r = [&](double a, double b, double w) {   // NiHu's alternative listing (b,a),(a,a),(a,b)
		q.push_back({b, a, w}); q.push_back({a, a, w}); q.push_back({a, b, w});
	};
	auto s111 = [&](double a, double b, double c3, double w) {
		q.push_back({a, b, w}); q.push_back({c3, a, w}); q.push_back({b, c3, w});
		q.push_back({b, a, w}); q.push_back({a, c3, w}); q.push_back({c3, b, w});
	};
	auto s111b = [&](double a, double b, double c3, double w) {   // (a,b),(b,c),(c,a),(b,a),(c,b),(a,c)
		q.push_back({a, b, w}); q.push_back({b, c3, w}); q.push_back({c3, a, w});
		q.push_back({b, a, w}); q.push_back({c3, b, w}); q.push_back({a, c3, w});
	};
	switch (degree) {
	case 4:
		s21(0.445948490915965, 0.108103018168070, 0.111690794839006);
		s21(0.091576213509771, 0.816847572980459, 0.054975871827661);
		break;
	case 5:
		q.push_back({c, c, 0.112500000000000});
		s21(0.470142064105115, 0.059715871789770, 0.066197076394253);
		s21(0.101286507323456, 0.797426985353087, 0.062969590272414);
		break;
	case 6:
		s21(0.249286745170910, 0.501426509658179, 0.058393137863189);
		s21(0.063089014491502, 0.873821971016996, 0.025422453185104);
		s111(0.310352451033784, 0.053145049844817, 0.636502499121399, 0.041425537809187);
		break;
	case 7:
		q.push_back({c, c, -0.074785022233841});
		s21(0.260345966079040, 0.479308067841920, 0.087807628716604);
		s21(0.065130102902216, 0.869739794195568, 0.026673617804419);
		s111(0.312865496004874, 0.048690315425316, 0.638444188569810, 0.038556880445128);
		break;
	case 8:
		q.push_back({c, c, 0.072157803838894});
		s21(0.459292588292723, 0.081414823414554, 0.047545817133642);
		s21(0.170569307751760, 0.658861384496480, 0.051608685267359);
		s21(0.050547228317031, 0.898905543365938, 0.016229248811599);
		s111(0.263112829634638, 0.008394777409958, 0.728492392955404, 0.013615157087218);
		break;
	case 9:
		q.push_back({c, c, 0.048567898141400});
		s21(0.489682519198738, 0.020634961602525, 0.015667350113570);
		s21(0.437089591492937, 0.125820817014127, 0.038913770502387);
		s21(0.188203535619033, 0.623592928761935, 0.039823869463605);
		s21(0.044729513394453, 0.910540973211095, 0.012788837829349);
		s111(0.221962989160766, 0.036838412054736, 0.741198598784498, 0.021641769688645);
		break;
	case 10:
		q.push_back({c, c, 0.045408995191377});
		s21r(0.485577633383657, 0.028844733232685, 0.018362978878233);
		s21r(0.109481575485037, 0.781036849029926, 0.022660529717764);
		s111b(0.141707219414880, 0.307939838764121, 0.550352941820999, 0.036378958422710);
		s111b(0.025003534762686, 0.246672560639903, 0.728323904597411, 0.014163621265528);
		s111b(0.009540815400299, 0.066803251012200, 0.923655933587500, 0.004710833481867);
		break;
	case 11: case 12:   // NiHu uses the degree-12 rule for 11
		s21r(0.488217389773805, 0.023565220452390, 0.012865533220227);
		s21r(0.439724392294460, 0.120551215411079, 0.021846272269019);
		s21r(0.271210385012116, 0.457579229975768, 0.031429112108943);
		s21r(0.127576145541586, 0.744847708916828, 0.017398056465355);
		s21r(0.021317350453210, 0.957365299093579, 0.003083130525780);
		s111b(0.115343494534698, 0.275713269685514, 0.608943235779788, 0.020185778883191);
		s111b(0.022838332222257, 0.281325580989940, 0.695836086787803, 0.011178386601152);
		s111b(0.025734050548330, 0.116251915907597, 0.858014033544073, 0.008658115554329);
		break;
	case 13:
		q.push_back({c, c, 0.026260461700401});
		s21r(0.495048184939705, 0.009903630120591, 0.005640072604665);
		s21r(0.468716635109574, 0.062566729780852, 0.015711759181227);
		s21r(0.414521336801277, 0.170957326397447, 0.023536251252097);
		s21r(0.229399572042831, 0.541200855914337, 0.023681793268178);
		s21r(0.114424495196330, 0.771151009607340, 0.015583764522897);
		s21r(0.024811391363459, 0.950377217273082, 0.003987885732537);
		s111b(0.094853828379579, 0.268794997058761, 0.636351174561660, 0.018424201364366);
		s111b(0.018100773278807, 0.291730066734288, 0.690169159986905, 0.008700731651911);
		s111b(0.022233076674090, 0.126357385491669, 0.851409537834241, 0.007760893419522);
		break;
	case 14:
		s21r(0.488963910362179, 0.022072179275643, 0.010941790684715);
		s21r(0.417644719340454, 0.164710561319092, 0.016394176772063);
		s21r(0.273477528308839, 0.453044943382323, 0.025887052253646);
		s21r(0.177205532412543, 0.645588935174913, 0.021081294368497);
		s21r(0.061799883090873, 0.876400233818255, 0.007216849834889);
		s21r(0.019390961248701, 0.961218077502598, 0.002461701801200);
		s111b(0.057124757403648, 0.172266687821356, 0.770608554774996, 0.012332876606282);
		s111b(0.092916249356972, 0.336861459796345, 0.570222290846683, 0.019285755393531);
		s111b(0.014646950055654, 0.298372882136258, 0.686980167808088, 0.007218154056767);
		s111b(0.001268330932872, 0.118974497696957, 0.879757171370171, 0.002505114419250);
		break;
	default:
		// nihu/core/gaussian_quadrature.hpp:508-510
		throw std::out_of_range("unsupported dunavant degree");
	}
	return q;
}

// Gauss-Legendre rule with N = degree/2 + 1 points on [-1, 1], ascending nodes
// (nihu/core/gaussian_quadrature.hpp:64-88,136-150).  Nodes are the roots of
// P_N found by Newton iteration; weights 2 / ((1-x^2) P_N'(x)^2).
static void legendre_eval(int n, double x, double &p, double &dp)
{
	double p0 = 1.0, p1 = x;
	if (n == 0) { p = 1.0; dp = 0.0; return; }
	for (int k = 2; k <= n; ++k) {
		double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
		p0 = p1; p1 = pk;
	}
	p = p1;
	dp = n * (x * p1 - p0) / (x * x - 1.0);
}

static std::vector<std::pair<double, double> > gauss_legendre_n(int N)
{
	std::vector<std::pair<double, double> > r(N);
	for (int i = 0; i < N; ++i) {
		double x = std::cos(PI * (i + 0.75) / (N + 0.5));
		for (int it = 0; it < 100; ++it) {
			double p, dp;
			legendre_eval(N, x, p, dp);
			double dx = p / dp;
			x -= dx;
			if (std::abs(dx) < 1e-17) break;
		}
		double p, dp;
		legendre_eval(N, x, p, dp);
		r[N - 1 - i] = {x, 2.0 / ((1.0 - x * x) * dp * dp)};
	}
	return r;
}

static const std::vector<std::pair<double, double> > &gauss_line(int degree)
{
	// static stores of the line rules NiHu uses (orders 7, 10, 40) plus any
	// order a test asks for
	static std::vector<std::pair<double, double> > pool[128];
	static bool filled[128] = {false};
	if (degree < 0 || degree >= 128)
		throw std::out_of_range("gauss_line: unsupported degree");
	#pragma omp critical(orc_gauss_line)
	if (!filled[degree]) { pool[degree] = gauss_legendre_n(degree / 2 + 1); filled[degree] = true; }
	return pool[degree];
}

static const std::vector<QP> &dunavant_store(int degree)
{
	// function-local static pool, as nihu/util/store_pattern.hpp:34-43
	static std::vector<QP> pool[15];
	static bool init = [] { for (int d = 0; d < 15; ++d) pool[d] = dunavant(d); return true; }();
	(void)init;
	if (degree < 0 || degree > 14)
		throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	return pool[degree];
}

// ---------------------------------------------------------------------------
// Geometry of a flat 3-node triangle (tria_1_elem)
// nihu/core/element.hpp:272-292 (centre, linear size estimate), :503-516
// (normal = dx/dxi x dx/deta); nihu/library/lib_shape.hpp:134-199 (shape set);
// nihu/library/lib_domain.cpp:59-71 (corners, edges, centre 1/3).
// ---------------------------------------------------------------------------
struct Elem {
	V3 X[3];
	V3 c;        // centre = x(1/3,1/3)
	V3 n;        // jacobian vector (X1-X0)x(X2-X0)
	double J;    // |n|
	V3 nu;       // unit normal
	double R;    // mean edge length
	double diam; // longest edge
	int id;
};

static inline V3 elem_x(Elem const &e, double xi, double eta)
{
	double N0 = 1.0 - xi - eta, N1 = xi, N2 = eta;
	return {e.X[0].x * N0 + e.X[1].x * N1 + e.X[2].x * N2,
		e.X[0].y * N0 + e.X[1].y * N1 + e.X[2].y * N2,
		e.X[0].z * N0 + e.X[1].z * N1 + e.X[2].z * N2};
}

static Elem make_elem(V3 X0, V3 X1, V3 X2, int id)
{
	Elem e;
	e.X[0] = X0; e.X[1] = X1; e.X[2] = X2; e.id = id;
	e.c = elem_x(e, 1. / 3., 1. / 3.);
	e.n = cross(X1 - X0, X2 - X0);
	e.J = norm(e.n);
	e.nu = e.n / e.J;
	double l0 = norm(X1 - X0), l1 = norm(X2 - X1), l2 = norm(X0 - X2);
	double d = 0.0;
	d += l0; d += l1; d += l2;
	e.R = d / 3.0;
	e.diam = std::max(l0, std::max(l1, l2));
	return e;
}

static std::vector<Elem> make_mesh(const double *xyz, int n_nodes, const int *tri, int n_tri)
{
	std::vector<Elem> m(n_tri);
	for (int e = 0; e < n_tri; ++e) {
		V3 X[3];
		for (int a = 0; a < 3; ++a) {
			int v = tri[3 * e + a];
			if (v < 0 || v >= n_nodes) throw std::out_of_range("node index out of range");
			X[a] = {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
		}
		m[e] = make_elem(X[0], X[1], X[2], e);
	}
	return m;
}

// ---------------------------------------------------------------------------
// Kernels: nihu/library/helmholtz_kernel.hpp:169-228,
// nihu/library/laplace_kernel.hpp:133-171,
// nihu/library/normal_derivative_kernel.hpp:114-257.  r = y - x.
// family 0 = Laplace, 1 = Helmholtz.  Laplace results have zero imaginary part.
// The wavenumber type K is double or complex (NiHu's WaveNumber template
// parameter, helmholtz_kernel.hpp:103); the formulas are the same.
// ---------------------------------------------------------------------------
template <class K>
static cplx kernel_eval(int family, int nx, int ny, K k, V3 x, V3 y, V3 nxv, V3 nyv)
{
	V3 rvec = y - x;
	double r = norm(rvec);
	const cplx I(0.0, 1.0);
	if (family == 0) {
		if (nx == 0 && ny == 0)
			return 1. / r / (4. * PI);
		if (nx == 0 && ny == 1) {
			double rdny = dot(rvec, nyv) / r;
			return (-1. / (r * r) / (4. * PI)) * rdny;
		}
		if (nx == 1 && ny == 0) {
			double rdnx = -dot(rvec, nxv) / r;
			return (-1. / (r * r) / (4. * PI)) * rdnx;
		}
		double rdny = dot(rvec, nyv) / r;
		double rdnx = -dot(rvec, nxv) / r;
		double g = 1. / (r * r * r) / (4. * PI);
		double f1 = -1. * g, f0 = 3. * g;
		return f0 * rdnx * rdny - f1 * dot(nxv, nyv);
	}
	if (nx == 0 && ny == 0) {
		cplx ikr = I * (k * r);
		return std::exp(-ikr) / r / (4. * PI);
	}
	if (nx + ny == 1) {
		cplx ikr = I * (k * r);
		cplx f = std::exp(-ikr) / (r * r) / (4. * PI) * (-ikr - 1.);
		double rd = (ny == 1) ? dot(rvec, nyv) / r : -dot(rvec, nxv) / r;
		return f * rd;
	}
	double rdny = dot(rvec, nyv) / r;
	double rdnx = -dot(rvec, nxv) / r;
	cplx ikr = I * (k * r);
	cplx g = std::exp(-ikr) / (r * r * r) / (4. * PI);
	cplx f1 = -g * (1. + ikr);
	cplx f0 = g * (3. + ikr * (3. + ikr));
	return f0 * rdnx * rdny - f1 * dot(nxv, nyv);
}

// ---------------------------------------------------------------------------
// Quadrature degree policy
// nihu/core/complexity_estimator.hpp:69-75, nihu/library/interval_estimator.hpp:51-62,
// nihu/tmp/interval.hpp:173-185 (first breakpoint with r < x),
// nihu/library/distance_kernel_intervals.hpp:116-155 (accuracy 3 tables),
// far-field behaviour inverse<1+Nx+Ny> (nihu/library/helmholtz_kernel.hpp,
// laplace_kernel.hpp kernel_traits_ns).
// ---------------------------------------------------------------------------
struct BP { double x; int deg; };
static const std::vector<BP> &interval_table(int order)
{
	static const std::vector<BP> inv1 = {{11. / 10., 6}, {18. / 10., 4}, {91. / 10., 2}, {1e9 / 1., 0}};
	static const std::vector<BP> inv2 = {{11. / 10., 8}, {15. / 10., 6}, {28. / 10., 4}, {158. / 10., 2}, {1e9 / 1., 0}};
	static const std::vector<BP> inv3 = {{10. / 10., 9}, {11. / 10., 9}, {14. / 10., 8}, {19. / 10., 6},
		{36. / 10., 4}, {223. / 10., 2}, {1e9 / 1., 0}};
	if (order == 1) return inv1;
	if (order == 2) return inv2;
	return inv3;
}

static int eval_interval(std::vector<BP> const &t, double r)
{
	for (auto const &bp : t)
		if (r < bp.x) return bp.deg;
	// nihu/library/lib_tmp.cpp:8-11
	throw std::out_of_range("cannot determine interval value");
}

// pair-distance convention shared with the GPU engine
static inline double dist2(V3 a, V3 b)
{
	double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
	return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

struct Policy {
	int rule;          // 0: NiHu adaptive degree (accuracy 3), 1: fixed Dunavant degree
	int fixed_degree;  // Dunavant degree for rule 1
	int near_mode;     // 0: d / R_mean_edge < limit (NiHu), 1: d / diameter < limit, 2: no near pass
	double near_limit; // 1.5 in NiHu
};

// nearly_singular_integral<...>::needed,
// nihu/library/helmholtz_nearly_singular_integrals.hpp:234-246 and the Laplace
// counterparts in laplace_nearly_singular_integrals.hpp.
static bool near_needed(Policy const &p, double d2, Elem const &trial)
{
	if (p.near_mode == 2) return false;
	double d = std::sqrt(d2);
	double S = p.near_mode == 0 ? trial.R : trial.diam;
	return d / S < p.near_limit;
}

static int regular_degree(Policy const &p, int field_complexity, int nx, int ny, double d2, Elem const &trial)
{
	if (p.rule == 1) return p.fixed_degree;
	double distance = std::sqrt(d2);
	return field_complexity + eval_interval(interval_table(1 + nx + ny), distance / trial.R);
}

// ---------------------------------------------------------------------------
// Regular collocational integral: sum_q K(x, y_q) * (J * w_q)
// nihu/core/double_integral.hpp:445-467 (eval_on_accelerator), :543-589
// ---------------------------------------------------------------------------
template <class K>
static cplx regular_integral(int family, int nx, int ny, K k, V3 x, V3 nxv, Elem const &trial, int degree)
{
	if (degree > 14)
		throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	cplx res = 0.0;
	for (auto const &q : dunavant_store(degree)) {
		V3 y = elem_x(trial, q.xi, q.eta);
		res += kernel_eval(family, nx, ny, k, x, y, nxv, trial.nu) * (1.0 * (trial.J * q.w));
	}
	return res;
}

// ---------------------------------------------------------------------------
// Diagonal (singular) integrals over a constant flat triangle, x0 = centre
// plane_element_helper: nihu/library/plane_element_helper.hpp:42-74
// Helmholtz SLP / HSP: nihu/library/helmholtz_singular_integrals.hpp:703-839
// Laplace SLP / HSP: nihu/library/laplace_singular_integrals.hpp:456-505
// DLP / DLPt: zero, nihu/library/normal_derivative_singular_integrals.hpp:13-58
// ---------------------------------------------------------------------------
static void plane_element_helper(Elem const &e, V3 x0, double r[3], double theta[3], double alpha[3])
{
	V3 R[3], C[3];
	for (int i = 0; i < 3; ++i) {
		R[i] = e.X[i] - x0;
		r[i] = norm(R[i]);
		R[i] = R[i] / r[i];
		C[i] = e.X[i] - e.X[(i + 1) % 3];
		C[i] = C[i] / norm(C[i]);
	}
	for (int i = 0; i < 3; ++i) {
		double cs = dot(R[i], R[(i + 1) % 3]);
		theta[i] = cs >= 1. ? 0. : std::acos(cs);
		alpha[i] = std::acos(dot(R[i], C[i]));
	}
}

template <class K>
static cplx singular_integral(int family, int nx, int ny, K k, Elem const &e)
{
	if (nx + ny == 1) return 0.0;
	V3 x0 = e.c;
	double r[3], theta[3], alpha[3];
	plane_element_helper(e, x0, r, theta, alpha);
	const cplx I(0.0, 1.0);
	if (family == 0) {
		double res = 0.0;
		if (nx == 0) {
			for (int i = 0; i < 3; ++i)
				res += r[i] * std::sin(alpha[i]) *
					std::log(std::tan((alpha[i] + theta[i]) / 2.) / std::tan(alpha[i] / 2.));
		} else {
			for (int i = 0; i < 3; ++i)
				res += (std::cos(alpha[i] + theta[i]) - std::cos(alpha[i])) / (r[i] * std::sin(alpha[i]));
		}
		return res / (4. * PI);
	}
	auto const &q7 = dunavant_store(7);
	if (nx == 0) {
		double I_stat = 0.0;
		for (int i = 0; i < 3; ++i)
			I_stat += r[i] * std::sin(alpha[i]) *
				std::log(std::tan((alpha[i] + theta[i]) / 2.0) / std::tan(alpha[i] / 2.0));
		cplx I_dyn = 0.0;
		for (auto const &q : q7) {
			double rq = norm(elem_x(e, q.xi, q.eta) - x0);
			double jac = e.J;
			cplx dyn = -I * k * std::exp(-I * k * rq / 2.0) * sinc(k * rq / 2.0);
			I_dyn += dyn * q.w * jac;
		}
		return (I_stat + I_dyn) / (4.0 * PI);
	}
	double IG0 = 0.0, IddG0 = 0.0;
	for (int i = 0; i < 3; ++i) {
		if (theta[i] < 1e-3) continue;
		IG0 += r[i] * std::sin(alpha[i]) * std::log(std::tan((alpha[i] + theta[i]) / 2.0) / std::tan(alpha[i] / 2.0));
		IddG0 += (std::cos(alpha[i] + theta[i]) - std::cos(alpha[i])) / (r[i] * std::sin(alpha[i]));
	}
	cplx I_acc = 0.0;
	for (auto const &q : q7) {
		double rq = norm(elem_x(e, q.xi, q.eta) - x0);
		double jac = e.J;
		cplx ikr = I * k * rq;
		cplx dyn;
		if (std::abs(rq) > 1e-3)
			dyn = (std::exp(-ikr) * (1.0 + ikr) - 1.0 + ikr * ikr / 2.0) / rq / rq / rq;
		else
			dyn = -I * k * k * k * (1.0 / 3.0 - ikr * (1.0 / 8.0 - ikr * (1.0 / 30.0 - ikr * (1.0 / 144.0 -
				ikr * (1.0 / 840.0 - ikr / 5760.0)))));
		I_acc += dyn * q.w * jac;
	}
	return (IddG0 + k * k / 2.0 * IG0 + I_acc) / (4.0 * PI);
}

// ---------------------------------------------------------------------------
// Near-singular integrals over a constant flat triangle
// Laplace static part (planar polar, radial integral in closed form):
//   nihu/library/laplace_nearly_singular_integrals.hpp:62-389
// plane_elem_transform, plane_elem_helper_mid:
//   nihu/library/plane_element_helper.hpp:76-128
// Helmholtz = Laplace static + Dunavant-14 (42 points) of (K_helm - K_lap):
//   nihu/library/helmholtz_nearly_singular_integrals.hpp:48-219
// ---------------------------------------------------------------------------
// The planar polygon with corners X[0..NC-1] (triangles: NC = 3; the
// linearised quadrilaterals of the near-singular quadrilateral integrals:
// NC = 4); the plane is spanned by the corners 0, 1, 2.
static double planar_near_static(int nx, int ny, const V3 *X, int NC, V3 x0_in, V3 nx_in, int line_degree)
{
	// T = [t1, t3 x t1, t3] with t1 = v1/|v1|, t3 = (v1 x v2)/|v1 x v2|
	V3 v1 = X[1] - X[0], v2 = X[2] - X[0];
	V3 T0 = normalized(v1);
	V3 T2 = normalized(cross(v1, v2));
	V3 T1 = normalized(cross(T2, T0));
	auto to_plane = [&](V3 p) -> V3 { return {dot(T0, p), dot(T1, p), dot(T2, p)}; };
	V3 C[4];
	for (int i = 0; i < NC; ++i) C[i] = to_plane(X[i]);
	V3 x0 = to_plane(x0_in);
	V3 nxl = to_plane(nx_in);

	double theta_lim[4], ref_distance[4], theta0[4];
	for (int n = 0; n < NC; ++n) {
		double c1x = C[n].x, c1y = C[n].y;
		double c2x = C[(n + 1) % NC].x, c2y = C[(n + 1) % NC].y;
		double lx = c2x - c1x, ly = c2y - c1y;
		double ln = std::sqrt(lx * lx + ly * ly);
		lx /= ln; ly /= ln;
		double d1x = c1x - x0.x, d1y = c1y - x0.y;
		double dl = d1x * lx + d1y * ly;
		double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		theta_lim[n] = std::atan2(d1y, d1x);
		theta0[n] = std::atan2(d0y, d0x);
		ref_distance[n] = std::sqrt(d0x * d0x + d0y * d0y);
	}
	double z = x0.z - C[0].z;

	bool hsp = (nx == 1 && ny == 1);
	// line quadrature order 10 (6 points) for SLP/DLP/DLPt, 40 (21 points) for HSP
	auto const &gl = gauss_line(line_degree > 0 ? line_degree : (hsp ? 40 : 10));
	double result = 0.0;
	for (int i = 0; i < NC; ++i) {
		double th1 = theta_lim[i], th2 = theta_lim[(i + 1) % NC];
		if (std::abs(th2 - th1) > PI) {
			if (th2 > th1) th1 += TWO_PI;
			else th2 += TWO_PI;
		}
		double jac = (th2 - th1) / 2.;
		for (auto const &g : gl) {
			double xi = g.first, w = g.second;
			double theta = th1 * (1. - xi) / 2. + th2 * (1. + xi) / 2.;
			double R = ref_distance[i] / std::cos(theta - theta0[i]);
			double r = std::sqrt(R * R + z * z);
			if (nx == 0 && ny == 0) {
				result += (r - std::abs(z)) * jac * w;
			} else if (nx == 0 && ny == 1) {
				result += (1 - std::abs(z) / r) * jac * w;
			} else if (nx == 1 && ny == 0) {
				V3 rvec = {R * std::cos(theta), R * std::sin(theta), -z};
				result += (-dot(nxl, rvec) / r + (nxl.x * std::cos(theta) + nxl.y * std::sin(theta)) * std::log(R + r)
					- nxl.z * sgn(z)) * jac * w;
			} else {
				V3 rvec = {R * std::cos(theta), R * std::sin(theta), -z};
				double rdnx = -dot(rvec, nxl) / r;
				double integrand;
				if (std::abs(z) > 1e-12)
					integrand = -R * R / (r * r) / z * rdnx;
				else
					integrand = -nxl.z / R;
				result += integrand * jac * w;
			}
		}
	}
	if (nx == 0 && ny == 1) return sgn(z) * result / (4. * PI);
	return result / (4. * PI);
}

static double laplace_near_static(int nx, int ny, Elem const &e, V3 x0_in, V3 nx_in, int line_degree = 0)
{
	return planar_near_static(nx, ny, e.X, 3, x0_in, nx_in, line_degree);
}

template <class K>
static cplx near_integral(int family, int nx, int ny, K k, Elem const &e, V3 x0, V3 nxv, int line_degree = 0)
{
	double res_static = laplace_near_static(nx, ny, e, x0, nxv, line_degree);
	if (family == 0) return res_static;
	cplx res_dynamic = 0.0;
	for (auto const &q : dunavant_store(14)) {
		V3 y = elem_x(e, q.xi, q.eta);
		double jac = e.J;
		res_dynamic += (kernel_eval(1, nx, ny, k, x0, y, nxv, e.nu) - kernel_eval(0, nx, ny, k, x0, y, nxv, e.nu))
			* q.w * jac;
	}
	return res_static + res_dynamic;
}

// ---------------------------------------------------------------------------
// Pair integral with classification
// nihu/core/double_integral.hpp:599-673 (WITH / WITHOUT singularity check),
// nihu/core/element_match.hpp:84-112 (collocation: face match = same id)
// ---------------------------------------------------------------------------
template <class K>
static cplx pair_integral(int family, int nx, int ny, K k, Policy const &p, Elem const &test, Elem const &trial,
	bool same_mesh)
{
	if (same_mesh && test.id == trial.id)
		return singular_integral(family, nx, ny, k, trial);
	double d2 = dist2(test.c, trial.c);
	if (near_needed(p, d2, trial))
		return near_integral(family, nx, ny, k, trial, test.c, test.nu);
	int degree = regular_degree(p, 0, nx, ny, d2, trial);
	return regular_integral(family, nx, ny, k, test.c, test.nu, trial, degree);
}

// ---------------------------------------------------------------------------
// Galerkin formalism, isoparametric linear triangles (nodal dofs)
// nihu/core/double_integral.hpp:181-206 (diadic regular rule, N-set weighting),
// :283-383 (singularity check); nihu/core/element_match.hpp:84-112 and
// nihu/core/element.hpp:375-413 (node overlap -> face / edge / vertex match);
// nihu/core/singular_accelerator.hpp:137-359 (canonical rules rotated to the
// shared corners); nihu/core/singular_galerkin_quadrature.hpp:197-455 (the
// Duffy-type 4D rules on tria x tria); order 7 (4-point Gauss-Legendre per
// axis) from singular_quadrature_order (helmholtz_kernel.hpp:69-71).
// ---------------------------------------------------------------------------
struct Overlap { int num, ind1, ind2; };

// element::get_overlapping, node by node in the same order
static Overlap get_overlapping(const int a[3], const int b[3])
{
	int num = 0, s1 = 0, s2 = 0;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			if (a[i] == b[j]) {
				if (num == 0) { s1 = i; s2 = j; }
				else {
					if (i < s1 || (i == 2 && s1 == 0 && num == 1)) s1 = i;
					if (j < s2 || (j == 2 && s2 == 0 && num == 1)) s2 = j;
				}
				++num;
			}
	return {num, s1, s2};
}

struct SPt { double txi, teta, sxi, seta, w; };

// canonical singular rule of a match dimension (2 face, 1 edge, 0 vertex)
static std::vector<SPt> make_singular_rule(int match)
{
	auto const &g = gauss_line(7);       // 4 points on [-1, 1]
	std::vector<std::pair<double, double> > b;
	for (auto const &p : g) b.push_back({(1.0 + p.first) / 2.0, p.second * 0.5});   // to [0, 1]
	std::vector<SPt> out;
	const int ndom = match == 2 ? 3 : match == 1 ? 6 : 1;
	const bool symmetric = match != 1;
	for (auto const &p1 : b)
	for (auto const &p2 : b)
	for (auto const &p3 : b)
	for (auto const &p4 : b)
	for (int idx = 0; idx < ndom; ++idx) {
		double x0 = p1.first, x1 = p2.first, x2 = p3.first, x3 = p4.first;
		double w = p1.second * p2.second * p3.second * p4.second;
		double mu1 = 0, mu2 = 0, xi1 = 0, xi2 = 0, J = 0;
		if (match == 2) {
			switch (idx) {
			case 0: mu1 = x0; mu2 = x0 * x1; xi1 = (1 - mu1) * x2; xi2 = xi1 * x3; J = (1 - mu1) * xi1; break;
			case 1: mu1 = x0 * x1; mu2 = x0 * (x1 - 1.0); xi1 = (1 - mu1 + mu2) * x2 - mu2; xi2 = (xi1 + mu2) * x3 - mu2;
				J = (1.0 - mu1 + mu2) * (xi1 + mu2); break;
			default: mu1 = x0 * x1; mu2 = x0; xi1 = (1 - mu2) * x2 + mu2 - mu1; xi2 = (xi1 - mu2 + mu1) * x3;
				J = (1 - mu2) * (xi1 - mu2 + mu1); break;
			}
			w *= J * x0;
		} else if (match == 1) {
			switch (idx) {
			case 0: mu1 = -x0 * x1; mu2 = -x0 * x1 * x2; xi1 = (1.0 - x0) * x3 + x0; xi2 = x0 * (1.0 - x1 + x1 * x2); break;
			case 1: mu1 = x0 * x1; mu2 = x0 * x1 * x2; xi1 = (1.0 - x0) * x3 + x0 * (1.0 - x1); xi2 = x0 * (1.0 - x1); break;
			case 2: mu1 = -x0 * x1 * x2; mu2 = x0 * x1 * (1.0 - x2); xi1 = (1.0 - x0) * x3 + x0; xi2 = x0 * (1.0 - x1); break;
			case 3: mu1 = x0 * x1 * x2; mu2 = x0 * x1 * (x2 - 1.0); xi1 = (1.0 - x0) * x3 + x0 * (1.0 - x1 * x2);
				xi2 = x0 * (1.0 - x1 * x2); break;
			case 4: mu1 = -x0 * x1 * x2; mu2 = -x0 * x1; xi1 = (1.0 - x0) * x3 + x0; xi2 = x0; break;
			default: mu1 = x0 * x1 * x2; mu2 = x0 * x1; xi1 = (1.0 - x0) * x3 + x0 * (1.0 - x1 * x2); xi2 = x0 * (1.0 - x1); break;
			}
			w *= x1 * x0 * x0 * (1.0 - x0);
		}
		double e1 = mu1 + xi1, e2 = mu2 + xi2;      // eta
		if (match == 0) {
			xi1 = x0; xi2 = xi1 * x1; e1 = xi1 * x2; e2 = e1 * x3;
			w *= x2 * x0 * x0 * x0;
		}
		SPt p = {xi1 - xi2, xi2, e1 - e2, e2, w};
		out.push_back(p);
		if (symmetric) out.push_back({p.sxi, p.seta, p.txi, p.teta, w});
	}
	return out;
}

static const std::vector<SPt> &singular_rule(int match)
{
	static const std::vector<SPt> r0 = make_singular_rule(0), r1 = make_singular_rule(1), r2 = make_singular_rule(2);
	return match == 0 ? r0 : match == 1 ? r1 : r2;
}

// point of the reference triangle mapped by the corner permutation c[]
// (quadrature::transform with isoparam_shape_set and these corners)
static inline void map_corners(double xi, double eta, const int c[3], double &oxi, double &oeta)
{
	static const double cx[3] = {0.0, 1.0, 0.0}, cy[3] = {0.0, 0.0, 1.0};
	double N0 = 1.0 - xi - eta, N1 = xi, N2 = eta;
	oxi = N0 * cx[c[0]] + N1 * cx[c[1]] + N2 * cx[c[2]];
	oeta = N0 * cy[c[0]] + N1 * cy[c[1]] + N2 * cy[c[2]];
}

// the fields of a Galerkin pair: linear = 0 constant x constant, 1
// isoparametric x isoparametric, 2 constant test x isoparametric trial, 3
// isoparametric test x constant trial
static inline bool linear_test(int linear) { return linear == 1 || linear == 3; }
static inline bool linear_trial(int linear) { return linear == 1 || linear == 2; }

// 3x3 Galerkin block of one element pair (not yet scaled); throws like NiHu.
// An isoparametric field has 3 nodal shape functions (field complexity 1), a
// constant field one shape function 1 (complexity 0); the pair's degree takes
// the larger complexity (nihu/core/complexity_estimator.hpp:39-75).  The
// block is [test shape][trial shape], [0][*] / [*][0] for a constant side.
template <class WN>
static void galerkin_block(int family, int nx, int ny, WN k, Policy const &p, Elem const &te, Elem const &se,
	const int *tn, const int *sn, bool same_mesh, int linear, cplx blk[3][3])
{
	for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) blk[a][b] = 0.0;
	const bool lt = linear_test(linear), ls = linear_trial(linear);
	int match = -1;
	Overlap ov = {0, 0, 0};
	if (same_mesh) {
		if (te.id == se.id) match = 2;
		else {
			ov = get_overlapping(tn, sn);
			if (ov.num > 1) match = 1;
			else if (ov.num == 1) match = 0;
		}
	}
	if (match >= 0) {
		// DLP / DLPt over a flat triangle vanish on the face match
		// (nihu/library/normal_derivative_singular_integrals.hpp:13-58)
		if (nx + ny == 1 && match == 2) return;
		// 1/r^3 needs 4 free dimensions (nihu/core/singular_accelerator.hpp:483-503)
		if (nx == 1 && ny == 1 && match >= 1) throw std::runtime_error("Invalid general singular quadrature");
		int tc[3] = {0, 1, 2}, scn[3] = {0, 1, 2};
		if (match < 2) {
			for (int q = 0; q < 3; ++q) tc[q] = (ov.ind1 + q) % 3;
			for (int q = 0; q < 3; ++q) scn[q] = match == 1 ? (ov.ind2 + 1 + 3 - q) % 3 : (ov.ind2 + q) % 3;
		}
		for (auto const &P : singular_rule(match)) {
			double txi, teta, sxi, seta;
			map_corners(P.txi, P.teta, tc, txi, teta);
			map_corners(P.sxi, P.seta, scn, sxi, seta);
			V3 x = elem_x(te, txi, teta), y = elem_x(se, sxi, seta);
			cplx K = kernel_eval(family, nx, ny, k, x, y, te.nu, se.nu);
			double Nt[3] = {1.0 - txi - teta, txi, teta}, Ns[3] = {1.0 - sxi - seta, sxi, seta};
			if (!lt) Nt[0] = 1.0;
			if (!ls) Ns[0] = 1.0;
			for (int a = 0; a < (lt ? 3 : 1); ++a)
				for (int b = 0; b < (ls ? 3 : 1); ++b)
					blk[a][b] += (Nt[a] * (te.J * P.w)) * K * (Ns[b] * (se.J * 1.0));
		}
		return;
	}
	double d2 = dist2(te.c, se.c);
	int degree = regular_degree(p, (lt || ls) ? 1 : 0, nx, ny, d2, se);
	if (degree > 14) throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	auto const &Q = dunavant_store(degree);
	for (auto const &qp : Q) {
		V3 x = elem_x(te, qp.xi, qp.eta);
		double Nt[3] = {1.0 - qp.xi - qp.eta, qp.xi, qp.eta};
		if (!lt) Nt[0] = 1.0;
		for (auto const &qs : Q) {
			V3 y = elem_x(se, qs.xi, qs.eta);
			double Ns[3] = {1.0 - qs.xi - qs.eta, qs.xi, qs.eta};
			if (!ls) Ns[0] = 1.0;
			cplx K = kernel_eval(family, nx, ny, k, x, y, te.nu, se.nu);
			for (int a = 0; a < (lt ? 3 : 1); ++a)
				for (int b = 0; b < (ls ? 3 : 1); ++b)
					blk[a][b] += (Nt[a] * (te.J * qp.w)) * K * (Ns[b] * (se.J * qs.w));
		}
	}
}

// ---------------------------------------------------------------------------
// Quadrilateral elements with constant fields, collocation at the centre
// quad_1_elem: 4-node bilinear, possibly warped.
// nihu/library/lib_shape.hpp:215-310 (shape functions, first and second
// derivatives; jacobian order 1), nihu/library/lib_domain.cpp:74-94 (corners
// (-1,-1), (1,-1), (1,1), (-1,1), edges 01 12 23 30, centre (0,0)),
// nihu/core/element.hpp:272-292 (centre, mean edge length), :613-616 (normal =
// dx/dxi x dx/deta).
// ---------------------------------------------------------------------------
struct QElem {
	V3 X[4];
	V3 c;        // x(0, 0)
	V3 nu;       // unit normal at the centre
	double R;    // mean edge length
	double diam; // largest corner-to-corner distance
	int id;
};

static const double QCX[4] = {-1.0, 1.0, 1.0, -1.0}, QCY[4] = {-1.0, -1.0, 1.0, 1.0};

static inline V3 q_x(QElem const &e, double xi, double eta)
{
	double N[4] = {(1.0 - xi) * (1.0 - eta) / 4.0, (1.0 + xi) * (1.0 - eta) / 4.0,
		(1.0 + xi) * (1.0 + eta) / 4.0, (1.0 - xi) * (1.0 + eta) / 4.0};
	V3 x = {0.0, 0.0, 0.0};
	for (int a = 0; a < 4; ++a) x = x + e.X[a] * N[a];
	return x;
}

static inline void q_dx(QElem const &e, double xi, double eta, V3 &dxi, V3 &deta)
{
	double Dxi[4] = {-(1.0 - eta) / 4.0, (1.0 - eta) / 4.0, (1.0 + eta) / 4.0, -(1.0 + eta) / 4.0};
	double Deta[4] = {-(1.0 - xi) / 4.0, -(1.0 + xi) / 4.0, (1.0 + xi) / 4.0, (1.0 - xi) / 4.0};
	dxi = {0.0, 0.0, 0.0}; deta = {0.0, 0.0, 0.0};
	for (int a = 0; a < 4; ++a) { dxi = dxi + e.X[a] * Dxi[a]; deta = deta + e.X[a] * Deta[a]; }
}

// d2x / dxi deta (the other second derivatives of the bilinear map vanish)
static inline V3 q_dxe(QElem const &e)
{
	return e.X[0] * 0.25 + e.X[1] * (-0.25) + e.X[2] * 0.25 + e.X[3] * (-0.25);
}

static inline V3 q_normal(QElem const &e, double xi, double eta)
{
	V3 a, b;
	q_dx(e, xi, eta, a, b);
	return cross(a, b);
}

static QElem make_qelem(const V3 X[4], int id)
{
	QElem e;
	for (int a = 0; a < 4; ++a) e.X[a] = X[a];
	e.id = id;
	e.c = q_x(e, 0.0, 0.0);
	e.nu = normalized(q_normal(e, 0.0, 0.0));
	double d = 0.0;
	for (int a = 0; a < 4; ++a) d += norm(q_x(e, QCX[(a + 1) % 4], QCY[(a + 1) % 4]) - q_x(e, QCX[a], QCY[a]));
	e.R = d / 4.0;
	e.diam = 0.0;
	for (int a = 0; a < 4; ++a)
		for (int b = a + 1; b < 4; ++b) e.diam = std::max(e.diam, norm(X[b] - X[a]));
	return e;
}

// A test point of either element type: the collocation point (centre) and
// the unit normal there.
struct TestPt { V3 x, nu; int id; };

// the elements of a mesh given as npe (3 or 4) node indices per element
static void read_elems(const double *xyz, int n_nodes, const int *el, int n_el, int npe,
	std::vector<Elem> *tris, std::vector<QElem> *quads, std::vector<TestPt> *pts)
{
	if (npe != 3 && npe != 4) throw std::invalid_argument("elements must have 3 or 4 nodes");
	for (int e = 0; e < n_el; ++e) {
		V3 X[4];
		for (int a = 0; a < npe; ++a) {
			int v = el[npe * e + a];
			if (v < 0 || v >= n_nodes) throw std::out_of_range("node index out of range");
			X[a] = {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
		}
		if (npe == 3) {
			Elem t = make_elem(X[0], X[1], X[2], e);
			if (tris) tris->push_back(t);
			if (pts) pts->push_back({t.c, t.nu, e});
		} else {
			QElem q = make_qelem(X, e);
			if (quads) quads->push_back(q);
			if (pts) pts->push_back({q.c, q.nu, e});
		}
	}
}

// Gauss-Legendre tensor rule on [-1, 1]^2 of a degree: N = degree/2 + 1 points
// per axis, xi outer, eta inner (nihu/core/gaussian_quadrature.hpp:152-207).
static std::vector<QP> gauss_quad(int degree)
{
	auto const &g = gauss_line(degree);
	std::vector<QP> q;
	for (auto const &a : g)
		for (auto const &b : g) q.push_back({a.first, b.first, a.second * b.second});
	return q;
}

static const std::vector<QP> &gauss_quad_store(int degree)
{
	static std::vector<QP> pool[128];
	static bool filled[128] = {false};
	if (degree < 0 || degree >= 128) throw std::out_of_range("quadrilateral rule degree out of range");
	#pragma omp critical(orc_gauss_quad)
	if (!filled[degree]) { pool[degree] = gauss_quad(degree); filled[degree] = true; }
	return pool[degree];
}

// kernel at a trial point of a quadrilateral times the Jacobian |n(xi)|
template <class K>
static inline cplx q_kernel_jac(int family, int nx, int ny, K k, V3 x, V3 nxv, QElem const &e, double xi,
	double eta)
{
	V3 y = q_x(e, xi, eta);
	V3 n = q_normal(e, xi, eta);
	double J = norm(n);
	return kernel_eval(family, nx, ny, k, x, y, nxv, n / J) * J;
}

// Regular pair: sum_q K(x, y_q) J(xi_q) w_q.  The degree is the field
// complexity (jacobian order 1 of quad_1: nihu/core/complexity_estimator.hpp:
// 48-62) plus the distance interval of the kernel (NiHu policy), or fixed.
template <class K>
static cplx quad_regular(int family, int nx, int ny, K k, V3 x, V3 nxv, QElem const &e, int degree)
{
	if (degree > 14)
		throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	cplx res = 0.0;
	for (auto const &q : gauss_quad_store(degree))
		res += q_kernel_jac(family, nx, ny, k, x, nxv, e, q.xi, q.eta) * q.w;
	return res;
}

// Duffy-type blind rule around a point of the reference square: the square is
// cut into the four triangles (xi0, corner s, corner s+1), each the image of
// the Gauss square rule of the degree under the bilinear map of the
// degenerate quadrilateral (xi0, xi0, corner s, corner s+1), weights times the
// signed Jacobian.  nihu/core/duffy_quadrature.hpp (on_face, duffy_impl),
// nihu/core/quadrature.hpp:105-127 (transform_inplace).
static std::vector<QP> quad_duffy_face(int degree, double x0, double y0)
{
	std::vector<QP> out;
	auto const &src = gauss_quad_store(degree);
	for (int s = 0; s < 4; ++s) {
		double cx[4] = {x0, x0, QCX[s], QCX[(s + 1) % 4]};
		double cy[4] = {y0, y0, QCY[s], QCY[(s + 1) % 4]};
		for (auto const &q : src) {
			double u = q.xi, v = q.eta;
			double N[4] = {(1.0 - u) * (1.0 - v) / 4.0, (1.0 + u) * (1.0 - v) / 4.0,
				(1.0 + u) * (1.0 + v) / 4.0, (1.0 - u) * (1.0 + v) / 4.0};
			double Du[4] = {-(1.0 - v) / 4.0, (1.0 - v) / 4.0, (1.0 + v) / 4.0, -(1.0 + v) / 4.0};
			double Dv[4] = {-(1.0 - u) / 4.0, -(1.0 + u) / 4.0, (1.0 + u) / 4.0, (1.0 - u) / 4.0};
			double px = 0, py = 0, a = 0, b = 0, c = 0, d = 0;
			for (int i = 0; i < 4; ++i) {
				px += N[i] * cx[i]; py += N[i] * cy[i];
				a += Du[i] * cx[i]; b += Du[i] * cy[i];
				c += Dv[i] * cx[i]; d += Dv[i] * cy[i];
			}
			out.push_back({px, py, q.w * (a * d - c * b)});
		}
	}
	return out;
}

// Singular (same element) SLP / DLP / DLPt: the blind Duffy rule of order 7
// (singular_quadrature_order, nihu/library/helmholtz_kernel.hpp:69-71) around
// the centre; nihu/core/singular_accelerator.hpp:363-456 and
// nihu/core/double_integral.hpp:470-510 (eval_singular_on_accelerator).
template <class K>
static cplx quad_singular_blind(int family, int nx, int ny, K k, QElem const &e)
{
	static const std::vector<QP> rule = quad_duffy_face(7, 0.0, 0.0);
	cplx res = 0.0;
	for (auto const &q : rule)
		res += q_kernel_jac(family, nx, ny, k, e.c, e.nu, e, q.xi, q.eta) * q.w;
	return res;
}

// Guiggiani's method for the hypersingular collocational integral (Guiggiani,
// Krishnasamy, Rudolphi, Rizzo, J. Appl. Mech. 59 (1992) 604-614) as NiHu
// applies it: nihu/library/guiggiani_1992.hpp:138-330 (the sides of the
// reference domain seen from the singular point in the distorted plane
// eta = T xi, where T makes the first-order distance expansion unit length;
// the Laurent terms subtracted under the radial integral and added back in
// closed form), with the Laurent coefficients of the Laplace HSP core,
// nihu/library/laplace_kernel.hpp:266-295 (also used for Helmholtz HSP, whose
// singular core is Laplace's).  Radial rule of degree rad_order, angular rule
// of degree tan_order (5 and 9 in NiHu's HSP shortcuts,
// nihu/library/helmholtz_singular_integrals.hpp:1190-1236).  A complex
// wavenumber leaves the Laurent terms as they are: it enters the expansion of
// the HSP kernel first at order 1 / r (the k^2 term), which the polar Jacobian
// rho makes bounded.
template <class K>
static cplx quad_guiggiani(int family, K k, QElem const &e, double xi0x, double xi0y, int rad_order,
	int tan_order)
{
	V3 dxi, deta;
	q_dx(e, xi0x, xi0y, dxi, deta);
	V3 dxe = q_dxe(e);
	V3 nJ = cross(dxi, deta);
	V3 n0 = normalized(nJ);
	V3 x0 = q_x(e, xi0x, xi0y);

	double la = norm(dxi), lb = norm(deta);
	double cosg = dot(dxi, deta) / la / lb;
	double u = lb / la;
	double T00 = la, T01 = la * cosg * u, T10 = 0.0, T11 = la * std::sqrt(1.0 - cosg * cosg) * u;
	double detT = T00 * T11 - T10 * T01;
	double Ti00 = T11 / detT, Ti01 = -T01 / detT, Ti10 = -T10 / detT, Ti11 = T00 / detT;
	double e0x = T00 * xi0x + T01 * xi0y, e0y = T10 * xi0x + T11 * xi0y;
	V3 J0 = nJ / detT;

	double theta_lim[4], theta0[4], ref_distance[4];
	for (int n = 0; n < 4; ++n) {
		double c1x = T00 * QCX[n] + T01 * QCY[n], c1y = T10 * QCX[n] + T11 * QCY[n];
		int m = (n + 1) % 4;
		double c2x = T00 * QCX[m] + T01 * QCY[m], c2y = T10 * QCX[m] + T11 * QCY[m];
		double lx = 1.0, ly = 0.0;
		double ln = std::sqrt((c2x - c1x) * (c2x - c1x) + (c2y - c1y) * (c2y - c1y));
		if (ln > 1e-3) { lx = (c2x - c1x) / ln; ly = (c2y - c1y) / ln; }
		double d1x = c1x - e0x, d1y = c1y - e0y;
		double dl = d1x * lx + d1y * ly;
		double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		theta_lim[n] = std::atan2(d1y, d1x);
		theta0[n] = std::atan2(d0y, d0x);
		ref_distance[n] = std::sqrt(d0x * d0x + d0y * d0y);
	}

	auto const &gt = gauss_line(tan_order);
	auto const &gr = gauss_line(rad_order);
	cplx I = 0.0;
	for (int n = 0; n < 4; ++n) {
		double t1 = theta_lim[n], t2 = theta_lim[(n + 1) % 4];
		if (std::abs(t2 - t1) < 1e-3) continue;
		if (std::abs(t2 - t1) > PI) t2 += TWO_PI;
		for (auto const &qt : gt) {
			double theta = ((1.0 - qt.first) * t1 + (1.0 + qt.first) * t2) / 2.0;
			double w_theta = qt.second * (t2 - t1) / 2.0;
			// direction in the reference square and the Taylor terms of the
			// distance (r1, r2) and of the Jacobian vector (J0, J1)
			double ct = std::cos(theta), st = std::sin(theta);
			double vx = Ti00 * ct + Ti01 * st, vy = Ti10 * ct + Ti11 * st;
			V3 r1 = dxi * vx + deta * vy;
			V3 r2 = dxe * (vx * vy);
			V3 J1 = (cross(dxi, dxe) * vx + cross(dxe, deta) * vy) / detT;
			// Laurent coefficients of the Laplace HSP core (constant field: N0 = 1, N1 = 0)
			V3 g1 = r1 * (dot(r2, J0) + dot(r1, J1));
			V3 b0 = J0 * -1.0;
			V3 b1 = g1 * 3.0 - J1;
			double a0 = dot(b0, n0);
			double a1 = dot(b1, n0);
			double Sm2 = -3.0 * dot(r1, r2);
			double Fm1 = -(Sm2 * a0 + a1) / (4.0 * PI);
			double Fm2 = -a0 / (4.0 * PI);

			double rho_lim = ref_distance[n] / std::cos(theta - theta0[n]);
			I += (Fm1 * std::log(rho_lim) - Fm2 / rho_lim) * w_theta;
			for (auto const &qr : gr) {
				double rho = (1.0 + qr.first) * rho_lim / 2.0;
				double w_rho = qr.second * rho_lim / 2.0;
				double ex = rho * ct, ey = rho * st;
				double xi = Ti00 * ex + Ti01 * ey + xi0x, eta = Ti10 * ex + Ti11 * ey + xi0y;
				V3 y = q_x(e, xi, eta);
				V3 nv = q_normal(e, xi, eta);
				double J = norm(nv);
				cplx F = kernel_eval(family, 1, 1, k, x0, y, n0, nv / J) * (J / detT * rho);
				double sing = (Fm1 + Fm2 / rho) / rho;
				I += (F - sing) * (w_theta * w_rho);
			}
		}
	}
	return I;
}

// Image of a point on the element by Newton-Raphson iteration on
// x(xi) + zeta n(xi) / sqrt|n(xi)| = x0 from (0, 0, 0); false if not
// converged.  nihu/core/inverse_mapping.hpp:37-118.
static bool quad_inverse_map(QElem const &e, V3 x0, double tol, int max_iter, double &xi, double &eta, double &zeta)
{
	double r[3] = {0.0, 0.0, 0.0};
	V3 dxe = q_dxe(e);
	for (int it = 0; it < max_iter; ++it) {
		V3 y = q_x(e, r[0], r[1]);
		V3 n = q_normal(e, r[0], r[1]);
		double sj = std::sqrt(norm(n));
		n = n / sj;
		V3 f = y + n * r[2] - x0;
		if (norm(f) < tol) { xi = r[0]; eta = r[1]; zeta = r[2]; return true; }
		V3 a, b;
		q_dx(e, r[0], r[1], a, b);
		V3 dnxi = cross(a, dxe) / sj;
		V3 dneta = cross(dxe, b) / sj;
		V3 c0 = a + dnxi * r[2], c1 = b + dneta * r[2], c2 = n;
		// solve [c0 c1 c2] d = f by Gaussian elimination with partial pivoting
		double M[3][4] = {{c0.x, c1.x, c2.x, f.x}, {c0.y, c1.y, c2.y, f.y}, {c0.z, c1.z, c2.z, f.z}};
		for (int col = 0; col < 3; ++col) {
			int piv = col;
			for (int i = col + 1; i < 3; ++i)
				if (std::abs(M[i][col]) > std::abs(M[piv][col])) piv = i;
			if (piv != col) for (int j = 0; j < 4; ++j) std::swap(M[piv][j], M[col][j]);
			for (int i = col + 1; i < 3; ++i) {
				double fct = M[i][col] / M[col][col];
				for (int j = col; j < 4; ++j) M[i][j] -= fct * M[col][j];
			}
		}
		double d[3];
		for (int i = 2; i >= 0; --i) {
			double sacc = M[i][3];
			for (int j = i + 1; j < 3; ++j) sacc -= M[i][j] * d[j];
			d[i] = sacc / M[i][i];
		}
		for (int i = 0; i < 3; ++i) r[i] -= d[i];
	}
	return false;
}

// the parallelogram tangent to the element at xi0 (element::get_linearized_elem,
// nihu/core/element.hpp:418-437)
static QElem quad_linearized(QElem const &e, double xi0, double eta0)
{
	V3 x0 = q_x(e, xi0, eta0);
	V3 a, b;
	q_dx(e, xi0, eta0, a, b);
	V3 X[4];
	for (int i = 0; i < 4; ++i) X[i] = x0 + a * (QCX[i] - xi0) + b * (QCY[i] - eta0);
	return make_qelem(X, e.id);
}

// Planar constant collocation integral over a planar quadrilateral (the
// linearised element): Laplace static part in the plane plus, for Helmholtz,
// the 8 x 8-point Gauss rule (degree 14) of K_helm - K_lap;
// nihu/library/helmholtz_nearly_singular_integrals.hpp:40-219 with
// elem_t = quad_1_elem, nihu/library/laplace_nearly_singular_integrals.hpp:62-389.
template <class K>
static cplx quad_planar_near(int family, int nx, int ny, K k, QElem const &lin, V3 x0, V3 nxv)
{
	V3 C[4];
	for (int i = 0; i < 4; ++i) C[i] = q_x(lin, QCX[i], QCY[i]);
	cplx res = planar_near_static(nx, ny, C, 4, x0, nxv, 0);
	if (family == 0) return res;
	V3 nyv = normalized(q_normal(lin, 0.0, 0.0));
	cplx dyn = 0.0;
	for (auto const &q : gauss_quad_store(14)) {
		V3 y = q_x(lin, q.xi, q.eta);
		double jac = norm(q_normal(lin, q.xi, q.eta));
		dyn += (kernel_eval(1, nx, ny, k, x0, y, nxv, nyv) - kernel_eval(0, nx, ny, k, x0, y, nxv, nyv)) * q.w * jac;
	}
	return res + dyn;
}

// Near-singular integral by polar integration around the image xi0 of the
// collocation point of K J - K_lin J_lin (the linearised element at xi0), plus
// the planar integral over the linearised element;
// nihu/library/nearly_singular_collocational.hpp:71-203 (inverse mapping
// tolerance 1e-5, 100 iterations).
template <class K>
static cplx quad_near_polar(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv, int rad_order,
	int tan_order)
{
	double xi0, eta0, zeta;
	if (!quad_inverse_map(e, x0, 1e-5, 100, xi0, eta0, zeta))
		throw std::runtime_error("Could not perform inverse mapping");
	QElem lin = quad_linearized(e, xi0, eta0);
	double theta_lim[4], theta0[4], ref_distance[4];
	for (int n = 0; n < 4; ++n) {
		int m = (n + 1) % 4;
		double c1x = QCX[n], c1y = QCY[n], c2x = QCX[m], c2y = QCY[m];
		double lx = 1.0, ly = 0.0;
		double ln = std::sqrt((c2x - c1x) * (c2x - c1x) + (c2y - c1y) * (c2y - c1y));
		if (ln > 1e-3) { lx = (c2x - c1x) / ln; ly = (c2y - c1y) / ln; }
		double d1x = c1x - xi0, d1y = c1y - eta0;
		double dl = d1x * lx + d1y * ly;
		double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		theta_lim[n] = std::atan2(d1y, d1x);
		theta0[n] = std::atan2(d0y, d0x);
		ref_distance[n] = std::sqrt(d0x * d0x + d0y * d0y);
	}
	auto const &gt = gauss_line(tan_order);
	auto const &gr = gauss_line(rad_order);
	cplx I = 0.0;
	for (int n = 0; n < 4; ++n) {
		double t1 = theta_lim[n], t2 = theta_lim[(n + 1) % 4];
		if (std::abs(t2 - t1) > PI) {
			if (t2 > t1) t1 += TWO_PI;
			else t2 += TWO_PI;
		}
		if (std::abs(t2 - t1) < 1e-3) continue;
		for (auto const &qt : gt) {
			double theta = ((1.0 - qt.first) * t1 + (1.0 + qt.first) * t2) / 2.0;
			double w_theta = qt.second * (t2 - t1) / 2.0;
			double rho_lim = ref_distance[n] / std::cos(theta - theta0[n]);
			for (auto const &qr : gr) {
				double rho = (1.0 + qr.first) * rho_lim / 2.0;
				double w_rho = qr.second * rho_lim / 2.0;
				double xi = rho * std::cos(theta) + xi0, eta = rho * std::sin(theta) + eta0;
				cplx F = q_kernel_jac(family, nx, ny, k, x0, nxv, e, xi, eta);
				cplx Fl = q_kernel_jac(family, nx, ny, k, x0, nxv, lin, xi, eta);
				I += (F - Fl) * (w_theta * w_rho * rho);
			}
		}
	}
	return I + quad_planar_near(family, nx, ny, k, lin, x0, nxv);
}

// Telles' cubic transformation of a Gauss rule towards xi0 (J. C. F. Telles,
// Int. J. Numer. Meth. Eng. 24 (1987) 959-973, eqs. (21) and (27)), applied
// per axis with the distance parameter d; nihu/library/telles_1987.hpp:30-96.
static std::vector<QP> telles_rule(int degree, double xi0, double eta0, double d)
{
	std::vector<QP> q = gauss_quad_store(degree);
	double r_bar;
	if (d < 0.05) r_bar = 0.0;
	else if (d <= 1.3) r_bar = 0.85 + 0.24 * std::log(d);
	else if (d <= 3.618) r_bar = 0.893 + 0.0832 * std::log(d);
	else r_bar = 1.0;
	for (int dim = 0; dim < 2; ++dim) {
		double e = dim == 0 ? xi0 : eta0;
		double f = 1.0 + 2.0 * r_bar;
		double qq = 1.0 / (2.0 * f) * ((e * (3.0 - 2.0 * r_bar) - 2.0 * e * e * e / f) * 1.0 / f - e);
		double p = 1.0 / (3.0 * f * f) * (4.0 * r_bar * (1.0 - r_bar) + 3.0 * (1.0 - e * e));
		double s = std::sqrt(qq * qq + p * p * p);
		double gb = std::cbrt(-qq + s) + std::cbrt(-qq - s) + e / f;
		double Q = 1.0 + 3.0 * gb * gb;
		double a = (1.0 - r_bar) / Q, b = -3.0 * (1.0 - r_bar) * gb / Q, c = (3.0 * gb * gb + r_bar) / Q, dd = -b;
		for (auto &pt : q) {
			double x = dim == 0 ? pt.xi : pt.eta;
			double nx = a * x * x * x + b * x * x + c * x + dd;
			double jac = 3.0 * a * x * x + 2.0 * b * x + c;
			if (dim == 0) pt.xi = nx; else pt.eta = nx;
			pt.w *= jac;
		}
	}
	return q;
}

// Near-singular integral with the Telles-transformed Gauss rule of degree 30
// towards the image of the collocation point (inverse mapping tolerance 1e-6);
// nihu/library/nearly_singular_collocational_telles.hpp:66-103.
template <class K>
static cplx quad_near_telles(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv, int degree)
{
	double xi0, eta0, zeta;
	if (!quad_inverse_map(e, x0, 1e-6, 100, xi0, eta0, zeta))
		throw std::runtime_error("Could not perform inverse mapping");
	cplx I = 0.0;
	for (auto const &q : telles_rule(degree, xi0, eta0, zeta))
		I += q_kernel_jac(family, nx, ny, k, x0, nxv, e, q.xi, q.eta) * q.w;
	return I;
}

// Near-singular method per kernel for collocation on elements other than
// constant triangles (d / R_mean_edge < 1.5):
// Helmholtz SLP polar (15, 15), DLP / DLPt / HSP Telles 30
// (nihu/library/helmholtz_nearly_singular_integrals.hpp:266-310,360-404,
// 461-505,560-604); Laplace SLP / DLP / DLPt polar (15, 15), HSP polar
// (100, 100) (nihu/library/laplace_nearly_singular_integrals.hpp:443-486,
// 539-582,634-677,730-773).
template <class K>
static cplx quad_near(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv)
{
	if (family == 1 && nx + ny > 0) return quad_near_telles(family, nx, ny, k, e, x0, nxv, 30);
	if (family == 0 && nx == 1 && ny == 1) return quad_near_polar(family, nx, ny, k, e, x0, nxv, 100, 100);
	return quad_near_polar(family, nx, ny, k, e, x0, nxv, 15, 15);
}

// Singular (same element): HSP by Guiggiani (5, 9) at the centre, the others
// by the blind Duffy rule (nihu/library/helmholtz_singular_integrals.hpp:
// 1190-1236, laplace_singular_integrals.hpp:1279-1318; DLP / DLPt are not
// zero on quadrilaterals: normal_derivative_singular_integrals.hpp:13-58
// covers tria_1 only).
template <class K>
static cplx quad_singular(int family, int nx, int ny, K k, QElem const &e)
{
	if (nx == 1 && ny == 1) return quad_guiggiani(family, k, e, 0.0, 0.0, 5, 9);
	return quad_singular_blind(family, nx, ny, k, e);
}

static bool near_needed_q(Policy const &p, double d2, QElem const &trial)
{
	if (p.near_mode == 2) return false;
	double d = std::sqrt(d2);
	double S = p.near_mode == 0 ? trial.R : trial.diam;
	return d / S < p.near_limit;
}

// Pair integral with classification for a quadrilateral trial element; the
// regular degree is 1 (field complexity) + the interval degree, or the fixed
// degree of the policy (a Gauss square rule of that degree).
template <class K>
static cplx quad_pair_integral(int family, int nx, int ny, K k, Policy const &p, TestPt const &t,
	QElem const &trial, bool same_mesh)
{
	if (same_mesh && t.id == trial.id) return quad_singular(family, nx, ny, k, trial);
	double d2 = dist2(t.x, trial.c);
	if (near_needed_q(p, d2, trial)) return quad_near(family, nx, ny, k, trial, t.x, t.nu);
	int degree = p.rule == 1 ? p.fixed_degree
		: 1 + eval_interval(interval_table(1 + nx + ny), std::sqrt(d2) / trial.R);
	return quad_regular(family, nx, ny, k, t.x, t.nu, trial, degree);
}

// ---------------------------------------------------------------------------
// Galerkin formalism on quadrilaterals (quad_1_elem): constant_view (one dof
// per element; field complexity = polynomial order 0 + jacobian order 1) or
// isoparametric_view (the four corner nodes with the bilinear shape
// functions; complexity 1 + 1), nihu/core/complexity_estimator.hpp:39-75,
// nihu/library/lib_shape.hpp:215-230.
//   regular pairs: the Gauss rule of one degree on both elements (diadic),
//     degree = complexity + the kernel's distance interval of d / R_trial, or
//     fixed (nihu/core/double_integral.hpp:283-320, :181-206);
//   node-sharing pairs on one mesh (nihu/core/element_match.hpp:84-112,
//     element.hpp:375-413): the quad x quad rules of
//     nihu/core/singular_galerkin_quadrature.hpp:457-625 with the subdomain
//     corners of nihu/library/lib_singular_galerkin_quadrature.cpp:6-31, of
//     singular_quadrature_order 7 (helmholtz_kernel.hpp:69-71), rotated to
//     the shared corners as nihu/core/singular_accelerator.hpp:283-325 does
//     (the rotations and reflections of the square, |J| = 1);
//   the hypersingular kernel has no such rule on face and edge matches
//     (singular_accelerator.hpp:483-503), as for triangles; the double layer
//     of a coincident quadrilateral is integrated (NiHu's zero shortcut,
//     normal_derivative_singular_integrals.hpp:13-58, is for tria_1 only).
// ---------------------------------------------------------------------------

// element::get_overlapping for two quadrilaterals
static Overlap get_overlapping_quads(const int a[4], const int b[4])
{
	int num = 0, s1 = 0, s2 = 0;
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j)
			if (a[i] == b[j]) {
				if (num == 0) { s1 = i; s2 = j; }
				else {
					if (i < s1 || (i == 3 && s1 == 0 && num == 1)) s1 = i;
					if (j < s2 || (j == 3 && s2 == 0 && num == 1)) s2 = j;
				}
				++num;
			}
	return {num, s1, s2};
}

// bilinear shape functions of the reference square, corner order
// (-1,-1), (1,-1), (1,1), (-1,1)
static inline void quad_shapes(double xi, double eta, double N[4])
{
	N[0] = (1.0 - xi) * (1.0 - eta) / 4.0;
	N[1] = (1.0 + xi) * (1.0 - eta) / 4.0;
	N[2] = (1.0 + xi) * (1.0 + eta) / 4.0;
	N[3] = (1.0 - xi) * (1.0 + eta) / 4.0;
}

// Subdomains of the plane of the difference variable mu = eta - xi, as four
// corners each (a repeated first corner makes a triangle fanned from the
// origin), for the face, edge and corner matches; the table of
// lib_singular_galerkin_quadrature.cpp:6-31.
static const double QG_FACE[4][4][2] = {
	{{0, 0}, {0, 0}, {2, -2}, {2, 0}}, {{0, 0}, {0, 0}, {2, 0}, {2, 2}},
	{{0, 0}, {0, 0}, {2, 2}, {0, 2}}, {{0, 0}, {0, 0}, {0, 2}, {-2, 2}}};
static const double QG_EDGE[6][4][2] = {
	{{0, 0}, {0, 0}, {-2, 0}, {-2, -2}}, {{0, 0}, {0, 0}, {-2, -2}, {0, -2}},
	{{0, 0}, {0, 0}, {0, -2}, {2, -2}}, {{0, 0}, {0, 0}, {2, -2}, {2, 0}},
	{{-2, -2}, {-2, -4}, {0, -4}, {0, -2}}, {{0, -2}, {0, -4}, {2, -4}, {2, -2}}};
static const double QG_CORNER[5][4][2] = {
	{{0, 0}, {0, 0}, {-2, 0}, {-2, -2}}, {{0, 0}, {0, 0}, {-2, -2}, {0, -2}},
	{{-4, 0}, {-4, -2}, {-2, -2}, {-2, 0}}, {{-4, -2}, {-4, -4}, {-2, -4}, {-2, -2}},
	{{-2, -2}, {-2, -4}, {0, -4}, {0, -2}}};

// the bilinear map of the reference square onto the plane quadrilateral of
// corners C: image point and signed Jacobian (quadrature::transform)
static inline void plane_bilinear(const double C[4][2], double xi, double eta, double &u, double &v, double &J)
{
	double N[4];
	quad_shapes(xi, eta, N);
	const double dxi[4] = {-(1.0 - eta) / 4.0, (1.0 - eta) / 4.0, (1.0 + eta) / 4.0, -(1.0 + eta) / 4.0};
	const double deta[4] = {-(1.0 - xi) / 4.0, -(1.0 + xi) / 4.0, (1.0 + xi) / 4.0, (1.0 - xi) / 4.0};
	double a = 0, b = 0, c = 0, d = 0;
	u = v = 0.0;
	for (int k = 0; k < 4; ++k) {
		u += N[k] * C[k][0]; v += N[k] * C[k][1];
		a += dxi[k] * C[k][0]; b += dxi[k] * C[k][1];
		c += deta[k] * C[k][0]; d += deta[k] * C[k][1];
	}
	J = a * d - b * c;
}

// the trial point of the canonical configuration from mu + xi: the identity
// (face), a reflection across the edge eta_2 = -1 (edge) or through the
// corner (-1, -1) (corner)
static inline void qg_trial(int match, double m0, double m1, double &e0, double &e1)
{
	if (match == 2) { e0 = m0; e1 = m1; }
	else if (match == 1) { e0 = m0; e1 = -2.0 - m1; }
	else { e0 = -2.0 - m0; e1 = -2.0 - m1; }
}

// canonical quad x quad singular rule of a match dimension and Gauss degree:
// for each subdomain, an outer rule over mu, and for each outer point an inner
// rule over the rectangle of test points xi with mu + xi inside the trial
// square (clamped to the test square); the face rule is symmetrised by
// swapping test and trial.  Points (txi, teta, sxi, seta, w).
static std::vector<SPt> make_quad_singular_rule(int match, int degree)
{
	const std::vector<QP> base = gauss_quad(degree);
	const double (*dom)[4][2] = match == 2 ? QG_FACE : match == 1 ? QG_EDGE : QG_CORNER;
	const int ndom = match == 2 ? 4 : match == 1 ? 6 : 5;
	double lo[2], hi[2];
	qg_trial(match, -1.0, -1.0, lo[0], lo[1]);
	qg_trial(match, 1.0, 1.0, hi[0], hi[1]);
	for (int j = 0; j < 2; ++j)
		if (lo[j] > hi[j]) std::swap(lo[j], hi[j]);
	std::vector<SPt> out;
	for (int d = 0; d < ndom; ++d)
		for (auto const &po : base) {
			double m0, m1, Jo;
			plane_bilinear(dom[d], po.xi, po.eta, m0, m1, Jo);
			const double wo = po.w * Jo;
			double C[4][2] = {{lo[0] - m0, lo[1] - m1}, {hi[0] - m0, lo[1] - m1}, {hi[0] - m0, hi[1] - m1},
				{lo[0] - m0, hi[1] - m1}};
			for (auto &c : C)
				for (double &v : c) v = std::max(std::min(v, 1.0), -1.0);
			for (auto const &pi : base) {
				double x0, x1, Ji, e0, e1;
				plane_bilinear(C, pi.xi, pi.eta, x0, x1, Ji);
				qg_trial(match, m0 + x0, m1 + x1, e0, e1);
				const double w = wo * (pi.w * Ji);
				out.push_back({x0, x1, e0, e1, w});
				if (match == 2) out.push_back({e0, e1, x0, x1, w});
			}
		}
	return out;
}

static const std::vector<SPt> &quad_singular_rule(int match)
{
	static const std::vector<SPt> r0 = make_quad_singular_rule(0, 7), r1 = make_quad_singular_rule(1, 7),
		r2 = make_quad_singular_rule(2, 7);
	return match == 0 ? r0 : match == 1 ? r1 : r2;
}

// point of the reference square mapped by the corner permutation c[]
// (quadrature::transform with isoparam_shape_set and these corners)
static inline void map_quad_corners(double xi, double eta, const int c[4], double &oxi, double &oeta)
{
	double N[4];
	quad_shapes(xi, eta, N);
	oxi = oeta = 0.0;
	for (int k = 0; k < 4; ++k) { oxi += N[k] * QCX[c[k]]; oeta += N[k] * QCY[c[k]]; }
}

// 4x4 Galerkin block of a quadrilateral pair (test shape a, trial shape b;
// [0][0] for constant fields); throws like NiHu
template <class WN>
static void galerkin_quad_block(int family, int nx, int ny, WN k, Policy const &p, QElem const &te,
	QElem const &se, const int *tn, const int *sn, bool same_mesh, int linear, cplx blk[4][4])
{
	for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) blk[a][b] = 0.0;
	const bool lt = linear_test(linear), ls = linear_trial(linear);
	// test point (N_a J w) and trial point (N_b J) of one point pair
	auto add = [&](double txi, double teta, double sxi, double seta, double wt, double ws) {
		const V3 x = q_x(te, txi, teta), y = q_x(se, sxi, seta);
		const V3 n1 = q_normal(te, txi, teta), n2 = q_normal(se, sxi, seta);
		const double J1 = norm(n1), J2 = norm(n2);
		const cplx K = kernel_eval(family, nx, ny, k, x, y, n1 / J1, n2 / J2);
		double Nt[4], Ns[4];
		quad_shapes(txi, teta, Nt);
		quad_shapes(sxi, seta, Ns);
		if (!lt) Nt[0] = 1.0;
		if (!ls) Ns[0] = 1.0;
		for (int a = 0; a < (lt ? 4 : 1); ++a)
			for (int b = 0; b < (ls ? 4 : 1); ++b) blk[a][b] += (Nt[a] * (J1 * wt)) * K * (Ns[b] * (J2 * ws));
	};
	int match = -1;
	Overlap ov = {0, 0, 0};
	if (same_mesh) {
		if (te.id == se.id) match = 2;
		else {
			ov = get_overlapping_quads(tn, sn);
			if (ov.num > 1) match = 1;
			else if (ov.num == 1) match = 0;
		}
	}
	if (match >= 0) {
		if (nx == 1 && ny == 1 && match >= 1) throw std::runtime_error("Invalid general singular quadrature");
		int tc[4] = {0, 1, 2, 3}, scn[4] = {0, 1, 2, 3};
		if (match < 2)
			for (int q = 0; q < 4; ++q) {
				tc[q] = (ov.ind1 + q) % 4;
				scn[q] = match == 1 ? (ov.ind2 + 1 + 4 - q) % 4 : (ov.ind2 + q) % 4;
			}
		for (auto const &P : quad_singular_rule(match)) {
			double txi, teta, sxi, seta;
			map_quad_corners(P.txi, P.teta, tc, txi, teta);
			map_quad_corners(P.sxi, P.seta, scn, sxi, seta);
			add(txi, teta, sxi, seta, P.w, 1.0);
		}
		return;
	}
	const double d2 = dist2(te.c, se.c);
	const int degree = p.rule == 1 ? p.fixed_degree
		: ((lt || ls) ? 2 : 1) + eval_interval(interval_table(1 + nx + ny), std::sqrt(d2) / se.R);
	if (degree > 14) throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	auto const &Q = gauss_quad_store(degree);
	for (auto const &qp : Q)
		for (auto const &qs : Q) add(qp.xi, qp.eta, qs.xi, qs.eta, qp.w, qs.w);
}

// the quadrilaterals of a mesh
static std::vector<QElem> make_qmesh(const double *xyz, int n_nodes, const int *quad, int n)
{
	std::vector<QElem> out;
	read_elems(xyz, n_nodes, quad, n, 4, nullptr, &out, nullptr);
	return out;
}


// ---------------------------------------------------------------------------
// quad_1_gauss_field (nihu/library/quad_1_gauss_field.hpp,
// quad_1_gauss_shape_set.hpp): a scalar field over a quadrilateral with four
// dofs at the 2 x 2 Gauss points, nodes in the order of
// nihu/library/lib_shape.cpp:128-134 -- (-,-), (+,-), (-,+), (+,+) at
// +-1/sqrt(3) -- and the bilinear Lagrange shape functions on those nodes,
//   N_b(xi, eta) = (1 + s_b sqrt3 xi)(1 + t_b sqrt3 eta) / 4.
// Collocation with it (dirac of the field) puts the collocation points at the
// nodes; each element pair is a 4 x 4 block (test node a, trial dof b) or,
// for a constant test field (field points), 1 x 4.
//   classification: element centres, as for constant fields
//     (helmholtz_nearly_singular_integrals.hpp:281-293, element_match.hpp:
//     84-112: only the same element is singular in collocation);
//   regular: degree 2 (polynomial order 1 + jacobian order 1 of the field,
//     nihu/core/complexity_estimator.hpp:47-75) + the distance-table degree,
//     one Gauss square rule for the whole block, N_b(xi_q) in the weights
//     (double_integral.hpp:445-467);
//   near-singular: per test node, polar with F = K J N(xi) - K_lin J_lin N(xi0)
//     plus the planar integral times N(xi0) (nearly_singular_collocational.hpp:
//     86-197), or Telles with N(xi) in the weights;
//   singular: per test node xi_a, HSP by Guiggiani (5, 9) at xi_a with the
//     shape-function terms of the Laurent coefficients
//     (helmholtz_singular_integrals.hpp:1189-1236, guiggiani_1992.hpp:180-330,
//     laplace_kernel.hpp:266-295), the others by the blind Duffy rule of
//     degree 7 around xi_a (singular_accelerator.hpp:400-456,
//     double_integral.hpp:483-503).
// ---------------------------------------------------------------------------
typedef std::array<cplx, 4> C4;
static const double GSX[4] = {-1.0, 1.0, -1.0, 1.0}, GSY[4] = {-1.0, -1.0, 1.0, 1.0};

static inline void gauss_node(int a, double &xi, double &eta)
{
	xi = GSX[a] * (std::sqrt(3.0) / 3.0);
	eta = GSY[a] * (std::sqrt(3.0) / 3.0);
}

static inline void gauss_shape(double xi, double eta, double N[4])
{
	const double s3 = std::sqrt(3.0);
	for (int b = 0; b < 4; ++b) N[b] = (1.0 + GSX[b] * s3 * xi) * (1.0 + GSY[b] * s3 * eta) / 4.0;
}

static inline void gauss_dshape(double xi, double eta, double dNx[4], double dNy[4])
{
	const double s3 = std::sqrt(3.0);
	for (int b = 0; b < 4; ++b) {
		dNx[b] = GSX[b] * s3 * (1.0 + GSY[b] * s3 * eta) / 4.0;
		dNy[b] = (1.0 + GSX[b] * s3 * xi) * GSY[b] * s3 / 4.0;
	}
}

static inline void c4_zero(C4 &v) { for (auto &z : v) z = 0.0; }

template <class K>
static void gauss_regular(int family, int nx, int ny, K k, V3 x, V3 nxv, QElem const &e, int degree, C4 &out)
{
	if (degree > 14)
		throw std::out_of_range("Too high quadrature degree selected for collocational integration");
	c4_zero(out);
	for (auto const &q : gauss_quad_store(degree)) {
		cplx v = q_kernel_jac(family, nx, ny, k, x, nxv, e, q.xi, q.eta) * q.w;
		double N[4];
		gauss_shape(q.xi, q.eta, N);
		for (int b = 0; b < 4; ++b) out[b] += v * N[b];
	}
}

template <class K>
static void gauss_near_polar(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv, int rad_order,
	int tan_order, C4 &out)
{
	double xi0, eta0, zeta;
	if (!quad_inverse_map(e, x0, 1e-5, 100, xi0, eta0, zeta))
		throw std::runtime_error("Could not perform inverse mapping");
	QElem lin = quad_linearized(e, xi0, eta0);
	double N0[4];
	gauss_shape(xi0, eta0, N0);
	double theta_lim[4], theta0[4], ref_distance[4];
	for (int n = 0; n < 4; ++n) {
		int m = (n + 1) % 4;
		double c1x = QCX[n], c1y = QCY[n], c2x = QCX[m], c2y = QCY[m];
		double lx = 1.0, ly = 0.0;
		double ln = std::sqrt((c2x - c1x) * (c2x - c1x) + (c2y - c1y) * (c2y - c1y));
		if (ln > 1e-3) { lx = (c2x - c1x) / ln; ly = (c2y - c1y) / ln; }
		double d1x = c1x - xi0, d1y = c1y - eta0;
		double dl = d1x * lx + d1y * ly;
		double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		theta_lim[n] = std::atan2(d1y, d1x);
		theta0[n] = std::atan2(d0y, d0x);
		ref_distance[n] = std::sqrt(d0x * d0x + d0y * d0y);
	}
	auto const &gt = gauss_line(tan_order);
	auto const &gr = gauss_line(rad_order);
	c4_zero(out);
	for (int n = 0; n < 4; ++n) {
		double t1 = theta_lim[n], t2 = theta_lim[(n + 1) % 4];
		if (std::abs(t2 - t1) > PI) {
			if (t2 > t1) t1 += TWO_PI;
			else t2 += TWO_PI;
		}
		if (std::abs(t2 - t1) < 1e-3) continue;
		for (auto const &qt : gt) {
			double theta = ((1.0 - qt.first) * t1 + (1.0 + qt.first) * t2) / 2.0;
			double w_theta = qt.second * (t2 - t1) / 2.0;
			double rho_lim = ref_distance[n] / std::cos(theta - theta0[n]);
			for (auto const &qr : gr) {
				double rho = (1.0 + qr.first) * rho_lim / 2.0;
				double w_rho = qr.second * rho_lim / 2.0;
				double xi = rho * std::cos(theta) + xi0, eta = rho * std::sin(theta) + eta0;
				cplx F = q_kernel_jac(family, nx, ny, k, x0, nxv, e, xi, eta);
				cplx Fl = q_kernel_jac(family, nx, ny, k, x0, nxv, lin, xi, eta);
				double N[4];
				gauss_shape(xi, eta, N);
				for (int b = 0; b < 4; ++b) out[b] += (F * N[b] - Fl * N0[b]) * (w_theta * w_rho * rho);
			}
		}
	}
	cplx planar = quad_planar_near(family, nx, ny, k, lin, x0, nxv);
	for (int b = 0; b < 4; ++b) out[b] += planar * N0[b];
}

template <class K>
static void gauss_near_telles(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv, int degree, C4 &out)
{
	double xi0, eta0, zeta;
	if (!quad_inverse_map(e, x0, 1e-6, 100, xi0, eta0, zeta))
		throw std::runtime_error("Could not perform inverse mapping");
	c4_zero(out);
	for (auto const &q : telles_rule(degree, xi0, eta0, zeta)) {
		cplx v = q_kernel_jac(family, nx, ny, k, x0, nxv, e, q.xi, q.eta) * q.w;
		double N[4];
		gauss_shape(q.xi, q.eta, N);
		for (int b = 0; b < 4; ++b) out[b] += v * N[b];
	}
}

// NiHu's near-singular method per kernel, as quad_near
template <class K>
static void gauss_near(int family, int nx, int ny, K k, QElem const &e, V3 x0, V3 nxv, C4 &out)
{
	if (family == 1 && nx + ny > 0) return gauss_near_telles(family, nx, ny, k, e, x0, nxv, 30, out);
	if (family == 0 && nx == 1 && ny == 1) return gauss_near_polar(family, nx, ny, k, e, x0, nxv, 100, 100, out);
	gauss_near_polar(family, nx, ny, k, e, x0, nxv, 15, 15, out);
}

// the blind Duffy rule of a degree around xi0 with the shape functions
template <class K>
static void gauss_duffy(int family, int nx, int ny, K k, QElem const &e, double xi0, double eta0, int degree, C4 &out)
{
	V3 x = q_x(e, xi0, eta0);
	V3 n = normalized(q_normal(e, xi0, eta0));
	c4_zero(out);
	for (auto const &q : quad_duffy_face(degree, xi0, eta0)) {
		cplx v = q_kernel_jac(family, nx, ny, k, x, n, e, q.xi, q.eta) * q.w;
		double N[4];
		gauss_shape(q.xi, q.eta, N);
		for (int b = 0; b < 4; ++b) out[b] += v * N[b];
	}
}

// Guiggiani's finite part at xi0 with the shape functions: per dof b the
// Laurent coefficients a0 N_b(xi0) and a1 N_b(xi0) + a0 (grad N_b(xi0) . v),
// v the direction in the reference square (laplace_kernel.hpp:272-292,
// guiggiani_1992.hpp:215-252); otherwise as quad_guiggiani.
template <class K>
static void gauss_guiggiani(int family, K k, QElem const &e, double xi0x, double xi0y, int rad_order, int tan_order,
	C4 &out)
{
	V3 dxi, deta;
	q_dx(e, xi0x, xi0y, dxi, deta);
	V3 dxe = q_dxe(e);
	V3 nJ = cross(dxi, deta);
	V3 n0 = normalized(nJ);
	V3 x0 = q_x(e, xi0x, xi0y);
	double N0[4], dNx[4], dNy[4];
	gauss_shape(xi0x, xi0y, N0);
	gauss_dshape(xi0x, xi0y, dNx, dNy);

	double la = norm(dxi), lb = norm(deta);
	double cosg = dot(dxi, deta) / la / lb;
	double u = lb / la;
	double T00 = la, T01 = la * cosg * u, T10 = 0.0, T11 = la * std::sqrt(1.0 - cosg * cosg) * u;
	double detT = T00 * T11 - T10 * T01;
	double Ti00 = T11 / detT, Ti01 = -T01 / detT, Ti10 = -T10 / detT, Ti11 = T00 / detT;
	double e0x = T00 * xi0x + T01 * xi0y, e0y = T10 * xi0x + T11 * xi0y;
	V3 J0 = nJ / detT;

	double theta_lim[4], theta0[4], ref_distance[4];
	for (int n = 0; n < 4; ++n) {
		double c1x = T00 * QCX[n] + T01 * QCY[n], c1y = T10 * QCX[n] + T11 * QCY[n];
		int m = (n + 1) % 4;
		double c2x = T00 * QCX[m] + T01 * QCY[m], c2y = T10 * QCX[m] + T11 * QCY[m];
		double lx = 1.0, ly = 0.0;
		double ln = std::sqrt((c2x - c1x) * (c2x - c1x) + (c2y - c1y) * (c2y - c1y));
		if (ln > 1e-3) { lx = (c2x - c1x) / ln; ly = (c2y - c1y) / ln; }
		double d1x = c1x - e0x, d1y = c1y - e0y;
		double dl = d1x * lx + d1y * ly;
		double d0x = d1x - lx * dl, d0y = d1y - ly * dl;
		theta_lim[n] = std::atan2(d1y, d1x);
		theta0[n] = std::atan2(d0y, d0x);
		ref_distance[n] = std::sqrt(d0x * d0x + d0y * d0y);
	}

	auto const &gt = gauss_line(tan_order);
	auto const &gr = gauss_line(rad_order);
	c4_zero(out);
	for (int n = 0; n < 4; ++n) {
		double t1 = theta_lim[n], t2 = theta_lim[(n + 1) % 4];
		if (std::abs(t2 - t1) < 1e-3) continue;
		if (std::abs(t2 - t1) > PI) t2 += TWO_PI;
		for (auto const &qt : gt) {
			double theta = ((1.0 - qt.first) * t1 + (1.0 + qt.first) * t2) / 2.0;
			double w_theta = qt.second * (t2 - t1) / 2.0;
			double ct = std::cos(theta), st = std::sin(theta);
			double vx = Ti00 * ct + Ti01 * st, vy = Ti10 * ct + Ti11 * st;
			V3 r1 = dxi * vx + deta * vy;
			V3 r2 = dxe * (vx * vy);
			V3 J1 = (cross(dxi, dxe) * vx + cross(dxe, deta) * vy) / detT;
			V3 g1 = r1 * (dot(r2, J0) + dot(r1, J1));
			V3 b0 = J0 * -1.0;
			V3 b1 = g1 * 3.0 - J1;
			double a0 = dot(b0, n0);
			double a1 = dot(b1, n0);
			double Sm2 = -3.0 * dot(r1, r2);
			double rho_lim = ref_distance[n] / std::cos(theta - theta0[n]);
			double Fm1[4], Fm2[4];
			for (int b = 0; b < 4; ++b) {
				double A0 = a0 * N0[b];
				double A1 = a1 * N0[b] + a0 * (dNx[b] * vx + dNy[b] * vy);
				Fm1[b] = -(Sm2 * A0 + A1) / (4.0 * PI);
				Fm2[b] = -A0 / (4.0 * PI);
				out[b] += (Fm1[b] * std::log(rho_lim) - Fm2[b] / rho_lim) * w_theta;
			}
			for (auto const &qr : gr) {
				double rho = (1.0 + qr.first) * rho_lim / 2.0;
				double w_rho = qr.second * rho_lim / 2.0;
				double ex = rho * ct, ey = rho * st;
				double xi = Ti00 * ex + Ti01 * ey + xi0x, eta = Ti10 * ex + Ti11 * ey + xi0y;
				V3 y = q_x(e, xi, eta);
				V3 nv = q_normal(e, xi, eta);
				double J = norm(nv);
				cplx F = kernel_eval(family, 1, 1, k, x0, y, n0, nv / J) * (J / detT * rho);
				double N[4];
				gauss_shape(xi, eta, N);
				for (int b = 0; b < 4; ++b) {
					double sing = (Fm1[b] + Fm2[b] / rho) / rho;
					out[b] += (F * N[b] - sing) * (w_theta * w_rho);
				}
			}
		}
	}
}

// the same element, collocation at node a
template <class K>
static void gauss_singular(int family, int nx, int ny, K k, QElem const &e, int a, C4 &out)
{
	double xi0, eta0;
	gauss_node(a, xi0, eta0);
	if (nx == 1 && ny == 1) return gauss_guiggiani(family, k, e, xi0, eta0, 5, 9, out);
	gauss_duffy(family, nx, ny, k, e, xi0, eta0, 7, out);
}

// One block row: collocation point x (unit normal nxv) of the test element
// with centre tc and id tid against a quadrilateral trial element with the
// gauss field; node a >= 0: x is the test element's node a (same-mesh
// singular check), else a constant test field.
template <class K>
static void gauss_pair_row(int family, int nx, int ny, K k, Policy const &p, V3 x, V3 nxv, V3 tc, int tid, int a,
	QElem const &trial, bool same_mesh, C4 &out)
{
	if (same_mesh && a >= 0 && tid == trial.id) return gauss_singular(family, nx, ny, k, trial, a, out);
	double d2 = dist2(tc, trial.c);
	if (near_needed_q(p, d2, trial)) return gauss_near(family, nx, ny, k, trial, x, nxv, out);
	int degree = p.rule == 1 ? p.fixed_degree
		: 2 + eval_interval(interval_table(1 + nx + ny), std::sqrt(d2) / trial.R);
	gauss_regular(family, nx, ny, k, x, nxv, trial, degree, out);
}

} // namespace orc

// ===========================================================================
// C interface (ctypes)
// ===========================================================================
using namespace orc;

static thread_local std::string g_err;

#define ORC_TRY try {
#define ORC_CATCH } catch (std::out_of_range const &ex) { g_err = ex.what(); return 2; } \
	catch (std::exception const &ex) { g_err = ex.what(); return 1; } return 0;

extern "C" {

typedef struct { int rule; int fixed_degree; int near_mode; double near_limit; } orc_policy;

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_num_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

int orc_dunavant(int degree, double *xi, double *eta, double *w, int cap, int *n_out)
{
	ORC_TRY
	auto q = dunavant(degree);
	if ((int)q.size() > cap) { g_err = "capacity"; return 1; }
	for (size_t i = 0; i < q.size(); ++i) { xi[i] = q[i].xi; eta[i] = q[i].eta; w[i] = q[i].w; }
	*n_out = (int)q.size();
	ORC_CATCH
}

int orc_gauss_legendre(int n, double *x, double *w)
{
	auto g = gauss_legendre_n(n);
	for (int i = 0; i < n; ++i) { x[i] = g[i].first; w[i] = g[i].second; }
	return 0;
}

int orc_geometry(const double *xyz, int n_nodes, const int *tri, int n_tri,
	double *centers, double *unit_normals, double *jac, double *R, double *diam)
{
	ORC_TRY
	auto m = make_mesh(xyz, n_nodes, tri, n_tri);
	for (int e = 0; e < n_tri; ++e) {
		centers[3 * e] = m[e].c.x; centers[3 * e + 1] = m[e].c.y; centers[3 * e + 2] = m[e].c.z;
		unit_normals[3 * e] = m[e].nu.x; unit_normals[3 * e + 1] = m[e].nu.y; unit_normals[3 * e + 2] = m[e].nu.z;
		jac[e] = m[e].J; R[e] = m[e].R; diam[e] = m[e].diam;
	}
	ORC_CATCH
}

int orc_kernel(int family, int nx, int ny, double k, const double *x, const double *y,
	const double *nxv, const double *nyv, double *out)
{
	ORC_TRY
	cplx v = kernel_eval(family, nx, ny, k, {x[0], x[1], x[2]}, {y[0], y[1], y[2]},
		{nxv[0], nxv[1], nxv[2]}, {nyv[0], nyv[1], nyv[2]});
	out[0] = v.real(); out[1] = v.imag();
	ORC_CATCH
}

// complex wavenumber k_re + i k_im
int orc_kernel_ck(int family, int nx, int ny, double k_re, double k_im, const double *x, const double *y,
	const double *nxv, const double *nyv, double *out)
{
	ORC_TRY
	cplx v = kernel_eval(family, nx, ny, cplx(k_re, k_im), {x[0], x[1], x[2]}, {y[0], y[1], y[2]},
		{nxv[0], nxv[1], nxv[2]}, {nyv[0], nyv[1], nyv[2]});
	out[0] = v.real(); out[1] = v.imag();
	ORC_CATCH
}

// kind: 0 regular (Dunavant degree), 1 near-singular, 2 singular (x0 = centre).
// For kind 1, degree > 0 replaces NiHu's line quadrature order of the static
// part (used by tests to converge the formula to reference golden values).
int orc_elem_integral(int kind, int family, int nx, int ny, double k, const double *coords9,
	const double *x0, const double *nx0, int degree, double *out)
{
	ORC_TRY
	Elem e = make_elem({coords9[0], coords9[1], coords9[2]}, {coords9[3], coords9[4], coords9[5]},
		{coords9[6], coords9[7], coords9[8]}, 0);
	V3 x = {x0[0], x0[1], x0[2]}, n = {nx0[0], nx0[1], nx0[2]};
	cplx v;
	if (kind == 0) v = regular_integral(family, nx, ny, k, x, n, e, degree);
	else if (kind == 1) v = near_integral(family, nx, ny, k, e, x, n, degree);
	else v = singular_integral(family, nx, ny, k, e);
	out[0] = v.real(); out[1] = v.imag();
	ORC_CATCH
}

// complex wavenumber k_re + i k_im
int orc_elem_integral_ck(int kind, int family, int nx, int ny, double k_re, double k_im, const double *coords9,
	const double *x0, const double *nx0, int degree, double *out)
{
	ORC_TRY
	Elem e = make_elem({coords9[0], coords9[1], coords9[2]}, {coords9[3], coords9[4], coords9[5]},
		{coords9[6], coords9[7], coords9[8]}, 0);
	V3 x = {x0[0], x0[1], x0[2]}, n = {nx0[0], nx0[1], nx0[2]};
	const cplx k(k_re, k_im);
	cplx v;
	if (kind == 0) v = regular_integral(family, nx, ny, k, x, n, e, degree);
	else if (kind == 1) v = near_integral(family, nx, ny, k, e, x, n, degree);
	else v = singular_integral(family, nx, ny, k, e);
	out[0] = v.real(); out[1] = v.imag();
	ORC_CATCH
}

// Pair classes for a block of rows: -1 diagonal, -2 near, otherwise the regular
// Dunavant degree.  cls is row-major [n_rows][n_trial].
int orc_classify(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh,
	int nx, int ny, const orc_policy *pol, const int *rows, int n_rows, int8_t *cls)
{
	ORC_TRY
	auto test = make_mesh(txyz, tn, ttri, tnt);
	auto trial = same_mesh ? test : make_mesh(sxyz, sn, stri, snt);
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	int nr = rows ? n_rows : tnt;
	#pragma omp parallel for schedule(dynamic, 8)
	for (int r = 0; r < nr; ++r) {
		int i = rows ? rows[r] : r;
		for (int j = 0; j < snt; ++j) {
			int8_t c;
			if (same_mesh && i == j) c = -1;
			else {
				double d2 = dist2(test[i].c, trial[j].c);
				if (near_needed(p, d2, trial[j])) c = -2;
				else c = (int8_t)regular_degree(p, 0, nx, ny, d2, trial[j]);
			}
			cls[(int64_t)r * snt + j] = c;
		}
	}
	ORC_CATCH
}

// result(r, j) (+)= alpha * integral(test row rows[r], trial j); complex
// interleaved, column-major with leading dimension ld (ld >= n_rows).
// nihu/core/assembly.hpp:61-97 (the pair loop, accumulating through
// nihu/util/matrix_block.hpp:59-65); the scalar multiplies each pair result
// (nihu/core/integral_operator.hpp:159-167).  OpenMP over rows is the oracle's
// own extension; with n_threads = 1 it is NiHu's serial loop.
} // extern "C"

template <class K>
static int assemble_impl(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh,
	int family, int nx, int ny, K k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	ORC_TRY
	auto test = make_mesh(txyz, tn, ttri, tnt);
	auto trial = same_mesh ? test : make_mesh(sxyz, sn, stri, snt);
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	int nr = rows ? n_rows : tnt;
	cplx alpha(alpha_re, alpha_im);
	if (n_threads <= 0) n_threads = orc_num_threads();
	std::string err;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
	for (int r = 0; r < nr; ++r) {
		try {
			int i = rows ? rows[r] : r;
			for (int j = 0; j < snt; ++j) {
				cplx v = alpha * pair_integral(family, nx, ny, k, p, test[i], trial[j], same_mesh != 0);
				double *o = out + 2 * ((int64_t)j * ld + r);
				if (accumulate) { o[0] += v.real(); o[1] += v.imag(); }
				else { o[0] = v.real(); o[1] = v.imag(); }
			}
		} catch (std::exception const &ex) {
			#pragma omp critical
			err = ex.what();
		}
	}
	if (!err.empty()) throw std::out_of_range(err);
	ORC_CATCH
}


extern "C" {

int orc_assemble(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh,
	int family, int nx, int ny, double k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	return assemble_impl(txyz, tn, ttri, tnt, sxyz, sn, stri, snt, same_mesh, family, nx, ny, k, alpha_re, alpha_im,
		pol, rows, n_rows, out, ld, accumulate, n_threads);
}

// the same with a complex wavenumber k_re + i k_im
int orc_assemble_ck(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh,
	int family, int nx, int ny, double k_re, double k_im, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	return assemble_impl(txyz, tn, ttri, tnt, sxyz, sn, stri, snt, same_mesh, family, nx, ny, cplx(k_re, k_im),
		alpha_re, alpha_im, pol, rows, n_rows, out, ld, accumulate, n_threads);
}

// identity operator: result(i,i) += alpha on the same mesh (local branch of
// nihu/core/assembly.hpp:78-96; value N_trial(centre) = 1,
// nihu/core/single_integral.hpp:146-178)
int orc_identity(int n, double alpha_re, double alpha_im, const int *rows, int n_rows, double *out, int64_t ld)
{
	int nr = rows ? n_rows : n;
	for (int r = 0; r < nr; ++r) {
		int i = rows ? rows[r] : r;
		double *o = out + 2 * ((int64_t)i * ld + r);
		o[0] += alpha_re; o[1] += alpha_im;
	}
	return 0;
}

// Burton-Miller collocation matrices in NiHu's composition order
// (nihu/applications/helmholtz/helmholtz_3d_coll.cpp:41-96):
//   B = Gs: G, then alpha*Ht, then alpha/2*I;   A = Hs: alpha*D, then -1/2*I, then H
// alpha = -i/k there.  Both outputs are overwritten (setZero first).
int orc_assemble_bm(const double *xyz, int nn, const int *tri, int nt, double k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int64_t ld, int n_threads)
{
	int nr = rows ? n_rows : nt;
	for (int j = 0; j < nt; ++j)
		for (int r = 0; r < nr; ++r) {
			double *a = A + 2 * ((int64_t)j * ld + r), *b = B + 2 * ((int64_t)j * ld + r);
			a[0] = a[1] = b[0] = b[1] = 0.0;
		}
	cplx al(alpha_re, alpha_im);
	int s;
	if ((s = orc_assemble(xyz, nn, tri, nt, xyz, nn, tri, nt, 1, 1, 0, 0, k, 1.0, 0.0, pol, rows, n_rows, B, ld, 1, n_threads))) return s;
	if ((s = orc_assemble(xyz, nn, tri, nt, xyz, nn, tri, nt, 1, 1, 1, 0, k, al.real(), al.imag(), pol, rows, n_rows, B, ld, 1, n_threads))) return s;
	orc_identity(nt, (al / 2.).real(), (al / 2.).imag(), rows, n_rows, B, ld);
	if ((s = orc_assemble(xyz, nn, tri, nt, xyz, nn, tri, nt, 1, 1, 1, 1, k, al.real(), al.imag(), pol, rows, n_rows, A, ld, 1, n_threads))) return s;
	orc_identity(nt, -.5, 0.0, rows, n_rows, A, ld);
	if ((s = orc_assemble(xyz, nn, tri, nt, xyz, nn, tri, nt, 1, 1, 0, 1, k, 1.0, 0.0, pol, rows, n_rows, A, ld, 1, n_threads))) return s;
	return 0;
}

// the same with a complex wavenumber k_re + i k_im (alpha as given)
int orc_assemble_bm_ck(const double *xyz, int nn, const int *tri, int nt, double k_re, double k_im, double alpha_re,
	double alpha_im, const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int64_t ld,
	int n_threads)
{
	int nr = rows ? n_rows : nt;
	for (int j = 0; j < nt; ++j)
		for (int r = 0; r < nr; ++r) {
			double *a = A + 2 * ((int64_t)j * ld + r), *b = B + 2 * ((int64_t)j * ld + r);
			a[0] = a[1] = b[0] = b[1] = 0.0;
		}
	cplx al(alpha_re, alpha_im);
	int s;
	auto op = [&](int nx, int ny, cplx sc, double *out) {
		return orc_assemble_ck(xyz, nn, tri, nt, xyz, nn, tri, nt, 1, 1, nx, ny, k_re, k_im, sc.real(), sc.imag(), pol,
			rows, n_rows, out, ld, 1, n_threads);
	};
	if ((s = op(0, 0, 1.0, B))) return s;
	if ((s = op(1, 0, al, B))) return s;
	orc_identity(nt, (al / 2.).real(), (al / 2.).imag(), rows, n_rows, B, ld);
	if ((s = op(1, 1, al, A))) return s;
	orc_identity(nt, -.5, 0.0, rows, n_rows, A, ld);
	if ((s = op(0, 1, 1.0, A))) return s;
	return 0;
}

// NiHu's own multithreaded pattern: one whole Burton-Miller assembly per
// wavenumber per thread (benchmark/EAA/sphere/homogeneous/bem/bem.cpp:91).
// Each frequency writes its own A_f, B_f (n_rows x nt blocks, ld = n_rows),
// the output arrays are [n_k][nt][n_rows] complex.
int orc_sweep_bm(const double *xyz, int nn, const int *tri, int nt, const double *ks, int n_k,
	const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int n_threads)
{
	int nr = rows ? n_rows : nt;
	int64_t blk = 2 * (int64_t)nr * nt;
	if (n_threads <= 0) n_threads = orc_num_threads();
	int status = 0;
	#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
	for (int f = 0; f < n_k; ++f) {
		double k = ks[f];
		int s = orc_assemble_bm(xyz, nn, tri, nt, k, 0.0, -1.0 / k, pol, rows, n_rows, A + f * blk, B + f * blk, nr, 1);
		if (s) {
			#pragma omp critical
			status = s;
		}
	}
	return status;
}


// Galerkin assembly, result(dof_a, dof_b) (+)= alpha * block (complex interleaved,
// column-major, ld >= test dofs).  linear = 1: isoparametric_view (dofs = mesh
// nodes, nihu/core/field.hpp:234-258); linear = 0: constant_view (dofs =
// elements).  family 2 = identity (numerical mass matrix,
// nihu/core/single_integral.hpp:65-140).
} // extern "C"

template <class K>
static int assemble_galerkin_impl(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh, int linear,
	int family, int nx, int ny, K k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld, int accumulate,
	int n_threads)
{
	ORC_TRY
	auto test = make_mesh(txyz, tn, ttri, tnt);
	auto trial = same_mesh ? test : make_mesh(sxyz, sn, stri, snt);
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	cplx alpha(alpha_re, alpha_im);
	const bool lt = linear_test(linear), ls = linear_trial(linear);
	const int ndt = lt ? 3 : 1, nds = ls ? 3 : 1;
	// dof of (element, local shape function): node id, or the element itself
	if (same_mesh) stri = ttri;
	auto tdof = [&](int e, int a) { return lt ? ttri[3 * e + a] : e; };
	auto sdof = [&](int e, int b) { return ls ? stri[3 * e + b] : e; };
	int nrows = lt ? tn : tnt;
	int ncols = ls ? (same_mesh ? tn : sn) : (same_mesh ? tnt : snt);
	if (!accumulate)
		for (int j = 0; j < ncols; ++j)
			for (int i = 0; i < nrows; ++i) { out[2 * ((int64_t)j * ld + i)] = 0; out[2 * ((int64_t)j * ld + i) + 1] = 0; }
	if (family == 2) {
		// identity: single integral of N_a N_b J with the Dunavant degree
		// test order + trial order + jacobian order
		if (!same_mesh) return 0;
		auto const &qm = dunavant_store((lt ? 1 : 0) + (ls ? 1 : 0));
		for (int e = 0; e < tnt; ++e) {
			for (int a = 0; a < ndt; ++a)
				for (int b = 0; b < nds; ++b) {
					double m = 0.0;
					for (auto const &q : qm) {
						double Nt[3] = {1.0 - q.xi - q.eta, q.xi, q.eta}, Ns[3] = {Nt[0], Nt[1], Nt[2]};
						if (!lt) Nt[0] = 1.0;
						if (!ls) Ns[0] = 1.0;
						m += Nt[a] * Ns[b] * test[e].J * q.w;
					}
					cplx v = alpha * m;
					double *o = out + 2 * ((int64_t)sdof(e, b) * ld + tdof(e, a));
					o[0] += v.real(); o[1] += v.imag();
				}
		}
		return 0;
	}
	if (n_threads <= 0) n_threads = orc_num_threads();
	std::string err;
	int errcode = 0;
	// test_elems: only these test elements contribute (rows of dofs whose
	// elements are all listed come out complete)
	const int n_loop = test_elems ? n_test_elems : tnt;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
	for (int ii = 0; ii < n_loop; ++ii) {
		const int i = test_elems ? test_elems[ii] : ii;
		try {
			for (int j = 0; j < snt; ++j) {
				cplx blk[3][3];
				galerkin_block(family, nx, ny, k, p, test[i], trial[j], ttri + 3 * i, stri + 3 * j, same_mesh != 0,
					linear, blk);
				for (int a = 0; a < ndt; ++a)
					for (int b = 0; b < nds; ++b) {
						cplx v = alpha * blk[a][b];
						double *o = out + 2 * ((int64_t)sdof(j, b) * ld + tdof(i, a));
						#pragma omp atomic
						o[0] += v.real();
						#pragma omp atomic
						o[1] += v.imag();
					}
			}
		} catch (std::out_of_range const &ex) {
			#pragma omp critical
			{ err = ex.what(); errcode = 2; }
		} catch (std::exception const &ex) {
			#pragma omp critical
			{ err = ex.what(); errcode = 1; }
		}
	}
	if (errcode == 2) throw std::out_of_range(err);
	if (errcode == 1) throw std::runtime_error(err);
	ORC_CATCH
}


extern "C" {

int orc_assemble_galerkin(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh, int linear,
	int family, int nx, int ny, double k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld, int accumulate,
	int n_threads)
{
	return assemble_galerkin_impl(txyz, tn, ttri, tnt, sxyz, sn, stri, snt, same_mesh, linear, family, nx, ny, k,
		alpha_re, alpha_im, pol, test_elems, n_test_elems, out, ld, accumulate, n_threads);
}

// the same with a complex wavenumber k_re + i k_im
int orc_assemble_galerkin_ck(const double *txyz, int tn, const int *ttri, int tnt,
	const double *sxyz, int sn, const int *stri, int snt, int same_mesh, int linear,
	int family, int nx, int ny, double k_re, double k_im, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld, int accumulate,
	int n_threads)
{
	return assemble_galerkin_impl(txyz, tn, ttri, tnt, sxyz, sn, stri, snt, same_mesh, linear, family, nx, ny,
		cplx(k_re, k_im), alpha_re, alpha_im, pol, test_elems, n_test_elems, out, ld, accumulate, n_threads);
}

// node overlap of two elements as NiHu's scan finds it: returns the count
int orc_overlap(const int *a, const int *b, int *ind1, int *ind2)
{
	Overlap o = get_overlapping(a, b);
	*ind1 = o.ind1;
	*ind2 = o.ind2;
	return o.num;
}

// canonical singular rule of a match dimension: n points (txi, teta, sxi, seta, w)
int orc_singular_rule(int match, double *pts, int cap, int *n_out)
{
	auto const &r = singular_rule(match);
	if ((int)r.size() > cap) { g_err = "capacity"; return 1; }
	for (size_t q = 0; q < r.size(); ++q) {
		pts[5 * q] = r[q].txi; pts[5 * q + 1] = r[q].teta; pts[5 * q + 2] = r[q].sxi;
		pts[5 * q + 3] = r[q].seta; pts[5 * q + 4] = r[q].w;
	}
	*n_out = (int)r.size();
	return 0;
}

// one Galerkin block of an element pair given by coordinates and node ids
int orc_galerkin_block(int family, int nx, int ny, double k, const orc_policy *pol, const double *tc9, const int *tn3,
	const double *sc9, const int *sn3, int same_mesh, int same_elem, int linear, double *blk18)
{
	ORC_TRY
	Elem te = make_elem({tc9[0], tc9[1], tc9[2]}, {tc9[3], tc9[4], tc9[5]}, {tc9[6], tc9[7], tc9[8]}, 0);
	Elem se = make_elem({sc9[0], sc9[1], sc9[2]}, {sc9[3], sc9[4], sc9[5]}, {sc9[6], sc9[7], sc9[8]}, same_elem ? 0 : 1);
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	cplx blk[3][3];
	galerkin_block(family, nx, ny, k, p, te, se, tn3, sn3, same_mesh != 0, linear, blk);
	for (int a = 0; a < 3; ++a)
		for (int b = 0; b < 3; ++b) { blk18[2 * (3 * a + b)] = blk[a][b].real(); blk18[2 * (3 * a + b) + 1] = blk[a][b].imag(); }
	ORC_CATCH
}


} // extern "C"

template <class K>
static int assemble_elems_impl(const double *txyz, int tn, const int *tel, int tnt, int tnpe,
	const double *sxyz, int sn, const int *sel, int snt, int snpe, int same_mesh,
	int family, int nx, int ny, K k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	ORC_TRY
	if (same_mesh && tnpe != snpe) throw std::invalid_argument("same mesh with different element types");
	std::vector<Elem> ttri, stri;
	std::vector<QElem> tq, sq;
	std::vector<TestPt> tp;
	read_elems(txyz, tn, tel, tnt, tnpe, &ttri, &tq, &tp);
	if (!same_mesh) read_elems(sxyz, sn, sel, snt, snpe, &stri, &sq, nullptr);
	std::vector<Elem> const &trit = same_mesh ? ttri : stri;
	std::vector<QElem> const &quadt = same_mesh ? tq : sq;
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	int nr = rows ? n_rows : tnt;
	cplx alpha(alpha_re, alpha_im);
	if (n_threads <= 0) n_threads = orc_num_threads();
	std::string err;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
	for (int r = 0; r < nr; ++r) {
		try {
			int i = rows ? rows[r] : r;
			for (int j = 0; j < snt; ++j) {
				cplx v;
				if (snpe == 4) v = quad_pair_integral(family, nx, ny, k, p, tp[i], quadt[j], same_mesh != 0);
				else if (tnpe == 3) v = pair_integral(family, nx, ny, k, p, ttri[i], trit[j], same_mesh != 0);
				else {
					// quadrilateral test point against a triangle: never the same element
					Elem const &trial = trit[j];
					double d2 = dist2(tp[i].x, trial.c);
					if (near_needed(p, d2, trial)) v = near_integral(family, nx, ny, k, trial, tp[i].x, tp[i].nu);
					else v = regular_integral(family, nx, ny, k, tp[i].x, tp[i].nu, trial,
						regular_degree(p, 0, nx, ny, d2, trial));
				}
				v *= alpha;
				double *o = out + 2 * ((int64_t)j * ld + r);
				if (accumulate) { o[0] += v.real(); o[1] += v.imag(); }
				else { o[0] = v.real(); o[1] = v.imag(); }
			}
		} catch (std::exception const &ex) {
			#pragma omp critical
			err = ex.what();
		}
	}
	if (!err.empty()) throw std::out_of_range(err);
	ORC_CATCH
}

extern "C" {

// Collocation with constant fields on meshes of triangles (npe = 3) or
// quadrilaterals (npe = 4); test and trial meshes may differ in type.
int orc_assemble_elems(const double *txyz, int tn, const int *tel, int tnt, int tnpe,
	const double *sxyz, int sn, const int *sel, int snt, int snpe, int same_mesh,
	int family, int nx, int ny, double k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	return assemble_elems_impl(txyz, tn, tel, tnt, tnpe, sxyz, sn, sel, snt, snpe, same_mesh, family, nx, ny, k,
		alpha_re, alpha_im, pol, rows, n_rows, out, ld, accumulate, n_threads);
}

// the same with a complex wavenumber k_re + i k_im
int orc_assemble_elems_ck(const double *txyz, int tn, const int *tel, int tnt, int tnpe,
	const double *sxyz, int sn, const int *sel, int snt, int snpe, int same_mesh,
	int family, int nx, int ny, double k_re, double k_im, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	return assemble_elems_impl(txyz, tn, tel, tnt, tnpe, sxyz, sn, sel, snt, snpe, same_mesh, family, nx, ny,
		cplx(k_re, k_im), alpha_re, alpha_im, pol, rows, n_rows, out, ld, accumulate, n_threads);
}

// Burton-Miller A and B (as orc_assemble_bm) on a mesh of npe-node elements
int orc_assemble_bm_elems(const double *xyz, int nn, const int *el, int ne, int npe, double k, double alpha_re,
	double alpha_im, const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int64_t ld, int n_threads)
{
	int nr = rows ? n_rows : ne;
	for (int j = 0; j < ne; ++j)
		for (int r = 0; r < nr; ++r) {
			double *a = A + 2 * ((int64_t)j * ld + r), *b = B + 2 * ((int64_t)j * ld + r);
			a[0] = a[1] = b[0] = b[1] = 0.0;
		}
	cplx al(alpha_re, alpha_im);
	int s;
	if ((s = orc_assemble_elems(xyz, nn, el, ne, npe, xyz, nn, el, ne, npe, 1, 1, 0, 0, k, 1.0, 0.0, pol, rows, n_rows,
		B, ld, 1, n_threads))) return s;
	if ((s = orc_assemble_elems(xyz, nn, el, ne, npe, xyz, nn, el, ne, npe, 1, 1, 1, 0, k, al.real(), al.imag(), pol,
		rows, n_rows, B, ld, 1, n_threads))) return s;
	orc_identity(ne, (al / 2.).real(), (al / 2.).imag(), rows, n_rows, B, ld);
	if ((s = orc_assemble_elems(xyz, nn, el, ne, npe, xyz, nn, el, ne, npe, 1, 1, 1, 1, k, al.real(), al.imag(), pol,
		rows, n_rows, A, ld, 1, n_threads))) return s;
	orc_identity(ne, -.5, 0.0, rows, n_rows, A, ld);
	if ((s = orc_assemble_elems(xyz, nn, el, ne, npe, xyz, nn, el, ne, npe, 1, 1, 0, 1, k, 1.0, 0.0, pol, rows, n_rows,
		A, ld, 1, n_threads))) return s;
	return 0;
}

// the same with a complex wavenumber k_re + i k_im (alpha as given)
int orc_assemble_bm_elems_ck(const double *xyz, int nn, const int *el, int ne, int npe, double k_re, double k_im,
	double alpha_re, double alpha_im, const orc_policy *pol, const int *rows, int n_rows, double *A, double *B,
	int64_t ld, int n_threads)
{
	int nr = rows ? n_rows : ne;
	for (int j = 0; j < ne; ++j)
		for (int r = 0; r < nr; ++r) {
			double *a = A + 2 * ((int64_t)j * ld + r), *b = B + 2 * ((int64_t)j * ld + r);
			a[0] = a[1] = b[0] = b[1] = 0.0;
		}
	cplx al(alpha_re, alpha_im);
	int s;
	auto op = [&](int nx, int ny, cplx sc, double *out) {
		return orc_assemble_elems_ck(xyz, nn, el, ne, npe, xyz, nn, el, ne, npe, 1, 1, nx, ny, k_re, k_im, sc.real(),
			sc.imag(), pol, rows, n_rows, out, ld, 1, n_threads);
	};
	if ((s = op(0, 0, 1.0, B))) return s;
	if ((s = op(1, 0, al, B))) return s;
	orc_identity(ne, (al / 2.).real(), (al / 2.).imag(), rows, n_rows, B, ld);
	if ((s = op(1, 1, al, A))) return s;
	orc_identity(ne, -.5, 0.0, rows, n_rows, A, ld);
	if ((s = op(0, 1, 1.0, A))) return s;
	return 0;
}

// Element-level quadrilateral integrals for the tests.  coords12: the four
// corners; x0, nx0: collocation point and unit normal (what 0-3, 7, 8);
// xi0: a point of the reference square (what 4-6).
//   0 regular, Gauss degree o1          4 Guiggiani HSP at xi0, orders (o1, o2)
//   1 singular at the centre            5 blind Duffy rule of degree o1 at xi0
//   2 polar near, orders (o1, o2)       6 inverse mapping of x0: (xi, eta, zeta, ok)
//   3 Telles near, degree o1            7 near dispatch (NiHu's choice per kernel)
//   8 planar integral over the element itself (planar elements only)
} // extern "C"

template <class K>
static int quad_elem_impl(int what, int family, int nx, int ny, K k, const double *coords12, const double *x0p,
	const double *nx0p, const double *xi0p, int o1, int o2, double *out)
{
	ORC_TRY
	V3 X[4];
	for (int a = 0; a < 4; ++a) X[a] = {coords12[3 * a], coords12[3 * a + 1], coords12[3 * a + 2]};
	QElem e = make_qelem(X, 0);
	V3 x0 = x0p ? V3{x0p[0], x0p[1], x0p[2]} : e.c;
	V3 n0 = nx0p ? V3{nx0p[0], nx0p[1], nx0p[2]} : e.nu;
	cplx v = 0.0;
	switch (what) {
	case 0: v = quad_regular(family, nx, ny, k, x0, n0, e, o1); break;
	case 1: v = quad_singular(family, nx, ny, k, e); break;
	case 2: v = quad_near_polar(family, nx, ny, k, e, x0, n0, o1, o2); break;
	case 3: v = quad_near_telles(family, nx, ny, k, e, x0, n0, o1); break;
	case 4: v = quad_guiggiani(family, k, e, xi0p[0], xi0p[1], o1, o2); break;
	case 5: {
		V3 xs = q_x(e, xi0p[0], xi0p[1]);
		V3 ns = normalized(q_normal(e, xi0p[0], xi0p[1]));
		for (auto const &q : quad_duffy_face(o1, xi0p[0], xi0p[1]))
			v += q_kernel_jac(family, nx, ny, k, xs, ns, e, q.xi, q.eta) * q.w;
		break;
	}
	case 6: {
		// no image: NaN coordinates beside the flag
		double xi = std::nan(""), eta = std::nan(""), zeta = std::nan("");
		bool ok = quad_inverse_map(e, x0, o1 > 0 ? std::pow(10.0, -o1) : 1e-5, 100, xi, eta, zeta);
		out[0] = xi; out[1] = eta; out[2] = zeta; out[3] = ok ? 1.0 : 0.0;
		return 0;
	}
	case 7: v = quad_near(family, nx, ny, k, e, x0, n0); break;
	case 8: v = quad_planar_near(family, nx, ny, k, e, x0, n0); break;
	default: throw std::invalid_argument("unknown element integral");
	}
	out[0] = v.real(); out[1] = v.imag();
	ORC_CATCH
}

extern "C" {

int orc_quad_elem(int what, int family, int nx, int ny, double k, const double *coords12, const double *x0p,
	const double *nx0p, const double *xi0p, int o1, int o2, double *out)
{
	return quad_elem_impl(what, family, nx, ny, k, coords12, x0p, nx0p, xi0p, o1, o2, out);
}

// the same with a complex wavenumber k_re + i k_im
int orc_quad_elem_ck(int what, int family, int nx, int ny, double k_re, double k_im, const double *coords12,
	const double *x0p, const double *nx0p, const double *xi0p, int o1, int o2, double *out)
{
	return quad_elem_impl(what, family, nx, ny, cplx(k_re, k_im), coords12, x0p, nx0p, xi0p, o1, o2, out);
}

// Guiggiani / Duffy / Telles rules for the tests: the Duffy rule of degree o1
// around xi0 as (xi, eta, w) triples
int orc_quad_duffy_rule(int degree, double xi0, double eta0, double *pts, int cap, int *n_out)
{
	ORC_TRY
	auto r = quad_duffy_face(degree, xi0, eta0);
	*n_out = (int)r.size();
	if ((int)r.size() > cap) throw std::invalid_argument("buffer too small");
	for (size_t i = 0; i < r.size(); ++i) { pts[3 * i] = r[i].xi; pts[3 * i + 1] = r[i].eta; pts[3 * i + 2] = r[i].w; }
	ORC_CATCH
}

int orc_telles_rule(int degree, double xi0, double eta0, double d, double *pts, int cap, int *n_out)
{
	ORC_TRY
	auto r = telles_rule(degree, xi0, eta0, d);
	*n_out = (int)r.size();
	if ((int)r.size() > cap) throw std::invalid_argument("buffer too small");
	for (size_t i = 0; i < r.size(); ++i) { pts[3 * i] = r[i].xi; pts[3 * i + 1] = r[i].eta; pts[3 * i + 2] = r[i].w; }
	ORC_CATCH
}

// quadrilateral geometry as the oracle computes it: centre, unit normal at the
// centre, mean edge, diameter (n x 3, n x 3, n, n)
int orc_geometry_quad(const double *xyz, int n_nodes, const int *quad, int n, double *c, double *nu, double *R,
	double *diam)
{
	ORC_TRY
	std::vector<QElem> q;
	read_elems(xyz, n_nodes, quad, n, 4, nullptr, &q, nullptr);
	for (int e = 0; e < n; ++e) {
		c[3 * e] = q[e].c.x; c[3 * e + 1] = q[e].c.y; c[3 * e + 2] = q[e].c.z;
		nu[3 * e] = q[e].nu.x; nu[3 * e + 1] = q[e].nu.y; nu[3 * e + 2] = q[e].nu.z;
		R[e] = q[e].R;
		diam[e] = q[e].diam;
	}
	ORC_CATCH
}

} // extern "C"

// ===========================================================================
// quad_1_gauss_field collocation
// ===========================================================================
template <class K>
static int assemble_gauss_impl(const double *txyz, int tn, const int *tel, int tnt, int tnpe,
	const double *sxyz, int sn, const int *sel, int snt, int same_mesh, int test_gauss,
	int family, int nx, int ny, K k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	ORC_TRY
	if (test_gauss && tnpe != 4) throw std::invalid_argument("a gauss test field needs quadrilaterals");
	if (same_mesh && !test_gauss) throw std::invalid_argument("same mesh needs the gauss test field");
	std::vector<Elem> ttri;
	std::vector<QElem> tq, sq;
	std::vector<TestPt> tp;
	read_elems(txyz, tn, tel, tnt, tnpe, &ttri, &tq, &tp);
	if (!same_mesh) read_elems(sxyz, sn, sel, snt, 4, nullptr, &sq, nullptr);
	std::vector<QElem> const &trial = same_mesh ? tq : sq;
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	const int n_test_rows = test_gauss ? 4 * tnt : tnt;
	int nr = rows ? n_rows : n_test_rows;
	cplx alpha(alpha_re, alpha_im);
	if (n_threads <= 0) n_threads = orc_num_threads();
	std::string err;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
	for (int r = 0; r < nr; ++r) {
		try {
			int row = rows ? rows[r] : r;
			if (row < 0 || row >= n_test_rows) throw std::out_of_range("row out of range");
			V3 x, nv;
			int a = -1, ie = row;
			if (test_gauss) {
				ie = row / 4;
				a = row % 4;
				double xi, eta;
				gauss_node(a, xi, eta);
				x = q_x(tq[ie], xi, eta);
				nv = normalized(q_normal(tq[ie], xi, eta));
			} else {
				x = tp[ie].x;
				nv = tp[ie].nu;
			}
			for (int j = 0; j < snt; ++j) {
				C4 v;
				gauss_pair_row(family, nx, ny, k, p, x, nv, tp[ie].x, tp[ie].id, a, trial[j], same_mesh != 0, v);
				for (int b = 0; b < 4; ++b) {
					cplx z = alpha * v[b];
					double *o = out + 2 * ((int64_t)(4 * j + b) * ld + r);
					if (accumulate) { o[0] += z.real(); o[1] += z.imag(); }
					else { o[0] = z.real(); o[1] = z.imag(); }
				}
			}
		} catch (std::exception const &ex) {
			#pragma omp critical
			err = ex.what();
		}
	}
	if (!err.empty()) throw std::out_of_range(err);
	ORC_CATCH
}

template <class K>
static int gauss_elem_impl(int what, int family, int nx, int ny, K k, const double *coords12, const double *x0p,
	const double *nx0p, const double *xi0p, int o1, int o2, double *out8)
{
	ORC_TRY
	V3 X[4];
	for (int a = 0; a < 4; ++a) X[a] = {coords12[3 * a], coords12[3 * a + 1], coords12[3 * a + 2]};
	QElem e = make_qelem(X, 0);
	V3 x0 = x0p ? V3{x0p[0], x0p[1], x0p[2]} : e.c;
	V3 n0 = nx0p ? V3{nx0p[0], nx0p[1], nx0p[2]} : e.nu;
	C4 v;
	switch (what) {
	case 0: gauss_regular(family, nx, ny, k, x0, n0, e, o1, v); break;
	case 1: gauss_singular(family, nx, ny, k, e, o1, v); break;
	case 2: gauss_near_polar(family, nx, ny, k, e, x0, n0, o1, o2, v); break;
	case 3: gauss_near_telles(family, nx, ny, k, e, x0, n0, o1, v); break;
	case 4: gauss_guiggiani(family, k, e, xi0p[0], xi0p[1], o1, o2, v); break;
	case 5: gauss_duffy(family, nx, ny, k, e, xi0p[0], xi0p[1], o1, v); break;
	case 7: gauss_near(family, nx, ny, k, e, x0, n0, v); break;
	default: throw std::invalid_argument("unknown element integral");
	}
	for (int b = 0; b < 4; ++b) { out8[2 * b] = v[b].real(); out8[2 * b + 1] = v[b].imag(); }
	ORC_CATCH
}

extern "C" {

// Collocation with the quadrilateral gauss field of the trial mesh (4 dofs
// per element, column 4 j + b) and, test_gauss = 1, the dirac of the test
// mesh's gauss field (rows 4 i + a: node a of test element i) or, 0, the
// dirac of its constant field (rows = test elements, any element type: field
// points).  k_im: imaginary part of the wavenumber.
int orc_assemble_gauss(const double *txyz, int tn, const int *tel, int tnt, int tnpe,
	const double *sxyz, int sn, const int *sel, int snt, int same_mesh, int test_gauss,
	int family, int nx, int ny, double k_re, double k_im, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *out, int64_t ld, int accumulate, int n_threads)
{
	if (k_im == 0.0)
		return assemble_gauss_impl(txyz, tn, tel, tnt, tnpe, sxyz, sn, sel, snt, same_mesh, test_gauss, family, nx,
			ny, k_re, alpha_re, alpha_im, pol, rows, n_rows, out, ld, accumulate, n_threads);
	return assemble_gauss_impl(txyz, tn, tel, tnt, tnpe, sxyz, sn, sel, snt, same_mesh, test_gauss, family, nx, ny,
		cplx(k_re, k_im), alpha_re, alpha_im, pol, rows, n_rows, out, ld, accumulate, n_threads);
}

// Burton-Miller A and B on a quadrilateral mesh with the gauss field (as
// orc_assemble_bm; the identity is the unit matrix over the 4 n dofs: the
// shape functions are Lagrange polynomials on the collocation nodes)
int orc_assemble_bm_gauss(const double *xyz, int nn, const int *quad, int ne, double k_re, double k_im,
	double alpha_re, double alpha_im, const orc_policy *pol, const int *rows, int n_rows, double *A, double *B,
	int64_t ld, int n_threads)
{
	const int nd = 4 * ne;
	int nr = rows ? n_rows : nd;
	for (int j = 0; j < nd; ++j)
		for (int r = 0; r < nr; ++r) {
			double *a = A + 2 * ((int64_t)j * ld + r), *b = B + 2 * ((int64_t)j * ld + r);
			a[0] = a[1] = b[0] = b[1] = 0.0;
		}
	cplx al(alpha_re, alpha_im);
	int s;
	auto op = [&](int nx, int ny, cplx sc, double *out) {
		return orc_assemble_gauss(xyz, nn, quad, ne, 4, xyz, nn, quad, ne, 1, 1, 1, nx, ny, k_re, k_im, sc.real(),
			sc.imag(), pol, rows, n_rows, out, ld, 1, n_threads);
	};
	if ((s = op(0, 0, 1.0, B))) return s;
	if ((s = op(1, 0, al, B))) return s;
	orc_identity(nd, (al / 2.).real(), (al / 2.).imag(), rows, n_rows, B, ld);
	if ((s = op(1, 1, al, A))) return s;
	orc_identity(nd, -.5, 0.0, rows, n_rows, A, ld);
	if ((s = op(0, 1, 1.0, A))) return s;
	return 0;
}

// Element-level gauss-field integrals for the tests (4 complex results, one
// per trial dof): 0 regular (Gauss degree o1) at (x0, nx0), 1 singular at
// node o1, 2 polar near (o1, o2), 3 Telles near (degree o1), 4 Guiggiani at
// xi0 (o1, o2), 5 Duffy rule of degree o1 around xi0, 7 near dispatch.
int orc_gauss_elem(int what, int family, int nx, int ny, double k_re, double k_im, const double *coords12,
	const double *x0p, const double *nx0p, const double *xi0p, int o1, int o2, double *out8)
{
	if (k_im == 0.0) return gauss_elem_impl(what, family, nx, ny, k_re, coords12, x0p, nx0p, xi0p, o1, o2, out8);
	return gauss_elem_impl(what, family, nx, ny, cplx(k_re, k_im), coords12, x0p, nx0p, xi0p, o1, o2, out8);
}

// shape functions and nodes of the gauss field (N[4], dN/dxi[4], dN/deta[4]
// at (xi, eta); nodes[8])
int orc_gauss_shape(double xi, double eta, double *N12, double *nodes8)
{
	gauss_shape(xi, eta, N12);
	gauss_dshape(xi, eta, N12 + 4, N12 + 8);
	for (int a = 0; a < 4; ++a) gauss_node(a, nodes8[2 * a], nodes8[2 * a + 1]);
	return 0;
}

} // extern "C"


// Galerkin assembly on quadrilateral meshes, result(dof_a, dof_b) (+)= alpha *
// block (complex interleaved, column-major, ld >= test dofs).  linear = 1:
// dofs = mesh nodes (isoparametric_view), 0: elements (constant_view).
// family 2 = identity: the single integral of N_a N_b J with the Gauss degree
// of the shape orders plus the jacobian order 1 (nihu/core/single_integral.hpp:
// 65-140).
template <class K>
static int assemble_galerkin_quad_impl(const double *txyz, int tn, const int *tq, int tnt,
	const double *sxyz, int sn, const int *sq, int snt, int same_mesh, int linear,
	int family, int nx, int ny, K k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld, int accumulate,
	int n_threads)
{
	ORC_TRY
	auto test = make_qmesh(txyz, tn, tq, tnt);
	auto trial = same_mesh ? test : make_qmesh(sxyz, sn, sq, snt);
	if (same_mesh) { sn = tn; sq = tq; snt = tnt; }
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	const cplx alpha(alpha_re, alpha_im);
	const bool lt = linear_test(linear), ls = linear_trial(linear);
	const int ndt = lt ? 4 : 1, nds = ls ? 4 : 1;
	auto tdof = [&](int e, int a) { return lt ? tq[4 * e + a] : e; };
	auto sdof = [&](int e, int b) { return ls ? sq[4 * e + b] : e; };
	const int nrows = lt ? tn : tnt, ncols = ls ? sn : snt;
	if (!accumulate)
		for (int j = 0; j < ncols; ++j)
			for (int i = 0; i < nrows; ++i) { out[2 * ((int64_t)j * ld + i)] = 0; out[2 * ((int64_t)j * ld + i) + 1] = 0; }
	if (family == 2) {
		if (!same_mesh) return 0;
		auto const &qm = gauss_quad_store((lt ? 1 : 0) + (ls ? 1 : 0) + 1);
		for (int e = 0; e < tnt; ++e)
			for (int a = 0; a < ndt; ++a)
				for (int b = 0; b < nds; ++b) {
					double m = 0.0;
					for (auto const &q : qm) {
						double Nt[4], Ns[4];
						quad_shapes(q.xi, q.eta, Nt);
						quad_shapes(q.xi, q.eta, Ns);
						if (!lt) Nt[0] = 1.0;
						if (!ls) Ns[0] = 1.0;
						m += Nt[a] * Ns[b] * norm(q_normal(test[e], q.xi, q.eta)) * q.w;
					}
					const cplx v = alpha * m;
					double *o = out + 2 * ((int64_t)sdof(e, b) * ld + tdof(e, a));
					o[0] += v.real(); o[1] += v.imag();
				}
		return 0;
	}
	if (n_threads <= 0) n_threads = orc_num_threads();
	std::string err;
	int errcode = 0;
	const int n_loop = test_elems ? n_test_elems : tnt;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
	for (int ii = 0; ii < n_loop; ++ii) {
		const int i = test_elems ? test_elems[ii] : ii;
		try {
			for (int j = 0; j < snt; ++j) {
				cplx blk[4][4];
				galerkin_quad_block(family, nx, ny, k, p, test[i], trial[j], tq + 4 * i, sq + 4 * j, same_mesh != 0,
					linear, blk);
				for (int a = 0; a < ndt; ++a)
					for (int b = 0; b < nds; ++b) {
						const cplx v = alpha * blk[a][b];
						double *o = out + 2 * ((int64_t)sdof(j, b) * ld + tdof(i, a));
						#pragma omp atomic
						o[0] += v.real();
						#pragma omp atomic
						o[1] += v.imag();
					}
			}
		} catch (std::out_of_range const &ex) {
			#pragma omp critical
			{ err = ex.what(); errcode = 2; }
		} catch (std::exception const &ex) {
			#pragma omp critical
			{ err = ex.what(); errcode = 1; }
		}
	}
	if (errcode == 2) throw std::out_of_range(err);
	if (errcode == 1) throw std::runtime_error(err);
	ORC_CATCH
}

extern "C" {

int orc_assemble_galerkin_quad(const double *txyz, int tn, const int *tq, int tnt,
	const double *sxyz, int sn, const int *sq, int snt, int same_mesh, int linear,
	int family, int nx, int ny, double k_re, double k_im, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld, int accumulate,
	int n_threads)
{
	if (k_im == 0.0)
		return assemble_galerkin_quad_impl(txyz, tn, tq, tnt, sxyz, sn, sq, snt, same_mesh, linear, family, nx, ny,
			k_re, alpha_re, alpha_im, pol, test_elems, n_test_elems, out, ld, accumulate, n_threads);
	return assemble_galerkin_quad_impl(txyz, tn, tq, tnt, sxyz, sn, sq, snt, same_mesh, linear, family, nx, ny,
		cplx(k_re, k_im), alpha_re, alpha_im, pol, test_elems, n_test_elems, out, ld, accumulate, n_threads);
}

// node overlap of two quadrilaterals as NiHu's scan finds it
int orc_overlap_quads(const int *a, const int *b, int *ind1, int *ind2)
{
	Overlap o = get_overlapping_quads(a, b);
	*ind1 = o.ind1;
	*ind2 = o.ind2;
	return o.num;
}

// canonical quad x quad singular rule of a match dimension and Gauss degree
int orc_quad_singular_rule(int match, int degree, double *pts, int cap, int *n_out)
{
	ORC_TRY
	if (match < 0 || match > 2) throw std::invalid_argument("match must be 0, 1 or 2");
	const std::vector<SPt> r = make_quad_singular_rule(match, degree);
	if ((int)r.size() > cap) throw std::invalid_argument("capacity");
	for (size_t q = 0; q < r.size(); ++q) {
		pts[5 * q] = r[q].txi; pts[5 * q + 1] = r[q].teta; pts[5 * q + 2] = r[q].sxi;
		pts[5 * q + 3] = r[q].seta; pts[5 * q + 4] = r[q].w;
	}
	*n_out = (int)r.size();
	ORC_CATCH
}

// one Galerkin block of a quadrilateral pair given by corner coordinates
// (12 each) and node ids; same_elem: the pair is one element
int orc_galerkin_quad_block(int family, int nx, int ny, double k_re, double k_im, const orc_policy *pol,
	const double *tc12, const int *tn4, const double *sc12, const int *sn4, int same_mesh, int same_elem,
	int linear, double *blk32)
{
	ORC_TRY
	V3 X[4], Y[4];
	for (int a = 0; a < 4; ++a) {
		X[a] = {tc12[3 * a], tc12[3 * a + 1], tc12[3 * a + 2]};
		Y[a] = {sc12[3 * a], sc12[3 * a + 1], sc12[3 * a + 2]};
	}
	QElem te = make_qelem(X, 0), se = make_qelem(Y, same_elem ? 0 : 1);
	Policy p = {pol->rule, pol->fixed_degree, pol->near_mode, pol->near_limit};
	cplx blk[4][4];
	if (k_im == 0.0) galerkin_quad_block(family, nx, ny, k_re, p, te, se, tn4, sn4, same_mesh != 0, linear, blk);
	else galerkin_quad_block(family, nx, ny, cplx(k_re, k_im), p, te, se, tn4, sn4, same_mesh != 0, linear, blk);
	for (int a = 0; a < 4; ++a)
		for (int b = 0; b < 4; ++b) {
			blk32[2 * (4 * a + b)] = blk[a][b].real();
			blk32[2 * (4 * a + b) + 1] = blk[a][b].imag();
		}
	ORC_CATCH
}

} // extern "C"
