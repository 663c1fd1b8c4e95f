// Known-answer tests of the C++ front end, written the way NiHu's own
// tutorial tests are (nihu/tutorial/helmholtz_3d_transparent_standalone.cpp:17-92,
// collocation and Galerkin; nihu/test/stallone_unit/helmholtz_3d_pulsating_sphere_test.cpp:28-73,
// Galerkin with constant fields), on the GPU through the C ABI, plus the same
// answers with isoparametric (nodal) fields.
//
//   test_facade <mesh.off> [quadrilateral mesh.off [mixed mesh.off]]
// Prints the errors and exits non-zero if one is above its bound.  With a
// quadrilateral mesh, also the collocation answers and the Burton-Miller
// composition (hypersingular diagonal by Guiggiani's method) on it; with a
// mesh of both element types (read with both tags), the same again.

#include "nihu_amd/nihu.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

using nihu_amd::cMatrix;
typedef std::complex<double> cplx;

// dense LU with partial pivoting: solves A x = b in place (A overwritten)
static std::vector<cplx> solve(cMatrix A, std::vector<cplx> b)
{
	const int64_t n = A.rows();
	for (int64_t k = 0; k < n; ++k) {
		int64_t p = k;
		for (int64_t i = k + 1; i < n; ++i)
			if (std::abs(A(i, k)) > std::abs(A(p, k))) p = i;
		if (p != k) {
			for (int64_t j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
			std::swap(b[k], b[p]);
		}
		for (int64_t i = k + 1; i < n; ++i) {
			const cplx f = A(i, k) / A(k, k);
			if (f == cplx(0)) continue;
			for (int64_t j = k; j < n; ++j) A(i, j) -= f * A(k, j);
			b[i] -= f * b[k];
		}
	}
	std::vector<cplx> x(n);
	for (int64_t i = n - 1; i >= 0; --i) {
		cplx s = b[i];
		for (int64_t j = i + 1; j < n; ++j) s -= A(i, j) * x[j];
		x[i] = s / A(i, i);
	}
	return x;
}

struct centre_normal {
	double c[3], n[3];
};

// sample points of a space's dofs: element centres and normals, or for
// nodal dofs the nodes with the sphere's radial normal
static std::vector<centre_normal> geometry(nihu_amd::isoparametric_space const &s)
{
	auto const &X = s.get_mesh().nodes();
	std::vector<centre_normal> g(s.get_num_dofs());
	for (size_t v = 0; v < g.size(); ++v) {
		double r = std::sqrt(X[3 * v] * X[3 * v] + X[3 * v + 1] * X[3 * v + 1] + X[3 * v + 2] * X[3 * v + 2]);
		for (int d = 0; d < 3; ++d) { g[v].c[d] = X[3 * v + d]; g[v].n[d] = X[3 * v + d] / r; }
	}
	return g;
}

template <class Space>
static std::vector<centre_normal> geometry(Space const &s)
{
	nihu_amd::mesh const &m = s.get_mesh();
	std::vector<centre_normal> g(m.num_elements());
	auto const &X = m.nodes();
	auto const &T = m.elements();
	for (int e = 0; e < m.num_elements(); ++e) {
		// node ids of element e: its row of the element table when the mesh
		// holds both types ([type id, n0, ..., n3])
		const int npe = m.nodes_per_element() ? m.nodes_per_element()
			: T[5 * e] == nihu_amd::tria_1_id ? 3 : 4;
		const int32_t *ids = m.nodes_per_element() ? &T[(size_t)npe * e] : &T[5 * (size_t)e + 1];
		const double *p[4];
		for (int a = 0; a < npe; ++a) p[a] = &X[3 * ids[a]];
		double u[3], v[3];
		for (int d = 0; d < 3; ++d) {
			if (npe == 4) {
				// bilinear quadrilateral: x(0, 0) and dx/dxi x dx/deta there
				g[e].c[d] = (p[0][d] + p[1][d] + p[2][d] + p[3][d]) / 4.0;
				u[d] = (p[1][d] + p[2][d] - p[0][d] - p[3][d]) / 4.0;
				v[d] = (p[2][d] + p[3][d] - p[0][d] - p[1][d]) / 4.0;
			} else {
				g[e].c[d] = (p[0][d] + p[1][d] + p[2][d]) / 3.0;
				u[d] = p[1][d] - p[0][d];
				v[d] = p[2][d] - p[0][d];
			}
		}
		double nv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
		double l = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
		for (int d = 0; d < 3; ++d) g[e].n[d] = nv[d] / l;
	}
	return g;
}

// transparent test: point sources inside the surface give exact p0, q0;
// solve (M - I/2) p = L q0 and compare with p0
template <class TestSpace, class TrialSpace>
static double tester(TestSpace const &test_space, TrialSpace const &trial_space)
{
	const double pi = 3.141592653589793238462643383279502884;
	double wn = 1.;
	auto I_op = nihu_amd::identity_integral_operator();
	auto L_op = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(wn));
	auto M_op = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLP_kernel<double>(wn));

	size_t N = test_space.get_num_dofs();
	std::vector<cplx> q0(N, 0.0), p0(N, 0.0);
	const double x0[2][3] = {{.2, .3, .1}, {.3, .2, .2}};
	const double A[2] = {1.0, -1.0};
	auto g = geometry(trial_space);
	for (size_t k = 0; k < N; ++k) {
		for (int s = 0; s < 2; ++s) {
			double rv[3], r = 0, rdn = 0;
			for (int d = 0; d < 3; ++d) { rv[d] = g[k].c[d] - x0[s][d]; r += rv[d] * rv[d]; }
			r = std::sqrt(r);
			for (int d = 0; d < 3; ++d) rdn += rv[d] * g[k].n[d] / r;
			cplx ikr(0., wn * r);
			p0[k] += A[s] * std::exp(-ikr) / r / (4. * pi);
			q0[k] += A[s] * -(std::exp(-ikr) / r / r) * rdn / (4. * pi) * (1. + ikr);
		}
	}

	cMatrix L(N, N), M(N, N), I(N, N);
	I << test_space * I_op[trial_space];
	L << test_space * L_op[trial_space];
	M << test_space * M_op[trial_space];

	cMatrix S(N, N);
	std::vector<cplx> rhs(N, 0.0);
	for (size_t i = 0; i < N; ++i)
		for (size_t j = 0; j < N; ++j) {
			S(i, j) = M(i, j) - .5 * I(i, j);
			rhs[i] += L(i, j) * q0[j];
		}
	auto p = solve(S, rhs);
	double e = 0, nrm = 0;
	for (size_t k = 0; k < N; ++k) { e += std::norm(p[k] - p0[k]); nrm += std::norm(p0[k]); }
	return std::sqrt(e / nrm);
}

// pulsating sphere: p = -r/(1+ikr) q for a unit normal velocity; the wave
// number type W is double or std::complex<double> (NiHu's WaveNumber)
template <class TestSpace, class TrialSpace, class W = double>
static double pulsating(TestSpace const &test_space, TrialSpace const &trial_space, W k = W(1.0))
{
	const size_t N = trial_space.get_num_dofs();
	cMatrix L(N, N), M(N, N), I(N, N);
	auto I_op = nihu_amd::identity_integral_operator();
	auto L_op = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<W>(k));
	auto M_op = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLP_kernel<W>(k));
	I << test_space * I_op[trial_space];
	L << test_space * L_op[trial_space];
	M << test_space * M_op[trial_space];
	cMatrix S(N, N);
	std::vector<cplx> rhs(N, 0.0);
	for (size_t i = 0; i < N; ++i)
		for (size_t j = 0; j < N; ++j) {
			S(i, j) = M(i, j) - .5 * I(i, j);
			rhs[i] += L(i, j);
		}
	auto p = solve(S, rhs);
	const cplx panal = -1.0 / (1.0 + cplx(0, 1) * cplx(k));
	double e = 0;
	for (auto const &v : p) e += std::norm(v - panal);
	return std::sqrt(e / N) / std::abs(panal);
}

// Burton-Miller composition as nihu/applications/helmholtz/helmholtz_3d_coll.cpp:63-96
// writes it, against the fused nb_assemble_bm of the same matrices
static double burton_miller(nihu_amd::mesh const &m)
{
	auto trial_space = nihu_amd::constant_view(m);
	auto test_space = nihu_amd::dirac(trial_space);
	const size_t N = trial_space.get_num_dofs();
	double k = 4.0;
	std::complex<double> alpha(0., -1. / k);
	auto Gop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(k));
	auto Hop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLP_kernel<double>(k));
	auto Htop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLPt_kernel<double>(k));
	auto Dop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_HSP_kernel<double>(k));
	auto Iop = nihu_amd::identity_integral_operator();
	cMatrix Gs(N, N), Hs(N, N);
	Gs << test_space * Gop[trial_space];
	Gs << test_space * (alpha * Htop)[trial_space];
	Gs << test_space * (alpha / 2. * Iop)[trial_space];
	Hs << test_space * (alpha * Dop)[trial_space];
	Hs << test_space * (-.5 * Iop)[trial_space];
	Hs << test_space * Hop[trial_space];
	cMatrix A(N, N), B(N, N);
	nb_out oa = {A.data(), (int64_t)N, (int64_t)N, (int64_t)N, NB_C128, NB_HOST, 0};
	nb_out ob = {B.data(), (int64_t)N, (int64_t)N, (int64_t)N, NB_C128, NB_HOST, 0};
	nihu_amd::check(nb_assemble_bm(m.handle(), k, nb_scalar{alpha.real(), alpha.imag()}, &nihu_amd::default_policy(),
		&oa, &ob, nullptr));
	double e = 0, nrm = 0;
	for (size_t i = 0; i < N; ++i)
		for (size_t j = 0; j < N; ++j) {
			e += std::norm(A(i, j) - Hs(i, j)) + std::norm(B(i, j) - Gs(i, j));
			nrm += std::norm(Hs(i, j)) + std::norm(Gs(i, j));
		}
	return std::sqrt(e / nrm);
}

// the same composition over the quadrilateral gauss field against
// nb_assemble_bm_gauss
static double burton_miller_gauss(nihu_amd::gauss_space const &trial_space)
{
	auto test_space = nihu_amd::dirac(trial_space);
	const size_t N = trial_space.get_num_dofs();
	double k = 4.0;
	std::complex<double> alpha(0., -1. / k);
	auto Gop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(k));
	auto Hop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLP_kernel<double>(k));
	auto Htop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_DLPt_kernel<double>(k));
	auto Dop = nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_HSP_kernel<double>(k));
	auto Iop = nihu_amd::identity_integral_operator();
	cMatrix Gs(N, N), Hs(N, N);
	Gs << test_space * Gop[trial_space];
	Gs << test_space * (alpha * Htop)[trial_space];
	Gs << test_space * (alpha / 2. * Iop)[trial_space];
	Hs << test_space * (alpha * Dop)[trial_space];
	Hs << test_space * (-.5 * Iop)[trial_space];
	Hs << test_space * Hop[trial_space];
	cMatrix A(N, N), B(N, N);
	nb_out oa = {A.data(), (int64_t)N, (int64_t)N, (int64_t)N, NB_C128, NB_HOST, 0};
	nb_out ob = {B.data(), (int64_t)N, (int64_t)N, (int64_t)N, NB_C128, NB_HOST, 0};
	nihu_amd::check(nb_assemble_bm_gauss(trial_space.get_mesh().handle(), nb_scalar{k, 0.0},
		nb_scalar{alpha.real(), alpha.imag()}, &nihu_amd::default_policy(), &oa, &ob, nullptr));
	double e = 0, nrm = 0;
	for (size_t i = 0; i < N; ++i)
		for (size_t j = 0; j < N; ++j) {
			e += std::norm(A(i, j) - Hs(i, j)) + std::norm(B(i, j) - Gs(i, j));
			nrm += std::norm(Hs(i, j)) + std::norm(Gs(i, j));
		}
	return std::sqrt(e / nrm);
}

int main(int argc, char *argv[])
{
	if (argc < 2) {
		std::fprintf(stderr, "Usage: %s OFFname\n", argv[0]);
		return 1;
	}
	auto mesh = nihu_amd::read_off_mesh(argv[1], nihu_amd::tria_1_tag());
	auto const &trial_space = nihu_amd::constant_view(mesh);
	int bad = 0;

	double coll_error = tester(nihu_amd::dirac(trial_space), trial_space);
	std::printf("Collocation transparent error: %g\n", coll_error);
	bad |= !(coll_error < 0.05);

	double puls = pulsating(nihu_amd::dirac(trial_space), trial_space);
	std::printf("Pulsating sphere relative error (collocation): %g\n", puls);
	bad |= !(puls < 0.05);

	double puls_ck = pulsating(nihu_amd::dirac(trial_space), trial_space, cplx(1.0, -0.3));
	std::printf("Pulsating sphere relative error (collocation, complex wave number): %g\n", puls_ck);
	bad |= !(puls_ck < 0.05);

	double gal_error = tester(trial_space, trial_space);
	std::printf("Galerkin transparent error: %g\n", gal_error);
	bad |= !(gal_error < 0.05);

	double gal_puls = pulsating(trial_space, trial_space);
	std::printf("Pulsating sphere relative error (Galerkin): %g\n", gal_puls);
	bad |= !(gal_puls < 0.05);

	auto const &iso_space = nihu_amd::isoparametric_view(mesh);
	double iso_error = tester(iso_space, iso_space);
	std::printf("Galerkin isoparametric transparent error: %g\n", iso_error);
	bad |= !(iso_error < 0.05);

	double iso_puls = pulsating(iso_space, iso_space);
	std::printf("Pulsating sphere relative error (Galerkin isoparametric): %g\n", iso_puls);
	bad |= !(iso_puls < 0.05);

	// a hypersingular Galerkin operator on one mesh has no valid singular rule
	try {
		nihu_amd::cMatrix D(iso_space.get_num_dofs(), iso_space.get_num_dofs());
		D << iso_space * nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_HSP_kernel<double>(1.0))[iso_space];
		std::printf("no exception for Galerkin HSP\n");
		bad = 1;
	} catch (std::runtime_error const &ex) {
		std::printf("runtime_error as expected: %s\n", ex.what());
	}

	// mixed field types: constant test x isoparametric trial and the reverse.
	// Gauss's identity on the closed surface: the double layer's row sums are
	// -1/2 times the mass matrix's; the single layer's two mixed matrices are
	// transposes of each other (a symmetric kernel) to quadrature error
	{
		const int ne = trial_space.get_num_dofs(), nn = iso_space.get_num_dofs();
		nihu_amd::cMatrix D(ne, nn), I(ne, nn), CL(ne, nn), LC(nn, ne);
		D << trial_space * nihu_amd::create_integral_operator(nihu_amd::laplace_3d_DLP_kernel())[iso_space];
		I << trial_space * nihu_amd::identity_integral_operator()[iso_space];
		CL << trial_space * nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(1.0))[iso_space];
		LC << iso_space * nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(1.0))[trial_space];
		double gi = 0.0, im = 0.0, tr = 0.0, cm = 0.0;
		for (int e = 0; e < ne; ++e) {
			cplx sd = 0.0, si = 0.0;
			for (int b = 0; b < nn; ++b) {
				sd += D(e, b);
				si += I(e, b);
				tr = std::max(tr, std::abs(CL(e, b) - LC(b, e)));
				cm = std::max(cm, std::abs(CL(e, b)));
			}
			gi = std::max(gi, std::abs(sd + 0.5 * si));
			im = std::max(im, std::abs(si));
		}
		std::printf("Mixed-field Galerkin: Gauss identity %g, transpose symmetry %g\n", gi / im, tr / cm);
		bad |= !(gi < 2e-3 * im && tr < 1e-3 * cm);
	}

	double bm = burton_miller(mesh);
	std::printf("Burton-Miller composition vs fused assembly: %g\n", bm);
	bad |= !(bm < 1e-6);

	if (argc > 2) {
		auto qmesh = nihu_amd::read_off_mesh(argv[2], nihu_amd::quad_1_tag());
		auto const &qspace = nihu_amd::constant_view(qmesh);
		double qt = tester(nihu_amd::dirac(qspace), qspace);
		std::printf("Quadrilateral collocation transparent error: %g\n", qt);
		bad |= !(qt < 0.05);
		double qp = pulsating(nihu_amd::dirac(qspace), qspace);
		std::printf("Pulsating sphere relative error (quadrilateral collocation): %g\n", qp);
		bad |= !(qp < 0.05);
		double qbm = burton_miller(qmesh);
		std::printf("Quadrilateral Burton-Miller composition vs fused assembly: %g\n", qbm);
		bad |= !(qbm < 1e-6);
		// the gauss field as helmholtz_3d_coll.cpp:228-240 builds it
		std::vector<unsigned> fields;
		for (int e = 0; e < qmesh.num_elements(); ++e) {
			fields.push_back(nihu_amd::quad_1_gauss_field_id);
			for (int c = 0; c < 4; ++c) fields.push_back((unsigned)qmesh.elements()[4 * e + c]);
			for (int c = 0; c < 4; ++c) fields.push_back((unsigned)(4 * e + c));
		}
		auto gspace = nihu_amd::create_function_space(qmesh.nodes(), fields, nihu_amd::quad_1_gauss_field_tag());
		double gp = pulsating(nihu_amd::dirac(gspace), gspace);
		std::printf("Pulsating sphere relative error (quadrilateral gauss field): %g\n", gp);
		bad |= !(gp < 0.05);
		double gbm = burton_miller_gauss(gspace);
		std::printf("Gauss-field Burton-Miller composition vs fused assembly: %g\n", gbm);
		bad |= !(gbm < 1e-6);
		// another dof numbering is refused
		std::swap(fields[5], fields[6]);
		try {
			nihu_amd::create_function_space(qmesh.nodes(), fields, nihu_amd::quad_1_gauss_field_tag());
			std::printf("no exception for a permuted gauss dof numbering\n");
			bad = 1;
		} catch (std::invalid_argument const &ex) {
			std::printf("gauss dof numbering: invalid_argument as expected: %s\n", ex.what());
		}
		// the Galerkin formalism on quadrilaterals, constant and isoparametric
		// (bilinear) fields
		double gqc = pulsating(qspace, qspace);
		auto const &qiso = nihu_amd::isoparametric_view(qmesh);
		double gql = pulsating(qiso, qiso);
		std::printf("Pulsating sphere relative error (quadrilateral Galerkin): constant %g, isoparametric %g\n",
			gqc, gql);
		bad |= !(gqc < 0.05 && gql < 0.05);
	}

	if (argc > 3) {
		// triangles and quadrilaterals in one mesh, read as NiHu's EAA
		// benchmark drivers read theirs
		auto mmesh = nihu_amd::read_off_mesh(argv[3], nihu_amd::quad_1_tag(), nihu_amd::tria_1_tag());
		auto const &mspace = nihu_amd::constant_view(mmesh);
		std::printf("mixed mesh: %d elements, nodes_per_element %d\n", mmesh.num_elements(),
			mmesh.nodes_per_element());
		bad |= mmesh.nodes_per_element() != 0;
		double mt = tester(nihu_amd::dirac(mspace), mspace);
		std::printf("Mixed-mesh collocation transparent error: %g\n", mt);
		bad |= !(mt < 0.05);
		double mp = pulsating(nihu_amd::dirac(mspace), mspace);
		std::printf("Pulsating sphere relative error (mixed-mesh collocation): %g\n", mp);
		bad |= !(mp < 0.05);
		double mbm = burton_miller(mmesh);
		std::printf("Mixed-mesh Burton-Miller composition vs fused assembly: %g\n", mbm);
		bad |= !(mbm < 1e-6);
		try {
			nihu_amd::cMatrix X(mspace.get_num_dofs(), mspace.get_num_dofs());
			X << mspace * nihu_amd::create_integral_operator(nihu_amd::helmholtz_3d_SLP_kernel<double>(1.0))[mspace];
			std::printf("no exception for mixed-mesh Galerkin\n");
			bad = 1;
		} catch (std::logic_error const &ex) {
			std::printf("mixed-mesh Galerkin: logic_error as expected: %s\n", ex.what());
		}
	}

	// error behaviour: a Dunavant degree above 14 throws std::out_of_range
	try {
		nihu_amd::set_policy(nb_policy{NB_RULE_FIXED, 15, NB_NEAR_MEAN_EDGE, 1.5});
		cMatrix X(trial_space.get_num_dofs(), trial_space.get_num_dofs());
		X << nihu_amd::dirac(trial_space) * nihu_amd::create_integral_operator(
			nihu_amd::helmholtz_3d_SLP_kernel<double>(1.0))[trial_space];
		std::printf("no exception for degree 15\n");
		bad = 1;
	} catch (std::out_of_range const &ex) {
		std::printf("out_of_range as expected: %s\n", ex.what());
	}
	return bad;
}
