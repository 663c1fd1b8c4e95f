// Host code under the sanitizers: the engine's host geometry, quadrature
// and Galerkin layout routines (nihu_amd/csrc/nbem_geom.cpp, nbem_layout.cpp)
// and, in the address / undefined-behaviour build, the FP64 oracle with
// OpenMP threads.  Every routine runs once on one thread for a reference,
// then four threads run all of them at once on the same read-only inputs;
// each thread's outputs must equal the reference bit for bit.  Built and run
// by tests/test_sanitizers.py (-fsanitize=address,undefined and, without the
// oracle, -fsanitize=thread); a report from a sanitizer fails the run.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

extern "C" {
int nbg_dunavant(int degree, double *xi, double *eta, double *w, int cap);
void nbg_gauss_legendre(int n, double *x, double *w);
int nbg_geometry(const double *xyz, int n_nodes, const int32_t *tri, int n, double *c, double *nu, double *J, double *R,
	double *diam, double *X);
double nbg_threshold(double scale, double limit);
int nbg_quad_table(int degree, int n, const double *c, const double *J, const double *X, float *out);
int nbg_geometry_quad(const double *xyz, int n_nodes, const int32_t *quad, int n, double *c, double *nu, double *J,
	double *R, double *diam, double *X);
int nbg_quad4_table(int N, int n, const double *c, const double *X, float *out);
int nbg_quad_duffy(int ng, double x0, double y0, double *out);
int nbg_singular_rule(int match, double *out, int cap);
int64_t nbg_singular_pairs(int n_nodes, int n, const int32_t *tri, int32_t *pe, int32_t *pf, int32_t *info);
int nbg_galerkin_layout(int n, const int32_t *tri, const double *c, int chunk, int32_t *order_out, int32_t *elem,
	uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin);
int nbg_quad_singular_rule(int match, double *out, int cap);
int64_t nbg_singular_pairs_n(int n_nodes, int n, const int32_t *el, int npe, int32_t *pe, int32_t *pf, int32_t *info);
int nbg_galerkin_layout_n(int n, const int32_t *tri, int npe, int cs, const double *c, int chunk,
	int32_t *order_out, int32_t *elem, uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin);
int nbg_quad_gal_table(int N, int n, const double *c, const double *X, float *out);
#ifdef WITH_ORACLE
typedef struct { int rule; int fixed_degree; int near_mode; double near_limit; } orc_policy;
int orc_assemble_bm(const double *xyz, int nn, const int *tri, int nt, double k, double alpha_re, double alpha_im,
	const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int64_t ld, int n_threads);
int orc_assemble_galerkin(const double *txyz, int tn, const int *ttri, int tnt, const double *sxyz, int sn,
	const int *stri, int snt, int same_mesh, int linear, int family, int nx, int ny, double k, double alpha_re,
	double alpha_im, const orc_policy *pol, const int *test_elems, int n_test_elems, double *out, int64_t ld,
	int accumulate, int n_threads);
int orc_assemble_bm_elems(const double *xyz, int nn, const int *el, int ne, int npe, double k, double alpha_re,
	double alpha_im, const orc_policy *pol, const int *rows, int n_rows, double *A, double *B, int64_t ld,
	int n_threads);
#endif
}

namespace {

// icosahedron subdivided `level` times, vertices on the unit sphere, radial
// perturbation from a fixed sequence (no library generator: reproducible)
void icosphere(int level, std::vector<double> &xyz, std::vector<int32_t> &tri)
{
	const double t = (1.0 + std::sqrt(5.0)) / 2.0;
	const double v0[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t}, {0, -1, -t},
		{0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
	const int f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
		{11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9}, {4, 9, 5},
		{2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
	xyz.clear();
	for (auto &p : v0) {
		const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
		xyz.insert(xyz.end(), {p[0] / r, p[1] / r, p[2] / r});
	}
	std::vector<int32_t> f(&f0[0][0], &f0[0][0] + 60);
	for (int l = 0; l < level; ++l) {
		std::map<std::pair<int, int>, int> mid;
		auto m = [&](int a, int b) {
			auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
			auto it = mid.find(key);
			if (it != mid.end()) return it->second;
			double p[3];
			for (int d = 0; d < 3; ++d) p[d] = 0.5 * (xyz[3 * a + d] + xyz[3 * b + d]);
			const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
			const int id = (int)(xyz.size() / 3);
			xyz.insert(xyz.end(), {p[0] / r, p[1] / r, p[2] / r});
			mid[key] = id;
			return id;
		};
		std::vector<int32_t> g;
		for (size_t e = 0; e < f.size(); e += 3) {
			const int a = f[e], b = f[e + 1], c = f[e + 2];
			const int ab = m(a, b), bc = m(b, c), ca = m(c, a);
			g.insert(g.end(), {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca});
		}
		f.swap(g);
	}
	for (size_t v = 0; v < xyz.size() / 3; ++v) {
		const double s = 1.0 + 0.01 * std::sin(1.7 * (double)v);
		for (int d = 0; d < 3; ++d) xyz[3 * v + d] *= s;
	}
	tri = f;
}

// all outputs of the host routines, flattened to bytes for comparison
std::vector<unsigned char> run_host(const std::vector<double> &xyz, const std::vector<int32_t> &tri,
	const std::vector<double> &qxyz, const std::vector<int32_t> &quad)
{
	std::vector<unsigned char> out;
	auto put = [&](const void *p, size_t n) {
		const unsigned char *b = (const unsigned char *)p;
		out.insert(out.end(), b, b + n);
	};
	const int nn = (int)(xyz.size() / 3), n = (int)(tri.size() / 3);
	for (int d = 0; d <= 14; ++d) {
		double xi[64], eta[64], w[64];
		const int q = nbg_dunavant(d, xi, eta, w, 64);
		put(&q, sizeof q);
		put(xi, q * sizeof(double));
		put(w, q * sizeof(double));
	}
	for (int g : {6, 21, 51}) {
		std::vector<double> x(g), w(g);
		nbg_gauss_legendre(g, x.data(), w.data());
		put(x.data(), g * sizeof(double));
	}
	std::vector<double> c(3 * n), nu(3 * n), J(n), R(n), diam(n), X(9 * n);
	if (nbg_geometry(xyz.data(), nn, tri.data(), n, c.data(), nu.data(), J.data(), R.data(), diam.data(), X.data()))
		std::printf("nbg_geometry failed\n");
	put(c.data(), c.size() * 8);
	put(nu.data(), nu.size() * 8);
	std::vector<double> thr(n);
	for (int e = 0; e < n; ++e) thr[e] = nbg_threshold(R[e], 1.5) + nbg_threshold(diam[e], 2.0);
	put(thr.data(), n * 8);
	std::vector<float> qt((size_t)n * 7 * 8);
	nbg_quad_table(5, n, c.data(), J.data(), X.data(), qt.data());
	put(qt.data(), qt.size() * 4);
	for (int m = 0; m < 3; ++m) {
		std::vector<double> r(5 * 1536);
		const int k = nbg_singular_rule(m, r.data(), 1536);
		put(r.data(), (size_t)5 * k * 8);
	}
	const int64_t np = nbg_singular_pairs(nn, n, tri.data(), nullptr, nullptr, nullptr);
	std::vector<int32_t> pe(np), pf(np), info(np);
	nbg_singular_pairs(nn, n, tri.data(), pe.data(), pf.data(), info.data());
	put(pe.data(), np * 4);
	put(info.data(), np * 4);
	const int tiles = nbg_galerkin_layout(n, tri.data(), c.data(), 64, nullptr, nullptr, nullptr, nullptr, nullptr,
		nullptr);
	std::vector<int32_t> order(n), elem((size_t)tiles * 64), rows((size_t)tiles * 64), col(4 * (size_t)n),
		fin((size_t)((n + 63) / 64) * 8);
	std::vector<uint32_t> contrib(2 * (size_t)tiles * 64);
	nbg_galerkin_layout(n, tri.data(), c.data(), 64, order.data(), elem.data(), contrib.data(), rows.data(),
		col.data(), fin.data());
	put(elem.data(), elem.size() * 4);
	put(col.data(), col.size() * 4);
	const int nq = (int)(quad.size() / 4), nqn = (int)(qxyz.size() / 3);
	std::vector<double> qc(3 * nq), qnu(3 * nq), qJ(nq), qR(nq), qd(nq), qX(12 * nq);
	nbg_geometry_quad(qxyz.data(), nqn, quad.data(), nq, qc.data(), qnu.data(), qJ.data(), qR.data(), qd.data(),
		qX.data());
	put(qc.data(), qc.size() * 8);
	std::vector<float> q4((size_t)nq * 9 * 8);
	nbg_quad4_table(3, nq, qc.data(), qX.data(), q4.data());
	put(q4.data(), q4.size() * 4);
	std::vector<double> du(3 * 64);
	const int nd = nbg_quad_duffy(4, 0.1, -0.2, du.data());
	put(du.data(), (size_t)3 * nd * 8);
	// Galerkin on quadrilaterals: rules, node-sharing pairs, point rows, and
	// the layouts of the node dofs and of the element dofs (every corner the
	// element itself)
	for (int m = 0; m < 3; ++m) {
		std::vector<double> r(5 * 2048);
		const int k = nbg_quad_singular_rule(m, r.data(), 2048);
		put(r.data(), (size_t)5 * k * 8);
	}
	const int64_t nqp = nbg_singular_pairs_n(nqn, nq, quad.data(), 4, nullptr, nullptr, nullptr);
	std::vector<int32_t> qpe(nqp), qpf(nqp), qinfo(nqp);
	nbg_singular_pairs_n(nqn, nq, quad.data(), 4, qpe.data(), qpf.data(), qinfo.data());
	put(qinfo.data(), nqp * 4);
	std::vector<float> gq((size_t)nq * 4 * 16);
	nbg_quad_gal_table(2, nq, qc.data(), qX.data(), gq.data());
	put(gq.data(), gq.size() * 4);
	std::vector<int32_t> edof(4 * (size_t)nq);
	for (int e = 0; e < nq; ++e)
		for (int a = 0; a < 4; ++a) edof[4 * e + a] = e;
	const std::vector<int32_t> *maps[2] = {&quad, &edof};
	for (const std::vector<int32_t> *dm : maps) {
		const int qt_tiles = nbg_galerkin_layout_n(nq, dm->data(), 4, 6, qc.data(), 2, nullptr, nullptr, nullptr,
			nullptr, nullptr, nullptr);
		std::vector<int32_t> qo(nq), qe((size_t)qt_tiles * 64), qr((size_t)qt_tiles * 128), qcol(5 * (size_t)nq),
			qfin((size_t)((nq + 1) / 2) * 6);
		std::vector<uint32_t> qcon((size_t)qt_tiles * 256);
		nbg_galerkin_layout_n(nq, dm->data(), 4, 6, qc.data(), 2, qo.data(), qe.data(), qcon.data(), qr.data(),
			qcol.data(), qfin.data());
		put(qr.data(), qr.size() * 4);
		put(qcol.data(), qcol.size() * 4);
	}
	return out;
}

} // namespace

int main()
{
	std::vector<double> xyz;
	std::vector<int32_t> tri;
	icosphere(2, xyz, tri);
	// a 2 x 2 patch of warped quadrilaterals
	std::vector<double> qxyz;
	for (int j = 0; j < 3; ++j)
		for (int i = 0; i < 3; ++i) qxyz.insert(qxyz.end(), {(double)i, (double)j, 0.1 * i * j});
	std::vector<int32_t> quad = {0, 1, 4, 3, 1, 2, 5, 4, 3, 4, 7, 6, 4, 5, 8, 7};
	const std::vector<unsigned char> ref = run_host(xyz, tri, qxyz, quad);
	std::vector<std::vector<unsigned char> > got(4);
	std::vector<std::thread> th;
	for (int t = 0; t < 4; ++t) th.emplace_back([&, t]() { got[t] = run_host(xyz, tri, qxyz, quad); });
	for (auto &t : th) t.join();
	int bad = 0;
	for (int t = 0; t < 4; ++t)
		if (got[t] != ref) { std::printf("thread %d: host routines differ from the single-thread run\n", t); ++bad; }
#ifdef WITH_ORACLE
	{
		const int n = (int)(tri.size() / 3), nn = (int)(xyz.size() / 3);
		const orc_policy pol = {0, 0, 0, 1.5};
		std::vector<double> A1(2 * (size_t)n * n), B1(A1.size()), A4(A1.size()), B4(A1.size());
		orc_assemble_bm(xyz.data(), nn, tri.data(), n, 4.0, 0.0, -0.25, &pol, nullptr, n, A1.data(), B1.data(), n, 1);
		orc_assemble_bm(xyz.data(), nn, tri.data(), n, 4.0, 0.0, -0.25, &pol, nullptr, n, A4.data(), B4.data(), n, 4);
		if (A1 != A4 || B1 != B4) { std::printf("oracle: 4 OpenMP threads differ from 1\n"); ++bad; }
		std::vector<double> L(2 * (size_t)nn * nn);
		orc_assemble_galerkin(xyz.data(), nn, tri.data(), n, xyz.data(), nn, tri.data(), n, 1, 1, 1, 0, 0, 3.0, 1.0,
			0.0, &pol, nullptr, 0, L.data(), nn, 0, 4);
		const int nq = (int)(quad.size() / 4), nqn = (int)(qxyz.size() / 3);
		std::vector<double> QA(2 * (size_t)nq * nq), QB(QA.size());
		orc_assemble_bm_elems(qxyz.data(), nqn, quad.data(), nq, 4, 2.0, 0.0, -0.5, &pol, nullptr, nq, QA.data(),
			QB.data(), nq, 4);
		for (double v : L) if (!std::isfinite(v)) { std::printf("oracle: Galerkin entry not finite\n"); ++bad; break; }
		for (double v : QA) if (!std::isfinite(v)) { std::printf("oracle: quadrilateral entry not finite\n"); ++bad; break; }
	}
#endif
	std::printf("%s\n", bad ? "FAILED" : "OK");
	return bad ? 1 : 0;
}
