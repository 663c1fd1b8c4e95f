// nbem_api.hip -- C ABI of the engine (include/nbem.h): mesh residency,
// per-mesh caches (classification thresholds, quadrature tables, near-pair
// lists, Galerkin layouts and singular pair lists) and the launch sequence of
// one assembly.
// Collocation:
//   1. regular pairs, FP32 (regular_kernel)
//   2. near-singular pairs, FP64 (fp64_pass_kernel<..., false>)
//   3. diagonal pairs on the same mesh, FP64 (fp64_pass_kernel<..., true>)
// The three passes write disjoint entries, so no atomics are used and every
// entry is computed by one lane in a fixed order (results do not depend on
// what else runs on the device).
// Galerkin:
//   1. regular pairs, FP32 (galerkin_regular_kernel)
//   2. coincident / edge / vertex pairs on the same mesh, FP64
//      (galerkin_singular_kernel)
// With nodal (isoparametric) dofs an entry sums the blocks of every element
// pair around its two nodes; those sums use atomics, so their rounding order
// (not their value beyond FP32 / FP64 rounding) depends on scheduling.

#include "../../include/nbem.h"
#include "nbem_kernels.hpp"
#include "nbem_sweep.hpp"
#include "nbem_galerkin.hpp"
#include "nbem_quad.hpp"
#include "nbem_gauss.hpp"
#include "nbem_galquad.hpp"
#include "nbem_solve.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <tuple>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

extern "C" {
int nbg_dunavant(int degree, double *xi, double *eta, double *w, int cap);
void nbg_gauss_legendre(int n, double *x, double *w);
int nbg_geometry(const double *xyz, int n_nodes, const int32_t *tri, int n, double *c, double *nu,
	double *J, double *R, double *diam, double *X);
double nbg_threshold(double scale, double limit);
int nbg_quad_table(int degree, int n, const double *c, const double *J, const double *X, float *out);
int nbg_quad_table64(int degree, int n, const double *c, const double *J, const double *X, double *out);
int nbg_singular_rule(int match, double *out, int cap);
int64_t nbg_singular_pairs(int n_nodes, int n, const int32_t *tri, int32_t *pe, int32_t *pf, int32_t *info);
int nbg_galerkin_layout(int n, const int32_t *tri, const double *c, int chunk, int32_t *order_out, int32_t *elem,
	uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin);
int nbg_geometry_quad(const double *xyz, int n_nodes, const int32_t *quad, int n, double *c, double *nu,
	double *J, double *R, double *diam, double *X);
int nbg_quad4_table(int N, int n, const double *c, const double *X, float *out);
int nbg_quad_duffy(int ng, double x0, double y0, double *out);
int nbg_gauss_nodes(int n, const double *X, double *x, double *nu);
int nbg_quad4_table64(int N, int n, const double *X, double *out);
int nbg_gauss_shapes(int N, double *out);
int nbg_quad_gal_table(int N, int n, const double *c, const double *X, float *out);
int nbg_quad_singular_rule(int match, double *out, int cap);
int64_t nbg_singular_pairs_n(int n_nodes, int n, const int32_t *el, int npe, int32_t *pe, int32_t *pf, int32_t *info);
int nbg_galerkin_layout_n(int n, const int32_t *tri, int npe, int cs, const double *c, int chunk,
	int32_t *order_out, int32_t *elem, uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin);
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
	g_err = msg;
	return code;
}

#define NB_HIP(expr)                                                                               \
	do {                                                                                           \
		hipError_t e_ = (expr);                                                                    \
		if (e_ != hipSuccess) {                                                                    \
			if (e_ == hipErrorOutOfMemory) return fail(NB_ERR_OOM, std::string(#expr) + ": out of device memory"); \
			return fail(NB_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));          \
		}                                                                                          \
	} while (0)

struct HipFree {
	void operator()(void *p) const { if (p) (void)hipFree(p); }
};
typedef std::unique_ptr<void, HipFree> DevPtr;

int dev_alloc(DevPtr &p, size_t bytes)
{
	void *q = nullptr;
	if (bytes == 0) bytes = 16;
	hipError_t e = hipMalloc(&q, bytes);
	if (e == hipErrorOutOfMemory) return fail(NB_ERR_OOM, "hipMalloc: out of device memory");
	if (e != hipSuccess) return fail(NB_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
	p.reset(q);
	return NB_OK;
}

template <class T>
int upload(DevPtr &p, const std::vector<T> &v)
{
	int s = dev_alloc(p, v.size() * sizeof(T));
	if (s) return s;
	NB_HIP(hipMemcpy(p.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
	return NB_OK;
}

struct NearList {
	uint64_t test_uid;
	int same, mode;
	double limit;
	int n;
	DevPtr pi, pj;
};

struct LaunchCfg {
	int tile_cols = 32;
	int k_batch = 4;
	int persistent = 4;
};
std::mutex g_cfg_mu;
LaunchCfg g_cfg;

// per-pass device timing (nb_profile)
constexpr int NPASS = 4;
std::mutex g_prof_mu;
bool g_prof_on = false;
struct ProfEv { int pass; hipEvent_t a, b; };
std::vector<ProfEv> g_prof_pending;
double g_prof_ms[NPASS] = {0, 0, 0, 0};
int64_t g_prof_n[NPASS] = {0, 0, 0, 0};

bool prof_enabled()
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	return g_prof_on;
}

// record the start of a pass; returns the event pair index or -1
int prof_begin(int pass, hipStream_t st)
{
	if (!prof_enabled()) return -1;
	ProfEv e;
	e.pass = pass;
	if (hipEventCreate(&e.a) != hipSuccess) return -1;
	if (hipEventCreate(&e.b) != hipSuccess) { (void)hipEventDestroy(e.a); return -1; }
	(void)hipEventRecord(e.a, st);
	std::lock_guard<std::mutex> lk(g_prof_mu);
	g_prof_pending.push_back(e);
	return (int)g_prof_pending.size() - 1;
}

void prof_end(int idx, hipStream_t st)
{
	if (idx < 0) return;
	hipEvent_t b;
	{
		std::lock_guard<std::mutex> lk(g_prof_mu);
		b = g_prof_pending[idx].b;
	}
	(void)hipEventRecord(b, st);
}

// per-device side stream: the FP64 passes run on it, forked from and joined
// back into the caller's stream with events, so that they overlap the regular
// pass (all passes write disjoint entries)
std::mutex g_side_mu;
std::map<int, hipStream_t> g_side;

int side_stream(hipStream_t *out)
{
	// the FP64 passes run on the caller's stream after the regular pass unless
	// NBEM_OVERLAP=1 puts them on a side stream that overlaps it (same result;
	// measured on MI355X the overlap gains nothing once the FP64 pass is cheap,
	// since both passes are VALU-bound)
	static const bool serial = [] {
		const char *e = std::getenv("NBEM_OVERLAP");
		return !(e && e[0] == '1');
	}();
	if (serial) { *out = nullptr; return NB_OK; }
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_side_mu);
	auto it = g_side.find(dev);
	if (it == g_side.end()) {
		hipStream_t s;
		NB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		it = g_side.emplace(dev, s).first;
	}
	*out = it->second;
	return NB_OK;
}

// per-device rule tables of the FP64 pass
std::mutex g_rules_mu;
std::map<int, DevPtr> g_rules;

int device_rules(const nbk::Rules **out)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_rules_mu);
	auto it = g_rules.find(dev);
	if (it == g_rules.end()) {
		nbk::Rules r;
		nbg_gauss_legendre(6, r.gl6x, r.gl6w);
		nbg_gauss_legendre(21, r.gl21x, r.gl21w);
		if (nbg_dunavant(7, r.d7xi, r.d7eta, r.d7w, 13) != 13) return fail(NB_ERR_RUNTIME, "Dunavant 7 table");
		if (nbg_dunavant(14, r.d14xi, r.d14eta, r.d14w, 42) != 42) return fail(NB_ERR_RUNTIME, "Dunavant 14 table");
		nbg_gauss_legendre(3, r.gl3x, r.gl3w);
		nbg_gauss_legendre(5, r.gl5x, r.gl5w);
		nbg_gauss_legendre(8, r.gl8x, r.gl8w);
		nbg_gauss_legendre(16, r.gl16x, r.gl16w);
		nbg_gauss_legendre(51, r.gl51x, r.gl51w);
		if (nbg_quad_duffy(4, 0.0, 0.0, &r.duffy[0][0]) != 64) return fail(NB_ERR_RUNTIME, "Duffy rule");
		for (int a = 0; a < 4; ++a) {
			const double g = std::sqrt(3.0) / 3.0;
			if (nbg_quad_duffy(4, (a & 1) ? g : -g, (a & 2) ? g : -g, &r.gduffy[a][0][0]) != 64)
				return fail(NB_ERR_RUNTIME, "Duffy rule");
		}
		for (int d = 0; d < nbk::NSLOT; ++d) {
			double xi[64], eta[64], w[64];
			const int q = d <= 14 ? nbg_dunavant(d, xi, eta, w, 64) : 0;
			if (q > nbk::REF_MAXQ) return fail(NB_ERR_RUNTIME, "Dunavant reference table");
			for (int p = 0; p < nbk::REF_MAXQ; ++p)
				r.ref[d][p] = p < q ? make_float2((float)xi[p], (float)eta[p]) : make_float2(0.0f, 0.0f);
		}
		DevPtr p;
		int s = dev_alloc(p, sizeof(r));
		if (s) return s;
		NB_HIP(hipMemcpy(p.get(), &r, sizeof(r), hipMemcpyHostToDevice));
		it = g_rules.emplace(dev, std::move(p)).first;
	}
	*out = (const nbk::Rules *)it->second.get();
	return NB_OK;
}

// the rule tables of the current device if they were built already (no allocation)
int device_rules_ready(const nbk::Rules **out)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_rules_mu);
	auto it = g_rules.find(dev);
	if (it == g_rules.end()) return fail(NB_ERR_ARG, "the device's quadrature rule tables are not resident yet");
	*out = (const nbk::Rules *)it->second.get();
	return NB_OK;
}

// per-device canonical singular rules of the Galerkin pass (vertex, edge, face)
struct SingRules {
	DevPtr r[3];
	int n[3];
};
std::mutex g_srules_mu;
std::map<int, std::unique_ptr<SingRules> > g_srules;

int device_sing_rules(const SingRules **out)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_srules_mu);
	auto it = g_srules.find(dev);
	if (it == g_srules.end()) {
		std::unique_ptr<SingRules> sr(new SingRules());
		std::vector<double> buf(5 * 1536);
		for (int m = 0; m < 3; ++m) {
			const int n = nbg_singular_rule(m, buf.data(), 1536);
			if (n <= 0) return fail(NB_ERR_RUNTIME, "singular Galerkin rule");
			std::vector<double> v(buf.begin(), buf.begin() + 5 * n);
			int s = upload(sr->r[m], v);
			if (s) return s;
			sr->n[m] = n;
		}
		it = g_srules.emplace(dev, std::move(sr)).first;
	}
	*out = it->second.get();
	return NB_OK;
}

// the quad x quad rules of the Galerkin singular pass, per device
std::mutex g_qsrules_mu;
std::map<int, std::unique_ptr<SingRules> > g_qsrules;

int device_quad_sing_rules(const SingRules **out)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_qsrules_mu);
	auto it = g_qsrules.find(dev);
	if (it == g_qsrules.end()) {
		std::unique_ptr<SingRules> sr(new SingRules());
		std::vector<double> buf(5 * 2048);
		for (int m = 0; m < 3; ++m) {
			const int n = nbg_quad_singular_rule(m, buf.data(), 2048);
			if (n <= 0) return fail(NB_ERR_RUNTIME, "singular Galerkin rule of quadrilaterals");
			std::vector<double> v(buf.begin(), buf.begin() + 5 * n);
			int s = upload(sr->r[m], v);
			if (s) return s;
			sr->n[m] = n;
		}
		it = g_qsrules.emplace(dev, std::move(sr)).first;
	}
	*out = it->second.get();
	return NB_OK;
}

} // namespace

struct nb_mesh {
	uint64_t uid = 0;   // unique over the process: cache key that survives address reuse
	int device = 0;
	int n_nodes = 0, n = 0;
	int npe = 3;                               // nodes per element: 3 triangles, 4 quadrilaterals
	std::vector<double> c, nu, J, R, diam, X;   // SoA host copies (FP64)
	DevPtr d_c, d_nuf, d_nu64, d_X, d_J;
	int thr_mode = -1;
	double thr_limit = 0.0;
	DevPtr d_thr, d_rec;
	DevPtr d_qt[nbk::NSLOT];
	int qn[nbk::NSLOT] = {0};
	DevPtr d_qt64[nbk::NSLOT];      // FP64 point tables of the FP64 precision policy
	DevPtr d_qq[nbk::QN_MAX + 1];              // quadrilaterals: Gauss square tables by order
	DevPtr d_qq64[nbk::QN_MAX + 1];            // the same in FP64 (y, w J, unit normal): FP64 precision policy
	// near-pair lists per (test mesh, near policy), oldest first; callers hold
	// a list by shared_ptr, so evicting it cannot free one in use
	std::vector<std::shared_ptr<NearList> > near_cache;
	std::vector<int32_t> tri;                  // element-major node ids (host)
	DevPtr d_tri;                              // node ids [npe][n] (quadrilaterals: uploaded on first use)
	// Galerkin, nodal dofs: element order, wave tiles, column-slot schedule
	// (glay[0]: node dofs; glay[1]: element dofs, the dof map of a constant
	// field in a pair with a nodal one)
	struct GalLayout {
		int tiles = 0, chunk = 0;
		DevPtr order, elem, contrib, rows, col, fin;
	} glay[2];
	DevPtr d_edof;                             // element dof map [npe][n]: e at every corner
	// Galerkin singular pairs against itself (coincident pairs first)
	int64_t sp_n = -1;
	DevPtr sp_e, sp_f, sp_info;
	// both element types (npe = 0): the triangle part, the quadrilateral part
	// and, per part, the element id (table row) of each of its elements
	std::unique_ptr<nb_mesh> part[2];
	DevPtr d_ids[2];
	// every element's collocation point and normal in table order (the test
	// side of the regular pass over all rows)
	std::unique_ptr<nb_mesh> view;
	// quadrilateral near pass: count of pairs without an inverse image
	DevPtr d_status;
	// the gauss field's nodes as test points (quadrilaterals): x(xi_a) [3][4n]
	// FP64, the unit normals there FP32 and FP64
	DevPtr d_gx, d_gnuf, d_gnu64;
	// Galerkin on quadrilaterals: point rows of the N x N Gauss rules
	DevPtr d_gq[nbk::GALQ_NMAX + 1];
	// the sweep's per-call buffers (wavenumber constants, FP64 static parts),
	// kept between calls and grown when a call needs more: a hipFree per call
	// would wait for every stream of the device
	DevPtr sweep_ws;
	size_t sweep_ws_bytes = 0;
	std::mutex mu;
	int64_t bytes = 0;
};

namespace {

// squared-distance thresholds of the trial mesh for a near-singular policy
int ensure_thresholds(nb_mesh *m, int mode, double limit)
{
	if (m->d_thr && m->thr_mode == mode && m->thr_limit == limit) return NB_OK;
	const int n = m->n;
	// breakpoints x = num / den of the accuracy-3 tables, evaluated as NiHu's
	// tmp::break_point::x_value does
	static const double bp[nbk::NTHR - 1] = {10. / 10., 11. / 10., 14. / 10., 15. / 10., 18. / 10., 19. / 10.,
		28. / 10., 36. / 10., 91. / 10., 158. / 10., 223. / 10.};
	std::vector<double> thr((size_t)nbk::NTHR * n);
	for (int e = 0; e < n; ++e) {
		double scale = mode == NB_NEAR_DIAMETER ? m->diam[e] : m->R[e];
		thr[e] = mode == NB_NEAR_NONE ? 0.0 : nbg_threshold(scale, limit);
		for (int t = 1; t < nbk::NTHR; ++t) thr[(size_t)t * n + e] = nbg_threshold(m->R[e], bp[t - 1]);
	}
	int s = upload(m->d_thr, thr);
	if (s) return s;
	// the same thresholds, the centre, normal and one-point weight per element
	// in one record (the regular pass reads it through the scalar cache); the
	// one-point weight assumes Dunavant degrees 0 and 1 are the centroid rule
	double xi[4], eta[4], w[4];
	for (int deg = 0; deg <= 1; ++deg)
		if (nbg_dunavant(deg, xi, eta, w, 4) != 1 || xi[0] != 1. / 3. || eta[0] != 1. / 3. || w[0] != 0.5)
			return fail(NB_ERR_RUNTIME, "Dunavant degrees 0 and 1 must be the centroid rule");
	const double four_pi = 4.0 * 3.141592653589793238462643383279502884;
	std::vector<nbk::TrialRec> rec(n);
	for (int e = 0; e < n; ++e) {
		nbk::TrialRec &r = rec[e];
		std::memset(&r, 0, sizeof(r));
		for (int t = 0; t < nbk::NTHR; ++t) r.thr[t] = thr[(size_t)t * n + e];
		for (int d = 0; d < 3; ++d) { r.c[d] = m->c[(size_t)d * n + e]; r.nu[d] = (float)m->nu[(size_t)d * n + e]; }
		// one-point rule: Dunavant centroid (weight 1/2) or the Gauss centre of [-1, 1]^2 (weight 4)
		r.w1 = (float)((m->npe == 4 ? 4.0 : 0.5) * m->J[e] / four_pi);
	}
	if ((s = upload(m->d_rec, rec))) return s;
	if (m->thr_mode < 0) m->bytes += (int64_t)n * (nbk::NTHR * 8 + sizeof(nbk::TrialRec));
	m->thr_mode = mode;
	m->thr_limit = limit;
	return NB_OK;
}

int ensure_quad(nb_mesh *m, int degree)
{
	if (degree < 0 || degree > 14) return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
	if (m->d_qt[degree]) return NB_OK;
	double xi[64], eta[64], w[64];
	int Q = nbg_dunavant(degree, xi, eta, w, 64);
	std::vector<float> tab((size_t)m->n * Q * 8);
	if (nbg_quad_table(degree, m->n, m->c.data(), m->J.data(), m->X.data(), tab.data()) != Q)
		return fail(NB_ERR_RUNTIME, "quadrature table");
	int s = upload(m->d_qt[degree], tab);
	if (s) return s;
	m->qn[degree] = Q;
	m->bytes += (int64_t)tab.size() * 4;
	return NB_OK;
}

// FP64 point table of one degree (offsets from the centre, w J / (4 pi))
int ensure_quad64(nb_mesh *m, int degree)
{
	int s;
	if ((s = ensure_quad(m, degree))) return s;
	if (m->d_qt64[degree]) return NB_OK;
	std::vector<double> tab((size_t)m->n * m->qn[degree] * 4);
	if (nbg_quad_table64(degree, m->n, m->c.data(), m->J.data(), m->X.data(), tab.data()) != m->qn[degree])
		return fail(NB_ERR_RUNTIME, "quadrature table");
	if ((s = upload(m->d_qt64[degree], tab))) return s;
	m->bytes += (int64_t)tab.size() * 8;
	return NB_OK;
}

// Gauss square table with N x N points of a quadrilateral mesh
int ensure_qq(nb_mesh *m, int N)
{
	if (N < 1 || N > nbk::QN_MAX) return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
	if (m->d_qq[N]) return NB_OK;
	std::vector<float> tab((size_t)m->n * N * N * 8);
	if (nbg_quad4_table(N, m->n, m->c.data(), m->X.data(), tab.data()) != N * N)
		return fail(NB_ERR_RUNTIME, "quadrilateral quadrature table");
	int s = upload(m->d_qq[N], tab);
	if (s) return s;
	m->bytes += (int64_t)tab.size() * 4;
	return NB_OK;
}

// FP64 Gauss square table of order N (the FP64 precision policy)
int ensure_qq64(nb_mesh *m, int N)
{
	if (N < 1 || N > nbk::QN_MAX) return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
	if (m->d_qq64[N]) return NB_OK;
	std::vector<double> tab((size_t)m->n * N * N * 8);
	if (nbg_quad4_table64(N, m->n, m->X.data(), tab.data()) != N * N)
		return fail(NB_ERR_RUNTIME, "quadrilateral quadrature table");
	int s = upload(m->d_qq64[N], tab);
	if (s) return s;
	m->bytes += (int64_t)tab.size() * 8;
	return NB_OK;
}

// the gauss field's nodes of a quadrilateral mesh (test points)
int ensure_gauss_nodes(nb_mesh *m)
{
	if (m->d_gx) return NB_OK;
	const size_t r = (size_t)4 * m->n;
	std::vector<double> x(3 * r), nu(3 * r);
	if (nbg_gauss_nodes(m->n, m->X.data(), x.data(), nu.data()))
		return fail(NB_ERR_ARG, "degenerate quadrilateral (zero normal at a Gauss point)");
	std::vector<float> nuf(nu.begin(), nu.end());
	int s;
	if ((s = upload(m->d_gx, x)) || (s = upload(m->d_gnu64, nu)) || (s = upload(m->d_gnuf, nuf))) return s;
	m->bytes += (int64_t)r * (24 + 24 + 12);
	return NB_OK;
}

// per-device shape-function tables of the gauss field by Gauss order
std::mutex g_gshape_mu;
std::map<int, DevPtr> g_gshape;

int device_gauss_shapes(const nbk::GaussShapes **out)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_gshape_mu);
	auto it = g_gshape.find(dev);
	if (it == g_gshape.end()) {
		std::unique_ptr<nbk::GaussShapes> g(new nbk::GaussShapes());
		std::memset(g.get(), 0, sizeof(nbk::GaussShapes));
		std::vector<double> v(4 * nbk::QN_MAX * nbk::QN_MAX);
		for (int N = 1; N <= nbk::QN_MAX; ++N) {
			if (nbg_gauss_shapes(N, v.data()) != N * N) return fail(NB_ERR_RUNTIME, "gauss field shape table");
			for (int q = 0; q < N * N; ++q)
				for (int b = 0; b < 4; ++b) {
					g->n[N][q][b] = (float)v[4 * q + b];
					g->n64[N][q][b] = v[4 * q + b];
				}
		}
		DevPtr p;
		int s = dev_alloc(p, sizeof(nbk::GaussShapes));
		if (s) return s;
		NB_HIP(hipMemcpy(p.get(), g.get(), sizeof(nbk::GaussShapes), hipMemcpyHostToDevice));
		it = g_gshape.emplace(dev, std::move(p)).first;
	}
	*out = (const nbk::GaussShapes *)it->second.get();
	return NB_OK;
}

// degrees the regular pass may select under a policy (kernel orders in km)
void policy_degrees(const nb_policy &p, int km, int field_complexity, std::vector<int> &deg)
{
	deg.clear();
	if (p.rule == NB_RULE_FIXED) { deg.push_back(p.fixed_degree); return; }
	static const int d1[] = {6, 4, 2, 0}, d2[] = {8, 6, 4, 2, 0}, d3[] = {9, 8, 6, 4, 2, 0};
	auto add = [&](const int *d, int n) { for (int t = 0; t < n; ++t) deg.push_back(field_complexity + d[t]); };
	if (km & nbk::SLP) add(d1, 4);
	if (km & (nbk::DLP | nbk::DLPT)) add(d2, 5);
	if (km & nbk::HSP) add(d3, 6);
	std::sort(deg.begin(), deg.end());
	deg.erase(std::unique(deg.begin(), deg.end()), deg.end());
}

// the cached near-pair list of (test mesh, near policy), or null
std::shared_ptr<NearList> find_near_list(const nb_mesh *trial, const nb_mesh *test, int same, const nb_policy &pol)
{
	for (auto &nl : trial->near_cache)
		if (nl->test_uid == test->uid && nl->same == same && nl->mode == pol.near_mode && nl->limit == pol.near_limit)
			return nl;
	return nullptr;
}

int ensure_near_list(nb_mesh *trial, const nb_mesh *test, int same, const nb_policy &pol, hipStream_t st,
	std::shared_ptr<NearList> *out)
{
	if ((*out = find_near_list(trial, test, same, pol))) return NB_OK;
	std::shared_ptr<NearList> nl(new NearList());
	nl->test_uid = test->uid; nl->same = same; nl->mode = pol.near_mode; nl->limit = pol.near_limit; nl->n = 0;
	if (pol.near_mode != NB_NEAR_NONE) {
		nbk::ListArgs a;
		a.n_rows = test->n; a.n_cols = trial->n; a.same_mesh = same;
		a.row_tiles = (test->n + 255) / 256;
		a.chunk = 1024;
		a.n_chunks = (trial->n + a.chunk - 1) / a.chunk;
		a.tn = test->n; a.sn = trial->n;
		a.tc = (const double *)test->d_c.get();
		a.sc = (const double *)trial->d_c.get();
		a.thr_near = (const double *)trial->d_thr.get();
		const size_t ncnt = (size_t)a.n_rows * a.n_chunks;
		DevPtr counts, offsets, tmp;
		int s;
		if ((s = dev_alloc(counts, ncnt * sizeof(int)))) return s;
		if ((s = dev_alloc(offsets, ncnt * sizeof(int)))) return s;
		a.counts = (int *)counts.get();
		a.offsets = (const int *)offsets.get();
		a.pi = a.pj = nullptr;
		const int grid = a.row_tiles * a.n_chunks;
		hipLaunchKernelGGL(nbk::near_list_kernel<false>, dim3(grid), dim3(256), 0, st, a);
		NB_HIP(hipGetLastError());
		size_t tmp_bytes = 0;
		NB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (int *)counts.get(), (int *)offsets.get(), (int)ncnt, st));
		if ((s = dev_alloc(tmp, tmp_bytes))) return s;
		NB_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tmp_bytes, (int *)counts.get(), (int *)offsets.get(), (int)ncnt, st));
		int last[2];
		NB_HIP(hipMemcpyAsync(&last[0], (int *)offsets.get() + ncnt - 1, sizeof(int), hipMemcpyDeviceToHost, st));
		NB_HIP(hipMemcpyAsync(&last[1], (int *)counts.get() + ncnt - 1, sizeof(int), hipMemcpyDeviceToHost, st));
		NB_HIP(hipStreamSynchronize(st));
		nl->n = last[0] + last[1];
		if ((s = dev_alloc(nl->pi, (size_t)nl->n * sizeof(int)))) return s;
		if ((s = dev_alloc(nl->pj, (size_t)nl->n * sizeof(int)))) return s;
		a.pi = (int *)nl->pi.get();
		a.pj = (int *)nl->pj.get();
		hipLaunchKernelGGL(nbk::near_list_kernel<true>, dim3(grid), dim3(256), 0, st, a);
		NB_HIP(hipGetLastError());
		NB_HIP(hipStreamSynchronize(st));   // counts/offsets are freed on return
		trial->bytes += (int64_t)nl->n * 8;
	}
	if (trial->near_cache.size() >= 8) {
		trial->bytes -= (int64_t)trial->near_cache.front()->n * 8;
		trial->near_cache.erase(trial->near_cache.begin());
	}
	trial->near_cache.push_back(nl);
	*out = std::move(nl);
	return NB_OK;
}

// trial-sequence chunk per work-group: as long as possible (the slot
// accumulators flush at every chunk end) while the grid keeps >= 2048
// work-groups for the 256 CUs
int gal_chunk(int n_tiles, int sn)
{
	int ch = 512;
	const long long wg = (n_tiles + 3) / 4;
	while (ch > 64 && wg * ((sn + ch - 1) / ch) < 2048) ch >>= 1;
	return ch;
}

// the nodal-dof layout of a mesh of triangles (GAL_CS slots per lane) or
// quadrilaterals (GALQ_CS slots)
// dof map of a mesh's corners: kind 0 the nodes, kind 1 the element itself
// (a constant field: the nodal kernels' sums over the corners then give the
// constant field's integrals, as the shape functions sum to one)
std::vector<int32_t> dof_map(const nb_mesh *m, int kind)
{
	if (kind == 0) return m->tri;
	std::vector<int32_t> d((size_t)m->npe * m->n);
	for (int e = 0; e < m->n; ++e)
		for (int a = 0; a < m->npe; ++a) d[(size_t)m->npe * e + a] = e;
	return d;
}

// the element dof map on the device, [npe][n]
int ensure_edof(nb_mesh *m)
{
	if (m->d_edof) return NB_OK;
	std::vector<int32_t> soa((size_t)m->npe * m->n);
	for (int e = 0; e < m->n; ++e)
		for (int a = 0; a < m->npe; ++a) soa[(size_t)a * m->n + e] = e;
	int s = upload(m->d_edof, soa);
	if (s) return s;
	m->bytes += (int64_t)soa.size() * 4;
	return NB_OK;
}

int ensure_gal_layout(nb_mesh *m, int kind = 0)
{
	nb_mesh::GalLayout &L = m->glay[kind];
	if (L.tiles > 0) return NB_OK;
	const int n = m->n, npe = m->npe;
	const std::vector<int32_t> dm = dof_map(m, kind);
	const int cs = npe == 4 ? nbk::GALQ_CS : nbk::GAL_CS;
	const int tiles = nbg_galerkin_layout_n(n, dm.data(), npe, cs, m->c.data(), 512, nullptr, nullptr, nullptr,
		nullptr, nullptr, nullptr);
	if (tiles <= 0) return fail(NB_ERR_RUNTIME, "Galerkin layout");
	const int chunk = gal_chunk(tiles, n);
	const size_t nl = (size_t)tiles * nbk::GAL_TILE, rpl = npe == 4 ? 2 : 1;
	std::vector<int32_t> order(n), elem(nl), rows(rpl * nl), col((size_t)(npe + 1) * n),
		fin((size_t)((n + chunk - 1) / chunk) * cs);
	std::vector<uint32_t> contrib(2 * rpl * nl);
	if (nbg_galerkin_layout_n(n, dm.data(), npe, cs, m->c.data(), chunk, order.data(), elem.data(),
		contrib.data(), rows.data(), col.data(), fin.data()) != tiles)
		return fail(NB_ERR_RUNTIME, "Galerkin layout");
	int s;
	if ((s = upload(L.order, order)) || (s = upload(L.elem, elem)) || (s = upload(L.contrib, contrib)) ||
		(s = upload(L.rows, rows)) || (s = upload(L.col, col)) || (s = upload(L.fin, fin)))
		return s;
	L.tiles = tiles;
	L.chunk = chunk;
	m->bytes += (int64_t)(order.size() + elem.size() + rows.size() + contrib.size() + col.size() + fin.size()) * 4;
	return NB_OK;
}

int ensure_sing_pairs(nb_mesh *m)
{
	if (m->sp_n >= 0) return NB_OK;
	const int64_t cnt = nbg_singular_pairs_n(m->n_nodes, m->n, m->tri.data(), m->npe, nullptr, nullptr, nullptr);
	if (cnt > 0x7fffffff) return fail(NB_ERR_RANGE, "too many singular element pairs");
	std::vector<int32_t> pe(cnt), pf(cnt), info(cnt);
	nbg_singular_pairs_n(m->n_nodes, m->n, m->tri.data(), m->npe, pe.data(), pf.data(), info.data());
	int s;
	if ((s = upload(m->sp_e, pe)) || (s = upload(m->sp_f, pf)) || (s = upload(m->sp_info, info))) return s;
	m->sp_n = cnt;
	m->bytes += cnt * 12;
	return NB_OK;
}

int dtype_bytes(int dt)
{
	switch (dt) {
	case NB_C64: return 8;
	case NB_C128: return 16;
	case NB_F32: return 4;
	case NB_F64: return 8;
	default: return 0;
	}
}

// zero_kernel over a rows x cols block: one grid row per column (up to 65535,
// the kernel strides beyond), enough blocks along x for a column of 16-byte words
void launch_zero(void *out, int64_t ld, int64_t rows, int64_t cols, int elem_bytes, hipStream_t st)
{
	if (rows <= 0 || cols <= 0) return;
	const int64_t words16 = (rows * elem_bytes + 15) / 16;
	const unsigned gx = (unsigned)std::min<int64_t>((words16 + 255) / 256, 64);
	const unsigned gy = (unsigned)std::min<int64_t>(cols, 65535);
	hipLaunchKernelGGL(nbk::zero_kernel, dim3(gx, gy), dim3(256), 0, st, out, (long long)ld, (long long)rows,
		(long long)cols, elem_bytes);
}

int check_out(const nb_out *o, int64_t rows, int64_t cols, bool complex_needed)
{
	if (!o) return fail(NB_ERR_ARG, "null output descriptor");
	if (!o->ptr) return fail(NB_ERR_ARG, "null output pointer");
	if (dtype_bytes(o->dtype) == 0) return fail(NB_ERR_ARG, "unknown output dtype");
	if (o->location != NB_HOST && o->location != NB_DEVICE) return fail(NB_ERR_ARG, "unknown output location");
	if (o->rows != rows || o->cols != cols)
		return fail(NB_ERR_ARG, "output size must be n_test x n_trial");
	if (o->ld < rows) return fail(NB_ERR_ARG, "leading dimension smaller than the number of rows");
	if (complex_needed && (o->dtype == NB_F32 || o->dtype == NB_F64))
		return fail(NB_ERR_ARG, "complex result into a real output");
	return NB_OK;
}

struct Job {
	const nb_mesh *test;
	nb_mesh *trial;
	int formalism;
	int same, family, km, mode;
	double k;
	double ki = 0.0;      // imaginary part of a complex wavenumber
	nb_scalar alpha;
	nb_policy pol;
	int tile_cols;
	// a block of a mesh with both element types: result row / column of each
	// test / trial element of the block (null: its own index)
	const int *rmap = nullptr, *cmap = nullptr;
	// passes to run: 1 regular pairs, 2 near-singular and diagonal pairs
	int passes = 3;
};

// element nodes [npe][n] on the device (quadrilateral meshes upload them
// on first use)
int ensure_elem_nodes(nb_mesh *m)
{
	if (m->d_tri) return NB_OK;
	const int n = m->n, npe = m->npe;
	std::vector<int32_t> soa((size_t)npe * n);
	for (int e = 0; e < n; ++e)
		for (int a = 0; a < npe; ++a) soa[(size_t)a * n + e] = m->tri[(size_t)npe * e + a];
	int s = upload(m->d_tri, soa);
	if (s) return s;
	m->bytes += (int64_t)soa.size() * 4;
	return NB_OK;
}

// Galerkin point rows of the N x N Gauss rule of a quadrilateral mesh
int ensure_gq(nb_mesh *m, int N)
{
	if (N < 1 || N > nbk::GALQ_NMAX) return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
	if (m->d_gq[N]) return NB_OK;
	std::vector<float> tab((size_t)m->n * N * N * nbk::GALQ_ROW);
	if (nbg_quad_gal_table(N, m->n, m->c.data(), m->X.data(), tab.data()) != N * N)
		return fail(NB_ERR_RUNTIME, "quadrilateral Galerkin table");
	int s = upload(m->d_gq[N], tab);
	if (s) return s;
	m->bytes += (int64_t)tab.size() * 4;
	return NB_OK;
}

// Galerkin assembly on quadrilateral meshes into a device buffer: the FP32
// regular pass over all pairs (node-sharing pairs skipped), then on one mesh
// the FP64 pass over the coincident, edge- and corner-adjacent pairs
int run_galerkin_quad(const Job &jb, void *out, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	nb_mesh *trial = jb.trial;
	nb_mesh *test = const_cast<nb_mesh *>(jb.test);
	// the nodal kernels serve every pair with an isoparametric side; a
	// constant side runs through its element dof map
	const bool lt = jb.formalism == NB_GALERKIN_LINEAR || jb.formalism == NB_GALERKIN_LINEAR_CONST;
	const bool ls = jb.formalism == NB_GALERKIN_LINEAR || jb.formalism == NB_GALERKIN_CONST_LINEAR;
	const bool nodal = lt || ls;
	const bool ck = jb.ki != 0.0 && jb.family == NB_HELMHOLTZ;
	const int64_t rows = lt ? test->n_nodes : test->n, cols = ls ? trial->n_nodes : trial->n;
	const int eb = dtype_bytes(dtype);
	const double four_pi = 4.0 * 3.141592653589793238462643383279502884;
	auto zero = [&]() -> int {
		launch_zero(out, ld, rows, cols, eb, st);
		NB_HIP(hipGetLastError());
		return NB_OK;
	};
	int s;
	if ((s = ensure_elem_nodes(test)) || (s = ensure_elem_nodes(trial))) return s;
	if ((!lt && nodal && (s = ensure_edof(test))) || (!ls && nodal && (s = ensure_edof(trial)))) return s;
	const int *t_map = (const int *)(lt ? test->d_tri.get() : test->d_edof.get());
	const int *s_map = (const int *)(ls ? trial->d_tri.get() : trial->d_edof.get());
	if (jb.family == NB_IDENTITY) {
		if (!accumulate && (s = zero())) return s;
		// local operator: nothing on different meshes (nihu/core/assembly.hpp:141-142)
		if (jb.same)
			hipLaunchKernelGGL(nbk::galq_identity_kernel, dim3((test->n + 255) / 256), dim3(256), 0, st, out,
				(long long)ld, test->n, (const double *)test->d_X.get(), t_map, s_map,
				nodal ? 1 : 0, dtype, jb.alpha.re, jb.alpha.im);
		NB_HIP(hipGetLastError());
		return NB_OK;
	}
	// 1/r^3 has no valid singular rule on coincident and edge-adjacent pairs
	// (nihu/core/singular_accelerator.hpp:483-503, double_integral.hpp:255-273)
	if (jb.same && jb.km == nbk::HSP) return fail(NB_ERR_RUNTIME, "Invalid general singular quadrature");
	const int complexity = nodal ? 2 : 1;
	int fixed_n = 0;
	std::vector<int> ns;
	if (jb.pol.rule == NB_RULE_FIXED) {
		if (jb.pol.fixed_degree > 14)
			return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
		fixed_n = jb.pol.fixed_degree / 2 + 1;
		ns.push_back(fixed_n);
	} else {
		if ((s = ensure_thresholds(trial, jb.pol.near_mode, jb.pol.near_limit))) return s;
		std::vector<int> degs;
		policy_degrees(jb.pol, jb.km, complexity, degs);
		for (int d : degs) ns.push_back(d / 2 + 1);
	}
	for (int N : ns)
		if ((s = ensure_gq(trial, N)) || (s = ensure_gq(test, N))) return s;
	const int kt = lt ? 0 : 1, ks = ls ? 0 : 1;   // layout kinds: node or element dofs
	if (nodal && ((s = ensure_gal_layout(test, kt)) || (s = ensure_gal_layout(trial, ks)))) return s;
	if (jb.same && (s = ensure_sing_pairs(trial))) return s;
	if (nodal && !accumulate && (s = zero())) return s;

	nbk::GalQArgs a;
	std::memset(&a, 0, sizeof(a));
	if (nodal) {
		a.n_tiles = test->glay[kt].tiles;
		a.chunk = trial->glay[ks].chunk;
		a.t_elem = (const int *)test->glay[kt].elem.get();
		a.t_contrib = (const uint2 *)test->glay[kt].contrib.get();
		a.t_rows = (const int *)test->glay[kt].rows.get();
		a.s_order = (const int *)trial->glay[ks].order.get();
		a.s_col = (const int *)trial->glay[ks].col.get();
		a.s_fin = (const int *)trial->glay[ks].fin.get();
	} else {
		a.n_tiles = (test->n + nbk::GAL_TILE - 1) / nbk::GAL_TILE;
		a.chunk = gal_chunk(a.n_tiles, trial->n);
	}
	a.tn = test->n; a.sn = trial->n; a.same_mesh = jb.same;
	a.t_quad = (const int *)test->d_tri.get();
	a.s_quad = (const int *)trial->d_tri.get();
	a.tc = (const double *)test->d_c.get();
	a.sc = (const double *)trial->d_c.get();
	a.thr = (const double *)trial->d_thr.get();
	for (int N = 1; N <= nbk::GALQ_NMAX; ++N) {
		a.tq[N] = (const float *)test->d_gq[N].get();
		a.sq[N] = (const float *)trial->d_gq[N].get();
	}
	a.fixed_n = fixed_n;
	a.complexity = complexity;
	const double two_pi = 6.283185307179586476925286766559005768;
	a.k = (float)jb.k;
	a.k_rev = jb.k / two_pi;
	a.k_revf = (float)a.k_rev;
	a.ki = (float)jb.ki;
	a.alpha = make_float2((float)(jb.alpha.re / four_pi), (float)(jb.alpha.im / four_pi));
	a.out = out; a.ld = ld; a.dtype = dtype; a.accumulate = accumulate;
	const long long grid = (long long)((a.n_tiles + 3) / 4) * ((trial->n + a.chunk - 1) / a.chunk);
	{
		using namespace nbk;
		const int pe = prof_begin(0, st);
#define NB_GQ(F, K) do { \
		if (F == HELM && ck) { \
			if (nodal) hipLaunchKernelGGL((galq_regular_kernel<F, K, 4, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
			else hipLaunchKernelGGL((galq_regular_kernel<F, K, 1, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
		} else if (nodal) hipLaunchKernelGGL((galq_regular_kernel<F, K, 4>), dim3(grid), dim3(256), 0, st, a); \
		else hipLaunchKernelGGL((galq_regular_kernel<F, K, 1>), dim3(grid), dim3(256), 0, st, a); } while (0)
		if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GQ(HELM, SLP); break;
			case DLP: NB_GQ(HELM, DLP); break;
			case DLPT: NB_GQ(HELM, DLPT); break;
			default: NB_GQ(HELM, HSP); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GQ(LAP, SLP); break;
			case DLP: NB_GQ(LAP, DLP); break;
			case DLPT: NB_GQ(LAP, DLPT); break;
			default: NB_GQ(LAP, HSP); break;
			}
		}
#undef NB_GQ
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	if (!jb.same) return NB_OK;

	// node-sharing pairs in FP64 (the double layer of a coincident
	// quadrilateral is integrated: it vanishes only on flat elements)
	const SingRules *sr = nullptr;
	if ((s = device_quad_sing_rules(&sr))) return s;
	nbk::GalQSingArgs g;
	std::memset(&g, 0, sizeof(g));
	g.n_pairs = (int)trial->sp_n;
	if (g.n_pairs <= 0) return NB_OK;
	g.pe = (const int *)trial->sp_e.get();
	g.pf = (const int *)trial->sp_f.get();
	g.pinfo = (const int *)trial->sp_info.get();
	g.n = trial->n;
	g.X = (const double *)trial->d_X.get();
	g.t_dof = t_map;
	g.s_dof = s_map;
	for (int m = 0; m < 3; ++m) { g.rule[m] = (const double *)sr->r[m].get(); g.rule_n[m] = sr->n[m]; }
	g.k = jb.k;
	g.ki = jb.ki;
	g.alpha = make_double2(jb.alpha.re, jb.alpha.im);
	g.out = out; g.ld = ld; g.dtype = dtype;
	{
		using namespace nbk;
		const int pe = prof_begin(2, st);
#define NB_GQS(F, K) do { \
		if (F == HELM && ck) { \
			if (nodal) hipLaunchKernelGGL((galq_singular_kernel<F, K, 4, (F == HELM)>), dim3(g.n_pairs), dim3(64), 0, st, g); \
			else hipLaunchKernelGGL((galq_singular_kernel<F, K, 1, (F == HELM)>), dim3(g.n_pairs), dim3(64), 0, st, g); \
		} else if (nodal) hipLaunchKernelGGL((galq_singular_kernel<F, K, 4>), dim3(g.n_pairs), dim3(64), 0, st, g); \
		else hipLaunchKernelGGL((galq_singular_kernel<F, K, 1>), dim3(g.n_pairs), dim3(64), 0, st, g); } while (0)
		if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GQS(HELM, SLP); break;
			case DLP: NB_GQS(HELM, DLP); break;
			default: NB_GQS(HELM, DLPT); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GQS(LAP, SLP); break;
			case DLP: NB_GQS(LAP, DLP); break;
			default: NB_GQS(LAP, DLPT); break;
			}
		}
#undef NB_GQS
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	return NB_OK;
}

// Galerkin assembly into a device buffer
int run_galerkin(const Job &jb, void *out, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	if (jb.pol.precision != NB_PREC_MIXED)
		return fail(NB_ERR_UNSUPPORTED, "the FP64 precision policy is implemented for collocation");
	if (jb.trial->npe == 4) return run_galerkin_quad(jb, out, ld, dtype, accumulate, st);
	nb_mesh *trial = jb.trial;
	nb_mesh *test = const_cast<nb_mesh *>(jb.test);
	// the nodal kernels serve every pair with an isoparametric side; a
	// constant side runs through its element dof map
	const bool lt = jb.formalism == NB_GALERKIN_LINEAR || jb.formalism == NB_GALERKIN_LINEAR_CONST;
	const bool ls = jb.formalism == NB_GALERKIN_LINEAR || jb.formalism == NB_GALERKIN_CONST_LINEAR;
	const bool nodal = lt || ls;
	const bool ck = jb.ki != 0.0 && jb.family == NB_HELMHOLTZ;
	const int64_t rows = lt ? test->n_nodes : test->n, cols = ls ? trial->n_nodes : trial->n;
	const int eb = dtype_bytes(dtype);
	auto zero = [&]() -> int {
		launch_zero(out, ld, rows, cols, eb, st);
		NB_HIP(hipGetLastError());
		return NB_OK;
	};
	int s;
	if ((!lt && nodal && (s = ensure_edof(test))) || (!ls && nodal && (s = ensure_edof(trial)))) return s;
	const int *t_map = (const int *)(lt ? test->d_tri.get() : test->d_edof.get());
	const int *s_map = (const int *)(ls ? trial->d_tri.get() : trial->d_edof.get());
	if (jb.family == NB_IDENTITY) {
		if (!accumulate && (s = zero())) return s;
		// local operator: nothing on different meshes (nihu/core/assembly.hpp:141-142)
		if (jb.same)
			hipLaunchKernelGGL(nbk::galerkin_identity_kernel, dim3((test->n + 255) / 256), dim3(256), 0, st, out,
				(long long)ld, test->n, (const double *)test->d_J.get(), t_map, s_map, nodal ? 1 : 0,
				dtype, jb.alpha.re, jb.alpha.im);
		NB_HIP(hipGetLastError());
		return NB_OK;
	}
	// 1/r^3 has no valid singular rule on coincident and edge-adjacent pairs
	// (nihu/core/singular_accelerator.hpp:483-503, double_integral.hpp:255-273)
	if (jb.same && jb.km == nbk::HSP) return fail(NB_ERR_RUNTIME, "Invalid general singular quadrature");
	if ((s = ensure_thresholds(trial, jb.pol.near_mode, jb.pol.near_limit))) return s;
	std::vector<int> degs;
	policy_degrees(jb.pol, jb.km, nodal ? 1 : 0, degs);
	for (int d : degs)
		if ((s = ensure_quad(trial, d)) || (s = ensure_quad(test, d))) return s;
	const int kt = lt ? 0 : 1, ks = ls ? 0 : 1;   // layout kinds: node or element dofs
	if (nodal && ((s = ensure_gal_layout(test, kt)) || (s = ensure_gal_layout(trial, ks)))) return s;
	if (jb.same && (s = ensure_sing_pairs(trial))) return s;
	if (nodal && !accumulate && (s = zero())) return s;

	const nbk::Rules *rules = nullptr;
	if ((s = device_rules(&rules))) return s;
	nbk::GalArgs a;
	std::memset(&a, 0, sizeof(a));
	a.ref = &rules->ref[0][0];     // device address (no host access)
	if (nodal) {
		a.n_tiles = test->glay[kt].tiles;
		a.chunk = trial->glay[ks].chunk;
		a.t_elem = (const int *)test->glay[kt].elem.get();
		a.t_contrib = (const uint2 *)test->glay[kt].contrib.get();
		a.t_rows = (const int *)test->glay[kt].rows.get();
		a.s_order = (const int *)trial->glay[ks].order.get();
		a.s_col = (const int *)trial->glay[ks].col.get();
		a.s_fin = (const int *)trial->glay[ks].fin.get();
	} else {
		a.n_tiles = (test->n + nbk::GAL_TILE - 1) / nbk::GAL_TILE;
		a.chunk = gal_chunk(a.n_tiles, trial->n);
	}
	a.tn = test->n; a.sn = trial->n; a.same_mesh = jb.same;
	a.t_tri = (const int *)test->d_tri.get();
	a.s_tri = (const int *)trial->d_tri.get();
	a.tc = (const double *)test->d_c.get();
	a.sc = (const double *)trial->d_c.get();
	a.tnu = (const float *)test->d_nuf.get();
	a.snu = (const float *)trial->d_nuf.get();
	a.thr = (const double *)trial->d_thr.get();
	for (int d = 0; d < nbk::NSLOT; ++d) {
		a.tq[d] = (const nbk::QPoint *)test->d_qt[d].get();
		a.sq[d] = (const nbk::QPoint *)trial->d_qt[d].get();
		a.qn[d] = trial->qn[d];
	}
	a.fixed_degree = jb.pol.rule == NB_RULE_FIXED ? jb.pol.fixed_degree : -1;
	const double two_pi = 6.283185307179586476925286766559005768;
	a.k = (float)jb.k;
	a.k_rev = jb.k / two_pi;
	a.k_revf = (float)a.k_rev;
	a.ki = (float)jb.ki;
	a.alpha = make_float2((float)jb.alpha.re, (float)jb.alpha.im);
	a.out = out; a.ld = ld; a.dtype = dtype; a.accumulate = accumulate;
	const long long grid = (long long)((a.n_tiles + 3) / 4) * ((trial->n + a.chunk - 1) / a.chunk);
	const bool fix6 = jb.pol.rule == NB_RULE_FIXED && trial->qn[jb.pol.fixed_degree] == 6;
	{
		using namespace nbk;
		const int pe = prof_begin(0, st);
#define NB_GAL(F, K) do { \
		if (F == HELM && ck) { \
			if (nodal) hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 3, 0, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
			else hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 1, 0, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
		} else if (nodal) { \
			if (F == HELM && fix6) hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 3, (F == HELM ? 6 : 0)>), dim3(grid), dim3(256), 0, st, a); \
			else hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 3, 0>), dim3(grid), dim3(256), 0, st, a); \
		} else { \
			if (F == HELM && fix6) hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 1, (F == HELM ? 6 : 0)>), dim3(grid), dim3(256), 0, st, a); \
			else hipLaunchKernelGGL((galerkin_regular_kernel<F, K, 1, 0>), dim3(grid), dim3(256), 0, st, a); \
		} } while (0)
		if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GAL(HELM, SLP); break;
			case DLP: NB_GAL(HELM, DLP); break;
			case DLPT: NB_GAL(HELM, DLPT); break;
			default: NB_GAL(HELM, HSP); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GAL(LAP, SLP); break;
			case DLP: NB_GAL(LAP, DLP); break;
			case DLPT: NB_GAL(LAP, DLPT); break;
			default: NB_GAL(LAP, HSP); break;
			}
		}
#undef NB_GAL
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	if (!jb.same) return NB_OK;

	// singular pairs in FP64; the double layer of a coincident flat pair is 0
	// (nihu/library/normal_derivative_singular_integrals.hpp:13-58)
	const SingRules *sr = nullptr;
	if ((s = device_sing_rules(&sr))) return s;
	nbk::GalSingArgs g;
	std::memset(&g, 0, sizeof(g));
	const int64_t first = (jb.km == nbk::DLP || jb.km == nbk::DLPT) ? trial->n : 0;
	g.n_pairs = (int)(trial->sp_n - first);
	if (g.n_pairs <= 0) return NB_OK;
	g.pe = (const int *)trial->sp_e.get() + first;
	g.pf = (const int *)trial->sp_f.get() + first;
	g.pinfo = (const int *)trial->sp_info.get() + first;
	g.n = trial->n;
	g.X = (const double *)trial->d_X.get();
	g.J = (const double *)trial->d_J.get();
	g.nu = (const double *)trial->d_nu64.get();
	g.t_dof = t_map;
	g.s_dof = s_map;
	for (int m = 0; m < 3; ++m) { g.rule[m] = (const double *)sr->r[m].get(); g.rule_n[m] = sr->n[m]; }
	g.k = jb.k;
	g.ki = jb.ki;
	g.alpha = make_double2(jb.alpha.re, jb.alpha.im);
	g.out = out; g.ld = ld; g.dtype = dtype;
	{
		using namespace nbk;
		const int pe = prof_begin(2, st);
#define NB_GSING(F, K) do { \
		if (F == HELM && ck) { \
			if (nodal) hipLaunchKernelGGL((galerkin_singular_kernel<F, K, 3, (F == HELM)>), dim3(g.n_pairs), dim3(64), 0, st, g); \
			else hipLaunchKernelGGL((galerkin_singular_kernel<F, K, 1, (F == HELM)>), dim3(g.n_pairs), dim3(64), 0, st, g); \
		} else if (nodal) hipLaunchKernelGGL((galerkin_singular_kernel<F, K, 3>), dim3(g.n_pairs), dim3(64), 0, st, g); \
		else hipLaunchKernelGGL((galerkin_singular_kernel<F, K, 1>), dim3(g.n_pairs), dim3(64), 0, st, g); } while (0)
		if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GSING(HELM, SLP); break;
			case DLP: NB_GSING(HELM, DLP); break;
			default: NB_GSING(HELM, DLPT); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GSING(LAP, SLP); break;
			case DLP: NB_GSING(LAP, DLP); break;
			default: NB_GSING(LAP, DLPT); break;
			}
		}
#undef NB_GSING
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	return NB_OK;
}

// one assembly into device buffers (dtype, ld, accumulate taken from outs)
// Collocation on a quadrilateral trial mesh: the regular pass (FP32) and the
// near-singular and diagonal passes (FP64) on a side stream, disjoint entries.
// The near-singular methods of quadrilaterals start from the image of the
// collocation point on the element (Newton from the centre), which is only
// reliable close to the element: NiHu applies them for d / R < 1.5.
int check_quad_policy(const nb_policy &p)
{
	if (p.near_mode == NB_NEAR_DIAMETER || (p.near_mode == NB_NEAR_MEAN_EDGE && p.near_limit > 1.5))
		return fail(NB_ERR_UNSUPPORTED, "quadrilaterals: the near-singular rule must be mean edge with limit <= 1.5 "
			"(NiHu's d / R < 1.5) or none");
	return NB_OK;
}

// Newton step limit of the quadrilateral inverse mapping: NiHu's 100
// (nihu/core/inverse_mapping.hpp:37-118).  NBEM_INVERSE_MAPPING_STEPS
// lowers it -- a test hook that makes the mapping fail on purpose, so that
// the error path runs with every kernel finishing normally.
int inverse_mapping_steps()
{
	const char *e = std::getenv("NBEM_INVERSE_MAPPING_STEPS");
	if (!e || !e[0]) return 100;
	const int v = std::atoi(e);
	return v < 0 ? 0 : (v > 100 ? 100 : v);
}

int run_quad(const Job &jb, void *out0, void *out1, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	nb_mesh *trial = jb.trial;
	const nb_mesh *test = jb.test;
	int s;
	// the FP64 precision policy: the regular pairs in FP64 too
	const bool fp64 = jb.pol.precision == NB_PREC_FP64;
	if ((s = check_quad_policy(jb.pol))) return s;
	if ((s = ensure_thresholds(trial, jb.pol.near_mode, jb.pol.near_limit))) return s;
	int fixed_n = 0;
	if (jb.pol.rule == NB_RULE_FIXED) {
		fixed_n = jb.pol.fixed_degree / 2 + 1;
		if ((s = fp64 ? ensure_qq64(trial, fixed_n) : ensure_qq(trial, fixed_n))) return s;
	} else {
		// degree 1 + {0, 2, 4, 6, 8, 9}: Gauss orders 1..6
		for (int N = 1; N <= 6; ++N)
			if ((s = fp64 ? ensure_qq64(trial, N) : ensure_qq(trial, N))) return s;
	}
	std::shared_ptr<NearList> nl;
	if ((jb.passes & 2) && (s = ensure_near_list(trial, test, jb.same, jb.pol, st, &nl))) return s;
	nbk::QuadArgs a;
	std::memset(&a, 0, sizeof(a));
	a.rmap = jb.rmap; a.cmap = jb.cmap;
	a.n_rows = test->n; a.n_cols = trial->n; a.same_mesh = jb.same;
	a.tile_cols = jb.tile_cols > 0 ? jb.tile_cols : 32;
	a.row_tiles = (test->n + 255) / 256;
	a.tn = test->n; a.sn = trial->n;
	a.tc = (const double *)test->d_c.get();
	a.tnu = (const float *)test->d_nuf.get();
	a.rec = (const nbk::TrialRec *)trial->d_rec.get();
	for (int N = 1; N <= nbk::QN_MAX; ++N) a.qt[N] = (const nbk::QuadPt *)trial->d_qq[N].get();
	a.fixed_n = fixed_n;
	const double two_pi = 6.283185307179586476925286766559005768;
	a.k = (float)jb.k;
	a.k_rev = jb.k / two_pi;
	a.k_revf = (float)a.k_rev;
	a.ki = jb.ki;
	a.alpha = make_float2((float)jb.alpha.re, (float)jb.alpha.im);
	a.out0 = out0; a.out1 = out1; a.ld = ld; a.dtype = dtype; a.accumulate = accumulate;
	const long long grid = (long long)a.row_tiles * ((trial->n + a.tile_cols - 1) / a.tile_cols);
	// a complex wavenumber (Helmholtz only: the Laplace job carries ki = 0)
	const bool ck = jb.ki != 0.0;

	hipStream_t side = st;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr;
	const bool has_fp64 = (jb.passes & 2) && (nl->n > 0 || jb.same);
	if (has_fp64) {
		if ((s = side_stream(&side))) return s;
		if (!side) side = st;
		NB_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
		NB_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
		NB_HIP(hipEventRecord(ev_fork, st));
		NB_HIP(hipStreamWaitEvent(side, ev_fork, 0));
	}
	if (grid > 0 && (jb.passes & 1) && fp64) {
		using namespace nbk;
		const int pe = prof_begin(0, st);
		Quad64Args b;
		std::memset(&b, 0, sizeof(b));
		b.n_rows = a.n_rows; b.n_cols = a.n_cols; b.same_mesh = a.same_mesh;
		b.row_tiles = a.row_tiles; b.tile_cols = a.tile_cols;
		b.rmap = jb.rmap; b.cmap = jb.cmap;
		b.tn = test->n; b.sn = trial->n;
		b.tc = (const double *)test->d_c.get();
		b.tnu = (const double *)test->d_nu64.get();
		b.sc = (const double *)trial->d_c.get();
		b.thr = (const double *)trial->d_thr.get();
		for (int N = 1; N <= QN_MAX; ++N) b.qt[N] = (const double *)trial->d_qq64[N].get();
		b.fixed_n = fixed_n;
		b.k = jb.k; b.ki = jb.ki;
		b.alpha = make_double2(jb.alpha.re, jb.alpha.im);
		b.out0 = out0; b.out1 = out1; b.ld = ld; b.dtype = dtype; b.accumulate = accumulate;
#define NB_Q64R(F, K, M) do { \
		if (F == HELM && ck) hipLaunchKernelGGL((quad_regular64_kernel<F, K, M, (F == HELM)>), dim3(grid), dim3(256), 0, st, b); \
		else hipLaunchKernelGGL((quad_regular64_kernel<F, K, M>), dim3(grid), dim3(256), 0, st, b); } while (0)
		if (jb.mode == OUT_BM) NB_Q64R(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_Q64R(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_Q64R(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_Q64R(HELM, DLPT, OUT_SINGLE); break;
			default: NB_Q64R(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_Q64R(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_Q64R(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_Q64R(LAP, DLPT, OUT_SINGLE); break;
			default: NB_Q64R(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_Q64R
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	} else if (grid > 0 && (jb.passes & 1)) {
		using namespace nbk;
		const int pe = prof_begin(0, st);
		const bool mapped = jb.rmap || jb.cmap;
#define NB_QR(F, K, M) do { \
		if (mapped) hipLaunchKernelGGL((quad_regular_kernel<F, K, M, true>), dim3(grid), dim3(256), 0, st, a); \
		else hipLaunchKernelGGL((quad_regular_kernel<F, K, M, false>), dim3(grid), dim3(256), 0, st, a); } while (0)
#define NB_QRC(K, M) do { \
		if (mapped) hipLaunchKernelGGL((quad_regular_kernel<HELM, K, M, true, true>), dim3(grid), dim3(256), 0, st, a); \
		else hipLaunchKernelGGL((quad_regular_kernel<HELM, K, M, false, true>), dim3(grid), dim3(256), 0, st, a); \
		} while (0)
		if (ck) {
			if (jb.mode == OUT_BM) NB_QRC(ALL4, OUT_BM);
			else switch (jb.km) {
			case SLP: NB_QRC(SLP, OUT_SINGLE); break;
			case DLP: NB_QRC(DLP, OUT_SINGLE); break;
			case DLPT: NB_QRC(DLPT, OUT_SINGLE); break;
			default: NB_QRC(HSP, OUT_SINGLE); break;
			}
		} else if (jb.mode == OUT_BM) NB_QR(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_QR(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_QR(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_QR(HELM, DLPT, OUT_SINGLE); break;
			default: NB_QR(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_QR(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_QR(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_QR(LAP, DLPT, OUT_SINGLE); break;
			default: NB_QR(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_QR
#undef NB_QRC
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	if (!(jb.passes & 2)) return NB_OK;
	const nbk::Rules *rules = nullptr;
	if ((s = device_rules(&rules))) return s;
	nbk::NearArgs na;
	std::memset(&na, 0, sizeof(na));
	na.rmap = jb.rmap; na.cmap = jb.cmap;
	na.tn = test->n; na.sn = trial->n;
	na.tc = (const double *)test->d_c.get();
	na.tnu = (const double *)test->d_nu64.get();
	na.X = (const double *)trial->d_X.get();
	na.sJ = (const double *)trial->d_J.get();
	na.snu = (const double *)trial->d_nu64.get();
	na.k = jb.k;
	na.ki = jb.ki;
	na.alpha = make_double2(jb.alpha.re, jb.alpha.im);
	na.out0 = out0; na.out1 = out1; na.ld = ld; na.dtype = dtype; na.accumulate = accumulate;
	na.rules = rules;
	// the near pass counts pairs whose collocation point has no image on the
	// element (NiHu throws "Could not perform inverse mapping"); the count lives
	// with the trial mesh, whose lock this call holds until it has read it
	const bool check_images = nl->n > 0;
	na.inv_iter = inverse_mapping_steps();
	if (check_images) {
		if (!trial->d_status && (s = dev_alloc(trial->d_status, 256))) return s;
		na.status = (int *)trial->d_status.get();
		launch_zero(na.status, 1, 1, 1, 4, side);
		NB_HIP(hipGetLastError());
	}
	for (int pass = 0; pass < 2; ++pass) {
		const bool diag = pass == 1;
		if (diag && !jb.same) continue;
		na.n_pairs = diag ? test->n : nl->n;
		na.pi = diag ? nullptr : (const int *)nl->pi.get();
		na.pj = diag ? nullptr : (const int *)nl->pj.get();
		if (na.n_pairs == 0) continue;
		// diagonal: one lane per element; near: QLPP lanes per pair
		const long long lanes = (long long)na.n_pairs * (diag ? 1 : nbk::QLPP);
		const int g = (int)((lanes + 63) / 64);
		const int pe = prof_begin(diag ? 2 : 1, side);
		using namespace nbk;
#define NB_Q64(F, K, M) do { \
		if (diag) hipLaunchKernelGGL((quad_diag_kernel<F, K, M>), dim3(g), dim3(64), 0, side, na); \
		else hipLaunchKernelGGL((quad_near_kernel<F, K, M>), dim3(g), dim3(64), 0, side, na); } while (0)
#define NB_Q64C(K, M) do { \
		if (diag) hipLaunchKernelGGL((quad_diag_kernel<HELM, K, M, true>), dim3(g), dim3(64), 0, side, na); \
		else hipLaunchKernelGGL((quad_near_kernel<HELM, K, M, true>), dim3(g), dim3(64), 0, side, na); } while (0)
		if (ck) {
			if (jb.mode == OUT_BM) NB_Q64C(ALL4, OUT_BM);
			else switch (jb.km) {
			case SLP: NB_Q64C(SLP, OUT_SINGLE); break;
			case DLP: NB_Q64C(DLP, OUT_SINGLE); break;
			case DLPT: NB_Q64C(DLPT, OUT_SINGLE); break;
			default: NB_Q64C(HSP, OUT_SINGLE); break;
			}
		} else if (jb.mode == OUT_BM) NB_Q64(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_Q64(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_Q64(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_Q64(HELM, DLPT, OUT_SINGLE); break;
			default: NB_Q64(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_Q64(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_Q64(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_Q64(LAP, DLPT, OUT_SINGLE); break;
			default: NB_Q64(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_Q64
#undef NB_Q64C
		NB_HIP(hipGetLastError());
		prof_end(pe, side);
	}
	if (has_fp64) {
		NB_HIP(hipEventRecord(ev_join, side));
		NB_HIP(hipStreamWaitEvent(st, ev_join, 0));
		NB_HIP(hipEventDestroy(ev_fork));
		NB_HIP(hipEventDestroy(ev_join));
	}
	if (check_images) {
		int failed = 0;
		NB_HIP(hipMemcpyAsync(&failed, na.status, sizeof(int), hipMemcpyDeviceToHost, st));
		NB_HIP(hipStreamSynchronize(st));
		if (failed)
			return fail(NB_ERR_RUNTIME, "Could not perform inverse mapping: " + std::to_string(failed) +
				" near-singular quadrilateral pair(s) (their entries hold NaN)");
	}
	return NB_OK;
}


// Collocation with the quadrilateral gauss field on the trial mesh (4 dofs
// per element, column 4 j + b): the regular pass (FP32), then the
// near-singular and diagonal passes (FP64), disjoint entries.  Test side:
// NB_COLLOCATION_GAUSS the dirac of the test mesh's gauss field (rows 4 i + a),
// NB_COLLOCATION_GAUSS_TRIAL the dirac of its constant field (rows = elements).
int run_gauss(const Job &jb, void *out0, void *out1, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	nb_mesh *trial = jb.trial;
	nb_mesh *test = const_cast<nb_mesh *>(jb.test);
	const bool tg = jb.formalism == NB_COLLOCATION_GAUSS;
	int s;
	if (trial->npe != 4) return fail(NB_ERR_UNSUPPORTED, "the gauss field lives on quadrilateral meshes");
	if (tg && test->npe != 4) return fail(NB_ERR_UNSUPPORTED, "the gauss test field needs a quadrilateral test mesh");
	if (!tg && test->npe == 0)
		return fail(NB_ERR_UNSUPPORTED, "field points of the gauss field: a test mesh of one element type");
	if (!tg && jb.same)
		return fail(NB_ERR_UNSUPPORTED, "a constant test field against the gauss field is taken on another mesh only");
	// the FP64 precision policy: the regular blocks in FP64 too
	const bool fp64 = jb.pol.precision == NB_PREC_FP64;
	if ((s = check_quad_policy(jb.pol))) return s;
	if ((s = ensure_thresholds(trial, jb.pol.near_mode, jb.pol.near_limit))) return s;
	int fixed_n = 0;
	if (jb.pol.rule == NB_RULE_FIXED) {
		fixed_n = jb.pol.fixed_degree / 2 + 1;
		if ((s = fp64 ? ensure_qq64(trial, fixed_n) : ensure_qq(trial, fixed_n))) return s;
	} else {
		// degree 2 + {0, 2, 4, 6, 8, 9}: Gauss orders 2..6
		for (int N = 2; N <= 6; ++N)
			if ((s = fp64 ? ensure_qq64(trial, N) : ensure_qq(trial, N))) return s;
	}
	if (tg && (s = ensure_gauss_nodes(test))) return s;
	const nbk::GaussShapes *gs = nullptr;
	if ((s = device_gauss_shapes(&gs))) return s;
	std::shared_ptr<NearList> nl;
	if ((s = ensure_near_list(trial, test, jb.same, jb.pol, st, &nl))) return s;
	const int nodes = tg ? 4 : 1;
	nbk::GaussArgs a;
	std::memset(&a, 0, sizeof(a));
	a.n_rows = nodes * test->n; a.n_cols = trial->n; a.same_mesh = jb.same; a.nodes = nodes;
	a.tile_cols = jb.tile_cols > 0 ? jb.tile_cols : 32;
	a.row_tiles = (a.n_rows + 255) / 256;
	a.tn = a.n_rows; a.te = test->n;
	a.tx = (const double *)(tg ? test->d_gx.get() : test->d_c.get());
	a.tnu = (const float *)(tg ? test->d_gnuf.get() : test->d_nuf.get());
	a.tc = (const double *)test->d_c.get();
	a.rec = (const nbk::TrialRec *)trial->d_rec.get();
	for (int N = 1; N <= nbk::QN_MAX; ++N) a.qt[N] = (const nbk::QuadPt *)trial->d_qq[N].get();
	a.gs = gs;
	a.fixed_n = fixed_n;
	const double two_pi = 6.283185307179586476925286766559005768;
	a.k = (float)jb.k;
	a.k_rev = jb.k / two_pi;
	a.k_revf = (float)a.k_rev;
	a.ki = jb.ki;
	a.alpha = make_float2((float)jb.alpha.re, (float)jb.alpha.im);
	a.out0 = out0; a.out1 = out1; a.ld = ld; a.dtype = dtype; a.accumulate = accumulate;
	const long long grid = (long long)a.row_tiles * ((trial->n + a.tile_cols - 1) / a.tile_cols);
	const bool ck = jb.ki != 0.0;
	if (grid > 0 && fp64) {
		using namespace nbk;
		const int pe = prof_begin(0, st);
		Gauss64Args b;
		std::memset(&b, 0, sizeof(b));
		b.n_rows = a.n_rows; b.n_cols = a.n_cols; b.same_mesh = a.same_mesh;
		b.row_tiles = a.row_tiles; b.tile_cols = a.tile_cols; b.nodes = nodes;
		b.tn = a.tn; b.te = a.te;
		b.tx = a.tx;
		b.tnu = (const double *)(tg ? test->d_gnu64.get() : test->d_nu64.get());
		b.tc = a.tc;
		b.sc = (const double *)trial->d_c.get();
		b.sn = trial->n;
		b.thr = (const double *)trial->d_thr.get();
		for (int N = 1; N <= QN_MAX; ++N) b.qt[N] = (const double *)trial->d_qq64[N].get();
		b.gs = gs;
		b.fixed_n = fixed_n;
		b.k = jb.k; b.ki = jb.ki;
		b.alpha = make_double2(jb.alpha.re, jb.alpha.im);
		b.out0 = out0; b.out1 = out1; b.ld = ld; b.dtype = dtype; b.accumulate = accumulate;
#define NB_GR64(F, K, M) do { \
		if (F == HELM && ck) hipLaunchKernelGGL((gauss_regular64_kernel<F, K, M, (F == HELM)>), dim3(grid), dim3(256), 0, st, b); \
		else hipLaunchKernelGGL((gauss_regular64_kernel<F, K, M>), dim3(grid), dim3(256), 0, st, b); } while (0)
		if (jb.mode == OUT_BM) NB_GR64(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GR64(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_GR64(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_GR64(HELM, DLPT, OUT_SINGLE); break;
			default: NB_GR64(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GR64(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_GR64(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_GR64(LAP, DLPT, OUT_SINGLE); break;
			default: NB_GR64(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_GR64
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	} else if (grid > 0) {
		using namespace nbk;
		const int pe = prof_begin(0, st);
#define NB_GR(F, K, M) do { \
		if (F == HELM && ck) hipLaunchKernelGGL((gauss_regular_kernel<F, K, M, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
		else hipLaunchKernelGGL((gauss_regular_kernel<F, K, M>), dim3(grid), dim3(256), 0, st, a); } while (0)
		if (jb.mode == OUT_BM) NB_GR(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_GR(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_GR(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_GR(HELM, DLPT, OUT_SINGLE); break;
			default: NB_GR(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_GR(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_GR(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_GR(LAP, DLPT, OUT_SINGLE); break;
			default: NB_GR(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_GR
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	const nbk::Rules *rules = nullptr;
	if ((s = device_rules(&rules))) return s;
	nbk::GaussNearArgs na;
	std::memset(&na, 0, sizeof(na));
	na.nodes = nodes;
	na.tn = a.n_rows;
	na.tx = (const double *)(tg ? test->d_gx.get() : test->d_c.get());
	na.tnu = (const double *)(tg ? test->d_gnu64.get() : test->d_nu64.get());
	na.sn = trial->n;
	na.X = (const double *)trial->d_X.get();
	na.k = jb.k;
	na.ki = jb.ki;
	na.alpha = make_double2(jb.alpha.re, jb.alpha.im);
	na.out0 = out0; na.out1 = out1; na.ld = ld; na.dtype = dtype; na.accumulate = accumulate;
	na.rules = rules;
	na.inv_iter = inverse_mapping_steps();
	const bool check_images = nl->n > 0;
	if (check_images) {
		if (!trial->d_status && (s = dev_alloc(trial->d_status, 256))) return s;
		na.status = (int *)trial->d_status.get();
		launch_zero(na.status, 1, 1, 1, 4, st);
		NB_HIP(hipGetLastError());
	}
	for (int pass = 0; pass < 2; ++pass) {
		const bool diag = pass == 1;
		if (diag && !jb.same) continue;
		na.n_pairs = diag ? test->n : nl->n;
		na.pi = diag ? nullptr : (const int *)nl->pi.get();
		na.pj = diag ? nullptr : (const int *)nl->pj.get();
		if (na.n_pairs == 0) continue;
		// diagonal: one lane per (element, node); near: QLPP lanes per (pair, node)
		const long long lanes = (long long)na.n_pairs * nodes * (diag ? 1 : nbk::QLPP);
		const int g = (int)((lanes + 63) / 64);
		const int pe = prof_begin(diag ? 2 : 1, st);
		using namespace nbk;
#define NB_G64(F, K, M) do { \
		if (F == HELM && ck) { \
			if (diag) hipLaunchKernelGGL((gauss_diag_kernel<F, K, M, (F == HELM)>), dim3(g), dim3(64), 0, st, na); \
			else hipLaunchKernelGGL((gauss_near_kernel<F, K, M, (F == HELM)>), dim3(g), dim3(64), 0, st, na); \
		} else if (diag) hipLaunchKernelGGL((gauss_diag_kernel<F, K, M>), dim3(g), dim3(64), 0, st, na); \
		else hipLaunchKernelGGL((gauss_near_kernel<F, K, M>), dim3(g), dim3(64), 0, st, na); } while (0)
		if (jb.mode == OUT_BM) NB_G64(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_G64(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_G64(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_G64(HELM, DLPT, OUT_SINGLE); break;
			default: NB_G64(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_G64(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_G64(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_G64(LAP, DLPT, OUT_SINGLE); break;
			default: NB_G64(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_G64
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}
	if (check_images) {
		int failed = 0;
		NB_HIP(hipMemcpyAsync(&failed, na.status, sizeof(int), hipMemcpyDeviceToHost, st));
		NB_HIP(hipStreamSynchronize(st));
		if (failed)
			return fail(NB_ERR_RUNTIME, "Could not perform inverse mapping: " + std::to_string(failed) +
				" near-singular quadrilateral pair(s) (their entries hold NaN)");
	}
	return NB_OK;
}

int run_device(const Job &jb, void *out0, void *out1, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	if (jb.formalism != NB_COLLOCATION) return run_galerkin(jb, out0, ld, dtype, accumulate, st);
	if (jb.trial->npe == 4) return run_quad(jb, out0, out1, ld, dtype, accumulate, st);
	nb_mesh *trial = jb.trial;
	const nb_mesh *test = jb.test;
	int s;
	if ((s = ensure_thresholds(trial, jb.pol.near_mode, jb.pol.near_limit))) return s;
	std::vector<int> degs;
	policy_degrees(jb.pol, jb.km, 0, degs);
	const bool fp64 = jb.pol.precision == NB_PREC_FP64;
	for (int d : degs)
		if ((s = fp64 ? ensure_quad64(trial, d) : ensure_quad(trial, d))) return s;
	std::shared_ptr<NearList> nl;
	if ((jb.passes & 2) && (s = ensure_near_list(trial, test, jb.same, jb.pol, st, &nl))) return s;

	// 1. regular pairs
	nbk::RegArgs a;
	std::memset(&a, 0, sizeof(a));
	a.rmap = jb.rmap; a.cmap = jb.cmap;
	a.n_rows = test->n; a.n_cols = trial->n; a.same_mesh = jb.same;
	a.tile_cols = jb.tile_cols > 0 ? jb.tile_cols : 32;
	a.row_tiles = (test->n + 255) / 256;
	a.tn = test->n; a.sn = trial->n;
	a.tc = (const double *)test->d_c.get();
	a.tnu = (const float *)test->d_nuf.get();
	a.sc = (const double *)trial->d_c.get();
	a.snu = (const float *)trial->d_nuf.get();
	a.thr = (const double *)trial->d_thr.get();
	a.rec = (const nbk::TrialRec *)trial->d_rec.get();
	for (int d = 0; d < nbk::NSLOT; ++d) { a.qt[d] = (const nbk::QPoint *)trial->d_qt[d].get(); a.qn[d] = trial->qn[d]; }
	a.fixed_degree = jb.pol.rule == NB_RULE_FIXED ? jb.pol.fixed_degree : -1;
	a.field_complexity = 0;
	const double two_pi = 6.283185307179586476925286766559005768;
	a.k = (float)jb.k;
	a.k_rev = jb.k / two_pi;
	a.k_revf = (float)a.k_rev;
	a.ki = (float)jb.ki;
	const bool ck = jb.ki != 0.0 && jb.family == NB_HELMHOLTZ;
	a.alpha = make_float2((float)jb.alpha.re, (float)jb.alpha.im);
	a.out0 = out0; a.out1 = out1; a.ld = ld; a.dtype = dtype; a.accumulate = accumulate;
	const long long col_tiles = (trial->n + a.tile_cols - 1) / a.tile_cols;
	const long long grid = (long long)a.row_tiles * col_tiles;

	// fork the FP64 passes onto the side stream
	hipStream_t side = st;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr;
	const bool has_fp64 = (jb.passes & 2) && (nl->n > 0 || jb.same);
	if (has_fp64) {
		if ((s = side_stream(&side))) return s;
		if (!side) side = st;
		NB_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
		NB_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
		NB_HIP(hipEventRecord(ev_fork, st));
		NB_HIP(hipStreamWaitEvent(side, ev_fork, 0));
	}
	if (grid > 0 && (jb.passes & 1) && fp64) {
		// every regular pair in FP64
		using namespace nbk;
		Reg64Args b;
		std::memset(&b, 0, sizeof(b));
		b.n_rows = a.n_rows; b.n_cols = a.n_cols; b.same_mesh = a.same_mesh; b.row_tiles = a.row_tiles;
		b.tile_cols = a.tile_cols; b.rmap = a.rmap; b.cmap = a.cmap; b.tn = a.tn; b.sn = a.sn;
		b.tc = a.tc; b.tnu = (const double *)test->d_nu64.get();
		b.sc = a.sc; b.snu = (const double *)trial->d_nu64.get();
		b.thr = a.thr;
		for (int d = 0; d < NSLOT; ++d) { b.qt[d] = (const double *)trial->d_qt64[d].get(); b.qn[d] = trial->qn[d]; }
		b.fixed_degree = a.fixed_degree;
		b.k = jb.k;
		b.ki = jb.ki;
		b.alpha = make_double2(jb.alpha.re, jb.alpha.im);
		b.out0 = out0; b.out1 = out1; b.ld = ld; b.dtype = dtype; b.accumulate = accumulate;
		const int pe = prof_begin(0, st);
#define NB_R64(F, K, M) do { \
		if (F == HELM && ck) hipLaunchKernelGGL((regular64_kernel<F, K, M, (F == HELM)>), dim3(grid), dim3(256), 0, st, b); \
		else hipLaunchKernelGGL((regular64_kernel<F, K, M>), dim3(grid), dim3(256), 0, st, b); } while (0)
		if (jb.mode == OUT_BM) NB_R64(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_R64(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_R64(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_R64(HELM, DLPT, OUT_SINGLE); break;
			default: NB_R64(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_R64(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_R64(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_R64(LAP, DLPT, OUT_SINGLE); break;
			default: NB_R64(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_R64
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	} else if (grid > 0 && (jb.passes & 1)) {
		using namespace nbk;
		const int pe = prof_begin(0, st);
		const bool fix7 = jb.pol.rule == NB_RULE_FIXED && trial->qn[jb.pol.fixed_degree] == 7;
		const bool mapped = jb.rmap || jb.cmap;
#define NB_RG(F, K, M, Q) do { \
		if (F == HELM && ck && mapped) \
			hipLaunchKernelGGL((regular_kernel<F, K, M, Q, true, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
		else if (F == HELM && ck) \
			hipLaunchKernelGGL((regular_kernel<F, K, M, Q, false, (F == HELM)>), dim3(grid), dim3(256), 0, st, a); \
		else if (mapped) hipLaunchKernelGGL((regular_kernel<F, K, M, Q, true>), dim3(grid), dim3(256), 0, st, a); \
		else hipLaunchKernelGGL((regular_kernel<F, K, M, Q, false>), dim3(grid), dim3(256), 0, st, a); } while (0)
		if (jb.mode == OUT_BM) {
			if (fix7) NB_RG(HELM, ALL4, OUT_BM, 7);
			else NB_RG(HELM, ALL4, OUT_BM, 0);
		} else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_RG(HELM, SLP, OUT_SINGLE, 0); break;
			case DLP: NB_RG(HELM, DLP, OUT_SINGLE, 0); break;
			case DLPT: NB_RG(HELM, DLPT, OUT_SINGLE, 0); break;
			default: NB_RG(HELM, HSP, OUT_SINGLE, 0); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_RG(LAP, SLP, OUT_SINGLE, 0); break;
			case DLP: NB_RG(LAP, DLP, OUT_SINGLE, 0); break;
			case DLPT: NB_RG(LAP, DLPT, OUT_SINGLE, 0); break;
			default: NB_RG(LAP, HSP, OUT_SINGLE, 0); break;
			}
		}
#undef NB_RG
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	}

	if (!(jb.passes & 2)) return NB_OK;
	// 2./3. near-singular and diagonal pairs in FP64
	const nbk::Rules *rules = nullptr;
	if ((s = device_rules(&rules))) return s;
	nbk::NearArgs na;
	std::memset(&na, 0, sizeof(na));
	na.rmap = jb.rmap; na.cmap = jb.cmap;
	na.tn = test->n; na.sn = trial->n;
	na.tc = (const double *)test->d_c.get();
	na.tnu = (const double *)test->d_nu64.get();
	na.X = (const double *)trial->d_X.get();
	na.sJ = (const double *)trial->d_J.get();
	na.snu = (const double *)trial->d_nu64.get();
	na.k = jb.k;
	na.ki = jb.ki;
	na.alpha = make_double2(jb.alpha.re, jb.alpha.im);
	na.out0 = out0; na.out1 = out1; na.ld = ld; na.dtype = dtype; na.accumulate = accumulate;
	na.rules = rules;
	for (int pass = 0; pass < 2; ++pass) {
		const bool diag = pass == 1;
		if (diag && !jb.same) continue;
		na.n_pairs = diag ? test->n : nl->n;
		na.pi = diag ? nullptr : (const int *)nl->pi.get();
		na.pj = diag ? nullptr : (const int *)nl->pj.get();
		if (na.n_pairs == 0) continue;
		const int g = (na.n_pairs + 63) / 64;
		const int pe = prof_begin(diag ? 2 : 1, side);
		using namespace nbk;
#define NB_FP64(F, K, M) do { \
		if (F == HELM && ck) { \
			if (diag) hipLaunchKernelGGL((fp64_pass_kernel<F, K, M, true, (F == HELM)>), dim3(g), dim3(64), 0, side, na); \
			else hipLaunchKernelGGL((fp64_pass_kernel<F, K, M, false, (F == HELM)>), dim3(g), dim3(64), 0, side, na); \
		} else if (diag) hipLaunchKernelGGL((fp64_pass_kernel<F, K, M, true>), dim3(g), dim3(64), 0, side, na); \
		else hipLaunchKernelGGL((fp64_pass_kernel<F, K, M, false>), dim3(g), dim3(64), 0, side, na); } while (0)
		if (jb.mode == OUT_BM) NB_FP64(HELM, ALL4, OUT_BM);
		else if (jb.family == NB_HELMHOLTZ) {
			switch (jb.km) {
			case SLP: NB_FP64(HELM, SLP, OUT_SINGLE); break;
			case DLP: NB_FP64(HELM, DLP, OUT_SINGLE); break;
			case DLPT: NB_FP64(HELM, DLPT, OUT_SINGLE); break;
			default: NB_FP64(HELM, HSP, OUT_SINGLE); break;
			}
		} else {
			switch (jb.km) {
			case SLP: NB_FP64(LAP, SLP, OUT_SINGLE); break;
			case DLP: NB_FP64(LAP, DLP, OUT_SINGLE); break;
			case DLPT: NB_FP64(LAP, DLPT, OUT_SINGLE); break;
			default: NB_FP64(LAP, HSP, OUT_SINGLE); break;
			}
		}
#undef NB_FP64
		NB_HIP(hipGetLastError());
		prof_end(pe, side);
	}
	if (has_fp64) {
		// join: the caller's stream waits for the FP64 passes
		NB_HIP(hipEventRecord(ev_join, side));
		NB_HIP(hipStreamWaitEvent(st, ev_join, 0));
		NB_HIP(hipEventDestroy(ev_fork));
		NB_HIP(hipEventDestroy(ev_join));
	}
	return NB_OK;
}

void host_add(void *dst, int64_t ld, const void *src, int64_t rows, int64_t cols, int dtype, bool accumulate)
{
	const int eb = dtype_bytes(dtype);
	unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nt; ++t) {
		th.emplace_back([=]() {
			for (int64_t j = t; j < cols; j += nt) {
				char *d = (char *)dst + j * ld * eb;
				const char *sp = (const char *)src + j * rows * eb;
				if (!accumulate) { std::memcpy(d, sp, rows * eb); continue; }
				switch (dtype) {
				case NB_C64: case NB_F32: {
					float *df = (float *)d; const float *sf = (const float *)sp;
					for (int64_t q = 0; q < rows * eb / 4; ++q) df[q] += sf[q];
					break;
				}
				default: {
					double *dd = (double *)d; const double *sd = (const double *)sp;
					for (int64_t q = 0; q < rows * eb / 8; ++q) dd[q] += sd[q];
					break;
				}
				}
			}
		});
	}
	for (auto &t : th) t.join();
}

// the single-type parts of a mesh and the element ids of their elements
// (null: the identity)
int mesh_parts(const nb_mesh *m, nb_mesh *parts[2], const int *ids[2])
{
	if (m->npe != 0) {
		parts[0] = const_cast<nb_mesh *>(m);
		ids[0] = nullptr;
		return 1;
	}
	for (int p = 0; p < 2; ++p) {
		parts[p] = m->part[p].get();
		ids[p] = (const int *)m->d_ids[p].get();
	}
	return 2;
}

// Collocation with a mesh that holds both element types (NiHu's mesh over
// tria_1 and quad_1, element ids in table order).  Each (test part, trial
// part) block runs the single-type passes, whose stores go through the parts'
// element ids to the result's rows and columns.  A block of one part against
// itself on the same mesh keeps the diagonal pass; the other blocks pair
// different elements, where NiHu's same-mesh loop finds no element match and
// does what its different-mesh loop does (near check, else regular:
// nihu/core/double_integral.hpp:599-673).  The blocks write disjoint entries.
int run_mixed(const Job &jb, void *out0, void *out1, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	nb_mesh *tp[2], *sp[2];
	const int *tid[2], *sid[2];
	const int nt = mesh_parts(jb.test, tp, tid), ns = mesh_parts(jb.trial, sp, sid);
	int s;
	// refuse before any block has run
	for (int b = 0; b < ns; ++b)
		if (sp[b]->npe == 4 && (s = check_quad_policy(jb.pol))) return s;
	// a test mesh with both types: the regular pairs of a trial part against
	// all rows at once (the test view, table order: each wave stores 64
	// consecutive rows; the diagonal is the row equal to the trial element's
	// id), then the near-singular and diagonal pairs per test part
	const bool all_rows = jb.test->npe == 0;
	for (int b = 0; b < ns; ++b) {
		if (all_rows) {
			Job sub = jb;
			sub.test = jb.test->view.get();
			sub.trial = sp[b];
			sub.cmap = sid[b];
			sub.passes = 1;
			if ((s = run_device(sub, out0, out1, ld, dtype, accumulate, st))) return s;
		}
		for (int a = 0; a < nt; ++a) {
			Job sub = jb;
			sub.test = tp[a];
			sub.trial = sp[b];
			sub.same = jb.same && tp[a] == sp[b];
			sub.rmap = tid[a];
			sub.cmap = sid[b];
			sub.passes = all_rows ? 2 : 3;
			// the parts are private to their mesh, whose lock the caller holds
			if ((s = run_device(sub, out0, out1, ld, dtype, accumulate, st))) return s;
		}
	}
	return NB_OK;
}

int run_any(const Job &jb, void *out0, void *out1, int64_t ld, int dtype, int accumulate, hipStream_t st)
{
	if (jb.formalism == NB_COLLOCATION_GAUSS || jb.formalism == NB_COLLOCATION_GAUSS_TRIAL)
		return run_gauss(jb, out0, out1, ld, dtype, accumulate, st);
	if (jb.test->npe == 0 || jb.trial->npe == 0) return run_mixed(jb, out0, out1, ld, dtype, accumulate, st);
	return run_device(jb, out0, out1, ld, dtype, accumulate, st);
}

// run a job into up to two outputs (both device, or both host)
int run_job(const Job &jb, const nb_out *o0, const nb_out *o1, hipStream_t st)
{
	if (o1 && (o0->location != o1->location || o0->dtype != o1->dtype || o0->accumulate != o1->accumulate))
		return fail(NB_ERR_ARG, "A and B must share location, dtype and accumulate");
	const int64_t rows = o0->rows, cols = o0->cols;
	if (o0->location == NB_DEVICE) {
		if (o1 && o0->ld != o1->ld) return fail(NB_ERR_ARG, "A and B must share the leading dimension");
		return run_any(jb, o0->ptr, o1 ? o1->ptr : nullptr, o0->ld, o0->dtype, o0->accumulate, st);
	}
	const size_t bytes = (size_t)rows * cols * dtype_bytes(o0->dtype);
	DevPtr t0, t1;
	int s;
	if ((s = dev_alloc(t0, bytes))) return s;
	if (o1 && (s = dev_alloc(t1, bytes))) return s;
	if ((s = run_any(jb, t0.get(), o1 ? t1.get() : nullptr, rows, o0->dtype, 0, st))) return s;
	std::vector<char> h(bytes);
	const nb_out *outs[2] = {o0, o1};
	void *tmps[2] = {t0.get(), o1 ? t1.get() : nullptr};
	for (int q = 0; q < 2; ++q) {
		if (!outs[q]) continue;
		NB_HIP(hipMemcpyAsync(h.data(), tmps[q], bytes, hipMemcpyDeviceToHost, st));
		NB_HIP(hipStreamSynchronize(st));
		host_add(outs[q]->ptr, outs[q]->ld, h.data(), rows, cols, outs[q]->dtype, outs[q]->accumulate != 0);
	}
	return NB_OK;
}

int check_policy(const nb_policy *p)
{
	if (!p) return fail(NB_ERR_ARG, "null policy");
	if (p->rule != NB_RULE_NIHU && p->rule != NB_RULE_FIXED) return fail(NB_ERR_ARG, "unknown quadrature rule");
	if (p->rule == NB_RULE_FIXED && (p->fixed_degree < 0 || p->fixed_degree > 14))
		return fail(NB_ERR_RANGE, "Too high quadrature degree selected for collocational integration");
	if (p->near_mode < NB_NEAR_MEAN_EDGE || p->near_mode > NB_NEAR_NONE) return fail(NB_ERR_ARG, "unknown near-singular mode");
	if (p->near_mode != NB_NEAR_NONE && !(p->near_limit > 0.0)) return fail(NB_ERR_ARG, "near limit must be positive");
	if (p->precision != NB_PREC_MIXED && p->precision != NB_PREC_FP64) return fail(NB_ERR_ARG, "unknown precision");
	return NB_OK;
}

int kernel_mask(int nx, int ny)
{
	if (nx == 0 && ny == 0) return nbk::SLP;
	if (nx == 0 && ny == 1) return nbk::DLP;
	if (nx == 1 && ny == 0) return nbk::DLPT;
	return nbk::HSP;
}

int current_device_matches(const nb_mesh *m)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	if (dev != m->device) return fail(NB_ERR_ARG, "mesh lives on another device");
	return NB_OK;
}

std::mutex g_tune_mu;
// tuned launch shapes by (elements, wavenumbers up to 64, quadrature rule)
std::map<std::tuple<int, int, int>, nb_launch> g_tuned;

// Burton-Miller sweep.  tune_block > 0: only the regular pass on the leading
// tune_block x tune_block block (autotuning); otherwise the full matrices.
// Per-call device buffers of a sweep: allocated for the call, or carved from
// a region given by the caller (256-byte aligned pieces; the region's size is
// checked before anything runs).
struct CallBufs {
	char *base = nullptr;
	size_t cap = 0, used = 0;
	std::vector<DevPtr> owned;
	static size_t piece(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
	int get(size_t bytes, void **out)
	{
		if (!base) {
			owned.emplace_back();
			int s = dev_alloc(owned.back(), bytes);
			*out = owned.back().get();
			return s;
		}
		if (used + piece(bytes) > cap) return fail(NB_ERR_RANGE, "scratch region exhausted");
		*out = base + used;
		used += piece(bytes);
		return NB_OK;
	}
	template <class T>
	int put(const std::vector<T> &v, hipStream_t st, void **out)
	{
		int s = get(v.size() * sizeof(T), out);
		if (s) return s;
		NB_HIP(hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
		return NB_OK;
	}
};

// scratch: null -> per-call buffers are allocated; otherwise they are carved
// from scratch[0, scratch_cap), every per-mesh table must already be
// resident, no launch shape is tuned and the FP64 passes stay on the caller's
// stream (nothing is allocated or created)
int sweep_impl(const nb_mesh *mesh, const double *k, int32_t n_k, const nb_policy *policy, void *A, void *B,
	int64_t ld, int64_t k_stride, const nb_launch *launch, void *stream, int tune_block, int regular_only,
	char *scratch = nullptr, size_t scratch_cap = 0)
{
	if (!mesh || !k || !A || !B) return fail(NB_ERR_ARG, "null argument");
	if (n_k <= 0) return fail(NB_ERR_ARG, "n_k must be positive");
	int s;
	if ((s = check_policy(policy))) return s;
	if ((s = current_device_matches(mesh))) return s;
	if (mesh->npe != 3 && tune_block > 0)
		return fail(NB_ERR_UNSUPPORTED, "sweep autotuning is implemented for triangle meshes only");
	const int n = mesh->n;
	const int nblk = tune_block > 0 ? tune_block : n;
	if (ld < nblk || k_stride < ld * (int64_t)nblk) return fail(NB_ERR_ARG, "ld / k_stride too small for the matrices");
	for (int f = 0; f < n_k; ++f)
		if (!(k[f] > 0.0)) return fail(NB_ERR_ARG, "wavenumbers must be positive");
	hipStream_t st = (hipStream_t)stream;
	nb_mesh *m = const_cast<nb_mesh *>(mesh);
	std::lock_guard<std::mutex> lk(m->mu);
	const double two_pi = 6.283185307179586476925286766559005768;

	const bool tri = mesh->npe == 3;
	const bool noalloc = scratch != nullptr;
	// NiHu's distance-driven degrees: every degree its tables may select
	const bool nihu = policy->rule == NB_RULE_NIHU;
	std::vector<int> degs;
	policy_degrees(*policy, nbk::ALL4, 0, degs);
	if (noalloc) {
		bool ready = tri && policy->precision == NB_PREC_MIXED;
		for (int d : degs) ready = ready && m->d_qt[d];
		if (ready && !nihu) ready = m->qn[policy->fixed_degree] == 7 || m->qn[policy->fixed_degree] == 1;
		if (!ready)
			return fail(NB_ERR_UNSUPPORTED, "the allocation-free sweep takes triangle meshes with NiHu's degrees or a "
				"fixed 1- or 7-point rule, whose tables are resident (run nb_assemble_sweep once on this mesh and "
				"policy first)");
	}
	const int Q = tri && !nihu ? (ensure_quad(m, policy->fixed_degree) ? -1 : m->qn[policy->fixed_degree]) : 0;
	if (Q < 0) return NB_ERR_RANGE;
	if (!tri || (!nihu && Q != 7 && Q != 1) || policy->precision != NB_PREC_MIXED) {
		// quadrilateral meshes, meshes with both element types, the other
		// fixed rules and the FP64 precision policy: one Burton-Miller
		// assembly per wavenumber
		if (tune_block > 0) return fail(NB_ERR_UNSUPPORTED, "this rule has no one-launch sweep to tune");
		if (!tri)
			for (const nb_mesh *p : {m, m->part[1].get()})
				if (p && p->npe == 4 && (s = check_quad_policy(*policy))) return s;
		for (int f = 0; f < n_k; ++f) {
			Job jb;
			jb.test = m; jb.trial = m; jb.formalism = NB_COLLOCATION; jb.same = 1; jb.family = NB_HELMHOLTZ;
			jb.km = nbk::ALL4; jb.mode = nbk::OUT_BM;
			jb.k = k[f]; jb.alpha = {0.0, -1.0 / k[f]}; jb.pol = *policy; jb.tile_cols = 32;
			if ((s = run_any(jb, (float2 *)A + f * k_stride, (float2 *)B + f * k_stride, ld, NB_C64, 0, st))) return s;
		}
		return NB_OK;
	}
	// the launch shape first: autotuning runs its own sweeps, which take the
	// mesh lock, so it is released meanwhile -- before any per-mesh cache is
	// fetched below, so that nothing fetched can change under this call
	nb_launch lc = {32, 4, 4};
	if (launch && launch->tile_cols > 0) lc = *launch;
	else {
		bool found = false;
		{
			std::lock_guard<std::mutex> lt(g_tune_mu);
			auto it = g_tuned.find(std::make_tuple(n, std::min<int>(n_k, 64), (int)policy->rule));
			if (it != g_tuned.end()) { lc = it->second; found = true; }
		}
		if (!found && tune_block == 0 && !noalloc) {
			m->mu.unlock();
			nb_launch best;
			s = nb_autotune(mesh, n_k, policy, &best);
			m->mu.lock();
			if (s) return s;
			lc = best;
		}
	}
	if (lc.k_batch != 2 && lc.k_batch != 4 && lc.k_batch != 8) return fail(NB_ERR_ARG, "k_batch must be 2, 4 or 8");
	if (lc.persistent <= 0 || lc.persistent > 32) return fail(NB_ERR_ARG, "persistent workgroups per CU must be 1..32");
	if ((long long)((nblk + 255) / 256) * ((nblk + lc.tile_cols - 1) / lc.tile_cols) *
		((n_k + lc.k_batch - 1) / lc.k_batch) > 0x7fffffffLL)
		return fail(NB_ERR_RANGE, "too many work items");
	std::shared_ptr<NearList> nl;
	const nbk::Rules *rules = nullptr;
	if (noalloc) {
		if (!m->d_thr || m->thr_mode != policy->near_mode || m->thr_limit != policy->near_limit ||
			!(nl = find_near_list(m, m, 1, *policy)))
			return fail(NB_ERR_UNSUPPORTED, "the allocation-free sweep needs the mesh's tables for this policy "
				"resident (run nb_assemble_sweep once on this mesh and policy first)");
		if (!regular_only && (s = device_rules_ready(&rules))) return s;
		// the whole per-call region, checked before anything runs
		const size_t np = (size_t)nl->n + n;
		const size_t need = 2 * CallBufs::piece(8 * (size_t)n_k) + 2 * CallBufs::piece(4 * (size_t)n_k) +
			CallBufs::piece(8 * (size_t)n_k) + CallBufs::piece(16 * (size_t)n_k) +
			(regular_only ? 0 : CallBufs::piece(np * 4 * sizeof(double)));
		if (need > scratch_cap)
			return fail(NB_ERR_RANGE, "the allocation-free sweep needs " + std::to_string(need) + " bytes of scratch (" +
				std::to_string(n_k) + " wavenumbers, " + std::to_string(np) + " FP64 pairs); the limit is " +
				std::to_string(scratch_cap) + " bytes");
	} else {
		for (int d : degs)
			if ((s = ensure_quad(m, d))) return s;
		if ((s = ensure_thresholds(m, policy->near_mode, policy->near_limit))) return s;
		if ((s = ensure_near_list(m, m, 1, *policy, st, &nl))) return s;
	}
	CallBufs bufs;
	if (noalloc) {
		bufs.base = scratch;
		bufs.cap = scratch_cap;
	} else {
		// the mesh's own workspace (this call holds the mesh lock and waits for
		// its kernels before it returns, so the next call may reuse it)
		const size_t np = (size_t)nl->n + n;
		const size_t need = 2 * CallBufs::piece(8 * (size_t)n_k) + 2 * CallBufs::piece(4 * (size_t)n_k) +
			CallBufs::piece(8 * (size_t)n_k) + CallBufs::piece(16 * (size_t)n_k) +
			(regular_only ? 0 : CallBufs::piece(np * 4 * sizeof(double)));
		if (m->sweep_ws_bytes < need) {
			m->bytes -= (int64_t)m->sweep_ws_bytes;
			m->sweep_ws.reset();
			m->sweep_ws_bytes = 0;
			if ((s = dev_alloc(m->sweep_ws, need))) return s;
			m->sweep_ws_bytes = need;
			m->bytes += (int64_t)need;
		}
		bufs.base = (char *)m->sweep_ws.get();
		bufs.cap = m->sweep_ws_bytes;
	}

	// per-wavenumber constants
	std::vector<double> krev(n_k), kd(n_k);
	std::vector<float> kf(n_k), krf(n_k);
	std::vector<float2> alf(n_k);
	std::vector<double2> ald(n_k);
	for (int f = 0; f < n_k; ++f) {
		kd[f] = k[f];
		krev[f] = k[f] / two_pi;
		kf[f] = (float)k[f];
		krf[f] = (float)krev[f];
		alf[f] = make_float2(0.0f, (float)(-1.0 / k[f]));
		ald[f] = make_double2(0.0, -1.0 / k[f]);
	}
	void *d_krev, *d_kd, *d_kf, *d_krf, *d_alf, *d_ald;
	if ((s = bufs.put(krev, st, &d_krev)) || (s = bufs.put(kd, st, &d_kd)) || (s = bufs.put(kf, st, &d_kf)) ||
		(s = bufs.put(krf, st, &d_krf)) || (s = bufs.put(alf, st, &d_alf)) || (s = bufs.put(ald, st, &d_ald)))
		return s;

	int n_cu = 256;
	{
		int dev = 0;
		NB_HIP(hipGetDevice(&dev));
		NB_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
	}

	// FP64 passes on the side stream (static once per pair, dynamic per wavenumber)
	hipStream_t side = st;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr;
	const bool forked = !regular_only && !noalloc;
	if (!regular_only) {
		if (!noalloc && (s = device_rules(&rules))) return s;
		if (forked) {
			if ((s = side_stream(&side))) return s;
			if (!side) side = st;
		}
		const int np = nl->n + n;
		void *d_static = nullptr;
		if ((s = bufs.get((size_t)np * 4 * sizeof(double), &d_static))) return s;
		if (forked) {
			NB_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
			NB_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
			NB_HIP(hipEventRecord(ev_fork, st));
			NB_HIP(hipStreamWaitEvent(side, ev_fork, 0));
		}
		nbk::StaticArgs sa;
		sa.n = n; sa.n_near = nl->n; sa.pi = (const int *)nl->pi.get(); sa.pj = (const int *)nl->pj.get();
		sa.c = (const double *)m->d_c.get(); sa.nu64 = (const double *)m->d_nu64.get();
		sa.X = (const double *)m->d_X.get(); sa.rules = rules; sa.st = (double *)d_static;
		const int pe1 = prof_begin(1, side);
		hipLaunchKernelGGL(nbk::fp64_static_kernel, dim3((np + 63) / 64), dim3(64), 0, side, sa);
		NB_HIP(hipGetLastError());
		nbk::DynArgs da;
		da.n = n; da.n_near = nl->n; da.n_k = n_k; da.pi = sa.pi; da.pj = sa.pj; da.c = sa.c; da.nu64 = sa.nu64;
		da.X = sa.X; da.J = (const double *)m->d_J.get(); da.rules = rules; da.st = sa.st;
		da.k = (const double *)d_kd; da.alpha = (const double2 *)d_ald;
		da.A = (float2 *)A; da.B = (float2 *)B; da.ld = ld; da.kstride = k_stride;
		// uniformly spaced wavenumbers (to 1e-12 relative): the near pairs rotate
		// e^{-ikr} from one wavenumber to the next instead of evaluating sincos
		bool uniform = n_k >= 2 && nl->n > 0;
		const double dk = n_k >= 2 ? (k[n_k - 1] - k[0]) / (n_k - 1) : 0.0;
		for (int f = 0; f < n_k && uniform; ++f)
			uniform = std::fabs(k[f] - (k[0] + f * dk)) <= 1e-12 * std::fabs(k[f]);
		da.p0 = uniform ? nl->n : 0;
		if (uniform) {
			const long long nr = (long long)nl->n * ((n_k + nbk::KBLK - 1) / nbk::KBLK);
			hipLaunchKernelGGL(nbk::fp64_dynamic_rec_kernel, dim3((nr + 63) / 64), dim3(64), 0, side, da, dk);
			NB_HIP(hipGetLastError());
		}
		const long long nt = (long long)(np - da.p0) * n_k;
		hipLaunchKernelGGL(nbk::fp64_dynamic_kernel, dim3((nt + 63) / 64), dim3(64), 0, side, da);
		NB_HIP(hipGetLastError());
		prof_end(pe1, side);
	}

	nbk::SweepArgs sw;
	sw.n = n; sw.nr = nblk; sw.nc = nblk; sw.tile_cols = lc.tile_cols; sw.n_k = n_k;
	sw.row_tiles = (nblk + 255) / 256;
	sw.col_tiles = (nblk + lc.tile_cols - 1) / lc.tile_cols;
	const long long kbt = (n_k + lc.k_batch - 1) / lc.k_batch;
	const long long items = (long long)sw.row_tiles * sw.col_tiles * kbt;
	if (items > 0x7fffffffLL) return fail(NB_ERR_RANGE, "too many work items");
	sw.n_items = (int)items;
	const long long grid = std::min<long long>(items, (long long)lc.persistent * n_cu);
	using namespace nbk;
	if (nihu) {
		// the regular pairs under NiHu's degrees, all wavenumbers in one launch
		SweepNihuArgs sn;
		std::memset(&sn, 0, sizeof(sn));
		RegArgs &a = sn.r;
		a.n_rows = nblk; a.n_cols = nblk; a.same_mesh = 1;
		a.tile_cols = lc.tile_cols;
		a.row_tiles = sw.row_tiles;
		a.tn = n; a.sn = n;
		a.tc = (const double *)m->d_c.get();
		a.tnu = (const float *)m->d_nuf.get();
		a.sc = a.tc;
		a.snu = a.tnu;
		a.thr = (const double *)m->d_thr.get();
		a.rec = (const TrialRec *)m->d_rec.get();
		for (int d = 0; d < NSLOT; ++d) { a.qt[d] = (const QPoint *)m->d_qt[d].get(); a.qn[d] = m->qn[d]; }
		a.fixed_degree = -1;
		a.field_complexity = 0;
		sn.n_k = n_k; sn.n_items = sw.n_items; sn.col_tiles = sw.col_tiles;
		sn.k_rev = (const double *)d_krev; sn.kf = (const float *)d_kf; sn.krevf = (const float *)d_krf;
		sn.alpha = (const float2 *)d_alf;
		sn.A = (float2 *)A; sn.B = (float2 *)B; sn.ld = ld; sn.kstride = k_stride;
		const int pe = prof_begin(3, st);
		if (lc.k_batch == 2) hipLaunchKernelGGL((sweep_nihu_kernel<2>), dim3(grid), dim3(256), 0, st, sn);
		else if (lc.k_batch == 4) hipLaunchKernelGGL((sweep_nihu_kernel<4>), dim3(grid), dim3(256), 0, st, sn);
		else hipLaunchKernelGGL((sweep_nihu_kernel<8>), dim3(grid), dim3(256), 0, st, sn);
		NB_HIP(hipGetLastError());
		prof_end(pe, st);
	} else {
	sw.c = (const double *)m->d_c.get(); sw.nu = (const float *)m->d_nuf.get(); sw.thr = (const double *)m->d_thr.get();
	sw.qt = (const nbk::QPoint *)m->d_qt[policy->fixed_degree].get();
	sw.k_rev = (const double *)d_krev; sw.kf = (const float *)d_kf; sw.krevf = (const float *)d_krf;
	sw.alpha = (const float2 *)d_alf;
	sw.A = (float2 *)A; sw.B = (float2 *)B; sw.ld = ld; sw.kstride = k_stride;
	const int pe = prof_begin(3, st);
#define NB_SW(KB, QQ) hipLaunchKernelGGL((sweep_regular_kernel<KB, QQ>), dim3(grid), dim3(256), 0, st, sw)
	if (Q == 7) {
		if (lc.k_batch == 2) NB_SW(2, 7); else if (lc.k_batch == 4) NB_SW(4, 7); else NB_SW(8, 7);
	} else {
		if (lc.k_batch == 2) NB_SW(2, 1); else if (lc.k_batch == 4) NB_SW(4, 1); else NB_SW(8, 1);
	}
#undef NB_SW
	NB_HIP(hipGetLastError());
	prof_end(pe, st);
	}
	if (forked) {
		NB_HIP(hipEventRecord(ev_join, side));
		NB_HIP(hipStreamWaitEvent(st, ev_join, 0));
		NB_HIP(hipEventDestroy(ev_fork));
		NB_HIP(hipEventDestroy(ev_join));
	}
	// the per-call buffers (wavenumber constants, static parts) are freed on
	// return, or the scratch region is handed to the next call: wait for the
	// kernels that read them
	NB_HIP(hipStreamSynchronize(st));
	return NB_OK;
}

// The allocation-free sweep's per-call region (wavenumber constants and the
// FP64 static parts): one array per device, resident until the process ends.
// Host code uses the address hipGetSymbolAddress gives for the current
// device, never the array's name.  Nothing in it has to start at zero: every
// piece is written before it is read within one call.
constexpr size_t NB_NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NB_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;
// calls on one device use its region one at a time (each call waits for its
// kernels before it returns)
std::map<int, std::unique_ptr<std::mutex> > g_scratch_use;

int scratch_region(char **out, std::mutex **use)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	auto it = g_scratch_ptr.find(dev);
	if (it == g_scratch_ptr.end()) {
		void *p = nullptr;
		NB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p) return fail(NB_ERR_DEVICE, "hipGetSymbolAddress gave no address for the scratch region");
		it = g_scratch_ptr.emplace(dev, (char *)p).first;
		g_scratch_use.emplace(dev, std::unique_ptr<std::mutex>(new std::mutex()));
	}
	*out = it->second;
	*use = g_scratch_use[dev].get();
	return NB_OK;
}

} // namespace

namespace {

int assemble_galerkin(const nb_mesh *test, const nb_mesh *trial, int same_mesh, int formalism, nb_kernel kernel,
	nb_scalar alpha, const nb_policy *policy, const nb_out *out, hipStream_t st)
{
	const bool lt = formalism == NB_GALERKIN_LINEAR || formalism == NB_GALERKIN_LINEAR_CONST;
	const bool ls = formalism == NB_GALERKIN_LINEAR || formalism == NB_GALERKIN_CONST_LINEAR;
	const bool ident = kernel.family == NB_IDENTITY;
	int s;
	if (!((test->npe == 3 && trial->npe == 3) || (test->npe == 4 && trial->npe == 4)))
		return fail(NB_ERR_UNSUPPORTED, "the Galerkin formalism takes meshes of one element type, triangles or "
			"quadrilaterals, on both sides");
	if (!ident && kernel.family == NB_HELMHOLTZ && !std::isfinite(kernel.k_im))
		return fail(NB_ERR_ARG, "imaginary part of the wavenumber must be finite");
	if (!ident) {
		if (kernel.family != NB_LAPLACE && kernel.family != NB_HELMHOLTZ) return fail(NB_ERR_ARG, "unknown kernel family");
		if ((kernel.nx != 0 && kernel.nx != 1) || (kernel.ny != 0 && kernel.ny != 1))
			return fail(NB_ERR_UNSUPPORTED, "normal derivative orders must be 0 or 1");
		if (kernel.family == NB_HELMHOLTZ && !(kernel.k > 0.0)) return fail(NB_ERR_ARG, "wavenumber must be positive");
		if ((s = check_policy(policy))) return s;
	}
	const bool cplx = (!ident && kernel.family == NB_HELMHOLTZ) || alpha.im != 0.0;
	if ((s = check_out(out, lt ? test->n_nodes : test->n, ls ? trial->n_nodes : trial->n, cplx))) return s;
	nb_mesh *tr = const_cast<nb_mesh *>(trial);
	nb_mesh *te = const_cast<nb_mesh *>(test);
	// the test mesh's tables and layout are built on first use too
	std::unique_lock<std::mutex> l1(tr->mu, std::defer_lock), l2(te->mu, std::defer_lock);
	if (te == tr) l1.lock();
	else std::lock(l1, l2);
	Job jb;
	jb.test = test; jb.trial = tr; jb.formalism = formalism; jb.same = same_mesh ? 1 : 0;
	jb.family = kernel.family;
	jb.km = ident ? 0 : kernel_mask(kernel.nx, kernel.ny);
	jb.mode = nbk::OUT_SINGLE;
	jb.k = kernel.family == NB_HELMHOLTZ ? kernel.k : 0.0;
	jb.ki = kernel.family == NB_HELMHOLTZ ? kernel.k_im : 0.0;
	jb.alpha = alpha;
	if (policy) jb.pol = *policy;
	else jb.pol = nb_policy{NB_RULE_NIHU, 0, NB_NEAR_MEAN_EDGE, 1.5};
	jb.tile_cols = 0;
	return run_job(jb, out, nullptr, st);
}


// collocation with the quadrilateral gauss field (NB_COLLOCATION_GAUSS,
// NB_COLLOCATION_GAUSS_TRIAL): rows = (4 or 1) x test elements, columns =
// 4 x trial elements
int assemble_gauss(const nb_mesh *test, const nb_mesh *trial, int same_mesh, int formalism, nb_kernel kernel,
	nb_scalar alpha, const nb_policy *policy, const nb_out *out, hipStream_t st)
{
	const bool tg = formalism == NB_COLLOCATION_GAUSS;
	const int64_t rows = (tg ? 4 : 1) * (int64_t)test->n, cols = 4 * (int64_t)trial->n;
	int s;
	if (trial->npe != 4) return fail(NB_ERR_UNSUPPORTED, "the gauss field lives on quadrilateral meshes");
	if (tg && test->npe != 4) return fail(NB_ERR_UNSUPPORTED, "the gauss test field needs a quadrilateral test mesh");
	if (kernel.family == NB_IDENTITY) {
		if ((s = check_out(out, rows, cols, alpha.im != 0.0))) return s;
		// local operator: the unit matrix over the dofs on the same mesh (the
		// shape functions are Lagrange polynomials on the collocation nodes),
		// nothing on different meshes (nihu/core/assembly.hpp:141-142)
		if (!same_mesh) return NB_OK;
		if (!tg) return fail(NB_ERR_UNSUPPORTED, "a constant test field against the gauss field is taken on another mesh only");
		if (out->location == NB_HOST) {
			const int eb = dtype_bytes(out->dtype);
			for (int64_t j = 0; j < out->cols; ++j) {
				char *col = (char *)out->ptr + j * out->ld * eb;
				if (!out->accumulate) std::memset(col, 0, out->rows * eb);
				char *e = col + j * eb;
				switch (out->dtype) {
				case NB_C64: ((float *)e)[0] += (float)alpha.re; ((float *)e)[1] += (float)alpha.im; break;
				case NB_C128: ((double *)e)[0] += alpha.re; ((double *)e)[1] += alpha.im; break;
				case NB_F32: *(float *)e += (float)alpha.re; break;
				default: *(double *)e += alpha.re; break;
				}
			}
			return NB_OK;
		}
		if (!out->accumulate) launch_zero(out->ptr, out->ld, out->rows, out->cols, dtype_bytes(out->dtype), st);
		hipLaunchKernelGGL(nbk::identity_kernel, dim3((int)((rows + 255) / 256)), dim3(256), 0, st, out->ptr,
			(long long)out->ld, (int)rows, out->dtype, alpha.re, alpha.im);
		NB_HIP(hipGetLastError());
		return NB_OK;
	}
	if (kernel.family != NB_LAPLACE && kernel.family != NB_HELMHOLTZ) return fail(NB_ERR_ARG, "unknown kernel family");
	if ((kernel.nx != 0 && kernel.nx != 1) || (kernel.ny != 0 && kernel.ny != 1))
		return fail(NB_ERR_UNSUPPORTED, "normal derivative orders must be 0 or 1");
	if (kernel.family == NB_HELMHOLTZ && !(kernel.k > 0.0)) return fail(NB_ERR_ARG, "wavenumber must be positive");
	if (kernel.family == NB_HELMHOLTZ && !std::isfinite(kernel.k_im))
		return fail(NB_ERR_ARG, "imaginary part of the wavenumber must be finite");
	if ((s = check_policy(policy))) return s;
	if ((s = check_out(out, rows, cols, kernel.family == NB_HELMHOLTZ || alpha.im != 0.0))) return s;
	if (rows * 16 > ((int64_t)1 << 32)) return fail(NB_ERR_RANGE, "too many rows");
	nb_mesh *tr = const_cast<nb_mesh *>(trial);
	nb_mesh *te = const_cast<nb_mesh *>(test);
	// the test mesh's node tables are built on first use too
	std::unique_lock<std::mutex> l1(tr->mu, std::defer_lock), l2(te->mu, std::defer_lock);
	if (te == tr) l1.lock();
	else std::lock(l1, l2);
	Job jb;
	jb.test = test; jb.trial = tr; jb.formalism = formalism; jb.same = same_mesh ? 1 : 0;
	jb.family = kernel.family;
	jb.km = kernel_mask(kernel.nx, kernel.ny);
	jb.mode = nbk::OUT_SINGLE;
	jb.k = kernel.family == NB_HELMHOLTZ ? kernel.k : 0.0;
	jb.ki = kernel.family == NB_HELMHOLTZ ? kernel.k_im : 0.0;
	jb.alpha = alpha;
	jb.pol = *policy;
	{
		std::lock_guard<std::mutex> lc(g_cfg_mu);
		jb.tile_cols = g_cfg.tile_cols;
	}
	return run_job(jb, out, nullptr, st);
}

} // namespace

extern "C" {

const char *nb_last_error(void) { return g_err.c_str(); }

int nb_version(void) { return 10000; }

int nb_device_info(char *name, int32_t name_len, int32_t *n_cu, int64_t *hbm_bytes)
{
	int dev = 0;
	NB_HIP(hipGetDevice(&dev));
	hipDeviceProp_t pr;
	NB_HIP(hipGetDeviceProperties(&pr, dev));
	if (name && name_len > 0) {
		std::strncpy(name, pr.gcnArchName, name_len - 1);
		name[name_len - 1] = 0;
	}
	if (n_cu) *n_cu = pr.multiProcessorCount;
	if (hbm_bytes) *hbm_bytes = (int64_t)pr.totalGlobalMem;
	return NB_OK;
}

int nb_mesh_create(const double *xyz, int32_t n_nodes, const int32_t *tri, int32_t n_tri, nb_mesh **out)
{
	if (!out) return fail(NB_ERR_ARG, "null output handle");
	*out = nullptr;
	if (!xyz || !tri || n_nodes <= 0 || n_tri <= 0) return fail(NB_ERR_ARG, "empty mesh");
	static std::atomic<uint64_t> next_uid(1);
	std::unique_ptr<nb_mesh> m(new nb_mesh());
	m->uid = next_uid++;
	const int n = n_tri;
	m->n_nodes = n_nodes;
	m->n = n;
	m->c.resize(3 * (size_t)n); m->nu.resize(3 * (size_t)n); m->J.resize(n); m->R.resize(n);
	m->diam.resize(n); m->X.resize(9 * (size_t)n);
	int g = nbg_geometry(xyz, n_nodes, tri, n, m->c.data(), m->nu.data(), m->J.data(), m->R.data(),
		m->diam.data(), m->X.data());
	if (g == 1) return fail(NB_ERR_ARG, "node index out of range");
	if (g == 2) return fail(NB_ERR_ARG, "degenerate element (zero area)");
	NB_HIP(hipGetDevice(&m->device));
	m->tri.assign(tri, tri + 3 * (size_t)n);
	std::vector<int32_t> tri_soa(3 * (size_t)n);
	for (int e = 0; e < n; ++e)
		for (int a = 0; a < 3; ++a) tri_soa[(size_t)a * n + e] = tri[3 * (size_t)e + a];
	std::vector<float> nuf(m->nu.begin(), m->nu.end());
	int s;
	if ((s = upload(m->d_c, m->c)) || (s = upload(m->d_nu64, m->nu)) || (s = upload(m->d_nuf, nuf)) ||
		(s = upload(m->d_X, m->X)) || (s = upload(m->d_J, m->J)) || (s = upload(m->d_tri, tri_soa)))
		return s;
	m->bytes = (int64_t)n * (3 * 8 + 3 * 8 + 3 * 4 + 9 * 8 + 8 + 3 * 4);
	*out = m.release();
	return NB_OK;
}

int nb_mesh_create_quads(const double *xyz, int32_t n_nodes, const int32_t *quad, int32_t n_quad, nb_mesh **out)
{
	if (!out) return fail(NB_ERR_ARG, "null output handle");
	*out = nullptr;
	if (!xyz || !quad || n_nodes <= 0 || n_quad <= 0) return fail(NB_ERR_ARG, "empty mesh");
	static std::atomic<uint64_t> next_uid(1ull << 62);
	std::unique_ptr<nb_mesh> m(new nb_mesh());
	m->uid = next_uid++;
	const int n = n_quad;
	m->n_nodes = n_nodes;
	m->n = n;
	m->npe = 4;
	m->c.resize(3 * (size_t)n); m->nu.resize(3 * (size_t)n); m->J.resize(n); m->R.resize(n);
	m->diam.resize(n); m->X.resize(12 * (size_t)n);
	int g = nbg_geometry_quad(xyz, n_nodes, quad, n, m->c.data(), m->nu.data(), m->J.data(), m->R.data(),
		m->diam.data(), m->X.data());
	if (g == 1) return fail(NB_ERR_ARG, "node index out of range");
	if (g == 2) return fail(NB_ERR_ARG, "degenerate element (zero normal)");
	NB_HIP(hipGetDevice(&m->device));
	m->tri.assign(quad, quad + 4 * (size_t)n);
	std::vector<float> nuf(m->nu.begin(), m->nu.end());
	int s;
	if ((s = upload(m->d_c, m->c)) || (s = upload(m->d_nu64, m->nu)) || (s = upload(m->d_nuf, nuf)) ||
		(s = upload(m->d_X, m->X)) || (s = upload(m->d_J, m->J)))
		return s;
	m->bytes = (int64_t)n * (3 * 8 + 3 * 8 + 3 * 4 + 12 * 8 + 8);
	*out = m.release();
	return NB_OK;
}

int nb_mesh_create_elements(const double *xyz, int32_t n_nodes, const int32_t *elem, int32_t n_elem, nb_mesh **out)
{
	if (!out) return fail(NB_ERR_ARG, "null output handle");
	*out = nullptr;
	if (!xyz || !elem || n_nodes <= 0 || n_elem <= 0) return fail(NB_ERR_ARG, "empty mesh");
	std::vector<int32_t> conn[2];
	std::vector<int32_t> ids[2];
	for (int32_t e = 0; e < n_elem; ++e) {
		const int32_t *row = elem + 5 * (size_t)e;
		const int p = row[0] == NB_TRIA_1_ID ? 0 : row[0] == NB_QUAD_1_ID ? 1 : -1;
		if (p < 0) return fail(NB_ERR_ARG, "unknown element type id (tria_1 32303 and quad_1 32404 are supported)");
		conn[p].insert(conn[p].end(), row + 1, row + (p == 0 ? 4 : 5));
		ids[p].push_back(e);
	}
	// one element type: the plain mesh of that type (ids are the part's own)
	if (ids[1].empty()) return nb_mesh_create(xyz, n_nodes, conn[0].data(), n_elem, out);
	if (ids[0].empty()) return nb_mesh_create_quads(xyz, n_nodes, conn[1].data(), n_elem, out);
	static std::atomic<uint64_t> next_uid(3ull << 61);
	std::unique_ptr<nb_mesh> m(new nb_mesh());
	m->uid = next_uid++;
	m->n_nodes = n_nodes;
	m->n = n_elem;
	m->npe = 0;
	NB_HIP(hipGetDevice(&m->device));
	nb_mesh *h = nullptr;
	int s;
	if ((s = nb_mesh_create(xyz, n_nodes, conn[0].data(), (int32_t)ids[0].size(), &h))) return s;
	m->part[0].reset(h);
	if ((s = nb_mesh_create_quads(xyz, n_nodes, conn[1].data(), (int32_t)ids[1].size(), &h))) return s;
	m->part[1].reset(h);
	if ((s = upload(m->d_ids[0], ids[0])) || (s = upload(m->d_ids[1], ids[1]))) return s;
	{
		std::unique_ptr<nb_mesh> v(new nb_mesh());
		v->uid = next_uid++;
		v->device = m->device;
		v->n_nodes = n_nodes;
		v->n = n_elem;
		v->c.resize(3 * (size_t)n_elem);
		v->nu.resize(3 * (size_t)n_elem);
		for (int p = 0; p < 2; ++p) {
			const nb_mesh *q = m->part[p].get();
			for (int e = 0; e < q->n; ++e)
				for (int d = 0; d < 3; ++d) {
					v->c[(size_t)d * n_elem + ids[p][e]] = q->c[(size_t)d * q->n + e];
					v->nu[(size_t)d * n_elem + ids[p][e]] = q->nu[(size_t)d * q->n + e];
				}
		}
		std::vector<float> nuf(v->nu.begin(), v->nu.end());
		if ((s = upload(v->d_c, v->c)) || (s = upload(v->d_nu64, v->nu)) || (s = upload(v->d_nuf, nuf))) return s;
		m->view = std::move(v);
	}
	m->tri.assign(elem, elem + 5 * (size_t)n_elem);
	m->bytes = m->part[0]->bytes + m->part[1]->bytes + 4 * (int64_t)n_elem;
	*out = m.release();
	return NB_OK;
}

int nb_mesh_destroy(nb_mesh *mesh)
{
	delete mesh;
	return NB_OK;
}

int nb_mesh_info(const nb_mesh *mesh, int32_t *n_nodes, int32_t *n_tri, int64_t *device_bytes)
{
	if (!mesh) return fail(NB_ERR_ARG, "null mesh");
	if (n_nodes) *n_nodes = mesh->n_nodes;
	if (n_tri) *n_tri = mesh->n;
	if (device_bytes) *device_bytes = mesh->bytes;
	return NB_OK;
}

int nb_assemble(const nb_mesh *test, const nb_mesh *trial, int32_t same_mesh, int32_t formalism,
	nb_kernel kernel, nb_scalar alpha, const nb_policy *policy, const nb_out *out, void *stream)
{
	if (!test || !trial) return fail(NB_ERR_ARG, "null mesh");
	if (same_mesh && test != trial) return fail(NB_ERR_ARG, "same_mesh requires test == trial");
	if (formalism != NB_COLLOCATION && formalism != NB_GALERKIN && formalism != NB_GALERKIN_LINEAR &&
		formalism != NB_COLLOCATION_GAUSS && formalism != NB_COLLOCATION_GAUSS_TRIAL &&
		formalism != NB_GALERKIN_CONST_LINEAR && formalism != NB_GALERKIN_LINEAR_CONST)
		return fail(NB_ERR_ARG, "unknown formalism");
	int s;
	if ((s = current_device_matches(test)) || (s = current_device_matches(trial))) return s;
	hipStream_t st = (hipStream_t)stream;
	if (formalism == NB_COLLOCATION_GAUSS || formalism == NB_COLLOCATION_GAUSS_TRIAL)
		return assemble_gauss(test, trial, same_mesh, formalism, kernel, alpha, policy, out, st);
	if (formalism != NB_COLLOCATION) return assemble_galerkin(test, trial, same_mesh, formalism, kernel, alpha, policy, out, st);
	if (kernel.family == NB_IDENTITY) {
		if ((s = check_out(out, test->n, trial->n, alpha.im != 0.0))) return s;
		// local operator: nothing on different meshes (nihu/core/assembly.hpp:141-142)
		if (!same_mesh) return NB_OK;
		if (out->location == NB_HOST) {
			const int eb = dtype_bytes(out->dtype);
			for (int64_t j = 0; j < out->cols; ++j) {
				char *col = (char *)out->ptr + j * out->ld * eb;
				if (!out->accumulate) std::memset(col, 0, out->rows * eb);
				char *e = col + j * eb;
				switch (out->dtype) {
				case NB_C64: ((float *)e)[0] += (float)alpha.re; ((float *)e)[1] += (float)alpha.im; break;
				case NB_C128: ((double *)e)[0] += alpha.re; ((double *)e)[1] += alpha.im; break;
				case NB_F32: *(float *)e += (float)alpha.re; break;
				default: *(double *)e += alpha.re; break;
				}
			}
			return NB_OK;
		}
		if (!out->accumulate) launch_zero(out->ptr, out->ld, out->rows, out->cols, dtype_bytes(out->dtype), st);
		hipLaunchKernelGGL(nbk::identity_kernel, dim3((test->n + 255) / 256), dim3(256), 0, st, out->ptr,
			(long long)out->ld, test->n, out->dtype, alpha.re, alpha.im);
		NB_HIP(hipGetLastError());
		return NB_OK;
	}
	if (kernel.family != NB_LAPLACE && kernel.family != NB_HELMHOLTZ) return fail(NB_ERR_ARG, "unknown kernel family");
	if ((kernel.nx != 0 && kernel.nx != 1) || (kernel.ny != 0 && kernel.ny != 1))
		return fail(NB_ERR_UNSUPPORTED, "normal derivative orders must be 0 or 1");
	if (kernel.family == NB_HELMHOLTZ && !(kernel.k > 0.0)) return fail(NB_ERR_ARG, "wavenumber must be positive");
	if (kernel.family == NB_HELMHOLTZ && !std::isfinite(kernel.k_im))
		return fail(NB_ERR_ARG, "imaginary part of the wavenumber must be finite");
	if ((s = check_policy(policy))) return s;
	bool cplx = kernel.family == NB_HELMHOLTZ || alpha.im != 0.0;
	if ((s = check_out(out, test->n, trial->n, cplx))) return s;
	nb_mesh *tr = const_cast<nb_mesh *>(trial);
	std::lock_guard<std::mutex> lk(tr->mu);
	Job jb;
	jb.test = test; jb.trial = tr; jb.formalism = NB_COLLOCATION; jb.same = same_mesh ? 1 : 0; jb.family = kernel.family;
	jb.km = kernel_mask(kernel.nx, kernel.ny); jb.mode = nbk::OUT_SINGLE;
	jb.k = kernel.family == NB_HELMHOLTZ ? kernel.k : 0.0;
	jb.ki = kernel.family == NB_HELMHOLTZ ? kernel.k_im : 0.0;
	jb.alpha = alpha; jb.pol = *policy;
	{
		std::lock_guard<std::mutex> lc(g_cfg_mu);
		jb.tile_cols = g_cfg.tile_cols;
	}
	return run_job(jb, out, nullptr, st);
}

int nb_assemble_bm(const nb_mesh *mesh, double k, nb_scalar alpha, const nb_policy *policy,
	const nb_out *A, const nb_out *B, void *stream)
{
	if (!mesh) return fail(NB_ERR_ARG, "null mesh");
	if (!(k > 0.0)) return fail(NB_ERR_ARG, "wavenumber must be positive");
	int s;
	if ((s = current_device_matches(mesh))) return s;
	if ((s = check_policy(policy))) return s;
	if ((s = check_out(A, mesh->n, mesh->n, true)) || (s = check_out(B, mesh->n, mesh->n, true))) return s;
	nb_mesh *m = const_cast<nb_mesh *>(mesh);
	std::lock_guard<std::mutex> lk(m->mu);
	Job jb;
	jb.test = m; jb.trial = m; jb.formalism = NB_COLLOCATION; jb.same = 1; jb.family = NB_HELMHOLTZ; jb.km = nbk::ALL4;
	jb.mode = nbk::OUT_BM;
	jb.k = k; jb.alpha = alpha; jb.pol = *policy;
	{
		std::lock_guard<std::mutex> lc(g_cfg_mu);
		jb.tile_cols = g_cfg.tile_cols;
	}
	return run_job(jb, A, B, (hipStream_t)stream);
}

int nb_assemble_bm_ck(const nb_mesh *mesh, nb_scalar k, nb_scalar alpha, const nb_policy *policy,
	const nb_out *A, const nb_out *B, void *stream)
{
	if (!mesh) return fail(NB_ERR_ARG, "null mesh");
	if (!(k.re > 0.0) || !std::isfinite(k.im)) return fail(NB_ERR_ARG, "wavenumber must have a positive real part");
	int s;
	if ((s = current_device_matches(mesh))) return s;
	if ((s = check_policy(policy))) return s;
	if ((s = check_out(A, mesh->n, mesh->n, true)) || (s = check_out(B, mesh->n, mesh->n, true))) return s;
	nb_mesh *m = const_cast<nb_mesh *>(mesh);
	std::lock_guard<std::mutex> lk(m->mu);
	Job jb;
	jb.test = m; jb.trial = m; jb.formalism = NB_COLLOCATION; jb.same = 1; jb.family = NB_HELMHOLTZ; jb.km = nbk::ALL4;
	jb.mode = nbk::OUT_BM;
	jb.k = k.re; jb.ki = k.im; jb.alpha = alpha; jb.pol = *policy;
	{
		std::lock_guard<std::mutex> lc(g_cfg_mu);
		jb.tile_cols = g_cfg.tile_cols;
	}
	return run_job(jb, A, B, (hipStream_t)stream);
}

int nb_assemble_bm_gauss(const nb_mesh *mesh, nb_scalar k, nb_scalar alpha, const nb_policy *policy,
	const nb_out *A, const nb_out *B, void *stream)
{
	if (!mesh) return fail(NB_ERR_ARG, "null mesh");
	if (!(k.re > 0.0) || !std::isfinite(k.im)) return fail(NB_ERR_ARG, "wavenumber must have a positive real part");
	if (mesh->npe != 4) return fail(NB_ERR_UNSUPPORTED, "the gauss field lives on quadrilateral meshes");
	int s;
	if ((s = current_device_matches(mesh))) return s;
	if ((s = check_policy(policy))) return s;
	const int64_t nd = 4 * (int64_t)mesh->n;
	if ((s = check_out(A, nd, nd, true)) || (s = check_out(B, nd, nd, true))) return s;
	if (nd * 16 > ((int64_t)1 << 32)) return fail(NB_ERR_RANGE, "too many rows");
	nb_mesh *m = const_cast<nb_mesh *>(mesh);
	std::lock_guard<std::mutex> lk(m->mu);
	Job jb;
	jb.test = m; jb.trial = m; jb.formalism = NB_COLLOCATION_GAUSS; jb.same = 1; jb.family = NB_HELMHOLTZ;
	jb.km = nbk::ALL4; jb.mode = nbk::OUT_BM;
	jb.k = k.re; jb.ki = k.im; jb.alpha = alpha; jb.pol = *policy;
	{
		std::lock_guard<std::mutex> lc(g_cfg_mu);
		jb.tile_cols = g_cfg.tile_cols;
	}
	return run_job(jb, A, B, (hipStream_t)stream);
}

int nb_classify(const nb_mesh *test, const nb_mesh *trial, int32_t same_mesh, int32_t nx, int32_t ny,
	const nb_policy *policy, int8_t *cls, void *stream)
{
	if (!test || !trial || !cls) return fail(NB_ERR_ARG, "null argument");
	if (same_mesh && test != trial) return fail(NB_ERR_ARG, "same_mesh requires test == trial");
	if (test->npe == 0 || trial->npe == 0)
		return fail(NB_ERR_UNSUPPORTED, "pair classes are reported for single-type meshes only");
	int s;
	if ((s = check_policy(policy))) return s;
	nb_mesh *tr = const_cast<nb_mesh *>(trial);
	std::lock_guard<std::mutex> lk(tr->mu);
	if ((s = ensure_thresholds(tr, policy->near_mode, policy->near_limit))) return s;
	nbk::ClassArgs a;
	a.n_rows = test->n; a.n_cols = trial->n; a.same_mesh = same_mesh ? 1 : 0;
	a.tn = test->n; a.sn = trial->n; a.order = 1 + nx + ny;
	a.fixed_degree = policy->rule == NB_RULE_FIXED ? policy->fixed_degree : -1;
	// the jacobian order of quad_1 adds 1 to the regular degree
	a.field_complexity = trial->npe == 4 ? 1 : 0;
	a.tc = (const double *)test->d_c.get(); a.sc = (const double *)trial->d_c.get();
	a.thr = (const double *)trial->d_thr.get();
	a.cls = cls;
	const long long n = (long long)a.n_rows * a.n_cols;
	hipLaunchKernelGGL(nbk::classify_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, a);
	NB_HIP(hipGetLastError());
	return NB_OK;
}

int nb_assemble_sweep(const nb_mesh *mesh, const double *k, int32_t n_k, const nb_policy *policy, void *A, void *B,
	int64_t ld, int64_t k_stride, const nb_launch *launch, void *stream)
{
	return sweep_impl(mesh, k, n_k, policy, A, B, ld, k_stride, launch, stream, 0, 0);
}

int nb_assemble_sweep_noalloc(const nb_mesh *mesh, const double *k, int32_t n_k, const nb_policy *policy, void *A,
	void *B, int64_t ld, int64_t k_stride, const nb_launch *launch, void *stream)
{
	char *region = nullptr;
	std::mutex *use = nullptr;
	int s;
	if ((s = scratch_region(&region, &use))) return s;
	std::lock_guard<std::mutex> lk(*use);
	return sweep_impl(mesh, k, n_k, policy, A, B, ld, k_stride, launch, stream, 0, 0, region, NB_NOALLOC_SCRATCH_BYTES);
}

int nb_autotune(const nb_mesh *mesh, int32_t n_k, const nb_policy *policy, nb_launch *best)
{
	if (!mesh || !policy || !best) return fail(NB_ERR_ARG, "null argument");
	if (n_k <= 0) return fail(NB_ERR_ARG, "n_k must be positive");
	int s;
	if ((s = check_policy(policy))) return s;
	if (mesh->npe != 3) return fail(NB_ERR_UNSUPPORTED, "sweep autotuning is implemented for triangle meshes only");
	if (policy->precision != NB_PREC_MIXED) {
		// the FP64 precision policy runs one single-wavenumber assembly per wavenumber
		*best = {32, 4, 4};
		return NB_OK;
	}
	if (policy->rule == NB_RULE_FIXED && policy->fixed_degree >= 0 && policy->fixed_degree < nbk::NSLOT) {
		nb_mesh *m = const_cast<nb_mesh *>(mesh);
		std::lock_guard<std::mutex> lk(m->mu);
		if ((s = ensure_quad(m, policy->fixed_degree))) return s;
		const int q = m->qn[policy->fixed_degree];
		if (q != 1 && q != 7) {
			// this rule runs one single-wavenumber assembly per wavenumber
			*best = {32, 4, 4};
			return NB_OK;
		}
	}
	const int n = mesh->n;
	const int nb = std::min(n, 4096);          // tune on a leading block of the matrix
	const int nk = std::min<int>(n_k, 8);
	DevPtr A, B;
	const size_t bytes = (size_t)nb * nb * nk * sizeof(float2);
	if ((s = dev_alloc(A, bytes)) || (s = dev_alloc(B, bytes))) return s;
	std::vector<double> ks(nk);
	for (int f = 0; f < nk; ++f) ks[f] = 1.0 + f;
	hipStream_t st = nullptr;
	hipEvent_t e0, e1;
	NB_HIP(hipEventCreate(&e0));
	NB_HIP(hipEventCreate(&e1));
	double best_ms = 1e30;
	nb_launch bl = {32, 4, 4};
	for (int tc : {16, 32, 64})
		for (int kb : {2, 4, 8})
			for (int pc : {2, 4, 8}) {
				nb_launch c = {tc, kb, pc};
				if ((s = sweep_impl(mesh, ks.data(), nk, policy, A.get(), B.get(), nb, (int64_t)nb * nb, &c, st, nb, 1)))
					return s;   // warm-up
				NB_HIP(hipEventRecord(e0, st));
				for (int rep = 0; rep < 3; ++rep)
					if ((s = sweep_impl(mesh, ks.data(), nk, policy, A.get(), B.get(), nb, (int64_t)nb * nb, &c, st, nb, 1)))
						return s;
				NB_HIP(hipEventRecord(e1, st));
				NB_HIP(hipEventSynchronize(e1));
				float ms = 0.0f;
				NB_HIP(hipEventElapsedTime(&ms, e0, e1));
				if (ms < best_ms) { best_ms = ms; bl = c; }
			}
	NB_HIP(hipEventDestroy(e0));
	NB_HIP(hipEventDestroy(e1));
	{
		std::lock_guard<std::mutex> lk(g_tune_mu);
		g_tuned[std::make_tuple(n, std::min<int>(n_k, 64), (int)policy->rule)] = bl;
	}
	*best = bl;
	return NB_OK;
}

int nb_set_launch(const nb_launch *launch)
{
	if (!launch) return fail(NB_ERR_ARG, "null launch");
	if (launch->tile_cols < 0 || launch->tile_cols > 4096) return fail(NB_ERR_ARG, "tile_cols out of range");
	std::lock_guard<std::mutex> lc(g_cfg_mu);
	if (launch->tile_cols) g_cfg.tile_cols = launch->tile_cols;
	if (launch->k_batch) g_cfg.k_batch = launch->k_batch;
	if (launch->persistent) g_cfg.persistent = launch->persistent;
	return NB_OK;
}

int nb_profile(int32_t enable)
{
	std::lock_guard<std::mutex> lk(g_prof_mu);
	g_prof_on = enable != 0;
	return NB_OK;
}

int nb_profile_read(double *ms, int64_t *launches, int32_t reset)
{
	std::vector<ProfEv> pend;
	{
		std::lock_guard<std::mutex> lk(g_prof_mu);
		pend.swap(g_prof_pending);
	}
	for (auto &e : pend) {
		NB_HIP(hipEventSynchronize(e.b));
		float t = 0.0f;
		NB_HIP(hipEventElapsedTime(&t, e.a, e.b));
		std::lock_guard<std::mutex> lk(g_prof_mu);
		g_prof_ms[e.pass] += t;
		g_prof_n[e.pass] += 1;
		(void)hipEventDestroy(e.a);
		(void)hipEventDestroy(e.b);
	}
	std::lock_guard<std::mutex> lk(g_prof_mu);
	for (int p = 0; p < NPASS; ++p) {
		if (ms) ms[p] = g_prof_ms[p];
		if (launches) launches[p] = g_prof_n[p];
		if (reset) { g_prof_ms[p] = 0.0; g_prof_n[p] = 0; }
	}
	return NB_OK;
}

int nb_gemv_workspace(int64_t rows, int64_t cols, int64_t *bytes)
{
	if (!bytes) return fail(NB_ERR_ARG, "null argument");
	if (rows < 0 || cols < 0) return fail(NB_ERR_ARG, "negative size");
	if (rows == 0 || cols == 0) { *bytes = 0; return NB_OK; }
	int rt, ch;
	long long cl;
	nbk::gemv_shape(rows, cols, &rt, &ch, &cl);
	*bytes = (int64_t)ch * rows * 16;
	return NB_OK;
}

int nb_gemv_c64(const void *A, int64_t rows, int64_t cols, int64_t ld, const void *x, void *y, void *work,
	int64_t work_bytes, void *stream)
{
	if (rows < 0 || cols < 0) return fail(NB_ERR_ARG, "negative size");
	if (rows == 0) return NB_OK;
	if (!A || !x || !y || (cols > 0 && !work)) return fail(NB_ERR_ARG, "null argument");
	if (ld < rows) return fail(NB_ERR_ARG, "leading dimension smaller than the number of rows");
	hipStream_t st = (hipStream_t)stream;
	if (cols == 0) {
		hipLaunchKernelGGL(nbk::zero_kernel, dim3(1, 1), dim3(256), 0, st, y, (long long)rows, (long long)rows, 1LL, 16);
		NB_HIP(hipGetLastError());
		return NB_OK;
	}
	nbk::GemvArgs g;
	nbk::gemv_shape(rows, cols, &g.row_tiles, &g.chunks, &g.clen);
	if ((int64_t)g.chunks * rows * 16 > work_bytes) return fail(NB_ERR_ARG, "workspace smaller than nb_gemv_workspace");
	const long long blocks = (long long)g.row_tiles * g.chunks;
	if (blocks > 0x7fffffffLL) return fail(NB_ERR_RANGE, "matrix too large");
	g.A = (const float2 *)A; g.rows = rows; g.cols = cols; g.ld = ld;
	g.x = (const double2 *)x; g.y = (double2 *)y; g.work = (double2 *)work;
	hipLaunchKernelGGL(nbk::gemv_partial_kernel, dim3((unsigned)blocks), dim3(256), 0, st, g);
	NB_HIP(hipGetLastError());
	hipLaunchKernelGGL(nbk::gemv_reduce_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, g);
	NB_HIP(hipGetLastError());
	return NB_OK;
}

int nb_eval_math(int32_t fn, const double *x, double *y, int64_t n, void *stream)
{
	if (fn < 0 || fn > 3) return fail(NB_ERR_ARG, "unknown math routine");
	if (n < 0 || (n > 0 && (!x || !y))) return fail(NB_ERR_ARG, "null argument");
	if (n == 0) return NB_OK;
	hipStream_t st = (hipStream_t)stream;
	const long long g = (n + 255) / 256;
	if (g > 0x7fffffffLL) return fail(NB_ERR_RANGE, "too many values");
	hipLaunchKernelGGL(nbk::math_kernel, dim3((unsigned)g), dim3(256), 0, st, fn, x, y, (long long)n);
	NB_HIP(hipGetLastError());
	return NB_OK;
}

int nb_synchronize(void *stream)
{
	NB_HIP(hipStreamSynchronize((hipStream_t)stream));
	return NB_OK;
}

} // extern "C"
