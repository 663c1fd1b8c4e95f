// nbem_layout.cpp -- host-side set-up of the Galerkin passes (FP64, g++).
//
// * canonical singular rules of coincident / edge / vertex triangle pairs
//   (nihu/core/singular_galerkin_quadrature.hpp:197-455) and quadrilateral
//   pairs (:457-625), the 4-point Gauss-Legendre rule per axis of
//   singular_quadrature_order 7 (nihu/library/helmholtz_kernel.hpp:69-71,
//   laplace_kernel.hpp:68-71);
// * the node-overlap scan that decides the match type and the start corners
//   (nihu/core/element.hpp:375-413, element_match.hpp:84-112);
// * the list of singular element pairs of a mesh;
// * the element order, wave tiles and column-slot schedule of the regular
//   Galerkin kernel for nodal (isoparametric) dofs of 3 or 4 corners.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" void nbg_gauss_legendre(int n, double *x, double *w);

namespace {

constexpr int TILE = 64;   // test elements per wave
constexpr int RMAX = 64;   // distinct nodes per wave tile
constexpr int KMAX = 8;    // elements of a wave tile sharing one node
constexpr int CS_MAX = 8;  // column-node slots per wave (8 triangles, 6 quadrilaterals)

// Morton key of a point quantised on a 1024^3 grid
uint64_t spread3(uint64_t v)
{
	v &= 0x3ff;
	v = (v | (v << 16)) & 0x030000ffull;
	v = (v | (v << 8)) & 0x0300f00full;
	v = (v | (v << 4)) & 0x030c30c3ull;
	v = (v | (v << 2)) & 0x09249249ull;
	return v;
}

struct Tiles {
	std::vector<int> elem, rows;        // [n_tiles * TILE]
	std::vector<uint32_t> contrib;      // [n_tiles * TILE * 2]
	int n_tiles = 0;
};

// greedy wave tiles over an element order: up to TILE elements whose npe
// corners use at most rpl * RMAX distinct nodes (rpl local rows per lane),
// no node in more than KMAX of them.  Local row r of a tile (lane r % 64 owns
// it) lists its contributors (lane * 4 + corner) in 8 bytes, 0xff = none.
Tiles make_tiles(const std::vector<int> &order, const int32_t *tri, int npe, int rpl)
{
	Tiles t;
	const int n = (int)order.size();
	int pos = 0;
	std::vector<int> rows;
	std::vector<std::vector<int> > con;
	while (pos < n) {
		rows.clear();
		con.clear();
		int cnt = 0;
		// a contributor byte lane * 4 + corner must not read 255 (none): with
		// four corners lane 63 stays idle
		const int max_elem = npe == 4 ? TILE - 1 : TILE;
		while (pos + cnt < n && cnt < max_elem) {
			const int e = order[pos + cnt];
			int l3[4];
			int extra = 0;
			bool full = false;
			for (int a = 0; a < npe; ++a) {
				const int v = tri[npe * e + a];
				auto it = std::find(rows.begin(), rows.end(), v);
				if (it == rows.end()) {
					l3[a] = -1;
					bool repeat = false;
					for (int b = 0; b < a; ++b) repeat = repeat || tri[npe * e + b] == v;
					if (!repeat) ++extra;
					continue;
				}
				l3[a] = (int)(it - rows.begin());
				if ((int)con[l3[a]].size() >= KMAX) full = true;
			}
			if (full || (int)rows.size() + extra > rpl * RMAX) break;
			for (int a = 0; a < npe; ++a) {
				if (l3[a] < 0) {
					// a dof repeated within the element (a constant field's
					// dof map names the element at every corner) takes its row once
					auto it = std::find(rows.begin(), rows.end(), tri[npe * e + a]);
					if (it != rows.end()) l3[a] = (int)(it - rows.begin());
					else {
						l3[a] = (int)rows.size();
						rows.push_back(tri[npe * e + a]);
						con.emplace_back();
					}
				}
				con[l3[a]].push_back(cnt * 4 + a);
			}
			++cnt;
		}
		for (int l = 0; l < TILE; ++l) t.elem.push_back(l < cnt ? order[pos + l] : -1);
		for (int l = 0; l < rpl * TILE; ++l) {
			t.rows.push_back(l < (int)rows.size() ? rows[l] : -1);
			uint32_t w[2] = {0xffffffffu, 0xffffffffu};
			if (l < (int)rows.size())
				for (int q = 0; q < (int)con[l].size(); ++q) {
					w[q >> 2] &= ~(0xffu << (8 * (q & 3)));
					w[q >> 2] |= (uint32_t)con[l][q] << (8 * (q & 3));
				}
			t.contrib.push_back(w[0]);
			t.contrib.push_back(w[1]);
		}
		++t.n_tiles;
		pos += cnt;
	}
	return t;
}

// least-recently-used assignment of the column nodes of a trial sequence to
// CS slots, restarted at every chunk; returns the number of misses.  Per
// sequence position i, col[(npe + 1) i + b] names the node that corner b's
// slot held before (-1: none) and col[(npe + 1) i + npe] packs the corners'
// slots in 5 bits each; fin[chunk * CS + s] = slot s's node at the chunk end.
int64_t schedule(const std::vector<int> &order, const int32_t *tri, int npe, int CS, int chunk, int *col, int *fin)
{
	const int n = (int)order.size();
	int64_t misses = 0;
	int node[CS_MAX], last[CS_MAX];
	for (int c0 = 0, ch = 0; c0 < n; c0 += chunk, ++ch) {
		for (int s = 0; s < CS; ++s) { node[s] = -1; last[s] = -1; }
		const int c1 = std::min(n, c0 + chunk);
		for (int i = c0; i < c1; ++i) {
			const int e = order[i];
			int slot[4] = {-1, -1, -1, -1}, victim[4] = {-1, -1, -1, -1};
			for (int b = 0; b < npe; ++b)
				for (int s = 0; s < CS; ++s)
					if (node[s] == tri[npe * e + b]) { slot[b] = s; last[s] = i; }
			for (int b = 0; b < npe; ++b) {
				if (slot[b] >= 0) continue;
				// a dof repeated within the element shares the slot just given to it
				for (int s = 0; s < CS; ++s)
					if (last[s] == i && node[s] == tri[npe * e + b]) slot[b] = s;
				if (slot[b] >= 0) continue;
				int best = -1;
				for (int s = 0; s < CS; ++s) {
					if (last[s] == i) continue;        // in use by this element
					if (best < 0 || last[s] < last[best]) best = s;
				}
				slot[b] = best;
				victim[b] = node[best];
				node[best] = tri[npe * e + b];
				last[best] = i;
				++misses;
			}
			if (col) {
				int packed = 0;
				for (int b = 0; b < npe; ++b) {
					col[(npe + 1) * i + b] = victim[b];
					packed |= slot[b] << (5 * b);
				}
				col[(npe + 1) * i + npe] = packed;
			}
		}
		if (fin)
			for (int s = 0; s < CS; ++s) fin[ch * CS + s] = node[s];
	}
	return misses;
}

} // namespace

extern "C" {

// Canonical singular rule of a match dimension (2 coincident, 1 common edge,
// 0 common vertex) on the reference triangle pair: rows (test xi, test eta,
// trial xi, trial eta, weight); returns the point count (1536, 1536, 512) or
// -1 if cap is too small.  The Duffy-type maps and their Jacobians are those
// of the cited helper (Taylor's decomposition of the 4D simplex product).
int nbg_singular_rule(int match, double *out, int cap)
{
	double gx[4], gw[4];
	nbg_gauss_legendre(4, gx, gw);
	double x[4], w[4];
	for (int q = 0; q < 4; ++q) { x[q] = (1.0 + gx[q]) / 2.0; w[q] = gw[q] * 0.5; }   // on [0, 1]
	const int ndom = match == 2 ? 3 : match == 1 ? 6 : 1;
	const bool symmetric = match != 1;
	int n = 0;
	for (int i1 = 0; i1 < 4; ++i1)
	for (int i2 = 0; i2 < 4; ++i2)
	for (int i3 = 0; i3 < 4; ++i3)
	for (int i4 = 0; i4 < 4; ++i4)
	for (int d = 0; d < ndom; ++d) {
		const double u0 = x[i1], u1 = x[i2], u2 = x[i3], u3 = x[i4];
		double wt = w[i1] * w[i2] * w[i3] * w[i4];
		double s1, s2, t1, t2;   // test (s1, s2), trial (t1, t2) before the simplex shift
		if (match == 2) {
			double m1, m2, jac;
			if (d == 0) {
				m1 = u0; m2 = u0 * u1;
				s1 = (1 - m1) * u2; s2 = s1 * u3;
				jac = (1 - m1) * s1;
			} else if (d == 1) {
				m1 = u0 * u1; m2 = u0 * (u1 - 1.0);
				s1 = (1 - m1 + m2) * u2 - m2; s2 = (s1 + m2) * u3 - m2;
				jac = (1.0 - m1 + m2) * (s1 + m2);
			} else {
				m1 = u0 * u1; m2 = u0;
				s1 = (1 - m2) * u2 + m2 - m1; s2 = (s1 - m2 + m1) * u3;
				jac = (1 - m2) * (s1 - m2 + m1);
			}
			wt *= jac * u0;
			t1 = m1 + s1; t2 = m2 + s2;
		} else if (match == 1) {
			double m1, m2;
			switch (d) {
			case 0: m1 = -u0 * u1; m2 = -u0 * u1 * u2; s1 = (1.0 - u0) * u3 + u0; s2 = u0 * (1.0 - u1 + u1 * u2); break;
			case 1: m1 = u0 * u1; m2 = u0 * u1 * u2; s1 = (1.0 - u0) * u3 + u0 * (1.0 - u1); s2 = u0 * (1.0 - u1); break;
			case 2: m1 = -u0 * u1 * u2; m2 = u0 * u1 * (1.0 - u2); s1 = (1.0 - u0) * u3 + u0; s2 = u0 * (1.0 - u1); break;
			case 3: m1 = u0 * u1 * u2; m2 = u0 * u1 * (u2 - 1.0); s1 = (1.0 - u0) * u3 + u0 * (1.0 - u1 * u2);
				s2 = u0 * (1.0 - u1 * u2); break;
			case 4: m1 = -u0 * u1 * u2; m2 = -u0 * u1; s1 = (1.0 - u0) * u3 + u0; s2 = u0; break;
			default: m1 = u0 * u1 * u2; m2 = u0 * u1; s1 = (1.0 - u0) * u3 + u0 * (1.0 - u1 * u2); s2 = u0 * (1.0 - u1); break;
			}
			wt *= u1 * u0 * u0 * (1.0 - u0);
			t1 = m1 + s1; t2 = m2 + s2;
		} else {
			s1 = u0; s2 = s1 * u1; t1 = s1 * u2; t2 = t1 * u3;
			wt *= u2 * u0 * u0 * u0;
		}
		const double p[4] = {s1 - s2, s2, t1 - t2, t2};
		for (int rep = 0; rep < (symmetric ? 2 : 1); ++rep) {
			if (n >= cap) return -1;
			double *o = out + 5 * n;
			if (rep == 0) { o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = p[3]; }
			else { o[0] = p[2]; o[1] = p[3]; o[2] = p[0]; o[3] = p[1]; }
			o[4] = wt;
			++n;
		}
	}
	return n;
}

// Number of nodes two elements of npe corners share, with the start corners
// of the shared edge or vertex as NiHu's scan finds them (test nodes a, trial
// nodes b); the second match moves a start back to the last corner when the
// first one was corner 0 (the edge that wraps around the node list).
int nbg_overlap_n(const int32_t *a, const int32_t *b, int npe, int *ind1, int *ind2)
{
	int num = 0, s1 = 0, s2 = 0;
	for (int i = 0; i < npe; ++i)
		for (int j = 0; j < npe; ++j) {
			if (a[i] != b[j]) continue;
			if (num == 0) { s1 = i; s2 = j; }
			else {
				if (i < s1 || (i == npe - 1 && s1 == 0 && num == 1)) s1 = i;
				if (j < s2 || (j == npe - 1 && s2 == 0 && num == 1)) s2 = j;
			}
			++num;
		}
	*ind1 = s1;
	*ind2 = s2;
	return num;
}

// the same for two triangles
int nbg_overlap(const int32_t *a, const int32_t *b, int *ind1, int *ind2)
{
	return nbg_overlap_n(a, b, 3, ind1, ind2);
}

// Canonical singular rule of two quadrilaterals (reference square [-1, 1]^2)
// that are one element (match 2), share an edge (1: the trial square is the
// test square's mirror image across eta_2 = -1) or share a corner (0: its
// image through (-1, -1)): rows (test xi, test eta, trial xi, trial eta,
// weight); returns the point count (2048, 1536, 1280) or -1 if cap is too
// small.  In the difference variable mu = eta - xi the pair domain is split
// into triangles fanned from mu = 0 (a quadrilateral with a repeated corner:
// the bilinear map's Jacobian vanishes at mu = 0 and cancels the 1/r
// singularity) and squares away from it; for each outer point mu the inner
// rule covers the rectangle of test points with mu + xi in the trial square.
// The face rule is symmetrised by swapping test and trial.  4 x 4 Gauss points
// (order 7) on both levels (nihu/core/singular_galerkin_quadrature.hpp:
// 457-625; the subdomain table of nihu/library/
// lib_singular_galerkin_quadrature.cpp:6-31).
int nbg_quad_singular_rule(int match, double *out, int cap)
{
	static const double face[4][8] = {
		{0, 0, 0, 0, 2, -2, 2, 0}, {0, 0, 0, 0, 2, 0, 2, 2}, {0, 0, 0, 0, 2, 2, 0, 2}, {0, 0, 0, 0, 0, 2, -2, 2}};
	static const double edge[6][8] = {
		{0, 0, 0, 0, -2, 0, -2, -2}, {0, 0, 0, 0, -2, -2, 0, -2}, {0, 0, 0, 0, 0, -2, 2, -2},
		{0, 0, 0, 0, 2, -2, 2, 0}, {-2, -2, -2, -4, 0, -4, 0, -2}, {0, -2, 0, -4, 2, -4, 2, -2}};
	static const double corner[5][8] = {
		{0, 0, 0, 0, -2, 0, -2, -2}, {0, 0, 0, 0, -2, -2, 0, -2}, {-4, 0, -4, -2, -2, -2, -2, 0},
		{-4, -2, -4, -4, -2, -4, -2, -2}, {-2, -2, -2, -4, 0, -4, 0, -2}};
	const double (*dom)[8] = match == 2 ? face : match == 1 ? edge : corner;
	const int ndom = match == 2 ? 4 : match == 1 ? 6 : 5;
	double g[4], gw[4];
	nbg_gauss_legendre(4, g, gw);
	// image of (u, v) in [-1, 1]^2 under the bilinear map onto corners C
	// (counter-clockwise from (-1, -1)) and its signed Jacobian
	auto bilinear = [](const double *C, double u, double v, double &x, double &y) -> double {
		const double a = 0.25 * (1 - u) * (1 - v), b = 0.25 * (1 + u) * (1 - v), c = 0.25 * (1 + u) * (1 + v),
			d = 0.25 * (1 - u) * (1 + v);
		x = a * C[0] + b * C[2] + c * C[4] + d * C[6];
		y = a * C[1] + b * C[3] + c * C[5] + d * C[7];
		const double xu = 0.25 * ((1 - v) * (C[2] - C[0]) + (1 + v) * (C[4] - C[6]));
		const double yu = 0.25 * ((1 - v) * (C[3] - C[1]) + (1 + v) * (C[5] - C[7]));
		const double xv = 0.25 * ((1 - u) * (C[6] - C[0]) + (1 + u) * (C[4] - C[2]));
		const double yv = 0.25 * ((1 - u) * (C[7] - C[1]) + (1 + u) * (C[5] - C[3]));
		return xu * yv - yu * xv;
	};
	// canonical trial point of mu + xi
	auto trial = [match](double a, double b, double &e0, double &e1) {
		e0 = match == 0 ? -2.0 - a : a;
		e1 = match == 2 ? b : -2.0 - b;
	};
	// the trial square in the variable before that map
	const double lo0 = match == 0 ? -3.0 : -1.0, hi0 = match == 0 ? -1.0 : 1.0;
	const double lo1 = match == 2 ? -1.0 : -3.0, hi1 = match == 2 ? 1.0 : -1.0;
	auto clamp1 = [](double v) { return v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v); };
	int n = 0;
	for (int d = 0; d < ndom; ++d)
		for (int i = 0; i < 16; ++i) {
			double m0, m1;
			const double wo = gw[i / 4] * gw[i % 4] * bilinear(dom[d], g[i / 4], g[i % 4], m0, m1);
			const double C[8] = {clamp1(lo0 - m0), clamp1(lo1 - m1), clamp1(hi0 - m0), clamp1(lo1 - m1),
				clamp1(hi0 - m0), clamp1(hi1 - m1), clamp1(lo0 - m0), clamp1(hi1 - m1)};
			for (int j = 0; j < 16; ++j) {
				double x0, x1, e0, e1;
				const double w = wo * (gw[j / 4] * gw[j % 4] * bilinear(C, g[j / 4], g[j % 4], x0, x1));
				trial(m0 + x0, m1 + x1, e0, e1);
				for (int rep = 0; rep < (match == 2 ? 2 : 1); ++rep) {
					if (n >= cap) return -1;
					double *o = out + 5 * n++;
					if (rep == 0) { o[0] = x0; o[1] = x1; o[2] = e0; o[3] = e1; }
					else { o[0] = e0; o[1] = e1; o[2] = x0; o[3] = x1; }
					o[4] = w;
				}
			}
		}
	return n;
}

// Singular pairs of a mesh of npe-corner elements against itself: coincident
// pairs (e, e) first, then every ordered pair of distinct elements sharing a
// node.  info = match | ind1 << 4 | ind2 << 8.  Returns the pair count; with
// null outputs only counts.
int64_t nbg_singular_pairs_n(int n_nodes, int n, const int32_t *el, int npe, int32_t *pe, int32_t *pf, int32_t *info)
{
	std::vector<int> deg(n_nodes + 1, 0);
	for (int e = 0; e < n; ++e)
		for (int a = 0; a < npe; ++a) ++deg[el[npe * e + a] + 1];
	for (int v = 0; v < n_nodes; ++v) deg[v + 1] += deg[v];
	std::vector<int> adj(deg[n_nodes]);
	std::vector<int> fill(deg.begin(), deg.end() - 1);
	for (int e = 0; e < n; ++e)
		for (int a = 0; a < npe; ++a) adj[fill[el[npe * e + a]]++] = e;
	int64_t cnt = 0;
	for (int e = 0; e < n; ++e) {
		if (pe) { pe[cnt] = e; pf[cnt] = e; info[cnt] = 2; }
		++cnt;
	}
	std::vector<int> cand;
	for (int e = 0; e < n; ++e) {
		cand.clear();
		for (int a = 0; a < npe; ++a) {
			const int v = el[npe * e + a];
			for (int q = deg[v]; q < deg[v + 1]; ++q)
				if (adj[q] != e) cand.push_back(adj[q]);
		}
		std::sort(cand.begin(), cand.end());
		cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
		for (int f : cand) {
			int i1, i2;
			const int num = nbg_overlap_n(el + npe * e, el + npe * f, npe, &i1, &i2);
			if (pe) { pe[cnt] = e; pf[cnt] = f; info[cnt] = (num > 1 ? 1 : 0) | (i1 << 4) | (i2 << 8); }
			++cnt;
		}
	}
	return cnt;
}

// the same for triangles
int64_t nbg_singular_pairs(int n_nodes, int n, const int32_t *tri, int32_t *pe, int32_t *pf, int32_t *info)
{
	return nbg_singular_pairs_n(n_nodes, n, tri, 3, pe, pf, info);
}

// Layout of the nodal-dof regular kernel for one mesh of npe-corner elements
// with cs column-node slots per lane.  Picks the element order (the mesh's
// own, or Morton order of the centres when that packs the wave tiles
// better), builds the wave tiles and the column-slot schedule of the trial
// sequence in chunks of `chunk` elements.  A tile holds up to 64 elements
// (63 quadrilaterals: the contributor byte of lane 63, corner 3 would read as
// the empty marker 255);
// triangles use up to 64 distinct nodes (one local row per lane),
// quadrilaterals up to 128 (two local rows per lane: 64 quadrilaterals touch
// about 81 nodes).  Call with null arrays to get n_tiles; then with arrays of
// n_tiles * 64 (elem), n_tiles * 64 * rpl (rows), n_tiles * 128 * rpl
// (contrib), n (order), (npe + 1) n (col) and ceil(n / chunk) * cs (fin),
// rpl = 1 for triangles and 2 for quadrilaterals.  Returns -1 for npe
// outside 3..4 or cs outside npe..8.
int nbg_galerkin_layout_n(int n, const int32_t *tri, int npe, int cs, const double *c, int chunk,
	int32_t *order_out, int32_t *elem, uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin)
{
	if (npe < 3 || npe > 4 || cs < npe || cs > CS_MAX) return -1;
	std::vector<int> mesh_order(n), morton(n);
	for (int e = 0; e < n; ++e) mesh_order[e] = e;
	double lo[3], hi[3];
	for (int d = 0; d < 3; ++d) {
		lo[d] = *std::min_element(c + d * n, c + (d + 1) * n);
		hi[d] = *std::max_element(c + d * n, c + (d + 1) * n);
	}
	std::vector<std::pair<uint64_t, int> > key(n);
	for (int e = 0; e < n; ++e) {
		uint64_t k = 0;
		for (int d = 0; d < 3; ++d) {
			const double span = hi[d] - lo[d];
			const uint64_t q = span > 0 ? (uint64_t)std::min(1023.0, (c[d * n + e] - lo[d]) / span * 1023.0) : 0;
			k |= spread3(q) << d;
		}
		key[e] = {k, e};
	}
	std::stable_sort(key.begin(), key.end(),
		[](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) { return a.first < b.first; });
	for (int e = 0; e < n; ++e) morton[e] = key[e].second;
	const int rpl = npe == 4 ? 2 : 1;
	Tiles t0 = make_tiles(mesh_order, tri, npe, rpl), t1 = make_tiles(morton, tri, npe, rpl);
	const int64_t m0 = schedule(mesh_order, tri, npe, cs, chunk, nullptr, nullptr);
	const int64_t m1 = schedule(morton, tri, npe, cs, chunk, nullptr, nullptr);
	const bool use_morton = t1.n_tiles < t0.n_tiles || (t1.n_tiles == t0.n_tiles && m1 < m0);
	const std::vector<int> &ord = use_morton ? morton : mesh_order;
	const Tiles &t = use_morton ? t1 : t0;
	if (!order_out) return t.n_tiles;
	std::copy(ord.begin(), ord.end(), order_out);
	std::copy(t.elem.begin(), t.elem.end(), elem);
	std::copy(t.contrib.begin(), t.contrib.end(), contrib);
	std::copy(t.rows.begin(), t.rows.end(), rows);
	schedule(ord, tri, npe, cs, chunk, col, fin);
	return t.n_tiles;
}

// the same for triangles with 8 slots: col has 4 n entries, fin 8 per chunk
int nbg_galerkin_layout(int n, const int32_t *tri, const double *c, int chunk,
	int32_t *order_out, int32_t *elem, uint32_t *contrib, int32_t *rows, int32_t *col, int32_t *fin)
{
	return nbg_galerkin_layout_n(n, tri, 3, 8, c, chunk, order_out, elem, contrib, rows, col, fin);
}

} // extern "C"
