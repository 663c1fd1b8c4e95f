// nihu.hpp -- C++ front end of the MI355X BEM engine with NiHu's expression API.
//
// The names, argument meaning and error behaviour follow NiHu's hot-path API
// (paths relative to the NiHu repository root):
//   create_mesh / read_off_mesh          nihu/core/mesh.hpp:137-193, nihu/interface/read_off_mesh.hpp
//   constant_view, isoparametric_view, dirac
//                                         nihu/core/function_space.hpp:248-266,353-359
//   helmholtz_3d_{SLP,DLP,DLPt,HSP}_kernel, laplace_3d_*_kernel
//                                         nihu/library/helmholtz_kernel.hpp, laplace_kernel.hpp
//   create_integral_operator, identity_integral_operator, scalar * operator
//                                         nihu/core/integral_operator.hpp:131-329
//   op[trial_space]                       nihu/core/integral_operator.hpp:79-89
//   test_space * transform, wr + wr, result << wr
//                                         nihu/core/weighted_residual.hpp:155-206
// `result << wr` accumulates into the caller's matrix, as NiHu's
// matrix_block::operator+= does.  The result may be any column-major matrix
// type with data(), rows() and cols() (Eigen::Matrix qualifies) holding
// std::complex<double>, std::complex<float>, double or float, on the host, or
// a nihu_amd::device_matrix on the GPU.  Implemented formalisms: collocation
// (dirac(constant_view) test, constant_view trial; the quadrilateral gauss
// field) and Galerkin (constant_view or isoparametric_view on either side,
// dofs = elements or mesh nodes, on meshes of triangles or of
// quadrilaterals); other test / trial combinations throw std::logic_error.  Errors map to NiHu's exception types:
// std::out_of_range (quadrature degree), std::invalid_argument,
// std::logic_error, std::bad_alloc and std::runtime_error.

#ifndef NIHU_AMD_NIHU_HPP_INCLUDED
#define NIHU_AMD_NIHU_HPP_INCLUDED

#include "../../../include/nbem.h"

#include <complex>
#include <cstdint>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace nihu_amd {

inline void check(int status)
{
	if (status == NB_OK) return;
	const std::string msg = nb_last_error();
	switch (status) {
	case NB_ERR_RANGE: throw std::out_of_range(msg);
	case NB_ERR_ARG: throw std::invalid_argument(msg);
	case NB_ERR_UNSUPPORTED: throw std::logic_error(msg);
	case NB_ERR_OOM: throw std::bad_alloc();
	default: throw std::runtime_error(msg);
	}
}

// ------------------------------------------------------------------ policy

// Quadrature policy of regular and near-singular pairs.  The default is
// NiHu's: per-pair Dunavant degree from the accuracy-3 distance tables and
// the near-singular shortcut below 1.5 mean edge lengths.
inline nb_policy &default_policy()
{
	static nb_policy p = {NB_RULE_NIHU, 5, NB_NEAR_MEAN_EDGE, 1.5};
	return p;
}
inline void set_policy(nb_policy const &p) { default_policy() = p; }

// ------------------------------------------------------------------ mesh

struct tria_1_tag {};
struct quad_1_tag {};
enum { tria_1_id = 32303 };   // element type id of tria_1_elem (dimension 3, tria domain, 3 nodes)
enum { quad_1_id = 32404 };   // quad_1_elem (dimension 3, quad domain, 4 nodes)

// A mesh of flat triangles (3 nodes per element), bilinear quadrilaterals
// (4) or both (0: built from NiHu's element table) resident on the current
// GPU.  Elements are numbered by insertion order; shared ownership lets
// spaces and operators refer to it.
class mesh {
public:
	mesh(std::vector<double> const &xyz, std::vector<int32_t> const &elems, int nodes_per_element = 3)
		: m_xyz(xyz), m_tri(elems), m_npe(nodes_per_element)
	{
		if (m_npe != 3 && m_npe != 4) throw std::invalid_argument("elements must have 3 or 4 nodes");
		if (xyz.size() % 3 || elems.size() % m_npe)
			throw std::invalid_argument("node array must be n x 3 and element array m x nodes_per_element");
		nb_mesh *h = nullptr;
		const int32_t nn = (int32_t)(m_xyz.size() / 3), ne = (int32_t)(m_tri.size() / m_npe);
		check(m_npe == 3 ? nb_mesh_create(m_xyz.data(), nn, m_tri.data(), ne, &h)
		                 : nb_mesh_create_quads(m_xyz.data(), nn, m_tri.data(), ne, &h));
		m_h.reset(h, [](nb_mesh *p) { nb_mesh_destroy(p); });
	}
	// NiHu's element table with both element types: m rows [type id, n0, n1,
	// n2, n3] (n3 unused by tria_1), element id = row
	struct element_table {};
	mesh(std::vector<double> const &xyz, std::vector<int32_t> const &table, element_table)
		: m_xyz(xyz), m_tri(table), m_npe(0)
	{
		if (xyz.size() % 3 || table.size() % 5)
			throw std::invalid_argument("node array must be n x 3 and the element table m x 5");
		nb_mesh *h = nullptr;
		check(nb_mesh_create_elements(m_xyz.data(), (int32_t)(m_xyz.size() / 3), m_tri.data(),
			(int32_t)(m_tri.size() / 5), &h));
		m_h.reset(h, [](nb_mesh *p) { nb_mesh_destroy(p); });
	}
	nb_mesh *handle() const { return m_h.get(); }
	int32_t num_elements() const { return (int32_t)(m_tri.size() / (m_npe ? m_npe : 5)); }
	int32_t num_nodes() const { return (int32_t)(m_xyz.size() / 3); }
	int nodes_per_element() const { return m_npe; }
	std::vector<double> const &nodes() const { return m_xyz; }
	std::vector<int32_t> const &elements() const { return m_tri; }

private:
	std::vector<double> m_xyz;
	std::vector<int32_t> m_tri;
	int m_npe;
	std::shared_ptr<nb_mesh> m_h;
};

namespace detail {
// NiHu's element table rows [type_id, n0, ..., n_{npe-1}] of one element type
inline mesh from_element_table(std::vector<double> const &nodes, std::vector<unsigned> const &elements,
	unsigned type_id, int npe, char const *what)
{
	if (elements.size() % (npe + 1)) throw std::invalid_argument(std::string("elements must be m x ") +
		std::to_string(npe + 1) + ": type id and the node indices");
	std::vector<int32_t> el;
	el.reserve(elements.size() / (npe + 1) * npe);
	for (size_t e = 0; e < elements.size(); e += npe + 1) {
		if (elements[e] != type_id) throw std::invalid_argument(std::string("only ") + what + " elements are supported");
		for (int a = 1; a <= npe; ++a) el.push_back((int32_t)elements[e + a]);
	}
	return mesh(nodes, el, npe);
}

// NiHu's element table of m x 5 rows [type id, n0, ..., n3] holding tria_1
// and quad_1 rows (a table of one type gives the mesh of that type)
inline mesh from_mixed_table(std::vector<double> const &nodes, std::vector<int32_t> const &table)
{
	if (table.size() % 5) throw std::invalid_argument("elements must be m x 5: type id and four node indices");
	bool tri = false, quad = false;
	for (size_t e = 0; e < table.size(); e += 5) {
		if (table[e] == tria_1_id) tri = true;
		else if (table[e] == quad_1_id) quad = true;
		else throw std::invalid_argument("only tria_1 and quad_1 elements are supported");
	}
	if (tri && quad) return mesh(nodes, table, mesh::element_table());
	const int npe = tri ? 3 : 4;
	std::vector<int32_t> el;
	for (size_t e = 0; e < table.size(); e += 5)
		for (int a = 1; a <= npe; ++a) el.push_back(table[e + a]);
	return mesh(nodes, el, npe);
}

// OFF reader with NiHu's error messages; every face must have npe vertices
// (npe = 0: triangles and quadrilaterals alike)
inline mesh read_off(std::string const &fname, int npe)
{
	std::ifstream is(fname);
	std::string header;
	if (!(is >> header) || header != "OFF") throw std::runtime_error("Possibly invalid off file");
	unsigned nn, ne, nedge;
	if (!(is >> nn >> ne >> nedge)) throw std::runtime_error("Error reading number of mesh entries");
	std::vector<double> xyz(3 * (size_t)nn);
	for (auto &v : xyz)
		if (!(is >> v)) throw std::runtime_error("Error reading mesh nodes");
	std::vector<int32_t> el;
	el.reserve((size_t)(npe ? npe : 5) * ne);
	for (unsigned e = 0; e < ne; ++e) {
		unsigned nv;
		if (!(is >> nv)) throw std::runtime_error("Error reading mesh elements");
		if (npe ? nv != (unsigned)npe : nv != 3 && nv != 4)
			throw std::runtime_error("Invalid number of vertices for current element tag vector");
		if (!npe) el.push_back(nv == 3 ? (int32_t)tria_1_id : (int32_t)quad_1_id);
		for (unsigned a = 0; a < nv; ++a) {
			unsigned v;
			if (!(is >> v)) throw std::runtime_error("Error reading mesh elements");
			el.push_back((int32_t)v);
		}
		if (!npe && nv == 3) el.push_back(0);
	}
	return npe ? mesh(xyz, el, npe) : from_mixed_table(xyz, el);
}
} // namespace detail

// create_mesh(nodes, elements, tag): nodes is an (n x 3) and elements an
// (m x 4) row-major table [32303, n0, n1, n2] (tria_1_tag) or (m x 5)
// [32404, n0, n1, n2, n3] (quad_1_tag), as NiHu's mesh(nodes, elements)
// constructor reads them.
inline mesh create_mesh(std::vector<double> const &nodes, std::vector<unsigned> const &elements, tria_1_tag)
{
	return detail::from_element_table(nodes, elements, tria_1_id, 3, "tria_1");
}
inline mesh create_mesh(std::vector<double> const &nodes, std::vector<unsigned> const &elements, quad_1_tag)
{
	return detail::from_element_table(nodes, elements, quad_1_id, 4, "quad_1");
}

// create_mesh(nodes, elements, tria_1_tag(), quad_1_tag()) (either tag
// order): an (m x 5) table whose rows are tria_1 [32303, n0, n1, n2, any] or
// quad_1 [32404, n0, n1, n2, n3], element id = row whatever the type
inline mesh create_mesh(std::vector<double> const &nodes, std::vector<unsigned> const &elements, tria_1_tag,
	quad_1_tag)
{
	return detail::from_mixed_table(nodes, std::vector<int32_t>(elements.begin(), elements.end()));
}
inline mesh create_mesh(std::vector<double> const &nodes, std::vector<unsigned> const &elements, quad_1_tag,
	tria_1_tag)
{
	return create_mesh(nodes, elements, tria_1_tag(), quad_1_tag());
}

// read_off_mesh(filename, tags...): OFF reader (nihu/interface/read_off_mesh.hpp:75-105);
// with both tags, faces of 3 and 4 vertices in one file
inline mesh read_off_mesh(std::string const &fname, tria_1_tag) { return detail::read_off(fname, 3); }
inline mesh read_off_mesh(std::string const &fname, quad_1_tag) { return detail::read_off(fname, 4); }
inline mesh read_off_mesh(std::string const &fname, tria_1_tag, quad_1_tag) { return detail::read_off(fname, 0); }
inline mesh read_off_mesh(std::string const &fname, quad_1_tag, tria_1_tag) { return detail::read_off(fname, 0); }

// ------------------------------------------------------------------ spaces

enum space_kind { constant_kind = 0, isoparametric_kind = 1, dirac_constant_kind = 2, gauss_kind = 3,
	dirac_gauss_kind = 4 };

// constant_view: one dof per element (the element id)
class constant_space {
public:
	static const int kind = constant_kind;
	explicit constant_space(mesh const &m) : m_mesh(m) {}
	mesh const &get_mesh() const { return m_mesh; }
	int32_t get_num_dofs() const { return m_mesh.num_elements(); }

private:
	mesh m_mesh;
};

// isoparametric_view: linear shape functions, one dof per mesh node
// (get_num_points, nihu/core/function_space.hpp:156-165)
class isoparametric_space {
public:
	static const int kind = isoparametric_kind;
	explicit isoparametric_space(mesh const &m) : m_mesh(m) {}
	mesh const &get_mesh() const { return m_mesh; }
	int32_t get_num_dofs() const { return m_mesh.num_nodes(); }

private:
	mesh m_mesh;
};

// dirac(space): collocation at the element centres of the underlying space
class dirac_space {
public:
	static const int kind = dirac_constant_kind;
	explicit dirac_space(constant_space const &s) : m_space(s) {}
	mesh const &get_mesh() const { return m_space.get_mesh(); }
	int32_t get_num_dofs() const { return m_space.get_num_dofs(); }

private:
	constant_space m_space;
};

// The quadrilateral gauss field (quad_1_gauss_field, nihu/library/
// quad_1_gauss_field.hpp): four dofs per quadrilateral at the 2 x 2 Gauss
// points, dof 4 e + c at node c in the order (-,-), (+,-), (-,+), (+,+)
// (nihu/library/lib_shape.cpp:128-134).
struct quad_1_gauss_field_tag {};
enum { quad_1_gauss_field_id = 666 };   // field_traits::id<quad_1_gauss_field>

class gauss_space {
public:
	static const int kind = gauss_kind;
	explicit gauss_space(mesh const &m) : m_mesh(m)
	{
		if (m.nodes_per_element() != 4) throw std::invalid_argument("the gauss field needs a quadrilateral mesh");
	}
	mesh const &get_mesh() const { return m_mesh; }
	int32_t get_num_dofs() const { return 4 * m_mesh.num_elements(); }

private:
	mesh m_mesh;
};

// dirac(gauss field): collocation at the field's nodes
class dirac_gauss_space {
public:
	static const int kind = dirac_gauss_kind;
	explicit dirac_gauss_space(gauss_space const &s) : m_space(s) {}
	mesh const &get_mesh() const { return m_space.get_mesh(); }
	int32_t get_num_dofs() const { return m_space.get_num_dofs(); }

private:
	gauss_space m_space;
};

inline constant_space constant_view(mesh const &m) { return constant_space(m); }
inline isoparametric_space isoparametric_view(mesh const &m) { return isoparametric_space(m); }
inline dirac_space dirac(constant_space const &s) { return dirac_space(s); }
inline dirac_gauss_space dirac(gauss_space const &s) { return dirac_gauss_space(s); }

// create_function_space(nodes, fields, quad_1_gauss_field_tag()) as
// nihu/core/function_space.hpp:613-617 takes it, with the field matrix rows
// that nihu/applications/helmholtz/helmholtz_3d_coll.cpp:228-237 and
// benchmark/EAA/sphere/homogeneous/bem/bem.cpp:209-219 build:
// [666, n0, n1, n2, n3, d0, d1, d2, d3] with dofs d_c = 4 e + c (the
// numbering this engine assembles; another one throws).
inline gauss_space create_function_space(std::vector<double> const &nodes, std::vector<unsigned> const &fields,
	quad_1_gauss_field_tag)
{
	if (fields.size() % 9) throw std::invalid_argument("field rows must be [666, 4 node ids, 4 dof ids]");
	const size_t m = fields.size() / 9;
	std::vector<int32_t> el(4 * m);
	for (size_t e = 0; e < m; ++e) {
		unsigned const *r = &fields[9 * e];
		if (r[0] != (unsigned)quad_1_gauss_field_id) throw std::invalid_argument("field type id must be 666");
		for (int c = 0; c < 4; ++c) {
			el[4 * e + c] = (int32_t)r[1 + c];
			if (r[5 + c] != 4 * e + c) throw std::invalid_argument("gauss field dofs must be numbered 4 e + c");
		}
	}
	return gauss_space(mesh(nodes, el, 4));
}

// a space as the assembly sees it: mesh and kind
struct space_ref {
	mesh m;
	int kind;
};

template <class Space>
space_ref make_space_ref(Space const &s) { return space_ref{s.get_mesh(), Space::kind}; }

// ------------------------------------------------------------------ kernels

template <int Family, int Nx, int Ny, class WaveNumber = double>
struct kernel_t {
	explicit kernel_t(WaveNumber k = WaveNumber(0)) : m_k(k) {}
	WaveNumber get_wave_number() const { return m_k; }
	nb_kernel abi() const
	{
		// a real or complex wave number (NiHu's WaveNumber template parameter)
		std::complex<double> const w(m_k);
		return nb_kernel{Family, Nx, Ny, w.real(), w.imag()};
	}
	WaveNumber m_k;
};

template <class WaveNumber> using helmholtz_3d_SLP_kernel = kernel_t<NB_HELMHOLTZ, 0, 0, WaveNumber>;
template <class WaveNumber> using helmholtz_3d_DLP_kernel = kernel_t<NB_HELMHOLTZ, 0, 1, WaveNumber>;
template <class WaveNumber> using helmholtz_3d_DLPt_kernel = kernel_t<NB_HELMHOLTZ, 1, 0, WaveNumber>;
template <class WaveNumber> using helmholtz_3d_HSP_kernel = kernel_t<NB_HELMHOLTZ, 1, 1, WaveNumber>;
typedef kernel_t<NB_LAPLACE, 0, 0> laplace_3d_SLP_kernel;
typedef kernel_t<NB_LAPLACE, 0, 1> laplace_3d_DLP_kernel;
typedef kernel_t<NB_LAPLACE, 1, 0> laplace_3d_DLPt_kernel;
typedef kernel_t<NB_LAPLACE, 1, 1> laplace_3d_HSP_kernel;

// ------------------------------------------------------------------ operators

// An integral operator: a kernel (or the identity) and a scalar multiplier
// applied to every pair result (scaled_integral_operator).
class integral_operator {
public:
	integral_operator(nb_kernel k, std::complex<double> s = 1.0) : m_kernel(k), m_scalar(s) {}
	nb_kernel const &kernel() const { return m_kernel; }
	std::complex<double> scalar() const { return m_scalar; }
	template <class Space> class integral_transform operator[](Space const &trial) const;

private:
	nb_kernel m_kernel;
	std::complex<double> m_scalar;
};

template <int F, int Nx, int Ny, class W>
integral_operator create_integral_operator(kernel_t<F, Nx, Ny, W> const &k)
{
	return integral_operator(k.abi());
}

inline integral_operator identity_integral_operator()
{
	return integral_operator(nb_kernel{NB_IDENTITY, 0, 0, 0.0});
}

template <class Scalar, class = typename std::enable_if<std::is_arithmetic<Scalar>::value ||
	std::is_same<Scalar, std::complex<double> >::value || std::is_same<Scalar, std::complex<float> >::value>::type>
integral_operator operator*(Scalar const &s, integral_operator const &op)
{
	return integral_operator(op.kernel(), std::complex<double>(s) * op.scalar());
}

class integral_transform {
public:
	integral_transform(integral_operator const &op, space_ref const &trial) : m_op(op), m_trial(trial) {}
	integral_operator const &op() const { return m_op; }
	space_ref const &trial() const { return m_trial; }

private:
	integral_operator m_op;
	space_ref m_trial;
};

template <class Space>
integral_transform integral_operator::operator[](Space const &trial) const
{
	static_assert(std::is_same<Space, constant_space>::value || std::is_same<Space, isoparametric_space>::value ||
		std::is_same<Space, gauss_space>::value, "trial spaces: constant_view, isoparametric_view or the gauss field");
	return integral_transform(*this, make_space_ref(trial));
}

// the ABI formalism of a test / trial pair of spaces, or -1
inline int formalism_of(int test_kind, int trial_kind)
{
	if (test_kind == dirac_constant_kind && trial_kind == constant_kind) return NB_COLLOCATION;
	if (test_kind == constant_kind && trial_kind == constant_kind) return NB_GALERKIN;
	if (test_kind == isoparametric_kind && trial_kind == isoparametric_kind) return NB_GALERKIN_LINEAR;
	if (test_kind == constant_kind && trial_kind == isoparametric_kind) return NB_GALERKIN_CONST_LINEAR;
	if (test_kind == isoparametric_kind && trial_kind == constant_kind) return NB_GALERKIN_LINEAR_CONST;
	if (test_kind == dirac_gauss_kind && trial_kind == gauss_kind) return NB_COLLOCATION_GAUSS;
	if (test_kind == dirac_constant_kind && trial_kind == gauss_kind) return NB_COLLOCATION_GAUSS_TRIAL;
	return -1;
}

// test * transform: a weighted residual; sums of residuals accumulate in order
class weighted_residual {
public:
	struct term {
		integral_transform tr;
		space_ref test;
	};
	weighted_residual(space_ref const &test, integral_transform const &tr)
	{
		if (formalism_of(test.kind, tr.trial().kind) < 0)
			throw std::logic_error("test and trial spaces: dirac(constant) x constant, constant or isoparametric "
				"x constant or isoparametric, dirac(gauss) x gauss or dirac(constant) x gauss");
		m_terms.push_back(term{tr, test});
	}
	std::vector<term> const &terms() const { return m_terms; }
	weighted_residual operator+(weighted_residual const &o) const
	{
		weighted_residual r(*this);
		r.m_terms.insert(r.m_terms.end(), o.m_terms.begin(), o.m_terms.end());
		return r;
	}

private:
	std::vector<term> m_terms;
};

template <class Space>
weighted_residual operator*(Space const &test, integral_transform const &tr)
{
	static_assert(std::is_same<Space, dirac_space>::value || std::is_same<Space, constant_space>::value ||
		std::is_same<Space, isoparametric_space>::value || std::is_same<Space, dirac_gauss_space>::value,
		"test spaces: dirac(constant_view), constant_view, isoparametric_view, dirac(gauss field)");
	return weighted_residual(make_space_ref(test), tr);
}

// ------------------------------------------------------------------ results

// A column-major matrix in device memory (complex<float> by default)
template <class T = std::complex<float> >
class device_matrix {
public:
	device_matrix(int64_t rows, int64_t cols, void *ptr, int64_t ld) : m_rows(rows), m_cols(cols), m_ld(ld), m_ptr(ptr) {}
	int64_t rows() const { return m_rows; }
	int64_t cols() const { return m_cols; }
	int64_t ld() const { return m_ld; }
	void *data() const { return m_ptr; }
	typedef T Scalar;

private:
	int64_t m_rows, m_cols, m_ld;
	void *m_ptr;
};

template <class T> struct dtype_of;
template <> struct dtype_of<std::complex<float> > { enum { value = NB_C64 }; };
template <> struct dtype_of<std::complex<double> > { enum { value = NB_C128 }; };
template <> struct dtype_of<float> { enum { value = NB_F32 }; };
template <> struct dtype_of<double> { enum { value = NB_F64 }; };

template <class Res>
nb_out describe_host(Res &res)
{
	typedef typename std::remove_cv<typename std::remove_reference<decltype(*res.data())>::type>::type T;
	nb_out o;
	o.ptr = (void *)res.data();
	o.rows = res.rows();
	o.cols = res.cols();
	o.ld = res.rows();
	o.dtype = dtype_of<T>::value;
	o.location = NB_HOST;
	o.accumulate = 1;
	return o;
}

template <class T>
nb_out describe(device_matrix<T> &res)
{
	nb_out o;
	o.ptr = res.data();
	o.rows = res.rows();
	o.cols = res.cols();
	o.ld = res.ld();
	o.dtype = dtype_of<T>::value;
	o.location = NB_DEVICE;
	o.accumulate = 1;
	return o;
}

template <class Res>
nb_out describe(Res &res) { return describe_host(res); }

// result << wr: evaluate every term, accumulating into result
template <class Res>
Res &operator<<(Res &result, weighted_residual const &wr)
{
	nb_out o = describe(result);
	for (auto const &t : wr.terms()) {
		const nb_mesh *test = t.test.m.handle();
		const nb_mesh *trial = t.tr.trial().m.handle();
		const int same = test == trial;
		const std::complex<double> s = t.tr.op().scalar();
		check(nb_assemble(test, trial, same, formalism_of(t.test.kind, t.tr.trial().kind), t.tr.op().kernel(),
			nb_scalar{s.real(), s.imag()}, &default_policy(), &o, nullptr));
	}
	if (o.location == NB_DEVICE) check(nb_synchronize(nullptr));
	return result;
}

// ------------------------------------------------------------------ a host matrix

// Minimal column-major host matrix for callers without Eigen.
template <class T>
class matrix {
public:
	matrix(int64_t r = 0, int64_t c = 0) : m_r(r), m_c(c), m_v((size_t)(r * c), T(0)) {}
	int64_t rows() const { return m_r; }
	int64_t cols() const { return m_c; }
	T *data() { return m_v.data(); }
	T const *data() const { return m_v.data(); }
	T &operator()(int64_t i, int64_t j) { return m_v[(size_t)(j * m_r + i)]; }
	T const &operator()(int64_t i, int64_t j) const { return m_v[(size_t)(j * m_r + i)]; }
	void setZero() { std::fill(m_v.begin(), m_v.end(), T(0)); }

private:
	int64_t m_r, m_c;
	std::vector<T> m_v;
};

typedef matrix<std::complex<double> > cMatrix;
typedef matrix<double> dMatrix;

} // namespace nihu_amd

#endif
