/*
 * nbem.h -- C ABI of the MI355X (gfx950) dense BEM assembly engine.
 *
 * The engine replaces the pair loop behind NiHu's expression
 *     result << test_space * op[trial_space]
 * for 3D Laplace / Helmholtz SLP, DLP, DLPt and HSP kernels on flat 3-node
 * triangles: collocation (dirac(constant_view) x constant_view) and Galerkin
 * (constant_view x constant_view, isoparametric_view x isoparametric_view);
 * and on 4-node bilinear quadrilaterals: collocation with constant fields
 * and with the gauss field, and Galerkin with constant and isoparametric
 * (bilinear) fields.
 * Every entry point returns an
 * nb_status; on failure nb_last_error() returns a message for the calling
 * thread.  Plain pointers and sizes only.
 *
 * Reference interfaces (paths relative to the NiHu repository root):
 *   nb_mesh_create   <- NiHu::create_mesh / mesh(nodes, elements)
 *                       nihu/core/mesh.hpp:137-193, nihu/library/lib_mesh.hpp
 *   nb_assemble      <- assembly<TestSpace, TrialSpace, OnSameMesh>::eval_into
 *                       nihu/core/assembly.hpp:136-157, reached from
 *                       integral_transform::test_on_into (integral_transform.hpp:150-192)
 *                       and weighted_residual's operator<< (weighted_residual.hpp:197-206)
 *   nb_assemble_bm   <- the Burton-Miller composition of six eval_into calls in
 *                       nihu/applications/helmholtz/helmholtz_3d_coll.cpp:63-96
 *   nb_assemble_sweep<- the per-frequency loop of
 *                       benchmark/EAA/sphere/homogeneous/bem/bem.cpp:91-115
 */
#ifndef NBEM_H_INCLUDED
#define NBEM_H_INCLUDED

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the C++ facade maps them to NiHu's exception types */
enum nb_status {
	NB_OK = 0,
	NB_ERR_ARG = 1,          /* std::invalid_argument */
	NB_ERR_UNSUPPORTED = 2,  /* std::logic_error: combination NiHu rejects or this engine lacks */
	NB_ERR_RANGE = 3,        /* std::out_of_range: e.g. quadrature degree > 14 */
	NB_ERR_OOM = 4,          /* std::bad_alloc: device memory */
	NB_ERR_DEVICE = 5,       /* std::runtime_error: HIP error */
	NB_ERR_RUNTIME = 6       /* std::runtime_error: other */
};

/* kernel families: nihu/library/laplace_kernel.hpp, helmholtz_kernel.hpp */
enum nb_family { NB_LAPLACE = 0, NB_HELMHOLTZ = 1, NB_IDENTITY = 2 };

/* regular-pair quadrature policy */
enum nb_rule {
	NB_RULE_NIHU = 0,   /* NiHu default: per-pair Dunavant degree from the accuracy-3 distance
	                       tables (nihu/library/distance_kernel_intervals.hpp:116-155) */
	NB_RULE_FIXED = 1   /* one Dunavant degree for every regular pair (5 -> 7 points) */
};

/* near-singular classification of a pair (test centre x, trial element j) */
enum nb_near {
	NB_NEAR_MEAN_EDGE = 0, /* |x - c_j| / R_j < limit, R_j mean edge (NiHu, limit 1.5) */
	NB_NEAR_DIAMETER = 1,  /* |x - c_j| / diam_j < limit, diam_j longest edge */
	NB_NEAR_NONE = 2       /* no near-singular pass */
};

enum nb_dtype { NB_C64 = 0, NB_C128 = 1, NB_F32 = 2, NB_F64 = 3 };
enum nb_location { NB_HOST = 0, NB_DEVICE = 1 };
/* test x trial spaces (nihu/core/formalism.hpp:58-63, function_space.hpp):
 *   NB_COLLOCATION      dirac(constant_view(mesh)) x constant_view(mesh):
 *                       rows and columns are elements
 *   NB_GALERKIN         constant_view x constant_view: rows / columns are elements
 *   NB_GALERKIN_LINEAR  isoparametric_view x isoparametric_view: rows / columns
 *                       are mesh nodes (nihu/core/field.hpp:234-258)
 * Galerkin pairs that share a node on the same mesh use the Duffy-type rules of
 * nihu/core/singular_galerkin_quadrature.hpp:197-455 (triangles) and :457-625
 * (quadrilaterals, the subdomains of nihu/library/
 * lib_singular_galerkin_quadrature.cpp:6-31), rotated to the shared corners
 * as nihu/core/singular_accelerator.hpp:283-325 does; the hypersingular kernel
 * has none for coincident and edge-adjacent pairs and fails with
 * NB_ERR_RUNTIME "Invalid general singular quadrature", as NiHu throws.  Both
 * meshes of a Galerkin assembly hold one element type, triangles or
 * quadrilaterals (otherwise NB_ERR_UNSUPPORTED). */
/* NB_COLLOCATION_GAUSS: collocation with the quadrilateral gauss field
 * (NiHu's quad_1_gauss_field, nihu/library/quad_1_gauss_field.hpp, as the
 * GAUSS builds of nihu/applications/helmholtz/helmholtz_3d_coll.cpp:221-240
 * -- GAUSS is defined there, :20 -- and benchmark/EAA/sphere/homogeneous/
 * bem/bem.cpp:203-222 create it): four
 * dofs per quadrilateral at the 2 x 2 Gauss points, dof 4 e + b at node b in
 * the order (-,-), (+,-), (-,+), (+,+) of nihu/library/lib_shape.cpp:128-134
 * (the field matrix rows [666, nodes, 4 e .. 4 e + 3] of those programs);
 * test = dirac of the test mesh's gauss field (rows 4 i + a), trial = the
 * trial mesh's gauss field (columns 4 j + b); both meshes of quadrilaterals.
 * NB_COLLOCATION_GAUSS_TRIAL: the dirac of the test mesh's constant field
 * (rows = its elements, any one element type: field points,
 * dirac(constant_view(field)) of bem.cpp:233) against the trial mesh's
 * gauss field, on different meshes.  NB_IDENTITY with NB_COLLOCATION_GAUSS on
 * the same mesh: the unit matrix over the dofs. */
/* NB_GALERKIN_CONST_LINEAR: constant_view test x isoparametric_view trial
 * (rows = test elements, columns = trial nodes); NB_GALERKIN_LINEAR_CONST:
 * the reverse (rows = test nodes, columns = trial elements).  NiHu's general
 * formalism with two field types: the pair's degree takes the larger field
 * complexity (nihu/core/complexity_estimator.hpp:39-75), the singular rules
 * are those of the element types (nihu/core/match_types.hpp:54-83). */
enum nb_formalism { NB_COLLOCATION = 0, NB_GALERKIN = 1, NB_GALERKIN_LINEAR = 2, NB_COLLOCATION_GAUSS = 3,
                    NB_COLLOCATION_GAUSS_TRIAL = 4, NB_GALERKIN_CONST_LINEAR = 5, NB_GALERKIN_LINEAR_CONST = 6 };

typedef struct nb_mesh nb_mesh;

/* kernel: family, normal-derivative orders (nx, ny) in {0,1}^2, wavenumber
 * k + i k_im (NiHu's helmholtz_kernel<space, WaveNumber> with a real or
 * complex WaveNumber, library/helmholtz_kernel.hpp:103; k > 0; a nonzero
 * k_im is taken in every formalism on every mesh type) */
typedef struct {
	int32_t family;
	int32_t nx;
	int32_t ny;
	double k;
	double k_im;     /* 0 when the struct is zero-initialised */
} nb_kernel;

/* scalar multiplier of scaled_integral_operator (nihu/core/integral_operator.hpp:131-194) */
typedef struct {
	double re;
	double im;
} nb_scalar;

/* arithmetic of the regular pairs (near-singular and singular pairs are FP64 in both) */
enum nb_precision {
	NB_PREC_MIXED = 0,     /* regular pairs in FP32 with the phase k r reduced in FP64 (default) */
	NB_PREC_FP64 = 1       /* every pair in FP64: NiHu's plain method (collocation on triangles,
	                          quadrilaterals, meshes of both and with the quadrilateral gauss
	                          field; the Galerkin formalism: NB_ERR_UNSUPPORTED) */
};

typedef struct {
	int32_t rule;          /* nb_rule */
	int32_t fixed_degree;  /* Dunavant degree for NB_RULE_FIXED, 0..14 */
	int32_t near_mode;     /* nb_near */
	double near_limit;     /* 1.5 for NiHu */
	int32_t precision;     /* nb_precision; 0 when the struct is zero-initialised */
} nb_policy;

/* Caller-owned result matrix, column-major: entry (i, j) at ptr[j * ld + i],
 * i = test dof, j = trial dof.  accumulate = 1 adds into it as NiHu's
 * matrix_block::operator+= does (nihu/util/matrix_block.hpp:59-65);
 * accumulate = 0 overwrites every entry. */
typedef struct {
	void *ptr;
	int64_t rows;
	int64_t cols;
	int64_t ld;
	int32_t dtype;       /* nb_dtype */
	int32_t location;    /* nb_location */
	int32_t accumulate;
} nb_out;

/* Launch shape of the regular-pair kernels (0 = tuned default). */
typedef struct {
	int32_t tile_cols;     /* trial columns per workgroup */
	int32_t k_batch;       /* wavenumbers evaluated per geometry pass (sweep) */
	int32_t persistent;    /* sweep: workgroups per CU of the persistent grid */
} nb_launch;

const char *nb_last_error(void);
int nb_version(void);
int nb_device_info(char *name, int32_t name_len, int32_t *n_cu, int64_t *hbm_bytes);

/* Upload a triangle mesh: xyz[n_nodes][3], tri[n_tri][3] (0-based node ids,
 * element ids = row numbers).  Geometry (centres, normals, Jacobians, sizes)
 * is computed in FP64 and kept resident on the current device. */
int nb_mesh_create(const double *xyz, int32_t n_nodes, const int32_t *tri, int32_t n_tri, nb_mesh **out);
/* Upload a mesh of 4-node bilinear quadrilaterals (quad_1_elem, possibly
 * warped): quad[n_quad][4] with corners ordered around the element, the
 * normal dx/dxi x dx/deta outwards.  Collocation with constant fields
 * (nb_assemble with NB_COLLOCATION, nb_assemble_bm; nb_assemble_sweep runs
 * one Burton-Miller assembly per wavenumber), the gauss field
 * (NB_COLLOCATION_GAUSS) and Galerkin (NB_GALERKIN, NB_GALERKIN_LINEAR with
 * the four corner nodes as dofs); sweep autotuning rejects these meshes
 * (NB_ERR_UNSUPPORTED).
 * Replaces create_mesh(nodes, elements, quad_1_tag()) and the quad_1 paths
 * of the collocational double integral: nihu/core/double_integral.hpp:599-673,
 * nihu/library/guiggiani_1992.hpp:138-330 (hypersingular diagonal),
 * nihu/library/nearly_singular_collocational.hpp:71-203 and
 * nearly_singular_collocational_telles.hpp:66-103 (near-singular pairs). */
int nb_mesh_create_quads(const double *xyz, int32_t n_nodes, const int32_t *quad, int32_t n_quad, nb_mesh **out);
/* Upload a mesh from NiHu's element table: elem[n_elem][5] rows
 * [type id, n0, n1, n2, n3], type id NB_TRIA_1_ID (n3 ignored) or
 * NB_QUAD_1_ID; element id = row number, whatever the type (NiHu's
 * mesh(nodes, elements) numbers elements in table order:
 * nihu/core/mesh.hpp:137-193, push_element at :150; the tables that
 * read_off_data builds, nihu/interface/read_off_mesh.hpp:75-105, and
 * create_mesh(nodes, elements, tria_1_tag(), quad_1_tag())).  A table of one
 * type gives the same mesh as nb_mesh_create / nb_mesh_create_quads.  With
 * both types, collocation (nb_assemble with NB_COLLOCATION, nb_assemble_bm)
 * assembles each (test type, trial type) block and places it at the element
 * ids (nb_assemble_sweep: one Burton-Miller assembly per wavenumber); the
 * Galerkin formalism, sweep autotuning and nb_classify reject such meshes
 * (NB_ERR_UNSUPPORTED). */
#define NB_TRIA_1_ID 32303
#define NB_QUAD_1_ID 32404
int nb_mesh_create_elements(const double *xyz, int32_t n_nodes, const int32_t *elem, int32_t n_elem, nb_mesh **out);
int nb_mesh_destroy(nb_mesh *mesh);
int nb_mesh_info(const nb_mesh *mesh, int32_t *n_nodes, int32_t *n_tri, int64_t *device_bytes);

/* result (+)= alpha * <test, K trial>.  same_mesh = 1 when test and trial are
 * the same mesh object (NiHu decides this by pointer equality,
 * integral_transform.hpp:150-162): then the singular (diagonal) pass runs and
 * the identity family is defined.  stream: a hipStream_t or NULL. */
int nb_assemble(const nb_mesh *test, const nb_mesh *trial, int32_t same_mesh, int32_t formalism,
                nb_kernel kernel, nb_scalar alpha, const nb_policy *policy, const nb_out *out, void *stream);

/* Burton-Miller collocation matrices in one fused pass over the pairs:
 *   A = alpha*D - I/2 + H,   B = G + alpha*Ht + (alpha/2) I
 * with G, H, Ht, D the Helmholtz SLP, DLP, DLPt, HSP operators. */
int nb_assemble_bm(const nb_mesh *mesh, double k, nb_scalar alpha, const nb_policy *policy,
                   const nb_out *A, const nb_out *B, void *stream);
/* The same for a complex wavenumber k.re + i k.im (k.re > 0): the
 * composition of helmholtz_3d_coll.cpp:63-96 with a complex WaveNumber, on
 * meshes of triangles, of quadrilaterals or of both. */
int nb_assemble_bm_ck(const nb_mesh *mesh, nb_scalar k, nb_scalar alpha, const nb_policy *policy,
                      const nb_out *A, const nb_out *B, void *stream);
/* Burton-Miller A and B (as nb_assemble_bm, real or complex k) with the
 * quadrilateral gauss field (NB_COLLOCATION_GAUSS) on a quadrilateral mesh:
 * 4 n x 4 n, the identity the unit matrix over the dofs; the Burton-Miller
 * composition of helmholtz_3d_coll.cpp:63-96 in its GAUSS build. */
int nb_assemble_bm_gauss(const nb_mesh *mesh, nb_scalar k, nb_scalar alpha, const nb_policy *policy,
                         const nb_out *A, const nb_out *B, void *stream);

/* Burton-Miller matrices for n_k wavenumbers on one mesh in one persistent
 * launch (alpha = -i/k each).  A and B are device c64 arrays holding n_k
 * matrices of n_tri x n_tri with leading dimension ld, matrix f at
 * A + f * k_stride elements.  launch may be NULL (autotuned).  Triangle
 * meshes take the one-launch path under NiHu's distance-driven degrees (each
 * pair's distance, class and degrees found once for a batch of wavenumbers;
 * every matrix equal to nb_assemble_bm's at its wavenumber, the regular
 * entries bit for bit) or a fixed 1- or 7-point rule; other fixed rules and
 * meshes with quadrilaterals take one nb_assemble_bm pass per wavenumber. */
int nb_assemble_sweep(const nb_mesh *mesh, const double *k, int32_t n_k, const nb_policy *policy,
                      void *A, void *B, int64_t ld, int64_t k_stride, const nb_launch *launch, void *stream);

/* nb_assemble_sweep without any allocation -- no counterpart in NiHu: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as nb_assemble_sweep, restricted to its one-launch path (triangle
 * meshes; NiHu's degrees or a fixed 1- or 7-point rule).  The per-call buffers (wavenumber
 * constants and the FP64 static parts of the near-singular and diagonal
 * pairs, 48 bytes per wavenumber plus 32 bytes per FP64 pair, each piece
 * rounded up to 256 bytes) live in one 32 MiB region per device, held until
 * the process ends; larger requests fail with NB_ERR_RANGE before anything
 * runs.  The mesh's tables for the policy (thresholds, quadrature tables,
 * near-pair list) and the device's rule tables must be resident already: one
 * nb_assemble_sweep call on the mesh with the policy builds them (else
 * NB_ERR_UNSUPPORTED).  launch NULL takes the shape tuned for this mesh size
 * if there is one, else the default; nothing is tuned.  The FP64 passes run
 * on the caller's stream.  Calls on one device share its region: they take
 * it one at a time and each returns only when its kernels are done, so no
 * call sees another's contents half-written; nothing in the region needs to
 * start at zero. */
int nb_assemble_sweep_noalloc(const nb_mesh *mesh, const double *k, int32_t n_k, const nb_policy *policy, void *A,
                              void *B, int64_t ld, int64_t k_stride, const nb_launch *launch, void *stream);

/* Pair classes as the engine sees them: cls[i * n_trial + j] = -1 diagonal,
 * -2 near-singular, otherwise the regular Dunavant degree of kernel (nx, ny).
 * cls is a device int8 array of n_test * n_trial. */
int nb_classify(const nb_mesh *test, const nb_mesh *trial, int32_t same_mesh, int32_t nx, int32_t ny,
                const nb_policy *policy, int8_t *cls, void *stream);

/* Choose the launch shape of nb_assemble_sweep for this mesh by timing
 * candidates on the device; the result is cached per (mesh size, number of
 * wavenumbers up to 64, quadrature rule).  Fixed rules without a one-launch
 * path return the default shape. */
int nb_autotune(const nb_mesh *mesh, int32_t n_k, const nb_policy *policy, nb_launch *best);
int nb_set_launch(const nb_launch *launch);

int nb_synchronize(void *stream);

/* y = A x for the solve step: A a column-major complex64 matrix (rows x
 * cols, leading dimension ld, e.g. an assembled Burton-Miller matrix), x and
 * y complex128 vectors, all products and sums in FP64 (the sum over columns
 * in a fixed order: the result does not depend on scheduling).  work is a
 * device buffer of at least nb_gemv_workspace(rows, cols) bytes.  Stream-
 * ordered; no counterpart in NiHu, whose Burton-Miller application hands the
 * complex128 matrix to Eigen's GMRES
 * (nihu/applications/helmholtz/helmholtz_3d_coll.cpp:99-108). */
int nb_gemv_workspace(int64_t rows, int64_t cols, int64_t *bytes);
int nb_gemv_c64(const void *A, int64_t rows, int64_t cols, int64_t ld, const void *x, void *y, void *work,
                int64_t work_bytes, void *stream);

/* Test support (no NiHu counterpart): evaluate the engine's own FP64 device
 * routines of the near-singular and singular passes on n device doubles x
 * into y[2 n]: fn 0 sin and cos (y[2i], y[2i+1]; Cody-Waite reduction and
 * Taylor series), 1 natural logarithm, 2 1/sqrt, 3 1/x (y[2i]; y[2i+1] = 0). */
int nb_eval_math(int32_t fn, const double *x, double *y, int64_t n, void *stream);

/* Per-pass device timing with HIP events recorded on the launch stream
 * (pass 0 regular pairs, 1 near-singular pairs, 2 diagonal, 3 sweep).
 * nb_profile(1) starts recording, nb_profile_read sums the elapsed times
 * since the last reset (it waits for the recorded events). */
int nb_profile(int32_t enable);
int nb_profile_read(double *ms, int64_t *launches, int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* NBEM_H_INCLUDED */
