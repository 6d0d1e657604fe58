/*
 * zombie_amd.h -- C-ABI of the MI355X (gfx950) Walk-on-Spheres / Walk-on-Stars engine.
 *
 * The reference solves at the granularity of
 *   WalkOnSpheres<T,DIM>::solve(pde, walkSettings, nWalks, samplePts, statistics)
 *   (include/zombie/point_estimation/walk_on_spheres.h:38-44, walk_on_stars.h:40-46)
 * with std::function callback tables for the geometry (core/geometric_queries.h:58-124)
 * and the PDE data (core/pde.h:19-45).  std::function closures cannot run on the GPU,
 * so this boundary sits at solve() granularity: the caller builds a device scene once
 * (the counterpart of FcpwDirichletBoundaryHandler::buildAccelerationStructure +
 * populateGeometricQueriesFor*Boundary, utils/fcpw_geometric_queries.h:569-586,1062-1304),
 * describes the PDE data with device-representable function descriptors, and calls one
 * solve per batch of points.
 *
 * Conventions
 *  - Every entry point returns zg_status (0 = ok) and never exits the process; the text of
 *    the last error on the calling thread is available through zg_last_error().
 *  - Array arguments of the query and solve entry points are DEVICE pointers (HBM) unless
 *    the parameter name ends in _host.  `stream` is a hipStream_t passed as void* (NULL =
 *    the default stream).  Solve and query calls are stream-ordered; entry points that fill
 *    a host-side report synchronise the stream before returning.
 *  - Statistics are caller-owned, FP64, structure-of-arrays, and accumulate across calls
 *    with Chan's parallel merge (point_estimation/common.h:268-310): k calls of W walks
 *    estimate the same quantity as one call of k*W walks.
 *  - Random numbers: PCG32 (O'Neill, PCG-XSH-RR 64/32).  Walk w of point p in a call uses
 *    PCG32 stream p (initseq = p) started at initstate = mix(seed, p, stream_offset + w),
 *    where mix is a chain of three SplitMix64 finalisers:
 *      h = f(seed ^ 0x9e3779b97f4a7c15); h = f(h ^ p * 0xd6e8feb86659fd93);
 *      initstate = f(h ^ (stream_offset + w) * 0xa0761d6478bd642f)   (64-bit wrap-around,
 *    the walk index taken modulo 2^32), f(z) = SplitMix64's output function.  The walks of a
 *    point therefore start at distinct, unrelated positions of the point's stream, results
 *    are reproducible and independent of the launch shape, and k calls with advancing
 *    stream_offset equal one long call.  The reference seeds one generator per point from
 *    the clock and a stream counter (core/sampling.h:36-41) and draws its walks from it in
 *    turn; parity is statistical.
 */
#ifndef ZOMBIE_AMD_H
#define ZOMBIE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t zg_status;
#define ZG_OK 0
#define ZG_ERR_INVALID_ARGUMENT 1
#define ZG_ERR_UNSUPPORTED 2
#define ZG_ERR_DEVICE 3
#define ZG_ERR_OUT_OF_MEMORY 4
#define ZG_ERR_NO_DEVICE 5

/* ------------------------------------------------------------------------------------ */
/* Settings: POD mirror of WalkSettings (point_estimation/common.h:19-91), same order.   */
/* ------------------------------------------------------------------------------------ */
typedef struct zg_walk_settings {
    float epsilon_shell_for_absorbing_boundary;
    float epsilon_shell_for_reflecting_boundary;
    float silhouette_precision;
    float russian_roulette_threshold;
    float splitting_threshold;
    int32_t max_walk_length;
    int32_t steps_before_applying_tikhonov;
    int32_t steps_before_using_maximal_spheres;
    uint8_t solve_double_sided;
    uint8_t use_gradient_control_variates;
    uint8_t use_gradient_antithetic_variates;
    uint8_t use_cosine_sampling_for_derivatives;
    uint8_t ignore_absorbing_boundary_contribution;
    uint8_t ignore_reflecting_boundary_contribution;
    uint8_t ignore_source_contribution;
    uint8_t print_logs;
} zg_walk_settings;

/* The 4-argument constructor of WalkSettings (common.h:21-40): silhouette precision 1e-3,
 * Russian roulette off (0), splitting off (FLT_MAX), Tikhonov at step 0, maximal spheres
 * after max_walk_length steps, gradient control/antithetic variates on. */
void zg_walk_settings_default(zg_walk_settings* ws, float epsilon_shell_absorbing,
                              float epsilon_shell_reflecting, int32_t max_walk_length,
                              int32_t solve_double_sided);

/* ------------------------------------------------------------------------------------ */
/* PDE data: device-representable replacement of the PDE<T,DIM> callbacks (pde.h:19-45). */
/* ------------------------------------------------------------------------------------ */
/* function kinds; params[] meaning per kind */
#define ZG_FN_ZERO 0            /* 0                                                        */
#define ZG_FN_CONSTANT 1        /* p0                                                       */
#define ZG_FN_EXP_COS 2         /* p0 * exp(p1*x) * cos(p1*y)           (harmonic)          */
#define ZG_FN_SIN_PRODUCT 3     /* p0 * prod_{i<DIM} sin(p1*x_i)        (-Lap u = DIM p1^2 u) */
#define ZG_FN_CUBIC_HARMONIC 4  /* p0 * (x^3 - 3 x y^2)                 (harmonic)          */
#define ZG_FN_EXP_LINEAR 5      /* p0 * exp(p1*x + p2*y + p3*z)         (Lap u = |a|^2 u)   */
#define ZG_FN_DENSE_GRID 6      /* values on a cell-centred grid (see zg_function.grid)    */

/* which quantity of the manufactured field u a descriptor evaluates */
#define ZG_EVAL_VALUE 0              /* u(x)                                                 */
#define ZG_EVAL_NORMAL_DERIVATIVE 1  /* s * grad u(x) . n, s = -1 on the normal-aligned side  */
#define ZG_EVAL_SCREENED_SOURCE 2    /* -Lap u(x) + absorption * u(x)                        */
#define ZG_EVAL_ROBIN 3              /* s * grad u(x) . n + params[5] * u(x): Robin data h for */
                                     /* du/dn + kappa u = h with kappa = params[5]            */

/* ZG_FN_DENSE_GRID: DenseGrid<float,1,DIM> (utils/dense_grid.h:30-356) with the semantics of
 * its getDenseGridCallback0: grid[] holds shape[0] x shape[1] (x shape[2]) cell values, the
 * last axis fastest, for cells centred at grid_min + (i + 1/2) (grid_max - grid_min) / shape;
 * interpolate = 0 returns the value of the cell containing x (clamped to the grid),
 * 1 bi/trilinear interpolation between cell centres, clamped to the edge values.  The value is
 * returned whatever `eval` says.  grid points to device memory for the GPU solvers (host memory
 * for a CPU caller), unused dimensions of a 2D grid have shape 1. */
typedef struct zg_function {
    int32_t kind;      /* ZG_FN_*   */
    int32_t eval;      /* ZG_EVAL_* */
    float params[6];
    const float* grid;
    int32_t shape[3];
    int32_t interpolate;
    float grid_min[3];
    float grid_max[3];
} zg_function;

typedef struct zg_pde {
    float absorption_coeff;                  /* pde.h:25, >= 0                        */
    uint8_t is_source_constant;              /* pde.h:26                              */
    uint8_t are_robin_conditions_pure_neumann; /* pde.h:27                            */
    uint8_t are_robin_coeffs_nonnegative;    /* pde.h:28                              */
    uint8_t reserved;
    zg_function source;                      /* pde.h:31  source(x)                   */
    zg_function dirichlet;                   /* pde.h:34  dirichlet(x, normalAligned)  */
    zg_function robin;                       /* pde.h:37  robin(x, n, normalAligned)   */
    zg_function robin_coeff;                 /* pde.h:38  robinCoeff(x, n, aligned)    */
} zg_pde;

/* ------------------------------------------------------------------------------------ */
/* Scene: boundary meshes + wide BVHs in HBM.                                            */
/* ------------------------------------------------------------------------------------ */
typedef struct zg_scene zg_scene;

typedef struct zg_build_options {
    int32_t bvh_width;  /* 4 or 8 children per node                                    */
    int32_t leaf_size;  /* max primitives per leaf, 1..8                                */
    int32_t sah_bins;   /* binned-SAH buckets per axis, 4..64                           */
    int32_t reserved;
} zg_build_options;

typedef struct zg_scene_info {
    int32_t dim;
    int32_t bvh_width;
    int64_t n_absorbing_prims, n_reflecting_prims;
    int64_t n_absorbing_nodes, n_reflecting_nodes;
    int32_t absorbing_depth, reflecting_depth;
    int64_t node_bytes;          /* bytes of one packed node                              */
    int64_t prim_bytes;          /* bytes of one packed primitive                         */
    int64_t device_bytes;        /* total HBM held by the scene                           */
    int64_t n_silhouette_vertices;  /* silhouette candidates: vertices (2D) / edges (3D) */
} zg_scene_info;

/* Builds the device scene from a host mesh (positions_host [n_vertices*dim], indices_host
 * [n_prims*dim]: segments in 2D, triangles in 3D).  is_reflecting_host (nullable, per
 * primitive) partitions the mesh into the absorbing (Dirichlet) part and the reflecting
 * (Neumann) part, the role of partitionBoundaryMesh (fcpw_geometric_queries.h:460-510).
 * domain_min/max_host give the bounding domain of outsideBoundingDomain
 * (geometric_queries.h:376-379).  Replaces FcpwDirichletBoundaryHandler /
 * FcpwNeumannBoundaryHandler::buildAccelerationStructure (fcpw_geometric_queries.h:537-679)
 * and populateGeometricQueriesFor{Dirichlet,Neumann}Boundary (:1062-1304); for the
 * reflecting part the silhouette candidates of computeSilhouettes (:631, :667) -- vertices in
 * 2D, edges in 3D -- are built with getIgnoreCandidateSilhouetteCallback(solve_double_sided,
 * silhouette_precision) (:1029-1038; 3D dihedral angle as :955-967 measures it). */
zg_status zg_scene_create(int32_t dim, const float* positions_host, int64_t n_vertices,
                          const int32_t* indices_host, int64_t n_prims,
                          const uint8_t* is_reflecting_host,
                          const float* domain_min_host, const float* domain_max_host,
                          int32_t solve_double_sided, float silhouette_precision,
                          const zg_build_options* opts, zg_scene** out);
zg_status zg_scene_destroy(zg_scene* scene);

/* SDF-grid Dirichlet geometry (utils/sdf_grid_geometric_queries.h:87-190 SdfGrid,
 * :247-272 populateGeometricQueriesForDirichletBoundary): the whole boundary is absorbing and
 * given by signed distances at the cell centres of a grid over [grid_min, grid_max]
 * (values[shape[0] x shape[1] (x shape[2])], last axis fastest, host memory; copied), read with
 * bi/trilinear interpolation clamped to the edge.  zg_closest_point returns |sdf|, the
 * steepest-descent projection onto the zero level set and its normal (prim -1);
 * zg_inside_domain tests sdf < 0; zg_intersect_ray sphere-traces the zero level set;
 * zg_wos_solve walks with radii |sdf|, gradient points and walk splitting included (no
 * control-variate gradients).  Walks leaving [domain_min, domain_max]
 * (NULL: the grid box) escape. */
zg_status zg_sdf_scene_create(int32_t dim, const float* values_host, const int32_t* shape_host,
                              const float* grid_min_host, const float* grid_max_host,
                              const float* domain_min_host, const float* domain_max_host, zg_scene** out);

/* Robin coefficient bounds of the reflecting primitives (FcpwRobinBoundaryHandler<2>, <3>::
 * buildAccelerationStructure / updateCoefficientValues, fcpw_geometric_queries.h:726-1000):
 * min_coeff_host[p] <= max_coeff_host[p] for every primitive p of the scene (arrays of
 * n_prims entries; entries of absorbing primitives are ignored); both NULL clears them
 * (Neumann conditions).  With bounds set, star radii (zg_star_radius, zg_wost_solve) also
 * respect the Robin bound of every segment or triangle with max coefficient > 0
 * (reflectance_boundary_bvh/geometry.h:86-236, robin_bounds.h:12-80).  The pointwise
 * coefficient of the walk throughput is pde.robin_coeff.  Mesh scenes. */
zg_status zg_scene_set_robin_coefficients(zg_scene* scene, const float* min_coeff_host,
                                          const float* max_coeff_host);
zg_status zg_scene_get_info(const zg_scene* scene, zg_scene_info* info_host);

/* ------------------------------------------------------------------------------------ */
/* Batched geometric queries (device arrays).                                            */
/* ------------------------------------------------------------------------------------ */
/* boundary: 0 = absorbing, 1 = reflecting */
/* computeDistToAbsorbingBoundary / projectToAbsorbingBoundary (geometric_queries.h:73,78;
 * fcpw_geometric_queries.h:1069-1096): unsigned distance, closest point, unit normal of the
 * closest primitive (segment: (s.y,-s.x)/|s|; triangle: (b-a)x(c-a)/|.|, geometry_helpers.h
 * :15-21,77-81) and the index of that primitive in the caller's mesh.  Outputs nullable
 * except dist.  dist = FLT_MAX when the boundary is empty. */
/* Copies the packed wide BVH of one boundary part (0 absorbing, 1 reflecting) to the host:
 * nodes_host gets n_*_nodes * node_bytes / 4 floats, prims_host n_*_prims * prim_bytes / 4
 * floats (zg_scene_info), in the device layout (lo[DIM][W], hi[DIM][W], child[W] as int32
 * bits, padded to 64-byte multiples; primitives in leaf order).  Either pointer may be NULL.
 * No counterpart in the reference (FCPW's BVH is internal): it lets a test or a profiler
 * traverse exactly what the GPU traverses. */
zg_status zg_scene_copy_bvh(const zg_scene* scene, int32_t boundary, float* nodes_host, int64_t node_floats,
                            float* prims_host, int64_t prim_floats);

zg_status zg_closest_point(const zg_scene* scene, int32_t boundary, int64_t n,
                           const float* pts, float* dist, float* closest, float* normal,
                           int32_t* prim, void* stream);

/* insideDomain (geometric_queries.h:357-371): a point is inside when, for every axis, the
 * ray along +axis hits the whole boundary (absorbing + reflecting) an odd number of times. */
zg_status zg_inside_domain(const zg_scene* scene, int64_t n, const float* pts,
                           uint8_t* inside, void* stream);

/* intersectReflectingBoundary / intersectAbsorbingBoundary first hit
 * (fcpw_geometric_queries.h:1097-1115,1178-1196): ray origin + t dir, 0 <= t <= t_max.
 * hit[i] = 1 when a primitive is hit; t, point, normal of the first hit (t = t_max and a zero
 * point on a miss).  On SDF-grid scenes (boundary 0) the ray is sphere-traced against the zero
 * level set (SdfGrid::intersectZeroLevelSet, sdf_grid_geometric_queries.h:147-205). */
zg_status zg_intersect_ray(const zg_scene* scene, int32_t boundary, int64_t n,
                           const float* origins, const float* dirs, const float* t_max,
                           uint8_t* hit, float* t, float* point, float* normal, void* stream);

/* computeStarRadiusForReflectingBoundary for a Neumann boundary (fcpw_geometric_queries.h
 * :1254-1268): the distance to the closest silhouette point of the reflecting boundary
 * (vertex in 2D, point on an edge in 3D) within max_radius, clamped below by min_radius.
 * Mesh scenes. */
zg_status zg_star_radius(const zg_scene* scene, int64_t n, const float* pts,
                         const float* min_radius, const float* max_radius,
                         float silhouette_precision, const uint8_t* flip_normal_orientation,
                         float* radius, void* stream);

/* intersectsWithReflectingBoundary's visibility test: hasLineOfSight of the reflecting
 * aggregate (fcpw_geometric_queries.h:1218-1226) between p1[i] and p2[i] (offsets, if any,
 * applied by the caller).  visible[i] = 1 when no reflecting segment (2D) or triangle (3D)
 * crosses the open segment.  Mesh scenes. */
zg_status zg_line_of_sight(const zg_scene* scene, int64_t n, const float* p1, const float* p2,
                           uint8_t* visible, void* stream);

/* sampleReflectingBoundary (fcpw_geometric_queries.h:1227-1242): a point on the reflecting
 * boundary near the ball |y - x| <= radius drawn with the DIM random numbers u[DIM*i ..] by a
 * stochastic descent of the reflecting BVH (children overlapping the ball weighted by
 * branchTraversalWeight of their box centre, primitives by that weight times their length or
 * area), then a uniform point on the primitive from the remaining DIM - 1 numbers.  ok[i] = 0
 * when nothing overlaps; pdf[i] is the density of the sample per unit length (2D) or area
 * (3D).  pt / normal / pdf nullable.  Mesh scenes. */
zg_status zg_sample_reflecting(const zg_scene* scene, int64_t n, const float* x, const float* radius,
                               const float* u, uint8_t* ok, float* pt, float* normal, float* pdf,
                               void* stream);

/* ------------------------------------------------------------------------------------ */
/* Solves.                                                                               */
/* ------------------------------------------------------------------------------------ */
/* SampleType (common.h:139-143) */
#define ZG_SAMPLE_IN_DOMAIN 0
#define ZG_SAMPLE_ON_ABSORBING_BOUNDARY 1
#define ZG_SAMPLE_ON_REFLECTING_BOUNDARY 2

/* EstimationQuantity (common.h:145-149) */
#define ZG_ESTIMATE_SOLUTION 0
#define ZG_ESTIMATE_SOLUTION_AND_GRADIENT 1
#define ZG_ESTIMATE_SKIP 2

/* A vector field sampled on a regular grid over [lo, hi] (res[0] x res[1] x res[2] nodes,
 * x fastest; 4 floats per node, the 4th unused), read with (bi/tri)linear interpolation
 * clamped to the box.  Used as a control-variate gradient field (see cv_gradient_grid). */
typedef struct zg_grid_field {
    const float* data;   /* device, res[0]*res[1]*res[2]*4 floats (res[2] = 1 in 2D) */
    int32_t res[3];
    float lo[3], hi[3];
} zg_grid_field;

/* SamplePoint (common.h:151-195), structure of arrays, device pointers. */
typedef struct zg_sample_points {
    int64_t n;
    const float* pt;                         /* [n*dim]                                   */
    const float* normal;                     /* [n*dim] nullable (zero)                   */
    const uint8_t* type;                     /* [n] nullable (all ZG_SAMPLE_IN_DOMAIN)     */
    const float* dist_to_absorbing_boundary; /* [n] required (see zg_closest_point)       */
    const float* dist_to_reflecting_boundary;/* [n] nullable                              */
    const uint8_t* estimate_boundary_normal_aligned; /* [n] nullable                     */
    float* first_sphere_radius;              /* [n] nullable; written when a point has no  */
                                             /*     previous estimates                     */
    const int32_t* n_walks;                  /* [n] walks per point                        */
    const float* cv_gradient;                /* [n*dim] nullable: gradient of a control     */
                                             /* variate c at the point; every walk adds the */
                                             /* zero-mean term -grad c(x0).(x1 - x0) of its */
                                             /* first step (x1 uniform on the first sphere) */
    const zg_grid_field* cv_gradient_grid;   /* host struct, nullable: gradient field G of a */
                                             /* control variate; every later step k adds    */
                                             /* -T_k G(x_k).(x_{k+1} - x_k), zero mean too  */
    int32_t n_walks_max;                     /* host-side upper bound of n_walks[]         */
    int32_t reserved;
    const uint8_t* estimation_quantity;      /* [n] nullable (all ZG_ESTIMATE_SOLUTION)     */
    const float* direction_for_derivative;   /* [n*dim] nullable (e_0): direction of the    */
                                             /* directional derivative of gradient points  */
} zg_sample_points;

/* SampleStatistics (common.h:200-395) in FP64, structure of arrays, device pointers. */
typedef struct zg_statistics {
    int64_t* n_solution_estimates;
    double* solution_mean;
    double* solution_m2;
    int64_t* total_walk_length;
    int64_t* total_splits;
    /* gradient members, required when a point asks for ZG_ESTIMATE_SOLUTION_AND_GRADIENT */
    int64_t* n_gradient_estimates;
    double* gradient_mean;                    /* [n*dim]                                    */
    double* gradient_m2;                      /* [n*dim]                                    */
    double* total_first_source_contribution;  /* [n]                                        */
    double* total_derivative_contribution;    /* [n]                                        */
} zg_statistics;

/* WalkCompletionCode (common.h:132-137) */
#define ZG_REACHED_ABSORBING_BOUNDARY 0
#define ZG_TERMINATED_WITH_RUSSIAN_ROULETTE 1
#define ZG_EXCEEDED_MAX_WALK_LENGTH 2
#define ZG_ESCAPED_DOMAIN 3

typedef struct zg_solve_options {
    uint64_t seed;            /* PCG32 initstate of every walk of the call              */
    uint32_t stream_offset;   /* added to the walk index in the PCG32 stream id          */
    int32_t walks_per_lane;   /* walks of one point a lane claims at a time (1..64)       */
    int32_t regroup_interval; /* iterations between wavefront ballots that refill idle lanes */
    int32_t step_interval;    /* iterations (one BVH node visit each) between walk steps   */
    int32_t waves_per_cu;     /* persistent grid size = CUs * waves_per_cu (0 = auto)    */
    int32_t count_visits;     /* 1: instrumented kernel counting the BVH node/primitive visits
                                 of its own queries; 2: instead, at every query point, those of
                                 an exact nearest-first traversal from the root (the query's
                                 algorithmic work; walk on spheres only)                   */
    float distance_tolerance; /* tau in [0, 0.5): 0 = exact closest-point radii.  tau > 0 lets
                                 the distance query stop once every unexplored box lies within
                                 (1 - tau) of the best distance and walk with the certified
                                 lower bound instead (a ball inside the domain keeps walk on
                                 spheres unbiased; walks get slightly longer).  Queries within
                                 2 epsilon of the boundary and terminal projections stay exact. */
    int32_t wost_interleave_queries; /* walk on stars: 0 (default) runs each query of a step (star
                                 radius, ray, visibility, distance) to completion inside the step;
                                 1 advances every lane's query one BVH item per loop iteration,
                                 interleaved across the lanes of a wave, with the step logic every
                                 step_interval iterations (same walks and estimates, slower on
                                 the measured scenes; not with splitting or gradient points).
                                 count_visits 1 in walk on stars uses the interleaved kernel's
                                 counts, which cover the reflecting and silhouette BVHs too. */
} zg_solve_options;

void zg_solve_options_default(zg_solve_options* opts);

typedef struct zg_solve_report {
    int64_t completion_codes[4]; /* walks per WalkCompletionCode                         */
    int64_t total_steps;         /* sum of walkLength over completed walks               */
    int64_t total_walks;         /* walks started                                        */
    int64_t node_visits;         /* count_visits only: BVH nodes visited (all queries; walk on
                                    stars: the absorbing BVH's)                           */
    int64_t prim_visits;         /* count_visits only: primitives tested (as node_visits) */
    int64_t queries;             /* count_visits only: closest-point / traversal queries */
    float kernel_ms;             /* walk-kernel time, HIP events on `stream`             */
    float reserved;
    /* walk on stars with count_visits 1 (interleaved kernel): nodes and primitives read on the
       reflecting boundary's BVH (ray, visibility, boundary sample, Robin bound) and on the
       silhouette BVH (star radius; primitives = silhouette candidates) */
    int64_t reflecting_node_visits, reflecting_prim_visits;
    int64_t silhouette_node_visits, silhouette_prim_visits;
} zg_solve_report;

/* WalkOnSpheres<float,DIM>::solve (walk_on_spheres.h:131-168 -> estimateSolution :320-414,
 * walk :228-291, getTerminalContribution :293-318) for every point in one persistent launch.
 * report_host nullable; when given, the call synchronises `stream` and fills it.  Without a
 * report (and without gradient points) the call returns as soon as its work is enqueued on
 * `stream`: the workspace is freed in stream order, the statistics are final once the stream
 * has reached that point, and device-side errors (a traversal-stack overflow, an instance index
 * outside the batch) are reported by the next zg_synchronize().
 * Points with estimation_quantity ZG_ESTIMATE_SKIP are left alone; points asking for
 * ZG_ESTIMATE_SOLUTION_AND_GRADIENT run estimateSolutionAndGradient (:416-593): first sphere of
 * radius 0.99 * dist_to_absorbing_boundary, Latin hypercube directions, antithetic pairs,
 * cosine-weighted boundary directions, gradient control variates (the point's solution mean and
 * mean first-step source term).  The pairs of a call run in stages [0,1), [1,2), [2,4), ...;
 * the control variates are read from the statistics when a stage starts (the reference updates
 * them before every pair).  With gradient points the call copies estimation_quantity to the host
 * and synchronises `stream`. */
zg_status zg_wos_solve(const zg_scene* scene, const zg_pde* pde_host,
                       const zg_walk_settings* settings_host, const zg_sample_points* points_host,
                       zg_statistics* stats, const zg_solve_options* opts_host,
                       zg_solve_report* report_host, void* stream);

/* No counterpart in the reference: zg_wos_solve for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as zg_wos_solve, but the per-point workspace (32 B per point,
 * 40 B with walk splitting, plus 128 B of counters) lives in a fixed 32 MiB device array that
 * the library holds on each GPU it runs on until the process ends; the call makes no
 * allocation.  Limits: solution estimates only (estimation_quantity must be NULL), at most
 * 1,048,572 points per call (838,857 with splitting_threshold < 1), otherwise
 * ZG_ERR_INVALID_ARGUMENT / ZG_ERR_UNSUPPORTED.  The call synchronises `stream` (the counters
 * come back to fill report_host, whose kernel_ms is 0: no timing events are created).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (other host threads,
 * other streams) race and must be serialised by the caller. */
zg_status zg_wos_solve_noalloc(const zg_scene* scene, const zg_pde* pde_host,
                               const zg_walk_settings* settings_host, const zg_sample_points* points_host,
                               zg_statistics* stats, const zg_solve_options* opts_host,
                               zg_solve_report* report_host, void* stream);

/* WalkOnStars<float,DIM>::solve (walk_on_stars.h:150-187 -> estimateSolution :518-662,
 * walk :345-490); mesh scenes (2D segments, 3D triangles), harmonic Green's function or, in
 * 3D with absorption > 0, the Yukawa ball (dampening on steps that miss the reflecting
 * boundary, Yukawa reflectance on those that hit it); Neumann or Robin reflecting conditions (h = pde.robin; with
 * are_robin_conditions_pure_neumann = 0 the throughput takes the Robin reflectance with
 * kappa = pde.robin_coeff, and zg_scene_set_robin_coefficients gives the per-primitive bounds
 * of the star radius).  Silhouette vertices (2D) or edges (3D) bound the star radius.
 * points_host->normal, type and
 * estimate_boundary_normal_aligned describe points on the reflecting boundary;
 * first_sphere_radius (nullable) records the star radius of each point's first step on the
 * first call and is reused while the point has estimates.  Gradient points run
 * estimateSolutionAndGradient (:664-848) as in zg_wos_solve, with a first ball of radius
 * 0.99 min(dist_to_absorbing_boundary, dist_to_reflecting_boundary) (the reflecting distance is
 * queried when dist_to_reflecting_boundary is NULL).  Walks split (applyWeightWindow,
 * walk_on_stars.h:309-343: parts continue from the walk's state with throughput =
 * splitting_threshold, the estimate is the sum of the parts, splits are counted in
 * total_splits) when splitting_threshold < 1, or with Robin coefficients that may be negative
 * (are_robin_coeffs_nonnegative = 0, reflectance clamped to [0, 2]) and a finite threshold;
 * splitting with gradient points is ZG_ERR_UNSUPPORTED. */
zg_status zg_wost_solve(const zg_scene* scene, const zg_pde* pde_host,
                        const zg_walk_settings* settings_host, const zg_sample_points* points_host,
                        zg_statistics* stats, const zg_solve_options* opts_host,
                        zg_solve_report* report_host, void* stream);

/* The per-point Latin hypercube tables of a gradient solve (generateStratifiedSamples<DIM-1>,
 * core/sampling.h:435-458): table t holds n_samples rows of `dims` floats and is drawn from
 * the generator of walk stream_offset of point t under the seed ~seed (see Random numbers
 * above).  A solve with seed s and stream offset o uses the
 * table of point p with 2 * n_pairs rows (n_pairs = n_walks / 2 with antithetic variates, else
 * n_walks).  out: device array of n_tables * n_samples * dims floats. */
zg_status zg_stratified_samples(int32_t dims, int64_t n_tables, int32_t n_samples, uint64_t seed,
                                uint32_t stream_offset, float* out, void* stream);

/* getSpatiallySortedPointIndices (utils/nearest_neighbor_finder.h:112-128): the index order
 * of a kd-tree over points_host[n*dim] with leaves of at most 10 points, built as nanoflann's
 * KDTreeSingleIndexAdaptor builds it (widest box dimension, midpoint clamped to the points'
 * range, ties kept together; nanoflann itself is not part of this library).  Leaves are
 * contiguous and in depth-first order, so consecutive points are spatial neighbours: sorting
 * a solve's points this way keeps a wave's walks close together.  Host arrays; runs on the
 * host (no device needed).  indices_host[n] receives a permutation of 0..n-1. */
zg_status zg_spatially_sorted_indices(int32_t dim, const float* points_host, int64_t n,
                                      int32_t* indices_host);

/* ------------------------------------------------------------------------------------ */
/* Boundary value caching (variance_reduction/boundary_value_caching.h): splatting cached  */
/* boundary and domain samples into evaluation points with the free-space kernels.         */
/* ------------------------------------------------------------------------------------ */
/* A sample cache (SamplePoint fields the splat reads), structure of arrays, device pointers. */
typedef struct zg_bvc_samples {
    int64_t n;
    const float* pt;                 /* [n*dim]                                               */
    const float* normal;             /* [n*dim] unit normals of boundary samples               */
    const uint8_t* type;             /* [n] ZG_SAMPLE_ON_ABSORBING_BOUNDARY, _ON_REFLECTING_    */
                                     /*     BOUNDARY, or ZG_SAMPLE_IN_DOMAIN (source sample)   */
    const uint8_t* normal_aligned;   /* [n] nullable: estimateBoundaryNormalAligned           */
    const float* pdf;                /* [n] sampling density of the sample                     */
    const float* solution;           /* [n] estimated (or known) solution at the sample         */
    const float* normal_derivative;  /* [n] estimated or known normal derivative               */
    const float* contribution;       /* [n] Robin data h (reflecting) or source value (domain)  */
    const float* robin_coeff;        /* [n] nullable (0)                                      */
} zg_bvc_samples;

/* Evaluation points (EvaluationPoint, boundary_value_caching.h:18-56), device pointers. */
typedef struct zg_bvc_eval_points {
    int64_t n;
    const float* pt;                 /* [n*dim]                                               */
    const uint8_t* type;             /* [n] nullable (all ZG_SAMPLE_IN_DOMAIN)                  */
    const float* dist_to_absorbing_boundary;   /* [n]                                        */
    const float* dist_to_reflecting_boundary;  /* [n] nullable (infinite)                     */
} zg_bvc_eval_points;

/* Per evaluation point and category, FP64, Chan-merged across calls; categories in the order
 * absorbing, absorbing normal-aligned, reflecting, reflecting normal-aligned, source (the five
 * SampleStatistics of EvaluationPoint).  The solution at a point is the sum of the five means
 * (getEstimatedSolution), its gradient the sum of the five gradient means. */
#define ZG_BVC_CATEGORIES 5
typedef struct zg_bvc_statistics {
    int64_t* n;                      /* [n_eval*5]                                            */
    double* solution_mean;           /* [n_eval*5]                                            */
    double* solution_m2;             /* [n_eval*5]                                            */
    double* gradient_mean;           /* [n_eval*5*dim]                                        */
    double* gradient_m2;             /* [n_eval*5*dim]                                        */
} zg_bvc_statistics;

/* BoundaryValueCaching::splat over every sample for every evaluation point
 * (boundary_value_caching.h:370-402,683-810): free-space Green's function G, Poisson kernel P
 * and their gradients at r = max(radius_clamp, |y - x|) -- harmonic, or with absorption > 0 the
 * 3D Yukawa kernels (YukawaGreensFnFreeSpace<3>, distributions.h:271-321; grad P is the
 * derivative of P, where the reference's formula at :307-317 differs) -- optional kernel regularization
 * (distributions.h:12-55), the Robin / Neumann / Dirichlet forms of the boundary estimate,
 * alpha = 2 for evaluation points on the boundary (gradient 0 there); evaluation points closer
 * than the cutoffs to the boundaries are skipped; non-finite kernels skip the sample.  One
 * thread per evaluation point, samples staged through LDS.  absorption >= 0; > 0 in 2D is
 * ZG_ERR_UNSUPPORTED (screened Poisson runs in 3D). */
zg_status zg_bvc_splat(int32_t dim, float absorption, const zg_bvc_samples* samples_host,
                       const zg_bvc_eval_points* eval_host, float radius_clamp, float kernel_regularization,
                       float robin_coeff_cutoff_for_normal_derivative, float cutoff_dist_to_absorbing_boundary,
                       float cutoff_dist_to_reflecting_boundary, zg_bvc_statistics* stats_host, void* stream);

/* Evaluates a PDE data descriptor at n points (the PDE<T,DIM> callbacks source / dirichlet /
 * robin of pde.h:30-44, as the solvers evaluate them): out[i] = f(pts[i], normals[i],
 * aligned[i]) with the absorption coefficient used by ZG_EVAL_SCREENED_SOURCE.  normals and
 * aligned nullable; device arrays. */
zg_status zg_evaluate_function(int32_t dim, const zg_function* f_host, int64_t n, const float* pts,
                               const float* normals, const uint8_t* aligned, float absorption, float* out,
                               void* stream);

/* Measurement utility (no counterpart in the reference): reads `bytes` of device memory at
 * `buf` `reps` times with float4 loads over every CU and returns the elapsed milliseconds.
 * A buffer that fits the L2 caches gives the cache read ceiling, a large one the HBM
 * ceiling, the two rooflines the walk kernels are reported against. */
zg_status zg_probe_read_bandwidth(const void* buf, int64_t bytes, int32_t reps, float* ms_out, void* stream);

/* ------------------------------------------------------------------------------------ */
/* Batched solves: many independent instances (scene + PDE each) walked by one persistent */
/* launch.  The reference solves one scene per call (walk_on_spheres.h:131-168); a batch  */
/* is the same per-point estimator with a per-point instance index.  3D scenes of one BVH */
/* width; the scenes must outlive the batch.                                              */
/* ------------------------------------------------------------------------------------ */
typedef struct zg_batch zg_batch;
typedef struct zg_graph zg_graph;

zg_status zg_batch_create(int32_t n_instances, const zg_scene* const* scenes_host, const zg_pde* pdes_host,
                          zg_batch** out);
zg_status zg_batch_destroy(zg_batch* batch);
/* points of all instances concatenated; instance_of_point [n] (device) names each point's instance */
zg_status zg_batch_wos_solve(const zg_batch* batch, const zg_walk_settings* settings_host,
                             const zg_sample_points* points_host, const int32_t* instance_of_point,
                             zg_statistics* stats, const zg_solve_options* opts_host,
                             zg_solve_report* report_host, void* stream);
/* The same solve captured once in a hipGraph (fill, walk and merge kernels over a workspace the
 * graph owns; no memset nodes).  Each zg_graph_launch replays it with fresh PCG32 streams (the
 * merge kernel advances a device-side stream offset by n_walks_max) and accumulates into the
 * same statistics arrays, which must stay allocated.  zg_graph_read_report synchronises
 * `stream` and returns the counters of the last replay (kernel_ms = 0). */
zg_status zg_batch_wos_graph_create(const zg_batch* batch, const zg_walk_settings* settings_host,
                                    const zg_sample_points* points_host, const int32_t* instance_of_point,
                                    zg_statistics* stats, const zg_solve_options* opts_host, zg_graph** out);
zg_status zg_graph_launch(zg_graph* graph, void* stream);
zg_status zg_graph_read_report(zg_graph* graph, zg_solve_report* report_host, void* stream);
zg_status zg_graph_destroy(zg_graph* graph);

/* ------------------------------------------------------------------------------------ */
/* Learned control variate: Fourier-feature MLP trained on the GPU.  No counterpart in the */
/* reference; the nearest one is the running-mean gradient control variate of            */
/* estimateSolutionAndGradient (walk_on_spheres.h:444-450).  Its input gradient feeds     */
/* zg_sample_points.cv_gradient.                                                          */
/* ------------------------------------------------------------------------------------ */
typedef struct zg_mlp zg_mlp;

typedef struct zg_mlp_config {
    int32_t in_dim;           /* 2 or 3                                                  */
    int32_t n_frequencies;    /* F Gaussian frequency vectors: features [sin(Bx), cos(Bx)] */
    int32_t hidden;           /* units per hidden layer (multiple of 64, <= 512)         */
    int32_t n_hidden_layers;  /* SiLU layers, 1..8                                       */
    float frequency_sigma;    /* B entries ~ N(0, sigma^2)                               */
    float reserved;
    uint64_t seed;            /* frequencies and initial weights                         */
} zg_mlp_config;

/* max_batch: largest row count of one train/evaluate call (work buffers are sized for it) */
zg_status zg_mlp_create(const zg_mlp_config* cfg, int64_t max_batch, zg_mlp** out);
zg_status zg_mlp_destroy(zg_mlp* mlp);
zg_status zg_mlp_parameter_count(const zg_mlp* mlp, int64_t* count);
/* fp32 master parameters, per layer l = 0..n_hidden_layers: W_l (out x in, row-major) then b_l */
zg_status zg_mlp_get_parameters(const zg_mlp* mlp, float* params_host, int64_t count);
zg_status zg_mlp_set_parameters(zg_mlp* mlp, const float* params_host, int64_t count);
/* frequency matrix B, n_frequencies x in_dim */
zg_status zg_mlp_get_frequencies(const zg_mlp* mlp, float* freq_host, int64_t count);
/* loss 0.5 mean (c(x) - y)^2 and its parameter gradient (device arrays x [batch*in_dim],
 * y [batch]); grad_host/loss_host nullable (copied back when given, which synchronises). */
zg_status zg_mlp_gradients(zg_mlp* mlp, const float* x, const float* y, int64_t batch, float* grad_host,
                           float* loss_host, void* stream);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the gradient of that loss */
zg_status zg_mlp_train_step(zg_mlp* mlp, const float* x, const float* y, int64_t batch, float learning_rate,
                            float* loss_host, void* stream);
/* c(x) and grad_x c(x) for n points (any n; processed in chunks of max_batch); outputs nullable */
zg_status zg_mlp_evaluate(zg_mlp* mlp, const float* x, int64_t n, float* value, float* gradient, void* stream);

/* ------------------------------------------------------------------------------------ */
/* Device memory helpers for callers without their own HIP runtime (ctypes, cgo, ...).   */
/* ------------------------------------------------------------------------------------ */
zg_status zg_device_count(int32_t* count);
zg_status zg_malloc(void** ptr, size_t bytes);
zg_status zg_free(void* ptr);
zg_status zg_memcpy_htod(void* dst, const void* src_host, size_t bytes);
zg_status zg_memcpy_dtoh(void* dst_host, const void* src, size_t bytes);
zg_status zg_memset(void* dst, int32_t value, size_t bytes);
/* Waits for the device; reports (and clears) the device-side errors of solves that returned
 * without synchronising. */
zg_status zg_synchronize(void);

/* Text of the last error on the calling thread (empty when none); returns its length. */
int32_t zg_last_error(char* buf, int32_t buf_len);
/* Library version "major.minor.patch". */
const char* zg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ZOMBIE_AMD_H */
