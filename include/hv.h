/*
 * hv.h -- C ABI of libhv.so, the MI355X (gfx950) raycast Voronoi path.
 *
 * The library replaces the data-parallel hot path of HighVoronoi.jl v1.4.2:
 * the raycast vertex walk (src/sysvoronoi.jl, src/raycast.jl), the half-space
 * nearest-neighbour / range queries behind it (src/searchtrees.jl, src/extended.jl,
 * src/NearestNeighborModified/) and, later, the volume / interface integration
 * (src/polyintegrator.jl, src/mcintegrator.jl).  Plain pointers and sizes only.
 *
 * Conventions (shared by every entry point):
 *  - Points are d x n column-major FP64 (Julia's Vector{SVector{d,Float64}} layout).
 *  - Keys are sorted, 1-based int64: generator i is i, boundary plane p (1-based) is
 *    n + p, exactly the reference's vertex keys (src/extended.jl:27-34).
 *  - Every call returns an int status: HV_OK (0) or a negative HV_ERR_* code; the
 *    message of the last failure on a context is hv_last_error(ctx).
 *  - A context owns one HIP stream and the device copies of its points; it is not
 *    thread-safe (one per host thread, like one RaycastIncircleSkip per Julia thread,
 *    src/raycast-types.jl:298-307).  Calls are synchronous on return.
 *  - The caller allocates every host output; variable-size outputs use the two-step
 *    counts -> export protocol.
 *  - A mesh keeps what it was built on (generators, boundary, periodic images), so a later
 *    hv_set_points on its context does not change it, and any context on the mesh's device
 *    can integrate it.  Distinct meshes may be exported, integrated or freed from different
 *    host threads at once; one mesh must not be freed while another call uses it.
 */
#ifndef HV_H
#define HV_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define HV_OK 0
#define HV_ERR_ARG -1          /* bad argument / shape */
#define HV_ERR_DOMAIN -2       /* a generator lies outside the domain (src/boundary.jl:437-441) */
#define HV_ERR_FEW_POINTS -3   /* n <= d (src/sysvoronoi.jl:25-27) */
#define HV_ERR_DEGENERATE -4   /* input not in general position (near-tie / co-spherical vertex) */
#define HV_ERR_DEVICE -5       /* HIP runtime error */
#define HV_ERR_CAPACITY -6     /* a device table exceeded its capacity */
#define HV_ERR_UNSUPPORTED -7  /* configuration outside what the GPU path implements */
#define HV_ERR_STATE -8        /* call order (e.g. no points set) */

#define HV_PREC_FP64 0         /* plain FP64 candidate scan */
#define HV_PREC_FP32 1         /* FP32 VALU candidate filter (coordinates relative to a leaf-group
                                  origin, error-bounded threshold) + FP64 re-verification of every
                                  survivor: same vertex set and coordinates as HV_PREC_FP64;
                                  leaves of 32 points */
#define HV_PREC_BF16 2         /* BF16 MFMA candidate filter (hi/lo split, error-bounded margin)
                                  + FP64 re-verification of every survivor: the vertex set and
                                  coordinates equal those of HV_PREC_FP64; leaves of 32 points */

typedef struct hv_ctx hv_ctx;
typedef struct hv_mesh hv_mesh;
typedef struct hv_integral hv_integral;

typedef struct hv_opts {
    int32_t precision;        /* HV_PREC_* */
    int32_t leaf_size;        /* generators per tree leaf (0 = default 32; HV_PREC_BF16 and
                                 HV_PREC_FP32 need 32) */
    int32_t rays_per_block;   /* scan tile: at most this many rays per packet (8, 16, 32 or 64; a
                                 packet is one wavefront); 0 = the autotuner's choice (hv_tune) or 64 */
    int32_t reserved0;        /* walk-level schedule: <= 0 (default) = one scan per tentative end
                                 point, the other rays of that end point wait and are carried to
                                 the next level if their edge stays open; 1 = every ray of a level
                                 in one scan; > 1 = that many chunks, edges registered after each */
    double plane_tolerance;   /* src/raycast-types.jl:229 (0 = 1e-12) */
    double ray_tol;           /* src/raycast-types.jl:230; ignored (it acts on non-general vertices,
                                 which the device walk reports as HV_ERR_DEGENERATE) */
    int64_t vertex_capacity;  /* device vertex-table capacity hint (0 = auto) */
    uint64_t seed;            /* draws the directions of the interior seeds' descents (the
                                 reference's random descent, src/raycast.jl:45-109, 211-239); the
                                 vertex set does not depend on it, the traversal does */
    int32_t packet_floor;     /* scan tile: packets are halved (down to 8 rays) while a level would
                                 give fewer packets than this; 0 = the autotuner's choice or 2048 */
    int32_t packet_min;       /* scan tile: packets are not halved below this many rays (1, 2, 4 or
                                 8); 0 = the autotuner's choice or 8 */
} hv_opts;

typedef struct hv_stats_t {
    int64_t rays;             /* walkrays cast (SRI_walkray analogue) */
    int64_t levels;           /* BFS levels of the device walk */
    int64_t vertices;         /* vertices in the mesh */
    int64_t duplicates;       /* rays whose vertex was found by another ray of its level */
    int64_t pairs_scanned;    /* (ray, generator) pairs evaluated by the candidate scan (BF16:
                                 the 16x16 MFMA tiles formed) */
    int64_t leaves_visited;   /* (packet, leaf) visits */
    int64_t near_ties;        /* rays resolved by the reference's near-tie rule */
    int64_t degenerate;       /* rays flagged non-general (key longer than d+1) */
    int64_t fallbacks;        /* mixed precision: rays whose filter passed a candidate the ray
                                 did not already hold (each such ray is re-verified in FP64) */
    int64_t fp64_rechecks;    /* mixed precision: (ray, generator) pairs re-evaluated in FP64 */
    int64_t retries;          /* rays re-cast with a larger certificate radius */
    int64_t unbounded;        /* rays without an end point */
    double max_t_relerr;      /* 0: the filter computes no t of its own (winners are chosen by
                                 the FP64 rule only) */
    double ms_total;          /* device time of the last hv_build_mesh */
    double ms_raycast;        /* device time spent in the candidate-scan kernel */
    int64_t rays_skipped;     /* rays resolved without a scan (their far end point was found by
                                 another ray of the level: same tentative key, or closed edge) */
} hv_stats_t;

/* Context and inputs ------------------------------------------------------ */
int hv_create(hv_ctx** out, int device_ordinal);
int hv_destroy(hv_ctx* ctx);
const char* hv_last_error(const hv_ctx* ctx);

/* Replaces building the searcher's point store: VoronoiNodes(xs) + Raycast(xs)
 * (src/voronoinodes.jl:14-21, src/raycast.jl:26).  xs is host memory, d x n. */
int hv_set_points(hv_ctx* ctx, int d, int64_t n, const double* xs);
/* Same, with xs already in device memory (HBM) on the context's device. */
int hv_set_points_device(hv_ctx* ctx, int d, int64_t n, const double* xs_dev);
/* Boundary planes (src/boundary.jl:22-29, Boundary(planes...) :109-131): base and outer normal,
 * d x nplanes each, normals of unit length (HV_ERR_ARG otherwise: the mirror image
 * x + 2<n, b - x> n of src/boundary.jl:206-211 assumes it); bc as the reference's Plane.BC
 * (0 Dirichlet, -1 Neumann, >0 periodic partner).  The builds take the planes of
 * cuboid(d; ...) (periodic pairs included), or Dirichlet / Neumann planes of any direction:
 * when they bound a convex polytope the walk also seeds from its vertices where exactly d
 * planes meet; when they bound nothing (no planes at all: Boundary(), src/boundary.jl:133-135;
 * or a direction with no plane in front) a ray may run to infinity: it ends without a vertex
 * (hv_stats_t.unbounded counts them) and the integrators refuse the mesh. */
int hv_set_boundary(hv_ctx* ctx, int nplanes, const double* base, const double* normal, const int16_t* bc);

/* No counterpart in the reference (tile sizes are a device matter): the autotuner.  Builds the
 * mesh of the context's current points and boundary (a non-periodic cuboid) with each candidate
 * scan tile (rays_per_block 32 or 64, packet_floor 1024, 2048 or 4096), keeps the fastest for
 * (dimension, precision, size class floor(log2 n)) in a process-wide table that later builds and
 * batched walks of that class use unless their hv_opts set the tiles, and appends it as one
 * JSON line to cache_path when that is not NULL (hv_create loads the file named by the
 * environment variable HV_TUNE_CACHE).  Returns the chosen tiles and their build time (ms). */
int hv_tune(hv_ctx* ctx, int precision, const char* cache_path, int32_t* rays_per_block, int32_t* packet_floor,
            double* ms);

/* Mesh -------------------------------------------------------------------- */
/* Replaces voronoi(mesh; searcher) with threading dispatch (src/sysvoronoi.jl:21-82):
 * the whole vertex set of the tessellation, computed on the device. */
int hv_build_mesh(hv_ctx* ctx, const hv_opts* opts, hv_mesh** out);
/* Replaces voronoi(mesh; Iter=TODO) -> _voronoi(mesh, TODO, ...) for a subset of cells
 * (src/sysvoronoi.jl:21-39), the call systematic_refine! makes with the new and the affected
 * cells (src/meshrefine.jl:199,213,215): the vertices of the cells `cells` (ncells 1-based
 * generator ids, any order, repeats ignored; in a periodic box, original ids 1..n) -- every
 * vertex whose key holds one of them, each exactly as hv_build_mesh lists it -- found by a
 * walk over those cells' edges only.  The polytope integrators integrate the listed cells of
 * such a mesh (the others get volume 0 and no rows: integrate(..., relevant, ...) of
 * src/polyintegrator.jl:33); the Monte-Carlo integrators refuse it (HV_ERR_UNSUPPORTED). */
int hv_build_mesh_cells(hv_ctx* ctx, const hv_opts* opts, int64_t ncells, const int64_t* cells, hv_mesh** out);
int hv_mesh_counts(const hv_mesh* mesh, int64_t* nvert, int32_t* max_sig, int64_t* nrays_unbounded);
/* sig: max_sig x nvert int64 (1-based, sorted, 0 padded); coords: d x nvert FP64.
 * Vertices are exported sorted lexicographically by key. */
int hv_mesh_export(const hv_mesh* mesh, int64_t* sig, double* coords);
int hv_mesh_stats(const hv_mesh* mesh, hv_stats_t* out);
int hv_mesh_free(hv_mesh* mesh);
/* Replaces refine!(VG, xs) -> systematic_refine!(domain, xs) (src/geometry.jl:404-438,
 * src/meshrefine.jl:183-217) for a non-periodic mesh of all cells: the mesh of the mesh's n
 * generators followed by the m new points xs_new (d x m host FP64), which get the ids
 * n+1..n+m as refine! numbers them publicly; boundary plane p becomes n+m+p.  A vertex of the
 * old mesh is kept when no new point lies in its circumsphere by the reference's rule
 * (clean_affected!, src/meshrefine.jl:126-150: kept iff |x_sig[1] - r| <= (1 + 1e-7) times the
 * distance from r to the nearest new point); every vertex holding a new point comes from a walk
 * of the new cells (voronoi(mesh; Iter = new cells), src/meshrefine.jl:199).  In general
 * position the result equals hv_build_mesh of the union.  Afterwards the context holds the
 * union and the old mesh's boundary; the old mesh is unchanged and stays valid.  removed (may
 * be NULL): the number of old vertices dropped.  A periodic mesh (src/meshrefine.jl on the
 * reference's ghost nodes): its n originals and the new points are meshed afresh, since the new
 * points bring their own images; removed then counts the old vertices, by their (original,
 * shift) labels, that the new mesh no longer holds (the same vertices the rule above drops,
 * new points' images included).  Meshes of some cells: HV_ERR_UNSUPPORTED. */
int hv_refine_mesh(hv_ctx* ctx, const hv_opts* opts, const hv_mesh* mesh, int64_t m, const double* xs_new,
                   hv_mesh** out, int64_t* removed);
/* The mesh in DatabaseVertexStorage's layout (HeapDataBase, src/hvdatabase.jl:34-110, under
 * VDBVertexCentral, src/vdbdatabaseref.jl:88-145): every vertex once, in hv_mesh_export order,
 * as int64 words [d+1, sig..., r...] (r as FP64 bit patterns), 2d+2 words per vertex in heap;
 * and per generator i = 1..n_ext (hv_mesh_images) the heap positions (1-based, of the vertex's
 * first word) of the vertices whose key holds i, in CSR form: refs[cell_ptr[i-1] ..
 * cell_ptr[i]), -pos where i is the key's first entry (the cell that stores the vertex, a
 * negative entry of its index list), +pos for the other generators (push_ref!).  cell_ptr has
 * n_ext+1 entries and is always written; heap and refs are written when not NULL (refs needs
 * cell_ptr[n_ext] entries, at most (d+1) nvert). */
int hv_mesh_export_db(const hv_mesh* mesh, int64_t* heap, int64_t* cell_ptr, int64_t* refs);

/* Periodic boundaries (cuboid(d; periodic=...), src/boundary.jl:510-534, where the reference
 * adds periodic copies of the nodes near the boundary, src/domain.jl:139-390).  A box whose
 * planes 2k-1 and 2k (1-based) name each other as periodic partners is periodic in direction
 * k.  hv_build_mesh then meshes the n originals together with their images ("ghosts") within
 * a margin of the box, and the mesh holds exactly the vertices of the originals' cells: every
 * vertex has at least one original generator, each in that cell's own frame, so a vertex of
 * the periodic tessellation appears once per original cell it belongs to, translated.
 * Generator ids in such a mesh: 1..n the originals, n+1..n_ext the ghosts, n_ext + p boundary
 * plane p (only non-periodic planes occur).  This call returns n_ext (n for a non-periodic
 * mesh) and, when orig / shift are not NULL, for every id 1..n_ext the original's id (n_ext
 * int64) and its shift in box lengths per direction (d x n_ext int8): generator id sits at
 * x[orig] + shift .* box size.  The integrators accept such a mesh (cells 1..n; a neighbour
 * id above n is an image); hv_raycast_batch and the queries reject periodic planes.
 * Replaces the reference's periodic node numbering (src/domain.jl:338-390) with an explicit
 * (index, shift) label per generator. */
int hv_mesh_images(const hv_mesh* mesh, int64_t* n_ext, int64_t* orig, int8_t* shift);

/* Many independent meshes in one call: mesh k has d[k] in 2..6, n[k] generators xs[k] (d[k] x n[k],
 * host memory) and the nplanes[k] boundary planes base[k], normal[k] (d[k] x nplanes[k] each)
 * with the conditions bc[k] (nplanes[k] values as in hv_set_boundary; bc == NULL or
 * bc[k] == NULL: Dirichlet), all built with *opts.  Replaces a host loop of VoronoiGeometry(xs_k, boundary_k)
 * calls (src/geometry.jl:156 -> voronoi, src/sysvoronoi.jl:21-82), one per mesh.
 * Fused walk (nworkers >= 0): the meshes in a cuboid of the reference's plane order without
 * periodic sides are walked together, one walk per dimension: a forest of per-mesh trees in
 * shared tables, global generator ids, every level one launch per kernel over the rays of all
 * meshes (packets of one mesh each), taken in runs of at most 24M expected vertices, three
 * runs at once, each on its own stream and context (those contexts, with up to 24 GiB of device
 * tables each, stay for the process and serve later batch calls on the device).  A mesh
 * with a generator outside its box or on a face gets HV_ERR_DOMAIN / HV_ERR_DEGENERATE alone.
 * Every other mesh (periodic sides, other planes; every mesh when nworkers < 0) is built on
 * |nworkers| host threads (0 = 8), each with its own context and HIP stream, taking meshes in
 * index order.  Each mesh equals its own hv_build_mesh bit for bit (keys and coordinates); its
 * hv_mesh_stats describe the walk it was part of.  out[k] receives the mesh (free each with
 * hv_mesh_free; fused meshes share one allocation, released with the last of them) or NULL,
 * status[k] its HV_* code; the return value is HV_OK or the first failing mesh's code. */
int hv_build_mesh_batch(int device_ordinal, int32_t nmesh, const int32_t* d, const int64_t* n,
                        const double* const* xs, const int32_t* nplanes, const double* const* base,
                        const double* const* normal, const int16_t* const* bc, const hv_opts* opts,
                        int32_t nworkers, hv_mesh** out, int32_t* status);

/* No counterpart in the reference: frees the contexts (streams and kept device tables) that
 * hv_build_mesh_batch's fused walks keep on the device between calls (device_ordinal < 0: on
 * every device), for callers that need the memory back.  Meshes already built stay valid; the
 * next batch call creates its contexts again.  Waits for a batch call running meanwhile. */
int hv_batch_release(int device_ordinal);

/* Ray-level entry point: one walkray per input (src/raycast.jl:125-164 via
 * raycast_des2, src/raycast.jl:794-970).  For ray k: sig[k*(d+1) ...] is the sorted
 * 1-based key of a vertex, drop[k] the 0-based position of the generator to walk
 * away from, r[k*d ...] the vertex.  Outputs: out_sig (d+1 per ray, sorted, 1-based),
 * out_r (d per ray), out_t (ray parameter), out_status (0 found, 1 unbounded,
 * 2 near-tie resolved, <0 degenerate). */
int hv_raycast_batch(hv_ctx* ctx, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                     const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status);

/* The same with the reference's degenerate inclusion (src/raycast.jl:926-950): out_sig holds
 * max_sig (>= d+1) ids per ray, 0 padded, and out_len the key's length.  A ray whose new vertex
 * has more than d + 1 generators and planes on its sphere (within 1e-10 relative; the reference
 * keeps every candidate within the rounding of its key) gets the whole key, status 3 and its
 * vertex r + t u in out_r; a longer key than max_sig leaves status < 0 and out_len -1.  The
 * edges of such a vertex (FastEdgeIterator, src/edgeiterate.jl:82-295) are not enumerated. */
int hv_raycast_batch_keys(hv_ctx* ctx, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                          const double* r, int32_t max_sig, int64_t* out_sig, int32_t* out_len, double* out_r,
                          double* out_t, int32_t* out_status);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  hv_raycast_batch
 * with every pointer in device memory on the context's device (sig, drop, r in; out_* out, the
 * same layouts), allocating nothing: its per-ray state lives in one static device array per GPU
 * (16 MiB, at most 32768 rays per call: HV_ERR_CAPACITY beyond), and it runs on the search tree
 * the context's last hv_build_mesh or hv_raycast_batch built over its current points and
 * boundary (HV_ERR_STATE when there is none, or none with the tables of this precision: a
 * BF16 or FP32 call needs a build or a ray batch of that precision first).  Synchronous.  Calls
 * on one GPU share the array: calls that overlap on one GPU (several host threads) race. */
int hv_raycast_batch_noalloc(hv_ctx* ctx, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                             const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status);

/* Half-space nearest neighbour (src/extended.jl:108-142 with the predicate of
 * src/raycast.jl:385-392): for query k, the generator minimising |x - q_k| among
 * those with <x, u_k> > c_k (c_k = -inf disables the filter).  idx is 1-based
 * (0 when none), dist the Euclidean distance (inf when none). */
int hv_nn_halfspace_batch(hv_ctx* ctx, int64_t nq, const double* q, const double* u, const double* c,
                          int64_t* idx, double* dist);

/* Range query (src/searchtrees.jl:245, inclusive radius): generators with
 * |x - q_k| <= radius_k.  Two-step: pass idx == NULL to receive counts in ptr[1..nq]
 * (ptr[0] = 0, prefix sums), then call again with idx of size ptr[nq]. */
int hv_inrange_batch(hv_ctx* ctx, int64_t nq, const double* q, const double* radius, int64_t* ptr, int64_t* idx);

/* Integration ------------------------------------------------------------ */
/* Exact cell volumes and interface areas of a mesh (of the generators and boundary it was
 * built on; ctx, any context on the mesh's device, lends its stream and scratch memory and
 * keeps its own points): the integrator=VI_POLYGON path, Polygon_Integrator + _integrate
 * (src/polyintegrator.jl:33, :80-468; driver src/integrate.jl:252-382; neighbours
 * src/neighbors.jl:214-258).  General position only (HV_ERR_DEGENERATE otherwise). */
int hv_integrate_polytope(hv_ctx* ctx, const hv_mesh* mesh, hv_integral** out);
/* The same with an integrand f: R^d -> R^nf (1 <= nf <= 8), the VoronoiGeometry(...;
 * integrand=f) path of the polytope integrators (src/polyintegrator.jl:174-353, the rule of
 * src/fastpolyintegrator.jl:272-476): bulk integrals of f over the cells and interface
 * integrals over each (cell, neighbour) row, by the reference's quadrature (trapezoid on
 * edges; chains of face centres; f at each facet centre and at the cell's generator; exact
 * for affine f).  f is called on the host, a few times per integration, with npts points
 * (d x npts column-major) and fills values (nf x npts); user is passed through. */
typedef void (*hv_integrand_fn)(void* user, int64_t npts, const double* pts, double* values);
int hv_integrate_polytope_f(hv_ctx* ctx, const hv_mesh* mesh, int32_t nf, hv_integrand_fn f, void* user,
                            hv_integral** out);
/* Directional Monte-Carlo volumes and areas, ndir rays per cell: the integrator=VI_MONTECARLO
 * path, Montecarlo_Integrator (src/mcintegrator.jl:29-226, src/cleanup_cell.jl:6-48), with
 * independent directions per cell drawn from a counter-based stream keyed by seed.  At most
 * 2^24 - 1 cells per call (HV_ERR_CAPACITY beyond). */
int hv_integrate_mc(hv_ctx* ctx, const hv_mesh* mesh, int64_t ndir, uint64_t seed, hv_integral** out);
/* Point-sample Monte-Carlo volumes and areas: the Monte-Carlo integration of the cells
 * (Montecarlo_Integrator, src/mcintegrator.jl:29-226, with cleanup_cell, src/cleanup_cell.jl:6-48)
 * computed from nsamples stratified uniform samples of the domain box instead of directions per
 * cell: each sample's nearest and second-nearest generators (tree search, FP64 distances) give
 * volume_i = V_box * (samples in cell i) / n and, for each interface, the samples c1 within
 * eps = slab * (V_box / n_generators)^(1/d) of it and c2 within 2 eps (bisector of the two
 * nearest generators, or a boundary plane): area = V_box * (2 c1 - c2 / 2) / (n * eps) (the
 * first-order slab bias extrapolated away), generator interfaces averaged over both sides.
 * slab <= 0 selects 0.03.  nsamples is rounded up to whole strata (multiples of 64), at most
 * 2^32 per call (HV_ERR_CAPACITY beyond).  Periodic boxes: samples are drawn in the box and
 * counted in the original cell they fall in modulo the period. */
int hv_integrate_mc_points(hv_ctx* ctx, const hv_mesh* mesh, int64_t nsamples, uint64_t seed, double slab,
                           hv_integral** out);
/* Voronoi_Integral's parity surface (src/integral.jl:11-21): volumes[ncells]; for cell i
 * its neighbours nbr[nbr_ptr[i] .. nbr_ptr[i+1]) sorted, 1-based (planes n+p), and the
 * interface areas aligned with them.  ms (optional) receives the device time. */
int hv_integral_counts(const hv_integral* I, int64_t* ncells, int64_t* nnbr);
/* No counterpart in the reference (a measurement): for hv_integrate_mc_points, the (sample,
 * generator) distances its sampling kernel evaluated and that kernel's device time in ms
 * (0 and 0 for the other integrators). */
int hv_integral_work(const hv_integral* I, int64_t* work, double* ms_kernel);
int hv_integral_export(const hv_integral* I, double* vol, int64_t* nbr_ptr, int64_t* nbr, double* area, double* ms);
/* Integrand results of hv_integrate_polytope_f (Voronoi_Integral's bulk_integral and
 * interface_integral, src/integral.jl:11-21): nf, bulk (nf x ncells) and iface (nf x nnbr,
 * aligned with nbr); nf = 0 for an integral computed without an integrand. */
int hv_integral_export_f(const hv_integral* I, int32_t* nf, double* bulk, double* iface);
int hv_integral_free(hv_integral* I);

#ifdef __cplusplus
}
#endif
#endif
