// hv_internal.h -- host-side structures shared by the C ABI and the walk driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <cstdlib>
#include <vector>
#include "../../include/hv.h"
#include "hv_device.h"

// Device buffer pool: tables survive between builds so a steady-state build does no
// hipMalloc; a finished mesh takes ownership of its vertex tables and returns them.
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
};

// One lock per pool: a mesh returns its tables to the pool of the context that built it
// (hv_mesh_free), possibly from another host thread than the one using that context.
struct DevPool {
    int device = 0;
    std::mutex mu;
    std::map<std::string, DBuf> bufs;
    ~DevPool();
    // returns nullptr (and keeps the old buffer) when hipMalloc fails
    void* get(const std::string& name, size_t bytes, bool keep_contents = false, hipStream_t s = nullptr);
    // at least `bytes`, allocated exactly (no growth margin), contents not kept
    void* reserve(const std::string& name, size_t bytes, hipStream_t s);
    DBuf take(const std::string& name);
    void give(const std::string& name, DBuf b);
    size_t total();   // bytes held
};

// A device allocation shared by a context and the meshes built from it (the generator
// set a mesh refers to stays alive as long as the mesh does).
struct DevArr {
    int device = 0;
    void* p = nullptr;
    size_t n = 0;
    // a view into another allocation (which it keeps alive) instead of an allocation of its own
    std::shared_ptr<DevArr> parent;
    ~DevArr();
};
using ImgOrig = std::shared_ptr<const std::vector<int32_t>>;
using ImgShift = std::shared_ptr<const std::vector<int8_t>>;

struct hv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 0;
    int64_t N = 0;
    std::shared_ptr<DevPool> pool;
    // the working generator set (N x D FP64): written in place only while no mesh holds it
    std::shared_ptr<DevArr> xw;
    // boundary (host copies)
    int P = 0;
    std::vector<double> pbase, pnormal;
    std::vector<int16_t> pbc;
    // the domain the planes bound (set by hv_domain from the current planes): the cuboid of
    // the reference's plane order, or a general bounded convex polytope with its bounding box
    // (the Morton frame and the error scales) and its simple vertices (d planes meeting there:
    // the walk's corner seeds; point ncorner x D, plane indices ncorner x D ascending)
    bool dom_general = false;
    bool dom_unbounded = false;   // no bounded domain (Boundary() or too few planes): rays may
                                  // run to infinity; the frame is the generators' bounding box
    double dom_lo[HV_MAXD] = {}, dom_hi[HV_MAXD] = {};
    std::vector<double> dom_corner;
    std::vector<int> dom_corner_pl;
    std::string err;
    hipEvent_t ev[4];
    // streams beside `stream` (created on first use): the trees of a forest are built on them
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    // periodic boundaries: a build on planes with periodic partners replaces the working point
    // set (N points) by the nreal originals followed by their images ("ghosts"); the originals
    // stay in xorig.
    int64_t nreal = 0;
    bool extended = false;
    std::shared_ptr<DevArr> xorig;
    double delta = 0.0;                  // margin of the extended set beyond the box
    double rmax = 0.0;                   // largest circumradius of the periodic mesh
    ImgOrig img_orig;                    // per working-set id: the original (0-based)
    ImgShift img_shift;                  // per working-set id: d shifts in box lengths
    // the search tree of the last tree build over the context's own points and planes (the
    // ray-level entry point without allocation runs on it); cleared when they change
    bool tree_ok = false;
    bool tree_on_plane = false;   // the last tree build found a generator on a boundary plane
    int tree_tables = 0;                 // HV_PREC_* tables built beside it
    int tree_leaf = 0;
    DevGeo tree_g{};
    std::shared_ptr<DevArr> tree_xw;
};

struct hv_mesh {
    std::shared_ptr<DevPool> pool;   // where the vertex tables go back on free (null: freed)
    int device = 0;
    int D = 0;
    int64_t N = 0;
    int64_t nv = 0;
    DBuf key, x, ids;   // vertex table (provisional ids) and the list of live ids
    hv_stats_t stats{};
    // what the mesh was built on: the working set and the caller's boundary, so that any
    // context on the same device can integrate it (hv_bind_mesh)
    std::shared_ptr<DevArr> xw;
    int P = 0;
    std::vector<double> pbase, pnormal;
    std::vector<int16_t> pbc;
    // periodic meshes: generators are the nreal originals and N - nreal ghosts (image table
    // below); every vertex has at least one original generator.  nreal == N otherwise.
    int64_t nreal = 0;
    bool extended = false;
    double delta = 0.0;
    double rmax = 0.0;
    ImgOrig img_orig;
    ImgShift img_shift;
    // a mesh of some cells only (hv_build_mesh_cells): their 0-based ids, sorted; null for all
    std::shared_ptr<const std::vector<int>> cells;
    // meshes of one batched walk share one allocation for their vertex tables (key, x, ids
    // point into it); null when the mesh owns its tables
    std::shared_ptr<DevArr> store;
};

// Meshes of one dimension walked together (hv_build_mesh_batch): inputs, the forest of their
// trees and the resulting meshes.  Generator ids are global over the meshes (mesh k holds ids
// goff[k] .. goff[k] + n[k] - 1) and mesh k's boundary plane p has id ntot + 2 D k + p.
struct Forest {
    int nmesh = 0;
    std::vector<const double*> xs;          // host points, d x n[k]
    std::vector<int> n;
    std::vector<const double*> base, normal;   // host planes, d x 2d each (a cuboid)
    int64_t ntot = 0;
    int mbits = 1;                          // bits of a mesh index
    std::vector<int> goff;
    std::shared_ptr<DevArr> xo;             // every mesh's points by global id
    std::vector<DevGeo> geo;                // host copy of each mesh's tree
    GeoSet F{};                             // device copies (per-mesh trees, generator -> mesh)
    std::vector<hv_mesh*> out;              // one mesh per input, null until built
};

// The context's working set for writing `bytes`: the current buffer when no mesh refers to
// it and it is large enough, a new one otherwise (nullptr when out of memory).
double* hv_points_for_write(hv_ctx* c, size_t bytes);
// Record on a new mesh what it was built on (points, boundary, periodic state).
void hv_mesh_snapshot(const hv_ctx* c, hv_mesh* m);

// Binds a context to the state a mesh was built on for the duration of one call (the
// integrators read points, planes and periodic images from the context) and restores the
// context's own state afterwards.
struct hv_bind_mesh {
    hv_ctx* c;
    int D;
    int64_t N, nreal;
    std::shared_ptr<DevArr> xw;
    int P;
    std::vector<double> pbase, pnormal;
    std::vector<int16_t> pbc;
    bool extended;
    double delta, rmax;
    ImgOrig img_orig;
    ImgShift img_shift;
    hv_bind_mesh(hv_ctx* ctx, const hv_mesh* m);
    ~hv_bind_mesh();
};

// planes of the context as an axis-aligned box: 0 and the mask of periodic directions
// (bit k: planes 2k and 2k+1 are each other's periodic partners), or an HV_ERR_* code
int hv_box_periodic_mask(hv_ctx* c, int D, int* mask);

int hv_fail(hv_ctx* c, int code, const std::string& msg);

// Tile sizes of the candidate scan: at most ppw_max rays per packet (8 .. 64), halved while a
// level would give fewer than min_packets packets.
struct HvTiles {
    int ppw_max = 64;
    int min_packets = 2048;
    int ppw_min = 8;   // packets are not halved below this many rays
};
// The tiles a build uses: the caller's (hv_opts.rays_per_block / packet_floor when set), else
// the autotuner's for (dimension, precision, size class of n), else the defaults.
HvTiles hv_tiles_for(const hv_ctx* c, const hv_opts& o, int D, int precision, int64_t n);
// per-device address of the scratch array of hv_raycast_batch_noalloc (resolved once per device)
int hv_scratch_init(int device);

template <int D>
int hv_build_mesh_impl(hv_ctx* c, const hv_opts* o, hv_mesh** out, const std::vector<int>* cells);
template <int D>
int hv_build_forest_meshes(hv_ctx* c, const hv_opts* o, Forest& fo);
// the domain of the context's planes (fills dom_*); HV_OK, or an HV_* code with *why set
int hv_domain(hv_ctx* c, int D, const char** why);
// expected vertices per generator of a mesh in the unit cube (sizes the walk's tables)
double vertices_per_point(int D);
// A batch walks the meshes of a dimension in runs of at most HV_FOREST_VERTICES expected
// vertices (one walk each: a level's ray tables grow with the vertices of the walk, and a
// bounded walk keeps a context's tables reusable), HV_BATCH_WALKERS runs at once, each on its
// own context and stream.  Measured on 512 meshes of 3..6D (ms per batch, tables kept between
// calls): one walker with runs of 16M / 24M / 32M / 40M vertices 541 / 514 / 499 / 583; runs
// of 32M with 1 / 2 / 3 / 4 walkers 475 / 408 / 393 / 386 (26 GB of tables per walker); 24M
// with 3 / 4 walkers 398 / 402 (20 GB each); 16M with 4 / 6 walkers 409 / 413 (13 GB each).
constexpr double HV_FOREST_VERTICES = 24e6;
constexpr int HV_BATCH_WALKERS = 3;
template <int D>
int hv_mc_points_impl(hv_ctx* c, int64_t nsamples, uint64_t seed, double slab, const int64_t* ptr,
                      const uint64_t* pkey, unsigned long long* ccnt, unsigned long long* pcnt, int64_t* nused,
                      double* eps_out, double* vbox_out, int64_t* pairs_out,
                      double* kernel_ms);
template <int D>
// full_sig / full_len (hv_raycast_batch_keys; NULL otherwise): max_sig ids per ray, every generator
// and plane on the new vertex's sphere for a ray out of general position
int hv_raycast_batch_impl(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                          const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status,
                          int32_t max_sig = 0, int64_t* full_sig = nullptr, int32_t* full_len = nullptr);
template <int D>
int hv_raycast_noalloc_impl(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                            const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status);
template <int D>
int hv_nn_halfspace_impl(hv_ctx* c, int64_t nq, const double* q, const double* u, const double* cc,
                         int64_t* idx, double* dist);
template <int D>
int hv_inrange_impl(hv_ctx* c, int64_t nq, const double* q, const double* radius, int64_t* ptr, int64_t* idx);
template <int D>
int hv_mesh_export_impl(const hv_mesh* m, int64_t* sig, double* coords);
// the refined mesh from an old mesh and the mesh S of the new cells on the union (hv_refine_mesh)
template <int D>
int hv_refine_merge(hv_ctx* c, const hv_mesh* old, const hv_mesh* S, hv_mesh** out, int64_t* removed);

// cell volumes and interface areas (hv_integrate.hip)
struct hv_integral;
template <int D>
int hv_integrate_polytope_impl(hv_ctx* c, const hv_mesh* mesh, hv_integral** out);
template <int D>
int hv_integrate_polytope_f_impl(hv_ctx* c, const hv_mesh* mesh, int nf, hv_integrand_fn fn, void* user,
                                 hv_integral** out);
int hv_integral_export_f_impl(const hv_integral* I, int32_t* nf, double* bulk, double* iface);
template <int D>
int hv_integrate_mc_impl(hv_ctx* c, const hv_mesh* mesh, int64_t ndir, uint64_t seed, hv_integral** out);
template <int D>
int hv_integrate_mc_points_impl(hv_ctx* c, const hv_mesh* mesh, int64_t nsamples, uint64_t seed, double slab,
                                hv_integral** out);
int hv_integral_counts_impl(const hv_integral* I, int64_t* ncells, int64_t* nnbr);
int hv_integral_work_impl(const hv_integral* I, int64_t* work, double* ms_kernel);
int hv_integral_export_impl(const hv_integral* I, double* vol, int64_t* nbr_ptr, int64_t* nbr, double* area,
                            double* ms);
void hv_integral_free_impl(hv_integral* I);
