/*
 * hv_oracle.h -- CPU oracle for the HighVoronoi raycast walk (TEST INFRASTRUCTURE).
 *
 * This is a plain-C FP64 restatement of the reference's default mesh algorithm
 * (HighVoronoi.jl v1.4.2, method RCNonGeneralHP, kd-tree search, single thread).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker / CPU baseline; the product path never calls it.
 *
 * Index conventions at this interface match the reference's Julia keys:
 * generators are 1..n, boundary plane p (1-based) is n+p, keys are sorted.
 */
#ifndef HV_ORACLE_H
#define HV_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;
typedef struct orc_mesh orc_mesh;

/* xs: d x n column-major (point i at xs[i*d]); base/normal: d x np likewise.
 * Mirrors src/raycast.jl:26 Raycast(xs; domain) + src/raycast-types.jl:412. */
int orc_ctx_create(int d, int64_t n, const double* xs, int np, const double* base,
                   const double* normal, orc_ctx** out);
void orc_ctx_free(orc_ctx* c);

/* Walk cells [cell_begin, cell_end) (0-based) in order, as src/sysvoronoi.jl:152-215
 * __voronoi does for Iter = cell_begin+1 : cell_end.  seed drives the descent RNG. */
int orc_voronoi(orc_ctx* c, int64_t cell_begin, int64_t cell_end, uint64_t seed, orc_mesh** out);

int64_t orc_mesh_nvert(const orc_mesh* m);
int orc_mesh_maxsig(const orc_mesh* m);
/* keys: maxsig x nvert (1-based, zero padded), coords: d x nvert. */
int orc_mesh_export(const orc_mesh* m, int64_t* keys, double* coords);
/* out[0..9]: walkrays, nn calls, inrange calls, descents, corrections (tolerance breach),
 * irreparable, duplicate pushes, fraud rejections, unbounded rays, hp full_mode casts */
int orc_mesh_stats(const orc_mesh* m, int64_t* out);
int orc_ctx_stats(const orc_ctx* c, int64_t* out);
void orc_mesh_free(orc_mesh* m);

/* One walkray from vertex (sig, r) along the edge that drops sig[drop] (0-based position),
 * explored from cell `cell` (1-based), as src/sysvoronoi.jl:490-525 + src/raycast.jl:125-164.
 * sig: 1-based sorted key of length d+1.  On success writes the new sorted 1-based key
 * (out_len entries, capacity 64) and its coordinates; returns 0, 1 for an unbounded ray,
 * negative on failure. */
int orc_walkray(orc_ctx* c, const int64_t* sig, int drop, const double* r, int64_t cell,
                int64_t* out_sig, int* out_len, double* out_r, double* out_u);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
