/*
 * hv_oracle.c -- CPU oracle: FP64 restatement of HighVoronoi.jl's default raycast walk.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/ (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the product.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * HighVoronoi.jl v1.4.2 source tree).  Third-party arithmetic on the path is
 * restated from its published algorithm:
 *   - StaticArrays qr (Householder) in u_qr          -> modified Gram-Schmidt with
 *     re-orthogonalisation; same unit vector up to rounding (u is unique).
 *   - IterativeSolvers cg! in _correct_vertex         -> Gaussian elimination with
 *     partial pivoting on the same d x d system (same solution up to rounding).
 *   - DoubleFloats Double64 in vertex_calculation_hp  -> long double accumulation.
 *   - Julia randn in descent/randray                  -> splitmix64 + Box-Muller
 *     (the descent only chooses WHICH vertex seeds a cell; the vertex set does not
 *     depend on it).
 * Parity of these four is pinned by the mathematics they compute, not bit-for-bit.
 */
#include "hv_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define MAXD 8
#define MAXSIG 64
#define ORC_NSTATS 12   /* the counters of oracle_py.STAT_NAMES */

static __thread char g_err[512];
static void set_err(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char* orc_last_error(void) { return g_err; }

/* Tolerances of RaycastParameter{Float64}: src/raycast-types.jl:226-230 */
static const double VARIANCE_TOL = 1.0e-15;
static const double BREAK_TOL = 1.0e-5;
static const double B_NODES_TOL = 1.0e-7;
static const double PLANE_TOL = 1.0e-12;

/* ------------------------------------------------------------------ */
/* small vector helpers                                                */
static inline double dotd(const double* a, const double* b, int d) {
    double s = 0; for (int i = 0; i < d; ++i) s += a[i] * b[i]; return s;
}
static inline double dist2d(const double* a, const double* b, int d) {
    double s = 0; for (int i = 0; i < d; ++i) { double t = a[i] - b[i]; s += t * t; } return s;
}
static inline double normd(const double* a, int d) { return sqrt(dotd(a, a, d)); }

/* ------------------------------------------------------------------ */
/* kd-tree: src/NearestNeighborModified/kd_tree.jl:105-158 (leafsize 10, split on the
 * widest side of the node's hyper-rectangle at the median, data reordered).       */
typedef struct {
    int dim;            /* -1 for a leaf */
    double split, lo, hi;
    int64_t left, right, begin, end;
} kdnode;

typedef struct {
    int64_t first, len;     /* key in pool */
    int64_t coord;          /* index into coords */
} vslot;

struct orc_mesh {
    int d;
    int64_t n;
    int64_t nv, cap;
    int64_t* koff; int* klen;
    int64_t* kpool; int64_t kpool_len, kpool_cap;
    double* coords;
    /* hash: open addressing on vertex ids */
    int64_t* htab; int64_t hmask;
    /* per-cell vertex lists */
    int64_t** cell_v; int64_t* cell_n; int64_t* cell_cap;
    int64_t stats[ORC_NSTATS];
};

struct orc_ctx {
    int d; int64_t n; int np;
    double* x;        /* (n+np) x d : generators then mirror slots (ExtendedNodes, src/extended.jl:9-63) */
    double* pb; double* pn;
    char* active;
    int64_t* perm; double* xr;
    kdnode* nodes; int64_t nnodes, nodes_cap;
    double rootlo[MAXD], roothi[MAXD];
    /* scratch */
    int64_t* ids; double* ts; int64_t scratch_cap, nids;
    /* HP vectors: searcher.hp_vars (src/raycast-types.jl:398-410) */
    double V[MAXD][MAXD]; double rhs_cg[MAXD];
    /* edge hash for one cell */
    int64_t* epool; int64_t epool_len, epool_cap;
    int64_t* ekey; int* elen; int64_t* ecell1; int64_t* ecell2; uint32_t* egen;
    int64_t emask, ecount; uint32_t gen;
    uint64_t rng;
    orc_mesh* mesh;   /* mesh being built (stats) */
    int64_t stats[ORC_NSTATS];
};

static void* xmalloc(size_t n) { void* p = malloc(n ? n : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }

static void kd_swap(int64_t* a, int64_t i, int64_t j) { int64_t t = a[i]; a[i] = a[j]; a[j] = t; }

/* quickselect (src/NearestNeighborModified/utilities.jl:29-57 select_spec!) */
static void kd_select(orc_ctx* c, int64_t lo, int64_t hi, int64_t k, int dim) {
    int d = c->d;
    while (hi > lo) {
        int64_t mid = lo + (hi - lo) / 2;
        double pv = c->x[c->perm[mid] * d + dim];
        kd_swap(c->perm, mid, hi);
        int64_t s = lo;
        for (int64_t i = lo; i < hi; ++i)
            if (c->x[c->perm[i] * d + dim] < pv) kd_swap(c->perm, i, s++);
        kd_swap(c->perm, s, hi);
        if (s == k) return;
        if (s < k) lo = s + 1; else hi = s - 1;
    }
}

static int64_t kd_build(orc_ctx* c, int64_t begin, int64_t end, double* blo, double* bhi) {
    if (c->nnodes == c->nodes_cap) {
        c->nodes_cap = c->nodes_cap ? 2 * c->nodes_cap : 1024;
        c->nodes = (kdnode*)realloc(c->nodes, sizeof(kdnode) * c->nodes_cap);
    }
    int64_t id = c->nnodes++;
    kdnode nd; memset(&nd, 0, sizeof nd);
    nd.begin = begin; nd.end = end; nd.dim = -1; nd.left = nd.right = -1;
    int64_t np_ = end - begin;
    if (np_ <= 10) { c->nodes[id] = nd; return id; }
    int sd = 0; double ms = -1.0;
    for (int k = 0; k < c->d; ++k) { double s = bhi[k] - blo[k]; if (s > ms) { ms = s; sd = k; } }
    int64_t mid = begin + np_ / 2;
    kd_select(c, begin, end - 1, mid, sd);
    double sv = c->x[c->perm[mid] * c->d + sd];
    nd.dim = sd; nd.split = sv; nd.lo = blo[sd]; nd.hi = bhi[sd];
    double save = bhi[sd]; bhi[sd] = sv;
    int64_t l = kd_build(c, begin, mid, blo, bhi);
    bhi[sd] = save;
    save = blo[sd]; blo[sd] = sv;
    int64_t r = kd_build(c, mid, end, blo, bhi);
    blo[sd] = save;
    nd.left = l; nd.right = r;
    c->nodes[id] = nd;
    return id;
}

/* half-space skip predicate: myskips, src/raycast.jl:385-392 */
typedef struct { const double* u; double c; int use; } skipper;
static inline int do_skip(const orc_ctx* c, const skipper* s, int64_t i) {
    if (!s->use) return 0;
    return dotd(c->x + i * c->d, s->u, c->d) <= s->c;
}

/* knn_kernel!, src/NearestNeighborModified/kd_tree.jl:497-538 + add_points_knn!, tree_ops.jl:134-154 */
static void kd_nn_rec(const orc_ctx* c, int64_t ni, const double* q, double min_dist, double* best, int64_t* bi, const skipper* sk) {
    const kdnode* nd = &c->nodes[ni];
    int d = c->d;
    if (nd->dim < 0) {
        for (int64_t z = nd->begin; z < nd->end; ++z) {
            double dd = dist2d(c->xr + z * d, q, d);
            if (dd <= *best) {
                int64_t idx = c->perm[z];
                if (do_skip(c, sk, idx)) continue;
                *best = dd; *bi = idx;
            }
        }
        return;
    }
    double pd = q[nd->dim];
    double sdiff = pd - nd->split;
    int64_t close_, far_; double ddiff;
    if (sdiff > 0) { close_ = nd->right; far_ = nd->left; ddiff = fmax(0.0, pd - nd->hi); }
    else { close_ = nd->left; far_ = nd->right; ddiff = fmax(0.0, nd->lo - pd); }
    kd_nn_rec(c, close_, q, min_dist, best, bi, sk);
    double new_min = min_dist - ddiff * ddiff + sdiff * sdiff;
    if (new_min < *best) kd_nn_rec(c, far_, q, new_min, best, bi, sk);
}

static double box_min_dist(const orc_ctx* c, const double* q) {
    double s = 0;
    for (int k = 0; k < c->d; ++k) {
        double t = 0;
        if (q[k] < c->rootlo[k]) t = c->rootlo[k] - q[k];
        else if (q[k] > c->roothi[k]) t = q[k] - c->roothi[k];
        s += t * t;
    }
    return s;
}

/* nn with mirrors: src/extended.jl:108-142 (strict < for mirrors); searchtrees.jl:238-242 */
static int64_t ext_nn(orc_ctx* c, const double* q, const skipper* sk, double* dist_out) {
    c->stats[1]++;
    double best = DBL_MAX; int64_t bi = -1;
    kd_nn_rec(c, 0, q, box_min_dist(c, q), &best, &bi, sk);
    double dist = (bi >= 0) ? sqrt(best) : INFINITY;
    for (int p = 0; p < c->np; ++p) {
        int64_t id = c->n + p;
        if (!c->active[p] || do_skip(c, sk, id)) continue;
        double dm = sqrt(dist2d(q, c->x + id * c->d, c->d));
        if (dm < dist) { bi = id; dist = dm; }
    }
    if (dist_out) *dist_out = dist;
    return bi;
}

static void push_id(orc_ctx* c, int64_t id) {
    if (c->nids == c->scratch_cap) {
        c->scratch_cap *= 2;
        c->ids = (int64_t*)realloc(c->ids, sizeof(int64_t) * c->scratch_cap);
        c->ts = (double*)realloc(c->ts, sizeof(double) * c->scratch_cap);
    }
    c->ids[c->nids++] = id;
}

/* inrange_kernel!, kd_tree.jl:693-743; add_points_inrange!, tree_ops.jl:183-195 (inclusive) */
static void kd_inrange_rec(orc_ctx* c, int64_t ni, const double* q, double r2, double min_dist) {
    if (min_dist > r2) return;
    const kdnode* nd = &c->nodes[ni];
    int d = c->d;
    if (nd->dim < 0) {
        for (int64_t z = nd->begin; z < nd->end; ++z)
            if (dist2d(c->xr + z * d, q, d) <= r2) push_id(c, c->perm[z]);
        return;
    }
    double pd = q[nd->dim];
    double sdiff = pd - nd->split;
    int64_t close_, far_; double ddiff;
    if (sdiff > 0) { close_ = nd->right; far_ = nd->left; ddiff = fmax(0.0, pd - nd->hi); }
    else { close_ = nd->left; far_ = nd->right; ddiff = fmax(0.0, nd->lo - pd); }
    kd_inrange_rec(c, close_, q, r2, min_dist);
    double new_min = min_dist - ddiff * ddiff + sdiff * sdiff;
    kd_inrange_rec(c, far_, q, r2, new_min);
}

/* _inrange with mirrors: src/extended.jl:265-277 (strict < for mirrors) */
static void ext_inrange(orc_ctx* c, const double* q, double rad) {
    c->stats[2]++;
    c->nids = 0;
    kd_inrange_rec(c, 0, q, rad * rad, box_min_dist(c, q));
    for (int p = 0; p < c->np; ++p) {
        int64_t id = c->n + p;
        if (!c->active[p]) continue;
        if (sqrt(dist2d(q, c->x + id * c->d, c->d)) < rad) push_id(c, id);
    }
}

/* ------------------------------------------------------------------ */
/* boundary: reflect src/boundary.jl:206-211; activate_cell/activate_mirror src/raycast.jl:354-375 */
static void activate_cell(orc_ctx* c, int64_t cell) {
    int d = c->d;
    const double* xc = c->x + cell * d;
    for (int p = 0; p < c->np; ++p) {
        const double* nrm = c->pn + p * d; const double* b = c->pb + p * d;
        double s = 0; for (int k = 0; k < d; ++k) s += nrm[k] * (b[k] - xc[k]);
        double* m = c->x + (c->n + p) * d;
        for (int k = 0; k < d; ++k) m[k] = xc[k] + nrm[k] * (2 * s);
        c->active[p] = 1;
    }
}

/* project, src/boundary.jl:461-473 */
static void project_pt(const orc_ctx* c, double* x) {
    int d = c->d;
    for (int p = 0; p < c->np; ++p) {
        const double* nrm = c->pn + p * d; const double* b = c->pb + p * d;
        double s = 0; for (int k = 0; k < d; ++k) s += (b[k] - x[k]) * nrm[k];
        if (s < 0) for (int k = 0; k < d; ++k) x[k] += s * nrm[k];
    }
}

/* ------------------------------------------------------------------ */
/* t formulas: src/raycast.jl:394-432 */
static double get_t_hp(const orc_ctx* c, const double* r, const double* u, const double* x0, const double* x) {
    int d = c->d; double num = 0, den = 0;
    for (int k = 0; k < d; ++k) { double dx = x[k] - x0[k]; num += dx * (x0[k] + x[k] - 2 * r[k]); den += u[k] * dx; }
    return num / (2 * den);
}
static double get_t_hp_(const orc_ctx* c, const double* r, const double* u, const double* x0, const double* x, double du, double* err) {
    int d = c->d; double Dx[MAXD]; double nn = 0;
    for (int k = 0; k < d; ++k) { Dx[k] = x[k] - x0[k]; nn += Dx[k] * Dx[k]; }
    nn = sqrt(nn);
    for (int k = 0; k < d; ++k) Dx[k] /= nn;
    double den = dotd(u, Dx, d);
    double v = 0; for (int k = 0; k < d; ++k) v += Dx[k] * (x0[k] + x[k] - 2 * r[k]);
    v = v / (2 * den);
    *err = (v * du + normd(r, d) * 1e-15) / den;
    return v;
}
/* get_scale, src/raycast.jl:530-540 */
static double get_scale(const orc_ctx* c, const double* u, const double* x0, const double* rr) {
    int d = c->d; double delta[MAXD];
    for (int k = 0; k < d; ++k) delta[k] = rr[k] - x0[k];
    double ref = dotd(delta, delta, d);
    double vert = dotd(u, delta, d); vert *= vert;
    double hori = vert > ref ? 0.0 : ref - vert;
    return sqrt(hori / ref);
}

/* prepare_vertex_calculation, src/raycast.jl:633-660 (vectors stored as rows here) */
static void prepare_vertex_calculation(orc_ctx* c, const int64_t* sig) {
    int d = c->d;
    memset(c->V, 0, sizeof c->V); memset(c->rhs_cg, 0, sizeof c->rhs_cg);
    const double* x0 = c->x + sig[0] * d;
    for (int i = 1; i < d; ++i) {
        const double* xi = c->x + sig[i] * d;
        double v[MAXD]; for (int k = 0; k < d; ++k) v[k] = xi[k] - x0[k];
        double nv = normd(v, d);
        double s = 0;
        for (int k = 0; k < d; ++k) { c->V[i - 1][k] = v[k] / nv; s += (xi[k] + x0[k]) * c->V[i - 1][k]; }
        c->rhs_cg[i - 1] = 0.5 * s;
    }
    for (int i = 1; i < d - 1; ++i) {
        double vi[MAXD]; memcpy(vi, c->V[i], sizeof(double) * d);
        for (int j = 0; j < i; ++j) {
            for (int rep = 0; rep < 2; ++rep) {
                double a = dotd(vi, c->V[j], d);
                for (int k = 0; k < d; ++k) vi[k] -= a * c->V[j][k];
                c->rhs_cg[i] -= a * c->rhs_cg[j];
            }
        }
        double vin = normd(vi, d);
        for (int k = 0; k < d; ++k) c->V[i][k] = vi[k] / vin;
        c->rhs_cg[i] /= vin;
    }
}

/* vertex_calculation_hp, src/raycast.jl:662-707 */
static void vertex_calculation_hp(orc_ctx* c, const int64_t* edge, const double* r, int64_t gen, double* out) {
    int d = c->d;
    const double* x0 = c->x + edge[0] * d;
    const double* xn = c->x + gen * d;
    double v[MAXD]; for (int k = 0; k < d; ++k) v[k] = xn[k] - x0[k];
    double nv = normd(v, d);
    double s = 0;
    for (int k = 0; k < d; ++k) { v[k] /= nv; s += (xn[k] + x0[k]) * v[k]; }
    c->rhs_cg[d - 1] = 0.5 * s;
    for (int j = 0; j < d - 1; ++j) {
        for (int rep = 0; rep < 2; ++rep) {
            double a = dotd(v, c->V[j], d);
            for (int k = 0; k < d; ++k) v[k] -= a * c->V[j][k];
            c->rhs_cg[d - 1] -= a * c->rhs_cg[j];
        }
    }
    double vin = normd(v, d);
    for (int k = 0; k < d; ++k) c->V[d - 1][k] = v[k] / vin;
    c->rhs_cg[d - 1] /= vin;
    long double myrhs[MAXD];
    for (int i = 0; i < d; ++i) myrhs[i] = (long double)c->rhs_cg[i] - (long double)dotd(c->V[i], r, d);
    for (int k = 0; k < d; ++k) {
        long double s2 = 0; for (int i = 0; i < d; ++i) s2 += (long double)c->V[i][k] * myrhs[i];
        out[k] = r[k] + (double)s2;
    }
}

/* vertex_variance, src/raycast.jl:320-330 */
static double vertex_variance(const orc_ctx* c, const int64_t* s, int cnt, const double* r) {
    int d = c->d; double dist[MAXSIG]; double m = 0;
    for (int k = 0; k < cnt; ++k) { dist[k] = dist2d(c->x + s[k] * d, r, d); m += dist[k]; }
    m = -m / cnt;
    double acc = 0; for (int k = 0; k < cnt; ++k) { double t = dist[k] + m; acc += t * t; }
    return acc / (m * m);
}

/* solve A x = b (d x d, row-major), partial pivoting; returns 0 on success */
static int gauss_solve(int d, double A[MAXD][MAXD], double* b, double* x) {
    int piv[MAXD];
    for (int i = 0; i < d; ++i) piv[i] = i;
    for (int col = 0; col < d; ++col) {
        int best = col; double bv = fabs(A[col][col]);
        for (int r = col + 1; r < d; ++r) if (fabs(A[r][col]) > bv) { bv = fabs(A[r][col]); best = r; }
        if (bv == 0.0) return -1;
        if (best != col) {
            for (int k = 0; k < d; ++k) { double t = A[col][k]; A[col][k] = A[best][k]; A[best][k] = t; }
            double t = b[col]; b[col] = b[best]; b[best] = t;
        }
        for (int r = col + 1; r < d; ++r) {
            double f = A[r][col] / A[col][col];
            for (int k = col; k < d; ++k) A[r][k] -= f * A[col][k];
            b[r] -= f * b[col];
        }
    }
    for (int i = d - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < d; ++k) s -= A[i][k] * x[k];
        x[i] = s / A[i][i];
    }
    (void)piv;
    return 0;
}

/* _correct_vertex, src/raycast.jl:287-318: the normal equations (V V^T) x = V b whose
 * solution is V^T x = b, rows x_{s_i}-x_{s_{d+1}}, b_i = (|x_{s_i}|^2-|x_{s_{d+1}}|^2)/2. */
static void correct_vertex(orc_ctx* c, const int64_t* s, double* r) {
    int d = c->d;
    double A[MAXD][MAXD], b[MAXD], x[MAXD];
    const double* xl = c->x + s[d] * d;
    double nl = dotd(xl, xl, d);
    for (int i = 0; i < d; ++i) {
        const double* xi = c->x + s[i] * d;
        for (int k = 0; k < d; ++k) A[i][k] = xi[k] - xl[k];
        b[i] = 0.5 * (dotd(xi, xi, d) - nl);
    }
    if (gauss_solve(d, A, b, x) == 0) memcpy(r, x, sizeof(double) * d);
}

/* walkray_correct_vertex, src/raycast.jl:242-279; returns success */
static int walkray_correct_vertex(orc_ctx* c, double* r, const int64_t* edge, int64_t gen) {
    int d = c->d;
    int64_t s[MAXD + 1];
    for (int i = 0; i < d; ++i) s[i] = edge[i];
    s[d] = gen;
    double vv = vertex_variance(c, s, d + 1, r);
    int i = 0;
    while (i < 3 && vv > 0.0001 * VARIANCE_TOL) {
        ++i; correct_vertex(c, s, r); vv = vertex_variance(c, s, d + 1, r);
    }
    if (vv > VARIANCE_TOL && vv < BREAK_TOL) {
        correct_vertex(c, s, r); vv = vertex_variance(c, s, d + 1, r);
        c->stats[4]++;
    }
    if (vv > BREAK_TOL) { c->stats[5]++; return 0; }
    return 1;
}

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b; return (x > y) - (x < y);
}

static int in_list(int64_t v, const int64_t* a, int n) { for (int i = 0; i < n; ++i) if (a[i] == v) return 1; return 0; }

/* get__r, src/raycast.jl:719-773 */
static int64_t get_r(orc_ctx* c, const double* vvv, const int64_t* edge, const skipper* sk, double t,
                     const double* u, const double* r, const double* x0, int full_mode, double first_t,
                     double* r_out, double* t_out) {
    int d = c->d;
    double _r[MAXD]; memcpy(_r, vvv, sizeof(double) * d);
    double t2buf = first_t;
    double my_dist = sqrt(dist2d(x0, vvv, d));
    int64_t ret_i = -1;
    int modified = 0;
    int ii = 1;
    for (;;) {
        ++ii; if (ii == 4) break;
        int64_t i = ext_nn(c, _r, sk, NULL);
        if (i == edge[0]) break;
        ret_i = i;
        if (i >= 0) {
            modified = 1;
            const double* x = c->x + i * d;
            if (sqrt(dist2d(x, _r, d)) >= my_dist) break;
            double t2 = get_t_hp(c, r, u, x0, x);
            t2buf = t2;
            double rt[MAXD]; for (int k = 0; k < d; ++k) rt[k] = r[k] + t2 * u[k];
            t2 += get_t_hp(c, rt, u, x0, x);
            t = t2;
        } else break;
        for (int k = 0; k < d; ++k) _r[k] = r[k] + t * u[k];
        if (full_mode && modified) { double tmp[MAXD]; vertex_calculation_hp(c, edge, _r, i, tmp); memcpy(_r, tmp, sizeof(double) * d); }
        double nd = sqrt(dist2d(c->x + i * d, _r, d));
        if (nd >= my_dist) break;
        my_dist = nd;
    }
    memcpy(r_out, _r, sizeof(double) * d);
    *t_out = t2buf;
    return ret_i;
}

/* raycast_des2(..., ::HPUnion), src/raycast.jl:794-970.
 * sig (lsig entries) is the edge to be extended; edge (ledge) = the edge (edge[0] its min);
 * origin = the key of the source vertex.  Returns winner (>=0) or -1 for t=Inf. */
static int64_t raycast_hp(orc_ctx* c, const int64_t* sig, int lsig, const double* r, const double* u,
                          const int64_t* edge, int ledge, const int64_t* origin, int lorigin,
                          int walkray_mode, double du, int64_t* sig_out, int* lsig_out,
                          double* t_out, double* r2_out) {
    int d = c->d;
    const double* x0 = c->x + edge[0] * d;
    double c1 = -DBL_MAX;
    for (int k = 0; k < lsig; ++k) { double v = dotd(c->x + sig[k] * d, u, d); if (v > c1) c1 = v; }
    skipper sk; sk.u = u; sk.c = c1 + fabs(c1) * PLANE_TOL; sk.use = 1;
    double vvv[MAXD];
    double a = 0; for (int k = 0; k < d; ++k) a += u[k] * (x0[k] - r[k]);
    for (int k = 0; k < d; ++k) vvv[k] = r[k] + u[k] * a;
    int64_t i = ext_nn(c, vvv, &sk, NULL);
    if (i < 0) { *t_out = INFINITY; return -1; }
    const double* x = c->x + i * d;
    double full_error;
    double t = get_t_hp_(c, r, u, x0, x, du, &full_error);
    double first_t = t;
    double _vvv[MAXD]; for (int k = 0; k < d; ++k) _vvv[k] = r[k] + t * u[k];
    double scale = get_scale(c, u, x0, _vvv);
    double relative_error = full_error / sqrt(dist2d(_vvv, r, d));
    int full_mode = walkray_mode && (relative_error > 1e-10 || full_error > 1e-8 / fmax(scale, 1e-4));
    if (full_mode) c->stats[9]++;
    double dummy;
    t += get_t_hp_(c, _vvv, u, x0, x, du, &dummy);
    double vtu[MAXD]; for (int k = 0; k < d; ++k) vtu[k] = _vvv[k] + t * u[k];
    if (full_mode) {
        prepare_vertex_calculation(c, sig);
        vertex_calculation_hp(c, edge, vtu, i, vvv);
    } else memcpy(vvv, vtu, sizeof(double) * d);
    double _r[MAXD];
    int64_t new_i = get_r(c, vvv, edge, &sk, t, u, r, x0, full_mode, first_t, _r, &t);
    if (new_i < 0) { *t_out = INFINITY; return -1; }
    double measure = 0;
    for (int k = 0; k < ledge; ++k) { double m = sqrt(dist2d(c->x + edge[k] * d, _r, d)); if (m > measure) measure = m; }
    double upper_t = t * 1.0000000001;
    ext_inrange(c, _r, (1 + fmax(relative_error, B_NODES_TOL * 10 * scale)) * measure);
    int64_t lidss = c->nids;
    if (lidss == 0) { *t_out = INFINITY; return -1; }
    int64_t* idss = c->ids; double* ts = c->ts;
    for (int64_t k = 0; k < lidss; ++k)
        ts[k] = in_list(idss[k], origin, lorigin) ? 0.0 : get_t_hp(c, r, u, x0, c->x + idss[k] * d);
    for (int64_t k = 0; k < lidss; ++k) {
        if (in_list(idss[k], origin, lorigin)) { idss[k] = INT64_MAX; continue; }
        if (ts[k] < PLANE_TOL || ts[k] > upper_t) ts[k] = 0.0;
        else if (ts[k] < upper_t) upper_t = ts[k];
    }
    upper_t += fmin(10e-8, (10 + d) * full_error);
    double max_dist = 0.0; int64_t generator = -1;
    int in_window = 0;
    for (int64_t k = 0; k < lidss; ++k) {
        if (ts[k] > upper_t) ts[k] = 0.0;
        else if (ts[k] > 0.0) {
            ++in_window;
            double s = 0; const double* xk = c->x + idss[k] * d;
            for (int q = 0; q < d; ++q) s += u[q] * (xk[q] - x0[q]);
            ts[k] = s;
            if (s > max_dist) { max_dist = s; generator = idss[k]; }
        }
    }
    /* a near tie: more than one candidate within the window, the winner by <u, x - x0> */
    if (walkray_mode && in_window > 1) c->stats[10]++;
    if (generator < 0) { *t_out = INFINITY; return -1; }
    t = get_t_hp(c, r, u, x0, c->x + generator * d);
    double r2[MAXD]; for (int k = 0; k < d; ++k) r2[k] = r[k] + t * u[k];
    if (walkray_mode && full_mode) { double tmp[MAXD]; vertex_calculation_hp(c, edge, r2, generator, tmp); memcpy(r2, tmp, sizeof(double) * d); }
    double minmeas, maxmeas;
    minmeas = maxmeas = sqrt(dist2d(c->x + generator * d, r2, d));
    for (int k = 0; k < lsig; ++k) {
        double n2 = sqrt(dist2d(c->x + sig[k] * d, r2, d));
        if (n2 < minmeas) minmeas = n2;
        if (n2 > maxmeas) maxmeas = n2;
    }
    double measure2 = maxmeas + scale * (maxmeas - minmeas);
    int lo = 0;
    for (int k = 0; k < lsig; ++k) sig_out[lo++] = sig[k];
    for (int64_t k = 0; k < lidss; ++k) {
        if (idss[k] == INT64_MAX) continue;
        if (sqrt(dist2d(c->x + idss[k] * d, r2, d)) > measure2) continue;
        if (lo >= MAXSIG) { set_err("degenerate vertex exceeds key capacity"); return -2; }
        sig_out[lo++] = idss[k];
    }
    qsort(sig_out, lo, sizeof(int64_t), cmp_i64);
    *lsig_out = lo;
    *t_out = t;
    memcpy(r2_out, r2, sizeof(double) * d);
    return generator;
}

/* u_qr, src/tools.jl:773-790 and delta_u, src/edgeiteratebase.jl:64-74.
 * sig has lsig = d+1 sorted entries, i = position of the dropped generator. */
static void u_qr(orc_ctx* c, const int64_t* sig, int lsig, int i, double* u, double* du) {
    int d = c->d;
    double X[MAXD][MAXD];  /* columns as rows */
    int col = 0;
    for (int j = 0; j < lsig; ++j) { if (j == i) continue; memcpy(X[col++], c->x + sig[j] * d, sizeof(double) * d); }
    double origin[MAXD]; memcpy(origin, X[d - 1], sizeof(double) * d);
    memcpy(X[d - 1], c->x + sig[i] * d, sizeof(double) * d);
    for (int j = 0; j < d; ++j) for (int k = 0; k < d; ++k) X[j][k] -= origin[k];
    /* orthonormal basis of the first d-1 columns, then project column d */
    double Q[MAXD][MAXD];
    for (int j = 0; j < d; ++j) {
        double v[MAXD]; memcpy(v, X[j], sizeof(double) * d);
        for (int rep = 0; rep < 2; ++rep)
            for (int l = 0; l < j; ++l) { double a = dotd(v, Q[l], d); for (int k = 0; k < d; ++k) v[k] -= a * Q[l][k]; }
        double nv = normd(v, d);
        for (int k = 0; k < d; ++k) Q[j][k] = v[k] / nv;
    }
    /* R[end,end] = <Q[:,end], X[:,end]> > 0 for this construction, so u = -Q[:,end] */
    for (int k = 0; k < d; ++k) u[k] = -Q[d - 1][k];
    double s = 0;
    for (int j = 0; j < d - 1; ++j) { double nx = normd(X[j], d); s += fabs(dotd(X[j], u, d) / nx); }
    *du = s;
}

/* randray, src/raycast.jl:211-239: random unit vector orthogonal to span{x_i - x_k} */
static double rng_u01(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
static double rng_normal(uint64_t* s) {
    double u1 = rng_u01(s), u2 = rng_u01(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static void randray(orc_ctx* c, const int64_t* pts, int k, double* u) {
    int d = c->d; double v[MAXD][MAXD];
    const double* xk = c->x + pts[k - 1] * d;
    for (int i = 0; i < k - 1; ++i) {
        for (int q = 0; q < d; ++q) v[i][q] = c->x[pts[i] * d + q] - xk[q];
        for (int rep = 0; rep < 2; ++rep) {
            for (int j = 0; j < i; ++j) { double a = dotd(v[i], v[j], d); for (int q = 0; q < d; ++q) v[i][q] -= a * v[j][q]; }
            double nv = normd(v[i], d); for (int q = 0; q < d; ++q) v[i][q] /= nv;
        }
    }
    for (int q = 0; q < d; ++q) u[q] = rng_normal(&c->rng);
    for (int rep = 0; rep < 2; ++rep) {
        for (int i = 0; i < k - 1; ++i) { double a = dotd(u, v[i], d); for (int q = 0; q < d; ++q) u[q] -= a * v[i][q]; }
        double nu = normd(u, d); for (int q = 0; q < d; ++q) u[q] /= nu;
    }
}

/* descent, src/raycast.jl:45-109 */
static int descent(orc_ctx* c, int64_t start, int64_t* sig_out, double* r_out) {
    int d = c->d;
    c->stats[3]++;
    for (int count = 1; count < 10; ++count) {
        int64_t sig[MAXSIG]; int lsig = 1; sig[0] = start;
        double r[MAXD]; memcpy(r, c->x + start * d, sizeof(double) * d);
        int64_t minimal_edge[MAXD + 1]; memset(minimal_edge, 0, sizeof minimal_edge); minimal_edge[0] = start;
        int ok = 1; int k = 1; int loop_counter = 0;
        while (k <= d) {
            ++loop_counter;
            double u[MAXD]; randray(c, minimal_edge, k, u);
            int64_t nsig[MAXSIG]; int nl; double t, r2[MAXD];
            int64_t g = raycast_hp(c, sig, lsig, r, u, sig, lsig, sig, lsig, 0, 1e-14, nsig, &nl, &t, r2);
            if (g == -2) return -2;
            if (!isfinite(t)) {
                for (int q = 0; q < d; ++q) u[q] = -u[q];
                g = raycast_hp(c, sig, lsig, r, u, sig, lsig, sig, lsig, 0, 1e-14, nsig, &nl, &t, r2);
                if (g == -2) return -2;
            }
            if (!isfinite(t)) {
                if (loop_counter <= 100) continue;
                set_err("descent: no vertex in both directions"); return -3;
            }
            memcpy(r, r2, sizeof(double) * d);
            memcpy(sig, nsig, sizeof(int64_t) * nl); lsig = nl;
            minimal_edge[k] = g;
            double vv = vertex_variance(c, minimal_edge, k + 1, r);
            if (vv > VARIANCE_TOL) { set_err("descent: variance"); return -4; }
            ++k; loop_counter = 0;
        }
        if (!ok) continue;
        if (vertex_variance(c, minimal_edge, d + 1, r) > VARIANCE_TOL) continue;
        qsort(sig, lsig, sizeof(int64_t), cmp_i64);
        project_pt(c, r);
        walkray_correct_vertex(c, r, minimal_edge, minimal_edge[d]);
        memcpy(sig_out, sig, sizeof(int64_t) * lsig);
        memcpy(r_out, r, sizeof(double) * d);
        return lsig;
    }
    set_err("descent failed");
    return -5;
}

/* ------------------------------------------------------------------ */
/* mesh DB (src/abstractmesh.jl:111-146 push!/haskey semantics; keyed by sorted sig) */
static uint64_t hash_key(const int64_t* k, int n) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < n; ++i) { h ^= (uint64_t)k[i]; h *= 1099511628211ull; }
    h ^= h >> 29; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 32;
    return h;
}

static orc_mesh* mesh_new(int d, int64_t n) {
    orc_mesh* m = (orc_mesh*)xcalloc(1, sizeof(orc_mesh));
    m->d = d; m->n = n; m->cap = 1024;
    m->koff = (int64_t*)xmalloc(sizeof(int64_t) * m->cap);
    m->klen = (int*)xmalloc(sizeof(int) * m->cap);
    m->coords = (double*)xmalloc(sizeof(double) * d * m->cap);
    m->kpool_cap = 1024 * (d + 1);
    m->kpool = (int64_t*)xmalloc(sizeof(int64_t) * m->kpool_cap);
    m->hmask = 4095;
    m->htab = (int64_t*)xmalloc(sizeof(int64_t) * (m->hmask + 1));
    for (int64_t i = 0; i <= m->hmask; ++i) m->htab[i] = -1;
    m->cell_v = (int64_t**)xcalloc(n, sizeof(int64_t*));
    m->cell_n = (int64_t*)xcalloc(n, sizeof(int64_t));
    m->cell_cap = (int64_t*)xcalloc(n, sizeof(int64_t));
    return m;
}

static int64_t mesh_find(const orc_mesh* m, const int64_t* key, int len) {
    uint64_t h = hash_key(key, len) & m->hmask;
    for (;;) {
        int64_t v = m->htab[h];
        if (v < 0) return -1;
        if (m->klen[v] == len && memcmp(m->kpool + m->koff[v], key, sizeof(int64_t) * len) == 0) return v;
        h = (h + 1) & m->hmask;
    }
}

static void mesh_rehash(orc_mesh* m) {
    m->hmask = 2 * (m->hmask + 1) - 1;
    free(m->htab);
    m->htab = (int64_t*)xmalloc(sizeof(int64_t) * (m->hmask + 1));
    for (int64_t i = 0; i <= m->hmask; ++i) m->htab[i] = -1;
    for (int64_t v = 0; v < m->nv; ++v) {
        uint64_t h = hash_key(m->kpool + m->koff[v], m->klen[v]) & m->hmask;
        while (m->htab[h] >= 0) h = (h + 1) & m->hmask;
        m->htab[h] = v;
    }
}

/* returns vertex id; 0 if inserted, 1 if it already existed */
static int mesh_push(orc_mesh* m, const int64_t* key, int len, const double* r) {
    if (mesh_find(m, key, len) >= 0) { m->stats[6]++; return 1; }
    int d = m->d;
    if (m->nv == m->cap) {
        m->cap *= 2;
        m->koff = (int64_t*)realloc(m->koff, sizeof(int64_t) * m->cap);
        m->klen = (int*)realloc(m->klen, sizeof(int) * m->cap);
        m->coords = (double*)realloc(m->coords, sizeof(double) * d * m->cap);
    }
    if (m->kpool_len + len > m->kpool_cap) {
        while (m->kpool_len + len > m->kpool_cap) m->kpool_cap *= 2;
        m->kpool = (int64_t*)realloc(m->kpool, sizeof(int64_t) * m->kpool_cap);
    }
    int64_t v = m->nv++;
    m->koff[v] = m->kpool_len; m->klen[v] = len;
    memcpy(m->kpool + m->kpool_len, key, sizeof(int64_t) * len); m->kpool_len += len;
    memcpy(m->coords + v * d, r, sizeof(double) * d);
    if (2 * m->nv > m->hmask) mesh_rehash(m);
    else {
        uint64_t h = hash_key(key, len) & m->hmask;
        while (m->htab[h] >= 0) h = (h + 1) & m->hmask;
        m->htab[h] = v;
    }
    for (int i = 0; i < len; ++i) {
        int64_t g = key[i];
        if (g >= m->n) continue;
        if (m->cell_n[g] == m->cell_cap[g]) {
            m->cell_cap[g] = m->cell_cap[g] ? 2 * m->cell_cap[g] : 16;
            m->cell_v[g] = (int64_t*)realloc(m->cell_v[g], sizeof(int64_t) * m->cell_cap[g]);
        }
        m->cell_v[g][m->cell_n[g]++] = v;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* per-cell edge hash: pushedge!, src/edgehashing.jl:66-115 (semantics of cell1/cell2).
 * Entries of older cells are recognised by a generation stamp, so clearing is O(1). */
static void edges_clear(orc_ctx* c) {
    c->gen++; c->epool_len = 0; c->ecount = 0;
    if (c->gen == 0) { memset(c->egen, 0, sizeof(uint32_t) * (c->emask + 1)); c->gen = 1; }
}

static void edges_rehash(orc_ctx* c) {
    int64_t om = c->emask, nm = 2 * (om + 1) - 1;
    int64_t* ok = c->ekey; int* ol = c->elen; int64_t* o1 = c->ecell1; int64_t* o2 = c->ecell2; uint32_t* og = c->egen;
    c->ekey = (int64_t*)xmalloc(sizeof(int64_t) * (nm + 1));
    c->elen = (int*)xmalloc(sizeof(int) * (nm + 1));
    c->ecell1 = (int64_t*)xmalloc(sizeof(int64_t) * (nm + 1));
    c->ecell2 = (int64_t*)xmalloc(sizeof(int64_t) * (nm + 1));
    c->egen = (uint32_t*)xcalloc(nm + 1, sizeof(uint32_t));
    c->emask = nm;
    for (int64_t i = 0; i <= om; ++i) {
        if (og[i] != c->gen) continue;
        uint64_t h = hash_key(c->epool + ok[i], ol[i]) & nm;
        while (c->egen[h] == c->gen) h = (h + 1) & nm;
        c->egen[h] = c->gen; c->ekey[h] = ok[i]; c->elen[h] = ol[i]; c->ecell1[h] = o1[i]; c->ecell2[h] = o2[i];
    }
    free(ok); free(ol); free(o1); free(o2); free(og);
}

/* returns 1 iff the edge already had both endpoints registered */
static int pushedge(orc_ctx* c, const int64_t* key, int len, int64_t cell, int mode) {
    if (2 * (c->ecount + 1) > c->emask) edges_rehash(c);
    uint64_t h = hash_key(key, len) & c->emask;
    for (;;) {
        if (c->egen[h] != c->gen) {
            if (c->epool_len + len > c->epool_cap) {
                while (c->epool_len + len > c->epool_cap) c->epool_cap *= 2;
                c->epool = (int64_t*)realloc(c->epool, sizeof(int64_t) * c->epool_cap);
            }
            c->egen[h] = c->gen; c->ekey[h] = c->epool_len; c->elen[h] = len;
            memcpy(c->epool + c->epool_len, key, sizeof(int64_t) * len); c->epool_len += len;
            c->ecell1[h] = cell; c->ecell2[h] = -1;
            c->ecount++;
            return 0;
        }
        if (c->elen[h] == len && memcmp(c->epool + c->ekey[h], key, sizeof(int64_t) * len) == 0) {
            if (c->ecell2[h] != -1) return 1;
            if (mode || c->ecell1[h] != cell) c->ecell2[h] = cell;
            return 0;
        }
        h = (h + 1) & c->emask;
    }
}

/* ------------------------------------------------------------------ */
int orc_ctx_create(int d, int64_t n, const double* xs, int np, const double* base, const double* normal, orc_ctx** out) {
    if (d < 2 || d > MAXD) { set_err("dimension must be in 2..8"); return -1; }
    if (n <= d) { set_err("There are not enough points to create a Voronoi tessellation"); return -1; }
    orc_ctx* c = (orc_ctx*)xcalloc(1, sizeof(orc_ctx));
    c->d = d; c->n = n; c->np = np;
    c->x = (double*)xmalloc(sizeof(double) * d * (n + np));
    memcpy(c->x, xs, sizeof(double) * d * n);
    memset(c->x + d * n, 0, sizeof(double) * d * np);
    c->pb = (double*)xmalloc(sizeof(double) * d * (np ? np : 1));
    c->pn = (double*)xmalloc(sizeof(double) * d * (np ? np : 1));
    if (np) { memcpy(c->pb, base, sizeof(double) * d * np); memcpy(c->pn, normal, sizeof(double) * d * np); }
    c->active = (char*)xcalloc(np ? np : 1, 1);
    /* check_boundary, src/boundary.jl:437-441 */
    for (int64_t i = 0; i < n; ++i) {
        const double* xi = xs + i * d;
        for (int k = 0; k < d; ++k) if (!isfinite(xi[k])) { set_err("non-finite coordinate"); orc_ctx_free(c); return -2; }
        for (int p = 0; p < np; ++p) {
            double s = 0; for (int k = 0; k < d; ++k) s += (c->pb[p * d + k] - xi[k]) * c->pn[p * d + k];
            if (s < 0) { set_err("node does not lie in the domain given by the boundary"); orc_ctx_free(c); return -3; }
        }
    }
    c->perm = (int64_t*)xmalloc(sizeof(int64_t) * n);
    for (int64_t i = 0; i < n; ++i) c->perm[i] = i;
    for (int k = 0; k < d; ++k) { c->rootlo[k] = DBL_MAX; c->roothi[k] = -DBL_MAX; }
    for (int64_t i = 0; i < n; ++i) for (int k = 0; k < d; ++k) {
        double v = xs[i * d + k]; if (v < c->rootlo[k]) c->rootlo[k] = v; if (v > c->roothi[k]) c->roothi[k] = v;
    }
    double blo[MAXD], bhi[MAXD]; memcpy(blo, c->rootlo, sizeof blo); memcpy(bhi, c->roothi, sizeof bhi);
    kd_build(c, 0, n, blo, bhi);
    c->xr = (double*)xmalloc(sizeof(double) * d * n);
    for (int64_t z = 0; z < n; ++z) memcpy(c->xr + z * d, xs + c->perm[z] * d, sizeof(double) * d);
    c->scratch_cap = 1024;
    c->ids = (int64_t*)xmalloc(sizeof(int64_t) * c->scratch_cap);
    c->ts = (double*)xmalloc(sizeof(double) * c->scratch_cap);
    c->emask = (1 << 16) - 1;
    c->epool_cap = 1 << 16;
    c->epool = (int64_t*)xmalloc(sizeof(int64_t) * c->epool_cap);
    c->ekey = (int64_t*)xmalloc(sizeof(int64_t) * (c->emask + 1));
    c->elen = (int*)xmalloc(sizeof(int) * (c->emask + 1));
    c->ecell1 = (int64_t*)xmalloc(sizeof(int64_t) * (c->emask + 1));
    c->ecell2 = (int64_t*)xmalloc(sizeof(int64_t) * (c->emask + 1));
    c->egen = (uint32_t*)xcalloc(c->emask + 1, sizeof(uint32_t));
    c->gen = 0;
    *out = c;
    return 0;
}

void orc_ctx_free(orc_ctx* c) {
    if (!c) return;
    free(c->x); free(c->pb); free(c->pn); free(c->active); free(c->perm); free(c->xr); free(c->nodes);
    free(c->ids); free(c->ts); free(c->epool); free(c->ekey); free(c->elen); free(c->ecell1); free(c->ecell2); free(c->egen);
    free(c);
}

/* General_EdgeIterator positions for a vertex explored from `cell`:
 * src/edgeiteratebase.jl:37-51 (sig[1]==cell: all; sig[2]==cell: only position 1). */
static int edge_positions(const int64_t* sig, int lsig, int64_t cell, int* a, int* b) {
    if (sig[0] == cell) { *a = 0; *b = lsig - 1; return 1; }
    if (sig[1] == cell) { *a = 0; *b = 0; return 1; }
    *a = 1; *b = 0; return 0;
}

/* queue_edges_general_position, src/edgeiteratebase.jl:128-149 */
static int queue_edges(orc_ctx* c, const int64_t* sig, int lsig, int64_t cell) {
    int d = c->d;
    if (lsig == d + 1 && sig[1] < cell) return 1;
    int a, b; edge_positions(sig, lsig, cell, &a, &b);
    int all = 1;
    int64_t e[MAXSIG];
    for (int s = a; s <= b; ++s) {
        int l = 0; for (int j = 0; j < lsig; ++j) if (j != s) e[l++] = sig[j];
        all &= pushedge(c, e, l, sig[s], 0);
    }
    return all;
}

/* walkray, src/raycast.jl:125-164 -> returns 0 new vertex (sig_out,r_out), 1 unbounded/failure */
static int walkray(orc_ctx* c, const int64_t* sig, int lsig, const double* r, int drop, const double* u, double du,
                   int64_t* sig_out, int* lout, double* r_out) {
    int d = c->d;
    c->stats[0]++;
    int64_t edge[MAXSIG]; int le = 0;
    for (int j = 0; j < lsig; ++j) if (j != drop) edge[le++] = sig[j];
    double t, r2[MAXD];
    int64_t g = raycast_hp(c, edge, le, r, u, edge, le, sig, lsig, 1, du, sig_out, lout, &t, r2);
    if (g == -2) return -2;
    if (t == 0.0 || !isfinite(t) || g < 0) { c->stats[8]++; return 1; }
    walkray_correct_vertex(c, r2, edge, g);
    memcpy(r_out, r2, sizeof(double) * d);
    return 0;
}

typedef struct { int64_t vid; } qitem;

/* fraud_vertex, src/sysvoronoi.jl:527-552 (only for keys longer than 2^d) */
static int fraud_vertex(orc_ctx* c, const int64_t* sig, int lsig, const double* r) {
    int d = c->d;
    if (lsig <= (1 << d)) return 0;
    double maxd = 0; for (int i = 0; i < lsig; ++i) { double v = sqrt(dist2d(r, c->x + sig[i] * d, d)); if (v > maxd) maxd = v; }
    /* verify_vertex: no generator strictly closer than the key's generators */
    double best; int64_t bi = ext_nn(c, r, &(skipper){0, 0, 0}, &best);
    if (bi >= 0 && best < maxd * (1 - 1e-10) && !in_list(bi, sig, lsig)) return 1;
    double dist = 0, maxdist = 0;
    for (int k = 0; k < lsig - 1; ++k) for (int i = k + 1; i < lsig; ++i) {
        double dd = sqrt(dist2d(c->x + sig[i] * d, c->x + sig[k] * d, d));
        dist += dd; if (dd > maxdist) maxdist = dd;
    }
    dist /= lsig * (lsig - 1) / 2.0;
    if (maxdist / maxd < 0.001 * lsig / (double)(1 << d) || dist / maxd < 1000 * PLANE_TOL) return 1;
    return 0;
}

/* systematic_explore_vertex, src/sysvoronoi.jl:490-525 */
static int explore_vertex(orc_ctx* c, orc_mesh* m, int64_t vid, int64_t cell, int64_t** queue, int64_t* qn, int64_t* qcap) {
    int d = c->d;
    int64_t sig[MAXSIG]; int lsig = m->klen[vid];
    memcpy(sig, m->kpool + m->koff[vid], sizeof(int64_t) * lsig);
    double r[MAXD]; memcpy(r, m->coords + vid * d, sizeof(double) * d);
    if (lsig != d + 1) { set_err("non-general vertex reached (FastEdgeIterator path is not restated)"); return -10; }
    int a, b; edge_positions(sig, lsig, cell, &a, &b);
    int k = 0;
    for (int s = a; s <= b; ++s) {
        int64_t e[MAXSIG]; int l = 0; for (int j = 0; j < lsig; ++j) if (j != s) e[l++] = sig[j];
        int bb = pushedge(c, e, l, cell, 1);
        if (e[0] != cell || bb) continue;
        double u[MAXD], du;
        u_qr(c, sig, lsig, s, u, &du);
        int64_t sig2[MAXSIG]; int l2; double r2[MAXD];
        int w = walkray(c, sig, lsig, r, s, u, du, sig2, &l2, r2);
        if (w < 0) return w;
        if (w == 1) continue;  /* pushray! (unbounded) */
        if (l2 <= d) continue;
        if (l2 <= d + 1 || mesh_find(m, sig2, l2) < 0) {
            if (fraud_vertex(c, sig2, l2, r2)) { m->stats[7]++; continue; }
            ++k;
            int existed = mesh_push(m, sig2, l2, r2);
            if (queue_edges(c, sig2, l2, cell)) continue;
            if (existed) continue;
            int64_t nv = mesh_find(m, sig2, l2);
            if (*qn == *qcap) { *qcap *= 2; *queue = (int64_t*)realloc(*queue, sizeof(int64_t) * *qcap); }
            (*queue)[(*qn)++] = nv;
        }
    }
    return k;
}

/* systematic_explore_cell, src/sysvoronoi.jl:384-442 */
static int explore_cell(orc_ctx* c, orc_mesh* m, int64_t cell, int64_t** queue, int64_t* qcap) {
    int d = c->d;
    memset(c->active, 0, c->np);
    activate_cell(c, cell);
    int64_t qn = 0;
    edges_clear(c);
    int64_t mm = m->cell_n[cell];
    if (mm > 0) {
        for (int64_t i = 0; i < mm; ++i) {
            int64_t v = m->cell_v[cell][i];
            queue_edges(c, m->kpool + m->koff[v], m->klen[v], cell);
        }
        for (int64_t i = 0; i < mm; ++i) {
            int rc = explore_vertex(c, m, m->cell_v[cell][i], cell, queue, &qn, qcap);
            if (rc < 0) return rc;
        }
    }
    if (qn == 0 && mm == 0) {
        int64_t sig2[MAXSIG]; double r2[MAXD];
        int l2 = descent(c, cell, sig2, r2);
        if (l2 < 0) return l2;
        mesh_push(m, sig2, l2, r2);
        int64_t nv = mesh_find(m, sig2, l2);
        (*queue)[qn++] = nv;
        queue_edges(c, sig2, l2, cell);
    }
    while (qn > 0) {
        int64_t v = (*queue)[--qn];
        int rc = explore_vertex(c, m, v, cell, queue, &qn, qcap);
        if (rc < 0) return rc;
    }
    (void)d;
    return 0;
}

int orc_voronoi(orc_ctx* c, int64_t cell_begin, int64_t cell_end, uint64_t seed, orc_mesh** out) {
    if (cell_begin < 0 || cell_end > c->n || cell_begin > cell_end) { set_err("bad cell range"); return -1; }
    orc_mesh* m = mesh_new(c->d, c->n);
    memset(c->stats, 0, sizeof c->stats);
    c->rng = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
    int64_t qcap = 1024; int64_t* queue = (int64_t*)xmalloc(sizeof(int64_t) * qcap);
    int rc = 0;
    for (int64_t cell = cell_begin; cell < cell_end; ++cell) {
        rc = explore_cell(c, m, cell, &queue, &qcap);
        if (rc < 0) break;
    }
    free(queue);
    for (int i = 0; i < ORC_NSTATS; ++i) m->stats[i] += c->stats[i];
    if (rc < 0) { orc_mesh_free(m); return rc; }
    *out = m;
    return 0;
}

int64_t orc_mesh_nvert(const orc_mesh* m) { return m->nv; }
int orc_mesh_maxsig(const orc_mesh* m) { int mx = 0; for (int64_t v = 0; v < m->nv; ++v) if (m->klen[v] > mx) mx = m->klen[v]; return mx; }
int orc_mesh_export(const orc_mesh* m, int64_t* keys, double* coords) {
    int ms = orc_mesh_maxsig(m);
    for (int64_t v = 0; v < m->nv; ++v) {
        for (int j = 0; j < ms; ++j) keys[v * ms + j] = j < m->klen[v] ? m->kpool[m->koff[v] + j] + 1 : 0;
        memcpy(coords + v * m->d, m->coords + v * m->d, sizeof(double) * m->d);
    }
    return 0;
}
int orc_mesh_stats(const orc_mesh* m, int64_t* out) { memcpy(out, m->stats, sizeof m->stats); return 0; }
/* the counters of the context's last orc_walkray */
int orc_ctx_stats(const orc_ctx* c, int64_t* out) { memcpy(out, c->stats, sizeof c->stats); return 0; }
void orc_mesh_free(orc_mesh* m) {
    if (!m) return;
    free(m->koff); free(m->klen); free(m->kpool); free(m->coords); free(m->htab);
    for (int64_t g = 0; g < m->n; ++g) free(m->cell_v[g]);
    free(m->cell_v); free(m->cell_n); free(m->cell_cap);
    free(m);
}

int orc_walkray(orc_ctx* c, const int64_t* sig1, int drop, const double* r, int64_t cell1,
                int64_t* out_sig, int* out_len, double* out_r, double* out_u) {
    int d = c->d;
    int64_t sig[MAXSIG];
    for (int j = 0; j <= d; ++j) sig[j] = sig1[j] - 1;
    memset(c->stats, 0, sizeof c->stats);
    int64_t cell = cell1 - 1;
    if (cell < 0 || cell >= c->n) { set_err("bad cell"); return -1; }
    memset(c->active, 0, c->np);
    activate_cell(c, cell);
    double u[MAXD], du;
    u_qr(c, sig, d + 1, drop, u, &du);
    if (out_u) memcpy(out_u, u, sizeof(double) * d);
    int64_t s2[MAXSIG]; int l2; double r2[MAXD];
    int w = walkray(c, sig, d + 1, r, drop, u, du, s2, &l2, r2);
    if (w != 0) return w;
    for (int j = 0; j < l2; ++j) out_sig[j] = s2[j] + 1;
    *out_len = l2;
    memcpy(out_r, r2, sizeof(double) * d);
    return 0;
}
