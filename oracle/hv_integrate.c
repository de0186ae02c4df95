/*
 * hv_integrate.c -- CPU oracle for cell volumes and interface areas (TEST INFRASTRUCTURE).
 *
 * FP64 restatements of the reference's two integrators, on a finished vertex set:
 *
 *  - orc_polytope: the exact polytope integrator (src/polyintegrator.jl:80-468).  For cell
 *    i and neighbour k, the interface F_ik is decomposed recursively: a k-face with apex at
 *    the midpoint of its vertices (src/polyintegrator.jl:494-507) is the union of cones over
 *    its (k-1)-faces, so |F| = (1/k) sum_G h(apex, aff G) |G|, down to edges whose measure
 *    is their length.  The reference evaluates the cone heights through signed minors of
 *    the chain of apexes (src/Leibnitzrule.jl:7-85); here the height is the distance from
 *    the apex to the sub-face's hyperplane inside the face's affine hull (projection onto
 *    the orthogonal complement of the face's constraint normals), the same quantity.
 *    Volumes: vol_i = sum_k A_ik * h_ik / d with h_ik = |x_i - x_k| / 2 for a generator and
 *    the distance to the plane for a boundary facet (src/polyintegrator.jl:296-312).
 *
 *  - orc_mc_dirs: the directional Monte-Carlo integrator (src/mcintegrator.jl:98-204,
 *    mc_raycast :207-226, symmetrisation :176-200).  Each cell draws ndir independent
 *    uniform directions; the reference draws one direction and repeats it (the loop at
 *    src/mcintegrator.jl:85-93 never re-enters after the first draw), which is fixed here.
 *    Directions come from a counter-based stream: normal number s of cell i, direction q
 *    is Box-Muller over splitmix64 outputs 2p, 2p+1 (p = s / 2) with s running over
 *    ((i * ndir + q) * dpad + k), dpad = d rounded up to even.
 *
 * Only tests/ may load this; the product path never calls it.
 * Index conventions: generators 1..n, boundary plane p is n+p (1-based), keys sorted.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IMAXD 8

typedef struct {
    int d;
    int64_t n;
    const double* xs;      /* n x d, point i at xs[i*d] */
    int np;
    const double* pbase;   /* np x d */
    const double* pnormal; /* np x d */
    const int64_t* keys;   /* nv x (d+1), 1-based */
    const double* coords;  /* nv x d */
} geo_t;

static double dot(const double* a, const double* b, int d) {
    double s = 0.0;
    for (int k = 0; k < d; ++k) s += a[k] * b[k];
    return s;
}

static int key_has(const geo_t* g, int64_t v, int64_t id) {
    const int64_t* k = g->keys + v * (g->d + 1);
    for (int j = 0; j <= g->d; ++j)
        if (k[j] == id) return 1;
    return 0;
}

/* constraint normal of member b of a face whose first real generator is a (1-based ids) */
static void grad(const geo_t* g, int64_t a, int64_t b, double* w) {
    const int d = g->d;
    if (b > g->n) {
        memcpy(w, g->pnormal + (b - g->n - 1) * d, sizeof(double) * d);
    } else {
        for (int k = 0; k < d; ++k) w[k] = g->xs[(b - 1) * d + k] - g->xs[(a - 1) * d + k];
    }
}

/* orthonormal basis (modified Gram-Schmidt, two passes) of the normals of S minus a */
static int face_basis(const geo_t* g, const int64_t* S, int m, int64_t a, double Q[IMAXD][IMAXD]) {
    const int d = g->d;
    int q = 0;
    for (int s = 0; s < m; ++s) {
        if (S[s] == a) continue;
        double w[IMAXD];
        grad(g, a, S[s], w);
        for (int pass = 0; pass < 2; ++pass)
            for (int r = 0; r < q; ++r) {
                const double c = dot(w, Q[r], d);
                for (int k = 0; k < d; ++k) w[k] -= c * Q[r][k];
            }
        const double nw = sqrt(dot(w, w, d));
        if (!(nw > 0)) return -1;
        for (int k = 0; k < d; ++k) Q[q][k] = w[k] / nw;
        ++q;
    }
    return q;
}

/* distance from p (a point of aff F_S) to aff F_{S+j} inside aff F_S */
static double sub_height(const geo_t* g, const double* p, int64_t a, int64_t j, double Q[IMAXD][IMAXD], int q) {
    const int d = g->d;
    double w[IMAXD];
    grad(g, a, j, w);
    for (int pass = 0; pass < 2; ++pass)
        for (int r = 0; r < q; ++r) {
            const double c = dot(w, Q[r], d);
            for (int k = 0; k < d; ++k) w[k] -= c * Q[r][k];
        }
    const double nr = sqrt(dot(w, w, d));
    double psi;
    if (j > g->n) {
        const double* nn = g->pnormal + (j - g->n - 1) * d;
        psi = dot(nn, g->pbase + (j - g->n - 1) * d, d) - dot(nn, p, d);
        return fabs(psi) / nr;
    }
    double da = 0.0, dj = 0.0;
    for (int k = 0; k < d; ++k) {
        const double ta = p[k] - g->xs[(a - 1) * d + k];
        const double tj = p[k] - g->xs[(j - 1) * d + k];
        da += ta * ta;
        dj += tj * tj;
    }
    return fabs(dj - da) / (2.0 * nr);
}

/* measure of the face S (m ids, sorted not required) with vertex list V (nvx ids) */
static double face_measure(const geo_t* g, int64_t* S, int m, const int64_t* V, int64_t nvx, int64_t* scratch,
                           int* bad) {
    const int d = g->d;
    const int k = d + 1 - m;
    if (k == 1) {
        if (nvx != 2) { *bad = 1; return 0.0; }
        double s = 0.0;
        for (int c = 0; c < d; ++c) {
            const double t = g->coords[V[0] * d + c] - g->coords[V[1] * d + c];
            s += t * t;
        }
        return sqrt(s);
    }
    double apex[IMAXD] = {0};
    for (int64_t v = 0; v < nvx; ++v)
        for (int c = 0; c < d; ++c) apex[c] += g->coords[V[v] * d + c];
    for (int c = 0; c < d; ++c) apex[c] /= (double)nvx;
    int64_t a = 0;
    for (int s = 0; s < m; ++s)
        if (S[s] <= g->n) { a = S[s]; break; }
    double Q[IMAXD][IMAXD];
    const int q = face_basis(g, S, m, a, Q);
    if (q < 0) { *bad = 1; return 0.0; }
    /* candidate sub-faces: generators of the vertex keys not in S, each taken once */
    double total = 0.0;
    int64_t* sub = scratch;
    int64_t* rest = scratch + nvx;
    for (int64_t v = 0; v < nvx; ++v) {
        const int64_t* key = g->keys + V[v] * (d + 1);
        for (int e = 0; e <= d; ++e) {
            const int64_t j = key[e];
            int inS = 0;
            for (int s = 0; s < m; ++s) inS |= (S[s] == j);
            if (inS) continue;
            /* first vertex of V carrying j handles it */
            int seen = 0;
            for (int64_t u = 0; u < v && !seen; ++u) seen = key_has(g, V[u], j);
            if (seen) continue;
            int64_t cnt = 0;
            for (int64_t u = v; u < nvx; ++u)
                if (key_has(g, V[u], j)) sub[cnt++] = V[u];
            if (cnt < k) continue;
            const double h = sub_height(g, apex, a, j, Q, q);
            S[m] = j;
            total += h * face_measure(g, S, m + 1, sub, cnt, rest, bad);
        }
    }
    return total / (double)k;
}

/*
 * cell_ptr / cell_vert: the vertices of each cell (CSR over 0-based cells, 0-based vertex
 * rows); nbr_ptr / nbr: each cell's neighbours, sorted, 1-based (planes n+p).  Writes
 * vol[n] and area[nbr_ptr[n]] aligned with nbr.  Returns 0, or -1 on a degenerate face.
 */
int orc_polytope(int d, int64_t n, const double* xs, int np, const double* pbase, const double* pnormal,
                 int64_t nv, const int64_t* keys, const double* coords, const int64_t* cell_ptr,
                 const int64_t* cell_vert, const int64_t* nbr_ptr, const int64_t* nbr, double* vol, double* area) {
    (void)nv;
    if (d < 2 || d > IMAXD - 1) return -2;
    geo_t g = {d, n, xs, np, pbase, pnormal, keys, coords};
    int64_t maxv = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = cell_ptr[i + 1] - cell_ptr[i];
        if (c > maxv) maxv = c;
    }
    int64_t* facet = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxv + 1));
    int64_t* scratch = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * (d + 2) * (maxv + 1)));
    int bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t* V = cell_vert + cell_ptr[i];
        const int64_t nvx = cell_ptr[i + 1] - cell_ptr[i];
        double V_i = 0.0;
        for (int64_t t = nbr_ptr[i]; t < nbr_ptr[i + 1]; ++t) {
            const int64_t k = nbr[t];
            int64_t cnt = 0;
            for (int64_t v = 0; v < nvx; ++v)
                if (key_has(&g, V[v], k)) facet[cnt++] = V[v];
            int64_t S[IMAXD + 1] = {i + 1, k};
            const double A = (cnt >= d) ? face_measure(&g, S, 2, facet, cnt, scratch, &bad) : 0.0;
            area[t] = A;
            double h;
            if (k > n) {
                const double* nn = pnormal + (k - n - 1) * d;
                h = dot(nn, pbase + (k - n - 1) * d, d) - dot(nn, xs + i * d, d);
            } else {
                double s = 0.0;
                for (int c = 0; c < d; ++c) {
                    const double tt = xs[i * d + c] - xs[(k - 1) * d + c];
                    s += tt * tt;
                }
                h = 0.5 * sqrt(s);
            }
            V_i += A * h / (double)d;
        }
        vol[i] = V_i;
    }
    free(facet);
    free(scratch);
    return bad ? -1 : 0;
}

/* ---- Monte-Carlo (directional) -------------------------------------------------- */

static uint64_t sm64(uint64_t seed, uint64_t s) {
    uint64_t z = seed + (s + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif(uint64_t seed, uint64_t s) { return (double)(sm64(seed, s) >> 11) * (1.0 / 9007199254740992.0); }

static double normal_at(uint64_t seed, uint64_t s) {
    const uint64_t p = s >> 1;
    double u1 = unif(seed, 2 * p);
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = unif(seed, 2 * p + 1);
    const double rad = sqrt(-2.0 * log(u1));
    return (s & 1) ? rad * sin(2.0 * M_PI * u2) : rad * cos(2.0 * M_PI * u2);
}

/* direction q of cell i (0-based): normalised Gaussian vector */
void orc_mc_direction(int d, uint64_t seed, int64_t i, int64_t ndir, int64_t q, double* u) {
    const int dpad = (d + 1) & ~1;
    const uint64_t base = ((uint64_t)i * (uint64_t)ndir + (uint64_t)q) * (uint64_t)dpad;
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
        u[k] = normal_at(seed, base + (uint64_t)k);
        s += u[k] * u[k];
    }
    s = sqrt(s);
    for (int k = 0; k < d; ++k) u[k] /= s;
}

/*
 * nbr_ptr / nbr as for orc_polytope.  vol[n], area aligned with nbr.  Boundary neighbours
 * act through their mirror image of x_i (src/boundary.jl:206-211).
 */
int orc_mc_dirs(int d, int64_t n, const double* xs, int np, const double* pbase, const double* pnormal,
                const int64_t* nbr_ptr, const int64_t* nbr, int64_t ndir, uint64_t seed, double* vol, double* area) {
    (void)np;
    if (d < 2 || d > IMAXD - 1 || ndir < 1) return -2;
    const double cvol = pow(M_PI, 0.5 * d) / tgamma(1.0 + 0.5 * d);
    const double carea = d * cvol;
    int64_t maxk = 0;
    for (int64_t i = 0; i < n; ++i)
        if (nbr_ptr[i + 1] - nbr_ptr[i] > maxk) maxk = nbr_ptr[i + 1] - nbr_ptr[i];
    double* ext = (double*)malloc(sizeof(double) * (size_t)(maxk * d + 1));
    double* nrm = (double*)malloc(sizeof(double) * (size_t)(maxk * d + 1));
    for (int64_t i = 0; i < n; ++i) {
        const double* x = xs + i * d;
        const int64_t b = nbr_ptr[i], K = nbr_ptr[i + 1] - b;
        for (int64_t t = 0; t < K; ++t) {
            const int64_t k = nbr[b + t];
            double* e = ext + t * d;
            if (k > n) {
                const double* nn = pnormal + (k - n - 1) * d;
                const double h = dot(nn, pbase + (k - n - 1) * d, d) - dot(nn, x, d);
                for (int c = 0; c < d; ++c) e[c] = x[c] + 2.0 * h * nn[c];
            } else {
                memcpy(e, xs + (k - 1) * d, sizeof(double) * d);
            }
            double s = 0.0;
            for (int c = 0; c < d; ++c) s += (e[c] - x[c]) * (e[c] - x[c]);
            s = sqrt(s);
            for (int c = 0; c < d; ++c) nrm[t * d + c] = (e[c] - x[c]) / s;
            area[b + t] = 0.0;
        }
        double V = 0.0;
        for (int64_t q = 0; q < ndir; ++q) {
            double u[IMAXD];
            orc_mc_direction(d, seed, i, ndir, q, u);
            const double c0 = dot(x, u, d);
            double ts = INFINITY;
            int64_t win = -1;
            for (int64_t t = 0; t < K; ++t) {
                const double* e = ext + t * d;
                if (dot(e, u, d) <= c0) continue;
                double num = 0.0, den = 0.0;
                for (int c = 0; c < d; ++c) {
                    num += (x[c] - e[c]) * (x[c] - e[c]);
                    den += u[c] * (e[c] - x[c]);
                }
                const double tt = num / (2.0 * den);
                if (tt > 0 && tt < ts) { ts = tt; win = t; }
            }
            V += pow(ts, d);
            if (win >= 0 && ts < INFINITY) area[b + win] += pow(ts, d - 1) / fabs(dot(nrm + win * d, u, d));
        }
        vol[i] = V * cvol / (double)ndir;
        for (int64_t t = 0; t < K; ++t) area[b + t] *= carea / (double)ndir;
    }
    /* symmetrisation with the neighbour's estimate, in cell order (src/mcintegrator.jl:176-200) */
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t t = nbr_ptr[i]; t < nbr_ptr[i + 1]; ++t) {
            const int64_t k = nbr[t];
            if (k > n || k - 1 >= i) continue;
            const int64_t j = k - 1;
            int64_t tj = -1;
            for (int64_t s = nbr_ptr[j]; s < nbr_ptr[j + 1]; ++s)
                if (nbr[s] == i + 1) { tj = s; break; }
            if (tj < 0) continue;
            const double old = area[t], neigh = area[tj];
            const double nw = fabs(neigh) < old * 1.0e-10 ? old : 0.5 * (old + neigh);
            double s = 0.0;
            for (int c = 0; c < d; ++c) s += (xs[i * d + c] - xs[j * d + c]) * (xs[i * d + c] - xs[j * d + c]);
            const double factor = 0.5 * sqrt(s) / (double)d;
            vol[j] += (nw - neigh) * factor;
            vol[i] += (nw - old) * factor;
            area[t] = nw;
            area[tj] = nw;
        }
    }
    free(ext);
    free(nrm);
    return 0;
}
