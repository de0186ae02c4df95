// hv_integrate.hip -- gfx950 kernels and host driver of cell volumes and interface areas.
//
// Exact polytope integrator (the reference's src/polyintegrator.jl:80-468 computes, cell by
// cell, each interface F_ik by a recursive cone decomposition and vol_i = sum_k A_ik h_ik/d,
// :296-312).  Here the same decomposition runs over all faces of the tessellation at once,
// one face dimension per pass:
//   * in general position every subset S of a vertex key that holds a generator is a
//     face; faces with |S| = m have dimension k = d + 1 - m and are identified by their
//     generator set, packed into 128 bits (21 bits per id);
//   * pass m: emit the (S, vertex) incidences of every vertex, radix-sort by S, run heads
//     give the faces and their vertex lists;
//   * edges (m = d) measure |v_a - v_b|; a face of dimension k > 1 is the union of cones
//     from one of its vertices p over its (k-1)-faces T = S + {j} that miss p:
//     |F_S| = (1/k) sum_T h(p, aff F_T) |F_T|, h measured inside aff F_S;
//   * a sub-face T is found from a vertex v of S that holds j: T is the subset of v's key
//     with S's positions and j's, whose incidence (v, subset) the pass m + 1 maps to its face
//     (one load; the map also says whether v is T's first vertex, which takes T once);
//   * facets (m = 2) are the interfaces A_ik; volumes take the generator as apex.
// 1, 16 or 64 lanes per face read its children in a fixed order: no atomics, the result does
// not depend on scheduling.
//
// Monte-Carlo integrator (src/mcintegrator.jl:98-204): one workgroup per cell casts ndir
// rays from the generator (mc_raycast, :207-226) against the cell's neighbours (boundary
// planes through their mirror image of the generator), accumulates t^d and the facet
// hits, then the neighbour estimates are symmetrised (:176-200).  Each cell draws its own
// directions (counter-based splitmix64 + Box-Muller); the reference reuses one direction
// for every sample (the draw loop at :85-93 never re-enters), which is not carried over.
#include "hv_device.h"
#include "hv_internal.h"
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <numeric>

#define HVCK(call)                                                                 \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return hv_fail(c, HV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
static T* iget(hv_ctx* c, const char* name, size_t count) {
    return (T*)c->pool->get(name, std::max<size_t>(1, count) * sizeof(T), false, c->stream);
}

constexpr uint64_t HV_F21 = 0x1FFFFFull;   // id field; all ones = empty slot
// face incidences one pass of a chunk may hold (a chunk of cells at a time beyond)
constexpr int64_t HV_INC_BUDGET = 1600000000ll;

__device__ __forceinline__ void hv_fkey_put(uint64_t& lo, uint64_t& hi, int s, uint64_t id) {
    if (s < 3) lo = (lo & ~(HV_F21 << (21 * s))) | (id << (21 * s));
    else hi = (hi & ~(HV_F21 << (21 * (s - 3)))) | (id << (21 * (s - 3)));
}
__device__ __forceinline__ int hv_fkey_get(uint64_t lo, uint64_t hi, int s) {
    return (int)(s < 3 ? (lo >> (21 * s)) & HV_F21 : (hi >> (21 * (s - 3))) & HV_F21);
}
// (hi, lo) lexicographic: the order the two-pass radix sort leaves
__device__ __forceinline__ bool hv_fkey_less(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
    return ah < bh || (ah == bh && al < bl);
}

// dense copy of the live vertex records
template <int D>
__global__ void k_int_gather(const int* __restrict__ ids, int nv, const int* __restrict__ key,
                             const double* __restrict__ x, int* __restrict__ vk, double* __restrict__ vc) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int s = ids[v];
#pragma unroll
    for (int j = 0; j <= D; ++j) vk[(int64_t)v * (D + 1) + j] = key[(int64_t)s * (D + 1) + j];
#pragma unroll
    for (int k = 0; k < D; ++k) vc[(int64_t)v * D + k] = x[(int64_t)s * D + k];
}

// packed key words of each vertex (21 bits per id): word 0 holds ids 0..2, word 1 ids 3..5,
// word 2 id 6; sorting by (w0, w1, w2) gives a numbering fixed by the vertex set alone
template <int D>
__global__ void k_vkey_words(const int* __restrict__ vk, int nv, uint64_t* __restrict__ w0, uint64_t* __restrict__ w1,
                             uint64_t* __restrict__ w2, int* __restrict__ idx) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint64_t a = 0, b = 0, c = 0;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const uint64_t id = (uint64_t)vk[(int64_t)v * (D + 1) + j];
        if (j < 3) a |= id << (21 * (2 - j));
        else if (j < 6) b |= id << (21 * (5 - j));
        else c = id;
    }
    w0[v] = a;
    w1[v] = b;
    w2[v] = c;
    idx[v] = v;
}

template <int D>
__global__ void k_int_permute(const int* __restrict__ perm, int nv, const int* __restrict__ vk, const double* __restrict__ vc,
                              int* __restrict__ vk2, double* __restrict__ vc2) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int s = perm[v];
#pragma unroll
    for (int j = 0; j <= D; ++j) vk2[(int64_t)v * (D + 1) + j] = vk[(int64_t)s * (D + 1) + j];
#pragma unroll
    for (int k = 0; k < D; ++k) vc2[(int64_t)v * D + k] = vc[(int64_t)s * D + k];
}

// k_face_emit numbers the m-subsets of a key by their position masks in increasing order: the
// c-th mask of popcount m (those below 2^(d+1) come first, so one list serves every d)
struct HvMaskList {
    unsigned char m[8][36];
    constexpr HvMaskList() : m{} {
        int cnt[8] = {};
        for (int mask = 0; mask < 128; ++mask) {
            int p = 0;
            for (int b = 0; b < 7; ++b) p += (mask >> b) & 1;
            m[p][cnt[p]++] = (unsigned char)mask;
        }
    }
};
__constant__ HvMaskList hv_mask_list = HvMaskList();

// (face, vertex) incidences of pass m, one thread each (contiguous stores): incidence
// o = v * C + c is the c-th m-subset of vertex v's key; one holding no cell of the chunk (cmask
// over the ids; ids below N are generators) gets the all-ones key, which sorts past the end
// (k_first_sent counts them after the sort: no same-address atomics here)
template <int D>
__global__ void k_face_emit(const int* __restrict__ vk, int nv, int m, int N, const uint8_t* __restrict__ cmask, int C,
                            uint64_t* __restrict__ lo, uint64_t* __restrict__ hi, int* __restrict__ idx) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < (int64_t)nv * C) {
        const int v = (int)(o / C), c = (int)(o - (int64_t)v * C);
        const int mask = hv_mask_list.m[m][c];
        const int* k = vk + (int64_t)v * (D + 1);
        uint64_t l = ~0ull, h = ~0ull;
        int s = 0;
        bool real = false;
#pragma unroll
        for (int j = 0; j <= D; ++j) {
            if ((mask >> j) & 1) {
                const int id = k[j];
                real = real || (id < N && cmask[id] != 0);
                hv_fkey_put(l, h, s, (uint64_t)id);
                ++s;
            }
        }
        if (!real) { l = ~0ull; h = ~0ull; }
        lo[o] = l;
        hi[o] = h;
        idx[o] = (int)o;
    }
}

// incidences sent past the end of a sorted pass (lo in sorted order: all ones exactly for them,
// a real key's first id is below all ones): n minus the first such position, by bisection
__global__ void k_first_sent(const uint64_t* __restrict__ lo, int64_t n, unsigned long long* __restrict__ nsent) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t a = 0, b = n;
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (lo[mid] == ~0ull) b = mid;
        else a = mid + 1;
    }
    *nsent = (unsigned long long)(n - a);
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const int* __restrict__ perm, int64_t n,
                             uint64_t* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// sorted incidences -> run heads; incidence o belongs to vertex o / C
__global__ void k_face_heads(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                             const int* __restrict__ perm, int64_t n, int C, int* __restrict__ vid,
                             uint8_t* __restrict__ head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vid[i] = perm[i] / C;
    head[i] = (i == 0 || lo[i] != lo[i - 1] || hi[i] != hi[i - 1]) ? 1 : 0;
}

__global__ void k_iota64(int* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int)i;
}

// face table of a pass: key, first incidence, apex vertex
__global__ void k_face_table(const int* __restrict__ fstart, int nface, const uint64_t* __restrict__ lo,
                             const uint64_t* __restrict__ hi, uint64_t* __restrict__ flo, uint64_t* __restrict__ fhi) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    flo[f] = lo[fstart[f]];
    fhi[f] = hi[fstart[f]];
}

// k_face_emit numbers the m-subsets of a key by their position masks in increasing order, so
// the subset with mask M is the rank of M among the masks of its popcount
struct HvMaskRank {
    unsigned char r[256];
    constexpr HvMaskRank() : r{} {
        int cnt[9] = {};
        for (int mask = 0; mask < 256; ++mask) {
            int p = 0;
            for (int b = 0; b < 8; ++b) p += (mask >> b) & 1;
            r[mask] = (unsigned char)cnt[p]++;
        }
    }
};
__constant__ HvMaskRank hv_mask_rank = HvMaskRank();

// face of each incidence of a pass (incidence o = vertex * C + subset): f for the face's first
// incidence (its smallest vertex), ~f for the others; the parent pass finds a sub-face
// S + {j} of a face S from one of its vertices with one load
__global__ void k_face_inc(const int* __restrict__ fstart, int nface, const int* __restrict__ perm,
                           int* __restrict__ foi) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int b = fstart[f], e = fstart[f + 1];
    foi[perm[b]] = f;
    for (int i = b + 1; i < e; ++i) foi[perm[i]] = ~f;
}

template <int D>
__global__ void k_edge_len(const int* __restrict__ fstart, int nface, const int* __restrict__ vid,
                           const double* __restrict__ vc, double* __restrict__ meas, int* __restrict__ bad) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int b = fstart[f], e = fstart[f + 1];
    if (e - b != 2) { atomicOr(bad, 1); meas[f] = 0.0; return; }
    const int va = vid[b], vb = vid[b + 1];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double t = vc[(int64_t)va * D + k] - vc[(int64_t)vb * D + k];
        s = fma(t, t, s);
    }
    meas[f] = sqrt(s);
}

struct IntGeo {
    int N, P;
    int nreal;          // cells integrated: 0 .. nreal-1 (N - nreal periodic images follow)
    const double* xo;   // N x D, original order
    const double* pn;   // P x D outer normals
    const double* ph;   // P offsets <n_p, b_p>
};

// constraint normal of member b of a face whose first generator is a
template <int D>
__device__ __forceinline__ void hv_face_grad(const IntGeo& g, int a, int b, double* w) {
    if (b >= g.N) {
#pragma unroll
        for (int k = 0; k < D; ++k) w[k] = g.pn[(b - g.N) * D + k];
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) w[k] = g.xo[(int64_t)b * D + k] - g.xo[(int64_t)a * D + k];
    }
}

// Sub-face references of the incidences of a pass, written in incidence order (vertex-major:
// a vertex's key and its row of the previous pass's incidence-to-face map are read once for
// all its subsets, contiguously), so that the measure of a face reads one record per vertex
// instead of the vertex's key and that row.  Entry q of incidence o (the q-th key position e
// not in the subset, ascending): the id j = key[e] in the low word, the pass-(m + 1) map entry
// of the subset + {e} in the high word; the ne = d + 1 - m entries of an incidence lie together
// at p[o * ne] (faces of m >= d - 3 ids: at most four entries, in the four key arrays' space).
struct HvRec {
    uint64_t* p;
    int ne;
};

template <int D>
__global__ void k_face_rec(const int* __restrict__ vk, int64_t tot, int m, int C, int C1, const int* __restrict__ foi1,
                           HvRec rec) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= tot) return;
    const int64_t v = o / C;
    const int mask = hv_mask_list.m[m][(int)(o - v * C)];
    const int* kv = vk + v * (D + 1);
    int q = 0;
#pragma unroll
    for (int e = 0; e <= D; ++e) {
        if ((mask >> e) & 1) continue;
        const int x = foi1[v * C1 + hv_mask_rank.r[mask | (1 << e)]];
        rec.p[o * rec.ne + q] = (uint64_t)(uint32_t)kv[e] | ((uint64_t)(uint32_t)x << 32);
        ++q;
    }
}

// Measure of each face of pass m from the faces of pass m + 1.  L lanes per face (1, 16 or 64),
// splitting the face's vertex list (faces of many vertices: the facets and ridges of cells in
// 5D and up), the partial sums reduced in a fixed order.  (A dispatch holds at most 2^32
// work-items, so the L > 1 forms walk the faces grid-stride.)
template <int D, int L>
__device__ __forceinline__ void hv_face_measure_one(int f, int lane, int step, const IntGeo& g, int m,
                                                    const int* __restrict__ fstart, const int* __restrict__ vid,
                                                    const uint64_t* __restrict__ flo, const uint64_t* __restrict__ fhi,
                                                    int nface1, int C1, const int* __restrict__ foi1,
                                                    const double* __restrict__ meas1,
                                                    const int* __restrict__ vk, const double* __restrict__ vc,
                                                    const int* __restrict__ perm, const HvRec& rec,
                                                    double* __restrict__ meas, int* __restrict__ bad) {
    const uint64_t lo = flo[f], hi = fhi[f];
    int S[D + 1];
#pragma unroll
    for (int s = 0; s <= D; ++s) S[s] = s < m ? hv_fkey_get(lo, hi, s) : -1;
    const int a = S[0];   // ids sorted: generators before planes, the face holds one
    // orthonormal basis of the face's constraint normals (two Gram-Schmidt passes)
    double Q[D][D];
    int q = 0;
#pragma unroll
    for (int s = 1; s < D + 1; ++s) {
        if (s >= m) break;
        double w[D];
        hv_face_grad<D>(g, a, S[s], w);
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int r = 0; r < D - 1; ++r) {   // r < q, unrolled: Q stays in registers
                if (r >= q) break;
                double cdot = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) cdot = fma(w[k], Q[r][k], cdot);
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = fma(-cdot, Q[r][k], w[k]);
            }
        double nw = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nw = fma(w[k], w[k], nw);
        nw = sqrt(nw);
        if (!(nw > 0.0)) {   // the same on every lane of a wavefront
            if (lane == 0) { atomicOr(bad, 2); meas[f] = 0.0; }
            return;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) Q[s - 1][k] = w[k] / nw;   // q == s - 1
        ++q;
    }
    const int b0 = fstart[f], e0 = fstart[f + 1];
    const int apex = vid[b0];
    const int* kp = vk + (int64_t)apex * (D + 1);
    double p[D], da = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        p[k] = vc[(int64_t)apex * D + k];
        const double t = p[k] - g.xo[(int64_t)a * D + k];
        da = fma(t, t, da);
    }
    double sum = 0.0;
    // the sub-face T = S + {j} of the face, met at one of its vertices; x: T's entry in the
    // incidence-to-face map of pass m + 1
    auto subface = [&](int j, int x) {
        bool skip = false;
#pragma unroll
        for (int s = 0; s <= D; ++s) skip = skip || (kp[s] == j);
        if (skip) return;   // the sub-face holds the apex (zero height)
        const int t1 = x >= 0 ? x : ~x;
        if (t1 >= nface1) { atomicOr(bad, 4); return; }
        // each sub-face once: from the first of its vertices (lists ascend by vertex)
        if (x < 0) return;
        double w[D];
        hv_face_grad<D>(g, a, j, w);
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int r = 0; r < D - 1; ++r) {   // r < q, unrolled: Q stays in registers
                if (r >= q) break;
                double cdot = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) cdot = fma(w[k], Q[r][k], cdot);
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = fma(-cdot, Q[r][k], w[k]);
            }
        double nr = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nr = fma(w[k], w[k], nr);
        nr = sqrt(nr);
        double h;
        if (j >= g.N) {
            double np_ = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) np_ = fma(g.pn[(j - g.N) * D + k], p[k], np_);
            h = fabs(g.ph[j - g.N] - np_) / nr;
        } else {
            double dj = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double t = p[k] - g.xo[(int64_t)j * D + k];
                dj = fma(t, t, dj);
            }
            h = fabs(dj - da) / (2.0 * nr);
        }
        sum = fma(h, meas1[t1], sum);
    };
    if (rec.p) {
        for (int i = b0 + lane; i < e0; i += step) {
            const uint64_t* ro = rec.p + (int64_t)perm[i] * rec.ne;
#pragma unroll
            for (int z = 0; z < 4; ++z) {
                if (z >= rec.ne) break;
                const uint64_t rw = ro[z];
                subface((int)(uint32_t)rw, (int)(uint32_t)(rw >> 32));
            }
        }
    } else {
        for (int i = b0 + lane; i < e0; i += step) {
            const int v = vid[i];
            const int* kv = vk + (int64_t)v * (D + 1);
            int kr[D + 1], smask = 0;
#pragma unroll
            for (int e = 0; e <= D; ++e) {
                kr[e] = kv[e];
                bool in = false;
#pragma unroll
                for (int s = 0; s <= D; ++s) in = in || (S[s] == kr[e]);
                smask |= in ? 1 << e : 0;
            }
            for (int e = 0; e <= D; ++e)
                if (!((smask >> e) & 1))   // j not in S: T is the subset smask + e in pass m + 1
                    subface(kr[e], foi1[(int64_t)v * C1 + hv_mask_rank.r[smask | (1 << e)]]);
        }
    }
    if constexpr (L > 1) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (lane != 0) return;
    }
    meas[f] = sum / (double)(D + 1 - m);
}

template <int D, int L>
__global__ __launch_bounds__(128) void k_face_measure(IntGeo g, int m, int nface, const int* __restrict__ fstart, const int* __restrict__ vid,
                               const uint64_t* __restrict__ flo, const uint64_t* __restrict__ fhi, int nface1, int C1,
                               const int* __restrict__ foi1, const double* __restrict__ meas1,
                               const int* __restrict__ vk, const double* __restrict__ vc,
                               const int* __restrict__ perm, HvRec rec, double* __restrict__ meas,
                               int* __restrict__ bad) {
    if constexpr (L > 1) {   // one wavefront per workgroup, 64 / L faces at a time
        for (int64_t f = (int64_t)blockIdx.x * (64 / L) + threadIdx.x / L; f < nface; f += (int64_t)gridDim.x * (64 / L))
            hv_face_measure_one<D, L>((int)f, (int)(threadIdx.x % L), L, g, m, fstart, vid, flo, fhi, nface1, C1,
                                      foi1, meas1, vk, vc, perm, rec, meas, bad);
    } else {
        const int f = blockIdx.x * blockDim.x + threadIdx.x;
        if (f < nface)
            hv_face_measure_one<D, 1>(f, 0, 1, g, m, fstart, vid, flo, fhi, nface1, C1, foi1, meas1, vk, vc, perm,
                                      rec, meas, bad);
    }
}

// ---- integrands (the reference's rule, src/polyintegrator.jl:174-353) -----------------------
// Every face G is cut into the simplices of the chains of face centres (the mean of a face's
// vertices) from G down to an edge; f enters only at the two ends of each edge:
//   J(edge) = |e| (f(v1) + f(v2)) / 2,   J(G) = sum over sub-faces H of h_H(c_G) J(H) / dim G,
// h_H(c_G) the distance within G from G's centre to H.  A facet F (dim d-1) then gets
//   I(F) = (d-1)/d J(F) + f(c_F) |F| / d,
// and a cell, with apex at its generator x,
//   bulk = f(x) vol / (d+1) + sum over facets F of dist(x, F) I(F) / (d+1).
// Exact for affine f.  nf <= HV_MAXF components per point.
constexpr int HV_MAXF = 8;

template <int D>
__global__ void k_face_center(const int* __restrict__ fstart, int nface, const int* __restrict__ vid,
                              const double* __restrict__ vc, double* __restrict__ cen) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int b = fstart[f], e = fstart[f + 1];
    double acc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] = 0.0;
    for (int i = b; i < e; ++i) {
        const int v = vid[i];
#pragma unroll
        for (int k = 0; k < D; ++k) acc[k] += vc[(int64_t)v * D + k];
    }
    const double inv = 1.0 / (double)(e - b);
#pragma unroll
    for (int k = 0; k < D; ++k) cen[(int64_t)f * D + k] = acc[k] * inv;
}

__global__ void k_edge_int(const int* __restrict__ fstart, int nface, const int* __restrict__ vid,
                           const double* __restrict__ meas, const double* __restrict__ fv, int nf,
                           double* __restrict__ J) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    if (fstart[f + 1] - fstart[f] != 2) {   // flagged by k_edge_len; the integration fails
        for (int c = 0; c < nf; ++c) J[(int64_t)f * nf + c] = 0.0;
        return;
    }
    const int va = vid[fstart[f]], vb = vid[fstart[f] + 1];
    for (int c = 0; c < nf; ++c)
        J[(int64_t)f * nf + c] = meas[f] * 0.5 * (fv[(int64_t)va * nf + c] + fv[(int64_t)vb * nf + c]);
}

// J of each face of pass m from the faces of pass m + 1, apex at the face's centre (L lanes per
// face, as k_face_measure)
template <int D, int L>
__device__ __forceinline__ void hv_face_int_one(int f, int lane, int step, const IntGeo& g, int m,
                                                const int* __restrict__ fstart, const int* __restrict__ vid,
                                                const uint64_t* __restrict__ flo, const uint64_t* __restrict__ fhi,
                                                const double* __restrict__ cen, int nface1, int C1,
                                                const int* __restrict__ foi1,
                                                const double* __restrict__ J1, int nf, const int* __restrict__ vk,
                                                const int* __restrict__ perm, const HvRec& rec,
                                                double* __restrict__ J, int* __restrict__ bad) {
    const uint64_t lo = flo[f], hi = fhi[f];
    int S[D + 1];
#pragma unroll
    for (int s = 0; s <= D; ++s) S[s] = s < m ? hv_fkey_get(lo, hi, s) : -1;
    const int a = S[0];
    double Q[D][D];
    int q = 0;
#pragma unroll
    for (int s = 1; s < D + 1; ++s) {
        if (s >= m) break;
        double w[D];
        hv_face_grad<D>(g, a, S[s], w);
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int r = 0; r < D - 1; ++r) {   // r < q, unrolled: Q stays in registers
                if (r >= q) break;
                double cdot = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) cdot = fma(w[k], Q[r][k], cdot);
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = fma(-cdot, Q[r][k], w[k]);
            }
        double nw = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nw = fma(w[k], w[k], nw);
        nw = sqrt(nw);
        if (!(nw > 0.0)) {
            if (lane == 0) atomicOr(bad, 2);
            return;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) Q[s - 1][k] = w[k] / nw;   // q == s - 1
        ++q;
    }
    double p[D], da = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        p[k] = cen[(int64_t)f * D + k];
        const double t = p[k] - g.xo[(int64_t)a * D + k];
        da = fma(t, t, da);
    }
    double sum[HV_MAXF];
    for (int c = 0; c < nf; ++c) sum[c] = 0.0;
    const int b0 = fstart[f], e0 = fstart[f + 1];
    // the sub-face T = S + {j}; x: T's entry in the incidence-to-face map of pass m + 1
    auto subface = [&](int j, int x) {
        const int t1 = x >= 0 ? x : ~x;
        if (t1 >= nface1) { atomicOr(bad, 4); return; }
        if (x < 0) return;   // each sub-face once
        double w[D];
        hv_face_grad<D>(g, a, j, w);
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int r = 0; r < D - 1; ++r) {   // r < q, unrolled: Q stays in registers
                if (r >= q) break;
                double cdot = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) cdot = fma(w[k], Q[r][k], cdot);
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = fma(-cdot, Q[r][k], w[k]);
            }
        double nr = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nr = fma(w[k], w[k], nr);
        nr = sqrt(nr);
        double h;
        if (j >= g.N) {
            double np_ = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) np_ = fma(g.pn[(j - g.N) * D + k], p[k], np_);
            h = fabs(g.ph[j - g.N] - np_) / nr;
        } else {
            double dj = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double t = p[k] - g.xo[(int64_t)j * D + k];
                dj = fma(t, t, dj);
            }
            h = fabs(dj - da) / (2.0 * nr);
        }
        for (int c = 0; c < nf; ++c) sum[c] = fma(h, J1[(int64_t)t1 * nf + c], sum[c]);
    };
    if (rec.p) {
        for (int i = b0 + lane; i < e0; i += step) {
            const uint64_t* ro = rec.p + (int64_t)perm[i] * rec.ne;
#pragma unroll
            for (int z = 0; z < 4; ++z) {
                if (z >= rec.ne) break;
                const uint64_t rw = ro[z];
                subface((int)(uint32_t)rw, (int)(uint32_t)(rw >> 32));
            }
        }
    } else {
        for (int i = b0 + lane; i < e0; i += step) {
            const int v = vid[i];
            const int* kv = vk + (int64_t)v * (D + 1);
            int kr[D + 1], smask = 0;
#pragma unroll
            for (int e = 0; e <= D; ++e) {
                kr[e] = kv[e];
                bool in = false;
#pragma unroll
                for (int s = 0; s <= D; ++s) in = in || (S[s] == kr[e]);
                smask |= in ? 1 << e : 0;
            }
            for (int e = 0; e <= D; ++e)
                if (!((smask >> e) & 1))
                    subface(kr[e], foi1[(int64_t)v * C1 + hv_mask_rank.r[smask | (1 << e)]]);
        }
    }
    if constexpr (L > 1) {
        for (int c = 0; c < nf; ++c) {
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1) sum[c] += __shfl_xor(sum[c], off);
        }
        if (lane != 0) return;
    }
    for (int c = 0; c < nf; ++c) J[(int64_t)f * nf + c] = sum[c] / (double)(D + 1 - m);
}

template <int D, int L>
__global__ __launch_bounds__(128) void k_face_int(IntGeo g, int m, int nface, const int* __restrict__ fstart, const int* __restrict__ vid,
                           const uint64_t* __restrict__ flo, const uint64_t* __restrict__ fhi,
                           const double* __restrict__ cen, int nface1, int C1, const int* __restrict__ foi1,
                           const double* __restrict__ J1, int nf, const int* __restrict__ vk,
                           const int* __restrict__ perm, HvRec rec, double* __restrict__ J, int* __restrict__ bad) {
    if constexpr (L > 1) {
        for (int64_t f = (int64_t)blockIdx.x * (64 / L) + threadIdx.x / L; f < nface; f += (int64_t)gridDim.x * (64 / L))
            hv_face_int_one<D, L>((int)f, (int)(threadIdx.x % L), L, g, m, fstart, vid, flo, fhi, cen, nface1, C1,
                                  foi1, J1, nf, vk, perm, rec, J, bad);
    } else {
        const int f = blockIdx.x * blockDim.x + threadIdx.x;
        if (f < nface)
            hv_face_int_one<D, 1>(f, 0, 1, g, m, fstart, vid, flo, fhi, cen, nface1, C1, foi1, J1, nf, vk, perm, rec,
                                  J, bad);
    }
}

// lanes per face of pass m, by the face dimension k = d + 1 - m (polygons and 3-faces hold
// tens of vertices, 4-faces hundreds, 5-faces more).  Measured per pass at 6D N=50k periodic
// (ms per chunk, L = 1 / 16 / 64): k = 2: 130 / 455 / 1270, k = 3: 231 / 304 / 686,
// k = 4: 272 / 188 / 206, k = 5: 160 / 87 / 55; at 5D N=100k: k = 3: 56 / 59 / 126,
// k = 4: 49 / 29 / 31.  A function of (d, m) only, so that every mesh and chunk sums a face's
// cones in the same order.
__host__ inline int hv_face_lanes(int D, int m) {
    const int k = D + 1 - m;
    return k >= 5 ? 64 : (k == 4 ? 16 : 1);
}

// launches k_face_measure or k_face_int (K) with L lanes per face
template <int L, typename K, typename... A>
static void hv_face_launch(K kern, int nface, hipStream_t s, A... args) {
    if (L == 1) kern<<<nblk(std::max(nface, 1), 128), 128, 0, s>>>(args...);
    else kern<<<(unsigned)std::min<int64_t>(std::max<int64_t>(((int64_t)nface * L + 63) / 64, 1), 1 << 20), 64, 0, s>>>(args...);
}

// facets: I(F) = (d-1)/d J(F) + f(c_F) |F| / d
template <int D>
__global__ void k_facet_int(int nface, const double* __restrict__ J, const double* __restrict__ fc,
                            const double* __restrict__ meas, int nf, double* __restrict__ I) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    for (int c = 0; c < nf; ++c)
        I[(int64_t)f * nf + c] = (double)(D - 1) / (double)D * J[(int64_t)f * nf + c] +
                                 fc[(int64_t)f * nf + c] * meas[f] / (double)D;
}

// row t's interface integral from its facet (rows sorted with their facet numbers)
__global__ void k_row_int(int64_t npair, const int* __restrict__ pface, const double* __restrict__ I, int nf,
                          double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npair) return;
    for (int c = 0; c < nf; ++c) out[t * nf + c] = I[(int64_t)pface[t] * nf + c];
}

// cells 0 .. nreal-1: f(x) vol / (d+1) + sum over rows of dist * I(row) / (d+1)
template <int D>
__global__ void k_cell_bulk(IntGeo g, const uint64_t* __restrict__ pkey, const int64_t* __restrict__ ptr,
                            const double* __restrict__ vol, const double* __restrict__ rint,
                            const double* __restrict__ fx, int nf, double* __restrict__ bulk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nreal) return;
    double y[HV_MAXF];
    for (int c = 0; c < nf; ++c) y[c] = fx[(int64_t)i * nf + c] * vol[i];
    for (int64_t t = ptr[i]; t < ptr[i + 1]; ++t) {
        const int k = (int)(pkey[t] & 0xffffffffull);
        double h;
        if (k >= g.N) {
            double nx = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) nx = fma(g.pn[(k - g.N) * D + c], g.xo[(int64_t)i * D + c], nx);
            h = g.ph[k - g.N] - nx;
        } else {
            double s2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const double t2 = g.xo[(int64_t)i * D + c] - g.xo[(int64_t)k * D + c];
                s2 = fma(t2, t2, s2);
            }
            h = 0.5 * sqrt(s2);
        }
        for (int c = 0; c < nf; ++c) y[c] = fma(h, rint[t * nf + c], y[c]);
    }
    for (int c = 0; c < nf; ++c) bulk[(int64_t)i * nf + c] = y[c] / (double)(D + 1);
}

// facets {a, b} -> (cell, neighbour, area) rows for both sides
// rows (cell, neighbour) of the facets: one per side that is a cell of the chunk
__global__ void k_facet_pairs(int nface, const uint64_t* __restrict__ flo, const double* __restrict__ meas, int N,
                              const uint8_t* __restrict__ cmask, uint64_t* __restrict__ pkey, double* __restrict__ parea,
                              int* __restrict__ pface, unsigned long long* __restrict__ np) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    const int a = hv_fkey_get(flo[f], 0, 0), b = hv_fkey_get(flo[f], 0, 1);
    const bool ra = a < N && cmask[a] != 0, rb = b < N && cmask[b] != 0;
    const int n = (ra ? 1 : 0) + (rb ? 1 : 0);
    if (!n) return;
    unsigned long long o = atomicAdd(np, (unsigned long long)n);
    if (ra) {
        pkey[o] = ((uint64_t)a << 32) | (uint64_t)b;
        parea[o] = meas[f];
        pface[o] = f;
        ++o;
    }
    if (rb) {
        pkey[o] = ((uint64_t)b << 32) | (uint64_t)a;
        parea[o] = meas[f];
        pface[o] = f;
    }
}

__global__ void k_set_mask(const int* __restrict__ ids, int n, uint8_t* __restrict__ mask) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) mask[ids[t]] = 1;
}

// vertices holding a cell of the chunk
template <int D>
__global__ void k_vert_in_chunk(const int* __restrict__ vk, int nv, int N, const uint8_t* __restrict__ cmask,
                                uint8_t* __restrict__ flag, int* __restrict__ iota) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    bool in = false;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const int id = vk[(int64_t)v * (D + 1) + j];
        in = in || (id < N && cmask[id] != 0);
    }
    flag[v] = in ? 1 : 0;
    iota[v] = v;
}

template <int D>
__global__ __launch_bounds__(256) void k_count_in_chunk(const int* __restrict__ vk, int nv, int N,
                                                        const uint8_t* __restrict__ cmask, unsigned long long* __restrict__ cnt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (v < nv)
#pragma unroll
        for (int j = 0; j <= D; ++j) {
            const int id = vk[(int64_t)v * (D + 1) + j];
            in = in || (id < N && cmask[id] != 0);
        }
    hv_wave_count(cnt, in);
}

template <int D>
__global__ void k_int_subset(const int* __restrict__ sel, int n, const int* __restrict__ vk, const double* __restrict__ vc,
                             const double* __restrict__ fv, int nf, int* __restrict__ vk2, double* __restrict__ vc2,
                             double* __restrict__ fv2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int v = sel[t];
#pragma unroll
    for (int j = 0; j <= D; ++j) vk2[(int64_t)t * (D + 1) + j] = vk[(int64_t)v * (D + 1) + j];
#pragma unroll
    for (int k = 0; k < D; ++k) vc2[(int64_t)t * D + k] = vc[(int64_t)v * D + k];
    if (fv)
        for (int f = 0; f < nf; ++f) fv2[(int64_t)t * nf + f] = fv[(int64_t)v * nf + f];
}

// row payloads in the sorted order of the rows
__global__ void k_gather_rows(const int* __restrict__ perm, int64_t n, const double* __restrict__ ra,
                              double* __restrict__ parea, const double* __restrict__ ri, double* __restrict__ rint, int nf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm[t];
    parea[t] = ra[s];
    if (ri)
        for (int f = 0; f < nf; ++f) rint[t * nf + f] = ri[(int64_t)s * nf + f];
}

template <int D>
__global__ void k_cell_volumes(IntGeo g, const uint64_t* __restrict__ pkey, const double* __restrict__ parea,
                               int64_t npair, double* __restrict__ vol, int64_t* __restrict__ ptr) {
    // one thread per cell: its rows are contiguous after the sort; binary search the range
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > g.N) return;
    int64_t lo = 0, hi = npair;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int)(pkey[mid] >> 32) < i) lo = mid + 1;
        else hi = mid;
    }
    ptr[i] = lo;
    if (i == g.N) return;
    double V = 0.0;
    for (int64_t t = lo; t < npair && (int)(pkey[t] >> 32) == i; ++t) {
        const int k = (int)(pkey[t] & 0xffffffffull);
        double h;
        if (k >= g.N) {
            double nx = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) nx = fma(g.pn[(k - g.N) * D + c], g.xo[(int64_t)i * D + c], nx);
            h = g.ph[k - g.N] - nx;
        } else {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const double t2 = g.xo[(int64_t)i * D + c] - g.xo[(int64_t)k * D + c];
                s = fma(t2, t2, s);
            }
            h = 0.5 * sqrt(s);
        }
        V = fma(parea[t], h / (double)D, V);
    }
    vol[i] = V;
}

// ---- Monte-Carlo ------------------------------------------------------------------------

__device__ __forceinline__ uint64_t hv_sm64(uint64_t seed, uint64_t s) {
    uint64_t z = seed + (s + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double hv_unif(uint64_t seed, uint64_t s) {
    return (double)(hv_sm64(seed, s) >> 11) * (1.0 / 9007199254740992.0);
}

constexpr int HV_MC_MAXK = 512;

template <int D>
__global__ __launch_bounds__(256) void k_mc_cell(IntGeo g, const int64_t* __restrict__ ptr,
                                                 const uint64_t* __restrict__ pkey, int64_t ndir, uint64_t seed,
                                                 double cvol, double carea, double* __restrict__ vol,
                                                 double* __restrict__ area) {
    __shared__ double ext[HV_MC_MAXK][D], nrm[HV_MC_MAXK][D];
    __shared__ double acc[HV_MC_MAXK];
    __shared__ double vred[256];
    const int i = blockIdx.x;
    const int64_t b = ptr[i];
    const int K = (int)(ptr[i + 1] - b);
    double x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = g.xo[(int64_t)i * D + c];
    for (int t = threadIdx.x; t < K; t += blockDim.x) {
        const int k = (int)(pkey[b + t] & 0xffffffffull);
        double e[D];
        if (k >= g.N) {
            // mirror image of the generator in the boundary plane (src/boundary.jl:206-211)
            double nx = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) nx = fma(g.pn[(k - g.N) * D + c], x[c], nx);
            const double h = g.ph[k - g.N] - nx;
#pragma unroll
            for (int c = 0; c < D; ++c) e[c] = x[c] + 2.0 * h * g.pn[(k - g.N) * D + c];
        } else {
#pragma unroll
            for (int c = 0; c < D; ++c) e[c] = g.xo[(int64_t)k * D + c];
        }
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) s = fma(e[c] - x[c], e[c] - x[c], s);
        s = sqrt(s);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            ext[t][c] = e[c];
            nrm[t][c] = (e[c] - x[c]) / s;
        }
        acc[t] = 0.0;
    }
    __syncthreads();
    constexpr int DP = (D + 1) & ~1;
    double V = 0.0;
    for (int64_t q = threadIdx.x; q < ndir; q += blockDim.x) {
        const uint64_t base = ((uint64_t)i * (uint64_t)ndir + (uint64_t)q) * (uint64_t)DP;
        double u[D], uu = 0.0;
#pragma unroll
        for (int c = 0; c < D; c += 2) {
            const uint64_t p = (base + c) >> 1;
            const double u1 = fmax(hv_unif(seed, 2 * p), 1e-300), u2 = hv_unif(seed, 2 * p + 1);
            const double rad = sqrt(-2.0 * log(u1));
            double sn, cs;
            sincos(2.0 * M_PI * u2, &sn, &cs);
            u[c] = rad * cs;
            if (c + 1 < D) u[c + 1] = rad * sn;
        }
#pragma unroll
        for (int c = 0; c < D; ++c) uu = fma(u[c], u[c], uu);
        uu = sqrt(uu);
        double c0 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            u[c] /= uu;
            c0 = fma(x[c], u[c], c0);
        }
        double ts = INFINITY;
        int win = -1;
        for (int t = 0; t < K; ++t) {
            double eu = 0.0, num = 0.0, den = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const double e = ext[t][c];
                eu = fma(e, u[c], eu);
                num = fma(x[c] - e, x[c] - e, num);
                den = fma(u[c], e - x[c], den);
            }
            if (eu <= c0) continue;
            const double tt = num / (2.0 * den);
            if (tt > 0.0 && tt < ts) { ts = tt; win = t; }
        }
        double tp = 1.0;
#pragma unroll
        for (int c = 0; c < D - 1; ++c) tp *= ts;
        V += tp * ts;
        if (win >= 0 && ts < INFINITY) {
            double cn = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) cn = fma(nrm[win][c], u[c], cn);
            atomicAdd(&acc[win], tp / fabs(cn));
        }
    }
    vred[threadIdx.x] = V;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) vred[threadIdx.x] += vred[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) vol[i] = vred[0] * cvol / (double)ndir;
    for (int t = threadIdx.x; t < K; t += blockDim.x) area[b + t] = acc[t] * carea / (double)ndir;
}

// symmetrise each generator-generator interface with the neighbour's estimate (pairs k < i)
template <int D>
__global__ void k_mc_sym(IntGeo g, const int64_t* __restrict__ ptr, const uint64_t* __restrict__ pkey, int64_t npair,
                         double* __restrict__ vol, double* __restrict__ area) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npair) return;
    const int i = (int)(pkey[t] >> 32), k = (int)(pkey[t] & 0xffffffffull);
    if (k >= g.N || k >= i || i >= g.nreal) return;
    int64_t lo = ptr[k], hi = ptr[k + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int)(pkey[mid] & 0xffffffffull) < i) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= ptr[k + 1] || (int)(pkey[lo] & 0xffffffffull) != i) return;
    const double old = area[t], neigh = area[lo];
    const double nw = fabs(neigh) < old * 1.0e-10 ? old : 0.5 * (old + neigh);
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const double d2 = g.xo[(int64_t)i * D + c] - g.xo[(int64_t)k * D + c];
        s = fma(d2, d2, s);
    }
    const double factor = 0.5 * sqrt(s) / (double)D;
    atomicAdd(&vol[k], (nw - neigh) * factor);
    atomicAdd(&vol[i], (nw - old) * factor);
    area[t] = nw;
    area[lo] = nw;
}

// ---- dense ids -----------------------------------------------------------------------------
// Face keys pack 21 bits per id.  The face passes run on a dense numbering of the ids that
// occur in the mesh's keys (order-preserving, so sorted keys stay sorted): with periodic images
// the working set can exceed 2^21 generators while far fewer of them touch a walked cell.
template <int D>
__global__ void k_mark_used(const int* __restrict__ vk, int nv, int* __restrict__ used) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
#pragma unroll
    for (int j = 0; j <= D; ++j) used[vk[(int64_t)v * (D + 1) + j]] = 1;
}

// dense[i] (exclusive prefix sum of used) -> gid[dense] = i for the used ids
__global__ void k_dense_inverse(const int* __restrict__ used, const int* __restrict__ dense, int n, int* __restrict__ gid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && used[i]) gid[dense[i]] = i;
}

template <int D>
__global__ void k_remap_keys(int* __restrict__ vk, int nv, const int* __restrict__ dense) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
#pragma unroll
    for (int j = 0; j <= D; ++j) vk[(int64_t)v * (D + 1) + j] = dense[vk[(int64_t)v * (D + 1) + j]];
}

// generator rows and planes of the dense numbering: dense ids below ngen are generators
template <int D>
__global__ void k_dense_geo(const int* __restrict__ gid, int nused, int ngen, int N, const double* __restrict__ xo,
                            const double* __restrict__ pn, const double* __restrict__ ph, double* __restrict__ xd,
                            double* __restrict__ pnd, double* __restrict__ phd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nused) return;
    const int o = gid[i];
    if (i < ngen) {
#pragma unroll
        for (int k = 0; k < D; ++k) xd[(int64_t)i * D + k] = xo[(int64_t)o * D + k];
    } else {
        const int p = o - N, q = i - ngen;
#pragma unroll
        for (int k = 0; k < D; ++k) pnd[q * D + k] = pn[p * D + k];
        phd[q] = ph[p];
    }
}

// rows (cell << 32 | neighbour) back to the caller's ids; the map preserves the order
__global__ void k_rows_to_ids(uint64_t* __restrict__ rows, int64_t n, const int* __restrict__ gid) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint64_t r = rows[t];
    rows[t] = ((uint64_t)gid[(int)(r >> 32)] << 32) | (uint64_t)(uint32_t)gid[(int)(r & 0xffffffffull)];
}

// ---- host driver -------------------------------------------------------------------------

struct hv_integral {
    int64_t N = 0;
    std::vector<double> vol, area;
    std::vector<int64_t> ptr, nbr;
    double ms = 0.0;
    int64_t work = 0;                  // point samples: (sample, generator) distances evaluated
    double ms_kernel = 0.0;            // point samples: device time of the sampling kernel
    int nf = 0;                        // integrand components (0: none)
    std::vector<double> bulk, iface;   // nf x N, nf x rows
};

static int64_t binom(int n, int k) {
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

struct Integrand {
    int nf;
    hv_integrand_fn fn;
    void* user;
};

// f at npts device points (D x npts) -> device values (nf x npts), through the caller's function
static int eval_integrand(hv_ctx* c, const Integrand* fi, int D, const double* dpts, int64_t npts, double* dvals) {
    if (npts == 0) return HV_OK;
    std::vector<double> hp((size_t)npts * D), hv((size_t)npts * fi->nf, 0.0);
    HVCK(hipMemcpyAsync(hp.data(), dpts, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    fi->fn(fi->user, npts, hp.data(), hv.data());
    HVCK(hipMemcpyAsync(dvals, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    return HV_OK;
}

// faces of the tessellation, all passes; leaves the sorted (cell, neighbour) rows with areas.
// fi (optional): also the integrals of the rows' interfaces (*rint, npair x nf) by the
// reference's rule (see k_face_int)
template <int D>
static int facet_rows(hv_ctx* c, const hv_mesh* mesh, const IntGeo& g0, uint64_t*& pkey, double*& parea, int64_t& npair,
                      const Integrand* fi = nullptr, double** rint = nullptr) {
    const int nv = (int)mesh->nv;
    int* vk = iget<int>(c, "i_vk", (size_t)nv * (D + 1));
    double* vc = iget<double>(c, "i_vc", (size_t)nv * D);
    int* bad = iget<int>(c, "i_bad", 1);
    unsigned long long* cnt = iget<unsigned long long>(c, "i_cnt", 2);
    if (!vk || !vc || !bad || !cnt) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
    HVCK(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    IntGeo g = g0;          // the face passes' geometry, in dense ids
    int* gid = nullptr;     // dense id -> the caller's id
    int* used = nullptr;    // caller's id -> occurs in a key
    int* dense = nullptr;   // caller's id -> dense id (exclusive prefix sum of used)
    const int nid = g0.N + g0.P;
    int ngen = 0, nused = 0;
    {
        // the live vertices in a numbering fixed by the vertex set (not by the walk's
        // insertion order), so that apexes and summation orders are reproducible
        int* vk0 = iget<int>(c, "i_vk0", (size_t)nv * (D + 1));
        double* vc0 = iget<double>(c, "i_vc0", (size_t)nv * D);
        uint64_t* w0 = iget<uint64_t>(c, "i_w0", nv);
        uint64_t* w1 = iget<uint64_t>(c, "i_w1", nv);
        uint64_t* w2 = iget<uint64_t>(c, "i_w2", nv);
        uint64_t* ws = iget<uint64_t>(c, "i_ws", nv);
        uint64_t* wso = iget<uint64_t>(c, "i_wso", nv);
        int* pa = iget<int>(c, "i_vpa", nv);
        int* pb = iget<int>(c, "i_vpb", nv);
        if (!vk0 || !vc0 || !w0 || !w1 || !w2 || !ws || !wso || !pa || !pb)
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
        k_int_gather<D><<<nblk(nv, 256), 256, 0, c->stream>>>((const int*)mesh->ids.p, nv, (const int*)mesh->key.p,
                                                              (const double*)mesh->x.p, vk0, vc0);
        HVCK(hipGetLastError());
        // dense ids of the generators and planes that occur in the keys
        used = iget<int>(c, "i_used", (size_t)nid + 1);
        dense = iget<int>(c, "i_dense", (size_t)nid + 1);
        if (!used || !dense) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
        HVCK(hipMemsetAsync(used, 0, sizeof(int) * ((size_t)nid + 1), c->stream));
        if (nv) {
            k_mark_used<D><<<nblk(nv, 256), 256, 0, c->stream>>>(vk0, nv, used);
            HVCK(hipGetLastError());
        }
        size_t tbs = 0;
        HVCK(hipcub::DeviceScan::ExclusiveSum(nullptr, tbs, used, dense, nid + 1, c->stream));
        void* tmps = c->pool->get("i_stmp", tbs, false, c->stream);
        if (!tmps) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (scan)");
        HVCK(hipcub::DeviceScan::ExclusiveSum(tmps, tbs, used, dense, nid + 1, c->stream));
        int hcnt[2] = {0, 0};   // used ids below N (generators), used ids in all
        HVCK(hipMemcpyAsync(&hcnt[0], dense + g0.N, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipMemcpyAsync(&hcnt[1], dense + nid, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        ngen = hcnt[0];
        nused = hcnt[1];
        if (nused >= (int)HV_F21)
            return hv_fail(c, HV_ERR_UNSUPPORTED, "more than 2^21 - 2 generators and planes occur in the mesh (21-bit face keys)");
        gid = iget<int>(c, "i_gid", (size_t)std::max(1, nused));
        double* xd = iget<double>(c, "i_xd", (size_t)std::max(1, ngen) * D);
        double* pnd = iget<double>(c, "i_pnd", (size_t)std::max(1, nused - ngen) * D);
        double* phd = iget<double>(c, "i_phd", (size_t)std::max(1, nused - ngen));
        if (!gid || !xd || !pnd || !phd) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
        k_dense_inverse<<<nblk(nid, 256), 256, 0, c->stream>>>(used, dense, nid, gid);
        HVCK(hipGetLastError());
        if (nv) {
            k_remap_keys<D><<<nblk(nv, 256), 256, 0, c->stream>>>(vk0, nv, dense);
            HVCK(hipGetLastError());
        }
        if (nused) {
            k_dense_geo<D><<<nblk(nused, 256), 256, 0, c->stream>>>(gid, nused, ngen, g0.N, g0.xo, g0.pn, g0.ph, xd, pnd, phd);
            HVCK(hipGetLastError());
        }
        g.N = ngen;
        g.P = nused - ngen;
        g.nreal = g0.nreal;
        g.xo = xd;
        g.pn = pnd;
        g.ph = phd;
        k_vkey_words<D><<<nblk(nv, 256), 256, 0, c->stream>>>(vk0, nv, w0, w1, w2, pa);
        HVCK(hipGetLastError());
        size_t tb = 0;
        HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ws, wso, pa, pb, nv, 0, 64, c->stream));
        void* tmp = c->pool->get("i_vtmp", tb, false, c->stream);
        if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
        // least significant word first; each later pass is stable
        const uint64_t* words[3] = {w2, w1, w0};
        int* src = pa;
        int* dst = pb;
        for (int wi = (D >= 6 ? 0 : 1); wi < 3; ++wi) {
            k_gather_u64<<<nblk(nv, 256), 256, 0, c->stream>>>(words[wi], src, nv, ws);
            HVCK(hipGetLastError());
            HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, ws, wso, src, dst, nv, 0, 64, c->stream));
            std::swap(src, dst);
        }
        k_int_permute<D><<<nblk(nv, 256), 256, 0, c->stream>>>(src, nv, vk0, vc0, vk, vc);
        HVCK(hipGetLastError());
    }
    const int nf = fi ? fi->nf : 0;
    double* fv = nullptr;   // f at the vertices (edge ends)
    if (fi) {
        fv = iget<double>(c, "i_fv", (size_t)nv * nf);
        if (!fv) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
        int rc = eval_integrand(c, fi, D, vc, nv, fv);
        if (rc) return rc;
    }
    // Cells to integrate (dense ids), taken a chunk at a time.  A chunk's face passes hold the
    // vertices of its cells only: a face with one of the chunk's cells has all its vertices
    // there, and so do its sub-faces.  A facet of cells in two chunks is measured in both from
    // the same vertex list in the same order (the canonical vertex order, restricted), so the
    // two values are identical.  One chunk unless a pass would exceed HV_INC_BUDGET incidences.
    std::vector<int> hcell;
    {
        std::vector<int> hd((size_t)g0.nreal + 1);
        HVCK(hipMemcpyAsync(hd.data(), dense, sizeof(int) * ((size_t)g0.nreal + 1), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        const auto add = [&](int i) -> bool {
            if (hd[(size_t)i + 1] == hd[(size_t)i]) return false;   // no vertex holds cell i
            hcell.push_back(hd[(size_t)i]);
            return true;
        };
        if (mesh->cells) {
            for (int i : *mesh->cells)
                if (!add(i)) return hv_fail(c, HV_ERR_STATE, "a listed cell of the mesh has no vertex");
        } else {
            for (int i = 0; i < g0.nreal; ++i)
                if (!add(i)) return hv_fail(c, HV_ERR_STATE, "a cell of the mesh has no vertex");
        }
    }
    const int ncell = (int)hcell.size();
    int64_t cmax = 1;
    for (int m = 2; m <= D; ++m) cmax = std::max(cmax, binom(D + 1, m));
    int nchunk = 1;
    if ((int64_t)nv * cmax > HV_INC_BUDGET)
        nchunk = (int)std::min<int64_t>(std::max(ncell, 1), (int64_t)std::ceil(1.3 * (double)nv * (double)cmax / (double)HV_INC_BUDGET));
    if (nchunk > 1) {
        // chunks of nearby cells (Morton order of the generators), so that few vertices are in two
        std::vector<double> hx((size_t)std::max(1, ngen) * D);
        HVCK(hipMemcpyAsync(hx.data(), g.xo, sizeof(double) * (size_t)ngen * D, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        double lo[D], hi[D];
        for (int k = 0; k < D; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
        for (int i : hcell)
            for (int k = 0; k < D; ++k) { lo[k] = std::min(lo[k], hx[(size_t)i * D + k]); hi[k] = std::max(hi[k], hx[(size_t)i * D + k]); }
        constexpr int B = 60 / D;
        std::vector<std::pair<uint64_t, int>> code(ncell);
        for (int q = 0; q < ncell; ++q) {
            const int i = hcell[q];
            uint64_t z = 0;
            uint32_t u[D];
            for (int k = 0; k < D; ++k) {
                const double w = hi[k] > lo[k] ? (hx[(size_t)i * D + k] - lo[k]) / (hi[k] - lo[k]) : 0.0;
                u[k] = (uint32_t)std::min(std::max(w, 0.0) * (double)((1u << B) - 1), (double)((1u << B) - 1));
            }
            for (int b = B - 1; b >= 0; --b)
                for (int k = 0; k < D; ++k) z = (z << 1) | ((u[k] >> b) & 1u);
            code[q] = {z, i};
        }
        std::sort(code.begin(), code.end());
        for (int q = 0; q < ncell; ++q) hcell[q] = code[q].second;
    }
    uint8_t* cmask = iget<uint8_t>(c, "i_cmask", (size_t)std::max(1, nused));
    int* dchunk = iget<int>(c, "i_chunk", (size_t)std::max(1, ncell));
    if (!cmask || !dchunk) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
    if (ncell) HVCK(hipMemcpyAsync(dchunk, hcell.data(), sizeof(int) * ncell, hipMemcpyHostToDevice, c->stream));
    const bool whole = nchunk == 1 && !mesh->cells;   // every vertex takes part: no selection
    auto set_chunk = [&](int q, int nch) -> int {
        const int cb = (int)((int64_t)ncell * q / nch), ce = (int)((int64_t)ncell * (q + 1) / nch);
        HVCK(hipMemsetAsync(cmask, 0, (size_t)std::max(1, nused), c->stream));
        if (ce > cb) {
            k_set_mask<<<nblk(ce - cb, 256), 256, 0, c->stream>>>(dchunk + cb, ce - cb, cmask);
            HVCK(hipGetLastError());
        }
        return HV_OK;
    };
    int64_t nvc_max = nv;   // vertices of the largest chunk
    if (!whole && nchunk > 1) {
        // a vertex of cells in several chunks is in each of them: more chunks until every
        // chunk's passes fit (at most 2^31 - 1 incidences)
        unsigned long long* vcnt = iget<unsigned long long>(c, "i_vcnt", 1);
        if (!vcnt) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
        for (;;) {
            int64_t worst = 0;
            for (int q = 0; q < nchunk; ++q) {
                int rc = set_chunk(q, nchunk);
                if (rc) return rc;
                HVCK(hipMemsetAsync(vcnt, 0, sizeof(unsigned long long), c->stream));
                k_count_in_chunk<D><<<nblk(nv, 256), 256, 0, c->stream>>>(vk, nv, ngen, cmask, vcnt);
                HVCK(hipGetLastError());
                unsigned long long h = 0;
                HVCK(hipMemcpyAsync(&h, vcnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
                HVCK(hipStreamSynchronize(c->stream));
                worst = std::max<int64_t>(worst, (int64_t)h);
            }
            if (worst * cmax < (1ll << 31)) {
                nvc_max = worst;
                break;
            }
            if (nchunk >= ncell) return hv_fail(c, HV_ERR_CAPACITY, "too many face incidences for one cell");
            nchunk = std::min(ncell, 2 * nchunk);
        }
    }
    {
        // the per-incidence buffers of the face passes at their largest size, once: growing them
        // between passes and chunks frees and maps tens of GB while the device idles
        const int64_t tmax = nvc_max * cmax;
        if (tmax >= (1ll << 31)) return hv_fail(c, HV_ERR_CAPACITY, "too many face incidences for one pass");
        const size_t T = (size_t)std::max<int64_t>(tmax, 1);
        bool ok = true;
        ok = ok && c->pool->reserve("i_keys", 32 * T, c->stream);   // lo, hi, lo2, hi2 of a pass
        for (const char* nmv : {"i_pa", "i_pb", "i_iota", "i_vida", "i_vidb", "i_foia", "i_foib"})
            ok = ok && c->pool->reserve(nmv, 4 * T, c->stream);
        ok = ok && c->pool->reserve("i_head", T, c->stream);
        ok = ok && c->pool->reserve("i_fsta", 4 * (T + 1), c->stream) && c->pool->reserve("i_fstb", 4 * (T + 1), c->stream);
        if (!whole) {
            ok = ok && c->pool->reserve("i_vkc", sizeof(int) * (size_t)std::max<int64_t>(nvc_max, 1) * (D + 1), c->stream);
            ok = ok && c->pool->reserve("i_vcc", sizeof(double) * (size_t)std::max<int64_t>(nvc_max, 1) * D, c->stream);
        }
        if (!ok) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (faces)");
        size_t tb = 0, tb2 = 0;
        uint64_t* k0 = (uint64_t*)c->pool->get("i_keys", 0);
        HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k0, (int*)nullptr, (int*)nullptr, (int)tmax, 0, 64, c->stream));
        HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb2, (int*)nullptr, (uint8_t*)nullptr, (int*)nullptr, cnt + 1, (int)tmax, c->stream));
        if (!c->pool->reserve("i_tmp", tb, c->stream) || !c->pool->reserve("i_tmp2", tb2, c->stream))
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
    }
    int64_t ntot = 0;                                  // rows of the chunks so far
    for (int q = 0; q < nchunk; ++q) {
        {
            int rc = set_chunk(q, nchunk);
            if (rc) return rc;
        }
        const int* vkc = vk;
        const double* vcc = vc;
        const double* fvc = fv;
        int nvc = nv;
        if (!whole) {
            // the vertices holding a cell of the chunk, in the canonical order
            uint8_t* vflag = iget<uint8_t>(c, "i_vflag", (size_t)std::max(1, nv));
            int* viota = iget<int>(c, "i_viota", (size_t)std::max(1, nv));
            int* vsel = iget<int>(c, "i_vsel", (size_t)std::max(1, nv));
            if (!vflag || !viota || !vsel) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
            if (nv) {
                k_vert_in_chunk<D><<<nblk(nv, 256), 256, 0, c->stream>>>(vk, nv, ngen, cmask, vflag, viota);
                HVCK(hipGetLastError());
            }
            size_t tbf = 0;
            HVCK(hipcub::DeviceSelect::Flagged(nullptr, tbf, viota, vflag, vsel, cnt + 1, nv, c->stream));
            void* tmpf = c->pool->get("i_tmpf", tbf, false, c->stream);
            if (!tmpf) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
            HVCK(hipcub::DeviceSelect::Flagged(tmpf, tbf, viota, vflag, vsel, cnt + 1, nv, c->stream));
            unsigned long long hsel = 0;
            HVCK(hipMemcpyAsync(&hsel, cnt + 1, sizeof(hsel), hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            nvc = (int)hsel;
            int* vk2 = iget<int>(c, "i_vkc", (size_t)std::max(1, nvc) * (D + 1));
            double* vc2 = iget<double>(c, "i_vcc", (size_t)std::max(1, nvc) * D);
            double* fv2 = fi ? iget<double>(c, "i_fvc", (size_t)std::max(1, nvc) * nf) : nullptr;
            if (!vk2 || !vc2 || (fi && !fv2)) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
            if (nvc) {
                k_int_subset<D><<<nblk(nvc, 256), 256, 0, c->stream>>>(vsel, nvc, vk, vc, fv, nf, vk2, vc2, fv2);
                HVCK(hipGetLastError());
            }
            vkc = vk2;
            vcc = vc2;
            fvc = fv2;
        }
        // per pass: sorted incidences (vid), face starts, keys, measures; "1" = pass m + 1
        int nface1 = 0;
        for (int m = D; m >= 2; --m) {
            const int C = (int)binom(D + 1, m);
            const int64_t tot = (int64_t)nvc * C;
            if (tot >= (1ll << 31)) return hv_fail(c, HV_ERR_CAPACITY, "too many face incidences for one pass");
            const char* sfx = (m & 1) ? "a" : "b";
            auto nm = [&](const char* s) { return std::string(s) + sfx; };
            // the pass's four key arrays in one allocation (their space holds the sub-face
            // records once the face table is out)
            uint64_t* keys = iget<uint64_t>(c, "i_keys", 4 * (size_t)tot);
            uint64_t* lo = keys;
            uint64_t* hi = keys ? keys + tot : nullptr;
            uint64_t* lo2 = keys ? keys + 2 * tot : nullptr;
            uint64_t* hi2 = keys ? keys + 3 * tot : nullptr;
            int* pa = iget<int>(c, "i_pa", tot);
            int* pb = iget<int>(c, "i_pb", tot);
            uint8_t* head = iget<uint8_t>(c, "i_head", tot);
            if (!lo || !hi || !lo2 || !hi2 || !pa || !pb || !head) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (faces)");
            // faces with a cell of the chunk (the others may be incomplete here)
            k_face_emit<D><<<nblk(std::max<int64_t>(tot, 1), 256), 256, 0, c->stream>>>(vkc, nvc, m, g.N, cmask, C, lo, hi, pa);
            HVCK(hipGetLastError());
            // two stable radix passes: by lo, then by hi -> (hi, lo) order.  Only the id fields a
            // pass uses: the empty ones are all ones in every real key, and a real key's last
            // used field is below all ones, so the incidences sent past the end stay there
            const int lo_bits = 21 * std::min(m, 3), hi_bits = 21 * std::max(m - 3, 0);
            size_t tb = 0, tbh = 0;
            HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, lo, lo2, pa, pb, (int)tot, 0, lo_bits, c->stream));
            if (m > 3)
                HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tbh, hi2, hi, pb, pa, (int)tot, 0, hi_bits, c->stream));
            void* tmp = c->pool->get("i_tmp", std::max(tb, tbh), false, c->stream);
            if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
            HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, lo, lo2, pa, pb, (int)tot, 0, lo_bits, c->stream));
            int* perm = pb;
            if (m > 3) {
                k_gather_u64<<<nblk(tot, 256), 256, 0, c->stream>>>(hi, pb, tot, hi2);
                HVCK(hipGetLastError());
                HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tbh, hi2, hi, pb, pa, (int)tot, 0, hi_bits, c->stream));
                perm = pa;
                k_gather_u64<<<nblk(tot, 256), 256, 0, c->stream>>>(lo, perm, tot, lo2);
                HVCK(hipGetLastError());
            }
            // (m <= 3: every key's hi word is all ones, no id past the third field, so hi needs
            // no reordering)
            // lo2 / hi: sorted keys; perm: sorted incidence numbers
            k_first_sent<<<1, 64, 0, c->stream>>>(lo2, tot, cnt);
            HVCK(hipGetLastError());
            unsigned long long hsent = 0;
            HVCK(hipMemcpyAsync(&hsent, cnt, sizeof(hsent), hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            const int64_t nvalid = tot - (int64_t)hsent;
            int* vid = iget<int>(c, nm("i_vid").c_str(), nvalid);
            int* fstart = iget<int>(c, nm("i_fst").c_str(), nvalid + 1);
            int* iota = iget<int>(c, "i_iota", nvalid);
            if (!vid || !fstart || !iota) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (faces)");
            k_face_heads<<<nblk(nvalid, 256), 256, 0, c->stream>>>(lo2, hi, perm, nvalid, C, vid, head);
            HVCK(hipGetLastError());
            k_iota64<<<nblk(nvalid, 256), 256, 0, c->stream>>>(iota, nvalid);
            HVCK(hipGetLastError());
            size_t tb2 = 0;
            HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb2, iota, head, fstart, cnt + 1, (int)nvalid, c->stream));
            void* tmp2 = c->pool->get("i_tmp2", tb2, false, c->stream);
            if (!tmp2) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
            HVCK(hipcub::DeviceSelect::Flagged(tmp2, tb2, iota, head, fstart, cnt + 1, (int)nvalid, c->stream));
            unsigned long long hnf = 0;
            HVCK(hipMemcpyAsync(&hnf, cnt + 1, sizeof(hnf), hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            const int nface = (int)hnf;
            const int endv = (int)nvalid;
            HVCK(hipMemcpyAsync(fstart + nface, &endv, sizeof(int), hipMemcpyHostToDevice, c->stream));
            uint64_t* flo = iget<uint64_t>(c, nm("i_flo").c_str(), nface);
            uint64_t* fhi = iget<uint64_t>(c, nm("i_fhi").c_str(), nface);
            double* meas = iget<double>(c, nm("i_meas").c_str(), nface);
            if (!flo || !fhi || !meas) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (faces)");
            k_face_table<<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, lo2, hi, flo, fhi);
            HVCK(hipGetLastError());
            if (m > 2) {   // the next pass finds its sub-faces here
                int* foi = iget<int>(c, nm("i_foi").c_str(), tot);
                if (!foi) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (faces)");
                HVCK(hipMemsetAsync(foi, 0x80, sizeof(int) * tot, c->stream));   // no face: decodes past the end
                k_face_inc<<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, perm, foi);
                HVCK(hipGetLastError());
            }
            double* J = fi ? iget<double>(c, nm("i_J").c_str(), (size_t)nface * nf) : nullptr;
            if (fi && !J) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
            if (m == D) {
                k_edge_len<D><<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, vid, vcc, meas, bad);
                if (fi) {
                    HVCK(hipGetLastError());
                    k_edge_int<<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, vid, meas, fvc, nf, J);
                }
            } else {
                const char* s1 = (m & 1) ? "b" : "a";
                auto n1 = [&](const char* s) { return std::string(s) + s1; };
                const int C1 = (int)binom(D + 1, m + 1);
                const int* foi1 = (const int*)c->pool->get(n1("i_foi"), 0);
                const double* meas1 = (const double*)c->pool->get(n1("i_meas"), 0);
                if (!foi1 || !meas1) return hv_fail(c, HV_ERR_DEVICE, "face tables of the previous pass missing");
                const int L = hv_face_lanes(D, m);
                // sub-face records in the key arrays' space, free once the face table is out
                HvRec rec{nullptr, 0};
                if (D - m <= 3) {   // facets in 6D (five entries) read the keys directly
                    rec.p = keys;
                    rec.ne = D + 1 - m;
                    k_face_rec<D><<<nblk(std::max<int64_t>(tot, 1), 256), 256, 0, c->stream>>>(
                        (const int*)vkc, tot, m, C, C1, foi1, rec);
                    HVCK(hipGetLastError());
                }
#define HV_FM(LL) hv_face_launch<LL>(k_face_measure<D, LL>, nface, c->stream, g, m, nface, fstart, vid, flo, fhi, \
                                     nface1, C1, foi1, meas1, (const int*)vkc, (const double*)vcc, (const int*)perm, \
                                     rec, meas, bad)
                if (L == 64) HV_FM(64);
                else if (L == 16) HV_FM(16);
                else HV_FM(1);
#undef HV_FM
                if (fi) {
                    HVCK(hipGetLastError());
                    double* cen = iget<double>(c, "i_cen", (size_t)nface * D);
                    const double* J1 = (const double*)c->pool->get(n1("i_J"), 0);
                    if (!cen || !J1) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
                    k_face_center<D><<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, vid, vcc, cen);
                    HVCK(hipGetLastError());
#define HV_FI(LL) hv_face_launch<LL>(k_face_int<D, LL>, nface, c->stream, g, m, nface, fstart, vid, flo, fhi, \
                                     (const double*)cen, nface1, C1, foi1, J1, nf, (const int*)vkc, \
                                     (const int*)perm, rec, J, bad)
                    if (L == 64) HV_FI(64);
                    else if (L == 16) HV_FI(16);
                    else HV_FI(1);
#undef HV_FI
                }
            }
            HVCK(hipGetLastError());
            double* Iface = nullptr;   // facet integrals (m == 2)
            if (fi && m == 2) {
                double* cen = iget<double>(c, "i_cen", (size_t)nface * D);
                double* fc = iget<double>(c, "i_fc", (size_t)nface * nf);
                Iface = iget<double>(c, "i_If", (size_t)nface * nf);
                if (!cen || !fc || !Iface) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
                if (m == D) {   // d = 2: the facets are the edges, no centres yet
                    k_face_center<D><<<nblk(nface, 256), 256, 0, c->stream>>>(fstart, nface, vid, vcc, cen);
                    HVCK(hipGetLastError());
                }
                int rc = eval_integrand(c, fi, D, cen, nface, fc);
                if (rc) return rc;
                k_facet_int<D><<<nblk(nface, 256), 256, 0, c->stream>>>(nface, J, fc, meas, nf, Iface);
                HVCK(hipGetLastError());
            }
            nface1 = nface;
            if (m == 2) {
                pkey = iget<uint64_t>(c, "i_pkey", (size_t)2 * nface);
                uint64_t* pkey2 = iget<uint64_t>(c, "i_pkey2", (size_t)2 * nface);
                parea = iget<double>(c, "i_parea", (size_t)2 * nface);
                double* parea2 = iget<double>(c, "i_parea2", (size_t)2 * nface);
                int* pface2 = iget<int>(c, "i_pface2", (size_t)2 * nface);
                if (!pkey || !pkey2 || !parea || !parea2 || !pface2)
                    return hv_fail(c, HV_ERR_DEVICE, "out of device memory (facets)");
                HVCK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
                k_facet_pairs<<<nblk(nface, 256), 256, 0, c->stream>>>(nface, flo, meas, g.N, cmask, pkey2, parea2, pface2, cnt);
                HVCK(hipGetLastError());
                unsigned long long hp = 0;
                HVCK(hipMemcpyAsync(&hp, cnt, sizeof(hp), hipMemcpyDeviceToHost, c->stream));
                HVCK(hipStreamSynchronize(c->stream));
                npair = (int64_t)hp;
                size_t tb3 = 0;
                HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, pkey2, pkey, parea2, parea, (int)npair, 0, 64, c->stream));
                void* tmp3 = c->pool->get("i_tmp3", tb3, false, c->stream);
                if (!tmp3) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
                HVCK(hipcub::DeviceRadixSort::SortPairs(tmp3, tb3, pkey2, pkey, parea2, parea, (int)npair, 0, 64, c->stream));
                if (fi) {
                    // the same stable sort carrying the facet numbers, then the rows' integrals
                    uint64_t* pkey3 = iget<uint64_t>(c, "i_pkey3", (size_t)2 * nface);
                    int* pface = iget<int>(c, "i_pface", (size_t)2 * nface);
                    *rint = iget<double>(c, "i_rint", (size_t)std::max<int64_t>(npair, 1) * nf);
                    if (!pkey3 || !pface || !*rint) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
                    size_t tb4 = 0;
                    HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb4, pkey2, pkey3, pface2, pface, (int)npair, 0, 64, c->stream));
                    void* tmp4 = c->pool->get("i_tmp4", tb4, false, c->stream);
                    if (!tmp4) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
                    HVCK(hipcub::DeviceRadixSort::SortPairs(tmp4, tb4, pkey2, pkey3, pface2, pface, (int)npair, 0, 64, c->stream));
                    if (npair > 0) {
                        k_row_int<<<nblk(npair, 256), 256, 0, c->stream>>>(npair, pface, Iface, nf, *rint);
                        HVCK(hipGetLastError());
                    }
                }
            }
        }
        if (!whole || nchunk > 1) {
            // this chunk's rows after the rows of the chunks before
            uint64_t* rk = (uint64_t*)c->pool->get("i_rk", sizeof(uint64_t) * (size_t)std::max<int64_t>(ntot + npair, 1), true, c->stream);
            double* ra = (double*)c->pool->get("i_ra", sizeof(double) * (size_t)std::max<int64_t>(ntot + npair, 1), true, c->stream);
            double* ri = fi ? (double*)c->pool->get("i_ri", sizeof(double) * (size_t)std::max<int64_t>(ntot + npair, 1) * nf, true, c->stream)
                            : nullptr;
            if (!rk || !ra || (fi && !ri)) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (rows)");
            if (npair) {
                HVCK(hipMemcpyAsync(rk + ntot, pkey, sizeof(uint64_t) * npair, hipMemcpyDeviceToDevice, c->stream));
                HVCK(hipMemcpyAsync(ra + ntot, parea, sizeof(double) * npair, hipMemcpyDeviceToDevice, c->stream));
                if (fi) HVCK(hipMemcpyAsync(ri + ntot * nf, *rint, sizeof(double) * npair * nf, hipMemcpyDeviceToDevice, c->stream));
            }
            ntot += npair;
        }
    }
    if (!whole || nchunk > 1) {
        // rows of all chunks, sorted by (cell, neighbour)
        uint64_t* rk = (uint64_t*)c->pool->get("i_rk", 0);
        double* ra = (double*)c->pool->get("i_ra", 0);
        double* ri = fi ? (double*)c->pool->get("i_ri", 0) : nullptr;
        npair = ntot;
        pkey = iget<uint64_t>(c, "i_pkey", (size_t)npair);
        parea = iget<double>(c, "i_parea", (size_t)npair);
        int* ia = iget<int>(c, "i_rpa", (size_t)npair);
        int* ib = iget<int>(c, "i_rpb", (size_t)npair);
        if (fi) *rint = iget<double>(c, "i_rint", (size_t)std::max<int64_t>(npair, 1) * nf);
        if (!rk || !ra || !pkey || !parea || !ia || !ib || (fi && (!ri || !*rint)))
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (rows)");
        if (npair > 0) {
            k_iota64<<<nblk(npair, 256), 256, 0, c->stream>>>(ia, npair);
            HVCK(hipGetLastError());
            size_t tbr = 0;
            HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tbr, rk, pkey, ia, ib, (int)npair, 0, 64, c->stream));
            void* tmpr = c->pool->get("i_tmpr", tbr, false, c->stream);
            if (!tmpr) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
            HVCK(hipcub::DeviceRadixSort::SortPairs(tmpr, tbr, rk, pkey, ia, ib, (int)npair, 0, 64, c->stream));
            k_gather_rows<<<nblk(npair, 256), 256, 0, c->stream>>>(ib, npair, ra, parea, ri, fi ? *rint : nullptr, nf);
            HVCK(hipGetLastError());
        }
    }
    if (npair > 0) {
        k_rows_to_ids<<<nblk(npair, 256), 256, 0, c->stream>>>(pkey, npair, gid);
        HVCK(hipGetLastError());
    }
    int hbad = 0;
    HVCK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    if (hbad) return hv_fail(c, HV_ERR_DEGENERATE, "a face is not in general position (integration)");
    return HV_OK;
}

// ---- neighbour rows without face measures (the Monte-Carlo integrators) -------------------
// Cells i and k are neighbours when i and k occur together in a vertex key (src/neighbors.jl:
// 219-251; general position).  Every unordered pair (a < b, a a generator) of every live
// vertex key goes into a device hash set; the set is compacted, mirrored to (a, b) and (b, a)
// rows and sorted, which gives the rows facet_rows leaves (keys (cell << 32) | neighbour).
constexpr uint64_t HV_PAIR_EMPTY = ~0ull;

template <int D>
__global__ void k_pair_insert(const int* __restrict__ ids, int nv, const int* __restrict__ key, int N,
                              unsigned long long* __restrict__ tab, uint64_t mask, int* __restrict__ full) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int* k = key + (int64_t)ids[t] * (D + 1);
    int kk[D + 1];
#pragma unroll
    for (int j = 0; j <= D; ++j) kk[j] = k[j];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        if (kk[a] >= N) break;   // planes sort last
#pragma unroll
        for (int b = a + 1; b <= D; ++b) {
            const unsigned long long pk = ((unsigned long long)kk[a] << 32) | (unsigned)kk[b];
            uint64_t h = hv_mix64(pk) & mask;
            bool done = false;
            for (uint64_t probe = 0; probe <= mask && !done; ++probe) {
                const unsigned long long cur = tab[h];
                if (cur == pk) { done = true; break; }
                if (cur == HV_PAIR_EMPTY) {
                    const unsigned long long old = atomicCAS(&tab[h], HV_PAIR_EMPTY, pk);
                    if (old == HV_PAIR_EMPTY || old == pk) { done = true; break; }
                }
                h = (h + 1) & mask;
            }
            if (!done) atomicExch(full, 1);
        }
    }
}

__global__ void k_pair_mirror(const unsigned long long* __restrict__ u, int64_t nu, int N,
                              uint64_t* __restrict__ rows, int* __restrict__ count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = t < nu;
    const unsigned long long pk = ok ? u[t] : 0ull;
    const int a = (int)(pk >> 32), b = (int)(pk & 0xffffffffull);
    const bool both = ok && b < N;
    int o = hv_wave_alloc(count, ok);
    if (ok) rows[o] = pk;
    o = hv_wave_alloc(count, both);
    if (both) rows[o] = ((uint64_t)b << 32) | (unsigned)a;
}

__global__ void k_row_ptr(const uint64_t* __restrict__ rows, int64_t nrow, int N, int64_t* __restrict__ ptr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    const uint64_t want = (uint64_t)i << 32;
    int64_t lo = 0, hi = nrow;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (rows[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    ptr[i] = lo;
}

struct PairNotEmpty {
    __host__ __device__ bool operator()(const unsigned long long& v) const { return v != HV_PAIR_EMPTY; }
};

// neighbours per cell, a generous bound for the hash-set size (d = 2..6, unit-cube densities)
static int64_t neighbours_bound(int D) {
    static const int nb[7] = {0, 0, 8, 24, 64, 160, 400};
    return nb[std::min(std::max(D, 2), 6)];
}

template <int D>
static int neighbor_rows(hv_ctx* c, const hv_mesh* mesh, uint64_t*& pkey, int64_t& npair, int64_t*& ptr) {
    const int N = (int)c->N;
    const int nv = (int)mesh->nv;
    int* full = iget<int>(c, "n_full", 1);
    int* cnt = iget<int>(c, "n_cnt", 2);
    ptr = iget<int64_t>(c, "i_ptr", (size_t)N + 1);
    if (!full || !cnt || !ptr) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbours)");
    // rows of the walked cells (images add the pairs of their vertices' other generators)
    const int64_t nrows = (int64_t)mesh->nreal * neighbours_bound(D) * (mesh->nreal < N ? 2 : 1);
    uint64_t tsize = 1024;
    while ((int64_t)tsize < nrows + 1024) tsize <<= 1;
    if (tsize > (1ull << 30)) return hv_fail(c, HV_ERR_CAPACITY, "too many cells for the neighbour set");
    unsigned long long* tab = nullptr;
    for (int attempt = 0;; ++attempt) {
        tab = iget<unsigned long long>(c, "n_tab", tsize);
        if (!tab) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbour set)");
        HVCK(hipMemsetAsync(tab, 0xff, tsize * sizeof(unsigned long long), c->stream));
        HVCK(hipMemsetAsync(full, 0, sizeof(int), c->stream));
        if (nv > 0) {
            k_pair_insert<D><<<nblk(nv, 256), 256, 0, c->stream>>>((const int*)mesh->ids.p, nv, (const int*)mesh->key.p,
                                                                   N, tab, tsize - 1, full);
            HVCK(hipGetLastError());
        }
        int hfull = 0;
        HVCK(hipMemcpyAsync(&hfull, full, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        if (!hfull) break;
        if (attempt >= 4 || tsize >= (1ull << 29)) return hv_fail(c, HV_ERR_CAPACITY, "neighbour set overflow");
        tsize <<= 2;
    }
    // compact the set, mirror, sort
    unsigned long long* u = iget<unsigned long long>(c, "n_u", tsize);
    if (!u) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbours)");
    size_t tb = 0;
    HVCK(hipcub::DeviceSelect::If(nullptr, tb, tab, u, cnt, (int)tsize, PairNotEmpty(), c->stream));
    void* tmp = c->pool->get("n_tmp", tb, false, c->stream);
    if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbours)");
    HVCK(hipcub::DeviceSelect::If(tmp, tb, tab, u, cnt, (int)tsize, PairNotEmpty(), c->stream));
    int nu = 0;
    HVCK(hipMemcpyAsync(&nu, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    uint64_t* rows = iget<uint64_t>(c, "n_rows", 2 * (size_t)nu);
    uint64_t* rows2 = iget<uint64_t>(c, "n_rows2", 2 * (size_t)nu);
    if (!rows || !rows2) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbours)");
    HVCK(hipMemsetAsync(cnt, 0, sizeof(int), c->stream));
    if (nu > 0) {
        k_pair_mirror<<<nblk(nu, 256), 256, 0, c->stream>>>(u, nu, N, rows, cnt);
        HVCK(hipGetLastError());
    }
    int nr = 0;
    HVCK(hipMemcpyAsync(&nr, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    tb = 0;
    HVCK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, rows, rows2, nr, 0, 64, c->stream));
    tmp = c->pool->get("n_tmp", tb, false, c->stream);
    if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (neighbours)");
    HVCK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, rows, rows2, nr, 0, 64, c->stream));
    k_row_ptr<<<nblk((int64_t)N + 1, 256), 256, 0, c->stream>>>(rows2, nr, N, ptr);
    HVCK(hipGetLastError());
    pkey = rows2;
    npair = nr;
    return HV_OK;
}

template <int D>
static int make_geo(hv_ctx* c, const hv_mesh* mesh, IntGeo& g) {
    // the context is bound to the mesh's own working set and boundary (hv_bind_mesh)
    if (!mesh || mesh->D != D || mesh->N != c->N || !c->xw || c->xw != mesh->xw)
        return hv_fail(c, HV_ERR_STATE, "the context is not bound to the mesh's generators");
    {
        const char* why = nullptr;
        const int rc = hv_domain(c, D, &why);
        if (rc) return hv_fail(c, rc, why);
        if (c->dom_unbounded) return hv_fail(c, HV_ERR_UNSUPPORTED, "the integrators need a bounded domain");
    }
    g.N = (int)c->N;
    g.nreal = (int)mesh->nreal;
    g.P = c->P;
    g.xo = (const double*)c->xw->p;
    double* pn = iget<double>(c, "i_pn", (size_t)std::max(1, c->P) * D);
    double* ph = iget<double>(c, "i_ph", (size_t)std::max(1, c->P));
    if (!g.xo || !pn || !ph) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integration)");
    std::vector<double> hph(std::max(1, c->P), 0.0);
    for (int p = 0; p < c->P; ++p)
        for (int k = 0; k < D; ++k) hph[p] += c->pnormal[p * D + k] * c->pbase[p * D + k];
    if (c->P) {
        HVCK(hipMemcpyAsync(pn, c->pnormal.data(), sizeof(double) * c->P * D, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(ph, hph.data(), sizeof(double) * c->P, hipMemcpyHostToDevice, c->stream));
    }
    g.pn = pn;
    g.ph = ph;
    return HV_OK;
}

template <int D>
static int finish(hv_ctx* c, int64_t ncell, const uint64_t* pkey, const double* area_d, const double* vol_d,
                  const int64_t* ptr_d, int64_t npair, float ms, hv_integral** out, int nf = 0,
                  const double* bulk_d = nullptr, const double* rint_d = nullptr) {
    // cells 0 .. ncell-1 (the originals) and their rows, which come first in the sorted order
    hv_integral* I = new hv_integral();
    I->N = ncell;
    I->vol.resize(ncell);
    I->ptr.resize(ncell + 1);
    if (ncell < c->N) {
        if (hipMemcpyAsync(&npair, ptr_d + ncell, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            delete I;
            return hv_fail(c, HV_ERR_DEVICE, "integration copy-out failed");
        }
    }
    I->area.resize(npair);
    I->nbr.resize(npair);
    std::vector<uint64_t> hk(npair);
    hipError_t e = hipSuccess;
    if (npair) {
        e = hipMemcpyAsync(hk.data(), pkey, sizeof(uint64_t) * npair, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(I->area.data(), area_d, sizeof(double) * npair, hipMemcpyDeviceToHost, c->stream);
    }
    if (nf > 0) {
        I->nf = nf;
        I->bulk.resize((size_t)ncell * nf);
        I->iface.resize((size_t)npair * nf);
        if (e == hipSuccess && ncell)
            e = hipMemcpyAsync(I->bulk.data(), bulk_d, sizeof(double) * I->bulk.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && npair)
            e = hipMemcpyAsync(I->iface.data(), rint_d, sizeof(double) * I->iface.size(), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(I->vol.data(), vol_d, sizeof(double) * ncell, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(I->ptr.data(), ptr_d, sizeof(int64_t) * (ncell + 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        delete I;
        return hv_fail(c, HV_ERR_DEVICE, std::string("integration copy-out: ") + hipGetErrorString(e));
    }
    for (int64_t t = 0; t < npair; ++t) I->nbr[t] = (int64_t)(hk[t] & 0xffffffffull) + 1;
    I->ms = ms;
    *out = I;
    return HV_OK;
}

template <int D>
int hv_integrate_polytope_impl(hv_ctx* c, const hv_mesh* mesh, hv_integral** out) {
    return hv_integrate_polytope_f_impl<D>(c, mesh, 0, nullptr, nullptr, out);
}

template <int D>
int hv_integrate_polytope_f_impl(hv_ctx* c, const hv_mesh* mesh, int nf, hv_integrand_fn fn, void* user,
                                 hv_integral** out) {
    if (nf < 0 || nf > HV_MAXF || (nf > 0 && !fn)) return hv_fail(c, HV_ERR_ARG, "integrand: 0 < nf <= 8 and a function");
    IntGeo g;
    int rc = make_geo<D>(c, mesh, g);
    if (rc) return rc;
    HVCK(hipEventRecord(c->ev[0], c->stream));
    uint64_t* pkey = nullptr;
    double* parea = nullptr;
    int64_t npair = 0;
    Integrand spec{nf, fn, user};
    double* rint = nullptr;
    rc = facet_rows<D>(c, mesh, g, pkey, parea, npair, nf > 0 ? &spec : nullptr, &rint);
    if (rc) return rc;
    double* vol = iget<double>(c, "i_vol", c->N);
    int64_t* ptr = iget<int64_t>(c, "i_ptr", c->N + 1);
    if (!vol || !ptr) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (volumes)");
    k_cell_volumes<D><<<nblk(c->N + 1, 256), 256, 0, c->stream>>>(g, pkey, parea, npair, vol, ptr);
    HVCK(hipGetLastError());
    double* bulk = nullptr;
    if (nf > 0) {
        // f at the generators (the cells' apexes), then the bulk integrals
        double* fx = iget<double>(c, "i_fx", (size_t)g.nreal * nf);
        bulk = iget<double>(c, "i_bulk", (size_t)g.nreal * nf);
        if (!fx || !bulk) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (integrand)");
        rc = eval_integrand(c, &spec, D, g.xo, g.nreal, fx);
        if (rc) return rc;
        k_cell_bulk<D><<<nblk(g.nreal, 256), 256, 0, c->stream>>>(g, pkey, ptr, vol, rint, fx, nf, bulk);
        HVCK(hipGetLastError());
    }
    HVCK(hipEventRecord(c->ev[1], c->stream));
    HVCK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HVCK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    return finish<D>(c, g.nreal, pkey, parea, vol, ptr, npair, ms, out, nf, bulk, rint);
}

template <int D>
int hv_integrate_mc_impl(hv_ctx* c, const hv_mesh* mesh, int64_t ndir, uint64_t seed, hv_integral** out) {
    if (ndir < 1 || ndir > (1ll << 40)) return hv_fail(c, HV_ERR_ARG, "ndir must be positive");
    IntGeo g;
    int rc = make_geo<D>(c, mesh, g);
    if (rc) return rc;
    HVCK(hipEventRecord(c->ev[0], c->stream));
    uint64_t* pkey = nullptr;
    int64_t npair = 0;
    int64_t* ptr = nullptr;
    rc = neighbor_rows<D>(c, mesh, pkey, npair, ptr);
    if (rc) return rc;
    double* vol = iget<double>(c, "i_vol", c->N);
    double* area = iget<double>(c, "i_marea", npair);
    if (!vol || !area) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (volumes)");
    std::vector<int64_t> hptr(c->N + 1);
    HVCK(hipMemcpyAsync(hptr.data(), ptr, sizeof(int64_t) * (c->N + 1), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    // one workgroup of 256 per cell, and a dispatch holds at most 2^32 work-items
    if ((int64_t)g.nreal > (1ll << 24) - 1)
        return hv_fail(c, HV_ERR_CAPACITY, "directional Monte Carlo: at most 2^24 - 1 cells per call");
    for (int64_t i = 0; i < g.nreal; ++i)
        if (hptr[i + 1] - hptr[i] > HV_MC_MAXK) return hv_fail(c, HV_ERR_CAPACITY, "a cell has more neighbours than the Monte-Carlo kernel holds");
    const double cvol = std::pow(M_PI, 0.5 * D) / std::tgamma(1.0 + 0.5 * D);
    HVCK(hipMemsetAsync(area, 0, sizeof(double) * std::max<int64_t>(npair, 1), c->stream));
    k_mc_cell<D><<<(unsigned)g.nreal, 256, 0, c->stream>>>(g, ptr, pkey, ndir, seed, cvol, D * cvol, vol, area);
    HVCK(hipGetLastError());
    k_mc_sym<D><<<nblk(npair, 256), 256, 0, c->stream>>>(g, ptr, pkey, npair, vol, area);
    HVCK(hipGetLastError());
    HVCK(hipEventRecord(c->ev[1], c->stream));
    HVCK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HVCK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    return finish<D>(c, g.nreal, pkey, area, vol, ptr, npair, ms, out);
}

// Point-sample counts -> volumes (sample fraction of the box) and interface areas.  A slab of
// width w beside an interface holds w * A - O(w^2) of volume (samples near the interface's
// own boundary see a third generator), so from the counts c1 (width eps) and c2 (width
// 2 eps): A = 2 A(eps) - A(2 eps) = V (2 c1 - c2 / 2) / (n eps), first order removed; a
// generator-generator interface averages the counts of its two sides.
__global__ void k_mcp_finish(int N, const int64_t* __restrict__ ptr, const uint64_t* __restrict__ pkey, int64_t npair,
                             const unsigned long long* __restrict__ ccnt, const unsigned long long* __restrict__ pcnt,
                             double n, double eps, double vbox, double* __restrict__ vol, double* __restrict__ area) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) vol[t] = vbox * (double)ccnt[t] / n;
    if (t >= npair) return;
    const int i = (int)(pkey[t] >> 32), k = (int)(pkey[t] & 0xffffffffull);
    double c1 = (double)pcnt[2 * t], c2 = (double)pcnt[2 * t + 1];
    if (k < N) {
        int64_t lo = ptr[k], hi = ptr[k + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int)(pkey[mid] & 0xffffffffull) < i) lo = mid + 1;
            else hi = mid;
        }
        if (lo < ptr[k + 1] && (int)(pkey[lo] & 0xffffffffull) == i) {
            c1 = 0.5 * (c1 + (double)pcnt[2 * lo]);
            c2 = 0.5 * (c2 + (double)pcnt[2 * lo + 1]);
        }
    }
    area[t] = fmax(0.0, vbox * (2.0 * c1 - 0.5 * c2) / (n * eps));
}

template <int D>
int hv_integrate_mc_points_impl(hv_ctx* c, const hv_mesh* mesh, int64_t nsamples, uint64_t seed, double slab,
                                hv_integral** out) {
    if (nsamples < 1 || nsamples > (1ll << 40)) return hv_fail(c, HV_ERR_ARG, "nsamples must be positive");
    if (!(slab > 0.0)) slab = 0.03;
    IntGeo g;
    int rc = make_geo<D>(c, mesh, g);
    if (rc) return rc;
    HVCK(hipEventRecord(c->ev[0], c->stream));
    uint64_t* pkey = nullptr;
    int64_t npair = 0;
    int64_t* ptr = nullptr;
    rc = neighbor_rows<D>(c, mesh, pkey, npair, ptr);
    if (rc) return rc;
    double* vol = iget<double>(c, "i_vol", c->N);
    double* area = iget<double>(c, "i_marea", std::max<int64_t>(npair, 1));
    unsigned long long* ccnt = iget<unsigned long long>(c, "i_ccnt", c->N);
    unsigned long long* pcnt = iget<unsigned long long>(c, "i_pcnt", 2 * std::max<int64_t>(npair, 1));
    if (!vol || !area || !ccnt || !pcnt) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (volumes)");
    HVCK(hipMemsetAsync(ccnt, 0, sizeof(unsigned long long) * c->N, c->stream));
    HVCK(hipMemsetAsync(pcnt, 0, sizeof(unsigned long long) * 2 * std::max<int64_t>(npair, 1), c->stream));
    int64_t n = 0;
    double eps = 0.0, vbox = 0.0;
    int64_t work = 0;
    double kms = 0.0;
    rc = hv_mc_points_impl<D>(c, nsamples, seed, slab, ptr, pkey, ccnt, pcnt, &n, &eps, &vbox, &work, &kms);
    if (rc) return rc;
    k_mcp_finish<<<nblk(std::max<int64_t>(c->N, npair), 256), 256, 0, c->stream>>>(
        g.nreal, ptr, pkey, npair, ccnt, pcnt, (double)n, eps, vbox, vol, area);
    HVCK(hipGetLastError());
    HVCK(hipEventRecord(c->ev[1], c->stream));
    HVCK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HVCK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    rc = finish<D>(c, g.nreal, pkey, area, vol, ptr, npair, ms, out);
    if (rc == HV_OK) {
        (*out)->work = work;
        (*out)->ms_kernel = kms;
    }
    return rc;
}

int hv_integral_work_impl(const hv_integral* I, int64_t* work, double* ms_kernel) {
    if (!I) return HV_ERR_ARG;
    if (work) *work = I->work;
    if (ms_kernel) *ms_kernel = I->ms_kernel;
    return HV_OK;
}

int hv_integral_counts_impl(const hv_integral* I, int64_t* ncells, int64_t* nnbr) {
    if (!I) return HV_ERR_ARG;
    if (ncells) *ncells = I->N;
    if (nnbr) *nnbr = (int64_t)I->nbr.size();
    return HV_OK;
}

int hv_integral_export_f_impl(const hv_integral* I, int32_t* nf, double* bulk, double* iface) {
    if (!I) return HV_ERR_ARG;
    if (nf) *nf = I->nf;
    if (bulk) std::copy(I->bulk.begin(), I->bulk.end(), bulk);
    if (iface) std::copy(I->iface.begin(), I->iface.end(), iface);
    return HV_OK;
}

int hv_integral_export_impl(const hv_integral* I, double* vol, int64_t* nbr_ptr, int64_t* nbr, double* area,
                            double* ms) {
    if (!I) return HV_ERR_ARG;
    if (vol) std::copy(I->vol.begin(), I->vol.end(), vol);
    if (nbr_ptr) std::copy(I->ptr.begin(), I->ptr.end(), nbr_ptr);
    if (nbr) std::copy(I->nbr.begin(), I->nbr.end(), nbr);
    if (area) std::copy(I->area.begin(), I->area.end(), area);
    if (ms) *ms = I->ms;
    return HV_OK;
}

void hv_integral_free_impl(hv_integral* I) { delete I; }

#define HV_IINST(D)                                                                                 \
    template int hv_integrate_polytope_impl<D>(hv_ctx*, const hv_mesh*, hv_integral**);          \
    template int hv_integrate_polytope_f_impl<D>(hv_ctx*, const hv_mesh*, int, hv_integrand_fn, void*, hv_integral**); \
    template int hv_integrate_mc_impl<D>(hv_ctx*, const hv_mesh*, int64_t, uint64_t, hv_integral**); \
    template int hv_integrate_mc_points_impl<D>(hv_ctx*, const hv_mesh*, int64_t, uint64_t, double, hv_integral**);
HV_IINST(2)
HV_IINST(3)
HV_IINST(4)
HV_IINST(5)
HV_IINST(6)
