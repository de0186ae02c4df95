// hv_device.h -- device-side data layout and helpers of the gfx950 raycast walk.
//
// HBM layout (all tables live for one hv_build_mesh call):
//   generators  xo[g*D + k]       FP64 records in original index order (gathers by key)
//               xs[j*(D+1) + k]   FP64 records in Morton order, slot D = |x_j|^2 (leaf
//                                 scans and neighbour gathers: one record per point)
//               sidx[j]           sorted position -> original index
//   tree        complete binary tree in heap order over leaves of L sorted points:
//               node 1 is the root, node P2+l is leaf l; boxes blo/bhi[k*NN + node]
//   vertices    key[v*(D+1) + j] sorted int32 (generator g or N+p for plane p),
//               x[v*D + k] FP64; ids are provisional: one slot per ray of a level
//   vertex hash int32 slots (-1 empty) holding vertex ids, probed linearly
//   edge hash   uint64 slots: ((v << 4) | drop) + 1 of the first vertex that
//               registered the edge, bit 62 set once its second end point is known
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HV_MAXD 6
#define HV_EDGE_DONE (1ull << 62)
// edge-hash slot: bits 0..34 (vertex << 4 | position) + 1, bits 35..61 a fingerprint of the
// edge (the top 27 bits of its 64-bit hash), bit 62 = both end points known
#define HV_EDGE_VMASK ((1ull << 35) - 1ull)
#define HV_EDGE_FMASK (((1ull << 27) - 1ull) << 35)
#define HV_EDGE_FP(hash64) (((hash64) >> 37) << 35)

struct DevGeo {
    // N: generator ids are below N and boundary plane p is N + p; n: the points of this tree
    // (sorted positions 0 .. n-1).  Equal for one point set; for a forest of meshes in one walk
    // each mesh's tree has its own n and its plane ids above every mesh's generators.
    int N, n, P, L, nleaf, P2, NN;
    const double* __restrict__ xo;   // N x D
    const double* __restrict__ xs;   // N x (D+1), slot D = |x|^2
    const int* __restrict__ sidx;
    const double* __restrict__ pn;   // P x D normals (row per plane)
    const double* __restrict__ ph;   // P offsets <n_p, b_p>
    const double* __restrict__ blo;  // D x NN
    const double* __restrict__ bhi;
    const float* __restrict__ cb;    // 8-ary tree: child boxes of internal node n at cb[(n*8+c)*2D ..], lo then hi
    int h;                           // 8-ary tree height (leaves at level h, 8^h padded leaves)
    const int* __restrict__ rank;    // original index -> sorted position
    const int* __restrict__ knn;     // HV_KNN nearest generators of each sorted position (sorted indices)
    double dlo[HV_MAXD], dhi[HV_MAXD];  // domain box (Morton frame)
    double xmax;                        // bound on |x| over the domain (rounding bounds)
    int unbounded;                      // no bounded domain: a ray may have no plane ahead
    // mixed-precision scan (leaves of 32): per sorted point a 32-slot BF16 K-vector relative
    // to the origin of its group of 8 leaves (16 dwords); per group the origin (slots
    // 0..D-1) and the radius of the group about it (slot 7)
    const uint32_t* __restrict__ pkv;
    const double* __restrict__ lorg;
    // FP32 scan (leaves of 32): per sorted point 8 floats, its coordinates relative to the same
    // group origin, then |x - o|^2 (slot D; +1e30 for padding records past N), zeros after
    const float* __restrict__ pxf;
    // per sorted point its coordinates in FP32, 8-float rows (zero padded): the ray start's
    // pre-ranking of its candidates
    const float* __restrict__ xf;
};

#define HV_KNN 16

// A walk over a forest of meshes (batched meshes of one dimension): generator ids of all meshes
// in one range, each mesh with its own tree and planes (its DevGeo), mesh of a generator by table.
struct GeoSet {
    const DevGeo* __restrict__ gs;   // per mesh
    const int* __restrict__ mof;     // generator id -> mesh
    int nmesh;
};
// The tree and planes a ray or vertex is worked in: the one tree (B false), or the tree of the
// mesh its generator `gen` belongs to (B true).
template <bool B>
__device__ __forceinline__ const DevGeo& hv_geo(const DevGeo& g, const GeoSet& F, int gen) {
    if constexpr (B) return F.gs[F.mof[gen]];
    else return g;
}
template <bool B>
__device__ __forceinline__ const DevGeo& hv_geo_mesh(const DevGeo& g, const GeoSet& F, int mesh) {
    if constexpr (B) return F.gs[mesh];
    else return g;
}

typedef unsigned int hv_u32x4 __attribute__((ext_vector_type(4)));
typedef float hv_f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 hv_bf16x8 __attribute__((ext_vector_type(8)));

// BF16 round-to-nearest-even of a finite float, as the 16-bit pattern
__device__ __forceinline__ uint32_t hv_bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ double hv_bf16_val(uint32_t h) { return (double)__uint_as_float(h << 16); }
// v ~ h + m with two BF16 terms: |v - h - m| <= (2^-16 + 2^-23) |v|
__device__ __forceinline__ void hv_split2(double v, uint32_t& h, uint32_t& m) {
    h = hv_bf16_rn((float)v);
    m = hv_bf16_rn((float)(v - hv_bf16_val(h)));
}
// one 16x16x32 BF16 tile product, FP32 accumulate from zero: lane l receives rows
// 4*(l/16)+e (e = 0..3) of column l%16
__device__ __forceinline__ hv_f32x4 hv_mfma_bf16(hv_u32x4 a, hv_u32x4 b) {
    const hv_f32x4 z = {0.f, 0.f, 0.f, 0.f};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(hv_bf16x8, a), __builtin_bit_cast(hv_bf16x8, b),
                                                   z, 0, 0, 0);
}
// K-slot layout of the mixed-precision membership form (x' = x - o, w = -2 (c - o)):
//   point: [0,D) xh | [D,2D) xh | [2D,3D) xm | 3D: sh | 3D+1: sm | 3D+2: 1 | 3D+3: 1
//   ray:   [0,D) wh | [D,2D) wm | [2D,3D) wh | 3D: 1  | 3D+1: 1  | 3D+2: ch | 3D+3: cm
// so the dot product is |x'|^2 - 2<x', c - o> + (|c - o|^2 - R^2 - margin) up to the split terms.
template <int D>
__device__ __forceinline__ hv_u32x4 hv_pack_slots(const uint32_t* sl, int q) {
    hv_u32x4 r;
    r[0] = sl[8 * q + 0] | (sl[8 * q + 1] << 16);
    r[1] = sl[8 * q + 2] | (sl[8 * q + 3] << 16);
    r[2] = sl[8 * q + 4] | (sl[8 * q + 5] << 16);
    r[3] = sl[8 * q + 6] | (sl[8 * q + 7] << 16);
    return r;
}

// Ray K-vector for one group of leaves: ball centre c, radius^2 R2, FP64 filter slack feps,
// group origin o and radius rad (every point of the group within rad of o).  The margin bounds the BF16 split
// error (2^-16 per split, three omitted cross terms per coordinate) and the FP32 accumulation
// of at most 22 exact products (21 roundings of 2^-23), so that acc < 0 for every generator
// with f(x) = |x - c|^2 - R2 <= feps.
template <int D>
__device__ __forceinline__ void hv_ray_kvec(const double* c, double R2, double feps, const double* o, double rad,
                                            hv_u32x4* row) {
    double beta[D], bb = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        beta[k] = c[k] - o[k];
        bb = fma(beta[k], beta[k], bb);
    }
    const double fcp = bb - R2;
    const double margin = 1.2e-4 * rad * sqrt(bb) + 2.5e-5 * (rad * rad + fabs(fcp)) + feps +
                          1e-13 * (bb + R2 + rad * rad);
    uint32_t sl[24];
#pragma unroll
    for (int s = 0; s < 24; ++s) sl[s] = 0u;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        uint32_t h, m;
        hv_split2(-2.0 * beta[k], h, m);
        sl[k] = h;
        sl[D + k] = m;
        sl[2 * D + k] = h;
    }
    sl[3 * D] = 0x3f80u;
    sl[3 * D + 1] = 0x3f80u;
    uint32_t ch, cm;
    hv_split2(fcp - margin, ch, cm);
    sl[3 * D + 2] = ch;
    sl[3 * D + 3] = cm;
    row[0] = hv_pack_slots<D>(sl, 0);
    row[1] = hv_pack_slots<D>(sl, 1);
    row[2] = hv_pack_slots<D>(sl, 2);
}

struct DevRays {
    int64_t* slot;   // edge-hash slot of the walked edge
    int* v;          // source vertex id
    int* drop;       // position in the source key of the generator walked away from
    double* u;       // D per ray
    double* c;       // half-space threshold (src/raycast.jl:802-805)
    double* du;      // orthogonality defect of u (src/edgeiteratebase.jl:64-74)
    double* tinit;   // unused
    uint64_t* code;  // Morton code of the initial certificate-ball centre (packet order)
    double* t1;      // best two candidates after the start phase (planes + nearest neighbours)
    double* t2;
    int* i1;
    int* i2;
};

// Wavefront-aggregated atomics: one atomic per wave instead of one per lane (counters
// and slot allocation hit a single address; per-lane atomics there serialise).  Both
// must be reached by every lane of the wave (no divergent early return before them).
// Workgroups are handed to the 8 XCDs round-robin (block b -> XCD b % 8), each XCD with its
// own L2.  Renumber so that XCD x works on one contiguous range of blocks: neighbouring
// packets (Morton order) then share an L2 instead of spreading every region over all eight.
__device__ __forceinline__ int hv_xcd_block_of(int nb) {
    const int b = blockIdx.x;
    const int x = b & 7, per = nb >> 3, extra = nb & 7;
    return x * per + min(x, extra) + (b >> 3);
}
__device__ __forceinline__ int hv_xcd_block() { return hv_xcd_block_of(gridDim.x); }
// The same for the first nb blocks of a larger grid (the count known only on the device):
// blocks 0 .. nb-1 still go round-robin to all eight XCDs; -1 for the blocks past nb.
__device__ __forceinline__ int hv_xcd_block_first(int nb) { return (int)blockIdx.x < nb ? hv_xcd_block_of(nb) : -1; }

__device__ __forceinline__ int hv_wave_alloc(int* counter, bool want) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    const unsigned long long below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    return base + __popcll(below);
}

// Lanes of a wavefront add 1 each to counters[key] with one atomic per distinct key (lanes of a
// wavefront mostly share a key: rays and vertices come grouped by mesh); returns each active
// lane's slot, the counter's old value plus its rank among the lanes of its key.  Every lane of
// the wavefront must reach the call.
__device__ __forceinline__ int hv_wave_add_by_key(int* counters, int key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(active);
    int slot = -1;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int k = __shfl(key, leader);
        const unsigned long long same = __ballot(active && key == k) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[k], __popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) slot = base + __popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return slot;
}

__device__ __forceinline__ void hv_wave_count(unsigned long long* ctr, bool cond) {
    const unsigned long long m = __ballot(cond);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    if (lane == __ffsll((long long)m) - 1) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

__host__ __device__ __forceinline__ uint64_t hv_mix64(uint64_t z) {
    z ^= z >> 33; z *= 0xff51afd7ed558ccdULL;
    z ^= z >> 33; z *= 0xc4ceb9fe1a85ec53ULL;
    z ^= z >> 33;
    return z;
}

template <int K>
__device__ __forceinline__ uint64_t hv_hash_ints(const int* k) {
    uint64_t h = 0x9E3779B97F4A7C15ULL;
#pragma unroll
    for (int i = 0; i < K; ++i) h = hv_mix64(h ^ (uint64_t)(uint32_t)k[i]);
    return h;
}

// edge of vertex key `key` (D+1 entries) without position `pos`
// (static indices + selects: a runtime index into a register array would go to scratch)
template <int D>
__device__ __forceinline__ void hv_edge_of(const int* key, int pos, int* e) {
#pragma unroll
    for (int j = 0; j < D; ++j) e[j] = (j < pos) ? key[j] : key[j + 1];
}

template <int D>
__device__ __forceinline__ int hv_key_at(const int* key, int pos) {
    int s = key[0];
#pragma unroll
    for (int j = 1; j <= D; ++j) s = (j == pos) ? key[j] : s;
    return s;
}

// Row j (0..D-1) of the circumcentre system of a sorted key, relative to x0 = x_{key[0]}:
//   generator e:  2 <x_e - x0, y> = |x_e - x0|^2        plane p:  <n_p, y> = h_p - <n_p, x0>
template <int D>
__device__ __forceinline__ void hv_vertex_row(const DevGeo& g, int e, const double* x0, double* row, double& rhs) {
    const int N = g.N;
    if (e < N) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double dk = g.xo[(int64_t)(e) * D + k] - x0[k];
            row[k] = 2.0 * dk;
            s = fma(dk, dk, s);
        }
        rhs = s;
    } else {
        const int p = e - N;
        double s = g.ph[p];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double nk = g.pn[p * D + k];
            row[k] = nk;
            s = fma(-nk, x0[k], s);
        }
        rhs = s;
    }
}

// Circumcentre of a sorted key (D+1 elements; generators first, planes N+p after).
// The constraint rows a_j are orthonormalised (modified Gram-Schmidt, two passes) while
// the right-hand side follows the same transformation, so y = sum_j beta_j q_j is the
// point of the affine set {a_j . y = b_j} -- the projection the reference applies in
// prepare_vertex_calculation / vertex_calculation_hp (src/raycast.jl:633-707).  One step
// of iterative refinement uses the rows rebuilt from memory.  False if singular.
template <int D>
__device__ __forceinline__ bool hv_vertex_coords(const DevGeo& g, const int* key, double* out) {
    const int N = g.N;
    double x0[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x0[k] = g.xo[(int64_t)(key[0]) * D + k];
    double Q[D][D], Rl[D][D], inv[D], beta[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double rhs;
        hv_vertex_row<D>(g, key[j + 1], x0, Q[j], rhs);
#pragma unroll
        for (int i = 0; i < D; ++i) Rl[j][i] = 0.0;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int i = 0; i < j; ++i) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) a = fma(Q[j][k], Q[i][k], a);
#pragma unroll
                for (int k = 0; k < D; ++k) Q[j][k] = fma(-a, Q[i][k], Q[j][k]);
                rhs = fma(-a, beta[i], rhs);
                Rl[j][i] += a;
            }
        }
        double nn = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nn = fma(Q[j][k], Q[j][k], nn);
        if (!(nn > 0.0)) return false;
        const double s = 1.0 / sqrt(nn);
        inv[j] = s;
#pragma unroll
        for (int k = 0; k < D; ++k) Q[j][k] *= s;
        beta[j] = rhs * s;
    }
    double y[D];
#pragma unroll
    for (int k = 0; k < D; ++k) y[k] = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int k = 0; k < D; ++k) y[k] = fma(beta[j], Q[j][k], y[k]);
    // refinement: residual r_j = b_j - a_j . y, transformed like b
    double bet2[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double row[D], rhs;
        hv_vertex_row<D>(g, key[j + 1], x0, row, rhs);
#pragma unroll
        for (int k = 0; k < D; ++k) rhs = fma(-row[k], y[k], rhs);
#pragma unroll
        for (int i = 0; i < j; ++i) rhs = fma(-Rl[j][i], bet2[i], rhs);
        bet2[j] = rhs * inv[j];
    }
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int k = 0; k < D; ++k) y[k] = fma(bet2[j], Q[j][k], y[k]);
#pragma unroll
    for (int k = 0; k < D; ++k) out[k] = x0[k] + y[k];
    return true;
}

// Morton code of a point (bits interleaved, 60/D bits per coordinate) in the domain box.
template <int D>
__device__ __forceinline__ uint64_t hv_morton(const DevGeo& g, const double* p) {
    constexpr int B = 60 / D;
    constexpr double S = (double)((1ull << B) - 1);
    uint32_t q[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double w = g.dhi[k] - g.dlo[k];
        double f = (p[k] - g.dlo[k]) / (w > 0 ? w : 1.0);
        f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        q[k] = (uint32_t)(f * S);
    }
    uint64_t code = 0;
#pragma unroll
    for (int b = B - 1; b >= 0; --b) {
#pragma unroll
        for (int k = 0; k < D; ++k) code = (code << 1) | ((q[k] >> b) & 1u);
    }
    return code;
}
