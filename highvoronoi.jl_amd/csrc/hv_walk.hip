// hv_walk.hip -- gfx950 kernels and host driver of the device-resident raycast walk.
//
// What the reference does one ray at a time on the CPU (src/sysvoronoi.jl:384-525,
// src/raycast.jl:125-164,794-970), this file does as a breadth-first walk over the
// whole vertex graph at once:
//   level k:  register the d+1 edges of every vertex found in level k-1 in a device
//             edge hash; an edge is walked only when exactly one end point is known
//             (the cell1/cell2 rule of pushedge!, src/edgehashing.jl:66-115);
//             emit one ray per such edge (direction of u_qr, src/tools.jl:773-790);
//             sort rays by the Morton code of their expected end point;
//             raycast in packets of 64 rays per wavefront over a kd-style tree of
//             Morton-ordered generator leaves: argmin of the equidistance parameter t
//             with the reference's winner rule (src/raycast.jl:877-916);
//             insert the new keys into a device vertex hash; the winners are level k.
// Seeds are the 2^d corners of the box: each is the vertex {nearest generator} u
// {the d planes through it}.
#include "hv_device.h"
#include "hv_internal.h"
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#define HVCK(call)                                                                 \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return hv_fail(c, HV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ray outcome codes of the scan
#define HV_ST_NONE 0
#define HV_ST_FOUND 1
#define HV_ST_TIE 2
#define HV_ST_SINGULAR -2
#define HV_ST_DROPPED -1   // a seed candidate that was not kept (counted nowhere)

// ---------------------------------------------------------------------------------
// points and tree

// Over a forest: the mesh of item t, for items numbered mesh after mesh (off[m] .. off[m+1])
__device__ __forceinline__ int hv_mesh_of_item(const int* __restrict__ off, int nmesh, int t) {
    int lo = 0, hi = nmesh;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// doubles as unsigned keys in the same order (min / max by integer atomics)
__host__ __device__ inline unsigned long long hv_double_key(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double hv_key_double(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, b);
}

// bounding box of n points (n x D): key[k] min, key[D + k] max; per workgroup in LDS, then one
// atomic per coordinate and workgroup
template <int D>
__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ x, int n, unsigned long long* __restrict__ key) {
    __shared__ unsigned long long smin[D], smax[D];
    if (threadIdx.x < D) { smin[threadIdx.x] = ~0ull; smax[threadIdx.x] = 0ull; }
    __syncthreads();
    double lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double v = x[(int64_t)i * D + k];
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
#pragma unroll
    for (int k = 0; k < D; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], off));
        }
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < D; ++k) {
            atomicMin(&smin[k], hv_double_key(lo[k]));
            atomicMax(&smax[k], hv_double_key(hi[k]));
        }
    __syncthreads();
    if (threadIdx.x < D) {
        atomicMin(&key[threadIdx.x], smin[threadIdx.x]);
        atomicMax(&key[D + threadIdx.x], smax[threadIdx.x]);
    }
}

template <int D>
__global__ void k_points_in(const double* __restrict__ xin, int N, DevGeo g, double* __restrict__ xo,
                            uint64_t* __restrict__ code, int* __restrict__ idx, int* __restrict__ bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double p[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        p[k] = xin[(int64_t)i * D + k];
        xo[(int64_t)i * D + k] = p[k];
    }
    // check_boundary (src/boundary.jl:427-441): <b_p - x, n_p> >= 0 for every plane (bit 0 of
    // *bad when not).  Bit 1: a generator on a plane (within 1e-12 relative), which coincides
    // with its own mirror image: the walk of its cell meets vertices with more than d + 1
    // generators on their sphere (the reference's non-general path), reported as degenerate.
    bool ok = true, on = false;
    for (int q = 0; q < g.P; ++q) {
        double s = g.ph[q];
#pragma unroll
        for (int k = 0; k < D; ++k) s -= g.pn[q * D + k] * p[k];
        ok = ok && (s >= 0.0);
        on = on || (s <= 1e-12 * (1.0 + fabs(g.ph[q])));
    }
#pragma unroll
    for (int k = 0; k < D; ++k) ok = ok && isfinite(p[k]);
    if (!ok) atomicOr(bad, 1);
    else if (on) atomicOr(bad, 2);
    code[i] = hv_morton<D>(g, p);
    idx[i] = i;
}

// xo and rank: this tree's rows (local index); sidx receives global ids (goff + local)
template <int D>
__global__ void k_points_sorted(const double* __restrict__ xo, const int* __restrict__ idx, int N, int goff,
                                double* __restrict__ xs, int* __restrict__ sidx, int* __restrict__ rank) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    int gi = idx[j];
    sidx[j] = goff + gi;
    rank[gi] = j;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double v = xo[(int64_t)gi * D + k];
        xs[(int64_t)j * (D + 1) + k] = v;
        s = fma(v, v, s);
    }
    xs[(int64_t)j * (D + 1) + D] = s;
}

template <int D>
__global__ void k_leaf_boxes(DevGeo g, double* __restrict__ blo, double* __restrict__ bhi) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= g.P2) return;
    int node = g.P2 + l;
    int j0 = l * g.L, j1 = min(j0 + g.L, g.n);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double lo = 1e300, hi = -1e300;
        for (int j = j0; j < j1; ++j) {
            double v = g.xs[(int64_t)(j) * (D + 1) + k];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        blo[(int64_t)k * g.NN + node] = lo;
        bhi[(int64_t)k * g.NN + node] = hi;
    }
}

template <int D>
__global__ void k_node_boxes(int first, int count, int NN, double* __restrict__ blo, double* __restrict__ bhi) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int node = first + t;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        blo[(int64_t)k * NN + node] = fmin(blo[(int64_t)k * NN + 2 * node], blo[(int64_t)k * NN + 2 * node + 1]);
        bhi[(int64_t)k * NN + node] = fmax(bhi[(int64_t)k * NN + 2 * node], bhi[(int64_t)k * NN + 2 * node + 1]);
    }
}

// 8-ary tree over the same leaves (level h, 8^h slots, node (l,i) has children (l+1, 8i+c)).
// Internal node (l,i) has global id (8^l - 1)/7 + i and stores the boxes of its 8 children
// contiguously, in FP32 rounded outward, so one load per visited node tests 8 children.
// An empty child is the point box lo = hi = 1e30: its distance^2 overflows to +inf for every
// query, with the clamp form p - med3(p, lo, hi) as with max(lo - p, p - hi, 0).
template <int D>
__global__ void k_tree8_level(DevGeo g, int l, int nodes_at_l, float* __restrict__ cb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes_at_l * 8) return;
    const int i = t >> 3, c = t & 7;
    int off = 0, pw = 1;
    for (int q = 0; q < l; ++q) { off += pw; pw *= 8; }
    const int gid = off + i;
    float lo[D], hi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { lo[k] = 1e30f; hi[k] = -1e30f; }
    const int child = 8 * i + c;
    if (l + 1 == g.h) {
        const int j0 = child * g.L, j1 = min(j0 + g.L, g.n);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double mn = 1e300, mx = -1e300;
            for (int j = j0; j < j1; ++j) {
                const double v = g.xs[(int64_t)(j) * (D + 1) + k];
                mn = fmin(mn, v);
                mx = fmax(mx, v);
            }
            if (j1 > j0) { lo[k] = __double2float_rd(mn); hi[k] = __double2float_ru(mx); }
        }
    } else {
        const int cg = off + pw + child;   // global id of the child node at level l+1
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float* b = cb + ((int64_t)cg * 8 + q) * (2 * D);
            if (!(b[0] < 1e30f)) continue;   // empty grandchild
#pragma unroll
            for (int k = 0; k < D; ++k) { lo[k] = fminf(lo[k], b[k]); hi[k] = fmaxf(hi[k], b[D + k]); }
        }
    }
    const bool empty = !(lo[0] <= hi[0]);
    float* o = cb + ((int64_t)gid * 8 + c) * (2 * D);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        o[k] = empty ? 1e30f : lo[k];
        o[D + k] = empty ? 1e30f : hi[k];
    }
}

// Mixed-precision scan tables (leaves of 32 points): the origin of each group of 8 leaves
// (the children of one last-level tree node; centre of the group's bounding box) with the
// radius of the group about it, then one BF16 K-vector per sorted point relative to its
// group origin.  Records past N are padding whose form is
// +1e30 for every ray.
// One wavefront per group of 8 leaves (256 points), the lanes over its points (min, max and
// the largest distance do not depend on the order they are combined in).
template <int D, bool B>
__global__ __launch_bounds__(64) void k_leaf_origin(DevGeo g0, GeoSet F, const int* __restrict__ off,
                                                    double* __restrict__ lorg0) {
    int leaf = blockIdx.x;
    const int lane = threadIdx.x;
    int mk = 0;
    if constexpr (B) {
        if (leaf >= off[F.nmesh]) return;
        mk = hv_mesh_of_item(off, F.nmesh, leaf);
        leaf -= off[mk];
    }
    const DevGeo& g = hv_geo_mesh<B>(g0, F, mk);
    double* lorg = B ? const_cast<double*>(g.lorg) : lorg0;
    if (leaf >= (g.nleaf + 7) / 8) return;
    const int j0 = leaf * 256, j1 = min(j0 + 256, g.n);
    double o[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double mn = 1e300, mx = -1e300;
        for (int j = j0 + lane; j < j1; j += 64) {
            const double v = g.xs[(int64_t)j * (D + 1) + k];
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
#pragma unroll
        for (int s2 = 32; s2 > 0; s2 >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, s2));
            mx = fmax(mx, __shfl_xor(mx, s2));
        }
        o[k] = 0.5 * (mn + mx);
    }
    double r2 = 0.0;
    for (int j = j0 + lane; j < j1; j += 64) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double t = g.xs[(int64_t)j * (D + 1) + k] - o[k];
            s = fma(t, t, s);
        }
        r2 = fmax(r2, s);
    }
#pragma unroll
    for (int s2 = 32; s2 > 0; s2 >>= 1) r2 = fmax(r2, __shfl_xor(r2, s2));
    if (lane < 8) {   // the origin in slots 0..D-1, the radius in slot 7 (D <= 6)
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) v = lane == k ? o[k] : v;
        if (lane == 7) v = sqrt(r2) * (1.0 + 1e-12) + 1e-300;
        lorg[(int64_t)leaf * 8 + lane] = v;
    }
}

template <int D, bool B>
__global__ void k_point_kvec(DevGeo g0, GeoSet F, const int* __restrict__ off, const double* __restrict__ lorg0,
                             hv_u32x4* __restrict__ pkv0) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int mk = 0;
    if constexpr (B) {
        if (j >= off[F.nmesh]) return;
        mk = hv_mesh_of_item(off, F.nmesh, j);
        j -= off[mk];
    }
    const DevGeo& g = hv_geo_mesh<B>(g0, F, mk);
    const double* lorg = B ? g.lorg : lorg0;
    hv_u32x4* pkv = B ? (hv_u32x4*)const_cast<uint32_t*>(g.pkv) : pkv0;
    if (j >= g.nleaf * 32) return;
    uint32_t sl[32];
#pragma unroll
    for (int s = 0; s < 32; ++s) sl[s] = 0u;
    if (j < g.n) {
        const int leaf = j >> 8;
        double xp[D], s2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            xp[k] = g.xs[(int64_t)j * (D + 1) + k] - lorg[(int64_t)leaf * 8 + k];
            s2 = fma(xp[k], xp[k], s2);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            uint32_t h, m;
            hv_split2(xp[k], h, m);
            sl[k] = h;
            sl[D + k] = h;
            sl[2 * D + k] = m;
        }
        uint32_t sh, sm;
        hv_split2(s2, sh, sm);
        sl[3 * D] = sh;
        sl[3 * D + 1] = sm;
        sl[3 * D + 2] = 0x3f80u;
        sl[3 * D + 3] = 0x3f80u;
    } else {
        sl[3 * D] = hv_bf16_rn(1e30f);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) pkv[(int64_t)j * 4 + q] = hv_pack_slots<D>(sl, q);
}

template <int D>
__global__ void k_point_rows32(DevGeo g, float* __restrict__ xf) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.n) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) xf[(int64_t)j * 8 + k] = k < D ? (float)g.xs[(int64_t)j * (D + 1) + k] : 0.f;
}

template <int D, bool B>
__global__ void k_point_f32(DevGeo g0, GeoSet F, const int* __restrict__ off, const double* __restrict__ lorg0,
                            float* __restrict__ pxf0) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int mk = 0;
    if constexpr (B) {
        if (j >= off[F.nmesh]) return;
        mk = hv_mesh_of_item(off, F.nmesh, j);
        j -= off[mk];
    }
    const DevGeo& g = hv_geo_mesh<B>(g0, F, mk);
    const double* lorg = B ? g.lorg : lorg0;
    float* pxf = B ? const_cast<float*>(g.pxf) : pxf0;
    if (j >= g.nleaf * 32) return;
    float row[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) row[k] = 0.f;
    if (j < g.n) {
        const int grp = j >> 8;
        double s2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double xp = g.xs[(int64_t)j * (D + 1) + k] - lorg[(int64_t)grp * 8 + k];
            row[k] = (float)xp;
            s2 = fma(xp, xp, s2);
        }
        row[D] = (float)s2;
    } else {
        row[D] = 1e30f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) pxf[(int64_t)j * 8 + k] = row[k];
}

// ---------------------------------------------------------------------------------
// seeds: one block per corner of the domain, brute-force nearest generator (ties -> smaller
// index): the 2^D box corners, or (cx, cpl: point and ascending plane indices per corner) the
// vertices of a general polytope where exactly D planes meet.  Over a forest (B), block b is
// box corner b % 2^D of mesh b / 2^D.

template <int D, bool B>
__global__ __launch_bounds__(256) void k_seed_corners(DevGeo g0, GeoSet F, int* __restrict__ vkey, double* __restrict__ vx,
                                                      int64_t base, int* __restrict__ status,
                                                      const double* __restrict__ cx = nullptr,
                                                      const int* __restrict__ cpl = nullptr) {
    __shared__ double sd[256];
    __shared__ int si[256];
    const int corner = B ? (int)(blockIdx.x & ((1u << D) - 1u)) : (int)blockIdx.x;
    const DevGeo& g = B ? F.gs[blockIdx.x >> D] : g0;
    double cp[D];
    int planes[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (cx) {
            cp[k] = cx[(int64_t)corner * D + k];
            planes[k] = g.N + cpl[(int64_t)corner * D + k];
        } else {
            bool hi = (corner >> k) & 1;
            cp[k] = hi ? g.dhi[k] : g.dlo[k];
            planes[k] = g.N + 2 * k + (hi ? 0 : 1);   // plane 2k: x_k = hi, 2k+1: x_k = lo
        }
    }
    double best = 1e300;
    int bi = 0x7fffffff;
    for (int j = threadIdx.x; j < g.n; j += blockDim.x) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double t = g.xs[(int64_t)(j) * (D + 1) + k] - cp[k];
            s = fma(t, t, s);
        }
        const int id = g.sidx[j];
        if (s < best || (s == best && id < bi)) { best = s; bi = id; }
    }
    sd[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double o = sd[threadIdx.x + w];
            int oi = si[threadIdx.x + w];
            if (o < sd[threadIdx.x] || (o == sd[threadIdx.x] && oi < si[threadIdx.x])) {
                sd[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int64_t id = base + blockIdx.x;
        vkey[id * (D + 1)] = si[0];
        // planes are already ascending in k
#pragma unroll
        for (int k = 0; k < D; ++k) vkey[id * (D + 1) + 1 + k] = planes[k];
#pragma unroll
        for (int k = 0; k < D; ++k) vx[id * D + k] = cp[k];
        status[blockIdx.x] = 1;
    }
}

// ---------------------------------------------------------------------------------
// vertex hash insert: winners of a level (status > 0) become live vertices

template <int D>
__device__ __forceinline__ bool keys_equal(const int* __restrict__ a, const int* __restrict__ b) {
    bool eq = true;
#pragma unroll
    for (int j = 0; j <= D; ++j) eq = eq && (a[j] == b[j]);
    return eq;
}

// Adds the workgroup's counts of four conditions to counters 0..3 (counter k at
// ctr[k * stride + the workgroup's copy]: with stride > 1 the workgroups spread over that many
// copies): partial sums in LDS, then at most one global atomic per counter and workgroup
// (same-address atomics serialise).  Every thread of the workgroup calls it.
__device__ __forceinline__ void hv_block_count4(unsigned long long* ctr, bool c0, bool c1, bool c2, bool c3,
                                                int stride = 1) {
    __shared__ unsigned int part[4];
    if (threadIdx.x < 4) part[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned n0 = __popcll(__ballot(c0)), n1 = __popcll(__ballot(c1));
    const unsigned n2 = __popcll(__ballot(c2)), n3 = __popcll(__ballot(c3));
    if ((threadIdx.x & 63) == 0) {
        if (n0) atomicAdd(&part[0], n0);
        if (n1) atomicAdd(&part[1], n1);
        if (n2) atomicAdd(&part[2], n2);
        if (n3) atomicAdd(&part[3], n3);
    }
    __syncthreads();
    if (threadIdx.x < 4 && part[threadIdx.x])
        atomicAdd(&ctr[threadIdx.x * stride + (stride > 1 ? blockIdx.x % (unsigned)stride : 0u)],
                  (unsigned long long)part[threadIdx.x]);
}

template <int D>
__global__ __launch_bounds__(256) void k_insert(const int* __restrict__ status, int n, const int* __restrict__ n_dev,
                                                const int* __restrict__ vkey, int64_t base, int* __restrict__ htab,
                                                uint64_t hmask, uint8_t* __restrict__ win,
                                                unsigned long long* __restrict__ ctr, int cstride) {
    // n_dev (when given): the candidate count in device memory; flags up to n are written
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = n_dev ? min(n, *n_dev) : n;
    bool dup = false, full = false, none = false, singular = false;
    if (t >= nv && t < n) win[t] = 0;
    if (t < nv) {
        uint8_t won = 0;
        none = status[t] == HV_ST_NONE;
        singular = status[t] == HV_ST_SINGULAR;
        if (status[t] > 0) {
            const int id = (int)(base + t);
            const int* key = vkey + (int64_t)id * (D + 1);
            uint64_t h = hv_hash_ints<D + 1>(key) & hmask;
            full = true;
            for (uint64_t probe = 0; probe <= hmask; ++probe) {
                int cur = htab[h];
                if (cur < 0) {
                    const int old = atomicCAS(&htab[h], -1, id);
                    if (old == -1) { won = 1; full = false; break; }
                    cur = old;
                }
                if (keys_equal<D>(key, vkey + (int64_t)cur * (D + 1))) { dup = true; full = false; break; }
                h = (h + 1) & hmask;
            }
        }
        win[t] = won;
    }
    hv_block_count4(ctr, dup, full, none, singular, cstride);
}

template <int D>
__global__ void k_rehash_vertices(const int* __restrict__ ids, int64_t n, const int* __restrict__ vkey,
                                  int* __restrict__ htab, uint64_t hmask) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int id = ids[t];
    uint64_t h = hv_hash_ints<D + 1>(vkey + (int64_t)id * (D + 1)) & hmask;
    for (;;) {
        if (atomicCAS(&htab[h], -1, id) == -1) return;
        h = (h + 1) & hmask;
    }
}

// ---------------------------------------------------------------------------------
// edges: registration (pushedge! semantics) and ray emission

// Edge bookkeeping, per level and in two small tables instead of one table of every edge
// (which would be a random access into ~1 GB per lookup):
//   ray-edge table:  the edges of this level's rays (value: ray index + 1);
//   new-edge table:  the edges of this level's new vertices (value: (vertex << 4 | pos) + 1).
// An edge of a new vertex is either the edge of a ray of this level (its other end point is
// that ray's source: closed, and the ray's entry is marked), or shared with another new
// vertex of the level (closed), or open: then its owner emits a ray in the next level.  An
// open edge of an older vertex always has a ray (scanned, waiting or carried) in the level
// that finds its other end point, so nothing older needs to be kept.  Slots: bits 0..34 the
// value, bits 35..61 a fingerprint of the edge, bit 62 closed.
template <int D>
__global__ __launch_bounds__(256) void k_rt_build(DevRays R, int nr, const int* __restrict__ vkey,
                                                  unsigned long long* __restrict__ rt, uint64_t rmask,
                                                  unsigned long long* __restrict__ ctr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool full = false;
    if (t < nr) {
        int e[D];
        hv_edge_of<D>(vkey + (int64_t)R.v[t] * (D + 1), R.drop[t], e);
        const uint64_t hfull = hv_hash_ints<D>(e);
        const unsigned long long fp = HV_EDGE_FP(hfull);
        const unsigned long long mine = fp | ((unsigned long long)t + 1ull);
        uint64_t h = hfull & rmask;
        int64_t out = -1;
        for (uint64_t probe = 0; probe <= rmask; ++probe) {
            unsigned long long cur = rt[h];
            if (cur == 0ull) {
                const unsigned long long old = atomicCAS(&rt[h], 0ull, mine);
                if (old == 0ull) { out = (int64_t)h; break; }
                cur = old;
            }
            if ((cur & HV_EDGE_FMASK) == fp) {
                const int r2 = (int)((cur & HV_EDGE_VMASK) - 1ull);
                int e2[D];
                hv_edge_of<D>(vkey + (int64_t)R.v[r2] * (D + 1), R.drop[r2], e2);
                bool eq = true;
#pragma unroll
                for (int j = 0; j < D; ++j) eq = eq && (e[j] == e2[j]);
                if (eq) { out = (int64_t)h; break; }
            }
            h = (h + 1) & rmask;
        }
        full = out < 0;
        R.slot[t] = out < 0 ? 0 : out;
    }
    hv_wave_count(&ctr[1], full);
}

// Walked edges: those with a generator below nwalk (all edges when nwalk is the generator
// count), or, with a cell mask over the ngen generators, those holding a marked generator.
template <int D>
__device__ __forceinline__ bool hv_edge_walked(const int* e, int nwalk, const uint8_t* __restrict__ cmask, int ngen) {
    if (!cmask) return e[0] < nwalk;
    bool w = false;
#pragma unroll
    for (int j = 0; j < D; ++j) w = w || (e[j] < ngen && cmask[e[j]] != 0);
    return w;
}

template <int D>
__global__ __launch_bounds__(256) void k_register(const int* __restrict__ frontier, int nf,
                                                  const int* __restrict__ nf_dev, const int* __restrict__ vkey,
                                                  int N, const uint8_t* __restrict__ cmask, int ngen, DevRays R,
                                                  unsigned long long* __restrict__ rt, uint64_t rmask,
                                                  unsigned long long* __restrict__ nt, uint64_t nmask,
                                                  int64_t* __restrict__ reg, unsigned long long* __restrict__ ctr) {
    // nf_dev (when given): the vertex count in device memory, nf then only bounds the grid
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = nf_dev ? min(nf, *nf_dev) : nf;
    bool full = false;
    if (t < nv * (D + 1)) {
        const int fi = t / (D + 1), pos = t % (D + 1);
        const int v = frontier[fi];
        const int* key = vkey + (int64_t)v * (D + 1);
        int e[D];
        hv_edge_of<D>(key, pos, e);
        int64_t out = -1;   // an edge of planes only is a corner point, not a line
        if (hv_edge_walked<D>(e, N, cmask, ngen)) {
            const uint64_t hfull = hv_hash_ints<D>(e);
            const unsigned long long fp = HV_EDGE_FP(hfull);
            // the edge of a ray of this level: its other end point is the ray's source
            bool ray_edge = false;
            {
                uint64_t h = hfull & rmask;
                for (uint64_t probe = 0; probe <= rmask; ++probe) {
                    const unsigned long long cur = rt[h];
                    if (cur == 0ull) break;
                    if ((cur & HV_EDGE_FMASK) == fp) {
                        const int r2 = (int)((cur & HV_EDGE_VMASK) - 1ull);
                        int e2[D];
                        hv_edge_of<D>(vkey + (int64_t)R.v[r2] * (D + 1), R.drop[r2], e2);
                        bool eq = true;
#pragma unroll
                        for (int j = 0; j < D; ++j) eq = eq && (e[j] == e2[j]);
                        if (eq) {
                            atomicOr(&rt[h], HV_EDGE_DONE);
                            ray_edge = true;
                            break;
                        }
                    }
                    h = (h + 1) & rmask;
                }
            }
            if (!ray_edge) {
                const unsigned long long mine = fp | (((((unsigned long long)v) << 4) | (unsigned)pos) + 1ull);
                uint64_t h = hfull & nmask;
                out = -2;
                for (uint64_t probe = 0; probe <= nmask; ++probe) {
                    unsigned long long cur = nt[h];
                    if (cur == 0ull) {
                        const unsigned long long old = atomicCAS(&nt[h], 0ull, mine);
                        if (old == 0ull) { out = (int64_t)h; break; }
                        cur = old;
                    }
                    // another edge's slot unless the fingerprints agree (then compare exactly)
                    if ((cur & HV_EDGE_FMASK) != fp) { h = (h + 1) & nmask; continue; }
                    const unsigned long long raw = (cur & HV_EDGE_VMASK) - 1ull;
                    int e2[D];
                    hv_edge_of<D>(vkey + (int64_t)(raw >> 4) * (D + 1), (int)(raw & 15ull), e2);
                    bool eq = true;
#pragma unroll
                    for (int j = 0; j < D; ++j) eq = eq && (e[j] == e2[j]);
                    if (eq) {
                        atomicOr(&nt[h], HV_EDGE_DONE);
                        out = (int64_t)h;
                        break;
                    }
                    h = (h + 1) & nmask;
                }
                full = (out == -2);
            }
        }
        reg[t] = out;
    }
    hv_wave_count(&ctr[1], full);
}

// Direction of the walk along the edge key\{key[pos]} (u_qr, src/tools.jl:773-790):
// u = -normalize(P_perp w), P_perp the projection onto the orthogonal complement of the
// edge's affine hull; w = x_s - x0 for a dropped generator s, n_s for a dropped plane
// (its mirror image of x0 lies on the +n_s side).  Planes enter the span via their
// normals.  du is the orthogonality defect of src/edgeiteratebase.jl:64-74 and c the
// half-space threshold of src/raycast.jl:802-805.
template <int D>
__device__ __forceinline__ void hv_span_vec(const DevGeo& g, int el, const double* x0, double* out) {
    if (el < g.N) {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = g.xo[(int64_t)(el) * D + k] - x0[k];
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = g.pn[(el - g.N) * D + k];
    }
}

template <int D>
__device__ __forceinline__ void hv_ray_setup(const DevGeo& g, const int* key, int pos, const double* r, double plane_tol,
                             double* u, double& c, double& du, double& rho) {
    const int N = g.N;
    int e[D];
    hv_edge_of<D>(key, pos, e);
    const int s = hv_key_at<D>(key, pos);
    double x0[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x0[k] = g.xo[(int64_t)(e[0]) * D + k];
    double q[D - 1][D];
    // slots from the top so that plane normals (largest indices) are orthogonalised first
#pragma unroll
    for (int a = 0; a < D - 1; ++a) {
        hv_span_vec<D>(g, e[D - 1 - a], x0, q[a]);
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
#pragma unroll
            for (int b = 0; b < a; ++b) {
                double dp = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) dp = fma(q[a][k], q[b][k], dp);
#pragma unroll
                for (int k = 0; k < D; ++k) q[a][k] = fma(-dp, q[b][k], q[a][k]);
            }
        }
        double nn = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nn = fma(q[a][k], q[a][k], nn);
        nn = 1.0 / sqrt(nn);
#pragma unroll
        for (int k = 0; k < D; ++k) q[a][k] *= nn;
    }
    double w[D];
    hv_span_vec<D>(g, s, x0, w);
#pragma unroll
    for (int rep = 0; rep < 2; ++rep) {
#pragma unroll
        for (int b = 0; b < D - 1; ++b) {
            double dp = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) dp = fma(w[k], q[b][k], dp);
#pragma unroll
            for (int k = 0; k < D; ++k) w[k] = fma(-dp, q[b][k], w[k]);
        }
    }
    double nw = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) nw = fma(w[k], w[k], nw);
    nw = -1.0 / sqrt(nw);
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = w[k] * nw;
    du = 0.0;
#pragma unroll
    for (int a = 0; a < D - 1; ++a) {
        double raw[D];
        hv_span_vec<D>(g, e[D - 1 - a], x0, raw);
        double dp = 0.0, nn = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) { dp = fma(raw[k], u[k], dp); nn = fma(raw[k], raw[k], nn); }
        du += fabs(dp) / sqrt(nn);
    }
    double ux0 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) ux0 = fma(u[k], x0[k], ux0);
    double cmax = ux0;
#pragma unroll
    for (int j = 1; j < D; ++j) {
        const int el = e[j];
        double v;
        if (el < N) {
            v = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) v = fma(u[k], g.xo[(int64_t)(el) * D + k], v);
        } else {
            const int p = el - N;
            double dist = g.ph[p], nu = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) { dist = fma(-g.pn[p * D + k], x0[k], dist); nu = fma(g.pn[p * D + k], u[k], nu); }
            v = ux0 + 2.0 * dist * nu;   // <mirror of x0 across p, u>
        }
        cmax = fmax(cmax, v);
    }
    c = cmax + fabs(cmax) * plane_tol;
    double rr = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) { double t = r[k] - x0[k]; rr = fma(t, t, rr); }
    rho = sqrt(rr);
}

// One thread per 4 edge slots (t = block*1024 + j*256 + thread, coalesced per j), and one
// counter atomic per workgroup: a wavefront-level allocation made this pass bound by the
// serialised same-address atomics (one per 64 slots) instead of by its loads.
constexpr int HV_EMIT_PER = 4;

template <int D>
__global__ __launch_bounds__(256) void k_emit(const int* __restrict__ frontier, int nf, const int* __restrict__ vkey,
                                              const double* __restrict__ vx, DevGeo g,
                                              const unsigned long long* __restrict__ nt,
                                              const int64_t* __restrict__ reg, DevRays R, int* __restrict__ nrays,
                                              double plane_tol) {
    __shared__ int wsum[4];
    __shared__ int wbase[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nslot = (int64_t)nf * (D + 1);
    const int64_t t0 = (int64_t)blockIdx.x * (256 * HV_EMIT_PER) + threadIdx.x;
    int64_t h[HV_EMIT_PER];
    unsigned long long m[HV_EMIT_PER];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < HV_EMIT_PER; ++j) {
        const int64_t t = t0 + j * 256;
        bool want = false;
        h[j] = -1;
        if (t < nslot) {
            h[j] = reg[t];
            if (h[j] >= 0) {
                const unsigned long long cur = nt[h[j]];
                const int fi = (int)(t / (D + 1)), pos = (int)(t % (D + 1));
                const unsigned long long mine =
                    ((((unsigned long long)frontier[fi]) << 4) | (unsigned)pos) + 1ull;
                want = !(cur & HV_EDGE_DONE) && ((cur & HV_EDGE_VMASK) == mine);
            }
        }
        m[j] = __ballot(want);
        cnt += __popcll(m[j]);
    }
    if (lane == 0) wsum[wv] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        int b = tot ? atomicAdd(nrays, tot) : 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) { wbase[w] = b; b += wsum[w]; }
    }
    __syncthreads();
    int base = wbase[wv];
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
#pragma unroll
    for (int j = 0; j < HV_EMIT_PER; ++j) {
        if ((m[j] >> lane) & 1ull) {
            const int idx = base + __popcll(m[j] & below);
            const int64_t t = t0 + j * 256;
            R.slot[idx] = h[j];
            R.v[idx] = frontier[(int)(t / (D + 1))];
            R.drop[idx] = (int)(t % (D + 1));
        }
        base += __popcll(m[j]);
    }
}

// Direction, half-space threshold and orthogonality defect of rays [0, n) (dense: one
// thread per emitted ray).
template <int D, bool B>
__global__ __launch_bounds__(256) void k_ray_dir(DevGeo g0, GeoSet F, DevRays R, int n, const int* __restrict__ vkey,
                                                 const double* __restrict__ vx, double plane_tol) {
    const int t = hv_xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int v = R.v[t], pos = R.drop[t];
    // key read straight from global memory: a private copy indexed by `pos` would live in scratch
    const int* key = vkey + (int64_t)v * (D + 1);
    const DevGeo& g = hv_geo<B>(g0, F, key[0]);
    double r[D];
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = vx[(int64_t)v * D + k];
    double u[D], c, du, rho;
    hv_ray_setup<D>(g, key, pos, r, plane_tol, u, c, du, rho);
#pragma unroll
    for (int k = 0; k < D; ++k) R.u[(int64_t)t * D + k] = u[k];
    R.c[t] = c;
    R.du[t] = du;
    (void)rho;
}

// ---------------------------------------------------------------------------------
// The candidate scan: one wavefront walks the generator tree for 64 rays at once.


template <int D>
__device__ __forceinline__ double hv_box_dist2(const DevGeo& g, int node, const double* p) {
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double lo = g.blo[(int64_t)k * g.NN + node], hi = g.bhi[(int64_t)k * g.NN + node];
        const double a = fmax(fmax(lo - p[k], p[k] - hi), 0.0);
        d2 = fma(a, a, d2);
    }
    return d2;
}

__device__ __forceinline__ void hv_wave_sync() {
    // lanes of one wavefront: LDS operations complete in order, this only stops the
    // compiler from moving LDS accesses across the point
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr double HV_W = 1e-7;
// Rays are ordered by the top 32 of the 60 Morton bits of their ball centre (2^(32/d) cells
// per axis: far finer than a level's rays are spread), half the radix passes of all 60.
constexpr int HV_RAY_SORT_LO = 28;


// read-only tables seen through the constant address space: uniform loads become SMEM
typedef const __attribute__((address_space(4))) double hv_cdouble;
typedef const __attribute__((address_space(4))) float hv_cfloat;

// a value known to be equal on every lane, moved to scalar registers
__device__ __forceinline__ double hv_uniform(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(bits & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(bits >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}   // widest near-tie window of the reference (src/raycast.jl:902)

// Per-ray geometry shared by the start phase and the packet scan.
template <int D>
struct RayGeo {
    double u[D], r[D], x0[D];
    int e[D];
    int s;
    double cth, du, ux0, K;
};

template <int D>
__device__ __forceinline__ void hv_load_ray(const DevGeo& g, const DevRays& R, int ray, const int* __restrict__ vkey,
                                            const double* __restrict__ vx, RayGeo<D>& G) {
    const int v = R.v[ray], pos = R.drop[ray];
    // key read straight from global memory: a private copy indexed by `pos` would live in scratch
    const int* key = vkey + (int64_t)v * (D + 1);
    hv_edge_of<D>(key, pos, G.e);
    G.s = hv_key_at<D>(key, pos);
    double rx0 = 0.0, x0x0 = 0.0;
    G.ux0 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        G.u[k] = R.u[(int64_t)ray * D + k];
        G.r[k] = vx[(int64_t)v * D + k];
        G.x0[k] = g.xo[(int64_t)(G.e[0]) * D + k];
        G.ux0 = fma(G.u[k], G.x0[k], G.ux0);
        rx0 = fma(G.r[k], G.x0[k], rx0);
        x0x0 = fma(G.x0[k], G.x0[k], x0x0);
    }
    G.K = 2.0 * rx0 - x0x0;
    G.cth = R.c[ray];
    G.du = R.du[ray];
}

// Top-2 candidates; selects only (branchy updates get sunk into a scratch store).
struct Top2 {
    double t1 = 1e300, t2 = 1e300;
    int i1 = -1, i2 = -1;
    __device__ __forceinline__ void offer(double t, int id) {
        const bool b1 = t < t1;
        const bool b2 = !b1 && (t < t2) && (id != i1);
        const double nt2 = b1 ? t1 : (b2 ? t : t2);
        const int ni2 = b1 ? i1 : (b2 ? id : i2);
        t1 = b1 ? t : t1;
        i1 = b1 ? id : i1;
        t2 = nt2;
        i2 = ni2;
    }
};

// exact t of generator j (sorted index) with coordinates xk, |x|^2 = xxj (src/raycast.jl:424-431).
// Every candidate ahead of the edge competes, also one closer than plane_tol: the reference
// drops such a candidate, but its window (upper_t, from the nearest generator ahead) then holds
// nothing else and the ray ends without a vertex (src/raycast.jl:883-894), which the winner
// rule decides once the nearest candidate is known.
template <int D>
__device__ __forceinline__ void hv_consider(const RayGeo<D>& G, Top2& T, const double* xk, double xxj, int j,
                                            double plane_tol) {
    (void)plane_tol;
    double ux = 0.0, rxj = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ux = fma(G.u[k], xk[k], ux);
        rxj = fma(G.r[k], xk[k], rxj);
    }
    if (!(ux > G.cth)) return;
    const double den2 = 2.0 * (ux - G.ux0);
    const double num = fma(-2.0, rxj, xxj + G.K);
    if (num < den2 * (T.t1 + HV_W)) T.offer(num / den2, j);
}

// box distance helpers for the per-thread (stackless, heap-order) tree walks
template <int D>
__device__ __forceinline__ double hv_box_d2(const DevGeo& g, int node, const double* q) {
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double lo = g.blo[(int64_t)k * g.NN + node], hi = g.bhi[(int64_t)k * g.NN + node];
        const double a = fmax(fmax(lo - q[k], q[k] - hi), 0.0);
        d2 = fma(a, a, d2);
    }
    return d2;
}

__device__ __forceinline__ int hv_next_node(int i) {
    while (i & 1) i >>= 1;
    return i == 0 ? 0 : i + 1;
}

// A ray whose start candidates all lie behind it in a domain without a bounded boundary: the
// best candidate ahead of it from the binary tree (boxes with a point ahead, <u, x> > c, and
// within the ball of the best t so far); none: the ray runs to infinity (the reference's
// unbounded ray, t = Inf, src/sysvoronoi.jl:504-511)
template <int D>
__device__ void hv_probe(const DevGeo& g, const RayGeo<D>& G, Top2& T, double plane_tol) {
    int node = 1;
    while (node) {
        double umax = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double lo = g.blo[(int64_t)k * g.NN + node], hi = g.bhi[(int64_t)k * g.NN + node];
            umax = fma(G.u[k], G.u[k] > 0.0 ? hi : lo, umax);
        }
        bool enter = umax > G.cth;
        if (enter && T.i1 >= 0) {
            const double tau = T.t1 + HV_W;
            double q[D], R2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                q[k] = fma(tau, G.u[k], G.r[k]);
                const double a = q[k] - G.x0[k];
                R2 = fma(a, a, R2);
            }
            enter = hv_box_d2<D>(g, node, q) <= R2 * (1.0 + 1e-9);
        }
        if (enter) {
            if (node < g.P2) {
                node = 2 * node;
                continue;
            }
            const int j0 = (node - g.P2) * g.L, j1 = min(j0 + g.L, g.n);
            for (int j = j0; j < j1; ++j) {
                double xk[D];
#pragma unroll
                for (int k = 0; k < D; ++k) xk[k] = g.xs[(int64_t)j * (D + 1) + k];
                hv_consider<D>(G, T, xk, g.xs[(int64_t)j * (D + 1) + D], j, plane_tol);
            }
        }
        node = hv_next_node(node);
    }
}

// Start phase, one thread per ray: boundary planes (mirror pseudo-generators,
// src/raycast.jl:354-375, src/extended.jl:131-141) and the HV_KNN nearest generators of
// x0 (the winner is a Delaunay neighbour of x0, usually among them).  The Morton code
// of the resulting certificate-ball centre orders the rays into packets.
template <int D, bool B>
__global__ __launch_bounds__(256) void k_ray_init(DevGeo g0, GeoSet F, DevRays R, int nrays, const int* __restrict__ vkey,
                                                  const double* __restrict__ vx, double plane_tol) {
    const int ray = hv_xcd_block() * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    const int gen0 = vkey[(int64_t)R.v[ray] * (D + 1)];   // a generator: planes sort last
    const DevGeo& g = hv_geo<B>(g0, F, gen0);
    const int N = g.N;
    RayGeo<D> G;
    hv_load_ray<D>(g, R, ray, vkey, vx, G);
    Top2 T;
    for (int p = 0; p < g.P; ++p) {
        const int id = N + p;
        bool member = (id == G.s);
#pragma unroll
        for (int j = 0; j < D; ++j) member = member || (G.e[j] == id);
        if (member) continue;
        double nu = 0.0, dist = g.ph[p], nr = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double nk = g.pn[p * D + k];
            nu = fma(nk, G.u[k], nu);
            dist = fma(-nk, G.x0[k], dist);
            nr = fma(nk, G.r[k], nr);
        }
        if (!(G.ux0 + 2.0 * dist * nu > G.cth)) continue;
        T.offer((g.ph[p] - nr) / nu, id);   // one closer than plane_tol ends the ray (hv_consider)
    }
    // the winner is a Delaunay neighbour of every generator of the edge: scan their lists.
    // FP32 pre-ranking on the difference form t = (|v|^2 + 2<v, x0 - r>) / (2<u, v>), v = x - x0
    // (small magnitudes), then the best three in FP64 (hv_consider): the start only needs
    // exact t of real candidates, which candidates win it changes nothing downstream.
    float uf[D], wf[D], x0f[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        uf[k] = (float)G.u[k];
        wf[k] = (float)(G.x0[k] - G.r[k]);
        x0f[k] = (float)G.x0[k];
    }
    const float cf = (float)(G.cth - G.ux0);
    // the edge's own generators lie on the bisector (their FP32 <u, v> is rounding noise)
    int er[D];
#pragma unroll
    for (int a = 0; a < D; ++a) er[a] = G.e[a] < N ? g.rank[G.e[a]] : -1;
    float b0 = INFINITY, b1 = INFINITY, b2 = INFINITY;
    int j0b = -1, j1b = -1, j2b = -1;
    const float4* xf4 = (const float4*)g.xf;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const int ea = G.e[a];
        if (ea >= N) break;   // planes sort last
        const int* nb = g.knn + (int64_t)g.rank[ea] * HV_KNN;
        int jv[HV_KNN];   // the list's ids in 16-byte loads, ahead of the dependent row gathers
#pragma unroll
        for (int q = 0; q < HV_KNN / 4; ++q) {
            const int4 w = ((const int4*)nb)[q];
            jv[4 * q] = w.x;
            jv[4 * q + 1] = w.y;
            jv[4 * q + 2] = w.z;
            jv[4 * q + 3] = w.w;
        }
#pragma unroll
        for (int m = 0; m < HV_KNN; ++m) {
            const int j = jv[m];
            bool edge = false;
#pragma unroll
            for (int a2 = 0; a2 < D; ++a2) edge = edge || (j == er[a2]);
            if (edge) continue;
            const float4 A = xf4[(int64_t)j * 2];
            float x[8] = {A.x, A.y, A.z, A.w, 0.f, 0.f, 0.f, 0.f};
            if constexpr (D > 4) {
                const float4 Bq = xf4[(int64_t)j * 2 + 1];
                x[4] = Bq.x; x[5] = Bq.y; x[6] = Bq.z; x[7] = Bq.w;
            }
            float vu = 0.f, vv = 0.f, vw = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float v = x[k] - x0f[k];
                vu = fmaf(uf[k], v, vu);
                vv = fmaf(v, v, vv);
                vw = fmaf(v, wf[k], vw);
            }
            if (!(vu > cf)) continue;
            const float tt = __fdividef(fmaf(2.f, vw, vv), 2.f * vu);
            if (j == j0b || j == j1b || j == j2b) continue;
            const bool c0 = tt < b0, c1 = !c0 && tt < b1, c2 = !c0 && !c1 && tt < b2;
            b2 = (c0 || c1) ? b1 : (c2 ? tt : b2);
            j2b = (c0 || c1) ? j1b : (c2 ? j : j2b);
            b1 = c0 ? b0 : (c1 ? tt : b1);
            j1b = c0 ? j0b : (c1 ? j : j1b);
            b0 = c0 ? tt : b0;
            j0b = c0 ? j : j0b;
        }
    }
    {
        const int jb[3] = {j0b, j1b, j2b};
#pragma unroll
        for (int z = 0; z < 3; ++z) {
            const int j = jb[z];
            if (j < 0) continue;
            double xk[D];
#pragma unroll
            for (int k = 0; k < D; ++k) xk[k] = g.xs[(int64_t)(j) * (D + 1) + k];
            hv_consider<D>(G, T, xk, g.xs[(int64_t)(j) * (D + 1) + D], j, plane_tol);
        }
    }
    if (g.unbounded && T.i1 < 0) hv_probe<D>(g, G, T, plane_tol);
    R.t1[ray] = T.t1;
    R.t2[ray] = T.t2;
    R.i1[ray] = T.i1;
    R.i2[ray] = T.i2;
    double bc[D];
#pragma unroll
    for (int k = 0; k < D; ++k) bc[k] = G.r[k] + (T.t1 + HV_W) * G.u[k];
    // over a forest: rays of one mesh together (mesh above the top Morton bits the sort uses)
    if constexpr (B) R.code[ray] = ((uint64_t)F.mof[gen0] << 32) | (hv_morton<D>(g, bc) >> HV_RAY_SORT_LO);
    else R.code[ray] = hv_morton<D>(g, bc);
}

// One wavefront serves 64 rays at once (a packet).  Every generator x with t(x) < tau
// lies strictly inside the certificate ball B(tau) = {x : |x - p|^2 < |x0 - p|^2},
// p = r + tau u: for x ahead of the edge, |x-p|^2 - |x0-p|^2 = 2<u,x-x0>(t(x) - tau), and
// B(tau) holds no generator behind the edge.  So
//   * a tree node is skipped when its box misses every live ray's ball, and
//   * the leaf scan is the membership test f(x) = |x|^2 - 2<x,p> + |p|^2 - R^2 < eps,
//     exact t only for members.
// The walk is depth-first, children nearer to the packet centre first; each visited leaf
// is staged in LDS by the whole wavefront (coalesced) and scanned by the lanes whose
// ball reaches it.  The two smallest t are kept; the reference's near-tie rule
// (src/raycast.jl:877-916) is applied to them at the end.
// One wavefront per workgroup: packets take very different times, and a workgroup of
// several wavefronts holds its slots until its slowest packet is done.
// The HV_KNN nearest generators of every generator, 64 Morton-consecutive queries per
// wavefront: the packet's own 64 points first, then the 8-ary tree nearest-first, entering a
// child when its box is within the current 16th distance of some query (FP32 test, widened);
// each visited leaf is staged in LDS and scanned by every lane into a sorted top-16.
template <int D, bool B>
__global__ __launch_bounds__(64) void k_knn64(DevGeo g0, GeoSet F, const int* __restrict__ boff, int* __restrict__ knn0) {
    __shared__ int stk[128], stkl[128];
    __shared__ float cbx[16 * D];
    __shared__ unsigned long long hms[8];
    __shared__ float px[D][64];
    const int lane = threadIdx.x;
    int blk = hv_xcd_block(), mk = 0;
    if constexpr (B) {   // block of the whole grid -> (mesh, block of the mesh)
        mk = __builtin_amdgcn_readfirstlane(hv_mesh_of_item(boff, F.nmesh, blk));
        blk -= boff[mk];
    }
    const DevGeo& g = hv_geo_mesh<B>(g0, F, mk);
    int* knn = B ? const_cast<int*>(g.knn) : knn0;
    const int N = g.n;   // sorted positions
    const int base = blk * 64;
    const int j = base + lane;
    const bool act = j < N;
    const int jj = act ? j : N - 1;
    // FP32 distances: the lists only supply candidates that are always evaluated in FP64
    float pf[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pf[k] = (float)g.xs[(int64_t)jj * (D + 1) + k];
    float bd[HV_KNN];
    int bi[HV_KNN];
#pragma unroll
    for (int m = 0; m < HV_KNN; ++m) { bd[m] = 3.0e38f; bi[m] = jj; }
    auto consider = [&](int q, float d2) {
        if (!(d2 < bd[HV_KNN - 1]) || q == jj) return;
        int pos = 0;
#pragma unroll
        for (int m = 0; m < HV_KNN; ++m) pos += (bd[m] <= d2) ? 1 : 0;
#pragma unroll
        for (int m = HV_KNN - 1; m > 0; --m) {
            bd[m] = (m > pos) ? bd[m - 1] : ((m == pos) ? d2 : bd[m]);
            bi[m] = (m > pos) ? bi[m - 1] : ((m == pos) ? q : bi[m]);
        }
        bd[0] = (pos == 0) ? d2 : bd[0];
        bi[0] = (pos == 0) ? q : bi[0];
    };
    // scan the npts points staged in px (global ids q0 + s)
    auto scan_staged = [&](int q0, int npts, bool skip_own) {
        for (int s = 0; s < npts; ++s) {
            const int q = q0 + s;
            if (skip_own && q >= base && q < base + 64) continue;
            float d2 = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float a = px[k][s] - pf[k];
                d2 = fmaf(a, a, d2);
            }
            consider(q, d2);
        }
    };
    {
        const int np = min(64, N - base);
#pragma unroll
        for (int k = 0; k < D; ++k) px[k][lane] = pf[k];
        hv_wave_sync();
        scan_staged(base, np, false);
        hv_wave_sync();
    }
    // packet centre (uniform) orders the children
    float pcf[D];
    {
        double c[D], cnt = act ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) c[k] = act ? (double)pf[k] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < D; ++k) c[k] += __shfl_xor(c[k], off);
            cnt += __shfl_xor(cnt, off);
        }
#pragma unroll
        for (int k = 0; k < D; ++k)
            pcf[k] = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint((float)(c[k] / cnt))));
    }
    auto child_d2f = [&](int c, const float* q) {
        float d2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float a = q[k] - __builtin_amdgcn_fmed3f(q[k], cbx[c * 2 * D + k], cbx[c * 2 * D + D + k]);
            d2 = fmaf(a, a, d2);
        }
        return d2;
    };
    const float slack = 1e-6f * (float)(g.xmax + 1.0);
    auto reach = [&]() {   // widened radius^2 of the current 16th neighbour
        const float r = sqrtf(bd[HV_KNN - 1]) + slack;
        return fminf(r * r * (1.0f + 1e-5f), 1e30f);
    };
    // children ordered by distance to the packet centre (ties: lower index first): rank r
    // holds child (ord >> 4r) & 7
    uint32_t ord = 0u;
    auto to_rank = [&](unsigned cm) {
        unsigned rr = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) rr |= ((cm >> ((ord >> (4 * q)) & 7u)) & 1u) << q;
        return rr;
    };
    int sp = 1;
    stk[0] = 0;
    stkl[0] = 0;
    while (sp > 0) {
        --sp;
        hv_wave_sync();
        const int gid = __builtin_amdgcn_readfirstlane(stk[sp]);
        const int l = __builtin_amdgcn_readfirstlane(stkl[sp]);
        {
            const float* src = g.cb + (int64_t)gid * (16 * D);
            for (int q = lane; q < 16 * D; q += 64) cbx[q] = src[q];
        }
        hv_wave_sync();
        const float R2f = reach();
        unsigned rem = 0u;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const unsigned long long m = __ballot(act && (child_d2f(c, pf) <= R2f));
            if (lane == 0) hms[c] = m;
            rem |= (m != 0ull) ? (1u << c) : 0u;
        }
        {
            const uint32_t kb = __float_as_uint(child_d2f(lane & 7, pcf));   // non-negative: orders as unsigned
            const int me = lane & 7;
            int rk = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const uint32_t kc = __builtin_amdgcn_readlane(kb, c);
                rk += (kc < kb || (kc == kb && c < me)) ? 1 : 0;
            }
            uint32_t o = 0u;
#pragma unroll
            for (int c = 0; c < 8; ++c) o |= (uint32_t)c << (4 * __builtin_amdgcn_readlane(rk, c));
            ord = o;
        }
        rem = __builtin_amdgcn_readfirstlane(rem);
        hv_wave_sync();
        if (rem == 0u) continue;
        const int off = ((1 << (3 * l)) - 1) / 7;
        const int i = gid - off;
        if (l + 1 == g.h) {
            unsigned rr = to_rank(rem);
            while (rr) {
                const int cc = (ord >> (4 * __builtin_ctz(rr))) & 7u;
                rr &= rr - 1u;
                const int leaf = 8 * i + cc;
                const int j0 = leaf * g.L, npts = max(0, min(g.L, N - j0));
                if (j0 >= base && j0 + npts <= base + 64) continue;   // the packet's own points
                const bool hit = act && ((hms[cc] >> lane) & 1ull) && (child_d2f(cc, pf) <= reach());
                if (__ballot(hit) == 0ull) continue;
                for (int s0 = 0; s0 < npts; s0 += 64) {
                    const int cnt = min(64, npts - s0);
                    if (lane < cnt) {
#pragma unroll
                        for (int k = 0; k < D; ++k) px[k][lane] = (float)g.xs[(int64_t)(j0 + s0 + lane) * (D + 1) + k];
                    }
                    hv_wave_sync();
                    if (hit) scan_staged(j0 + s0, cnt, true);
                    hv_wave_sync();
                }
            }
        } else {
            const int cbase = off + (1 << (3 * l)) + 8 * i;
            unsigned rr = to_rank(rem);
            while (rr) {
                const int q = 31 - __builtin_clz(rr);
                rr &= ~(1u << q);
                const int c = (ord >> (4 * q)) & 7u;
                stk[sp] = cbase + c;
                stkl[sp] = l + 1;
                ++sp;
            }
        }
    }
    if (act) {
#pragma unroll
        for (int m = 0; m < HV_KNN; ++m) knn[(int64_t)j * HV_KNN + m] = bi[m];
    }
}

// Point-sample Monte Carlo: wavefront w holds samples 64w .. 64w+63, all in stratum
// floor(64w / S) of a G^d grid over the domain box (S samples per stratum, a multiple of 64),
// each uniform in its stratum from a counter-based stream.  A packet walk as in k_knn64 finds
// each sample's nearest and second-nearest generators (FP64 distances to the staged leaves).
// Accumulated: the samples of each nearest cell (one atomic per distinct cell of the
// wavefront), and per (cell, neighbour) row the samples that lie within eps (counter 2 row)
// and within 2 eps (counter 2 row + 1) of the bisector of their two nearest generators, or of
// a boundary plane, on the cell's side.
__device__ __forceinline__ double hv_sample_unif(uint64_t seed, uint64_t idx) {
    return (double)(hv_mix64(seed ^ ((idx + 1ull) * 0x9E3779B97F4A7C15ull)) >> 11) * (1.0 / 9007199254740992.0);
}

struct HvBox {
    double lo[6], hi[6];
};

__device__ __forceinline__ void hv_row_add(const int64_t* __restrict__ ptr, const uint64_t* __restrict__ pkey,
                                           unsigned long long* __restrict__ pcnt, int cell, int nb, bool inner) {
    int64_t lo = ptr[cell], hi = ptr[cell + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int)(pkey[mid] & 0xffffffffull) < nb) lo = mid + 1;
        else hi = mid;
    }
    if (lo < ptr[cell + 1] && (int)(pkey[lo] & 0xffffffffull) == nb) {
        atomicAdd(&pcnt[2 * lo + 1], 1ull);
        if (inner) atomicAdd(&pcnt[2 * lo], 1ull);
    }
}

constexpr int HV_MCP_SPREAD = 256;   // copies of the sample kernel's work counter

template <int D>
__global__ __launch_bounds__(64) void k_mc_points(DevGeo g, int G, int64_t S, uint64_t seed, double eps,
                                                  const int64_t* __restrict__ ptr, const uint64_t* __restrict__ pkey,
                                                  unsigned long long* __restrict__ ccnt,
                                                  unsigned long long* __restrict__ pcnt, HvBox box, int pmask,
                                                  int nreal, const int* __restrict__ porig,
                                                  const uint32_t* __restrict__ pshift,
                                                  const unsigned long long* __restrict__ pkeys,
                                                  const int* __restrict__ pvals, uint64_t pmaskh,
                                                  unsigned long long* __restrict__ npairs) {
    // samples are drawn in `box`.  Periodic directions (pmask): the working set holds images
    // of the nreal originals (porig / pshift per id: the original and the shift, 4 bits per
    // direction, shift + 8).  A sample whose nearest generator a is an image lies, moved by
    // -shift(a), in the cell of a's original; it is counted there, and its second generator b
    // is renamed to the image of b's original at shift(b) - shift(a), looked up in the table
    // (pkeys, pvals): key (original << 32) | shift code.
    __shared__ int stk[128], stkl[128];
    __shared__ float cbx[16 * D];
    __shared__ unsigned long long hms[8];
    __shared__ float px[D][64];
    const int lane = threadIdx.x;
    const int N = g.N;
    const int64_t smp = (int64_t)hv_xcd_block() * 64 + lane;
    double p[D];
    float pf[D];
    {
        int64_t b = smp / S;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int ck = (int)(b % G);
            b /= G;
            const double wk = (box.hi[k] - box.lo[k]) / (double)G;
            p[k] = box.lo[k] + ((double)ck + hv_sample_unif(seed, (uint64_t)smp * D + k)) * wk;
            pf[k] = (float)p[k];
        }
    }
    double bd0 = 1e300, bd1 = 1e300;
    int bi0 = -1, bi1 = -1;
    auto consider = [&](int q, double d2) {
        const bool f = d2 < bd0, s2 = !f && d2 < bd1;
        bd1 = f ? bd0 : (s2 ? d2 : bd1);
        bi1 = f ? bi0 : (s2 ? q : bi1);
        bd0 = f ? d2 : bd0;
        bi0 = f ? q : bi0;
    };
    // packet centre (uniform) orders the children
    float pcf[D];
    {
        double cc[D];
#pragma unroll
        for (int k = 0; k < D; ++k) cc[k] = p[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < D; ++k) cc[k] += __shfl_xor(cc[k], off);
        }
#pragma unroll
        for (int k = 0; k < D; ++k)
            pcf[k] = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint((float)(cc[k] * (1.0 / 64.0)))));
    }
    auto child_d2f = [&](int c, const float* q) {
        float d2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float a = q[k] - __builtin_amdgcn_fmed3f(q[k], cbx[c * 2 * D + k], cbx[c * 2 * D + D + k]);
            d2 = fmaf(a, a, d2);
        }
        return d2;
    };
    const double slack = 1e-6 * (g.xmax + 1.0);
    auto reach = [&]() {   // widened FP32 radius^2 of the current second neighbour
        const double r = sqrt(bd1) + slack;
        return (float)fmin(r * r * (1.0 + 1e-5), 1e30);
    };
    uint32_t ord = 0u;   // children by distance to the packet centre: rank r holds (ord >> 4r) & 7
    auto to_rank = [&](unsigned cm) {
        unsigned rr = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) rr |= ((cm >> ((ord >> (4 * q)) & 7u)) & 1u) << q;
        return rr;
    };
    int sp = 1;
    stk[0] = 0;
    stkl[0] = 0;
    unsigned work = 0u;   // (sample, generator) distances this lane evaluated
    while (sp > 0) {
        --sp;
        hv_wave_sync();
        const int gid = __builtin_amdgcn_readfirstlane(stk[sp]);
        const int l = __builtin_amdgcn_readfirstlane(stkl[sp]);
        {
            const float* src = g.cb + (int64_t)gid * (16 * D);
            for (int q = lane; q < 16 * D; q += 64) cbx[q] = src[q];
        }
        hv_wave_sync();
        const float R2f = reach();
        unsigned rem = 0u;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const unsigned long long m = __ballot(child_d2f(c, pf) <= R2f);
            if (lane == 0) hms[c] = m;
            rem |= (m != 0ull) ? (1u << c) : 0u;
        }
        {
            const uint32_t kb = __float_as_uint(child_d2f(lane & 7, pcf));
            const int me = lane & 7;
            int rk = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const uint32_t kc = __builtin_amdgcn_readlane(kb, c);
                rk += (kc < kb || (kc == kb && c < me)) ? 1 : 0;
            }
            uint32_t o = 0u;
#pragma unroll
            for (int c = 0; c < 8; ++c) o |= (uint32_t)c << (4 * __builtin_amdgcn_readlane(rk, c));
            ord = o;
        }
        rem = __builtin_amdgcn_readfirstlane(rem);
        hv_wave_sync();
        if (rem == 0u) continue;
        const int off = ((1 << (3 * l)) - 1) / 7;
        const int i = gid - off;
        if (l + 1 == g.h) {
            unsigned rr = to_rank(rem);
            while (rr) {
                const int cc = (ord >> (4 * __builtin_ctz(rr))) & 7u;
                rr &= rr - 1u;
                const int leaf = 8 * i + cc;
                const int j0 = leaf * g.L, npts = max(0, min(g.L, N - j0));
                const bool hit = ((hms[cc] >> lane) & 1ull) && (child_d2f(cc, pf) <= reach());
                if (__ballot(hit) == 0ull) continue;
                for (int s0 = 0; s0 < npts; s0 += 64) {
                    const int cnt = min(64, npts - s0);
                    if (lane < cnt) {   // FP32 rows; the two nearest are measured again in FP64
#pragma unroll
                        for (int k = 0; k < D; ++k) px[k][lane] = g.xf[(int64_t)(j0 + s0 + lane) * 8 + k];
                    }
                    hv_wave_sync();
                    if (hit) {
                        work += (unsigned)cnt;
                        for (int q = 0; q < cnt; ++q) {
                            float d2 = 0.f;
#pragma unroll
                            for (int k = 0; k < D; ++k) {
                                const float a = px[k][q] - pf[k];
                                d2 = fmaf(a, a, d2);
                            }
                            consider(j0 + s0 + q, (double)d2);
                        }
                    }
                    hv_wave_sync();
                }
            }
        } else {
            const int cbase = off + (1 << (3 * l)) + 8 * i;
            unsigned rr = to_rank(rem);
            while (rr) {
                const int q = 31 - __builtin_clz(rr);
                rr &= ~(1u << q);
                const int c = (ord >> (4 * q)) & 7u;
                stk[sp] = cbase + c;
                stkl[sp] = l + 1;
                ++sp;
            }
        }
    }
    {
        unsigned long long w = work;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
        // spread over HV_MCP_SPREAD copies (one same-address atomic per wavefront serialises)
        if (lane == 0) atomicAdd(&npairs[blockIdx.x % HV_MCP_SPREAD], w);
    }
    // the two nearest in FP64 (the FP32 scan may order a near tie either way)
    auto d2_64 = [&](int j) {
        double s2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = g.xs[(int64_t)j * (D + 1) + k] - p[k];
            s2 = fma(a, a, s2);
        }
        return s2;
    };
    bd0 = d2_64(bi0);   // bi0 >= 0: the walk starts with an unbounded radius and N > d
    if (bi1 >= 0) {
        bd1 = d2_64(bi1);
        if (bd1 < bd0) {
            const double td = bd0; bd0 = bd1; bd1 = td;
            const int ti = bi0; bi0 = bi1; bi1 = ti;
        }
    }
    int cell = g.sidx[bi0];
    int nb1 = bi1 >= 0 ? g.sidx[bi1] : -1;
    if (pmask && cell >= nreal) {
        const uint32_t ca = pshift[cell];
        cell = porig[cell];
        if (nb1 >= 0) {
            const uint32_t cb = pshift[nb1];
            uint32_t code = 0u;
            bool ok = true;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const int sh = (int)((cb >> (4 * k)) & 15u) - (int)((ca >> (4 * k)) & 15u);
                ok = ok && sh >= -8 && sh <= 7;
                code |= (uint32_t)((sh + 8) & 15) << (4 * k);
            }
            const int ob = porig[nb1];
            int found = -1;
            if (ok) {
                const unsigned long long key = ((unsigned long long)(unsigned)ob << 32) | code;
                uint64_t h = hv_mix64(key) & pmaskh;
                for (uint64_t probe = 0; probe <= pmaskh; ++probe) {
                    const unsigned long long cur = pkeys[h];
                    if (cur == key) { found = pvals[h]; break; }
                    if (cur == ~0ull) break;
                    h = (h + 1) & pmaskh;
                }
            }
            nb1 = found;   // -1: that image is not in the working set (then no interface count)
        }
    }
    // samples per nearest cell: one atomic per distinct cell of the wavefront
    unsigned long long todo = __ballot(true);
    while (todo) {
        const int l0 = __ffsll((long long)todo) - 1;
        const int cl = __builtin_amdgcn_readlane(cell, l0);
        const unsigned long long same = __ballot(cell == cl) & todo;
        if (lane == l0) atomicAdd(&ccnt[cl], (unsigned long long)__popcll(same));
        todo &= ~same;
    }
    // slabs: the bisector of the two nearest generators, the boundary planes
    if (bi1 >= 0) {
        double dd = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = g.xs[(int64_t)bi0 * (D + 1) + k] - g.xs[(int64_t)bi1 * (D + 1) + k];
            dd = fma(a, a, dd);
        }
        const double gap = bd1 - bd0, w = 2.0 * eps * sqrt(dd);   // distance to the bisector = gap / (2|xi - xj|)
        if (gap < 2.0 * w && nb1 >= 0) hv_row_add(ptr, pkey, pcnt, cell, nb1, gap < w);
    }
    for (int q = 0; q < g.P; ++q) {
        if ((pmask >> (q >> 1)) & 1) continue;   // periodic planes bound no cell
        double dist = g.ph[q];
#pragma unroll
        for (int k = 0; k < D; ++k) dist = fma(-g.pn[q * D + k], p[k], dist);
        if (dist < 2.0 * eps) hv_row_add(ptr, pkey, pcnt, cell, N + q, dist < eps);
    }
}

// MODE 0: FP64 leaf scan.  MODE 1: BF16 MFMA filter, MODE 2: FP32 filter, each with FP64
// re-verification of the survivors (leaves of 32 points).  MODE 1 is held to 128 VGPRs
// (4 wavefronts per SIMD): the few values that spill are the rarely used FP64 ray data.
// A packet holds `ppw` <= 64 rays (lanes ppw..63 idle): small levels use fewer rays per
// packet, so that there are enough packets to fill the chip and each one walks fewer leaves.
// Over a forest (B) the rays are ordered by mesh and a packet holds rays of one mesh: mesh m
// holds the list positions seg[m] .. seg[m+1] and the packets pk_off[m] .. pk_off[m+1].
template <int D, int MODE, bool B>
__global__ __launch_bounds__(64, MODE == 1 ? 4 : 1) void k_raycast64(DevGeo g0, GeoSet F, const int* __restrict__ pk_off,
                                                  const int* __restrict__ seg, DevRays R, const int* __restrict__ order, int nrays,
                                                  const int* __restrict__ nrays_dev, int ppw,
                                                  const int* __restrict__ vkey, const double* __restrict__ vx,
                                                  int* __restrict__ status, int* __restrict__ win,
                                                  double* __restrict__ tout, double plane_tol,
                                                  unsigned long long* __restrict__ ctr0, int cstride) {
    __shared__ int stk[128], stkl[128];
    __shared__ float cbx[16 * D];
    // statistics counter k of this workgroup: ctr[k * cstride]; with cstride > 1 the
    // workgroups spread over cstride copies (same-address atomics from every packet, one per
    // wavefront, cost about 1 ms per build at 5D N=100k)
    unsigned long long* __restrict__ ctr = ctr0 + (cstride > 1 ? blockIdx.x % (unsigned)cstride : 0u);
    // MODE 0: the staged leaf (FP64).  MODE 1: ray K-vectors of the current node (MFMA B
    // operand source, 80-byte rows).  MODE 2: the staged leaf (FP32 rows relative to the
    // group origin).  MODE 1 and 2: the sorted ids of each ray's edge generators (the
    // candidates it already holds) and the group origin.
    __shared__ double leafx[D + 1][MODE == 0 ? 64 : 1];
    __shared__ hv_u32x4 rk4[MODE == 1 ? 64 : 1][5];
    __shared__ float4 leaf32[MODE == 2 ? 32 : 1][2];
    __shared__ int kid[MODE >= 1 ? 64 : 1][D];
    __shared__ double nog[MODE >= 1 ? 8 : 1];
    const int lane = threadIdx.x & 63;
    // nrays_dev (when given): the ray count in device memory, nrays then only bounds the grid
    if (nrays_dev) nrays = min(nrays, *nrays_dev);
    int start, cnt, mk = 0;
    if constexpr (B) {
        const int pk = hv_xcd_block_first(pk_off[F.nmesh]);
        if (pk < 0) return;   // past the last packet
        int lo = 0, hi = F.nmesh;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pk_off[mid] <= pk) lo = mid;
            else hi = mid;
        }
        mk = __builtin_amdgcn_readfirstlane(lo);
        start = seg[mk] + (pk - pk_off[mk]) * ppw;
        cnt = min(ppw, seg[mk + 1] - start);
    } else {
        const int pk = hv_xcd_block_first((nrays + ppw - 1) / ppw);
        if (pk < 0) return;   // the whole packet is past the last ray
        start = pk * ppw;
        cnt = min(ppw, nrays - start);
    }
    const DevGeo& g = hv_geo_mesh<B>(g0, F, mk);
    const int rid = start + lane;
    const bool act = lane < cnt;
    const int ray = act ? order[rid] : order[start];
    const int N = g.N;
    RayGeo<D> G;
    hv_load_ray<D>(g, R, ray, vkey, vx, G);
    Top2 T;
    T.t1 = R.t1[ray];
    T.t2 = R.t2[ray];
    T.i1 = R.i1[ray];
    T.i2 = R.i2[ray];
    // certificate ball B(tau) (centre r + tau u) and the membership form
    // f(x) = |x|^2 - 2<x, bc> + fC; bcf / R2f are an FP32 centre and a radius^2 widened so
    // that the FP32 box test never drops a box the FP64 ball reaches.  MODE 1 keeps no FP64
    // centre: it is formed again where needed (registers bound the occupancy).
    const int e0 = G.e[0];
    double bc[MODE == 0 ? D : 1], tball, R2, fC, feps;
    float bcf[D], R2f;
    const double slackf = 1e-6 * (g.xmax + 1.0);
    auto set_ball = [&](double tau) {
        tball = tau;
        R2 = 0.0;
        double pp = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double b = fma(tau, G.u[k], G.r[k]);
            if constexpr (MODE == 0) bc[k] = b;
            const double t = b - g.xo[(int64_t)e0 * D + k];
            R2 = fma(t, t, R2);
            pp = fma(b, b, pp);
            bcf[k] = (float)b;
        }
        fC = pp - R2;
        // rounding of the expanded form, with room: |x| <= xmax for every generator
        const double sc = sqrt(pp) + g.xmax;
        feps = 64.0 * 2.220446049250313e-16 * sc * sc;
        R2 = R2 * (1.0 + 1e-9) + 1e-300;
        const double rr = sqrt(R2) + slackf;
        R2f = (float)(rr * rr * (1.0 + 1e-5));
    };
    // a ray with nothing ahead (unbounded domains: its start searched the whole half-space)
    // keeps an empty ball at its own vertex and ends without one
    set_ball(T.t1 < 1e299 ? T.t1 + HV_W : 0.0);
    // packet centre: mean ball centre over the active lanes (orders the children)
    float pcf[D];
    {
        const double wgt = act ? 1.0 : 0.0;
        double cnt = wgt, pc[D];
#pragma unroll
        for (int k = 0; k < D; ++k) pc[k] = wgt * (double)bcf[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < D; ++k) pc[k] += __shfl_xor(pc[k], off);
            cnt += __shfl_xor(cnt, off);
        }
        const double ic = cnt > 0 ? 1.0 / cnt : 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k)
            pcf[k] = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint((float)(pc[k] * ic))));
    }
    // packet sphere about pcf holding every lane's starting ball (FP32, padded for rounding).
    // Valid candidates lie ahead of their edge, where later (smaller) balls are nested in
    // the first one, so a child box beyond this sphere holds no candidate of any lane.
    float rp2f;
    {
        float dd = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float a = bcf[k] - pcf[k];
            dd = fmaf(a, a, dd);
        }
        float rl = act ? sqrtf(dd) + sqrtf(R2f) : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) rl = fmaxf(rl, __shfl_xor(rl, off));
        rl = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(rl))) * (1.0f + 1e-5f) + 1e-30f;
        rp2f = rl * rl * (1.0f + 1e-5f);
    }
    // per-packet counts in 32 bits: at most 2^25 leaves (N < 2^30) and 8 tiles per leaf
    unsigned long long pairs = 0, rechecks = 0;
    uint32_t leaves = 0, tiles = 0;
    bool flagged = false;
    // candidates the ray already holds: its edge generators (sorted ids) and the best two
    if constexpr (MODE >= 1) {
#pragma unroll
        for (int z = 0; z < D; ++z) kid[lane][z] = G.e[z] < N ? g.rank[G.e[z]] : -1;
    }
    if constexpr (MODE == 1) rk4[lane][3] = hv_u32x4{0u, 0u, 0u, 0u};   // K-slots 24..31
    auto leaf_known = [&](int leaf) -> uint32_t {
        uint32_t km = 0u;
#pragma unroll
        for (int z = 0; z < D + 2; ++z) {
            const int id = z < D ? kid[lane][z] : (z == D ? T.i1 : T.i2);
            km |= (id >= 0 && (id >> 5) == leaf) ? (1u << (id & 31)) : 0u;
        }
        return km;
    };
    // box of child c of the staged node (FP32), distance^2 to a point
    auto child_d2f = [&](int c, const float* p) {
        float d2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float a = p[k] - __builtin_amdgcn_fmed3f(p[k], cbx[c * 2 * D + k], cbx[c * 2 * D + D + k]);
            d2 = fmaf(a, a, d2);
        }
        return d2;
    };
    // children of the staged node ordered by distance to the packet centre (ties: lower
    // index first): rank r holds child (ord >> 4r) & 7.  Distances are non-negative floats
    // (or +inf), which order as unsigned integers.
    uint32_t key[8];
    uint32_t ord = 0u;
    // child mask -> rank mask
    auto to_rank = [&](unsigned cm) {
        unsigned rr = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) rr |= ((cm >> ((ord >> (4 * q)) & 7u)) & 1u) << q;
        return rr;
    };
    auto load_leaf = [&](int leaf, double* pf) {
        const int j0 = leaf * g.L, npts = max(0, min(g.L, g.n - j0));
        if (lane < npts) {
#pragma unroll
            for (int k = 0; k <= D; ++k) pf[k] = g.xs[(int64_t)(j0 + lane) * (D + 1) + k];
        }
    };
    // The child boxes of the node walked next (and, MODE 1, the origin of its leaf group)
    // are loaded into registers while the current node is worked on.
    float cbr0 = 0.f, cbr1 = 0.f;
    double lorr = 0.0;
    int pref = -1;
    auto prefetch = [&](int node, int lvl) {
        const float* src = g.cb + (int64_t)node * (16 * D);
        if (lane < 16 * D) cbr0 = src[lane];
        if (lane < 16 * D - 64) cbr1 = src[lane + 64];
        if constexpr (MODE >= 1) {
            if (lvl + 1 == g.h && lane < 8) lorr = g.lorg[(int64_t)(node - ((1 << (3 * lvl)) - 1) / 7) * 8 + lane];   // index in its level
        }
        pref = node;
    };
    int sp = 1;
    stk[0] = 0;
    stkl[0] = 0;
    prefetch(0, 0);
    while (sp > 0) {
        --sp;
        hv_wave_sync();
        const int gid = __builtin_amdgcn_readfirstlane(stk[sp]);
        const int l = __builtin_amdgcn_readfirstlane(stkl[sp]);
        if (gid != pref) prefetch(gid, l);
        // stage the 8 child boxes of the node (one contiguous block)
        if (lane < 16 * D) cbx[lane] = cbr0;
        if (lane < 16 * D - 64) cbx[lane + 64] = cbr1;
        if constexpr (MODE >= 1) {
            if (l + 1 == g.h && lane < 8) nog[lane] = lorr;
        }
        // a node whose children are leaves pushes nothing: the stack top is walked next
        if (l + 1 == g.h && sp > 0) {
            const int nx = __builtin_amdgcn_readfirstlane(stk[sp - 1]);
            const int nl = __builtin_amdgcn_readfirstlane(stkl[sp - 1]);
            prefetch(nx, nl);
        }
        hv_wave_sync();
        // child distances to the packet centre (order, and the packet-sphere prefilter), then
        // the children each lane's ball reaches (kept per lane) among those the sphere reaches
        unsigned near = 0u;
        {
            const uint32_t kb = __float_as_uint(child_d2f(lane & 7, pcf));
            const int me = lane & 7;
            int rk = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                key[c] = __builtin_amdgcn_readlane(kb, c);
                near |= (key[c] <= __float_as_uint(rp2f)) ? (1u << c) : 0u;   // non-negative floats
                rk += (key[c] < kb || (key[c] == kb && c < me)) ? 1 : 0;
            }
            uint32_t o = 0u;
#pragma unroll
            for (int c = 0; c < 8; ++c) o |= (uint32_t)c << (4 * __builtin_amdgcn_readlane(rk, c));
            ord = o;
        }
        unsigned rem = 0u, mine = 0u;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!((near >> c) & 1u)) continue;
            const bool h = (child_d2f(c, bcf) <= R2f) & act;
            mine |= h ? (1u << c) : 0u;
            rem |= (__ballot(h) != 0ull) ? (1u << c) : 0u;
        }
        const double tnode = tball;   // a ball that shrinks later in the node is re-tested
        rem = __builtin_amdgcn_readfirstlane(rem);
        if (rem == 0u) continue;
        int off = 0, pw = 1;
        for (int q = 0; q < l; ++q) { off += pw; pw *= 8; }
        const int i = gid - off;
        if (l + 1 == g.h) {
            if constexpr (MODE == 0) {
                // children are leaves: nearest first, the next one loaded while this one is scanned
                double pf[D + 1];
                unsigned rr = to_rank(rem);
                int cc = (ord >> (4 * __builtin_ctz(rr))) & 7u;
                rr &= rr - 1u;
                load_leaf(8 * i + cc, pf);
                while (cc >= 0) {
                    const int cn = rr ? (int)((ord >> (4 * __builtin_ctz(rr))) & 7u) : -1;
                    rr &= rr - 1u;
                    const int leaf = 8 * i + cc;
                    const int j0 = leaf * g.L, npts = max(0, min(g.L, g.n - j0));
                    // balls only shrink: re-test with the current one
                    const bool hit = act && ((mine >> cc) & 1u) && (tball == tnode || child_d2f(cc, bcf) <= R2f);
                    const unsigned long long m = __ballot(hit);
                    if (m && lane < npts) {
#pragma unroll
                        for (int k = 0; k <= D; ++k) leafx[k][lane] = pf[k];
                    }
                    if (cn >= 0) load_leaf(8 * i + cn, pf);
                    if (m) {
                        hv_wave_sync();
                        leaves += 1;
                        pairs += (unsigned long long)__popcll(m) * (unsigned long long)npts;
                        if (hit) {
                            const double before = T.t1;
                            // membership f = |x|^2 + fC - 2<x, bc>, four independent chains
                            int q = 0;
                            for (; q + 4 <= npts; q += 4) {
                                double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
#pragma unroll
                                for (int k = 0; k < D; ++k) {
                                    f0 = fma(leafx[k][q], bc[k], f0);
                                    f1 = fma(leafx[k][q + 1], bc[k], f1);
                                    f2 = fma(leafx[k][q + 2], bc[k], f2);
                                    f3 = fma(leafx[k][q + 3], bc[k], f3);
                                }
                                f0 = fma(-2.0, f0, leafx[D][q] + fC);
                                f1 = fma(-2.0, f1, leafx[D][q + 1] + fC);
                                f2 = fma(-2.0, f2, leafx[D][q + 2] + fC);
                                f3 = fma(-2.0, f3, leafx[D][q + 3] + fC);
                                if (fmin(fmin(f0, f1), fmin(f2, f3)) < feps) {
                                    const double fs[4] = {f0, f1, f2, f3};
#pragma unroll
                                    for (int z = 0; z < 4; ++z) {
                                        if (fs[z] < feps) {
                                            double xk[D];
#pragma unroll
                                            for (int k = 0; k < D; ++k) xk[k] = leafx[k][q + z];
                                            hv_consider<D>(G, T, xk, leafx[D][q + z], j0 + q + z, plane_tol);
                                        }
                                    }
                                }
                            }
                            for (; q < npts; ++q) {
                                double f = 0.0;
#pragma unroll
                                for (int k = 0; k < D; ++k) f = fma(leafx[k][q], bc[k], f);
                                f = fma(-2.0, f, leafx[D][q] + fC);
                                if (f < feps) {
                                    double xk[D];
#pragma unroll
                                    for (int k = 0; k < D; ++k) xk[k] = leafx[k][q];
                                    hv_consider<D>(G, T, xk, leafx[D][q], j0 + q, plane_tol);
                                }
                            }
                            if (T.t1 < before) set_ball(T.t1 + HV_W);
                        }
                        hv_wave_sync();
                    }
                    cc = cn;
                }
            } else if constexpr (MODE == 2) {
                // Children are leaves of 32 points, staged in LDS as FP32 rows relative to the
                // group origin o.  Per ray, once per node: b = bc - o (FP64, then FP32) and the
                // test |x-o|^2 - 2<x-o, b> + (|b|^2 - R2) < thr in FP32, where thr bounds the
                // FP32 rounding of every term (relative coordinates: all terms are of the size
                // of the group and the ball) plus the FP64 filter's own slack 2 feps, so every
                // point the FP64 scan would consider survives.  A ball that shrinks inside the
                // node keeps the older (larger) form.
                float m2b[D], cf, thrf;
                {
                    double bb = 0.0;
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double b = fma(tball, G.u[k], G.r[k]) - nog[k];
                        bb = fma(b, b, bb);
                        m2b[k] = (float)(-2.0 * b);
                    }
                    const double rad = nog[7], C = bb - R2;
                    cf = (float)C;
                    const double bound = 2.0 * (D + 4) * 5.9604644775390625e-08 *
                                             (rad * rad + fabs(C) + 2.0 * rad * sqrt(bb)) +
                                         2.0 * feps + 1e-13 * (bb + R2 + rad * rad);
                    thrf = __double2float_ru(bound);
                }
                float4 pa = make_float4(0.f, 0.f, 0.f, 0.f), pb = pa;
                auto load_leaf2 = [&](int leaf, float4& a, float4& b) {
                    if (lane < 32) {
                        const float4* src = (const float4*)(g.pxf) + ((int64_t)leaf * 32 + lane) * 2;
                        a = src[0];
                        b = src[1];
                    }
                };
                unsigned rr = to_rank(rem);
                int cc = (ord >> (4 * __builtin_ctz(rr))) & 7u;
                rr &= rr - 1u;
                load_leaf2(8 * i + cc, pa, pb);
                while (cc >= 0) {
                    const int cn = rr ? (int)((ord >> (4 * __builtin_ctz(rr))) & 7u) : -1;
                    rr &= rr - 1u;
                    const int leaf = 8 * i + cc;
                    const int j0 = leaf * 32, npts = max(0, min(32, g.n - j0));
                    const bool hit = act && ((mine >> cc) & 1u) && (tball == tnode || child_d2f(cc, bcf) <= R2f);
                    const unsigned long long m = __ballot(hit);
                    if (m && lane < 32) {
                        leaf32[lane][0] = pa;
                        leaf32[lane][1] = pb;
                    }
                    if (cn >= 0) load_leaf2(8 * i + cn, pa, pb);
                    if (m) {
                        hv_wave_sync();
                        leaves += 1;
                        pairs += (unsigned long long)__popcll(m) * (unsigned long long)npts;
                        if (hit) {
                            uint32_t surv = 0u;
#pragma unroll 4
                            for (int q = 0; q < 32; ++q) {
                                const float4 a = leaf32[q][0], b = leaf32[q][1];
                                const float xr[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                                float f = xr[D] + cf;
#pragma unroll
                                for (int k = 0; k < D; ++k) f = fmaf(xr[k], m2b[k], f);
                                surv |= (f < thrf) ? (1u << q) : 0u;
                            }
                            surv &= (npts >= 32 ? 0xffffffffu : ((1u << npts) - 1u)) & ~leaf_known(leaf);
                            if (surv) {
                                // survivors that the ray does not already hold, in FP64
                                const double before = T.t1;
                                flagged = true;
                                while (surv) {
                                    const int q = __builtin_ctz(surv);
                                    surv &= surv - 1u;
                                    const int j = j0 + q;
                                    double xk[D];
#pragma unroll
                                    for (int k = 0; k < D; ++k) xk[k] = g.xs[(int64_t)j * (D + 1) + k];
                                    hv_consider<D>(G, T, xk, g.xs[(int64_t)j * (D + 1) + D], j, plane_tol);
                                    ++rechecks;
                                }
                                if (T.t1 < before) set_ball(T.t1 + HV_W);
                            }
                        }
                        hv_wave_sync();
                    }
                    cc = cn;
                }
            } else {
                // Children are leaves of 32 points (two 16-point tiles).  Each ray's K-vector
                // is formed once per node, relative to the node origin that the point vectors
                // were built with; the B operands of the four 16-ray tiles stay in registers.
                // A ball that shrinks inside the node keeps the older (larger) vector, which
                // only lets more points through the filter.
                hv_u32x4 pa0, pa1;
                auto load_leaf1 = [&](int leaf, hv_u32x4& a0, hv_u32x4& a1) {
                    const hv_u32x4* base = (const hv_u32x4*)(g.pkv) + ((int64_t)leaf * 32 + (lane & 15)) * 4 + (lane >> 4);
                    a0 = base[0];
                    a1 = base[16 * 4];
                };
                unsigned rr = to_rank(rem);
                int cc = (ord >> (4 * __builtin_ctz(rr))) & 7u;
                rr &= rr - 1u;
                load_leaf1(8 * i + cc, pa0, pa1);
                {
                    double no[D], cb_[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        no[k] = nog[k];
                        cb_[k] = fma(tball, G.u[k], G.r[k]);
                    }
                    hv_ray_kvec<D>(cb_, R2, feps, no, nog[7], rk4[lane]);
                }
                hv_wave_sync();
                while (cc >= 0) {
                    const int cn = rr ? (int)((ord >> (4 * __builtin_ctz(rr))) & 7u) : -1;
                    rr &= rr - 1u;
                    const int leaf = 8 * i + cc;
                    const int j0 = leaf * 32, npts = max(0, min(32, g.n - j0));
                    const bool hit = act && ((mine >> cc) & 1u) && (tball == tnode || child_d2f(cc, bcf) <= R2f);
                    const unsigned long long m = __ballot(hit);
                    const hv_u32x4 ca0 = pa0, ca1 = pa1;
                    if (cn >= 0) load_leaf1(8 * i + cn, pa0, pa1);
                    if (m) {
                        leaves += 1;
                        // survivors of this leaf for the lane's own ray (bit q = point j0 + q)
                        uint32_t sv = 0u;
#pragma unroll
                        for (int g4 = 0; g4 < 4; ++g4) {
                            if (((m >> (16 * g4)) & 0xffffull) == 0ull) continue;
                            // both tiles issued before either result is read, so their latencies
                            // overlap (the second tile of a leaf with at most 16 points reads the
                            // padding records, whose form is positive for every ray)
                            const hv_u32x4 bop = rk4[16 * g4 + (lane & 15)][lane >> 4];
                            const hv_f32x4 acc0 = hv_mfma_bf16(ca0, bop);
                            const hv_f32x4 acc1 = hv_mfma_bf16(ca1, bop);
#pragma unroll
                            for (int t = 0; t < 2; ++t) {
                                if (16 * t >= npts) continue;
                                ++tiles;
                                const hv_f32x4 acc = t == 0 ? acc0 : acc1;
                                // lane holds points 16t + 4*(lane/16) + e of ray 16*g4 + lane%16
                                const uint32_t ob = __float_as_uint(acc[0]) | __float_as_uint(acc[1]) |
                                                    __float_as_uint(acc[2]) | __float_as_uint(acc[3]);
                                if (__ballot((int)ob < 0) == 0ull) continue;
                                uint32_t nib = 0u;
#pragma unroll
                                for (int e = 0; e < 4; ++e) nib |= ((int)__float_as_uint(acc[e]) < 0) ? (1u << e) : 0u;
                                // the ray's lane 16*g4 + c gathers the nibbles of lanes c, c+16, c+32, c+48
                                uint32_t col = 0u;
#pragma unroll
                                for (int q = 0; q < 4; ++q)
                                    col |= (uint32_t)__shfl((int)nib, (lane & 15) + 16 * q) << (4 * q);
                                if ((lane >> 4) == g4) sv |= col << (16 * t);
                            }
                        }
                        // survivors the ray does not already hold go to FP64, every ray at once
                        if (sv) sv &= ~leaf_known(leaf) & (npts >= 32 ? 0xffffffffu : ((1u << npts) - 1u));
                        if (__ballot(sv != 0u)) {
                            if (sv) {
                                const double before = T.t1;
                                flagged = true;
                                while (sv) {
                                    const int q = __builtin_ctz(sv);
                                    sv &= sv - 1u;
                                    const int j = j0 + q;
                                    double xk[D];
#pragma unroll
                                    for (int k = 0; k < D; ++k) xk[k] = g.xs[(int64_t)j * (D + 1) + k];
                                    hv_consider<D>(G, T, xk, g.xs[(int64_t)j * (D + 1) + D], j, plane_tol);
                                    ++rechecks;
                                }
                                if (T.t1 < before) set_ball(T.t1 + HV_W);
                            }
                        }
                    }
                    cc = cn;
                }
            }
        } else {
            // internal children: farthest pushed first so the nearest is walked next
            const int cbase = off + pw + 8 * i;
            int last = -1;
            unsigned rr = to_rank(rem);
            while (rr) {
                const int q = 31 - __builtin_clz(rr);
                rr &= ~(1u << q);
                const int c = (ord >> (4 * q)) & 7u;
                stk[sp] = cbase + c;
                stkl[sp] = l + 1;
                ++sp;
                last = cbase + c;
            }
            prefetch(last, l + 1);
        }
    }
    if (lane == 0) {
        // BF16: every MFMA tile forms 16 x 16 (ray, point) pairs
        atomicAdd(&ctr[0], MODE == 1 ? 256ull * tiles : pairs);
        atomicAdd(&ctr[cstride], (unsigned long long)leaves);
    }
    if constexpr (MODE >= 1) {
        // FP64 re-verifications and the rays that needed any
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) rechecks += __shfl_xor(rechecks, off);
        if (lane == 0 && rechecks) atomicAdd(&ctr[6 * cstride], rechecks);
        hv_wave_count(&ctr[7 * cstride], flagged && act);
    }
    bool tie = false;
    if (act) {
        if (T.i1 < 0) {
            status[rid] = HV_ST_NONE;
        } else {
            // original indices and the reference's form of t (src/raycast.jl:424-431)
            double x0t[D];
#pragma unroll
            for (int k = 0; k < D; ++k) x0t[k] = g.xo[(int64_t)e0 * D + k];
            auto orig = [&](int i) { return i < N ? g.sidx[i] : i; };
            auto t_ref = [&](int w, double tscan, double& score, double& cosang) -> double {
                if (w >= N) {
                    const int p = w - N;
                    double nu = 0.0, dist = g.ph[p];
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        nu = fma(g.pn[p * D + k], G.u[k], nu);
                        dist = fma(-g.pn[p * D + k], x0t[k], dist);
                    }
                    score = 2.0 * dist * nu;
                    cosang = nu;
                    return tscan;
                }
                double num = 0.0, den = 0.0, dd = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double xk = g.xo[(int64_t)(w) * D + k];
                    const double dx = xk - x0t[k];
                    num = fma(dx, x0t[k] + xk - 2.0 * G.r[k], num);
                    den = fma(G.u[k], dx, den);
                    dd = fma(dx, dx, dd);
                }
                score = den;
                cosang = den / sqrt(dd);
                return num / (2.0 * den);
            };
            int w1 = orig(T.i1);
            double s1, c1;
            double tw = t_ref(w1, T.t1, s1, c1);
            int st = HV_ST_FOUND;
            // the nearest candidate closer than plane_tol: the reference's window is empty, no
            // vertex (src/raycast.jl:883-894 drops it and upper_t excludes every other)
            if (!(tw >= plane_tol)) st = HV_ST_NONE;
            else if (T.i2 >= 0) {
                const int w2 = orig(T.i2);
                double s2, c2;
                const double tw2 = t_ref(w2, T.t2, s2, c2);
                double rn = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) rn = fma(G.r[k], G.r[k], rn);
                const double full_error = (tw * R.du[ray] + sqrt(rn) * 1e-15) / c1;
                const double window = fmin(1e-7, (10.0 + D) * full_error);
                double to = tw2, so = s2;   // the candidate that does not win
                if (tw2 < tw || tw2 <= tw + window) {
                    // near tie: the larger <u, x - x0> wins (src/raycast.jl:903-916)
                    st = HV_ST_TIE;
                    tie = true;
                    if (s2 > s1) { to = tw; so = s1; w1 = w2; tw = tw2; }
                }
                // The reference's key takes every candidate whose distance from the new vertex is
                // within rounding of the key's own (src/raycast.jl:926-950): the other candidate
                // inside the chosen sphere (to < tw), or on it -- |x_o - r|^2 - R^2 = 2 (to - tw)
                // <u, x_o - x0> within a few ulps of R^2 -- makes a key longer than d + 1, a
                // vertex out of general position, which the device walk reports instead.
                double R2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double a = x0t[k] - fma(tw, G.u[k], G.r[k]);
                    R2 = fma(a, a, R2);
                }
                if (to < tw || (to - tw) * so <= 16.0 * 2.220446049250313e-16 * R2) st = HV_ST_SINGULAR;
            }
            win[rid] = w1;
            status[rid] = st;
            tout[rid] = tw;
        }
    }
    hv_wave_count(&ctr[2 * cstride], tie);
}

// New key = edge + winner (sorted) and its circumcentre, written to the provisional slot.
// Packets of at most ppw_max rays (64 = a wavefront), halved (down to 8) while the level would
// give fewer than min_packets packets: a packet's walk is serial, and a few long walks leave
// the chip idle.  Both are tile sizes of the autotuner (hv_tune; defaults 64 and 2048).
// one interior seed vertex per this many generators (Morton order)
constexpr int HV_SEED_STRIDE = 8;          // large point sets: one descent per 8 generators
constexpr int HV_SEED_ALL_MAX = 1 << 17;   // up to this many generators, and in 2D, every one
static inline int hv_packet_rays(int64_t n, const HvTiles& t = HvTiles{}) {
    int p = t.ppw_max;
    while (p > t.ppw_min && n < (int64_t)p * t.min_packets) p >>= 1;
    return p;
}

// ppw rays per packet in nblocks workgroups (over a forest: the packet plan pk_off / seg)
template <int D, bool B>
static void launch_scan(int precision, const DevGeo& g, const GeoSet& F, const int* pk_off, const int* seg, int ppw,
                        unsigned nblocks, const DevRays& R, const int* list, int n, const int* n_dev,
                        const int* vkey, const double* vx, int* status, int* win, double* tout, double plane_tol,
                        unsigned long long* ctr, hipStream_t st, int cstride = 1) {
    if (precision == HV_PREC_BF16)
        k_raycast64<D, 1, B><<<nblocks, 64, 0, st>>>(g, F, pk_off, seg, R, list, n, n_dev, ppw, vkey, vx, status, win, tout,
                                                    plane_tol, ctr, cstride);
    else if (precision == HV_PREC_FP32)
        k_raycast64<D, 2, B><<<nblocks, 64, 0, st>>>(g, F, pk_off, seg, R, list, n, n_dev, ppw, vkey, vx, status, win, tout,
                                                    plane_tol, ctr, cstride);
    else
        k_raycast64<D, 0, B><<<nblocks, 64, 0, st>>>(g, F, pk_off, seg, R, list, n, n_dev, ppw, vkey, vx, status, win, tout,
                                                    plane_tol, ctr, cstride);
}

template <int D>
static void launch_scan1(int precision, const DevGeo& g, const DevRays& R, const int* list, int n, const int* n_dev,
                         const int* vkey, const double* vx, int* status, int* win, double* tout, double plane_tol,
                         unsigned long long* ctr, hipStream_t st, const HvTiles& tiles = HvTiles{}, int cstride = 1) {
    const int ppw = hv_packet_rays(n, tiles);
    launch_scan<D, false>(precision, g, GeoSet{}, nullptr, nullptr, ppw, nblk(n, ppw), R, list, n, n_dev, vkey, vx, status,
                          win, tout, plane_tol, ctr, st, cstride);
}

template <int D, bool B>
__global__ __launch_bounds__(256) void k_finalize(DevGeo g0, GeoSet F, DevRays R, const int* __restrict__ order, int nrays,
                           const int* __restrict__ nrays_dev, int* __restrict__ vkey, double* __restrict__ vx,
                           int64_t base, int* __restrict__ status, const int* __restrict__ win) {
    if (nrays_dev) nrays = min(nrays, *nrays_dev);
    const int blk = hv_xcd_block_first((nrays + (int)blockDim.x - 1) / (int)blockDim.x);
    if (blk < 0) return;
    const int rid = blk * blockDim.x + threadIdx.x;
    if (rid >= nrays) return;
    if (status[rid] <= 0) return;
    const int ray = order[rid];
    const int v = R.v[ray], pos = R.drop[ray];
    // key read straight from global memory: a private copy indexed by `pos` would live in scratch
    const int* key = vkey + (int64_t)v * (D + 1);
    const DevGeo& g = hv_geo<B>(g0, F, key[0]);
    int e[D];
    hv_edge_of<D>(key, pos, e);
    const int w1 = win[rid];
    int ins = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) ins += (e[j] < w1) ? 1 : 0;
    int nk[D + 1];
#pragma unroll
    for (int j = 0; j <= D; ++j) nk[j] = (j < ins) ? e[j] : (j == ins ? w1 : e[j > 0 ? j - 1 : 0]);
    double xn[D];
    if (!hv_vertex_coords<D>(g, nk, xn)) { status[rid] = HV_ST_SINGULAR; return; }
    const int64_t id = base + rid;
#pragma unroll
    for (int j = 0; j <= D; ++j) vkey[id * (D + 1) + j] = nk[j];
#pragma unroll
    for (int k = 0; k < D; ++k) vx[id * D + k] = xn[k];
}

__global__ void k_iota(int* __restrict__ a, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// ---------------------------------------------------------------------------------
// host driver

constexpr int HV_DESC_LANES = 4;   // lanes per interior seed (k_seed_descent)
// The descent and its vertex solve take about 200 VGPRs in 5D (2 wavefronts per SIMD, 6 250
// wavefronts for 100k seeds): held to 128 registers (4 per SIMD) with some spills it runs
// faster (5D N=100k mixture 56.9 -> 56.2 ms per build, 512 meshes 356 -> 350 ms).
#define HV_SEED_OCC __attribute__((amdgpu_waves_per_eu(4, 8)))
template <int D, bool B>
__global__ __launch_bounds__(64) HV_SEED_OCC void k_seed_descent(DevGeo g0, GeoSet F, int nseed, int stride, int nwalk,
                                                     const int* __restrict__ cells, int* __restrict__ vkey,
                                                     double* __restrict__ vx, int64_t base, int* __restrict__ status,
                                                     uint64_t seed);
template <int D>
__global__ __launch_bounds__(256) void k_seed_exact(DevGeo g, const int* __restrict__ cells, int ncell,
                                                    int* __restrict__ vkey, double* __restrict__ vx, int64_t base,
                                                    int* __restrict__ status);
template <int D>
__global__ __launch_bounds__(256) void k_mark_cells(const int* __restrict__ ids, int nv, const int* __restrict__ vkey,
                                                    int nreal, int* __restrict__ mark);

namespace {

struct Tree {
    DevGeo g{};
};

template <typename T>
T* pget(hv_ctx* c, const char* name, size_t count, bool keep = false) {
    return (T*)c->pool->get(name, count * sizeof(T), keep, c->stream);
}

// Points -> SoA + Morton order + leaf boxes + complete binary tree of boxes.  tables: the
// per-point scan tables to build besides the tree (HV_PREC_BF16: BF16 K-vectors,
// HV_PREC_FP32: FP32 relative rows; both with the leaf-group origins).
// Where the tables of one tree go: the context's pool tables (one point set), or one mesh's
// segments of a forest's tables (batched meshes walked together).  Generator ids are global
// over a forest: the mesh's points are ids goff .. goff + n - 1, row goff + i of xo_base.
struct TreeOut {
    const double* xin = nullptr;   // the n points, original order
    int n = 0, goff = 0;
    int Nid = 0;                   // generator ids are below Nid; plane p of this tree is Nid + p
    int P = 0;
    const double* pnormal = nullptr;   // host planes (P x D)
    const double* pbase = nullptr;
    double *pn = nullptr, *ph = nullptr;
    double* xo_base = nullptr;
    double* xs = nullptr;
    int* sidx = nullptr;
    int* rank_base = nullptr;      // indexed by global id
    int* knn = nullptr;
    float* xf = nullptr;
    double *blo = nullptr, *bhi = nullptr;
    float* cb = nullptr;
    double* lorg = nullptr;
    hv_u32x4* pkv = nullptr;
    float* pxf = nullptr;
    int* bad = nullptr;            // points outside the planes are counted here
    bool planes_ready = false;     // pn / ph already hold the planes
    const double* dom_lo = nullptr;    // a general polytope's bounding box (null: axis planes)
    const double* dom_hi = nullptr;
    bool unbounded = false;        // no bounded domain (dom_lo / dom_hi: the generators' box)
    bool defer_tables = false;     // leave the 16-NN lists and the scan tables to a forest-wide pass
};

// Table sizes of a tree over n points in leaves of L
struct TreeDims {
    int nleaf, P2, NN, h;
    int64_t ninternal;
};
static TreeDims tree_dims(int n, int L) {
    TreeDims t{};
    t.nleaf = (n + L - 1) / L;
    t.P2 = 1;
    while (t.P2 < t.nleaf) t.P2 <<= 1;
    t.NN = 2 * t.P2;
    t.h = 1;
    int64_t cap = 8;
    while (cap < t.nleaf) { cap *= 8; ++t.h; }
    int64_t pw = 1;
    t.ninternal = 0;
    for (int l = 0; l < t.h; ++l) { t.ninternal += pw; pw *= 8; }
    return t;
}

// Points -> Morton order + leaf boxes + complete binary tree of boxes + 8-ary child-box tree +
// 16-NN lists; tables: the per-point scan tables to build besides (HV_PREC_BF16: BF16
// K-vectors, HV_PREC_FP32: FP32 relative rows; both with the leaf-group origins).  Launches
// only (the domain check is left in *o.bad).
template <int D>
int build_tree_core(hv_ctx* c, int L, int tables, const TreeOut& o, Tree& T, hipStream_t st = nullptr, int slot = -1) {
    // st: the stream to build on (the context's by default); slot >= 0: temporaries of that
    // slot, so that trees of a forest can be built on several streams at once
    if (!st) st = c->stream;
    const std::string sfx = slot >= 0 ? "#" + std::to_string(slot) : std::string();
    const int N = o.n;
    DevGeo& g = T.g;
    memset(&g, 0, sizeof g);
    g.N = o.Nid;
    g.n = N;
    g.P = o.P;
    g.L = L;
    const TreeDims td = tree_dims(N, L);
    g.nleaf = td.nleaf;
    g.P2 = td.P2;
    g.NN = td.NN;
    // domain box: a general polytope's bounding box, else from the axis-aligned planes
    for (int k = 0; k < D; ++k) { g.dlo[k] = 0.0; g.dhi[k] = 1.0; }
    if (o.dom_lo) {
        for (int k = 0; k < D; ++k) { g.dlo[k] = o.dom_lo[k]; g.dhi[k] = o.dom_hi[k]; }
    } else {
        for (int p = 0; p < o.P; ++p)
            for (int k = 0; k < D; ++k) {
                double nk = o.pnormal[p * D + k];
                if (nk > 0.5) g.dhi[k] = o.pbase[p * D + k];
                if (nk < -0.5) g.dlo[k] = o.pbase[p * D + k];
            }
    }
    g.unbounded = o.unbounded ? 1 : 0;
    {
        double s = 0.0;
        for (int k = 0; k < D; ++k) {
            const double a = std::max(std::fabs(g.dlo[k]), std::fabs(g.dhi[k]));
            s += a * a;
        }
        g.xmax = std::sqrt(s);
    }
    std::vector<double> hph(std::max(1, o.P));
    for (int p = 0; p < o.P; ++p) {
        double s = 0;
        for (int k = 0; k < D; ++k) s += o.pnormal[p * D + k] * o.pbase[p * D + k];
        hph[p] = s;
    }
    if (o.P && !o.planes_ready) {
        HVCK(hipMemcpyAsync(o.pn, o.pnormal, sizeof(double) * o.P * D, hipMemcpyHostToDevice, st));
        HVCK(hipMemcpyAsync(o.ph, hph.data(), sizeof(double) * o.P, hipMemcpyHostToDevice, st));
        HVCK(hipStreamSynchronize(st));   // hph is a local
    }
    g.pn = o.pn;
    g.ph = o.ph;
    uint64_t* code = pget<uint64_t>(c, ("code" + sfx).c_str(), N);
    uint64_t* code2 = pget<uint64_t>(c, ("code2" + sfx).c_str(), N);
    int* idx = pget<int>(c, ("idx" + sfx).c_str(), N);
    int* idx2 = pget<int>(c, ("idx2" + sfx).c_str(), N);
    if (!code || !code2 || !idx || !idx2) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (tree)");
    double* xo = o.xo_base + (int64_t)o.goff * D;
    k_points_in<D><<<nblk(N, 256), 256, 0, st>>>(o.xin, N, g, xo, code, idx, o.bad);
    HVCK(hipGetLastError());
    size_t tmpb = 0;
    HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpb, code, code2, idx, idx2, N, 0, 64, st));
    void* tmp = c->pool->get("cubtmp" + sfx, tmpb, false, st);
    if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
    HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tmpb, code, code2, idx, idx2, N, 0, 64, st));
    k_points_sorted<D><<<nblk(N, 256), 256, 0, st>>>(xo, idx2, N, o.goff, o.xs, o.sidx, o.rank_base + o.goff);
    HVCK(hipGetLastError());
    g.xo = o.xo_base;
    g.xs = o.xs;
    g.sidx = o.sidx;
    g.rank = o.rank_base;
    g.knn = o.knn;
    k_point_rows32<D><<<nblk(N, 256), 256, 0, st>>>(g, o.xf);
    HVCK(hipGetLastError());
    g.xf = o.xf;
    g.blo = o.blo;
    g.bhi = o.bhi;
    k_leaf_boxes<D><<<nblk(g.P2, 256), 256, 0, st>>>(g, o.blo, o.bhi);
    HVCK(hipGetLastError());
    for (int first = g.P2 / 2; first >= 1; first /= 2) {
        k_node_boxes<D><<<nblk(first, 256), 256, 0, st>>>(first, first, g.NN, o.blo, o.bhi);
        HVCK(hipGetLastError());
    }
    // 8-ary tree of child boxes (levels bottom-up)
    g.h = td.h;
    g.cb = o.cb;
    for (int l = g.h - 1; l >= 0; --l) {
        int nodes = 1;
        for (int q = 0; q < l; ++q) nodes *= 8;
        k_tree8_level<D><<<nblk((int64_t)nodes * 8, 256), 256, 0, st>>>(g, l, nodes, o.cb);
        HVCK(hipGetLastError());
    }
    if (tables != HV_PREC_FP64 && L != 32) return hv_fail(c, HV_ERR_UNSUPPORTED, "the mixed-precision scans use leaves of 32 points");
    if (tables != HV_PREC_FP64) g.lorg = o.lorg;
    if (tables == HV_PREC_BF16) g.pkv = (const uint32_t*)o.pkv;
    if (tables == HV_PREC_FP32) g.pxf = o.pxf;
    if (o.defer_tables) return HV_OK;
    k_knn64<D, false><<<nblk(N, 64), 64, 0, st>>>(g, GeoSet{}, nullptr, o.knn);
    HVCK(hipGetLastError());
    if (tables != HV_PREC_FP64) {
        k_leaf_origin<D, false><<<(g.nleaf + 7) / 8, 64, 0, st>>>(g, GeoSet{}, nullptr, o.lorg);
        HVCK(hipGetLastError());
        if (tables == HV_PREC_BF16) {
            k_point_kvec<D, false><<<nblk((int64_t)g.nleaf * 32, 256), 256, 0, st>>>(g, GeoSet{}, nullptr, o.lorg, o.pkv);
            HVCK(hipGetLastError());
        } else {
            k_point_f32<D, false><<<nblk((int64_t)g.nleaf * 32, 256), 256, 0, st>>>(g, GeoSet{}, nullptr, o.lorg, o.pxf);
            HVCK(hipGetLastError());
        }
    }
    return HV_OK;
}

// The tree over the context's working set, its tables in the context's pool.
template <int D>
int build_tree(hv_ctx* c, int L, Tree& T, int tables = HV_PREC_FP64) {
    const int N = (int)c->N;
    if (!c->xw) return hv_fail(c, HV_ERR_STATE, "no points set");
    const TreeDims td = tree_dims(N, L);
    TreeOut o;
    o.xin = (const double*)c->xw->p;
    o.n = N;
    o.goff = 0;
    o.Nid = N;
    o.P = c->P;
    o.pnormal = c->pnormal.data();
    o.pbase = c->pbase.data();
    {
        const char* why = nullptr;
        if (hv_domain(c, D, &why) == HV_OK && c->dom_general) {
            if (c->dom_unbounded) {   // the frame: the generators' bounding box
                unsigned long long* kb = pget<unsigned long long>(c, "bbox", 2 * D);
                if (!kb) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (tree)");
                std::vector<unsigned long long> init(2 * D);
                for (int k = 0; k < D; ++k) { init[k] = ~0ull; init[D + k] = 0ull; }
                HVCK(hipMemcpyAsync(kb, init.data(), sizeof(unsigned long long) * 2 * D, hipMemcpyHostToDevice, c->stream));
                if (N > 0) {
                    k_bbox<D><<<std::min<unsigned>(nblk(N, 256), 1024u), 256, 0, c->stream>>>((const double*)c->xw->p, N, kb);
                    HVCK(hipGetLastError());
                }
                HVCK(hipMemcpyAsync(init.data(), kb, sizeof(unsigned long long) * 2 * D, hipMemcpyDeviceToHost, c->stream));
                HVCK(hipStreamSynchronize(c->stream));
                for (int k = 0; k < D; ++k) {
                    c->dom_lo[k] = hv_key_double(init[k]);
                    c->dom_hi[k] = hv_key_double(init[D + k]);
                    if (!(c->dom_hi[k] >= c->dom_lo[k])) { c->dom_lo[k] = 0.0; c->dom_hi[k] = 1.0; }
                }
                o.unbounded = true;
            }
            o.dom_lo = c->dom_lo;
            o.dom_hi = c->dom_hi;
        }
    }
    o.pn = pget<double>(c, "pn", std::max(1, c->P * D));
    o.ph = pget<double>(c, "ph", std::max(1, c->P));
    o.xo_base = pget<double>(c, "xo", (size_t)N * D);
    o.xs = pget<double>(c, "xs", (size_t)N * (D + 1));
    o.sidx = pget<int>(c, "sidx", N);
    o.rank_base = pget<int>(c, "rank", N);
    o.knn = pget<int>(c, "knn", (size_t)N * HV_KNN);
    o.xf = pget<float>(c, "xf", (size_t)N * 8);
    o.blo = pget<double>(c, "blo", (size_t)td.NN * D);
    o.bhi = pget<double>(c, "bhi", (size_t)td.NN * D);
    o.cb = pget<float>(c, "cb8", (size_t)td.ninternal * 8 * 2 * D);
    o.bad = pget<int>(c, "bad", 1);
    if (!o.pn || !o.ph || !o.xo_base || !o.xs || !o.sidx || !o.rank_base || !o.knn || !o.xf || !o.blo || !o.bhi ||
        !o.cb || !o.bad)
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (tree)");
    if (tables != HV_PREC_FP64) {
        o.lorg = pget<double>(c, "lorg", (size_t)((td.nleaf + 7) / 8) * 8);
        if (tables == HV_PREC_BF16) o.pkv = pget<hv_u32x4>(c, "pkv", (size_t)td.nleaf * 32 * 4);
        else o.pxf = pget<float>(c, "pxf", (size_t)td.nleaf * 32 * 8);
        if (!o.lorg || (tables == HV_PREC_BF16 ? !o.pkv : !o.pxf))
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (mixed-precision tables)");
    }
    HVCK(hipMemsetAsync(o.bad, 0, sizeof(int), c->stream));
    int rc = build_tree_core<D>(c, L, tables, o, T);
    if (rc) return rc;
    int hbad = 0;
    HVCK(hipMemcpyAsync(&hbad, o.bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    if (hbad & 1) return hv_fail(c, HV_ERR_DOMAIN, "a node does not lie in the domain given by the boundary");
    c->tree_on_plane = hbad != 0;   // the walk refuses it; the queries do not mind
    return HV_OK;
}

// axis-aligned box check: planes 2k (x_k = hi, +e_k) and 2k+1 (x_k = lo, -e_k)
// the tree just built over the context's own points and planes, for the entry point without
// allocation (its tables stay in the pool until the next tree build replaces them)
void record_tree(hv_ctx* c, const DevGeo& g, int tables, int L) {
    c->tree_ok = true;
    c->tree_tables = tables;
    c->tree_leaf = L;
    c->tree_g = g;
    c->tree_xw = c->xw;
}

// the walk's domain: a cuboid or a bounded convex polytope, Dirichlet / Neumann planes
int check_box(hv_ctx* c, int D) {
    const char* why = nullptr;
    const int rc = hv_domain(c, D, &why);
    if (rc) return hv_fail(c, rc, why);
    for (int p = 0; p < c->P; ++p)
        if (c->pbc[p] > 0) return hv_fail(c, HV_ERR_UNSUPPORTED, "periodic planes are not handled by the device walk");
    return HV_OK;
}

// listed vertices per walked cell of a periodic box: the mean vertex count of a
// Poisson-Voronoi cell over d + 1 (2D 6, 3D 27.1, 4D 192.6, 6D ~21 600: 3 100 measured at 6D
// N=50k), with room for the vertices listed again in a neighbour's frame
double vertices_per_periodic_cell(int D) {
    switch (D) {
        case 2: return 2.4;
        case 3: return 7.5;
        case 4: return 42.0;
        case 5: return 350.0;
        default: return 3300.0;
    }
}


}  // namespace

// The domain {x : <n_p, x> <= <n_p, b_p> for every plane p} (src/boundary.jl:427-441).
// The cuboid of the reference's plane order (cuboid(d; ...), src/boundary.jl:510-534): plane
// 2k is x_k = hi_k with normal +e_k, plane 2k+1 is x_k = lo_k with normal -e_k.  Any other set
// of planes bounds a convex polytope or nothing bounded: it is unbounded when there are at
// most d planes, when the normals do not span R^d, or when some direction y != 0 has
// <n_p, y> <= 0 for every plane (such a direction, if any, is the null vector of d - 1
// independent normals: an extreme ray of the cone {y : N y <= 0}).  A bounded polytope's
// vertices (d independent planes meeting, no plane violated) give its bounding box; those
// where exactly d planes meet are the walk's corner seeds.  Unbounded (Boundary() without
// planes among them, src/boundary.jl:133-135): rays may leave every plane behind, and the
// frame is the generators' bounding box (set by the tree build).
static bool hv_solve_small(int D, double* A, double* b) {   // A x = b in place (x -> b), pivoting
    for (int col = 0; col < D; ++col) {
        int piv = col;
        for (int r = col + 1; r < D; ++r)
            if (std::fabs(A[r * D + col]) > std::fabs(A[piv * D + col])) piv = r;
        if (!(std::fabs(A[piv * D + col]) > 1e-10)) return false;   // unit normals: well scaled
        if (piv != col) {
            for (int k = 0; k < D; ++k) std::swap(A[col * D + k], A[piv * D + k]);
            std::swap(b[col], b[piv]);
        }
        for (int r = col + 1; r < D; ++r) {
            const double f = A[r * D + col] / A[col * D + col];
            for (int k = col; k < D; ++k) A[r * D + k] -= f * A[col * D + k];
            b[r] -= f * b[col];
        }
    }
    for (int r = D - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < D; ++k) s -= A[r * D + k] * b[k];
        b[r] = s / A[r * D + r];
    }
    return true;
}

// next combination of `m` indices out of `n` (ascending); false after the last
static bool hv_next_comb(std::vector<int>& a, int n) {
    const int m = (int)a.size();
    int i = m - 1;
    while (i >= 0 && a[i] == n - m + i) --i;
    if (i < 0) return false;
    ++a[i];
    for (int j = i + 1; j < m; ++j) a[j] = a[j - 1] + 1;
    return true;
}

int hv_domain(hv_ctx* c, int D, const char** why) {
    const int P = c->P;
    const double* nrm = c->pnormal.data();
    const double* bas = c->pbase.data();
    c->dom_general = false;
    c->dom_unbounded = false;
    c->dom_corner.clear();
    c->dom_corner_pl.clear();
    bool cub = P == 2 * D;
    for (int p = 0; cub && p < P; ++p)
        for (int q = 0; q < D; ++q)
            cub = cub && nrm[p * D + q] == ((q == p / 2) ? ((p & 1) ? -1.0 : 1.0) : 0.0);
    if (cub) {
        for (int k = 0; k < D; ++k) {
            c->dom_hi[k] = bas[(2 * k) * D + k];
            c->dom_lo[k] = bas[(2 * k + 1) * D + k];
        }
        return HV_OK;
    }
    c->dom_general = true;
    std::vector<double> h(P);
    for (int p = 0; p < P; ++p) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += nrm[p * D + k] * bas[p * D + k];
        h[p] = s;
    }
    double ncomb = 1.0;
    for (int i = 0; i < D; ++i) ncomb = ncomb * (double)(P - i) / (double)(i + 1);
    if (P <= D) {
        c->dom_unbounded = true;
        return HV_OK;
    }
    if (ncomb > 2e6) {
        *why = "too many boundary planes for a general polytope (more than 2e6 d-subsets)";
        return HV_ERR_UNSUPPORTED;
    }
    // the normals span R^d?
    {
        std::vector<double> Q;
        int r = 0;
        for (int p = 0; p < P && r < D; ++p) {
            std::vector<double> w(nrm + p * D, nrm + p * D + D);
            for (int pass = 0; pass < 2; ++pass)
                for (int a = 0; a < r; ++a) {
                    double dp = 0.0;
                    for (int k = 0; k < D; ++k) dp += w[k] * Q[a * D + k];
                    for (int k = 0; k < D; ++k) w[k] -= dp * Q[a * D + k];
                }
            double nw = 0.0;
            for (int k = 0; k < D; ++k) nw += w[k] * w[k];
            if (nw > 1e-18) {
                nw = std::sqrt(nw);
                for (int k = 0; k < D; ++k) Q.push_back(w[k] / nw);
                ++r;
            }
        }
        if (r < D) {
            c->dom_unbounded = true;
            return HV_OK;
        }
    }
    // no direction of recession: the null vector of every d - 1 independent normals has a
    // plane facing it on both sides
    {
        std::vector<int> S(D - 1);
        for (int i = 0; i < D - 1; ++i) S[i] = i;
        do {
            std::vector<double> Q;
            bool indep = true;
            for (int i = 0; i < D - 1 && indep; ++i) {
                std::vector<double> w(nrm + S[i] * D, nrm + S[i] * D + D);
                for (int pass = 0; pass < 2; ++pass)
                    for (int a = 0; a < i; ++a) {
                        double dp = 0.0;
                        for (int k = 0; k < D; ++k) dp += w[k] * Q[a * D + k];
                        for (int k = 0; k < D; ++k) w[k] -= dp * Q[a * D + k];
                    }
                double nw = 0.0;
                for (int k = 0; k < D; ++k) nw += w[k] * w[k];
                if (!(nw > 1e-18)) { indep = false; break; }
                nw = std::sqrt(nw);
                for (int k = 0; k < D; ++k) Q.push_back(w[k] / nw);
            }
            if (!indep) continue;
            // y: the coordinate axis with the largest part orthogonal to the d - 1 normals
            double y[HV_MAXD] = {}, best = -1.0;
            for (int e = 0; e < D; ++e) {
                double w[HV_MAXD] = {};
                w[e] = 1.0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int a = 0; a < D - 1; ++a) {
                        double dp = 0.0;
                        for (int k = 0; k < D; ++k) dp += w[k] * Q[a * D + k];
                        for (int k = 0; k < D; ++k) w[k] -= dp * Q[a * D + k];
                    }
                double nw = 0.0;
                for (int k = 0; k < D; ++k) nw += w[k] * w[k];
                if (nw > best) {
                    best = nw;
                    for (int k = 0; k < D; ++k) y[k] = w[k];
                }
            }
            best = std::sqrt(best);
            for (int k = 0; k < D; ++k) y[k] /= best;
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                bool recedes = true;
                for (int q = 0; q < P && recedes; ++q) {
                    double dp = 0.0;
                    for (int k = 0; k < D; ++k) dp += nrm[q * D + k] * y[k];
                    recedes = sgn * dp <= 1e-12;
                }
                if (recedes) {
                    c->dom_unbounded = true;
                    return HV_OK;
                }
            }
        } while (hv_next_comb(S, P));
    }
    // vertices: bounding box, and the simple ones as corner seeds
    for (int k = 0; k < D; ++k) {
        c->dom_lo[k] = INFINITY;
        c->dom_hi[k] = -INFINITY;
    }
    std::vector<int> S(D);
    for (int i = 0; i < D; ++i) S[i] = i;
    int nvert = 0;
    do {
        double A[HV_MAXD * HV_MAXD], x[HV_MAXD];
        for (int i = 0; i < D; ++i) {
            for (int k = 0; k < D; ++k) A[i * D + k] = nrm[S[i] * D + k];
            x[i] = h[S[i]];
        }
        if (!hv_solve_small(D, A, x)) continue;
        bool inside = true;
        int active = 0;
        for (int q = 0; q < P && inside; ++q) {
            double s = -h[q];
            for (int k = 0; k < D; ++k) s += nrm[q * D + k] * x[k];
            const double tol = 1e-10 * (1.0 + std::fabs(h[q]));
            if (s > tol) inside = false;
            else if (s >= -tol) ++active;
        }
        if (!inside) continue;
        ++nvert;
        for (int k = 0; k < D; ++k) {
            c->dom_lo[k] = std::min(c->dom_lo[k], x[k]);
            c->dom_hi[k] = std::max(c->dom_hi[k], x[k]);
        }
        if (active == D) {
            for (int k = 0; k < D; ++k) c->dom_corner.push_back(x[k]);
            for (int i = 0; i < D; ++i) c->dom_corner_pl.push_back(S[i]);
        }
    } while (hv_next_comb(S, P));
    if (nvert == 0) {
        *why = "the boundary planes do not enclose a domain";
        return HV_ERR_DOMAIN;
    }
    return HV_OK;
}

// vertex count per generator in the unit cube, measured with the CPU oracle (d = 2..6)
double vertices_per_point(int D) {
    switch (D) {
        case 2: return 2.2;
        case 3: return 7.0;
        case 4: return 29.0;
        case 5: return 140.0;
        default: return 700.0;
    }
}

// Tentative keys: after the start phase each ray knows a candidate vertex
// key = edge + nearest-neighbour winner.  Rays of one level with the same tentative key
// are grouped; only the first (the primary) is scanned.  When the primary's scan confirms
// the key, it is a Voronoi vertex containing every grouped ray's edge, so (general
// position: an edge has two end points) it is the far end point of all of them and the
// other rays need no scan.  Unconfirmed groups are scanned in a second pass.
template <int D>
__device__ __forceinline__ bool hv_tent_key(const DevGeo& g, const DevRays& R, const int* __restrict__ vkey, int ray,
                                            int* tk) {
    const int i1 = R.i1[ray];
    if (i1 < 0) return false;
    const int w = i1 < g.N ? g.sidx[i1] : i1;
    int e[D];
    hv_edge_of<D>(vkey + (int64_t)R.v[ray] * (D + 1), R.drop[ray], e);
    int ins = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) ins += (e[j] < w) ? 1 : 0;
#pragma unroll
    for (int j = 0; j <= D; ++j) tk[j] = (j < ins) ? e[j] : (j == ins ? w : e[j > 0 ? j - 1 : 0]);
    return true;
}

template <int D, bool B>
__global__ __launch_bounds__(256) void k_tent_group(DevGeo g0, GeoSet F, DevRays R, int nr, int nfresh,
                                                    const int* __restrict__ vkey, int* __restrict__ ttab, uint64_t tmask,
                                                    int* __restrict__ prim_of, uint8_t* __restrict__ primary) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr) return;
    const DevGeo& g = hv_geo<B>(g0, F, vkey[(int64_t)R.v[t] * (D + 1)]);
    int tk[D + 1];
    // carried rays (t >= nfresh) already waited once: scanned, and they claim no key
    if (t >= nfresh || !hv_tent_key<D>(g, R, vkey, t, tk)) { prim_of[t] = t; primary[t] = 1; return; }
    uint64_t h = hv_hash_ints<D + 1>(tk) & tmask;
    for (uint64_t probe = 0; probe <= tmask; ++probe) {
        int cur = ttab[h];
        if (cur < 0) {
            const int old = atomicCAS(&ttab[h], -1, t);
            if (old == -1) { prim_of[t] = t; primary[t] = 1; return; }
            cur = old;
        }
        // a claimer is a fresh ray with a key; over a forest it may belong to another mesh,
        // whose tree maps its start candidate (a sorted position or a plane id of that mesh)
        int ok[D + 1];
        const DevGeo& gc = hv_geo<B>(g0, F, vkey[(int64_t)R.v[cur] * (D + 1)]);
        hv_tent_key<D>(gc, R, vkey, cur, ok);
        if (keys_equal<D>(tk, ok)) { prim_of[t] = cur; primary[t] = 0; return; }
        h = (h + 1) & tmask;
    }
    prim_of[t] = t;
    primary[t] = 1;
}

__global__ void k_gather_flags(const int* __restrict__ ord, int nr, const uint8_t* __restrict__ by_ray,
                               uint8_t* __restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nr) keep[i] = by_ray[ord[i]];
}

// Rays that share a tentative end point with an earlier ray of the level wait: after the
// level's vertices are registered, a waiting ray whose edge got its second end point is done
// (the sequential walk would not walk that edge either); the others are carried into the
// next level's rays, where they are scanned.
template <int D>
__global__ void k_carry_out(DevRays R, int nr, const uint8_t* __restrict__ primary,
                            const unsigned long long* __restrict__ rt, int64_t* __restrict__ cs, int* __restrict__ cv,
                            int* __restrict__ cd, double* __restrict__ cu, double* __restrict__ cc,
                            double* __restrict__ cdu, int* __restrict__ count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool carry = t < nr && !primary[t] && !(rt[R.slot[t]] & HV_EDGE_DONE);
    const int o = hv_wave_alloc(count, carry);
    if (!carry) return;
    cs[o] = R.slot[t];
    cv[o] = R.v[t];
    cd[o] = R.drop[t];
#pragma unroll
    for (int k = 0; k < D; ++k) cu[(int64_t)o * D + k] = R.u[(int64_t)t * D + k];
    cc[o] = R.c[t];
    cdu[o] = R.du[t];
}

template <int D>
__global__ void k_carry_in(const int64_t* __restrict__ cs, const int* __restrict__ cv, const int* __restrict__ cd,
                           const double* __restrict__ cu, const double* __restrict__ cc, const double* __restrict__ cdu,
                           int n, DevRays R, int base) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int o = base + t;
    R.slot[o] = cs[t];
    R.v[o] = cv[t];
    R.drop[o] = cd[t];
#pragma unroll
    for (int k = 0; k < D; ++k) R.u[(int64_t)o * D + k] = cu[(int64_t)t * D + k];
    R.c[o] = cc[t];
    R.du[o] = cdu[t];
}

// Chunk selection (64-ray blocks interleaved over K chunks); a ray of a later chunk is
// dropped when its edge got its second end point from a vertex found by an earlier
// chunk of the same level (the sequential walk's edge hash would not walk it either).
__global__ void k_chunk(const int* __restrict__ ord, int nr, int c, int K, const int64_t* __restrict__ slot,
                        const unsigned long long* __restrict__ rt, int* __restrict__ cand,
                        uint8_t* __restrict__ keep) {
    // chunk c holds the sorted 64-ray blocks c, c+K, c+2K, ...: packets stay spatially
    // coherent while rays that end at the same vertex (neighbours in the sort) tend to
    // fall into different chunks
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pos = ((((int64_t)(i >> 6)) * K + c) << 6) + (i & 63);
    if (pos >= nr) return;
    const int ray = ord[pos];
    cand[i] = ray;
    keep[i] = (c == 0) ? 1 : ((rt[slot[ray]] & HV_EDGE_DONE) ? 0 : 1);
}

// ---- a forest of meshes walked together ----------------------------------------------------
// Rays of the list per mesh.  The list is a subsequence of the rays sorted by their codes,
// which hold the mesh above 32 bits, so each mesh's rays form one segment: its bounds by binary
// search, one thread per mesh.
__global__ void k_pkt_count(const int* __restrict__ list, int n, const int* __restrict__ n_dev,
                            const uint64_t* __restrict__ code, int nmesh, int* __restrict__ cnt) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmesh) return;
    if (n_dev) n = min(n, *n_dev);
    auto first_of = [&](int key) {   // first list position whose mesh is >= key
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)(code[list[mid]] >> 32) < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    cnt[m] = first_of(m + 1) - first_of(m);
}

// Prefix sums over the meshes (a wavefront, 64 meshes per step): seg = first list position of
// each mesh's rays, pk_off = first packet of each mesh (packets of ppw rays within a mesh).
__global__ void k_pkt_plan(const int* __restrict__ cnt, int nmesh, int ppw, int* __restrict__ seg, int* __restrict__ pk_off) {
    const int lane = threadIdx.x;
    int s0 = 0, p0 = 0;
    for (int b = 0; b < nmesh; b += 64) {
        const int m = b + lane;
        const int cm = m < nmesh ? cnt[m] : 0;
        int s = cm, p = (cm + ppw - 1) / ppw;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int s2 = __shfl_up(s, off), p2 = __shfl_up(p, off);
            if (lane >= off) { s += s2; p += p2; }
        }
        if (m < nmesh) {
            seg[m] = s0 + s - cm;
            pk_off[m] = p0 + p - (cm + ppw - 1) / ppw;
        }
        s0 += __shfl(s, 63);
        p0 += __shfl(p, 63);
    }
    if (lane == 0) {
        seg[nmesh] = s0;
        pk_off[nmesh] = p0;
    }
}

constexpr int HV_FOREST_STREAMS = 4;   // the box's hardware queues per process

// Every mesh's tree into segments of one set of tables: sorted positions of mesh k start at a
// multiple of 256 (whole groups of 8 leaves), generator ids at goff[k].
template <int D>
static int build_forest(hv_ctx* c, int L, int tables, Forest& fo) {
    const int nm = fo.nmesh;
    const int64_t ntot = fo.ntot;
    if (ntot + (int64_t)2 * D * nm >= (1ll << 30)) return hv_fail(c, HV_ERR_UNSUPPORTED, "too many generators in one batch");
    std::vector<int64_t> soff(nm + 1, 0), bo(nm + 1, 0), cbo(nm + 1, 0);
    std::vector<TreeDims> td(nm);
    for (int k = 0; k < nm; ++k) {
        td[k] = tree_dims(fo.n[k], L);
        soff[k + 1] = soff[k] + ((int64_t)fo.n[k] + 255) / 256 * 256;
        bo[k + 1] = bo[k] + (int64_t)td[k].NN * D;
        cbo[k + 1] = cbo[k] + td[k].ninternal * 8 * 2 * D;
    }
    const int64_t S = soff[nm];
    // points: packed on the host, one copy, kept by the meshes (their generator sets)
    auto xo = std::make_shared<DevArr>();
    xo->device = c->device;
    if (hipMalloc(&xo->p, sizeof(double) * (size_t)std::max<int64_t>(ntot, 1) * D) != hipSuccess) {
        (void)hipGetLastError();
        xo->p = nullptr;
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (batch points)");
    }
    xo->n = sizeof(double) * (size_t)ntot * D;
    fo.xo = xo;
    double* dxo = (double*)xo->p;
    {
        std::vector<double> hx((size_t)ntot * D);
        for (int k = 0; k < nm; ++k)
            std::memcpy(&hx[(size_t)fo.goff[k] * D], fo.xs[k], sizeof(double) * (size_t)fo.n[k] * D);
        HVCK(hipMemcpyAsync(dxo, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
    }
    const int P = 2 * D;
    double* pn = pget<double>(c, "f_pn", (size_t)nm * P * D);
    double* ph = pget<double>(c, "f_ph", (size_t)nm * P);
    double* xs = pget<double>(c, "f_xs", (size_t)S * (D + 1));
    int* sidx = pget<int>(c, "f_sidx", (size_t)S);
    int* rank = pget<int>(c, "f_rank", (size_t)ntot);
    int* knn = pget<int>(c, "f_knn", (size_t)S * HV_KNN);
    float* xf = pget<float>(c, "f_xf", (size_t)S * 8);
    double* blo = pget<double>(c, "f_blo", (size_t)bo[nm]);
    double* bhi = pget<double>(c, "f_bhi", (size_t)bo[nm]);
    float* cb = pget<float>(c, "f_cb", (size_t)cbo[nm]);
    double* lorg = tables != HV_PREC_FP64 ? pget<double>(c, "f_lorg", (size_t)(S / 256) * 8) : nullptr;
    hv_u32x4* pkv = tables == HV_PREC_BF16 ? pget<hv_u32x4>(c, "f_pkv", (size_t)S * 4) : nullptr;
    float* pxf = tables == HV_PREC_FP32 ? pget<float>(c, "f_pxf", (size_t)S * 8) : nullptr;
    int* bad = pget<int>(c, "f_bad", (size_t)nm);
    int* mof = pget<int>(c, "f_mof", (size_t)ntot);
    DevGeo* gs = pget<DevGeo>(c, "f_geo", (size_t)nm);
    if (!pn || !ph || !xs || !sidx || !rank || !knn || !xf || !blo || !bhi || !cb || !bad || !mof || !gs ||
        (tables != HV_PREC_FP64 && !lorg) || (tables == HV_PREC_BF16 && !pkv) || (tables == HV_PREC_FP32 && !pxf))
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (batch trees)");
    std::vector<double> hpn((size_t)nm * P * D), hph((size_t)nm * P);
    std::vector<int> hmof((size_t)ntot);
    for (int k = 0; k < nm; ++k) {
        for (int p = 0; p < P; ++p) {
            double h = 0.0;
            for (int q = 0; q < D; ++q) {
                hpn[((size_t)k * P + p) * D + q] = fo.normal[k][p * D + q];
                h += fo.normal[k][p * D + q] * fo.base[k][p * D + q];
            }
            hph[(size_t)k * P + p] = h;
        }
        std::fill(hmof.begin() + fo.goff[k], hmof.begin() + fo.goff[k] + fo.n[k], k);
    }
    HVCK(hipMemcpyAsync(pn, hpn.data(), sizeof(double) * hpn.size(), hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(ph, hph.data(), sizeof(double) * hph.size(), hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(mof, hmof.data(), sizeof(int) * hmof.size(), hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemsetAsync(bad, 0, sizeof(int) * nm, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    fo.geo.assign(nm, DevGeo{});
    // the meshes' trees on HV_FOREST_STREAMS streams at once (small meshes leave the chip
    // idle one at a time), each stream with its own temporaries, sized up front for the
    // largest mesh so that nothing is reallocated while another stream works
    constexpr int NS = HV_FOREST_STREAMS;
    while ((int)c->side.size() < NS) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return hv_fail(c, HV_ERR_DEVICE, "stream");
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(st);
            return hv_fail(c, HV_ERR_DEVICE, "event");
        }
        c->side.push_back(st);
        c->side_ev.push_back(ev);
    }
    {
        int nmax = 1;
        for (int k = 0; k < nm; ++k) nmax = std::max(nmax, fo.n[k]);
        size_t tmpb = 0;
        HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpb, (uint64_t*)nullptr, (uint64_t*)nullptr, (int*)nullptr,
                                                (int*)nullptr, nmax, 0, 64, c->stream));
        for (int q = 0; q < NS; ++q) {
            const std::string sfx = "#" + std::to_string(q);
            if (!pget<uint64_t>(c, ("code" + sfx).c_str(), nmax) || !pget<uint64_t>(c, ("code2" + sfx).c_str(), nmax) ||
                !pget<int>(c, ("idx" + sfx).c_str(), nmax) || !pget<int>(c, ("idx2" + sfx).c_str(), nmax) ||
                !c->pool->get("cubtmp" + sfx, tmpb, false, c->stream))
                return hv_fail(c, HV_ERR_DEVICE, "out of device memory (batch trees)");
        }
        HVCK(hipStreamSynchronize(c->stream));
    }
    for (int k = 0; k < nm; ++k) {
        TreeOut o;
        o.xin = dxo + (int64_t)fo.goff[k] * D;
        o.n = fo.n[k];
        o.goff = fo.goff[k];
        o.Nid = (int)(ntot + (int64_t)P * k);
        o.P = P;
        o.pnormal = fo.normal[k];
        o.pbase = fo.base[k];
        o.pn = pn + (size_t)k * P * D;
        o.ph = ph + (size_t)k * P;
        o.planes_ready = true;
        o.xo_base = dxo;
        o.xs = xs + soff[k] * (D + 1);
        o.sidx = sidx + soff[k];
        o.rank_base = rank;
        o.knn = knn + soff[k] * HV_KNN;
        o.xf = xf + soff[k] * 8;
        o.blo = blo + bo[k];
        o.bhi = bhi + bo[k];
        o.cb = cb + cbo[k];
        o.lorg = lorg ? lorg + (soff[k] / 256) * 8 : nullptr;
        o.pkv = pkv ? pkv + soff[k] * 4 : nullptr;
        o.pxf = pxf ? pxf + soff[k] * 8 : nullptr;
        o.bad = bad + k;
        o.defer_tables = true;
        Tree T;
        int rc = build_tree_core<D>(c, L, tables, o, T, c->side[k % NS], k % NS);
        if (rc) return rc;
        fo.geo[k] = T.g;
    }
    // the walk (on the context's stream) starts when every tree is built
    for (int q = 0; q < NS; ++q) {
        HVCK(hipEventRecord(c->side_ev[q], c->side[q]));
        HVCK(hipStreamWaitEvent(c->stream, c->side_ev[q], 0));
    }
    HVCK(hipMemcpyAsync(gs, fo.geo.data(), sizeof(DevGeo) * nm, hipMemcpyHostToDevice, c->stream));
    // 16-NN lists and scan tables of every mesh in one launch each (items numbered mesh after mesh)
    {
        std::vector<int> hoff(3 * (nm + 1), 0);
        int* ob = hoff.data();             // blocks of 64 points (16-NN)
        int* og = ob + (nm + 1);           // groups of 8 leaves (origins)
        int* op = og + (nm + 1);           // padded points of whole leaves (K-vectors / FP32 rows)
        for (int k = 0; k < nm; ++k) {
            ob[k + 1] = ob[k] + (fo.n[k] + 63) / 64;
            og[k + 1] = og[k] + (fo.geo[k].nleaf + 7) / 8;
            op[k + 1] = op[k] + fo.geo[k].nleaf * 32;
        }
        int* doff = pget<int>(c, "f_off", hoff.size());
        if (!doff) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (batch trees)");
        HVCK(hipMemcpyAsync(doff, hoff.data(), sizeof(int) * hoff.size(), hipMemcpyHostToDevice, c->stream));
        GeoSet Ft{gs, mof, nm};
        k_knn64<D, true><<<ob[nm], 64, 0, c->stream>>>(DevGeo{}, Ft, doff, nullptr);
        HVCK(hipGetLastError());
        if (tables != HV_PREC_FP64) {
            k_leaf_origin<D, true><<<og[nm], 64, 0, c->stream>>>(DevGeo{}, Ft, doff + (nm + 1), nullptr);
            HVCK(hipGetLastError());
            if (tables == HV_PREC_BF16)
                k_point_kvec<D, true><<<nblk(op[nm], 256), 256, 0, c->stream>>>(DevGeo{}, Ft, doff + 2 * (nm + 1), nullptr,
                                                                               nullptr);
            else
                k_point_f32<D, true><<<nblk(op[nm], 256), 256, 0, c->stream>>>(DevGeo{}, Ft, doff + 2 * (nm + 1), nullptr,
                                                                              nullptr);
            HVCK(hipGetLastError());
        }
        HVCK(hipStreamSynchronize(c->stream));   // hoff is a local
    }
    std::vector<int> hbad(nm);
    HVCK(hipMemcpyAsync(hbad.data(), bad, sizeof(int) * nm, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < nm; ++k) {
        if (hbad[k] & 1) return hv_fail(c, HV_ERR_DOMAIN, "a node does not lie in the domain given by the boundary");
        if (hbad[k]) return hv_fail(c, HV_ERR_DEGENERATE, "a node lies on a boundary plane (it coincides with its mirror image)");
    }
    fo.F.gs = gs;
    fo.F.mof = mof;
    fo.F.nmesh = nm;
    return HV_OK;
}

// Vertices of a forest walk -> the meshes: the live vertices sorted by mesh (stable: each
// mesh keeps the walk's order), then copied with keys in the mesh's own ids (generator
// goff + i -> i, plane ntot + 2 D k + p -> n_k + p).  A key's first entry is a generator of
// its mesh (planes sort last).
template <int D>
__global__ __launch_bounds__(256) void k_split_key(const int* __restrict__ ids, int nv, const int* __restrict__ vkey,
                                                   const int* __restrict__ mof, int* __restrict__ mkey) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nv) mkey[t] = mof[vkey[(int64_t)ids[t] * (D + 1)]];
}

// off[m] = first sorted position of mesh m (binary search in the sorted mesh keys), off[nm] = nv
__global__ void k_split_offsets(const int* __restrict__ mkey, int nv, int nm, int64_t* __restrict__ off) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nm) return;
    int lo = 0, hi = nv;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (mkey[mid] < m) lo = mid + 1;
        else hi = mid;
    }
    off[m] = lo;
}

template <int D>
__global__ __launch_bounds__(256) void k_split_copy(const int* __restrict__ ids, const int* __restrict__ mkey, int nv,
                                                    const int* __restrict__ vkey, const double* __restrict__ vx,
                                                    const int* __restrict__ goff, const int* __restrict__ nper,
                                                    int64_t ntot, int* __restrict__ okey, double* __restrict__ ox) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int v = ids[t], m = mkey[t];
    const int* key = vkey + (int64_t)v * (D + 1);
    const int64_t pbase = ntot + (int64_t)2 * D * m;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const int e = key[j];
        okey[t * (D + 1) + j] = e < ntot ? e - goff[m] : (int)(e - pbase) + nper[m];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) ox[t * D + k] = vx[(int64_t)v * D + k];
}

// The walk over the context's working set.  Only edges with a generator below nwalk are
// followed and only those generators seed it, so the mesh holds the vertices of the cells
// 0 .. nwalk-1 (all of them when nwalk == N).  With a list of cells (sorted, distinct, each
// below nwalk), only edges holding one of them are followed and only they seed the walk: the
// mesh holds exactly the vertices of those cells (each cell's edge graph is connected, and an
// edge's end points are vertices of every cell the edge holds).  A listed cell that neither a
// seed nor the walk reached gets an exact descent, and the walk continues from there.
// B: one walk over a forest of meshes of dimension D (Forest: every mesh's tree, generator ids
// global over the meshes, plane ids above all of them); its vertices go to the forest's meshes.
template <int D, bool B>
int hv_walk_impl(hv_ctx* c, const hv_opts* o, hv_mesh** out, int nwalk, const std::vector<int>* cells, Forest* fo) {
    hv_opts opt{};
    if (o) opt = *o;
    if (opt.precision != HV_PREC_FP64 && opt.precision != HV_PREC_BF16 && opt.precision != HV_PREC_FP32)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "precision: HV_PREC_FP64, HV_PREC_FP32 or HV_PREC_BF16");
    const int prec = opt.precision;
    if (prec != HV_PREC_FP64 && opt.leaf_size > 0 && opt.leaf_size != 32)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "the mixed-precision scans use leaves of 32 points");
    const int L = opt.leaf_size > 0 ? opt.leaf_size : 32;
    const double plane_tol = opt.plane_tolerance > 0 ? opt.plane_tolerance : 1e-12;
    const int KCH = opt.reserved0 > 1 ? opt.reserved0 : 1;   // chunks per level (0: tentative-key grouping)
    // scan tiles: the caller's, else the autotuner's for this dimension, precision and size
    const HvTiles tiles = hv_tiles_for(c, opt, D, prec, B ? fo->ntot : c->N);
    int rc = HV_OK;
    if constexpr (!B) {
        rc = check_box(c, D);
        if (rc) return rc;
    }
    const int N = B ? (int)fo->ntot : (int)c->N;
    const bool subset = cells != nullptr;
    const int ncell = subset ? (int)cells->size() : 0;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1], r0 = c->ev[2], r1 = c->ev[3];
    HVCK(hipEventRecord(e0, c->stream));
    Tree T;
    GeoSet F{};
    if constexpr (B) {
        rc = build_forest<D>(c, L, prec, *fo);
        if (rc) return rc;
        F = fo->F;
    } else {
        rc = build_tree<D>(c, L, T, prec);
        if (rc) return rc;
        if (c->tree_on_plane)
            return hv_fail(c, HV_ERR_DEGENERATE, "a node lies on a boundary plane (it coincides with its mirror image)");
        if (!c->extended) record_tree(c, T.g, prec, L);
    }
    DevGeo& g = T.g;   // over a forest: unused (every kernel takes its mesh's tree from F)
    const int nmesh = B ? fo->nmesh : 1;
    // box corners, or the simple vertices of a general polytope (single meshes only)
    const bool poly = !B && c->dom_general;
    const int nseed = poly ? (int)(c->dom_corner_pl.size() / D) : nmesh << D;
    double* ccx = nullptr;
    int* ccpl = nullptr;
    if (poly && nseed > 0) {
        ccx = pget<double>(c, "corner_x", (size_t)nseed * D);
        ccpl = pget<int>(c, "corner_pl", (size_t)nseed * D);
        if (!ccx || !ccpl) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (corners)");
        HVCK(hipMemcpyAsync(ccx, c->dom_corner.data(), sizeof(double) * nseed * D, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(ccpl, c->dom_corner_pl.data(), sizeof(int) * nseed * D, hipMemcpyHostToDevice, c->stream));
        HVCK(hipStreamSynchronize(c->stream));   // the host vectors may change with the planes
    }
    // the listed cells (device copy) and their mask over the generators
    uint8_t* cmask = nullptr;
    int* dcells = nullptr;
    if (subset) {
        cmask = pget<uint8_t>(c, "cmask", (size_t)N);
        dcells = pget<int>(c, "cells", (size_t)std::max(1, ncell));
        if (!cmask || !dcells) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (cells)");
        std::vector<uint8_t> hm((size_t)N, 0);
        for (int i : *cells) hm[(size_t)i] = 1;
        HVCK(hipMemcpyAsync(cmask, hm.data(), (size_t)N, hipMemcpyHostToDevice, c->stream));
        if (ncell) HVCK(hipMemcpyAsync(dcells, cells->data(), sizeof(int) * ncell, hipMemcpyHostToDevice, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
    }
    // initial sizes from the expected vertex count (or the caller's hint); all tables grow on demand
    const double est = opt.vertex_capacity > 0 ? (double)opt.vertex_capacity
                       : B ? vertices_per_point(D) * N
                       : subset ? vertices_per_periodic_cell(D) * ncell
                       : (nwalk < N ? vertices_per_periodic_cell(D) : vertices_per_point(D)) * nwalk;
    int64_t vcap = (int64_t)(2.0 * est) + 4 * nseed + 1024;
    auto hpow = [](int64_t n) { uint64_t p = 1024; while ((int64_t)p < n) p <<= 1; return p; };
    uint64_t hsize = hpow((int64_t)(2.4 * est) + 4 * nseed);
    int* vkey = pget<int>(c, "vkey", (size_t)vcap * (D + 1));
    double* vx = pget<double>(c, "vx", (size_t)vcap * D);
    int* ids = pget<int>(c, "ids", (size_t)vcap);
    int* htab = pget<int>(c, "htab", hsize);
    unsigned long long* ctr = pget<unsigned long long>(c, "ctr", 16);
    // the scan's statistics (0 pairs, 1 leaves, 2 ties, 6 FP64 rechecks, 7 rays rechecked), each
    // in HV_SCAN_SPREAD copies that the packets share out
    constexpr int HV_SCAN_SPREAD = 256;
    unsigned long long* sctr = pget<unsigned long long>(c, "sctr", 12 * HV_SCAN_SPREAD);
    // k_insert's counters (8 duplicates, 9 vertex-hash overflow, 10 rays without a vertex,
    // 11 singular vertices) likewise, after the scan's
    unsigned long long* ictr = sctr + 8 * HV_SCAN_SPREAD;
    int* dcount = pget<int>(c, "dcount", 4);
    if (!vkey || !vx || !ids || !htab || !ctr || !sctr || !dcount)
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (tables)");
    HVCK(hipMemsetAsync(htab, 0xff, hsize * sizeof(int), c->stream));
    HVCK(hipMemsetAsync(ctr, 0, 16 * sizeof(unsigned long long), c->stream));
    HVCK(hipMemsetAsync(sctr, 0, 12 * HV_SCAN_SPREAD * sizeof(unsigned long long), c->stream));
    hv_stats_t st{};
    unsigned long long hctr[16];
    int64_t vnext = 0;   // next provisional vertex id
    int64_t nlive = 0;   // live vertices: ids[0 .. nlive)

    // counters in ctr[]: 0 pairs, 1 leaves, 2 ties, 4 unique edges, 5 edge overflow, 8 dup, 9 vertex overflow
    auto read_ctr = [&]() -> int {
        HVCK(hipMemcpyAsync(hctr, ctr, sizeof hctr, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        return HV_OK;
    };
    // per-level edge tables (see k_register): the ray-edge table of the current level and the
    // new-edge table, cleared at the first registration of a level (after that level's rays
    // were emitted from it)
    unsigned long long* rtab = nullptr;
    uint64_t rsize = 0;
    unsigned long long* ntab = nullptr;
    uint64_t nsize = 0;
    bool nt_fresh = true;   // the next registration starts a new level's table
    auto build_rt = [&](const DevRays& R, int nr) -> int {
        rsize = hpow(2 * (int64_t)nr + 64);
        rtab = pget<unsigned long long>(c, "rtab", rsize);
        if (!rtab) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (ray-edge table)");
        HVCK(hipMemsetAsync(rtab, 0, rsize * sizeof(unsigned long long), c->stream));
        if (nr > 0) {
            k_rt_build<D><<<nblk(nr, 256), 256, 0, c->stream>>>(R, nr, vkey, rtab, rsize - 1, ctr + 4);
            HVCK(hipGetLastError());
        }
        return HV_OK;
    };
    // vertex table and vertex hash room for `incoming` more candidates
    auto ensure_vertices = [&](int64_t incoming) -> int {
        if (vnext + incoming > (int64_t)INT32_MAX) return hv_fail(c, HV_ERR_CAPACITY, "more than 2^31 vertex slots");
        if (vnext + incoming > vcap) {
            const int64_t ncap = std::max<int64_t>(vcap * 3 / 2, vnext + incoming + 1024);
            int* nk2 = (int*)c->pool->get("vkey", (size_t)ncap * (D + 1) * sizeof(int), true, c->stream);
            double* nx2 = (double*)c->pool->get("vx", (size_t)ncap * D * sizeof(double), true, c->stream);
            int* ni2 = (int*)c->pool->get("ids", (size_t)ncap * sizeof(int), true, c->stream);
            if (!nk2 || !nx2 || !ni2) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (vertex table)");
            vkey = nk2; vx = nx2; ids = ni2; vcap = ncap;
        }
        if ((nlive + incoming) * 2 > (int64_t)hsize) {
            uint64_t nsz = hsize;
            while ((nlive + incoming) * 2 > (int64_t)nsz) nsz <<= 1;
            int* nh = (int*)c->pool->get("htab", nsz * sizeof(int), false, c->stream);
            if (!nh) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (vertex hash)");
            HVCK(hipMemsetAsync(nh, 0xff, nsz * sizeof(int), c->stream));
            if (nlive) {
                k_rehash_vertices<D><<<nblk(nlive, 256), 256, 0, c->stream>>>(ids, nlive, vkey, nh, nsz - 1);
                HVCK(hipGetLastError());
            }
            htab = nh;
            hsize = nsz;
        }
        return HV_OK;
    };
    // candidates at provisional ids base .. base+n with status > 0 -> hash -> new ids appended at ids[nlive]
    // cnt_dev: leave the count in device memory (no host wait) instead of returning it in nsel
    // n_dev: the candidate count in device memory (n then bounds it)
    auto insert_and_select = [&](const int* status, int n, int64_t base, int& nsel, int* cnt_dev = nullptr,
                                 const int* n_dev = nullptr) -> int {
        uint8_t* win = pget<uint8_t>(c, "win", n);
        if (!win) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (win)");
        k_insert<D><<<nblk(n, 256), 256, 0, c->stream>>>(status, n, n_dev, vkey, base, htab, hsize - 1, win, ictr,
                                                         HV_SCAN_SPREAD);
        HVCK(hipGetLastError());
        size_t tb = 0;
        hipcub::CountingInputIterator<int> it((int)base);
        int* cnt = cnt_dev ? cnt_dev : dcount;
        HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, win, ids + nlive, cnt, n, c->stream));
        void* tmp = c->pool->get("cubtmp2", tb, false, c->stream);
        if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
        HVCK(hipcub::DeviceSelect::Flagged(tmp, tb, it, win, ids + nlive, cnt, n, c->stream));
        if (cnt_dev) return HV_OK;
        HVCK(hipMemcpyAsync(&nsel, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        return HV_OK;
    };
    // edges of new vertices ids[first .. first+count); `bound` caps the new vertices the
    // level may still register (sizes the new-edge table at the level's first registration)
    // count_dev: the vertex count is in device memory and `count` only bounds it
    auto register_new = [&](const DevRays& R, int64_t first, int count, int64_t bound, int64_t* reg,
                            const int* count_dev = nullptr) -> int {
        if (nt_fresh) {
            nsize = hpow(2 * (D + 1) * std::max<int64_t>(bound, count) + 64);
            ntab = pget<unsigned long long>(c, "ntab", nsize);
            if (!ntab) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (new-edge table)");
            HVCK(hipMemsetAsync(ntab, 0, nsize * sizeof(unsigned long long), c->stream));
            nt_fresh = false;
        }
        if (count == 0) return HV_OK;
        k_register<D><<<nblk((int64_t)count * (D + 1), 256), 256, 0, c->stream>>>(
            ids + first, count, count_dev, vkey, nwalk, cmask, N, R, rtab, rsize - 1, ntab, nsize - 1, reg, ctr + 4);
        HVCK(hipGetLastError());
        return HV_OK;
    };

    // seeds: the 2^d box corners and interior vertices found by descent from generators in
    // Morton order (more fronts, fewer levels): from every generator up to HV_SEED_ALL_MAX of
    // them and in 2D, where fewer levels pay most (5D N=100k: 11 levels instead of 15 with
    // one per 8, 58 against 59 ms; 512 small meshes 8% faster); from one per HV_SEED_STRIDE
    // in larger sets, where the descents' own cost outgrows the levels they save (3D N=5M:
    // 91 against 102 ms).  Every generator's cell when only some cells are walked: their
    // union need not be connected through walked edges.
    const bool part = nwalk < N || subset;
    // over a forest every generator seeds (meshes of a batch are small)
    const int stride = (B || part || D == 2 || N <= HV_SEED_ALL_MAX) ? 1 : HV_SEED_STRIDE;
    const int ndesc = subset ? ncell : N / stride;
    const int nstart = nseed + ndesc;
    int64_t* reg_cur = pget<int64_t>(c, "regA", (size_t)nstart * (D + 1));
    if (!reg_cur) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    {
        rc = ensure_vertices(nstart);
        if (rc) return rc;
        int* status = pget<int>(c, "status", nstart);
        if (!status) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
        if (part) {   // a corner belongs to a cell that is not walked: status HV_ST_DROPPED (-1)
            if (nseed > 0) HVCK(hipMemsetAsync(status, 0xff, sizeof(int) * nseed, c->stream));
        } else if (nseed > 0) {
            k_seed_corners<D, B><<<nseed, 256, 0, c->stream>>>(g, F, vkey, vx, vnext, status, ccx, ccpl);
            HVCK(hipGetLastError());
        }
        if (ndesc > 0) {
            k_seed_descent<D, B><<<nblk((int64_t)ndesc * HV_DESC_LANES, 64), 64, 0, c->stream>>>(g, F, ndesc, stride, nwalk, dcells, vkey, vx,
                                                                          vnext + nseed, status + nseed,
                                                                          o ? o->seed : 0);
            HVCK(hipGetLastError());
        }
        int nsel = 0;
        rc = insert_and_select(status, nstart, vnext, nsel);
        if (rc) return rc;
        DevRays none{};
        rc = build_rt(none, 0);
        if (rc) return rc;
        rc = register_new(none, nlive, nsel, nsel, reg_cur);
        if (rc) return rc;
        vnext += nstart;
        nlive += nsel;
    }
    bool reg_is_a = true;
    int64_t fbeg = 0, fcount = nlive;   // frontier = ids[fbeg .. fbeg+fcount)
    double ms_ray = 0.0;
    int64_t levels = 0, skipped = 0;
    int ncar = 0;   // rays carried from the previous level (waiting rays whose edge stayed open)
    for (int phase = 0;; ++phase) {
    while (fcount > 0 || ncar > 0) {
        if (++levels > 1000000) return hv_fail(c, HV_ERR_CAPACITY, "walk did not terminate");
        const int64_t nslots = fcount * (D + 1) + ncar;
        if (nslots > (int64_t)INT32_MAX) return hv_fail(c, HV_ERR_CAPACITY, "level too large");
        DevRays R;
        R.slot = pget<int64_t>(c, "ray_slot", nslots);
        R.v = pget<int>(c, "ray_v", nslots);
        R.drop = pget<int>(c, "ray_drop", nslots);
        R.u = pget<double>(c, "ray_u", nslots * D);
        R.c = pget<double>(c, "ray_c", nslots);
        R.du = pget<double>(c, "ray_du", nslots);
        R.tinit = nullptr;
        R.code = pget<uint64_t>(c, "ray_code", nslots);
        R.t1 = pget<double>(c, "ray_t1", nslots);
        R.t2 = pget<double>(c, "ray_t2", nslots);
        R.i1 = pget<int>(c, "ray_i1", nslots);
        R.i2 = pget<int>(c, "ray_i2", nslots);
        uint64_t* code2 = pget<uint64_t>(c, "ray_code2", nslots);
        int* ord0 = pget<int>(c, "ray_ord0", nslots);
        int* ord = pget<int>(c, "ray_ord", nslots);
        int* cand = pget<int>(c, "ray_cand", nslots);
        int* ordc = pget<int>(c, "ray_ordc", nslots);
        uint8_t* keep = pget<uint8_t>(c, "ray_keep", nslots);
        int* status = pget<int>(c, "status", nslots);
        int* rwin = pget<int>(c, "ray_win", nslots);
        double* rt = pget<double>(c, "ray_t", nslots);
        int64_t* reg_next = pget<int64_t>(c, reg_is_a ? "regB" : "regA", nslots * (D + 1));
        if (!R.slot || !R.v || !R.drop || !R.u || !R.c || !R.du || !R.code || !code2 || !ord0 || !ord || !cand ||
            !R.t1 || !R.t2 || !R.i1 || !R.i2 ||
            !ordc || !keep || !status || !rwin || !rt || !reg_next)
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (rays)");
        HVCK(hipMemsetAsync(dcount, 0, sizeof(int), c->stream));
        if (fcount > 0) {
            k_emit<D><<<nblk(fcount * (D + 1), 256 * HV_EMIT_PER), 256, 0, c->stream>>>(
                ids + fbeg, (int)fcount, vkey, vx, g, ntab, reg_cur, R, dcount, plane_tol);
            HVCK(hipGetLastError());
        }
        int nr = 0;
        HVCK(hipMemcpyAsync(&nr, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        rc = read_ctr();
        if (rc) return rc;
        if (hctr[5]) return hv_fail(c, HV_ERR_CAPACITY, "edge hash overflow");
        const int nfresh = nr;
        if (nfresh > 0) {
            k_ray_dir<D, B><<<nblk(nfresh, 256), 256, 0, c->stream>>>(g, F, R, nfresh, vkey, vx, plane_tol);
            HVCK(hipGetLastError());
        }
        if (ncar > 0) {
            k_carry_in<D><<<nblk(ncar, 256), 256, 0, c->stream>>>(
                (const int64_t*)c->pool->get("car_slot", 0), (const int*)c->pool->get("car_v", 0),
                (const int*)c->pool->get("car_drop", 0), (const double*)c->pool->get("car_u", 0),
                (const double*)c->pool->get("car_c", 0), (const double*)c->pool->get("car_du", 0), ncar, R, nr);
            HVCK(hipGetLastError());
            nr += ncar;
            ncar = 0;
        }
        if (nr == 0) break;
        rc = build_rt(R, nr);
        if (rc) return rc;
        nt_fresh = true;
        // start phase, then sort rays by the Morton code of their certificate-ball centre
        k_ray_init<D, B><<<nblk(nr, 256), 256, 0, c->stream>>>(g, F, R, nr, vkey, vx, plane_tol);
        HVCK(hipGetLastError());
        {
            // over a forest the code holds the mesh above 32 Morton bits
            const int sb0 = B ? 0 : HV_RAY_SORT_LO, sb1 = B ? 32 + (B ? fo->mbits : 0) : 60;
            size_t tb = 0;
            HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, R.code, code2, ord0, ord, nr, sb0, sb1, c->stream));
            void* tmp = c->pool->get("cubtmp", tb, false, c->stream);
            if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
            k_iota<<<nblk(nr, 256), 256, 0, c->stream>>>(ord0, nr);
            HVCK(hipGetLastError());
            HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, R.code, code2, ord0, ord, nr, sb0, sb1, c->stream));
        }
        const int64_t level_begin = nlive;
        // scan a list of rays, then new vertices -> hash -> registered edges
        // defer: the new-vertex count stays in dcount[1] (the caller reads it with its next
        // wait and then advances nlive and the scan time)
        // n_dev (defer only): the ray count is in device memory and n bounds it
        auto process = [&](const int* list, int n, bool defer, const int* n_dev = nullptr) -> int {
            int r_ = ensure_vertices(n);
            if (r_) return r_;
            if constexpr (B) {
                // packets of one mesh each: rays and packets per mesh of the (mesh-ordered) list
                const int ppw = hv_packet_rays(n, tiles);
                int* pcnt = pget<int>(c, "pk_cnt", (size_t)nmesh);
                int* seg = pget<int>(c, "pk_seg", (size_t)nmesh + 1);
                int* pko = pget<int>(c, "pk_off", (size_t)nmesh + 1);
                if (!pcnt || !seg || !pko) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (packets)");
                k_pkt_count<<<nblk(nmesh, 64), 64, 0, c->stream>>>(list, n, n_dev, R.code, nmesh, pcnt);
                HVCK(hipGetLastError());
                k_pkt_plan<<<1, 64, 0, c->stream>>>(pcnt, nmesh, ppw, seg, pko);
                HVCK(hipGetLastError());
                HVCK(hipEventRecord(r0, c->stream));
                launch_scan<D, true>(prec, g, F, pko, seg, ppw, nblk(n, ppw) + (unsigned)nmesh, R, list, n, n_dev, vkey, vx,
                                     status, rwin, rt, plane_tol, sctr, c->stream, HV_SCAN_SPREAD);
            } else {
                HVCK(hipEventRecord(r0, c->stream));
                launch_scan1<D>(prec, g, R, list, n, n_dev, vkey, vx, status, rwin, rt, plane_tol, sctr, c->stream, tiles,
                                HV_SCAN_SPREAD);
            }
            HVCK(hipGetLastError());
            HVCK(hipEventRecord(r1, c->stream));
            k_finalize<D, B><<<nblk(n, 256), 256, 0, c->stream>>>(g, F, R, list, n, n_dev, vkey, vx, vnext, status, rwin);
            HVCK(hipGetLastError());
            int nsel = 0;
            if (defer) {
                r_ = insert_and_select(status, n, vnext, nsel, dcount + 1, n_dev);
                if (r_) return r_;
                r_ = register_new(R, nlive, n, n, reg_next + (nlive - level_begin) * (D + 1), dcount + 1);
                if (r_) return r_;
                vnext += n;   // provisional slots of all n; st.rays and nlive follow the caller's wait
                return HV_OK;
            }
            r_ = insert_and_select(status, n, vnext, nsel);
            if (r_) return r_;
            float ms = 0.f;
            HVCK(hipEventElapsedTime(&ms, r0, r1));
            ms_ray += ms;
            if (nsel || nt_fresh) {
                r_ = register_new(R, nlive, nsel, KCH > 1 ? (int64_t)nr : nsel,
                                  reg_next + (nlive - level_begin) * (D + 1));
                if (r_) return r_;
            }
            st.rays += n;
            nlive += nsel;
            vnext += n;
            return HV_OK;
        };
        auto select = [&](const int* in, const uint8_t* flags, int n, int* outl, int& nout) -> int {
            size_t tb = 0;
            HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb, in, flags, outl, dcount, n, c->stream));
            void* tmp = c->pool->get("cubtmp2", tb, false, c->stream);
            if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
            HVCK(hipcub::DeviceSelect::Flagged(tmp, tb, in, flags, outl, dcount, n, c->stream));
            HVCK(hipMemcpyAsync(&nout, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            return HV_OK;
        };
        if (opt.reserved0 <= 0) {
            // one scan per tentative end point (the ray start's best candidate); the other
            // rays of that end point wait and are carried to the next level if still open
            uint64_t tsz = 1024;
            while ((int64_t)tsz < 2 * (int64_t)nr) tsz <<= 1;
            int* ttab = pget<int>(c, "ttab", tsz);
            int* prim_of = pget<int>(c, "prim_of", nr);
            uint8_t* pflag = pget<uint8_t>(c, "pflag", nr);
            int64_t* cs = pget<int64_t>(c, "car_slot", nr);
            int* cv = pget<int>(c, "car_v", nr);
            int* cd = pget<int>(c, "car_drop", nr);
            double* cu = pget<double>(c, "car_u", (size_t)nr * D);
            double* cc = pget<double>(c, "car_c", nr);
            double* cdu = pget<double>(c, "car_du", nr);
            if (!ttab || !prim_of || !pflag || !cs || !cv || !cd || !cu || !cc || !cdu)
                return hv_fail(c, HV_ERR_DEVICE, "out of device memory (groups)");
            HVCK(hipMemsetAsync(ttab, 0xff, tsz * sizeof(int), c->stream));
            k_tent_group<D, B><<<nblk(nr, 256), 256, 0, c->stream>>>(g, F, R, nr, nfresh, vkey, ttab, tsz - 1, prim_of, pflag);
            HVCK(hipGetLastError());
            k_gather_flags<<<nblk(nr, 256), 256, 0, c->stream>>>(ord, nr, pflag, keep);
            HVCK(hipGetLastError());
            // the primary rays (count in dcount[2]) are scanned without a host wait: the kernels
            // read the count and nr bounds their grids
            {
                size_t tb = 0;
                HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb, ord, keep, ordc, dcount + 2, nr, c->stream));
                void* tmp = c->pool->get("cubtmp2", tb, false, c->stream);
                if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
                HVCK(hipcub::DeviceSelect::Flagged(tmp, tb, ord, keep, ordc, dcount + 2, nr, c->stream));
            }
            rc = process(ordc, nr, true, dcount + 2);
            if (rc) return rc;
            HVCK(hipMemsetAsync(dcount, 0, sizeof(int), c->stream));
            k_carry_out<D><<<nblk(nr, 256), 256, 0, c->stream>>>(R, nr, pflag, rtab, cs, cv, cd, cu, cc, cdu, dcount);
            HVCK(hipGetLastError());
            // one wait: carried rays (dcount[0]), new vertices (dcount[1]), scanned rays (dcount[2])
            int hc[3] = {0, 0, 0};
            HVCK(hipMemcpyAsync(hc, dcount, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            ncar = hc[0];
            const int np = hc[2];
            nlive += hc[1];
            st.rays += np;
            {
                float ms = 0.f;
                HVCK(hipEventElapsedTime(&ms, r0, r1));
                ms_ray += ms;
            }
            skipped += (int64_t)(nr - np) - ncar;
        } else {
            // interleaved chunks: 64-ray blocks over K chunks, edge registration after each
            const int K = std::min(KCH, nr);
            const int nfull = nr >> 6, rem = nr & 63;
            for (int ch = 0; ch < K; ++ch) {
                const int nb = nfull > ch ? (nfull - ch + K - 1) / K : 0;
                const int nc = nb * 64 + ((nfull % K == ch) ? rem : 0);
                if (nc <= 0) continue;
                k_chunk<<<nblk(nc, 256), 256, 0, c->stream>>>(ord, nr, ch, K, R.slot, rtab, cand, keep);
                HVCK(hipGetLastError());
                int n = nc;
                const int* rays_c = cand;
                if (ch > 0) {
                    rc = select(cand, keep, nc, ordc, n);
                    if (rc) return rc;
                    rays_c = ordc;
                    skipped += nc - n;
                }
                if (n == 0) continue;
                rc = process(rays_c, n, false);
                if (rc) return rc;
            }
        }
        reg_cur = reg_next;
        reg_is_a = !reg_is_a;
        fbeg = level_begin;
        fcount = nlive - level_begin;
    }
    if (B || !subset || ncell == 0) break;
    // listed cells without a vertex: an exact descent each (phase 0), or give up (phase 1)
    std::vector<int> miss;
    {
        int* mark = pget<int>(c, "cmark", (size_t)N);
        if (!mark) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (cells)");
        HVCK(hipMemsetAsync(mark, 0, sizeof(int) * N, c->stream));
        if (nlive) {
            k_mark_cells<D><<<nblk(nlive, 256), 256, 0, c->stream>>>(ids, (int)nlive, vkey, N, mark);
            HVCK(hipGetLastError());
        }
        std::vector<int> hmark((size_t)N);
        HVCK(hipMemcpyAsync(hmark.data(), mark, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        for (int i : *cells)
            if (!hmark[(size_t)i]) miss.push_back(i);
    }
    if (miss.empty()) break;
    if (phase > 0) return hv_fail(c, HV_ERR_DEGENERATE, "a listed cell has no certified vertex (input not in general position?)");
    {
        const int nm = (int)miss.size();
        rc = ensure_vertices(nm);
        if (rc) return rc;
        int* status = pget<int>(c, "status", nm);
        int* dmiss = pget<int>(c, "cmiss", nm);
        if (!status || !dmiss) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (cells)");
        HVCK(hipMemcpyAsync(dmiss, miss.data(), sizeof(int) * nm, hipMemcpyHostToDevice, c->stream));
        if constexpr (!B) k_seed_exact<D><<<nm, 256, 0, c->stream>>>(g, dmiss, nm, vkey, vx, vnext, status);
        HVCK(hipGetLastError());
        int nsel = 0;
        rc = insert_and_select(status, nm, vnext, nsel);
        if (rc) return rc;
        DevRays none{};
        rc = build_rt(none, 0);
        if (rc) return rc;
        nt_fresh = true;
        reg_cur = pget<int64_t>(c, reg_is_a ? "regA" : "regB", (size_t)std::max(1, nsel) * (D + 1));
        if (!reg_cur) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
        rc = register_new(none, nlive, nsel, nsel, reg_cur);
        if (rc) return rc;
        vnext += nm;
        fbeg = nlive;
        fcount = nsel;
        nlive += nsel;
    }
    }
    HVCK(hipEventRecord(e1, c->stream));
    rc = read_ctr();
    if (rc) return rc;
    {
        std::vector<unsigned long long> hs(12 * HV_SCAN_SPREAD);
        HVCK(hipMemcpyAsync(hs.data(), sctr, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        for (int k : {0, 1, 2, 6, 7, 8, 9, 10, 11}) {
            hctr[k] = 0;
            for (int q = 0; q < HV_SCAN_SPREAD; ++q) hctr[k] += hs[(size_t)k * HV_SCAN_SPREAD + q];
        }
    }
    float ms_all = 0.f;
    HVCK(hipEventElapsedTime(&ms_all, e0, e1));
    if (hctr[9]) return hv_fail(c, HV_ERR_CAPACITY, "vertex hash overflow");
    if (hctr[5]) return hv_fail(c, HV_ERR_CAPACITY, "edge hash overflow");
    if (hctr[11])
        return hv_fail(c, HV_ERR_DEGENERATE, "a new vertex is not in general position (more than d + 1 generators on its "
                                             "sphere within rounding, or a singular circumcentre system)");
    st.levels = levels;
    st.vertices = nlive;
    st.pairs_scanned = (int64_t)hctr[0];
    st.leaves_visited = (int64_t)hctr[1];
    st.near_ties = (int64_t)hctr[2];
    st.duplicates = (int64_t)hctr[8];
    st.rays_skipped = skipped;
    st.fallbacks = (int64_t)hctr[7];
    st.fp64_rechecks = (int64_t)hctr[6];
    st.retries = 0;
    st.unbounded = (int64_t)hctr[10];
    st.degenerate = (int64_t)hctr[11];
    st.max_t_relerr = 0.0;
    st.ms_total = ms_all;
    st.ms_raycast = ms_ray;
    if constexpr (B) {
        // the forest's meshes: vertices by mesh, keys in each mesh's own ids, one allocation
        const int nm = fo->nmesh;
        const int nl = (int)std::max<int64_t>(nlive, 1);
        int* mk0 = pget<int>(c, "s_mkey0", (size_t)nl);
        int* mk1 = pget<int>(c, "s_mkey1", (size_t)nl);
        int* sid = pget<int>(c, "s_ids", (size_t)nl);
        int64_t* doff = pget<int64_t>(c, "s_off", (size_t)nm + 1);
        int* dgoff = pget<int>(c, "s_goff", (size_t)nm);
        int* dn = pget<int>(c, "s_n", (size_t)nm);
        if (!mk0 || !mk1 || !sid || !doff || !dgoff || !dn) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (split)");
        if (nlive) {
            k_split_key<D><<<nblk(nlive, 256), 256, 0, c->stream>>>(ids, (int)nlive, vkey, F.mof, mk0);
            HVCK(hipGetLastError());
            size_t tb = 0;
            HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, mk0, mk1, ids, sid, (int)nlive, 0, fo->mbits, c->stream));
            void* tmp = c->pool->get("cubtmp", tb, false, c->stream);
            if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (split sort)");
            HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, mk0, mk1, ids, sid, (int)nlive, 0, fo->mbits, c->stream));
        }
        k_split_offsets<<<nblk(nm + 1, 256), 256, 0, c->stream>>>(mk1, (int)nlive, nm, doff);
        HVCK(hipGetLastError());
        std::vector<int64_t> hoff(nm + 1, 0);
        HVCK(hipMemcpyAsync(hoff.data(), doff, sizeof(int64_t) * (nm + 1), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        std::vector<int> hc(nm);
        int maxc = 1;
        for (int k = 0; k < nm; ++k) {
            hc[k] = (int)(hoff[k + 1] - hoff[k]);
            maxc = std::max(maxc, hc[k]);
        }
        const size_t kb = ((size_t)nlive * (D + 1) * sizeof(int) + 255) / 256 * 256;
        const size_t xb = ((size_t)nlive * D * sizeof(double) + 255) / 256 * 256;
        auto store = std::make_shared<DevArr>();
        store->device = c->device;
        if (hipMalloc(&store->p, kb + xb + sizeof(int) * (size_t)maxc) != hipSuccess) {
            (void)hipGetLastError();
            store->p = nullptr;
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory (batch meshes)");
        }
        store->n = kb + xb + sizeof(int) * (size_t)maxc;
        int* okey = (int*)store->p;
        double* ox = (double*)((char*)store->p + kb);
        int* oid = (int*)((char*)store->p + kb + xb);
        HVCK(hipMemcpyAsync(dgoff, fo->goff.data(), sizeof(int) * nm, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(dn, fo->n.data(), sizeof(int) * nm, hipMemcpyHostToDevice, c->stream));
        if (nlive) {
            k_split_copy<D><<<nblk(nlive, 256), 256, 0, c->stream>>>(sid, mk1, (int)nlive, vkey, vx, dgoff, dn, fo->ntot,
                                                                     okey, ox);
            HVCK(hipGetLastError());
        }
        k_iota<<<nblk(maxc, 256), 256, 0, c->stream>>>(oid, maxc);
        HVCK(hipGetLastError());
        HVCK(hipStreamSynchronize(c->stream));
        for (int k = 0; k < nm; ++k) {
            hv_mesh* mk = new hv_mesh();
            mk->pool = nullptr;
            mk->store = store;
            mk->device = c->device;
            mk->D = D;
            mk->N = fo->n[k];
            mk->nv = hc[k];
            mk->key = DBuf{okey + hoff[k] * (D + 1), sizeof(int) * (size_t)hc[k] * (D + 1)};
            mk->x = DBuf{ox + hoff[k] * D, sizeof(double) * (size_t)hc[k] * D};
            mk->ids = DBuf{oid, sizeof(int) * (size_t)hc[k]};
            mk->stats = st;
            mk->stats.vertices = hc[k];
            auto xw = std::make_shared<DevArr>();
            xw->device = c->device;
            xw->p = (double*)fo->xo->p + (int64_t)fo->goff[k] * D;
            xw->n = sizeof(double) * (size_t)fo->n[k] * D;
            xw->parent = fo->xo;
            mk->xw = xw;
            mk->P = 2 * D;
            mk->pbase.assign(fo->base[k], fo->base[k] + 2 * D * D);
            mk->pnormal.assign(fo->normal[k], fo->normal[k] + 2 * D * D);
            mk->pbc.assign(2 * D, 0);
            mk->nreal = fo->n[k];
            mk->extended = false;
            fo->out[k] = mk;
        }
        return HV_OK;
    }
    hv_mesh* m = new hv_mesh();
    m->pool = c->pool;
    m->device = c->device;
    m->D = D;
    m->N = N;
    m->nv = nlive;
    m->key = c->pool->take("vkey");
    m->x = c->pool->take("vx");
    m->ids = c->pool->take("ids");
    m->stats = st;
    hv_mesh_snapshot(c, m);
    if (subset) m->cells = std::make_shared<const std::vector<int>>(*cells);
    *out = m;
    return HV_OK;
}

template <int D>
int hv_walk_mesh(hv_ctx* c, const hv_opts* o, hv_mesh** out, int nwalk, const std::vector<int>* cells = nullptr) {
    return hv_walk_impl<D, false>(c, o, out, nwalk, cells, nullptr);
}

// Meshes of dimension D walked together: one forest of trees, one breadth-first walk whose
// levels hold the rays of every mesh, packets of one mesh each, in runs of at most
// HV_FOREST_VERTICES expected vertices.

template <int D>
int hv_build_forest_meshes(hv_ctx* c, const hv_opts* o, Forest& all) {
    all.out.assign(all.nmesh, nullptr);
    int rc = HV_OK;
    for (int k0 = 0; k0 < all.nmesh && rc == HV_OK;) {
        Forest fo;
        double est = 0.0;
        int k1 = k0;
        while (k1 < all.nmesh && (k1 == k0 || est + vertices_per_point(D) * all.n[k1] <= HV_FOREST_VERTICES)) {
            est += vertices_per_point(D) * all.n[k1];
            fo.xs.push_back(all.xs[k1]);
            fo.n.push_back(all.n[k1]);
            fo.base.push_back(all.base[k1]);
            fo.normal.push_back(all.normal[k1]);
            ++k1;
        }
        fo.nmesh = k1 - k0;
        fo.goff.assign(fo.nmesh, 0);
        fo.ntot = 0;
        for (int k = 0; k < fo.nmesh; ++k) {
            fo.goff[k] = (int)fo.ntot;
            fo.ntot += fo.n[k];
        }
        if (fo.ntot >= (1ll << 30)) { rc = hv_fail(c, HV_ERR_UNSUPPORTED, "too many generators in one batch"); break; }
        fo.mbits = 1;
        while ((1 << fo.mbits) < fo.nmesh) ++fo.mbits;
        fo.out.assign(fo.nmesh, nullptr);
        hv_mesh* none = nullptr;
        rc = hv_walk_impl<D, true>(c, o, &none, (int)fo.ntot, nullptr, &fo);
        for (int k = 0; k < fo.nmesh; ++k) all.out[k0 + k] = fo.out[k];
        k0 = k1;
    }
    if (rc)
        for (auto& m : all.out)
            if (m) { hv_mesh_free(m); m = nullptr; }
    return rc;
}

// ---- periodic boundaries ------------------------------------------------------------------
// A periodic box is meshed as the originals plus their images within a margin delta of the
// box ("ghosts"), in the box widened by delta in the periodic directions (Dirichlet), walking
// only the originals' cells.  The result is the periodic tessellation's cells of the
// originals when no vertex of them has a circumsphere reaching beyond the widened box (then
// every image that could lie inside it is present) and no circumradius exceeds delta (then the
// nearest generator of every point of the box is present, which the point-sample integrator
// relies on).  k_periodic_reach returns the largest of those two amounts over the mesh; the
// build is repeated with a wider margin while it exceeds delta.

template <int D>
__global__ __launch_bounds__(256) void k_periodic_reach(const int* __restrict__ ids, int nv, const int* __restrict__ vkey,
                                                        const double* __restrict__ vx, const double* __restrict__ xin,
                                                        HvBox box, int mask, unsigned long long* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double need = 0.0;
    if (t < nv) {
        const int v = ids[t];
        const int g0 = vkey[(int64_t)v * (D + 1)];   // the smallest id: a generator, never a plane
        double r2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = vx[(int64_t)v * D + k] - xin[(int64_t)g0 * D + k];
            r2 = fma(a, a, r2);
        }
        const double r = sqrt(r2);
        need = r;
#pragma unroll
        for (int k = 0; k < D; ++k)
            if ((mask >> k) & 1) {
                const double x = vx[(int64_t)v * D + k];
                need = fmax(need, fmax(x + r - box.hi[k], box.lo[k] - (x - r)));
            }
    }
    double rr = need;   // the circumradius alone (need starts from it)
    if (t < nv) {
        const int v = ids[t];
        const int g0 = vkey[(int64_t)v * (D + 1)];
        double r2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = vx[(int64_t)v * D + k] - xin[(int64_t)g0 * D + k];
            r2 = fma(a, a, r2);
        }
        rr = sqrt(r2);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        need = fmax(need, __shfl_xor(need, off));
        rr = fmax(rr, __shfl_xor(rr, off));
    }
    if ((threadIdx.x & 63) == 0 && need > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(need));
    if ((threadIdx.x & 63) == 0 && rr > 0.0) atomicMax(out + 1, (unsigned long long)__double_as_longlong(rr));
}

// every original must own a vertex: mark the generators below nreal found in the keys
template <int D>
__global__ __launch_bounds__(256) void k_mark_cells(const int* __restrict__ ids, int nv, const int* __restrict__ vkey,
                                                    int nreal, int* __restrict__ mark) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int* key = vkey + (int64_t)ids[t] * (D + 1);
#pragma unroll
    for (int a = 0; a <= D; ++a)
        if (key[a] < nreal) mark[key[a]] = 1;
}

__global__ __launch_bounds__(256) void k_count_unmarked(const int* __restrict__ mark, int n,
                                                        unsigned long long* __restrict__ cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    hv_wave_count(cnt, t < n && mark[t] == 0);
}

template <int D>
static int periodic_mesh(hv_ctx* c, const hv_opts* o, hv_mesh** out, int mask, const std::vector<int>* cells) {
    const int64_t n = c->nreal;
    const size_t obytes = sizeof(double) * D * n;
    if (!c->extended) {   // keep the originals: the working set is about to be replaced
        if (!c->xw) return hv_fail(c, HV_ERR_STATE, "no points set");
        c->xorig = c->xw;
    }
    if (!c->xorig) return hv_fail(c, HV_ERR_STATE, "the originals of the periodic set are gone");
    std::vector<double> hx((size_t)n * D);
    HVCK(hipMemcpyAsync(hx.data(), c->xorig->p, obytes, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    HvBox box{};
    double vol = 1.0;
    for (int k = 0; k < D; ++k) {   // planes 2k: x_k = hi, 2k+1: x_k = lo (checked by hv_box_periodic_mask)
        box.hi[k] = c->pbase[(2 * k) * D + k];
        box.lo[k] = c->pbase[(2 * k + 1) * D + k];
        vol *= box.hi[k] - box.lo[k];
    }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < D; ++k) {
            const double x = hx[(size_t)i * D + k];
            if (!(x >= box.lo[k] && x <= box.hi[k]))
                return hv_fail(c, HV_ERR_DOMAIN, "a node does not lie in the domain given by the boundary");
        }
    // first margin: about three times the mean circumradius of a Poisson-Delaunay simplex at this
    // density (n omega_d r^d / vol is Gamma(d, 1) distributed, mean d)
    const double omega = std::pow(M_PI, 0.5 * D) / std::tgamma(1.0 + 0.5 * D);
    double delta = 2.9 * std::pow((double)D * vol / (omega * (double)n), 1.0 / D);
    const std::vector<double> pbase = c->pbase;
    const std::vector<int16_t> pbc = c->pbc;
    for (int attempt = 0;; ++attempt) {
        // the extended set: originals (shift 0), then per original its images in [lo - delta, hi + delta]
        // expected size: the box widened by delta in the periodic directions, at the same density
        double grow = 1.0;
        for (int k = 0; k < D; ++k)
            if ((mask >> k) & 1) grow *= 1.0 + 2.0 * delta / (box.hi[k] - box.lo[k]);
        const size_t expect = (size_t)std::min(1.1 * grow * (double)n + 1024.0, (double)(1ll << 30));
        // admissible shifts per original and direction form one interval [lo_s, lo_s + cnt_s);
        // labels are int8 and the point-sample integrator packs them in 4 bits: at most 7 box
        // lengths each way
        const auto shifts_of = [&](const double* x, int* lo_s, int* cnt_s) -> int64_t {
            int64_t combos = 1;
            for (int k = 0; k < D; ++k) {
                lo_s[k] = 0;
                cnt_s[k] = 1;
                if (!((mask >> k) & 1)) continue;
                const double L = box.hi[k] - box.lo[k];
                const int m = (int)std::ceil(delta / L) + 1;
                int a = 1, b = -1;
                for (int s = -m; s <= m; ++s) {
                    const double y = x[k] + s * L;
                    if (y >= box.lo[k] - delta && y <= box.hi[k] + delta) { a = std::min(a, s); b = std::max(b, s); }
                }
                lo_s[k] = a;
                cnt_s[k] = b - a + 1;
                combos *= cnt_s[k];
            }
            return combos;
        };
        for (int k = 0; k < D; ++k)
            if (((mask >> k) & 1) && std::ceil(delta / (box.hi[k] - box.lo[k])) > 7.0)
                return hv_fail(c, HV_ERR_CAPACITY, "periodic images beyond 7 box lengths (too few points for this box)");
        int lo_s[6], cnt_s[6];
        int64_t total = n;
        for (int64_t i = 0; i < n; ++i) {
            total += shifts_of(&hx[(size_t)i * D], lo_s, cnt_s) - 1;
            if (total >= (1ll << 30) - 64) return hv_fail(c, HV_ERR_CAPACITY, "too many periodic images for 32-bit keys");
        }
        std::vector<double> ext;
        std::vector<int32_t> iorig;
        std::vector<int8_t> ishift;
        ext.reserve((size_t)total * D);
        iorig.reserve((size_t)total);
        ishift.reserve((size_t)total * D);
        ext.assign(hx.begin(), hx.end());
        iorig.resize((size_t)n);
        ishift.assign((size_t)n * D, 0);
        for (int64_t i = 0; i < n; ++i) iorig[(size_t)i] = (int32_t)i;
        for (int64_t i = 0; i < n; ++i) {
            const double* x = &hx[(size_t)i * D];
            const int64_t combos = shifts_of(x, lo_s, cnt_s);
            for (int64_t q = 0; q < combos; ++q) {
                int sh[6];
                int64_t r = q;
                bool zero = true;
                for (int k = 0; k < D; ++k) {
                    sh[k] = lo_s[k] + (int)(r % cnt_s[k]);
                    r /= cnt_s[k];
                    zero = zero && sh[k] == 0;
                }
                if (zero) continue;
                for (int k = 0; k < D; ++k) {
                    ext.push_back(x[k] + sh[k] * (box.hi[k] - box.lo[k]));
                    ishift.push_back((int8_t)sh[k]);
                }
                iorig.push_back((int32_t)i);
            }
        }
        const int64_t next = (int64_t)iorig.size();
        double* xin = hv_points_for_write(c, sizeof(double) * D * next);
        if (!xin) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (periodic images)");
        HVCK(hipMemcpyAsync(xin, ext.data(), sizeof(double) * D * next, hipMemcpyHostToDevice, c->stream));
        c->N = next;
        c->extended = true;
        c->delta = delta;
        auto img_o = std::make_shared<const std::vector<int32_t>>(std::move(iorig));
        auto img_s = std::make_shared<const std::vector<int8_t>>(std::move(ishift));
        c->img_orig = img_o;
        c->img_shift = img_s;
        // the widened box, Dirichlet, for the walk; the caller's planes come back afterwards
        for (int k = 0; k < D; ++k)
            if ((mask >> k) & 1) {
                c->pbase[(2 * k) * D + k] = box.hi[k] + delta;
                c->pbase[(2 * k + 1) * D + k] = box.lo[k] - delta;
                c->pbc[2 * k] = c->pbc[2 * k + 1] = 0;
            }
        hv_mesh* m = nullptr;
        int rc = hv_walk_mesh<D>(c, o, &m, (int)n, cells);
        c->pbase = pbase;
        c->pbc = pbc;
        if (rc) return rc;
        if (m->nv == 0) {
            hv_mesh_free(m);
            return hv_fail(c, HV_ERR_DEGENERATE, "no seed vertex of the periodic cells was found");
        }
        unsigned long long* rb = pget<unsigned long long>(c, "preach", 2);
        if (!rb) {
            hv_mesh_free(m);
            return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
        }
        HVCK(hipMemsetAsync(rb, 0, 2 * sizeof(unsigned long long), c->stream));
        if (m->nv > 0) {
            k_periodic_reach<D><<<nblk(m->nv, 256), 256, 0, c->stream>>>((const int*)m->ids.p, (int)m->nv,
                                                                         (const int*)m->key.p, (const double*)m->x.p,
                                                                         xin, box, mask, rb);
            HVCK(hipGetLastError());
        }
        unsigned long long hb[2] = {0, 0};
        HVCK(hipMemcpyAsync(hb, rb, sizeof hb, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        double need = 0.0;
        std::memcpy(&need, &hb[0], sizeof need);
        std::memcpy(&c->rmax, &hb[1], sizeof need);
        if (need <= delta * (1.0 - 1e-9) && cells) {
            // listed cells: the walk has given each of them a vertex
            hv_mesh_snapshot(c, m);
            *out = m;
            return HV_OK;
        }
        if (need <= delta * (1.0 - 1e-9)) {
            // a cell whose own seed was dropped is reached through a walked edge shared with
            // another original's cell; one with no such neighbour would be missing: refuse
            int* mark = pget<int>(c, "pmark", (size_t)n);
            unsigned long long* miss = pget<unsigned long long>(c, "pmiss", 1);
            if (!mark || !miss) {
                hv_mesh_free(m);
                return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
            }
            HVCK(hipMemsetAsync(mark, 0, sizeof(int) * n, c->stream));
            HVCK(hipMemsetAsync(miss, 0, sizeof(unsigned long long), c->stream));
            k_mark_cells<D><<<nblk(m->nv, 256), 256, 0, c->stream>>>((const int*)m->ids.p, (int)m->nv,
                                                                     (const int*)m->key.p, (int)n, mark);
            HVCK(hipGetLastError());
            k_count_unmarked<<<nblk(n, 256), 256, 0, c->stream>>>(mark, (int)n, miss);
            HVCK(hipGetLastError());
            unsigned long long hmiss = 0;
            HVCK(hipMemcpyAsync(&hmiss, miss, sizeof hmiss, hipMemcpyDeviceToHost, c->stream));
            HVCK(hipStreamSynchronize(c->stream));
            if (hmiss) {
                hv_mesh_free(m);
                return hv_fail(c, HV_ERR_DEGENERATE, "a periodic cell was not reached by the walk");
            }
            hv_mesh_snapshot(c, m);
            *out = m;
            return HV_OK;
        }
        hv_mesh_free(m);
        if (attempt >= 8) return hv_fail(c, HV_ERR_CAPACITY, "the periodic margin did not settle");
        delta = std::max(1.3 * delta, 1.05 * need);
    }
}

int hv_box_periodic_mask(hv_ctx* c, int D, int* mask) {
    *mask = 0;
    const char* why = nullptr;
    const int drc = hv_domain(c, D, &why);
    if (drc) return hv_fail(c, drc, why);
    for (int k = 0; k < D && k < c->P / 2; ++k) {
        const int a = c->pbc[2 * k], b = c->pbc[2 * k + 1];
        if (a <= 0 && b <= 0) continue;
        // cuboid's pairing (1-based): plane 2k-1 <-> 2k
        if (a != 2 * k + 2 || b != 2 * k + 1)
            return hv_fail(c, HV_ERR_UNSUPPORTED, "periodic planes must be the pairs of cuboid(d; periodic=...)");
        *mask |= 1 << k;
    }
    if (*mask && c->dom_general)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "periodic planes must be the pairs of cuboid(d; periodic=...)");
    if (!*mask) return check_box(c, D);
    std::vector<int16_t> keep = c->pbc;
    std::fill(c->pbc.begin(), c->pbc.end(), (int16_t)0);
    const int rc = check_box(c, D);
    c->pbc = keep;
    return rc;
}

template <int D>
int hv_build_mesh_impl(hv_ctx* c, const hv_opts* o, hv_mesh** out, const std::vector<int>* cells) {
    int mask = 0;
    const int rc = hv_box_periodic_mask(c, D, &mask);
    if (rc) return rc;
    if (!mask) return hv_walk_mesh<D>(c, o, out, (int)c->N, cells);
    return periodic_mesh<D>(c, o, out, mask, cells);
}

template <int D>
int hv_mesh_export_impl(const hv_mesh* m, int64_t* sig, double* coords) {
    const int64_t nv = m->nv;
    std::vector<int> ids(nv);
    if (hipMemcpy(ids.data(), m->ids.p, sizeof(int) * nv, hipMemcpyDeviceToHost) != hipSuccess) return HV_ERR_DEVICE;
    int64_t maxid = 0;
    for (int v : ids) maxid = std::max<int64_t>(maxid, v);
    std::vector<int> key((size_t)(maxid + 1) * (D + 1));
    std::vector<double> x((size_t)(maxid + 1) * D);
    if (hipMemcpy(key.data(), m->key.p, key.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return HV_ERR_DEVICE;
    if (hipMemcpy(x.data(), m->x.p, x.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return HV_ERR_DEVICE;
    std::vector<int64_t> order(nv);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        const int* ka = &key[(size_t)ids[a] * (D + 1)];
        const int* kb = &key[(size_t)ids[b] * (D + 1)];
        return std::lexicographical_compare(ka, ka + D + 1, kb, kb + D + 1);
    });
    for (int64_t i = 0; i < nv; ++i) {
        const int v = ids[order[i]];
        for (int j = 0; j <= D; ++j) sig[i * (D + 1) + j] = (int64_t)key[(size_t)v * (D + 1) + j] + 1;
        for (int k = 0; k < D; ++k) coords[i * D + k] = x[(size_t)v * D + k];
    }
    return HV_OK;
}

// ---------------------------------------------------------------------------------
// ray-level and query entry points (hv_raycast_batch, hv_nn_halfspace_batch, hv_inrange_batch)

template <int D>
__global__ __launch_bounds__(256) void k_setup_rays(int n, const int* __restrict__ vkey, const double* __restrict__ vx,
                             const int* __restrict__ drop, DevGeo g, DevRays R, double plane_tol) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    // key read straight from global memory: a private copy indexed by `pos` would live in scratch
    const int* key = vkey + (int64_t)t * (D + 1);
    double r[D];
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = vx[(int64_t)t * D + k];
    double u[D], c, du, rho;
    hv_ray_setup<D>(g, key, drop[t], r, plane_tol, u, c, du, rho);
    R.v[t] = t;
    R.drop[t] = drop[t];
#pragma unroll
    for (int k = 0; k < D; ++k) R.u[(int64_t)t * D + k] = u[k];
    R.c[t] = c;
    R.du[t] = du;
    double pe[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pe[k] = r[k] + rho * u[k];
    R.code[t] = hv_morton<D>(g, pe);
}


template <int D>
__global__ void k_nn_halfspace(DevGeo g, int nq, const double* __restrict__ q, const double* __restrict__ u,
                               const double* __restrict__ cc, int64_t* __restrict__ idx, double* __restrict__ dist) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    double p[D], uu[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { p[k] = q[(int64_t)t * D + k]; uu[k] = u[(int64_t)t * D + k]; }
    const double c = cc[t];
    double best = INFINITY;
    int bi = -1;
    int node = 1;
    while (node) {
        if (hv_box_d2<D>(g, node, p) <= best) {
            if (node < g.P2) { node = 2 * node; continue; }
            const int leaf = node - g.P2;
            const int j0 = leaf * g.L, j1 = min(j0 + g.L, g.n);
            for (int j = j0; j < j1; ++j) {
                double d2 = 0.0, ux = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double xk = g.xs[(int64_t)(j) * (D + 1) + k];
                    const double a = xk - p[k];
                    d2 = fma(a, a, d2);
                    ux = fma(uu[k], xk, ux);
                }
                if (!(ux > c)) continue;
                const int gi = g.sidx[j];
                if (d2 < best || (d2 == best && gi < bi)) { best = d2; bi = gi; }
            }
        }
        node = hv_next_node(node);
    }
    idx[t] = bi >= 0 ? (int64_t)bi + 1 : 0;
    dist[t] = bi >= 0 ? sqrt(best) : INFINITY;
}

template <int D>
__global__ void k_inrange(DevGeo g, int nq, const double* __restrict__ q, const double* __restrict__ rad,
                          const int64_t* __restrict__ ptr, int64_t* __restrict__ out, int64_t* __restrict__ counts) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    double p[D];
#pragma unroll
    for (int k = 0; k < D; ++k) p[k] = q[(int64_t)t * D + k];
    const double r2 = rad[t] * rad[t];
    int64_t cnt = 0;
    const int64_t o = out ? ptr[t] : 0;
    int node = 1;
    while (node) {
        if (hv_box_d2<D>(g, node, p) <= r2) {
            if (node < g.P2) { node = 2 * node; continue; }
            const int leaf = node - g.P2;
            const int j0 = leaf * g.L, j1 = min(j0 + g.L, g.n);
            for (int j = j0; j < j1; ++j) {
                double d2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double a = g.xs[(int64_t)(j) * (D + 1) + k] - p[k];
                    d2 = fma(a, a, d2);
                }
                if (d2 <= r2) {
                    if (out) out[o + cnt] = (int64_t)g.sidx[j] + 1;
                    ++cnt;
                }
            }
        }
        node = hv_next_node(node);
    }
    if (counts) counts[t] = cnt;
}

// Interior seed vertices (the reference's descent, src/raycast.jl:45-109, without its random
// restarts): from generator i = the one at sorted position (t + 1/2) * stride, d steps, each
// moving along a direction inside the current face (orthogonal to x_s - x_i for the members s
// found so far) to the first generator that becomes equidistant, its candidates the 16 nearest
// neighbours of the members.  A seed is kept only if it is certain: away from every boundary
// plane, and exactly its d + 1 generators within (1 + 1e-9) times its radius (tree count), so
// the walk, reaching the same vertex, gives it the same key; its coordinates are computed
// from the key as every other vertex's.  status[t] = HV_ST_FOUND kept, HV_ST_DROPPED not.
// HV_DESC_LANES lanes work on one seed: each takes every HV_DESC_LANES-th candidate of a list
// and every HV_DESC_LANES-th point of a leaf, and the group combines them with shuffles; the
// rest of a seed's state is held alike by its lanes, so the group's control flow is uniform.
template <int D, bool B>
__global__ __launch_bounds__(64) HV_SEED_OCC void k_seed_descent(DevGeo g0, GeoSet F, int nseed, int stride, int nwalk,
                                                     const int* __restrict__ cells, int* __restrict__ vkey,
                                                     double* __restrict__ vx, int64_t base, int* __restrict__ status,
                                                     uint64_t seed) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = gt / HV_DESC_LANES, sub = gt % HV_DESC_LANES;
    if (t >= nseed) return;   // a seed's lanes leave together
    // over a forest (B): from every generator t, in its mesh's tree
    const DevGeo& g = hv_geo<B>(g0, F, t);
    const int N = g.N;
    int S[D + 1];
    // from the listed cells, or from every stride-th generator in Morton order
    const int js = B ? g.rank[t] : (cells ? g.rank[cells[t]] : min(g.n - 1, t * stride + stride / 2));
    S[0] = g.sidx[js];
    if (S[0] >= nwalk) {   // only cells of generators below nwalk are walked
        if (sub == 0) status[t] = HV_ST_DROPPED;
        return;
    }
    double xi[D], p[D], q[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k) { xi[k] = g.xs[(int64_t)js * (D + 1) + k]; p[k] = xi[k]; }
    bool ok = true;
#pragma unroll
    for (int step = 0; step < D; ++step) {
        // direction: a pseudo-random vector drawn from (seed, t, step), orthogonalised against
        // the face's normals
        double u[D];
#pragma unroll
        for (int k = 0; k < D; ++k)
            u[k] = (double)(hv_mix64(((uint64_t)t << 8) + (uint64_t)(step * D + k) + 0x5eedull +
                                     seed * 0x9e3779b97f4a7c15ull) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
#pragma unroll
            for (int a = 0; a < step; ++a) {
                double dp = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) dp = fma(u[k], q[a][k], dp);
#pragma unroll
                for (int k = 0; k < D; ++k) u[k] = fma(-dp, q[a][k], u[k]);
            }
        }
        double nu = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nu = fma(u[k], u[k], nu);
        if (!(nu > 1e-12)) { ok = false; break; }
        nu = 1.0 / sqrt(nu);
#pragma unroll
        for (int k = 0; k < D; ++k) u[k] *= nu;
        double pi2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) pi2 = fma(p[k] - xi[k], p[k] - xi[k], pi2);
        double best = INFINITY;
        int bo = -1, bj = -1;
        for (int a = 0; a <= step; ++a) {
            const int* nb = g.knn + (int64_t)g.rank[S[a]] * HV_KNN;
            for (int m = sub; m < HV_KNN; m += HV_DESC_LANES) {
                const int j = nb[m];
                const int o = g.sidx[j];
                bool member = false;
                for (int b = 0; b <= step; ++b) member = member || (S[b] == o);
                if (member) continue;
                double den = 0.0, px2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double x = g.xs[(int64_t)j * (D + 1) + k];
                    den = fma(u[k], x - xi[k], den);
                    px2 = fma(p[k] - x, p[k] - x, px2);
                }
                if (!(den > 0.0)) continue;
                const double tt = (px2 - pi2) / (2.0 * den);
                if (tt < best || (tt == best && o < bo)) { best = tt; bo = o; bj = j; }
            }
        }
        // the group's (t, id) minimum: the order of the candidates does not change it
#pragma unroll
        for (int off = 1; off < HV_DESC_LANES; off <<= 1) {
            const double b2 = __shfl_xor(best, off, HV_DESC_LANES);
            const int o2 = __shfl_xor(bo, off, HV_DESC_LANES);
            const int j2 = __shfl_xor(bj, off, HV_DESC_LANES);
            if (b2 < best || (b2 == best && o2 >= 0 && (bo < 0 || o2 < bo))) { best = b2; bo = o2; bj = j2; }
        }
        if (bo < 0 || !(best > 0.0) || !(best < INFINITY)) { ok = false; break; }
        // a boundary plane reached first: not an interior seed
        for (int pl = 0; pl < g.P; ++pl) {
            double nu2 = 0.0, np = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) { nu2 = fma(g.pn[pl * D + k], u[k], nu2); np = fma(g.pn[pl * D + k], p[k], np); }
            if (nu2 > 0.0 && (g.ph[pl] - np) / nu2 <= best) ok = false;
        }
        if (!ok) break;
#pragma unroll
        for (int k = 0; k < D; ++k) p[k] = fma(best, u[k], p[k]);
        S[step + 1] = bo;
        if (step + 1 < D) {
            // normal of the new bisector, orthonormalised into the face basis
            double w[D];
#pragma unroll
            for (int k = 0; k < D; ++k) w[k] = g.xs[(int64_t)bj * (D + 1) + k] - xi[k];
#pragma unroll
            for (int rep = 0; rep < 2; ++rep) {
#pragma unroll
                for (int a = 0; a < step; ++a) {
                    double dp = 0.0;
#pragma unroll
                    for (int k = 0; k < D; ++k) dp = fma(w[k], q[a][k], dp);
#pragma unroll
                    for (int k = 0; k < D; ++k) w[k] = fma(-dp, q[a][k], w[k]);
                }
            }
            double nw = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) nw = fma(w[k], w[k], nw);
            if (!(nw > 1e-24)) { ok = false; break; }
            nw = 1.0 / sqrt(nw);
#pragma unroll
            for (int k = 0; k < D; ++k) q[step][k] = w[k] * nw;
        }
    }
    int key[D + 1];
    double xn[D];
    if (ok) {
        // sorted key (original ids), then the vertex exactly as the walk computes it
#pragma unroll
        for (int a = 0; a <= D; ++a) key[a] = S[a];
        for (int a = 1; a <= D; ++a)
            for (int b = a; b > 0 && key[b - 1] > key[b]; --b) { const int x = key[b]; key[b] = key[b - 1]; key[b - 1] = x; }
        ok = hv_vertex_coords<D>(g, key, xn);
    }
    if (ok) {
        // well inside the box, and exactly the key's generators on or within the sphere
        double r2 = 0.0, ext = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = xn[k] - g.xo[(int64_t)key[0] * D + k];
            r2 = fma(a, a, r2);
            ext = fmax(ext, g.dhi[k] - g.dlo[k]);
        }
        for (int pl = 0; pl < g.P && ok; ++pl) {
            double np = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) np = fma(g.pn[pl * D + k], xn[k], np);
            if (!(g.ph[pl] - np > 1e-9 * ext + sqrt(r2) * 1e-6)) ok = false;
        }
        const double rr = r2 * (1.0 + 2e-9);
        int cnt = 0;
        int node = 1;
        while (node && ok) {
            if (hv_box_d2<D>(g, node, xn) <= rr) {
                if (node < g.P2) { node = 2 * node; continue; }
                const int leaf = node - g.P2;
                const int j0 = leaf * g.L, j1 = min(j0 + g.L, g.n);   // sorted positions of this tree
                int c = 0;
                for (int j = j0 + sub; j < j1; j += HV_DESC_LANES) {
                    double d2 = 0.0;
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double a = g.xs[(int64_t)j * (D + 1) + k] - xn[k];
                        d2 = fma(a, a, d2);
                    }
                    c += (d2 <= rr) ? 1 : 0;
                }
#pragma unroll
                for (int off = 1; off < HV_DESC_LANES; off <<= 1) c += __shfl_xor(c, off, HV_DESC_LANES);
                cnt += c;
                if (cnt > D + 1) ok = false;
            }
            node = hv_next_node(node);
        }
        ok = ok && cnt == D + 1;
    }
    if (sub != 0) return;
    if (ok) {
        const int64_t id = base + t;
#pragma unroll
        for (int a = 0; a <= D; ++a) vkey[id * (D + 1) + a] = key[a];
#pragma unroll
        for (int k = 0; k < D; ++k) vx[id * D + k] = xn[k];
    }
    status[t] = ok ? HV_ST_FOUND : HV_ST_DROPPED;
}

// (t, id) minimum over a workgroup, smaller id first on equal t; every thread gets the result
__device__ __forceinline__ void hv_block_argmin(double& t, int& id, double* st, int* sid) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double t2 = __shfl_xor(t, off);
        const int i2 = __shfl_xor(id, off);
        if (t2 < t || (t2 == t && i2 < id)) { t = t2; id = i2; }
    }
    if ((tid & 63) == 0) { st[tid >> 6] = t; sid[tid >> 6] = id; }
    __syncthreads();
    const int nw = (int)(blockDim.x >> 6);
    t = st[0];
    id = sid[0];
    for (int w = 1; w < nw; ++w)
        if (st[w] < t || (st[w] == t && sid[w] < id)) { t = st[w]; id = sid[w]; }
    __syncthreads();
}

// Exact descent for a cell that the fast seeds and the walk left without a vertex
// (k_seed_descent takes its candidates from the 16 nearest neighbours, which can miss the
// first equidistant generator, and gives up when a boundary plane comes first): the same
// moves, each step's first equidistant generator found by a scan over every generator (one
// workgroup per cell), and boundary planes admitted as members, as the reference's descent
// casts its rays against the mirrors as well (src/raycast.jl:45-109).  The result is certified
// as in k_seed_descent: inside the box, away from the planes not in its key, and exactly its
// generators on or within (1 + 1e-9) times its radius (counted over every generator).
// status[t] = HV_ST_FOUND, or HV_ST_DROPPED when no vertex could be certified.
template <int D>
__global__ __launch_bounds__(256) void k_seed_exact(DevGeo g, const int* __restrict__ cells, int ncell,
                                                    int* __restrict__ vkey, double* __restrict__ vx, int64_t base,
                                                    int* __restrict__ status) {
    __shared__ double s_t[4];
    __shared__ int s_i[4];
    __shared__ int s_cnt;
    const int t = blockIdx.x;
    if (t >= ncell) return;   // uniform over the workgroup
    const int N = g.N, P = g.P, tid = threadIdx.x;
    const int cell = cells[t];
    int S[D + 1];
#pragma unroll
    for (int a = 0; a <= D; ++a) S[a] = cell;
    double xi[D], p[D], q[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k) { xi[k] = g.xo[(int64_t)cell * D + k]; p[k] = xi[k]; }
    // every thread computes the same moves from the same reduced results: `ok` stays uniform
    bool ok = true;
#pragma unroll
    for (int step = 0; step < D; ++step) {
        if (!ok) break;
        double u[D];
#pragma unroll
        for (int k = 0; k < D; ++k)
            u[k] = (double)(hv_mix64(((uint64_t)cell << 8) + (uint64_t)(step * D + k) + 0xe5acull) >> 11) *
                       (1.0 / 9007199254740992.0) - 0.5;
#pragma unroll
        for (int rep = 0; rep < 2; ++rep)
#pragma unroll
            for (int a = 0; a < step; ++a) {
                double dp = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) dp = fma(u[k], q[a][k], dp);
#pragma unroll
                for (int k = 0; k < D; ++k) u[k] = fma(-dp, q[a][k], u[k]);
            }
        double nu = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) nu = fma(u[k], u[k], nu);
        if (!(nu > 1e-12)) { ok = false; break; }
        nu = 1.0 / sqrt(nu);
#pragma unroll
        for (int k = 0; k < D; ++k) u[k] *= nu;
        double pi2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) pi2 = fma(p[k] - xi[k], p[k] - xi[k], pi2);
        double best = INFINITY;
        int bid = INT_MAX;
        for (int j = tid; j < g.n; j += (int)blockDim.x) {
            const int o = g.sidx[j];
            bool member = false;
#pragma unroll
            for (int b = 0; b <= D; ++b) member = member || (b <= step && S[b] == o);
            if (member) continue;
            double den = 0.0, px2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double x = g.xs[(int64_t)j * (D + 1) + k];
                den = fma(u[k], x - xi[k], den);
                px2 = fma(p[k] - x, p[k] - x, px2);
            }
            if (!(den > 0.0)) continue;
            const double tt = (px2 - pi2) / (2.0 * den);
            if (tt < best || (tt == best && o < bid)) { best = tt; bid = o; }
        }
        if (tid == 0)
            for (int pl = 0; pl < P; ++pl) {
                bool member = false;
#pragma unroll
                for (int b = 0; b <= D; ++b) member = member || (b <= step && S[b] == N + pl);
                if (member) continue;
                double nu2 = 0.0, np = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) { nu2 = fma(g.pn[pl * D + k], u[k], nu2); np = fma(g.pn[pl * D + k], p[k], np); }
                if (!(nu2 > 0.0)) continue;
                const double tt = (g.ph[pl] - np) / nu2;
                if (tt < best || (tt == best && N + pl < bid)) { best = tt; bid = N + pl; }
            }
        hv_block_argmin(best, bid, s_t, s_i);
        if (bid == INT_MAX || !(best > 0.0) || !(best < INFINITY)) { ok = false; break; }
#pragma unroll
        for (int k = 0; k < D; ++k) p[k] = fma(best, u[k], p[k]);
        S[step + 1] = bid;
        if (step + 1 < D) {
            // the new member's normal (bisector or plane), orthonormalised into the face basis
            double w[D];
#pragma unroll
            for (int k = 0; k < D; ++k) w[k] = bid < N ? g.xo[(int64_t)bid * D + k] - xi[k] : g.pn[(bid - N) * D + k];
#pragma unroll
            for (int rep = 0; rep < 2; ++rep)
#pragma unroll
                for (int a = 0; a < step; ++a) {
                    double dp = 0.0;
#pragma unroll
                    for (int k = 0; k < D; ++k) dp = fma(w[k], q[a][k], dp);
#pragma unroll
                    for (int k = 0; k < D; ++k) w[k] = fma(-dp, q[a][k], w[k]);
                }
            double nw = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) nw = fma(w[k], w[k], nw);
            if (!(nw > 1e-24)) { ok = false; break; }
            nw = 1.0 / sqrt(nw);
#pragma unroll
            for (int k = 0; k < D; ++k) q[step][k] = w[k] * nw;
        }
    }
    int key[D + 1];
    double xn[D];
    int ngm = 0;
    if (ok) {
        // sorted key: generators first, planes (ids N + p) after
#pragma unroll
        for (int a = 0; a <= D; ++a) key[a] = S[a];
        for (int a = 1; a <= D; ++a)
            for (int b = a; b > 0 && key[b - 1] > key[b]; --b) { const int x = key[b]; key[b] = key[b - 1]; key[b - 1] = x; }
#pragma unroll
        for (int a = 0; a <= D; ++a) ngm += key[a] < N ? 1 : 0;
        ok = hv_vertex_coords<D>(g, key, xn);
    }
    if (ok) {
        double r2 = 0.0, ext = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double a = xn[k] - g.xo[(int64_t)key[0] * D + k];
            r2 = fma(a, a, r2);
            ext = fmax(ext, g.dhi[k] - g.dlo[k]);
        }
        for (int pl = 0; pl < P && ok; ++pl) {
            bool in_key = false;
#pragma unroll
            for (int a = 0; a <= D; ++a) in_key = in_key || key[a] == N + pl;
            if (in_key) continue;
            double np = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) np = fma(g.pn[pl * D + k], xn[k], np);
            if (!(g.ph[pl] - np > 1e-9 * ext + sqrt(r2) * 1e-6)) ok = false;
        }
        const double rr = r2 * (1.0 + 2e-9);
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        int cnt = 0;
        for (int j = tid; j < g.n; j += (int)blockDim.x) {
            double d2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double a = g.xs[(int64_t)j * (D + 1) + k] - xn[k];
                d2 = fma(a, a, d2);
            }
            cnt += (d2 <= rr) ? 1 : 0;
        }
        if (cnt) atomicAdd(&s_cnt, cnt);
        __syncthreads();
        ok = ok && s_cnt == ngm;
    }
    if (tid == 0) {
        if (ok) {
            const int64_t id = base + t;
#pragma unroll
            for (int a = 0; a <= D; ++a) vkey[id * (D + 1) + a] = key[a];
#pragma unroll
            for (int k = 0; k < D; ++k) vx[id * D + k] = xn[k];
        }
        status[t] = ok ? HV_ST_FOUND : HV_ST_DROPPED;
    }
}

// Degenerate inclusion (src/raycast.jl:926-950): for a ray out of general position (status
// HV_ST_SINGULAR), every generator and boundary plane on the sphere of its new vertex
// r2 = r + t u, the reference's key of more than d + 1 entries.  A generator belongs when its
// distance from r2 is within 1e-10 (relative) of the edge's, a plane when r2 lies on it within
// 1e-10 of that radius.  Per ray, up to `maxk` sorted ids (key, 0-based) and their count
// (-1: more than maxk).
template <int D>
__global__ __launch_bounds__(128) void k_full_key(DevGeo g, DevRays R, const int* __restrict__ order, int nrays,
                                                  const int* __restrict__ vkey, const double* __restrict__ vx,
                                                  const int* __restrict__ status, const double* __restrict__ tout,
                                                  int maxk, int* __restrict__ okey, int* __restrict__ olen,
                                                  double* __restrict__ ox) {
    const int rid = blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= nrays) return;
    if (status[rid] != HV_ST_SINGULAR) { olen[rid] = 0; return; }
    const int ray = order[rid];
    const int v = R.v[ray];
    const int e0 = vkey[(int64_t)v * (D + 1) + (R.drop[ray] == 0 ? 1 : 0)];   // the edge's first generator
    const double t = tout[rid];
    double r2[D], R2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        r2[k] = fma(t, R.u[(int64_t)ray * D + k], vx[(int64_t)v * D + k]);
        const double a = g.xo[(int64_t)e0 * D + k] - r2[k];
        R2 = fma(a, a, R2);
        ox[(int64_t)rid * D + k] = r2[k];
    }
    const double lim = R2 * (1.0 + 2e-10);
    int* key = okey + (int64_t)rid * maxk;
    int n = 0;
    bool over = false;
    int node = 1;
    while (node) {
        if (hv_box_d2<D>(g, node, r2) <= lim) {
            if (node < g.P2) { node = 2 * node; continue; }
            const int leaf = node - g.P2;
            const int j0 = leaf * g.L, j1 = min(j0 + g.L, g.n);
            for (int j = j0; j < j1; ++j) {
                double d2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double a = g.xs[(int64_t)j * (D + 1) + k] - r2[k];
                    d2 = fma(a, a, d2);
                }
                if (d2 <= lim) {
                    if (n < maxk) key[n] = g.sidx[j];
                    else over = true;
                    ++n;
                }
            }
        }
        node = hv_next_node(node);
    }
    const double rad = sqrt(R2);
    for (int p = 0; p < g.P; ++p) {
        double dist = g.ph[p];
#pragma unroll
        for (int k = 0; k < D; ++k) dist = fma(-g.pn[p * D + k], r2[k], dist);
        if (fabs(dist) <= 1e-10 * rad) {
            if (n < maxk) key[n] = g.N + p;
            else over = true;
            ++n;
        }
    }
    // insertion sort (a key holds a few tens of ids at most)
    const int m = min(n, maxk);
    for (int a = 1; a < m; ++a)
        for (int b = a; b > 0 && key[b - 1] > key[b]; --b) { const int x = key[b]; key[b] = key[b - 1]; key[b - 1] = x; }
    olen[rid] = over ? -1 : n;
}

template <int D>
int hv_raycast_batch_impl(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                          const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status,
                          int32_t max_sig, int64_t* full_sig, int32_t* full_len) {
    if (precision != HV_PREC_FP64 && precision != HV_PREC_BF16 && precision != HV_PREC_FP32)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "precision: HV_PREC_FP64, HV_PREC_FP32 or HV_PREC_BF16");
    if (nrays == 0) return HV_OK;
    if (nrays > (1 << 29)) return hv_fail(c, HV_ERR_ARG, "too many rays");
    const int N = (int)c->N;
    const int P = c->P;
    int rc = check_box(c, D);
    if (rc) return rc;
    // host validation: sorted keys, an edge with a real generator
    std::vector<int> hkey((size_t)nrays * (D + 1));
    std::vector<int> hdrop(nrays);
    for (int64_t k = 0; k < nrays; ++k) {
        for (int j = 0; j <= D; ++j) {
            const int64_t e = sig[k * (D + 1) + j];
            if (e < 1 || e > (int64_t)N + P) return hv_fail(c, HV_ERR_ARG, "key element out of range");
            if (j && e <= sig[k * (D + 1) + j - 1]) return hv_fail(c, HV_ERR_ARG, "keys must be sorted and distinct");
            hkey[k * (D + 1) + j] = (int)(e - 1);
        }
        if (drop[k] < 0 || drop[k] > D) return hv_fail(c, HV_ERR_ARG, "drop position out of range");
        hdrop[k] = drop[k];
        const int first = (drop[k] == 0) ? hkey[k * (D + 1) + 1] : hkey[k * (D + 1)];
        if (first >= N) return hv_fail(c, HV_ERR_ARG, "the walked edge must contain a generator");
    }
    Tree T;
    rc = build_tree<D>(c, 32, T, precision);
    if (rc) return rc;
    record_tree(c, T.g, precision, 32);
    DevGeo& g = T.g;
    int* vkey = pget<int>(c, "b_vkey", (size_t)2 * nrays * (D + 1));
    double* vx = pget<double>(c, "b_vx", (size_t)2 * nrays * D);
    int* dd = pget<int>(c, "b_drop", nrays);
    DevRays R;
    R.v = pget<int>(c, "b_rv", nrays);
    R.drop = pget<int>(c, "b_rd", nrays);
    R.u = pget<double>(c, "b_ru", (size_t)nrays * D);
    R.c = pget<double>(c, "b_rc", nrays);
    R.du = pget<double>(c, "b_rdu", nrays);
    R.tinit = nullptr;
    R.code = pget<uint64_t>(c, "b_code", nrays);
    R.t1 = pget<double>(c, "b_t1", nrays);
    R.t2 = pget<double>(c, "b_t2", nrays);
    R.i1 = pget<int>(c, "b_i1", nrays);
    R.i2 = pget<int>(c, "b_i2", nrays);
    uint64_t* code2 = pget<uint64_t>(c, "b_code2", nrays);
    int* ord0 = pget<int>(c, "b_ord0", nrays);
    int* ord = pget<int>(c, "b_ord", nrays);
    int* status = pget<int>(c, "b_status", nrays);
    double* tout = pget<double>(c, "b_t", nrays);
    int* rwin = pget<int>(c, "b_win", nrays);
    unsigned long long* ctr = pget<unsigned long long>(c, "b_ctr", 16);
    if (!vkey || !vx || !dd || !R.v || !R.drop || !R.u || !R.c || !R.du || !R.code || !code2 || !ord0 || !ord ||
        !status || !tout || !ctr || !rwin || !R.t1 || !R.t2 || !R.i1 || !R.i2)
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    HVCK(hipMemcpyAsync(vkey, hkey.data(), sizeof(int) * hkey.size(), hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(vx, r, sizeof(double) * nrays * D, hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(dd, hdrop.data(), sizeof(int) * nrays, hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemsetAsync(ctr, 0, 16 * sizeof(unsigned long long), c->stream));
    HVCK(hipMemsetAsync(status, 0, sizeof(int) * nrays, c->stream));
    const double plane_tol = 1e-12;
    k_setup_rays<D><<<nblk(nrays, 256), 256, 0, c->stream>>>((int)nrays, vkey, vx, dd, g, R, plane_tol);
    HVCK(hipGetLastError());
    k_ray_init<D, false><<<nblk(nrays, 256), 256, 0, c->stream>>>(g, GeoSet{}, R, (int)nrays, vkey, vx, plane_tol);
    HVCK(hipGetLastError());
    size_t tb = 0;
    HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, R.code, code2, ord0, ord, (int)nrays, HV_RAY_SORT_LO, 60, c->stream));
    void* tmp = c->pool->get("cubtmp", tb, false, c->stream);
    if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (sort)");
    k_iota<<<nblk(nrays, 256), 256, 0, c->stream>>>(ord0, (int)nrays);
    HVCK(hipGetLastError());
    HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, R.code, code2, ord0, ord, (int)nrays, HV_RAY_SORT_LO, 60, c->stream));
    launch_scan1<D>(precision, g, R, ord, (int)nrays, nullptr, vkey, vx, status, rwin, tout, plane_tol, ctr, c->stream);
    HVCK(hipGetLastError());
    k_finalize<D, false><<<nblk(nrays, 256), 256, 0, c->stream>>>(g, GeoSet{}, R, ord, (int)nrays, nullptr, vkey, vx, nrays,
                                                                  status, rwin);
    HVCK(hipGetLastError());
    // full keys of the rays out of general position (hv_raycast_batch_keys)
    std::vector<int> hfk, hfl;
    std::vector<double> hfx;
    if (full_sig) {
        int* fk = pget<int>(c, "b_fkey", (size_t)nrays * max_sig);
        int* fl = pget<int>(c, "b_flen", nrays);
        double* fx = pget<double>(c, "b_fx", (size_t)nrays * D);
        if (!fk || !fl || !fx) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (full keys)");
        k_full_key<D><<<nblk(nrays, 128), 128, 0, c->stream>>>(g, R, ord, (int)nrays, vkey, vx, status, tout, max_sig,
                                                               fk, fl, fx);
        HVCK(hipGetLastError());
        hfk.resize((size_t)nrays * max_sig);
        hfl.resize(nrays);
        hfx.resize((size_t)nrays * D);
        HVCK(hipMemcpyAsync(hfk.data(), fk, sizeof(int) * hfk.size(), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipMemcpyAsync(hfl.data(), fl, sizeof(int) * nrays, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipMemcpyAsync(hfx.data(), fx, sizeof(double) * hfx.size(), hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<int> hord(nrays), hst(nrays), hk((size_t)nrays * (D + 1));
    std::vector<double> hx((size_t)nrays * D), ht(nrays);
    HVCK(hipMemcpyAsync(hord.data(), ord, sizeof(int) * nrays, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipMemcpyAsync(hst.data(), status, sizeof(int) * nrays, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipMemcpyAsync(hk.data(), vkey + (size_t)nrays * (D + 1), sizeof(int) * hk.size(), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipMemcpyAsync(hx.data(), vx + (size_t)nrays * D, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipMemcpyAsync(ht.data(), tout, sizeof(double) * nrays, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    for (int64_t rid = 0; rid < nrays; ++rid) {
        const int k = hord[rid];
        const int s = hst[rid];
        out_status[k] = s == HV_ST_FOUND ? 0 : (s == HV_ST_TIE ? 2 : (s == HV_ST_NONE ? 1 : -1));
        for (int j = 0; j <= D; ++j) out_sig[(int64_t)k * (D + 1) + j] = s > 0 ? (int64_t)hk[rid * (D + 1) + j] + 1 : 0;
        for (int q = 0; q < D; ++q) out_r[(int64_t)k * D + q] = s > 0 ? hx[rid * D + q] : NAN;
        out_t[k] = s > 0 ? ht[rid] : INFINITY;
        if (full_sig) {
            int64_t* fs = full_sig + (int64_t)k * max_sig;
            const int fl = hfl[rid];
            full_len[k] = s > 0 ? D + 1 : (s == HV_ST_SINGULAR ? fl : 0);
            for (int j = 0; j < max_sig; ++j) fs[j] = 0;
            if (s > 0) {
                for (int j = 0; j <= D && j < max_sig; ++j) fs[j] = out_sig[(int64_t)k * (D + 1) + j];
            } else if (s == HV_ST_SINGULAR && fl > 0) {
                for (int j = 0; j < fl; ++j) fs[j] = (int64_t)hfk[(size_t)rid * max_sig + j] + 1;
                for (int q = 0; q < D; ++q) out_r[(int64_t)k * D + q] = hfx[(size_t)rid * D + q];
                out_t[k] = ht[rid];
                out_status[k] = 3;
            }
        }
    }
    return HV_OK;
}

// ---- ray-level entry point without allocation -----------------------------------------------
// Device scratch of hv_raycast_batch_noalloc: per-ray state for at most HV_NA_MAX_RAYS rays (in
// any dimension up to 6) and the ray sort's temporary storage.  One copy per GPU, held until the
// process ends; host code takes its address from hipGetSymbolAddress (per device, once, when a
// context is created there), never from the symbol.  Calls on one GPU share it: two calls that
// overlap on one GPU race.
constexpr int HV_NA_MAX_RAYS = 32768;
constexpr size_t HV_NA_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[HV_NA_BYTES];

static std::mutex hv_na_mu;
static void* hv_na_ptr[64] = {};   // per device ordinal: the address of `scratch`

int hv_scratch_init(int device) {
    if (device < 0 || device >= 64) return HV_ERR_ARG;
    std::lock_guard<std::mutex> lk(hv_na_mu);
    if (hv_na_ptr[device]) return HV_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return HV_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return HV_ERR_DEVICE;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return HV_ERR_DEVICE;
    }
    hv_na_ptr[device] = p;
    return HV_OK;
}

// carve the scratch: 256-byte aligned pieces; false when they do not fit
struct NaCarve {
    unsigned char* base;
    size_t off = 0, cap;
    template <typename T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base + off);
        off += n * sizeof(T);
        return p;
    }
    bool ok() const { return off <= cap; }
};

// caller's rays (1-based int64 keys, drop positions, vertices) -> the walk's records, checked
template <int D>
__global__ void k_na_rays_in(int n, const int64_t* __restrict__ sig, const int32_t* __restrict__ drop,
                             const double* __restrict__ r, int N, int P, int* __restrict__ vkey, double* __restrict__ vx,
                             int* __restrict__ dd, int* __restrict__ bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    bool ok = true;
    int64_t prev = 0;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const int64_t e = sig[(int64_t)t * (D + 1) + j];
        ok = ok && e >= 1 && e <= (int64_t)N + P && e > prev;
        prev = e;
        vkey[(int64_t)t * (D + 1) + j] = ok ? (int)(e - 1) : 0;
    }
    const int p = drop[t];
    ok = ok && p >= 0 && p <= D;
    // the walked edge holds a generator: its smallest element
    if (ok) ok = (p == 0 ? sig[(int64_t)t * (D + 1) + 1] : sig[(int64_t)t * (D + 1)]) <= N;
    dd[t] = ok ? p : 0;
#pragma unroll
    for (int k = 0; k < D; ++k) vx[(int64_t)t * D + k] = r[(int64_t)t * D + k];
    if (!ok) atomicOr(bad, 1);
}

// results in the caller's ray order (ord: sorted position -> ray), as hv_raycast_batch returns them
template <int D>
__global__ void k_na_out(int n, const int* __restrict__ ord, const int* __restrict__ status, const int* __restrict__ vkey,
                         const double* __restrict__ vx, const double* __restrict__ tout, int64_t* __restrict__ out_sig,
                         double* __restrict__ out_r, double* __restrict__ out_t, int32_t* __restrict__ out_status) {
    const int rid = blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= n) return;
    const int k = ord[rid];
    const int s = status[rid];
    out_status[k] = s == HV_ST_FOUND ? 0 : (s == HV_ST_TIE ? 2 : (s == HV_ST_NONE ? 1 : -1));
#pragma unroll
    for (int j = 0; j <= D; ++j)
        out_sig[(int64_t)k * (D + 1) + j] = s > 0 ? (int64_t)vkey[((int64_t)n + rid) * (D + 1) + j] + 1 : 0;
#pragma unroll
    for (int q = 0; q < D; ++q) out_r[(int64_t)k * D + q] = s > 0 ? vx[((int64_t)n + rid) * D + q] : (double)NAN;
    out_t[k] = s > 0 ? tout[rid] : (double)INFINITY;
}

template <int D>
int hv_raycast_noalloc_impl(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                            const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status) {
    if (precision != HV_PREC_FP64 && precision != HV_PREC_BF16 && precision != HV_PREC_FP32)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "precision: HV_PREC_FP64, HV_PREC_FP32 or HV_PREC_BF16");
    if (nrays > HV_NA_MAX_RAYS) return hv_fail(c, HV_ERR_CAPACITY, "at most 32768 rays per call");
    int rc = check_box(c, D);
    if (rc) return rc;
    if (!c->tree_ok || c->tree_xw != c->xw || c->tree_leaf != 32 ||
        (precision != HV_PREC_FP64 && c->tree_tables != precision))
        return hv_fail(c, HV_ERR_STATE, "no search tree of these points for this precision: call hv_build_mesh or "
                                        "hv_raycast_batch with it first");
    if (nrays == 0) return HV_OK;
    void* sp = hv_na_ptr[c->device];
    if (!sp) return hv_fail(c, HV_ERR_STATE, "scratch not resolved on this device");
    const int n = (int)nrays;
    const DevGeo& g = c->tree_g;
    NaCarve cv{(unsigned char*)sp, 0, HV_NA_BYTES};
    int* vkey = cv.take<int>((size_t)2 * n * (D + 1));
    double* vx = cv.take<double>((size_t)2 * n * D);
    int* dd = cv.take<int>(n);
    DevRays R{};
    R.v = cv.take<int>(n);
    R.drop = cv.take<int>(n);
    R.u = cv.take<double>((size_t)n * D);
    R.c = cv.take<double>(n);
    R.du = cv.take<double>(n);
    R.code = cv.take<uint64_t>(n);
    R.t1 = cv.take<double>(n);
    R.t2 = cv.take<double>(n);
    R.i1 = cv.take<int>(n);
    R.i2 = cv.take<int>(n);
    uint64_t* code2 = cv.take<uint64_t>(n);
    int* ord0 = cv.take<int>(n);
    int* ord = cv.take<int>(n);
    int* status = cv.take<int>(n);
    double* tout = cv.take<double>(n);
    int* rwin = cv.take<int>(n);
    unsigned long long* ctr = cv.take<unsigned long long>(16);
    int* bad = cv.take<int>(1);
    size_t tb = 0;
    HVCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, R.code, code2, ord0, ord, n, HV_RAY_SORT_LO, 60, c->stream));
    void* tmp = cv.take<unsigned char>(tb);
    if (!cv.ok()) return hv_fail(c, HV_ERR_CAPACITY, "the rays do not fit the scratch array");
    HVCK(hipMemsetAsync(ctr, 0, 16 * sizeof(unsigned long long), c->stream));
    HVCK(hipMemsetAsync(status, 0, sizeof(int) * n, c->stream));
    HVCK(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    k_na_rays_in<D><<<nblk(n, 256), 256, 0, c->stream>>>(n, sig, drop, r, g.N, g.P, vkey, vx, dd, bad);
    HVCK(hipGetLastError());
    int hbad = 0;
    HVCK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    if (hbad) return hv_fail(c, HV_ERR_ARG, "a key is unsorted or out of range, a drop position is out of range, or "
                                           "a walked edge holds no generator");
    const double plane_tol = 1e-12;
    k_setup_rays<D><<<nblk(n, 256), 256, 0, c->stream>>>(n, vkey, vx, dd, g, R, plane_tol);
    HVCK(hipGetLastError());
    k_ray_init<D, false><<<nblk(n, 256), 256, 0, c->stream>>>(g, GeoSet{}, R, n, vkey, vx, plane_tol);
    HVCK(hipGetLastError());
    k_iota<<<nblk(n, 256), 256, 0, c->stream>>>(ord0, n);
    HVCK(hipGetLastError());
    HVCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, R.code, code2, ord0, ord, n, HV_RAY_SORT_LO, 60, c->stream));
    launch_scan1<D>(precision, g, R, ord, n, nullptr, vkey, vx, status, rwin, tout, plane_tol, ctr, c->stream);
    HVCK(hipGetLastError());
    k_finalize<D, false><<<nblk(n, 256), 256, 0, c->stream>>>(g, GeoSet{}, R, ord, n, nullptr, vkey, vx, nrays, status, rwin);
    HVCK(hipGetLastError());
    k_na_out<D><<<nblk(n, 256), 256, 0, c->stream>>>(n, ord, status, vkey, vx, tout, out_sig, out_r, out_t, out_status);
    HVCK(hipGetLastError());
    HVCK(hipStreamSynchronize(c->stream));
    return HV_OK;
}

template <int D>
int hv_nn_halfspace_impl(hv_ctx* c, int64_t nq, const double* q, const double* u, const double* cc,
                         int64_t* idx, double* dist) {
    if (nq == 0) return HV_OK;
    Tree T;
    int rc = build_tree<D>(c, 32, T);
    if (rc) return rc;
    double* dq = pget<double>(c, "q_q", (size_t)nq * D);
    double* du_ = pget<double>(c, "q_u", (size_t)nq * D);
    double* dc = pget<double>(c, "q_c", nq);
    int64_t* di = pget<int64_t>(c, "q_i", nq);
    double* dd = pget<double>(c, "q_d", nq);
    if (!dq || !du_ || !dc || !di || !dd) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    HVCK(hipMemcpyAsync(dq, q, sizeof(double) * nq * D, hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(du_, u, sizeof(double) * nq * D, hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(dc, cc, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
    k_nn_halfspace<D><<<nblk(nq, 128), 128, 0, c->stream>>>(T.g, (int)nq, dq, du_, dc, di, dd);
    HVCK(hipGetLastError());
    HVCK(hipMemcpyAsync(idx, di, sizeof(int64_t) * nq, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipMemcpyAsync(dist, dd, sizeof(double) * nq, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    return HV_OK;
}

template <int D>
int hv_inrange_impl(hv_ctx* c, int64_t nq, const double* q, const double* radius, int64_t* ptr, int64_t* idx) {
    ptr[0] = 0;
    if (nq == 0) return HV_OK;
    Tree T;
    int rc = build_tree<D>(c, 32, T);
    if (rc) return rc;
    double* dq = pget<double>(c, "q_q", (size_t)nq * D);
    double* dr = pget<double>(c, "q_r", nq);
    int64_t* dcnt = pget<int64_t>(c, "q_cnt", nq);
    int64_t* dptr = pget<int64_t>(c, "q_ptr", nq + 1);
    if (!dq || !dr || !dcnt || !dptr) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    HVCK(hipMemcpyAsync(dq, q, sizeof(double) * nq * D, hipMemcpyHostToDevice, c->stream));
    HVCK(hipMemcpyAsync(dr, radius, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
    if (!idx) {
        k_inrange<D><<<nblk(nq, 128), 128, 0, c->stream>>>(T.g, (int)nq, dq, dr, nullptr, nullptr, dcnt);
        HVCK(hipGetLastError());
        std::vector<int64_t> cnt(nq);
        HVCK(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int64_t) * nq, hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        for (int64_t k = 0; k < nq; ++k) ptr[k + 1] = ptr[k] + cnt[k];
        return HV_OK;
    }
    const int64_t total = ptr[nq];
    int64_t* dout = pget<int64_t>(c, "q_out", std::max<int64_t>(total, 1));
    if (!dout) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    HVCK(hipMemcpyAsync(dptr, ptr, sizeof(int64_t) * (nq + 1), hipMemcpyHostToDevice, c->stream));
    k_inrange<D><<<nblk(nq, 128), 128, 0, c->stream>>>(T.g, (int)nq, dq, dr, dptr, dout, nullptr);
    HVCK(hipGetLastError());
    if (total) HVCK(hipMemcpyAsync(idx, dout, sizeof(int64_t) * total, hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < nq; ++k) std::sort(idx + ptr[k], idx + ptr[k + 1]);
    return HV_OK;
}

// Point-sample Monte Carlo counts (k_mc_points) for the rows (ptr, pkey) of the context's
// mesh.  nsamples is rounded up to a whole number of strata of S samples (a multiple of 64);
// the number used, the slab width and the box volume are returned.
template <int D>
int hv_mc_points_impl(hv_ctx* c, int64_t nsamples, uint64_t seed, double slab, const int64_t* ptr,
                      const uint64_t* pkey, unsigned long long* ccnt, unsigned long long* pcnt, int64_t* nused,
                      double* eps_out, double* vbox_out, int64_t* pairs_out,
                      double* kernel_ms) {
    int pmask = 0;
    int rc = hv_box_periodic_mask(c, D, &pmask);
    if (rc) return rc;
    if (c->dom_general)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "point samples are drawn in a box: the planes must be those of cuboid(d; ...)");
    if (pmask && (!c->extended || !c->img_orig || !c->img_shift))
        return hv_fail(c, HV_ERR_STATE, "a periodic box needs its mesh built first");
    HvBox box{};
    double vbox = 1.0;
    for (int k = 0; k < D; ++k) {
        box.hi[k] = c->pbase[(2 * k) * D + k];
        box.lo[k] = c->pbase[(2 * k + 1) * D + k];
        vbox *= box.hi[k] - box.lo[k];
    }
    // the tree over the working set; with periodic images, in the box widened as for the walk
    const std::vector<double> pbase = c->pbase;
    const std::vector<int16_t> pbc = c->pbc;
    for (int k = 0; k < D; ++k)
        if ((pmask >> k) & 1) {
            c->pbase[(2 * k) * D + k] = box.hi[k] + c->delta;
            c->pbase[(2 * k + 1) * D + k] = box.lo[k] - c->delta;
            c->pbc[2 * k] = c->pbc[2 * k + 1] = 0;
        }
    Tree T;
    rc = build_tree<D>(c, 32, T);
    c->pbase = pbase;
    c->pbc = pbc;
    if (rc) return rc;
    const DevGeo& g = T.g;
    const uint32_t* pshift = nullptr;
    const int* porig = nullptr;
    const unsigned long long* pkeys = nullptr;
    const int* pvals = nullptr;
    uint64_t hmask = 0;
    const double eps = slab * std::pow(vbox / (double)c->nreal, 1.0 / D);
    if (pmask) {
        // a sample's second generator lies within rmax + 4 eps of it when it counts for an
        // interface; the working set must hold it
        if (c->rmax + 4.0 * eps > c->delta)
            return hv_fail(c, HV_ERR_UNSUPPORTED, "slab too wide for the margin of the periodic images");
        const int64_t ne = c->N;
        std::vector<uint32_t> code((size_t)ne);
        for (int64_t i = 0; i < ne; ++i) {
            uint32_t v = 0u;
            for (int k = 0; k < D; ++k) {
                const int sh = (*c->img_shift)[(size_t)i * D + k];
                if (sh < -8 || sh > 7) return hv_fail(c, HV_ERR_CAPACITY, "periodic image shift beyond 7 box lengths");
                v |= (uint32_t)(sh + 8) << (4 * k);
            }
            code[(size_t)i] = v;
        }
        // (original, shift code) -> id, open addressing at load <= 1/2
        uint64_t hs = 1024;
        while ((int64_t)hs < 2 * ne) hs <<= 1;
        std::vector<unsigned long long> hk(hs, ~0ull);
        std::vector<int> hv(hs, -1);
        for (int64_t i = 0; i < ne; ++i) {
            const unsigned long long key = ((unsigned long long)(unsigned)(*c->img_orig)[(size_t)i] << 32) | code[(size_t)i];
            uint64_t h = hv_mix64(key) & (hs - 1);
            while (hk[h] != ~0ull) h = (h + 1) & (hs - 1);
            hk[h] = key;
            hv[h] = (int)i;
        }
        uint32_t* dcode = pget<uint32_t>(c, "pshift", (size_t)ne);
        int* dorig = pget<int>(c, "porig", (size_t)ne);
        unsigned long long* dkeys = pget<unsigned long long>(c, "pkeys", hs);
        int* dvals = pget<int>(c, "pvals", hs);
        if (!dcode || !dorig || !dkeys || !dvals) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
        HVCK(hipMemcpyAsync(dcode, code.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(dorig, c->img_orig->data(), sizeof(int) * ne, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(dkeys, hk.data(), sizeof(unsigned long long) * hs, hipMemcpyHostToDevice, c->stream));
        HVCK(hipMemcpyAsync(dvals, hv.data(), sizeof(int) * hs, hipMemcpyHostToDevice, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
        pshift = dcode;
        porig = dorig;
        pkeys = dkeys;
        pvals = dvals;
        hmask = hs - 1;
    }
    int G = std::max(1, (int)std::floor(std::pow((double)nsamples / 256.0, 1.0 / D)));
    G = std::min(G, (int)std::floor(std::pow(2.0, 40.0 / D)));   // G^d strata stay below 2^40
    int64_t B = 1;
    for (int k = 0; k < D; ++k) B *= G;
    int64_t S = (nsamples + B - 1) / B;
    S = (S + 63) / 64 * 64;
    const int64_t n = B * S;
    // one work-item per sample, and a dispatch holds at most 2^32 work-items
    if (n > (1ll << 32) - 64) return hv_fail(c, HV_ERR_CAPACITY, "too many samples for one call (at most 2^32)");
    unsigned long long* npairs = pget<unsigned long long>(c, "mcp_pairs", HV_MCP_SPREAD);
    if (!npairs) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    HVCK(hipMemsetAsync(npairs, 0, sizeof(unsigned long long) * HV_MCP_SPREAD, c->stream));
    HVCK(hipEventRecord(c->ev[2], c->stream));
    k_mc_points<D><<<(unsigned)(n / 64), 64, 0, c->stream>>>(g, G, S, seed, eps, ptr, pkey, ccnt, pcnt, box, pmask,
                                                             (int)c->nreal, porig, pshift, pkeys, pvals, hmask, npairs);
    HVCK(hipGetLastError());
    HVCK(hipEventRecord(c->ev[3], c->stream));
    unsigned long long hpv[HV_MCP_SPREAD];
    HVCK(hipMemcpyAsync(hpv, npairs, sizeof(hpv), hipMemcpyDeviceToHost, c->stream));
    HVCK(hipStreamSynchronize(c->stream));
    unsigned long long hp = 0;
    for (int q = 0; q < HV_MCP_SPREAD; ++q) hp += hpv[q];
    float kms = 0.f;
    HVCK(hipEventElapsedTime(&kms, c->ev[2], c->ev[3]));
    *pairs_out = (int64_t)hp;
    *kernel_ms = kms;
    *nused = n;
    *eps_out = eps;
    *vbox_out = vbox;
    return HV_OK;
}

// ---------------------------------------------------------------------------------
// Refinement: the mesh of a point set X with new points Y appended (refine!,
// src/geometry.jl:404-438 -> systematic_refine!, src/meshrefine.jl:183-217).  A vertex of X's
// mesh stays when no new point lies inside its circumsphere, by the reference's rule: kept iff
// |x_sig[1] - r| <= (1 + 1e-7) * (distance from r to the nearest new point) (clean_affected!,
// src/meshrefine.jl:126-150, which tests only vertices whose generators are all affected
// cells; in general position a new point inside a circumsphere makes every generator of that
// vertex a neighbour of the new cell, so the tested set is the same).  Every other vertex of
// the refined mesh holds a new point and comes from the walk of the new cells.

// keep[t] = 1 when vertex ids[t] of the old mesh has no new point (sorted position j with
// sidx[j] >= nold) strictly inside its circumsphere shrunk by 1 + 1e-7
template <int D>
__global__ __launch_bounds__(128) void k_refine_keep(DevGeo g, int nv, const int* __restrict__ ids,
                                                     const int* __restrict__ vkey, const double* __restrict__ vx,
                                                     const double* __restrict__ xo, int nold,
                                                     uint8_t* __restrict__ keep) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int64_t v = ids[t];
    const int64_t g0 = vkey[v * (D + 1)];   // the smallest id of a key is a generator
    double r[D], R2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        r[k] = vx[v * D + k];
        const double a = xo[g0 * D + k] - r[k];
        R2 = fma(a, a, R2);
    }
    const double f = 1.0 + 1e-7;
    const double lim = R2 / (f * f);
    bool hit = false;
    int node = 1;
    while (node && !hit) {
        if (hv_box_d2<D>(g, node, r) < lim) {
            if (node < g.P2) { node = 2 * node; continue; }
            const int leaf = node - g.P2;
            const int j0 = leaf * g.L, j1 = min(j0 + g.L, g.n);
            for (int j = j0; j < j1; ++j) {
                if (g.sidx[j] < nold) continue;
                double d2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const double a = g.xs[(int64_t)j * (D + 1) + k] - r[k];
                    d2 = fma(a, a, d2);
                }
                if (d2 < lim) { hit = true; break; }
            }
        }
        node = hv_next_node(node);
    }
    keep[t] = hit ? 0 : 1;
}

// rows base .. base + n of the refined mesh from vertices sel[0 .. n) of a source table; ids at
// or above `from` (the old boundary planes) move up by `shift`
template <int D>
__global__ __launch_bounds__(256) void k_refine_copy(int n, const int* __restrict__ sel, const int* __restrict__ vkey,
                                                     const double* __restrict__ vx, int from, int shift, int base,
                                                     int* __restrict__ okey, double* __restrict__ ox,
                                                     int* __restrict__ oids) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t v = sel[t], o = (int64_t)base + t;
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const int k = vkey[v * (D + 1) + j];
        okey[o * (D + 1) + j] = k >= from ? k + shift : k;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) ox[o * D + k] = vx[v * D + k];
    oids[o] = (int)o;
}

// The context holds the union (the old mesh's nold generators, then the new ones) and S is
// the mesh of the new cells on it.  Out: the old mesh's kept vertices (plane ids shifted by the
// number of new points) followed by S's vertices, in tables of its own.
template <int D>
int hv_refine_merge(hv_ctx* c, const hv_mesh* old, const hv_mesh* S, hv_mesh** out, int64_t* removed) {
    const int nold = (int)old->N;
    const int shift = (int)(c->N - old->N);
    Tree T;
    int rc = build_tree<D>(c, 32, T);
    if (rc) return rc;
    const int nv0 = (int)old->nv;
    uint8_t* keep = pget<uint8_t>(c, "rf_keep", std::max(nv0, 1));
    int* sel = pget<int>(c, "rf_sel", std::max(nv0, 1));
    int* dcnt = pget<int>(c, "rf_cnt", 1);
    if (!keep || !sel || !dcnt) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (refine)");
    int nkeep = 0;
    if (nv0 > 0) {
        k_refine_keep<D><<<nblk(nv0, 128), 128, 0, c->stream>>>(T.g, nv0, (const int*)old->ids.p, (const int*)old->key.p,
                                                               (const double*)old->x.p, (const double*)c->xw->p, nold,
                                                               keep);
        HVCK(hipGetLastError());
        size_t tb = 0;
        HVCK(hipcub::DeviceSelect::Flagged(nullptr, tb, (const int*)old->ids.p, keep, sel, dcnt, nv0, c->stream));
        void* tmp = c->pool->get("cubtmp2", tb, false, c->stream);
        if (!tmp) return hv_fail(c, HV_ERR_DEVICE, "out of device memory (select)");
        HVCK(hipcub::DeviceSelect::Flagged(tmp, tb, (const int*)old->ids.p, keep, sel, dcnt, nv0, c->stream));
        HVCK(hipMemcpyAsync(&nkeep, dcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HVCK(hipStreamSynchronize(c->stream));
    }
    const int64_t nvs = S->nv, tot = (int64_t)nkeep + nvs;
    if (tot >= (1ll << 31)) return hv_fail(c, HV_ERR_CAPACITY, "refined mesh exceeds 2^31 vertices");
    hv_mesh* M = new hv_mesh;
    M->device = c->device;
    M->D = D;
    M->N = c->N;
    M->nv = tot;
    const size_t nk = (size_t)std::max<int64_t>(tot, 1);
    M->key.n = nk * (D + 1) * sizeof(int);
    M->x.n = nk * D * sizeof(double);
    M->ids.n = nk * sizeof(int);
    if (hipMalloc(&M->key.p, M->key.n) != hipSuccess || hipMalloc(&M->x.p, M->x.n) != hipSuccess ||
        hipMalloc(&M->ids.p, M->ids.n) != hipSuccess) {
        for (DBuf* b : {&M->key, &M->x, &M->ids})
            if (b->p) (void)hipFree(b->p);
        delete M;
        (void)hipGetLastError();
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (refined mesh)");
    }
    if (nkeep > 0)
        k_refine_copy<D><<<nblk(nkeep, 256), 256, 0, c->stream>>>(nkeep, sel, (const int*)old->key.p,
                                                                  (const double*)old->x.p, nold, shift, 0,
                                                                  (int*)M->key.p, (double*)M->x.p, (int*)M->ids.p);
    if (nvs > 0)
        k_refine_copy<D><<<nblk(nvs, 256), 256, 0, c->stream>>>((int)nvs, (const int*)S->ids.p, (const int*)S->key.p,
                                                                (const double*)S->x.p, INT_MAX, 0, nkeep,
                                                                (int*)M->key.p, (double*)M->x.p, (int*)M->ids.p);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        hv_mesh_free(M);
        return hv_fail(c, HV_ERR_DEVICE, "refine copy failed");
    }
    M->stats = S->stats;
    M->stats.vertices = tot;
    hv_mesh_snapshot(c, M);
    if (removed) *removed = nv0 - nkeep;
    *out = M;
    return HV_OK;
}

#define HV_INST(D)                                                                                 \
    template int hv_mc_points_impl<D>(hv_ctx*, int64_t, uint64_t, double, const int64_t*,         \
                                      const uint64_t*, unsigned long long*, unsigned long long*,   \
                                      int64_t*, double*, double*, int64_t*, double*);                                 \
    template int hv_build_mesh_impl<D>(hv_ctx*, const hv_opts*, hv_mesh**, const std::vector<int>*);   \
    template int hv_build_forest_meshes<D>(hv_ctx*, const hv_opts*, Forest&);                        \
    template int hv_mesh_export_impl<D>(const hv_mesh*, int64_t*, double*);                        \
    template int hv_refine_merge<D>(hv_ctx*, const hv_mesh*, const hv_mesh*, hv_mesh**, int64_t*);   \
    template int hv_raycast_batch_impl<D>(hv_ctx*, int, int64_t, const int64_t*, const int32_t*,   \
                                          const double*, int64_t*, double*, double*, int32_t*,     \
                                          int32_t, int64_t*, int32_t*);                            \
    template int hv_raycast_noalloc_impl<D>(hv_ctx*, int, int64_t, const int64_t*, const int32_t*, \
                                            const double*, int64_t*, double*, double*, int32_t*);  \
    template int hv_nn_halfspace_impl<D>(hv_ctx*, int64_t, const double*, const double*,           \
                                         const double*, int64_t*, double*);                        \
    template int hv_inrange_impl<D>(hv_ctx*, int64_t, const double*, const double*, int64_t*, int64_t*);
HV_INST(2)
HV_INST(3)
HV_INST(4)
HV_INST(5)
HV_INST(6)
