// hv_api.hip -- the extern "C" surface of libhv.so (declared in include/hv.h).
#include "hv_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <fstream>
#include <mutex>
#include <map>
#include <cstring>
#include <string>

// ---------------------------------------------------------------- device pool
DevPool::~DevPool() {
    (void)hipSetDevice(device);
    for (auto& kv : bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
}

void* DevPool::get(const std::string& name, size_t bytes, bool keep_contents, hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu);
    DBuf& b = bufs[name];
    if (bytes <= b.n && b.p) return b.p;
    if (bytes == 0) return b.p;
    size_t want = bytes + bytes / 4 + 256;
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (keep_contents && b.p && b.n) {
        if (hipMemcpyAsync(p, b.p, b.n, hipMemcpyDeviceToDevice, s) != hipSuccess) {
            (void)hipFree(p);
            return nullptr;
        }
        (void)hipStreamSynchronize(s);
    }
    if (b.p) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(b.p);
    }
    b.p = p;
    b.n = want;
    return p;
}

void* DevPool::reserve(const std::string& name, size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu);
    DBuf& b = bufs[name];
    if (bytes <= b.n && b.p) return b.p;
    if (b.p) {   // the old one first: the pair may not fit together
        (void)hipStreamSynchronize(s);
        (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    b.p = p;
    b.n = std::max<size_t>(bytes, 256);
    return p;
}

size_t DevPool::total() {
    std::lock_guard<std::mutex> lk(mu);
    size_t t = 0;
    for (auto& kv : bufs) t += kv.second.n;
    return t;
}

DBuf DevPool::take(const std::string& name) {
    std::lock_guard<std::mutex> lk(mu);
    DBuf b;
    auto it = bufs.find(name);
    if (it != bufs.end()) {
        b = it->second;
        bufs.erase(it);
    }
    return b;
}

void DevPool::give(const std::string& name, DBuf b) {
    if (!b.p) return;
    std::lock_guard<std::mutex> lk(mu);
    DBuf& cur = bufs[name];
    if (!cur.p) {
        cur = b;
    } else if (cur.n < b.n) {
        (void)hipFree(cur.p);
        cur = b;
    } else {
        (void)hipFree(b.p);
    }
}

DevArr::~DevArr() {
    if (!p || parent) return;
    int cur = 0;
    const bool had = hipGetDevice(&cur) == hipSuccess;
    (void)hipSetDevice(device);
    (void)hipFree(p);
    if (had) (void)hipSetDevice(cur);
}

double* hv_points_for_write(hv_ctx* c, size_t bytes) {
    if (c->xw && c->xw.use_count() == 1 && c->xw->n >= bytes) return (double*)c->xw->p;
    auto a = std::make_shared<DevArr>();
    a->device = c->device;
    if (hipMalloc(&a->p, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();
        a->p = nullptr;
        return nullptr;
    }
    a->n = bytes;
    c->xw = std::move(a);
    return (double*)c->xw->p;
}

void hv_mesh_snapshot(const hv_ctx* c, hv_mesh* m) {
    m->xw = c->xw;
    m->P = c->P;
    m->pbase = c->pbase;
    m->pnormal = c->pnormal;
    m->pbc = c->pbc;
    m->nreal = c->nreal;
    m->extended = c->extended;
    m->delta = c->delta;
    m->rmax = c->rmax;
    m->img_orig = c->img_orig;
    m->img_shift = c->img_shift;
}

hv_bind_mesh::hv_bind_mesh(hv_ctx* ctx, const hv_mesh* m)
    : c(ctx), D(ctx->D), N(ctx->N), nreal(ctx->nreal), xw(ctx->xw), P(ctx->P), pbase(ctx->pbase),
      pnormal(ctx->pnormal), pbc(ctx->pbc), extended(ctx->extended), delta(ctx->delta), rmax(ctx->rmax),
      img_orig(ctx->img_orig), img_shift(ctx->img_shift) {
    c->tree_ok = false;
    c->D = m->D;
    c->N = m->N;
    c->nreal = m->nreal;
    c->xw = m->xw;
    c->P = m->P;
    c->pbase = m->pbase;
    c->pnormal = m->pnormal;
    c->pbc = m->pbc;
    c->extended = m->extended;
    c->delta = m->delta;
    c->rmax = m->rmax;
    c->img_orig = m->img_orig;
    c->img_shift = m->img_shift;
}

hv_bind_mesh::~hv_bind_mesh() {
    c->tree_ok = false;   // a tree built meanwhile belongs to the mesh's generators
    c->D = D;
    c->N = N;
    c->nreal = nreal;
    c->xw = std::move(xw);
    c->P = P;
    c->pbase = std::move(pbase);
    c->pnormal = std::move(pnormal);
    c->pbc = std::move(pbc);
    c->extended = extended;
    c->delta = delta;
    c->rmax = rmax;
    c->img_orig = std::move(img_orig);
    c->img_shift = std::move(img_shift);
}

// HV_LOG_ERRORS set: every failure is also written to stderr (the contexts of a batch call are
// internal, so their messages are not otherwise visible)
int hv_fail(hv_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    static const bool log = std::getenv("HV_LOG_ERRORS") != nullptr;
    if (log) std::fprintf(stderr, "libhv: error %d: %s\n", code, msg.c_str());
    return code;
}

// ---------------------------------------------------------------- scan tiles (autotuner)
// Process-wide table (dimension, precision, size class) -> tiles, filled by hv_tune and from the
// JSON lines file named by HV_TUNE_CACHE (one object per line, the keys hv_tune writes).
static std::mutex tune_mu;
static std::map<std::tuple<int, int, int>, HvTiles> tune_table;
static bool tune_loaded = false;

static int size_class(int64_t n) {
    int b = 0;
    while (b < 62 && (1ll << (b + 1)) <= n) ++b;
    return b;
}

static bool json_int(const std::string& line, const char* key, long long& v) {
    const std::string k = std::string("\"") + key + "\":";
    const size_t p = line.find(k);
    if (p == std::string::npos) return false;
    return std::sscanf(line.c_str() + p + k.size(), "%lld", &v) == 1;
}

static void tune_load_locked() {
    if (tune_loaded) return;
    tune_loaded = true;
    const char* path = std::getenv("HV_TUNE_CACHE");
    if (!path || !*path) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        long long d, prec, cls, ppw, floor_;
        if (!json_int(line, "d", d) || !json_int(line, "precision", prec) || !json_int(line, "size_class", cls) ||
            !json_int(line, "rays_per_block", ppw) || !json_int(line, "packet_floor", floor_))
            continue;
        if (ppw != 8 && ppw != 16 && ppw != 32 && ppw != 64) continue;
        if (floor_ < 1 || floor_ > (1 << 20)) continue;
        tune_table[{(int)d, (int)prec, (int)cls}] = HvTiles{(int)ppw, (int)floor_};
    }
}

HvTiles hv_tiles_for(const hv_ctx* c, const hv_opts& o, int D, int precision, int64_t n) {
    (void)c;
    HvTiles t;
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        tune_load_locked();
        auto it = tune_table.find({D, precision, size_class(n)});
        if (it != tune_table.end()) t = it->second;
    }
    if (o.rays_per_block == 8 || o.rays_per_block == 16 || o.rays_per_block == 32 || o.rays_per_block == 64)
        t.ppw_max = o.rays_per_block;
    if (o.packet_floor > 0) t.min_packets = std::min(o.packet_floor, 1 << 20);
    if (o.packet_min == 1 || o.packet_min == 2 || o.packet_min == 4 || o.packet_min == 8) t.ppw_min = o.packet_min;
    t.ppw_min = std::min(t.ppw_min, t.ppw_max);
    return t;
}

// ---------------------------------------------------------------- dispatch helpers
#define HV_DISPATCH(D, fn, ...)                               \
    switch (D) {                                              \
        case 2: return fn<2>(__VA_ARGS__);                    \
        case 3: return fn<3>(__VA_ARGS__);                    \
        case 4: return fn<4>(__VA_ARGS__);                    \
        case 5: return fn<5>(__VA_ARGS__);                    \
        case 6: return fn<6>(__VA_ARGS__);                    \
        default: return HV_ERR_UNSUPPORTED;                   \
    }

// the refined mesh of hv_refine_mesh: a walk of the new cells on the union, merged with the old mesh
template <int D>
static int refine_d(hv_ctx* c, const hv_opts* o, const hv_mesh* old, hv_mesh** out, int64_t* removed) {
    std::vector<int> list;
    for (int64_t i = old->N; i < c->N; ++i) list.push_back((int)i);
    hv_mesh* S = nullptr;
    int rc = hv_build_mesh_impl<D>(c, o, &S, &list);
    if (rc) return rc;
    rc = hv_refine_merge<D>(c, old, S, out, removed);
    hv_mesh_free(S);
    return rc;
}

// Vertex labels of a mesh: per vertex its key with every generator id replaced by
// (original, shift) and every plane id by its plane, sorted: the same physical vertex in the
// same cell frame has the same label in meshes of different working sets.
static std::vector<std::vector<uint64_t>> mesh_labels(const hv_mesh* m, int* rc) {
    std::vector<std::vector<uint64_t>> out;
    const int D = m->D;
    const int64_t nv = m->nv;
    std::vector<int64_t> sig((size_t)std::max<int64_t>(nv, 1) * (D + 1));
    std::vector<double> x((size_t)std::max<int64_t>(nv, 1) * D);
    *rc = nv ? hv_mesh_export(m, sig.data(), x.data()) : HV_OK;
    if (*rc) return out;
    const bool per = m->img_orig && m->img_shift;
    out.resize((size_t)nv);
    for (int64_t v = 0; v < nv; ++v) {
        auto& row = out[(size_t)v];
        row.resize(D + 1);
        for (int a = 0; a <= D; ++a) {
            const int64_t id = sig[(size_t)v * (D + 1) + a] - 1;   // 0-based
            uint64_t code;
            if (id >= m->N) {
                code = (1ull << 62) | (uint64_t)(id - m->N);   // plane
            } else {
                const int64_t o = per ? (*m->img_orig)[(size_t)id] : id;
                uint64_t sh = 0;
                for (int k = 0; k < D; ++k)
                    sh |= (uint64_t)((per ? (*m->img_shift)[(size_t)id * D + k] : 0) + 8) << (4 * k);
                code = ((uint64_t)o << 24) | sh;
            }
            row[a] = code;
        }
        std::sort(row.begin(), row.end());
    }
    std::sort(out.begin(), out.end());
    return out;
}

extern "C" {

int hv_create(hv_ctx** out, int device_ordinal) {
    if (!out) return HV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HV_ERR_DEVICE;
    if (device_ordinal < 0 || device_ordinal >= ndev) return HV_ERR_ARG;
    if (hipSetDevice(device_ordinal) != hipSuccess) return HV_ERR_DEVICE;
    hv_ctx* c = new hv_ctx();
    c->device = device_ordinal;
    c->pool = std::make_shared<DevPool>();
    c->pool->device = device_ordinal;
    if (hv_scratch_init(device_ordinal) != HV_OK) {
        delete c;
        return HV_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return HV_ERR_DEVICE;
    }
    for (int i = 0; i < 4; ++i)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            delete c;
            return HV_ERR_DEVICE;
        }
    *out = c;
    return HV_OK;
}

int hv_destroy(hv_ctx* c) {
    if (!c) return HV_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(c->ev[i]);
    for (auto s : c->side) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (auto e : c->side_ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    c->xw.reset();
    c->xorig.reset();
    c->pool.reset();
    delete c;
    return HV_OK;
}

const char* hv_last_error(const hv_ctx* c) { return c ? c->err.c_str() : "null context"; }

static int set_points_common(hv_ctx* c, int d, int64_t n, const double* xs, hipMemcpyKind kind) {
    if (!c || !xs) return hv_fail(c, HV_ERR_ARG, "null argument");
    if (d < 2 || d > 6) return hv_fail(c, HV_ERR_UNSUPPORTED, "dimension must be in 2..6");
    if (n <= d) return hv_fail(c, HV_ERR_FEW_POINTS, "There are not enough points to create a Voronoi tessellation");
    if (n >= (1ll << 30)) return hv_fail(c, HV_ERR_UNSUPPORTED, "too many generators for 32-bit keys");
    (void)hipSetDevice(c->device);
    void* p = hv_points_for_write(c, sizeof(double) * d * n);
    if (!p) return hv_fail(c, HV_ERR_DEVICE, "out of device memory");
    if (hipMemcpyAsync(p, xs, sizeof(double) * d * n, kind, c->stream) != hipSuccess)
        return hv_fail(c, HV_ERR_DEVICE, "copy of the points failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return hv_fail(c, HV_ERR_DEVICE, "stream sync failed");
    if (c->D != d) {
        c->P = 0;
        c->pbase.clear();
        c->pnormal.clear();
        c->pbc.clear();
    }
    c->D = d;
    c->N = n;
    c->nreal = n;
    c->extended = false;
    c->tree_ok = false;
    c->xorig.reset();
    c->img_orig.reset();
    c->img_shift.reset();
    c->delta = c->rmax = 0.0;
    return HV_OK;
}

// back from the extended set of a periodic build to the originals
static int restore_originals(hv_ctx* c) {
    if (!c->extended) return HV_OK;
    if (!c->xorig) return hv_fail(c, HV_ERR_STATE, "the originals of the periodic set are gone");
    c->xw = std::move(c->xorig);
    c->tree_ok = false;
    c->N = c->nreal;
    c->extended = false;
    c->img_orig.reset();
    c->img_shift.reset();
    c->delta = c->rmax = 0.0;
    return HV_OK;
}

int hv_set_points(hv_ctx* c, int d, int64_t n, const double* xs) {
    return set_points_common(c, d, n, xs, hipMemcpyHostToDevice);
}

int hv_set_points_device(hv_ctx* c, int d, int64_t n, const double* xs_dev) {
    return set_points_common(c, d, n, xs_dev, hipMemcpyDeviceToDevice);
}

int hv_set_boundary(hv_ctx* c, int nplanes, const double* base, const double* normal, const int16_t* bc) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "hv_set_points must come first");
    if (nplanes < 0 || (nplanes > 0 && (!base || !normal))) return hv_fail(c, HV_ERR_ARG, "bad planes");
    const int D = c->D;
    (void)hipSetDevice(c->device);
    const int rc = restore_originals(c);
    if (rc) return rc;
    // unit normals: a plane's mirror image x + 2<n, b - x> n (src/boundary.jl:206-211) and its
    // ray intersections assume them
    for (int i = 0; i < nplanes; ++i) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += normal[(size_t)i * D + k] * normal[(size_t)i * D + k];
        if (!(std::fabs(s - 1.0) <= 1e-12)) return hv_fail(c, HV_ERR_ARG, "boundary normals must have unit length");
    }
    c->tree_ok = false;
    c->P = nplanes;
    c->pbase.assign(base, base + (size_t)nplanes * D);
    c->pnormal.assign(normal, normal + (size_t)nplanes * D);
    c->pbc.assign(nplanes, 0);
    if (bc)
        for (int i = 0; i < nplanes; ++i) c->pbc[i] = bc[i];
    return HV_OK;
}

int hv_build_mesh(hv_ctx* c, const hv_opts* o, hv_mesh** out) {
    if (!c || !out) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_build_mesh_impl, c, o, out, nullptr)
}

int hv_tune(hv_ctx* c, int precision, const char* cache_path, int32_t* rays_per_block, int32_t* packet_floor,
            double* ms) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (precision != HV_PREC_FP64 && precision != HV_PREC_FP32 && precision != HV_PREC_BF16)
        return hv_fail(c, HV_ERR_UNSUPPORTED, "precision: HV_PREC_FP64, HV_PREC_FP32 or HV_PREC_BF16");
    int mask = 0;
    (void)hipSetDevice(c->device);
    int rc = hv_box_periodic_mask(c, c->D, &mask);
    if (rc) return rc;
    if (mask) return hv_fail(c, HV_ERR_UNSUPPORTED, "the autotuner takes a non-periodic box");
    const int ppws[2] = {32, 64};
    const int floors[3] = {1024, 2048, 4096};
    HvTiles best;
    double best_ms = 1e300;
    hv_opts o{};
    o.precision = precision;
    for (int warm = 0; warm < 2; ++warm)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 3; ++b) {
                o.rays_per_block = ppws[a];
                o.packet_floor = floors[b];
                hv_mesh* m = nullptr;
                rc = hv_build_mesh(c, &o, &m);
                if (rc) return rc;
                const double t = m->stats.ms_total;
                hv_mesh_free(m);
                if (warm == 0) continue;   // the first pass pages the tables in
                if (t < best_ms) {
                    best_ms = t;
                    best = HvTiles{ppws[a], floors[b]};
                }
            }
    const int cls = size_class(c->N);
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        tune_load_locked();
        tune_table[{c->D, precision, cls}] = best;
    }
    if (cache_path && *cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        if (!f) return hv_fail(c, HV_ERR_ARG, "cannot write the tune cache");
        f << "{\"d\":" << c->D << ",\"precision\":" << precision << ",\"size_class\":" << cls
          << ",\"rays_per_block\":" << best.ppw_max << ",\"packet_floor\":" << best.min_packets
          << ",\"ms\":" << best_ms << "}\n";
    }
    if (rays_per_block) *rays_per_block = best.ppw_max;
    if (packet_floor) *packet_floor = best.min_packets;
    if (ms) *ms = best_ms;
    return HV_OK;
}

int hv_build_mesh_cells(hv_ctx* c, const hv_opts* o, int64_t ncells, const int64_t* cells, hv_mesh** out) {
    if (!c || !out || ncells < 0 || (ncells > 0 && !cells)) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    const int64_t n = c->nreal;   // the caller's generators (originals of a periodic box)
    std::vector<int> list;
    list.reserve((size_t)ncells);
    for (int64_t k = 0; k < ncells; ++k) {
        if (cells[k] < 1 || cells[k] > n) return hv_fail(c, HV_ERR_ARG, "cell index out of range (1..n)");
        list.push_back((int)(cells[k] - 1));
    }
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_build_mesh_impl, c, o, out, &list)
}

// refine! of a periodic mesh (src/meshrefine.jl:183-217 on the reference's ghost nodes): the
// originals and the new points are meshed afresh (the new points' images change the working
// set); `removed` counts the old vertices, by label, that the new mesh does not hold
static int refine_periodic(hv_ctx* c, const hv_opts* o, const hv_mesh* old, int64_t m, const double* xs_new,
                           hv_mesh** out, int64_t* removed) {
    const int D = old->D;
    const int64_t n = old->nreal, nu = n + m;
    if (nu >= (1ll << 30)) return hv_fail(c, HV_ERR_UNSUPPORTED, "too many generators for 32-bit keys");
    double* u = nullptr;
    if (hipMalloc(&u, sizeof(double) * D * nu) != hipSuccess) {
        (void)hipGetLastError();
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (refine)");
    }
    int rc = HV_OK;
    // the originals lead the mesh's working set
    if (hipMemcpyAsync(u, old->xw->p, sizeof(double) * D * n, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(u + D * n, xs_new, sizeof(double) * D * m, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = hv_fail(c, HV_ERR_DEVICE, "copy of the points failed");
    if (!rc) rc = set_points_common(c, D, nu, u, hipMemcpyDeviceToDevice);
    (void)hipFree(u);
    if (rc) return rc;
    rc = hv_set_boundary(c, old->P, old->P ? old->pbase.data() : nullptr, old->P ? old->pnormal.data() : nullptr,
                         old->pbc.empty() ? nullptr : old->pbc.data());
    if (rc) return rc;
    hv_mesh* M = nullptr;
    rc = hv_build_mesh(c, o, &M);
    if (rc) return rc;
    if (removed) {
        int lrc = HV_OK;
        const auto a = mesh_labels(old, &lrc);
        const auto b = lrc ? decltype(a)() : mesh_labels(M, &lrc);
        if (lrc) {
            hv_mesh_free(M);
            return hv_fail(c, lrc, "export of the meshes failed");
        }
        int64_t gone = 0;
        size_t j = 0;
        for (const auto& row : a) {
            while (j < b.size() && b[j] < row) ++j;
            if (j == b.size() || b[j] != row) ++gone;
        }
        *removed = gone;
    }
    *out = M;
    return HV_OK;
}

int hv_refine_mesh(hv_ctx* c, const hv_opts* o, const hv_mesh* old, int64_t m, const double* xs_new, hv_mesh** out,
                   int64_t* removed) {
    if (!c || !old || !out || m < 1 || !xs_new) return hv_fail(c, HV_ERR_ARG, "null argument or no new points");
    *out = nullptr;
    if (old->device != c->device) return hv_fail(c, HV_ERR_ARG, "the mesh lives on another device");
    if (old->cells) return hv_fail(c, HV_ERR_UNSUPPORTED, "refinement of a mesh of some cells");
    if (!old->xw) return hv_fail(c, HV_ERR_STATE, "the mesh holds no generators");
    (void)hipSetDevice(c->device);
    if (old->extended || old->img_orig) return refine_periodic(c, o, old, m, xs_new, out, removed);
    if (!old->xw) return hv_fail(c, HV_ERR_STATE, "the mesh holds no generators");
    const int D = old->D;
    const int64_t n = old->N, nu = n + m;
    if (nu >= (1ll << 30)) return hv_fail(c, HV_ERR_UNSUPPORTED, "too many generators for 32-bit keys");
    (void)hipSetDevice(c->device);
    // the union in device memory: the mesh's generators, then the new points
    double* u = nullptr;
    if (hipMalloc(&u, sizeof(double) * D * nu) != hipSuccess) {
        (void)hipGetLastError();
        return hv_fail(c, HV_ERR_DEVICE, "out of device memory (refine)");
    }
    int rc = HV_OK;
    if (hipMemcpyAsync(u, old->xw->p, sizeof(double) * D * n, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(u + D * n, xs_new, sizeof(double) * D * m, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = hv_fail(c, HV_ERR_DEVICE, "copy of the points failed");
    if (!rc) rc = set_points_common(c, D, nu, u, hipMemcpyDeviceToDevice);
    (void)hipFree(u);
    if (rc) return rc;
    rc = hv_set_boundary(c, old->P, old->P ? old->pbase.data() : nullptr, old->P ? old->pnormal.data() : nullptr,
                         old->pbc.empty() ? nullptr : old->pbc.data());
    if (rc) return rc;
    HV_DISPATCH(D, refine_d, c, o, old, out, removed)
}

// A mesh of a batch that the fused walk takes: a cuboid in the reference's plane order without
// periodic sides, n > d, every point inside the box (else its status is set and it is left out).
static int fusable(int d, int64_t n, const double* xs, int npl, const double* base, const double* normal,
                   const int16_t* bc, bool& periodic) {
    periodic = false;
    if (d < 2 || d > 6) return HV_ERR_UNSUPPORTED;
    if (!xs || (npl > 0 && (!base || !normal))) return HV_ERR_ARG;
    if (n <= d) return HV_ERR_FEW_POINTS;
    if (n >= (1ll << 30)) return HV_ERR_UNSUPPORTED;
    if (npl != 2 * d) return HV_ERR_UNSUPPORTED;
    double lo[6], hi[6];
    for (int k = 0; k < d; ++k)
        for (int side = 0; side < 2; ++side) {
            const int p = 2 * k + side;
            for (int q = 0; q < d; ++q)
                if (normal[p * d + q] != ((q == k) ? (side == 0 ? 1.0 : -1.0) : 0.0)) return HV_ERR_UNSUPPORTED;
            if (bc && bc[p] > 0) periodic = true;
            (side == 0 ? hi : lo)[k] = base[p * d + k];
        }
    if (periodic) return HV_OK;
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k) {
            const double x = xs[i * d + k];
            if (!(x >= lo[k] && x <= hi[k])) return HV_ERR_DOMAIN;
            // on a face (the walk's rule for a generator on a plane): out of general position
            if (hi[k] - x <= 1e-12 * (1.0 + std::fabs(hi[k])) || x - lo[k] <= 1e-12 * (1.0 + std::fabs(lo[k])))
                return HV_ERR_DEGENERATE;
        }
    return HV_OK;
}

// the fused batch's walker contexts per (device, walker), kept between calls
static std::mutex batch_mu;
static std::map<std::pair<int, int>, hv_ctx*> batch_ctx;

int hv_batch_release(int device_ordinal) {
    std::lock_guard<std::mutex> lk(batch_mu);
    for (auto it = batch_ctx.begin(); it != batch_ctx.end();) {
        if (it->first.first == device_ordinal || device_ordinal < 0) {
            if (it->second) hv_destroy(it->second);
            it = batch_ctx.erase(it);
        } else {
            ++it;
        }
    }
    return HV_OK;
}

int hv_build_mesh_batch(int device_ordinal, int32_t nmesh, const int32_t* d, const int64_t* n,
                        const double* const* xs, const int32_t* nplanes, const double* const* base,
                        const double* const* normal, const int16_t* const* bc, const hv_opts* opts,
                        int32_t nworkers, hv_mesh** out, int32_t* status) {
    if (nmesh < 0 || (nmesh > 0 && (!d || !n || !xs || !nplanes || !base || !normal || !out || !status)))
        return HV_ERR_ARG;
    for (int k = 0; k < nmesh; ++k) {
        out[k] = nullptr;
        status[k] = HV_ERR_DEVICE;   // overwritten by the path that builds it
    }
    if (nmesh == 0) return HV_OK;
    // fused: the cuboid meshes of each dimension in one walk (nworkers >= 0); the others
    // (periodic sides, other boundaries; every mesh when nworkers < 0) on worker threads
    std::vector<int> rest;
    std::vector<std::vector<int>> group(7);
    for (int k = 0; k < nmesh; ++k) {
        if (nworkers < 0) { rest.push_back(k); continue; }
        bool periodic = false;
        const int rc = fusable(d[k], n[k], xs[k], nplanes[k], base[k], normal[k], bc ? bc[k] : nullptr, periodic);
        if (rc == HV_OK && !periodic) group[d[k]].push_back(k);
        else if (rc == HV_OK || rc == HV_ERR_UNSUPPORTED) rest.push_back(k);   // the single path decides
        else status[k] = rc;
    }
    bool any = false;
    for (int dd = 2; dd <= 6; ++dd) any = any || !group[dd].empty();
    if (any) {
        // Fused walks: the meshes of each dimension in runs of at most HV_FOREST_VERTICES
        // expected vertices (one walk each; the runs do not depend on how they are scheduled),
        // taken largest first by a few walkers, each with its own stream and context, so that
        // one walk's small levels and host waits overlap another's work.  The walkers'
        // contexts stay for the process (streams, events, and device tables up to
        // HV_BATCH_KEEP bytes each, reused by the next call; the meshes own what they keep);
        // batch calls on one device take turns.
        std::lock_guard<std::mutex> lk(batch_mu);
        struct Job {
            int dd;
            std::vector<int> ks;
            double est;
        };
        std::vector<Job> jobs;
        for (int dd = 2; dd <= 6; ++dd)
            for (size_t k0 = 0; k0 < group[dd].size();) {
                Job j{dd, {}, 0.0};
                size_t k1 = k0;
                while (k1 < group[dd].size() &&
                       (k1 == k0 || j.est + vertices_per_point(dd) * n[group[dd][k1]] <= HV_FOREST_VERTICES)) {
                    j.est += vertices_per_point(dd) * n[group[dd][k1]];
                    j.ks.push_back(group[dd][k1]);
                    ++k1;
                }
                jobs.push_back(std::move(j));
                k0 = k1;
            }
        std::stable_sort(jobs.begin(), jobs.end(), [](const Job& a, const Job& b) { return a.est > b.est; });
        const int nwk = std::max(1, std::min<int>(HV_BATCH_WALKERS, (int)jobs.size()));
        std::vector<hv_ctx*> ctxs(nwk, nullptr);
        std::vector<int> crc(nwk, HV_OK);
        for (int w = 0; w < nwk; ++w) {
            hv_ctx*& c = batch_ctx[{device_ordinal, w}];
            if (!c && (crc[w] = hv_create(&c, device_ordinal)) != HV_OK) c = nullptr;
            ctxs[w] = c;
        }
        std::atomic<int> next{0};
        auto walker = [&](int w) {
            hv_ctx* c = ctxs[w];
            if (c) (void)hipSetDevice(c->device);
            for (;;) {
                const int q = next.fetch_add(1);
                if (q >= (int)jobs.size()) break;
                const Job& j = jobs[q];
                if (!c) {
                    for (int k : j.ks) status[k] = crc[w];
                    continue;
                }
                Forest fo;
                fo.nmesh = (int)j.ks.size();
                for (int k : j.ks) {
                    fo.xs.push_back(xs[k]);
                    fo.n.push_back((int)n[k]);
                    fo.base.push_back(base[k]);
                    fo.normal.push_back(normal[k]);
                }
                c->D = j.dd;
                int rc = HV_OK;
                switch (j.dd) {
                    case 2: rc = hv_build_forest_meshes<2>(c, opts, fo); break;
                    case 3: rc = hv_build_forest_meshes<3>(c, opts, fo); break;
                    case 4: rc = hv_build_forest_meshes<4>(c, opts, fo); break;
                    case 5: rc = hv_build_forest_meshes<5>(c, opts, fo); break;
                    default: rc = hv_build_forest_meshes<6>(c, opts, fo); break;
                }
                for (size_t q2 = 0; q2 < j.ks.size(); ++q2) {
                    const int k = j.ks[q2];
                    out[k] = rc == HV_OK ? fo.out[q2] : nullptr;
                    status[k] = rc;
                }
            }
            if (c) {
                (void)hipStreamSynchronize(c->stream);
                c->tree_ok = false;
                constexpr size_t HV_BATCH_KEEP = (size_t)24 << 30;
                if (c->pool->total() > HV_BATCH_KEEP) {
                    auto fresh = std::make_shared<DevPool>();
                    fresh->device = c->device;
                    c->pool = fresh;   // the previous pool frees its tables here
                }
            }
        };
        std::vector<std::thread> th;
        for (int w = 1; w < nwk; ++w) {
            try {
                th.emplace_back(walker, w);
            } catch (...) {
                break;   // fewer walkers: the ones running take every run
            }
        }
        walker(0);
        for (auto& t : th) t.join();
    }
    if (!rest.empty()) {
        const int nw = std::max(1, std::min<int>(nworkers > 0 ? nworkers : (nworkers < 0 ? -nworkers : 8), (int)rest.size()));
        std::atomic<int> next{0};
        auto work = [&]() {
            hv_ctx* c = nullptr;
            if (hv_create(&c, device_ordinal) != HV_OK) return;   // the other workers take its share
            for (;;) {
                const int q = next.fetch_add(1);
                if (q >= (int)rest.size()) break;
                const int k = rest[q];
                int rc = hv_set_points(c, d[k], n[k], xs[k]);
                if (rc == HV_OK) rc = hv_set_boundary(c, nplanes[k], base[k], normal[k], bc ? bc[k] : nullptr);
                if (rc == HV_OK) rc = hv_build_mesh(c, opts, &out[k]);
                // the mesh owns its vertex tables outright: the worker's pool goes with the worker
                if (rc == HV_OK) out[k]->pool.reset();
                status[k] = rc;
            }
            hv_destroy(c);
        };
        std::vector<std::thread> th;
        try {
            th.reserve(nw);
            for (int w = 0; w < nw; ++w) th.emplace_back(work);
        } catch (...) {
            // fewer workers than asked (at least the ones started run every mesh); none: fail
            if (th.empty()) return HV_ERR_DEVICE;
        }
        for (auto& t : th) t.join();
    }
    for (int k = 0; k < nmesh; ++k)
        if (status[k] != HV_OK) return status[k];
    return HV_OK;
}

int hv_mesh_counts(const hv_mesh* m, int64_t* nvert, int32_t* max_sig, int64_t* nrays_unbounded) {
    if (!m) return HV_ERR_ARG;
    if (nvert) *nvert = m->nv;
    if (max_sig) *max_sig = m->D + 1;
    if (nrays_unbounded) *nrays_unbounded = m->stats.unbounded;
    return HV_OK;
}

int hv_mesh_export(const hv_mesh* m, int64_t* sig, double* coords) {
    if (!m || !sig || !coords) return HV_ERR_ARG;
    (void)hipSetDevice(m->device);
    HV_DISPATCH(m->D, hv_mesh_export_impl, m, sig, coords)
}

int hv_mesh_export_db(const hv_mesh* m, int64_t* heap, int64_t* cell_ptr, int64_t* refs) {
    if (!m || !cell_ptr) return HV_ERR_ARG;
    if (m->D < 2 || m->D > 6) return HV_ERR_UNSUPPORTED;
    const int D = m->D;
    const int64_t nv = m->nv, N = m->N;
    std::vector<int64_t> sig((size_t)std::max<int64_t>(nv, 1) * (D + 1));
    std::vector<double> x((size_t)std::max<int64_t>(nv, 1) * D);
    if (nv > 0) {
        const int rc = hv_mesh_export(m, sig.data(), x.data());
        if (rc) return rc;
    }
    // per generator the number of vertices holding it
    std::fill(cell_ptr, cell_ptr + N + 1, 0);
    for (int64_t v = 0; v < nv; ++v)
        for (int j = 0; j <= D; ++j) {
            const int64_t g = sig[(size_t)v * (D + 1) + j];
            if (g <= N) ++cell_ptr[g];
        }
    for (int64_t i = 0; i < N; ++i) cell_ptr[i + 1] += cell_ptr[i];
    const int64_t W = 2 * D + 2;   // words per vertex: length, d + 1 ids, d coordinates
    if (heap)
        for (int64_t v = 0; v < nv; ++v) {
            int64_t* h = heap + v * W;
            h[0] = D + 1;
            for (int j = 0; j <= D; ++j) h[1 + j] = sig[(size_t)v * (D + 1) + j];
            std::memcpy(h + D + 2, &x[(size_t)v * D], sizeof(double) * D);
        }
    if (refs) {
        std::vector<int64_t> fill(cell_ptr, cell_ptr + N);
        for (int64_t v = 0; v < nv; ++v) {
            const int64_t pos = v * W + 1;   // 1-based position of the vertex's length word
            for (int j = 0; j <= D; ++j) {
                const int64_t g = sig[(size_t)v * (D + 1) + j];
                if (g <= N) refs[fill[g - 1]++] = j == 0 ? -pos : pos;
            }
        }
    }
    return HV_OK;
}

int hv_mesh_images(const hv_mesh* m, int64_t* n_ext, int64_t* orig, int8_t* shift) {
    if (!m || !n_ext) return HV_ERR_ARG;
    *n_ext = m->N;
    const bool per = m->img_orig && m->img_shift;
    for (int64_t i = 0; i < m->N; ++i) {
        if (orig) orig[i] = (per ? (int64_t)(*m->img_orig)[i] : i) + 1;
        if (shift)
            for (int k = 0; k < m->D; ++k) shift[i * m->D + k] = per ? (*m->img_shift)[i * m->D + k] : 0;
    }
    return HV_OK;
}

int hv_mesh_stats(const hv_mesh* m, hv_stats_t* out) {
    if (!m || !out) return HV_ERR_ARG;
    *out = m->stats;
    return HV_OK;
}

int hv_mesh_free(hv_mesh* m) {
    if (!m) return HV_ERR_ARG;
    (void)hipSetDevice(m->device);
    if (m->store) {
        // tables in the batch's shared allocation: released with the last mesh of the batch
    } else if (m->pool) {
        m->pool->give("vkey", m->key);
        m->pool->give("vx", m->x);
        m->pool->give("ids", m->ids);
    } else {
        for (DBuf* b : {&m->key, &m->x, &m->ids})
            if (b->p) (void)hipFree(b->p);
    }
    delete m;
    return HV_OK;
}

int hv_raycast_batch(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                     const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (nrays < 0 || (nrays > 0 && (!sig || !drop || !r || !out_sig || !out_r || !out_t || !out_status)))
        return hv_fail(c, HV_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_raycast_batch_impl, c, precision, nrays, sig, drop, r, out_sig, out_r, out_t, out_status)
}

int hv_raycast_batch_keys(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                          const double* r, int32_t max_sig, int64_t* out_sig, int32_t* out_len, double* out_r,
                          double* out_t, int32_t* out_status) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (max_sig < c->D + 1 || max_sig > 1024) return hv_fail(c, HV_ERR_ARG, "max_sig must be in d+1..1024");
    if (nrays < 0 ||
        (nrays > 0 && (!sig || !drop || !r || !out_sig || !out_len || !out_r || !out_t || !out_status)))
        return hv_fail(c, HV_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    std::vector<int64_t> key((size_t)std::max<int64_t>(nrays, 1) * (c->D + 1));
    HV_DISPATCH(c->D, hv_raycast_batch_impl, c, precision, nrays, sig, drop, r, key.data(), out_r, out_t, out_status,
                max_sig, out_sig, out_len)
}

int hv_raycast_batch_noalloc(hv_ctx* c, int precision, int64_t nrays, const int64_t* sig, const int32_t* drop,
                             const double* r, int64_t* out_sig, double* out_r, double* out_t, int32_t* out_status) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (nrays < 0 || (nrays > 0 && (!sig || !drop || !r || !out_sig || !out_r || !out_t || !out_status)))
        return hv_fail(c, HV_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_raycast_noalloc_impl, c, precision, nrays, sig, drop, r, out_sig, out_r, out_t, out_status)
}

int hv_nn_halfspace_batch(hv_ctx* c, int64_t nq, const double* q, const double* u, const double* cc,
                          int64_t* idx, double* dist) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (nq < 0 || (nq > 0 && (!q || !u || !cc || !idx || !dist))) return hv_fail(c, HV_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_nn_halfspace_impl, c, nq, q, u, cc, idx, dist)
}

int hv_inrange_batch(hv_ctx* c, int64_t nq, const double* q, const double* radius, int64_t* ptr, int64_t* idx) {
    if (!c) return HV_ERR_ARG;
    if (c->D == 0) return hv_fail(c, HV_ERR_STATE, "no points set");
    if (nq < 0 || (nq > 0 && (!q || !radius || !ptr))) return hv_fail(c, HV_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    HV_DISPATCH(c->D, hv_inrange_impl, c, nq, q, radius, ptr, idx)
}

int hv_integrate_polytope(hv_ctx* c, const hv_mesh* mesh, hv_integral** out) {
    if (!c || !mesh || !out) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->device != c->device) return hv_fail(c, HV_ERR_STATE, "the mesh lives on another device");
    (void)hipSetDevice(c->device);
    hv_bind_mesh bind(c, mesh);
    HV_DISPATCH(mesh->D, hv_integrate_polytope_impl, c, mesh, out)
}

int hv_integrate_polytope_f(hv_ctx* c, const hv_mesh* mesh, int32_t nf, hv_integrand_fn f, void* user,
                            hv_integral** out) {
    if (!c || !mesh || !out) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (nf < 1 || !f) return hv_fail(c, HV_ERR_ARG, "integrand: nf >= 1 and a function");
    if (mesh->device != c->device) return hv_fail(c, HV_ERR_STATE, "the mesh lives on another device");
    (void)hipSetDevice(c->device);
    hv_bind_mesh bind(c, mesh);
    HV_DISPATCH(mesh->D, hv_integrate_polytope_f_impl, c, mesh, nf, f, user, out)
}

int hv_integral_export_f(const hv_integral* I, int32_t* nf, double* bulk, double* iface) {
    return hv_integral_export_f_impl(I, nf, bulk, iface);
}

int hv_integrate_mc(hv_ctx* c, const hv_mesh* mesh, int64_t ndir, uint64_t seed, hv_integral** out) {
    if (!c || !mesh || !out) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->device != c->device) return hv_fail(c, HV_ERR_STATE, "the mesh lives on another device");
    if (mesh->cells) return hv_fail(c, HV_ERR_UNSUPPORTED, "the Monte-Carlo integrators take a mesh of all cells");
    (void)hipSetDevice(c->device);
    hv_bind_mesh bind(c, mesh);
    HV_DISPATCH(mesh->D, hv_integrate_mc_impl, c, mesh, ndir, seed, out)
}

int hv_integrate_mc_points(hv_ctx* c, const hv_mesh* mesh, int64_t nsamples, uint64_t seed, double slab,
                           hv_integral** out) {
    if (!c || !mesh || !out) return hv_fail(c, HV_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->device != c->device) return hv_fail(c, HV_ERR_STATE, "the mesh lives on another device");
    if (mesh->cells) return hv_fail(c, HV_ERR_UNSUPPORTED, "the Monte-Carlo integrators take a mesh of all cells");
    (void)hipSetDevice(c->device);
    hv_bind_mesh bind(c, mesh);
    HV_DISPATCH(mesh->D, hv_integrate_mc_points_impl, c, mesh, nsamples, seed, slab, out)
}

int hv_integral_counts(const hv_integral* I, int64_t* ncells, int64_t* nnbr) {
    return hv_integral_counts_impl(I, ncells, nnbr);
}

int hv_integral_work(const hv_integral* I, int64_t* work, double* ms_kernel) {
    return hv_integral_work_impl(I, work, ms_kernel);
}

int hv_integral_export(const hv_integral* I, double* vol, int64_t* nbr_ptr, int64_t* nbr, double* area, double* ms) {
    return hv_integral_export_impl(I, vol, nbr_ptr, nbr, area, ms);
}

int hv_integral_free(hv_integral* I) {
    if (!I) return HV_ERR_ARG;
    hv_integral_free_impl(I);
    return HV_OK;
}

}  // extern "C"
