// C ABI of the MI355X RNS multiple-precision path (include/mpres_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpres_hip.h"
#include "batch_plan.hpp"
#include "internal.hpp"
#include "ops.hpp"

using namespace mph;

const Ops *mph::find_variant(int N, int L) {
#define MPH_CASE(VN, VL) \
    if (N == VN && L == VL) return ops_n##VN##_l##VL();
    MPH_VARIANTS(MPH_CASE)
#undef MPH_CASE
    return nullptr;
}

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

int mph::set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(MPRES_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// operation classes of mpres_set_lanes; 6 is a timing slot only (gemm residue products)
enum OpKind { OP_ELEM = 0, OP_DOT = 1, OP_GEMV = 2, OP_GEMM = 3, OP_SPMV = 4, OP_AXPY = 5, OP_MFMA = 7, OP_COUNT = 8 };
bool lanes_op(int op) { return op >= 0 && op < OP_COUNT && op != 6; }
// extra timing slots: the residue products of the split-form gemm, the precision form,
// batches, and the split-form gemv's residue and chain kernels together (two streams)
constexpr int PROF_GEMM_RESID = 6, PROF_GEMM_MFMA = 7, PROF_BATCH = 8, PROF_GEMV_PAIR = 9, PROF_SLOTS = 10;

const Ops *find_ops(int n, int L) { return find_variant(n, L); }

// Built-in lanes-per-number choice for each (moduli count, operation), from
// measurements on MI355X (dot and axpy 2^24 at 8 moduli, gemv 4096^2 and gemm
// 512^3 at 32 moduli, SpMV 10^6 rows at 16 moduli, precision-form gemm 2048^3 at
// 8 and 32 moduli); mpres_set_lanes overrides. Element-parallel work with
// enough numbers prefers one lane per number (no lane repeats a number's scalar
// logic); reductions with few independent chains prefer more lanes.
// Lanes by call size measured on MI355X (mpres-blas_amd/lanes_tune.py, the table in
// mpres-blas_amd/tuning/lanes_gfx950.json): a call of size <= bound[i] uses lanes[i], a
// larger one the last entry. Installed at context creation for these sets; mpres_set_lanes
// replaces an operation's table by a fixed choice, mpres_set_lanes(op, 0) restores it.
struct BuiltinLanes {
    const char *set;
    int op, count;
    long long bound[4];
    int lanes[4];
};
const BuiltinLanes kBuiltinLanes[] = {
    {"8_2double", OP_DOT, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {8, 8, 8, 8}},
    {"8_2double", OP_AXPY, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {8, 2, 1, 1}},
    {"8_2double", OP_GEMV, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {8, 8, 8, 1}},
    {"8_2double", OP_GEMM, 4, {741455LL, 5931641LL, 47453132LL, 134217728LL}, {8, 8, 8, 1}},
    {"8_2double", OP_SPMV, 3, {31622LL, 316227LL, 1000000LL}, {8, 8, 8}},
    {"8_2double", OP_ELEM, 3, {16384LL, 262144LL, 1048576LL}, {8, 2, 2}},
    {"16_4double", OP_DOT, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {16, 16, 4, 4}},
    {"16_4double", OP_AXPY, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {16, 4, 1, 1}},
    {"16_4double", OP_GEMV, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {16, 16, 16, 16}},
    {"16_4double", OP_GEMM, 4, {741455LL, 5931641LL, 47453132LL, 134217728LL}, {16, 16, 4, 4}},
    {"16_4double", OP_SPMV, 3, {31622LL, 316227LL, 1000000LL}, {2, 4, 4}},
    {"16_4double", OP_ELEM, 3, {16384LL, 262144LL, 1048576LL}, {16, 4, 4}},
    {"32_8double", OP_DOT, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {32, 8, 8, 8}},
    {"32_8double", OP_AXPY, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {8, 8, 1, 1}},
    {"32_8double", OP_GEMV, 4, {16384LL, 262144LL, 4194304LL, 16777216LL}, {32, 4, 8, 8}},
    {"32_8double", OP_GEMM, 4, {741455LL, 5931641LL, 47453132LL, 134217728LL}, {8, 8, 8, 8}},
    {"32_8double", OP_SPMV, 3, {31622LL, 316227LL, 1000000LL}, {1, 2, 4}},
    {"32_8double", OP_ELEM, 3, {16384LL, 262144LL, 1048576LL}, {32, 8, 8}},
};

int default_lanes(int n, int op) {
    switch (n) {
        case 8: return (op == OP_AXPY || op == OP_MFMA) ? 1 : 8;
        case 16: return op == OP_SPMV ? 16 : 4;
        default: return 8;
    }
}

}  // namespace

struct mpres_ctx {
    int device = 0;
    int set_id = -1;
    DevConsts host{};
    DevConsts *dev = nullptr;
    int *dpow2 = nullptr;
    int bits = 0;  // bit length of the largest modulus
    const Ops *ops[OP_COUNT] = {};
    // per operation, an optional table of lane counts by call size (mpres_set_lanes_table):
    // entry i serves calls of size <= bound[i], the last entry larger ones
    struct LanesTable {
        int count = 0;
        long long bound[MPRES_LANES_TABLE_MAX] = {};
        const Ops *ops[MPRES_LANES_TABLE_MAX] = {};
    } lt[OP_COUNT];
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // a second stream (non-blocking, default priority) for work of one call that does not
    // depend on the rest (the gemm residue path), with its fork and join events
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // the pool's last use: an event recorded on the stream of the last call that used it
    hipEvent_t ws_ev = nullptr;
    hipStream_t ws_stream = nullptr;
    bool ws_used = false;
    // optional timing of each operation's main kernel (mpres_profile)
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_free;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_live;
    double prof_ms[PROF_SLOTS] = {};
    long prof_count[PROF_SLOTS] = {};
};

namespace {

size_t mpsize(const mpres_ctx *c) { return (size_t)(4 * c->host.n + 40); }
// the built-in lanes table of operation op for the context's moduli set, if there is one
// (every listed lane count is compiled; an entry that is not leaves the table out)
void install_builtin_lanes(mpres_ctx *c, int op) {
    auto &t = c->lt[op];
    t.count = 0;
    const char *name = get_set(c->set_id)->name;
    for (const BuiltinLanes &b : kBuiltinLanes) {
        if (b.op != op || std::strcmp(b.set, name) != 0) continue;
        for (int i = 0; i < b.count; i++) {
            t.bound[i] = b.bound[i];
            t.ops[i] = find_ops(c->host.n, b.lanes[i]);
            if (!t.ops[i]) return;
        }
        t.count = b.count;
        return;
    }
}
// the kernel variant of operation op for a call of the given size (mpres_set_lanes_table)
const Ops *sized(const mpres_ctx *c, int op, long long size) {
    const auto &t = c->lt[op];
    for (int i = 0; i < t.count; i++)
        if (size <= t.bound[i]) return t.ops[i];
    return t.count ? t.ops[t.count - 1] : c->ops[op];
}
// the split-form SpMV's workspace: the REDO count, 3 x SP_NB class counters, the row order and
// the REDO row list
size_t spmv_ws_bytes(int m) { return 4 * (1 + 3 * (size_t)SP_NB + 2 * (size_t)m); }

SoA soa_of(const mp_array_t *a) { return soa_view(a); }

// the split and precision-form gemm kernels read an element's residues as 16-byte vectors
bool digits_aligned(const mp_array_t *a) { return ((uintptr_t)a->digits & 15) == 0; }

// workspace: caller's buffer or the context pool. Calls on one stream use the pool in
// stream order; a call on another stream first waits (on the device) for the pool's
// last use, and the pool grows only after that use has ended, so calls on different
// streams never overwrite each other's partial results. WsUse records the new last use
// once the calling operation has enqueued its kernels.
struct WsUse {
    mpres_ctx *c;
    hipStream_t st;
    bool pool = false;
    WsUse(mpres_ctx *c_, hipStream_t st_) : c(c_), st(st_) {}
    int get(void *ws, size_t ws_bytes, size_t need, char **out) {
        if (need == 0) {
            *out = nullptr;
            return MPRES_OK;
        }
        if (ws) {
            if (ws_bytes < need)
                return fail(MPRES_ERR_WORKSPACE, "workspace of %zu bytes, %zu needed", ws_bytes, need);
            *out = (char *)ws;
            return MPRES_OK;
        }
        if (!c->ws_ev) HIPCHK(hipEventCreateWithFlags(&c->ws_ev, hipEventDisableTiming));
        if (c->ws_bytes < need) {
            if (c->ws) {
                if (c->ws_used) HIPCHK(hipEventSynchronize(c->ws_ev));
                HIPCHK(hipFree(c->ws));
                c->ws = nullptr;
                c->ws_bytes = 0;
            }
            HIPCHK(hipMalloc(&c->ws, need));
            c->ws_bytes = need;
        } else if (c->ws_used && c->ws_stream != st) {
            HIPCHK(hipStreamWaitEvent(st, c->ws_ev, 0));
        }
        pool = true;
        *out = (char *)c->ws;
        return MPRES_OK;
    }
    ~WsUse() {
        if (!pool) return;
        if (hipEventRecord(c->ws_ev, st) == hipSuccess) {
            c->ws_stream = st;
            c->ws_used = true;
        }
    }
};

// Runs f(stream) on the context's side stream, ordered after the work already queued on
// st; join() makes st wait for it. Without a side stream (its creation failed) f runs on st.
struct Fork {
    mpres_ctx *c;
    hipStream_t st, s1;
    Fork(mpres_ctx *c_, hipStream_t st_) : c(c_), st(st_), s1(st_) {
        if (!c->side) {
            if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) c->side = nullptr;
            if (c->side && (hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
                            hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess)) {
                (void)hipStreamDestroy(c->side);
                c->side = nullptr;
            }
        }
        if (c->side && hipEventRecord(c->fork_ev, st) == hipSuccess &&
            hipStreamWaitEvent(c->side, c->fork_ev, 0) == hipSuccess)
            s1 = c->side;
    }
    hipError_t join() {
        if (s1 == st) return hipSuccess;
        hipError_t e = hipEventRecord(c->join_ev, s1);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, c->join_ev, 0);
        return e;
    }
};

unsigned next_pow2(unsigned v) {  // src/common.cuh:92-101
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}

hipStream_t strm(void *st) { return (hipStream_t)st; }

// Brackets the main kernel of operation `op` with events when profiling is on.
struct ProfScope {
    mpres_ctx *c;
    int op;
    hipStream_t st;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    bool on = false;
    ProfScope(mpres_ctx *c_, int op_, hipStream_t st_) : c(c_), op(op_), st(st_) {
        if (!c->prof || op < 0) return;
        if (c->prof_free.empty()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) {
                (void)hipEventDestroy(a);
                return;
            }
            c->prof_free.push_back({a, b});
        }
        ev = c->prof_free.back();
        c->prof_free.pop_back();
        on = hipEventRecord(ev.first, st) == hipSuccess;
    }
    ~ProfScope() {
        if (!on) return;
        (void)hipEventRecord(ev.second, st);
        c->prof_live.push_back({op, ev});
    }
};

}  // namespace

int mph::ctx_device(const mpres_ctx *c) { return c ? c->device : -1; }

extern "C" {

const char *mpres_last_error(void) { return g_err.c_str(); }
int mpres_num_moduli_sets(void) { return num_sets(); }
const char *mpres_moduli_set_name(int id) {
    const ModuliSet *s = get_set(id);
    return s ? s->name : nullptr;
}
int mpres_moduli_set_id(const char *name) {
    if (!name) return -1;
    for (int i = 0; i < num_sets(); i++)
        if (std::strcmp(get_set(i)->name, name) == 0) return i;
    return -1;
}

int mpres_ctx_create(mpres_ctx_t **out, int set_id, int device) {
    if (!out) return fail(MPRES_ERR_ARG, "null output pointer");
    *out = nullptr;
    const ModuliSet *s = get_set(set_id);
    if (!s) return fail(MPRES_ERR_ARG, "unknown moduli set %d", set_id);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return fail(MPRES_ERR_NO_DEVICE, "no HIP device %d", device);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(MPRES_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    HIPCHK(hipSetDevice(device));
    if (int rc = noalloc_init(device)) return rc;
    auto *c = new mpres_ctx();
    c->device = device;
    c->set_id = set_id;
    std::vector<int> pow2;
    if (!compute_consts(*s, c->host, pow2)) {
        delete c;
        return fail(MPRES_ERR_ARG, "moduli set %s: inconsistent RNS_MODULI_PRODUCT_LOG2", s->name);
    }
    for (int i = 0; i < s->n; i++)
        while ((1LL << c->bits) <= s->m[i]) c->bits++;
    for (int op = 0; op < OP_COUNT; op++) {
        c->ops[op] = find_ops(s->n, default_lanes(s->n, op));
        if (!c->ops[op]) {
            delete c;
            return fail(MPRES_ERR_UNSUPPORTED, "no kernels compiled for %d moduli", s->n);
        }
        install_builtin_lanes(c, op);
    }
    if (hipMalloc(&c->dpow2, pow2.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->dev, sizeof(DevConsts)) != hipSuccess) {
        (void)hipFree(c->dpow2);
        delete c;
        return fail(MPRES_ERR_HIP, "allocating the constant tables failed");
    }
    c->host.pow2 = c->dpow2;
    hipError_t e = hipMemcpy(c->dpow2, pow2.data(), pow2.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->dev, &c->host, sizeof(DevConsts), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(c->dev);
        (void)hipFree(c->dpow2);
        delete c;
        return fail(MPRES_ERR_HIP, "copying the constant tables: %s", hipGetErrorString(e));
    }
    *out = c;
    return MPRES_OK;
}

int mpres_ctx_destroy(mpres_ctx_t *c) {
    if (!c) return MPRES_OK;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    if (c->ws) (void)hipFree(c->ws);
    if (c->ws_ev) (void)hipEventDestroy(c->ws_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    for (auto &e : c->prof_live) c->prof_free.push_back(e.second);
    for (auto &e : c->prof_free) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    (void)hipFree(c->dev);
    (void)hipFree(c->dpow2);
    delete c;
    return MPRES_OK;
}

int mpres_ctx_info(const mpres_ctx_t *c, int *n, int *precision, int *mp_h, int *mp_j, int *log2m) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (n) *n = c->host.n;
    if (precision) *precision = c->host.mp_precision;
    if (mp_h) *mp_h = c->host.mp_h;
    if (mp_j) *mp_j = c->host.mp_j;
    if (log2m) *log2m = c->host.log2m;
    return MPRES_OK;
}

size_t mpres_mp_float_size(const mpres_ctx_t *c) { return c ? mpsize(c) : 0; }

static int table_of(const DevConsts &h, const std::vector<int> &pow2, int what, void *out, size_t bytes) {
    const int n = h.n;
    std::vector<char> buf;
    auto put = [&](const void *p, size_t sz) { buf.insert(buf.end(), (const char *)p, (const char *)p + sz); };
    switch (what) {
        case 0: put(h.m, n * sizeof(int)); break;
        case 1: put(h.w, n * sizeof(int)); break;
        case 2: put(pow2.data(), pow2.size() * sizeof(int)); break;
        case 3: put(h.mpow2, sizeof h.mpow2); break;
        case 4:
            for (int j = 0; j < P2T; j++) put(h.h[j], n * sizeof(int));
            break;
        case 5:
            for (int j = 0; j < P2T; j++) put(h.pow2inv[j], n * sizeof(int));
            break;
        case 6:
            for (int j = 0; j < n; j++) put(h.mrcinv[j], n * sizeof(int));
            break;
        case 7:
            put(h.rrd, n * sizeof(double));
            put(h.rru, n * sizeof(double));
            put(&h.acc, sizeof(double));
            break;
        case 8: {
            const ErC *u[4] = {&h.unit_low, &h.unit_upp, &h.inv_unit_low, &h.inv_unit_upp};
            for (auto *e : u) {
                er_float_t v{e->frac, (long)e->exp};
                put(&v, sizeof v);
            }
            break;
        }
        case 9: {
            const int v[5] = {h.mp_precision, h.mp_h, h.mp_j, h.ref_factor, h.log2m};
            put(v, sizeof v);
            break;
        }
        default: return fail(MPRES_ERR_ARG, "unknown table %d", what);
    }
    if (bytes < buf.size()) return fail(MPRES_ERR_ARG, "table %d needs %zu bytes", what, buf.size());
    std::memcpy(out, buf.data(), buf.size());
    return MPRES_OK;
}

int mpres_ctx_table(const mpres_ctx_t *c, int what, void *out, size_t bytes) {
    if (!c || !out) return fail(MPRES_ERR_ARG, "null argument");
    std::vector<int> p2((size_t)(c->host.log2m + 1) * c->host.n);
    if (hipMemcpy(p2.data(), c->dpow2, p2.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(MPRES_ERR_HIP, "reading RNS_POW2");
    return table_of(c->host, p2, what, out, bytes);
}

int mpres_moduli_set_table(int set_id, int what, void *out, size_t bytes) {
    const ModuliSet *s = get_set(set_id);
    if (!s || !out) return fail(MPRES_ERR_ARG, "invalid argument");
    DevConsts h;
    std::vector<int> p2;
    if (!compute_consts(*s, h, p2)) return fail(MPRES_ERR_ARG, "inconsistent moduli set");
    return table_of(h, p2, what, out, bytes);
}

// ---------------------------------------------------------------- arrays
int mpres_array_init(mpres_ctx_t *c, mp_array_t *a, size_t size) {
    if (!c || !a || size == 0) return fail(MPRES_ERR_ARG, "invalid array request");
    std::memset(a, 0, sizeof *a);
    const size_t n = (size_t)c->host.n;
    const int len = (int)size;
    hipError_t e = hipMalloc(&a->digits, n * size * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&a->sign, size * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&a->exp, size * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&a->eval, 2 * size * sizeof(er_float_t));
    if (e == hipSuccess) e = hipMalloc(&a->buf, size * sizeof(mpres_int4_t));
    if (e == hipSuccess) e = hipMalloc(&a->len, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(a->len, &len, sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mpres_array_clear(c, a);  // frees what was allocated (hipFree(NULL) is a no-op)
        return fail(MPRES_ERR_HIP, "allocating an array of %zu elements: %s", size, hipGetErrorString(e));
    }
    return MPRES_OK;
}

int mpres_array_clear(mpres_ctx_t *c, mp_array_t *a) {
    (void)c;
    if (!a) return MPRES_OK;
    (void)hipFree(a->digits);
    (void)hipFree(a->sign);
    (void)hipFree(a->exp);
    (void)hipFree(a->eval);
    (void)hipFree(a->buf);
    (void)hipFree(a->len);
    std::memset(a, 0, sizeof *a);
    return MPRES_OK;
}

int mpres_array_host2device(mpres_ctx_t *c, mp_array_t *dst, const void *src, size_t size, void *st) {
    if (!c || !dst || (!src && size)) return fail(MPRES_ERR_ARG, "null argument");
    const int n = c->host.n;
    const size_t S = mpsize(c);
    std::vector<int> dig(size * n), sg(size), ex(size);
    std::vector<er_float_t> ev(2 * size);
    const char *p = (const char *)src;
    for (size_t i = 0; i < size; i++) {
        std::memcpy(&dig[i * n], p + i * S, n * sizeof(int));
        std::memcpy(&sg[i], p + i * S + 4 * n, sizeof(int));
        std::memcpy(&ex[i], p + i * S + 4 * n + 4, sizeof(int));
        std::memcpy(&ev[i], p + i * S + 4 * n + 8, sizeof(er_float_t));
        std::memcpy(&ev[size + i], p + i * S + 4 * n + 24, sizeof(er_float_t));
    }
    // upper bounds live at eval + len; len is the array's actual length
    int len = 0;
    HIPCHK(hipMemcpy(&len, dst->len, sizeof(int), hipMemcpyDeviceToHost));
    if ((size_t)len < size) return fail(MPRES_ERR_ARG, "array of length %d, %zu elements given", len, size);
    // the copies read the staging vectors: every path waits for them before the vectors go away
    hipError_t e = hipMemcpyAsync(dst->digits, dig.data(), dig.size() * sizeof(int), hipMemcpyHostToDevice, strm(st));
    if (e == hipSuccess) e = hipMemcpyAsync(dst->sign, sg.data(), size * sizeof(int), hipMemcpyHostToDevice, strm(st));
    if (e == hipSuccess) e = hipMemcpyAsync(dst->exp, ex.data(), size * sizeof(int), hipMemcpyHostToDevice, strm(st));
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst->eval, ev.data(), size * sizeof(er_float_t), hipMemcpyHostToDevice, strm(st));
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst->eval + len, ev.data() + size, size * sizeof(er_float_t), hipMemcpyHostToDevice,
                           strm(st));
    const hipError_t w = hipStreamSynchronize(strm(st));
    if (e == hipSuccess) e = w;
    if (e != hipSuccess) return fail(MPRES_ERR_HIP, "copying an array to the device: %s", hipGetErrorString(e));
    return MPRES_OK;
}

int mpres_array_device2host(mpres_ctx_t *c, void *dst, const mp_array_t *src, size_t size, void *st) {
    if (!c || !src || (!dst && size)) return fail(MPRES_ERR_ARG, "null argument");
    const int n = c->host.n;
    const size_t S = mpsize(c);
    int len = 0;
    HIPCHK(hipMemcpy(&len, src->len, sizeof(int), hipMemcpyDeviceToHost));
    if ((size_t)len < size) return fail(MPRES_ERR_ARG, "array of length %d, %zu elements asked", len, size);
    std::vector<int> dig(size * n), sg(size), ex(size);
    std::vector<er_float_t> ev(2 * size);
    // the copies write the staging vectors: every path waits for them before the vectors go away
    hipError_t e = hipMemcpyAsync(dig.data(), src->digits, dig.size() * sizeof(int), hipMemcpyDeviceToHost, strm(st));
    if (e == hipSuccess) e = hipMemcpyAsync(sg.data(), src->sign, size * sizeof(int), hipMemcpyDeviceToHost, strm(st));
    if (e == hipSuccess) e = hipMemcpyAsync(ex.data(), src->exp, size * sizeof(int), hipMemcpyDeviceToHost, strm(st));
    if (e == hipSuccess)
        e = hipMemcpyAsync(ev.data(), src->eval, size * sizeof(er_float_t), hipMemcpyDeviceToHost, strm(st));
    if (e == hipSuccess)
        e = hipMemcpyAsync(ev.data() + size, src->eval + len, size * sizeof(er_float_t), hipMemcpyDeviceToHost,
                           strm(st));
    const hipError_t w = hipStreamSynchronize(strm(st));
    if (e == hipSuccess) e = w;
    if (e != hipSuccess) return fail(MPRES_ERR_HIP, "copying an array to the host: %s", hipGetErrorString(e));
    char *p = (char *)dst;
    for (size_t i = 0; i < size; i++) {
        std::memcpy(p + i * S, &dig[i * n], n * sizeof(int));
        std::memcpy(p + i * S + 4 * n, &sg[i], sizeof(int));
        std::memcpy(p + i * S + 4 * n + 4, &ex[i], sizeof(int));
        std::memcpy(p + i * S + 4 * n + 8, &ev[i], sizeof(er_float_t));
        std::memcpy(p + i * S + 4 * n + 24, &ev[size + i], sizeof(er_float_t));
    }
    return MPRES_OK;
}

int mpres_array_set_d(mpres_ctx_t *c, mp_array_t *dst, const double *x, int count, void *st) {
    if (!c || !dst || (!x && count > 0)) return fail(MPRES_ERR_ARG, "null argument");
    if (count <= 0) return MPRES_OK;
    HIPCHK(c->ops[OP_ELEM]->set_d(c->dev, x, count, soa_of(dst), nullptr, true, strm(st)));
    return MPRES_OK;
}

int mpres_set_d(mpres_ctx_t *c, void *dst, const double *x, int count, void *st) {
    if (!c || (!dst && count > 0) || (!x && count > 0)) return fail(MPRES_ERR_ARG, "null argument");
    if (count <= 0) return MPRES_OK;
    SoA none{};
    HIPCHK(c->ops[OP_ELEM]->set_d(c->dev, x, count, none, (char *)dst, false, strm(st)));
    return MPRES_OK;
}

int mpres_aos_to_array(mpres_ctx_t *c, mp_array_t *dst, const void *src, int count, void *st) {
    if (!c || !dst || (!src && count > 0)) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->aos_to_soa(c->dev, (const char *)src, soa_of(dst), count, strm(st)));
    return MPRES_OK;
}

int mpres_array_to_aos(mpres_ctx_t *c, void *dst, const mp_array_t *src, int count, void *st) {
    if (!c || !src || (!dst && count > 0)) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->soa_to_aos(c->dev, soa_of(src), (char *)dst, count, strm(st)));
    return MPRES_OK;
}

int mpres_elementwise(mpres_ctx_t *c, int op, int count, const void *x, const void *y, void *r, void *st) {
    if (!c || op < 0 || op > 4) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    if (count <= 0) return MPRES_OK;
    if (!x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(sized(c, OP_ELEM, count)->elem(c->dev, op, count, (const char *)x, (const char *)y, (const double *)y,
                                          (char *)r, strm(st)));
    return MPRES_OK;
}

int mpres_cmp(mpres_ctx_t *c, int mode, int count, const void *x, const void *y, int *r, void *st) {
    if (!c || (mode != MPRES_CMP && mode != MPRES_CMP_ABS)) return fail(MPRES_ERR_ARG, "invalid comparison %d", mode);
    if (count <= 0) return MPRES_OK;
    if (!x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->cmp(c->dev, mode, count, (const char *)x, (const char *)y, nullptr, nullptr, r, strm(st)));
    return MPRES_OK;
}

int mpres_rns_cmp(mpres_ctx_t *c, int count, const int *x, const int *y, int *r, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->cmp(c->dev, 2, count, nullptr, nullptr, x, y, r, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- export
int mpres_export_limbs(const mpres_ctx_t *c) {
    if (!c) return 0;
    return (31 * c->host.n + 31) / 32 + 1;  // export_limbs<N>() of the kernels
}

int mpres_get_d(mpres_ctx_t *c, int count, const void *x, double *out, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !out) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->exp_out(c->dev, 0, count, (const char *)x, SoA{}, false, out, nullptr, nullptr, nullptr,
                                    strm(st)));
    return MPRES_OK;
}

int mpres_array_get_d(mpres_ctx_t *c, int count, const mp_array_t *x, double *out, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !out) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->exp_out(c->dev, 0, count, nullptr, soa_of(x), true, out, nullptr, nullptr, nullptr,
                                    strm(st)));
    return MPRES_OK;
}

int mpres_get_integer(mpres_ctx_t *c, int count, const void *x, unsigned *limbs, int *sign, int *exp, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !limbs || !sign || !exp) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->exp_out(c->dev, 1, count, (const char *)x, SoA{}, false, nullptr, limbs, sign, exp,
                                    strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- dot
size_t mpres_dot_workspace(const mpres_ctx_t *c, int n, int g3, int b3) {
    (void)n;
    if (!c || g3 <= 0 || b3 <= 0) return 0;
    return mpsize(c) * ((size_t)g3 * b3 + (size_t)g3 + (size_t)b3 + 1);
}

static int dot_common(mpres_ctx_t *c, WsUse &wu, int mode, int n, SoA xs, int incx, SoA ys, int incy,
                      const char *xa, const char *ya, int g3, int b3, void *ws, size_t wsb, hipStream_t st,
                      char **result) {
    const size_t S = mpsize(c);
    char *w;
    int rc = wu.get(ws, wsb, mpres_dot_workspace(c, n, g3, b3), &w);
    if (rc) return rc;
    const long long G = (long long)g3 * b3;
    char *p1 = w, *p2 = p1 + S * G, *p3 = p2 + S * g3, *p4 = p3 + S * b3;
    const Ops *o = sized(c, OP_DOT, n);
    const unsigned P = next_pow2((unsigned)b3);
    // pass 1: strided fold over the products, halving tree per block of b3 (src/mpreduct.cuh:38-74)
    {
        ProfScope ps(c, OP_DOT, st);
        HIPCHK(o->fold(c->dev, mode, 0, n, G, xs, incx, ys, incy, xa, ya, p1, st));
    }
    HIPCHK(o->fold(c->dev, mode, 1, n, G, xs, incx, ys, incy, xa, ya, p1, st));
    HIPCHK(o->tree(c->dev, p1, b3, P, g3, p2, st));
    // pass 2: one block of b3 over the g3 block results (src/mpreduct.cuh:81-112)
    SoA none{};
    HIPCHK(o->fold(c->dev, 2, 0, g3, b3, none, 1, none, 1, p2, nullptr, p3, st));
    HIPCHK(o->tree(c->dev, p3, b3, P, 1, p4, st));
    *result = p4;
    return MPRES_OK;
}

int mpres_dot(mpres_ctx_t *c, int n, const mp_array_t *x, int incx, const mp_array_t *y, int incy, mp_array_t *r,
              int g3, int b3, void *ws, size_t wsb, void *st) {
    if (!c || !x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;  // src/blas/dot.cuh:88-90
    if (g3 <= 0 || b3 <= 0) return fail(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g3, b3);
    char *res;
    WsUse wu(c, strm(st));
    int rc = dot_common(c, wu, 0, n, soa_of(x), incx, soa_of(y), incy, nullptr, nullptr, g3, b3, ws, wsb, strm(st), &res);
    if (rc) return rc;
    HIPCHK(c->ops[OP_DOT]->aos_to_soa(c->dev, res, soa_of(r), 1, strm(st)));
    return MPRES_OK;
}

size_t mpres_dot_v2_workspace(const mpres_ctx_t *c, int n, int blocks, int threads) {
    return mpres_dot_workspace(c, n, blocks, threads);
}

int mpres_dot_v2(mpres_ctx_t *c, int n, const void *x, const void *y, void *r, int blocks, int threads, void *ws,
                 size_t wsb, void *st) {
    if (!c || !x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;
    if (blocks <= 0 || threads <= 0) return fail(MPRES_ERR_ARG, "invalid launch shape %d x %d", blocks, threads);
    char *res;
    SoA none{};
    WsUse wu(c, strm(st));
    int rc = dot_common(c, wu, 1, n, none, 1, none, 1, (const char *)x, (const char *)y, blocks, threads, ws, wsb, strm(st),
                        &res);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(r, res, mpsize(c), hipMemcpyDeviceToDevice, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- norm2 (v2)
size_t mpres_norm2_v2_workspace(const mpres_ctx_t *c, int n, int blocks, int threads) {
    if (!c) return 0;
    return mpres_dot_workspace(c, n, blocks, threads) + 16;
}

// sum of squares in the order of mp_sumsq_kernel1/2 (the v2 dot of x with itself: the squares
// are nonnegative, and mp_add_abs of nonnegative operands is mp_add), then mp_dsqrt; the
// double lands in *dres (device) and, after a stream synchronisation, in *result
static int norm2_into(mpres_ctx_t *c, int n, const void *x, int blocks, int threads, char *w, size_t wb, double *dres,
                      double *result, hipStream_t st) {
    char *res;
    WsUse inner(c, st);
    SoA none{};
    int rc = dot_common(c, inner, 1, n, none, 1, none, 1, (const char *)x, (const char *)x, blocks, threads, w, wb, st,
                        &res);
    if (rc) return rc;
    HIPCHK(c->ops[OP_ELEM]->hostop(c->dev, 2, 1, res, nullptr, nullptr, dres, st));
    HIPCHK(hipMemcpyAsync(result, dres, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MPRES_OK;
}

int mpres_norm2_v2(mpres_ctx_t *c, int n, const void *x, int blocks, int threads, double *result, void *ws, size_t wsb,
                   void *st) {
    if (!c || !x || !result) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0 || blocks <= 0 || threads <= 0) return fail(MPRES_ERR_ARG, "invalid size or launch shape");
    const size_t dw = mpres_dot_workspace(c, n, blocks, threads);
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, dw + 16, &w);
    if (rc) return rc;
    return norm2_into(c, n, x, blocks, threads, w, dw, (double *)(w + dw), result, strm(st));
}

// ---------------------------------------------------------------- CG / PCG (CSR)
size_t mpres_cg_workspace(const mpres_ctx_t *c, int n, int threads, int blocks_reduce) {
    if (!c || n <= 0 || threads <= 0 || blocks_reduce <= 0) return 0;
    return mpsize(c) * (4 * (size_t)n + 5) + mpres_dot_workspace(c, n, blocks_reduce, threads) + 16 +
           spmv_ws_bytes(n);
}

static int spmv_common(mpres_ctx_t *c, int fmt, int m, const SpMat &A, int tpr, const void *x, void *y, void *st,
                       int *list);

// y = A x in the solver's format (the scalar SpMV kernels: one thread per row in the reference);
// list: the solver workspace's room for the CSR kernels' long-row list
static int solver_spmv(mpres_ctx_t *c, int n, const mpres_sparse_t *A, const void *x, void *y, void *st, int *list) {
    switch (A->format) {
        case MPRES_SPARSE_CSR:
            if (!A->csr.as || !A->csr.ja || !A->csr.irp) return fail(MPRES_ERR_ARG, "null argument");
            return spmv_common(c, SP_CSR, n, SpMat{A->csr.as, A->csr.ja, A->csr.irp, nullptr, nullptr, 0, 0}, 1, x,
                               y, st, list);
        case MPRES_SPARSE_ELL:
            if (A->maxnzr < 0) return fail(MPRES_ERR_ARG, "maxnzr %d", A->maxnzr);
            return spmv_common(c, SP_ELL, n, SpMat{A->ell.as, A->ell.ja, nullptr, nullptr, nullptr, A->maxnzr, 0}, 1,
                               x, y, st, list);
        case MPRES_SPARSE_JAD:
            if (A->maxnzr < 0) return fail(MPRES_ERR_ARG, "maxnzr %d", A->maxnzr);
            return spmv_common(c, SP_JAD, n, SpMat{A->jad.as, A->jad.ja, A->jad.jcp, A->jad.perm, nullptr, A->maxnzr, 0},
                               1, x, y, st, list);
        case MPRES_SPARSE_DIA:
            if (A->ndiag < 0) return fail(MPRES_ERR_ARG, "ndiag %d", A->ndiag);
            return spmv_common(c, SP_DIA, n, SpMat{A->dia.as, nullptr, nullptr, nullptr, A->dia.offset, A->ndiag, n}, 1,
                               x, y, st, list);
        default: return fail(MPRES_ERR_ARG, "unknown sparse format %d", A->format);
    }
}

int mpres_cg_csr(mpres_ctx_t *c, int n, const csr_t *A, const void *b, double tol, int maxit, void *x,
                 const double *precond, double *resvec, int *iters, int threads, int blocks_reduce, void *ws,
                 size_t wsb, void *st) {
    if (!A) return fail(MPRES_ERR_ARG, "null argument");
    mpres_sparse_t S{};
    S.format = MPRES_SPARSE_CSR;
    S.csr = *A;
    return mpres_cg(c, n, &S, b, tol, maxit, x, precond, resvec, iters, threads, blocks_reduce, ws, wsb, st);
}

int mpres_cg(mpres_ctx_t *c, int n, const mpres_sparse_t *A, const void *b, double tol, int maxit, void *x,
             const double *precond, double *resvec, int *iters, int threads, int blocks_reduce, void *ws, size_t wsb,
             void *st) {
    if (!c || !A || !b || !x || !resvec || !iters) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0 || maxit < 0 || threads <= 0 || blocks_reduce <= 0) return fail(MPRES_ERR_ARG, "invalid arguments");
    const hipStream_t s = strm(st);
    const size_t S = mpsize(c), dw = mpres_dot_workspace(c, n, blocks_reduce, threads);
    char *w;
    WsUse wu(c, s);
    int rc = wu.get(ws, wsb, mpres_cg_workspace(c, n, threads, blocks_reduce), &w);
    if (rc) return rc;
    char *r = w, *p = r + S * n, *q = p + S * n, *z = q + S * n, *rho = z + S * n, *rhop = rho + S, *alpha = rhop + S,
         *beta = alpha + S, *pq = beta + S, *dws = pq + S;
    double *dres = (double *)(dws + dw);
    int *list = (int *)(dws + dw + 16);
    char *zz = precond ? z : r;  // z = M^-1 r (diagonal preconditioner) or r itself
#define CGCHK(call)                   \
    do {                              \
        if (int e_ = (call)) return e_; \
    } while (0)
    // r = b - A x (src/sparse/solver/cg_csr.cuh:78-80)
    CGCHK(solver_spmv(c, n, A, x, r, st, list));
    CGCHK(mpres_diff_v2(c, n, b, r, r, st));
    double norm0 = 0, normk = 0;
    CGCHK(norm2_into(c, n, r, blocks_reduce, threads, dws, dw, dres, &norm0, s));
    const double eps = norm0 * tol;
    normk = norm0;
    int k = 0;
    resvec[0] = normk / norm0;
    while (normk > eps && k < maxit) {  // :88-114 (pcg_csr.cuh:88-127)
        if (precond) CGCHK(mpres_prod_d_v2(c, n, r, precond, z, st));
        HIPCHK(hipMemcpyAsync(rhop, rho, S, hipMemcpyDeviceToDevice, s));
        CGCHK(mpres_dot_v2(c, n, r, zz, rho, blocks_reduce, threads, dws, dw, st));
        if (k == 0) {
            HIPCHK(hipMemcpyAsync(p, zz, S * n, hipMemcpyDeviceToDevice, s));
        } else {
            CGCHK(mpres_div(c, 1, rho, rhop, beta, st));
            CGCHK(mpres_axpy_v2(c, n, beta, p, zz, p, st));
        }
        CGCHK(solver_spmv(c, n, A, p, q, st, list));
        CGCHK(mpres_dot_v2(c, n, p, q, pq, blocks_reduce, threads, dws, dw, st));
        CGCHK(mpres_div(c, 1, rho, pq, alpha, st));
        CGCHK(mpres_axpy_v2(c, n, alpha, p, x, x, st));
        CGCHK(mpres_maxpy_v2(c, n, alpha, q, r, r, st));
        CGCHK(norm2_into(c, n, r, blocks_reduce, threads, dws, dw, dres, &normk, s));
        resvec[k + 1] = normk / norm0;
        k++;
    }
#undef CGCHK
    *iters = k;
    return MPRES_OK;
}

// ---------------------------------------------------------------- axpy
int mpres_axpy(mpres_ctx_t *c, int n, const mp_array_t *alpha, const mp_array_t *x, int incx, mp_array_t *y, int incy,
               void *st) {
    if (!c || !alpha || !x || !y) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;  // src/blas/axpy.cuh:49-51
    ProfScope ps(c, OP_AXPY, strm(st));
    HIPCHK(sized(c, OP_AXPY, n)->axpy(c->dev, n, soa_of(alpha), soa_of(x), incx, soa_of(y), incy, strm(st)));
    return MPRES_OK;
}

int mpres_axpy_v2(mpres_ctx_t *c, int n, const void *alpha, const void *x, const void *y, void *r, void *st) {
    if (!c || !alpha || !x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;
    HIPCHK(sized(c, OP_AXPY, n)->axpy_v2(c->dev, n, (const char *)alpha, (const char *)x, (const char *)y,
                                         (char *)r, false, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- v2 helpers of the solvers
int mpres_maxpy_v2(mpres_ctx_t *c, int n, const void *alpha, const void *x, const void *y, void *r, void *st) {
    if (!c || !alpha || !x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;
    HIPCHK(sized(c, OP_AXPY, n)->axpy_v2(c->dev, n, (const char *)alpha, (const char *)x, (const char *)y,
                                         (char *)r, true, strm(st)));
    return MPRES_OK;
}

int mpres_diff_v2(mpres_ctx_t *c, int n, const void *x, const void *y, void *r, void *st) {
    return mpres_elementwise(c, MPRES_OP_SUB, n, x, y, r, st);
}

int mpres_prod_d_v2(mpres_ctx_t *c, int n, const void *x, const double *d, void *r, void *st) {
    return mpres_elementwise(c, MPRES_OP_MUL_D, n, x, d, r, st);
}

// ---------------------------------------------------------------- the reference's host MPFR operations
int mpres_div(mpres_ctx_t *c, int count, const void *x, const void *y, void *r, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !y || !r) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->hostop(c->dev, 0, count, (const char *)x, (const char *)y, (char *)r, nullptr, strm(st)));
    return MPRES_OK;
}

int mpres_sqrt(mpres_ctx_t *c, int count, const void *x, void *r, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !r) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->hostop(c->dev, 1, count, (const char *)x, nullptr, (char *)r, nullptr, strm(st)));
    return MPRES_OK;
}

int mpres_dsqrt(mpres_ctx_t *c, int count, const void *x, double *out, void *st) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    if (count <= 0) return MPRES_OK;
    if (!x || !out) return fail(MPRES_ERR_ARG, "null argument");
    HIPCHK(c->ops[OP_ELEM]->hostop(c->dev, 2, count, (const char *)x, nullptr, nullptr, out, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- gemv
size_t mpres_gemv_workspace(const mpres_ctx_t *c, int trans, int m, int n, int b3) {
    if (!c || m <= 0 || n <= 0 || b3 <= 0) return 0;
    const bool tr = trans != MPRES_NO_TRANS;
    const size_t lx = tr ? m : n, ly = tr ? n : m;
    // alpha*x, the ly*b3 partials, the ly row sums, then the main kernel's auxiliary data
    return mpsize(c) * (lx + ly * (size_t)b3 + ly) + c->ops[OP_GEMV]->gemv_aux(tr, m, n);
}

int mpres_gemv(mpres_ctx_t *c, int trans, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
               const mp_array_t *x, int incx, const mp_array_t *beta, mp_array_t *y, int incy, int b3, void *ws,
               size_t wsb, void *st) {
    if (!c || !alpha || !A || !x || !beta || !y) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/gemv.cuh:155-157
    if (incx == 0 || incy == 0 || lda < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid increments or leading dimension");  // :159-161
    if (b3 <= 0) return fail(MPRES_ERR_ARG, "invalid blockDim3 %d", b3);
    if (trans != MPRES_NO_TRANS && trans != MPRES_TRANS && trans != MPRES_CONJ_TRANS)
        return fail(MPRES_ERR_ARG, "invalid trans %d", trans);
    const bool tr = trans != MPRES_NO_TRANS;
    const int lx = tr ? m : n, ly = tr ? n : m;
    const size_t S = mpsize(c);
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_gemv_workspace(c, trans, m, n, b3), &w);
    if (rc) return rc;
    char *bx = w, *part = bx + S * lx, *sums = part + S * (size_t)ly * b3, *aux = sums + S * (size_t)ly;
    const Ops *o = sized(c, OP_GEMV, (long long)m * n);
    SoA ys = soa_of(y);
    // buffer1 = alpha * x, y = beta * y (src/blas/gemv.cuh:175-190)
    HIPCHK(o->scal(c->dev, lx, soa_of(alpha), soa_of(x), incx, bx, false, strm(st)));
    HIPCHK(o->scal(c->dev, ly, soa_of(beta), ys, incy, nullptr, true, strm(st)));
    // per-output strided sums of A .* buffer1 and the halving tree (:199-218, :44-76).
    // Split form: the residues (pass 1) run on the side stream while the scalar chains
    // (pass 2) run on st.
    const long long nvt = (long long)ly * b3;
    HIPCHK(o->gemv_fold(c->dev, tr, 0, m, n, soa_of(A), lda, bx, b3, nvt, part, aux, strm(st)));
    if (o->gemv_split(tr, b3)) {
        ProfScope pair(c, PROF_GEMV_PAIR, strm(st));
        Fork fk(c, strm(st));
        {
            ProfScope ps(c, OP_GEMV, fk.s1);
            HIPCHK(o->gemv_fold(c->dev, tr, 1, m, n, soa_of(A), lda, bx, b3, nvt, part, aux, fk.s1));
        }
        HIPCHK(o->gemv_fold(c->dev, tr, 2, m, n, soa_of(A), lda, bx, b3, nvt, part, aux, strm(st)));
        HIPCHK(fk.join());
    } else {
        ProfScope ps(c, OP_GEMV, strm(st));
        HIPCHK(o->gemv_fold(c->dev, tr, 1, m, n, soa_of(A), lda, bx, b3, nvt, part, aux, strm(st)));
    }
    HIPCHK(o->gemv_fold(c->dev, tr, 3, m, n, soa_of(A), lda, bx, b3, nvt, part, aux, strm(st)));
    HIPCHK(o->tree(c->dev, part, b3, next_pow2((unsigned)b3), ly, sums, strm(st)));
    HIPCHK(o->gemv_final(c->dev, ly, ys, incy, sums, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- gemm
size_t mpres_gemm_workspace(const mpres_ctx_t *c, int m, int n, int k) {
    if (!c || m <= 0 || n <= 0 || k < 0) return 0;
    // the m*n dot-product sums, and over 16/32 moduli the pre-scaled residue planes
    return c->ops[OP_GEMM]->gemm_ws(m, n, k);
}

int mpres_gemm(mpres_ctx_t *c, int ta, int tb, int m, int n, int k, const mp_array_t *alpha, const mp_array_t *A,
               int lda, const mp_array_t *B, int ldb, const mp_array_t *beta, mp_array_t *C, int ldc, void *ws,
               size_t wsb, void *st) {
    if (!c || !alpha || !A || !B || !beta || !C) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0 || k <= 0) return MPRES_OK;  // src/blas/gemm.cuh:75-77
    if ((ta == MPRES_NO_TRANS && lda < std::max(1, m)) || (ta != MPRES_NO_TRANS && lda < std::max(1, k)) ||
        (tb == MPRES_NO_TRANS && ldb < std::max(1, k)) || (tb != MPRES_NO_TRANS && ldb < std::max(1, n)) ||
        ldc < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid leading dimension");  // :82-96
    if (ta != MPRES_NO_TRANS || tb != MPRES_NO_TRANS)
        return fail(MPRES_ERR_UNSUPPORTED, "only non-transposed matrices are supported");  // :114-125
    if (!digits_aligned(A) || !digits_aligned(B))
        return fail(MPRES_ERR_ARG, "digits of A and B must be 16-byte aligned");
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_gemm_workspace(c, m, n, k), &w);
    if (rc) return rc;
    const Ops *o = sized(c, OP_GEMM, (long long)m * n * k);
    // the residue path (pass 0) runs on the side stream while the chains (pass 1) run on st
    Fork fk(c, strm(st));
    {
        ProfScope ps(c, PROF_GEMM_RESID, fk.s1);
        HIPCHK(o->gemm(c->dev, 0, c->bits, m, n, k, soa_of(alpha), soa_of(A), lda, soa_of(B), ldb, soa_of(beta),
                       soa_of(C), ldc, w, fk.s1));
    }
    {
        ProfScope ps(c, OP_GEMM, strm(st));
        HIPCHK(o->gemm(c->dev, 1, c->bits, m, n, k, soa_of(alpha), soa_of(A), lda, soa_of(B), ldb, soa_of(beta),
                       soa_of(C), ldc, w, strm(st)));
    }
    HIPCHK(fk.join());
    for (int pass = 2; pass < 5; pass++)
        HIPCHK(o->gemm(c->dev, pass, c->bits, m, n, k, soa_of(alpha), soa_of(A), lda, soa_of(B), ldb, soa_of(beta),
                       soa_of(C), ldc, w, strm(st)));
    return MPRES_OK;
}

size_t mpres_gemm_mfma_workspace(const mpres_ctx_t *c, int m, int n, int k) {
    if (!c || m <= 0 || n <= 0 || k < 0) return 0;
    return c->ops[OP_MFMA]->gemm_mfma_ws(c->bits, m, n, k);
}

int mpres_gemm_mfma(mpres_ctx_t *c, int ta, int tb, int m, int n, int k, const mp_array_t *alpha,
                    const mp_array_t *A, int lda, const mp_array_t *B, int ldb, const mp_array_t *beta, mp_array_t *C,
                    int ldc, void *ws, size_t wsb, void *st) {
    if (!c || !alpha || !A || !B || !beta || !C) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;
    const bool tra = ta == MPRES_TRANS || ta == MPRES_CONJ_TRANS, trb = tb == MPRES_TRANS || tb == MPRES_CONJ_TRANS;
    if (k < 0 || lda < std::max(1, tra ? k : m) || ldb < std::max(1, trb ? n : k) || ldc < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid dimensions");
    if (k > MPRES_MFMA_MAX_K) return fail(MPRES_ERR_UNSUPPORTED, "k = %d exceeds %d", k, MPRES_MFMA_MAX_K);
    if (!digits_aligned(A) || !digits_aligned(B))
        return fail(MPRES_ERR_ARG, "digits of A and B must be 16-byte aligned");
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_gemm_mfma_workspace(c, m, n, k), &w);
    if (rc) return rc;
    ProfScope ps(c, PROF_GEMM_MFMA, strm(st));
    HIPCHK(sized(c, OP_MFMA, (long long)m * n * k)->gemm_mfma(c->dev, c->bits, c->host.log2m, tra, trb, m, n, k, soa_of(alpha), soa_of(A), lda,
                                      soa_of(B), ldb, soa_of(beta), soa_of(C), ldc, w, strm(st)));
    return MPRES_OK;
}

int mpres_gemm_v2(mpres_ctx_t *c, int ta, int tb, int m, int n, int k, const void *alpha, const void *A, int lda,
                  const void *B, int ldb, const void *beta, void *C, int ldc, void *ws, size_t wsb, void *st) {
    if (!c || !alpha || !A || !B || !beta || !C) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;
    if (k < 0 || lda <= 0 || ldb <= 0 || ldc < std::max(1, m)) return fail(MPRES_ERR_ARG, "invalid dimensions");
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_gemm_workspace(c, m, n, k), &w);
    if (rc) return rc;
    const int tra = ta == MPRES_TRANS || ta == MPRES_CONJ_TRANS, trb = tb == MPRES_TRANS || tb == MPRES_CONJ_TRANS;
    // as in mpres_gemm: the residue path (pass 0) on the side stream, the chains on st
    const Ops *o = sized(c, OP_GEMM, (long long)m * n * k);
    Fork fk(c, strm(st));
    for (int pass = 0; pass < 5; pass++) {
        if (pass == 2) HIPCHK(fk.join());
        HIPCHK(o->gemm_v2(c->dev, pass, c->bits, tra, trb, m, n, k, (const char *)alpha, (const char *)A,
                                        lda, (const char *)B, ldb, (const char *)beta, (char *)C, ldc, w,
                                        pass == 0 ? fk.s1 : strm(st)));
    }
    return MPRES_OK;
}

// ---------------------------------------------------------------- SpMV
// one SpMV in any storage format: the main kernel, then the fixup of the rows it marked
// list: spmv_ws_bytes(m) for the split form's row order and lists (NULL: from the context's pool)
static int spmv_common(mpres_ctx_t *c, int fmt, int m, const SpMat &A, int tpr, const void *x, void *y, void *st,
                       int *list) {
    if (m <= 0) return MPRES_OK;
    if (tpr < 1 || tpr > 32 || (tpr & (tpr - 1))) return fail(MPRES_ERR_ARG, "threadsPerRow %d not in 1..32", tpr);
    const Ops *o = sized(c, OP_SPMV, m);
    if (tpr > 256 / o->L) {
        // the tree of a row must fit one workgroup: fall back to a narrower group
        for (int L = o->L / 2; L >= 1; L /= 2) {
            const Ops *q = find_ops(c->host.n, L);
            if (q && tpr <= 256 / L) {
                o = q;
                break;
            }
        }
        if (tpr > 256 / o->L) return fail(MPRES_ERR_UNSUPPORTED, "threadsPerRow %d with %d lanes", tpr, o->L);
    }
    // the split form's row order and row lists
    char *w = (char *)list;
    WsUse wu(c, strm(st));
    if (!w) {
        const int rc = wu.get(nullptr, 0, spmv_ws_bytes(m), &w);
        if (rc) return rc;
    }
    {
        ProfScope ps(c, OP_SPMV, strm(st));
        HIPCHK(o->spmv(c->dev, 0, fmt, m, A, (const char *)x, (char *)y, tpr, (int *)w, strm(st)));
        if (o->spmv_two(fmt)) {
            // the long rows' residues on the side stream beside the short rows' (disjoint rows of y)
            Fork fk(c, strm(st));
            HIPCHK(o->spmv(c->dev, 3, fmt, m, A, (const char *)x, (char *)y, tpr, (int *)w, fk.s1));
            HIPCHK(o->spmv(c->dev, 2, fmt, m, A, (const char *)x, (char *)y, tpr, (int *)w, strm(st)));
            HIPCHK(fk.join());
        }
    }
    HIPCHK(o->spmv(c->dev, 1, fmt, m, A, (const char *)x, (char *)y, tpr, (int *)w, strm(st)));
    return MPRES_OK;
}

int mpres_spmv_csrv(mpres_ctx_t *c, int m, const csr_t *csr, int tpr, const void *x, void *y, void *st) {
    if (!c || !csr || !x || !y) return fail(MPRES_ERR_ARG, "null argument");
    return spmv_common(c, SP_CSR, m, SpMat{csr->as, csr->ja, csr->irp, nullptr, nullptr, 0, 0}, tpr, x, y, st, nullptr);
}

int mpres_spmv_csr(mpres_ctx_t *c, int m, const csr_t *csr, const void *x, void *y, void *st) {
    return mpres_spmv_csrv(c, m, csr, 1, x, y, st);
}

int mpres_spmv_ell(mpres_ctx_t *c, int m, int maxnzr, const ell_t *ell, const void *x, void *y, void *st) {
    if (!c || !ell || !x || !y) return fail(MPRES_ERR_ARG, "null argument");
    if (maxnzr < 0) return fail(MPRES_ERR_ARG, "maxnzr %d", maxnzr);
    return spmv_common(c, SP_ELL, m, SpMat{ell->as, ell->ja, nullptr, nullptr, nullptr, maxnzr, 0}, 1, x, y, st, nullptr);
}

int mpres_spmv_jadv(mpres_ctx_t *c, int m, int maxnzr, const jad_t *jad, int tpr, const void *x, void *y, void *st) {
    if (!c || !jad || !x || !y) return fail(MPRES_ERR_ARG, "null argument");
    if (maxnzr < 0) return fail(MPRES_ERR_ARG, "maxnzr %d", maxnzr);
    return spmv_common(c, SP_JAD, m, SpMat{jad->as, jad->ja, jad->jcp, jad->perm, nullptr, maxnzr, 0}, tpr, x, y,
                       st, nullptr);
}

int mpres_spmv_jad(mpres_ctx_t *c, int m, int maxnzr, const jad_t *jad, const void *x, void *y, void *st) {
    return mpres_spmv_jadv(c, m, maxnzr, jad, 1, x, y, st);
}

int mpres_spmv_dia(mpres_ctx_t *c, int m, int n, int ndiag, const dia_t *dia, const void *x, void *y, void *st) {
    if (!c || !dia || !x || !y) return fail(MPRES_ERR_ARG, "null argument");
    if (ndiag < 0 || n < 0) return fail(MPRES_ERR_ARG, "invalid dimensions");
    return spmv_common(c, SP_DIA, m, SpMat{dia->as, nullptr, nullptr, nullptr, dia->offset, ndiag, n}, 1, x, y, st, nullptr);
}

// ---------------------------------------------------------------- profiling
int mpres_profile(mpres_ctx_t *c, int enable) {
    if (!c) return fail(MPRES_ERR_ARG, "null context");
    c->prof = enable != 0;
    return MPRES_OK;
}

int mpres_profile_read(mpres_ctx_t *c, int op, double *ms, long *launches, int reset) {
    if (!c || op < 0 || op >= PROF_SLOTS) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    for (auto &e : c->prof_live) {
        float t = 0;
        HIPCHK(hipEventSynchronize(e.second.second));
        HIPCHK(hipEventElapsedTime(&t, e.second.first, e.second.second));
        c->prof_ms[e.first] += t;
        c->prof_count[e.first] += 1;
        c->prof_free.push_back(e.second);
    }
    c->prof_live.clear();
    if (ms) *ms = c->prof_ms[op];
    if (launches) *launches = c->prof_count[op];
    if (reset) {
        c->prof_ms[op] = 0;
        c->prof_count[op] = 0;
    }
    return MPRES_OK;
}

// ---------------------------------------------------------------- batched form
}  // extern "C"

struct mpres_batch {
    mpres_ctx *ctx = nullptr;
    const Ops *ops = nullptr;
    char *dmem = nullptr;  // item table, then the queue counter
    int count = 0;
    char *ws = nullptr;
    bool own_ws = false;
    double fmas = 0;
    mpres_batch_tuning_t used{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int batch_resolve(const mpres_ctx *c, const mpres_batch_tuning_t *t, mpres_batch_tuning_t &u, const Ops **ops) {
    // built-in geometry: the autotuner's best on the batching workload at 16 moduli
    // (mpres-blas_amd/tuning/batch_gfx950.json): the gemv lanes (4 at 16 moduli), 4
    // workgroups per CU, 128 gemv rows per item
    u.lanes = (t && t->lanes) ? t->lanes : c->ops[OP_GEMV]->L;
    u.wg_per_cu = (t && t->wg_per_cu) ? t->wg_per_cu : 4;
    u.gemv_rows = (t && t->gemv_rows) ? t->gemv_rows : 128;
    if (u.wg_per_cu < 0 || u.wg_per_cu > 32 || u.gemv_rows < 0)
        return fail(MPRES_ERR_ARG, "invalid batch tuning (%d workgroups per CU, %d rows)", u.wg_per_cu, u.gemv_rows);
    *ops = find_ops(c->host.n, u.lanes);
    if (!*ops) return fail(MPRES_ERR_UNSUPPORTED, "%d lanes per number not compiled for %d moduli", u.lanes, c->host.n);
    return MPRES_OK;
}

void batch_free(mpres_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->dmem) (void)hipFree(b->dmem);
    if (b->own_ws && b->ws) (void)hipFree(b->ws);
    delete b;
}

}  // namespace

extern "C" {

size_t mpres_batch_workspace(const mpres_ctx_t *c, const mpres_batch_problem_t *probs, int count,
                             const mpres_batch_tuning_t *tuning) {
    if (!c) return 0;
    mpres_batch_tuning_t u;
    const Ops *o;
    if (batch_resolve(c, tuning, u, &o)) return 0;
    BatchLayout lay;
    if (batch_layout(probs, count, c->host.n, u.gemv_rows, lay)) return 0;
    return lay.ws_bytes;
}

int mpres_batch_plan(mpres_ctx_t *c, const mpres_batch_problem_t *probs, int count,
                     const mpres_batch_tuning_t *tuning, void *ws, size_t wsb, mpres_batch_t **plan) {
    if (!c || !plan) return fail(MPRES_ERR_ARG, "null argument");
    *plan = nullptr;
    mpres_batch_tuning_t u;
    const Ops *o;
    int rc = batch_resolve(c, tuning, u, &o);
    if (rc) return rc;
    BatchLayout lay;
    rc = batch_layout(probs, count, c->host.n, u.gemv_rows, lay);
    if (rc) return fail(rc, "%s", lay.err.c_str());
    if (ws && wsb < lay.ws_bytes)
        return fail(MPRES_ERR_WORKSPACE, "workspace of %zu bytes, %zu needed", wsb, lay.ws_bytes);
    auto *b = new mpres_batch();
    b->ctx = c;
    b->ops = o;
    b->count = (int)lay.items.size();
    b->fmas = lay.fmas;
    b->used = u;
    *plan = b;
    if (b->count == 0) return MPRES_OK;
    auto bail = [&](const char *what, hipError_t e) {
        batch_free(b);
        *plan = nullptr;
        return fail(MPRES_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    if (ws) {
        b->ws = (char *)ws;
    } else if (lay.ws_bytes) {
        if ((e = hipMalloc(&b->ws, lay.ws_bytes)) != hipSuccess) return bail("hipMalloc(batch scratch)", e);
        b->own_ws = true;
    }
    const size_t tbl = lay.items.size() * sizeof(BatchItem);
    const size_t qoff = (tbl + 255) / 256 * 256;
    if ((e = hipMalloc(&b->dmem, qoff + 256)) != hipSuccess) return bail("hipMalloc(batch items)", e);
    if ((e = hipMemcpy(b->dmem, lay.items.data(), tbl, hipMemcpyHostToDevice)) != hipSuccess)
        return bail("hipMemcpy(batch items)", e);
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device)) != hipSuccess)
        return bail("hipDeviceGetAttribute", e);
    const int grid = std::max(1, std::min(b->count, cus * std::max(1, u.wg_per_cu)));
    // capture the queue reset and the persistent kernel once
    hipStream_t cs;
    if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return bail("hipStreamCreate", e);
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = o->batch(c->dev, (const BatchItem *)b->dmem, b->count, (int *)(b->dmem + qoff), b->ws,
                                       grid, cs);
        e = hipStreamEndCapture(cs, &b->graph);
        if (le != hipSuccess) e = le;
    }
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return bail("capturing the batch graph", e);
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
        return bail("hipGraphInstantiate", e);
    return MPRES_OK;
}

int mpres_batch_run(mpres_batch_t *b, void *st) {
    if (!b) return fail(MPRES_ERR_ARG, "null plan");
    if (b->count == 0) return MPRES_OK;
    ProfScope ps(b->ctx, PROF_BATCH, strm(st));
    HIPCHK(hipGraphLaunch(b->exec, strm(st)));
    return MPRES_OK;
}

int mpres_batch_info(const mpres_batch_t *b, double *fmas, int *items, mpres_batch_tuning_t *used) {
    if (!b) return fail(MPRES_ERR_ARG, "null plan");
    if (fmas) *fmas = b->fmas;
    if (items) *items = b->count;
    if (used) *used = b->used;
    return MPRES_OK;
}

int mpres_batch_destroy(mpres_batch_t *b) {
    if (!b) return MPRES_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    return MPRES_OK;
}

// ---------------------------------------------------------------- the rest of the v1 BLAS
// Each reference element kernel triple (esi, digits, round) is one k_v1map pass here; the
// passes run in the reference's order, so every intermediate the reference writes to a
// caller buffer is written there too.
static int v1map(mpres_ctx_t *c, int op, int sf, int m, int n, const mp_array_t *S, int incs, const mp_array_t *X,
                 int ldx, int incx, const mp_array_t *Y, int ldy, int incy, mp_array_t *W, int ldw, int incw,
                 hipStream_t st) {
    HIPCHK(c->ops[OP_ELEM]->v1map(c->dev, op, sf, m, n, S ? soa_of(S) : SoA{}, incs, soa_of(X), ldx, incx,
                                  Y ? soa_of(Y) : SoA{}, ldy, incy, soa_of(W), ldw, incw, st));
    return MPRES_OK;
}
#define V1(...)                            \
    do {                                   \
        const int rc_ = v1map(__VA_ARGS__); \
        if (rc_) return rc_;               \
    } while (0)

int mpres_scal(mpres_ctx_t *c, int n, const mp_array_t *alpha, mp_array_t *x, int incx, void *st) {
    if (!c || !alpha || !x) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0 || incx <= 0) return MPRES_OK;  // src/blas/scal.cuh:46-48
    V1(c, V1_MUL, SF_SCALAR, n, 1, alpha, 1, x, 0, incx, nullptr, 0, 0, x, 0, incx, strm(st));
    return MPRES_OK;
}

int mpres_rot(mpres_ctx_t *c, int n, mp_array_t *x, int incx, mp_array_t *y, int incy, const mp_array_t *cc,
              const mp_array_t *s, mp_array_t *buffer1, mp_array_t *buffer2, void *st) {
    if (!c || !x || !y || !cc || !s || !buffer1 || !buffer2) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;  // src/blas/rot.cuh:52-54
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, n, 1, s, 1, x, 0, incx, nullptr, 0, 0, buffer1, 0, 1, q);  // buffer1 = s x
    V1(c, V1_MUL, SF_SCALAR, n, 1, s, 1, y, 0, incy, nullptr, 0, 0, buffer2, 0, 1, q);  // buffer2 = s y
    // x = c x and y = c y with increment 1 whatever incx and incy are (src/blas/rot.cuh:78-81 call mp_scal so)
    V1(c, V1_MUL, SF_SCALAR, n, 1, cc, 1, x, 0, 1, nullptr, 0, 0, x, 0, 1, q);
    V1(c, V1_MUL, SF_SCALAR, n, 1, cc, 1, y, 0, 1, nullptr, 0, 0, y, 0, 1, q);
    V1(c, V1_ADD, 0, n, 1, nullptr, 0, x, 0, incx, buffer2, 0, 1, x, 0, incx, q);  // x = x + buffer2
    V1(c, V1_SUB, 0, n, 1, nullptr, 0, y, 0, incy, buffer1, 0, 1, y, 0, incy, q);  // y = y - buffer1
    return MPRES_OK;
}

int mpres_waxpby(mpres_ctx_t *c, int n, const mp_array_t *alpha, const mp_array_t *x, int incx,
                 const mp_array_t *beta, const mp_array_t *y, int incy, mp_array_t *w, int incw, mp_array_t *buffer,
                 void *st) {
    if (!c || !alpha || !x || !beta || !y || !w || !buffer) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;  // src/blas/waxpby.cuh:54-56
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, n, 1, alpha, 1, x, 0, incx, nullptr, 0, 0, buffer, 0, 1, q);  // buffer = alpha x
    V1(c, V1_MUL, SF_SCALAR, n, 1, beta, 1, y, 0, incy, nullptr, 0, 0, w, 0, incw, q);     // w = beta y
    V1(c, V1_ADD, 0, n, 1, nullptr, 0, w, 0, incw, buffer, 0, 1, w, 0, incw, q);           // w = w + buffer
    return MPRES_OK;
}

int mpres_axpy_dot(mpres_ctx_t *c, int n, const mp_array_t *alpha, mp_array_t *w, int incw, const mp_array_t *v,
                   int incv, const mp_array_t *u, int incu, mp_array_t *r, mp_array_t *buffer, int g3, int b3,
                   void *ws, size_t wsb, void *st) {
    if (!c || !alpha || !w || !v || !u || !r || !buffer) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;  // src/blas/axpydot.cuh:57-59
    if (g3 <= 0 || b3 <= 0) return fail(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g3, b3);
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, n, 1, alpha, 1, v, 0, incv, nullptr, 0, 0, buffer, 0, 1, q);  // buffer = alpha v
    V1(c, V1_SUB, 0, n, 1, nullptr, 0, w, 0, incw, buffer, 0, 1, w, 0, incw, q);           // w = w - buffer
    return mpres_dot(c, n, u, incu, w, incw, r, g3, b3, ws, wsb, st);                       // r = u . w
}

// ---- sum and maximum of absolute values: the fold over grid1 x block1 virtual threads,
// the tree per block, then one block of block1 over the grid1 block results
size_t mpres_asum_workspace(const mpres_ctx_t *c, int n, int g1, int b1) { return mpres_dot_workspace(c, n, g1, b1); }

// maximum |x_i| of an SoA vector into *res (AoS, sign 0); w holds (g1 b1 + g1 + b1 + 1) numbers
static int maxabs_into(mpres_ctx_t *c, int n, SoA x, int incx, int g1, int b1, char *w, hipStream_t st, char **res) {
    const size_t S = mpsize(c);
    const long long G = (long long)g1 * b1;
    char *p1 = w, *p2 = p1 + S * G, *p3 = p2 + S * g1, *p4 = p3 + S * b1;
    const Ops *o = c->ops[OP_DOT];
    const unsigned P = next_pow2((unsigned)b1);
    // src/mpreduct.cuh:199-238 over x, then :246-286 over the block results
    HIPCHK(o->maxabs(c->dev, 0, n, G, x, incx, nullptr, p1, b1, P, g1, p2, st));
    HIPCHK(o->maxabs(c->dev, 1, g1, b1, SoA{}, 1, p2, p3, b1, P, 1, p4, st));
    *res = p4;
    return MPRES_OK;
}

int mpres_asum(mpres_ctx_t *c, int n, const mp_array_t *x, int incx, mp_array_t *r, int g1, int b1, void *ws,
               size_t wsb, void *st) {
    if (!c || !x || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0 || incx <= 0) return MPRES_OK;  // src/blas/asum.cuh:42-44
    if (g1 <= 0 || b1 <= 0) return fail(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g1, b1);
    const size_t S = mpsize(c);
    char *w;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_asum_workspace(c, n, g1, b1), &w);
    if (rc) return rc;
    const long long G = (long long)g1 * b1;
    char *p1 = w, *p2 = p1 + S * G, *p3 = p2 + S * g1, *p4 = p3 + S * b1;
    const Ops *o = c->ops[OP_DOT];
    const unsigned P = next_pow2((unsigned)b1);
    // src/mpreduct.cuh:120-153: |x_i| summed by mp_add_abs, which on the nonnegative terms and
    // partial sums is mp_add; :161-192 over the block results
    HIPCHK(o->fold_abs(c->dev, n, G, soa_of(x), incx, p1, strm(st)));
    HIPCHK(o->tree(c->dev, p1, b1, P, g1, p2, strm(st)));
    SoA none{};
    HIPCHK(o->fold(c->dev, 2, 0, g1, b1, none, 1, none, 1, p2, nullptr, p3, strm(st)));
    HIPCHK(o->tree(c->dev, p3, b1, P, 1, p4, strm(st)));
    HIPCHK(o->aos_to_soa(c->dev, p4, soa_of(r), 1, strm(st)));
    return MPRES_OK;
}

int mpres_norm(mpres_ctx_t *c, int norm, int n, const mp_array_t *x, int incx, mp_array_t *r, int g1, int b1,
               void *ws, size_t wsb, void *st) {
    if (norm == MPRES_ONE_NORM) return mpres_asum(c, n, x, incx, r, g1, b1, ws, wsb, st);  // src/blas/norm.cuh:44-47
    if (!c || !x || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (norm != MPRES_INF_NORM) return fail(MPRES_ERR_ARG, "invalid norm %d", norm);
    if (n <= 0 || incx <= 0) return MPRES_OK;  // :49-51
    if (g1 <= 0 || b1 <= 0) return fail(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g1, b1);
    char *w, *res;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_asum_workspace(c, n, g1, b1), &w);
    if (rc) return rc;
    rc = maxabs_into(c, n, soa_of(x), incx, g1, b1, w, strm(st), &res);
    if (rc) return rc;
    HIPCHK(c->ops[OP_DOT]->aos_to_soa(c->dev, res, soa_of(r), 1, strm(st)));
    return MPRES_OK;
}

// ---- general matrices (column-major, SoA)
int mpres_ge_add(mpres_ctx_t *c, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
                 const mp_array_t *beta, const mp_array_t *B, int ldb, mp_array_t *C, int ldc, mp_array_t *buffer,
                 void *st) {
    if (!c || !alpha || !A || !beta || !B || !C || !buffer) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/geadd.cuh:61-63
    if (lda < std::max(1, m) || ldb < std::max(1, m) || ldc < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid leading dimension");  // :65-67
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, m, n, alpha, 1, A, lda, 1, nullptr, 0, 0, buffer, m, 1, q);  // buffer = alpha A
    V1(c, V1_MUL, SF_SCALAR, m, n, beta, 1, B, ldb, 1, nullptr, 0, 0, C, ldc, 1, q);      // C = beta B
    V1(c, V1_ADD, 0, m, n, nullptr, 0, C, ldc, 1, buffer, m, 1, C, ldc, 1, q);            // C = C + buffer
    return MPRES_OK;
}

int mpres_ge_acc(mpres_ctx_t *c, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
                 const mp_array_t *beta, mp_array_t *B, int ldb, mp_array_t *buffer, void *st) {
    if (!c || !alpha || !A || !beta || !B || !buffer) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/geacc.cuh:60-62
    if (lda < std::max(1, m) || ldb < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid leading dimension");  // :64-66
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, m, n, alpha, 1, A, lda, 1, nullptr, 0, 0, buffer, m, 1, q);  // buffer = alpha A
    V1(c, V1_MUL, SF_SCALAR, m, n, beta, 1, B, ldb, 1, nullptr, 0, 0, B, ldb, 1, q);      // B = beta B
    V1(c, V1_ADD, 0, m, n, nullptr, 0, B, ldb, 1, buffer, m, 1, B, ldb, 1, q);            // B = B + buffer
    return MPRES_OK;
}

int mpres_ge_diag_scale(mpres_ctx_t *c, int side, int m, int n, const mp_array_t *D, int incd, mp_array_t *A, int lda,
                        void *st) {
    if (!c || !D || !A) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/gediagscale.cuh:57-59
    if (incd == 0 || lda < std::max(1, m)) return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");
    if (side != MPRES_LEFT && side != MPRES_RIGHT) return fail(MPRES_ERR_ARG, "invalid side %d", side);
    // right: column j times D_j (A = A D, :75-81); left: row i times D_i (A = D A, :83-89)
    V1(c, V1_MUL, side == MPRES_RIGHT ? SF_COL : SF_ROW, m, n, D, incd, A, lda, 1, nullptr, 0, 0, A, lda, 1, strm(st));
    return MPRES_OK;
}

int mpres_ge_lr_scale(mpres_ctx_t *c, int m, int n, const mp_array_t *DL, int incdl, const mp_array_t *DR, int incdr,
                      mp_array_t *A, int lda, void *st) {
    if (!c || !DL || !DR || !A) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/gelrscale.cuh:59-61
    if (incdl == 0 || incdr == 0 || lda < std::max(1, m))
        return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");  // :63-65
    // A = DL A, then A = A DR (:78-92)
    V1(c, V1_MUL, SF_ROW, m, n, DL, incdl, A, lda, 1, nullptr, 0, 0, A, lda, 1, strm(st));
    V1(c, V1_MUL, SF_COL, m, n, DR, incdr, A, lda, 1, nullptr, 0, 0, A, lda, 1, strm(st));
    return MPRES_OK;
}

size_t mpres_ge_norm_workspace(const mpres_ctx_t *c, int norm, int m, int n, int g1, int b1) {
    if (!c || g1 <= 0 || b1 <= 0) return 0;
    const size_t outs = (size_t)std::max(norm == MPRES_ONE_NORM ? n : m, 0);
    return mpsize(c) * (outs * b1 + outs) + mpres_dot_workspace(c, 0, g1, b1);
}

int mpres_ge_norm(mpres_ctx_t *c, int norm, int m, int n, const mp_array_t *A, int lda, mp_array_t *r,
                  mp_array_t *buffer, int g1, int b1, void *ws, size_t wsb, void *st) {
    if (!c || !A || !r || !buffer) return fail(MPRES_ERR_ARG, "null argument");
    if (norm != MPRES_ONE_NORM && norm != MPRES_INF_NORM) return fail(MPRES_ERR_ARG, "invalid norm %d", norm);
    if (m <= 0 || n <= 0) return MPRES_OK;  // src/blas/genorm.cuh:144-146
    if (lda < std::max(1, m)) return fail(MPRES_ERR_ARG, "invalid leading dimension");  // :148-150
    if (g1 <= 0 || b1 <= 0) return fail(MPRES_ERR_ARG, "invalid reduction shape %d x %d", g1, b1);
    const bool cols = norm == MPRES_ONE_NORM;
    const int outs = cols ? n : m;
    const size_t S = mpsize(c);
    char *w, *res;
    WsUse wu(c, strm(st));
    int rc = wu.get(ws, wsb, mpres_ge_norm_workspace(c, norm, m, n, g1, b1), &w);
    if (rc) return rc;
    char *part = w, *sums = part + S * (size_t)outs * b1, *rest = sums + S * (size_t)outs;
    const Ops *o = c->ops[OP_DOT];
    // column (one-norm) or row (infinity-norm) sums of |a| into buffer (:39-139, :154-163),
    // then the maximum of buffer as mp_norm's infinity norm with (grid1, block1)
    HIPCHK(o->matsum(c->dev, cols, m, n, soa_of(A), lda, b1, part, strm(st)));
    HIPCHK(o->tree(c->dev, part, b1, next_pow2((unsigned)b1), outs, sums, strm(st)));
    HIPCHK(o->aos_to_soa(c->dev, sums, soa_of(buffer), outs, strm(st)));
    rc = maxabs_into(c, outs, soa_of(buffer), 1, g1, b1, rest, strm(st), &res);
    if (rc) return rc;
    HIPCHK(o->aos_to_soa(c->dev, res, soa_of(r), 1, strm(st)));
    return MPRES_OK;
}

int mpres_ger(mpres_ctx_t *c, int m, int n, const mp_array_t *alpha, const mp_array_t *x, int incx,
              const mp_array_t *y, int incy, mp_array_t *A, int lda, mp_array_t *buffer1, mp_array_t *buffer2,
              void *st) {
    if (!c || !alpha || !x || !y || !A || !buffer1 || !buffer2) return fail(MPRES_ERR_ARG, "null argument");
    if (m < 0 || n < 0 || lda < std::max(1, m) || incx == 0 || incy == 0)
        return fail(MPRES_ERR_ARG, "invalid size, increment or leading dimension");  // src/blas/ger.cuh:158-163
    if (m == 0 || n == 0) return MPRES_OK;                                            // :165-167
    const hipStream_t q = strm(st);
    V1(c, V1_MUL, SF_SCALAR, n, 1, alpha, 1, y, 0, incy, nullptr, 0, 0, buffer1, 0, 1, q);  // buffer1 = alpha y
    // buffer2 = x buffer1^T, m x n with leading dimension m (:190-199)
    V1(c, V1_MUL, SF_COL, m, n, buffer1, 1, x, 0, incx, nullptr, 0, 0, buffer2, m, 1, q);
    V1(c, V1_ADD, 0, m, n, nullptr, 0, A, lda, 1, buffer2, m, 1, A, lda, 1, q);  // A = A + buffer2
    return MPRES_OK;
}
#undef V1

// ---------------------------------------------------------------- the rest of the v2 BLAS
int mpres_scal_v2(mpres_ctx_t *c, int n, const void *alpha, const void *x, void *r, void *st) {
    if (!c || !alpha || !x || !r) return fail(MPRES_ERR_ARG, "null argument");
    if (n <= 0) return MPRES_OK;
    HIPCHK(c->ops[OP_ELEM]->v2map(c->dev, V2_SCAL, n, 1, 0, (const char *)alpha, (const char *)x, 1, nullptr, 1,
                                  nullptr, 0, (char *)r, strm(st)));
    return MPRES_OK;
}

int mpres_prod_v2(mpres_ctx_t *c, int n, const void *x, const void *y, void *r, void *st) {
    return mpres_elementwise(c, MPRES_OP_MUL, n, x, y, r, st);
}

int mpres_ger_v2(mpres_ctx_t *c, int m, int n, const void *alpha, const void *x, int incx, const void *y, int incy,
                 void *A, int lda, void *st) {
    if (!c || !alpha || !x || !y || !A) return fail(MPRES_ERR_ARG, "null argument");
    if (m <= 0 || n <= 0) return MPRES_OK;
    if (incx == 0 || incy == 0 || lda < std::max(1, m)) return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");
    HIPCHK(c->ops[OP_ELEM]->v2map(c->dev, V2_GER, m, n, 0, (const char *)alpha, (const char *)x, incx,
                                  (const char *)y, incy, (char *)A, lda, nullptr, strm(st)));
    return MPRES_OK;
}

static int uplo_flag(int uplo, int *upper) {
    if (uplo != MPRES_UPPER && uplo != MPRES_LOWER) return fail(MPRES_ERR_ARG, "invalid uplo %d", uplo);
    *upper = uplo == MPRES_UPPER;
    return MPRES_OK;
}

int mpres_syr_v2(mpres_ctx_t *c, int uplo, int n, const void *alpha, const void *x, int incx, void *A, int lda,
                 void *st) {
    if (!c || !alpha || !x || !A) return fail(MPRES_ERR_ARG, "null argument");
    int up, rc = uplo_flag(uplo, &up);
    if (rc) return rc;
    if (n <= 0) return MPRES_OK;
    if (incx == 0 || lda < std::max(1, n)) return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");
    HIPCHK(c->ops[OP_ELEM]->v2map(c->dev, V2_SYR, n, n, up, (const char *)alpha, (const char *)x, incx, nullptr, 1,
                                  (char *)A, lda, nullptr, strm(st)));
    return MPRES_OK;
}

int mpres_syr2_v2(mpres_ctx_t *c, int uplo, int n, const void *alpha, const void *x, int incx, const void *y, int incy,
                  void *A, int lda, void *st) {
    if (!c || !alpha || !x || !y || !A) return fail(MPRES_ERR_ARG, "null argument");
    int up, rc = uplo_flag(uplo, &up);
    if (rc) return rc;
    if (n <= 0) return MPRES_OK;
    if (incx == 0 || incy == 0 || lda < std::max(1, n))
        return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");
    HIPCHK(c->ops[OP_ELEM]->v2map(c->dev, V2_SYR2, n, n, up, (const char *)alpha, (const char *)x, incx,
                                  (const char *)y, incy, (char *)A, lda, nullptr, strm(st)));
    return MPRES_OK;
}

int mpres_symv_v2(mpres_ctx_t *c, int uplo, int n, const void *alpha, const void *A, int lda, const void *x, int incx,
                  const void *beta, void *y, int incy, void *st) {
    if (!c || !alpha || !A || !x || !beta || !y) return fail(MPRES_ERR_ARG, "null argument");
    int up, rc = uplo_flag(uplo, &up);
    if (rc) return rc;
    if (n <= 0) return MPRES_OK;
    if (incx == 0 || incy == 0 || lda < std::max(1, n))
        return fail(MPRES_ERR_ARG, "invalid increment or leading dimension");
    HIPCHK(c->ops[OP_ELEM]->symv_v2(c->dev, up, n, (const char *)alpha, (const char *)A, lda, (const char *)x, incx,
                                    (const char *)beta, (char *)y, incy, strm(st)));
    return MPRES_OK;
}

int mpres_syrk_v2(mpres_ctx_t *c, int uplo, int trans, int n, int k, const void *alpha, const void *A, int lda,
                  const void *beta, void *C, int ldc, void *st) {
    if (!c || !alpha || !A || !beta || !C) return fail(MPRES_ERR_ARG, "null argument");
    int up, rc = uplo_flag(uplo, &up);
    if (rc) return rc;
    if (trans != MPRES_NO_TRANS && trans != MPRES_TRANS && trans != MPRES_CONJ_TRANS)
        return fail(MPRES_ERR_ARG, "invalid trans %d", trans);
    if (n <= 0 || k < 0) return MPRES_OK;
    if (ldc < std::max(1, n) || lda < std::max(1, trans == MPRES_TRANS ? k : n))
        return fail(MPRES_ERR_ARG, "invalid leading dimension");
    // only mblas_trans reads A transposed (src/blas/v2/syrk_v2.cuh:69-73)
    HIPCHK(c->ops[OP_ELEM]->syrk_v2(c->dev, up, trans == MPRES_TRANS, n, k, (const char *)alpha, (const char *)A, lda,
                                    (const char *)beta, (char *)C, ldc, strm(st)));
    return MPRES_OK;
}

// ---------------------------------------------------------------- tuning
int mpres_lanes_supported(const mpres_ctx_t *c, int lanes) {
    return (c && find_ops(c->host.n, lanes) != nullptr) ? 1 : 0;
}

int mpres_set_lanes(mpres_ctx_t *c, int op, int lanes) {
    if (!c || !lanes_op(op)) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    const Ops *o = find_ops(c->host.n, lanes == 0 ? default_lanes(c->host.n, op) : lanes);
    if (!o) return fail(MPRES_ERR_UNSUPPORTED, "%d lanes per number not compiled for %d moduli", lanes, c->host.n);
    c->ops[op] = o;
    // a fixed choice replaces the size table; 0 restores the built-in one
    if (lanes == 0)
        install_builtin_lanes(c, op);
    else
        c->lt[op].count = 0;
    return MPRES_OK;
}

int mpres_get_lanes(const mpres_ctx_t *c, int op) {
    if (!c || !lanes_op(op)) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    return c->ops[op]->L;
}

int mpres_set_lanes_table(mpres_ctx_t *c, int op, int count, const long long *bounds, const int *lanes) {
    if (!c || !lanes_op(op)) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    if (count < 0 || count > MPRES_LANES_TABLE_MAX) return fail(MPRES_ERR_ARG, "table of %d entries", count);
    if (count > 0 && (!bounds || !lanes)) return fail(MPRES_ERR_ARG, "null argument");
    const Ops *ops[MPRES_LANES_TABLE_MAX];
    for (int i = 0; i < count; i++) {
        if (i > 0 && bounds[i] <= bounds[i - 1]) return fail(MPRES_ERR_ARG, "bounds must increase");
        ops[i] = find_ops(c->host.n, lanes[i]);
        if (!ops[i])
            return fail(MPRES_ERR_UNSUPPORTED, "%d lanes per number not compiled for %d moduli", lanes[i], c->host.n);
    }
    auto &t = c->lt[op];
    t.count = count;
    for (int i = 0; i < count; i++) {
        t.bound[i] = bounds[i];
        t.ops[i] = ops[i];
    }
    return MPRES_OK;
}

int mpres_get_lanes_for(const mpres_ctx_t *c, int op, long long size) {
    if (!c || !lanes_op(op)) return fail(MPRES_ERR_ARG, "invalid operation %d", op);
    return sized(c, op, size)->L;
}

}  // extern "C"
