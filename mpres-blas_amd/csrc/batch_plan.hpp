// Host side of the batched form: checks the problems as their single calls do,
// splits them into work items and lays out the items' scratch. Pure C++ (no HIP
// calls), so the host tests compile it as it is.
#pragma once

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mpres_hip.h"
#include "ops.hpp"

namespace mph {

inline SoA soa_view(const mp_array_t *a) {
    SoA s;
    s.d = a->digits;
    s.s = a->sign;
    s.e = a->exp;
    s.ev = (ErMem *)a->eval;
    s.lenp = a->len;
    s.len = 0;
    return s;
}

struct BatchLayout {
    std::vector<BatchItem> items;  // in launch order (descending work)
    size_t ws_bytes = 0;
    double fmas = 0;               // multiply-accumulates of the batch
    std::string err;
};

constexpr size_t BATCH_WS_ALIGN = 256;

// Returns MPRES_OK, or MPRES_ERR_ARG with layout.err naming the first bad problem.
// Problems the single call would skip (dot n <= 0, gemv m or n <= 0) give no item.
inline int batch_layout(const mpres_batch_problem_t *probs, int count, int n_moduli, int gemv_rows,
                        BatchLayout &out) {
    out = BatchLayout{};
    const long long S = 4LL * n_moduli + 40;
    char buf[160];
    auto bad = [&](int i, const char *what) {
        std::snprintf(buf, sizeof buf, "batch problem %d: %s", i, what);
        out.err = buf;
        return MPRES_ERR_ARG;
    };
    if (count < 0 || (count > 0 && !probs)) {
        out.err = "invalid problem list";
        return MPRES_ERR_ARG;
    }
    if (gemv_rows <= 0) {
        out.err = "gemv rows per item must be positive";
        return MPRES_ERR_ARG;
    }
    struct Work {
        BatchItem it;
        double work;
        size_t bytes;
    };
    std::vector<Work> v;
    for (int i = 0; i < count; i++) {
        const mpres_batch_problem_t &p = probs[i];
        if (p.op == MPRES_BATCH_DOT) {
            // mpres_dot: src/blas/dot.cuh:84-107
            if (!p.x || !p.y || !p.r) return bad(i, "null argument");
            if (p.n <= 0) continue;
            if (p.grid3 <= 0 || p.block3 <= 0) return bad(i, "invalid reduction shape");
            Work w{};
            w.it.kind = BATCH_DOT;
            w.it.n = p.n;
            w.it.incx = p.incx;
            w.it.incy = p.incy;
            w.it.g3 = p.grid3;
            w.it.b3 = p.block3;
            w.it.x = soa_view(p.x);
            w.it.y = soa_view(p.y);
            w.it.r = soa_view(p.r);
            const long long G = (long long)p.grid3 * p.block3;
            w.bytes = (size_t)((G + p.block3) * S);
            w.work = (double)p.n + (double)G + (double)p.block3;
            out.fmas += p.n;
            v.push_back(w);
        } else if (p.op == MPRES_BATCH_GEMV) {
            // mpres_gemv: src/blas/gemv.cuh:150-268
            if (!p.alpha || !p.A || !p.x || !p.beta || !p.y) return bad(i, "null argument");
            if (p.m <= 0 || p.n <= 0) continue;
            if (p.incx == 0 || p.incy == 0 || p.lda < std::max(1, p.m))
                return bad(i, "invalid increments or leading dimension");
            if (p.block3 <= 0) return bad(i, "invalid blockDim3");
            if (p.trans != MPRES_NO_TRANS && p.trans != MPRES_TRANS && p.trans != MPRES_CONJ_TRANS)
                return bad(i, "invalid trans");
            const bool tr = p.trans != MPRES_NO_TRANS;
            const int lx = tr ? p.m : p.n, ly = tr ? p.n : p.m;
            out.fmas += (double)p.m * p.n;
            for (int r0 = 0; r0 < ly; r0 += gemv_rows) {
                const int r1 = std::min(ly, r0 + gemv_rows), h = r1 - r0;
                Work w{};
                w.it.kind = tr ? BATCH_GEMV_T : BATCH_GEMV_N;
                w.it.r0 = r0;
                w.it.r1 = r1;
                w.it.m = p.m;
                w.it.n = p.n;
                w.it.lda = p.lda;
                w.it.incx = p.incx;
                w.it.incy = p.incy;
                w.it.b3 = p.block3;
                w.it.alpha = soa_view(p.alpha);
                w.it.A = soa_view(p.A);
                w.it.x = soa_view(p.x);
                w.it.beta = soa_view(p.beta);
                w.it.y = soa_view(p.y);
                w.bytes = (size_t)(((long long)lx + (long long)h * p.block3) * S);
                w.work = (double)h * lx + (double)lx + (double)h * p.block3;
                v.push_back(w);
            }
        } else {
            return bad(i, "unknown operation");
        }
    }
    std::stable_sort(v.begin(), v.end(), [](const Work &a, const Work &b) { return a.work > b.work; });
    size_t off = 0;
    out.items.reserve(v.size());
    for (auto &w : v) {
        w.it.ws = (long long)off;
        off += (w.bytes + BATCH_WS_ALIGN - 1) / BATCH_WS_ALIGN * BATCH_WS_ALIGN;
        out.items.push_back(w.it);
    }
    out.ws_bytes = off;
    return MPRES_OK;
}

}  // namespace mph
