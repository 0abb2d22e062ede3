// CPU emulation of the library's reduction and gemm kernels for one lane per
// number (L = 1): the kernel templates of mpres-blas_amd/csrc/kernels.hpp are
// compiled with g++ against a host stand-in of the HIP header and every
// (block, thread) is run in turn. This checks the kernels' own logic (fast
// path, marking, fixup, epilogues) against the oracle on the CPU, where a
// mistake can be found with ordinary tools. Built by tests/test_kernels_emu.py.
#include <pthread.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../mpres-blas_amd/csrc/internal.hpp"
#include "../../mpres-blas_amd/csrc/kernels.hpp"

using namespace mph;

namespace {

struct Ctx {
    DevConsts c;
    std::vector<int> pow2;
};

bool make_ctx(int set_id, Ctx &x) {
    const ModuliSet *s = get_set(set_id);
    if (!s || !compute_consts(*s, x.c, x.pow2)) return false;
    x.c.pow2 = x.pow2.data();
    return true;
}

// run kernel body f over `groups` groups of one lane each, BLK threads per block
template <class F>
void run(long long groups, F &&f) {
    const long long blocks = (groups + BLK - 1) / BLK;
    for (long long b = 0; b < blocks; b++)
        for (unsigned t = 0; t < (unsigned)BLK; t++) {
            blockIdx = EmuIdx{(unsigned)b, 0, 0};
            threadIdx = EmuIdx{t, 0, 0};
            f();
        }
}

// run a kernel that synchronises its workgroup: the BLK threads of each block run as
// concurrent host threads with a barrier behind __syncthreads, blocks one after another
pthread_barrier_t g_bar;
void bar_wait() { pthread_barrier_wait(&g_bar); }
template <class F>
void run_wg(long long groups, F &&f) {
    const long long blocks = (groups + BLK - 1) / BLK;
    emu_barrier = bar_wait;
    for (long long b = 0; b < blocks; b++) {
        pthread_barrier_init(&g_bar, nullptr, BLK);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < (unsigned)BLK; t++)
            th.emplace_back([&, t, b] {
                blockIdx = EmuIdx{(unsigned)b, 0, 0};
                threadIdx = EmuIdx{t, 0, 0};
                f();
            });
        for (auto &t : th) t.join();
        pthread_barrier_destroy(&g_bar);
    }
    emu_barrier = nullptr;
}

// one workgroup (block b of nb, T threads) of a kernel that synchronises, its threads as
// concurrent host threads with a barrier behind __syncthreads
template <class F>
void run_block(unsigned b, unsigned nb, unsigned T, F &&f) {
    emu_barrier = bar_wait;
    pthread_barrier_init(&g_bar, nullptr, T);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            blockIdx = EmuIdx{b, 0, 0};
            threadIdx = EmuIdx{t, 0, 0};
            gridDim = EmuIdx{nb, 1, 1};
            f();
        });
    for (auto &t : th) t.join();
    pthread_barrier_destroy(&g_bar);
    emu_barrier = nullptr;
}

template <int N>
int gemm_v2_n(const Ctx &x, int ta, int tb, int m, int n, int k, char *alpha, char *A, int lda, char *B, int ldb,
              char *beta, char *Cm, int ldc, char *ws, int split) {
    const long long e = (long long)m * n;
    const AoSRef<N> a{A}, b{B}, w{ws};
    if (split) {
        run(e, [&] { k_gemm_sum<N, 1, false, false, AoSRef<N>>(&x.c, ta, tb, m, n, k, a, lda, b, ldb, w); });
        run(e, [&] { k_gemm_sum<N, 1, true, true, AoSRef<N>>(&x.c, ta, tb, m, n, k, a, lda, b, ldb, w); });
    } else {
        run(e, [&] { k_gemm_sum<N, 1, true, false, AoSRef<N>>(&x.c, ta, tb, m, n, k, a, lda, b, ldb, w); });
    }
    run(e, [&] {
        k_gemm_fin<N, 1, true, AoSRef<N>>(&x.c, m, n, AoSRef<N>{alpha}, AoSRef<N>{beta}, AoSRef<N>{Cm}, ldc, w);
    });
    return 0;
}

// split form: minima, pre-scaled planes and chains through the kernels; the
// residue products by exact 64-bit arithmetic, digits through resid_digit; then
// the fixup and epilogue kernels
template <int N>
int gemm_v2_split_n(const Ctx &x, int ta, int tb, int m, int n, int k, char *alpha, char *A, int lda, char *B, int ldb,
                    char *beta, char *Cm, int ldc, char *ws) {
    const long long mn = (long long)m * n;
    const AoSRef<N> a{A}, b{B}, w{ws};
    std::vector<int> EA(m), EB(n);
    // the exponent minima kernels reduce through LDS across a workgroup: restated here
    for (int row = 0; row < m; row++) {
        int e = 0x7fffffff;
        for (int i = 0; i < k; i++) e = std::min(e, *a.exp(op_index<0>(ta, row, i, lda)));
        EA[row] = k > 0 ? e : 0;
    }
    for (int col = 0; col < n; col++) {
        int e = 0x7fffffff;
        for (int i = 0; i < k; i++) e = std::min(e, *b.exp(op_index<1>(tb, col, i, ldb)));
        EB[col] = k > 0 ? e : 0;
    }
    // int8 limb planes through the kernel; the INT8 MFMA products restated in 64-bit integers
    int bits = 0;
    for (int r = 0; r < N; r++)
        while ((1LL << bits) <= x.c.m[r]) bits++;
    const int nl = bits <= 4 * LIMB_BITS ? 4 : 5;
    const int MP = (m + 63) / 64 * 64, NP = (n + 63) / 64 * 64, KP = ((k > 0 ? k : 1) + LIMB_T - 1) / LIMB_T * LIMB_T;
    std::vector<signed char> PA((size_t)N * nl * MP * KP, 99), PB((size_t)N * nl * NP * KP, 99);
    run((long long)MP * KP,
        [&] { k_gemm_limbs<N, 1, AoSRef<N>, 0>(&x.c, m, MP, k, KP, nl, ta, a, lda, EA.data(), PA.data()); });
    run((long long)NP * KP,
        [&] { k_gemm_limbs<N, 1, AoSRef<N>, 1>(&x.c, n, NP, k, KP, nl, tb, b, ldb, EB.data(), PB.data()); });
    run(mn, [&] { k_gemm_chain<N, 1, AoSRef<N>>(&x.c, ta, tb, m, n, k, a, lda, b, ldb, w); });
    auto val = [&](const std::vector<signed char> &P, int JP, int r, int j, int i) {
        unsigned long long v = 0;
        for (int q = 0; q < nl; q++) {
            // k-step-major plane: [i / LIMB_T][j][i % LIMB_T]
            const int limb = P[((size_t)r * nl + q) * JP * KP + ((size_t)(i / LIMB_T) * JP + j) * LIMB_T + i % LIMB_T];
            if (limb < 0) return ~0ULL;  // not a limb: the planes are not fully written
            v |= (unsigned long long)limb << (LIMB_BITS * q);
        }
        return v;
    };
    std::vector<int> S((size_t)N * mn);
    for (int r = 0; r < N; r++) {
        const unsigned long long mod = x.c.m[r];
        for (int col = 0; col < n; col++)
            for (int row = 0; row < m; row++) {
                unsigned long long acc = 0;
                for (int i = 0; i < KP; i++) {
                    const unsigned long long va = val(PA, MP, r, row, i), vb = val(PB, NP, r, col, i);
                    if (va >= mod || vb >= mod) return -2;  // padding must be zero, values residues
                    acc = (acc + va * vb % mod) % mod;
                }
                S[(size_t)r * mn + row + (size_t)m * col] = (int)acc;
            }
        for (int j = m; j < MP; j++)
            for (int i = 0; i < KP; i++)
                if (val(PA, MP, r, j, i) != 0) return -2;
    }
    run_wg(mn, [&] { k_gemm_digits<N, 1>(&x.c, m, n, S.data(), EA.data(), EB.data(), w); });
    run(mn, [&] { k_gemm_sum<N, 1, true, true, AoSRef<N>>(&x.c, ta, tb, m, n, k, a, lda, b, ldb, w); });
    run(mn, [&] {
        k_gemm_fin<N, 1, true, AoSRef<N>>(&x.c, m, n, AoSRef<N>{alpha}, AoSRef<N>{beta}, AoSRef<N>{Cm}, ldc, w);
    });
    return 0;
}

template <int N>
int dot_v2_pass1_n(const Ctx &x, long long cnt, long long nvt, char *X, char *Y, char *out, int split) {
    SoA none{};
    const AoSRef<N> xa{X}, ya{Y}, o{out};
    if (split) {
        run(nvt, [&] { k_fold<N, 1, 1, false, false>(&x.c, cnt, nvt, none, 1, none, 1, xa, ya, o); });
        run(nvt, [&] { k_fold<N, 1, 1, true, true>(&x.c, cnt, nvt, none, 1, none, 1, xa, ya, o); });
    } else {
        run(nvt, [&] { k_fold<N, 1, 1, true, false>(&x.c, cnt, nvt, none, 1, none, 1, xa, ya, o); });
    }
    return 0;
}

// CSR SpMV of the 16/32-moduli split form with one lane per row (the tree has no
// levels): scalar chains, residues, then the fused kernel's fixup of marked rows
template <int N>
int spmv_split_n(const Ctx &x, int m, const int *irp, const int *ja, const double *as, char *X, char *Y,
                 int *counts) {
    const SpMat A{as, ja, irp, nullptr, nullptr, 0, 0};
    const AoSRef<N> xa{X}, ya{Y};
    std::vector<int> ws(1 + 3 * SP_NB + 2 * (size_t)m, 0);
    int *rcnt = ws.data(), *hist = rcnt + 1, *off = hist + SP_NB, *cur = off + SP_NB;
    int *perm = cur + SP_NB, *rlist = perm + m;
    const int *nlong = off + SP_LONG_CLASS;
    const unsigned gm = (unsigned)((m + BLK * SP_RPT - 1) / (BLK * SP_RPT));
    for (unsigned b = 0; b < gm; b++) run_block(b, gm, BLK, [&] { k_sp_bin_count<N, 1>(m, irp, hist); });
    run(1, [&] { k_sp_bin_scan<N, 1>(hist, off); });
    for (unsigned b = 0; b < gm; b++) run_block(b, gm, BLK, [&] { k_sp_bin_scatter<N, 1>(m, irp, off, cur, perm); });
    run(m, [&] { k_spmv_chain<N, SP_CSR, SP_PF1>(&x.c, m, A, xa, ya, 1, perm, rcnt, rlist); });
    run((long long)m * (N / SGR) * SP_TS, [&] { k_spmv_resid<N, SP_CSR>(&x.c, m, A, xa, ya, perm, nlong); });
    // the long rows, two workgroups (fewer groups than rows: each takes several)
    for (unsigned b = 0; b < 2; b++) run_block(b, 2, BLK, [&] { k_spmv_resid_wave<N>(&x.c, A, xa, ya, perm, nlong); });
    // the fixup over the marked rows, two workgroups (each takes several list entries)
    for (unsigned b = 0; b < 2; b++) run_block(b, 2, BLK, [&] { k_spmv<N, 1, true, true, SP_CSR>(&x.c, m, A, xa, ya, 1, rcnt, rlist); });
    counts[0] = *nlong;
    counts[1] = *rcnt;
    return 0;
}

// the v1 (SoA) chain kernel and the generic AoS chain kernel over the same operands (A m x k,
// B k x n, column-major): W1 from k_gemm_chain, W2 from k_gemm_chain_nn (UB or not); the
// sums' sign, exponent and evaluation, or REDO
template <int N>
int chain_nn_n(const Ctx &x, int m, int n, int k, char *Ab, char *Bb, char *W1, char *W2, int ub) {
    const AoSRef<N> a{Ab}, b{Bb}, w1{W1}, w2{W2};
    struct Soa {
        std::vector<int> d, s, e;
        std::vector<ErMem> ev;
        int len;
        SoA view() { return SoA{d.data(), s.data(), e.data(), ev.data(), &len, len}; }
    };
    auto soa = [&](const AoSRef<N> &v, int cnt, Soa &o) {
        o.len = cnt;
        o.d.resize((size_t)cnt * N);
        o.s.resize(cnt);
        o.e.resize(cnt);
        o.ev.resize(2 * (size_t)cnt);
        for (int i = 0; i < cnt; i++) {
            for (int r = 0; r < N; r++) o.d[(size_t)i * N + r] = v.digits(i)[r];
            o.s[i] = *v.sign(i);
            o.e[i] = *v.exp(i);
            o.ev[i] = v.eval(i)[0];
            o.ev[(size_t)cnt + i] = v.eval(i)[1];
        }
    };
    Soa sa, sb;
    soa(a, m * k, sa);
    soa(b, k * n, sb);
    const SoA SA = sa.view(), SB = sb.view();
    run((long long)m * n, [&] { k_gemm_chain<N, 1, AoSRef<N>>(&x.c, 0, 0, m, n, k, a, m, b, k, w1); });
    if (ub)
        run((long long)m * n, [&] { k_gemm_chain_nn<N, 1, true>(&x.c, m, n, k, SA, m, SB, k, w2); });
    else
        run((long long)m * n, [&] { k_gemm_chain_nn<N, 1, false>(&x.c, m, n, k, SA, m, SB, k, w2); });
    return 0;
}

}  // namespace

extern "C" {

int emu_gemm_chain_nn(int set_id, int m, int n, int k, void *A, void *B, void *W1, void *W2, int ub) {
    Ctx x;
    if (!make_ctx(set_id, x)) return -1;
    switch (x.c.n) {
        case 16: return chain_nn_n<16>(x, m, n, k, (char *)A, (char *)B, (char *)W1, (char *)W2, ub);
        case 32: return chain_nn_n<32>(x, m, n, k, (char *)A, (char *)B, (char *)W1, (char *)W2, ub);
    }
    return -1;
}

// counts: the number of long rows (2^SP_LONG_BITS terms or more) and the marked-row list length
int emu_spmv_split(int set_id, int m, const int *irp, const int *ja, const double *as, void *X, void *Y, int *counts) {
    Ctx x;
    if (!make_ctx(set_id, x)) return -1;
    switch (x.c.n) {
        case 16: return spmv_split_n<16>(x, m, irp, ja, as, (char *)X, (char *)Y, counts);
        case 32: return spmv_split_n<32>(x, m, irp, ja, as, (char *)X, (char *)Y, counts);
    }
    return -1;
}


// v2 gemm (AoS) on the CPU; ws holds m*n mp_float_t; split = 1: fast kernel + fixup epilogue
int emu_gemm_v2(int set_id, int ta, int tb, int m, int n, int k, void *alpha, void *A, int lda, void *B, int ldb,
                void *beta, void *Cm, int ldc, void *ws, int split) {
    Ctx x;
    if (!make_ctx(set_id, x)) return -1;
    switch (x.c.n) {
        case 8: return gemm_v2_n<8>(x, ta, tb, m, n, k, (char *)alpha, (char *)A, lda, (char *)B, ldb, (char *)beta, (char *)Cm, ldc, (char *)ws, split);
        case 16: return gemm_v2_n<16>(x, ta, tb, m, n, k, (char *)alpha, (char *)A, lda, (char *)B, ldb, (char *)beta, (char *)Cm, ldc, (char *)ws, split);
        case 32: return gemm_v2_n<32>(x, ta, tb, m, n, k, (char *)alpha, (char *)A, lda, (char *)B, ldb, (char *)beta, (char *)Cm, ldc, (char *)ws, split);
    }
    return -1;
}

// v2 gemm in the split form (16 and 32 moduli); ws holds m*n mp_float_t
int emu_gemm_v2_split(int set_id, int ta, int tb, int m, int n, int k, void *alpha, void *A, int lda, void *B, int ldb,
                      void *beta, void *Cm, int ldc, void *ws) {
    Ctx x;
    if (!make_ctx(set_id, x)) return -1;
    switch (x.c.n) {
        case 16: return gemm_v2_split_n<16>(x, ta, tb, m, n, k, (char *)alpha, (char *)A, lda, (char *)B, ldb, (char *)beta, (char *)Cm, ldc, (char *)ws);
        case 32: return gemm_v2_split_n<32>(x, ta, tb, m, n, k, (char *)alpha, (char *)A, lda, (char *)B, ldb, (char *)beta, (char *)Cm, ldc, (char *)ws);
    }
    return -1;
}

// first pass of the v2 dot: the nvt strided partial sums of x_i * y_i
int emu_dot_v2_pass1(int set_id, long long cnt, long long nvt, void *X, void *Y, void *out, int split) {
    Ctx x;
    if (!make_ctx(set_id, x)) return -1;
    switch (x.c.n) {
        case 8: return dot_v2_pass1_n<8>(x, cnt, nvt, (char *)X, (char *)Y, (char *)out, split);
        case 16: return dot_v2_pass1_n<16>(x, cnt, nvt, (char *)X, (char *)Y, (char *)out, split);
        case 32: return dot_v2_pass1_n<32>(x, cnt, nvt, (char *)X, (char *)Y, (char *)out, split);
    }
    return -1;
}

}  // extern "C"
