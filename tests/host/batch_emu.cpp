// CPU emulation of the batched form (mpres_batch_run) for one lane per number
// (L = 1): the item layout of mpres-blas_amd/csrc/batch_plan.hpp and the
// persistent kernel of mpres-blas_amd/csrc/kernels_batch.hpp, compiled with g++
// against the host stand-in of the HIP header. The threads of one workgroup run
// as concurrent host threads with a real barrier behind __syncthreads; the
// workgroups run one after another and share the queue counter, as on the GPU.
// Operands are mp_array_t whose pointers are host memory. Built by
// tests/test_batch_cpu.py.
#include <pthread.h>

#include <thread>
#include <vector>

#include "../../mpres-blas_amd/csrc/batch_plan.hpp"
#include "../../mpres-blas_amd/csrc/internal.hpp"
#include "../../mpres-blas_amd/csrc/kernels_batch.hpp"

using namespace mph;

namespace {

pthread_barrier_t g_bar;
void bar_wait() { pthread_barrier_wait(&g_bar); }

template <int N>
int batch_n(const DevConsts *c, const mpres_batch_problem_t *probs, int count, int rows, int wgs, char *ws,
            size_t wsb) {
    BatchLayout lay;
    const int rc = batch_layout(probs, count, N, rows, lay);
    if (rc) return rc;
    if (lay.ws_bytes > wsb) return MPRES_ERR_WORKSPACE;
    int queue = 0;
    emu_barrier = bar_wait;
    for (int b = 0; b < wgs; b++) {
        pthread_barrier_init(&g_bar, nullptr, BLK);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < (unsigned)BLK; t++)
            th.emplace_back([&, t, b] {
                blockIdx = EmuIdx{(unsigned)b, 0, 0};
                threadIdx = EmuIdx{t, 0, 0};
                k_batch<N, 1>(c, lay.items.data(), (int)lay.items.size(), &queue, ws);
            });
        for (auto &t : th) t.join();
        pthread_barrier_destroy(&g_bar);
    }
    emu_barrier = nullptr;
    return MPRES_OK;
}

}  // namespace

extern "C" {

// scratch bytes of the batch, or a negative status
long long emu_batch_workspace(int set_id, const mpres_batch_problem_t *probs, int count, int rows) {
    const ModuliSet *s = get_set(set_id);
    if (!s) return -1;
    BatchLayout lay;
    const int rc = batch_layout(probs, count, s->n, rows, lay);
    return rc ? rc : (long long)lay.ws_bytes;
}

// item kinds, row ranges and scratch offsets in launch order (up to cap items); returns the item count
int emu_batch_items(int set_id, const mpres_batch_problem_t *probs, int count, int rows, int *kind, int *r0, int *r1,
                    long long *off, int cap) {
    const ModuliSet *s = get_set(set_id);
    if (!s) return -1;
    BatchLayout lay;
    const int rc = batch_layout(probs, count, s->n, rows, lay);
    if (rc) return rc;
    for (int i = 0; i < (int)lay.items.size() && i < cap; i++) {
        kind[i] = lay.items[i].kind;
        r0[i] = lay.items[i].r0;
        r1[i] = lay.items[i].r1;
        off[i] = lay.items[i].ws;
    }
    return (int)lay.items.size();
}

int emu_batch_run(int set_id, const mpres_batch_problem_t *probs, int count, int rows, int wgs, void *ws,
                  size_t wsb) {
    const ModuliSet *s = get_set(set_id);
    if (!s) return -1;
    DevConsts c;
    std::vector<int> pow2;
    if (!compute_consts(*s, c, pow2)) return -1;
    c.pow2 = pow2.data();
    switch (s->n) {
        case 8: return batch_n<8>(&c, probs, count, rows, wgs, (char *)ws, wsb);
        case 16: return batch_n<16>(&c, probs, count, rows, wgs, (char *)ws, wsb);
        case 32: return batch_n<32>(&c, probs, count, rows, wgs, (char *)ws, wsb);
    }
    return -1;
}

}  // extern "C"
