// Per-moduli-set constants, computed on the host once per context and kept
// in device memory (one block per context, read through a kernel argument).
// These replace the reference's __constant__/__device__ symbols filled by
// rns_const_init / mp_const_init (src/rns.cuh:102-118, 324-442;
// src/arith/arith_utils.cuh:37-85). A context owns its constants, so several
// moduli sets can be live in one process.
#pragma once

#include <stdint.h>

namespace mph {

constexpr int MAXN = 32;
constexpr int P2T = 30;  // RNS_P2_SCALING_THRESHOLD (src/params.h)

struct ErC {
    double frac;
    int exp;
    int pad_;
};

struct DevConsts {
    int n;
    int log2m;        // RNS_MODULI_PRODUCT_LOG2: RNS_POW2 has log2m + 1 rows
    int mp_h;         // MP_H
    int mp_j;         // MP_J
    int ref_factor;   // RNS_EVAL_REF_FACTOR
    int mp_precision; // MP_PRECISION
    double acc;       // RNS_EVAL_ACCURACY
    ErC unit_low, unit_upp, inv_unit_low, inv_unit_upp;  // RNS_EVAL_UNIT / RNS_EVAL_INV_UNIT
    double ul_rh, ul_rl, uu_rh, uu_rl;  // two-double reciprocals of unit_low.frac, unit_upp.frac (recip2)
    int m[MAXN];           // moduli
    int w[MAXN];           // (M/m_i)^-1 mod m_i
    double md[MAXN];       // m_i as double
    double minv[MAXN];     // 1.0 / m_i rounded to nearest
    double rrd[MAXN];      // RNS_MODULI_RECIP_RD
    double rru[MAXN];      // RNS_MODULI_RECIP_RU
    int mpow2[P2T];        // M mod 2^(j+1)
    int h[P2T][MAXN];      // M_i mod 2^(j+1)
    int pow2inv[P2T][MAXN];// 2^-(j+1) mod m_i
    int mrcinv[MAXN][MAXN];// m_i^-1 mod m_j (j > i)
    const int *pow2;       // device pointer: RNS_POW2, (log2m + 1) x n, row-major
};

}  // namespace mph
