/*
 * Oracle internals: context layout and shared scalar helpers.
 * TEST INFRASTRUCTURE ONLY (see mpres_oracle.h).
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include "mpres_oracle.h"

/* DBL_PHI1 = 2^-53 (1 + 2^-52) and DBL_ETA = 2^-1074, src/bitwise.cuh:51-61 */
#define ORC_PHI1 (ldexp(1.0, -53) * (1 + 2 * ldexp(1.0, -53)))
#define ORC_ETA (ldexp(1.0, -1074))
/* the same two values as literals: 2^-53 + 2^-105 and 2^-1074 */
#define ORC_PHI1_C (0x1.0000000000001p-53)
#define ORC_ETA_C (0x1p-1074)

struct orc_ctx {
    int n;
    int log2m;          /* RNS_MODULI_PRODUCT_LOG2: RNS_POW2 has log2m+1 rows */
    int log2m_floor;    /* floor(log2 M) */
    int m[32];
    int w[32];          /* RNS_PART_MODULI_PRODUCT_INVERSE */
    int *pow2;          /* RNS_POW2, (log2m+1) x n */
    int mpow2[30];      /* RNS_MODULI_PRODUCT_POW2_RESIDUES */
    int h[30][32];      /* RNS_PART_MODULI_PRODUCT_POW2_RESIDUES */
    int pow2inv[30][32];/* RNS_POW2_INVERSE */
    double recip_rd[32], recip_ru[32];
    double eval_acc;
    int ref_factor;
    orc_er_t unit_low, unit_upp, inv_unit_low, inv_unit_upp;
    int mrc_inv[32][32];
    int mp_precision, mp_h, mp_j;
};

union orc_du { double d; uint64_t u; };

#endif
