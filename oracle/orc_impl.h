/*
 * Oracle implementation body, included once per (ORC_N, ORC_POL) pair by
 * orc_api.c. TEST INFRASTRUCTURE ONLY (see mpres_oracle.h).
 *
 * ORC_POL 0 = policy E (reference CPU code), 1 = policy D (reference CUDA code).
 * Every function names the reference lines it restates.
 *
 * Pinning (what fixes each function independently of this restatement,
 * tests/test_oracle_*.py):
 *   constants            - exact big-integer / rational recomputation in Python
 *   dadd/dmul/ddiv D     - exact rational rounding (fractions) on random and edge inputs
 *   dadd/dmul/ddiv E     - bit for bit against the definition RN(c) -/+ (phi|c| + eta) with phi, eta
 *                          derived from src/bitwise.cuh:51-61, on edge inputs; enclosure
 *   psum                 - tree shape: equals the balanced pairwise tree computed in
 *                          Python with the pinned directed adds
 *   mod ops, mrc, mrd    - brute force with Python integers (CRT)
 *   rank                 - exact rank floor(sum s_i/m_i) with Python fractions
 *   scale2pow            - closed form floor(X / 2^D) by CRT
 *   eval (full/fast)     - bits recomposed from exact directed rounding and the CRT on inputs
 *                          forcing every branch; fast == full without ambiguity; the width
 *                          bound of the sums; enclosure low <= X/M <= upp; the reference's
 *                          walk-down sweep from M-1 (tests/arith/test_rns_eval_verify.cu)
 *   set_d                - closed form: X = trimmed significand, exp, sign
 *   mul / mul_d          - closed form X*Y (no wrap) and exponent / sign rules
 *   add / sub            - closed form of the aligned residue sum and the exponent /
 *                          zeroing rules; the reference's verify operands
 *                          (tests/arith/test_mp_add_verify.cu:47-48, test_mp_mul_verify.cu:42-43)
 *   round                - closed form floor(X/2^bits), exp += bits
 *   compositions         - exact dyadic result within the a-priori bound of the
 *                          reference's accuracy tests; E and D agree on sign/exp/digits
 */

#define ORC_CAT_(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT_(a, b)
#if ORC_POL == 0
#define ORC_SFX ORC_CAT(_e, ORC_N)
#else
#define ORC_SFX ORC_CAT(_d, ORC_N)
#endif
#define F(x) ORC_CAT(x, ORC_SFX)
#define N ORC_N
#define MPF F(mpf)

typedef struct {
    int digits[N];
    int sign;
    int exp;
    orc_er_t eval[2];
} MPF;

/* ------------------------------------------------------------------ */
/* Directed rounding                                                  */
/* ------------------------------------------------------------------ */
#if ORC_POL == 0
/* src/dinterval.cuh:43-165: RN result widened by phi*|c| + eta, fmac unfused */
static inline double F(dadd_rd)(double a, double b) { double c = a + b; return c - (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(dadd_ru)(double a, double b) { double c = a + b; return c + (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(dsub_rd)(double a, double b) { double c = a - b; return c - (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(dsub_ru)(double a, double b) { double c = a - b; return c + (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(dmul_rd)(double a, double b) { double c = a * b; return c - (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(dmul_ru)(double a, double b) { double c = a * b; return c + (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(ddiv_rd)(double a, double b) { double c = a / b; return c - (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
static inline double F(ddiv_ru)(double a, double b) { double c = a / b; return c + (ORC_PHI1_C * fabs(c) + ORC_ETA_C); }
#else
/* IEEE-754 directed rounding (__dadd_rd, __dmul_ru, ... of the CUDA path) */
#define ORC_DOP(name, mode, expr)                                        \
    static inline double F(name)(double a, double b) {                   \
        volatile double va = a, vb = b;                                  \
        fesetround(mode);                                                \
        volatile double r = (expr);                                      \
        fesetround(FE_TONEAREST);                                        \
        return r;                                                        \
    }
ORC_DOP(dadd_rd, FE_DOWNWARD, va + vb)
ORC_DOP(dadd_ru, FE_UPWARD, va + vb)
ORC_DOP(dsub_rd, FE_DOWNWARD, va - vb)
ORC_DOP(dsub_ru, FE_UPWARD, va - vb)
ORC_DOP(dmul_rd, FE_DOWNWARD, va * vb)
ORC_DOP(dmul_ru, FE_UPWARD, va * vb)
ORC_DOP(ddiv_rd, FE_DOWNWARD, va / vb)
ORC_DOP(ddiv_ru, FE_UPWARD, va / vb)
#undef ORC_DOP
#endif

/* Pairwise sums of N (a power of two) values: the balanced tree of
 * psum<N> followed by the root addition of 0 that psum(2N)<N> performs
 * (CPU src/pairwise.cuh:402-479, CUDA :855-931). Under policy E that
 * root addition widens the bound; under IEEE rounding it is exact. */
static double F(tree_rd)(const double *x, int k) {
    if (k == 2) return F(dadd_rd)(x[0], x[1]);
    return F(dadd_rd)(F(tree_rd)(x, k / 2), F(tree_rd)(x + k / 2, k / 2));
}
static double F(tree_ru)(const double *x, int k) {
    if (k == 2) return F(dadd_ru)(x[0], x[1]);
    return F(dadd_ru)(F(tree_ru)(x, k / 2), F(tree_ru)(x + k / 2, k / 2));
}
static inline double F(psum_rd)(const double *x) { return F(dadd_rd)(F(tree_rd)(x, N), 0); }
static inline double F(psum_ru)(const double *x) { return F(dadd_ru)(F(tree_ru)(x, N), 0); }

/* ------------------------------------------------------------------ */
/* Extended-range floats, src/extrange.cuh                             */
/* ------------------------------------------------------------------ */
static inline double F(fast_scalbn)(double x, long n) { /* src/bitwise.cuh:92-97 */
    union orc_du u;
    u.d = x;
    u.u += (uint64_t)n << 52;
    return u.d * (x != 0 && n >= -1023);
}
static inline double F(scal)(double x, long n) {
#if ORC_POL == 0
    return F(fast_scalbn)(x, n);   /* CPU er_add, src/extrange.cuh:158 */
#else
    return scalbn(x, (int)n);      /* CUDA er_add, src/extrange.cuh:436 */
#endif
}
static inline void F(er_adjust)(orc_er_t *x) { /* src/extrange.cuh:135-144 (CUDA :412-421 identical) */
    union orc_du u;
    u.d = x->frac;
    x->exp = (x->exp + (long)((u.u & ~((uint64_t)1 << 63)) >> 52) - 1023) * (x->frac != 0);
    u.u = (u.u & ((((uint64_t)1 << 52) - 1) | ((uint64_t)1 << 63))) | ((uint64_t)1023 << 52);
    x->frac = u.d * (x->frac != 0);
}
static inline void F(er_set_d)(orc_er_t *r, double x) { /* src/extrange.cuh:46-58 */
    union orc_du u;
    u.d = x;
    if (x != 0) {
        r->exp = (long)((u.u & ~((uint64_t)1 << 63)) >> 52) - 1023;
        u.u = (u.u & ((((uint64_t)1 << 52) - 1) | ((uint64_t)1 << 63))) | ((uint64_t)1023 << 52);
        r->frac = u.d;
    } else {
        r->exp = 0;
        r->frac = 0;
    }
}
/* er_add_rd / er_add_ru, src/extrange.cuh:169-209 (CUDA :450-490) */
static inline orc_er_t F(er_add_rd)(orc_er_t x, orc_er_t y) {
    x.exp *= (x.frac != 0);
    y.exp *= (y.frac != 0);
    long dexp = (x.exp - y.exp) * (x.frac != 0) * (y.frac != 0);
    orc_er_t r;
    if (dexp > 0) {
        r.exp = x.exp;
        r.frac = F(dadd_rd)(x.frac, F(scal)(y.frac, -dexp));
    } else if (dexp < 0) {
        r.exp = y.exp;
        r.frac = F(dadd_rd)(y.frac, F(scal)(x.frac, dexp));
    } else {
        r.exp = (x.exp == 0) ? y.exp : x.exp;
        r.frac = F(dadd_rd)(x.frac, y.frac);
    }
    F(er_adjust)(&r);
    return r;
}
static inline orc_er_t F(er_add_ru)(orc_er_t x, orc_er_t y) {
    x.exp *= (x.frac != 0);
    y.exp *= (y.frac != 0);
    long dexp = (x.exp - y.exp) * (x.frac != 0) * (y.frac != 0);
    orc_er_t r;
    if (dexp > 0) {
        r.exp = x.exp;
        r.frac = F(dadd_ru)(x.frac, F(scal)(y.frac, -dexp));
    } else if (dexp < 0) {
        r.exp = y.exp;
        r.frac = F(dadd_ru)(y.frac, F(scal)(x.frac, dexp));
    } else {
        r.exp = (x.exp == 0) ? y.exp : x.exp;
        r.frac = F(dadd_ru)(x.frac, y.frac);
    }
    F(er_adjust)(&r);
    return r;
}
/* er_md_rd / er_md_ru: x*y/z, src/extrange.cuh:331-348 (CUDA :627-644) */
static inline orc_er_t F(er_md_rd)(orc_er_t x, orc_er_t y, orc_er_t z) {
    orc_er_t r;
    r.exp = x.exp + y.exp - z.exp;
    r.frac = F(ddiv_rd)(F(dmul_rd)(x.frac, y.frac), z.frac);
    F(er_adjust)(&r);
    return r;
}
static inline orc_er_t F(er_md_ru)(orc_er_t x, orc_er_t y, orc_er_t z) {
    orc_er_t r;
    r.exp = x.exp + y.exp - z.exp;
    r.frac = F(ddiv_ru)(F(dmul_ru)(x.frac, y.frac), z.frac);
    F(er_adjust)(&r);
    return r;
}
/* er_ucmp, src/extrange.cuh:393-399 (CUDA :671-677) */
static inline int F(er_ucmp)(orc_er_t x, orc_er_t y) {
    long diff = x.exp - y.exp;
    if (diff == 0) return x.frac == y.frac ? 0 : (x.frac > y.frac ? 1 : -1);
    return x.frac != 0 && (diff > 0 || (diff < 0 && y.frac == 0)) ? 1 : -1;
}

/* ------------------------------------------------------------------ */
/* Modular arithmetic, src/modular.cuh                                 */
/* ------------------------------------------------------------------ */
static inline int F(mod_mul_exact)(int x, int y, int m) { /* :78-82 / :150-154 */
    long r = (long)x * (long)y;
    return (int)(r % (long)m);
}
static inline int F(mod_psub)(int x, int y, int m) { /* :68-72 / :140-144 */
    long r = (long)x - (long)y + (long)m;
    return (int)(r % (long)m);
}
#if ORC_POL == 0
static inline int F(mod_mul_q)(int x, int y, int m, double d) { /* :90-95, double-reciprocal quotient */
    long r = (long)x * (long)y;
    double q = (double)r * d;
    int i = (int)q;
    return (int)(r - (long)i * (long)m);
}
#endif
/* rns_mul: CPU uses the quotient form (:218-222), CUDA the exact form (:250-256) */
static inline void F(rns_mul)(const orc_ctx *c, int *r, const int *x, const int *y) {
    for (int i = 0; i < N; i++) {
#if ORC_POL == 0
        r[i] = F(mod_mul_q)(x[i], y[i], c->m[i], 1.0 / c->m[i]);
#else
        r[i] = F(mod_mul_exact)(x[i], y[i], c->m[i]);
#endif
    }
}
static inline int F(check_zero)(const int *x) {
    for (int i = 0; i < N; i++)
        if (x[i] != 0) return 0;
    return 1;
}

/* ------------------------------------------------------------------ */
/* Mixed-radix conversion, src/rns.cuh:495-557 (CUDA :570-630)         */
/* ------------------------------------------------------------------ */
static void F(mrc)(const orc_ctx *c, int *mr, const int *x) {
    for (int i = 0; i < N; i++) {
        mr[i] = x[i];
        for (int j = 0; j < i; j++) {
            if (mr[i] < mr[j])
                mr[i] = (int)((long)c->m[i] - (long)mr[j] + (long)mr[i]);
            else
                mr[i] = mr[i] - mr[j];
            mr[i] = F(mod_mul_exact)(mr[i], c->mrc_inv[j][i], c->m[i]);
        }
    }
}
static inline int F(mrd)(const orc_ctx *c, const int *x) {
    int mr[N];
    F(mrc)(c, mr, x);
    return mr[N - 1];
}
static int F(mrc_compare)(const orc_ctx *c, const int *x, const int *y) {
    int mx[N], my[N];
    F(mrc)(c, mx, x);
    F(mrc)(c, my, y);
    for (int i = N - 1; i >= 0; i--) {
        if (mx[i] > my[i]) return 1;
        if (my[i] > mx[i]) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Interval evaluation, src/rns.cuh:645-781 (CUDA :797-933)            */
/* ------------------------------------------------------------------ */
static inline int F(refine_k)(const orc_ctx *c, double sumu) {
    double t = -(ceil(log2(sumu)) + 1);
    return (int)(t > c->ref_factor ? t : c->ref_factor);
}
static void F(eval_compute)(const orc_ctx *c, orc_er_t *low, orc_er_t *upp, const int *x, int fast) {
    int s[N];
    double fl[N], fu[N];
    double suml, sumu;
#if ORC_POL == 0
    if (F(check_zero)(x)) {  /* :652 */
        low->frac = upp->frac = 0;
        low->exp = upp->exp = 0;
        return;
    }
#endif
    F(rns_mul)(c, s, x, c->w);
    for (int i = 0; i < N; i++) {
        fl[i] = F(dmul_rd)((double)s[i], c->recip_rd[i]);
        fu[i] = F(dmul_ru)((double)s[i], c->recip_ru[i]);
    }
    suml = F(psum_rd)(fl);
    sumu = F(psum_ru)(fu);
#if ORC_POL == 1
    if (suml == 0 && sumu == 0) {  /* :814 */
        low->frac = upp->frac = 0;
        low->exp = upp->exp = 0;
        return;
    }
#endif
    unsigned whl = (unsigned)suml, whu = (unsigned)sumu;
#if ORC_POL == 0
    suml = suml - whl;
    sumu = sumu - whu;
#else
    suml = F(dsub_rd)(suml, (double)whl);
    sumu = F(dsub_ru)(sumu, (double)whu);
#endif
    if (fast) {  /* :734-741 / :895-903 */
        if (sumu >= c->eval_acc) {
            F(er_set_d)(low, suml);
            F(er_set_d)(upp, sumu);
            return;
        }
    } else {  /* :672-692 / :824-845 */
        F(er_set_d)(low, suml);
        F(er_set_d)(upp, sumu);
        int mr = -1;
        if (whl != whu) mr = F(mrd)(c, x);
        if (mr > 0) {
            *upp = c->inv_unit_upp;
            return;
        }
        if (mr == 0) *low = c->unit_low;
        if (sumu >= c->eval_acc) return;
    }
    /* refinement loop */
    int K = 0;
    while (sumu < c->eval_acc) {
        int k = F(refine_k)(c, sumu);
        F(rns_mul)(c, s, s, &c->pow2[k * N]);
        for (int i = 0; i < N; i++) {
#if ORC_POL == 0
            fu[i] = F(ddiv_ru)((double)s[i], (double)c->m[i]);
#else
            fu[i] = F(dmul_ru)((double)s[i], c->recip_ru[i]);
#endif
        }
        sumu = F(psum_ru)(fu);
#if ORC_POL == 0
        sumu -= (unsigned)sumu;
#else
        sumu = F(dsub_ru)(sumu, (double)(unsigned)sumu);
#endif
        K += k;
    }
    for (int i = 0; i < N; i++) {
#if ORC_POL == 0
        fl[i] = F(ddiv_rd)((double)s[i], (double)c->m[i]);
#else
        fl[i] = F(dmul_rd)((double)s[i], c->recip_rd[i]);
#endif
    }
    suml = F(psum_rd)(fl);
#if ORC_POL == 0
    suml -= (unsigned)suml;
#else
    suml = F(dsub_rd)(suml, (double)(unsigned)suml);
#endif
    F(er_set_d)(low, suml);
    F(er_set_d)(upp, sumu);
    low->exp -= K;
    upp->exp -= K;
}

/* ------------------------------------------------------------------ */
/* Power-of-two scaling, src/rns.cuh:946-1048 (CUDA :1061-1160)        */
/* ------------------------------------------------------------------ */
static int F(rank_full)(const orc_ctx *c, const int *x, const int *s) {
    double fl[N], fu[N];
    for (int i = 0; i < N; i++) {
        fl[i] = F(dmul_rd)((double)s[i], c->recip_rd[i]);
        fu[i] = F(dmul_ru)((double)s[i], c->recip_ru[i]);
    }
    double suml = F(psum_rd)(fl), sumu = F(psum_ru)(fu);
    unsigned whl = (unsigned)suml, whu = (unsigned)sumu;
    if (whl == whu) return (int)whl;
    return F(mrd)(c, x) == 0 ? (int)whu : (int)whl;
}
static int F(rank_fast)(const orc_ctx *c, const int *s) {
    double fu[N];
    for (int i = 0; i < N; i++) fu[i] = F(dmul_ru)((double)s[i], c->recip_ru[i]);
    return (int)F(psum_ru)(fu);
}
static void F(scaling_step)(const orc_ctx *c, int *y, int k, unsigned j, int pow2j, const int *x, const int *cc) {
    long residue = 0;
    for (int i = 0; i < N; i++) residue += (long)F(mod_mul_exact)(c->h[j - 1][i], cc[i], pow2j);
    long temp = (long)k * (long)c->mpow2[j - 1];
    residue = (residue - temp) % pow2j;
    if (residue < 0) residue += pow2j;
#if ORC_POL == 0
    int mult[N];
    for (int i = 0; i < N; i++) mult[i] = (int)(residue % c->m[i]);
    for (int i = 0; i < N; i++) mult[i] = F(mod_psub)(x[i], mult[i], c->m[i]);
    F(rns_mul)(c, y, mult, c->pow2inv[j - 1]);
#else
    for (int i = 0; i < N; i++) {
        int mult = (int)(residue % c->m[i]);
        mult = F(mod_psub)(x[i], mult, c->m[i]);
        y[i] = F(mod_mul_exact)(mult, c->pow2inv[j - 1][i], c->m[i]);
    }
#endif
}
static void F(scale2pow)(const orc_ctx *c, int *result, const int *xin, unsigned D) {
    int x[N], cc[N];
    for (int i = 0; i < N; i++) x[i] = xin[i];
    for (int i = 0; i < N; i++) result[i] = x[i];
    int t = (int)(D / 30);
    int k;
    if (t > 0) {
        F(rns_mul)(c, cc, x, c->w);
        k = F(rank_full)(c, x, cc);
        F(scaling_step)(c, result, k, 30, 1 << 30, x, cc);
        t -= 1;
    }
    while (t > 0) {
        F(rns_mul)(c, cc, result, c->w);
        k = F(rank_fast)(c, cc);
        F(scaling_step)(c, result, k, 30, 1 << 30, result, cc);
        t -= 1;
    }
    unsigned d = D % 30;
    if (d > 0) {
        F(rns_mul)(c, cc, result, c->w);
#if ORC_POL == 0
        k = d < D ? F(rank_fast)(c, cc) : F(rank_full)(c, result, cc);   /* :1045 */
#else
        k = d < D ? F(rank_full)(c, result, cc) : F(rank_fast)(c, cc);   /* :1157 */
#endif
        F(scaling_step)(c, result, k, d, 1 << d, result, cc);
    }
}

/* ------------------------------------------------------------------ */
/* mp_float_t arithmetic, src/arith                                    */
/* ------------------------------------------------------------------ */
static const MPF F(MP_ZERO);  /* all fields zero, src/arith/arith_utils.cuh:64-72 */

/* mp_round, arith_utils.cuh:138-146 (CUDA :176-183) with the guard of its callers */
static inline void F(round_if_needed)(const orc_ctx *c, MPF *r) {
    if (r->eval[1].frac != 0 && r->eval[1].exp >= c->mp_h) {
        int bits = (int)(r->eval[1].exp - c->mp_h + 1);
        if (bits > 0) {
            r->exp += bits;
            F(scale2pow)(c, r->digits, r->digits, (unsigned)bits);
            F(eval_compute)(c, &r->eval[0], &r->eval[1], r->digits, 1);
        }
    }
}

/* sign_estimate, arith_utils.cuh:152-162 (CUDA :199-207) */
static int F(sign_estimate)(const orc_ctx *c, const int *dx, const int *dy, int sx, int sy, int gamma, int theta,
                            int nzx, int nzy) {
    int lx[N], ly[N];
    for (int i = 0; i < N; i++) {
        lx[i] = F(mod_mul_exact)(dx[i], c->pow2[gamma * N + i] * nzx, c->m[i]);
        ly[i] = F(mod_mul_exact)(dy[i], c->pow2[theta * N + i] * nzy, c->m[i]);
    }
    int cmp = F(mrc_compare)(c, lx, ly);
    return (cmp < 0 ? sy : sx) * (cmp != 0);
}

/* RNS_POW2 row index: rows beyond RNS_MODULI_PRODUCT_LOG2 are read out of bounds
 * by the reference when an operand is zero (its digits are all zero, so the row
 * value does not reach the result); clamp to the last row. */
static inline int F(prow)(const orc_ctx *c, int g) { return g > c->log2m ? c->log2m : g; }

/* mp_add: CPU src/arith/add.cuh:31-114; CUDA mp_add_common :126-184 + round */
static void F(add)(const orc_ctx *c, MPF *r, MPF x, MPF y) {
    int ex = x.exp, ey = y.exp, sx = x.sign, sy = y.sign;
    int dexp = ex - ey;
    int gamma = dexp * (dexp > 0);
    int theta = -dexp * (dexp < 0);
    int nzx = ((y.eval[1].frac == 0) || (theta + y.eval[1].exp) < c->mp_j);
    int nzy = ((x.eval[1].frac == 0) || (gamma + x.eval[1].exp) < c->mp_j);
    gamma = gamma * nzy;
    theta = theta * nzx;
    ex = (ex - gamma) * nzx;
    ey = (ey - theta) * nzy;
    sx *= nzx;
    sy *= nzy;
    int fx = (1 - 2 * sx) * nzx;
    int fy = (1 - 2 * sy) * nzy;
    x.eval[0].exp += gamma;
    x.eval[1].exp += gamma;
    y.eval[0].exp += theta;
    y.eval[1].exp += theta;
    x.eval[0].frac *= fx;
    x.eval[1].frac *= fx;
    y.eval[0].frac *= fy;
    y.eval[1].frac *= fy;
    r->eval[0] = F(er_add_rd)(x.eval[sx], y.eval[sy]);
    r->eval[1] = F(er_add_ru)(x.eval[1 - sx], y.eval[1 - sy]);
    const int pg = F(prow)(c, gamma), pt = F(prow)(c, theta);
    int sr;
#if ORC_POL == 0
    if (r->eval[0].frac * r->eval[1].frac >= 0) {  /* :79 */
        sr = (r->eval[0].frac < 0);
    } else {
        sr = F(sign_estimate)(c, x.digits, y.digits, sx, sy, pg, pt, nzx, nzy);
        r->eval[sr].frac = c->unit_low.frac * (1 - 2 * sr);
        r->eval[sr].exp = c->unit_low.exp;
    }
#else
    sr = r->eval[0].frac < 0;  /* :164-169 */
    if (sr != (r->eval[1].frac < 0)) {
        sr = F(sign_estimate)(c, x.digits, y.digits, sx, sy, pg, pt, nzx, nzy);
        r->eval[sr].frac = c->unit_low.frac * (1 - 2 * sr);
        r->eval[sr].exp = c->unit_low.exp;
    }
#endif
    r->sign = sr;
    r->exp = (ex == 0) ? ey : ex;
    for (int i = 0; i < N; i++) {
        long a = (long)c->pow2[pg * N + i] * fx, b = (long)c->pow2[pt * N + i] * fy;
        long v = (a * (long)x.digits[i] + b * (long)y.digits[i]) % (long)c->m[i];
        int res = (int)v;
        r->digits[i] = res < 0 ? res + c->m[i] : res;
    }
    if (sr == 1) {
        for (int i = 0; i < N; i++) r->digits[i] = (c->m[i] - r->digits[i]) % c->m[i];
        orc_er_t tmp = r->eval[0];
        r->eval[0].frac = -1 * r->eval[1].frac;
        r->eval[0].exp = r->eval[1].exp;
        r->eval[1].frac = -1 * tmp.frac;
        r->eval[1].exp = tmp.exp;
    }
    F(round_if_needed)(c, r);
}

/* mp_add_abs (|x| + |y|): CPU src/arith/addabs.cuh:31-74; CUDA mp_add_abs_common :81-116
 * and mp_add_abs :118-127. Both align as mp_add does, add the bounds and the scaled
 * residues without signs, and round. */
static void F(add_abs)(const orc_ctx *c, MPF *r, MPF x, MPF y) {
    int ex = x.exp, ey = y.exp;
    int dexp = ex - ey;
    int gamma = dexp * (dexp > 0);
    int theta = -dexp * (dexp < 0);
    int nzx = ((y.eval[1].frac == 0) || (theta + y.eval[1].exp) < c->mp_j);
    int nzy = ((x.eval[1].frac == 0) || (gamma + x.eval[1].exp) < c->mp_j);
    gamma = gamma * nzy;
    theta = theta * nzx;
    ex = (ex - gamma) * nzx;
    ey = (ey - theta) * nzy;
    x.eval[0].exp += gamma;
    x.eval[1].exp += gamma;
    y.eval[0].exp += theta;
    y.eval[1].exp += theta;
    x.eval[0].frac *= nzx;
    x.eval[1].frac *= nzx;
    y.eval[0].frac *= nzy;
    y.eval[1].frac *= nzy;
    r->eval[0] = F(er_add_rd)(x.eval[0], y.eval[0]);
    r->eval[1] = F(er_add_ru)(x.eval[1], y.eval[1]);
    r->sign = 0;
    r->exp = (ex == 0) ? ey : ex;
    const int pg = F(prow)(c, gamma), pt = F(prow)(c, theta);
    for (int i = 0; i < N; i++) {
        long v = ((long)c->pow2[pg * N + i] * nzx * (long)x.digits[i] +
                  (long)c->pow2[pt * N + i] * nzy * (long)y.digits[i]) % (long)c->m[i];
        r->digits[i] = (int)v;
    }
    F(round_if_needed)(c, r);
}

/* mp_cmp (abs_only = 0) and mp_cmp_abs (abs_only = 1): CPU src/arith/cmp.cuh:31-67 and
 * cmpabs.cuh:31-64; CUDA mp_cmp_common cmp.cuh:79-114 and mp_cmp_abs_common cmpabs.cuh:76-109.
 * The operands are aligned to the larger exponent (the smaller one is dropped when the larger's
 * aligned evaluation reaches 2^MP_J), then compared as RNS magnitudes with the aligned
 * evaluations: rns_cmp(x, exl, exu, y, eyl, eyu), src/rns.cuh:1208-1225 (CUDA :1275-1289).
 * Both policies compute the same thing: residues by exact products, no directed rounding. */
static int F(cmp)(const orc_ctx *c, MPF x, MPF y, int abs_only) {
    int dexp = x.exp - y.exp;
    int gamma = dexp * (dexp > 0);
    int theta = -dexp * (dexp < 0);
    int nzx = ((y.eval[1].frac == 0) || (theta + y.eval[1].exp) < c->mp_j);
    int nzy = ((x.eval[1].frac == 0) || (gamma + x.eval[1].exp) < c->mp_j);
    gamma = gamma * nzy;
    theta = theta * nzx;
    x.eval[0].exp += gamma;
    x.eval[1].exp += gamma;
    y.eval[0].exp += theta;
    y.eval[1].exp += theta;
    x.eval[0].frac *= nzx;
    x.eval[1].frac *= nzx;
    y.eval[0].frac *= nzy;
    y.eval[1].frac *= nzy;
    int lx[N], ly[N];
    const int pg = F(prow)(c, gamma), pt = F(prow)(c, theta);
    for (int i = 0; i < N; i++) {
        lx[i] = nzx * F(mod_mul_exact)(x.digits[i], c->pow2[pg * N + i], c->m[i]);
        ly[i] = nzy * F(mod_mul_exact)(y.digits[i], c->pow2[pt * N + i], c->m[i]);
    }
    int cmp;
    if (F(er_ucmp)(x.eval[0], y.eval[1]) > 0) {
        cmp = 1;
    } else if (F(er_ucmp)(y.eval[0], x.eval[1]) > 0) {
        cmp = -1;
    } else {
        int eq = 1;
        for (int i = 0; i < N; i++)
            if (lx[i] != ly[i]) {
                eq = 0;
                break;
            }
        cmp = eq ? 0 : F(mrc_compare)(c, lx, ly);
    }
    if (abs_only) return cmp;
    const int sx = x.sign, sy = y.sign;
    const int greater = (sx == 0 && sy == 1) || (sx == 0 && sy == 0 && cmp == 1) || (sx == 1 && sy == 1 && cmp == -1);
    const int less = (sx == 1 && sy == 0) || (sx == 0 && sy == 0 && cmp == -1) || (sx == 1 && sy == 1 && cmp == 1);
    return greater ? 1 : less ? -1 : 0;
}

/* mp_mul: CPU src/arith/mul.cuh:31-42; CUDA :53-74 (and the esi/digits/round
 * vector kernels src/mpvector.cuh:45-122,139-216 compute the same fields) */
static void F(mul)(const orc_ctx *c, MPF *r, MPF x, MPF y) {
    r->exp = x.exp + y.exp;
    r->sign = x.sign ^ y.sign;
    r->eval[0] = F(er_md_rd)(x.eval[0], y.eval[0], c->unit_upp);
    r->eval[1] = F(er_md_ru)(x.eval[1], y.eval[1], c->unit_low);
    for (int i = 0; i < N; i++) r->digits[i] = F(mod_mul_exact)(x.digits[i], y.digits[i], c->m[i]);
    F(round_if_needed)(c, r);
}

/* mp_mul_d: CPU src/arith/muld.cuh:31-47; CUDA :59-73 */
static void F(mul_d)(const orc_ctx *c, MPF *r, MPF x, double y) {
    union orc_du u;
    u.d = y;
    long significand = (long)((u.u & (((uint64_t)1 << 52) - 1)) | ((uint64_t)1 << 52)) * (y != 0);
    int yexp = (int)((u.u >> 52) & 2047) - (1023 + 52) * (y != 0);
    int ysign = (int)(u.u >> 63);
    r->exp = x.exp + yexp;
    r->sign = x.sign ^ ysign;
    r->eval[0].exp = x.eval[0].exp;
    r->eval[1].exp = x.eval[1].exp;
    r->eval[0].frac = F(dmul_rd)(x.eval[0].frac, (double)significand);
    r->eval[1].frac = F(dmul_ru)(x.eval[1].frac, (double)significand);
    F(er_adjust)(&r->eval[0]);
    F(er_adjust)(&r->eval[1]);
    for (int i = 0; i < N; i++)
        r->digits[i] = F(mod_mul_exact)(x.digits[i], (int)(significand % c->m[i]), c->m[i]);
    F(round_if_needed)(c, r);
}

/* mp_set_d, src/arith/assign.cuh:54-81 (+ mp_set :41-49). The reference has this
 * conversion on the CPU only; the policy-D variant evaluates with the CUDA eval. */
static void F(set_d)(const orc_ctx *c, MPF *r, double x) {
    union orc_du u;
    u.d = x;
    int sign = (int)(u.u >> 63);
    int exp = (int)((u.u >> 52) & 2047);
    long significand;
    if (exp == 0)
        significand = (long)(u.u & (((uint64_t)1 << 52) - 1));
    else
        significand = (long)((u.u & (((uint64_t)1 << 52) - 1)) | ((uint64_t)1 << 52));
    exp = exp - 1023 - 52;
    if (significand != 0) {
        while ((significand & 1) == 0) {
            significand >>= 1;
            exp++;
        }
    } else {
        exp = 0;
    }
    r->sign = sign;
    r->exp = exp;
    for (int i = 0; i < N; i++) r->digits[i] = (int)(significand % (long)c->m[i]);
    F(eval_compute)(c, &r->eval[0], &r->eval[1], r->digits, 0);
    F(round_if_needed)(c, r);
}

#if ORC_POL == 1
/* v1 element-wise addition: mp_vector_add_esi_kernel + mp_vector_add_digits_kernel
 * + mp_vector_round_kernel (src/mpvector.cuh:234-472, 669-714); the matrix kernels
 * (src/mpmatrix.cuh:47-172, 410-430) compute the same. No MRC on a straddling
 * interval, residues by a double-reciprocal quotient. */
static void F(vadd)(const orc_ctx *c, MPF *r, MPF x, MPF y) {
    int ex = x.exp, ey = y.exp, sx = x.sign, sy = y.sign;
    int dexp = ex - ey;
    int gamma = dexp * (dexp > 0);
    int theta = -dexp * (dexp < 0);
    int nzx = ((y.eval[1].frac == 0) || (theta + y.eval[1].exp) < c->mp_j);
    int nzy = ((x.eval[1].frac == 0) || (gamma + x.eval[1].exp) < c->mp_j);
    gamma = gamma * nzy;
    theta = theta * nzx;
    ex = (ex - gamma) * nzx;
    ey = (ey - theta) * nzy;
    sx *= nzx;
    sy *= nzy;
    int fx = (1 - 2 * sx) * nzx;
    int fy = (1 - 2 * sy) * nzy;
    x.eval[0].exp += gamma;
    x.eval[1].exp += gamma;
    y.eval[0].exp += theta;
    y.eval[1].exp += theta;
    x.eval[0].frac *= fx;
    x.eval[1].frac *= fx;
    y.eval[0].frac *= fy;
    y.eval[1].frac *= fy;
    r->eval[0] = F(er_add_rd)(x.eval[sx], y.eval[sy]);
    r->eval[1] = F(er_add_ru)(x.eval[1 - sx], y.eval[1 - sy]);
    r->exp = (ex == 0) ? ey : ex;
    int minus = r->eval[0].frac < 0 && r->eval[1].frac < 0;
    r->sign = minus;
    if (minus) {
        orc_er_t tmp = r->eval[0];
        r->eval[0].frac = -r->eval[1].frac;
        r->eval[0].exp = r->eval[1].exp;
        r->eval[1].frac = -tmp.frac;
        r->eval[1].exp = tmp.exp;
    }
    const int pg = F(prow)(c, gamma), pt = F(prow)(c, theta);
    for (int i = 0; i < N; i++) {
        const int m = c->m[i];
        const double d = 1.0 / m;
        long rr = (long)(fx * x.digits[i]) * (long)c->pow2[pg * N + i] + (long)(fy * y.digits[i]) * (long)c->pow2[pt * N + i];
        double q = (double)rr * d;
        /* with ~31-bit moduli the quotient can leave the int range; the CUDA
         * conversion saturates (PTX cvt clamps float-to-integer results), where
         * a plain C cast on x86 would give INT_MIN */
        int qi = (q >= 2147483648.0) ? 2147483647 : (q <= -2147483649.0) ? (-2147483647 - 1) : (int)q;
        int res = (int)(rr - (long)qi * (long)m);
        if (r->sign == 1) res = (int)(((long)m - (long)res) % (long)m);
        r->digits[i] = res < 0 ? res + m : res;
    }
    /* mp_vector_round_kernel */
    int bits = (int)((r->eval[1].exp - c->mp_h + 1) * (r->eval[1].frac != 0));
    if (bits > 0) {
        r->exp += bits;
        F(scale2pow)(c, r->digits, r->digits, (unsigned)bits);
        F(eval_compute)(c, &r->eval[0], &r->eval[1], r->digits, 1);
    }
}
#endif

/* ------------------------------------------------------------------ */
/* Compositions in the reference's operation order                     */
/* ------------------------------------------------------------------ */
static inline long F(bidx)(long i, long n, int inc) { /* BLAS stride rule, e.g. src/mpvector.cuh:65-67 */
    return inc > 0 ? i * inc : (-n + i + 1) * inc;
}
static unsigned F(next_pow2)(unsigned v) { /* src/common.cuh:92-101 */
    v--;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}
/* halving tree of mp_array_reduce_sum_kernel1/2 (src/mpreduct.cuh:60-68) */
static void F(tree)(const orc_ctx *c, MPF *s, int bsize) {
    unsigned p = F(next_pow2)((unsigned)bsize);
    for (unsigned i = p >> 1; i >= 1; i >>= 1) {
        for (unsigned tid = 0; tid < i; tid++)
            if (tid + i < (unsigned)bsize) F(add)(c, &s[tid], s[tid], s[tid + i]);
    }
}

/* src/blas/dot.cuh:41-107 */
static void F(dot)(const orc_ctx *c, int n, const MPF *x, int incx, const MPF *y, int incy, int g3, int b3, MPF *res) {
    if (n <= 0) return;
    MPF *prod = (MPF *)malloc(sizeof(MPF) * (size_t)n);
    MPF *part = (MPF *)malloc(sizeof(MPF) * (size_t)g3);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) F(mul)(c, &prod[i], x[F(bidx)(i, n, incx)], y[F(bidx)(i, n, incy)]);
    const long G = (long)g3 * b3;
#pragma omp parallel for schedule(dynamic, 1)
    for (int bid = 0; bid < g3; bid++) {
        MPF *s = (MPF *)malloc(sizeof(MPF) * (size_t)b3);
        for (int tid = 0; tid < b3; tid++) {
            s[tid] = F(MP_ZERO);
            for (long i = (long)bid * b3 + tid; i < n; i += G) F(add)(c, &s[tid], s[tid], prod[i]);
        }
        F(tree)(c, s, b3);
        part[bid] = s[0];
        free(s);
    }
    MPF *s = (MPF *)malloc(sizeof(MPF) * (size_t)b3);
    for (int tid = 0; tid < b3; tid++) {
        s[tid] = F(MP_ZERO);
        for (int i = tid; i < g3; i += b3) F(add)(c, &s[tid], s[tid], part[i]);
    }
    F(tree)(c, s, b3);
    *res = s[0];
    free(s);
    free(part);
    free(prod);
}

/* src/blas/axpy.cuh:45-74: buffer = alpha*x (+round), y = buffer + y (+round) */
static void F(axpy)(const orc_ctx *c, int n, const MPF *alpha, const MPF *x, int incx, MPF *y, int incy) {
    if (n <= 0) return;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        MPF b;
        long iy = F(bidx)(i, n, incy);
        F(mul)(c, &b, x[F(bidx)(i, n, incx)], *alpha);
#if ORC_POL == 1
        F(vadd)(c, &y[iy], b, y[iy]);
#else
        F(add)(c, &y[iy], b, y[iy]);
#endif
    }
}

/* src/blas/v2/axpy_v2.cuh:39-52 */
static void F(axpy_v2)(const orc_ctx *c, int n, const MPF *alpha, const MPF *x, const MPF *y, MPF *r) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        MPF ax;
        F(mul)(c, &ax, *alpha, x[i]);
        F(add)(c, &r[i], ax, y[i]);
    }
}

/* src/blas/v2/maxpy_v2.cuh: r = (-alpha) x + y (mp_mul, then mp_add) */
static void F(maxpy_v2)(const orc_ctx *c, int n, const MPF *alpha, const MPF *x, const MPF *y, MPF *r) {
    MPF a = *alpha;
    a.sign ^= 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        MPF ax;
        F(mul)(c, &ax, a, x[i]);
        F(add)(c, &r[i], ax, y[i]);
    }
}

/* halving tree of mp_sumsq_kernel1/2 with mp_add_abs (src/blas/v2/norm2_v2.cuh:50-58) */
static void F(tree_abs)(const orc_ctx *c, MPF *s, int bsize) {
    unsigned p = F(next_pow2)((unsigned)bsize);
    for (unsigned i = p >> 1; i >= 1; i >>= 1) {
        for (unsigned tid = 0; tid < i; tid++)
            if (tid + i < (unsigned)bsize) F(add_abs)(c, &s[tid], s[tid], s[tid + i]);
    }
}
/* sum of squares of mp_norm2<blocks, threads> (src/blas/v2/norm2_v2.cuh:36-120):
 * mp_sumsq_kernel1 (thread tid of block bid folds x_i * x_i, i = bid*threads + tid, step
 * blocks*threads, by mp_add_abs from zero; halving tree), then mp_sumsq_kernel2 (one block
 * of `threads` over the `blocks` block sums) */
static void F(sumsq)(const orc_ctx *c, int n, const MPF *x, int blocks, int threads, MPF *res) {
    MPF *part = (MPF *)malloc(sizeof(MPF) * (size_t)blocks);
    const long G = (long)blocks * threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (int bid = 0; bid < blocks; bid++) {
        MPF *s = (MPF *)malloc(sizeof(MPF) * (size_t)threads);
        for (int tid = 0; tid < threads; tid++) {
            s[tid] = F(MP_ZERO);
            for (long i = (long)bid * threads + tid; i < n; i += G) {
                MPF p;
                F(mul)(c, &p, x[i], x[i]);
                F(add_abs)(c, &s[tid], s[tid], p);
            }
        }
        F(tree_abs)(c, s, threads);
        part[bid] = s[0];
        free(s);
    }
    MPF *s = (MPF *)malloc(sizeof(MPF) * (size_t)threads);
    for (int tid = 0; tid < threads; tid++) {
        s[tid] = F(MP_ZERO);
        for (int i = tid; i < blocks; i += threads) F(add_abs)(c, &s[tid], s[tid], part[i]);
    }
    F(tree_abs)(c, s, threads);
    *res = s[0];
    free(s);
    free(part);
}

#if ORC_POL == 0
/* ------------------------------------------------------------------ */
/* Host MPFR operations of the reference (policy-independent: the CUDA  */
/* mp_div / mp_sqrt / mp_dsqrt copy the operands to the host and run    */
/* these, src/arith/div.cuh:56-65, src/arith/sqrt.cuh:65-89)            */
/* ------------------------------------------------------------------ */
/* mp_get_mpfr, src/arith/assign.cuh:185-211: X = sum digits_i * RNS_ORTHOGONAL_BASE_i mod M,
 * then (-1)^sign X 2^exp rounded to r's precision to nearest (mpfr_set_str of the binary
 * string there, mpfr_set_z_2exp here: both round the exact value correctly) */
static void F(get_mpfr)(const orc_ctx *c, mpfr_t r, const MPF *x) {
    mpz_t X, M, t;
    mpz_inits(X, M, t, NULL);
    mpz_set_ui(M, 1);
    for (int i = 0; i < N; i++) mpz_mul_ui(M, M, (unsigned long)c->m[i]);
    for (int i = 0; i < N; i++) {
        mpz_divexact_ui(t, M, (unsigned long)c->m[i]);  /* M_i */
        mpz_mul_ui(t, t, (unsigned long)c->w[i]);       /* M_i * w_i */
        mpz_mul_si(t, t, (long)x->digits[i]);
        mpz_add(X, X, t);
    }
    mpz_mod(X, X, M);
    mpfr_set_z_2exp(r, X, (long)x->exp, MPFR_RNDN);
    if (x->sign == 1) mpfr_neg(r, r, MPFR_RNDN);  /* "-" prefix: a zero becomes -0, as the string does */
    mpz_clears(X, M, t, NULL);
}
/* mp_set_mpfr, src/arith/assign.cuh:86-127: the significand string of mpfr_get_str (base 2,
 * all digits), trailing zeros dropped, residues of its magnitude, the CPU evaluation */
static void F(set_mpfr)(const orc_ctx *c, MPF *r, mpfr_srcptr x) {
    mpfr_exp_t e;
    char *str = mpfr_get_str(NULL, &e, 2, 0, x, MPFR_RNDN);
    size_t len = strlen(str);
    size_t num = 0;
    for (size_t i = len; i-- > 0;)
        if (str[i] != '0') {
            num = i;
            break;
        }
    mpz_t mant, rem;
    mpz_inits(mant, rem, NULL);
    size_t used = (num != len - 1) ? num + 1 : len;
    char *sub = (char *)malloc(used + 1);
    memcpy(sub, str, used);
    sub[used] = 0;
    if (mpz_set_str(mant, sub, 2) != 0) mpz_set_ui(mant, 0);  /* "-" alone: mant keeps 0 */
    r->exp = (int)(e - (long)used);
    if (str[0] == '-') {
        r->sign = 1;
        r->exp++;
        mpz_neg(mant, mant);
    } else {
        r->sign = 0;
    }
    for (int i = 0; i < N; i++) r->digits[i] = (int)mpz_fdiv_r_ui(rem, mant, (unsigned long)c->m[i]);
    F(eval_compute)(c, &r->eval[0], &r->eval[1], r->digits, 0);
    free(sub);
    mpfr_free_str(str);
    mpz_clears(mant, rem, NULL);
}
/* mp_div, src/arith/div.cuh:31-44 */
static void F(div)(const orc_ctx *c, MPF *r, const MPF *x, const MPF *y) {
    mpfr_t a, b;
    mpfr_init2(a, c->mp_precision);
    mpfr_init2(b, c->mp_precision);
    F(get_mpfr)(c, a, x);
    F(get_mpfr)(c, b, y);
    mpfr_div(a, a, b, MPFR_RNDN);
    F(set_mpfr)(c, r, a);
    mpfr_clear(a);
    mpfr_clear(b);
}
/* mp_sqrt and mp_dsqrt, src/arith/sqrt.cuh:33-57 */
static void F(sqrt)(const orc_ctx *c, MPF *r, const MPF *x) {
    mpfr_t a;
    mpfr_init2(a, c->mp_precision);
    F(get_mpfr)(c, a, x);
    mpfr_sqrt(a, a, MPFR_RNDN);
    F(set_mpfr)(c, r, a);
    mpfr_clear(a);
}
static double F(dsqrt)(const orc_ctx *c, const MPF *x) {
    mpfr_t a;
    mpfr_init2(a, c->mp_precision);
    F(get_mpfr)(c, a, x);
    mpfr_sqrt(a, a, MPFR_RNDN);
    double d = mpfr_get_d(a, MPFR_RNDN);
    mpfr_clear(a);
    return d;
}
#endif

/* src/blas/gemv.cuh:44-268 */
static void F(gemv)(const orc_ctx *c, int trans, int m, int n, const MPF *alpha, const MPF *A, int lda, const MPF *x,
                    int incx, const MPF *beta, MPF *y, int incy, int b3, const int *rows, int nrows) {
    if (m <= 0 || n <= 0 || incx == 0 || incy == 0 || lda < (m > 1 ? m : 1)) return;
    const int lx = trans ? m : n;   /* length of x */
    const int ly = trans ? n : m;   /* length of y */
    MPF *bx = (MPF *)malloc(sizeof(MPF) * (size_t)lx);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < lx; j++) F(mul)(c, &bx[j], x[F(bidx)(j, lx, incx)], *alpha);
    const int cnt = rows ? nrows : ly;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < cnt; q++) {
        const int i = rows ? rows[q] : q;   /* output index */
        MPF *s = (MPF *)malloc(sizeof(MPF) * (size_t)b3);
        const int len = trans ? m : n;
        for (int t = 0; t < b3; t++) {
            s[t] = F(MP_ZERO);
            for (int j = t; j < len; j += b3) {
                MPF p;
                const MPF *a = trans ? &A[(long)i * lda + j] : &A[(long)j * lda + i];
                F(mul)(c, &p, *a, bx[j]);
                F(add)(c, &s[t], s[t], p);
            }
        }
        F(tree)(c, s, b3);
        long iy = F(bidx)(i, ly, incy);
        MPF yb;
        F(mul)(c, &yb, y[iy], *beta);
        F(add)(c, &y[iy], yb, s[0]);
        free(s);
    }
    free(bx);
}

/* src/blas/gemm.cuh:39-167 (no-trans) */
static void F(gemm)(const orc_ctx *c, int m, int n, int k, const MPF *alpha, const MPF *A, int lda, const MPF *B,
                    int ldb, const MPF *beta, MPF *C, int ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int col = 0; col < n; col++)
        for (int row = 0; row < m; row++) {
            MPF sum = F(MP_ZERO), mul, cb;
            for (int i = 0; i < k; i++) {
                F(mul)(c, &mul, A[(long)lda * i + row], B[(long)col * ldb + i]);
                F(add)(c, &sum, sum, mul);
            }
            F(mul)(c, &mul, sum, *alpha);
            MPF *cc = &C[(long)col * ldc + row];
            F(mul)(c, &cb, *cc, *beta);
#if ORC_POL == 1
            F(vadd)(c, cc, cb, mul);
#else
            F(add)(c, cc, cb, mul);
#endif
        }
}

/* src/blas/v2/gemm_v2.cuh:55-81 */
static void F(gemm_v2)(const orc_ctx *c, int ta, int tb, int m, int n, int k, const MPF *alpha, const MPF *A, int lda,
                       const MPF *B, int ldb, const MPF *beta, MPF *C, int ldc, const int *entries, int nentries) {
    const long cnt = entries ? nentries : (long)m * n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long q = 0; q < cnt; q++) {
        int row, col;
        if (entries) {
            row = entries[2 * q];
            col = entries[2 * q + 1];
        } else {
            row = (int)(q % m);
            col = (int)(q / m);
        }
        MPF sum = F(MP_ZERO), mul;
        for (int i = 0; i < k; i++) {
            long ia = ta ? (long)i + (long)lda * row : (long)row + (long)lda * i;
            long ib = tb ? (long)col + (long)ldb * i : (long)i + (long)ldb * col;
            F(mul)(c, &mul, A[ia], B[ib]);
            F(add)(c, &sum, sum, mul);
        }
        F(mul)(c, &sum, sum, *alpha);
        MPF *cc = &C[(long)row + (long)col * ldc];
        F(mul)(c, &mul, *cc, *beta);
        F(add)(c, cc, sum, mul);
    }
}

/* src/sparse/spmv/spmv_csr.cuh:48-64 (tpr == 1) and spmv_csrv.cuh:48-87 */
static void F(spmv)(const orc_ctx *c, int m, const int *irp, const int *ja, const double *as, const MPF *x, MPF *y,
                    int tpr, const int *rows, int nrows) {
    const int cnt = rows ? nrows : m;
#pragma omp parallel for schedule(dynamic, 64)
    for (int q = 0; q < cnt; q++) {
        const int row = rows ? rows[q] : q;
        MPF s[32], p;
        for (int l = 0; l < tpr; l++) {
            s[l] = F(MP_ZERO);
            for (int i = irp[row] + l; i < irp[row + 1]; i += tpr) {
                F(mul_d)(c, &p, x[ja[i]], as[i]);
                F(add)(c, &s[l], s[l], p);
            }
        }
        for (int h = 16; h >= 1; h >>= 1)
            if (tpr >= 2 * h)
                for (int l = 0; l < h; l++) F(add)(c, &s[l], s[l], s[l + h]);
        y[q] = s[0];
    }
}

/* the ELLPACK, jagged-diagonal and diagonal kernels (src/sparse/spmv/spmv_ell.cuh:50-66,
 * spmv_jad.cuh:50-66, spmv_jadv.cuh:50-87, spmv_dia.cuh:50-66): per row, the lane-strided
 * terms in the stored order from zero, then the 16/8/4/2/1 tree (tpr > 1: jadv only).
 * fmt 1 ELL (as, ja: m x w column-major; zero values skipped), 2 JAD (jcp, perm; the row's
 * result goes to y[perm[row]]), 3 DIA (offset; w diagonals, ncols columns) */
static void F(spmv_fmt)(const orc_ctx *c, int fmt, int m, int w, int ncols, const double *as, const int *ja,
                        const int *jcp, const int *perm, const int *offset, const MPF *x, MPF *y, int tpr) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int row = 0; row < m; row++) {
        MPF s[32], p;
        for (int l = 0; l < tpr; l++) {
            s[l] = F(MP_ZERO);
            for (int t = l; t < w; t += tpr) {
                int col;
                double v;
                if (fmt == 1) {
                    col = ja[(long)t * m + row];
                    v = as[(long)t * m + row];
                    if (v == 0) continue;
                } else if (fmt == 2) {
                    long q = (long)row + jcp[t];
                    if (q >= jcp[t + 1]) continue;
                    col = ja[q];
                    v = as[q];
                } else {
                    long cc = (long)row + offset[t];
                    v = as[(long)m * t + row];
                    if (cc < 0 || cc >= ncols || v == 0) continue;
                    col = (int)cc;
                }
                F(mul_d)(c, &p, x[col], v);
                F(add)(c, &s[l], s[l], p);
            }
        }
        for (int h = 16; h >= 1; h >>= 1)
            if (tpr >= 2 * h)
                for (int l = 0; l < h; l++) F(add)(c, &s[l], s[l], s[l + h]);
        y[fmt == 2 ? perm[row] : row] = s[0];
    }
}

#undef MPF
#undef N
#undef F
#undef ORC_SFX
#undef ORC_CAT
#undef ORC_CAT_
