// Host build of the product's scalar numerics (mpres-blas_amd/csrc/numerics.hpp)
// so that tests can compare them with the oracle on the CPU.
#include <vector>

#include "../../mpres-blas_amd/csrc/internal.hpp"
#include "../../mpres-blas_amd/csrc/numerics.hpp"
#include "../../mpres-blas_amd/csrc/bigint.hpp"

extern "C" {
double nh_add(int up, double a, double b) { return up ? mph::add_ru(a, b) : mph::add_rd(a, b); }
double nh_sub(int up, double a, double b) { return up ? mph::sub_ru(a, b) : mph::sub_rd(a, b); }
double nh_mul(int up, double a, double b) { return up ? mph::mul_ru(a, b) : mph::mul_rd(a, b); }
double nh_div(int up, double a, double b) { return up ? mph::div_ru(a, b) : mph::div_rd(a, b); }
// directed division through the two-double reciprocal of z (recip2, divk_rd / divk_ru)
double nh_divk(int up, double a, double z) {
    double zh, zl;
    mph::recip2(z, zh, zl);
    return up ? mph::divk_ru(a, z, zh, zl) : mph::divk_rd(a, z, zh, zl);
}
// the interval evaluation's unit fractions of a moduli set and their stored reciprocals
int nh_unit(int set_id, double *out) {
    const mph::ModuliSet *s = mph::get_set(set_id);
    if (!s) return -1;
    mph::DevConsts c;
    std::vector<int> pow2;
    if (!mph::compute_consts(*s, c, pow2)) return -1;
    out[0] = c.unit_low.frac;
    out[1] = c.unit_upp.frac;
    out[2] = c.ul_rh;
    out[3] = c.ul_rl;
    out[4] = c.uu_rh;
    out[5] = c.uu_rl;
    return 0;
}
int nh_ceil_log2(double x) { return mph::ceil_log2_rn(x); }
int nh_mulmod(int a, int b, int m) { return mph::mulmod(a, b, (double)m, 1.0 / (double)m); }
int nh_modd(double v, int m) { return mph::modd(v, (double)m, 1.0 / (double)m); }
void nh_er_add(int up, double xf, int xe, double yf, int ye, double *rf, int *re) {
    mph::Er r = mph::er_add_dir({xf, xe}, {yf, ye}, up != 0);
    *rf = r.frac;
    *re = r.exp;
}
void nh_er_adjust(double f, int e, double *rf, int *re) {
    mph::Er r = mph::er_adjust(f, e);
    *rf = r.frac;
    *re = r.exp;
}
// fixed-width integers (bigint.hpp) at 24 limbs, the exact finish's common accumulator
constexpr int BW = 24;
static mph::Big<BW> big_in(const uint32_t *a) {
    mph::Big<BW> r;
    for (int i = 0; i < BW; i++) r.l[i] = a[i];
    return r;
}
static void big_out(const mph::Big<BW> &r, uint32_t *o) {
    for (int i = 0; i < BW; i++) o[i] = r.l[i];
}
void nh_big_shl(const uint32_t *a, int s, uint32_t *o) { big_out(mph::big_shl(big_in(a), s), o); }
void nh_big_shr(const uint32_t *a, int s, uint32_t *o) { big_out(mph::big_shr(big_in(a), s), o); }
// bit length, trailing zeros, bit i, any bit below i
void nh_big_bits(const uint32_t *a, int i, int *o) {
    const mph::Big<BW> x = big_in(a);
    o[0] = mph::big_bitlen(x);
    o[1] = mph::big_ctz(x);
    o[2] = mph::big_bit(x, i);
    o[3] = mph::big_any_below(x, i) ? 1 : 0;
}
int nh_big_cmp(const uint32_t *a, const uint32_t *b) { return mph::big_cmp(big_in(a), big_in(b)); }
int nh_big_round(const uint32_t *a, int p, uint32_t *o) {
    mph::Big<BW> x = big_in(a);
    int e = 0;
    mph::big_round(x, e, p);
    big_out(x, o);
    return e;
}
// a mod m through FP64 steps (big_mod_d) and through 64-bit integer division (big_mod_u32)
void nh_big_mod(const uint32_t *a, int m, int *o) {
    const mph::Big<BW> x = big_in(a);
    o[0] = mph::big_mod_d(x, (double)m, 1.0 / (double)m);
    o[1] = (int)mph::big_mod_u32(x, (uint32_t)m);
}
}
