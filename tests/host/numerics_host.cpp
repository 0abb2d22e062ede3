// Host build of the product's scalar numerics (mpres-blas_amd/csrc/numerics.hpp)
// so that tests can compare them with the oracle on the CPU.
#include <vector>

#include "../../mpres-blas_amd/csrc/internal.hpp"
#include "../../mpres-blas_amd/csrc/numerics.hpp"

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
}
