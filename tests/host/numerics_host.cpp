// Host build of the product's scalar numerics (mpres-blas_amd/csrc/numerics.hpp)
// so that tests can compare them with the oracle on the CPU.
#include "../../mpres-blas_amd/csrc/numerics.hpp"

extern "C" {
double nh_add(int up, double a, double b) { return up ? mph::add_ru(a, b) : mph::add_rd(a, b); }
double nh_sub(int up, double a, double b) { return up ? mph::sub_ru(a, b) : mph::sub_rd(a, b); }
double nh_mul(int up, double a, double b) { return up ? mph::mul_ru(a, b) : mph::mul_rd(a, b); }
double nh_div(int up, double a, double b) { return up ? mph::div_ru(a, b) : mph::div_rd(a, b); }
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
