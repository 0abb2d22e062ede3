// Matrix Market input for the sparse formats (host code, no GPU work): the reference's
// src/sparse/utils/mtx_reader.cuh step for step. The order in which entries land in each
// row fixes the summation order of the SpMV kernels, so the conversions keep the
// reference's own sequence of operations: the file's entries first, then (symmetric
// files) the mirrored ones, then std::sort by row, which is not stable - rows are
// ordered within by whatever std::sort does with that sequence, as in the reference.
// All arrays are host arrays the caller allocates with the sizes mpres_mtx_info gives.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mpres_hip.h"
#include "internal.hpp"

namespace {

using mph::set_error;

// the file after its comment lines, optionally past the size line
bool open_body(std::ifstream &f, const char *path, bool skip_size) {
    f.open(path);
    if (!f) return false;
    while (f.peek() == '%') f.ignore(2048, '\n');
    if (skip_size) f.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    return true;
}

// sorted, deduplicated offsets col - row of the file's entries; mirrored ones appended
// for symmetric files (calc_ndiag and build_dia, mtx_reader.cuh:92-117, 254-275)
bool diag_offsets(const char *path, int lines, bool symmetric, std::vector<int> &off) {
    std::ifstream f;
    if (!open_body(f, path, true)) return false;
    off.assign(lines, 0);
    for (int l = 0; l < lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        off[l] = col - row;
    }
    std::sort(off.begin(), off.end());
    off.erase(std::unique(off.begin(), off.end()), off.end());
    if (symmetric) {
        const int nd = (int)off.size();
        for (int i = nd - 2; i > -1; i--) off.push_back(-off[i]);
    }
    return true;
}

// convert_to_coo with sort_coo_rows, mtx_reader.cuh:119-137, 188-224
bool read_coo(const char *path, int lines, bool symmetric, std::vector<double> &as, std::vector<int> &ia,
              std::vector<int> &ja) {
    std::ifstream f;
    if (!open_body(f, path, true)) return false;
    int nzmd = 0;
    std::vector<double> a(lines);
    std::vector<int> r(lines), c(lines);
    std::vector<double> ma;
    std::vector<int> mr, mc;
    for (int l = 0; l < lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        a[l] = v;
        c[l] = col - 1;
        r[l] = row - 1;
        if (col == row) nzmd++;
        if (symmetric && row != col) {
            ma.push_back(v);
            mc.push_back(row - 1);
            mr.push_back(col - 1);
        }
    }
    a.insert(a.end(), ma.begin(), ma.end());
    r.insert(r.end(), mr.begin(), mr.end());
    c.insert(c.end(), mc.begin(), mc.end());
    struct coo {
        double data;
        int cols;
        int rows;
    };
    std::vector<coo> x(a.size());
    for (size_t i = 0; i < a.size(); i++) x[i] = coo{a[i], c[i], r[i]};
    std::sort(x.begin(), x.end(), [](const coo p, const coo q) { return p.rows < q.rows; });
    as.resize(x.size());
    ia.resize(x.size());
    ja.resize(x.size());
    for (size_t i = 0; i < x.size(); i++) {
        as[i] = x[i].data;
        ia[i] = x[i].rows;
        ja[i] = x[i].cols;
    }
    (void)nzmd;
    return true;
}

// make_irp_array, mtx_reader.cuh:159-167
void make_irp(int m, int nnz, const int *ia, int *irp) {
    int p = 0;
    for (int i = 0; i < m + 1; i++) {
        while (p < nnz && i > ia[p]) p++;
        irp[i] = p;
    }
}

int bad_file(const char *path) { return set_error(MPRES_ERR_ARG, "cannot read Matrix Market file %s", path); }

}  // namespace

extern "C" {

// collect_mtx_stats, mtx_reader.cuh:52-90
int mpres_mtx_info(const char *path, int *m, int *n, int *lines, int *maxnzr, int *nzmd, int *symmetric) {
    if (!path || !m || !n || !lines || !maxnzr || !nzmd || !symmetric) return set_error(MPRES_ERR_ARG, "null argument");
    std::ifstream f(path);
    if (!f) return bad_file(path);
    std::string head, type, formats, dtype, sym;
    f >> head >> type >> formats >> dtype >> sym;
    if (head != "%%MatrixMarket" || (type != "matrix" && type != "tensor") || (sym != "general" && sym != "symmetric"))
        return set_error(MPRES_ERR_ARG, "%s: unsupported Matrix Market header", path);
    *symmetric = sym == "symmetric";
    f.seekg(0, std::ios::beg);
    while (f.peek() == '%') f.ignore(2048, '\n');
    f >> *m >> *n >> *lines;
    if (!f || *m <= 0 || *n <= 0 || *lines < 0) return bad_file(path);
    std::vector<int> nz(*m, 0);
    *nzmd = 0;
    for (int l = 0; l < *lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        if (!f || row < 1 || row > *m || col < 1 || col > *n) return bad_file(path);
        nz[row - 1]++;
        if (row == col) (*nzmd)++;
        if (*symmetric && row != col) nz[col - 1]++;
    }
    *maxnzr = *std::max_element(nz.begin(), nz.end());
    return MPRES_OK;
}

// number of stored entries once mirrored: (lines - nzmd) * 2 + nzmd for symmetric files
int mpres_mtx_nnz(int lines, int nzmd, int symmetric) { return symmetric ? (lines - nzmd) * 2 + nzmd : lines; }

int mpres_mtx_ndiag(const char *path, int lines, int symmetric, int *ndiag) {
    if (!path || !ndiag) return set_error(MPRES_ERR_ARG, "null argument");
    std::vector<int> off;
    if (!diag_offsets(path, lines, symmetric != 0, off)) return bad_file(path);
    *ndiag = (int)off.size();
    return MPRES_OK;
}

// build_csr, mtx_reader.cuh:227-232: as, ja of nnz entries, irp of m + 1
int mpres_mtx_csr(const char *path, int m, int nnz, int lines, int symmetric, double *as, int *ja, int *irp) {
    if (!path || !as || !ja || !irp) return set_error(MPRES_ERR_ARG, "null argument");
    std::vector<double> a;
    std::vector<int> r, c;
    if (!read_coo(path, lines, symmetric != 0, a, r, c)) return bad_file(path);
    if ((int)a.size() != nnz) return set_error(MPRES_ERR_ARG, "%s holds %zu entries, nnz %d given", path, a.size(), nnz);
    std::copy(a.begin(), a.end(), as);
    std::copy(c.begin(), c.end(), ja);
    make_irp(m, nnz, r.data(), irp);
    return MPRES_OK;
}

// build_ell, mtx_reader.cuh:334-366: m x maxnzr column-major, padding ja = -1, as = 0
int mpres_mtx_ell(const char *path, int m, int maxnzr, int lines, int symmetric, double *as, int *ja) {
    if (!path || !as || !ja) return set_error(MPRES_ERR_ARG, "null argument");
    std::fill(ja, ja + (size_t)m * maxnzr, -1);
    std::fill(as, as + (size_t)m * maxnzr, 0.0);
    std::ifstream f;
    if (!open_body(f, path, true)) return bad_file(path);
    std::vector<int> cnum(m, 0);
    for (int l = 0; l < lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        if (cnum[row - 1] >= maxnzr) return set_error(MPRES_ERR_ARG, "%s: row %d exceeds maxnzr", path, row);
        as[(size_t)cnum[row - 1] * m + (row - 1)] = v;
        ja[(size_t)cnum[row - 1] * m + (row - 1)] = col - 1;
        cnum[row - 1]++;
        if (symmetric && row != col) {
            if (cnum[col - 1] >= maxnzr) return set_error(MPRES_ERR_ARG, "%s: row %d exceeds maxnzr", path, col);
            as[(size_t)cnum[col - 1] * m + (col - 1)] = v;
            ja[(size_t)cnum[col - 1] * m + (col - 1)] = row - 1;
            cnum[col - 1]++;
        }
    }
    return MPRES_OK;
}

// build_jad, mtx_reader.cuh:294-331: rows permuted by decreasing length (std::sort, not
// stable), jagged diagonals of the CSR rows
int mpres_mtx_jad(const char *path, int m, int maxnzr, int nnz, int lines, int symmetric, double *as, int *ja, int *jcp,
                  int *perm) {
    if (!path || !as || !ja || !jcp || !perm) return set_error(MPRES_ERR_ARG, "null argument");
    std::vector<double> cas(nnz);
    std::vector<int> cja(nnz), irp(m + 1);
    if (int rc = mpres_mtx_csr(path, m, nnz, lines, symmetric, cas.data(), cja.data(), irp.data())) return rc;
    std::vector<int> nz(m);
    for (int i = 0; i < m; i++) {
        perm[i] = i;
        nz[i] = irp[i + 1] - irp[i];
    }
    struct coo {
        int cols;
        int rows;
    };
    std::vector<coo> x(m);
    for (int i = 0; i < m; i++) x[i] = coo{perm[i], nz[i]};
    std::sort(x.begin(), x.end(), [](const coo p, const coo q) { return p.rows > q.rows; });  // sort_perm_rows
    for (int i = 0; i < m; i++) {
        nz[i] = x[i].rows;
        perm[i] = x[i].cols;
    }
    jcp[0] = 0;
    int count = 0;
    for (int j = 0; j < maxnzr; j++) {
        for (int i = 0; i < m; i++)
            if (nz[i] > 0) {
                nz[i]--;
                count++;
            }
        jcp[j + 1] = count;
    }
    int index = 0, j = 0;
    while (index < nnz && j < maxnzr) {
        for (int i = 0; i < jcp[j + 1] - jcp[j]; i++) {
            as[index] = cas[j + irp[perm[i]]];
            ja[index] = cja[j + irp[perm[i]]];
            index++;
        }
        j++;
    }
    return MPRES_OK;
}

// build_dia, mtx_reader.cuh:235-291: m x ndiag column-major, offsets sorted (mirrored
// ones appended for symmetric files)
int mpres_mtx_dia(const char *path, int m, int lines, int symmetric, int ndiag, double *as, int *offset) {
    if (!path || !as || !offset) return set_error(MPRES_ERR_ARG, "null argument");
    std::vector<int> off;
    if (!diag_offsets(path, lines, symmetric != 0, off)) return bad_file(path);
    if ((int)off.size() != ndiag) return set_error(MPRES_ERR_ARG, "%s has %zu diagonals, %d given", path, off.size(), ndiag);
    std::fill(as, as + (size_t)m * ndiag, 0.0);
    std::copy(off.begin(), off.end(), offset);
    std::ifstream f;
    if (!open_body(f, path, true)) return bad_file(path);
    for (int l = 0; l < lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        int idx = (int)(std::find(off.begin(), off.end(), col - row) - off.begin());
        as[(size_t)m * idx + (row - 1)] = v;
        if (symmetric) {
            idx = (int)(std::find(off.begin(), off.end(), row - col) - off.begin());
            as[(size_t)m * idx + (col - 1)] = v;
        }
    }
    return MPRES_OK;
}

// build_dense, mtx_reader.cuh:369-385: m x n column-major
int mpres_mtx_dense(const char *path, int m, int n, int lines, int symmetric, double *a) {
    if (!path || !a) return set_error(MPRES_ERR_ARG, "null argument");
    std::fill(a, a + (size_t)m * n, 0.0);
    std::ifstream f;
    if (!open_body(f, path, true)) return bad_file(path);
    for (int l = 0; l < lines; l++) {
        double v = 0.0;
        int row = 0, col = 0;
        f >> row >> col >> v;
        a[(size_t)(col - 1) * m + row - 1] = v;
        if (symmetric && row != col) a[(size_t)(row - 1) * m + col - 1] = v;
    }
    return MPRES_OK;
}

// read_vector, mtx_reader.cuh:388-395
int mpres_read_vector(const char *path, int n, double *v) {
    if (!path || !v) return set_error(MPRES_ERR_ARG, "null argument");
    std::ifstream f(path);
    if (!f) return bad_file(path);
    for (int l = 0; l < n; l++) f >> v[l];
    return MPRES_OK;
}

}  // extern "C"
