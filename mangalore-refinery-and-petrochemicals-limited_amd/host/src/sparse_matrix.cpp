// sparse_matrix.cpp -- compressed formats built by a two-pass counting sort.
//
// Semantics kept from the reference (src/sparse_matrix.cpp:59-110, 187-207,
// 226-278): entries ordered by (major, minor) index, duplicates summed,
// exact zeros dropped after summing, out-of-range indices throw
// std::out_of_range, a malformed matrix throws std::invalid_argument.
// Duplicates are summed in insertion order.
#include <bharatopt/sparse_matrix.hpp>

#include <algorithm>
#include <cmath>

namespace bharatopt {

namespace {

// Compress (major, minor, val) triplets: bucket by major index (stable), sort
// each bucket by minor index (stable), merge duplicates, drop zeros.
void compress(size_t n_major, size_t n_minor, const std::vector<COOTriplet>& t, bool row_major,
              std::vector<index_t>& ptr, std::vector<index_t>& idx, std::vector<real_t>& val) {
    auto major = [&](const COOTriplet& e) { return row_major ? e.row : e.col; };
    auto minor = [&](const COOTriplet& e) { return row_major ? e.col : e.row; };
    for (const auto& e : t) {
        if (e.row < 0 || e.col < 0 || (n_major > 0 && static_cast<size_t>(major(e)) >= n_major) ||
            (n_minor > 0 && static_cast<size_t>(minor(e)) >= n_minor))
            throw std::out_of_range("COO triplet index out of range during compression.");
    }
    std::vector<size_t> start(n_major + 1, 0);
    for (const auto& e : t) start[static_cast<size_t>(major(e)) + 1]++;
    for (size_t i = 0; i < n_major; ++i) start[i + 1] += start[i];
    std::vector<size_t> order(t.size());
    {
        std::vector<size_t> fill(start.begin(), start.end() - 1);
        for (size_t k = 0; k < t.size(); ++k) order[fill[static_cast<size_t>(major(t[k]))]++] = k;
    }
    ptr.assign(n_major + 1, 0);
    idx.clear();
    val.clear();
    idx.reserve(t.size());
    val.reserve(t.size());
    for (size_t i = 0; i < n_major; ++i) {
        auto b = order.begin() + static_cast<std::ptrdiff_t>(start[i]);
        auto e = order.begin() + static_cast<std::ptrdiff_t>(start[i + 1]);
        std::stable_sort(b, e, [&](size_t a, size_t c) { return minor(t[a]) < minor(t[c]); });
        for (auto it = b; it != e;) {
            const index_t key = minor(t[*it]);
            real_t s = t[*it].val;
            auto jt = it + 1;
            while (jt != e && minor(t[*jt]) == key) {
                s += t[*jt].val;
                ++jt;
            }
            if (s != 0.0) {
                idx.push_back(key);
                val.push_back(s);
            }
            it = jt;
        }
        ptr[i + 1] = static_cast<index_t>(idx.size());
    }
}

bool valid_compressed(size_t n_major, size_t n_minor, const std::vector<index_t>& ptr,
                      const std::vector<index_t>& idx, const std::vector<real_t>& val) {
    if (ptr.size() != n_major + 1 || ptr.front() != 0) return false;
    if (idx.size() != val.size() || ptr.back() != static_cast<index_t>(val.size())) return false;
    for (size_t i = 0; i < n_major; ++i) {
        if (ptr[i] > ptr[i + 1]) return false;
        index_t last = -1;
        for (index_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            const index_t j = idx[static_cast<size_t>(k)];
            if (j < 0 || static_cast<size_t>(j) >= n_minor || j <= last) return false;
            last = j;
        }
    }
    return true;
}

real_t lookup(const std::vector<index_t>& ptr, const std::vector<index_t>& idx, const std::vector<real_t>& val,
              index_t major, index_t minor) {
    for (index_t k = ptr[static_cast<size_t>(major)]; k < ptr[static_cast<size_t>(major) + 1]; ++k)
        if (idx[static_cast<size_t>(k)] == minor) return val[static_cast<size_t>(k)];
    return 0.0;
}

// y[major] = sum over the segment, in stored order
void gather_product(size_t n_major, const std::vector<index_t>& ptr, const std::vector<index_t>& idx,
                    const std::vector<real_t>& val, const real_t* x, real_t* y) {
    for (size_t i = 0; i < n_major; ++i) {
        real_t s = 0.0;
        for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
            s += val[static_cast<size_t>(k)] * x[static_cast<size_t>(idx[static_cast<size_t>(k)])];
        y[i] = s;
    }
}

// y[minor] += val * x[major], majors in ascending order
void scatter_product(size_t n_major, size_t n_minor, const std::vector<index_t>& ptr,
                     const std::vector<index_t>& idx, const std::vector<real_t>& val, const real_t* x, real_t* y) {
    std::fill(y, y + n_minor, 0.0);
    for (size_t i = 0; i < n_major; ++i) {
        const real_t xi = x[i];
        for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
            y[static_cast<size_t>(idx[static_cast<size_t>(k)])] += val[static_cast<size_t>(k)] * xi;
    }
}

}  // namespace

COOMatrix::COOMatrix(size_t rows, size_t cols) : rows_(rows), cols_(cols) {}

void COOMatrix::add_entry(index_t row, index_t col, real_t val) {
    if (row < 0 || (rows_ > 0 && static_cast<size_t>(row) >= rows_))
        throw std::out_of_range("COO row index out of range: " + std::to_string(row));
    if (col < 0 || (cols_ > 0 && static_cast<size_t>(col) >= cols_))
        throw std::out_of_range("COO col index out of range: " + std::to_string(col));
    entries_.push_back({row, col, val});
}

void COOMatrix::clear() {
    rows_ = cols_ = 0;
    entries_.clear();
}

COOMatrix COOMatrix::from_lp_model(const LPModel& model) {
    COOMatrix coo(model.num_constraints(), model.num_variables());
    const auto& cons = model.constraints();
    for (size_t i = 0; i < cons.size(); ++i)
        for (const auto& t : cons[i].terms) coo.add_entry(static_cast<index_t>(i), t.first, t.second);
    return coo;
}

CSRMatrix::CSRMatrix(size_t rows, size_t cols, std::vector<index_t> row_offsets, std::vector<index_t> col_indices,
                     std::vector<real_t> values)
    : rows_(rows), cols_(cols), ptr_(std::move(row_offsets)), idx_(std::move(col_indices)), val_(std::move(values)) {
    if (!is_valid()) throw std::invalid_argument("Attempted to construct invalid CSRMatrix.");
}

CSRMatrix CSRMatrix::from_coo(const COOMatrix& coo) {
    std::vector<index_t> p, i;
    std::vector<real_t> v;
    compress(coo.rows(), coo.cols(), coo.triplets(), true, p, i, v);
    return CSRMatrix(coo.rows(), coo.cols(), std::move(p), std::move(i), std::move(v));
}

CSRMatrix CSRMatrix::from_csc(const CSCMatrix& csc) {
    COOMatrix coo(csc.rows(), csc.cols());
    for (size_t j = 0; j < csc.cols(); ++j)
        for (index_t k = csc.col_offsets()[j]; k < csc.col_offsets()[j + 1]; ++k)
            coo.add_entry(csc.row_indices()[static_cast<size_t>(k)], static_cast<index_t>(j),
                          csc.values()[static_cast<size_t>(k)]);
    return from_coo(coo);
}

real_t CSRMatrix::get(index_t row, index_t col) const {
    if (row < 0 || static_cast<size_t>(row) >= rows_ || col < 0 || static_cast<size_t>(col) >= cols_)
        throw std::out_of_range("CSR index out of range.");
    return lookup(ptr_, idx_, val_, row, col);
}

std::vector<real_t> CSRMatrix::multiply(const std::vector<real_t>& x) const {
    if (x.size() != cols_)
        throw std::invalid_argument("SpMV dimension mismatch: x size " + std::to_string(x.size()) + " != cols " +
                                    std::to_string(cols_));
    std::vector<real_t> y(rows_, 0.0);
    multiply(x.data(), y.data());
    return y;
}

void CSRMatrix::multiply(const real_t* x, real_t* y) const { gather_product(rows_, ptr_, idx_, val_, x, y); }

std::vector<real_t> CSRMatrix::multiply_transpose(const std::vector<real_t>& x) const {
    if (x.size() != rows_)
        throw std::invalid_argument("Transpose SpMV dimension mismatch: x size " + std::to_string(x.size()) +
                                    " != rows " + std::to_string(rows_));
    std::vector<real_t> y(cols_, 0.0);
    multiply_transpose(x.data(), y.data());
    return y;
}

void CSRMatrix::multiply_transpose(const real_t* x, real_t* y) const {
    scatter_product(rows_, cols_, ptr_, idx_, val_, x, y);
}

bool CSRMatrix::is_valid() const { return valid_compressed(rows_, cols_, ptr_, idx_, val_); }

CSCMatrix::CSCMatrix(size_t rows, size_t cols, std::vector<index_t> col_offsets, std::vector<index_t> row_indices,
                     std::vector<real_t> values)
    : rows_(rows), cols_(cols), ptr_(std::move(col_offsets)), idx_(std::move(row_indices)), val_(std::move(values)) {
    if (!is_valid()) throw std::invalid_argument("Attempted to construct invalid CSCMatrix.");
}

CSCMatrix CSCMatrix::from_coo(const COOMatrix& coo) {
    std::vector<index_t> p, i;
    std::vector<real_t> v;
    compress(coo.cols(), coo.rows(), coo.triplets(), false, p, i, v);
    return CSCMatrix(coo.rows(), coo.cols(), std::move(p), std::move(i), std::move(v));
}

CSCMatrix CSCMatrix::from_csr(const CSRMatrix& csr) {
    COOMatrix coo(csr.rows(), csr.cols());
    for (size_t i = 0; i < csr.rows(); ++i)
        for (index_t k = csr.row_offsets()[i]; k < csr.row_offsets()[i + 1]; ++k)
            coo.add_entry(static_cast<index_t>(i), csr.col_indices()[static_cast<size_t>(k)],
                          csr.values()[static_cast<size_t>(k)]);
    return from_coo(coo);
}

real_t CSCMatrix::get(index_t row, index_t col) const {
    if (row < 0 || static_cast<size_t>(row) >= rows_ || col < 0 || static_cast<size_t>(col) >= cols_)
        throw std::out_of_range("CSC index out of range.");
    return lookup(ptr_, idx_, val_, col, row);
}

std::vector<real_t> CSCMatrix::multiply(const std::vector<real_t>& x) const {
    if (x.size() != cols_)
        throw std::invalid_argument("SpMV dimension mismatch: x size " + std::to_string(x.size()) + " != cols " +
                                    std::to_string(cols_));
    std::vector<real_t> y(rows_, 0.0);
    multiply(x.data(), y.data());
    return y;
}

void CSCMatrix::multiply(const real_t* x, real_t* y) const { scatter_product(cols_, rows_, ptr_, idx_, val_, x, y); }

std::vector<real_t> CSCMatrix::multiply_transpose(const std::vector<real_t>& x) const {
    if (x.size() != rows_)
        throw std::invalid_argument("Transpose SpMV dimension mismatch: x size " + std::to_string(x.size()) +
                                    " != rows " + std::to_string(rows_));
    std::vector<real_t> y(cols_, 0.0);
    multiply_transpose(x.data(), y.data());
    return y;
}

void CSCMatrix::multiply_transpose(const real_t* x, real_t* y) const {
    gather_product(cols_, ptr_, idx_, val_, x, y);
}

bool CSCMatrix::is_valid() const { return valid_compressed(cols_, rows_, ptr_, idx_, val_); }

}  // namespace bharatopt
