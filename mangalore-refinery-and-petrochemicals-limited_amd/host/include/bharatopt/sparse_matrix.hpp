// sparse_matrix.hpp -- host sparse formats with the public interface of the
// reference's include/bharatopt/sparse_matrix.hpp:16-119.  The GPU kernels take
// these arrays (row_offsets / col_indices / values) as they are.
#ifndef BHARATOPT_SPARSE_MATRIX_HPP
#define BHARATOPT_SPARSE_MATRIX_HPP

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include <bharatopt/config.hpp>
#include <bharatopt/lp_model.hpp>

namespace bharatopt {

class CSRMatrix;
class CSCMatrix;

struct COOTriplet {
    index_t row{0};
    index_t col{0};
    real_t val{0.0};
};

class COOMatrix {
public:
    COOMatrix(size_t rows = 0, size_t cols = 0);
    size_t rows() const { return rows_; }
    size_t cols() const { return cols_; }
    size_t nnz() const { return entries_.size(); }
    void add_entry(index_t row, index_t col, real_t val);
    void clear();
    const std::vector<COOTriplet>& triplets() const { return entries_; }
    static COOMatrix from_lp_model(const LPModel& model);

private:
    size_t rows_{0};
    size_t cols_{0};
    std::vector<COOTriplet> entries_;
};

class CSRMatrix {
public:
    CSRMatrix() = default;
    CSRMatrix(size_t rows, size_t cols, std::vector<index_t> row_offsets, std::vector<index_t> col_indices,
              std::vector<real_t> values);
    static CSRMatrix from_coo(const COOMatrix& coo);
    static CSRMatrix from_csc(const CSCMatrix& csc);

    size_t rows() const { return rows_; }
    size_t cols() const { return cols_; }
    size_t nnz() const { return val_.size(); }
    const std::vector<index_t>& row_offsets() const { return ptr_; }
    const std::vector<index_t>& col_indices() const { return idx_; }
    const std::vector<real_t>& values() const { return val_; }
    real_t get(index_t row, index_t col) const;

    std::vector<real_t> multiply(const std::vector<real_t>& x) const;
    void multiply(const real_t* x, real_t* y) const;
    std::vector<real_t> multiply_transpose(const std::vector<real_t>& x) const;
    void multiply_transpose(const real_t* x, real_t* y) const;
    bool is_valid() const;

private:
    size_t rows_{0};
    size_t cols_{0};
    std::vector<index_t> ptr_;
    std::vector<index_t> idx_;
    std::vector<real_t> val_;
};

class CSCMatrix {
public:
    CSCMatrix() = default;
    CSCMatrix(size_t rows, size_t cols, std::vector<index_t> col_offsets, std::vector<index_t> row_indices,
              std::vector<real_t> values);
    static CSCMatrix from_coo(const COOMatrix& coo);
    static CSCMatrix from_csr(const CSRMatrix& csr);

    size_t rows() const { return rows_; }
    size_t cols() const { return cols_; }
    size_t nnz() const { return val_.size(); }
    const std::vector<index_t>& col_offsets() const { return ptr_; }
    const std::vector<index_t>& row_indices() const { return idx_; }
    const std::vector<real_t>& values() const { return val_; }
    real_t get(index_t row, index_t col) const;

    std::vector<real_t> multiply(const std::vector<real_t>& x) const;
    void multiply(const real_t* x, real_t* y) const;
    std::vector<real_t> multiply_transpose(const std::vector<real_t>& x) const;
    void multiply_transpose(const real_t* x, real_t* y) const;
    bool is_valid() const;

private:
    size_t rows_{0};
    size_t cols_{0};
    std::vector<index_t> ptr_;
    std::vector<index_t> idx_;
    std::vector<real_t> val_;
};

}  // namespace bharatopt

#endif
