#ifndef IRLBA_HIP_TYPES_HPP
#define IRLBA_HIP_TYPES_HPP

// Column-major dense containers used where the reference's API takes Eigen
// types (Eigen::VectorXd / Eigen::MatrixXd).  Any type with the same members
// (size/rows/cols, data, resize, element access) can be used in their place.

#include <cstdint>
#include <vector>

namespace irlba {

typedef std::int64_t Index;

class VectorXd {
public:
    typedef double Scalar;
    VectorXd() = default;
    explicit VectorXd(Index n) : my_data(static_cast<size_t>(n), 0.0) {}
    Index size() const { return static_cast<Index>(my_data.size()); }
    Index rows() const { return size(); }
    void resize(Index n) { my_data.resize(static_cast<size_t>(n)); }
    double* data() { return my_data.data(); }
    const double* data() const { return my_data.data(); }
    double& operator[](Index i) { return my_data[static_cast<size_t>(i)]; }
    double operator[](Index i) const { return my_data[static_cast<size_t>(i)]; }
    double& coeffRef(Index i) { return (*this)[i]; }
    double coeff(Index i) const { return (*this)[i]; }
    auto begin() { return my_data.begin(); }
    auto end() { return my_data.end(); }
    auto begin() const { return my_data.begin(); }
    auto end() const { return my_data.end(); }

private:
    std::vector<double> my_data;
};

class MatrixXd {
public:
    typedef double Scalar;
    MatrixXd() = default;
    MatrixXd(Index nr, Index nc) : my_nr(nr), my_nc(nc), my_data(static_cast<size_t>(nr * nc), 0.0) {}
    Index rows() const { return my_nr; }
    Index cols() const { return my_nc; }
    Index size() const { return my_nr * my_nc; }
    void resize(Index nr, Index nc) {
        my_nr = nr;
        my_nc = nc;
        my_data.resize(static_cast<size_t>(nr * nc));
    }
    double* data() { return my_data.data(); }
    const double* data() const { return my_data.data(); }
    double& operator()(Index r, Index c) { return my_data[static_cast<size_t>(r + c * my_nr)]; }
    double operator()(Index r, Index c) const { return my_data[static_cast<size_t>(r + c * my_nr)]; }
    double* col(Index c) { return my_data.data() + c * my_nr; }
    const double* col(Index c) const { return my_data.data() + c * my_nr; }

private:
    Index my_nr = 0, my_nc = 0;
    std::vector<double> my_data;
};

}  // namespace irlba

#endif
