#ifndef IRLBA_HIP_OPTIONS_HPP
#define IRLBA_HIP_OPTIONS_HPP

// irlba::Options (include/irlba/Options.hpp:19-83 of the reference): the same
// fields, types and defaults.

#include <optional>
#include <random>

#include "types.hpp"

namespace irlba {

template <class EigenVector_ = VectorXd>
struct Options {
    double invariant_subspace_tolerance = -1;
    double convergence_tolerance = 1e-5;
    double singular_value_ratio_tolerance = -1;
    int extra_work = 7;
    int max_iterations = 1000;
    bool exact_for_small_matrix = true;
    bool exact_for_large_number = true;
    bool cap_number = false;
    typename std::mt19937_64::result_type seed = std::mt19937_64::default_seed;
    std::optional<EigenVector_> initial;
};

}  // namespace irlba

#endif
