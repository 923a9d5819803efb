// common.hpp -- enums, small utilities and the assertion macros of the drop-in API.
// Names and enumerator order follow common/lss_enumerations.hpp:15-129, common/lss_utility.hpp:17-196
// and common/lss_macros.hpp:37-53 of the reference, so user code written against those headers
// compiles unchanged.  The assertion macros keep the reference's behaviour: print and terminate.
#pragma once

#include <cstdlib>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <tuple>
#include <vector>

#define LSS_ASSERT(condition, message)                                                                     \
    {                                                                                                      \
        do {                                                                                               \
            if (!(condition)) {                                                                            \
                std::cerr << "Assertion `" #condition "` failed in " << __FILE__ << " line " << __LINE__   \
                          << ": " << message << std::endl;                                                 \
                std::terminate();                                                                          \
            }                                                                                              \
        } while (false);                                                                                   \
    }

#define LSS_VERIFY(variable, message)                                                                      \
    {                                                                                                      \
        LSS_ASSERT(variable, message);                                                                     \
    }

namespace lss_enumerations {

enum class by_enum { Row, Column };
enum class factorization_enum { QRMethod, LUMethod, CholeskyMethod, None };
enum class splitting_method_enum { DouglasRachford, CraigSneyd, ModifiedCraigSneyd, HundsdorferVerwer };
enum class tridiagonal_method_enum { CUDASolver, DoubleSweepSolver, SORSolver, ThomasLUSolver };
enum class dimension_enum { One, Two };
enum class grid_enum { Uniform, Nonuniform };
enum class traverse_direction_enum { Forward, Backward };
enum class memory_space_enum { Host, Device };
enum class flat_matrix_sort_enum { RowMajor, ColumnMajor };
enum class implicit_pde_schemes_enum { Euler, Theta_30, CrankNicolson, Theta_80 };
enum class explicit_pde_schemes_enum { Euler, ADEBarakatClark, ADESaulyev };
enum class dirichlet_side_enum { Left, Up, Right, Bottom };

}  // namespace lss_enumerations

namespace lss_utility {

template <typename T> static T const pi() { return static_cast<T>(3.14159265358979); }
template <typename T> static constexpr T NaN() { return std::numeric_limits<T>::quiet_NaN(); }
template <typename T> using sptr_t = std::shared_ptr<T>;
template <typename T> using uptr_t = std::unique_ptr<T>;

// closed interval [lower, upper] (lss_utility.hpp:130-188)
template <typename T> class range {
    T l_, u_;

  public:
    explicit range(T lower, T upper) : l_{lower}, u_{upper} {}
    explicit range() : range(T{}, T{}) {}
    T lower() const { return l_; }
    T upper() const { return u_; }
    T spread() const { return (u_ - l_); }
    T mid_point() const { return T(0.5) * (l_ + u_); }
};
template <typename T> using range_ptr = sptr_t<range<T>>;

}  // namespace lss_utility
