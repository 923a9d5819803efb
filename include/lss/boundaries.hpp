// boundaries.hpp -- boundary condition types of the drop-in API.
// Same class names, template parameters and value conventions as the reference's boundaries/
// (lss_boundary.hpp:11-40, lss_dirichlet_boundary.hpp:20-41, lss_neumann_boundary.hpp:21-40,
// lss_robin_boundary.hpp:21-45): a Neumann value g means u' = -g, a Robin pair (l, g) means
// u' + l u + g = 0 (the line solvers fold them that way).
#pragma once

#include "common.hpp"

namespace lss_boundary {

using lss_utility::sptr_t;

template <typename fp_type, typename... fp_space_types> class boundary {
  protected:
    std::function<fp_type(fp_type, fp_space_types...)> linear_;
    std::function<fp_type(fp_type, fp_space_types...)> const_;

  public:
    typedef fp_type value_type;
    explicit boundary() = delete;
    explicit boundary(const std::function<fp_type(fp_type, fp_space_types...)> &linear,
                      const std::function<fp_type(fp_type, fp_space_types...)> &constant)
        : linear_{linear}, const_{constant}
    {
    }
    virtual ~boundary() {}
    virtual const fp_type value(fp_type time, fp_space_types... space_args) const = 0;
};

template <typename fp_type, typename... fp_space_types>
class dirichlet_boundary : public boundary<fp_type, fp_space_types...> {
  public:
    dirichlet_boundary(const std::function<fp_type(fp_type, fp_space_types...)> &value)
        : boundary<fp_type, fp_space_types...>(nullptr, value)
    {
    }
    const fp_type value(fp_type time, fp_space_types... space_args) const override
    {
        return this->const_(time, space_args...);
    }
};

template <typename fp_type, typename... fp_space_types>
class neumann_boundary : public boundary<fp_type, fp_space_types...> {
  public:
    neumann_boundary(const std::function<fp_type(fp_type, fp_space_types...)> &value)
        : boundary<fp_type, fp_space_types...>(nullptr, value)
    {
    }
    const fp_type value(fp_type time, fp_space_types... space_args) const override
    {
        return this->const_(time, space_args...);
    }
};

template <typename fp_type, typename... fp_space_types>
class robin_boundary : public boundary<fp_type, fp_space_types...> {
  public:
    robin_boundary(const std::function<fp_type(fp_type, fp_space_types...)> &linear_value,
                   const std::function<fp_type(fp_type, fp_space_types...)> &value)
        : boundary<fp_type, fp_space_types...>(linear_value, value)
    {
    }
    const fp_type linear_value(fp_type time, fp_space_types... space_args) const
    {
        return this->linear_(time, space_args...);
    }
    const fp_type value(fp_type time, fp_space_types... space_args) const override
    {
        return this->const_(time, space_args...);
    }
};

template <typename fp_type, typename... fp_space_types>
using boundary_ptr = sptr_t<boundary<fp_type, fp_space_types...>>;
template <typename fp_type> using boundary_1d_ptr = sptr_t<boundary<fp_type>>;
template <typename fp_type> using boundary_2d_ptr = sptr_t<boundary<fp_type, fp_type>>;
template <typename fp_type, typename... fp_space_types>
using boundary_pair = std::pair<boundary_ptr<fp_type, fp_space_types...>, boundary_ptr<fp_type, fp_space_types...>>;
template <typename fp_type> using boundary_1d_pair = std::pair<boundary_1d_ptr<fp_type>, boundary_1d_ptr<fp_type>>;
template <typename fp_type> using boundary_2d_pair = std::pair<boundary_2d_ptr<fp_type>, boundary_2d_ptr<fp_type>>;

template <typename fp_type> using dirichlet_boundary_1d = dirichlet_boundary<fp_type>;
template <typename fp_type> using dirichlet_boundary_2d = dirichlet_boundary<fp_type, fp_type>;
template <typename fp_type> using dirichlet_boundary_1d_ptr = sptr_t<dirichlet_boundary_1d<fp_type>>;
template <typename fp_type> using dirichlet_boundary_2d_ptr = sptr_t<dirichlet_boundary_2d<fp_type>>;
template <typename fp_type> using neumann_boundary_1d = neumann_boundary<fp_type>;
template <typename fp_type> using neumann_boundary_2d = neumann_boundary<fp_type, fp_type>;
template <typename fp_type> using neumann_boundary_1d_ptr = sptr_t<neumann_boundary_1d<fp_type>>;
template <typename fp_type> using neumann_boundary_2d_ptr = sptr_t<neumann_boundary_2d<fp_type>>;
template <typename fp_type> using robin_boundary_1d = robin_boundary<fp_type>;
template <typename fp_type> using robin_boundary_2d = robin_boundary<fp_type, fp_type>;
template <typename fp_type> using robin_boundary_1d_ptr = sptr_t<robin_boundary_1d<fp_type>>;
template <typename fp_type> using robin_boundary_2d_ptr = sptr_t<robin_boundary_2d<fp_type>>;

}  // namespace lss_boundary
