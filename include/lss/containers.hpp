// containers.hpp -- container_2d of the drop-in API (Row- or Column-ordered matrix of fp values).
// Interface as containers/lss_container_2d.hpp:22-390 of the reference (element and whole-row /
// whole-column access, conversion between the two orders, flat row-wise data()).  Storage here is
// one contiguous vector, which is also what the GPU path copies.
#pragma once

#include "common.hpp"

namespace lss_containers {

using lss_enumerations::by_enum;
using lss_utility::sptr_t;

template <by_enum by, typename fp_type, template <typename, typename> typename container, typename allocator>
class container_2d;

template <typename fp_type, template <typename, typename> typename container, typename allocator>
class container_2d<by_enum::Row, fp_type, container, allocator> {
    std::size_t rows_ = 0, columns_ = 0;
    std::vector<fp_type> data_;

  public:
    typedef fp_type value_type;
    typedef container<fp_type, allocator> element_type;

    explicit container_2d(std::size_t rows, std::size_t columns)
        : rows_{rows}, columns_{columns}, data_(rows * columns)
    {
    }
    explicit container_2d(std::size_t rows, std::size_t columns, fp_type value)
        : rows_{rows}, columns_{columns}, data_(rows * columns, value)
    {
    }
    container_2d(container_2d<by_enum::Column, fp_type, container, allocator> const &copy)
        : rows_{copy.rows()}, columns_{copy.columns()}, data_(copy.rows() * copy.columns())
    {
        for (std::size_t r = 0; r < rows_; ++r)
            for (std::size_t c = 0; c < columns_; ++c)
                data_[r * columns_ + c] = copy(r, c);
    }
    container_2d(container_2d const &) = default;
    container_2d(container_2d &&) noexcept = default;
    container_2d &operator=(container_2d const &) = default;
    container_2d &operator=(container_2d &&) noexcept = default;
    container_2d &operator=(container_2d<by_enum::Column, fp_type, container, allocator> const &copy)
    {
        *this = container_2d(copy);
        return *this;
    }

    std::size_t rows() const { return rows_; }
    std::size_t columns() const { return columns_; }
    std::size_t total_size() const { return rows_ * columns_; }

    fp_type operator()(std::size_t row_idx, std::size_t col_idx) const
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        return data_[row_idx * columns_ + col_idx];
    }
    container<fp_type, allocator> operator()(std::size_t row_idx) const
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        return container<fp_type, allocator>(data_.begin() + row_idx * columns_,
                                             data_.begin() + (row_idx + 1) * columns_);
    }
    fp_type at(std::size_t row_idx, std::size_t col_idx) const { return operator()(row_idx, col_idx); }
    container<fp_type, allocator> at(std::size_t row_idx) const { return operator()(row_idx); }
    void operator()(std::size_t row_idx, container<fp_type, allocator> const &cont)
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(cont.size() == columns_, "Outside of column range");
        std::copy(cont.begin(), cont.end(), data_.begin() + row_idx * columns_);
    }
    void operator()(std::size_t row_idx, std::size_t col_idx, fp_type value)
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        data_[row_idx * columns_ + col_idx] = value;
    }
    std::vector<fp_type, allocator> const data() const
    {
        return std::vector<fp_type, allocator>(data_.begin(), data_.end());
    }
    void from_data(container<fp_type, allocator> const &data)
    {
        LSS_ASSERT(rows_ * columns_ == data.size(), "Data are not of the correct size");
        std::copy(data.begin(), data.end(), data_.begin());
    }
    // contiguous row-major storage (drop-in extension used by the GPU path)
    fp_type *raw() { return data_.data(); }
    const fp_type *raw() const { return data_.data(); }
};

template <typename fp_type, template <typename, typename> typename container, typename allocator>
class container_2d<by_enum::Column, fp_type, container, allocator> {
    std::size_t rows_ = 0, columns_ = 0;
    std::vector<fp_type> data_;  // column-major

  public:
    typedef fp_type value_type;
    typedef container<fp_type, allocator> element_type;

    explicit container_2d(std::size_t rows, std::size_t columns)
        : rows_{rows}, columns_{columns}, data_(rows * columns)
    {
    }
    explicit container_2d(std::size_t rows, std::size_t columns, fp_type value)
        : rows_{rows}, columns_{columns}, data_(rows * columns, value)
    {
    }
    container_2d(container_2d<by_enum::Row, fp_type, container, allocator> const &copy)
        : rows_{copy.rows()}, columns_{copy.columns()}, data_(copy.rows() * copy.columns())
    {
        for (std::size_t r = 0; r < rows_; ++r)
            for (std::size_t c = 0; c < columns_; ++c)
                data_[c * rows_ + r] = copy(r, c);
    }
    container_2d(container_2d const &) = default;
    container_2d(container_2d &&) noexcept = default;
    container_2d &operator=(container_2d const &) = default;
    container_2d &operator=(container_2d &&) noexcept = default;

    std::size_t rows() const { return rows_; }
    std::size_t columns() const { return columns_; }
    std::size_t total_size() const { return rows_ * columns_; }

    fp_type operator()(std::size_t row_idx, std::size_t col_idx) const
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        return data_[col_idx * rows_ + row_idx];
    }
    container<fp_type, allocator> operator()(std::size_t col_idx) const
    {
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        return container<fp_type, allocator>(data_.begin() + col_idx * rows_, data_.begin() + (col_idx + 1) * rows_);
    }
    void operator()(std::size_t col_idx, container<fp_type, allocator> const &cont)
    {
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        LSS_ASSERT(cont.size() == rows_, "Outside of row range");
        std::copy(cont.begin(), cont.end(), data_.begin() + col_idx * rows_);
    }
    void operator()(std::size_t row_idx, std::size_t col_idx, fp_type value)
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        data_[col_idx * rows_ + row_idx] = value;
    }
    // flat column-wise data, as the reference's Column container returns
    std::vector<fp_type, allocator> const data() const
    {
        return std::vector<fp_type, allocator>(data_.begin(), data_.end());
    }
};

template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
using rcontainer_2d = container_2d<by_enum::Row, fp_type, container, allocator>;
template <typename fp_type, template <typename, typename> typename container = std::vector,
          typename allocator = std::allocator<fp_type>>
using ccontainer_2d = container_2d<by_enum::Column, fp_type, container, allocator>;

// container_3d: layers of container_2d (containers/lss_container_3d.hpp:19-200 of the reference): element,
// whole-layer access and replacement.
template <by_enum by, typename fp_type, template <typename, typename> typename container, typename allocator>
class container_3d {
    std::size_t rows_ = 0, columns_ = 0, layers_ = 0;
    std::vector<container_2d<by, fp_type, container, allocator>> rdata_;

  public:
    typedef fp_type value_type;
    typedef container_2d<by, fp_type, container, allocator> element_type;

    explicit container_3d(std::size_t rows, std::size_t columns, std::size_t layers)
        : rows_{rows}, columns_{columns}, layers_{layers}
    {
        for (std::size_t l = 0; l < layers; ++l)
            rdata_.emplace_back(element_type(rows, columns));
    }
    explicit container_3d(std::size_t rows, std::size_t columns, std::size_t layers, fp_type value)
        : rows_{rows}, columns_{columns}, layers_{layers}
    {
        for (std::size_t l = 0; l < layers; ++l)
            rdata_.emplace_back(element_type(rows, columns, value));
    }
    std::size_t rows() const { return rows_; }
    std::size_t columns() const { return columns_; }
    std::size_t layers() const { return layers_; }
    std::size_t total_size() const { return rows_ * columns_ * layers_; }

    fp_type operator()(std::size_t row_idx, std::size_t col_idx, std::size_t lay_idx) const
    {
        LSS_ASSERT(row_idx < rows_, "Outside of row range");
        LSS_ASSERT(col_idx < columns_, "Outside of column range");
        LSS_ASSERT(lay_idx < layers_, "Outside of layer range");
        return rdata_[lay_idx](row_idx, col_idx);
    }
    element_type operator()(std::size_t lay_idx) const
    {
        LSS_ASSERT(lay_idx < layers_, "Outside of layer range");
        return rdata_[lay_idx];
    }
    fp_type at(std::size_t row_idx, std::size_t col_idx, std::size_t lay_idx) const
    {
        return operator()(row_idx, col_idx, lay_idx);
    }
    element_type at(std::size_t lay_idx) const { return operator()(lay_idx); }
    void operator()(std::size_t lay_idx, element_type const &container2d)
    {
        LSS_ASSERT(lay_idx < layers_, "Outside of layer range");
        LSS_ASSERT(container2d.rows() == rows_, "Outside of row range");
        LSS_ASSERT(container2d.columns() == columns_, "Outside of column range");
        rdata_[lay_idx] = container2d;
    }
    void operator()(std::size_t row_idx, std::size_t col_idx, std::size_t lay_idx, fp_type value)
    {
        LSS_ASSERT(lay_idx < layers_, "Outside of layer range");
        rdata_[lay_idx](row_idx, col_idx, value);
    }
    // layer lay_idx itself (drop-in extension used by the GPU path)
    element_type &layer(std::size_t lay_idx) { return rdata_[lay_idx]; }
};

}  // namespace lss_containers
