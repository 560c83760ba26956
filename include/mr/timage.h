// kipl::base::TImage<T,N> -- the image container of MuhRec's module API
// (core/kipl/kipl/include/base/timage.h:18-162), restated with the same
// semantics: copies are shallow and share one reference-counted buffer,
// Clone() detaches, Size(0) is the fastest axis, GetLinePtr(y,z) points at the
// start of row y of plane z.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

namespace kipl {
namespace base {

template <typename T, size_t N>
class TImage {
public:
    TImage() { std::fill(dims_, dims_ + N, size_t(0)); }
    explicit TImage(const size_t *dims) { Resize(dims); }

    // Reallocates (values are zeroed) unless the shape is unchanged.
    void Resize(const size_t *dims)
    {
        size_t n = 1;
        for (size_t i = 0; i < N; ++i) n *= dims[i];
        if (buf_ && n == Size() && std::equal(dims, dims + N, dims_)) return;
        std::copy(dims, dims + N, dims_);
        buf_ = std::make_shared<std::vector<T>>(n, T(0));
    }

    size_t Size() const
    {
        if (!buf_) return 0;
        return buf_->size();
    }
    size_t Size(size_t i) const { return i < N ? dims_[i] : 1; }
    const size_t *Dims() const { return dims_; }

    T *GetDataPtr() { return buf_ ? buf_->data() : nullptr; }
    const T *GetDataPtr() const { return buf_ ? buf_->data() : nullptr; }
    T *GetLinePtr(size_t y, size_t z = 0)
    {
        return GetDataPtr() + y * Size(0) + (N > 2 ? z * Size(0) * Size(1) : 0);
    }
    const T *GetLinePtr(size_t y, size_t z = 0) const
    {
        return GetDataPtr() + y * Size(0) + (N > 2 ? z * Size(0) * Size(1) : 0);
    }

    // Detaches from other owners of the buffer (deep copy).
    void Clone()
    {
        if (buf_ && buf_.use_count() > 1) buf_ = std::make_shared<std::vector<T>>(*buf_);
    }

    T &operator[](size_t i) { return (*buf_)[i]; }
    const T &operator[](size_t i) const { return (*buf_)[i]; }

    TImage &operator=(T v)
    {
        if (buf_) std::fill(buf_->begin(), buf_->end(), v);
        return *this;
    }
    TImage &operator*=(T v)
    {
        if (buf_)
            for (auto &x : *buf_) x *= v;
        return *this;
    }

private:
    size_t dims_[N];
    std::shared_ptr<std::vector<T>> buf_;
};

}  // namespace base
}  // namespace kipl
