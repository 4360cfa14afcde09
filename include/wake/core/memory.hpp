// memory.hpp -- 64-byte aligned, zero-initialised host array (the
// reference's AlignedBuffer<T>, include/wake/core/memory.hpp:53-159: size(),
// resize(), data(), operator[], iteration; copies are deep).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <new>

namespace wake {

inline constexpr std::size_t SIMD_ALIGNMENT = 64;

template <class T>
class AlignedBuffer {
 public:
  AlignedBuffer() = default;
  explicit AlignedBuffer(std::size_t n) { resize(n); }
  AlignedBuffer(const AlignedBuffer& o) { *this = o; }
  AlignedBuffer(AlignedBuffer&& o) noexcept { swap(o); }
  ~AlignedBuffer() { release(); }

  AlignedBuffer& operator=(const AlignedBuffer& o) {
    if (this != &o) {
      resize(o.size_);
      if (size_) std::memcpy(data_, o.data_, size_ * sizeof(T));
    }
    return *this;
  }
  AlignedBuffer& operator=(AlignedBuffer&& o) noexcept {
    swap(o);
    return *this;
  }

  // New elements are zero; existing ones are kept up to the new size.
  void resize(std::size_t n) {
    if (n == size_) return;
    T* p = nullptr;
    if (n) {
      p = static_cast<T*>(::operator new[](n * sizeof(T), std::align_val_t(SIMD_ALIGNMENT)));
      std::memset(static_cast<void*>(p), 0, n * sizeof(T));
      if (data_) std::memcpy(static_cast<void*>(p), data_, std::min(n, size_) * sizeof(T));
    }
    release();
    data_ = p;
    size_ = n;
  }
  void fill(const T& v) { std::fill(data_, data_ + size_, v); }

  std::size_t size() const noexcept { return size_; }
  bool empty() const noexcept { return size_ == 0; }
  T* data() noexcept { return data_; }
  const T* data() const noexcept { return data_; }
  T& operator[](std::size_t i) noexcept { return data_[i]; }
  const T& operator[](std::size_t i) const noexcept { return data_[i]; }
  T* begin() noexcept { return data_; }
  T* end() noexcept { return data_ + size_; }
  const T* begin() const noexcept { return data_; }
  const T* end() const noexcept { return data_ + size_; }

 private:
  void release() {
    if (data_) ::operator delete[](data_, std::align_val_t(SIMD_ALIGNMENT));
    data_ = nullptr;
    size_ = 0;
  }
  void swap(AlignedBuffer& o) noexcept {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
  }
  T* data_ = nullptr;
  std::size_t size_ = 0;
};

}  // namespace wake
