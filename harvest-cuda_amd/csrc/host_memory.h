// Process-wide host scratch arena (counterpart of the reference's
// src/host_memory.h:1-55 and HostLease src/engine.cpp:29-41).  One fixed
// block is lent to one call at a time (try_lock); concurrent callers and
// requests that do not fit use ordinary heap storage, so the arena never
// grows and never blocks.
#pragma once
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>

namespace harvest {

constexpr size_t kHostArenaBytes = size_t(20) << 20;

struct HostArena {
  unsigned char* next = nullptr;
  size_t remaining = 0;
  void* take(size_t bytes, size_t alignment) {
    if (!next) return nullptr;
    uintptr_t p = reinterpret_cast<uintptr_t>(next);
    size_t pad = (alignment - p % alignment) % alignment;
    if (pad > remaining || bytes > remaining - pad) return nullptr;
    void* out = next + pad;
    next += pad + bytes;
    remaining -= pad + bytes;
    return out;
  }
};

class HostCache {
 public:
  static HostCache& instance() {
    static HostCache cache;
    return cache;
  }
  void initialize() {
    std::lock_guard<std::mutex> lock(mutex_);
    if (!storage_) storage_.reset(new unsigned char[kHostArenaBytes]);
  }
  void clear() {
    std::lock_guard<std::mutex> lock(mutex_);
    storage_.reset();
  }
  bool allocated() {
    std::lock_guard<std::mutex> lock(mutex_);
    return storage_ != nullptr;
  }

  // Borrow the arena for one call if nobody else holds it.
  class Lease {
   public:
    Lease() : lock_(instance().mutex_, std::try_to_lock) {
      if (lock_.owns_lock()) {
        auto& c = instance();
        if (!c.storage_) c.storage_.reset(new unsigned char[kHostArenaBytes]);
        arena_.next = c.storage_.get();
        arena_.remaining = kHostArenaBytes;
      }
    }
    HostArena* arena() { return &arena_; }
   private:
    std::unique_lock<std::mutex> lock_;
    HostArena arena_;
  };

 private:
  std::mutex mutex_;
  std::unique_ptr<unsigned char[]> storage_;
};

// Uninitialised scratch array: arena storage when it fits, heap otherwise.
template <class T> class HostArray {
  static_assert(std::is_arithmetic<T>::value, "scratch holds numbers only");
 public:
  explicit HostArray(HostArena* arena = nullptr) : arena_(arena) {}
  void allocate(size_t n) {
    owned_.reset();
    data_ = nullptr;
    size_ = 0;
    if (!n) return;
    if (n > std::numeric_limits<size_t>::max() / sizeof(T)) throw std::bad_array_new_length();
    if (arena_) data_ = static_cast<T*>(arena_->take(n * sizeof(T), alignof(T)));
    if (!data_) {
      owned_.reset(new T[n]);
      data_ = owned_.get();
    }
    size_ = n;
  }
  T* data() { return data_; }
  const T* data() const { return data_; }
  size_t size() const { return size_; }
  T& operator[](size_t i) { return data_[i]; }
  const T& operator[](size_t i) const { return data_[i]; }
  bool in_arena() const { return data_ && !owned_; }

 private:
  HostArena* arena_;
  std::unique_ptr<T[]> owned_;
  T* data_ = nullptr;
  size_t size_ = 0;
};

}  // namespace harvest
