// One backend workspace per GPU, lent to one call at a time (counterpart of
// the reference's src/workspace_cache.h:10-71).  clear() bumps a generation
// and drops the cached slots; a workspace in use survives until its lease
// ends, creation racing a clear() is never cached, and destruction happens
// outside the cache lock.  Creation and destruction go through a replaceable
// function table so that tests can inject faults on the CPU.
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>

namespace harvest {

template <class Workspace> class WorkspaceCache {
 public:
  using Create = std::function<Workspace*(int)>;
  using Destroy = std::function<void(Workspace*)>;

  WorkspaceCache(Create create, Destroy destroy)
      : create_(std::move(create)), destroy_(std::move(destroy)) {}
  WorkspaceCache(const WorkspaceCache&) = delete;
  WorkspaceCache& operator=(const WorkspaceCache&) = delete;

 private:
  struct Slot {
    explicit Slot(Destroy d) : destroy(std::move(d)) {}
    ~Slot() {
      if (workspace) destroy(workspace);
    }
    std::mutex mutex;
    Workspace* workspace = nullptr;
    Destroy destroy;
  };

 public:
  class Lease {
   public:
    Lease(WorkspaceCache& cache, int device) {
      if (device < 0) throw std::invalid_argument("device must be a nonnegative visible GPU index");
      uint64_t generation;
      {
        std::lock_guard<std::mutex> lock(cache.mutex_);
        generation = cache.generation_;
        auto it = cache.slots_.find(device);
        if (it != cache.slots_.end()) slot_ = it->second;
      }
      if (!slot_) {
        auto created = std::make_shared<Slot>(cache.destroy_);
        created->workspace = cache.create_(device);
        std::lock_guard<std::mutex> lock(cache.mutex_);
        // A clear() during creation means the new workspace is used once and
        // then discarded; a concurrent creator that won keeps its own.
        if (generation == cache.generation_)
          slot_ = cache.slots_.emplace(device, created).first->second;
        else
          slot_ = created;
      }
      lock_ = std::unique_lock<std::mutex>(slot_->mutex);
    }
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    ~Lease() {
      // Unlock before the last owner (possibly this lease) destroys the slot.
      if (lock_.owns_lock()) lock_.unlock();
    }
    Workspace* get() const noexcept { return slot_->workspace; }

   private:
    std::shared_ptr<Slot> slot_;
    std::unique_lock<std::mutex> lock_;
  };

  void clear() {
    std::map<int, std::shared_ptr<Slot>> retired;
    {
      std::lock_guard<std::mutex> lock(mutex_);
      ++generation_;
      retired.swap(slots_);
    }
    // `retired` is destroyed here, outside the lock; slots still leased are
    // destroyed when their last lease ends.
  }

  size_t size() {
    std::lock_guard<std::mutex> lock(mutex_);
    return slots_.size();
  }

 private:
  Create create_;
  Destroy destroy_;
  std::mutex mutex_;
  std::map<int, std::shared_ptr<Slot>> slots_;
  uint64_t generation_ = 0;
};

}  // namespace harvest
