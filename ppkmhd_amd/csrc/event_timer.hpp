// Per-stage GPU timers of a solver handle: hipEvent pairs recorded on the handle's stream
// around the launches of the selected stages, resolved into totals when read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace ppk {

// hipEvent pairs per stage; events are pooled and resolved when the times are read
class EventTimer : public StageTimer {
 public:
  explicit EventTimer(hipStream_t s) : stream_(s) {}
  ~EventTimer() override {
    for (auto e : pool_) (void)hipEventDestroy(e);
  }
  void set_mask(uint32_t m) {
    mask_ = m;
    clear();
  }
  uint32_t mask() const { return mask_; }
  void mark(int stage, bool start) override {
    if (!(mask_ & (1u << stage))) return;
    hipEvent_t e = take();
    if (!e) return;
    (void)hipEventRecord(e, stream_);
    if (start) {
      open_[stage] = e;
    } else if (open_[stage]) {
      recs_.push_back({stage, open_[stage], e});
      open_[stage] = nullptr;
    }
  }
  // resolve all recorded pairs into totals; recycles the events
  hipError_t resolve() {
    hipError_t err = hipStreamSynchronize(stream_);
    if (err != hipSuccess) return err;
    for (auto &r : recs_) {
      float ms = 0.f;
      err = hipEventElapsedTime(&ms, r.a, r.b);
      if (err != hipSuccess) return err;
      total_ms_[r.stage] += ms;
      count_[r.stage] += 1;
      free_.push_back(r.a);
      free_.push_back(r.b);
    }
    recs_.clear();
    return hipSuccess;
  }
  double total(int st) const { return total_ms_[st]; }
  long long count(int st) const { return count_[st]; }

 private:
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  hipEvent_t take() {
    if (!free_.empty()) {
      hipEvent_t e = free_.back();
      free_.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pool_.push_back(e);
    return e;
  }
  void clear() {
    for (auto &r : recs_) {
      free_.push_back(r.a);
      free_.push_back(r.b);
    }
    recs_.clear();
    for (int i = 0; i < ST_COUNT; ++i) {
      total_ms_[i] = 0.0;
      count_[i] = 0;
      open_[i] = nullptr;
    }
  }
  hipStream_t stream_;
  uint32_t mask_ = 0;
  std::vector<hipEvent_t> pool_, free_;
  std::vector<Rec> recs_;
  hipEvent_t open_[ST_COUNT] = {};
  double total_ms_[ST_COUNT] = {};
  long long count_[ST_COUNT] = {};
};

}  // namespace ppk
