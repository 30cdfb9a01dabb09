// Device-time accounting of the main kernel of each stage (pbagpu_timing_* in include/pbagpu.h).
// When enabled, a stage records one event before and one after its kernel launch on the
// caller's stream and adds the elapsed time once the stage has synchronised that stream.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/pbagpu.h"

namespace pbagpu {

bool timing_enabled();
void timing_add(int stage, double ms);
void timing_add_units(int stage, double units);

class KernelTimer {
  public:
    explicit KernelTimer(int stage) : stage_(stage), on_(timing_enabled()) {
        if (on_ && (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess)) on_ = false;
    }
    void start(hipStream_t s) { if (on_ && hipEventRecord(a_, s) != hipSuccess) on_ = false; }
    void stop(hipStream_t s) {
        if (on_ && hipEventRecord(b_, s) == hipSuccess) armed_ = true;
    }
    // call after the stream has been synchronised
    void collect() {
        float ms = 0.f;
        if (on_ && armed_ && hipEventElapsedTime(&ms, a_, b_) == hipSuccess) timing_add(stage_, (double)ms);
        armed_ = false;
    }
    ~KernelTimer() {
        if (a_) (void)hipEventDestroy(a_);
        if (b_) (void)hipEventDestroy(b_);
    }

  private:
    int stage_;
    bool on_;
    bool armed_ = false;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace pbagpu
