// Shared helpers of the porescale HIP library: error reporting, HIP checks,
// wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "porescale.h"

namespace psx {

void set_error(const char *fmt, ...);

// Ensures a HIP device exists and is current; returns PSX_OK or PSX_EHIP.
int require_device();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Device buffers and events that live for one library call, released on
// every way out of it (HIP errors return early).  hipFree waits for the
// device, so work still queued on them completes first.
class CallScope {
  public:
    CallScope() = default;
    CallScope(const CallScope &) = delete;
    CallScope &operator=(const CallScope &) = delete;
    ~CallScope() {
        for (hipEvent_t e : events_) (void)hipEventDestroy(e);
        for (void *b : bufs_) (void)hipFree(b);
    }
    void buffer(void *b) {
        if (b) bufs_.push_back(b);
    }
    void event(hipEvent_t e) {
        if (e) events_.push_back(e);
    }

  private:
    std::vector<void *> bufs_;
    std::vector<hipEvent_t> events_;
};

// Owning pointer to device memory (hipFree on reset or scope exit).
struct DeviceFree {
    void operator()(void *b) const {
        if (b) (void)hipFree(b);
    }
};
template <class T>
using DeviceBuf = std::unique_ptr<T, DeviceFree>;

// Owns a half-built handle inside a create call: an early return (a failed
// HIP call, an argument error) or a host exception leaving the call destroys
// it with the handle's own destroy function; release() hands it over on
// success.  The destroy functions accept partly initialised handles.
template <class T>
class CreateGuard {
  public:
    CreateGuard(T *h, void (*destroy)(T *)) : h_(h), destroy_(destroy) {}
    ~CreateGuard() {
        if (h_) destroy_(h_);
    }
    CreateGuard(const CreateGuard &) = delete;
    CreateGuard &operator=(const CreateGuard &) = delete;
    T *release() {
        T *h = h_;
        h_ = nullptr;
        return h;
    }

  private:
    T *h_;
    void (*destroy_)(T *);
};

// Remaps a linear block id so that each XCD (blocks are dealt round-robin to
// the 8 XCDs) works on one contiguous slice of the index space; neighbouring
// nodes then share an XCD's L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nblocks) {
    const int64_t nx = 8;
    int64_t per = nblocks / nx;
    int64_t rem = nblocks % nx;
    int64_t x = b % nx, k = b / nx;
    // blocks of XCD x cover [start_x, start_x + per + (x < rem))
    int64_t start = x * per + (x < rem ? x : rem);
    int64_t cnt = per + (x < rem ? 1 : 0);
    return k < cnt ? start + k : b;  // the tail (k >= cnt) cannot happen; keep b
}

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace psx

#define PSX_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            psx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                           __LINE__);                                                     \
            return PSX_EHIP;                                                              \
        }                                                                                 \
    } while (0)

#define PSX_CHECK_ARG(cond, msg)              \
    do {                                      \
        if (!(cond)) {                        \
            psx::set_error("%s", msg);        \
            return PSX_EINVAL;                \
        }                                     \
    } while (0)

#define PSX_GUARD_BEGIN try {
#define PSX_GUARD_END                                          \
    }                                                          \
    catch (const std::bad_alloc &) {                           \
        psx::set_error("host allocation failed");              \
        return PSX_ENOMEM;                                     \
    }                                                          \
    catch (const std::exception &ex) {                         \
        psx::set_error("%s", ex.what());                       \
        return PSX_EINVAL;                                     \
    }
