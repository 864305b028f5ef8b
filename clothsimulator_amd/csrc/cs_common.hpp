// cs_common.hpp — host-side plumbing shared by the library's translation units:
// HIP error propagation, grow-only device buffers and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace cs {

// A HIP failure inside the library is raised as HipError and turned into
// CS_E_HIP (with the message kept for cs_last_error) at the C-ABI boundary.
struct HipError : std::runtime_error {
    explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct StatusError : std::runtime_error {
    int status;
    StatusError(int st, const std::string& s) : std::runtime_error(s), status(st) {}
};

#define CS_HIP(expr)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            throw ::cs::HipError(std::string(#expr) + ": " + hipGetErrorString(_e) + " at " __FILE__); \
    } while (0)

#define CS_LAUNCH_CHECK() CS_HIP(hipGetLastError())

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Grow-only device allocation.  Growing frees the old block (hipFree waits for
// the device), so growth is only done between stream phases.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b) {
        if (b <= bytes && p) return;
        release();
        size_t want = b + b / 4 + 256;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            throw StatusError(3, "device allocation of " + std::to_string(want) + " bytes failed");
        }
        bytes = want;
    }
    template <class T> T* as(size_t n) {
        ensure(n * sizeof(T) > 0 ? n * sizeof(T) : 1);
        return static_cast<T*>(p);
    }
    template <class T> T* ptr() const { return static_cast<T*>(p); }
    // Grow to at least `need` bytes keeping the first `used` bytes (stream-ordered copy).
    void grow_keep(size_t need, size_t used, hipStream_t s) {
        if (need <= bytes && p) return;
        size_t want = need + need / 2 + 256;
        void* q = nullptr;
        if (hipMalloc(&q, want) != hipSuccess) throw StatusError(3, "device allocation of " + std::to_string(want) + " bytes failed");
        if (p && used) {
            hipError_t e = hipMemcpyAsync(q, p, used, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) { (void)hipFree(q); throw HipError(std::string("grow_keep copy: ") + hipGetErrorString(e)); }
        }
        release();
        p = q;
        bytes = want;
    }
};

}  // namespace cs
