// common.hpp -- error handling, device-memory accounting and the context object
// shared by the host side of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bioshell_gpu.h"

namespace bs {

// An error that crosses the C ABI as a status code plus a message.
struct Error : std::runtime_error {
    bs_status code;
    Error(bs_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void fail(bs_status code, const char* fmt, ...);
void set_last_error(const std::string& msg);
void clear_last_error();

#define BS_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::bs::fail(BS_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                              \
    } while (0)

#define BS_REQUIRE(cond, ...)                                  \
    do {                                                       \
        if (!(cond)) ::bs::fail(BS_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

// Runs `f` and maps exceptions to a status; nothing unwinds past the ABI.
template <class F>
bs_status guard(F&& f)
{
    try {
        clear_last_error();
        f();
        return BS_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return BS_ERR_INTERNAL;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return BS_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return BS_ERR_INTERNAL;
    }
}

// Device memory charged against a byte budget shared by a context and the
// libraries / servers made from it.
class Budget {
public:
    explicit Budget(size_t cap) : cap_(cap) {}
    void* alloc(size_t bytes, const char* what);
    void free(void* p, size_t bytes);
    size_t in_use() const { return used_; }
    size_t cap() const { return cap_; }
    void set_cap(size_t c) { cap_ = c; }
    bool fits(size_t bytes) const { return cap_ == 0 || used_ + bytes <= cap_; }
    // Reserve / release accounting without allocating (admission control).
    bool try_reserve(size_t bytes);
    void release(size_t bytes);

private:
    mutable std::mutex mu_;
    size_t cap_;
    size_t used_ = 0;
};

// A grow-only device buffer charged to a Budget.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void bind(Budget* b) { budget_ = b; }
    // Ensures capacity >= bytes; contents are not preserved on growth.
    void* ensure(size_t bytes, const char* what);
    void reset();
    void* get() const { return ptr_; }
    size_t size() const { return bytes_; }
    // Allocation count of this buffer (0: never allocated): a changed value
    // means memory of unknown contents, even at the same address and size.
    uint64_t generation() const { return gen_; }
    template <class T>
    T* as() const { return static_cast<T*>(ptr_); }

private:
    Budget* budget_ = nullptr;
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    uint64_t gen_ = 0;
};

// Copies a list of (dst, src, bytes) spans with up to 16 threads.
struct Span {
    void* dst;
    const void* src;
    size_t bytes;
};
void parallel_copy(const std::vector<Span>& spans);

// Makes `dev` the current device for a scope.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        BS_HIP(hipGetDevice(&prev));
        if (prev != dev) BS_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;
};

// Event pair used to time a stretch of work on one stream.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    void create();
    void destroy();
    void start(hipStream_t s) { BS_HIP(hipEventRecord(a, s)); }
    void stop(hipStream_t s) { BS_HIP(hipEventRecord(b, s)); }
    float ms() const;  // synchronises on b
};

}  // namespace bs
