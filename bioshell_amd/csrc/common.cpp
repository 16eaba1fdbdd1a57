// common.cpp -- error state, device-memory budget, buffers and timers.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

namespace bs {

static thread_local std::string t_last_error;

void set_last_error(const std::string& msg) { t_last_error = msg; }
void clear_last_error() { t_last_error.clear(); }
const char* last_error_cstr() { return t_last_error.c_str(); }

void fail(bs_status code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

void* Budget::alloc(size_t bytes, const char* what)
{
    if (bytes == 0) return nullptr;
    {
        std::lock_guard<std::mutex> g(mu_);
        if (cap_ != 0 && used_ + bytes > cap_)
            fail(BS_ERR_OUT_OF_BUDGET, "%s: %zu bytes would exceed the HBM budget (%zu of %zu in use)",
                 what, bytes, used_, cap_);
        used_ += bytes;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> g(mu_);
        used_ -= bytes;
        fail(e == hipErrorOutOfMemory ? BS_ERR_OUT_OF_BUDGET : BS_ERR_HIP, "%s: hipMalloc(%zu) failed: %s",
             what, bytes, hipGetErrorString(e));
    }
    return p;
}

void Budget::free(void* p, size_t bytes)
{
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> g(mu_);
    used_ -= bytes;
}

bool Budget::try_reserve(size_t bytes)
{
    std::lock_guard<std::mutex> g(mu_);
    if (cap_ != 0 && used_ + bytes > cap_) return false;
    used_ += bytes;
    return true;
}

void Budget::release(size_t bytes)
{
    std::lock_guard<std::mutex> g(mu_);
    used_ -= bytes;
}

void* DevBuf::ensure(size_t bytes, const char* what)
{
    if (bytes <= bytes_ && ptr_) return ptr_;
    reset();
    if (bytes == 0) return nullptr;
    // grow geometrically so repeated slightly larger calls do not reallocate
    const size_t want = bytes + bytes / 8;
    ptr_ = budget_->alloc(want, what);
    bytes_ = want;
    ++gen_;
    return ptr_;
}

void DevBuf::reset()
{
    if (ptr_ && budget_) budget_->free(ptr_, bytes_);
    ptr_ = nullptr;
    bytes_ = 0;
}

void Timer::create()
{
    BS_HIP(hipEventCreate(&a));
    BS_HIP(hipEventCreate(&b));
}

void Timer::destroy()
{
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    a = b = nullptr;
}

float Timer::ms() const
{
    BS_HIP(hipEventSynchronize(b));
    float v = 0.0f;
    BS_HIP(hipEventElapsedTime(&v, a, b));
    return v;
}

// Copies a list of (dst, src, bytes) spans with up to 16 threads.
void parallel_copy(const std::vector<Span>& spans)
{
    size_t total = 0;
    for (const Span& s : spans) total += s.bytes;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<size_t>(std::min(16u, hw), std::max<size_t>(1, total >> 24));
    if (nt <= 1) {
        for (const Span& s : spans) std::memcpy(s.dst, s.src, s.bytes);
        return;
    }
    // thread t copies bytes [t total / nt, (t+1) total / nt) of the concatenation
    auto work = [&](int t) {
        const size_t lo = total * t / nt, hi = total * (t + 1) / nt;
        size_t pos = 0;
        for (const Span& s : spans) {
            const size_t a = std::max(lo, pos), b = std::min(hi, pos + s.bytes);
            if (a < b)
                std::memcpy(static_cast<char*>(s.dst) + (a - pos),
                            static_cast<const char*>(s.src) + (a - pos), b - a);
            pos += s.bytes;
            if (pos >= hi) break;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
}

}  // namespace bs
