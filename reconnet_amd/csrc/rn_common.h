// Shared definitions for the MI355X (gfx950) ReconNet projector library.
#pragma once

#include <cstdio>
#include <utility>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace rn {

// Error raised inside the library; converted to a status + message at the
// C-ABI boundary (nothing is thrown across extern "C").
struct Error : public std::runtime_error {
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

void set_last_error(const std::string& msg);
const char* last_error();
void clear_last_error();

#define RN_HIP_CHECK(expr)                                                              \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            throw ::rn::Error(std::string(#expr) + " -> " + hipGetErrorString(_e));    \
    } while (0)

#define RN_REQUIRE(cond, msg)                                                           \
    do {                                                                                \
        if (!(cond)) throw ::rn::Error(std::string("invalid argument: ") + (msg));      \
    } while (0)

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Test hook: RECONNET_FAULT=alloc makes every device allocation fail, so the
// error path of every entry point can be exercised without a GPU fault.
bool fault_injected(const char* what);

// RAII device buffer (hipMalloc / hipFree), optionally stream-ordered.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes == 0) return;
        if (fault_injected("alloc")) throw Error("device allocation failed (fault injection)");
        RN_HIP_CHECK(hipMalloc(&p_, bytes));
        n_ = bytes;
    }
    void ensure(size_t bytes) { if (bytes > n_) alloc(bytes); }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr; n_ = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    size_t size() const { return n_; }
private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Stream owned by one ABI call (the reference ABI is synchronous; each call
// gets its own stream so that concurrent host threads do not serialise on
// the legacy default stream).
class CallStream {
public:
    CallStream() { RN_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
    ~CallStream() { if (s_) (void)hipStreamDestroy(s_); }
    hipStream_t get() const { return s_; }
    void sync() const { RN_HIP_CHECK(hipStreamSynchronize(s_)); }
private:
    hipStream_t s_ = nullptr;
};

// Run f for the C-ABI: exceptions never cross extern "C"; a failure prints
// "<name> failed: <reason>" on stderr and becomes the thread's last error.
template <class F>
void guarded(const char* name, F&& f)
{
    try {
        clear_last_error();
        f();
    } catch (const std::exception& e) {
        std::string m = std::string(name) + " failed: " + e.what();
        std::fprintf(stderr, "%s\n", m.c_str());
        set_last_error(m);
    } catch (...) {
        std::string m = std::string(name) + " failed: unknown error";
        std::fprintf(stderr, "%s\n", m.c_str());
        set_last_error(m);
    }
}

template <class F>
int guarded_status(const char* name, F&& f)
{
    guarded(name, std::forward<F>(f));
    return last_error()[0] == '\0' ? 0 : -1;
}

inline void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e));
}

} // namespace rn
