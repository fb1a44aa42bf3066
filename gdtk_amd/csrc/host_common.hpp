// host_common.hpp -- error handling shared by the host sides of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <string>

#include "../../include/gdtk_hip.h"

namespace gdtk {

// Message of the last failing C-ABI call on this thread (defined in sim.cpp).
extern thread_local std::string g_last_error;

struct Error : std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) throw Error(GDTK_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Device buffer of `count` elements of T, zero-filled on allocation.  The fill runs on the
// null stream, which the library's non-blocking streams do not wait for, so the device is
// synchronised before the buffer is handed out.
template <class T>
struct DevArr {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        free();
        n = count;
        HIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        HIPCHK(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
        HIPCHK(hipDeviceSynchronize());
    }
    // Grows (never shrinks) to at least `count` elements; contents are not kept.
    void reserve(size_t count) {
        if (count > n || !p) alloc(count);
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevArr() { free(); }
};

// Runs f, mapping exceptions to a GDTK_ERR_* status and the thread's last-error message.
template <class F>
static int guarded(F&& f) {
    try {
        return f();
    } catch (Error& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (std::exception& e) {
        g_last_error = e.what();
        return GDTK_ERR_ARG;
    }
}

}  // namespace gdtk
