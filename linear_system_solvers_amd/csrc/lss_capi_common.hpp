// lss_capi_common.hpp -- host-side helpers shared by the translation units that implement the
// C-ABI of liblss_amd.so: the error type every entry point turns into an lss_status, the per-thread
// message lss_last_error() returns, and an owning device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "lss_amd.h"

namespace lssamd_host {

// message of the calling thread's last failed entry point (defined in lss_adi_capi.hip)
std::string &last_error();

struct lss_error : std::runtime_error {
    int code;
    lss_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LSS_HIPCHK(expr)                                                                                       \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            throw ::lssamd_host::lss_error(LSS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// runs f and maps what it throws to an lss_status (no exception crosses the C-ABI)
template <typename F> int guarded(F &&f)
{
    try {
        f();
        return LSS_OK;
    } catch (const lss_error &e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        last_error() = "host allocation failed";
        return LSS_ERR_ALLOC;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return LSS_ERR_INVALID;
    }
}

inline void require(bool ok, const std::string &msg)
{
    if (!ok)
        throw lss_error(LSS_ERR_INVALID, msg);
}

struct dev_buffer {
    void *p = nullptr;
    dev_buffer() = default;
    dev_buffer(const dev_buffer &) = delete;
    dev_buffer &operator=(const dev_buffer &) = delete;
    ~dev_buffer()
    {
        if (p)
            (void)hipFree(p);
    }
    void alloc(size_t bytes)
    {
        if (bytes == 0)
            bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            throw lss_error(LSS_ERR_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
    }
};

}  // namespace lssamd_host
