// Shared host-side helpers: error propagation, device buffers, timers.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace emtd {

// An error carrying one of the EMTD_ERR_* codes of emtd.h.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string& s);

#define EMTD_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::emtd::Error(0x0080, std::string("HIP error ") + hipGetErrorString(e_) +   \
                                            " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            hipError_t e = hipMalloc(&p, count * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr;
                n = 0;
                throw Error(0x0001, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
            }
        }
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s = 0) {
        if (count > n) throw Error(0x0004, "upload larger than buffer");
        if (count) EMTD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s = 0) {
        if (v.size() > n || !p) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    void download(T* h, size_t count, hipStream_t s = 0) const {
        if (count) EMTD_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

template <typename T>
DevBuf<T> to_device(const std::vector<T>& v, hipStream_t s = 0) {
    DevBuf<T> b(v.size());
    b.upload(v.data(), v.size(), s);
    return b;
}

inline double wall_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Reads the device ordinal from an argument, -1 meaning "current device".
inline void select_device(int device) {
    if (device >= 0) EMTD_HIP(hipSetDevice(device));
}

}  // namespace emtd
