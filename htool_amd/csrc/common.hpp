// Shared host-side definitions of libhtool_amd: error handling, device arrays,
// the leaf record.
#pragma once

#include "../../include/htool_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace htg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HTG_HIP(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::htg::Error(e_ == hipErrorOutOfMemory ? HTG_ERR_OOM : HTG_ERR_HIP,                 \
                               std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_));  \
    } while (0)

// Owning device array (hipMalloc / hipFree); uploads are synchronous.
template <typename T>
struct DevArray {
    T *p     = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { reset(); }
    void reset() {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        reset();
        if (count == 0)
            count = 1;
        HTG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    // at least `count` elements, contents not kept; grows only
    void reserve(size_t count) {
        if (count > n || !p)
            alloc(count);
    }
    void zero() {
        if (p)
            HTG_HIP(hipMemset(p, 0, n * sizeof(T)));
    }
    void upload(const T *src, size_t count) {
        alloc(count);
        if (count)
            HTG_HIP(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
    void take(DevArray &o) {
        reset();
        p   = o.p;
        n   = o.n;
        o.p = nullptr;
        o.n = 0;
    }
};

inline int pad4(int v) { return (v + 3) & ~3; }

// Message of the last failed entry point on this thread (htg_last_error).
std::string &last_error();

// Runs an entry point's body; exceptions become a status and a message.
template <typename F>
int guard(F &&f) {
    try {
        f();
        return HTG_OK;
    } catch (const Error &e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        last_error() = "host allocation failed";
        return HTG_ERR_OOM;
    } catch (const std::invalid_argument &e) {
        last_error() = std::string("[Htool error] ") + e.what();
        return HTG_ERR_INVALID;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return HTG_ERR_HIP;
    }
}

// One leaf of the H-matrix in cluster numbering.
struct Leaf {
    int oi, m, oj, n, k;           // k = -1: dense
    uint64_t src_u = 0, src_v = 0; // build path: U (m x k col-major) and V (k x n row-major) in the store
    int k_hi = 0;                  // low-rank: terms stored in class 0 (the rest in class 1)
    std::vector<double> sig;       // low-rank: per-term scale (empty: all 1)
    bool mirror = false;           // symmetric storage: transposed copy of a kept off-diagonal leaf
};

// Stored row length of a low-rank leaf's V: the global columns
// [oj & ~3, round4(oj + n)), so rows and the matching x entries are 16-byte aligned.
inline int v_ld(const Leaf &L) { return pad4(L.oj + L.n) - (L.oj & ~3); }

} // namespace htg
