// common.hpp -- shared definitions of libtopo_hip (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include "../../include/topo.h"

namespace topo {

// thread-local error message behind topo_last_error()
void set_error(const char* fmt, ...);

struct TopoError {
    int code;
};

#define TOPO_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            ::topo::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),    \
                              __FILE__, __LINE__);                                       \
            throw ::topo::TopoError{_e == hipErrorOutOfMemory ? TOPO_E_OOM : TOPO_E_HIP}; \
        }                                                                               \
    } while (0)

#define TOPO_CHECK(cond, code, ...)                \
    do {                                           \
        if (!(cond)) {                             \
            ::topo::set_error(__VA_ARGS__);        \
            throw ::topo::TopoError{code};         \
        }                                          \
    } while (0)

// Structured grid of one MG level. Device element order ("zyx"): ez fastest,
// then ey, then ex: ei = (ex*nely + ey)*nelz + ez. Node/DOF order is the
// reference's (iz fastest): n = (ix*(nely+1) + iy)*(nelz+1) + iz.
struct Grid {
    int nelx = 0, nely = 0, nelz = 0;
    __host__ __device__ int nnx() const { return nelx + 1; }
    __host__ __device__ int nny() const { return nely + 1; }
    __host__ __device__ int nnz() const { return nelz + 1; }
    __host__ __device__ int64_t nel() const { return (int64_t)nelx * nely * nelz; }
    __host__ __device__ int64_t nnodes() const { return (int64_t)nnx() * nny() * nnz(); }
    __host__ __device__ int64_t ndof() const { return 3 * nnodes(); }
};

// host helpers shared by the C-ABI files
void k0_matrix(double nu, double* ke576);
bool edofs_structured(const int32_t* edofs, int64_t ndof, Grid* g);

}  // namespace topo

// run body, translating exceptions into status codes at the C boundary
#define TOPO_API_BEGIN try {
#define TOPO_API_END                                  \
    }                                                 \
    catch (const ::topo::TopoError& e) {              \
        return e.code;                                \
    }                                                 \
    catch (const std::bad_alloc&) {                   \
        ::topo::set_error("host allocation failed");  \
        return TOPO_E_OOM;                            \
    }                                                 \
    catch (...) {                                     \
        ::topo::set_error("unexpected C++ exception"); \
        return TOPO_E_HIP;                            \
    }
