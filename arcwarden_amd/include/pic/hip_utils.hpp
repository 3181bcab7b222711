// pic/hip_utils.hpp -- error handling of the C++ facade over libarcwarden_hip.
//
// The reference throws arc::CudaError (a std::runtime_error carrying
// "file:line: expr: name (string)", include/pic/cuda_utils.hpp:29-77); the
// facade rethrows a failed aw_status as arc::HipError with the library's
// message (call name, file:line, HIP / rocFFT error).  hipStream_t is the
// HIP runtime's opaque stream handle, declared here exactly as the HIP
// headers declare it so this header compiles with a plain C++ compiler.
#ifndef ARC_PIC_HIP_UTILS_HPP
#define ARC_PIC_HIP_UTILS_HPP

#include "aw.h"

#include <stdexcept>
#include <string>

struct ihipStream_t;
typedef struct ihipStream_t* hipStream_t;

namespace arc {

struct HipError : std::runtime_error {
    aw_status status;
    HipError(aw_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

inline std::string last_error() {
    char buf[1024];
    aw_last_error(buf, sizeof(buf));
    return buf;
}

// Throws HipError when st is not AW_OK.
inline void check(aw_status st) {
    if (st != AW_OK) throw HipError(st, last_error());
}

inline aw_stream to_aw(hipStream_t s) { return reinterpret_cast<aw_stream>(s); }

inline void stream_synchronize(hipStream_t s) { check(aw_stream_synchronize(to_aw(s))); }

}  // namespace arc

#define AW_CHECK(expr) ::arc::check(expr)

#endif  // ARC_PIC_HIP_UTILS_HPP
