// Shared declarations of the HIP G0W0/evGW library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/quack_amd.h"

namespace qk {

// Thrown inside the library, turned into a status code at the C-ABI edge.
struct Error {
    int code;
    std::string msg;
};

#define QK_HIP(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::qk::Error{QK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " at " + \
                                            __FILE__ + ":" + std::to_string(__LINE__)};               \
    } while (0)

#define QK_CHECK(cond, code, msg)                              \
    do {                                                       \
        if (!(cond)) throw ::qk::Error{(code), std::string(msg)}; \
    } while (0)

// Batched addressing: batch z = z0 + zdiv * z1; operand offset = z0*s0 + z1*s1.
// upper != 0: only output tiles that meet the upper triangle (row <= column)
// are computed (for a symmetric product; the rest of C is left as it was).
struct Batch {
    long long sa0 = 0, sa1 = 0, sb0 = 0, sb1 = 0, sc0 = 0, sc1 = 0;
    int zdiv = 1;
    int count = 1;
    int upper = 0;
};

// True while the stream records a graph capture (no host timing or
// synchronisation may run then).
inline bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Process-wide GEMM tile autotuner (on by default): see qk_gemm.hip.
void set_autotune(bool on);
// The library FP64 GEMM as an autotuner candidate for large plain products (off by default).
void set_library_gemm(bool on);
bool autotune_on();
std::string autotune_report();
std::string autotune_report_bf16();
// forget every tuned shape (the next large call of each shape tunes again)
void autotune_reset();
void autotune_reset_bf16();

// C(MxN) = alpha * op(A)(MxK) * op(B)(KxN) + beta * C, column-major, FP64 MFMA.
// Returns the number of kernel launches it issued (more than one while tuning).
int dgemm(hipStream_t s, bool transA, bool transB, int M, int N, int K, double alpha, const double* A,
           long long lda, const double* B, long long ldb, double beta, double* C, long long ldc,
           const Batch& b = Batch());

}  // namespace qk
