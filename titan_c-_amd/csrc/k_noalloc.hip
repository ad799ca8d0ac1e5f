// k_noalloc.hip -- titan_chlld_batch_noalloc: the batch Riemann solve of titan_chlld_batch on
// caller-owned device arrays, for callers that cannot allocate device memory (code under HIP graph
// capture, tight per-step loops, harnesses that forbid hipMalloc).
//
// The call's only working memory beyond the caller's inputs and outputs is `scratch`, one
// file-scope device array per GPU (256 bytes, held until the process ends): word 0 is the flag the
// solve kernel raises on the reference's fatal branch, word 1 the flag of an out-of-range method or
// wave-speed code.  The host never passes the array's name to a kernel: at a device's first call,
// with that device current, hipGetSymbolAddress gives the device address of that GPU's copy, which
// is kept per device and used for every kernel argument and copy.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (two host threads, or two streams without an order
// between them) race on the flags.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/titan_gpu.h"
#include "launch.h"

namespace titan {

constexpr int kNoallocWords = 64;                 // 256 bytes
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint32_t scratch[kNoallocWords];

static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[kMaxDevices];      // device address of each GPU's scratch, or null

// The current device's scratch address (looked up once per device).
static int scratch_address(uint32_t **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TITAN_E_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TITAN_E_HIP;
        g_scratch_ptr[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch_ptr[dev];
    return TITAN_OK;
}

// Clears the flags (a kernel rather than hipMemsetAsync, so a captured graph replays it).
__global__ void k_scratch_clear(uint32_t *w, int n) {
    if ((int)threadIdx.x < n) w[threadIdx.x] = 0u;
}

// Raises w[1] when a state's method or wave-speed code is outside 0..3 (titan_chlld_batch rejects
// those on the host before any launch).
__global__ void __launch_bounds__(256) k_chlld_codes(int64_t n, const int8_t *__restrict__ n_state,
                                                     const int8_t *__restrict__ n_disc, uint32_t *w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = n_state[i], d = n_disc ? n_disc[i] : 0;
    if (s < 0 || s > 3 || d < 0 || d > 3) atomicOr(&w[1], 1u);
}

}  // namespace titan

using namespace titan;

extern "C" int titan_chlld_batch_noalloc(int64_t n, const int8_t *n_state, const int8_t *n_disc,
                                         const uint8_t *flags, const double *normal, const double *w,
                                         const double *qL, const double *qR, int K, const double *kL,
                                         const double *kR, double *q, double *k, double *ds, uint8_t *branch,
                                         int precision, void *stream) {
    if (n < 0 || n > TITAN_NOALLOC_MAX_STATES || K < 0 || K > 4 || !n_state || !normal || !qL || !qR || !q)
        return TITAN_E_ARG;
    if (K > 0 && (!kL || !kR || !k)) return TITAN_E_ARG;
    if (precision != TITAN_PREC_FP64 && precision != TITAN_PREC_MIXED) return TITAN_E_ARG;
    if (n == 0) return TITAN_OK;
    uint32_t *sw = nullptr;
    const int rc = scratch_address(&sw);
    if (rc != TITAN_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(64), 0, s, sw, 2);
    hipLaunchKernelGGL(k_chlld_codes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, n_state, n_disc, sw);
    launch_chlld_batch(s, precision, n, K, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch,
                       reinterpret_cast<int *>(sw));
    if (hipGetLastError() != hipSuccess) return TITAN_E_HIP;
    uint32_t h[2] = {0u, 0u};
    if (hipMemcpyAsync(h, sw, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return TITAN_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return TITAN_E_HIP;
    if (h[1]) return TITAN_E_ARG;
    return h[0] ? TITAN_E_RIEMANN : TITAN_OK;
}
