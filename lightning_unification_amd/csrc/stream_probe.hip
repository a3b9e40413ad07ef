// Streaming-bandwidth probe of the device, for the benchmark's roofline
// argument (bench.py --mode stream).  Not part of the engine's interface: it
// is built as its own small library (libpl_stream.so) and only the benchmark
// loads it.  The kernels move complex128 amplitudes with the same
// non-temporal loads and stores, XCD-aware block order and one-shot grid as
// the gate kernels, so their rates are the ceilings those kernels work
// against:
//   mode 0  read only        (re + im of each amplitude compared with a
//                            run-time value no sum matches)
//   mode 1  write only       (a constant)
//   mode 2  read-modify-write in place, one amplitude per thread
//   mode 3  read-modify-write in place, one pair (i, i + 2^p) per thread:
//           the access pattern of a dense 1-qubit gate on target bit p
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint64_t xcd_block() {
    const uint64_t b = blockIdx.x, nb = gridDim.x;
    if (nb & 7) {
        return b;
    }
    return (b & 7) * (nb >> 3) + (b >> 3);
}

__device__ __forceinline__ double2 ld(const double2 *p) {
    double2 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
__device__ __forceinline__ void st(double2 *p, double2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// `never` is a run-time value no sum matches, so the store is never taken,
// but the compiler cannot drop the loads that feed the comparison
__global__ __launch_bounds__(kB) void k_read(const double2 *__restrict__ a, uint64_t n, double never,
                                             double *__restrict__ out) {
    const uint64_t i = xcd_block() * kB + threadIdx.x;
    double s = 0;
    if (i < n) {
        const double2 v = ld(a + i);
        s = v.x + v.y;
    }
    if (s == never) {
        out[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kB) void k_write(double2 *__restrict__ a, uint64_t n) {
    const uint64_t i = xcd_block() * kB + threadIdx.x;
    if (i < n) {
        st(a + i, double2{0.5, -0.25});
    }
}

__global__ __launch_bounds__(kB) void k_rmw(double2 *__restrict__ a, uint64_t n, double c, double s) {
    const uint64_t i = xcd_block() * kB + threadIdx.x;
    if (i < n) {
        const double2 v = ld(a + i);
        st(a + i, double2{c * v.x - s * v.y, s * v.x + c * v.y});
    }
}

__global__ __launch_bounds__(kB) void k_rmw_pair(double2 *__restrict__ a, uint64_t npairs, int p, double c,
                                                 double s) {
    const uint64_t g = xcd_block() * kB + threadIdx.x;
    if (g >= npairs) {
        return;
    }
    const uint64_t lo = g & ((uint64_t{1} << p) - 1);
    const uint64_t i0 = ((g >> p) << (p + 1)) | lo, i1 = i0 | (uint64_t{1} << p);
    const double2 u = ld(a + i0), v = ld(a + i1);
    // a real rotation [[c, -s], [s, c]]
    st(a + i0, double2{c * u.x - s * v.x, c * u.y - s * v.y});
    st(a + i1, double2{s * u.x + c * v.x, s * u.y + c * v.y});
}

thread_local const char *g_msg = "";

} // namespace

extern "C" {

const char *sp_last_error(void) { return g_msg; }

// `reps` launches of `mode` over the n complex128 amplitudes at dev (device
// memory of at least 16 n bytes, n a power of two >= 512; `scratch` holds one
// double per workgroup for mode 0).  *ms receives the mean time per launch,
// measured with events on a stream of its own.
int sp_run(void *dev, uint64_t n, int mode, int p, int reps, void *scratch, double *ms) {
    if (!dev || !ms || n < 512 || (n & (n - 1)) || reps < 1 || mode < 0 || mode > 3 ||
        (mode == 3 && (p < 0 || (uint64_t{2} << p) > n)) || (mode == 0 && !scratch)) {
        g_msg = "invalid argument";
        return 1;
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        g_msg = "stream or event creation failed";
        return 3;
    }
    auto *a = static_cast<double2 *>(dev);
    const uint64_t work = mode == 3 ? n / 2 : n;
    const dim3 grid(static_cast<unsigned>((work + kB - 1) / kB));
    hipError_t err = hipEventRecord(e0, st);
    for (int r = 0; r < reps && err == hipSuccess; r++) {
        switch (mode) {
        case 0: hipLaunchKernelGGL(k_read, grid, dim3(kB), 0, st, a, n, -12345.678, static_cast<double *>(scratch)); break;
        case 1: hipLaunchKernelGGL(k_write, grid, dim3(kB), 0, st, a, n); break;
        case 2: hipLaunchKernelGGL(k_rmw, grid, dim3(kB), 0, st, a, n, 0.6, 0.8); break;
        default: hipLaunchKernelGGL(k_rmw_pair, grid, dim3(kB), 0, st, a, work, p, 0.6, 0.8); break;
        }
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        err = hipEventRecord(e1, st);
    }
    if (err == hipSuccess) {
        err = hipEventSynchronize(e1);
    }
    float t = 0;
    if (err == hipSuccess) {
        err = hipEventElapsedTime(&t, e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (err != hipSuccess) {
        g_msg = hipGetErrorString(err);
        return 3;
    }
    *ms = static_cast<double>(t) / reps;
    return 0;
}

} // extern "C"
