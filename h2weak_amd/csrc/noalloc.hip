// noalloc.hip -- the dense product for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  h2w_dense_matvec stages x in tree order and the row list in
// buffers it allocates per call; h2w_dense_matvec_noalloc stages them in one
// fixed device array per GPU instead, reached through the address that
// hipGetSymbolAddress gives for the device the handle was created on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "h2w_handle.hpp"

namespace h2w {
namespace {

// Staging of one call: x in tree order (kNoAllocMaxN doubles), the rows in tree order
// (kNoAllocMaxRows int32), a count of out-of-range row indices (one int32 at the end).
constexpr int64_t kNoAllocMaxN = (int64_t)1 << 20;
constexpr int64_t kNoAllocMaxRows = (int64_t)1 << 20;
constexpr size_t kXtBytes = (size_t)kNoAllocMaxN * 8;
constexpr size_t kRowBytes = (size_t)kNoAllocMaxRows * 4;
constexpr size_t kScratchBytes = kXtBytes + kRowBytes + 256;

// One copy per GPU, held until the process ends.  In host code `scratch` is only a
// stand-in for the symbol: its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// out[q] = tree index of input row rows[q] (rows == NULL: row q); indices outside [0, N)
// map to row 0 and are counted in *bad
__global__ __launch_bounds__(256) void k_rows_to_tree(const int64_t *__restrict__ rows, int64_t nrows,
                                                      const int32_t *__restrict__ iperm, int64_t N,
                                                      int32_t *__restrict__ out, int32_t *__restrict__ bad) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nrows; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = rows ? rows[q] : q;
        if (r >= 0 && r < N) out[q] = iperm[r];
        else {
            out[q] = 0;
            atomicAdd(bad, 1);
        }
    }
}

__global__ void k_clear_flag(int32_t *p) { *p = 0; }

}  // namespace

// Device address of this GPU's staging array, looked up once per device with that device
// current (called when a handle is attached to its device).
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error{H2W_EINVAL, "device ordinal out of range"};
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error{H2W_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch[device] = p;
}

}  // namespace h2w

using h2w::Error;

extern "C" {

int h2w_dense_matvec_noalloc(h2w_handle *h, const double *x, double *y, const int64_t *rows, int64_t nrows) {
    try {
        if (!h || !x || !y) throw Error{H2W_EINVAL, "NULL argument"};
        if (h->opts.plan_only || !h->stream) throw Error{H2W_ESTATE, "handle was created with plan_only"};
        if (h->cplx) throw Error{H2W_EINVAL, "complex kernels: use h2w_dense_matvec"};
        const int64_t N = h->N;
        if (!rows) nrows = N;
        if (nrows < 0) throw Error{H2W_EINVAL, "nrows < 0"};
        if (N > h2w::kNoAllocMaxN)
            throw Error{H2W_EINVAL, "h2w_dense_matvec_noalloc: N = " + std::to_string(N) + " exceeds its staging limit of " +
                                        std::to_string(h2w::kNoAllocMaxN) + " points (use h2w_dense_matvec)"};
        if (nrows > h2w::kNoAllocMaxRows)
            throw Error{H2W_EINVAL, "h2w_dense_matvec_noalloc: " + std::to_string(nrows) + " rows exceed its limit of " +
                                        std::to_string(h2w::kNoAllocMaxRows) + " per call"};
        void *base = nullptr;
        {
            std::lock_guard<std::mutex> lk(h2w::g_mu);
            const int dev = h->opts.device;
            base = dev >= 0 && dev < h2w::kMaxDevices ? h2w::g_scratch[dev] : nullptr;
        }
        if (!base) throw Error{H2W_ESTATE, "staging array of the handle's device not set up"};
        H2W_HIP(hipSetDevice(h->opts.device));
        hipStream_t s = h->stream;
        double *xt = (double *)base;
        int32_t *tr = (int32_t *)((unsigned char *)base + h2w::kXtBytes);
        int32_t *bad = (int32_t *)((unsigned char *)base + h2w::kXtBytes + h2w::kRowBytes);
        hipLaunchKernelGGL(h2w::k_clear_flag, dim3(1), dim3(1), 0, s, bad);
        h2w::launch_permute_in(x, N, xt, N, h->d_perm.p, N, 1, s);
        if (nrows > 0) {
            const int grid = (int)std::min<int64_t>((nrows + 255) / 256, 4096);
            hipLaunchKernelGGL(h2w::k_rows_to_tree, dim3(grid), dim3(256), 0, s, rows, nrows, h->d_iperm.p, N, tr,
                               bad);
        }
        h2w::launch_dense_rows(tr, nrows, xt, y, N, h->kp, s);
        H2W_HIP(hipGetLastError());
        int32_t nbad = 0;
        H2W_HIP(hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, s));
        H2W_HIP(hipStreamSynchronize(s));
        if (nbad > 0) throw Error{H2W_EINVAL, std::to_string(nbad) + " row indices out of range"};
        return H2W_OK;
    } catch (const Error &e) {
        h2w::set_error(e.msg);
        return e.code;
    }
}

}  // extern "C"
