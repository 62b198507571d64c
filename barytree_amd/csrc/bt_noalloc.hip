// bt_noalloc.hip -- device scratch without allocation (gfx950).
//
// One file-scope __device__ array holds the scratch of the allocation-free
// entry points; each GPU holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel (in host code it is a
// stand-in, not a device address): the device address is looked up once per
// GPU with hipGetSymbolAddress, with that GPU current, and kept.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "bt_noalloc.h"

static __device__ __align__(256) unsigned char scratch[btk::NOALLOC_SCRATCH_BYTES];

namespace btk {

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};

__global__ void k_fill_neg(int m, int *pos)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) pos[k] = -1;
}

__global__ void k_find_sel(int b, int e, const int *__restrict__ order, int m, const int *__restrict__ sel,
                           int *__restrict__ pos)
{
    for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        const int c = order[i];
        int lo = 0, hi = m;  // first k with sel[k] >= c
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sel[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        for (int k = lo; k < m && sel[k] == c; k++) pos[k] = i;
    }
}

__global__ void k_gather_pos(int m, const int *__restrict__ pos, const double *__restrict__ x,
                             const double *__restrict__ y, const double *__restrict__ z,
                             const double *__restrict__ q, double *__restrict__ ox, double *__restrict__ oy,
                             double *__restrict__ oz, double *__restrict__ oq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int p = pos[k];
    if (p < 0) {
        ox[k] = oy[k] = oz[k] = oq[k] = NAN;
        return;
    }
    ox[k] = x[p]; oy[k] = y[p]; oz[k] = z[p]; oq[k] = q[p];
}

}  // namespace

hipError_t noalloc_scratch(void **ptr, size_t *bytes)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *ptr = g_ptr[dev];
    *bytes = NOALLOC_SCRATCH_BYTES;
    return hipSuccess;
}

hipError_t launch_fill_neg(int m, int *pos, hipStream_t st)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_neg, dim3((m + 255) / 256), dim3(256), 0, st, m, pos);
    return hipGetLastError();
}

hipError_t launch_find_sel(int b, int e, const int *order, int m, const int *sel, int *pos, hipStream_t st)
{
    if (e <= b || m <= 0) return hipSuccess;
    const long long n = (long long)e - b;
    const int grid = (int)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_find_sel, dim3(grid), dim3(256), 0, st, b, e, order, m, sel, pos);
    return hipGetLastError();
}

hipError_t launch_gather_pos(int m, const int *pos, const double *x, const double *y, const double *z,
                             const double *q, double *ox, double *oy, double *oz, double *oq, hipStream_t st)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_pos, dim3((m + 255) / 256), dim3(256), 0, st, m, pos, x, y, z, q, ox, oy, oz, oq);
    return hipGetLastError();
}

}  // namespace btk
