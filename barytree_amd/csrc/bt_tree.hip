// bt_tree.hip -- device tree build: exact bounding boxes and stable
// midpoint partitions (gfx950).
//
// One block of 256 threads (4 waves of 64) per chunk of up to TREE_CHUNK
// elements.  Coordinates are read as coalesced SoA streams; a partition
// writes the lower and the upper part of a chunk as two contiguous runs, so
// its stores are coalesced as well.  The kernels are HBM-bound: a bounding
// box reads 24 B per element, a partition pass reads and writes 28 B.
#include <hip/hip_runtime.h>

#include "bt_tree.h"

namespace btk {

namespace {

constexpr int TB = 256;  // threads per block
constexpr int WAVES = TB / 64;

__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

__global__ __launch_bounds__(TB) void k_tree_bbox(const TreeChunk *__restrict__ chunks,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ z,
                                                  double *__restrict__ part)
{
    const TreeChunk c = chunks[blockIdx.x];
    // every chunk holds at least one element: start from it
    double xn = x[c.beg], xx = xn, yn = y[c.beg], yx = yn, zn = z[c.beg], zx = zn;
    for (int i = c.beg + threadIdx.x; i < c.beg + c.cnt; i += TB) {
        const double a = x[i], b = y[i], d = z[i];
        xn = a < xn ? a : xn; xx = a > xx ? a : xx;
        yn = b < yn ? b : yn; yx = b > yx ? b : yx;
        zn = d < zn ? d : zn; zx = d > zx ? d : zx;
    }
    xn = wave_min(xn); xx = wave_max(xx);
    yn = wave_min(yn); yx = wave_max(yx);
    zn = wave_min(zn); zx = wave_max(zx);
    __shared__ double s[WAVES][6];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        s[w][0] = xn; s[w][1] = xx; s[w][2] = yn; s[w][3] = yx; s[w][4] = zn; s[w][5] = zx;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double v = s[0][k];
        for (int j = 1; j < WAVES; j++) {
            const double u = s[j][k];
            v = (k & 1) ? (u > v ? u : v) : (u < v ? u : v);
        }
        part[6 * (size_t)blockIdx.x + k] = v;
    }
}

// lower part: coord <= mid, or coord < mid for a one-element range
__device__ inline bool goes_low(const TreeSeg &s, const double *x, const double *y, const double *z, int i)
{
    if (s.dim < 0) return true;
    const double a = s.dim == 0 ? x[i] : (s.dim == 1 ? y[i] : z[i]);
    return s.strict ? (a < s.mid) : (a <= s.mid);
}

__global__ __launch_bounds__(TB) void k_tree_count(const TreeChunk *__restrict__ chunks,
                                                   const TreeSeg *__restrict__ segs,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z, int *__restrict__ low)
{
    const TreeChunk c = chunks[blockIdx.x];
    const TreeSeg s = segs[c.seg];
    int n = 0;
    for (int i = c.beg + threadIdx.x; i < c.beg + c.cnt; i += TB) n += goes_low(s, x, y, z, i) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    __shared__ int sn[WAVES];
    if ((threadIdx.x & 63) == 0) sn[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int j = 0; j < WAVES; j++) t += sn[j];
        low[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(TB) void k_tree_scatter(const TreeChunk *__restrict__ chunks,
                                                     const TreeSeg *__restrict__ segs,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     const int *__restrict__ idx,
                                                     double *__restrict__ ox, double *__restrict__ oy,
                                                     double *__restrict__ oz, int *__restrict__ oidx)
{
    const TreeChunk c = chunks[blockIdx.x];
    const TreeSeg s = segs[c.seg];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    // destinations: the lower part of the segment starts at s.beg, the upper
    // part at s.beg + s.lowtot; this chunk's elements follow those of the
    // segment's earlier chunks in each part
    int lo_next = s.beg + c.lowoff;
    int hi_next = s.beg + s.lowtot + (c.beg - s.beg - c.lowoff);
    __shared__ int sl[WAVES];
    for (int base = c.beg; base < c.beg + c.cnt; base += TB) {
        const int i = base + threadIdx.x;
        const bool valid = i < c.beg + c.cnt;
        const bool low = valid && goes_low(s, x, y, z, i);
        const unsigned long long bl = __ballot(low);
        const unsigned long long bv = __ballot(valid);
        if (lane == 0) sl[w] = __popcll(bl);
        __syncthreads();
        int lo_before = 0, lo_all = 0;
        for (int j = 0; j < WAVES; j++) {
            const int v = sl[j];
            lo_before += j < w ? v : 0;
            lo_all += v;
        }
        __syncthreads();
        // rank of this element among the earlier valid elements of the block
        const int valid_before_wave = w * 64;  // every wave but the last block iteration's tail is full
        const int nvalid_here = __popcll(bv & below);
        const int lrank = lo_before + __popcll(bl & below);
        const int pos = valid_before_wave + nvalid_here;  // position within this iteration
        if (valid) {
            const int dst = low ? lo_next + lrank : hi_next + (pos - lrank);
            ox[dst] = x[i]; oy[dst] = y[i]; oz[dst] = z[i]; oidx[dst] = idx[i];
        }
        const int nvalid = min(TB, c.beg + c.cnt - base);
        lo_next += lo_all;
        hi_next += nvalid - lo_all;
    }
}

__global__ void k_tree_gather(int n, const int *__restrict__ idx, const double *__restrict__ a,
                              double *__restrict__ out_a, const double *__restrict__ b, double fill_b,
                              double *__restrict__ out_b)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int j = idx[i];
        out_a[i] = a[j];
        if (out_b) out_b[i] = b ? b[j] : fill_b;
    }
}

__global__ void k_differ(int n, const double *__restrict__ a, const double *__restrict__ b, int *flag)
{
    bool d = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d |= __double_as_longlong(a[i]) != __double_as_longlong(b[i]);
    if (__any(d) && (threadIdx.x & 63) == 0) *flag = 1;  // every writer stores the same value
}

__global__ void k_iota(int n, int *__restrict__ idx)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) idx[i] = i;
}

__global__ void k_invert(int n, const int *__restrict__ order, int *__restrict__ inv)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) inv[order[i]] = i;
}

__global__ void k_gather_sel(int m, const int *__restrict__ sel, const int *__restrict__ inv,
                             const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const double *__restrict__ q,
                             double *__restrict__ ox, double *__restrict__ oy, double *__restrict__ oz,
                             double *__restrict__ oq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int j = inv[sel[k]];
    ox[k] = x[j]; oy[k] = y[j]; oz[k] = z[j]; oq[k] = q[j];
}

int grid_for(long long n) { return (int)std::min<long long>((n + TB - 1) / TB, 8192); }

}  // namespace

hipError_t launch_tree_bbox(int nchunks, const TreeChunk *chunks, const double *x, const double *y,
                            const double *z, double *part, hipStream_t st)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_bbox, dim3(nchunks), dim3(TB), 0, st, chunks, x, y, z, part);
    return hipGetLastError();
}

hipError_t launch_tree_count(int nchunks, const TreeChunk *chunks, const TreeSeg *segs, const double *x,
                             const double *y, const double *z, int *low, hipStream_t st)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_count, dim3(nchunks), dim3(TB), 0, st, chunks, segs, x, y, z, low);
    return hipGetLastError();
}

hipError_t launch_tree_scatter(int nchunks, const TreeChunk *chunks, const TreeSeg *segs, const double *x,
                               const double *y, const double *z, const int *idx, double *ox, double *oy,
                               double *oz, int *oidx, hipStream_t st)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_scatter, dim3(nchunks), dim3(TB), 0, st, chunks, segs, x, y, z, idx, ox, oy, oz,
                       oidx);
    return hipGetLastError();
}

hipError_t launch_tree_gather(int n, const int *idx, const double *a, double *out_a, const double *b,
                              double fill_b, double *out_b, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_gather, dim3(grid_for(n)), dim3(TB), 0, st, n, idx, a, out_a, b, fill_b, out_b);
    return hipGetLastError();
}

hipError_t launch_differ(int n, const double *a, const double *b, int *flag, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_differ, dim3(grid_for(n)), dim3(TB), 0, st, n, a, b, flag);
    return hipGetLastError();
}

hipError_t launch_iota(int n, int *idx, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(TB), 0, st, n, idx);
    return hipGetLastError();
}

hipError_t launch_invert(int n, const int *order, int *inv, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_invert, dim3(grid_for(n)), dim3(TB), 0, st, n, order, inv);
    return hipGetLastError();
}

hipError_t launch_gather_sel(int m, const int *sel, const int *inv, const double *x, const double *y,
                             const double *z, const double *q, double *ox, double *oy, double *oz, double *oq,
                             hipStream_t st)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_sel, dim3((m + TB - 1) / TB), dim3(TB), 0, st, m, sel, inv, x, y, z, q, ox, oy, oz,
                       oq);
    return hipGetLastError();
}

}  // namespace btk
