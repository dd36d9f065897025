// bp_noalloc.hip -- the stand-alone neighbour grid without allocation: the
// hash, stable sort and cell ranges of bp_neighbor_grid (updateGrid,
// particlesolver.cuh:712-765) over device buffers the caller owns, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops).  Its only scratch memory -- the sort's second key and value
// buffers, per-block digit counts and digit totals -- is the file-scope array
// `scratch` below, one copy per GPU, held until the process ends.  The
// kernels take the device address that hipGetSymbolAddress returns for the
// current GPU, never the array's name (which in host code is only a host-side
// stand-in).
#include <mutex>
#include <string>

#include "bp_kernels.h"
#include "bp_solver_impl.h"

namespace bp {
namespace {

using bp_host::fail;

// Largest particle count, and the scratch it needs: keys and values of the
// sort's second buffer (n ints each), digit counts (at most 1024 tiles of 256
// digits below 2^20 keys, radix_counts_len) and 256 digit totals.
constexpr int kNoAllocMaxN = 1 << 20;
constexpr size_t kCountsInts = (size_t)1024 * 256;
constexpr size_t kScratchInts = 2 * (size_t)kNoAllocMaxN + kCountsInts + 256;

static __device__ __align__(256) int scratch[kScratchInts];

// updateGridId (grid.cuh:30-47) of a given position: the same cell as
// k_predict_hash computes for x* = x + v * 0 + 0 (x + 0 is x for every finite x)
__global__ __launch_bounds__(kBlock) void k_hash_positions(const float4 *__restrict__ x, int *__restrict__ keys,
                                                           int n, GridDesc G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int g[3];
    G.pos(xyz(x[i]), g);
    keys[i] = G.addr(g[0], g[1], g[2]);
}

// -1 in both range arrays (a kernel rather than a memset: the call may be
// captured into a graph)
__global__ __launch_bounds__(kBlock) void k_clear_ranges(int *__restrict__ start, int *__restrict__ end, int cells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    if (start) start[c] = -1;
    if (end) end[c] = -1;
}

__global__ __launch_bounds__(kBlock) void k_copy_pairs(const int *__restrict__ ka, const int *__restrict__ va,
                                                       int *__restrict__ kb, int *__restrict__ vb, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kb[i] = ka[i];
    vb[i] = va[i];
}

// findStartEndId (grid.cuh:70-99) into separate start and end arrays
__global__ __launch_bounds__(kBlock) void k_find_ranges_split(const int *__restrict__ sk, int n,
                                                              int *__restrict__ start, int *__restrict__ end) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int u = sk[i];
    if (i == 0 || sk[i - 1] != u) {
        if (start) start[u] = i;
        if (i > 0 && end) end[sk[i - 1]] = i;
    }
    if (i == n - 1 && end) end[u] = n;
}

// the scratch array's device address on each GPU, looked up once per GPU
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int *g_scratch_ptr[kMaxDevices] = {};

bp_status scratch_for_current_device(int **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BP_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (int *)p;
    }
    *out = g_scratch_ptr[dev];
    return BP_OK;
}

}  // namespace
}  // namespace bp

using namespace bp;
using namespace bp_host;

extern "C" bp_status bp_neighbor_grid_noalloc(const float *positions, int n, const float cell_size[3],
                                              const int grid_dims[3], int *keys, int *perm, int *cell_start,
                                              int *cell_end, void *hip_stream) {
    if (n < 0 || !cell_size || !grid_dims) return fail(BP_ERR_INVALID_ARG, "invalid arguments");
    if (n > kNoAllocMaxN)
        return fail(BP_ERR_CAPACITY, "bp_neighbor_grid_noalloc takes at most 2^20 positions (" + std::to_string(n) +
                                         " given); use bp_neighbor_grid");
    if (n > 0 && (!positions || !keys || !perm))
        return fail(BP_ERR_INVALID_ARG, "positions, keys and perm are required when n > 0");
    GridDesc G{};
    for (int a = 0; a < 3; ++a) {
        if (!is_pow2(grid_dims[a]) || grid_dims[a] < 4 || grid_dims[a] > 1024)
            return fail(BP_ERR_INVALID_ARG, "grid_dims must be powers of two in [4, 1024]");
        if (!(cell_size[a] > 0.0f)) return fail(BP_ERR_INVALID_ARG, "cell_size must be > 0");
        G.cs[a] = cell_size[a];
        G.dims[a] = grid_dims[a];
        G.mask[a] = (uint32_t)grid_dims[a] - 1u;
    }
    G.stride_x = (uint32_t)(grid_dims[1] * grid_dims[2]);
    G.shift_y = (uint32_t)ilog2((unsigned)grid_dims[2]);
    G.shift_x = (uint32_t)ilog2((unsigned)(grid_dims[1] * grid_dims[2]));
    const long long cells = (long long)grid_dims[0] * grid_dims[1] * grid_dims[2];
    if (cells > (1LL << 30)) return fail(BP_ERR_CAPACITY, "at most 2^30 cells");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(BP_ERR_NO_DEVICE, "no HIP device visible");
    int *S = nullptr;
    bp_status r = scratch_for_current_device(&S);
    if (r != BP_OK) return r;
    if (radix_counts_len(n > 0 ? n : 1) > kCountsInts) return fail(BP_ERR_CAPACITY, "digit counts exceed the scratch");
    hipStream_t st = (hipStream_t)hip_stream;
    int *kb = S, *vb = S + kNoAllocMaxN, *counts = S + 2 * (size_t)kNoAllocMaxN, *totals = counts + kCountsInts;
    if (cell_start || cell_end)
        hipLaunchKernelGGL(k_clear_ranges, dim3(grid_blocks(cells)), dim3(kBlock), 0, st, cell_start, cell_end,
                           (int)cells);
    if (n > 0) {
        hipLaunchKernelGGL(k_hash_positions, dim3(grid_blocks(n)), dim3(kBlock), 0, st, (const float4 *)positions,
                           keys, n, G);
        int *sk = nullptr, *pm = nullptr;
        // keys and perm are the sort's first buffers, the scratch its second
        launch_radix_sort(keys, kb, perm, vb, counts, totals, n, ilog2((unsigned)cells), &sk, &pm, st);
        if (sk != keys) hipLaunchKernelGGL(k_copy_pairs, dim3(grid_blocks(n)), dim3(kBlock), 0, st, sk, pm, keys, perm, n);
        if (cell_start || cell_end)
            hipLaunchKernelGGL(k_find_ranges_split, dim3(grid_blocks(n)), dim3(kBlock), 0, st, keys, n, cell_start,
                               cell_end);
    }
    HIP_TRY(hipGetLastError());
    return BP_OK;
}
