// noalloc.hip -- the allocation-free variant of the CFL time step
// (RiemannBase.compute_time_step, riemann.pyx:120-188), for callers that
// cannot allocate device memory (harnesses that forbid hipMalloc, serving
// loops that must not allocate).  It needs no handle: the per-block partial
// minima live in a file-scope device array of fixed size, resolved per
// device with hipGetSymbolAddress on the first call there.
//
// The minimum is formed in two deterministic passes (a per-block minimum,
// then one block over the partials), with the same per-cell bound and the
// same bit-pattern ordering as phd_compute_time_step, so both return the
// same bits.  Every call on one GPU shares that GPU's array: calls that
// overlap on one GPU race.
#include <mutex>
#include <string.h>
#include "flux_common.h"

// partial minima (bit patterns) of up to NOALLOC_MAX_BLOCKS blocks of
// PHD_BLOCK cells, then the final minimum: 512 KiB per GPU, held until the
// process ends
#define NOALLOC_MAX_BLOCKS 65536
static __device__ __align__(256) unsigned long long scratch[NOALLOC_MAX_BLOCKS + 1];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;

std::mutex g_scratch_mutex;
unsigned long long* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// the device address of `scratch` on the current device (resolved once per
// device), or nullptr when the lookup fails
unsigned long long* scratch_on(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
    }
    return g_scratch_ptr[dev];
}

// pass 1: each block's minimum CFL bound over its real cells (a wave fmin,
// then across the block's waves), as a bit pattern into part[blockIdx.x]
template <int D>
__global__ __launch_bounds__(PHD_BLOCK) void k_cfl_partial(phd_cells c, double gamma, int boost,
                                                           unsigned long long* __restrict__ part)
{
    const long long i = (long long)blockIdx.x * PHD_BLOCK + threadIdx.x;
    double v = __longlong_as_double((long long)INF_BITS);
    if (i < c.n_cells && c.tag[i] == 0) v = cfl_value<D>(c, i, gamma, boost);
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    __shared__ double wmin[PHD_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmin[0];
        for (int w = 1; w < PHD_BLOCK / 64; w++) b = fmin(b, wmin[w]);
        part[blockIdx.x] = (unsigned long long)__double_as_longlong(b);
    }
}

// pass 2: one block folds the n partials (unsigned minimum of the bit
// patterns: non-negative doubles order as their bits) into part[MAX]
__global__ __launch_bounds__(PHD_BLOCK) void k_cfl_final(unsigned long long* __restrict__ part,
                                                         int n)
{
    unsigned long long m = INF_BITS;
    for (int b = threadIdx.x; b < n; b += PHD_BLOCK) {
        const unsigned long long x = part[b];
        m = x < m ? x : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    __shared__ unsigned long long wmin[PHD_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < PHD_BLOCK / 64; w++) m = wmin[w] < m ? wmin[w] : m;
        m = wmin[0] < m ? wmin[0] : m;
        part[NOALLOC_MAX_BLOCKS] = m;
    }
}

} // namespace

extern "C" {

int phd_noalloc_max_cells(void) { return NOALLOC_MAX_BLOCKS * PHD_BLOCK; }

int phd_compute_time_step_noalloc(const phd_cells* c, double gamma, double cfl, int boost,
                                  void* stream, double* dt_out)
{
    if (!c || (c->dim != 2 && c->dim != 3) || c->n_cells <= 0 || !dt_out || !c->tag)
        return PHD_E_ARG;
    const long long nb = ((long long)c->n_cells + PHD_BLOCK - 1) / PHD_BLOCK;
    if (nb > NOALLOC_MAX_BLOCKS) return PHD_E_SIZE;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PHD_E_HIP;
    unsigned long long* part = scratch_on(dev);
    if (!part) return PHD_E_HIP;
    hipStream_t s = (hipStream_t)stream;
    if (c->dim == 2) k_cfl_partial<2><<<(unsigned)nb, PHD_BLOCK, 0, s>>>(*c, gamma, boost, part);
    else k_cfl_partial<3><<<(unsigned)nb, PHD_BLOCK, 0, s>>>(*c, gamma, boost, part);
    k_cfl_final<<<1, PHD_BLOCK, 0, s>>>(part, (int)nb);
    if (hipPeekAtLastError() != hipSuccess) { (void)hipGetLastError(); return PHD_E_HIP; }
    unsigned long long bits = 0;
    if (hipMemcpyAsync(&bits, part + NOALLOC_MAX_BLOCKS, sizeof(bits), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return PHD_E_HIP;
    double m;
    memcpy(&m, &bits, sizeof(m));
    *dt_out = cfl * m;
    return 0;
}

} // extern "C"
