// Launch-shape helpers shared by the column-marching kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>

namespace swh {

// Compute units of the current device (cached per device).
inline int device_cus()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Rows per strip for a column-marching kernel over one grid (each wavefront owns 'own'
// columns and marches its strip row by row, recomputing 'halo' rows next to it). Such a
// kernel is latency-bound, so it wants as many resident wavefronts as fit in one round
// of workgroups: the fewest rows whose launch fills about 70 % of the device's
// resident-wave capacity for this kernel (workgroups per CU from its register use,
// cached in occ[block_x / 64]), at least 2 * halo rows.
inline int occupancy_rows(const void *kernel, std::atomic<int> *occ, int nx, int ny, int block_x, int own, int halo)
{
    const int wpb = block_x / 64;
    int nb = occ[wpb].load(std::memory_order_relaxed);
    if (nb <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block_x, 0) != hipSuccess || nb <= 0) nb = 1;
        occ[wpb].store(nb, std::memory_order_relaxed);
    }
    const double target = 0.7 * nb * wpb * device_cus();
    const double waves_x = (double)((nx + own - 1) / own);
    const int rows = (int)std::ceil(ny * waves_x / target);
    return std::min(std::max(rows, std::max(4, 2 * halo)), 1024);
}

}  // namespace swh
