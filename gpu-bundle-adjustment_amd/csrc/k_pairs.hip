// k_pairs.hip -- the Schur assembly's observation-pair lists, built on the
// device from the sorted layout (replaces the host loops that enumerate the
// reference's Phase 2 pairs, src/lm.cu:515-559: for each point, each pair of
// its observations e, f with camera(f) <= camera(e)).
//
// Task t = (a, lo, hi) owns the pairs (e, f) with camera(e) = a and camera(f)
// in [lo, hi); its slots qoff[t] .. qoff[t] + n_a - 1 are camera a's
// observations in index order (slot j <-> cam_list[cam_off[a] + j]), so the
// pairs of a task come grouped by e (ascending point) and ordered by f.  Two passes over the same loops, one thread
// per slot: count per slot, exclusive scan (write positions), fill.  Points
// with at least heavy_k observations (heavy_k > 0) have no pairs: their share
// of S is formed densely (k_heavy.hip).
#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "launchers.hpp"
#include "solver.hpp"

namespace gba_amd {

namespace {

// One thread per slot q (the task found by binary search in qoff).  kFill =
// false: cnt[q] = number of pairs of slot q (and cnt[nslots] = 0, the scan's
// trailing element); kFill = true: writes slot q's pairs from position pos[q].
template <bool kFill>
__global__ __launch_bounds__(256) void k_pairs(int ntask, const int4* __restrict__ tq,
                                               const long long* __restrict__ qoff, long long nslots,
                                               const int* __restrict__ cam_off, const int* __restrict__ cam_list,
                                               const int* __restrict__ spt, const int* __restrict__ pt_off,
                                               const int* __restrict__ scam, int heavy_k, long long* __restrict__ cnt,
                                               const long long* __restrict__ pos, int2* __restrict__ pairs) {
    const long long q = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (!kFill && q == nslots) cnt[nslots] = 0;
    if (q >= nslots) return;
    // the last task t with qoff[t] <= q (empty tasks share their successor's
    // first slot and are skipped); qoff[ntask] = nslots > q
    int lo = 0, hi = ntask;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (qoff[mid] <= q) lo = mid;
        else hi = mid;
    }
    const int4 d = tq[lo];  // camera a, lo, hi
    const int e = cam_list[cam_off[d.x] + static_cast<int>(q - qoff[lo])];
    const int p = spt[e];
    const int f1 = pt_off[p + 1];
    long long w = kFill ? pos[q] : 0;
    const int f0 = (heavy_k > 0 && f1 - pt_off[p] >= heavy_k) ? f1 : pt_off[p];
    // the point's observations are ordered by camera
    for (int f = f0; f < f1; ++f) {
        const int b = scam[f];
        if (b >= d.z) break;
        if (b < d.y) continue;
        if (kFill) pairs[w] = make_int2(e, f);
        ++w;
    }
    if (!kFill) cnt[q] = w;
}

__global__ void k_pair_bounds(int ntask, const long long* __restrict__ qoff, const long long* __restrict__ pos,
                              long long* __restrict__ bounds) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= ntask) bounds[t] = pos[qoff[t]];
}

}  // namespace

size_t pair_scan_work(long long n) {
    size_t bytes = 0;
    GBA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const long long*>(nullptr),
                                             static_cast<long long*>(nullptr), static_cast<int>(n)));
    return bytes;
}

void launch_pair_count(hipStream_t st, int ntask, const int4* tq, const long long* qoff, long long nslots,
                       const int* cam_off, const int* cam_list, const int* spt, const int* pt_off,
                       const int* scam, int heavy_k, long long* cnt) {
    if (ntask <= 0) return;
    const unsigned blocks = static_cast<unsigned>((nslots + 1 + 255) / 256);
    hipLaunchKernelGGL(k_pairs<false>, dim3(blocks), dim3(256), 0, st, ntask, tq, qoff, nslots, cam_off, cam_list,
                       spt, pt_off, scam, heavy_k, cnt, static_cast<const long long*>(nullptr),
                       static_cast<int2*>(nullptr));
}

void launch_pair_scan(hipStream_t st, const long long* cnt, long long* pos, long long n, void* work,
                      size_t work_bytes) {
    size_t bytes = work_bytes;
    GBA_HIP(hipcub::DeviceScan::ExclusiveSum(work, bytes, cnt, pos, static_cast<int>(n), st));
}

void launch_pair_bounds(hipStream_t st, int ntask, const long long* qoff, const long long* pos,
                        long long* bounds) {
    hipLaunchKernelGGL(k_pair_bounds, dim3((ntask + 1 + 255) / 256), dim3(256), 0, st, ntask, qoff, pos, bounds);
}

void launch_pair_fill(hipStream_t st, int ntask, const int4* tq, const long long* qoff, long long nslots,
                      const int* cam_off, const int* cam_list, const int* spt, const int* pt_off, const int* scam,
                      int heavy_k, const long long* pos, int2* pairs) {
    if (ntask <= 0 || nslots <= 0) return;
    const unsigned blocks = static_cast<unsigned>((nslots + 255) / 256);
    hipLaunchKernelGGL(k_pairs<true>, dim3(blocks), dim3(256), 0, st, ntask, tq, qoff, nslots, cam_off, cam_list,
                       spt, pt_off, scam, heavy_k, static_cast<long long*>(nullptr), pos, pairs);
}

}  // namespace gba_amd
