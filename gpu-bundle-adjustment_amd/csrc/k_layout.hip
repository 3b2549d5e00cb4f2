// k_layout.hip -- the point-major observation layout built on the device
// (replaces the reference's host point -> edge lists, src/lm.cu:378-381, with
// a CSR over observations sorted by (point, camera, index), plus a camera
// index over the sorted observations).
//
//   perm    : sorted position k -> caller's observation index e, ordered by
//             (point, camera, e): a radix sort of the keys point:camera with
//             the indices as values (the sort is stable, so ties keep e's
//             order);
//   scam/spt/spx : camera, point, pixels of sorted observation k;
//   pt_off  : the point CSR (P + 1), from the boundaries of the sorted points;
//   cam_list/cam_off : sorted positions k grouped by camera, ascending k
//             within a camera (a second stable radix sort, keys scam[k]),
//             and the camera CSR (C + 1) from the sorted cameras.
#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "launchers.hpp"
#include "solver.hpp"

namespace gba_amd {

namespace {

__global__ void k_layout_keys(int O, int cbits, const int* __restrict__ cam, const int* __restrict__ pt,
                              unsigned long long* __restrict__ keys, int* __restrict__ idx) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= O) return;
    keys[e] = (static_cast<unsigned long long>(pt[e]) << cbits) | static_cast<unsigned long long>(cam[e]);
    idx[e] = e;
}

// CSR offsets of n sorted keys in [0, nkeys): off[c] = first position whose
// key is >= c, off[nkeys] = n.  Thread k in [0, n] writes the offsets of the
// keys in (key[k-1], key[k]], so every entry is written once.
__global__ void k_csr_offsets(int n, int nkeys, const int* __restrict__ key, int* __restrict__ off) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const int prev = k == 0 ? -1 : key[k - 1];
    const int cur = k == n ? nkeys : key[k];
    for (int c = prev + 1; c <= cur; ++c) off[c] = k;
}

__global__ void k_layout_gather(int O, const int* __restrict__ perm, const int* __restrict__ cam,
                                const int* __restrict__ pt, const double* __restrict__ px,
                                int* __restrict__ scam, int* __restrict__ spt, double2* __restrict__ spx,
                                int* __restrict__ iota) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= O) return;
    const int e = perm[k];
    scam[k] = cam[e];
    spt[k] = pt[e];
    if (spx) spx[k] = make_double2(px[2 * static_cast<size_t>(e)], px[2 * static_cast<size_t>(e) + 1]);
    iota[k] = k;
}

// Problem-local camera / point indices of a batch's concatenated
// observations to batch indices: observation e of problem q (o_off[q] <= e <
// o_off[q + 1], found by binary search) gets + c0[q], + p0[q].
__global__ void k_offset_obs(int O, int count, const long long* __restrict__ o_off, const int* __restrict__ c0,
                             const int* __restrict__ p0, int* __restrict__ cam, int* __restrict__ pt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= O) return;
    int lo = 0, hi = count;  // o_off[lo] <= e < o_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (o_off[mid] <= e) lo = mid;
        else hi = mid;
    }
    cam[e] += c0[lo];
    pt[e] += p0[lo];
}

// Sets *flag when two adjacent sorted keys are equal (a point observed twice
// by one camera).
__global__ void k_repeated_keys(int O, const unsigned long long* __restrict__ key, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 1 && k < O && key[k] == key[k - 1]) atomicOr(flag, 1);
}

int bits_for(int n) {
    int b = 1;
    while ((1LL << b) < n) ++b;
    return b;
}

}  // namespace

void build_layout_device(hipStream_t st, int C, int P, int O, const int* d_cam, const int* d_pt,
                         const double* d_px, DevBuf<int>& perm, DevBuf<int>& scam, DevBuf<int>& spt,
                         DevBuf<double2>* spx, DevBuf<int>& pt_off, DevBuf<int>& cam_off, DevBuf<int>& cam_list,
                         int* repeated) {
    const int cbits = bits_for(C), pbits = bits_for(P);
    if (cbits + pbits > 62) throw Error("layout: camera and point indices too wide for the sort key");
    DevBuf<unsigned long long> keys, keys_sorted;
    DevBuf<int> idx, iota, cam_keys_sorted, flag;
    keys.alloc(static_cast<size_t>(O));
    keys_sorted.alloc(static_cast<size_t>(O));
    idx.alloc(static_cast<size_t>(O));
    iota.alloc(static_cast<size_t>(O));
    cam_keys_sorted.alloc(static_cast<size_t>(O));
    perm.alloc(static_cast<size_t>(O));
    scam.alloc(static_cast<size_t>(O));
    spt.alloc(static_cast<size_t>(O));
    if (spx) spx->alloc(static_cast<size_t>(O));
    pt_off.alloc(static_cast<size_t>(P) + 1);
    cam_off.alloc(static_cast<size_t>(C) + 1);
    cam_list.alloc(static_cast<size_t>(O));
    // scratch of the sorts
    size_t b1 = 0, b2 = 0;
    GBA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, keys.p, keys_sorted.p, idx.p, perm.p, O, 0,
                                               cbits + pbits, st));
    GBA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b2, scam.p, cam_keys_sorted.p, iota.p, cam_list.p, O, 0,
                                               cbits, st));
    const size_t wb = std::max(b1, b2);
    DevBuf<unsigned char> work;
    work.alloc(std::max<size_t>(wb, 1));
    const int blocks = (O + 255) / 256, blocks1 = (O + 1 + 255) / 256;
    hipLaunchKernelGGL(k_layout_keys, dim3(blocks), dim3(256), 0, st, O, cbits, d_cam, d_pt, keys.p, idx.p);
    size_t bytes = wb;
    GBA_HIP(hipcub::DeviceRadixSort::SortPairs(work.p, bytes, keys.p, keys_sorted.p, idx.p, perm.p, O, 0,
                                               cbits + pbits, st));
    hipLaunchKernelGGL(k_layout_gather, dim3(blocks), dim3(256), 0, st, O, static_cast<const int*>(perm.p), d_cam,
                       d_pt, spx ? d_px : nullptr, scam.p, spt.p, spx ? spx->p : nullptr, iota.p);
    bytes = wb;
    GBA_HIP(hipcub::DeviceRadixSort::SortPairs(work.p, bytes, scam.p, cam_keys_sorted.p, iota.p, cam_list.p, O, 0,
                                               cbits, st));
    flag.alloc(1);
    launch_fill_int(st, flag.p, 1, 0);
    hipLaunchKernelGGL(k_repeated_keys, dim3(blocks), dim3(256), 0, st, O,
                       static_cast<const unsigned long long*>(keys_sorted.p), flag.p);
    hipLaunchKernelGGL(k_csr_offsets, dim3(blocks1), dim3(256), 0, st, O, P, static_cast<const int*>(spt.p),
                       pt_off.p);
    hipLaunchKernelGGL(k_csr_offsets, dim3(blocks1), dim3(256), 0, st, O, C,
                       static_cast<const int*>(cam_keys_sorted.p), cam_off.p);
    int h_flag = 0;
    GBA_HIP(hipMemcpyAsync(&h_flag, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st));  // the temporaries are released on return
    if (repeated) *repeated = h_flag;
}

void launch_offset_obs(hipStream_t st, int O, int count, const long long* o_off, const int* c0, const int* p0,
                       int* cam, int* pt) {
    if (O <= 0) return;
    hipLaunchKernelGGL(k_offset_obs, dim3((O + 255) / 256), dim3(256), 0, st, O, count, o_off, c0, p0, cam, pt);
}

__global__ void k_cam_order_copies(int O, const int* __restrict__ cam_list, const int* __restrict__ spt,
                                   const double2* __restrict__ spx, int* __restrict__ cspt,
                                   double2* __restrict__ cspx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= O) return;
    const int e = cam_list[i];
    cspt[i] = spt[e];
    if (cspx) cspx[i] = spx[e];
}

void build_cam_order_copies(hipStream_t st, int O, const int* cam_list, const int* spt, const double2* spx,
                            DevBuf<int>& cspt, DevBuf<double2>& cspx) {
    cspt.alloc(std::max(1, O));
    if (spx) cspx.alloc(std::max(1, O));
    if (O > 0)
        hipLaunchKernelGGL(k_cam_order_copies, dim3((O + 255) / 256), dim3(256), 0, st, O, cam_list, spt, spx, cspt.p,
                           spx ? cspx.p : nullptr);
    GBA_HIP(hipGetLastError());
}

}  // namespace gba_amd
