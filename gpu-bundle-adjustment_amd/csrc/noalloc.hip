// noalloc.hip -- dense SPD factor + solve on caller-owned device buffers with
// no allocation at all: the variant of gba_dense_spd_solve for callers that
// may not call hipMalloc (code under HIP graph capture, serving loops).
//
// The factor and the triangular solves need device memory beyond the inputs
// and outputs (the padded row-major copy of A that the tile factor overwrites
// with L and L^T, the diagonal tiles and their inverses, the forward-solve
// vector, the tile-row flags and tickets, the failure flags).  Here it is one
// file-scope device array, `scratch`, sized for n <= kNoAllocMaxN; each GPU
// holds its own copy for the life of the process.  Host code never passes
// the array's name to a kernel (in host code it is only a stand-in symbol):
// on the first call on a device, with that device current, its device
// address is taken with hipGetSymbolAddress and kept per device, and that
// address is what the launches receive.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// device (two streams, two host threads) race and must be serialised by the
// caller.  Each call clears the flags, tickets and padding it relies on with
// kernels (no hipMemsetAsync), so nothing carries over from the previous call.
#include <mutex>

#include "device_common.hpp"
#include "gba_internal.hpp"
#include "launchers.hpp"
#include "solver.hpp"

namespace gba_amd {

namespace {

constexpr int kNoAllocMaxPad = (kNoAllocMaxN + kTile - 1) / kTile * kTile;
constexpr int kNoAllocMaxBlk = kNoAllocMaxPad / kTile;

// scratch layout (byte offsets, each a multiple of 256)
constexpr size_t kOffS = 0;                                                     // npad^2 doubles
constexpr size_t kOffD = kOffS + sizeof(double) * kNoAllocMaxPad * kNoAllocMaxPad;  // 2 nblk tiles
constexpr size_t kOffRhs = kOffD + sizeof(double) * 2 * kNoAllocMaxBlk * kTile * kTile;
constexpr size_t kOffY = kOffRhs + sizeof(double) * kNoAllocMaxPad;
constexpr size_t kOffX = kOffY + sizeof(double) * kNoAllocMaxPad;
constexpr size_t kOffSync = kOffX + sizeof(double) * kNoAllocMaxPad;            // 2 nblk + 2 ints
constexpr size_t kOffFlags = kOffSync + 256;                                    // TrialFlags
constexpr size_t kScratchBytes = kOffFlags + 256;
static_assert(sizeof(int) * (2 * kNoAllocMaxBlk + 2) <= 256, "sync words exceed their slot");
static_assert(sizeof(TrialFlags) <= 256, "flags exceed their slot");

}  // namespace

// the allocation-free entry point's scratch (one copy per GPU)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's scratch array (looked up once per device).
unsigned char* scratch_address() {
    int dev = -1;
    GBA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("gba_dense_spd_solve_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(std::string("gba_dense_spd_solve_noalloc: hipGetSymbolAddress failed: ") +
                        hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch[dev]);
}

// S (npad x npad, row-major) from column-major A (n x n, lower triangle
// read): one 64x64 tile per workgroup, read down A's columns, written along
// S's rows through LDS; the padding is the identity, the upper triangle of
// the n x n part is not read (the factor reads only S's lower triangle).
__global__ __launch_bounds__(256) void k_na_pack(int n, int npad, const double* __restrict__ A,
                                                  double* __restrict__ S) {
    __shared__ double t[kTile][kTile + 1];
    const int nb = npad / kTile;
    const int bi = blockIdx.x / nb, bj = blockIdx.x % nb;  // S tile (row block, column block)
    const int r0 = bi * kTile, c0 = bj * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += 256) {
        const int lr = q & 63, lc = q >> 6;  // consecutive threads walk down a column of A
        const int r = r0 + lr, c = c0 + lc;
        double v;
        if (r < n && c < n) v = (c <= r) ? A[static_cast<size_t>(c) * n + r] : 0.0;
        else v = (r == c) ? 1.0 : 0.0;
        t[lr][lc] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kTile * kTile; q += 256) {
        const int lr = q >> 6, lc = q & 63;
        S[static_cast<size_t>(r0 + lr) * npad + c0 + lc] = t[lr][lc];
    }
}

__global__ void k_na_rhs(int n, int npad, const double* __restrict__ b, double* __restrict__ rhs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < npad) rhs[i] = i < n ? b[i] : 0.0;
}

// ok = 1 and b = x after a successful solve; ok = 0 (b unchanged) when A is
// not positive definite; ok = -1 (b unchanged) when a triangular solve gave up
// waiting for a tile row (never expected).
__global__ void k_na_finish(int n, const double* __restrict__ x, const TrialFlags* __restrict__ flags,
                            double* __restrict__ b, int* __restrict__ ok) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int status = flags->chol_fail ? 0 : (flags->solve_err ? -1 : 1);
    if (i == 0) *ok = status;
    if (status == 1 && i < n) b[i] = x[i];
}

}  // namespace

void dense_spd_solve_noalloc(int n, const double* A, double* b, int* ok, hipStream_t st) {
    if (n <= 0 || n > kNoAllocMaxN)
        throw Error("gba_dense_spd_solve_noalloc: n must be in 1.." + std::to_string(kNoAllocMaxN) +
                    " (the size of the library's static scratch), got " + std::to_string(n));
    unsigned char* base = scratch_address();
    const int npad = (n + kTile - 1) / kTile * kTile, nblk = npad / kTile;
    double* S = reinterpret_cast<double*>(base + kOffS);
    double* D = reinterpret_cast<double*>(base + kOffD);
    double* rhs = reinterpret_cast<double*>(base + kOffRhs);
    double* y = reinterpret_cast<double*>(base + kOffY);
    double* x = reinterpret_cast<double*>(base + kOffX);
    int* sync = reinterpret_cast<int*>(base + kOffSync);
    TrialFlags* flags = reinterpret_cast<TrialFlags*>(base + kOffFlags);
    hipLaunchKernelGGL(k_na_pack, dim3(nblk * nblk), dim3(256), 0, st, n, npad, A, S);
    hipLaunchKernelGGL(k_na_rhs, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, static_cast<const double*>(b),
                       rhs);
    launch_clear_flags(st, flags);
    // the factor (S holds L and L^T afterwards, D the diagonal tiles and
    // their inverses) and both substitutions; launch_chol_solve clears its
    // flags and tickets with a fill kernel
    launch_cholesky(st, S, nullptr, npad, nblk, D, nullptr, flags);
    launch_chol_solve(st, S, nullptr, npad, nblk, D, nullptr, rhs, y, x, sync, flags);
    hipLaunchKernelGGL(k_na_finish, dim3((n + 255) / 256), dim3(256), 0, st, n, static_cast<const double*>(x),
                       static_cast<const TrialFlags*>(flags), b, ok);
    GBA_HIP(hipGetLastError());
}

}  // namespace gba_amd
