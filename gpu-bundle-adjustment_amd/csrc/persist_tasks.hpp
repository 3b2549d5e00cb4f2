// persist_tasks.hpp -- task list and dependency rules of the one-launch tile
// Cholesky (k_chol_persistent in k_dense.hip).  Plain C++ shared by the
// kernel and by the host-thread emulator that checks the protocol on the CPU
// (tests/cpp/persist_emulator.cpp), so both run exactly these rules.
//
// Matrix of nblk x nblk tiles (lower triangle).  Workgroup 0 runs the
// critical chain, one step per tile column k:
//   step k (k > 0): TRSM of tile (k, k-1) with Linv_{k-1}, then the update of
//   the diagonal tile (k, k) by that column, then (every k) the factor and
//   inverse of tile (k, k).
// Every other workgroup takes tasks through one atomic ticket, in this order,
// column by column (m = nblk - 1 - k tiles below the diagonal):
//   TRSM(i, k)      i = k+2 .. nblk-1        L_ik = A_ik Linv_kk^T
//   UPD(i, j, k)    k+1 <= j <= i, (i, j) != (k+1, k+1), row-major
//                                            A_ij -= L_ik L_jk^T
//   inverse roles of column k (FP32 form only), X = L^-1 row by row:
//     XDIAG(k)      X_kk = Linv_kk
//     XROW(k, j)    j < k: X_kj = -Linv_kk (L_{k,k-1} X_{k-1,j} + R_kj)
//     R(i, j, k)    i > k, j < k: R_ij (+)= L_{i,k-1} X_{k-1,j}
//   (R_ij holds sum_{m = j}^{k-1} L_im X_mj after the R roles of columns
//   j+1 .. k.)
// Progress flags (int arrays, cleared before each factorisation):
//   DIAG[k]        1 when L_kk and Linv_kk are stored
//   TRSM[i*n+k]    1 when L_ik (and L_ik^T) are stored
//   VER[i*n+j]     number of columns applied to tile (i, j)
//   XDONE[k*n+j]   1 when X_kj is stored
//   RVER[i*n+j]    number of terms summed into R_ij
// Every task waits only on the critical chain or on tasks with smaller
// tickets, and the critical step k waits only on tasks of columns < k - 1,
// which have smaller tickets than any task that waits on step k; a ticket is
// only ever held by a running workgroup, so the launch completes whether or
// not all its workgroups are resident at once.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define GBA_HD __host__ __device__
#else
#define GBA_HD
#endif

namespace gba_amd {

enum PTaskType { kPTrsm = 0, kPUpd = 1, kPXDiag = 2, kPXRow = 3, kPR = 4 };
enum PFlagArray { kPDiag = 0, kPTrsmDone = 1, kPVer = 2, kPXDone = 3, kPRVer = 4, kPNumArrays = 5 };

struct PTask {
    int type, i, j, k;
};

struct PFlag {
    int arr, idx, want;  // wait: flags[arr][idx] >= want; publish: flags[arr][idx] = want
};

// inverse roles of column k: the diagonal tile, k tiles of row k of X, and
// (nblk - 1 - k) * k R tiles
GBA_HD inline int ptask_inv_roles(int k, int nblk) { return (k + 1) + (nblk - 1 - k) * k; }

GBA_HD inline int ptask_col_trsm(int k, int nblk) {
    const int m = nblk - 1 - k;
    return m >= 1 ? m - 1 : 0;
}
GBA_HD inline int ptask_col_upd(int k, int nblk) {
    const int m = nblk - 1 - k;
    return m >= 1 ? m * (m + 1) / 2 - 1 : 0;
}
GBA_HD inline long long ptask_col_count(int k, int nblk, bool inverse) {
    return static_cast<long long>(ptask_col_trsm(k, nblk)) + ptask_col_upd(k, nblk) +
           (inverse ? ptask_inv_roles(k, nblk) : 0);
}

// (ii, jj), jj <= ii, of index u in the row-major order of a lower triangle
GBA_HD inline void ptask_tri(long long u, int& ii, int& jj) {
    // r (r + 1) / 2 <= u < (r + 1) (r + 2) / 2: a floating-point estimate,
    // then corrected by integer steps
    int r = static_cast<int>((sqrt(8.0 * static_cast<double>(u) + 1.0) - 1.0) * 0.5);
    while (r > 0 && static_cast<long long>(r) * (r + 1) / 2 > u) --r;
    while (static_cast<long long>(r + 1) * (r + 2) / 2 <= u) ++r;
    ii = r;
    jj = static_cast<int>(u - static_cast<long long>(r) * (r + 1) / 2);
}

// Task of ticket t; col_start[k] is the first ticket of column k
// (col_start[nblk] = the task count).
GBA_HD inline PTask ptask_decode(long long t, const long long* col_start, int nblk, bool inverse) {
    int lo = 0, hi = nblk;  // col_start[lo] <= t < col_start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (col_start[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int k = lo;
    long long local = t - col_start[k];
    const int nt = ptask_col_trsm(k, nblk), nu = ptask_col_upd(k, nblk);
    PTask p{0, 0, 0, k};
    if (local < nt) {
        p.type = kPTrsm;
        p.i = k + 2 + static_cast<int>(local);
        return p;
    }
    local -= nt;
    if (local < nu) {
        int ii, jj;
        ptask_tri(local + 1, ii, jj);  // index 0, tile (k+1, k+1), is the critical chain's
        p.type = kPUpd;
        p.i = k + 1 + ii;
        p.j = k + 1 + jj;
        return p;
    }
    local -= nu;
    const int r = static_cast<int>(local);
    if (r == k) {
        p.type = kPXDiag;
        p.i = p.j = k;
    } else if (r < k) {
        p.type = kPXRow;
        p.i = k;
        p.j = r;
    } else {
        const int b = r - (k + 1);
        p.type = kPR;
        p.i = k + 1 + b / k;
        p.j = b % k;
    }
    return p;
}

// What task p waits for (at most 5 conditions); returns their number.
GBA_HD inline int ptask_deps(const PTask& p, int nblk, PFlag* d) {
    const int n = nblk, k = p.k;
    int c = 0;
    switch (p.type) {
        case kPTrsm:
            d[c++] = PFlag{kPDiag, k, 1};
            d[c++] = PFlag{kPVer, p.i * n + k, k};
            break;
        case kPUpd:
            d[c++] = PFlag{kPTrsmDone, p.i * n + k, 1};
            d[c++] = PFlag{kPTrsmDone, p.j * n + k, 1};
            d[c++] = PFlag{kPVer, p.i * n + p.j, k};
            break;
        case kPXDiag:
            d[c++] = PFlag{kPDiag, k, 1};
            break;
        case kPXRow:  // i == k
            d[c++] = PFlag{kPDiag, k, 1};
            d[c++] = PFlag{kPTrsmDone, k * n + (k - 1), 1};
            d[c++] = PFlag{kPXDone, (k - 1) * n + p.j, 1};
            if (p.j < k - 1) d[c++] = PFlag{kPRVer, k * n + p.j, k - 1 - p.j};
            break;
        case kPR:
            d[c++] = PFlag{kPTrsmDone, p.i * n + (k - 1), 1};
            d[c++] = PFlag{kPXDone, (k - 1) * n + p.j, 1};
            if (p.j < k - 1) d[c++] = PFlag{kPRVer, p.i * n + p.j, k - 1 - p.j};
            break;
        default:
            break;
    }
    return c;
}

// The flag task p sets when its results are stored.
GBA_HD inline PFlag ptask_publish(const PTask& p, int nblk) {
    const int n = nblk, k = p.k;
    switch (p.type) {
        case kPTrsm: return PFlag{kPTrsmDone, p.i * n + k, 1};
        case kPUpd: return PFlag{kPVer, p.i * n + p.j, k + 1};
        case kPXDiag: return PFlag{kPXDone, k * n + k, 1};
        case kPXRow: return PFlag{kPXDone, k * n + p.j, 1};
        default: return PFlag{kPRVer, p.i * n + p.j, k - p.j};
    }
}

// Critical step k waits for tiles (k, k-1) and (k, k) to hold every column
// before k - 1 (the step applies column k - 1 to (k, k) itself).
GBA_HD inline PFlag pcrit_wait_offdiag(int k, int nblk) { return PFlag{kPVer, k * nblk + (k - 1), k - 1}; }
GBA_HD inline PFlag pcrit_wait_diag(int k, int nblk) { return PFlag{kPVer, k * nblk + k, k - 1}; }
GBA_HD inline PFlag pcrit_publish_trsm(int k, int nblk) { return PFlag{kPTrsmDone, k * nblk + (k - 1), 1}; }
GBA_HD inline PFlag pcrit_publish_diag(int k) { return PFlag{kPDiag, k, 1}; }

// Host: first ticket of each column (nblk + 1 entries).
inline void ptask_col_start(int nblk, bool inverse, long long* col_start) {
    col_start[0] = 0;
    for (int k = 0; k < nblk; ++k) col_start[k + 1] = col_start[k] + ptask_col_count(k, nblk, inverse);
}

}  // namespace gba_amd
