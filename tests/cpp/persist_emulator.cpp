// persist_emulator.cpp -- the one-launch tile Cholesky's protocol run by host
// threads (CPU check of k_chol_persistent before any GPU run).
//
// One thread plays workgroup 0 (the critical chain), W threads play the other
// workgroups: they take tickets from one atomic counter, decode them, wait on
// the flags ptask_deps() names and publish what ptask_publish() names -- the
// same rules, from the same header, as the kernel.  The arithmetic is the
// kernel's on tiles of tb x tb in double: S row-major with L_ik written below
// the diagonal and L_ik^T into the upper tile (k, i), diagonal tiles and their
// inverses in D / DinvT, X = L^-1 row by row with the R partial sums.  Every
// wait has a deadline; random yields shuffle the interleavings.
//
//   persist_emulator nblk tb workers seed inverse out.bin
// writes L (npad x npad, lower) and, with inverse = 1, X, both row-major
// doubles; exit 0 on success, 3 on a wait past its deadline (a deadlock).
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "persist_tasks.hpp"

using namespace gba_amd;

namespace {

int nblk, tb, npad;
bool inverse;
std::vector<double> S, D, DinvT, X, R;
std::vector<std::atomic<int>> flags[kPNumArrays];
std::atomic<long long> ticket{0};
std::atomic<int> failed{0};
std::vector<long long> col_start;

double* tile(std::vector<double>& m, int i, int j) { return m.data() + static_cast<size_t>(i) * tb * npad + j * tb; }
double& at(double* t, int r, int c) { return t[static_cast<size_t>(r) * npad + c]; }

void jitter(std::mt19937& rng) {
    if (rng() % 4 == 0) std::this_thread::yield();
    if (rng() % 64 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
}

bool wait(const PFlag& f) {
    const auto t0 = std::chrono::steady_clock::now();
    while (flags[f.arr][f.idx].load(std::memory_order_acquire) < f.want) {
        if (failed.load()) return false;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
            std::fprintf(stderr, "wait past its deadline: array %d index %d want %d have %d\n", f.arr, f.idx, f.want,
                         flags[f.arr][f.idx].load());
            failed.store(1);
            return false;
        }
        std::this_thread::yield();
    }
    return true;
}

void publish(const PFlag& f) { flags[f.arr][f.idx].store(f.want, std::memory_order_release); }

// L_ik = A_ik Linv_kk^T (Linv from DinvT[k]); L_ik below, L_ik^T in tile (k, i)
void trsm(int i, int k) {
    double* a = tile(S, i, k);
    const double* it = DinvT.data() + static_cast<size_t>(k) * tb * tb;  // it[c][q] = Linv[q][c]
    std::vector<double> l(static_cast<size_t>(tb) * tb);
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            double s = 0.0;
            for (int q = 0; q < tb; ++q) s += at(a, r, q) * it[q * tb + c];
            l[r * tb + c] = s;
        }
    double* up = tile(S, k, i);
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            at(a, r, c) = l[r * tb + c];
            at(up, c, r) = l[r * tb + c];
        }
}

// A_ij -= L_ik L_jk^T, both operands from the transposed tiles (k, i), (k, j)
void update(int i, int j, int k) {
    const double* ti = tile(S, k, i);
    const double* tj = tile(S, k, j);
    double* a = tile(S, i, j);
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            double s = 0.0;
            for (int q = 0; q < tb; ++q) s += ti[static_cast<size_t>(q) * npad + r] * tj[static_cast<size_t>(q) * npad + c];
            at(a, r, c) -= s;
        }
}

// factor + invert diagonal tile k in place: D[k] = L_kk, DinvT[k] = Linv^T
bool factor_diag(int k) {
    double* a = tile(S, k, k);
    std::vector<double> L(static_cast<size_t>(tb) * tb, 0.0), Li(static_cast<size_t>(tb) * tb, 0.0);
    for (int j = 0; j < tb; ++j) {
        double d = at(a, j, j);
        for (int q = 0; q < j; ++q) d -= L[j * tb + q] * L[j * tb + q];
        if (!(d > 0.0)) return false;
        L[j * tb + j] = std::sqrt(d);
        for (int r = j + 1; r < tb; ++r) {
            double s = at(a, r, j);
            for (int q = 0; q < j; ++q) s -= L[r * tb + q] * L[j * tb + q];
            L[r * tb + j] = s / L[j * tb + j];
        }
    }
    for (int c = 0; c < tb; ++c)
        for (int r = c; r < tb; ++r) {
            double s = (r == c) ? 1.0 : 0.0;
            for (int q = c; q < r; ++q) s -= L[r * tb + q] * Li[q * tb + c];
            Li[r * tb + c] = s / L[r * tb + r];
        }
    double* dd = D.data() + static_cast<size_t>(k) * tb * tb;
    double* di = DinvT.data() + static_cast<size_t>(k) * tb * tb;
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            dd[r * tb + c] = L[r * tb + c];
            di[r * tb + c] = Li[c * tb + r];
            at(a, r, c) = c <= r ? L[r * tb + c] : 0.0;
        }
    return true;
}

void critical(unsigned seed) {
    std::mt19937 rng(seed);
    for (int k = 0; k < nblk && !failed.load(); ++k) {
        if (k > 0) {
            if (!wait(pcrit_wait_offdiag(k, nblk))) return;
            trsm(k, k - 1);
            publish(pcrit_publish_trsm(k, nblk));
            jitter(rng);
            if (!wait(pcrit_wait_diag(k, nblk))) return;
            update(k, k, k - 1);
        }
        if (!factor_diag(k)) {
            std::fprintf(stderr, "not positive definite at tile %d\n", k);
            failed.store(2);
            return;
        }
        publish(pcrit_publish_diag(k));
        jitter(rng);
    }
}

// X / R roles (L_{i,k-1} read from its transposed copy in tile (k-1, i))
void inv_role(const PTask& p) {
    const int k = p.k, i = p.i, j = p.j;
    if (p.type == kPXDiag) {
        const double* it = DinvT.data() + static_cast<size_t>(k) * tb * tb;
        double* x = tile(X, k, k);
        for (int r = 0; r < tb; ++r)
            for (int c = 0; c < tb; ++c) at(x, r, c) = it[c * tb + r];
        return;
    }
    const double* lt = tile(S, k - 1, i);
    const double* xm = tile(X, k - 1, j);
    std::vector<double> acc(static_cast<size_t>(tb) * tb);
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            double s = 0.0;
            for (int q = 0; q < tb; ++q) s += lt[static_cast<size_t>(q) * npad + r] * xm[static_cast<size_t>(q) * npad + c];
            if (j < k - 1) s += at(tile(R, i, j), r, c);
            acc[r * tb + c] = s;
        }
    if (p.type == kPR) {
        for (int r = 0; r < tb; ++r)
            for (int c = 0; c < tb; ++c) at(tile(R, i, j), r, c) = acc[r * tb + c];
        return;
    }
    const double* it = DinvT.data() + static_cast<size_t>(k) * tb * tb;  // it[c][q] = Linv[q][c]
    double* x = tile(X, k, j);
    for (int r = 0; r < tb; ++r)
        for (int c = 0; c < tb; ++c) {
            double s = 0.0;
            for (int q = 0; q < tb; ++q) s += it[q * tb + r] * acc[q * tb + c];
            at(x, r, c) = -s;
        }
}

void worker(unsigned seed) {
    std::mt19937 rng(seed);
    const long long ntasks = col_start[nblk];
    for (;;) {
        if (failed.load()) return;
        const long long t = ticket.fetch_add(1);
        if (t >= ntasks) return;
        const PTask p = ptask_decode(t, col_start.data(), nblk, inverse);
        PFlag d[5];
        const int nd = ptask_deps(p, nblk, d);
        for (int q = 0; q < nd; ++q)
            if (!wait(d[q])) return;
        jitter(rng);
        switch (p.type) {
            case kPTrsm: trsm(p.i, p.k); break;
            case kPUpd: update(p.i, p.j, p.k); break;
            default: inv_role(p); break;
        }
        publish(ptask_publish(p, nblk));
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: persist_emulator nblk tb workers seed inverse out.bin\n");
        return 1;
    }
    nblk = std::atoi(argv[1]);
    tb = std::atoi(argv[2]);
    const int workers = std::atoi(argv[3]);
    const unsigned seed = static_cast<unsigned>(std::atoi(argv[4]));
    inverse = std::atoi(argv[5]) != 0;
    npad = nblk * tb;
    const size_t nn = static_cast<size_t>(npad) * npad;
    // SPD input: M^T M / n + I/2 from a seeded generator, lower triangle used
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> M(nn);
    for (auto& v : M) v = u(g);
    S.assign(nn, 0.0);
    for (int r = 0; r < npad; ++r)
        for (int c = 0; c <= r; ++c) {
            double s = 0.0;
            for (int q = 0; q < npad; ++q) s += M[static_cast<size_t>(q) * npad + r] * M[static_cast<size_t>(q) * npad + c];
            S[static_cast<size_t>(r) * npad + c] = s / npad + (r == c ? 0.5 : 0.0);
        }
    const std::vector<double> A0 = S;
    D.assign(static_cast<size_t>(nblk) * tb * tb, 0.0);
    DinvT = D;
    X.assign(nn, 0.0);
    R.assign(nn, 0.0);
    for (auto& f : flags) {
        std::vector<std::atomic<int>> v(static_cast<size_t>(nblk) * nblk);
        for (auto& x : v) x.store(0);
        f.swap(v);
    }
    col_start.resize(static_cast<size_t>(nblk) + 1);
    ptask_col_start(nblk, inverse, col_start.data());
    std::vector<std::thread> th;
    th.emplace_back(critical, seed * 7919u + 1u);
    for (int w = 0; w < workers; ++w) th.emplace_back(worker, seed * 104729u + 17u * w + 3u);
    for (auto& t : th) t.join();
    if (failed.load() == 1) return 3;
    if (failed.load() == 2) return 4;
    FILE* fp = std::fopen(argv[6], "wb");
    if (!fp) return 1;
    std::vector<double> L(nn, 0.0);
    for (int r = 0; r < npad; ++r)
        for (int c = 0; c <= r; ++c) L[static_cast<size_t>(r) * npad + c] = S[static_cast<size_t>(r) * npad + c];
    std::fwrite(A0.data(), sizeof(double), nn, fp);
    std::fwrite(L.data(), sizeof(double), nn, fp);
    std::fwrite(X.data(), sizeof(double), nn, fp);
    std::fclose(fp);
    std::printf("tasks %lld\n", col_start[nblk]);
    return 0;
}
