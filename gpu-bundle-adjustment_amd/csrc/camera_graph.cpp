// camera_graph.cpp -- the envelope (profile) layout of the reduced camera
// system for problems whose cameras see limited parts of the scene.
//
// Block (a, b) of S = U - W V^-1 W^T is non-zero only when cameras a and b
// observe a common point.  When that camera graph is sparse, the cameras are
// renumbered by reverse Cuthill-McKee (breadth-first from a pseudo-peripheral
// camera, neighbours by increasing degree, order reversed) so the non-zero
// blocks gather near the diagonal, and the dense tile factor then works only
// inside the profile: tile row i of S is zero left of tile column first[i],
// and a Cholesky factor keeps the row profile of its matrix (no fill left of
// the first non-zero of a row), so the tiles (i, j), first[i] <= j <= i, hold
// all of L.  Storage stays n_pad x n_pad (HBM holds a dense S up to C of
// several thousand); the work falls from n^3/3 to about n b^2 for a profile
// of width b.  The reference factors S with a dense Eigen LLT and names a
// sparse factor as its own next step (src/lm.cu:13-17).
#include <algorithm>
#include <cstdint>
#include <queue>
#include <vector>

#include "solver.hpp"

namespace gba_amd {

namespace {

struct Bitmap {
    int n;
    std::vector<uint64_t> w;
    explicit Bitmap(int n_) : n(n_), w((static_cast<size_t>(n_) * n_ + 63) / 64, 0) {}
    bool set(int a, int b) {  // true when newly set
        const size_t q = static_cast<size_t>(a) * n + b;
        const uint64_t m = uint64_t(1) << (q & 63);
        if (w[q >> 6] & m) return false;
        w[q >> 6] |= m;
        return true;
    }
    bool get(int a, int b) const {
        const size_t q = static_cast<size_t>(a) * n + b;
        return (w[q >> 6] >> (q & 63)) & 1;
    }
};

// Breadth-first levels from s over the unvisited cameras of its component;
// returns the last level's cameras.
std::vector<int> bfs_last_level(const std::vector<std::vector<int>>& adj, int s, int& depth) {
    std::vector<int> level(adj.size(), -1), cur{s}, last;
    level[s] = 0;
    depth = 0;
    while (!cur.empty()) {
        last = cur;
        std::vector<int> nxt;
        for (int a : cur)
            for (int b : adj[a])
                if (level[b] < 0) {
                    level[b] = level[a] + 1;
                    nxt.push_back(b);
                }
        if (!nxt.empty()) ++depth;
        cur.swap(nxt);
    }
    return last;
}

}  // namespace

bool plan_envelope(const gba_problem_view& prob, double max_density, EnvelopeLayout& out) {
    const int C = prob.num_cameras, P = prob.num_points, O = prob.num_observations;
    if (C < 8 || C > 40000) return false;  // tiny: nothing to gain; huge: the bitmap would not pay
    // cameras of each point (counting sort by point)
    std::vector<int> off(static_cast<size_t>(P) + 1, 0), cams(static_cast<size_t>(O));
    for (int e = 0; e < O; ++e) off[prob.obs_point[e] + 1]++;
    for (int p = 0; p < P; ++p) off[p + 1] += off[p];
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int e = 0; e < O; ++e) cams[fill[prob.obs_point[e]]++] = prob.obs_camera[e];
    }
    // the camera graph, given up as soon as it is too dense to pay
    Bitmap g(C);
    const double budget = max_density * C * static_cast<double>(C);
    double edges = 0.0;
    for (int p = 0; p < P; ++p)
        for (int x = off[p]; x < off[p + 1]; ++x)
            for (int y = off[p]; y < off[p + 1]; ++y) {
                const int a = cams[x], b = cams[y];
                if (a != b && g.set(a, b) && (edges += 1.0) > budget) return false;
            }
    std::vector<std::vector<int>> adj(static_cast<size_t>(C));
    for (int a = 0; a < C; ++a)
        for (int b = 0; b < C; ++b)
            if (g.get(a, b)) adj[a].push_back(b);
    std::vector<int> deg(static_cast<size_t>(C));
    for (int a = 0; a < C; ++a) deg[a] = static_cast<int>(adj[a].size());
    for (auto& v : adj) std::sort(v.begin(), v.end(), [&](int x, int y) { return deg[x] != deg[y] ? deg[x] < deg[y] : x < y; });
    // Cuthill-McKee over every component, each from a pseudo-peripheral camera
    std::vector<int> order;
    order.reserve(static_cast<size_t>(C));
    std::vector<char> seen(static_cast<size_t>(C), 0);
    for (;;) {
        int s = -1;
        for (int a = 0; a < C; ++a)
            if (!seen[a] && (s < 0 || deg[a] < deg[s])) s = a;
        if (s < 0) break;
        // pseudo-peripheral start: move to a minimum-degree camera of the
        // farthest level while the eccentricity grows
        int depth = 0;
        for (int it = 0; it < 8; ++it) {
            std::vector<int> last = bfs_last_level(adj, s, depth);
            int t = *std::min_element(last.begin(), last.end(), [&](int x, int y) { return deg[x] < deg[y]; });
            int d2 = 0;
            bfs_last_level(adj, t, d2);
            if (d2 <= depth) break;
            s = t;
        }
        std::queue<int> q;
        q.push(s);
        seen[s] = 1;
        while (!q.empty()) {
            const int a = q.front();
            q.pop();
            order.push_back(a);
            for (int b : adj[a])
                if (!seen[b]) {
                    seen[b] = 1;
                    q.push(b);
                }
        }
    }
    std::reverse(order.begin(), order.end());
    out.cam_of_pos = order;
    out.pos_of_cam.assign(static_cast<size_t>(C), 0);
    for (int i = 0; i < C; ++i) out.pos_of_cam[order[i]] = i;
    // tile profile of the renumbered S
    const int n = 9 * C, npad = (n + kTile - 1) / kTile * kTile, nblk = npad / kTile;
    std::vector<int> lo(static_cast<size_t>(C));  // lowest neighbour position, self included
    for (int i = 0; i < C; ++i) {
        int m = i;
        for (int b : adj[order[i]]) m = std::min(m, out.pos_of_cam[b]);
        lo[i] = m;
    }
    out.first.assign(static_cast<size_t>(nblk), 0);
    for (int t = 0; t < nblk; ++t) {
        int f = t;  // a padding-only tile row is the identity
        for (int r = t * kTile; r < std::min(n, (t + 1) * kTile); ++r) f = std::min(f, (9 * lo[r / 9]) / kTile);
        out.first[t] = f;
    }
    // work in tile products: envelope vs dense
    double env = 0.0, dense = 0.0;
    for (int k = 0; k + 1 < nblk; ++k) {
        const int m = nblk - 1 - k;
        dense += m + 0.5 * m * (m + 1);
        for (int i = k + 1; i < nblk; ++i) {
            if (out.first[i] > k) continue;
            env += 1.0;  // TRSM
            for (int j = k + 1; j <= i; ++j) env += out.first[j] <= k ? 1.0 : 0.0;
        }
    }
    out.work_ratio = dense > 0.0 ? env / dense : 1.0;
    out.density = edges / (static_cast<double>(C) * C);
    return true;
}

void envelope_factor_tables(const std::vector<int>& first, std::vector<int2>& trsm, std::vector<int>& trsm_off,
                            std::vector<int2>& upd, std::vector<int>& upd_off) {
    const int nblk = static_cast<int>(first.size());
    trsm.clear();
    upd.clear();
    trsm_off.assign(1, 0);
    upd_off.assign(1, 0);
    for (int k = 0; k + 1 < nblk; ++k) {
        const int m = nblk - 1 - k;
        for (int i = k + 1; i < nblk; ++i)
            if (first[i] <= k) trsm.push_back(make_int2(0, i - k - 1));
        // tile 0 (k+1, k+1) always: the update workgroup that owns it factors
        // it for the next column (its product is zero when row k+1 starts
        // after column k: the transposed tile it reads is never written and
        // stays zero)
        upd.push_back(make_int2(0, 0));
        for (int ii = 1; ii < m; ++ii) {
            const int i = k + 1 + ii;
            if (first[i] > k) continue;
            for (int jj = 0; jj <= ii; ++jj)
                if (first[k + 1 + jj] <= k) upd.push_back(make_int2(0, ii * (ii + 1) / 2 + jj));
        }
        trsm_off.push_back(static_cast<int>(trsm.size()));
        upd_off.push_back(static_cast<int>(upd.size()));
    }
}

}  // namespace gba_amd
