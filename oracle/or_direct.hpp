// or_direct.hpp — test infrastructure (CPU oracle): a direct sparse LL^T of a 3x3-block SPD
// system with a fill-reducing order, the counterpart of the reference's CPU solve
// (Eigen::SimplicialLLT, src/Cloth.cu:686-688) at sizes where the banded solve of
// or_physics.hpp (RCM order) is far too slow (a 513^2 cloth: band ~3,000 unknowns).
//
//   order:   nested dissection of the block graph by level structures: within a connected
//            set, a breadth-first level structure from a pseudo-peripheral node, its middle
//            level as the separator, the two sides dissected the same way; sets of at most
//            `leaf` nodes are not split; elimination: both sides, then the separator;
//   factor:  multifrontal in FP64 over the dissection tree: a front holds its separator's
//            unknowns and the later unknowns they couple to (the children's update sets and
//            the separator's own neighbours); it receives the matrix entries of its separator
//            columns and its children's update matrices (extend-add), factors its separator
//            columns (dense right-looking Cholesky in panels of 16 columns) and passes its
//            Schur complement up;
//   solve:   forward and backward substitution front by front.
// Exact up to rounding (pinned against the banded solve and a dense numpy solve in
// tests/test_oracle_direct.py).  Used only by tests and by bench.py's CPU reference column.
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <queue>
#include <vector>

namespace orc {

struct NdFront {
    std::vector<int> sep;       // block nodes eliminated here, in elimination order
    std::vector<int> upd;       // later block nodes coupled to them (ascending elimination position)
    std::vector<int> children;  // fronts whose update matrices are added here
    std::vector<double> L;      // factor columns: 3|sep| columns of 3(|sep|+|upd|) rows, column-major
};

class NdCholesky {
  public:
    // rowPtr/col: block pattern with both triangles; returns false when a pivot is not positive
    template <class T>
    bool factor(const std::vector<int>& rowPtr, const std::vector<int>& col, const std::vector<T>& val,
                int leaf = 64) {
        n_ = (int)rowPtr.size() - 1;
        fronts_.clear();
        roots_.clear();
        order(rowPtr, col, leaf);
        symbolic(rowPtr, col);
        return numeric(rowPtr, col, val);
    }

    template <class T> void solve(const std::vector<T>& b, std::vector<T>& x) const {
        std::vector<double> y(3 * (size_t)n_);
        for (size_t i = 0; i < y.size(); i++) y[i] = (double)b[i];
        for (int f : post_) {  // forward: L y = b
            const NdFront& F = fronts_[f];
            const int k = 3 * (int)F.sep.size(), m = k + 3 * (int)F.upd.size();
            std::vector<int> g = dofs(F);
            for (int j = 0; j < k; j++) {
                const double* cj = &F.L[(size_t)j * m];
                const double yj = y[g[j]] / cj[j];
                y[g[j]] = yj;
                for (int i = j + 1; i < m; i++) y[g[i]] -= cj[i] * yj;
            }
        }
        for (auto it = post_.rbegin(); it != post_.rend(); ++it) {  // backward: L^T x = y
            const NdFront& F = fronts_[*it];
            const int k = 3 * (int)F.sep.size(), m = k + 3 * (int)F.upd.size();
            std::vector<int> g = dofs(F);
            for (int j = k - 1; j >= 0; j--) {
                const double* cj = &F.L[(size_t)j * m];
                double s = y[g[j]];
                for (int i = j + 1; i < m; i++) s -= cj[i] * y[g[i]];
                y[g[j]] = s / cj[j];
            }
        }
        x.assign(y.size(), T(0));
        for (size_t i = 0; i < y.size(); i++) x[i] = (T)y[i];
    }

    double factor_flops() const { return flops_; }
    size_t factor_entries() const {
        size_t e = 0;
        for (const NdFront& F : fronts_) e += F.L.size();
        return e;
    }

  private:
    int n_ = 0;
    std::vector<NdFront> fronts_;
    std::vector<int> roots_, post_, pos_;
    double flops_ = 0.0;

    std::vector<int> dofs(const NdFront& F) const {
        std::vector<int> g;
        g.reserve(3 * (F.sep.size() + F.upd.size()));
        for (int v : F.sep) for (int a = 0; a < 3; a++) g.push_back(3 * v + a);
        for (int v : F.upd) for (int a = 0; a < 3; a++) g.push_back(3 * v + a);
        return g;
    }

    // ---- nested dissection
    void order(const std::vector<int>& rp, const std::vector<int>& ci, int leaf) {
        std::vector<int> mark(n_, 0), level(n_, -1);
        int stamp = 0;
        std::vector<int> all(n_);
        for (int i = 0; i < n_; i++) all[i] = i;
        // mark[v] == stamp: v belongs to the set being processed
        auto bfs = [&](const std::vector<int>& set, int src, int st, std::vector<int>& out, int& depth) {
            for (int v : set) level[v] = -1;
            out.clear();
            level[src] = 0;
            out.push_back(src);
            for (size_t h = 0; h < out.size(); h++) {
                const int v = out[h];
                for (int k = rp[v]; k < rp[v + 1]; k++) {
                    const int u = ci[k];
                    if (mark[u] == st && level[u] < 0) {
                        level[u] = level[v] + 1;
                        out.push_back(u);
                    }
                }
            }
            depth = level[out.back()];
        };
        struct Task { std::vector<int> set; int parent; };
        std::vector<Task> stack;
        stack.push_back({all, -1});
        std::vector<int> comp, lv;
        while (!stack.empty()) {
            Task t = std::move(stack.back());
            stack.pop_back();
            // split into connected components (visited: level >= 0)
            ++stamp;
            for (int v : t.set) mark[v] = stamp;
            const int st = stamp;
            std::vector<std::vector<int>> comps;
            for (int v : t.set) level[v] = -1;
            for (int v : t.set) {
                if (level[v] >= 0) continue;
                bfs_component(rp, ci, mark, st, level, v, comp);
                comps.push_back(comp);
            }
            for (std::vector<int>& cset : comps) {
                const int id = (int)fronts_.size();
                fronts_.emplace_back();
                if (t.parent >= 0) fronts_[t.parent].children.push_back(id);
                else roots_.push_back(id);
                if ((int)cset.size() <= leaf) {
                    fronts_[id].sep = cset;
                    continue;
                }
                ++stamp;
                for (int v : cset) mark[v] = stamp;
                const int cs = stamp;
                // pseudo-peripheral node: repeat BFS from the farthest node while the depth grows
                int src = cset[0], depth = -1;
                for (int rep = 0; rep < 4; rep++) {
                    int d;
                    bfs(cset, src, cs, lv, d);
                    if (d <= depth) break;
                    depth = d;
                    src = lv.back();
                }
                bfs(cset, src, cs, lv, depth);
                const int mid = depth / 2;
                std::vector<int> a, b, s;
                for (int v : lv) (level[v] < mid ? a : level[v] > mid ? b : s).push_back(v);
                if (a.empty() || b.empty()) {  // a set this thin is not split further
                    fronts_[id].sep = cset;
                    continue;
                }
                fronts_[id].sep = s;
                stack.push_back({std::move(a), id});
                stack.push_back({std::move(b), id});
            }
        }
        // post-order of the dissection tree; elimination positions
        post_.clear();
        std::vector<std::pair<int, int>> st2;
        for (int r : roots_) st2.push_back({r, 0});
        while (!st2.empty()) {
            auto& top = st2.back();
            const int f = top.first;
            if (top.second < (int)fronts_[f].children.size()) {
                const int c = fronts_[f].children[top.second++];
                st2.push_back({c, 0});
            } else {
                post_.push_back(f);
                st2.pop_back();
            }
        }
        pos_.assign(n_, -1);
        int p = 0;
        for (int f : post_)
            for (int v : fronts_[f].sep) pos_[v] = p++;
    }

    static void bfs_component(const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<int>& mark,
                              int st, std::vector<int>& level, int src, std::vector<int>& out) {
        out.clear();
        level[src] = 0;
        out.push_back(src);
        for (size_t h = 0; h < out.size(); h++) {
            const int v = out[h];
            for (int k = rp[v]; k < rp[v + 1]; k++) {
                const int u = ci[k];
                if (mark[u] == st && level[u] < 0) {
                    level[u] = 0;
                    out.push_back(u);
                }
            }
        }
    }

    // ---- update sets
    void symbolic(const std::vector<int>& rp, const std::vector<int>& ci) {
        std::vector<int> seen(n_, -1);
        for (int f : post_) {
            NdFront& F = fronts_[f];
            std::sort(F.sep.begin(), F.sep.end(), [&](int a, int b) { return pos_[a] < pos_[b]; });
            const int last = pos_[F.sep.back()];
            std::vector<int> u;
            auto add = [&](int v) {
                if (pos_[v] > last && seen[v] != f) {
                    seen[v] = f;
                    u.push_back(v);
                }
            };
            for (int v : F.sep)
                for (int k = rp[v]; k < rp[v + 1]; k++) add(ci[k]);
            for (int c : F.children)
                for (int v : fronts_[c].upd) add(v);
            std::sort(u.begin(), u.end(), [&](int a, int b) { return pos_[a] < pos_[b]; });
            F.upd = std::move(u);
        }
    }

    // ---- numeric factorisation
    template <class T>
    bool numeric(const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<T>& val) {
        flops_ = 0.0;
        std::vector<std::vector<double>> U(fronts_.size());  // update matrices, column-major
        std::vector<int> local(n_, -1);
        // a factorisation that runs longer than 15 s reports its progress on stderr
        const auto t0 = std::chrono::steady_clock::now();
        double lastReport = 0.0;
        int doneFronts = 0;
        for (int f : post_) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el - lastReport > 15.0) {
                std::fprintf(stderr, "nested-dissection LL^T: %d of %zu fronts factored, %.0f s\n", doneFronts,
                             post_.size(), el);
                std::fflush(stderr);
                lastReport = el;
            }
            doneFronts++;
            NdFront& F = fronts_[f];
            const int ns = (int)F.sep.size(), nu = (int)F.upd.size();
            const int k = 3 * ns, m = 3 * (ns + nu);
            for (int a = 0; a < ns; a++) local[F.sep[a]] = a;
            for (int a = 0; a < nu; a++) local[F.upd[a]] = ns + a;
            std::vector<double> W((size_t)m * m, 0.0);  // column-major, lower triangle used
            // original entries of the separator columns: A(u, v) for v in sep, pos(u) >= pos(v)
            for (int v : F.sep) {
                const int lv = local[v];
                for (int q = rp[v]; q < rp[v + 1]; q++) {
                    const int u = ci[q];
                    if (pos_[u] < pos_[v]) continue;
                    const int lu = local[u];
                    for (int a = 0; a < 3; a++)      // row 3 lu + a
                        for (int c = 0; c < 3; c++)  // column 3 lv + c: A_uv[a][c] = A_vu[c][a]
                            W[(size_t)(3 * lv + c) * m + 3 * lu + a] = (double)val[9 * (size_t)q + 3 * c + a];
                }
            }
            // extend-add of the children's update matrices
            for (int c : F.children) {
                const NdFront& C = fronts_[c];
                const int mc = 3 * (int)C.upd.size();
                std::vector<int> map(mc);
                for (int a = 0; a < (int)C.upd.size(); a++)
                    for (int d = 0; d < 3; d++) map[3 * a + d] = 3 * local[C.upd[a]] + d;
                const std::vector<double>& Uc = U[c];
                for (int jj = 0; jj < mc; jj++)
                    for (int ii = jj; ii < mc; ii++) {
                        const int r = map[ii], s = map[jj];
                        if (r >= s) W[(size_t)s * m + r] += Uc[(size_t)jj * mc + ii];
                        else W[(size_t)r * m + s] += Uc[(size_t)jj * mc + ii];
                    }
                U[c].clear();
                U[c].shrink_to_fit();
            }
            // dense right-looking Cholesky of the first k columns, panels of PW columns
            constexpr int PW = 16;
            for (int j0 = 0; j0 < k; j0 += PW) {
                const int j1 = std::min(k, j0 + PW);
                for (int j = j0; j < j1; j++) {
                    double* cj = &W[(size_t)j * m];
                    if (!(cj[j] > 0.0)) return false;
                    const double d = std::sqrt(cj[j]);
                    cj[j] = d;
                    const double inv = 1.0 / d;
                    for (int i = j + 1; i < m; i++) cj[i] *= inv;
                    for (int l = j + 1; l < j1; l++) {  // within the panel
                        double* cl = &W[(size_t)l * m];
                        const double f = cj[l];
                        for (int i = l; i < m; i++) cl[i] -= cj[i] * f;
                    }
                }
                // trailing columns: W(:, l) -= sum over the panel of L(:, p) L(l, p), each
                // column read and written once per panel
                const int np = j1 - j0;
                for (int l = j1; l < m; l++) {
                    double* cl = &W[(size_t)l * m];
                    double f[PW];
                    for (int p = 0; p < np; p++) f[p] = W[(size_t)(j0 + p) * m + l];
                    const double* c0 = &W[(size_t)j0 * m];
                    if (np == PW) {
                        for (int i = l; i < m; i++) {
                            double acc = 0.0;
                            for (int p = 0; p < PW; p++) acc += c0[(size_t)p * m + i] * f[p];
                            cl[i] -= acc;
                        }
                    } else {
                        for (int i = l; i < m; i++) {
                            double acc = 0.0;
                            for (int p = 0; p < np; p++) acc += c0[(size_t)p * m + i] * f[p];
                            cl[i] -= acc;
                        }
                    }
                }
            }
            flops_ += (double)k * (double)m * (double)m - (double)k * k * m + (double)k * k * k / 3.0;
            // keep the factor columns; the Schur complement becomes the update matrix
            F.L.assign(W.begin(), W.begin() + (size_t)k * m);
            const int mu = m - k;
            std::vector<double> Uf((size_t)mu * mu);
            for (int jj = 0; jj < mu; jj++)
                for (int ii = jj; ii < mu; ii++) Uf[(size_t)jj * mu + ii] = W[(size_t)(k + jj) * m + k + ii];
            U[f] = std::move(Uf);
            for (int v : F.sep) local[v] = -1;
            for (int v : F.upd) local[v] = -1;
        }
        return true;
    }
};

}  // namespace orc
