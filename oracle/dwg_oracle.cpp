// DWG oracle: a CPU restatement of the reference's diffusing-winding-gradients
// path, used ONLY as the checker by tests/, by __graft_entry__.smoke() and as
// the CPU-baseline leg of bench.py.  Nothing in the shipped library links or
// calls this file.
//
// Precision: FP64 throughout (the reference's CPU path, libigl's
// fast_winding_number, runs on Eigen::MatrixXd, DWGProcess.h:678, 722).
// Where a test needs the GPU's exact float decisions (Barnes-Hut near/far
// tests, kNN distances) a float mode reproduces them bit for bit.
//
// Function → reference map (all paths relative to the reference root):
//   orc_tree_build        deps/libigl/include/igl/octree.cpp:171-335 (octreeDWG)
//   orc_tree_precompute   deps/libigl/include/igl/fast_winding_number.cpp:9-120 (order 0)
//   orc_tree_eval         fast_winding_number.cpp:134-312 (BFS near list, β>0)
//                         src/TurboLibiGWN.cu:513-812 (screened kernel semantics)
//   orc_direct            fast_winding_number.cpp:313-323 (β≤0, double accumulator)
//                         src/TurboLibiGWN.cu:544-585 (screened term)
//   orc_direct_grad       NEW (no reference counterpart): exact ∇_q of the plain and screened terms
//   orc_tree_eval_grad    NEW: the Barnes-Hut term set of orc_tree_eval with each term's gradient
//   orc_voxel_grid        deps/libigl/include/igl/voxel_grid.cpp:15-87, grid.cpp:14-45
//   orc_mc_*              src/TurboMC.cu:373-478 (corners, lerp, table, negated normal)
//   orc_knn_brute         cudaKDTree/cukd/knn.h:277-292, 403-414; math.h:210-211
//   orc_radii             include/DWGProcess.h:325-343
//   orc_dwg_iterate       src/TurboDWG.cu:516-652 (σ/iso schedule, scatter, normalize, zero fill)
//   orc_gauss_map         src/TurboNormalInit.cu:3-27
//
// Pinning (tests/test_oracle.py): single-dipole closed form, sphere inside/
// outside limits, β→∞ tree == direct, analytic vs finite-difference gradient,
// voxel-grid closed form, MC planar/linear fields and table invariants,
// kNN vs an independent numpy brute force, radii on a regular lattice.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "../dwg_amd/csrc/mc_table.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// igl::parallel_for (parallel_for.h:117-160): contiguous slices over all
// hardware threads, serial below min_parallel iterations.
int g_threads = 0;

int thread_count() {
    if (g_threads > 0) return g_threads;
    unsigned h = std::thread::hardware_concurrency();
    return h == 0 ? 1 : int(h);
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body,
                  int64_t min_parallel = 1000) {
    int nt = thread_count();
    if (n < min_parallel || nt <= 1) {
        body(0, n);
        return;
    }
    int64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &body]() { body(b, e); });
    }
    for (auto& x : th) x.join();
}

struct Tree {
    // octreeDWG output
    std::vector<std::array<int, 8>> ch;       // CH (−1 = leaf)
    std::vector<std::vector<int>> pi;         // point indices per node
    std::vector<std::array<double, 3>> cn;    // cell centres
    std::vector<double> w;                    // cell widths
    std::vector<int> depth;
    // order-0 precompute
    std::vector<std::array<double, 3>> cm, ec;
    std::vector<double> r;
    // float copies (what a GPU would hold)
    std::vector<std::array<float, 3>> cm_f, ec_f;
    std::vector<float> r_f;
    const double* P = nullptr;  // borrowed points (n×3)
    int64_t n = 0;
};

inline float fdist2(float dx, float dy, float dz) {
    // the one float distance formula shared with the GPU kernels:
    // fma(dz, dz, fma(dy, dy, dx*dx))
    return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads() { return thread_count(); }

// ---------------------------------------------------------------------------
// Octree: octreeDWG restated.  Root cube = AABB centre, width = max extent.
// A node splits when it holds more than `leaf_size` points (1 in the
// reference) and depth < max_depth; all 8 children are created.  Octant bit
// 1: x ≥ cx, 2: y ≥ cy, 4: z ≥ cz.  Children are numbered in creation order
// (pre-order over split nodes), children of one node contiguous.
// ---------------------------------------------------------------------------
void* orc_tree_build(const double* P, int64_t n, int max_depth, int leaf_size) {
    auto* t = new Tree();
    t->P = P;
    t->n = n;
    std::vector<int> all(n);
    for (int64_t i = 0; i < n; ++i) all[i] = int(i);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], P[3 * i + d]);
            hi[d] = std::max(hi[d], P[3 * i + d]);
        }
    t->pi.push_back(all);
    t->ch.push_back({-1, -1, -1, -1, -1, -1, -1, -1});
    t->cn.push_back({(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2});
    t->w.push_back(std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]}));
    t->depth.push_back(0);
    std::function<void(int, int)> split = [&](int idx, int depth) {
        if (!(int64_t(t->pi[idx].size()) > leaf_size && depth < max_depth)) return;
        int m = int(t->ch.size());
        double h = t->w[idx] / 2;
        std::array<double, 3> c = t->cn[idx];
        for (int k = 0; k < 8; ++k) {
            t->ch[idx][k] = m + k;
            double q = h / 2;
            std::array<double, 3> cc = {c[0] + ((k & 1) ? q : -q), c[1] + ((k & 2) ? q : -q),
                                        c[2] + ((k & 4) ? q : -q)};
            t->ch.push_back({-1, -1, -1, -1, -1, -1, -1, -1});
            t->pi.emplace_back();
            t->cn.push_back(cc);
            t->w.push_back(h);
            t->depth.push_back(depth + 1);
        }
        for (int p : t->pi[idx]) {
            int o = (P[3 * p] >= c[0] ? 1 : 0) + (P[3 * p + 1] >= c[1] ? 2 : 0) +
                    (P[3 * p + 2] >= c[2] ? 4 : 0);
            t->pi[m + o].push_back(p);
        }
        for (int k = 0; k < 8; ++k) split(m + k, depth + 1);
    };
    split(0, 0);
    return t;
}

void orc_tree_free(void* h) { delete static_cast<Tree*>(h); }

int64_t orc_tree_num_nodes(void* h) { return int64_t(static_cast<Tree*>(h)->ch.size()); }

// Copies node topology: children (nodes×8), point counts, depths.
void orc_tree_topology(void* h, int* children, int* counts, int* depths) {
    auto* t = static_cast<Tree*>(h);
    for (size_t i = 0; i < t->ch.size(); ++i) {
        for (int k = 0; k < 8; ++k) children[8 * i + k] = t->ch[i][k];
        counts[i] = int(t->pi[i].size());
        depths[i] = t->depth[i];
    }
}

// Points of node `node` in stored order.
int64_t orc_tree_node_points(void* h, int64_t node, int* out) {
    auto* t = static_cast<Tree*>(h);
    auto& v = t->pi[node];
    if (out) std::copy(v.begin(), v.end(), out);
    return int64_t(v.size());
}

// Order-0 precompute (fast_winding_number.cpp:54-119): CM = ΣA·p/ΣA,
// EC = ΣA·n, R = max |p − CM|, all FP64.  Float copies are what a GPU holds.
void orc_tree_precompute(void* h, const double* N, const double* A) {
    auto* t = static_cast<Tree*>(h);
    size_t m = t->ch.size();
    t->cm.assign(m, {0, 0, 0});
    t->ec.assign(m, {0, 0, 0});
    t->r.assign(m, 0);
    t->cm_f.assign(m, {0, 0, 0});
    t->ec_f.assign(m, {0, 0, 0});
    t->r_f.assign(m, 0);
    const double* P = t->P;
    parallel_for(int64_t(m), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double at = 0, c[3] = {0, 0, 0}, ecv[3] = {0, 0, 0};
            for (int p : t->pi[i]) {
                at += A[p];
                for (int d = 0; d < 3; ++d) {
                    c[d] += A[p] * P[3 * p + d];
                    ecv[d] += A[p] * N[3 * p + d];
                }
            }
            for (int d = 0; d < 3; ++d) c[d] /= at;
            double rmax = 0;
            for (int p : t->pi[i]) {
                double dx = P[3 * p] - c[0], dy = P[3 * p + 1] - c[1], dz = P[3 * p + 2] - c[2];
                rmax = std::max(rmax, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
            for (int d = 0; d < 3; ++d) {
                t->cm[i][d] = c[d];
                t->ec[i][d] = ecv[d];
                t->cm_f[i][d] = float(c[d]);
                t->ec_f[i][d] = float(ecv[d]);
            }
            t->r[i] = rmax;
            t->r_f[i] = float(rmax);
        }
    });
}

// Refresh of EC only (TurboLibiGWN.cu:924-952): EC = Σ A n over the subtree.
void orc_tree_refresh_ec(void* h, const double* N, const double* A) {
    auto* t = static_cast<Tree*>(h);
    size_t m = t->ch.size();
    parallel_for(int64_t(m), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double ecv[3] = {0, 0, 0};
            for (int p : t->pi[i])
                for (int d = 0; d < 3; ++d) ecv[d] += A[p] * N[3 * p + d];
            for (int d = 0; d < 3; ++d) {
                t->ec[i][d] = ecv[d];
                t->ec_f[i][d] = float(ecv[d]);
            }
        }
    });
}

void orc_tree_moments(void* h, double* cm, double* r, double* ec) {
    auto* t = static_cast<Tree*>(h);
    for (size_t i = 0; i < t->ch.size(); ++i) {
        for (int d = 0; d < 3; ++d) {
            if (cm) cm[3 * i + d] = t->cm[i][d];
            if (ec) ec[3 * i + d] = t->ec[i][d];
        }
        if (r) r[i] = t->r[i];
    }
}

}  // extern "C"

namespace {

// Plain term (fast_winding_number.cpp:182-184): (d·An)/(4π|d|³ + 1e-8), d = p − q.
inline double plain_term(double dx, double dy, double dz, double ax, double ay, double az) {
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    return (dx * ax + dy * ay + dz * az) / (4.0 * kPi * r * r * r + 1e-8);
}

// Screened term (TurboLibiGWN.cu:544-585): with ρ = max(s, r),
// (σρ + 1)(d·An) / (e^{σρ}·4πρ³ + 1e-7); skipped when r ≤ 1e-10.
inline double screened_term(double dx, double dy, double dz, double ax, double ay, double az,
                            double s, double sigma) {
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (!(r > 1e-10)) return 0.0;
    double rho = s > r ? s : r;
    double dot = dx * ax + dy * ay + dz * az;
    return (rho * sigma + 1.0) * dot / (std::exp(rho * sigma) * (4.0 * kPi * rho * rho * rho) + 1e-7);
}

}  // namespace

extern "C" {

// Barnes-Hut evaluation, order 0.
//   kind 0: plain libigl kernel (fast_winding_number.cpp:246-312)
//   kind 1: screened DWG kernel (TurboLibiGWN.cu:513-812); far nodes use s = 0.0015,
//           points use radii[p] when use_radii else 0.0015.
// Term set: for each non-empty child of an opened node: if far, a leaf child
// is summed point by point and an internal child contributes one dipole;
// otherwise the child is opened.  The root is always opened.
// float_decisions = 1 takes the far test exactly as the GPU kernels do:
// fma-chained float |CM−q|² > (β·R)² with float CM, R, q.
// terms_out (nullable) receives point terms + dipole terms per query.
void orc_tree_eval(void* h, const double* N, const double* A, const double* radii,
                   const double* Q, int64_t nq, double beta, int kind, double sigma,
                   int use_radii, int float_decisions, double* out, int64_t* terms_out) {
    auto* t = static_cast<Tree*>(h);
    const double* P = t->P;
    const float beta_f = float(beta);
    parallel_for(nq, [&](int64_t b, int64_t e) {
        std::vector<int> near, next;
        for (int64_t qi = b; qi < e; ++qi) {
            const double q[3] = {Q[3 * qi], Q[3 * qi + 1], Q[3 * qi + 2]};
            const float qf[3] = {float(q[0]), float(q[1]), float(q[2])};
            double wn = 0;
            int64_t terms = 0;
            auto point_sum = [&](int node) {
                for (int p : t->pi[node]) {
                    double dx = P[3 * p] - q[0], dy = P[3 * p + 1] - q[1], dz = P[3 * p + 2] - q[2];
                    double ax = A[p] * N[3 * p], ay = A[p] * N[3 * p + 1], az = A[p] * N[3 * p + 2];
                    if (kind == 0)
                        wn += plain_term(dx, dy, dz, ax, ay, az);
                    else
                        wn += screened_term(dx, dy, dz, ax, ay, az, use_radii ? radii[p] : 0.0015, sigma);
                    ++terms;
                }
            };
            near.assign(1, 0);
            while (!near.empty()) {
                next.clear();
                for (int node : near) {
                    if (t->ch[node][0] == -1) {
                        point_sum(node);
                        continue;
                    }
                    for (int k = 0; k < 8; ++k) {
                        int c = t->ch[node][k];
                        if (t->pi[c].empty()) continue;
                        bool far;
                        if (float_decisions) {
                            float dx = t->cm_f[c][0] - qf[0], dy = t->cm_f[c][1] - qf[1],
                                  dz = t->cm_f[c][2] - qf[2];
                            float br = beta_f * t->r_f[c];
                            far = fdist2(dx, dy, dz) > br * br;
                        } else {
                            double dx = t->cm[c][0] - q[0], dy = t->cm[c][1] - q[1], dz = t->cm[c][2] - q[2];
                            far = std::sqrt(dx * dx + dy * dy + dz * dz) > beta * t->r[c];
                        }
                        if (far) {
                            if (t->ch[c][0] == -1) {
                                point_sum(c);
                            } else {
                                double dx = t->cm[c][0] - q[0], dy = t->cm[c][1] - q[1],
                                       dz = t->cm[c][2] - q[2];
                                if (kind == 0)
                                    wn += plain_term(dx, dy, dz, t->ec[c][0], t->ec[c][1], t->ec[c][2]);
                                else
                                    wn += screened_term(dx, dy, dz, t->ec[c][0], t->ec[c][1], t->ec[c][2],
                                                        0.0015, sigma);
                                ++terms;
                            }
                        } else {
                            next.push_back(c);
                        }
                    }
                }
                std::swap(near, next);
            }
            out[qi] = wn;
            if (terms_out) terms_out[qi] = terms;
        }
    });
}

// Direct all-pairs sums in FP64 (fast_winding_number.cpp:313-323).
// kind 0 plain, 1 screened (same term as orc_tree_eval).
void orc_direct(const double* P, const double* N, const double* A, const double* radii, int64_t n,
                const double* Q, int64_t nq, int kind, double sigma, int use_radii, double* out) {
    parallel_for(nq, [&](int64_t b, int64_t e) {
        for (int64_t qi = b; qi < e; ++qi) {
            double wn = 0;
            for (int64_t j = 0; j < n; ++j) {
                double dx = P[3 * j] - Q[3 * qi], dy = P[3 * j + 1] - Q[3 * qi + 1],
                       dz = P[3 * j + 2] - Q[3 * qi + 2];
                double ax = A[j] * N[3 * j], ay = A[j] * N[3 * j + 1], az = A[j] * N[3 * j + 2];
                if (kind == 0)
                    wn += plain_term(dx, dy, dz, ax, ay, az);
                else
                    wn += screened_term(dx, dy, dz, ax, ay, az, use_radii ? radii[j] : 0.0015, sigma);
            }
            out[qi] = wn;
        }
    });
}

}  // extern "C"

namespace {

// Value and ∇_q of one term (new; the reference has no gradient), FP64,
// d = p − q, an = A·n:
//   plain     w = (d·an)/D, D = 4π r³ + 1e-8:  ∇w = −an/D + 12π r (d·an) d / D²
//             (∂(d·an)/∂q = −an, ∂r³/∂q = −3 r d)
//   screened  w = g(ρ)(d·an), g = (σρ + 1)/D, D = e^{σρ} 4π ρ³ + 1e-7,
//             ρ = max(s, r), skipped when r ≤ 1e-10:
//             ∇w = −g an − (d·an) g'(ρ) d / r for r ≥ s (∂ρ/∂q = −d/r), −g an for r < s;
//             g' = (σ D − (σρ + 1) e^{σρ} 4π (σρ³ + 3ρ²)) / D²
inline void term_grad(int kind, const double d[3], const double an[3], double s, double sigma, double& w,
                      double g[3]) {
    const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dot = d[0] * an[0] + d[1] * an[1] + d[2] * an[2];
    if (kind == 0) {
        const double D = 4.0 * kPi * r * r * r + 1e-8;
        w = dot / D;
        for (int k = 0; k < 3; ++k) g[k] = -an[k] / D + 12.0 * kPi * r * dot * d[k] / (D * D);
        return;
    }
    w = 0;
    g[0] = g[1] = g[2] = 0;
    if (!(r > 1e-10)) return;
    const double rho = s > r ? s : r;
    const double E = std::exp(sigma * rho);
    const double D = E * 4.0 * kPi * rho * rho * rho + 1e-7;
    const double num = sigma * rho + 1.0;
    const double gv = num / D;
    w = gv * dot;
    const double gp = (sigma * D - num * E * 4.0 * kPi * (sigma * rho * rho * rho + 3.0 * rho * rho)) / (D * D);
    const double c = (s > r) ? 0.0 : -dot * gp / r;
    for (int k = 0; k < 3; ++k) g[k] = -gv * an[k] + c * d[k];
}

}  // namespace

extern "C" {

// FP64 all-pairs value and gradient; kind 0 plain, 1 screened (radii[j] when
// use_radii, else 0.0015).
void orc_direct_grad(const double* P, const double* N, const double* A, const double* radii, int64_t n,
                     const double* Q, int64_t nq, int kind, double sigma, int use_radii, double* out_w,
                     double* out_g) {
    parallel_for(nq, [&](int64_t b, int64_t e) {
        for (int64_t qi = b; qi < e; ++qi) {
            double wn = 0, g[3] = {0, 0, 0};
            for (int64_t j = 0; j < n; ++j) {
                const double d[3] = {P[3 * j] - Q[3 * qi], P[3 * j + 1] - Q[3 * qi + 1], P[3 * j + 2] - Q[3 * qi + 2]};
                const double an[3] = {A[j] * N[3 * j], A[j] * N[3 * j + 1], A[j] * N[3 * j + 2]};
                double w, gj[3];
                term_grad(kind, d, an, use_radii ? radii[j] : 0.0015, sigma, w, gj);
                wn += w;
                for (int k = 0; k < 3; ++k) g[k] += gj[k];
            }
            out_w[qi] = wn;
            for (int k = 0; k < 3; ++k) out_g[3 * qi + k] = g[k];
        }
    });
}

// Barnes-Hut value and gradient: the term set of orc_tree_eval (float far
// decisions), each term with its gradient; far nodes are dipoles at CM with
// moment EC and s = 0.0015.
void orc_tree_eval_grad(void* h, const double* N, const double* A, const double* radii, const double* Q,
                        int64_t nq, double beta, int kind, double sigma, int use_radii, double* out_w,
                        double* out_g) {
    auto* t = static_cast<Tree*>(h);
    const double* P = t->P;
    const float beta_f = float(beta);
    parallel_for(nq, [&](int64_t b, int64_t e) {
        std::vector<int> near, next;
        for (int64_t qi = b; qi < e; ++qi) {
            const double q[3] = {Q[3 * qi], Q[3 * qi + 1], Q[3 * qi + 2]};
            const float qf[3] = {float(q[0]), float(q[1]), float(q[2])};
            double wn = 0, g[3] = {0, 0, 0};
            auto add = [&](const double d[3], const double an[3], double s) {
                double w, gj[3];
                term_grad(kind, d, an, s, sigma, w, gj);
                wn += w;
                for (int k = 0; k < 3; ++k) g[k] += gj[k];
            };
            auto point_sum = [&](int node) {
                for (int p : t->pi[node]) {
                    const double d[3] = {P[3 * p] - q[0], P[3 * p + 1] - q[1], P[3 * p + 2] - q[2]};
                    const double an[3] = {A[p] * N[3 * p], A[p] * N[3 * p + 1], A[p] * N[3 * p + 2]};
                    add(d, an, use_radii ? radii[p] : 0.0015);
                }
            };
            near.assign(1, 0);
            while (!near.empty()) {
                next.clear();
                for (int node : near) {
                    if (t->ch[node][0] == -1) {
                        point_sum(node);
                        continue;
                    }
                    for (int k = 0; k < 8; ++k) {
                        const int c = t->ch[node][k];
                        if (t->pi[c].empty()) continue;
                        const float dx = t->cm_f[c][0] - qf[0], dy = t->cm_f[c][1] - qf[1], dz = t->cm_f[c][2] - qf[2];
                        const float br = beta_f * t->r_f[c];
                        if (fdist2(dx, dy, dz) > br * br) {
                            if (t->ch[c][0] == -1) {
                                point_sum(c);
                            } else {
                                const double d[3] = {t->cm[c][0] - q[0], t->cm[c][1] - q[1], t->cm[c][2] - q[2]};
                                add(d, t->ec[c].data(), 0.0015);
                            }
                        } else {
                            next.push_back(c);
                        }
                    }
                }
                std::swap(near, next);
            }
            out_w[qi] = wn;
            for (int k = 0; k < 3; ++k) out_g[3 * qi + k] = g[k];
        }
    });
}

// igl::voxel_grid(V, offset, s, pad=1) restated in closed form
// (voxel_grid.cpp:15-87, grid.cpp:14-45), FP64.  Grid point (i,j,k) =
// centre + ((i,j,k) − (res−1)/2)·h, index i + rx·(j + ry·k).
void orc_voxel_grid(const double* V, int64_t n, double offset, int s, double* origin, double* h,
                    int* res) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], V[3 * i + d]);
            hi[d] = std::max(hi[d], V[3 * i + d]);
        }
    double ext[3], ctr[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] -= offset;
        hi[d] += offset;
        ext[d] = hi[d] - lo[d];
        ctr[d] = (lo[d] + hi[d]) / 2;
    }
    int si = 0;
    for (int d = 1; d < 3; ++d)
        if (ext[d] > ext[si]) si = d;
    const int pad = 1;
    double ss = s - 2 * pad;
    double side[3];
    for (int d = 0; d < 3; ++d) side[d] = (d == si) ? ss : std::ceil(ss * ext[d] / ext[si]);
    for (int d = 0; d < 3; ++d) side[d] += 2 * pad;
    double amax = -1;
    int ai = 0;
    for (int d = 0; d < 3; ++d) {
        double ps = pad / (side[d] - 1.0);
        double a = ext[d] / (1.0 - 2.0 * ps);
        if (a > amax) {
            amax = a;
            ai = d;
        }
    }
    double hh = amax / (side[ai] - 1.0);
    for (int d = 0; d < 3; ++d) {
        res[d] = int(side[d]);
        origin[d] = ctr[d] - (side[d] - 1.0) / 2.0 * hh;
    }
    *h = hh;
}

// Marching cubes (TurboMC.cu:383-478) over a field sampled at
// origin + (i,j,k)·h, index i + rx(j + ry k).  Cubes are visited in x-fastest
// order and triangles appended in table order, so the output order is
// deterministic.  Returns the triangle count; the arrays (nullable) receive
// vertices (T×9), centroids (T×3) and negated unit normals (T×3).
int64_t orc_mc(const double* field, const int* res, const double* origin, double h, double iso,
               int64_t capacity, double* tris, double* cents, double* nrms) {
    static const dwg_mc::PackedTable tab = dwg_mc::make_packed_table();
    const int rx = res[0], ry = res[1], rz = res[2];
    int64_t count = 0;
    for (int z = 0; z + 1 < rz; ++z)
        for (int y = 0; y + 1 < ry; ++y)
            for (int x = 0; x + 1 < rx; ++x) {
                double val[8], pos[8][3];
                int cube = 0;
                for (int i = 0; i < 8; ++i) {
                    int mx = x + dwg_mc::corner_dx(i), my = y + dwg_mc::corner_dy(i), mz = z + dwg_mc::corner_dz(i);
                    val[i] = field[int64_t(mx) + int64_t(rx) * (my + int64_t(ry) * mz)];
                    pos[i][0] = origin[0] + mx * h;
                    pos[i][1] = origin[1] + my * h;
                    pos[i][2] = origin[2] + mz * h;
                    if (val[i] > iso) cube |= 1 << i;
                }
                uint64_t row = tab.row[cube];
                int nt = int(row >> 60);
                for (int tt = 0; tt < nt; ++tt) {
                    double v[3][3];
                    for (int k = 0; k < 3; ++k) {
                        int e = int((row >> (4 * (3 * tt + k))) & 0xF);
                        int a = dwg_mc::edge_a(e), b = dwg_mc::edge_b(e);
                        // P = P1 − ell (P2 − P1), ell = (a1 − iso)/(a2 − a1 + 1e-10)  (TurboMC.cu:373-381)
                        double ell = (val[a] - iso) / (val[b] - val[a] + 1e-10);
                        for (int d = 0; d < 3; ++d) v[k][d] = pos[a][d] - ell * (pos[b][d] - pos[a][d]);
                    }
                    if (count < capacity) {
                        if (tris)
                            for (int k = 0; k < 3; ++k)
                                for (int d = 0; d < 3; ++d) tris[9 * count + 3 * k + d] = v[k][d];
                        if (cents)
                            for (int d = 0; d < 3; ++d) cents[3 * count + d] = (v[0][d] + v[1][d] + v[2][d]) / 3.0;
                        if (nrms) {
                            double u[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
                            double w[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
                            double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
                            double s = -1.0 / (std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) + 1e-8);
                            for (int d = 0; d < 3; ++d) nrms[3 * count + d] = c[d] * s;
                        }
                    }
                    ++count;
                }
            }
    return count;
}

// Packed case table (for table-invariant tests): row[c] as in mc_table.h.
void orc_mc_table(uint64_t* rows) {
    static const dwg_mc::PackedTable tab = dwg_mc::make_packed_table();
    for (int c = 0; c < 256; ++c) rows[c] = tab.row[c];
}

// Exact kNN by brute force in float, ordered by (d², index) with
// d² = fma(dz,dz,fma(dy,dy,dx·dx)), d = q − p (cukd math.h:210-211, knn.h:403-414).
// skip (nullable): points with skip[j] != 0 are not candidates.
void orc_knn_brute(const float* P, int64_t n, const float* Q, int64_t nq, int k, const uint8_t* skip,
                   int* out_idx, float* out_d2) {
    parallel_for(nq, [&](int64_t b, int64_t e) {
        std::vector<uint64_t> best;
        for (int64_t qi = b; qi < e; ++qi) {
            best.assign(k, ~uint64_t(0));
            for (int64_t j = 0; j < n; ++j) {
                if (skip && skip[j]) continue;
                float d2 = fdist2(Q[3 * qi] - P[3 * j], Q[3 * qi + 1] - P[3 * j + 1], Q[3 * qi + 2] - P[3 * j + 2]);
                uint32_t bits;
                std::memcpy(&bits, &d2, 4);
                uint64_t key = (uint64_t(bits) << 32) | uint32_t(j);
                if (key < best[k - 1]) {
                    int pos = k - 1;
                    while (pos > 0 && best[pos - 1] > key) {
                        best[pos] = best[pos - 1];
                        --pos;
                    }
                    best[pos] = key;
                }
            }
            for (int i = 0; i < k; ++i) {
                out_idx[qi * k + i] = best[i] == ~uint64_t(0) ? -1 : int(uint32_t(best[i]));
                if (out_d2) {
                    uint32_t bits = uint32_t(best[i] >> 32);
                    float f;
                    std::memcpy(&f, &bits, 4);
                    out_d2[qi * k + i] = f;
                }
            }
        }
    });
}

}  // extern "C"

namespace {

// Exact kNN over a uniform cell grid (same (d², index) order as the brute
// force; used inside the oracle's DWG loop where brute force is too slow).
struct CellGrid {
    float lo[3];
    float cell;
    int dim[3];
    std::vector<int> start, items;
    const float* P;
    int64_t n;

    void build(const float* pts, int64_t np, double target_per_cell) {
        P = pts;
        n = np;
        float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int d = 0; d < 3; ++d) lo[d] = INFINITY;
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], P[3 * i + d]);
                hi[d] = std::max(hi[d], P[3 * i + d]);
            }
        double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6f});
        // surface-like data: occupancy grows with cell², so pick cell from N
        double c = ext / std::max(1.0, std::sqrt(double(n) / target_per_cell));
        cell = float(c);
        for (int d = 0; d < 3; ++d) dim[d] = std::max(1, int((hi[d] - lo[d]) / cell) + 1);
        int64_t nc = int64_t(dim[0]) * dim[1] * dim[2];
        start.assign(nc + 1, 0);
        std::vector<int64_t> cid(n);
        for (int64_t i = 0; i < n; ++i) {
            cid[i] = cell_of(P + 3 * i);
            start[cid[i] + 1]++;
        }
        for (int64_t c2 = 0; c2 < nc; ++c2) start[c2 + 1] += start[c2];
        items.assign(n, 0);
        std::vector<int> cur(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) items[cur[cid[i]]++] = int(i);
    }
    int coord(float v, int d) const {
        int c = int(std::floor((v - lo[d]) / cell));
        return std::min(std::max(c, 0), dim[d] - 1);
    }
    int64_t cell_of(const float* p) const {
        return coord(p[0], 0) + int64_t(dim[0]) * (coord(p[1], 1) + int64_t(dim[1]) * coord(p[2], 2));
    }
    void knn(const float* q, int k, const uint8_t* skip, std::vector<uint64_t>& best) const {
        best.assign(k, ~uint64_t(0));
        int c0[3] = {coord(q[0], 0), coord(q[1], 1), coord(q[2], 2)};
        int maxr = std::max({dim[0], dim[1], dim[2]});
        for (int r = 0; r <= maxr; ++r) {
            for (int z = c0[2] - r; z <= c0[2] + r; ++z) {
                if (z < 0 || z >= dim[2]) continue;
                for (int y = c0[1] - r; y <= c0[1] + r; ++y) {
                    if (y < 0 || y >= dim[1]) continue;
                    for (int x = c0[0] - r; x <= c0[0] + r; ++x) {
                        if (x < 0 || x >= dim[0]) continue;
                        if (std::max({std::abs(x - c0[0]), std::abs(y - c0[1]), std::abs(z - c0[2])}) != r) continue;
                        int64_t c = x + int64_t(dim[0]) * (y + int64_t(dim[1]) * z);
                        for (int s = start[c]; s < start[c + 1]; ++s) {
                            int j = items[s];
                            if (skip && skip[j]) continue;
                            float d2 = fdist2(q[0] - P[3 * j], q[1] - P[3 * j + 1], q[2] - P[3 * j + 2]);
                            uint32_t bits;
                            std::memcpy(&bits, &d2, 4);
                            uint64_t key = (uint64_t(bits) << 32) | uint32_t(j);
                            if (key < best[k - 1]) {
                                int pos = k - 1;
                                while (pos > 0 && best[pos - 1] > key) {
                                    best[pos] = best[pos - 1];
                                    --pos;
                                }
                                best[pos] = key;
                            }
                        }
                    }
                }
            }
            // searched block covers cells c0±r; any unsearched point is at
            // least this far away (per axis distance to the block's faces)
            if (best[k - 1] != ~uint64_t(0)) {
                double dmin = INFINITY;
                for (int d = 0; d < 3; ++d) {
                    double blo = lo[d] + double(c0[d] - r) * cell, bhi = lo[d] + double(c0[d] + r + 1) * cell;
                    if (c0[d] - r > 0) dmin = std::min(dmin, double(q[d]) - blo);
                    if (c0[d] + r + 1 < dim[d]) dmin = std::min(dmin, bhi - double(q[d]));
                }
                uint32_t bits = uint32_t(best[k - 1] >> 32);
                float kd2;
                std::memcpy(&kd2, &bits, 4);
                if (dmin == INFINITY) break;
                if (dmin > 0 && double(kd2) < dmin * dmin * (1.0 - 1e-6)) break;
            }
        }
    }
};

}  // namespace

extern "C" {

void orc_knn_grid(const float* P, int64_t n, const float* Q, int64_t nq, int k, const uint8_t* skip,
                  int* out_idx) {
    CellGrid g;
    g.build(P, n, std::max(2.0, k / 2.0));
    parallel_for(nq, [&](int64_t b, int64_t e) {
        std::vector<uint64_t> best;
        for (int64_t qi = b; qi < e; ++qi) {
            g.knn(Q + 3 * qi, k, skip, best);
            for (int i = 0; i < k; ++i) out_idx[qi * k + i] = best[i] == ~uint64_t(0) ? -1 : int(uint32_t(best[i]));
        }
    });
}

// Radii (DWGProcess.h:325-343): r_i = clamp(sqrt(Σ_{j∈kNN(i), self incl.} |p_i − p_j|² / k),
// rmin, rmax), FP64 accumulation over float kNN indices.
void orc_radii(const float* P, int64_t n, int k, double rmin, double rmax, double* out) {
    std::vector<int> idx(size_t(n) * k);
    orc_knn_grid(P, n, P, n, k, nullptr, idx.data());
    for (int64_t i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < k; ++j) {
            int p = idx[i * k + j];
            for (int d = 0; d < 3; ++d) {
                double df = double(P[3 * i + d]) - double(P[3 * p + d]);
                s += df * df;
            }
        }
        s = std::sqrt(s / k);
        out[i] = std::max(std::min(s, rmax), rmin);
    }
}

// Gauss-map initialisation (TurboNormalInit.cu:3-27): n = (p − c)/(|p − c| + 1e-8),
// c = centre of the points' AABB.
void orc_gauss_map(const double* P, int64_t n, double* N) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], P[3 * i + d]);
            hi[d] = std::max(hi[d], P[3 * i + d]);
        }
    for (int64_t i = 0; i < n; ++i) {
        double v[3];
        for (int d = 0; d < 3; ++d) v[d] = P[3 * i + d] - (lo[d] + hi[d]) / 2;
        double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int d = 0; d < 3; ++d) N[3 * i + d] = v[d] / (l + 1e-8);
    }
}

// σ schedule (TurboDWG.cu:526-533)
double orc_sigma_schedule(int it, double sigma_in) {
    double s = sigma_in;
    if (it < 4) s = 20.0;
    if (it > 0 && it % 7 == 0) s = 10.0;
    return s;
}

// The DWG iteration loop (TurboDWG.cu:516-652) over a fixed source set:
//   σ_it schedule → EC refresh → screened BH GWN at sources and grid →
//   iso = mean source WN (×0.5 when it%5==0, it≠0) → marching cubes →
//   15-NN of every triangle centroid among the sources → acc_i = Σ face
//   normals → n_i = acc/(|acc|+1e-8) unless |acc| < 1e-10 (then kept and
//   counted zero) → if more than 10 zeros: n_i = normalise(Σ of the 10 NN
//   non-zero normals).
// The float-atomic scatter of the reference is replaced by an exact sum.
// Per-iteration statistics go to w_mean[it], ntri[it], nzero[it].
void orc_dwg_iterate(const double* P, double* N, const double* A, const double* radii, int64_t n,
                     const int* res, const double* origin, double h, int max_depth, int leaf_size,
                     double beta, double sigma_in, int iters, int use_radii, double* w_mean,
                     int64_t* ntri, int64_t* nzero) {
    Tree* t = static_cast<Tree*>(orc_tree_build(P, n, max_depth, leaf_size));
    orc_tree_precompute(t, N, A);
    int64_t G = int64_t(res[0]) * res[1] * res[2];
    std::vector<double> grid(3 * G);
    for (int64_t k = 0; k < res[2]; ++k)
        for (int64_t j = 0; j < res[1]; ++j)
            for (int64_t i = 0; i < res[0]; ++i) {
                int64_t g = i + res[0] * (j + res[1] * k);
                grid[3 * g] = origin[0] + i * h;
                grid[3 * g + 1] = origin[1] + j * h;
                grid[3 * g + 2] = origin[2] + k * h;
            }
    std::vector<float> Pf(3 * n);
    for (int64_t i = 0; i < 3 * n; ++i) Pf[i] = float(P[i]);
    CellGrid cg;
    cg.build(Pf.data(), n, 7.5);
    std::vector<double> wsrc(n), wgrid(G);
    for (int it = 0; it < iters; ++it) {
        double sigma = orc_sigma_schedule(it, sigma_in);
        orc_tree_refresh_ec(t, N, A);
        orc_tree_eval(t, N, A, radii, P, n, beta, 1, sigma, use_radii, 1, wsrc.data(), nullptr);
        orc_tree_eval(t, N, A, radii, grid.data(), G, beta, 1, sigma, use_radii, 1, wgrid.data(), nullptr);
        double s = 0;
        for (int64_t i = 0; i < n; ++i) s += wsrc[i];
        double iso = s / double(n);
        w_mean[it] = iso;
        if (it % 5 == 0 && it != 0) iso *= 0.5;
        int64_t T = orc_mc(wgrid.data(), res, origin, h, iso, 0, nullptr, nullptr, nullptr);
        std::vector<double> cents(3 * T), nrms(3 * T);
        orc_mc(wgrid.data(), res, origin, h, iso, T, nullptr, cents.data(), nrms.data());
        ntri[it] = T;
        std::vector<float> cf(3 * T);
        for (int64_t i = 0; i < 3 * T; ++i) cf[i] = float(cents[i]);
        const int K = 15;
        std::vector<int> knn(size_t(T) * K);
        parallel_for(T, [&](int64_t b, int64_t e) {
            std::vector<uint64_t> best;
            for (int64_t ti = b; ti < e; ++ti) {
                cg.knn(cf.data() + 3 * ti, K, nullptr, best);
                for (int i = 0; i < K; ++i) knn[ti * K + i] = int(uint32_t(best[i]));
            }
        });
        std::vector<double> acc(3 * n, 0.0);
        for (int64_t ti = 0; ti < T; ++ti)
            for (int i = 0; i < K; ++i) {
                int p = knn[ti * K + i];
                for (int d = 0; d < 3; ++d) acc[3 * p + d] += nrms[3 * ti + d];
            }
        std::vector<uint8_t> zero(n, 0);
        int64_t nz = 0;
        for (int64_t i = 0; i < n; ++i) {
            double l = std::sqrt(acc[3 * i] * acc[3 * i] + acc[3 * i + 1] * acc[3 * i + 1] + acc[3 * i + 2] * acc[3 * i + 2]);
            if (l < 1e-10) {
                zero[i] = 1;
                ++nz;
            } else {
                for (int d = 0; d < 3; ++d) N[3 * i + d] = acc[3 * i + d] / (l + 1e-8);
            }
        }
        nzero[it] = nz;
        if (nz > 10) {
            const int KZ = 10;
            std::vector<double> fill(3 * n, 0.0);
            for (int64_t i = 0; i < n; ++i) {
                if (!zero[i]) continue;
                std::vector<uint64_t> best;
                cg.knn(Pf.data() + 3 * i, KZ, zero.data(), best);
                double v[3] = {0, 0, 0};
                for (int j = 0; j < KZ; ++j) {
                    if (best[j] == ~uint64_t(0)) continue;
                    int p = int(uint32_t(best[j]));
                    for (int d = 0; d < 3; ++d) v[d] += N[3 * p + d];
                }
                double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                for (int d = 0; d < 3; ++d) fill[3 * i + d] = v[d] / (l + 1e-8);
            }
            for (int64_t i = 0; i < n; ++i)
                if (zero[i])
                    for (int d = 0; d < 3; ++d) N[3 * i + d] = fill[3 * i + d];
        }
    }
    delete t;
}

}  // extern "C"
