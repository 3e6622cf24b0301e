// Host construction of an island batch.  Follows the reference's frame and
// Laplacian conventions:
//   param mode: param.py:19-56 (local_frames, cot_weights, _lap_triplets),
//               soft pin weight max|W| per island (param.py:393-396),
//               LSCM boundary terms and pins (param.py:142-179)
//   batch mode: gpu.py:88-142 (1e-30 clamps, thin-triangle mask, batch-global
//               pin weight at every island's first vertex)
#include "setup.hpp"

#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace uva {
namespace {

struct Edge {
    int a, b;        // a < b
    double w;        // half-cotangent weight of one triangle edge
    int dir;         // +1 for a->b as the triangle's half-edge, -1 for b->a
};

struct Island {
    int nv = 0, nt = 0;
    std::vector<double> x1, x2, y2, area, g;   // g: 6 per triangle (corner c: g[6f+2c], g[6f+2c+1])
    std::vector<unsigned char> ok;
    std::vector<double> wabs_max;              // one value: max |w|
    std::vector<int> row_ptr, row_col;         // unique neighbours (local ids), sorted
    std::vector<double> row_val, row_b;        // -sum w, LSCM coupling B_ij
    std::vector<double> diag;
    std::vector<unsigned char> bnd;            // endpoint of an edge used by one triangle
    std::vector<int> inc_ptr, inc_tri, inc_corner;
    int lp0 = -1, lp1 = -1;
    double ldist = 0.0;
    double wmax = 0.0;
    int64_t nnz = 0;
};

inline void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// perm (or NULL): the caller's local id of each vertex of P/T, which fixes the
// order in which the farthest-boundary-pair search visits vertices (the
// reference's ascending vertex order).
void build_island(Island& I, int nv, int nt, const double* P, const int64_t* T, const double* X,
                  const double* area, const double* cots, bool batch, const int64_t* lpins, const int* perm) {
    I.nv = nv;
    I.nt = nt;
    I.x1.resize(nt); I.x2.resize(nt); I.y2.resize(nt); I.area.resize(nt); I.ok.resize(nt);
    I.g.assign(6 * (size_t)nt, 0.0);
    std::vector<double> w(3 * (size_t)nt);
    double wmax = 0.0;
    for (int f = 0; f < nt; ++f) {
        double Xf[3][2];
        double ar;
        double cf[3];
        bool okf;
        if (X && !batch) {
            for (int c = 0; c < 3; ++c) { Xf[c][0] = X[6 * (size_t)f + 2 * c]; Xf[c][1] = X[6 * (size_t)f + 2 * c + 1]; }
            ar = area[f];
        } else {
            const double* p0 = P + 3 * T[3 * (size_t)f];
            const double* p1 = P + 3 * T[3 * (size_t)f + 1];
            const double* p2 = P + 3 * T[3 * (size_t)f + 2];
            double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            double nrm[3];
            cross3(e1, e2, nrm);
            double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            double tiny = batch ? 1e-30 : 1e-300;
            if (batch) l1 = std::max(l1, 1e-30);
            double lden = batch ? l1 : std::max(l1, 1e-300);
            double ex[3] = {e1[0] / lden, e1[1] / lden, e1[2] / lden};
            double nd = std::max(nn, tiny);
            double nh[3] = {nrm[0] / nd, nrm[1] / nd, nrm[2] / nd};
            double ey[3];
            cross3(nh, ex, ey);
            Xf[0][0] = 0.0; Xf[0][1] = 0.0;
            Xf[1][0] = l1; Xf[1][1] = 0.0;
            Xf[2][0] = e2[0] * ex[0] + e2[1] * ex[1] + e2[2] * ex[2];
            Xf[2][1] = e2[0] * ey[0] + e2[1] * ey[1] + e2[2] * ey[2];
            ar = 0.5 * nn;
            if (batch) {
                okf = (nn > 1e-14 * std::max(l1 * l1, 1e-300)) && std::isfinite(nn);
                if (!okf) for (int c = 0; c < 3; ++c) Xf[c][0] = Xf[c][1] = 0.0;
            }
        }
        // cotangent of the angle at corner k (opposite edge k+1 -> k+2)
        for (int k = 0; k < 3; ++k) {
            int i = (k + 1) % 3, j = (k + 2) % 3;
            double ax = Xf[i][0] - Xf[k][0], ay = Xf[i][1] - Xf[k][1];
            double bx = Xf[j][0] - Xf[k][0], by = Xf[j][1] - Xf[k][1];
            double cr = std::fabs(ax * by - ay * bx);
            double dt = ax * bx + ay * by;
            double c = dt / std::max(cr, batch ? 1e-30 : 1e-300);
            cf[k] = std::min(std::max(c, -1e4), 1e4);
        }
        if (cots && !batch) for (int k = 0; k < 3; ++k) cf[k] = cots[3 * (size_t)f + k];
        if (batch) {
            I.x1[f] = std::max(Xf[1][0], 1e-30);
            I.x2[f] = Xf[2][0];
            I.y2[f] = std::fabs(Xf[2][1]) > 1e-30 ? Xf[2][1] : 1e-30;
            I.ok[f] = okf;
            for (int k = 0; k < 3; ++k) w[3 * (size_t)f + k] = okf ? 0.5 * cf[k] : 0.0;
        } else {
            I.x1[f] = Xf[1][0];
            I.x2[f] = Xf[2][0];
            I.y2[f] = Xf[2][1];
            I.ok[f] = std::fabs(Xf[1][0] * Xf[2][1]) >= 1e-300;
            for (int k = 0; k < 3; ++k) w[3 * (size_t)f + k] = 0.5 * cf[k];
        }
        I.area[f] = ar;
        for (int k = 0; k < 3; ++k) wmax = std::max(wmax, std::fabs(w[3 * (size_t)f + k]));
        // rhs corner vectors: corner c gathers w_{c-1} d_{c-1} - w_{c+1} d_{c+1},
        // d_k = X_{k+1} - X_{k+2}   (param.py:457-468 scatter, regrouped per corner)
        double d[3][2];
        for (int k = 0; k < 3; ++k) {
            d[k][0] = Xf[(k + 1) % 3][0] - Xf[(k + 2) % 3][0];
            d[k][1] = Xf[(k + 1) % 3][1] - Xf[(k + 2) % 3][1];
        }
        for (int c = 0; c < 3; ++c) {
            int km = (c + 2) % 3, kp = (c + 1) % 3;
            double wm = w[3 * (size_t)f + km], wp = w[3 * (size_t)f + kp];
            I.g[6 * (size_t)f + 2 * c] = wm * d[km][0] - wp * d[kp][0];
            I.g[6 * (size_t)f + 2 * c + 1] = wm * d[km][1] - wp * d[kp][1];
        }
    }
    I.wmax = wmax;

    // unique undirected edges with summed weights, use counts and boundary
    // coupling.  The edge opposite corner k, taken from corner k+1 to k+2, is
    // one of the triangle's directed half-edges (0->1, 1->2, 2->0), so the
    // directions within a group of equal keys count the half-edges each way.
    // (the edge arrays are per-thread buffers kept across islands, so only
    // their growth touches new memory)
    thread_local std::vector<Edge> E, S;
    thread_local std::vector<int> pos;
    E.resize(3 * (size_t)nt);
    for (int f = 0; f < nt; ++f) {
        for (int k = 0; k < 3; ++k) {
            const int a = (int)T[3 * (size_t)f + (k + 1) % 3], b = (int)T[3 * (size_t)f + (k + 2) % 3];
            Edge& e = E[3 * (size_t)f + k];
            e.a = std::min(a, b);
            e.b = std::max(a, b);
            e.w = w[3 * (size_t)f + k];
            e.dir = a < b ? 1 : -1;
        }
    }
    // sort by (a, b): a counting sort by a (stable), then an insertion sort of
    // each vertex's few edges to larger vertices
    pos.assign(nv + 1, 0);
    for (const Edge& e : E) pos[e.a + 1]++;
    for (int v = 0; v < nv; ++v) pos[v + 1] += pos[v];
    S.resize(E.size());
    for (const Edge& e : E) S[pos[e.a]++] = e;     // pos[v] ends at the start of v + 1
    for (int v = nv; v > 0; --v) pos[v] = pos[v - 1];
    pos[0] = 0;
    for (int v = 0; v < nv; ++v)
        for (int q = pos[v] + 1; q < pos[v + 1]; ++q) {
            const Edge x = S[q];
            int r = q - 1;
            for (; r >= pos[v] && S[r].b > x.b; --r) S[r + 1] = S[r];
            S[r + 1] = x;
        }
    std::vector<int> ua, ubv;
    std::vector<double> uw, ub;
    std::vector<int> ucnt;
    std::vector<int> bsrc;       // sources of boundary half-edges
    ua.reserve(S.size() / 2 + 16); ubv.reserve(S.size() / 2 + 16);
    uw.reserve(S.size() / 2 + 16); ub.reserve(S.size() / 2 + 16); ucnt.reserve(S.size() / 2 + 16);
    for (size_t q = 0; q < S.size();) {
        size_t r = q;
        double ws = 0.0;
        int nab = 0, nba = 0;    // half-edges a->b and b->a (a < b)
        while (r < S.size() && S[r].a == S[q].a && S[r].b == S[q].b) {
            ws += S[r].w;
            if (S[r].dir > 0) ++nab;
            else ++nba;
            ++r;
        }
        const int a = S[q].a, b = S[q].b;
        // every boundary half-edge (its reverse absent) a->b adds B_ab -= 1/2,
        // B_ba += 1/2 (param.py:142-150)
        double beta = 0.0;
        if (nab && !nba) { beta -= 0.5 * nab; bsrc.push_back(a); }
        if (nba && !nab) { beta += 0.5 * nba; bsrc.push_back(b); }
        ua.push_back(a);
        ubv.push_back(b);
        uw.push_back(ws);
        ub.push_back(beta);
        ucnt.push_back((int)(r - q));
        q = r;
    }
    // rows
    std::vector<int> deg(nv, 0);
    for (size_t u = 0; u < ua.size(); ++u) { deg[ua[u]]++; deg[ubv[u]]++; }
    I.row_ptr.assign(nv + 1, 0);
    for (int v = 0; v < nv; ++v) I.row_ptr[v + 1] = I.row_ptr[v] + deg[v];
    I.row_col.resize(I.row_ptr[nv]);
    I.row_val.resize(I.row_ptr[nv]);
    I.row_b.resize(I.row_ptr[nv]);
    I.diag.assign(nv, 0.0);
    I.bnd.assign(nv, 0);
    std::vector<int> fill(I.row_ptr.begin(), I.row_ptr.end() - 1);
    for (size_t u = 0; u < ua.size(); ++u) {
        const int a = ua[u], b = ubv[u];
        int pa = fill[a]++, pb = fill[b]++;
        I.row_col[pa] = b; I.row_val[pa] = -uw[u]; I.row_b[pa] = ub[u];
        I.row_col[pb] = a; I.row_val[pb] = -uw[u]; I.row_b[pb] = -ub[u];
        I.diag[a] += uw[u];
        I.diag[b] += uw[u];
        if (ucnt[u] == 1) { I.bnd[a] = 1; I.bnd[b] = 1; }
    }
    I.nnz = I.row_ptr[nv];
    // rows are sorted: keys (b, a) with b < a precede keys (a, c) with c > a
    // corner incidence (vertex -> (triangle, corner)) in triangle order
    I.inc_ptr.assign(nv + 1, 0);
    for (int f = 0; f < nt; ++f)
        for (int c = 0; c < 3; ++c) I.inc_ptr[T[3 * (size_t)f + c] + 1]++;
    for (int v = 0; v < nv; ++v) I.inc_ptr[v + 1] += I.inc_ptr[v];
    I.inc_tri.resize(3 * (size_t)nt);
    I.inc_corner.resize(3 * (size_t)nt);
    std::vector<int> ip(I.inc_ptr.begin(), I.inc_ptr.end() - 1);
    for (int f = 0; f < nt; ++f)
        for (int c = 0; c < 3; ++c) {
            int v = (int)T[3 * (size_t)f + c];
            I.inc_tri[ip[v]] = f;
            I.inc_corner[ip[v]] = c;
            ip[v]++;
        }
    // LSCM pins: farthest boundary pair (param.py:172-179) unless given;
    // candidates in the caller's ascending vertex order
    std::sort(bsrc.begin(), bsrc.end());
    bsrc.erase(std::unique(bsrc.begin(), bsrc.end()), bsrc.end());
    if (perm) std::sort(bsrc.begin(), bsrc.end(), [&](int a, int b) { return perm[a] < perm[b]; });
    if (lpins) {
        I.lp0 = (int)lpins[0];
        I.lp1 = (int)lpins[1];
    } else if (!bsrc.empty()) {
        double c[3] = {0, 0, 0};
        for (int v : bsrc) for (int k = 0; k < 3; ++k) c[k] += P[3 * (size_t)v + k];
        for (int k = 0; k < 3; ++k) c[k] /= (double)bsrc.size();
        auto far_from = [&](const double* o) {
            int best = bsrc[0];
            double bd = -1.0;
            for (int v : bsrc) {
                double dx = P[3 * (size_t)v] - o[0], dy = P[3 * (size_t)v + 1] - o[1], dz = P[3 * (size_t)v + 2] - o[2];
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > bd) { bd = d2; best = v; }
            }
            return best;
        };
        int i = far_from(c);
        int j = far_from(P + 3 * (size_t)i);
        if (i == j) j = bsrc[0] != i ? bsrc[0] : bsrc.back();
        I.lp0 = i;
        I.lp1 = j;
    }
    if (I.lp0 >= 0) {
        const double* a = P + 3 * (size_t)I.lp0;
        const double* b = P + 3 * (size_t)I.lp1;
        I.ldist = std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) + (b[2] - a[2]) * (b[2] - a[2]));
    }
}

// Vertex order of an island inside the plan: ascending 3-D Morton code of
// the positions (21 bits per axis over the island's bounding box), ties in
// the caller's order.  Consecutive 64-vertex slices are then compact patches
// of the surface, which makes them the aggregates of the coarse level.
std::vector<int> morton_order(int nv, const double* P) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            const double x = P[3 * (size_t)v + k];
            if (std::isfinite(x)) { lo[k] = std::min(lo[k], x); hi[k] = std::max(hi[k], x); }
        }
    double span = 0.0;
    for (int k = 0; k < 3; ++k) span = std::max(span, hi[k] - lo[k]);
    const double sc = span > 0.0 ? 2097151.0 / span : 0.0;
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    std::vector<std::pair<uint64_t, int>> key(nv);
    for (int v = 0; v < nv; ++v) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            double q = (P[3 * (size_t)v + k] - lo[k]) * sc;
            q = std::isfinite(q) ? std::min(std::max(q, 0.0), 2097151.0) : 0.0;
            code |= spread((uint64_t)q) << k;
        }
        key[v] = {code, v};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(nv);
    for (int v = 0; v < nv; ++v) perm[v] = key[v].second;
    return perm;
}

bool spd_inverse(std::vector<double>& A, int n, std::vector<double>& S);

// Inverse of the coarse operator Ac = Z^T A Z of one island (A = Laplacian
// + pin weight; Z = indicator of the aggregates, g consecutive 64-row slices
// each), symmetric, in FP32.  Returns false if Ac is not numerically positive
// definite.
template <class VP, class VC, class VV, class VD>
bool coarse_inverse(int nv, int g, const VP& row_ptr, const VC& row_col, const VV& row_val, const VD& diag, int pin,
                    double wpin, float* out) {
    const int rows = kSlice * g;
    const int ns = (nv + rows - 1) / rows;
    std::vector<double> A((size_t)ns * ns, 0.0);
    for (int i = 0; i < nv; ++i) {
        const int a = i / rows;
        A[(size_t)a * ns + a] += diag[i] + (i == pin ? wpin : 0.0);
        for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) A[(size_t)a * ns + row_col[q] / rows] += row_val[q];
    }
    std::vector<double> S;
    if (!spd_inverse(A, ns, S)) return false;
    for (size_t k = 0; k < S.size(); ++k) out[k] = (float)S[k];
    return true;
}

// Slices per coarse aggregate of an island of ns slices: the smallest g whose
// coarse level fits beside the resident CG's vectors (v2 / v bytes per row and
// component) in the LDS of every member of every team the schedule may give
// the island (at least ceil(ns / res_slices) members, each holding at most
// res_slices slices; kMaxTeam at most), with the members' slice shares and
// aggregate counts as the kernel computes them.
int aggregate_slices(int ns, int res_slices, int v2, int v, bool coupled) {
    if (ns <= 1) return 1;
    const int m_min = std::max(1, (ns + res_slices - 1) / res_slices);
    const int m_max = std::max(m_min, std::min(kMaxTeam, ns));
    for (int g = 1; g < ns; ++g) {
        const int nc = (ns + g - 1) / g;
        bool ok = true;
        for (int m = m_min; m <= m_max && ok; ++m) {
            int nsl = 0, na = 0;
            for (int k = 0; k < m; ++k) {
                const int lo = (int)((long long)ns * k / m), hi = (int)((long long)ns * (k + 1) / m);
                nsl = std::max(nsl, hi - lo);
                if (hi > lo) na = std::max(na, (hi - 1) / g - lo / g + 1);
            }
            ok = res_vec_bytes_of(nsl * kSlice, m > 1 ? halo_rows(v2) : 0, v2, v) + res_coarse_bytes(nc, na, nsl, coupled) <=
                 kLdsVec;
        }
        if (ok) return g;
    }
    return ns;
}

// Inverse of a dense SPD matrix A (n x n, row-major, overwritten by its
// Cholesky factor) into S (full, exactly symmetric).  Returns false if A is
// not numerically positive definite or the inverse is not finite.
bool spd_inverse(std::vector<double>& A, int n, std::vector<double>& S) {
    for (int j = 0; j < n; ++j) {
        double* aj = &A[(size_t)j * n];
        double dj = aj[j];
        for (int k = 0; k < j; ++k) dj -= aj[k] * aj[k];
        if (!(dj > 0.0) || !std::isfinite(dj)) return false;
        const double ljj = std::sqrt(dj);
        aj[j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double* ai = &A[(size_t)i * n];
            double v = ai[j];
            for (int k = 0; k < j; ++k) v -= ai[k] * aj[k];
            ai[j] = v / ljj;
        }
    }
    // M = L^-1 (lower, stored transposed: Mt[c][i] = M[i][c]), then S = M^T M
    std::vector<double> Mt((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        double* mc = &Mt[(size_t)c * n];
        mc[c] = 1.0 / A[(size_t)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            const double* ai = &A[(size_t)i * n];
            double v = 0.0;
            for (int k = c; k < i; ++k) v -= ai[k] * mc[k];
            mc[i] = v / ai[i];
        }
    }
    S.assign((size_t)n * n, 0.0);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b) {
            const double* ma = &Mt[(size_t)a * n];
            const double* mb = &Mt[(size_t)b * n];
            double v = 0.0;
            for (int k = a; k < n; ++k) v += ma[k] * mb[k];
            if (!std::isfinite(v)) return false;
            S[(size_t)a * n + b] = v;
            S[(size_t)b * n + a] = v;
        }
    return true;
}

// Coarse level of the LSCM operator of one island.  The operator acts on
// (u, v) as [[K, B], [-B, K]] (K = Laplacian with its diagonal, B the
// antisymmetric boundary coupling), restricted to the rows and columns that
// are not hard pins: it is the real form of the Hermitian matrix K - iB.
// With Z the indicator of the aggregates (g consecutive 64-row slices, pin
// rows left out) for u and v alike, Ac = Z^T A Z = [[Kc, Bc], [-Bc, Kc]] has
// the same form, and so has its inverse [[Pc, Qc], [-Qc, Pc]].  Writes row a
// of [Pc | Qc] (2 nc FP32 values, Pc symmetric, Qc antisymmetric) per
// aggregate.  Returns false if Ac is not numerically positive definite.
bool lscm_coarse_inverse(int nv, int g, const std::vector<int>& row_ptr, const std::vector<int>& row_col,
                         const std::vector<double>& row_val, const std::vector<double>& row_b,
                         const std::vector<double>& diag, int lp0, int lp1, float* out) {
    const int rows = kSlice * g;
    const int nc = (nv + rows - 1) / rows;
    const int n2 = 2 * nc;
    std::vector<double> A((size_t)n2 * n2, 0.0);
    auto pinned = [&](int i) { return i == lp0 || i == lp1; };
    for (int i = 0; i < nv; ++i) {
        if (pinned(i)) continue;
        const int a = i / rows;
        A[(size_t)a * n2 + a] += diag[i];
        for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
            const int j = row_col[q];
            if (pinned(j)) continue;
            const int b = j / rows;
            A[(size_t)a * n2 + b] += row_val[q];
            A[(size_t)a * n2 + nc + b] += row_b[q];
        }
    }
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            A[(size_t)(nc + a) * n2 + nc + b] = A[(size_t)a * n2 + b];
            A[(size_t)(nc + a) * n2 + b] = -A[(size_t)a * n2 + nc + b];
        }
    std::vector<double> S;
    if (!spd_inverse(A, n2, S)) return false;
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            const double p = 0.5 * (S[(size_t)a * n2 + b] + S[(size_t)(nc + a) * n2 + nc + b]);
            const double q = 0.5 * (S[(size_t)a * n2 + nc + b] - S[(size_t)(nc + a) * n2 + b]);
            out[(size_t)a * n2 + b] = (float)p;
            out[(size_t)a * n2 + nc + b] = (float)q;
        }
    return true;
}

// Threads of the host plan build: the hardware threads, at most 16, or
// UVART_HOST_THREADS.
int host_threads() {
    static const int n = [] {
        const char* e = std::getenv("UVART_HOST_THREADS");
        if (e && std::atoi(e) > 0) return std::atoi(e);
        return (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
    }();
    return n;
}

template <class F>
void parallel_for(int n, F fn, const std::vector<int>* order = nullptr) {
    // (order: the items in the order threads take them, e.g. largest first,
    // so one large item is not left for last)
    int nth = host_threads();
    nth = std::min(nth, std::max(1, n));
    std::atomic<int> next(0);
    auto work = [&]() {
        for (int i = next++; i < n; i = next++) fn(order ? (*order)[i] : i);
    };
    if (nth <= 1) { work(); return; }
    std::vector<std::thread> th;
    for (int k = 0; k < nth; ++k) th.emplace_back(work);
    for (auto& t : th) t.join();
}

}  // namespace

std::string build_host_plan(HostPlan& hp, int64_t n_islands, const int64_t* vert_off, const int64_t* tri_off,
                            const double* P, const int64_t* T, const int64_t* pins, const double* X,
                            const double* area, const double* cots, bool batch, const int64_t* lscm_pins,
                            int wave_max_rows, int slots, int n_xcd, bool teams, bool reorder, bool coarse) {
    if (n_islands <= 0) return "n_islands must be positive";
    if (vert_off[0] != 0 || tri_off[0] != 0) return "offsets must start at 0";
    for (int64_t i = 0; i < n_islands; ++i) {
        if (vert_off[i + 1] < vert_off[i] || tri_off[i + 1] < tri_off[i]) return "offsets must be non-decreasing";
        if (vert_off[i + 1] == vert_off[i]) return "island " + std::to_string(i) + " has no vertices";
    }
    int64_t nv_all = vert_off[n_islands], nt_all = tri_off[n_islands];
    if (nv_all >= (int64_t(1) << 31) || 3 * nt_all >= (int64_t(1) << 31)) return "batch too large for 32-bit indices";
    for (int64_t i = 0; i < n_islands; ++i) {
        int64_t nv = vert_off[i + 1] - vert_off[i];
        for (int64_t f = tri_off[i]; f < tri_off[i + 1]; ++f)
            for (int c = 0; c < 3; ++c) {
                int64_t v = T[3 * f + c];
                if (v < 0 || v >= nv) return "triangle " + std::to_string(f) + " has a vertex outside its island";
            }
        if (pins && (pins[i] < 0 || pins[i] >= nv)) return "pin outside its island";
        if (lscm_pins && (lscm_pins[2 * i] < 0 || lscm_pins[2 * i] >= nv || lscm_pins[2 * i + 1] < 0 ||
                          lscm_pins[2 * i + 1] >= nv))
            return "LSCM pin outside its island";
    }
    int n = (int)n_islands;
    std::vector<Island> isl(n);
    // threads take the islands largest first
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return tri_off[a + 1] - tri_off[a] > tri_off[b + 1] - tri_off[b];
    });
    // per island: perm[v] = caller's local id of plan vertex v, inv its inverse
    std::vector<std::vector<int>> perm(n), inv(n);
    parallel_for(n, [&](int i) {
        int64_t v0 = vert_off[i], t0 = tri_off[i];
        const int nv = (int)(vert_off[i + 1] - v0), nt = (int)(tri_off[i + 1] - t0);
        if (reorder) perm[i] = morton_order(nv, P + 3 * v0);
        else { perm[i].resize(nv); for (int v = 0; v < nv; ++v) perm[i][v] = v; }
        inv[i].resize(nv);
        for (int v = 0; v < nv; ++v) inv[i][perm[i][v]] = v;
        std::vector<double> Pp(3 * (size_t)nv);
        for (int v = 0; v < nv; ++v)
            for (int k = 0; k < 3; ++k) Pp[3 * (size_t)v + k] = P[3 * (v0 + perm[i][v]) + k];
        std::vector<int64_t> Tp(3 * (size_t)nt);
        for (size_t q = 0; q < 3 * (size_t)nt; ++q) Tp[q] = inv[i][T[3 * t0 + (int64_t)q]];
        int64_t lp[2];
        if (lscm_pins) { lp[0] = inv[i][lscm_pins[2 * i]]; lp[1] = inv[i][lscm_pins[2 * i + 1]]; }
        build_island(isl[i], nv, nt, Pp.data(), Tp.data(), X ? X + 6 * t0 : nullptr, area ? area + t0 : nullptr,
                     cots ? cots + 3 * t0 : nullptr, batch, lscm_pins ? lp : nullptr, perm[i].data());
    }, &order);
    hp.n_isl = n;
    hp.n_vert = (int)nv_all;
    hp.n_tri = (int)nt_all;
    hp.batch = batch;
    double wglob = 0.0;
    for (auto& I : isl) wglob = std::max(wglob, I.wmax);
    wglob = std::max(wglob, 1e-12);
    // island table and slice layout
    hp.isl_v0.resize(n); hp.isl_nv.resize(n); hp.isl_t0.resize(n); hp.isl_nt.resize(n);
    hp.isl_s0.resize(n); hp.isl_ns.resize(n); hp.isl_pin.resize(n); hp.isl_wpin.resize(n);
    hp.isl_lpin0.resize(n); hp.isl_lpin1.resize(n); hp.isl_ldist.resize(n); hp.isl_bw.assign(n, 0);
    int64_t nslices = 0;
    std::vector<int64_t> slice_start(n);
    std::vector<std::vector<int>> widths(n);
    std::vector<int64_t> isl_pad(n);
    parallel_for(n, [&](int i) {
        const Island& I = isl[i];
        const int ns = (I.nv + kSlice - 1) / kSlice;
        widths[i].resize(ns);
        int64_t pad = 0;
        for (int s = 0; s < ns; ++s) {
            int wmx = 0;
            for (int r = s * kSlice; r < std::min(I.nv, (s + 1) * kSlice); ++r)
                wmx = std::max(wmx, I.row_ptr[r + 1] - I.row_ptr[r]);
            widths[i][s] = wmx;
            pad += (int64_t)wmx * kSlice;
        }
        isl_pad[i] = pad;
    }, &order);
    int64_t nnz_pad = 0;
    for (int i = 0; i < n; ++i) {
        slice_start[i] = nslices;
        nslices += (int64_t)widths[i].size();
        nnz_pad += isl_pad[i];
        hp.nnz_true += isl[i].nnz;
    }
    if (nnz_pad >= (int64_t(1) << 31)) return "batch too large for 32-bit matrix offsets";
    hp.sl_ptr.assign(nslices + 1, 0);
    // kSlice zero entries after the last slice: the resident SpMV reads them
    // for the steps past a row's width (weight 0, no select after the load);
    // every other entry is written by the threaded fill below
    hp.m_col.resize(nnz_pad + kSlice); hp.m_hi.resize(nnz_pad + kSlice);
    hp.m_lo.resize(nnz_pad + kSlice);
    // the LSCM boundary coupling (Batch-convention plans run no LSCM)
    if (!batch) hp.m_b.resize(nnz_pad + kSlice);
    for (int64_t o = nnz_pad; o < nnz_pad + kSlice; ++o) {
        hp.m_col[o] = 0; hp.m_hi[o] = 0.0f; hp.m_lo[o] = 0.0f;
        if (!batch) hp.m_b[o] = 0.0f;
    }
    hp.diag.resize(nv_all); hp.dinv32.resize(nv_all); hp.vbnd.resize(nv_all); hp.vperm.resize(nv_all);
    // coarse level: aggregates of g consecutive slices, as small as the
    // resident CG's LDS holds (FP32 ARAP vectors; FP64 LSCM vectors with the
    // coupled rows)
    hp.isl_cofs.assign(n, -1);
    hp.isl_cg.assign(n, 1);
    hp.isl_lcg.assign(n, 1);
    {
        int64_t c = 0, c2 = 0;
        hp.isl_lcofs.assign(n, -1);
        for (int i = 0; i < n; ++i) {
            const int ns = (int)widths[i].size();
            hp.isl_cg[i] = aggregate_slices(ns, kResSlices, 8, 4, false);
            hp.isl_lcg[i] = aggregate_slices(ns, kResSlicesF64, 16, 8, true);
            const int64_t nc = (ns + hp.isl_cg[i] - 1) / hp.isl_cg[i];
            const int64_t ncl = (ns + hp.isl_lcg[i] - 1) / hp.isl_lcg[i];
            if (coarse && nc >= 2) { hp.isl_cofs[i] = c; c += nc * nc; }
            // the LSCM operator's coarse level (islands with a boundary): rows of [Pc | Qc]
            if (coarse && ncl >= 2 && isl[i].lp0 >= 0) { hp.isl_lcofs[i] = c2; c2 += 2 * ncl * ncl; }
        }
        hp.cinv.assign(std::max<int64_t>(c, 1), 0.0f);
        hp.lcinv.assign(std::max<int64_t>(c2, 1), 0.0f);
    }
    hp.tv0.resize(nt_all); hp.tv1.resize(nt_all); hp.tv2.resize(nt_all);
    hp.fx1.resize(nt_all); hp.fx2.resize(nt_all); hp.fy2.resize(nt_all); hp.farea.resize(nt_all); hp.fok.resize(nt_all);
    hp.inc_ptr.resize(nv_all + 1);
    hp.inc_ptr[nv_all] = 3 * (int)nt_all;
    hp.inc_tri.resize(3 * nt_all);
    hp.inc_g.resize(6 * nt_all);
    // the triangle-major copy serves the step backward, which Batch-convention
    // plans do not run
    if (!batch) hp.tri_g.resize(6 * nt_all);
    hp.adj_ptr.resize(nv_all + 1);
    int64_t adj_total = 0;
    for (auto& I : isl) adj_total += I.nnz;
    hp.adj_idx.resize(adj_total);
    {
        int64_t s = 0, off = 0;
        for (int i = 0; i < n; ++i)
            for (int w : widths[i]) { hp.sl_ptr[s] = (int)off; off += (int64_t)w * kSlice; ++s; }
        hp.sl_ptr[nslices] = (int)off;
        // first row of every slice (islands' rows are contiguous, so the next
        // slice's first row ends this one)
        hp.sl_row.assign(nslices + 1, (int)nv_all);
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < (int)widths[i].size(); ++q)
                hp.sl_row[slice_start[i] + q] = (int)vert_off[i] + q * kSlice;
    }
    // prefix of adjacency: island bases, then each island's rows in parallel
    {
        std::vector<int64_t> abase(n + 1, 0);
        for (int i = 0; i < n; ++i) abase[i + 1] = abase[i] + isl[i].nnz;
        parallel_for(n, [&](int i) {
            const Island& I = isl[i];
            for (int v = 0; v < I.nv; ++v) hp.adj_ptr[vert_off[i] + v] = (int)(abase[i] + I.row_ptr[v]);
        }, &order);
        hp.adj_ptr[nv_all] = (int)abase[n];
    }
    parallel_for(n, [&](int i) {
        const Island& I = isl[i];
        int v0 = (int)vert_off[i], t0 = (int)tri_off[i];
        hp.isl_v0[i] = v0; hp.isl_nv[i] = I.nv; hp.isl_t0[i] = t0; hp.isl_nt[i] = I.nt;
        hp.isl_s0[i] = (int)slice_start[i]; hp.isl_ns[i] = (int)widths[i].size();
        int lp = inv[i][pins ? (int)pins[i] : 0];
        hp.isl_pin[i] = v0 + lp;
        for (int v = 0; v < I.nv; ++v) hp.vperm[v0 + v] = v0 + perm[i][v];
        double wp = batch ? wglob : (I.nt ? I.wmax : 1.0);
        hp.isl_wpin[i] = wp;
        hp.isl_lpin0[i] = I.lp0 >= 0 ? v0 + I.lp0 : -1;
        hp.isl_lpin1[i] = I.lp1 >= 0 ? v0 + I.lp1 : -1;
        hp.isl_ldist[i] = I.ldist;
        for (int f = 0; f < I.nt; ++f) {
            hp.tv0[t0 + f] = v0 + inv[i][T[3 * ((size_t)t0 + f)]];
            hp.tv1[t0 + f] = v0 + inv[i][T[3 * ((size_t)t0 + f) + 1]];
            hp.tv2[t0 + f] = v0 + inv[i][T[3 * ((size_t)t0 + f) + 2]];
            hp.fx1[t0 + f] = I.x1[f]; hp.fx2[t0 + f] = I.x2[f]; hp.fy2[t0 + f] = I.y2[f];
            hp.farea[t0 + f] = I.area[f]; hp.fok[t0 + f] = I.ok[f];
            if (!batch)
                for (int j = 0; j < 6; ++j) hp.tri_g[6 * ((size_t)t0 + f) + j] = I.g[6 * (size_t)f + j];
        }
        // incidences: global offsets are 3 * t0 + local offsets
        for (int v = 0; v <= I.nv; ++v) hp.inc_ptr[v0 + v] = 3 * t0 + I.inc_ptr[v];
        for (int q = 0; q < 3 * I.nt; ++q) {
            int f = I.inc_tri[q], c = I.inc_corner[q];
            hp.inc_tri[3 * (size_t)t0 + q] = t0 + f;
            hp.inc_g[2 * (3 * (size_t)t0 + q)] = I.g[6 * (size_t)f + 2 * c];
            hp.inc_g[2 * (3 * (size_t)t0 + q) + 1] = I.g[6 * (size_t)f + 2 * c + 1];
        }
        for (int v = 0; v < I.nv; ++v) {
            double dg = I.diag[v];
            hp.diag[v0 + v] = dg;
            double dp = dg + (v == lp ? wp : 0.0);
            hp.dinv32[v0 + v] = (float)(dp > 1e-30 ? 1.0 / dp : 1.0);
            hp.vbnd[v0 + v] = I.bnd[v];
            // neighbours in the caller's ascending order (the Jacobi repair sums them)
            const int a0 = hp.adj_ptr[v0 + v];
            int a = a0;
            for (int q = I.row_ptr[v]; q < I.row_ptr[v + 1]; ++q) {
                hp.adj_idx[a++] = v0 + I.row_col[q];
                hp.isl_bw[i] = std::max(hp.isl_bw[i], std::abs(I.row_col[q] - v));
            }
            // (a vertex has a few neighbours: insertion sort by caller id)
            const int* pi = perm[i].data();
            for (int q = a0 + 1; q < a; ++q) {
                const int x = hp.adj_idx[q], kx = pi[x - v0];
                int r = q - 1;
                for (; r >= a0 && pi[hp.adj_idx[r] - v0] > kx; --r) hp.adj_idx[r + 1] = hp.adj_idx[r];
                hp.adj_idx[r + 1] = x;
            }
        }
        if (hp.isl_cofs[i] >= 0 && !coarse_inverse(I.nv, hp.isl_cg[i], I.row_ptr, I.row_col, I.row_val, I.diag, lp,
                                                   wp, hp.cinv.data() + hp.isl_cofs[i]))
            hp.isl_cofs[i] = -1;
        if (hp.isl_lcofs[i] >= 0 &&
            !lscm_coarse_inverse(I.nv, hp.isl_lcg[i], I.row_ptr, I.row_col, I.row_val, I.row_b, I.diag, I.lp0, I.lp1,
                                 hp.lcinv.data() + hp.isl_lcofs[i]))
            hp.isl_lcofs[i] = -1;
        // sliced ELL
        for (int s = 0; s < (int)widths[i].size(); ++s) {
            int base = hp.sl_ptr[slice_start[i] + s];
            int wdt = widths[i][s];
            for (int l = 0; l < kSlice; ++l) {
                int r = s * kSlice + l;
                int cnt = r < I.nv ? I.row_ptr[r + 1] - I.row_ptr[r] : 0;
                for (int k = 0; k < wdt; ++k) {
                    size_t o = (size_t)base + (size_t)k * kSlice + l;
                    if (k < cnt) {
                        int q = I.row_ptr[r] + k;
                        double v = I.row_val[q];
                        float hi = (float)v;
                        hp.m_col[o] = v0 + I.row_col[q];
                        hp.m_hi[o] = hi;
                        hp.m_lo[o] = (float)(v - (double)hi);
                        if (!batch) hp.m_b[o] = (float)I.row_b[q];
                    } else {
                        hp.m_col[o] = v0 + std::min(r, I.nv - 1);
                        hp.m_hi[o] = 0.0f; hp.m_lo[o] = 0.0f;
                        if (!batch) hp.m_b[o] = 0.0f;
                    }
                }
            }
        }
    }, &order);
    // task order: most expensive first (cost ~ rows^1.5 for Jacobi-PCG on a 2-D Laplacian)
    std::vector<std::pair<double, int>> wv;
    std::vector<int> g;
    for (int i = 0; i < n; ++i) {
        if (isl[i].nv > wave_max_rows) g.push_back(i);
        else wv.push_back({-std::pow((double)isl[i].nv, 1.5), i});
    }
    std::sort(wv.begin(), wv.end());
    for (auto& p : wv) hp.order_wave.push_back(p.second);
    hp.group_isl = g;
    hp.sched_slots = slots;
    hp.sched_teams = teams;
    plan_teams(hp, g, slots, n_xcd, teams, nullptr);
    return "";
}

// Workgroups claim single tasks from one list per XCD unless UVART_SCHED is
// "static" (slot lists planned on the host).
bool claim_tasks_by_default() {
    const char* e = std::getenv("UVART_SCHED");
    return !(e && std::strcmp(e, "static") == 0);
}

void plan_teams(HostPlan& hp, const std::vector<int>& group_islands, int slots, int n_xcd, bool teams,
                const double* est_iters, int res_slices) {
    hp.order_group.clear(); hp.grp_mem.clear(); hp.grp_m.clear(); hp.grp_team.clear(); hp.slot_ptr.clear();
    hp.n_teams = 0;
    hp.max_m = 1;
    n_xcd = std::max(1, std::min(n_xcd, kMaxXcd));
    hp.n_xcd = n_xcd;
    const int S = std::max(1, slots / n_xcd);   // workgroup slots per XCD
    hp.slots_per_xcd = S;
    const int mcap = std::max(1, std::min(kMaxTeam, S));
    // CG iterations per global step (measured ones when given) and the time
    // of one island task on m members: per CG iteration a fixed part (team
    // barriers and latency) and a part per slice of the member
    auto iters = [&](int i) {
        return est_iters ? std::max(1.0, est_iters[i]) : 3.2 * std::pow((double)hp.isl_nv[i], 0.4);
    };
    auto t_of = [&](int i, int m) {
        const int per = (hp.isl_ns[i] + m - 1) / m;
        return iters(i) * ((m > 1 ? hp.cost[1] : hp.cost[0]) + hp.cost[2] * per);
    };
    const int n = (int)group_islands.size();
    std::vector<int> m(n, 1);
    if (teams) {
        double work = 0.0;
        for (int k = 0; k < n; ++k) {
            const int i = group_islands[k];
            m[k] = std::max(1, std::min(mcap, (hp.isl_ns[i] + res_slices - 1) / res_slices));
            work += m[k] * t_of(i, m[k]);
        }
        // islands longer than the ideal makespan get members while that pays
        const double target = work / (double)(S * n_xcd);
        for (int k = 0; k < n; ++k) {
            const int i = group_islands[k];
            while (m[k] < mcap && t_of(i, m[k]) > target && t_of(i, m[k] + 1) < 0.9 * t_of(i, m[k])) ++m[k];
        }
    }
    struct Item {
        double t;   // estimated time on its members
        int isl, m;
    };
    std::vector<Item> items(n);
    for (int k = 0; k < n; ++k) items[k] = {t_of(group_islands[k], m[k]), group_islands[k], m[k]};
    std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return a.t > b.t || (a.t == b.t && a.isl < b.isl);
    });
    // islands to the least-loaded XCD (load = workgroup time)
    std::vector<double> load(n_xcd, 0.0);
    std::vector<std::vector<Item>> lists(n_xcd);
    for (const Item& it : items) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += it.t * it.m;
        lists[x].push_back(it);
    }
    hp.claim_tasks = claim_tasks_by_default();
    if (hp.claim_tasks) {
        // one list per XCD, longest first, a team's members consecutive: the
        // workgroups claim its tasks one at a time
        hp.slot_ptr.assign(n_xcd + 1, 0);
        for (int x = 0; x < n_xcd; ++x) {
            for (const Item& it : lists[x]) {
                const int team = it.m > 1 ? hp.n_teams++ : -1;
                hp.max_m = std::max(hp.max_m, it.m);
                for (int k = 0; k < it.m; ++k) {
                    hp.order_group.push_back(it.isl);
                    hp.grp_mem.push_back(k);
                    hp.grp_m.push_back(it.m);
                    hp.grp_team.push_back(team);
                }
            }
            hp.slot_ptr[x + 1] = (int)hp.order_group.size();
        }
        return;
    }
    // per XCD: longest first, each island's members on the slots that free
    // up first, starting together; every slot keeps the global order
    struct Task {
        int isl, mem, m, team;
    };
    std::vector<std::vector<Task>> st((size_t)S * n_xcd);
    std::vector<int> idx(S);
    for (int x = 0; x < n_xcd; ++x) {
        std::vector<double> free(S, 0.0);
        for (const Item& it : lists[x]) {
            for (int k = 0; k < S; ++k) idx[k] = k;
            std::partial_sort(idx.begin(), idx.begin() + it.m, idx.end(), [&](int a, int b) {
                return free[a] < free[b] || (free[a] == free[b] && a < b);
            });
            std::sort(idx.begin(), idx.begin() + it.m);
            double start = 0.0;
            for (int k = 0; k < it.m; ++k) start = std::max(start, free[idx[k]]);
            const int team = it.m > 1 ? hp.n_teams++ : -1;
            hp.max_m = std::max(hp.max_m, it.m);
            for (int k = 0; k < it.m; ++k) {
                st[(size_t)x * S + idx[k]].push_back({it.isl, k, it.m, team});
                free[idx[k]] = start + it.t;
            }
        }
    }
    hp.slot_ptr.assign(st.size() + 1, 0);
    for (size_t q = 0; q < st.size(); ++q) {
        for (const Task& t : st[q]) {
            hp.order_group.push_back(t.isl);
            hp.grp_mem.push_back(t.mem);
            hp.grp_m.push_back(t.m);
            hp.grp_team.push_back(t.team);
        }
        hp.slot_ptr[q + 1] = (int)hp.order_group.size();
    }
}

std::string chart_regroup(int64_t n, const int64_t* faces, const int64_t* CV, int64_t nv, const int64_t* chart,
                          int64_t nchart, const int64_t* mesh_F, int64_t n_mesh_f, const double* mesh_V,
                          int64_t n_mesh_v, int64_t* keep, int64_t* nk, int64_t* vert_off, int64_t* tri_off, double* P,
                          int64_t* T, int64_t* vorder, int64_t* forder) {
    if (n < 0 || nv < 0 || nchart < 0) return "negative size";
    // range checks in parallel blocks of faces
    constexpr int64_t kBlk = 1 << 16;
    const int nb = (int)((n + kBlk - 1) / kBlk);
    std::vector<char> bad(std::max(nb, 1), 0);
    parallel_for(nb, [&](int b) {
        const int64_t f0 = b * kBlk, f1 = std::min(n, f0 + kBlk);
        char e = 0;
        for (int64_t f = f0; f < f1; ++f) {
            for (int k = 0; k < 3; ++k) e |= (CV[3 * f + k] < 0 || CV[3 * f + k] >= nv) ? 1 : 0;
            e |= (chart[f] < 0 || chart[f] >= nchart) ? 2 : 0;
            if (faces[f] < 0 || faces[f] >= n_mesh_f) e |= 4;
            else
                for (int k = 0; k < 3; ++k)
                    e |= (mesh_F[3 * faces[f] + k] < 0 || mesh_F[3 * faces[f] + k] >= n_mesh_v) ? 8 : 0;
        }
        bad[b] = e;
    });
    char e = 0;
    for (char x : bad) e |= x;
    if (e & 1) return "CV holds vertex ids outside [0, nv)";
    if (e & 2) return "chart ids outside [0, nchart)";
    if (e & 4) return "face ids outside the mesh";
    if (e & 8) return "a mesh face holds a vertex id outside the mesh";
    // each local vertex's chart and mesh vertex from the last of its faces
    // (as an in-order scatter leaves them): the last face of every vertex by
    // an atomic maximum, then that face writes (one writer per vertex)
    std::unique_ptr<std::atomic<int64_t>[]> last(new std::atomic<int64_t>[std::max<int64_t>(nv, 1)]);
    const int nbv0 = (int)((nv + kBlk - 1) / kBlk);
    parallel_for(nbv0, [&](int b) {
        const int64_t v1 = std::min(nv, (int64_t)(b + 1) * kBlk);
        for (int64_t v = (int64_t)b * kBlk; v < v1; ++v) last[v].store(-1, std::memory_order_relaxed);
    });
    parallel_for(nb, [&](int b) {
        const int64_t f1 = std::min(n, (int64_t)(b + 1) * kBlk);
        for (int64_t f = (int64_t)b * kBlk; f < f1; ++f)
            for (int k = 0; k < 3; ++k) {
                std::atomic<int64_t>& l = last[CV[3 * f + k]];
                int64_t cur = l.load(std::memory_order_relaxed);
                while (cur < f && !l.compare_exchange_weak(cur, f, std::memory_order_relaxed)) {
                }
            }
    });
    std::vector<int64_t> cv(nv, -1), corner(nv, 0);
    parallel_for(nb, [&](int b) {
        const int64_t f1 = std::min(n, (int64_t)(b + 1) * kBlk);
        for (int64_t f = (int64_t)b * kBlk; f < f1; ++f)
            for (int k = 0; k < 3; ++k) {
                const int64_t v = CV[3 * f + k];
                if (last[v].load(std::memory_order_relaxed) == f) {
                    cv[v] = chart[f];
                    corner[v] = mesh_F[3 * faces[f] + k];
                }
            }
    });
    for (int64_t v = 0; v < nv; ++v)
        if (cv[v] < 0) return "every vertex of the subset must belong to one of its faces";
    // stable counting sorts by chart: vertices (ascending local id inside a
    // chart) and faces (subset order inside a chart)
    std::vector<int64_t> vpos(nchart + 1, 0), fpos(nchart + 1, 0);
    for (int64_t v = 0; v < nv; ++v) vpos[cv[v] + 1]++;
    for (int64_t f = 0; f < n; ++f) fpos[chart[f] + 1]++;
    for (int64_t c = 0; c < nchart; ++c) {
        vpos[c + 1] += vpos[c];
        fpos[c + 1] += fpos[c];
    }
    int64_t k = 0;
    vert_off[0] = 0;
    tri_off[0] = 0;
    for (int64_t c = 0; c < nchart; ++c)
        if (fpos[c + 1] > fpos[c]) {      // a chart without faces has no vertices either
            keep[k] = c;
            vert_off[k + 1] = vpos[c + 1];
            tri_off[k + 1] = fpos[c + 1];
            ++k;
        }
    *nk = k;
    std::vector<int64_t> local(nv);
    {
        std::vector<int64_t> at(vpos.begin(), vpos.end() - 1);
        for (int64_t v = 0; v < nv; ++v) {
            const int64_t q = at[cv[v]]++;
            vorder[q] = v;
            local[v] = q - vpos[cv[v]];
        }
        std::vector<int64_t> ft(fpos.begin(), fpos.end() - 1);
        for (int64_t f = 0; f < n; ++f) forder[ft[chart[f]]++] = f;
    }
    // gathers: positions in island order, faces in island-local ids
    const int nbv = (int)((nv + kBlk - 1) / kBlk);
    parallel_for(nbv, [&](int b) {
        const int64_t i1 = std::min(nv, (int64_t)(b + 1) * kBlk);
        for (int64_t i = (int64_t)b * kBlk; i < i1; ++i)
            for (int q = 0; q < 3; ++q) P[3 * i + q] = mesh_V[3 * corner[vorder[i]] + q];
    });
    parallel_for(nb, [&](int b) {
        const int64_t i1 = std::min(n, (int64_t)(b + 1) * kBlk);
        for (int64_t i = (int64_t)b * kBlk; i < i1; ++i)
            for (int q = 0; q < 3; ++q) T[3 * i + q] = local[CV[3 * forder[i] + q]];
    });
    return "";
}

std::string build_matrix_plan(HostPlan& hp, int64_t n, const int64_t* indptr, const int64_t* indices,
                              const double* data) {
    if (n <= 0) return "matrix must have at least one row";
    if (n >= (int64_t(1) << 31)) return "matrix too large for 32-bit indices";
    if (indptr[0] != 0) return "indptr must start at 0";
    for (int64_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) return "indptr must be non-decreasing";
    for (int64_t q = 0; q < indptr[n]; ++q)
        if (indices[q] < 0 || indices[q] >= n) return "column index out of range";
    int nv = (int)n;
    hp = HostPlan();
    hp.n_isl = 1;
    hp.n_vert = nv;
    hp.n_tri = 0;
    hp.isl_v0 = {0}; hp.isl_nv = {nv}; hp.isl_t0 = {0}; hp.isl_nt = {0}; hp.isl_s0 = {0};
    int ns = (nv + kSlice - 1) / kSlice;
    hp.isl_ns = {ns};
    hp.isl_pin = {-1}; hp.isl_wpin = {0.0}; hp.isl_lpin0 = {-1}; hp.isl_lpin1 = {-1}; hp.isl_ldist = {0.0};
    hp.isl_bw = {0};
    // rows: off-diagonal entries with duplicates summed, columns ascending
    std::vector<std::vector<std::pair<int, double>>> rows(nv);
    hp.diag.assign(nv, 0.0);
    for (int i = 0; i < nv; ++i) {
        auto& r = rows[i];
        for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q) {
            int j = (int)indices[q];
            if (j == i) hp.diag[i] += data[q];
            else r.push_back({j, data[q]});
        }
        std::sort(r.begin(), r.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
            return a.first < b.first;
        });
        size_t w = 0;
        for (size_t k = 0; k < r.size(); ++k) {
            if (w && r[w - 1].first == r[k].first) r[w - 1].second += r[k].second;
            else r[w++] = r[k];
        }
        r.resize(w);
        hp.nnz_true += (int64_t)w;
        for (const auto& e : r) hp.isl_bw[0] = std::max(hp.isl_bw[0], std::abs(e.first - i));
    }
    hp.sl_ptr.assign(ns + 1, 0);
    int64_t off = 0;
    std::vector<int> widths(ns);
    for (int s = 0; s < ns; ++s) {
        int wmx = 0;
        for (int r = s * kSlice; r < std::min(nv, (s + 1) * kSlice); ++r) wmx = std::max(wmx, (int)rows[r].size());
        widths[s] = wmx;
        hp.sl_ptr[s] = (int)off;
        off += (int64_t)wmx * kSlice;
    }
    if (off >= (int64_t(1) << 31)) return "matrix too large for 32-bit offsets";
    hp.sl_ptr[ns] = (int)off;
    hp.m_col.assign(off + kSlice, 0); hp.m_hi.assign(off + kSlice, 0.0f);
    hp.m_lo.assign(off + kSlice, 0.0f); hp.m_b.assign(off + kSlice, 0.0f);
    for (int s = 0; s < ns; ++s)
        for (int l = 0; l < kSlice; ++l) {
            int r = s * kSlice + l;
            for (int k = 0; k < widths[s]; ++k) {
                size_t o = (size_t)hp.sl_ptr[s] + (size_t)k * kSlice + l;
                if (r < nv && k < (int)rows[r].size()) {
                    double v = rows[r][k].second;
                    float hi = (float)v;
                    hp.m_col[o] = rows[r][k].first;
                    hp.m_hi[o] = hi;
                    hp.m_lo[o] = (float)(v - (double)hi);
                } else {
                    hp.m_col[o] = std::min(r, nv - 1);
                    hp.m_hi[o] = 0.0f;
                    hp.m_lo[o] = 0.0f;
                }
            }
        }
    hp.dinv32.resize(nv);
    hp.vbnd.assign(nv, 0);
    hp.adj_ptr.assign(nv + 1, 0);
    for (int i = 0; i < nv; ++i) {
        double g = hp.diag[i];
        hp.dinv32[i] = (float)(std::fabs(g) > 1e-300 ? 1.0 / g : 1.0);
        hp.adj_ptr[i + 1] = hp.adj_ptr[i] + (int)rows[i].size();
    }
    hp.adj_idx.reserve(hp.adj_ptr[nv]);
    for (int i = 0; i < nv; ++i)
        for (auto& e : rows[i]) hp.adj_idx.push_back(e.first);
    hp.inc_ptr.assign(nv + 1, 0);
    hp.vperm.resize(nv);
    for (int i = 0; i < nv; ++i) hp.vperm[i] = i;
    hp.sl_row.assign(ns + 1, nv);
    for (int q = 0; q < ns; ++q) hp.sl_row[q] = q * kSlice;
    hp.isl_lcofs.assign(1, -1);   // no LSCM system on a general matrix
    hp.lcinv.assign(1, 0.0f);
    // coarse level over the 64-row slices of the given order
    hp.isl_cofs.assign(1, -1);
    const int g = (ns + kCoarseN - 1) / kCoarseN;
    const int nc = (ns + g - 1) / g;
    hp.isl_cg.assign(1, g);
    hp.isl_lcg.assign(1, g);
    hp.cinv.assign(1, 0.0f);
    if (nc >= 2) {
        std::vector<double> val(hp.adj_idx.size());
        for (int i = 0, q = 0; i < nv; ++i)
            for (auto& e : rows[i]) val[q++] = e.second;
        hp.cinv.assign((size_t)nc * nc, 0.0f);
        if (coarse_inverse(nv, g, hp.adj_ptr, hp.adj_idx, val, hp.diag, -1, 0.0, hp.cinv.data())) hp.isl_cofs[0] = 0;
    }
    if (nv > kWaveMaxRows) {
        hp.group_isl = {0};
        plan_teams(hp, {0}, 1, 1, false, nullptr);
    }
    else hp.order_wave.push_back(0);
    return "";
}

// param.py:478-516: move each non-boundary corner of a triangle with signed
// area <= 0 to the mean of its unique 1-ring neighbours, in triangle order,
// in place; stop after a round without such triangles.
void repair_sequential(double* U, int64_t n, const int64_t* T, int64_t t, int rounds) {
    std::vector<int64_t> keys;
    keys.reserve(6 * (size_t)t);
    for (int64_t f = 0; f < t; ++f)
        for (int k = 0; k < 3; ++k) {
            int64_t a = T[3 * f + k], b = T[3 * f + (k + 1) % 3];
            keys.push_back(a * n + b);
            keys.push_back(b * n + a);
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<int64_t> ptr(n + 1, 0);
    for (int64_t k : keys) ptr[k / n + 1]++;
    for (int64_t v = 0; v < n; ++v) ptr[v + 1] += ptr[v];
    // boundary: endpoints of directed half-edges whose reverse is absent
    std::vector<int64_t> dk;
    dk.reserve(3 * (size_t)t);
    for (int64_t f = 0; f < t; ++f)
        for (int k = 0; k < 3; ++k) dk.push_back(T[3 * f + k] * n + T[3 * f + (k + 1) % 3]);
    std::vector<int64_t> sdk(dk);
    std::sort(sdk.begin(), sdk.end());
    std::vector<unsigned char> bnd(n, 0);
    for (int64_t f = 0; f < t; ++f)
        for (int k = 0; k < 3; ++k) {
            int64_t a = T[3 * f + k], b = T[3 * f + (k + 1) % 3];
            if (!std::binary_search(sdk.begin(), sdk.end(), b * n + a)) { bnd[a] = 1; bnd[b] = 1; }
        }
    for (int r = 0; r < rounds; ++r) {
        int64_t bad = 0;
        for (int64_t f = 0; f < t; ++f) {
            int64_t a = T[3 * f], b = T[3 * f + 1], c = T[3 * f + 2];
            double ar = (U[2 * b] - U[2 * a]) * (U[2 * c + 1] - U[2 * a + 1]) -
                        (U[2 * c] - U[2 * a]) * (U[2 * b + 1] - U[2 * a + 1]);
            if (ar > 0) continue;
            ++bad;
            for (int k = 0; k < 3; ++k) {
                int64_t v = T[3 * f + k];
                if (bnd[v]) continue;
                double sx = 0.0, sy = 0.0;
                int64_t cnt = 0;
                for (int64_t q = ptr[v]; q < ptr[v + 1]; ++q) {
                    int64_t w = keys[q] % n;
                    sx += U[2 * w];
                    sy += U[2 * w + 1];
                    ++cnt;
                }
                if (cnt > 0) { U[2 * v] = sx / cnt; U[2 * v + 1] = sy / cnt; }
            }
        }
        if (bad == 0) break;
    }
}

}  // namespace uva
