// Periodic Voronoi geometry producer (host, C++17 + OpenMP): the mesh record the hydro step reads (include/luna_mesh.h).
//
// Each cell is built on its own: a box around the generator is clipped by the bisector planes of candidate neighbours in
// order of distance until no remaining candidate can cut it (a candidate at distance |d| cuts only if |d| < 2 r_max, r_max
// the cell's farthest vertex). 3D cells are convex polyhedra kept as their dual triangulation: every vertex is the triple of
// planes meeting there, oriented counter-clockwise on the Gauss sphere, so a clip removes the vertices outside the plane
// and closes the hole with one new vertex per boundary edge, placed on that polyhedron edge. 2D cells are convex polygons.
// A face is kept when both of its cells see each other; its data come from the owner (the lower index). Nearest periodic
// images only: every cell must stay within a quarter of the box, which the build checks.
#include "luna_mesh.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_error;

constexpr int MAXP = 160;  // planes a cell may collect (6 box planes + clipping neighbours)
constexpr int MAXT = 320;  // live dual triangles (polyhedron vertices)
constexpr int BITS = 30;   // snapping grid: 2^-30 of the longest side

struct Candidate {
    double d2;
    int j;
    double d[3];
};

struct FaceOut {
    int j;
    double area;      // |A|
    double fc[3];     // face centroid relative to this cell's generator
};

// The topology of one face of a cell's polyhedron (polygon) for the geometry's derivative: the neighbour behind it, the
// sign that makes its area positive, and its vertices in cycle order, each as the other generators whose bisectors meet
// there (3D: va and vb, two; 2D: va, one per end, the neighbour of the adjacent edge).
struct FaceTopo {
    int j, sigma;
    std::vector<int> va, vb;
};

struct CellOut {
    double volume = 0, centroid[3] = {0, 0, 0};
    std::vector<FaceOut> faces;
    bool want_topo = false;
    std::vector<FaceTopo> topo;
};

double wrap(double d, double side) { return d - side * std::round(d / side); }

// Uniform bins of the box, about one generator each.
struct Bins {
    int dim = 3, nb[3] = {1, 1, 1};
    double size[3] = {1, 1, 1};
    std::vector<int> start, items;

    void build(int D, int N, const double* x, const double* side) {
        dim = D;
        double volume = 1;
        for (int a = 0; a < D; ++a) volume *= side[a];
        const double h = std::pow(volume / N, 1.0 / D);
        long total = 1;
        for (int a = 0; a < D; ++a) {
            nb[a] = std::max(1, static_cast<int>(std::floor(side[a] / h)));
            size[a] = side[a] / nb[a];
            total *= nb[a];
        }
        for (int a = D; a < 3; ++a) nb[a] = 1;
        std::vector<int> bin(N), count(total + 1, 0);
        for (int i = 0; i < N; ++i) {
            bin[i] = index_of(x + static_cast<size_t>(i) * D);
            ++count[bin[i] + 1];
        }
        for (long b = 0; b < total; ++b) count[b + 1] += count[b];
        start = count;
        items.assign(N, 0);
        std::vector<int> fill(count.begin(), count.end() - 1);
        for (int i = 0; i < N; ++i) items[fill[bin[i]]++] = i;
    }
    int coord(double x, int a) const {
        int c = static_cast<int>(std::floor(x / size[a]));
        return std::min(std::max(c, 0), nb[a] - 1);
    }
    int index_of(const double* p) const {
        int b = 0;
        for (int a = dim - 1; a >= 0; --a) b = b * nb[a] + coord(p[a], a);
        return b;
    }
};

// Candidates of cell i within `r` bins (each other generator once, at its nearest image), sorted by distance.
// Returns the radius within which the list is complete.
double gather(const Bins& bins, int D, const double* x, const double* side, int i, int r, std::vector<Candidate>& out) {
    out.clear();
    const double* xi = x + static_cast<size_t>(i) * D;
    int c[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool all = false;
    for (int a = 0; a < D; ++a) {
        c[a] = bins.coord(xi[a], a);
        if (2 * r + 1 >= bins.nb[a]) all = true;
        lo[a] = c[a] - r;
        hi[a] = c[a] + r;
    }
    auto visit = [&](int b) {
        for (int k = bins.start[b]; k < bins.start[b + 1]; ++k) {
            const int j = bins.items[k];
            if (j == i) continue;
            Candidate cand;
            cand.j = j;
            cand.d[2] = 0;
            double d2 = 0;
            for (int a = 0; a < D; ++a) {
                cand.d[a] = wrap(x[static_cast<size_t>(j) * D + a] - xi[a], side[a]);
                d2 += cand.d[a] * cand.d[a];
            }
            cand.d2 = d2;
            out.push_back(cand);
        }
    };
    double complete = INFINITY;
    if (all) {
        const int total = bins.nb[0] * bins.nb[1] * bins.nb[2];
        for (int b = 0; b < total; ++b) visit(b);
    } else {
        for (int a = 0; a < D; ++a) complete = std::min(complete, r * bins.size[a]);
        for (int z = (D == 3 ? lo[2] : 0); z <= (D == 3 ? hi[2] : 0); ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int xx = lo[0]; xx <= hi[0]; ++xx) {
                    const int bx = (xx % bins.nb[0] + bins.nb[0]) % bins.nb[0];
                    const int by = (y % bins.nb[1] + bins.nb[1]) % bins.nb[1];
                    const int bz = D == 3 ? (z % bins.nb[2] + bins.nb[2]) % bins.nb[2] : 0;
                    visit((bz * bins.nb[1] + by) * bins.nb[0] + bx);
                }
    }
    std::sort(out.begin(), out.end(), [](const Candidate& a, const Candidate& b) { return a.d2 < b.d2 || (a.d2 == b.d2 && a.j < b.j); });
    return complete;
}

// A convex polyhedron around the origin as its dual triangulation.
struct Poly3 {
    int np = 0, nt = 0;
    double pn[MAXP][3], pb[MAXP];  // plane k: pn . x <= pb
    int pj[MAXP];                  // the neighbour behind plane k, -1 for the box
    uint8_t t[MAXT][3];
    double v[MAXT][3];
    uint16_t edge[MAXP][MAXP];     // edge[p][q]: the triangle holding the directed edge p -> q
    bool out[MAXT];

    void box(double H) {
        np = 6;
        for (int k = 0; k < 6; ++k) {
            pn[k][0] = pn[k][1] = pn[k][2] = 0;
            pn[k][k / 2] = k % 2 == 0 ? 1.0 : -1.0;
            pb[k] = H;
            pj[k] = -1;
        }
        nt = 0;
        for (int sx = 0; sx < 2; ++sx)
            for (int sy = 0; sy < 2; ++sy)
                for (int sz = 0; sz < 2; ++sz) {
                    int a = sx, b = 2 + sy, c = 4 + sz;  // planes +x/-x, +y/-y, +z/-z
                    // counter-clockwise on the Gauss sphere: det(n_a, n_b, n_c) > 0
                    const double det = pn[a][0] * (pn[b][1] * pn[c][2] - pn[b][2] * pn[c][1]) -
                                       pn[a][1] * (pn[b][0] * pn[c][2] - pn[b][2] * pn[c][0]) +
                                       pn[a][2] * (pn[b][0] * pn[c][1] - pn[b][1] * pn[c][0]);
                    if (det < 0) std::swap(b, c);
                    t[nt][0] = static_cast<uint8_t>(a);
                    t[nt][1] = static_cast<uint8_t>(b);
                    t[nt][2] = static_cast<uint8_t>(c);
                    v[nt][0] = sx == 0 ? H : -H;
                    v[nt][1] = sy == 0 ? H : -H;
                    v[nt][2] = sz == 0 ? H : -H;
                    ++nt;
                }
        index();
    }
    void index() {
        for (int i = 0; i < nt; ++i)
            for (int e = 0; e < 3; ++e) edge[t[i][e]][t[i][(e + 1) % 3]] = static_cast<uint16_t>(i);
    }
    double rmax2() const {
        double r = 0;
        for (int i = 0; i < nt; ++i) r = std::max(r, v[i][0] * v[i][0] + v[i][1] * v[i][1] + v[i][2] * v[i][2]);
        return r;
    }
    // Clips by d . x <= |d|^2 / 2, the bisector towards the neighbour at d. Returns -1 on overflow, 0 if nothing changed.
    int clip(const double* d, double d2, int j) {
        const double b = 0.5 * d2;
        double s[MAXT];
        int removed = 0;
        for (int i = 0; i < nt; ++i) {
            s[i] = d[0] * v[i][0] + d[1] * v[i][1] + d[2] * v[i][2] - b;
            out[i] = s[i] > 0;
            removed += out[i];
        }
        if (removed == 0) return 0;
        if (removed == nt || np >= MAXP) return -1;
        const int k = np++;
        pn[k][0] = d[0];
        pn[k][1] = d[1];
        pn[k][2] = d[2];
        pb[k] = b;
        pj[k] = j;
        int m = nt;
        for (int i = 0; i < nt; ++i) {
            if (!out[i]) continue;
            for (int e = 0; e < 3; ++e) {
                const int p = t[i][e], q = t[i][(e + 1) % 3];
                const int u = edge[q][p];
                if (out[u]) continue;
                if (m >= MAXT) return -1;
                const double f = s[u] / (s[u] - s[i]);  // where the plane cuts the edge from v[u] to v[i]
                for (int a = 0; a < 3; ++a) v[m][a] = v[u][a] + f * (v[i][a] - v[u][a]);
                t[m][0] = static_cast<uint8_t>(p);
                t[m][1] = static_cast<uint8_t>(q);
                t[m][2] = static_cast<uint8_t>(k);
                out[m] = false;
                ++m;
            }
        }
        int w = 0;
        for (int i = 0; i < m; ++i) {
            if (out[i]) continue;
            if (w != i) {
                std::memcpy(t[w], t[i], 3);
                std::memcpy(v[w], v[i], sizeof v[w]);
            }
            out[w] = false;
            ++w;
        }
        nt = w;
        index();
        return 1;
    }
};

// A convex polygon around the origin, counter-clockwise; label[i]: the plane of the edge from vertex i to i + 1.
struct Poly2 {
    int m = 0, np = 0;
    double v[MAXT][2];
    int label[MAXT];
    int pj[MAXP];

    void box(double H) {
        m = 4;
        np = 4;
        const double xs[4] = {H, -H, -H, H}, ys[4] = {H, H, -H, -H};
        for (int i = 0; i < 4; ++i) {
            v[i][0] = xs[i];
            v[i][1] = ys[i];
            label[i] = i;  // box edges
            pj[i] = -1;
        }
    }
    double rmax2() const {
        double r = 0;
        for (int i = 0; i < m; ++i) r = std::max(r, v[i][0] * v[i][0] + v[i][1] * v[i][1]);
        return r;
    }
    int clip(const double* d, double d2, int j) {
        const double b = 0.5 * d2;
        double s[MAXT];
        int removed = 0;
        for (int i = 0; i < m; ++i) {
            s[i] = d[0] * v[i][0] + d[1] * v[i][1] - b;
            removed += s[i] > 0;
        }
        if (removed == 0) return 0;
        if (removed == m || np >= MAXP) return -1;
        const int k = np++;
        pj[k] = j;
        double nv[MAXT][2];
        int nl[MAXT], n = 0;
        for (int i = 0; i < m; ++i) {
            const int u = (i + 1) % m;
            if (n + 2 >= MAXT) return -1;
            if (s[i] <= 0) {
                nv[n][0] = v[i][0];
                nv[n][1] = v[i][1];
                nl[n++] = label[i];
                if (s[u] > 0) {
                    const double f = s[i] / (s[i] - s[u]);
                    nv[n][0] = v[i][0] + f * (v[u][0] - v[i][0]);
                    nv[n][1] = v[i][1] + f * (v[u][1] - v[i][1]);
                    nl[n++] = k;
                }
            } else if (s[u] <= 0) {
                const double f = s[i] / (s[i] - s[u]);
                nv[n][0] = v[i][0] + f * (v[u][0] - v[i][0]);
                nv[n][1] = v[i][1] + f * (v[u][1] - v[i][1]);
                nl[n++] = label[i];
            }
        }
        m = n;
        std::memcpy(v, nv, sizeof(double) * 2 * n);
        std::memcpy(label, nl, sizeof(int) * n);
        return 1;
    }
};

// Faces, volume and centroid of a finished 3D cell (coordinates relative to its generator).
bool finish3(Poly3& P, const std::vector<Candidate>& cands, CellOut& cell, std::string& why) {
    (void)cands;
    cell.faces.clear();
    cell.topo.clear();
    double volume = 0, moment[3] = {0, 0, 0};
    static thread_local std::vector<int> seen;
    static thread_local std::vector<int> other_a, other_b;
    seen.assign(P.np, 0);
    for (int s = 0; s < P.nt; ++s)
        for (int e = 0; e < 3; ++e) {
            const int k = P.t[s][e];
            if (seen[k]) continue;
            seen[k] = 1;
            if (k < 6) {
                why = "cell not closed by its neighbours";
                return false;
            }
            // Walk the face of plane k: from triangle (k, x, y) on to the triangle holding the edge k -> y.
            double poly[MAXT][3];
            int count = 0, cur = s;
            other_a.clear();
            other_b.clear();
            do {
                int r = 0;
                while (P.t[cur][r] != k) ++r;
                const int y = P.t[cur][(r + 2) % 3];
                if (cell.want_topo) {
                    other_a.push_back(P.pj[P.t[cur][(r + 1) % 3]]);
                    other_b.push_back(P.pj[y]);
                }
                std::memcpy(poly[count++], P.v[cur], sizeof poly[0]);
                cur = P.edge[k][y];
                if (count >= MAXT) {
                    why = "face walk did not close";
                    return false;
                }
            } while (cur != s);
            const double* d = P.pn[k];
            const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const double n[3] = {d[0] / dn, d[1] / dn, d[2] / dn};
            double area = 0, c[3] = {0, 0, 0};
            for (int q = 1; q + 1 < count; ++q) {
                double a[3], b[3];
                for (int x = 0; x < 3; ++x) {
                    a[x] = poly[q][x] - poly[0][x];
                    b[x] = poly[q + 1][x] - poly[0][x];
                }
                const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
                const double tri = 0.5 * (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2]);
                area += tri;
                for (int x = 0; x < 3; ++x) c[x] += tri * (poly[0][x] + poly[q][x] + poly[q + 1][x]) * (1.0 / 3.0);
            }
            const int sigma = area < 0 ? -1 : 1;
            if (area < 0) {
                area = -area;
                for (double& x : c) x = -x;
            }
            if (!(area > 0)) continue;  // a face of no extent: no face
            if (cell.want_topo) cell.topo.push_back(FaceTopo{P.pj[k], sigma, other_a, other_b});
            FaceOut f;
            f.j = P.pj[k];
            f.area = area;
            for (int x = 0; x < 3; ++x) f.fc[x] = c[x] / area;
            const double pyramid = area * dn / 6.0;  // |A| |d| / 6: the face at |d| / 2 from the generator
            volume += pyramid;
            for (int x = 0; x < 3; ++x) moment[x] += pyramid * 0.75 * f.fc[x];
            cell.faces.push_back(f);
        }
    cell.volume = volume;
    for (int x = 0; x < 3; ++x) cell.centroid[x] = moment[x] / volume;
    return true;
}

bool finish2(Poly2& P, CellOut& cell, std::string& why) {
    cell.faces.clear();
    cell.topo.clear();
    double volume = 0, moment[2] = {0, 0};
    for (int i = 0; i < P.m; ++i) {
        const int k = P.label[i];
        if (k < 4) {
            why = "cell not closed by its neighbours";
            return false;
        }
        const int u = (i + 1) % P.m;
        const double ex = P.v[u][0] - P.v[i][0], ey = P.v[u][1] - P.v[i][1];
        const double length = std::sqrt(ex * ex + ey * ey);
        if (!(length > 0)) continue;
        const double mx = 0.5 * (P.v[u][0] + P.v[i][0]), my = 0.5 * (P.v[u][1] + P.v[i][1]);
        // distance of the edge's line from the generator: the bisector at |d| / 2
        const double h = std::fabs(P.v[i][0] * ey - P.v[i][1] * ex) / length;
        const double tri = 0.5 * length * h;
        if (cell.want_topo) {
            const double cross = P.v[i][0] * P.v[u][1] - P.v[i][1] * P.v[u][0];
            cell.topo.push_back(FaceTopo{P.pj[k], cross < 0 ? -1 : 1, {P.pj[P.label[(i + P.m - 1) % P.m]], P.pj[P.label[u]]}, {}});
        }
        FaceOut f;
        f.j = P.pj[k];
        f.area = length;
        f.fc[0] = mx;
        f.fc[1] = my;
        f.fc[2] = 0;
        volume += tri;
        moment[0] += tri * (2.0 / 3.0) * mx;
        moment[1] += tri * (2.0 / 3.0) * my;
        cell.faces.push_back(f);
    }
    cell.volume = volume;
    cell.centroid[0] = moment[0] / volume;
    cell.centroid[1] = moment[1] / volume;
    cell.centroid[2] = 0;
    return true;
}

// Every cell's polyhedron (polygon) from the snapped positions x; with `topo`, also each face's vertex topology. Sets
// g_error and returns false when a cell cannot be built.
bool build_cells(int D, int N, const double* xp, const double* side, bool topo, std::vector<CellOut>& cells) {
    Bins bins;
    bins.build(D, N, xp, side);
    for (CellOut& c : cells) c.want_topo = topo;
    std::atomic<int> failed{-1};
    std::string why;
    const double quarter = 0.25 * *std::min_element(side, side + D);
#pragma omp parallel
    {
        std::vector<Candidate> cands;
        Poly3* p3 = D == 3 ? new Poly3 : nullptr;
        Poly2* p2 = D == 2 ? new Poly2 : nullptr;
        std::string local;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < N; ++i) {
            if (failed.load() >= 0) continue;
            bool ok = false;
            for (int r = 2; r <= 64 && !ok; r *= 2) {
                const double complete = gather(bins, D, xp, side, i, r, cands);
                const double H = 0.5 * *std::min_element(side, side + D);
                if (D == 3)
                    p3->box(H);
                else
                    p2->box(H);
                double r2 = D == 3 ? p3->rmax2() : p2->rmax2();
                bool overflow = false;
                size_t c = 0;
                for (; c < cands.size(); ++c) {
                    if (cands[c].d2 > 4 * r2) break;
                    const int res = D == 3 ? p3->clip(cands[c].d, cands[c].d2, cands[c].j) : p2->clip(cands[c].d, cands[c].d2, cands[c].j);
                    if (res < 0) {
                        overflow = true;
                        break;
                    }
                    if (res > 0) r2 = D == 3 ? p3->rmax2() : p2->rmax2();
                }
                if (overflow) {
                    local = "cell " + std::to_string(i) + ": clipping overflow (degenerate or far too many neighbours)";
                    break;
                }
                if (4 * r2 > complete * complete) continue;  // a generator beyond the searched bins could still cut
                if (std::sqrt(r2) > quarter) {
                    local = "cell " + std::to_string(i) + ": larger than a quarter of the box (nearest images only)";
                    break;
                }
                ok = D == 3 ? finish3(*p3, cands, cells[i], local) : finish2(*p2, cells[i], local);
                if (!ok) local = "cell " + std::to_string(i) + ": " + local;
                break;
            }
            if (!ok) {
                if (local.empty()) local = "cell " + std::to_string(i) + ": neighbour search did not converge";
                int expected = -1;
                if (failed.compare_exchange_strong(expected, i)) {
#pragma omp critical
                    why = local;
                }
            }
        }
        delete p3;
        delete p2;
    }
    if (failed.load() >= 0) {
        g_error = "luna_mesh_build: " + why;
        return false;
    }
    return true;
}

template <class T>
T* copy_out(const std::vector<T>& v) {
    T* p = new T[v.empty() ? 1 : v.size()];
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" const char* luna_mesh_last_error(void) { return g_error.c_str(); }

extern "C" void luna_mesh_release(luna_mesh_record* r) {
    if (!r) return;
    delete[] r->owner;
    delete[] r->neighbour;
    delete[] r->start;
    delete[] r->list;
    delete[] r->area;
    delete[] r->face_centroid;
    delete[] r->centroid;
    delete[] r->volume;
    delete[] r->positions;
    std::memset(r, 0, sizeof *r);
}

extern "C" int luna_mesh_build(int D, int N, const double* positions, const double* side_in, luna_mesh_record* out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!out || !positions || !side_in) {
        g_error = "luna_mesh_build: null argument";
        return 1;
    }
    std::memset(out, 0, sizeof *out);
    if (D != 2 && D != 3) {
        g_error = "luna_mesh_build: dim must be 2 or 3";
        return 1;
    }
    if (N < 8) {
        g_error = "luna_mesh_build: at least 8 generators";
        return 1;
    }
    double side[3] = {0, 0, 0}, longest = 0;
    for (int a = 0; a < D; ++a) {
        side[a] = side_in[a];
        if (!(side[a] > 0)) {
            g_error = "luna_mesh_build: box sides must be positive";
            return 1;
        }
        longest = std::max(longest, side[a]);
    }
    // Snap to the grid of 2^-30 of the longest side, wrapped into the box.
    const double step = longest / static_cast<double>(1LL << BITS);
    long long period[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) period[a] = std::llround(side[a] / step);
    std::vector<double> x(static_cast<size_t>(N) * D);
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < D; ++a) {
            long long k = std::llround(positions[static_cast<size_t>(i) * D + a] * (1.0 / step)) % period[a];
            if (k < 0) k += period[a];
            x[static_cast<size_t>(i) * D + a] = static_cast<double>(k) * step;
        }
    std::vector<CellOut> cells(N);
    if (!build_cells(D, N, x.data(), side, false, cells)) return 2;
    // Faces: a pair is kept when both cells see each other; numbered by owner (lower index), then neighbour.
    std::vector<int> owned(N + 1, 0);
    auto sees = [&](int j, int i) {
        for (const FaceOut& f : cells[j].faces)
            if (f.j == i) return true;
        return false;
    };
    std::vector<std::vector<int>> keep(N);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < N; ++i)
        std::sort(cells[i].faces.begin(), cells[i].faces.end(), [](const FaceOut& a, const FaceOut& b) { return a.j < b.j; });
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < N; ++i) {
        const auto& faces = cells[i].faces;
        for (int k = 0; k < static_cast<int>(faces.size()); ++k)
            if (faces[k].j > i && sees(faces[k].j, i)) keep[i].push_back(k);
        owned[i + 1] = static_cast<int>(keep[i].size());
    }
    for (int i = 0; i < N; ++i) owned[i + 1] += owned[i];
    const int F = owned[N];
    std::vector<int> owner(F), neighbour(F), degree(N + 1, 0);
    std::vector<double> area(static_cast<size_t>(F) * D), fc(static_cast<size_t>(F) * D);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < N; ++i) {
        int f = owned[i];
        for (int k : keep[i]) {
            const FaceOut& face = cells[i].faces[k];
            const int j = face.j;
            double d[3] = {0, 0, 0}, d2 = 0;
            for (int a = 0; a < D; ++a) {
                d[a] = wrap(x[static_cast<size_t>(j) * D + a] - x[static_cast<size_t>(i) * D + a], side[a]);
                d2 += d[a] * d[a];
            }
            const double inv = 1.0 / std::sqrt(d2);
            owner[f] = i;
            neighbour[f] = j;
            for (int a = 0; a < D; ++a) {
                area[static_cast<size_t>(f) * D + a] = face.area * (d[a] * inv);
                fc[static_cast<size_t>(f) * D + a] = face.fc[a];
            }
            ++f;
        }
    }
    for (int f = 0; f < F; ++f) {
        ++degree[owner[f] + 1];
        ++degree[neighbour[f] + 1];
    }
    for (int i = 0; i < N; ++i) degree[i + 1] += degree[i];
    std::vector<int> list(degree[N]);
    {
        std::vector<int> fill(degree.begin(), degree.end() - 1);
        for (int f = 0; f < F; ++f) {  // ascending f: each cell's list comes out in ascending face order
            list[fill[owner[f]]++] = f;
            list[fill[neighbour[f]]++] = ~f;
        }
    }
    std::vector<double> volume(N), centroid(static_cast<size_t>(N) * D);
    for (int i = 0; i < N; ++i) {
        volume[i] = cells[i].volume;
        for (int a = 0; a < D; ++a) centroid[static_cast<size_t>(i) * D + a] = cells[i].centroid[a];
    }
    out->dim = D;
    out->cells = N;
    out->faces = F;
    out->owner = copy_out(owner);
    out->neighbour = copy_out(neighbour);
    out->start = copy_out(degree);
    out->list = copy_out(list);
    out->area = copy_out(area);
    out->face_centroid = copy_out(fc);
    out->centroid = copy_out(centroid);
    out->volume = copy_out(volume);
    out->positions = copy_out(x);
    for (int a = 0; a < 3; ++a) {
        out->side[a] = a < D ? side[a] : 0.0;
        out->period[a] = a < D ? period[a] : 0;
    }
    out->step = step;
    out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// ---------------------------------------------------------------------------------------------------------------------
// The geometry's derivative in the generator positions at fixed topology (the role of voronoi-cpp's GeometryDerivative,
// mesh_derivative.hpp:21-46 of the reference): the derivative of the same formulas as above with every vertex the meeting
// point of its bisectors. A 3D vertex v of cell i (relative to x_i) on the planes of neighbours a, b, c solves
// d_a . v = |d_a|^2 / 2 (d_a = x_a - x_i, nearest image) for all three, so d_a . dv = dd_a . (d_a - v); a 2D vertex the
// same with two. Faces: tri_q = (w_q - w_0) x (w_q+1 - w_0) . n / 2 (n = d / |d|), |A| = sigma sum tri_q,
// x_f = sigma sum tri_q (w_0 + w_q + w_q+1) / 3 / |A|, A = |A| n; 2D: |A| = |w_1 - w_0|, x_f the midpoint. Cells:
// V = sum |A| |d| / 6 (2D: sum |w_0 x w_1| / 2), c = sum 3/4 (2D: 2/3) pyramid x_f / V. The face rows are the owner's.

namespace {

using V3 = std::array<double, 3>;
inline V3 add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 mul(V3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot3(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 a, V3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
inline double cross2(V3 a, V3 b) { return a[0] * b[1] - a[1] * b[0]; }

// Solves M y = r for the D x D matrix of rows m[0..D-1] (Gaussian elimination with partial pivoting); transpose: M^T y = r.
void solve(int D, const V3* m, const double* r, double* y, bool transpose) {
    double a[3][4];
    for (int p = 0; p < D; ++p) {
        for (int q = 0; q < D; ++q) a[p][q] = transpose ? m[q][p] : m[p][q];
        a[p][3] = r[p];
    }
    for (int c = 0; c < D; ++c) {
        int piv = c;
        for (int p = c + 1; p < D; ++p)
            if (std::fabs(a[p][c]) > std::fabs(a[piv][c])) piv = p;
        if (piv != c)
            for (int q = 0; q < 4; ++q) std::swap(a[c][q], a[piv][q]);
        for (int p = c + 1; p < D; ++p) {
            const double f = a[p][c] / a[c][c];
            for (int q = c; q < 4; ++q) a[p][q] -= f * a[c][q];
        }
    }
    for (int c = D - 1; c >= 0; --c) {
        double s = a[c][3];
        for (int q = c + 1; q < D; ++q) s -= a[c][q] * y[q];
        y[c] = s / a[c][c];
    }
}

}  // namespace

struct luna_mesh_derivative {
    int D = 3, N = 0;
    double side[3] = {0, 0, 0};
    std::vector<double> x;                 // snapped positions
    std::vector<std::vector<FaceTopo>> topo;
    std::vector<std::vector<int>> face_id;  // per cell and topo face: the mesh's face row it gives (owner's kept faces), or -1

    V3 sep(int i, int j) const {
        V3 d{0, 0, 0};
        for (int a = 0; a < D; ++a) d[a] = wrap(x[static_cast<size_t>(j) * D + a] - x[static_cast<size_t>(i) * D + a], side[a]);
        return d;
    }
    V3 row(const double* p, int i) const {
        V3 r{0, 0, 0};
        for (int a = 0; a < D; ++a) r[a] = p[static_cast<size_t>(i) * D + a];
        return r;
    }
    // The vertex of cell i on the bisectors of js[0..D-1] (relative to x_i), with the rows it solved.
    V3 vertex(int i, const int* js, V3* m) const {
        double r[3];
        for (int p = 0; p < D; ++p) {
            m[p] = sep(i, js[p]);
            r[p] = 0.5 * dot3(m[p], m[p]);
        }
        double y[3] = {0, 0, 0};
        solve(D, m, r, y, false);
        return {y[0], y[1], D == 3 ? y[2] : 0.0};
    }
    // dv from dx (tangent).
    V3 vertex_tangent(int i, const int* js, const V3* m, V3 v, const double* dx) const {
        double s[3];
        for (int p = 0; p < D; ++p) s[p] = dot3(sub(row(dx, js[p]), row(dx, i)), sub(m[p], v));
        double y[3] = {0, 0, 0};
        solve(D, m, s, y, false);
        return {y[0], y[1], D == 3 ? y[2] : 0.0};
    }
    // vbar -> the bars of x_i and the js (adjoint), into acc.
    void vertex_adjoint(int i, const int* js, const V3* m, V3 v, V3 vbar, std::vector<std::pair<int, V3>>& acc) const {
        const double r[3] = {vbar[0], vbar[1], vbar[2]};
        double sbar[3] = {0, 0, 0};
        solve(D, m, r, sbar, true);
        for (int p = 0; p < D; ++p) {
            const V3 g = mul(sub(m[p], v), sbar[p]);
            acc.push_back({js[p], g});
            acc.push_back({i, mul(g, -1.0)});
        }
    }
};

extern "C" int luna_mesh_derivative_create(const luna_mesh_record* mesh, luna_mesh_derivative** out) {
    if (!mesh || !out) {
        g_error = "luna_mesh_derivative_create: null argument";
        return 1;
    }
    *out = nullptr;
    auto* g = new luna_mesh_derivative;
    g->D = mesh->dim;
    g->N = mesh->cells;
    for (int a = 0; a < 3; ++a) g->side[a] = mesh->side[a];
    g->x.assign(mesh->positions, mesh->positions + static_cast<size_t>(g->N) * g->D);
    std::vector<CellOut> cells(g->N);
    if (!build_cells(g->D, g->N, g->x.data(), g->side, true, cells)) {
        delete g;
        return 2;
    }
    g->topo.resize(g->N);
    g->face_id.resize(g->N);
    for (int i = 0; i < g->N; ++i) {
        g->topo[i] = std::move(cells[i].topo);
        g->face_id[i].assign(g->topo[i].size(), -1);
        for (size_t t = 0; t < g->topo[i].size(); ++t) {
            const int j = g->topo[i][t].j;
            if (j <= i) continue;
            for (int k = mesh->start[i]; k < mesh->start[i + 1]; ++k) {
                const int e = mesh->list[k];
                if (e >= 0 && mesh->neighbour[e] == j) g->face_id[i][t] = e;
            }
        }
    }
    *out = g;
    return 0;
}

extern "C" void luna_mesh_derivative_destroy(luna_mesh_derivative* g) { delete g; }

extern "C" int luna_mesh_derivative_tangent(const luna_mesh_derivative* g, const double* dx, double* dV, double* dc, double* dA, double* dxf) {
    if (!g || !dx || !dV || !dc || !dA || !dxf) {
        g_error = "luna_mesh_derivative_tangent: null argument";
        return 1;
    }
    const int D = g->D;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < g->N; ++i) {
        double V = 0, dVi = 0;
        V3 moment{0, 0, 0}, dmoment{0, 0, 0};
        const auto& faces = g->topo[i];
        for (size_t t = 0; t < faces.size(); ++t) {
            const FaceTopo& F = faces[t];
            const V3 d = g->sep(i, F.j), dd = sub(g->row(dx, F.j), g->row(dx, i));
            const double dist = std::sqrt(dot3(d, d)), ddist = dot3(d, dd) / dist;
            const V3 n = mul(d, 1.0 / dist), dn = mul(sub(dd, mul(n, dot3(n, dd))), 1.0 / dist);
            const int m = static_cast<int>(F.va.size());
            std::vector<V3> w(m), dw(m);
            for (int q = 0; q < m; ++q) {
                const int js[3] = {F.j, F.va[q], D == 3 ? F.vb[q] : 0};
                V3 rows[3];
                w[q] = g->vertex(i, js, rows);
                dw[q] = g->vertex_tangent(i, js, rows, w[q], dx);
            }
            double area = 0, darea = 0, pyr, dpyr;
            V3 fc, dfc;
            if (D == 3) {
                V3 cnum{0, 0, 0}, dcnum{0, 0, 0};
                for (int q = 1; q + 1 < m; ++q) {
                    const V3 a = sub(w[q], w[0]), b = sub(w[q + 1], w[0]), da = sub(dw[q], dw[0]), db = sub(dw[q + 1], dw[0]);
                    const V3 cr = cross(a, b), dcr = add(cross(da, b), cross(a, db));
                    const double tri = 0.5 * dot3(cr, n), dtri = 0.5 * (dot3(dcr, n) + dot3(cr, dn));
                    const V3 gq = mul(add(add(w[0], w[q]), w[q + 1]), 1.0 / 3.0), dgq = mul(add(add(dw[0], dw[q]), dw[q + 1]), 1.0 / 3.0);
                    area += tri;
                    darea += dtri;
                    cnum = add(cnum, mul(gq, tri));
                    dcnum = add(dcnum, add(mul(gq, dtri), mul(dgq, tri)));
                }
                area *= F.sigma;
                darea *= F.sigma;
                cnum = mul(cnum, F.sigma);
                dcnum = mul(dcnum, F.sigma);
                fc = mul(cnum, 1.0 / area);
                dfc = mul(sub(dcnum, mul(fc, darea)), 1.0 / area);
                pyr = area * dist / 6.0;
                dpyr = (darea * dist + area * ddist) / 6.0;
                moment = add(moment, mul(fc, 0.75 * pyr));
                dmoment = add(dmoment, mul(add(mul(dfc, pyr), mul(fc, dpyr)), 0.75));
            } else {
                const V3 e = sub(w[1], w[0]), de = sub(dw[1], dw[0]);
                area = std::sqrt(dot3(e, e));
                darea = dot3(e, de) / area;
                fc = mul(add(w[0], w[1]), 0.5);
                dfc = mul(add(dw[0], dw[1]), 0.5);
                pyr = 0.5 * F.sigma * cross2(w[0], w[1]);
                dpyr = 0.5 * F.sigma * (cross2(dw[0], w[1]) + cross2(w[0], dw[1]));
                moment = add(moment, mul(fc, (2.0 / 3.0) * pyr));
                dmoment = add(dmoment, mul(add(mul(dfc, pyr), mul(fc, dpyr)), 2.0 / 3.0));
            }
            V += pyr;
            dVi += dpyr;
            const int f = g->face_id[i][t];
            if (f >= 0) {
                const V3 dAv = add(mul(n, darea), mul(dn, area));
                for (int a = 0; a < D; ++a) {
                    dA[static_cast<size_t>(f) * D + a] = dAv[a];
                    dxf[static_cast<size_t>(f) * D + a] = dfc[a];
                }
            }
        }
        dV[i] = dVi;
        const V3 c = mul(moment, 1.0 / V);
        const V3 dci = mul(sub(dmoment, mul(c, dVi)), 1.0 / V);
        for (int a = 0; a < D; ++a) dc[static_cast<size_t>(i) * D + a] = dci[a];
    }
    return 0;
}

extern "C" int luna_mesh_derivative_adjoint(const luna_mesh_derivative* g, const double* Vbar, const double* cbar, const double* Abar,
                                            const double* xfbar, double* xbar) {
    if (!g || !Vbar || !cbar || !Abar || !xfbar || !xbar) {
        g_error = "luna_mesh_derivative_adjoint: null argument";
        return 1;
    }
    const int D = g->D, N = g->N;
    std::vector<std::vector<std::pair<int, V3>>> acc(N);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < N; ++i) {
        auto& out = acc[i];
        const auto& faces = g->topo[i];
        // Forward: V, moment (the cell's own faces), then c = moment / V.
        double V = 0;
        V3 moment{0, 0, 0};
        struct Face {
            V3 d, n, fc;
            double dist, area, pyr;
            std::vector<V3> w, rows;
        };
        std::vector<Face> fs(faces.size());
        for (size_t t = 0; t < faces.size(); ++t) {
            const FaceTopo& F = faces[t];
            Face& s = fs[t];
            s.d = g->sep(i, F.j);
            s.dist = std::sqrt(dot3(s.d, s.d));
            s.n = mul(s.d, 1.0 / s.dist);
            const int m = static_cast<int>(F.va.size());
            s.w.resize(m);
            s.rows.resize(static_cast<size_t>(m) * 3);
            for (int q = 0; q < m; ++q) {
                const int js[3] = {F.j, F.va[q], D == 3 ? F.vb[q] : 0};
                s.w[q] = g->vertex(i, js, &s.rows[static_cast<size_t>(q) * 3]);
            }
            if (D == 3) {
                double area = 0;
                V3 cnum{0, 0, 0};
                for (int q = 1; q + 1 < m; ++q) {
                    const double tri = 0.5 * dot3(cross(sub(s.w[q], s.w[0]), sub(s.w[q + 1], s.w[0])), s.n);
                    area += tri;
                    cnum = add(cnum, mul(add(add(s.w[0], s.w[q]), s.w[q + 1]), tri / 3.0));
                }
                s.area = area * F.sigma;
                s.fc = mul(cnum, F.sigma / s.area);
                s.pyr = s.area * s.dist / 6.0;
                moment = add(moment, mul(s.fc, 0.75 * s.pyr));
            } else {
                const V3 e = sub(s.w[1], s.w[0]);
                s.area = std::sqrt(dot3(e, e));
                s.fc = mul(add(s.w[0], s.w[1]), 0.5);
                s.pyr = 0.5 * F.sigma * cross2(s.w[0], s.w[1]);
                moment = add(moment, mul(s.fc, (2.0 / 3.0) * s.pyr));
            }
            V += s.pyr;
        }
        const V3 c = mul(moment, 1.0 / V);
        // Reverse: c = moment / V.
        const V3 cb = g->row(cbar, i);
        const V3 momentbar = mul(cb, 1.0 / V);
        const double Vb = Vbar[i] - dot3(cb, c) / V;
        const double share = D == 3 ? 0.75 : 2.0 / 3.0;
        for (size_t t = 0; t < faces.size(); ++t) {
            const FaceTopo& F = faces[t];
            const Face& s = fs[t];
            const int m = static_cast<int>(s.w.size());
            double pyrbar = Vb + share * dot3(momentbar, s.fc);
            V3 fcbar = mul(momentbar, share * s.pyr);
            double areabar = 0, distbar = 0;
            V3 nbar{0, 0, 0};
            const int f = g->face_id[i][t];
            if (f >= 0) {
                const V3 Ab = g->row(Abar, f);
                areabar += dot3(Ab, s.n);
                nbar = add(nbar, mul(Ab, s.area));
                fcbar = add(fcbar, g->row(xfbar, f));
            }
            std::vector<V3> wbar(m, V3{0, 0, 0});
            if (D == 3) {
                areabar += pyrbar * s.dist / 6.0;
                distbar += pyrbar * s.area / 6.0;
                // fc = cnum / area
                const V3 cnumbar = mul(fcbar, 1.0 / s.area);
                areabar -= dot3(fcbar, s.fc) / s.area;
                for (int q = 1; q + 1 < m; ++q) {
                    const V3 a = sub(s.w[q], s.w[0]), b = sub(s.w[q + 1], s.w[0]);
                    const V3 cr = cross(a, b);
                    const double tri = 0.5 * dot3(cr, s.n);
                    const V3 gq = mul(add(add(s.w[0], s.w[q]), s.w[q + 1]), 1.0 / 3.0);
                    // area = sigma sum tri, cnum = sigma sum tri gq
                    const double tribar = F.sigma * (areabar + dot3(cnumbar, gq));
                    const V3 gbar = mul(cnumbar, F.sigma * tri / 3.0);
                    wbar[0] = add(wbar[0], gbar);
                    wbar[q] = add(wbar[q], gbar);
                    wbar[q + 1] = add(wbar[q + 1], gbar);
                    // tri = cr . n / 2
                    const V3 crbar = mul(s.n, 0.5 * tribar);
                    nbar = add(nbar, mul(cr, 0.5 * tribar));
                    const V3 abar = cross(b, crbar), bbar = cross(crbar, a);
                    wbar[q] = add(wbar[q], abar);
                    wbar[q + 1] = add(wbar[q + 1], bbar);
                    wbar[0] = sub(wbar[0], add(abar, bbar));
                }
            } else {
                // pyr = sigma w0 x w1 / 2, fc = (w0 + w1) / 2, area = |w1 - w0|
                const double pb = 0.5 * F.sigma * pyrbar;
                wbar[0] = add(wbar[0], V3{pb * s.w[1][1], -pb * s.w[1][0], 0});
                wbar[1] = add(wbar[1], V3{-pb * s.w[0][1], pb * s.w[0][0], 0});
                wbar[0] = add(wbar[0], mul(fcbar, 0.5));
                wbar[1] = add(wbar[1], mul(fcbar, 0.5));
                const V3 e = sub(s.w[1], s.w[0]);
                const V3 eb = mul(e, areabar / s.area);
                wbar[1] = add(wbar[1], eb);
                wbar[0] = sub(wbar[0], eb);
            }
            // n = d / |d|, |d|
            V3 dbar = add(mul(sub(nbar, mul(s.n, dot3(s.n, nbar))), 1.0 / s.dist), mul(s.n, distbar));
            out.push_back({F.j, dbar});
            out.push_back({i, mul(dbar, -1.0)});
            for (int q = 0; q < m; ++q) {
                const int js[3] = {F.j, F.va[q], D == 3 ? F.vb[q] : 0};
                g->vertex_adjoint(i, js, &s.rows[static_cast<size_t>(q) * 3], s.w[q], wbar[q], out);
            }
        }
    }
    for (int i = 0; i < N; ++i)  // in cell order: the same sums every run
        for (const auto& p : acc[i])
            for (int a = 0; a < D; ++a) xbar[static_cast<size_t>(p.first) * D + a] += p.second[a];
    return 0;
}
