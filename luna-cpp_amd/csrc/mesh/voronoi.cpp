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

struct CellOut {
    double volume = 0, centroid[3] = {0, 0, 0};
    std::vector<FaceOut> faces;
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
    double volume = 0, moment[3] = {0, 0, 0};
    static thread_local std::vector<int> seen;
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
            do {
                int r = 0;
                while (P.t[cur][r] != k) ++r;
                const int y = P.t[cur][(r + 2) % 3];
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
            if (area < 0) {
                area = -area;
                for (double& x : c) x = -x;
            }
            if (!(area > 0)) continue;  // a face of no extent: no face
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
    Bins bins;
    bins.build(D, N, x.data(), side);
    std::vector<CellOut> cells(N);
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
                const double complete = gather(bins, D, x.data(), side, i, r, cands);
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
        return 2;
    }
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
