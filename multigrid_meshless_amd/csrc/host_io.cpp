// host_io.cpp -- configuration, Gmsh v2 I/O and the seeded jittered-cloud
// generator.  Host code only: none of it needs a GPU.
//
// The reader keeps the reference's semantics (hf/mesh_functions.h:172-476):
// boundary tags from line elements (2D) or triangles/quads (3D), corner tags
// from type-15 points (and lines in 3D), area-weighted normals accumulated per
// element, and removal of the first num_corners nodes.  It parses with a
// hand-written tokenizer instead of fscanf, which is what makes 1M-node files
// readable in seconds.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mgm_internal.h"

namespace mgm {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
const char *last_error() { return g_err.c_str(); }

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

// Monomial exponent table, x outer, y, z inner (hf/mesh_functions.h:131-170).
int poly_terms(int dim, int deg, short *px, short *py, short *pz)
{
    int c = 0;
    if (dim == 2) {
        for (int i = 0; i <= deg; i++)
            for (int j = 0; j <= deg - i; j++) {
                if (px) { px[c] = (short)i; py[c] = (short)j; pz[c] = 0; }
                c++;
            }
    } else {
        for (int i = 0; i <= deg; i++)
            for (int j = 0; j <= deg - i; j++)
                for (int k = 0; k <= deg - i - j; k++) {
                    if (px) { px[c] = (short)i; py[c] = (short)j; pz[c] = (short)k; }
                    c++;
                }
    }
    return c;
}

// ---------------------------------------------------------------------------
// Tokenizer over a whole file held in memory.
namespace {
struct Tok {
    std::vector<char> buf;
    size_t pos = 0;
    bool load(const char *path)
    {
        FILE *f = fopen(path, "rb");
        if (!f) return false;
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf.resize((size_t)n + 1);
        size_t got = fread(buf.data(), 1, (size_t)n, f);
        fclose(f);
        buf[got] = 0;
        buf.resize(got + 1);
        return true;
    }
    void skipws() { while (buf[pos] && (unsigned char)buf[pos] <= ' ') pos++; }
    bool word(std::string &w)
    {
        skipws();
        if (!buf[pos]) return false;
        size_t s = pos;
        while (buf[pos] && (unsigned char)buf[pos] > ' ') pos++;
        w.assign(&buf[s], pos - s);
        return true;
    }
    bool integer(long &v)
    {
        skipws();
        if (!buf[pos]) return false;
        char *e;
        v = strtol(&buf[pos], &e, 0);  // %i semantics
        if (e == &buf[pos]) return false;
        pos = (size_t)(e - buf.data());
        return true;
    }
    bool real(double &v)
    {
        skipws();
        if (!buf[pos]) return false;
        char *e;
        v = strtod(&buf[pos], &e);
        if (e == &buf[pos]) return false;
        pos = (size_t)(e - buf.data());
        return true;
    }
    void skipline()
    {
        while (buf[pos] && buf[pos] != '\n') pos++;
        if (buf[pos]) pos++;
    }
};
}  // namespace

int read_gmsh(const char *path, int dim, PointSet &o)
{
    Tok t;
    if (!t.load(path)) MGM_FAIL(-2, std::string("cannot open mesh file ") + path);
    std::string w;
    for (;;) {
        if (!t.word(w)) MGM_FAIL(-3, std::string("no $Nodes section in ") + path);
        if (w == "$Nodes") break;
    }
    long n;
    if (!t.integer(n) || n <= 0) MGM_FAIL(-3, "bad node count");
    std::vector<double> x(n), y(n), z(n), nx(n, 0.0), ny(n, 0.0), nz(n, 0.0);
    std::vector<uint8_t> bt(n, 0), ct(n, 0);
    double xmin = 1e6, ymin = 1e6, zmin = 1e6, xmax = -1e6, ymax = -1e6, zmax = -1e6;
    for (long i = 0; i < n; i++) {
        long id;
        if (!t.integer(id) || !t.real(x[i]) || !t.real(y[i]) || !t.real(z[i]))
            MGM_FAIL(-3, "truncated $Nodes section");
        if (x[i] < xmin) xmin = x[i];
        if (y[i] < ymin) ymin = y[i];
        if (z[i] < zmin) zmin = z[i];
        if (x[i] > xmax) xmax = x[i];
        if (y[i] > ymax) ymax = y[i];
        if (z[i] > zmax) zmax = z[i];
    }
    for (;;) {
        if (!t.word(w)) MGM_FAIL(-3, "no $Elements section");
        if (w == "$Elements") break;
    }
    long ne;
    if (!t.integer(ne) || ne < 0) MGM_FAIL(-3, "bad element count");
    o.num_elem = (int)ne;
    if (dim == 2)
        o.d_avg = sqrt((xmax - xmin) * (ymax - ymin) / (int)ne);
    else
        o.d_avg = cbrt((xmax - xmin) * (ymax - ymin) * (zmax - zmin) / (int)ne);
    auto node = [&](long v) -> long {
        if (v < 1 || v > n) return -1;
        return v - 1;
    };
    for (long ie = 0; ie < ne; ie++) {
        long id, et, ntag, t1, t2;
        if (!t.integer(id) || !t.integer(et)) MGM_FAIL(-3, "truncated $Elements section");
        bool bnd2 = (dim == 2 && et == 1), bnd3 = (dim == 3 && (et == 2 || et == 3));
        bool corner_line = (dim == 3 && et == 1), corner_pt = (et == 15);
        if (!(bnd2 || bnd3 || corner_line || corner_pt)) { t.skipline(); continue; }
        // two tags assumed, as the reference reads them (:295-297)
        if (!t.integer(ntag) || !t.integer(t1) || !t.integer(t2)) MGM_FAIL(-3, "bad element tags");
        if (corner_pt) {
            long a;
            if (!t.integer(a) || node(a) < 0) MGM_FAIL(-3, "bad point element");
            ct[a - 1] = 1;
        } else if (corner_line) {
            long a, b;
            if (!t.integer(a) || !t.integer(b) || node(a) < 0 || node(b) < 0) MGM_FAIL(-3, "bad line element");
            ct[a - 1] = 1; ct[b - 1] = 1;
        } else if (bnd2) {
            long a, b;
            if (!t.integer(a) || !t.integer(b) || node(a) < 0 || node(b) < 0) MGM_FAIL(-3, "bad line element");
            a--; b--;
            bt[a] = 1; bt[b] = 1;
            double dx = x[b] - x[a], dy = y[b] - y[a];
            nx[a] += -dy; ny[a] += dx;
            nx[b] += -dy; ny[b] += dx;
        } else {
            int nn = (et == 2) ? 3 : 4;
            long nd[4];
            for (int q = 0; q < nn; q++) {
                if (!t.integer(nd[q]) || node(nd[q]) < 0) MGM_FAIL(-3, "bad surface element");
                nd[q]--;
                bt[nd[q]] = 1;
            }
            long a = nd[0], b = nd[1], c = nd[2];
            double dx = x[b] - x[a], dy = y[b] - y[a], dz = z[b] - z[a];
            double dx1 = x[c] - x[a], dy1 = y[c] - y[a], dz1 = z[c] - z[a];
            double v1 = dy * dz1 - dz * dy1, v2 = dz * dx1 - dx * dz1, v3 = dx * dy1 - dy * dx1;
            for (int q = 0; q < nn; q++) nx[nd[q]] += v1;
            for (int q = 0; q < nn; q++) ny[nd[q]] += v2;
            for (int q = 0; q < nn; q++) nz[nd[q]] += v3;
        }
    }
    long nbt = 0, nc = 0;
    for (long i = 0; i < n; i++) { nbt += bt[i]; nc += ct[i]; }
    long N = n - nc;
    if (N <= 0) MGM_FAIL(-3, "mesh has no nodes after corner removal");
    auto cut = [&](auto &v) { v.erase(v.begin(), v.begin() + nc); };
    cut(x); cut(y); cut(z); cut(nx); cut(ny); cut(nz); cut(bt); cut(ct);
    o.N = (int)N;
    o.nb = (int)(nbt - nc);
    o.num_corners_removed = (int)nc;
    o.x.swap(x); o.y.swap(y); o.z.swap(z); o.nx.swap(nx); o.ny.swap(ny); o.nz.swap(nz);
    o.btag.swap(bt); o.ctag.swap(ct);
    return 0;
}

// ---------------------------------------------------------------------------
// Jittered Cartesian cloud on the unit square/cube, n points per axis,
// x_i = i/(n-1) + U(-jitter h, +jitter h).  Interior nodes jitter in every
// axis, boundary nodes only tangentially, corners (and 3D edges) not at all.
// Random draws come from SplitMix64(seed) in grid order (i fastest), then the
// boundary group and the interior group are each shuffled (Fisher-Yates with
// the same stream) so that an insertion-ordered kd-tree over file order stays
// shallow.  Node order written: corners, [3D edge nodes], boundary, interior,
// which is the order the reader's corner removal and the kNN's boundary-first
// indexing assume.
//
// Elements written (two tags each): type 15 per corner; boundary lines
// (2D: type 1 along each side; 3D: type 1 along each edge, which the reader
// tags as corners); 3D: two type-2 triangles per face cell with one winding
// per face; interior: two triangles per cell (2D) or, if requested, six
// tetrahedra per cell (3D) so that the bbox/num_elem spacing estimate is h.
int generate_cloud(int dim, int n, uint64_t seed, double jitter, int with_volume_elements,
                   PointSet &P, const char *gmsh_path)
{
    if (n < 3) MGM_FAIL(-1, "generate_cloud: need at least 3 points per axis");
    if (dim != 2 && dim != 3) MGM_FAIL(-1, "generate_cloud: dimension must be 2 or 3");
    if (!(jitter >= 0 && jitter < 0.5)) MGM_FAIL(-1, "generate_cloud: jitter must be in [0, 0.5)");
    const double h = 1.0 / (n - 1);
    uint64_t s = seed;
    const int nz_ = (dim == 3) ? n : 1;
    const size_t tot = (size_t)n * n * nz_;
    std::vector<double> X(tot), Y(tot), Z(tot);
    std::vector<int8_t> cls(tot);  // 0 corner, 1 edge(3D), 2 boundary, 3 interior
    auto at = [&](int i, int j, int k) { return ((size_t)k * n + j) * n + i; };
    for (int k = 0; k < nz_; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int ex = (i == 0 || i == n - 1), ey = (j == 0 || j == n - 1);
                int ez = (dim == 3) && (k == 0 || k == n - 1);
                int nbd = ex + ey + ez;
                size_t q = at(i, j, k);
                double px = i * h, py = j * h, pz = (dim == 3) ? k * h : 0.0;
                if (i == n - 1) px = 1.0;
                if (j == n - 1) py = 1.0;
                if (dim == 3 && k == n - 1) pz = 1.0;
                int c;
                if (dim == 2) c = (nbd == 2) ? 0 : (nbd == 1 ? 2 : 3);
                else c = (nbd == 3) ? 0 : (nbd == 2 ? 1 : (nbd == 1 ? 2 : 3));
                cls[q] = (int8_t)c;
                if (c >= 2) {
                    if (!ex) px += (2 * uniform01(s) - 1) * jitter * h;
                    if (!ey) py += (2 * uniform01(s) - 1) * jitter * h;
                    if (dim == 3 && !ez) pz += (2 * uniform01(s) - 1) * jitter * h;
                }
                X[q] = px; Y[q] = py; Z[q] = pz;
            }
    std::vector<size_t> grp[4];
    for (size_t q = 0; q < tot; q++) grp[cls[q]].push_back(q);
    for (int g = 2; g <= 3; g++) {
        auto &v = grp[g];
        for (size_t i = v.size(); i > 1; i--) {
            size_t j = (size_t)(splitmix64(s) % i);
            std::swap(v[i - 1], v[j]);
        }
    }
    std::vector<int> fid(tot);  // 1-based file id per grid point
    int next = 1;
    for (int g = 0; g < 4; g++)
        for (size_t q : grp[g]) fid[q] = next++;
    const int ncorner_removed = (int)(grp[0].size() + (dim == 3 ? grp[1].size() : 0));

    // In-memory reader result (what read_gmsh returns for the written file).
    const int nfile = next - 1;
    std::vector<double> x(nfile), y(nfile), z(nfile), nx(nfile, 0.0), ny(nfile, 0.0), nz(nfile, 0.0);
    std::vector<uint8_t> bt(nfile, 0), ct(nfile, 0);
    for (size_t q = 0; q < tot; q++) { int f = fid[q] - 1; x[f] = X[q]; y[f] = Y[q]; z[f] = Z[q]; }

    FILE *fo = nullptr;
    if (gmsh_path) {
        fo = fopen(gmsh_path, "w");
        if (!fo) MGM_FAIL(-2, std::string("cannot write ") + gmsh_path);
        fprintf(fo, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", nfile);
        for (int f = 0; f < nfile; f++) fprintf(fo, "%d %.17g %.17g %.17g\n", f + 1, x[f], y[f], z[f]);
        fprintf(fo, "$EndNodes\n");
    }
    // Element list, built in memory then counted/written.
    struct El { int type; int nn; int v[4]; };
    std::vector<El> els;
    for (size_t q : grp[0]) els.push_back({15, 1, {fid[q], 0, 0, 0}});
    if (dim == 2) {
        // sides traversed counter-clockwise so all normals of a side agree
        for (int i = 0; i < n - 1; i++) els.push_back({1, 2, {fid[at(i, 0, 0)], fid[at(i + 1, 0, 0)], 0, 0}});
        for (int j = 0; j < n - 1; j++) els.push_back({1, 2, {fid[at(n - 1, j, 0)], fid[at(n - 1, j + 1, 0)], 0, 0}});
        for (int i = n - 1; i > 0; i--) els.push_back({1, 2, {fid[at(i, n - 1, 0)], fid[at(i - 1, n - 1, 0)], 0, 0}});
        for (int j = n - 1; j > 0; j--) els.push_back({1, 2, {fid[at(0, j, 0)], fid[at(0, j - 1, 0)], 0, 0}});
        if (with_volume_elements)
            for (int j = 0; j < n - 1; j++)
                for (int i = 0; i < n - 1; i++) {
                    int a = fid[at(i, j, 0)], b = fid[at(i + 1, j, 0)], c = fid[at(i + 1, j + 1, 0)], d = fid[at(i, j + 1, 0)];
                    els.push_back({2, 3, {a, b, c, 0}});
                    els.push_back({2, 3, {a, c, d, 0}});
                }
    } else {
        // 12 edges as line chains (tagged corner by the reader)
        auto edge = [&](int ax, int c1, int c2) {
            for (int t = 0; t < n - 1; t++) {
                int i0[3], i1[3];
                int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                i0[ax] = t; i1[ax] = t + 1; i0[o1] = i1[o1] = c1; i0[o2] = i1[o2] = c2;
                els.push_back({1, 2, {fid[at(i0[0], i0[1], i0[2])], fid[at(i1[0], i1[1], i1[2])], 0, 0}});
            }
        };
        for (int ax = 0; ax < 3; ax++)
            for (int c1 : {0, n - 1})
                for (int c2 : {0, n - 1}) edge(ax, c1, c2);
        // faces: axis ax fixed at 0 or n-1; two triangles per cell
        for (int ax = 0; ax < 3; ax++)
            for (int side : {0, n - 1}) {
                int u = (ax + 1) % 3, v = (ax + 2) % 3;
                for (int b = 0; b < n - 1; b++)
                    for (int a = 0; a < n - 1; a++) {
                        int c00[3], c10[3], c11[3], c01[3];
                        c00[ax] = c10[ax] = c11[ax] = c01[ax] = side;
                        c00[u] = a; c00[v] = b; c10[u] = a + 1; c10[v] = b;
                        c11[u] = a + 1; c11[v] = b + 1; c01[u] = a; c01[v] = b + 1;
                        int p00 = fid[at(c00[0], c00[1], c00[2])], p10 = fid[at(c10[0], c10[1], c10[2])];
                        int p11 = fid[at(c11[0], c11[1], c11[2])], p01 = fid[at(c01[0], c01[1], c01[2])];
                        els.push_back({2, 3, {p00, p10, p11, 0}});
                        els.push_back({2, 3, {p00, p11, p01, 0}});
                    }
            }
        if (with_volume_elements)
            for (int k = 0; k < n - 1; k++)
                for (int j = 0; j < n - 1; j++)
                    for (int i = 0; i < n - 1; i++) {
                        int v[8];
                        for (int c = 0; c < 8; c++) v[c] = fid[at(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2))];
                        static const int T6[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                                     {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
                        for (auto &tt : T6) els.push_back({4, 4, {v[tt[0]], v[tt[1]], v[tt[2]], v[tt[3]]}});
                    }
    }
    if (fo) {
        fprintf(fo, "$Elements\n%zu\n", els.size());
        size_t eid = 1;
        for (auto &e : els) {
            fprintf(fo, "%zu %d 2 0 1", eid++, e.type);
            for (int q = 0; q < e.nn; q++) fprintf(fo, " %d", e.v[q]);
            fprintf(fo, "\n");
        }
        fprintf(fo, "$EndElements\n");
        if (fclose(fo) != 0) MGM_FAIL(-2, "error writing mesh file");
    }
    // Apply the reader's tagging to the in-memory copy (same rules as read_gmsh).
    for (auto &e : els) {
        if (e.type == 15) { ct[e.v[0] - 1] = 1; continue; }
        if (dim == 3 && e.type == 1) { ct[e.v[0] - 1] = 1; ct[e.v[1] - 1] = 1; continue; }
        if (dim == 2 && e.type == 1) {
            int a = e.v[0] - 1, b = e.v[1] - 1;
            bt[a] = bt[b] = 1;
            double dx = x[b] - x[a], dy = y[b] - y[a];
            nx[a] += -dy; ny[a] += dx; nx[b] += -dy; ny[b] += dx;
            continue;
        }
        if (dim == 3 && e.type == 2) {
            int a = e.v[0] - 1, b = e.v[1] - 1, c = e.v[2] - 1;
            bt[a] = bt[b] = bt[c] = 1;
            double dx = x[b] - x[a], dy = y[b] - y[a], dz = z[b] - z[a];
            double dx1 = x[c] - x[a], dy1 = y[c] - y[a], dz1 = z[c] - z[a];
            double v1 = dy * dz1 - dz * dy1, v2 = dz * dx1 - dx * dz1, v3 = dx * dy1 - dy * dx1;
            nx[a] += v1; nx[b] += v1; nx[c] += v1;
            ny[a] += v2; ny[b] += v2; ny[c] += v2;
            nz[a] += v3; nz[b] += v3; nz[c] += v3;
        }
    }
    double xmin = 1e6, ymin = 1e6, zmin = 1e6, xmax = -1e6, ymax = -1e6, zmax = -1e6;
    for (int f = 0; f < nfile; f++) {
        xmin = std::min(xmin, x[f]); xmax = std::max(xmax, x[f]);
        ymin = std::min(ymin, y[f]); ymax = std::max(ymax, y[f]);
        zmin = std::min(zmin, z[f]); zmax = std::max(zmax, z[f]);
    }
    const int ne = (int)els.size();
    P.num_elem = ne;
    P.d_avg = (dim == 2) ? sqrt((xmax - xmin) * (ymax - ymin) / ne)
                         : cbrt((xmax - xmin) * (ymax - ymin) * (zmax - zmin) / ne);
    long nbt = 0, nc = 0;
    for (int f = 0; f < nfile; f++) { nbt += bt[f]; nc += ct[f]; }
    if (nc != ncorner_removed) MGM_FAIL(-4, "generate_cloud: corner bookkeeping mismatch");
    auto cut = [&](auto &v) { v.erase(v.begin(), v.begin() + nc); };
    cut(x); cut(y); cut(z); cut(nx); cut(ny); cut(nz); cut(bt); cut(ct);
    P.N = nfile - (int)nc;
    P.nb = (int)(nbt - nc);
    P.num_corners_removed = (int)nc;
    P.x.swap(x); P.y.swap(y); P.z.swap(z); P.nx.swap(nx); P.ny.swap(ny); P.nz.swap(nz);
    P.btag.swap(bt); P.ctag.swap(ct);
    return 0;
}

}  // namespace mgm
