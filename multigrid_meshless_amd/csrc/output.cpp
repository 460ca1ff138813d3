// output.cpp -- the drivers' result files from a resident problem: the
// solution of a level in reference node order (mg_heat_conduction.c:89-93,
// mg_NS_solver.c:84-90) and a per-step convergence history
// (mg_heat_conduction.c:77, mg_NS_solver.c:77-78).  Values are printed with
// the drivers' own formats ("%lf", "%e"); rows are formatted on several host
// threads and written in order.
#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "problem.hpp"

using namespace mgm;

namespace {

// Rows [r0, r1) of the columns, "%lf" joined by ", ", one line per row.
void format_rows(const std::vector<const double *> &cols, int r0, int r1, std::string &out)
{
    char buf[64];
    out.clear();
    out.reserve((size_t)(r1 - r0) * cols.size() * 12);
    for (int r = r0; r < r1; r++) {
        for (size_t c = 0; c < cols.size(); c++) {
            const int n = snprintf(buf, sizeof buf, c ? ", %lf" : "%lf", cols[c][r]);
            out.append(buf, (size_t)n);
        }
        out.push_back('\n');
    }
}

}  // namespace

extern "C" {

int mgm_write_solution(const mgm_problem *pb, int lev, const char *path, int with_z, int nfields,
                       const int *fields)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    if (!pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (lev < 0 || lev >= pb->L) MGM_FAIL(-1, "level out of range");
    if (!path) MGM_FAIL(-1, "null path");
    if (nfields < 0 || nfields > 16 || (nfields && !fields)) MGM_FAIL(-1, "write_solution: 0..16 fields");
    for (int k = 0; k < nfields; k++)
        if (fields[k] < 0 || fields[k] >= MGM_NUM_FIELDS) MGM_FAIL(-1, "write_solution: unknown field");
    const Level &L = pb->lv[lev];
    const int N = L.N;
    std::vector<std::vector<double>> vals((size_t)nfields, std::vector<double>(N));
    for (int k = 0; k < nfields; k++) {
        int rc = mgm_get_field(pb, lev, fields[k], vals[k].data());
        if (rc) return rc;
    }
    std::vector<const double *> cols{L.x.data(), L.y.data()};
    if (with_z) cols.push_back(L.z.data());
    for (auto &v : vals) cols.push_back(v.data());

    FILE *f = fopen(path, "w");
    if (!f) MGM_FAIL(-2, "write_solution: cannot open the file");
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int chunk = 65536;
    std::vector<std::string> part(nt);
    bool ok = true;
    for (int r0 = 0; r0 < N && ok; r0 += chunk * nt) {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) {
            const int a = std::min(N, r0 + t * chunk), b = std::min(N, a + chunk);
            th.emplace_back(format_rows, std::cref(cols), a, b, std::ref(part[t]));
        }
        for (auto &t : th) t.join();
        for (auto &s : part)
            if (!s.empty() && fwrite(s.data(), 1, s.size(), f) != s.size()) ok = false;
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) MGM_FAIL(-2, "write_solution: write failed");
    return 0;
}

int mgm_write_history(const char *path, int n, const double *values, int scientific)
{
    if (!path) MGM_FAIL(-1, "null path");
    if (n < 0 || (n && !values)) MGM_FAIL(-1, "write_history: bad values");
    FILE *f = fopen(path, "w");
    if (!f) MGM_FAIL(-2, "write_history: cannot open the file");
    bool ok = true;
    for (int i = 0; i < n && ok; i++)
        ok = (scientific ? fprintf(f, "%d, %e\n", i, values[i]) : fprintf(f, "%d, %lf\n", i, values[i])) > 0;
    if (fclose(f) != 0) ok = false;
    if (!ok) MGM_FAIL(-2, "write_history: write failed");
    return 0;
}

}  // extern "C"
