// Sanitizer driver for libmgm's host-only code (csrc/host_io.cpp): the
// jittered-cloud generator and Gmsh writer, the Gmsh reader, the monomial
// table and SplitMix64.  Built with -fsanitize=address,undefined by
// tests/sanitize/Makefile; exits non-zero on any mismatch.
//   host_io_driver <scratch dir>
#include <cstdio>
#include <string>

#include "mgm_internal.h"

using namespace mgm;

static int check(bool ok, const char *what)
{
    if (!ok) fprintf(stderr, "host_io_driver: %s\n", what);
    return ok ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    int bad = 0;
    for (int dim = 2; dim <= 3; dim++)
        for (int vol = 0; vol <= 1; vol++) {
            const std::string path = std::string(argv[1]) + "/c" + std::to_string(dim) + std::to_string(vol) + ".msh";
            PointSet a, b;
            const int n = dim == 2 ? 23 : 9;
            bad += check(generate_cloud(dim, n, 7 + dim, 0.2, vol, a, path.c_str()) == 0, "generate");
            bad += check(read_gmsh(path.c_str(), dim, b) == 0, "read back");
            bad += check(a.N == b.N && a.nb == b.nb && a.x == b.x && a.y == b.y && a.z == b.z && a.nx == b.nx &&
                             a.btag == b.btag && a.ctag == b.ctag,
                         "round trip");
        }
    PointSet bogus;
    bad += check(read_gmsh((std::string(argv[1]) + "/missing.msh").c_str(), 2, bogus) != 0, "missing file");
    for (int dim = 2; dim <= 3; dim++)
        for (int deg = 1; deg <= 6; deg++) {
            short px[128], py[128], pz[128];
            const int n = poly_terms(dim, deg, nullptr, nullptr, nullptr);
            bad += check(n <= 128 && poly_terms(dim, deg, px, py, pz) == n, "poly terms");
        }
    uint64_t s = 1;
    double acc = 0;
    for (int i = 0; i < 1000; i++) acc += uniform01(s);
    bad += check(acc > 400 && acc < 600, "uniform01");
    printf("host_io_driver: %s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
