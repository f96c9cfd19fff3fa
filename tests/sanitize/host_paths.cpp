// The library's host-only code (mesh generation and RCM, stencils, quadrature
// tables, host RBF coefficients, reference arrays, block tables, Hilbert tile
// order and relabelling, device layouts for every tile size, batch unions,
// polyMesh write and read) on small corner, Riemann and periodic meshes, for
// a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_cpu_host_sanitizers.py).
#include <cstdio>
#include <cstdlib>
#include <string>
#include "afar_host.h"
#include "foam_io.h"
#include "layout.h"
using namespace afar;
static std::string case_dir;
static void exercise(const char* name, const PolyMesh& pm, int fq, int forder) {
    Mesh m = build_mesh(pm, true);
    Schemes sc; sc.cOrder = 7; sc.fOrder = forder; sc.faceQuadrature = fq;
    Precomputed pc = precompute(m, sc);
    std::vector<double> W0(4 * (size_t)m.nc, 1.0);
    RefArrays ra = make_ref_arrays(m, pc, W0.data(), sc.R, sc.Cv);
    BlockTables bt = make_block_tables(ra);
    BatchIndex one; one.cell_off = {0, ra.nc}; one.face_off = {0, ra.nf}; one.bface_off = {0, ra.nb};
    for (int group : {0, 32, 64}) {
        Relabel r = compact_tile_order(ra, one, group);
        RefArrays rl = relabel(ra, r);
        for (int tile : {64, 128, 256}) DevLayout L = make_dev_layout(rl, tile);
    }
    std::vector<const RefArrays*> parts = {&ra, &ra, &ra};
    BatchIndex idx;
    RefArrays u = union_ref_arrays(parts, idx);
    DevLayout Lu = make_dev_layout(u, 64);
    int reread = -1;
    if (ra.nb > 0) {
        write_polymesh_2d(m, case_dir);
        PolyMesh back = read_polymesh_2d(case_dir, nullptr, nullptr);
        reread = build_mesh(back, false).nc;
    }
    printf("%s: nc %d nf %d nb %d maxS %d, union nc %d, reread nc %d\n", name, ra.nc, ra.nf, ra.nb, ra.maxS, u.nc, reread);
    fflush(stdout);
}
int main(int argc, char** argv) {
    if (argc < 2) return 2;
    case_dir = argv[1];
    const double inflow[4] = {1 / 1.4, 2.0, 0.0, 1.0};
    exercise("corner", generate_corner(30, 15, 1.0, 0.5, 0.15, 1, 0.5, 15.0, inflow, 2), 0, 2);
    exercise("riemann", generate_riemann(24, 0.2, 1, 0.5), 1, 2);
    exercise("periodic tri", generate_periodic(12, 10.0, 0.15, 0, 1.0), 1, 3);
    exercise("periodic quad", generate_periodic(12, 10.0, 0.15, 0, 0.0), 0, 2);
    return 0;
}
