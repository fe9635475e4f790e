// The library's host-only C++ (ag_host.cpp: MT19937, the uniform and R-MAT
// generators, edge-list file I/O; ag_csr_check.h: the CSR file check) built
// with AddressSanitizer + UndefinedBehaviorSanitizer, no HIP.  It reads the
// reference's fixture graphs committed under tests/golden/reference_graphs,
// round-trips edge lists through files, and feeds the CSR check well-formed
// and malformed arrays.  Exits non-zero on the first failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "accelgraph_amd.h"
#include "ag_csr_check.h"
#include "edgeList.h"
#include "mt19937.h"

namespace ag {
// ag_runtime.hip defines it in the library; this host-only test has no HIP.
void set_last_error(int, const char *) {}
} // namespace ag

static int failures = 0;
static void expect(bool ok, const char *what)
{
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) failures++;
}

int main(int argc, char **argv)
{
    const std::string fixtures = argc > 1 ? argv[1] : "tests/golden/reference_graphs";
    const std::string tmp = argc > 2 ? argv[2] : "/tmp";
    mt19937state st;
    initializeMersenneState(&st, 5489);
    expect(generateRandInt(&st) == 3499211612u, "MT19937 known answer");

    std::vector<uint32_t> s(100000), d(100000);
    expect(ag_gen_uniform(1, 4096, 100000, s.data(), d.data()) == AG_OK, "uniform generator");
    std::vector<uint32_t> rs(65536), rd(65536);
    expect(ag_gen_rmat(2, 12, 16, 0.57, 0.19, 0.19, 1, rs.data(), rd.data(), 2) == AG_OK,
           "R-MAT s12 ef16");

    for (const char *f : {"v51_e1021/graph", "v51_e1021/graph.bin", "v51_e1021/graph.wbin",
                          "graphbrew/graph.bin", "v300_e2730/graph.wbin"}) {
        const std::string path = fixtures + "/" + f;
        const bool bin = std::strstr(f, ".bin") != nullptr, wbin = std::strstr(f, ".wbin") != nullptr;
        struct EdgeList *el =
            readEdgeListsFile(path.c_str(), bin ? 1u : wbin ? 2u : 0u, 0, 0);
        expect(el != nullptr && el->num_edges > 0, ("read " + path).c_str());
        if (!el) continue;
        const std::string out = tmp + "/ag_asan_roundtrip.bin";
        expect(writeEdgeListToBinaryFile(el, out.c_str()) == 0, "write binary edge list");
        struct EdgeList *back = readEdgeListsbin(out.c_str(), 0, 0, 0);
        expect(back && back->num_edges == el->num_edges &&
                   std::memcmp(back->edges_array_src, el->edges_array_src, 4ull * el->num_edges) == 0 &&
                   std::memcmp(back->edges_array_dest, el->edges_array_dest, 4ull * el->num_edges) == 0,
               "binary round trip");
        freeEdgeList(back);
        freeEdgeList(el);
        std::remove(out.c_str());
    }
    expect(readEdgeListsFile((tmp + "/no_such_file_here").c_str(), 1, 0, 0) == nullptr,
           "missing file refused");

    // CSR check: a well-formed 4-vertex CSR, then one defect at a time
    const uint32_t N = 4, E = 5;
    uint32_t deg[4] = {2, 0, 3, 0}, idx[4] = {0, 0, 2, 0}, off[5] = {0, 2, 2, 5, 5};
    expect(ag::csr_direction_check(N, E, deg, idx, off) == nullptr, "well-formed CSR passes");
    expect(ag::csr_direction_check(N, E, deg, nullptr, nullptr) == nullptr, "degrees alone pass");
    uint32_t off_bad[5] = {0, 3, 2, 5, 5};
    expect(ag::csr_direction_check(N, E, nullptr, nullptr, off_bad) != nullptr, "decreasing offsets");
    uint32_t off_end[5] = {0, 2, 2, 5, 6};
    expect(ag::csr_direction_check(N, E, nullptr, nullptr, off_end) != nullptr, "offsets[N] != E");
    uint32_t idx_bad[4] = {0, 0, 3, 0};
    expect(ag::csr_direction_check(N, E, deg, idx_bad, nullptr) != nullptr, "segment past E");
    uint32_t deg_bad[4] = {2, 0, 2, 1};
    expect(ag::csr_direction_check(N, E, deg_bad, idx, off) != nullptr, "degree vs offsets");
    uint32_t deg_sum[4] = {2, 0, 4, 0};
    expect(ag::csr_direction_check(N, E, deg_sum, nullptr, nullptr) != nullptr, "degree sum");
    uint32_t idx_wrap[4] = {0, 0, 0xFFFFFFFEu, 0};
    expect(ag::csr_direction_check(N, E, deg, idx_wrap, nullptr) != nullptr, "index overflow");
    std::printf("%s host C++ under ASan/UBSan\n", failures ? "FAIL" : "PASS");
    return failures ? 1 : 0;
}
