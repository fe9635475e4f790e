// CSR binary serialization: a resident graph's CSR arrays to a file and back,
// so a serving process can load graphs without re-sorting their edge lists.
//
// The reference's binary CSR format (graphCSR-binary, README.md:295-306,
// 477-486; accel-graph.c:339-342) lives in the OpenGraph submodule, which is
// not available, so this layout is this library's own:
//
//   offset  size  field
//        0     8  magic "AGCSRv1\0"
//        8     4  N (vertices), uint32
//       12     4  E (edges), uint32
//       16     4  section mask, uint32 (bit i set: section i present)
//       20     4  reserved, 0
//       24     8  FNV-1a 64 of the payload bytes
//       32     .  payload: the present sections in bit order, uint32 each
//
//   bit  section (length)            bit  section (length)
//     0  sorted src (E)                5  inverse sorted src (E)
//     1  sorted dest (E)               6  inverse sorted dest (E)
//     2  out_degree (N)                7  inverse out_degree (N)
//     3  edges_idx (N)                 8  inverse edges_idx (N)
//     4  offsets (N + 1)               9  inverse offsets (N + 1)
//
// All integers are little-endian.  Loading checks the magic, the section
// lengths against the file size, the checksum, offsets[N] == E and that every
// vertex id is < N.  Host code only (hipMemcpy to and from the device).
#include <algorithm>
#include <cstring>
#include <fstream>

#include "ag_csr_check.h"
#include "ag_internal.h"

namespace ag {

namespace {

constexpr char kMagic[8] = {'A', 'G', 'C', 'S', 'R', 'v', '1', '\0'};
constexpr int kSections = 10;

struct FileHeader {
    char magic[8];
    uint32_t n, e, mask, reserved;
    uint64_t checksum;
};
static_assert(sizeof(FileHeader) == 32, "header layout");

uint64_t fnv1a(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

uint64_t section_len(int i, uint32_t N, uint32_t E)
{
    switch (i) {
    case 0: case 1: case 5: case 6: return E;
    case 4: case 9: return (uint64_t)N + 1;
    default: return N;
    }
}

DevBuf &section_buf(ag_graph *g, int i)
{
    switch (i) {
    case 0: return g->fwd_src;
    case 1: return g->fwd_dst;
    case 2: return g->fwd_outdeg;
    case 3: return g->fwd_idx;
    case 4: return g->fwd_off;
    case 5: return g->inv_src;
    case 6: return g->inv_dst;
    case 7: return g->inv_outdeg;
    case 8: return g->inv_idx;
    default: return g->inv_off;
    }
}

} // namespace
} // namespace ag

using namespace ag;

extern "C" {

int ag_graph_save(ag_ctx *ctx, const ag_graph *gc, const char *path)
{
    return guard([&] {
        PhaseRange phase_("ag_graph_save");
        check_ctx(ctx);
        if (!gc || !path) fail(AG_ERR_INVALID, "null argument");
        ag_graph *g = const_cast<ag_graph *>(gc); // only read
        FileHeader h{};
        std::memcpy(h.magic, kMagic, 8);
        h.n = g->N;
        h.e = g->E;
        std::vector<std::vector<uint32_t>> data(kSections);
        uint64_t sum = kFnvBasis;
        for (int i = 0; i < kSections; i++) {
            const DevBuf &b = section_buf(g, i);
            const uint64_t len = section_len(i, g->N, g->E);
            if (!b.p || b.bytes < len * 4) continue;
            h.mask |= 1u << i;
            data[i].resize(len);
            if (len)
                AG_HIP(hipMemcpyAsync(data[i].data(), b.p, len * 4, hipMemcpyDeviceToHost,
                                      ctx->stream));
        }
        AG_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < kSections; i++)
            if (h.mask & (1u << i)) sum = fnv1a(sum, data[i].data(), data[i].size() * 4);
        h.checksum = sum;
        std::ofstream f(path, std::ios::binary | std::ios::trunc);
        if (!f) fail(AG_ERR_IO, "cannot open %s for writing", path);
        f.write(reinterpret_cast<const char *>(&h), sizeof(h));
        for (int i = 0; i < kSections; i++)
            if (h.mask & (1u << i))
                f.write(reinterpret_cast<const char *>(data[i].data()),
                        (std::streamsize)(data[i].size() * 4));
        f.flush();
        if (!f) fail(AG_ERR_IO, "write to %s failed", path);
    });
}

int ag_graph_load(ag_ctx *ctx, const char *path, ag_graph **out)
{
    return guard([&] {
        PhaseRange phase_("ag_graph_load");
        check_ctx(ctx);
        if (!path || !out) fail(AG_ERR_INVALID, "null argument");
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) fail(AG_ERR_IO, "cannot open %s", path);
        const uint64_t fsize = (uint64_t)f.tellg();
        f.seekg(0);
        FileHeader h{};
        if (fsize < sizeof(h) || !f.read(reinterpret_cast<char *>(&h), sizeof(h)))
            fail(AG_ERR_INVALID, "%s: too short for a CSR file", path);
        if (std::memcmp(h.magic, kMagic, 8) != 0) fail(AG_ERR_INVALID, "%s: not an AGCSRv1 file", path);
        if (h.mask >> kSections) fail(AG_ERR_INVALID, "%s: unknown sections 0x%x", path, h.mask);
        uint64_t want = sizeof(h);
        for (int i = 0; i < kSections; i++)
            if (h.mask & (1u << i)) want += section_len(i, h.n, h.e) * 4;
        if (want != fsize)
            fail(AG_ERR_INVALID, "%s: %llu bytes, the header describes %llu", path,
                 (unsigned long long)fsize, (unsigned long long)want);
        const bool fwd = (h.mask & 0x1f) == 0x1f, inv = (h.mask & 0x3e0) == 0x3e0;
        const bool pr_only = (h.mask & 0x1c4) == 0x1c4; // out_degree + inverse degree/idx/dest
        if (!fwd && !pr_only) fail(AG_ERR_INVALID, "%s: neither a full CSR nor the PageRank arrays", path);
        std::unique_ptr<ag_graph> g(new ag_graph());
        g->N = h.n;
        g->E = h.e;
        // Every section is read and checked on the host before any upload: the
        // checksum guards against corruption only (anyone can recompute it),
        // so the per-vertex ranges are checked too (ag_csr_check.h).
        uint64_t sum = kFnvBasis;
        std::vector<std::vector<uint32_t>> host(kSections);
        for (int i = 0; i < kSections; i++) {
            if (!(h.mask & (1u << i))) continue;
            const uint64_t len = section_len(i, h.n, h.e);
            host[i].resize(len);
            if (len && !f.read(reinterpret_cast<char *>(host[i].data()), (std::streamsize)(len * 4)))
                fail(AG_ERR_IO, "%s: read failed", path);
            sum = fnv1a(sum, host[i].data(), len * 4);
        }
        if (sum != h.checksum) fail(AG_ERR_INVALID, "%s: checksum mismatch (corrupt file)", path);
        auto sec = [&](int i) { return (h.mask & (1u << i)) ? host[i].data() : nullptr; };
        for (int d = 0; d < 2; d++) {
            const char *why = csr_direction_check(h.n, h.e, sec(5 * d + 2), sec(5 * d + 3),
                                                  sec(5 * d + 4));
            if (why) fail(AG_ERR_INVALID, "%s: %s CSR: %s", path, d ? "inverse" : "forward", why);
        }
        for (int i = 0; i < kSections; i++) {
            if (!(h.mask & (1u << i))) continue;
            const uint64_t len = section_len(i, h.n, h.e);
            DevBuf &b = section_buf(g.get(), i);
            b.alloc(len * 4);
            if (len)
                AG_HIP(hipMemcpyAsync(b.p, host[i].data(), len * 4, hipMemcpyHostToDevice,
                                      ctx->stream));
        }
        AG_HIP(hipStreamSynchronize(ctx->stream));
        // vertex ids must index the vertex arrays
        for (int i : {0, 1, 5, 6}) {
            if (!(h.mask & (1u << i)) || h.e == 0) continue;
            const uint32_t m = device_max1(ctx, section_buf(g.get(), i).as<uint32_t>(), h.e);
            if (m >= h.n) fail(AG_ERR_INVALID, "%s: vertex id %u >= N = %u", path, m, h.n);
        }
        g->has_full_fwd = fwd;
        g->has_inverse_full = inv;
        *out = g.release();
    });
}

int ag_csr_file_info(const char *path, uint32_t *num_vertices, uint32_t *num_edges,
                     uint32_t *sections)
{
    return guard([&] {
        if (!path) fail(AG_ERR_INVALID, "null path");
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) fail(AG_ERR_IO, "cannot open %s", path);
        const uint64_t fsize = (uint64_t)f.tellg();
        f.seekg(0);
        FileHeader h{};
        if (fsize < sizeof(h) || !f.read(reinterpret_cast<char *>(&h), sizeof(h)))
            fail(AG_ERR_INVALID, "%s: too short for a CSR file", path);
        if (std::memcmp(h.magic, kMagic, 8) != 0) fail(AG_ERR_INVALID, "%s: not an AGCSRv1 file", path);
        uint64_t want = sizeof(h), sum = kFnvBasis;
        std::vector<char> buf(1 << 20);
        for (int i = 0; i < kSections; i++)
            if (h.mask & (1u << i)) want += section_len(i, h.n, h.e) * 4;
        if ((h.mask >> kSections) || want != fsize)
            fail(AG_ERR_INVALID, "%s: size or section mask does not match the header", path);
        for (uint64_t left = fsize - sizeof(h); left;) {
            const size_t n = (size_t)std::min<uint64_t>(left, buf.size());
            if (!f.read(buf.data(), (std::streamsize)n)) fail(AG_ERR_IO, "%s: read failed", path);
            sum = fnv1a(sum, buf.data(), n);
            left -= n;
        }
        if (sum != h.checksum) fail(AG_ERR_INVALID, "%s: checksum mismatch (corrupt file)", path);
        if (num_vertices) *num_vertices = h.n;
        if (num_edges) *num_edges = h.e;
        if (sections) *sections = h.mask;
    });
}

} // extern "C"
