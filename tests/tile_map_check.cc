// Host-side check of the tiled apply's tables and output-tile map
// (reapers_amd/csrc/rp_tables.h), built by tests/test_tile_map.py.
//
//   tile_map_check <bits.bin> <len> <tile_bits> <threads> <p>
//
// reads n uint64 SpinOp bit strings (coefficients 1), builds the
// complex<float> tables with p resident-workgroup bits and prints one JSON
// line: the map's size, whether index -> tile is a bijection of the grid
// 2^(len - tile_bits), whether each group's terms are contiguous (every xh
// once), and the runs of equal XCD class along the group list for the table
// order against the contiguous order (classes xh >> p along sorted xh).
#include "rp_tables.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <set>

namespace rp {
int set_error(int code, const char *fmt, ...)
{
    va_list a;
    va_start(a, fmt);
    vfprintf(stderr, fmt, a);
    va_end(a);
    fputc('\n', stderr);
    return code;
}
}  // namespace rp

using namespace rp;

int main(int argc, char **argv)
{
    if (argc != 6) return 2;
    const int len = atoi(argv[2]), tb = atoi(argv[3]), nt = atoi(argv[4]), p = atoi(argv[5]);
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint64_t> bits;
    uint64_t b;
    while (fread(&b, 8, 1, f) == 1) bits.push_back(b);
    fclose(f);
    std::vector<double> coeff(2 * bits.size());
    for (size_t i = 0; i < bits.size(); i++) coeff[2 * i] = 1.0;
    Tables<double> tab;
    if (build_tables<float, double>(len, tb, nt, bits.data(), coeff.data(), bits.size(), tab, p))
        return 3;
    const int nbits = len - tb;
    const TileMap &m = tab.tilemap;
    bool bij = true;
    if (m.n) {
        if (m.n != nbits || m.p != p) bij = false;
        std::vector<char> seen(1u << nbits, 0);
        const uint32_t grid = 1u << nbits;
        for (uint32_t blk = 0; blk < grid && bij; blk++) {
            // the kernel's index: position within the XCD, XCD, batch
            const uint32_t j = blk >> 3;
            const uint32_t idx = ((j >> p) << (p + 3)) | ((blk & 7u) << p) | (j & ((1u << p) - 1));
            uint32_t t = 0;
            for (int i = 0; i < m.n; i++)
                if (idx >> i & 1u) t ^= m.col[i];
            if (t >= grid || seen[t]) bij = false;
            else seen[t] = 1;
        }
    }
    std::set<uint32_t> xh;
    bool contiguous = true;
    for (const GroupRuns &g : tab.groups)
        if (!xh.insert(g.xh).second) contiguous = false;
    // XCD classes: the span of the map's first p columns, or xh >> p
    Gf2Span xcd;
    for (int i = 0; i < (m.n ? p : 0); i++) xcd.insert(m.col[i]);
    auto cls = [&](uint32_t v) { return m.n ? xcd.reduce(v) : v >> p; };
    size_t runs = 0, runs_contig = 0;
    for (size_t i = 0; i < tab.groups.size(); i++)
        runs += i == 0 || cls(tab.groups[i].xh) != cls(tab.groups[i - 1].xh);
    uint32_t prev = 0;
    size_t i = 0;
    for (uint32_t v : xh) runs_contig += i++ == 0 || (v >> p) != prev, prev = v >> p;
    printf("{\"n\": %d, \"groups\": %zu, \"bijective\": %s, \"contiguous\": %s, \"runs\": %zu, "
           "\"runs_contiguous_order\": %zu}\n",
           m.n, tab.groups.size(), bij ? "true" : "false", contiguous ? "true" : "false", runs,
           runs_contig);
    return 0;
}
