// Host-side check of the device lattice layout (glac_amd/csrc/lattice.hpp):
// every storage index decodes to in-range coordinates and back, the
// lexicographic, parity and time-slice enumerations are bijections, and
// every neighbour step stays in range.  Built with hipcc for the host only
// by tests/test_layout.py; prints "ok <dims>" per lattice, exits non-zero
// on the first violation.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../glac_amd/csrc/lattice.hpp"

using namespace glacx;

#define REQUIRE(cond, ...)                                                                      \
    do {                                                                                         \
        if (!(cond)) {                                                                           \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                    \
            std::printf(__VA_ARGS__);                                                            \
            std::printf("\n");                                                                   \
            std::exit(1);                                                                        \
        }                                                                                        \
    } while (0)

static void check(const int dims[4])
{
    const Geom g = make_geom(dims);
    const int V = g.v;
    std::vector<char> seen(V, 0);
    for (int idx = 0; idx < V; idx++) {
        const Coord c = decode(g, idx);
        for (int k = 0; k < 4; k++) REQUIRE(c.c[k] >= 0 && c.c[k] < dims[k], "coord %d of idx %d", k, idx);
        REQUIRE(index_of(g, c) == idx, "index_of(decode(%d)) = %d", idx, index_of(g, c));
        const int lex = lex_of(g, c);
        REQUIRE(lex >= 0 && lex < V && !seen[lex], "lex %d of idx %d", lex, idx);
        seen[lex] = 1;
        // the multiply-free neighbour steps agree with index_of(step(...)),
        // also for two steps in a row (the staple's x + mu - nu)
        const Site st = site_of(g, idx);
        REQUIRE(st.idx == idx && st.p == ((c.c[0] + c.c[1] + c.c[2] + c.c[3]) & 1), "site_of(%d)", idx);
        for (int mu = 0; mu < 4; mu++)
            for (int s = -1; s <= 1; s += 2) {
                const Site a = site_step(g, st, mu, s);
                const Coord cn = step(g, c, mu, s);
                REQUIRE(a.idx == index_of(g, cn), "site_step(%d, mu %d, s %d) = %d, want %d", idx, mu, s, a.idx,
                        index_of(g, cn));
                const Site ac = site_of(g, a.idx);
                REQUIRE(ac.x == a.x && ac.y == a.y && ac.lz == a.lz && ac.zb == a.zb && ac.lt == a.lt &&
                            ac.tb == a.tb && ac.p == a.p,
                        "site_step(%d, mu %d, s %d) coordinates", idx, mu, s);
                for (int nu = 0; nu < 4; nu++)
                    for (int r = -1; r <= 1; r += 2)
                        REQUIRE(site_step(g, a, nu, r).idx == index_of(g, step(g, cn, nu, r)),
                                "two steps from %d: mu %d s %d, nu %d r %d", idx, mu, s, nu, r);
            }
        for (int mu = 0; mu < 4; mu++)
            for (int s = -1; s <= 1; s += 2) {
                const Coord n = step(g, c, mu, s);
                for (int k = 0; k < 4; k++) {
                    const int want = k == mu ? (c.c[k] + s + dims[k]) % dims[k] : c.c[k];
                    REQUIRE(n.c[k] == want, "step mu %d s %d", mu, s);
                }
                const int ni = index_of(g, n);
                REQUIRE(ni >= 0 && ni < V, "neighbour index %d", ni);
                // the neighbour has the opposite parity
                REQUIRE(((n.c[0] + n.c[1] + n.c[2] + n.c[3]) & 1) != ((c.c[0] + c.c[1] + c.c[2] + c.c[3]) & 1),
                        "parity of neighbour");
            }
    }
    for (int p = 0; p < 2; p++) {
        std::vector<char> hit(V, 0);
        for (int h = 0; h < V / 2; h++) {
            const int idx = parity_index(g, p, h);
            REQUIRE(idx >= 0 && idx < V && !hit[idx], "parity_index(%d, %d) = %d", p, h, idx);
            hit[idx] = 1;
            const Coord c = decode(g, idx);
            REQUIRE(((c.c[0] + c.c[1] + c.c[2] + c.c[3]) & 1) == p, "parity of parity_index(%d, %d)", p, h);
        }
    }
    std::vector<char> hit(V, 0);
    for (int t = 0; t < dims[3]; t++)
        for (int r = 0; r < g.vs; r++) {
            const int idx = slice_index(g, t, r);
            REQUIRE(idx >= 0 && idx < V && !hit[idx], "slice_index(%d, %d) = %d", t, r, idx);
            hit[idx] = 1;
            REQUIRE(decode(g, idx).c[3] == t, "time of slice_index(%d, %d)", t, r);
        }
    // the measurement's chunking: block b -> run (blk, lt), chunk jc; lanes
    // (p, k) cover every site once, all on slice t; partial slots unique
    {
        std::vector<char> site(V, 0), slot((size_t)g.nzb * g.cpr * dims[3], 0);
        const int nblocks = dims[3] * g.nzb * g.cpr;
        for (int b = 0; b < nblocks; b++) {
            const int rest = fdiv(b, g.d_cpr), jc = b - rest * g.cpr;
            const int lt = rest & ((1 << g.lbt) - 1), blk = rest >> g.lbt;
            const int bt = fdiv(blk, g.d_nzb), zb = blk - bt * g.nzb;
            const int t = (bt << g.lbt) + lt, c = zb * g.cpr + jc;
            REQUIRE(t < dims[3] && c < g.nzb * g.cpr, "chunk %d -> t %d c %d", b, t, c);
            REQUIRE(!slot[(size_t)c * dims[3] + t], "partial slot of chunk %d", b);
            slot[(size_t)c * dims[3] + t] = 1;
            for (int lane = 0; lane < 64; lane++) {
                const int k = jc * 32 + (lane & 31);
                if (k >= g.run) continue;
                const int idx = run_index(g, blk, lane >> 5, lt, k);
                REQUIRE(idx >= 0 && idx < V && !site[idx], "run_index chunk %d lane %d = %d", b, lane, idx);
                site[idx] = 1;
                REQUIRE(decode(g, idx).c[3] == t, "time of chunk %d lane %d", b, lane);
            }
        }
        for (int i = 0; i < V; i++) REQUIRE(site[i], "site %d not covered by the measurement chunks", i);
    }
    // the flow stage's pairing of the h-th sites of both parities
    for (int h = 0; h < V / 2; h++)
        for (int p = 0; p < 2; p++) {
            const Coord c = decode(g, parity_index(g, p, h));
            REQUIRE(((c.c[0] + c.c[1] + c.c[2] + c.c[3]) & 1) == p, "flow pairing parity");
        }
    std::printf("ok %d %d %d %d\n", dims[0], dims[1], dims[2], dims[3]);
}

int main(int argc, char **argv)
{
    if (argc % 4 != 1) return 2;
    for (int a = 1; a < argc; a += 4) {
        int d[4];
        for (int k = 0; k < 4; k++) d[k] = std::atoi(argv[a + k]);
        check(d);
    }
    return 0;
}
