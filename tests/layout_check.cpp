// Host layout rules of the SpMV kernels against a CPU replay of the kernels' indexing (no GPU is used).
//
// Builds CSR matrices of several shapes, lays them out with the library's own make_layout (the code the device
// backends upload), and replays on the CPU exactly what each kernel reads and writes: the row blocks and their
// nonzero ranges, the packed (position, column) words of the column-sorted layouts, the LDS staging positions,
// the per-slab row pointers and partial sums, the chunks of long rows. Checks that every index stays in range,
// that every staged product is written once and read once, that blocks respect their nonzero and row caps, that
// the slab count and the one-piece/slab choice follow their rules, and that the result equals the CSR product
// in FP64 to 1e-10. Argument: "sorted" or "unsorted" (the layout family, as TPDHG_LAYOUT selects it).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../taral-lp_amd/csrc/hipcommon.hpp"

using namespace tpdhg;

static int fails = 0;
#define CHECK(c)                                                                    \
    do {                                                                            \
        if (!(c) && fails++ < 20) std::printf("FAIL %s (line %d)\n", #c, __LINE__); \
    } while (0)

struct Case {
    const char* name;
    int rows, cols, per_row, long_rows, long_len, blocks;  // blocks > 1: row i touches only its block's columns
    int want_slabs;                                        // expected slab count (0 = one piece)
    int empty_run = 0;  // > 0: rows i with i mod (2 empty_run) < empty_run hold no nonzeros
};

static int wg_of(int sb) { return sb / 16 < kWG ? kWG : (sb / 16 > 1024 ? 1024 : sb / 16); }

// Longest row of rows [r0, r1) under row pointers rp (absolute nonzero positions).
template <class RP>
static int64_t longest(int r0, int r1, RP&& rp) {
    int64_t m = 0;
    for (int r = r0; r < r1; ++r) m = std::max<int64_t>(m, int64_t(rp(r + 1)) - rp(r));
    return m;
}

static void run(const Case& c, bool sorted_mode, unsigned seed) {
    std::mt19937 rng(seed);
    std::vector<int32_t> ptr{0}, idx;
    std::vector<double> val;
    const int bw = c.cols / c.blocks;
    for (int i = 0; i < c.rows; ++i) {
        const bool lng = i < c.long_rows;
        const bool empty = c.empty_run > 0 && i % (2 * c.empty_run) < c.empty_run;
        const int k = empty ? 0 : lng ? c.long_len : 1 + int(rng() % unsigned(2 * c.per_row));
        const int blk = std::min(c.blocks - 1, int(int64_t(i) * c.blocks / c.rows));
        std::vector<int32_t> cs;
        for (int t = 0; t < k; ++t)
            cs.push_back(c.blocks > 1 && !lng ? blk * bw + int(rng() % unsigned(bw)) : int(rng() % unsigned(c.cols)));
        std::sort(cs.begin(), cs.end());
        cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
        for (int x : cs) idx.push_back(x), val.push_back(double(int(rng() % 2001) - 1000) / 1000.0);
        ptr.push_back(int32_t(idx.size()));
    }
    std::vector<double> v(static_cast<size_t>(c.cols));
    for (double& x : v) x = double(int(rng() % 2001) - 1000) / 997.0;
    std::vector<double> ref(size_t(c.rows), 0.0);
    for (int i = 0; i < c.rows; ++i)
        for (int k = ptr[size_t(i)]; k < ptr[size_t(i) + 1]; ++k) ref[size_t(i)] += val[size_t(k)] * v[size_t(idx[size_t(k)])];

    const HostLayout L = make_layout(c.rows, c.cols, ptr.data(), idx.data());
    const int64_t nnz = ptr[size_t(c.rows)];
    CHECK(int64_t(L.idx.size()) == nnz);
    CHECK(L.nslab == c.want_slabs);
    // values in the layout's order, as upload_values arranges them
    std::vector<double> h(static_cast<size_t>(nnz));
    for (int64_t k = 0; k < nnz; ++k) h[size_t(k)] = L.perm.empty() ? val[size_t(k)] : val[size_t(L.perm[size_t(k)])];
    std::vector<double> y(size_t(c.rows), NAN);
    const uint32_t mask = L.sorted ? (1u << L.cbits) - 1u : 0xffffffffu;

    if (!L.slab && !L.sorted) {  // k_rowblocks + k_long_rows
        for (int b = 0; b < L.nblk; ++b) {
            const int r0 = L.blk[size_t(b)], r1 = L.blk[size_t(b) + 1], k0 = L.ptr[size_t(r0)], k1 = L.ptr[size_t(r1)];
            CHECK(r1 > r0);
            if (k1 - k0 > kCap) {
                CHECK(r1 - r0 == 1);
                continue;
            }
            CHECK(r1 - r0 <= kWG);
            CHECK(chain_fits(r1 - r0, longest(r0, r1, [&](int r) { return L.ptr[size_t(r)]; }), kWG));
            for (int r = r0; r < r1; ++r) {
                double s = 0;
                for (int k = L.ptr[size_t(r)]; k < L.ptr[size_t(r) + 1]; ++k) {
                    CHECK(L.idx[size_t(k)] >= 0 && L.idx[size_t(k)] < c.cols);
                    s += h[size_t(k)] * v[size_t(L.idx[size_t(k)])];
                }
                y[size_t(r)] = s;
            }
        }
    } else if (!L.slab) {  // k_sorted
        CHECK(L.sb >= kCap && L.sb <= kSortPlain && L.cbits == 19);
        const int WG = wg_of(L.sb);
        for (int b = 0; b < L.nblk; ++b) {
            const int r0 = L.blk[size_t(b)], r1 = L.blk[size_t(b) + 1], k0 = L.ptr[size_t(r0)], k1 = L.ptr[size_t(r1)];
            CHECK(r1 > r0);
            if (k1 - k0 > kCap && r1 - r0 == 1) continue;  // chunks below
            CHECK(k1 - k0 <= L.sb && r1 - r0 <= 8 * WG);
            CHECK(chain_fits(r1 - r0, longest(r0, r1, [&](int r) { return L.ptr[size_t(r)]; }), WG));
            std::vector<double> sh(size_t(L.sb), NAN);
            for (int k = k0; k < k1; ++k) {
                const uint32_t p = uint32_t(L.idx[size_t(k)]), pos = p >> L.cbits, col = p & mask;
                CHECK(int(pos) < k1 - k0 && int(col) < c.cols && std::isnan(sh[pos]));
                sh[pos] = h[size_t(k)] * v[col];
            }
            for (int r = r0; r < r1; ++r) {
                double s = 0;
                for (int j = L.ptr[size_t(r)] - k0; j < L.ptr[size_t(r) + 1] - k0; ++j) {
                    CHECK(!std::isnan(sh[size_t(j)]));
                    s += sh[size_t(j)];
                }
                y[size_t(r)] = s;
            }
        }
    } else {  // k_sorted_slab or k_slab_partial, then k_combine_epi
        CHECK(L.sorted == sorted_mode);
        std::vector<double> part(size_t(L.nslab) * size_t(c.rows), NAN);
        const int WG = L.sorted ? wg_of(L.sb) : kWG;
        CHECK(int(L.sstart.size()) == L.nslab + 1);
        if (L.sorted) {  // block starts and 16-bit in-block row offsets replace the per-slab row pointers
            CHECK(L.ptr.empty() && L.bk.size() == L.blk.size() && L.roff.size() == size_t(L.nslab) * size_t(c.rows));
        } else {
            CHECK(L.ptr.size() == size_t(L.nslab) * (size_t(c.rows) + 1));
        }
        for (int q = 0; q < L.nslab; ++q) {
            const int32_t* blk = L.blk.data() + L.sboff[size_t(q)];
            const int32_t* bk = L.sorted ? L.bk.data() + L.sboff[size_t(q)] : nullptr;
            const uint16_t* roff = L.sorted ? L.roff.data() + size_t(q) * size_t(c.rows) : nullptr;
            const int32_t* qp = L.sorted ? nullptr : L.ptr.data() + size_t(q) * (size_t(c.rows) + 1);
            const int first = L.sorted ? L.sstart[size_t(q)] : 0;
            const int width = L.sorted ? L.sstart[size_t(q) + 1] - L.sstart[size_t(q)] : c.cols;
            CHECK(L.snblk[size_t(q)] <= L.maxblk && blk[0] == 0 && blk[L.snblk[size_t(q)]] == c.rows);
            if (L.sorted) CHECK(bk[0] >= 0 && bk[L.snblk[size_t(q)]] <= int(L.idx.size()));
            for (int lb = 0; lb < L.snblk[size_t(q)]; ++lb) {
                const int r0 = blk[lb], r1 = blk[lb + 1];
                const int k0 = L.sorted ? bk[lb] : qp[r0], k1 = L.sorted ? bk[lb + 1] : qp[r1];
                CHECK(r1 > r0 && k1 >= k0);
                // row r's nonzeros in the layout: [rs(r), rs(r + 1)), absolute
                auto rs = [&](int r) { return L.sorted ? (r < r1 ? k0 + int(roff[r]) : k1) : qp[r]; };
                if (L.sorted)
                    for (int r = r0; r < r1; ++r) CHECK(rs(r) <= rs(r + 1) && (r > r0 || roff[r] == 0));
                if (k1 - k0 > kCap && r1 - r0 == 1) {  // one long row: strided sums
                    double s = 0;
                    for (int k = k0; k < k1; ++k) {
                        const uint32_t col = uint32_t(L.idx[size_t(k)]) & mask;
                        CHECK(int(col) < width);
                        s += h[size_t(k)] * v[size_t(first) + col];
                    }
                    part[size_t(q) * size_t(c.rows) + size_t(r0)] = s;
                    continue;
                }
                CHECK(chain_fits(r1 - r0, longest(r0, r1, rs), WG));
                if (!L.sorted) {
                    CHECK(k1 - k0 <= kCap && r1 - r0 <= kWG);
                    for (int r = r0; r < r1; ++r) {
                        double s = 0;
                        for (int k = qp[r]; k < qp[r + 1]; ++k) {
                            CHECK(int(int64_t(L.idx[size_t(k)]) * L.nslab / c.cols) == q);
                            s += h[size_t(k)] * v[size_t(L.idx[size_t(k)])];
                        }
                        part[size_t(q) * size_t(c.rows) + size_t(r)] = s;
                    }
                    continue;
                }
                CHECK(k1 - k0 <= L.sb && r1 - r0 <= 8 * WG);
                std::vector<double> sh(size_t(L.sb), NAN);
                for (int k = k0; k < k1; ++k) {
                    const uint32_t p = uint32_t(L.idx[size_t(k)]), pos = p >> L.cbits, col = p & mask;
                    CHECK(int(pos) < k1 - k0 && int(col) < width && std::isnan(sh[pos]));
                    sh[pos] = h[size_t(k)] * v[size_t(first) + col];
                }
                for (int r = r0; r < r1; ++r) {
                    double s = 0;
                    for (int j = rs(r) - k0; j < rs(r + 1) - k0; ++j) {
                        CHECK(!std::isnan(sh[size_t(j)]));
                        s += sh[size_t(j)];
                    }
                    part[size_t(q) * size_t(c.rows) + size_t(r)] = s;
                }
            }
        }
        for (int r = 0; r < c.rows; ++r) {
            double s = 0;
            for (int q = 0; q < L.nslab; ++q) {
                const double p = part[size_t(q) * size_t(c.rows) + size_t(r)];
                CHECK(!std::isnan(p));
                s += p;
            }
            y[size_t(r)] = s;
        }
    }
    if (!L.slab) {  // long rows of the one-piece layouts: kCap-nonzero chunks summed in order
        std::vector<double> lp(L.chunk.size() / 3);
        for (size_t ch = 0; ch < lp.size(); ++ch) {
            double s = 0;
            CHECK(L.chunk[3 * ch + 2] - L.chunk[3 * ch + 1] <= kCap);
            for (int k = L.chunk[3 * ch + 1]; k < L.chunk[3 * ch + 2]; ++k) {
                const uint32_t col = uint32_t(L.idx[size_t(k)]) & mask;
                CHECK(int(col) < c.cols);
                s += h[size_t(k)] * v[col];
            }
            lp[ch] = s;
        }
        for (size_t i = 0; i < L.lrow.size() / 3; ++i) {
            double s = 0;
            for (int ch = 0; ch < L.lrow[3 * i + 2]; ++ch) s += lp[size_t(L.lrow[3 * i + 1] + ch)];
            y[size_t(L.lrow[3 * i])] = s;
        }
    }
    double err = 0, scale = 0;
    for (int i = 0; i < c.rows; ++i) {
        CHECK(!std::isnan(y[size_t(i)]));
        err = std::max(err, std::fabs(y[size_t(i)] - ref[size_t(i)]));
        scale = std::max(scale, std::fabs(ref[size_t(i)]));
    }
    CHECK(err <= 1e-10 * (1 + scale));
    std::printf("%-26s rows %7d cols %7d nnz %9lld  slabs %2d sorted %d sb %5d  blocks %5d/%5d  err %.1e\n", c.name,
                c.rows, c.cols, (long long)nnz, L.nslab, int(L.sorted), L.sb, L.nblk, L.maxblk, err);
}

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "slab") == 0) {  // TPDHG_LAYOUT=slab: sorted slabs on small matrices too
        setenv("TPDHG_LAYOUT", "slab", 1);
        unsetenv("TPDHG_SLABS");
        const Case forced[] = {
            {"forced_plain", 3000, 5000, 10, 0, 0, 1, 8},
            {"forced_long_rows", 5000, 4000, 30, 3, 3900, 1, 8},
            {"forced_tiny_one_piece", 1, 10, 3, 0, 0, 1, 0},  // fewer than 16 columns: no slabs
            {"forced_block_structured", 16000, 110000, 16, 10, 10240, 64, 8},
            {"forced_runs_of_empty_rows", 12000, 60000, 20, 0, 0, 1, 8, 1500},
            {"forced_wide", 2000, 1200000, 30, 0, 0, 1, 16},
        };
        unsigned seed = 101;
        for (const Case& c : forced) run(c, c.want_slabs > 0, seed++);
        std::printf("fails %d\n", fails);
        return fails != 0;
    }
    const bool sorted_mode = argc > 1 && std::strcmp(argv[1], "sorted") == 0;
    setenv("TPDHG_LAYOUT", sorted_mode ? "sorted" : "unsorted", 1);
    unsetenv("TPDHG_SLABS");
    const Case cases[] = {
        {"plain", 3000, 5000, 10, 0, 0, 1, 0},
        {"plain_long_rows", 5000, 4000, 30, 3, 3900, 1, 0},  // rows of ~3,900 nonzeros: kCap-nonzero chunks
        // rows of 1,000 nonzeros between runs of 50 empty rows: blocks cut to bound the row-sum chains
        {"medium_rows_and_empty_runs", 4000, 20000, 10, 400, 1000, 1, 0, 50},
        {"slab8_medium_rows_and_empty_runs", 20000, 700000, 10, 400, 1000, 1, 8, 50},
        {"one_row", 1, 10, 3, 0, 0, 1, 0},
        {"slab8", 40000, 600000, 20, 0, 0, 1, 8},
        {"slab8_long_rows", 20000, 700000, 20, 2, 30000, 1, 8},
        {"slab16", 300000, 1200000, 15, 0, 0, 1, 16},
        {"block_structured_one_piece", 160000, 1100000, 16, 100, 10240, 64, 0},
        {"slab8_runs_of_empty_rows", 120000, 600000, 20, 0, 0, 1, 8, 15000},  // exercises the rows-per-block cap
    };
    unsigned seed = 1;
    for (const Case& c : cases) run(c, sorted_mode, seed++);
    if (sorted_mode) {  // 16,384-nonzero slab blocks on 1,024 threads need >= 2^24 nonzeros
        run({"slab16_largest_blocks", 900000, 1200000, 20, 0, 0, 1, 16}, true, seed++);
    }
    setenv("TPDHG_SLABS", "4", 1);
    run({"slab4_forced", 40000, 600000, 20, 0, 0, 1, 4}, sorted_mode, seed++);
    std::printf("fails %d\n", fails);
    return fails != 0;
}
