// Test probe: reads an MPS file with the REFERENCE reader (src/mps.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/, never committed) and prints the model in a canonical text form, so the
// tests can compare the library's reader against it line for line. TEST INFRASTRUCTURE ONLY.
#include <cmath>
#include <cstdio>

#include "gpu/pdhg_input.hpp"
#include "src/taral.hpp"

static void num(double v) {
    if (std::isinf(v)) std::printf(v > 0 ? " inf" : " -inf");
    else std::printf(" %.17g", v);
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    Model md;
    try {
        md = read_mps(argv[1]);
    } catch (const ParseError& e) {
        std::printf("parse_error %s\n", e.what());
        return 0;
    }
    const char* gate = pdhg_input_error(md);
    std::printf("gate %s\n", gate ? gate : "ok");
    std::printf("name %s\nmaximize %d\nobj_const", md.name.c_str(), md.maximize ? 1 : 0);
    num(md.obj_const);
    std::printf("\nsentinels %zu\nrows %zu\n", md.sentinel_bounds, md.row_names.size());
    for (size_t i = 0; i < md.row_names.size(); ++i) {
        std::printf("r %s", md.row_names[i].c_str());
        num(md.row_lo[i]), num(md.row_up[i]);
        std::printf("\n");
    }
    std::printf("cols %zu\n", md.col_names.size());
    for (size_t j = 0; j < md.col_names.size(); ++j) {
        std::printf("c %s %d", md.col_names[j].c_str(), int(md.is_int[j]));
        num(md.cost[j]), num(md.col_lo[j]), num(md.col_up[j]);
        for (const Entry& e : md.cols[j]) std::printf(" %d", e.index), num(e.value);
        std::printf("\n");
    }
    std::printf("q %zu\n", md.qobj.size());
    for (const QEntry& q : md.qobj) std::printf("q %d %d", q.row, q.col), num(q.value), std::printf("\n");
    return 0;
}
