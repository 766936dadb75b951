// Host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer (built by tests/test_sanitize.py
// with g++; no device code): the MPS reader on every file given on the command line (parse errors are expected
// for the err_* and parser_* fixtures and must throw cleanly), the input gate, the scaled problem, the
// original-model recheck, and full solves on the CPU backend through the PDHG driver (FP64 and FP32, one and
// several threads), all of whose memory accesses the sanitizers watch. The GPU factories are replaced by
// stand-ins that throw: this binary never touches a device.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.hpp"
#include "device.hpp"
#include "driver.hpp"
#include "model.hpp"
#include "prep.hpp"
#include "recheck.hpp"

namespace tpdhg {
std::unique_ptr<Device> make_device(const Scaled&, const DeviceOptions&) { throw std::runtime_error("no device"); }
std::unique_ptr<FastDevice> make_fast_device(const Scaled&, const FastOptions&) { throw std::runtime_error("no device"); }
std::unique_ptr<BatchDevice> make_batch_device(const std::vector<const Scaled*>&, const BatchOptions&) {
    throw std::runtime_error("no device");
}
int device_count() { return 0; }
}  // namespace tpdhg

int main(int argc, char** argv) {
    using namespace tpdhg;
    int parsed = 0, refused = 0, solved = 0;
    for (int a = 1; a < argc; ++a) {
        Model md;
        try {
            md = read_mps(argv[a]);
        } catch (const std::exception&) {
            ++refused;
            continue;
        }
        ++parsed;
        const LP lp = flatten(md);
        if (input_error(lp)) continue;
        const Scaled P = build_scaled_csc(lp.m, lp.n, lp.colptr.data(), lp.rowidx.data(), lp.val.data(),
                                          lp.cost.data(), lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(),
                                          lp.row_up.data(), lp.maximize ? -1.0 : 1.0);
        std::vector<double> x(size_t(lp.n), 0.0), y(size_t(lp.m), 0.0);
        (void)recheck(lp, x.data(), y.data());
        for (int bits : {64, 32})
            for (int threads : {1, 3}) {
                Options o;
                o.tol = 1e-6, o.precision_bits = bits, o.cpu_threads = threads, o.max_iter = 20000;
                const Report r = solve_lp(lp, o, Timings{}, std::chrono::steady_clock::now());
                if (!std::isfinite(r.objective) && r.out.status == "near_optimal") return 3;
                ++solved;
            }
    }
    std::printf("parsed %d refused %d solved %d\n", parsed, refused, solved);
    return parsed > 0 ? 0 : 4;
}
