// extern "C" boundary of libtaral_pdhg.so (declared in include/taral_pdhg.h): handles, return codes and a
// per-thread error message around the C++ host code and the HIP device backend.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/taral_pdhg.h"
#include "driver.hpp"
#include "noalloc.hpp"
#include "mlp.hpp"
#include "recheck.hpp"

struct tpdhg_model {
    tpdhg::LP lp;
};
struct tpdhg_scaled {
    tpdhg::Scaled s;
};
struct tpdhg_dev {
    std::unique_ptr<tpdhg::Device> d;
};
struct tpdhg_fast {
    std::unique_ptr<tpdhg::FastDevice> d;
};
struct tpdhg_batch {
    std::unique_ptr<tpdhg::BatchSession> s;
    tpdhg::Options o;
};
struct tpdhg_mlp {
    std::unique_ptr<tpdhg::Mlp> net;
};
struct tpdhg_session {
    tpdhg::LP lp;  // the session's own copy; Session keeps a reference to it
    std::unique_ptr<tpdhg::Session> s;
};

namespace {
thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        g_error.clear();
        return f();
    } catch (const tpdhg::ParseError& e) {
        return fail(TPDHG_ERR_PARSE, e.what());
    } catch (const tpdhg::ParseTimeLimit& e) {
        return fail(TPDHG_ERR_TIME, e.what());
    } catch (const std::bad_alloc&) {
        return fail(TPDHG_ERR_ARG, "out of host memory");
    } catch (const std::runtime_error& e) {
        const std::string w = e.what();
        if (w.find("no HIP device") != std::string::npos) return fail(TPDHG_ERR_NO_DEVICE, w);
        return fail(TPDHG_ERR_DEVICE, w);
    } catch (const std::exception& e) {
        return fail(TPDHG_ERR_ARG, e.what());
    }
}

int status_code(const std::string& s) {
    if (s == "near_optimal") return TPDHG_NEAR_OPTIMAL;
    if (s == "time_limit") return TPDHG_TIME_LIMIT;
    if (s == "numerical_failure") return TPDHG_NUMERICAL_FAILURE;
    return TPDHG_ITERATION_LIMIT;
}

tpdhg::Options to_options(const tpdhg_options* o) {
    tpdhg::Options r;
    if (!o) return r;
    r.tol = o->tol, r.time_limit = o->time_limit, r.max_iter = o->max_iter, r.eval_freq = o->eval_freq;
    r.row_rel = o->row_rel, r.precision_bits = o->precision_bits, r.refine = o->refine != 0;
    r.max_refinements = o->max_refinements, r.graph = o->use_graph != 0;
    r.verbose = o->verbose != 0, r.profile_kernels = o->profile_kernels != 0;
    r.halpern = o->halpern != 0, r.rho = o->reflection;
    r.omega0 = o->primal_weight, r.warm_fallback = o->warm_fallback != 0, r.progress_s = o->progress_interval;
    r.omega_rule = o->omega_rule;
    r.reuse_norm = o->reuse_norm != 0;
    r.cpu_threads = o->cpu_threads;
    r.restart_metric = o->restart_metric;
    return r;
}

// A CSR the kernels can index without leaving its arrays: ptr[0] = 0, nondecreasing row pointers ending at the
// nonzero count (`nnz` when it is >= 0), every column index in [0, cols), no null array behind a nonzero. Returns
// the first violation, or nullptr. The kernels trust these arrays, so every entry point that takes a caller's CSR
// checks them on the host first.
const char* csr_error(int32_t rows, int32_t cols, const int32_t* ptr, const int32_t* idx, const double* val,
                      int64_t nnz) {
    if (rows < 0 || cols < 0) return "negative dimensions";
    if (!ptr) return "null row pointers";
    if (ptr[0] != 0) return "row pointers must start at 0";
    for (int32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return "row pointers must not decrease";
    if (nnz >= 0 && ptr[rows] != nnz) return "row pointers do not end at the nonzero count";
    if (int64_t(ptr[rows]) > tpdhg::kMaxNnz) return "too many nonzeros for 32-bit nonzero indices";
    if (ptr[rows] > 0 && (!idx || !val)) return "null column indices or values";
    for (int32_t k = 0; k < ptr[rows]; ++k)
        if (idx[k] < 0 || idx[k] >= cols) return "column index out of range";
    return nullptr;
}

// The scaled problem arrays of tpdhg_dev_create / tpdhg_fast_create: both CSRs and every vector.
const char* problem_error(const tpdhg_scaled_arrays* p) {
    if (p->m < 0 || p->n < 0 || p->nnz < 0) return "negative dimensions";
    if (const char* e = csr_error(p->m, p->n, p->a_ptr, p->a_idx, p->a_val, p->nnz)) return e;
    if (const char* e = csr_error(p->n, p->m, p->at_ptr, p->at_idx, p->at_val, p->nnz)) return e;
    if (p->n > 0 && (!p->c || !p->l || !p->u || !p->Dc)) return "null column vector";
    if (p->m > 0 && (!p->rl || !p->ru || !p->Dr)) return "null row vector";
    return nullptr;
}
}  // namespace

extern "C" {

const char* tpdhg_last_error(void) { return g_error.c_str(); }

const char* tpdhg_status_string(int status) {
    switch (status) {
        case TPDHG_NEAR_OPTIMAL: return "near_optimal";
        case TPDHG_TIME_LIMIT: return "time_limit";
        case TPDHG_NUMERICAL_FAILURE: return "numerical_failure";
        default: return "iteration_limit";
    }
}

int tpdhg_device_count(void) { return tpdhg::device_count(); }

int tpdhg_model_read_mps(const char* path, tpdhg_model** out) {
    if (!path || !out) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        auto m = std::make_unique<tpdhg_model>();
        m->lp = tpdhg::flatten(tpdhg::read_mps(path));
        *out = m.release();
        return TPDHG_OK;
    });
}

int tpdhg_model_read_mps_text(const char* text, int64_t len, tpdhg_model** out) {
    if (!text || !out || len < 0) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        auto m = std::make_unique<tpdhg_model>();
        m->lp = tpdhg::flatten(tpdhg::read_mps_text(std::string(text, size_t(len))));
        *out = m.release();
        return TPDHG_OK;
    });
}

int tpdhg_model_from_csc(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                         const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                         const double* row_up, double obj_const, int32_t maximize, tpdhg_model** out) {
    if (m < 0 || n < 0 || !colptr || !out || (n > 0 && (!cost || !col_lo || !col_up)) ||
        (m > 0 && (!row_lo || !row_up)))
        return fail(TPDHG_ERR_ARG, "bad model arrays");
    return guarded([&]() -> int {
        const int64_t nnz = colptr[n];
        if (colptr[0] != 0 || nnz < 0 || nnz > INT32_MAX || (nnz > 0 && (!rowidx || !val)))
            return fail(TPDHG_ERR_ARG, "bad column pointers");
        for (int j = 0; j < n; ++j) {
            if (colptr[j + 1] < colptr[j]) return fail(TPDHG_ERR_ARG, "column pointers decrease");
            for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
                if (rowidx[k] < 0 || rowidx[k] >= m) return fail(TPDHG_ERR_ARG, "row index out of range");
                if (k > colptr[j] && rowidx[k] <= rowidx[k - 1])
                    return fail(TPDHG_ERR_ARG, "row indices must ascend strictly within a column");
                if (!std::isfinite(val[k])) return fail(TPDHG_ERR_ARG, "matrix value not finite");
            }
        }
        auto md = std::make_unique<tpdhg_model>();
        tpdhg::LP& lp = md->lp;
        lp.name = "LP", lp.m = m, lp.n = n;
        lp.colptr.assign(colptr, colptr + n + 1);
        lp.rowidx.assign(rowidx, rowidx + nnz), lp.val.assign(val, val + nnz);
        lp.cost.assign(cost, cost + n), lp.col_lo.assign(col_lo, col_lo + n), lp.col_up.assign(col_up, col_up + n);
        lp.row_lo.assign(row_lo, row_lo + m), lp.row_up.assign(row_up, row_up + m);
        lp.is_int.assign(size_t(n), 0);
        lp.obj_const = obj_const, lp.maximize = maximize != 0;
        *out = md.release();
        return TPDHG_OK;
    });
}

void tpdhg_model_free(tpdhg_model* model) { delete model; }

int tpdhg_model_get_info(const tpdhg_model* model, tpdhg_model_info* info) {
    if (!model || !info) return fail(TPDHG_ERR_ARG, "null argument");
    const tpdhg::LP& lp = model->lp;
    info->m = lp.m, info->n = lp.n, info->nnz = lp.nnz(), info->maximize = lp.maximize;
    info->has_integers = 0;
    for (char c : lp.is_int) info->has_integers |= c != 0;
    info->qobj_terms = int64_t(lp.qobj_terms), info->sentinel_bounds = int64_t(lp.sentinel_bounds);
    info->obj_const = lp.obj_const;
    return TPDHG_OK;
}

int tpdhg_model_export(const tpdhg_model* model, int64_t* colptr, int32_t* rowidx, double* val, double* cost,
                       double* col_lo, double* col_up, double* row_lo, double* row_up, int8_t* is_int) {
    if (!model) return fail(TPDHG_ERR_ARG, "null model");
    const tpdhg::LP& lp = model->lp;
    auto cp = [](auto* dst, const auto& src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(colptr, lp.colptr), cp(rowidx, lp.rowidx), cp(val, lp.val), cp(cost, lp.cost), cp(col_lo, lp.col_lo);
    cp(col_up, lp.col_up), cp(row_lo, lp.row_lo), cp(row_up, lp.row_up);
    if (is_int)
        for (size_t j = 0; j < lp.is_int.size(); ++j) is_int[j] = int8_t(lp.is_int[j]);
    return TPDHG_OK;
}

const char* tpdhg_model_name_of(const tpdhg_model* model, int32_t is_row, int32_t i) {
    if (!model || i < 0) return "";
    const auto& names = is_row ? model->lp.row_names : model->lp.col_names;
    return size_t(i) < names.size() ? names[size_t(i)].c_str() : "";
}

const char* tpdhg_model_input_error(const tpdhg_model* model) { return model ? tpdhg::input_error(model->lp) : nullptr; }

int tpdhg_scaled_build(const tpdhg_model* model, tpdhg_scaled** out) {
    if (!model || !out) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        const tpdhg::LP& lp = model->lp;
        auto s = std::make_unique<tpdhg_scaled>();
        s->s = tpdhg::build_scaled_csc(lp.m, lp.n, lp.colptr.data(), lp.rowidx.data(), lp.val.data(), lp.cost.data(),
                                       lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(), lp.row_up.data(),
                                       lp.maximize ? -1.0 : 1.0);
        *out = s.release();
        return TPDHG_OK;
    });
}

int tpdhg_scaled_view(const tpdhg_scaled* s, tpdhg_scaled_arrays* v) {
    if (!s || !v) return fail(TPDHG_ERR_ARG, "null argument");
    const tpdhg::Scaled& P = s->s;
    v->m = P.m, v->n = P.n, v->nnz = P.A.nnz();
    v->a_ptr = P.A.ptr.data(), v->a_idx = P.A.idx.data(), v->a_val = P.A.val.data();
    v->at_ptr = P.AT.ptr.data(), v->at_idx = P.AT.idx.data(), v->at_val = P.AT.val.data();
    v->c = P.c.data(), v->l = P.l.data(), v->u = P.u.data(), v->rl = P.rl.data(), v->ru = P.ru.data();
    v->Dr = P.Dr.data(), v->Dc = P.Dc.data(), v->bnorm = P.bnorm, v->cnorm = P.cnorm;
    return TPDHG_OK;
}

void tpdhg_scaled_free(tpdhg_scaled* s) { delete s; }

int tpdhg_dev_create(const tpdhg_scaled_arrays* p, int32_t precision_bits, int32_t eval_freq, int32_t use_graph,
                     tpdhg_dev** out) {
    if (!p || !out) return fail(TPDHG_ERR_ARG, "bad problem");
    if (const char* e = problem_error(p)) return fail(TPDHG_ERR_ARG, std::string("bad problem: ") + e);
    return guarded([&]() -> int {
        tpdhg::DeviceOptions o;
        o.precision_bits = precision_bits, o.eval_freq = eval_freq, o.use_graph = use_graph != 0;
        auto d = std::make_unique<tpdhg_dev>();
        d->d = tpdhg::make_device_raw(p->m, p->n, p->nnz, p->a_ptr, p->a_idx, p->a_val, p->at_ptr, p->at_idx,
                                      p->at_val, p->c, p->l, p->u, p->rl, p->ru, p->Dr, p->Dc, o);
        *out = d.release();
        return TPDHG_OK;
    });
}

#define DEV_CALL(dev, body)                                           \
    do {                                                              \
        if (!(dev) || !(dev)->d) return fail(TPDHG_ERR_ARG, "null device"); \
        return guarded([&]() -> int { body; return TPDHG_OK; });               \
    } while (0)

int tpdhg_dev_power_norm(tpdhg_dev* dev, int32_t iters, double* norm) {
    if (!norm || iters < 1) return fail(TPDHG_ERR_ARG, "null norm or iters < 1");
    DEV_CALL(dev, *norm = dev->d->power_norm(iters));
}
int tpdhg_dev_set_steps(tpdhg_dev* dev, double tau, double sigma) { DEV_CALL(dev, dev->d->set_steps(tau, sigma)); }
int tpdhg_dev_start(tpdhg_dev* dev, const double* x0, const double* z0) { DEV_CALL(dev, dev->d->start(x0, z0)); }
int tpdhg_dev_iterate(tpdhg_dev* dev, int32_t steps) { DEV_CALL(dev, dev->d->iterate(steps)); }
int tpdhg_dev_eval(tpdhg_dev* dev, int32_t avg, double f, double out[10]) {
    if (!out) return fail(TPDHG_ERR_ARG, "null output");
    DEV_CALL(dev, if (avg) dev->d->eval2(f, nullptr, out); else dev->d->eval2(f, out, nullptr));
}
int tpdhg_dev_eval2(tpdhg_dev* dev, double f, double cur[10], double avg[10]) {
    if (!cur && !avg) return fail(TPDHG_ERR_ARG, "null outputs");
    DEV_CALL(dev, dev->d->eval2(f, cur, avg));
}
int tpdhg_dev_restart(tpdhg_dev* dev, int32_t avg, double f) { DEV_CALL(dev, dev->d->restart(avg != 0, f)); }
int tpdhg_dev_download(tpdhg_dev* dev, int32_t avg, double f, double* x, double* z) {
    DEV_CALL(dev, dev->d->download(avg != 0, f, x, z));
}
int tpdhg_dev_spmv(tpdhg_dev* dev, int32_t transpose, const double* v, double* y) {
    if (!v || !y) return fail(TPDHG_ERR_ARG, "null vector");
    DEV_CALL(dev, dev->d->spmv(transpose ? 1 : 0, v, y));
}
int tpdhg_dev_info(tpdhg_dev* dev, int64_t* mem_used, char* name_buf, int32_t name_len) {
    DEV_CALL(dev, {
        if (mem_used) *mem_used = int64_t(dev->d->mem_used());
        if (name_buf && name_len > 0) std::snprintf(name_buf, size_t(name_len), "%s", dev->d->name().c_str());
    });
}
int tpdhg_dev_set_profile(tpdhg_dev* dev, int32_t on) { DEV_CALL(dev, dev->d->set_profile(on != 0)); }
int tpdhg_dev_kernel_times(tpdhg_dev* dev, double* row_ms, double* col_ms, int64_t* row_launches,
                           int64_t* col_launches) {
    DEV_CALL(dev, {
        tpdhg::KernelTimes t = dev->d->kernel_times();
        if (row_ms) *row_ms = t.row_ms;
        if (col_ms) *col_ms = t.col_ms;
        if (row_launches) *row_launches = t.row_launches;
        if (col_launches) *col_launches = t.col_launches;
    });
}
void tpdhg_dev_destroy(tpdhg_dev* dev) { delete dev; }

int tpdhg_fast_create(const tpdhg_scaled_arrays* p, int32_t precision_bits, int32_t eval_freq, int32_t use_graph,
                      double reflection, tpdhg_fast** out) {
    if (!p || !out) return fail(TPDHG_ERR_ARG, "bad problem");
    if (const char* e = problem_error(p)) return fail(TPDHG_ERR_ARG, std::string("bad problem: ") + e);
    return guarded([&]() -> int {
        tpdhg::Scaled S;  // owned copy of the caller's arrays
        S.m = p->m, S.n = p->n;
        const size_t m = size_t(p->m), n = size_t(p->n), z = size_t(p->nnz);
        if (p->a_ptr[p->m] != p->nnz || p->at_ptr[p->n] != p->nnz) return fail(TPDHG_ERR_ARG, "CSR pointers do not match nnz");
        S.A.rows = p->m, S.A.ptr.assign(p->a_ptr, p->a_ptr + m + 1), S.A.idx.assign(p->a_idx, p->a_idx + z);
        S.A.val.assign(p->a_val, p->a_val + z);
        S.AT.rows = p->n, S.AT.ptr.assign(p->at_ptr, p->at_ptr + n + 1), S.AT.idx.assign(p->at_idx, p->at_idx + z);
        S.AT.val.assign(p->at_val, p->at_val + z);
        S.c.assign(p->c, p->c + n), S.l.assign(p->l, p->l + n), S.u.assign(p->u, p->u + n);
        S.rl.assign(p->rl, p->rl + m), S.ru.assign(p->ru, p->ru + m), S.Dr.assign(p->Dr, p->Dr + m);
        S.Dc.assign(p->Dc, p->Dc + n), S.bnorm = p->bnorm, S.cnorm = p->cnorm;
        tpdhg::FastOptions o;
        o.precision_bits = precision_bits, o.eval_freq = eval_freq, o.use_graph = use_graph != 0, o.rho = reflection;
        auto f = std::make_unique<tpdhg_fast>();
        f->d = tpdhg::make_fast_device(S, o);
        *out = f.release();
        return TPDHG_OK;
    });
}

#define FAST_CALL(f, body)                                                  \
    do {                                                                    \
        if (!(f) || !(f)->d) return fail(TPDHG_ERR_ARG, "null fast handle"); \
        return guarded([&]() -> int { body; return TPDHG_OK; });            \
    } while (0)

int tpdhg_fast_power_norm(tpdhg_fast* f, int32_t iters, double* norm) {
    if (!norm || iters < 1) return fail(TPDHG_ERR_ARG, "null norm or iters < 1");
    FAST_CALL(f, *norm = f->d->power_norm(iters));
}
int tpdhg_fast_set_base(tpdhg_fast* f, const double* x_b, const double* z_b) { FAST_CALL(f, f->d->set_base(x_b, z_b)); }
int tpdhg_fast_begin_round(tpdhg_fast* f) { FAST_CALL(f, f->d->begin_round()); }
int tpdhg_fast_set_steps(tpdhg_fast* f, double tau, double sigma) { FAST_CALL(f, f->d->set_steps(tau, sigma)); }
int tpdhg_fast_restart(tpdhg_fast* f) { FAST_CALL(f, f->d->restart()); }
int tpdhg_fast_iterate(tpdhg_fast* f, int32_t steps) { FAST_CALL(f, f->d->iterate(steps)); }
int tpdhg_fast_eval(tpdhg_fast* f, double out[10]) {
    if (!out) return fail(TPDHG_ERR_ARG, "null output");
    FAST_CALL(f, f->d->eval(out));
}
int tpdhg_fast_fixed_point(tpdhg_fast* f, double out[3]) {
    if (!out) return fail(TPDHG_ERR_ARG, "null output");
    FAST_CALL(f, f->d->fixed_point(out));
}
int tpdhg_fast_end_round(tpdhg_fast* f) { FAST_CALL(f, f->d->end_round()); }
int tpdhg_fast_download(tpdhg_fast* f, double* x, double* z) { FAST_CALL(f, f->d->download(x, z)); }
void tpdhg_fast_destroy(tpdhg_fast* f) { delete f; }

void tpdhg_options_default(tpdhg_options* o) {
    if (!o) return;
    tpdhg::Options d;
    o->tol = d.tol, o->time_limit = d.time_limit, o->max_iter = d.max_iter, o->eval_freq = d.eval_freq;
    o->row_rel = d.row_rel, o->precision_bits = d.precision_bits, o->refine = d.refine ? 1 : 0;
    o->inner_tol = 0, o->max_refinements = d.max_refinements, o->use_graph = d.graph ? 1 : 0;
    o->verbose = 0, o->profile_kernels = 0;
    o->halpern = d.halpern ? 1 : 0, o->reflection = d.rho;
    o->primal_weight = d.omega0, o->warm_fallback = d.warm_fallback ? 1 : 0, o->progress_interval = d.progress_s;
    o->omega_rule = d.omega_rule;
    o->reuse_norm = d.reuse_norm ? 1 : 0;
    o->cpu_threads = d.cpu_threads;
    o->restart_metric = d.restart_metric;
}

namespace {
void fill_result(const tpdhg::Report& r, const tpdhg::LP& lp, tpdhg_result* res, double* x, double* y) {
    std::memset(res, 0, sizeof(*res));
    res->status = status_code(r.out.status);
    res->iterations = r.out.iterations, res->restarts = r.out.restarts, res->refinements = r.out.refinements;
    res->from_average = r.out.from_average;
    res->objective = r.objective;
    res->rel_primal = r.out.rel_primal, res->rel_dual = r.out.rel_dual, res->rel_gap = r.out.rel_gap;
    res->max_row_violation = r.max_row_violation, res->max_bound_violation = r.max_bound_violation;
    res->setup_s = r.setup_s, res->parse_s = r.tm.parse, res->scale_s = r.tm.scale;
    res->device_init_s = r.tm.init, res->transfer_s = r.tm.transfer, res->solve_s = r.solve_s;
    res->power_iter_s = r.out.power_s, res->total_s = r.total_s;
    res->norm_a = r.out.norm_a, res->primal_weight = r.out.omega;
    res->device_mem_bytes = int64_t(r.device_mem_bytes);
    res->m = lp.m, res->n = lp.n, res->nnz = r.nnz;
    res->row_kernel_ms = r.out.kernels.row_ms, res->col_kernel_ms = r.out.kernels.col_ms;
    res->row_launches = r.out.kernels.row_launches, res->col_launches = r.out.kernels.col_launches;
    std::snprintf(res->device, sizeof res->device, "%s", r.device_name.c_str());
    res->warm_start_used = r.out.warm_used ? 1 : 0;
    res->warm_kkt = r.out.warm_kkt, res->cold_kkt = r.out.cold_kkt;
    res->threads = r.threads;
    if (x) std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
    if (y) std::memcpy(y, r.y.data(), r.y.size() * sizeof(double));
}

int check_options(const tpdhg::Options& o) {
    if (o.precision_bits != 32 && o.precision_bits != 64) return fail(TPDHG_ERR_ARG, "precision_bits must be 32 or 64");
    if (o.refine && !o.halpern) return fail(TPDHG_ERR_ARG, "refine runs on the Halpern path: set halpern = 1");
    if (o.halpern && !(o.rho >= 1 && o.rho < 2)) return fail(TPDHG_ERR_ARG, "reflection must be in [1, 2)");
    if (!(o.omega0 >= 0) || !std::isfinite(o.omega0)) return fail(TPDHG_ERR_ARG, "primal_weight must be finite and >= 0");
    if (o.omega_rule < 0 || o.omega_rule > 4) return fail(TPDHG_ERR_ARG, "omega_rule must be 0, 1, 2, 3 or 4");
    if (o.restart_metric < 0 || o.restart_metric > 1) return fail(TPDHG_ERR_ARG, "restart_metric must be 0 or 1");
    if (o.cpu_threads < -1 || o.cpu_threads > 4096) return fail(TPDHG_ERR_ARG, "cpu_threads must be -1 (GPU) or 0..4096");
    if (o.cpu_threads >= 0 && o.halpern)
        return fail(TPDHG_ERR_ARG, "the CPU backend runs the reference scheme: set halpern = 0 and refine = 0");
    return TPDHG_OK;
}
}  // namespace

int tpdhg_solve(const tpdhg_model* model, const tpdhg_options* opt, const double* x0, const double* y0,
                tpdhg_result* res, double* x, double* y) {
    if (!model || !res) return fail(TPDHG_ERR_ARG, "null argument");
    if (const char* why = tpdhg::input_error(model->lp)) return fail(TPDHG_ERR_UNSUPPORTED, why);
    tpdhg::Options o = to_options(opt);
    if (o.eval_freq < 1) o.eval_freq = 1;
    if (int rc = check_options(o)) return rc;
    return guarded([&]() -> int {
        const tpdhg::LP& lp = model->lp;
        tpdhg::Report r = tpdhg::solve_lp(lp, o, tpdhg::Timings{}, std::chrono::steady_clock::now(), x0, y0);
        fill_result(r, lp, res, x, y);
        return TPDHG_OK;
    });
}

int tpdhg_session_create(const tpdhg_model* model, const tpdhg_options* opt, tpdhg_session** out) {
    if (!model || !out) return fail(TPDHG_ERR_ARG, "null argument");
    if (const char* why = tpdhg::input_error(model->lp)) return fail(TPDHG_ERR_UNSUPPORTED, why);
    tpdhg::Options o = to_options(opt);
    if (o.eval_freq < 1) o.eval_freq = 1;
    if (int rc = check_options(o)) return rc;
    return guarded([&]() -> int {
        auto s = std::make_unique<tpdhg_session>();
        s->lp = model->lp;
        s->s = std::make_unique<tpdhg::Session>(s->lp, o);
        *out = s.release();
        return TPDHG_OK;
    });
}

int tpdhg_session_solve(tpdhg_session* s, const double* x0, const double* y0, tpdhg_result* res, double* x, double* y) {
    if (!s || !s->s || !res) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        tpdhg::Report r = s->s->solve(x0, y0);
        fill_result(r, s->lp, res, x, y);
        return TPDHG_OK;
    });
}

int tpdhg_session_set_primal_weight(tpdhg_session* s, double w) {
    if (!s) return fail(TPDHG_ERR_ARG, "null argument");
    if (!(w >= 0) || !std::isfinite(w)) return fail(TPDHG_ERR_ARG, "primal_weight must be finite and >= 0");
    s->s->set_omega0(w);
    return TPDHG_OK;
}

int tpdhg_session_set_vectors(tpdhg_session* s, const double* cost, const double* col_lo, const double* col_up,
                              const double* row_lo, const double* row_up) {
    if (!s || !s->s || !cost || !col_lo || !col_up || !row_lo || !row_up) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        s->s->set_vectors(s->lp, cost, col_lo, col_up, row_lo, row_up);
        return TPDHG_OK;
    });
}

void tpdhg_session_free(tpdhg_session* s) { delete s; }

int tpdhg_result_json(const tpdhg_model* model, const tpdhg_options* opt, const tpdhg_result* res, char* buf,
                      int32_t len) {
    if (!model || !res || !buf || len <= 0) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        tpdhg::Report r;
        r.out.status = tpdhg_status_string(res->status);
        r.out.iterations = res->iterations, r.out.restarts = res->restarts, r.out.refinements = res->refinements;
        r.out.from_average = res->from_average != 0;
        r.out.rel_primal = res->rel_primal, r.out.rel_dual = res->rel_dual, r.out.rel_gap = res->rel_gap;
        r.out.norm_a = res->norm_a, r.out.omega = res->primal_weight, r.out.power_s = res->power_iter_s;
        r.objective = res->objective, r.max_row_violation = res->max_row_violation;
        r.max_bound_violation = res->max_bound_violation;
        r.setup_s = res->setup_s, r.solve_s = res->solve_s, r.total_s = res->total_s;
        r.tm.parse = res->parse_s, r.tm.scale = res->scale_s, r.tm.init = res->device_init_s;
        r.tm.transfer = res->transfer_s;
        r.device_name = res->device, r.device_mem_bytes = size_t(res->device_mem_bytes), r.nnz = res->nnz;
        r.out.warm_used = res->warm_start_used != 0;
        r.threads = res->threads, r.backend = res->threads > 0 ? "cpu" : "gpu";
        std::string s = tpdhg::report_json(r, model->lp, to_options(opt));
        std::snprintf(buf, size_t(len), "%s", s.c_str());
        return TPDHG_OK;
    });
}

int tpdhg_recheck(const tpdhg_model* model, const double* x, const double* y, tpdhg_recheck_result* out) {
    if (!model || !x || !out) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        const tpdhg::Recheck c = tpdhg::recheck(model->lp, x, y);
        out->objective = c.objective, out->row_viol_rel = c.row_viol_rel, out->bound_viol = c.bound_viol;
        out->dual_checked = c.dual_checked, out->dual_bound_finite = c.dual_bound_finite, out->dual_sign = c.dual_sign;
        out->dual_bound = c.dual_bound, out->dual_obj_loose = c.dual_obj_loose;
        out->dual_infeas_rel = c.dual_infeas_rel, out->dual_gap_rel = c.dual_gap_rel;
        return TPDHG_OK;
    });
}

int tpdhg_batch_create(const tpdhg_model* const* models, int32_t count, const tpdhg_options* opt, int32_t cap,
                       tpdhg_batch** out) {
    if (!models || !out || count <= 0) return fail(TPDHG_ERR_ARG, "null argument or empty batch");
    tpdhg::Options o = to_options(opt);
    if (o.eval_freq < 1) o.eval_freq = 1;
    if (int rc = check_options(o)) return rc;
    if (!o.halpern) return fail(TPDHG_ERR_ARG, "batches run the fast path: set halpern = 1");
    if (cap != 0 && cap != 1024 && cap != 2048 && cap != 4096 && cap != 8192)
        return fail(TPDHG_ERR_ARG, "cap must be 0, 1024, 2048, 4096 or 8192");
    std::vector<const tpdhg::LP*> lps;
    for (int32_t i = 0; i < count; ++i) {
        if (!models[i]) return fail(TPDHG_ERR_ARG, "null model in batch");
        if (const char* why = tpdhg::input_error(models[i]->lp)) return fail(TPDHG_ERR_UNSUPPORTED, why);
        lps.push_back(&models[i]->lp);
    }
    return guarded([&]() -> int {
        auto b = std::make_unique<tpdhg_batch>();
        b->o = o;
        b->s = std::make_unique<tpdhg::BatchSession>(lps, o, cap);
        *out = b.release();
        return TPDHG_OK;
    });
}

int tpdhg_batch_get_info(const tpdhg_batch* b, tpdhg_batch_info* info) {
    if (!b || !info) return fail(TPDHG_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    tpdhg::BatchDevice& d = b->s->device();
    info->count = d.count(), info->unique_matrices = d.unique_matrices(), info->cap = d.cap();
    info->nnz_total = d.nnz_total(), info->nnz_stored = d.nnz_stored();
    info->device_mem_bytes = int64_t(d.mem_used()), info->setup_s = b->s->setup_s();
    const std::vector<float> t = d.tune_ms();
    for (size_t k = 0; k < t.size() && k < 4; ++k) info->tune_ms[k] = t[k];
    info->persist_max = d.persist_max();
    info->persistent_blocks = d.persistent_blocks(), info->lockstep_blocks = d.lockstep_blocks();
    info->stage_switches = d.stage_switches();
    return TPDHG_OK;
}

int tpdhg_batch_set_vectors(tpdhg_batch* b, int32_t i, const double* cost, const double* col_lo, const double* col_up,
                            const double* row_lo, const double* row_up) {
    if (!b || !cost || !col_lo || !col_up || !row_lo || !row_up) return fail(TPDHG_ERR_ARG, "null argument");
    if (i < 0 || i >= b->s->count()) return fail(TPDHG_ERR_ARG, "instance index out of range");
    return guarded([&]() -> int {
        b->s->set_vectors(i, cost, col_lo, col_up, row_lo, row_up);
        return TPDHG_OK;
    });
}

int tpdhg_batch_solve(tpdhg_batch* b, const double* const* x0, const double* const* y0, const double* primal_weight,
                      tpdhg_result* results, double* const* x, double* const* y) {
    if (!b || !results) return fail(TPDHG_ERR_ARG, "null argument");
    const int B = b->s->count();
    if (primal_weight)
        for (int i = 0; i < B; ++i)
            if (!(primal_weight[i] >= 0) || !std::isfinite(primal_weight[i]))
                return fail(TPDHG_ERR_ARG, "primal weights must be finite and >= 0");
    return guarded([&]() -> int {
        std::vector<const double*> xs, ys;
        std::vector<double> om;
        if (x0) xs.assign(x0, x0 + B);
        if (y0) ys.assign(y0, y0 + B);
        if (primal_weight) om.assign(primal_weight, primal_weight + B);
        std::vector<tpdhg::Report> reps = b->s->solve(xs, ys, om);
        for (int i = 0; i < B; ++i)
            fill_result(reps[size_t(i)], b->s->lp(i), &results[i], x ? x[i] : nullptr, y ? y[i] : nullptr);
        return TPDHG_OK;
    });
}

int tpdhg_batch_probe(tpdhg_batch* b, int32_t iters, double* row_ms, double* col_ms, int64_t* row_launches,
                      int64_t* col_launches) {
    if (!b || !row_ms || !col_ms || !row_launches || !col_launches) return fail(TPDHG_ERR_ARG, "null argument");
    if (iters < 1 || iters > 100000) return fail(TPDHG_ERR_ARG, "iters must be in [1, 100000]");
    return guarded([&]() -> int {
        const tpdhg::KernelTimes t = b->s->device().probe(iters);
        *row_ms = t.row_ms, *col_ms = t.col_ms, *row_launches = t.row_launches, *col_launches = t.col_launches;
        return TPDHG_OK;
    });
}

void tpdhg_batch_free(tpdhg_batch* b) { delete b; }

int tpdhg_mlp_create(int32_t n_in, int32_t h1, int32_t h2, int32_t n_out, uint64_t seed, tpdhg_mlp** out) {
    if (!out) return fail(TPDHG_ERR_ARG, "null argument");
    return guarded([&]() -> int {
        auto m = std::make_unique<tpdhg_mlp>();
        m->net = tpdhg::make_mlp(n_in, h1, h2, n_out, seed);
        *out = m.release();
        return TPDHG_OK;
    });
}
int64_t tpdhg_mlp_param_count(const tpdhg_mlp* net) { return net && net->net ? net->net->param_count() : -1; }
#define MLP_CALL(net, body)                                                  \
    do {                                                                     \
        if (!(net) || !(net)->net) return fail(TPDHG_ERR_ARG, "null network"); \
        return guarded([&]() -> int { body; return TPDHG_OK; });             \
    } while (0)
int tpdhg_mlp_set_params(tpdhg_mlp* net, const float* p) {
    if (!p) return fail(TPDHG_ERR_ARG, "null argument");
    MLP_CALL(net, net->net->set_params(p));
}
int tpdhg_mlp_get_params(tpdhg_mlp* net, float* p) {
    if (!p) return fail(TPDHG_ERR_ARG, "null argument");
    MLP_CALL(net, net->net->get_params(p));
}
int tpdhg_mlp_forward(tpdhg_mlp* net, const float* x, int32_t batch, float* y) {
    if (!x || !y || batch < 1) return fail(TPDHG_ERR_ARG, "null argument or batch < 1");
    MLP_CALL(net, net->net->forward(x, batch, y));
}
int tpdhg_mlp_loss_grad(tpdhg_mlp* net, const float* x, const float* t, int32_t batch, double* loss, float* grad) {
    if (!x || !t || batch < 1) return fail(TPDHG_ERR_ARG, "null argument or batch < 1");
    MLP_CALL(net, { const double l = net->net->loss_grad(x, t, batch, grad); if (loss) *loss = l; });
}
int tpdhg_mlp_train(tpdhg_mlp* net, const float* x, const float* t, int32_t n, int32_t batch, int32_t epochs, float lr,
                    uint64_t seed, double* loss) {
    if (!x || !t || n < 1 || batch < 1 || epochs < 0 || !(lr > 0))
        return fail(TPDHG_ERR_ARG, "null argument, or n, batch, epochs or lr out of range");
    MLP_CALL(net, { const double l = net->net->train(x, t, n, batch, epochs, lr, seed); if (loss) *loss = l; });
}
void tpdhg_mlp_destroy(tpdhg_mlp* net) { delete net; }

int tpdhg_spmv_bench(int32_t precision_bits, int32_t rows, int32_t cols, const int32_t* ptr, const int32_t* idx,
                     const double* val, const double* v, double* y, int32_t reps, double* mean_ms, double* median_ms) {
    if (rows < 0 || cols < 0 || !ptr || !v || !y || !mean_ms || !median_ms || reps < 1)
        return fail(TPDHG_ERR_ARG, "bad arguments");
    if (const char* e = csr_error(rows, cols, ptr, idx, val, -1)) return fail(TPDHG_ERR_ARG, std::string("bad CSR: ") + e);
    return guarded([&]() -> int {
        tpdhg::spmv_bench(precision_bits, rows, cols, ptr, idx, val, v, y, reps, mean_ms, median_ms);
        return TPDHG_OK;
    });
}

int tpdhg_spmv_noalloc(int32_t precision_bits, int32_t rows, int64_t nnz, const int32_t* ptr, const int32_t* idx,
                       const void* val, const void* v, void* y, void* stream) {
    if (precision_bits != 32 && precision_bits != 64) return fail(TPDHG_ERR_ARG, "precision must be 32 or 64 bits");
    if (rows < 0 || nnz < 0) return fail(TPDHG_ERR_ARG, "rows and nnz must be >= 0");
    if (nnz > tpdhg::spmv_noalloc_max_nnz())
        return fail(TPDHG_ERR_ARG, "nnz exceeds the " + std::to_string(tpdhg::spmv_noalloc_max_nnz()) +
                                       " nonzeros the allocation-free SpMV's static scratch array covers");
    if (rows > 0 && (!ptr || !y || (nnz > 0 && (!idx || !val || !v))))
        return fail(TPDHG_ERR_ARG, "null device array");
    return guarded([&]() -> int {
        tpdhg::spmv_noalloc(precision_bits, rows, nnz, ptr, idx, val, v, y, static_cast<hipStream_t>(stream));
        return TPDHG_OK;
    });
}

int64_t tpdhg_spmv_noalloc_max_nnz(void) { return tpdhg::spmv_noalloc_max_nnz(); }

int tpdhg_stream_bench(int64_t bytes, int32_t reps, double* read_gbs, double* copy_gbs) {
    if (!read_gbs || !copy_gbs || reps < 1 || reps > 1000 || bytes < (int64_t(1) << 20) || bytes > (int64_t(64) << 30))
        return fail(TPDHG_ERR_ARG, "bytes must be in [1 MiB, 64 GiB], reps in [1, 1000], outputs non-null");
    return guarded([&]() -> int {
        tpdhg::stream_bench(bytes, reps, read_gbs, copy_gbs);
        return TPDHG_OK;
    });
}

}  // extern "C"
