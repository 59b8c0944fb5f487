// Serving loop: a queue of independent OT solve requests answered by two lanes.
//
//  * 2D lane (one host thread, one BatchEngine = one HIP stream): continuous
//    batching at Newton-iteration granularity.  Between two ticks it admits
//    queued requests (FIFO, while the HBM budget allows), drops the ones that
//    finished and relayouts the resident set on the device; each tick advances
//    every resident request by one Newton iteration (one batched cell pass, one
//    batched PCG).  A request's weights are bitwise the same as alone: the
//    batch's grids, reductions and tie offsets are per problem.
//  * 3D lane (workers_3d host threads, each with its own device handle and
//    stream): one request per worker at a time through the single-problem
//    Newton solve, concurrent with the 2D batch on the device.
// Admission counts an estimated device footprint per request (bytes per dirac of
// its lane) plus a fixed per-worker reservation for the 3D handles' scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>

#include "admission.h"
#include "batch_engine.h"

using namespace sdot_amd;

namespace {

using Clock = std::chrono::steady_clock;

// rows at the wide capacity, pass outputs, PCG vectors, warm-start neighbours (128 B),
// the smoothers' FP32 copy of the level-0 values (512 B)
constexpr double BYTES_PER_DIRAC_3D = 4262.0;
constexpr double HANDLE_3D_BYTES = 1.6e9;        // thread-kernel scratch slots of one 3D handle

struct Request {
    int64_t id = 0;
    int dim = 2;
    int64_t n = 0;
    std::vector<double> pos, w0, tgt, w;
    double box[6] = {0, 0, 0, 1, 1, 1};   // lo[dim], hi[dim] packed at [0..dim) and [dim..2 dim)
    double bytes = 0;
    int state = 0;   // 0 queued, 1 running, 2 finished
    sdot_srv_result res{};
};

}  // namespace

struct sdot_srv {
    sdot_srv_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> q2, q3;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_id = 1;
    double resident = 0, peak = 0, budget = 0;
    int64_t running = 0, completed = 0, cells = 0, ticks = 0;
    bool stop = false;
    Clock::time_point t0;
    std::thread lane2;
    std::vector<std::thread> lane3;
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void finish(Request *r, int status, int iters, double err_rel, const double *w_src) {
        // caller holds mu
        r->res.status = status;
        r->res.newton_iters = iters;
        r->res.max_rel_mass_err = err_rel;
        r->res.t_done = now();
        if (w_src) r->w.assign(w_src, w_src + r->n);
        r->state = 2;
        resident -= r->bytes;
        --running;
        ++completed;
        cv_done.notify_all();
        cv_work.notify_all();   // budget freed
    }
};

namespace {

void lane2_main(sdot_srv *S) {
    std::unique_ptr<BatchEngine> eng;
    std::vector<Request *> resident;   // parallel to eng->probs
    std::vector<int> finished;
    TickParams tp;
    tp.tol = S->cfg.tol;
    tp.stop_rule = S->cfg.stop_rule;
    tp.max_iter = S->cfg.max_iter;
    bool layout_dirty = false;
    std::vector<int> keep;
    for (;;) {
        std::vector<Request *> add;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) return;
                // admission: FIFO while the footprint fits (a request always starts on an idle device)
                while (!S->q2.empty() && (int)(resident.size() + add.size()) < S->cfg.max_batch_problems) {
                    Request *r = S->q2.front();
                    if (!srv_admit(S->resident, r->bytes, S->budget, S->running)) break;
                    S->q2.pop_front();
                    S->resident += r->bytes;
                    S->peak = std::max(S->peak, S->resident);
                    r->state = 1;
                    r->res.t_start = S->now();
                    ++S->running;
                    add.push_back(r);
                }
                if (!resident.empty() || !add.empty()) break;
                S->cv_work.wait(lk);
            }
        }
        try {
            // no graph capture in the lanes: three host threads share the device
            if (!eng) eng.reset(new BatchEngine(S->cfg.precision, 0, S->cfg.diracs_per_cell_x100, false));
            if (layout_dirty || !add.empty()) {
                // every resident request still running stays, in its current order
                keep.clear();
                for (int k = 0; k < (int)resident.size(); ++k)
                    if (resident[k] && !eng->probs[k].done) keep.push_back(k);
                std::vector<NewProblem> np;
                for (Request *r : add) {
                    NewProblem q;
                    q.n = r->n;
                    q.pos = r->pos.data();
                    q.w0 = r->w0.data();
                    q.tgt = r->tgt.data();
                    q.box[0] = r->box[0];
                    q.box[1] = r->box[1];
                    q.box[2] = r->box[2];
                    q.box[3] = r->box[3];
                    q.tag = r->id;
                    np.push_back(q);
                }
                eng->relayout(keep, np);
                std::vector<Request *> nr;
                for (int k : keep) nr.push_back(resident[k]);
                for (Request *r : add) nr.push_back(r);
                resident.swap(nr);
                add.clear();   // now owned through `resident`
                layout_dirty = false;
            }
            finished.clear();
            TickStats ts;
            eng->tick(tp, finished, ts);
            std::vector<std::vector<double>> wout(finished.size());
            for (size_t f = 0; f < finished.size(); ++f) {
                wout[f].resize(eng->probs[finished[f]].n);
                eng->read_weights(finished[f], wout[f].data());
            }
            std::lock_guard<std::mutex> lk(S->mu);
            S->cells += ts.cells;
            ++S->ticks;
            for (size_t f = 0; f < finished.size(); ++f) {
                const BatchProblem &b = eng->probs[finished[f]];
                S->finish(resident[finished[f]], b.status, b.iters, b.err, wout[f].data());
                resident[finished[f]] = nullptr;   // the waiter may free it from now on
            }
            if (!finished.empty()) layout_dirty = true;   // drop them at the next tick boundary
        } catch (...) {
            // a failure of the batch fails every resident request; the engine starts over
            sdot_status st = SDOT_HIP_ERROR;
            std::string msg = "2D lane failure";
            try {
                throw;
            } catch (const StatusError &e) {
                st = e.st;
                msg = e.msg;
            } catch (const HipError &e) {
                msg = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.err);
            } catch (const OomError &) {
                st = SDOT_OOM;
                msg = "device allocation failed";
            } catch (const std::string &e) {
                st = SDOT_CAPACITY_OVERFLOW;
                msg = e;
            } catch (...) {
            }
            std::lock_guard<std::mutex> lk(S->mu);
            S->err = msg;
            for (Request *r : resident)
                if (r && r->state == 1) S->finish(r, st, 0, INFINITY, nullptr);
            for (Request *r : add)
                if (r->state == 1) S->finish(r, st, 0, INFINITY, nullptr);
            resident.clear();
            keep.clear();
            layout_dirty = false;
            eng.reset();
        }
    }
}

void lane3_main(sdot_srv *S) {
    sdot_pd *pd = nullptr;
    bool reserved = false;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) break;
                if (!S->q3.empty()) {
                    Request *h = S->q3.front();
                    const double need = h->bytes + (reserved ? 0.0 : HANDLE_3D_BYTES);
                    if (srv_admit(S->resident, need, S->budget, S->running)) {
                        S->q3.pop_front();
                        S->resident += need;
                        if (!reserved) reserved = true;
                        S->peak = std::max(S->peak, S->resident);
                        h->state = 1;
                        h->res.t_start = S->now();
                        ++S->running;
                        r = h;
                        break;
                    }
                }
                S->cv_work.wait(lk);
            }
        }
        if (!r) break;
        int status = SDOT_OK;
        sdot_solve_stats st{};
        std::vector<double> w = r->w0;
        if (!pd) {
            status = sdot_pd_create(&pd, 3, S->cfg.precision);
            if (status == SDOT_OK) pd->use_graph = 0;   // see the 2D lane
        }
        if (status == SDOT_OK) status = sdot_pd_clear_shapes(pd);
        if (status == SDOT_OK) status = sdot_pd_add_box_shape(pd, r->box, r->box + 3, 1.0, -1);
        if (status == SDOT_OK) status = sdot_pd_set_diracs(pd, r->pos.data(), r->w0.data(), r->n);
        if (status == SDOT_OK)
            status = sdot_ot_solve(pd, r->tgt.data(), S->cfg.tol, S->cfg.stop_rule, S->cfg.max_iter, w.data(), &st);
        std::lock_guard<std::mutex> lk(S->mu);
        if (status != SDOT_OK && status != SDOT_NOT_CONVERGED && pd) S->err = sdot_pd_last_error(pd);
        S->cells += (int64_t)st.newton_iters * r->n;
        S->finish(r, status, st.newton_iters, st.max_rel_mass_err,
                  status == SDOT_OK || status == SDOT_NOT_CONVERGED ? w.data() : nullptr);
    }
    if (pd) sdot_pd_destroy(pd);
    if (reserved) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->resident -= HANDLE_3D_BYTES;
    }
}

}  // namespace

extern "C" {

sdot_status sdot_srv_create(sdot_srv **out, const sdot_srv_config *cfg) {
    if (!out) return SDOT_INVALID_ARG;
    *out = nullptr;
    sdot_srv_config c = cfg ? *cfg : sdot_srv_config{};
    if (c.precision != SDOT_FP64 && c.precision != SDOT_MIXED) return SDOT_INVALID_ARG;
    if (c.hbm_budget_bytes < 0 || c.max_batch_problems < 0 || c.workers_3d < 0 || c.diracs_per_cell_x100 < 0 ||
        c.tol < 0 || (c.stop_rule != 0 && c.stop_rule != 1) || c.max_iter < 0)
        return SDOT_INVALID_ARG;
    if (c.max_batch_problems == 0) c.max_batch_problems = 4096;
    if (c.workers_3d == 0) c.workers_3d = 2;
    if (c.diracs_per_cell_x100 == 0) c.diracs_per_cell_x100 = 250;
    if (c.tol == 0) c.tol = 1e-6;
    if (c.max_iter == 0) c.max_iter = 150;
    sdot_srv *S = new (std::nothrow) sdot_srv;
    if (!S) return SDOT_OOM;
    S->cfg = c;
    S->budget = c.hbm_budget_bytes;
    if (S->budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            delete S;
            return SDOT_HIP_ERROR;
        }
        S->budget = 0.8 * (double)fr;
    }
    S->t0 = Clock::now();
    try {
        S->lane2 = std::thread(lane2_main, S);
        for (int k = 0; k < c.workers_3d; ++k) S->lane3.emplace_back(lane3_main, S);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv_work.notify_all();
        if (S->lane2.joinable()) S->lane2.join();
        for (auto &t : S->lane3) t.join();
        delete S;
        return SDOT_OOM;
    }
    *out = S;
    return SDOT_OK;
}

sdot_status sdot_srv_submit(sdot_srv *S, int dim, int64_t n, const double *positions, const double *weights,
                            const double *targets, const double *box, int64_t *request_id) {
    if (!S || (dim != 2 && dim != 3) || n < 1 || n > 0x7FFFFFFF || !positions || !targets || !request_id)
        return SDOT_INVALID_ARG;
    for (int64_t i = 0; i < n; ++i) {
        if (!(targets[i] > 0) || (weights && !std::isfinite(weights[i]))) return SDOT_INVALID_ARG;
        for (int d = 0; d < dim; ++d)
            if (!std::isfinite(positions[dim * i + d])) return SDOT_INVALID_ARG;
    }
    if (box)
        for (int d = 0; d < dim; ++d)
            if (!(box[dim + d] > box[d])) return SDOT_INVALID_ARG;
    std::unique_ptr<Request> r(new (std::nothrow) Request);
    if (!r) return SDOT_OOM;
    try {
        r->dim = dim;
        r->n = n;
        r->pos.assign(positions, positions + dim * n);
        r->w0 = weights ? std::vector<double>(weights, weights + n) : std::vector<double>(n, 0.0);
        r->tgt.assign(targets, targets + n);
    } catch (const std::bad_alloc &) {
        return SDOT_OOM;
    }
    for (int d = 0; d < 3; ++d) {
        r->box[d] = (box && d < dim) ? box[d] : 0.0;
        r->box[3 + d] = (box && d < dim) ? box[dim + d] : 1.0;
    }
    if (dim == 2) {   // BatchProblem box layout: x0, y0, x1, y1
        const double b4[4] = {r->box[0], r->box[1], r->box[3], r->box[4]};
        std::copy(b4, b4 + 4, r->box);
    }
    r->bytes = (double)n * (dim == 2 ? BatchEngine::bytes_per_dirac() : BYTES_PER_DIRAC_3D);
    std::lock_guard<std::mutex> lk(S->mu);
    if (srv_never_fits(r->bytes, dim == 3 ? HANDLE_3D_BYTES : 0.0, S->budget)) {
        S->err = "request footprint exceeds the HBM budget";
        return SDOT_OOM;
    }
    r->id = S->next_id++;
    r->res.t_submit = S->now();
    *request_id = r->id;
    Request *p = r.get();
    S->reqs[r->id] = std::move(r);
    (dim == 2 ? S->q2 : S->q3).push_back(p);
    S->cv_work.notify_all();
    return SDOT_OK;
}

sdot_status sdot_srv_poll(sdot_srv *S, int64_t id, int *state) {
    if (!S || !state) return SDOT_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return SDOT_INVALID_ARG;
    *state = it->second->state;
    return SDOT_OK;
}

sdot_status sdot_srv_wait(sdot_srv *S, int64_t id, double timeout_s, double *weights_out, sdot_srv_result *res) {
    if (!S) return SDOT_INVALID_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return SDOT_INVALID_ARG;
    Request *r = it->second.get();
    auto ready = [&] { return r->state == 2; };
    if (timeout_s < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return SDOT_PENDING;
    if (weights_out && !r->w.empty()) std::copy(r->w.begin(), r->w.end(), weights_out);
    if (res) *res = r->res;
    S->reqs.erase(it);
    return SDOT_OK;
}

sdot_status sdot_srv_get_stats(sdot_srv *S, sdot_srv_stats *st) {
    if (!S || !st) return SDOT_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    st->queued_2d = (int64_t)S->q2.size();
    st->queued_3d = (int64_t)S->q3.size();
    st->running = S->running;
    st->completed = S->completed;
    st->resident_bytes = S->resident;
    st->peak_resident_bytes = S->peak;
    st->budget_bytes = S->budget;
    st->cells_evaluated = S->cells;
    st->ticks_2d = S->ticks;
    st->seconds = S->now();
    return SDOT_OK;
}

const char *sdot_srv_last_error(const sdot_srv *S) { return S ? S->err.c_str() : "null handle"; }

sdot_status sdot_srv_destroy(sdot_srv *S) {
    if (!S) return SDOT_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->lane2.joinable()) S->lane2.join();
    for (auto &t : S->lane3)
        if (t.joinable()) t.join();
    delete S;
    return SDOT_OK;
}

}  // extern "C"
