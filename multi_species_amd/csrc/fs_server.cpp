// Serving loop: continuous batching of simulation decks on one device.
//
// Requests (a deck and a number of steps) arrive from any host thread. A
// builder thread makes each deck's meshes on the host while the device keeps
// stepping the decks already resident. A worker thread owns the context and
// the batch (an fs_sim): between two steps it admits built decks in arrival
// order as long as a slot is free and the estimated device footprint of all
// resident decks stays within the HBM budget, then steps every resident deck
// at once (one batched field launch per RK4 stage, one remesh launch per
// kernel; a HIP graph replay for non-adaptive batches), and retires the decks
// that have taken their steps, keeping their final state on the host. A
// deck's state never leaves the device between its admission and retirement,
// and its results do not depend on which other decks share its steps.
//
// Every step's wall time (host submit to completion) is recorded: it is the
// per-step latency every resident deck sees. The statistics report its
// median and 99th percentile beside the aggregate point-steps per second.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fs_sim_internal.h"

namespace {

using clock_type = std::chrono::steady_clock;

double now_s() {
    return std::chrono::duration<double>(clock_type::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    fs_sim_params sim{};
    std::vector<fs_species_params> species;
    int n_steps = 0;
    // a deck resumed from the files of step restart_iter (-1: built from f0)
    int restart_iter = -1;
    std::string restart_dir;
    std::vector<std::string> names;
    int state = FS_REQ_QUEUED;
    int steps_done = 0;
    std::string error;
    fs::sim::Deck* built = nullptr;
    int64_t points = 0;
    size_t footprint = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<fs::sim::SpeciesState> result;
};

}  // namespace

struct fs_server {
    fs_ctx* ctx = nullptr;
    fs_server_params prm{};
    fs_sim* sim = nullptr;
    std::mutex mu;
    std::condition_variable cv;       // work for the builder or the worker
    std::condition_variable done_cv;  // a request finished
    std::deque<Request*> to_build, ready;
    std::vector<Request*> active;     // in the batch's deck order
    std::map<int64_t, std::unique_ptr<Request>> all;
    int64_t next_id = 1;
    bool stop = false;
    std::thread builder, worker;
    int building = 0;
    // statistics
    std::vector<double> step_ms;
    int64_t point_steps = 0;
    int64_t steps_replayed = 0;  // steps taken by a graph replay
    double busy_s = 0;
    double t_first = -1, t_last = -1;
    size_t hbm_in_use = 0;
    int64_t submitted = 0, completed = 0, failed = 0, steps = 0;
    std::string worker_error;
};

namespace {

void builder_loop(fs_server* S) {
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || !S->to_build.empty(); });
            if (S->stop) return;
            r = S->to_build.front();
            S->to_build.pop_front();
            r->state = FS_REQ_BUILDING;
            ++S->building;
        }
        fs_deck d{&r->sim, r->species.data(), (int)r->species.size()};
        fs::sim::Deck* D = nullptr;
        int rc;
        if (r->restart_iter >= 0) {
            std::vector<const char*> nm;
            for (const auto& s : r->names) nm.push_back(s.c_str());
            rc = fs::sim::build_deck_restart_host(&d, r->restart_dir.c_str(), nm.data(),
                                                  r->restart_iter, &D);
        } else {
            rc = fs::sim::build_deck_host(&d, &D);
        }
        std::string err = rc == FS_OK ? "" : fs_last_error();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            --S->building;
            if (rc != FS_OK) {
                r->state = FS_REQ_FAILED;
                r->error = err;
                r->t_done = now_s();
                ++S->failed;
                S->done_cv.notify_all();
            } else {
                r->built = D;
                r->points = fs::sim::deck_points(D);
                r->footprint = fs::sim::deck_footprint(D, S->prm.precision);
                if (r->footprint > (size_t)S->prm.hbm_budget_bytes) {
                    fs::sim::free_deck(D);
                    r->built = nullptr;
                    r->state = FS_REQ_FAILED;
                    r->error = "deck larger than the HBM budget";
                    r->t_done = now_s();
                    ++S->failed;
                    S->done_cv.notify_all();
                } else {
                    r->state = FS_REQ_READY;
                    S->ready.push_back(r);
                }
            }
        }
        S->cv.notify_all();
    }
}

void fail_all_active(fs_server* S, const std::string& why) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request* r : S->active) {
        r->state = FS_REQ_FAILED;
        r->error = why;
        r->t_done = now_s();
        ++S->failed;
    }
    S->active.clear();
    S->hbm_in_use = 0;
    S->worker_error = why;
    S->done_cv.notify_all();
}

void worker_loop(fs_server* S) {
    for (;;) {
        std::vector<Request*> admit;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || !S->ready.empty() || !S->active.empty(); });
            if (S->stop) return;
            // admission: arrival order, free slot, footprint within the budget
            while (!S->ready.empty()) {
                Request* r = S->ready.front();
                const int slots = S->prm.max_active > 0 ? S->prm.max_active : 1 << 30;
                if ((int)(S->active.size() + admit.size()) >= slots) break;
                if (S->hbm_in_use + r->footprint > (size_t)S->prm.hbm_budget_bytes) break;
                S->ready.pop_front();
                S->hbm_in_use += r->footprint;
                admit.push_back(r);
            }
        }
        if (!admit.empty()) {
            std::vector<fs::sim::Deck*> decks;
            for (Request* r : admit) {
                decks.push_back(r->built);
                r->built = nullptr;
            }
            const int rc = fs::sim::attach_decks(S->sim, decks, nullptr);
            const double t = now_s();
            std::lock_guard<std::mutex> lk(S->mu);
            for (Request* r : admit) {
                if (rc == FS_OK) {
                    r->state = FS_REQ_ACTIVE;
                    r->t_admit = t;
                    S->active.push_back(r);
                } else {
                    r->state = FS_REQ_FAILED;
                    r->error = fs_last_error();
                    r->t_done = t;
                    S->hbm_in_use -= r->footprint;
                    ++S->failed;
                }
            }
            if (rc != FS_OK) S->done_cv.notify_all();
            if (S->t_first < 0) S->t_first = t;
        }
        std::vector<Request*> act;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            act = S->active;
        }
        if (act.empty()) continue;
        const double t0 = now_s();
        const int rc = fs_sim_step(S->sim);
        const double t1 = now_s();
        if (rc != FS_OK) {
            fail_all_active(S, std::string("step failed: ") + fs_last_error());
            continue;
        }
        std::vector<int> fin;
        int64_t pts = 0;
        int64_t replayed = 0;
        (void)fs_sim_step_counts(S->sim, &replayed, nullptr);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->steps_replayed = replayed;
            S->step_ms.push_back((t1 - t0) * 1e3);
            S->busy_s += t1 - t0;
            ++S->steps;
            for (size_t k = 0; k < S->active.size(); ++k) {
                Request* r = S->active[k];
                r->steps_done += 1;
                pts += r->points;
                if (r->steps_done >= r->n_steps) fin.push_back((int)k);
            }
            S->point_steps += pts;
            S->t_last = t1;
        }
        if (fin.empty()) continue;
        std::vector<std::vector<fs::sim::SpeciesState>> out;
        const int rd = fs::sim::detach_decks(S->sim, fin, out);
        if (rd != FS_OK) {
            fail_all_active(S, std::string("retire failed: ") + fs_last_error());
            continue;
        }
        const double t2 = now_s();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            std::vector<Request*> keep;
            size_t q = 0;
            for (size_t k = 0; k < S->active.size(); ++k) {
                Request* r = S->active[k];
                if (q < fin.size() && fin[q] == (int)k) {
                    r->result = std::move(out[q]);
                    r->state = FS_REQ_DONE;
                    r->t_done = t2;
                    S->hbm_in_use -= r->footprint;
                    ++S->completed;
                    ++q;
                } else {
                    keep.push_back(r);
                }
            }
            S->active = std::move(keep);
            S->t_last = t2;
        }
        S->done_cv.notify_all();
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

using namespace fs;

extern "C" {

int fs_server_create(fs_ctx* ctx, const fs_server_params* prm, fs_server** out) {
    if (!ctx || !prm || !out) return fail(FS_E_INVALID, "fs_server_create: null argument");
    *out = nullptr;
    if (!(prm->hbm_budget_bytes > 0)) return fail(FS_E_INVALID, "fs_server_create: budget <= 0");
    auto S = std::make_unique<fs_server>();
    S->ctx = ctx;
    S->prm = *prm;
    FS_TRY(sim::create_empty(ctx, prm->precision, &S->sim));
    if (!prm->use_graph) FS_TRY(fs_sim_set_options(S->sim, prm->precision, 0));
    fs_server* raw = S.release();
    raw->builder = std::thread(builder_loop, raw);
    raw->worker = std::thread(worker_loop, raw);
    *out = raw;
    return FS_OK;
}

int fs_server_submit(fs_server* S, const fs_deck* deck, int n_steps, int64_t* id) {
    if (!S || !deck || !deck->sim || !deck->species || !id)
        return fail(FS_E_INVALID, "fs_server_submit: null argument");
    if (n_steps < 1 || deck->n_species < 1 || deck->n_species > 64)
        return fail(FS_E_INVALID, "fs_server_submit: n_steps >= 1 and 1..64 species");
    auto r = std::make_unique<Request>();
    r->sim = *deck->sim;
    r->species.assign(deck->species, deck->species + deck->n_species);
    r->n_steps = n_steps;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        *id = r->id;
        S->to_build.push_back(r.get());
        S->all[r->id] = std::move(r);
        ++S->submitted;
    }
    S->cv.notify_all();
    return FS_OK;
}

int fs_server_submit_restart(fs_server* S, const fs_deck* deck, const char* sim_dir,
                             const char* const* species_names, int iter, int n_steps,
                             int64_t* id) {
    if (!S || !deck || !deck->sim || !deck->species || !id || !species_names)
        return fail(FS_E_INVALID, "fs_server_submit_restart: null argument");
    if (n_steps < 1 || deck->n_species < 1 || deck->n_species > 64)
        return fail(FS_E_INVALID, "fs_server_submit_restart: n_steps >= 1 and 1..64 species");
    if (iter < 0) return fail(FS_E_INVALID, "fs_server_submit_restart: negative step");
    auto r = std::make_unique<Request>();
    r->sim = *deck->sim;
    r->species.assign(deck->species, deck->species + deck->n_species);
    for (int k = 0; k < deck->n_species; ++k) {
        if (!species_names[k])
            return fail(FS_E_INVALID, "fs_server_submit_restart: null species name");
        r->names.emplace_back(species_names[k]);
    }
    r->restart_iter = iter;
    r->restart_dir = sim_dir ? sim_dir : "";
    r->n_steps = n_steps;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        *id = r->id;
        S->to_build.push_back(r.get());
        S->all[r->id] = std::move(r);
        ++S->submitted;
    }
    S->cv.notify_all();
    return FS_OK;
}

int fs_server_write_result(fs_server* S, int64_t id, const char* sim_dir,
                           const char* const* species_names) {
    if (!S || !species_names) return fail(FS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(FS_E_INVALID, "unknown request %lld", (long long)id);
    Request* r = it->second.get();
    if (r->state != FS_REQ_DONE) return fail(FS_E_INVALID, "request %lld not done", (long long)id);
    for (size_t k = 0; k < r->result.size(); ++k) {
        if (!species_names[k]) return fail(FS_E_INVALID, "null species name");
        const sim::SpeciesState& st = r->result[k];
        FS_TRY(io::write_species(sim_dir, species_names[k], st.iter, (int64_t)st.xs.size(),
                                 st.xs.data(), st.ps.data(), st.fs.data(), st.qw.data(),
                                 st.es.data(), (int64_t)(st.rec.size() / 20), st.rec.data()));
    }
    return FS_OK;
}

int fs_server_status(fs_server* S, int64_t id, int* state, int* steps_done) {
    if (!S) return fail(FS_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(FS_E_INVALID, "unknown request %lld", (long long)id);
    if (state) *state = it->second->state;
    if (steps_done) *steps_done = it->second->steps_done;
    return FS_OK;
}

int fs_server_wait(fs_server* S, int64_t id, double timeout_s) {
    if (!S) return fail(FS_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(FS_E_INVALID, "unknown request %lld", (long long)id);
    Request* r = it->second.get();
    auto done = [&] { return r->state == FS_REQ_DONE || r->state == FS_REQ_FAILED; };
    if (timeout_s > 0) {
        if (!S->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
            return fail(FS_E_RANGE, "request %lld not done within %.1f s", (long long)id, timeout_s);
    } else {
        S->done_cv.wait(lk, done);
    }
    if (r->state == FS_REQ_FAILED)
        return fail(FS_E_INTERNAL, "request %lld failed: %s", (long long)id, r->error.c_str());
    return FS_OK;
}

int fs_server_result(fs_server* S, int64_t id, int species, int64_t* n_points, double* xs,
                     double* ps, double* fs, double* q_ws, double* es, double diag[4]) {
    if (!S) return fail(FS_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(FS_E_INVALID, "unknown request %lld", (long long)id);
    Request* r = it->second.get();
    if (r->state != FS_REQ_DONE) return fail(FS_E_INVALID, "request %lld not done", (long long)id);
    if (species < 0 || species >= (int)r->result.size())
        return fail(FS_E_INVALID, "bad species index");
    const sim::SpeciesState& st = r->result[species];
    if (n_points) *n_points = (int64_t)st.xs.size();
    if (xs) std::copy(st.xs.begin(), st.xs.end(), xs);
    if (ps) std::copy(st.ps.begin(), st.ps.end(), ps);
    if (fs) std::copy(st.fs.begin(), st.fs.end(), fs);
    if (q_ws) std::copy(st.qw.begin(), st.qw.end(), q_ws);
    if (es) std::copy(st.es.begin(), st.es.end(), es);
    if (diag) std::copy(st.diag, st.diag + 4, diag);
    return FS_OK;
}

int fs_server_get_stats(fs_server* S, fs_server_stats* out) {
    if (!S || !out) return fail(FS_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    fs_server_stats st{};
    st.submitted = S->submitted;
    st.completed = S->completed;
    st.failed = S->failed;
    st.active = (int64_t)S->active.size();
    st.queued = (int64_t)(S->to_build.size() + S->ready.size()) + S->building;
    st.steps = S->steps;
    st.point_steps = S->point_steps;
    st.busy_s = S->busy_s;
    st.span_s = (S->t_first >= 0 && S->t_last > S->t_first) ? S->t_last - S->t_first : 0.0;
    st.point_steps_per_s = st.span_s > 0 ? (double)S->point_steps / st.span_s : 0.0;
    st.step_p50_ms = percentile(S->step_ms, 50);
    st.step_p99_ms = percentile(S->step_ms, 99);
    st.hbm_in_use = (int64_t)S->hbm_in_use;
    st.steps_replayed = S->steps_replayed;
    *out = st;
    return FS_OK;
}

int fs_server_destroy(fs_server* S) {
    if (!S) return FS_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->builder.joinable()) S->builder.join();
    if (S->worker.joinable()) S->worker.join();
    for (auto& kv : S->all)
        if (kv.second->built) sim::free_deck(kv.second->built);
    fs_sim_destroy(S->sim);
    delete S;
    return FS_OK;
}

}  // extern "C"
