// kh_server.hip -- continuous batching of independent SPM runs on one GPU.
//
// Clients submit runs (parameters, particles, step count, a host buffer for
// the per-step trajectory) from any thread; one scheduler thread owns the
// device side. It admits queued runs while a slot is free and their device
// footprint fits the HBM budget and keeps every admitted run resident until
// it completes. Runs of plain spheres of one size class share a batch
// (kh_batch.h): a run is created in a free slot of a batch of its class (a
// new batch of up to batch_slots arenas when none has room), takes its first
// step alone, then joins, and the batch steps all its joined runs a few
// steps at a time, one launch per fluid pass for all of them, as one
// captured graph per membership. Other runs step alone on their own streams
// (after the first step each step is one captured-graph launch). Batches of
// different sizes and lone runs overlap on the GPU. A run's
// trajectory is recorded on the device every step and copied to the
// client's buffer once, when the run completes; so are its Output_udf frames
// when the client asked for them. Latency is measured from submission to
// that copy.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/kapsel_hip.h"
#include "kh_batch.h"

int kh_set_error(int status, const char *msg);

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Request {
    long long                  id = 0;
    kh_params                  P{};
    std::vector<kh_species>    species;
    std::vector<Particle>      parts;
    int                        nsteps = 0;
    double                    *traj_out = nullptr;
    int                        frame_every = 0;   // Output_udf frames every this many steps (0: none)
    int                        nframes     = 0;
    double                    *frames_out  = nullptr;
    long long                  bytes    = 0;
    long long                  grid     = 0;
    clk::time_point            t_submit, t_admit, t_done;
    int                        status = 1;  // 1 queued, 2 active, 0 done, < 0 failed (kh_status)
    std::string                error;
    // device side (scheduler thread only)
    kh_run   *run      = nullptr;
    int       enqueued = 0;
    kh_batch *batch    = nullptr;  // the batch it was created in (nullptr: steps alone)
    int       slot     = -1;
    bool      joined   = false;
};

}  // namespace

struct kh_server {
    kh_server_config                       cfg{};
    int                                    device = 0;
    long long                              budget = 0;
    std::mutex                             mu;
    std::condition_variable                cv;
    std::deque<Request *>                  queue;
    std::map<long long, std::unique_ptr<Request>> all;
    std::vector<Request *>                 active;  // scheduler thread only
    std::vector<std::pair<kh_batch *, long long>> batches;  // with their arena bytes (scheduler thread only)
    long long                              used = 0, peak_bytes = 0, next_id = 1;
    int                                    peak_active = 0, nactive = 0;
    bool                                   stop = false, started = false;
    clk::time_point                        t_first, t_last;
    std::vector<double>                    latency;  // seconds, completed runs
    long long                              done_steps = 0, failed = 0;
    double                                 done_points = 0;
    std::thread                            worker;
};

namespace {

void finish(kh_server *S, Request *q, int status, const char *err) {
    if (q->run) {
        if (status == 0 && q->traj_out && kh_get_trajectory(q->run, q->traj_out, q->nsteps) != KH_OK) {
            status = KH_ERR_HIP;
            err    = kh_last_error();
        }
        if (status == 0 && q->frames_out && kh_get_output(q->run, q->frames_out, q->nframes) != KH_OK) {
            status = KH_ERR_HIP;
            err    = kh_last_error();
        }
        kh_run_destroy(q->run);
        q->run = nullptr;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    q->t_done = clk::now();
    q->status = status;
    if (err) q->error = err;
    S->used -= q->bytes;
    S->nactive--;
    if (status == 0) {
        S->latency.push_back(secs(q->t_submit, q->t_done));
        S->done_steps += q->nsteps;
        S->done_points += (double)q->grid * q->nsteps;
        S->t_last = q->t_done;
    } else {
        S->failed++;
    }
    S->cv.notify_all();
}

// Device side of admission: create the run (in its batch slot, if any),
// upload, initial constraint block; a batched run takes its first step alone.
int start(Request *q) {
    const int np = (int)q->parts.size();
    int       st = q->batch ? kh::batch_run_create(q->batch, q->slot, &q->P, np, &q->run)
                            : kh_run_create(&q->P, np, &q->run);
    if (st == KH_OK) st = kh_set_particles(q->run, q->parts.data(), np);
    if (st == KH_OK) st = kh_init_fluid(q->run);
    if (st == KH_OK) st = kh_set_trajectory(q->run, q->nsteps);
    if (st == KH_OK && q->frames_out) st = kh_set_output(q->run, q->nframes, q->frame_every);
    if (st == KH_OK && q->batch) {
        st = kh_step(q->run, 1);
        if (st == KH_OK) q->enqueued = 1;
    }
    return st;
}

// A run leaves its batch (if it was created in one) before its results are
// copied and it is destroyed; the slot is free afterwards.
int leave(Request *q) {
    if (!q->batch || !q->run) return KH_OK;
    const int st = kh::batch_leave(q->batch, q->run);
    q->batch     = nullptr;
    return st;
}

bool batchable(const kh_server *S, const Request *q) {
    if (S->cfg.batch_slots < 2 || (int)q->parts.size() > KH_BATCH_NP_CAP) return false;
    for (const kh_species &sp : q->species)
        if (sp.propulsion == KH_JP_SLIP) return false;
    return true;
}

void loop(kh_server *S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        bool worked = false;
        // admission: a free slot and room in the HBM budget, in arrival order
        for (;;) {
            Request  *q = nullptr;
            kh_batch *b = nullptr;
            int       slot = -1, new_slots = 0;
            long long arena = 0;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                if (S->queue.empty() || (int)S->active.size() >= S->cfg.max_active) break;
                Request  *h     = S->queue.front();
                const int np    = (int)h->parts.size();
                long long need  = h->bytes;
                // device bytes beyond the run's own buffers (trajectory, frames)
                const long long extra = h->bytes - kh_run_bytes(&h->P, np);
                if (batchable(S, h)) {
                    for (auto &e : S->batches)
                        if (kh::batch_accepts(e.first, h->P, np) && (slot = kh::batch_free_slot(e.first)) >= 0) {
                            b = e.first;
                            break;
                        }
                    if (b) {
                        need = extra;
                    } else {  // a new batch, as many slots as the budget leaves (two at least)
                        const long long stride = (long long)kh::batch_stride(h->P, KH_BATCH_NP_CAP);
                        const long long room   = S->budget - S->used - extra;
                        new_slots              = (int)std::min<long long>(S->cfg.batch_slots, room / stride);
                        if (new_slots >= 2) {
                            arena = stride * new_slots;
                            need  = arena + extra;
                            slot  = 0;
                        } else {
                            new_slots = 0;
                        }
                    }
                }
                if (S->used + need > S->budget && !S->active.empty()) break;
                S->queue.pop_front();
                S->used += need;
                h->bytes = (b || new_slots) ? extra : h->bytes;  // what finish() releases
                S->nactive++;
                S->peak_bytes  = std::max(S->peak_bytes, S->used);
                S->peak_active = std::max(S->peak_active, S->nactive);
                h->status      = 2;
                h->t_admit     = clk::now();
                q              = h;
            }
            if (new_slots) {
                const int bst = kh::batch_create(q->P, KH_BATCH_NP_CAP, new_slots, &b);
                if (bst != KH_OK) {
                    const std::string e = kh_last_error();
                    {
                        std::lock_guard<std::mutex> lk(S->mu);
                        S->used -= arena;
                    }
                    finish(S, q, bst, e.c_str());
                    worked = true;
                    continue;
                }
                S->batches.push_back({b, arena});
            }
            q->batch  = b;
            q->slot   = slot;
            const int st = start(q);
            if (st != KH_OK) {
                const std::string e = kh_last_error();
                leave(q);
                finish(S, q, st, e.c_str());
            } else {
                S->active.push_back(q);
            }
            worked = true;
        }
        // progress: every lone run whose last chunk has drained gets the next one
        for (size_t i = 0; i < S->active.size();) {
            Request *q = S->active[i];
            if (q->batch) {
                i++;
                continue;
            }
            const int idle = q->enqueued ? kh_query(q->run) : 1;
            if (idle == 0) {
                i++;
                continue;
            }
            if (idle < 0 || q->enqueued == q->nsteps) {
                const std::string msg = idle < 0 ? kh_last_error() : "";
                finish(S, q, idle < 0 ? idle : 0, idle < 0 ? msg.c_str() : nullptr);
                S->active.erase(S->active.begin() + i);
                worked = true;
                continue;
            }
            const int k  = std::min(S->cfg.chunk_steps, q->nsteps - q->enqueued);
            const int st = kh_step(q->run, k);
            if (st != KH_OK) {
                const std::string msg = kh_last_error();
                finish(S, q, st, msg.c_str());
                S->active.erase(S->active.begin() + i);
                worked = true;
                continue;
            }
            q->enqueued += k;
            worked = true;
            i++;
        }
        // batches: once a batch's last chunk has drained, finished runs leave,
        // runs past their first step join, and the members get the next chunk
        for (size_t bi = 0; bi < S->batches.size();) {
            kh_batch *b    = S->batches[bi].first;
            const int idle = kh::batch_query(b);
            if (idle == 0) {
                bi++;
                continue;
            }
            std::vector<Request *> members;
            int                    nruns = 0;
            for (size_t i = 0; i < S->active.size();) {
                Request *q = S->active[i];
                if (q->batch != b) {
                    i++;
                    continue;
                }
                nruns++;
                int st = idle < 0 ? idle : KH_OK;
                if (st == KH_OK && !q->joined) {
                    const int own = kh_query(q->run);  // its first step, alone
                    if (own == 0) {
                        i++;
                        continue;
                    }
                    if (own < 0) st = own;
                    if (st == KH_OK && q->enqueued < q->nsteps) {
                        st = kh::batch_join(b, q->run);
                        if (st == KH_OK) q->joined = true;
                    }
                }
                if (st != KH_OK || q->enqueued == q->nsteps) {
                    std::string msg = st != KH_OK ? kh_last_error() : "";
                    const int   ls  = leave(q);
                    if (st == KH_OK && ls != KH_OK) {
                        st  = ls;
                        msg = kh_last_error();
                    }
                    finish(S, q, st, st != KH_OK ? msg.c_str() : nullptr);
                    S->active.erase(S->active.begin() + i);
                    nruns--;
                    worked = true;
                    continue;
                }
                if (q->joined) members.push_back(q);
                i++;
            }
            if (!members.empty()) {
                int k = S->cfg.chunk_steps;
                for (Request *q : members) k = std::min(k, q->nsteps - q->enqueued);
                const int st = kh::batch_step(b, k);
                if (st != KH_OK) {
                    const std::string msg = kh_last_error();
                    for (Request *q : members) {
                        leave(q);
                        finish(S, q, st, msg.c_str());
                        S->active.erase(std::find(S->active.begin(), S->active.end(), q));
                        nruns--;
                    }
                } else {
                    for (Request *q : members) q->enqueued += k;
                }
                worked = true;
            }
            if (nruns == 0) {  // every slot free: the arenas go back to the budget
                const long long arena = S->batches[bi].second;
                kh::batch_destroy(b);
                S->batches.erase(S->batches.begin() + bi);
                std::lock_guard<std::mutex> lk(S->mu);
                S->used -= arena;
                continue;
            }
            bi++;
        }
        if (!worked) {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->queue.empty() && S->active.empty()) break;
            if (S->active.empty())
                S->cv.wait_for(lk, std::chrono::milliseconds(5));
            else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }
}

}  // namespace

extern "C" {

int kh_server_create(const kh_server_config *cfg, kh_server **out) {
    if (!cfg || !out) return kh_set_error(KH_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->chunk_steps < 1) return kh_set_error(KH_ERR_ARG, "max_active and chunk_steps must be >= 1");
    if (cfg->hbm_budget <= 0 && !(cfg->hbm_fraction > 0 && cfg->hbm_fraction <= 1))
        return kh_set_error(KH_ERR_ARG, "hbm_fraction must be in (0, 1] without an explicit budget");
    auto S = new kh_server();
    S->cfg = *cfg;
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return kh_set_error(KH_ERR_HIP, "no HIP device");
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        delete S;
        return kh_set_error(KH_ERR_HIP, "hipMemGetInfo failed");
    }
    S->budget = cfg->hbm_budget > 0 ? cfg->hbm_budget : (long long)(cfg->hbm_fraction * (double)fr);
    S->worker = std::thread(loop, S);
    *out      = S;
    return KH_OK;
}

static int submit(kh_server *S, const kh_params *P, const Particle *parts, int np, int nsteps, double *traj_out,
                  int frame_every, double *frames_out, long long *id) {
    if (!S || !P || (!parts && np > 0) || !id) return kh_set_error(KH_ERR_ARG, "NULL argument");
    if (nsteps < 1) return kh_set_error(KH_ERR_ARG, "nsteps must be >= 1");
    if (frames_out && frame_every < 1) return kh_set_error(KH_ERR_ARG, "frame_every must be >= 1");
    const int nframes = frames_out ? nsteps / frame_every + 1 : 0;
    const long long bytes = kh_run_bytes(P, np) + (long long)sizeof(double) * 6 * nsteps * np +
                            (long long)sizeof(double) * nframes * (1 + (long long)KH_FRAME_WORDS * np);
    if (bytes < 0) return kh_set_error(KH_ERR_ARG, "bad run parameters");
    if (bytes > S->budget) return kh_set_error(KH_ERR_NOMEM, "run needs more device memory than the server's budget");
    auto q     = std::make_unique<Request>();
    q->P       = *P;
    q->species.assign(P->species, P->species + P->nspecies);
    q->P.species = q->species.data();
    q->parts.assign(parts, parts + np);
    q->nsteps      = nsteps;
    q->traj_out    = traj_out;
    q->frame_every = frames_out ? frame_every : 0;
    q->nframes     = nframes;
    q->frames_out  = frames_out;
    q->bytes       = bytes;
    q->grid     = (long long)P->nx * P->ny * P->nz;
    q->t_submit = clk::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return kh_set_error(KH_ERR_STATE, "server is shutting down");
    q->id = S->next_id++;
    *id   = q->id;
    if (!S->started) {
        S->started = true;
        S->t_first = q->t_submit;
    }
    S->queue.push_back(q.get());
    S->all[q->id] = std::move(q);
    S->cv.notify_all();
    return KH_OK;
}

int kh_server_submit(kh_server *S, const kh_params *P, const Particle *parts, int np, int nsteps, double *traj_out,
                     long long *id) {
    return submit(S, P, parts, np, nsteps, traj_out, 0, nullptr, id);
}

int kh_server_submit_output(kh_server *S, const kh_params *P, const Particle *parts, int np, int nsteps,
                            double *traj_out, int frame_every, double *frames_out, long long *id) {
    if (!frames_out) return kh_set_error(KH_ERR_ARG, "NULL frames_out");
    return submit(S, P, parts, np, nsteps, traj_out, frame_every, frames_out, id);
}

int kh_server_poll(kh_server *S, long long id) {
    if (!S) return kh_set_error(KH_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return kh_set_error(KH_ERR_ARG, "unknown run id");
    const Request *q = it->second.get();
    if (q->status < 0) return kh_set_error(q->status, q->error.c_str());
    return q->status == 0 ? 1 : 0;
}

int kh_server_wait(kh_server *S, long long id, double timeout_s) {
    if (!S) return kh_set_error(KH_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return kh_set_error(KH_ERR_ARG, "unknown run id");
    const Request *q  = it->second.get();
    const auto     dl = clk::now() + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(timeout_s));
    while (q->status > 0)
        if (S->cv.wait_until(lk, dl) == std::cv_status::timeout && q->status > 0) return 0;
    if (q->status < 0) return kh_set_error(q->status, q->error.c_str());
    return 1;
}

int kh_server_run_times(kh_server *S, long long id, double *queued_s, double *run_s) {
    if (!S || !queued_s || !run_s) return kh_set_error(KH_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return kh_set_error(KH_ERR_ARG, "unknown run id");
    const Request *q = it->second.get();
    if (q->status != 0) return kh_set_error(KH_ERR_STATE, "run not complete");
    *queued_s = secs(q->t_submit, q->t_admit);
    *run_s    = secs(q->t_admit, q->t_done);
    return KH_OK;
}

int kh_server_get_stats(kh_server *S, kh_server_stats *o) {
    if (!S || !o) return kh_set_error(KH_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    memset(o, 0, sizeof *o);
    o->submitted   = (long long)S->all.size();
    o->completed   = (long long)S->latency.size();
    o->failed      = S->failed;
    o->queued      = (long long)S->queue.size();
    o->active      = S->nactive;
    o->peak_active = S->peak_active;
    o->peak_bytes  = S->peak_bytes;
    o->budget      = S->budget;
    if (!S->latency.empty()) {
        std::vector<double> l = S->latency;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {  // nearest rank
            size_t k = (size_t)std::max(1.0, std::ceil(p * (double)l.size())) - 1;
            return 1e3 * l[std::min(k, l.size() - 1)];
        };
        double sum = 0;
        for (double x : l) sum += x;
        o->p50_ms  = pct(0.50);
        o->p99_ms  = pct(0.99);
        o->mean_ms = 1e3 * sum / (double)l.size();
        o->max_ms  = 1e3 * l.back();
        o->wall_s  = secs(S->t_first, S->t_last);
        if (o->wall_s > 0) {
            o->runs_per_s             = (double)l.size() / o->wall_s;
            o->steps_per_s            = (double)S->done_steps / o->wall_s;
            o->grid_point_steps_per_s = S->done_points / o->wall_s;
        }
    }
    return KH_OK;
}

int kh_server_destroy(kh_server *S) {
    if (!S) return KH_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    delete S;
    return KH_OK;
}

}  // extern "C"
