// Serving form: a continuous-batching request server on one device-resident model.
//
// Requests (periodic or open structures of any size) are queued as they arrive.  `lanes`
// worker threads each own a batch workspace and a HIP stream on the shared model; a free
// lane takes the queued requests in arrival order up to its share of the HBM budget and its
// atom limit, evaluates them as one packed batch (neighbour list, descriptors, networks,
// forces), writes each result into the caller's buffers and marks the request done.  While
// one lane's batch is on the GPU, requests that arrive meanwhile are admitted by the next free
// lane, so batches overlap on the device.  A request whose estimated footprint exceeds a
// lane's budget is rejected at submission (NNP_EBUDGET).  Every kernel is per atom and every
// reduction is per structure in a fixed order, so a result does not depend on the batch it
// was evaluated in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nnp_capi.h"
#include "engine.hpp"

namespace atomnn {

static double steady_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id;
    int natoms;
    std::vector<int32_t> type;
    std::vector<double> pos;
    double cell[9];
    bool periodic;
    nnp_result* out;
    double t_submit = 0, t_done = 0;
    size_t est_bytes = 0;
    int status = NNP_EAGAIN;
    std::string error;
};

class Server {
  public:
    Server(Engine& owner, const nnp_server_cfg& cfg) : owner_(owner), cfg_(cfg) {
        if (cfg_.lanes <= 0) cfg_.lanes = 3;
        if (cfg_.hbm_budget <= 0) cfg_.hbm_budget = (int64_t)16 << 30;
        if (cfg_.max_batch_atoms <= 0) cfg_.max_batch_atoms = 1 << 18;
        if (cfg_.max_batch_structures <= 0) cfg_.max_batch_structures = 4096;
        lane_budget_ = (size_t)(cfg_.hbm_budget / cfg_.lanes);
        for (int k = 0; k < cfg_.lanes; ++k) {
            lanes_.emplace_back(new Engine(owner_, k));
            lanes_.back()->may_tune = false;
            lanes_.back()->defer_count = true;
        }
        for (int k = 0; k < cfg_.lanes; ++k) workers_.emplace_back(&Server::worker, this, k);
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_queue_.notify_all();
        for (auto& t : workers_) t.join();
    }

    // Estimated device bytes of a batch of `atoms` atoms with about `p` entries per atom:
    // neighbour rows (40 B), force terms (32 B) and per-atom arrays (descriptor and dE/dG rows, the
    // moment table of S NE 35 values).
    size_t estimate(size_t atoms, double p) const {
        const size_t maxp = (size_t)std::ceil(p * 1.5 + 16);
        const int ne = owner_.shape == 0 ? 2 : (owner_.shape == 1 ? 4 : 8);  // exponents of the grid shape
        const size_t mom = (size_t)owner_.pb.plan.T * ne * 35 * 8;           // FP64 bound
        return atoms * (maxp * 72 + (size_t)owner_.pb.plan.ldg * 16 + mom + 128) + (1 << 20);
    }

    int64_t submit(const nnp_structure& s, nnp_result* out) {
        if (s.natoms <= 0 || !s.type || !s.pos || !out) throw std::invalid_argument("nnp_server_submit: invalid structure");
        auto r = std::make_unique<Request>();
        r->natoms = s.natoms;
        r->type.assign(s.type, s.type + s.natoms);
        for (int t : r->type)
            if (t < 0 || t >= owner_.host.T) throw std::invalid_argument("atom type out of range");
        r->pos.assign(s.pos, s.pos + 3 * (size_t)s.natoms);
        r->periodic = s.cell != nullptr;
        if (s.cell) std::memcpy(r->cell, s.cell, sizeof(r->cell));
        else std::memset(r->cell, 0, sizeof(r->cell));
        double p = s.natoms;
        if (s.cell) {
            const double* a = s.cell;
            const double vol = std::fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                                         a[2] * (a[3] * a[7] - a[4] * a[6]));
            if (vol > 0) p = s.natoms / vol * 4.0 / 3.0 * M_PI * std::pow(owner_.host.rc, 3);
        }
        r->est_bytes = estimate(s.natoms, p);
        if (r->est_bytes > lane_budget_) throw budget_error("request does not fit the HBM budget of a serving lane");
        r->out = out;
        r->t_submit = steady_now();
        std::lock_guard<std::mutex> g(mu_);
        r->id = next_id_++;
        const int64_t id = r->id;
        queue_.push_back(r.get());
        reqs_[id] = std::move(r);
        cv_queue_.notify_one();
        return id;
    }

    int poll(int64_t id) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) throw std::invalid_argument("nnp_server: unknown request id");
        return it->second->status;
    }

    int wait(int64_t id, double* t_submit, double* t_done) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) throw std::invalid_argument("nnp_server: unknown request id");
        Request* r = it->second.get();
        cv_done_.wait(lk, [&] { return r->status != NNP_EAGAIN; });
        if (t_submit) *t_submit = r->t_submit;
        if (t_done) *t_done = r->t_done;
        const int st = r->status;
        if (st < 0) set_error(r->error);
        reqs_.erase(it);  // a request is waited for once
        return st;
    }

    void stats(int64_t* out, int n) {
        std::lock_guard<std::mutex> g(mu_);
        size_t bytes = 0;
        for (auto& l : lanes_) bytes += l->workspace_bytes();
        const int64_t v[] = {completed_, batches_, max_concurrent_, (int64_t)bytes, (int64_t)peak_bytes_,
                             (int64_t)queue_.size()};
        for (int k = 0; k < n && k < (int)(sizeof(v) / sizeof(v[0])); ++k) out[k] = v[k];
    }

    struct budget_error : std::runtime_error {
        using std::runtime_error::runtime_error;
    };

  private:
    void worker(int k) {
        Engine& E = *lanes_[k];
        for (;;) {
            std::vector<Request*> batch;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_queue_.wait(lk, [&] { return stop_ || !queue_.empty(); });
                if (stop_ && queue_.empty()) return;
                size_t atoms = 0, bytes = 0;
                while (!queue_.empty() && (int)batch.size() < cfg_.max_batch_structures) {
                    Request* r = queue_.front();
                    if (!batch.empty() && (atoms + r->natoms > (size_t)cfg_.max_batch_atoms ||
                                           bytes + r->est_bytes > lane_budget_))
                        break;
                    batch.push_back(r);
                    atoms += r->natoms;
                    bytes += r->est_bytes;
                    queue_.pop_front();
                }
                ++running_;
                max_concurrent_ = std::max<int64_t>(max_concurrent_, running_);
            }
            std::vector<nnp_structure> ss;
            std::vector<nnp_result> res;
            for (Request* r : batch) {
                ss.push_back(nnp_structure{r->natoms, r->type.data(), r->pos.data(), r->periodic ? r->cell : nullptr});
                res.push_back(*r->out);
            }
            int status = NNP_OK;
            std::string err;
            try {
                E.run(ss, cfg_.flags);
                E.fetch(ss, cfg_.flags, res.data());
            } catch (const std::bad_alloc&) {
                status = NNP_ENOMEM;
                err = "out of device memory";
            } catch (const std::exception& e) {
                status = NNP_EDEVICE;
                err = e.what();
            }
            size_t held = E.workspace_bytes();
            // keep the lane inside its share of the budget; after a failed allocation start the next
            // batch from empty buffers (the device may still be short of memory)
            if (held > lane_budget_ || status == NNP_ENOMEM) {
                (void)hipStreamSynchronize(E.stream);
                for (DevMem* m : {&E.d_nbrj, &E.d_meta, &E.d_nbrr, &E.d_X, &E.d_G, &E.d_dEdG, &E.d_mom}) m->release();
                E.mom_fresh = 0;
            }
            const double t = steady_now();
            {
                std::lock_guard<std::mutex> g(mu_);
                size_t total = 0;
                for (auto& l : lanes_) total += l->workspace_bytes();
                peak_bytes_ = std::max(peak_bytes_, std::max(total, held));
                for (size_t q = 0; q < batch.size(); ++q) {
                    Request* r = batch[q];
                    r->out->energy = res[q].energy;
                    r->t_done = t;
                    r->status = status;
                    r->error = err;
                }
                completed_ += (int64_t)batch.size();
                ++batches_;
                --running_;
            }
            cv_done_.notify_all();
        }
    }

    Engine& owner_;
    nnp_server_cfg cfg_;
    size_t lane_budget_ = 0;
    std::vector<std::unique_ptr<Engine>> lanes_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_queue_, cv_done_;
    std::deque<Request*> queue_;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs_;
    int64_t next_id_ = 0, completed_ = 0, batches_ = 0, max_concurrent_ = 0, running_ = 0;
    size_t peak_bytes_ = 0;
    bool stop_ = false;
};

Engine& engine_of(nnp_model* m);

}  // namespace atomnn

using namespace atomnn;

struct nnp_server {
    Server* srv;
};

template <typename F>
static int srv_guard(F&& f) {
    try {
        set_error("");
        return f();
    } catch (const Server::budget_error& e) {
        set_error(e.what());
        return NNP_EBUDGET;
    } catch (const std::bad_alloc&) {
        set_error("out of memory");
        return NNP_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return NNP_EINVAL;
    }
}

extern "C" {

double nnp_server_now(void) { return steady_now(); }

int nnp_server_create(nnp_model* m, const nnp_server_cfg* cfg, nnp_server** out) {
    return srv_guard([&]() {
        if (!m || !out) throw std::invalid_argument("nnp_server_create: null argument");
        nnp_server_cfg c{};
        if (cfg) c = *cfg;
        *out = new nnp_server{new Server(engine_of(m), c)};
        return NNP_OK;
    });
}

int64_t nnp_server_submit(nnp_server* s, const nnp_structure* st, nnp_result* r) {
    int64_t id = -1;
    const int rc = srv_guard([&]() {
        if (!s || !st) throw std::invalid_argument("nnp_server_submit: null argument");
        id = s->srv->submit(*st, r);
        return NNP_OK;
    });
    return rc < 0 ? rc : id;
}

int nnp_server_poll(nnp_server* s, int64_t id) {
    return srv_guard([&]() {
        if (!s) throw std::invalid_argument("nnp_server_poll: null server");
        return s->srv->poll(id);
    });
}

int nnp_server_wait(nnp_server* s, int64_t id, double* t_submit, double* t_done) {
    int st = NNP_OK;
    const int rc = srv_guard([&]() {
        if (!s) throw std::invalid_argument("nnp_server_wait: null server");
        st = s->srv->wait(id, t_submit, t_done);
        return NNP_OK;
    });
    return rc < 0 ? rc : st;
}

int nnp_server_stats(nnp_server* s, int64_t* out, int n) {
    return srv_guard([&]() {
        if (!s || !out) throw std::invalid_argument("nnp_server_stats: invalid argument");
        s->srv->stats(out, n);
        return NNP_OK;
    });
}

void nnp_server_destroy(nnp_server* s) {
    if (!s) return;
    delete s->srv;
    delete s;
}

}  // extern "C"
