// server.hip - a request loop over the kriging / SGS entry points with
// continuous batching: jobs wait in arrival order and a free worker slot
// starts the first waiting job whose estimated HBM footprint fits the
// unused part of the budget, so jobs that arrive while others run are
// admitted as soon as a slot and memory free up, and their work overlaps on
// the device (every worker thread has its own HIP stream).  A job larger
// than the whole budget runs when nothing else is running.  The head of the
// queue can be overtaken by smaller jobs only until it has waited
// kHeadPatienceMs; after that it is admitted next.
//
// Jobs run through the library's host-pointer entry points, so a result is
// the one the same call returns alone.  Covariance tables and templates are
// shared through the library's resident cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpgl.h"
#include "engine.h"

using namespace hpgl_amd;

namespace {

constexpr double kHeadPatienceMs = 200.0;

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

long long cells(const hpgl_shape_t &s) { return (long long)s.m_data[0] * s.m_data[1] * s.m_data[2]; }

long long box(const int r[3]) { return (2ll * r[0] + 1) * (2ll * r[1] + 1) * (2ll * r[2] + 1); }

// Device bytes a job may hold at once: its arrays, neighbour lists and
// solve state per target (including the solve's bucket sort and hand-over
// lists, about 20 bytes), bins, and the covariance / rank tables
// (cell-binned rank table over twice the radii), with 25% headroom.
long long footprint(const hpgl_amd_job_t &j) {
    const long long N = cells(j.shape);
    auto search = [&](const int r[3], int max_n, long long T) {
        return T * (4ll * max_n + 44) + N * 12 + box(r) * 8 * 10;
    };
    long long b = 0;
    switch (j.kind) {
        case HPGL_AMD_JOB_OK: {
            const auto *p = (const hpgl_ok_params_t *)j.params;
            b = N * 10 + search(p->m_radiuses, p->m_max_neighbours, N) + box(p->m_radiuses) * 8 * 64;
            break;
        }
        case HPGL_AMD_JOB_SK: {
            const auto *p = (const hpgl_sk_params_t *)j.params;
            b = N * 10 + search(p->m_radiuses, p->m_max_neighbours, N);
            break;
        }
        case HPGL_AMD_JOB_LVM: {
            const auto *p = (const hpgl_ok_params_t *)j.params;
            b = N * 14 + search(p->m_radiuses, p->m_max_neighbours, N);
            break;
        }
        case HPGL_AMD_JOB_IK: {
            const auto *p = (const hpgl_ik_params_t *)j.params;
            long long s = 0;
            for (int k = 0; k < j.indicator_count; ++k)
                s = std::max(s, search(p[k].m_radiuses, p[k].m_max_neighbours, N));
            b = N * 4 + N * 4ll * j.indicator_count + s;
            break;
        }
        case HPGL_AMD_JOB_MEDIAN_IK: {
            const auto *p = (const hpgl_median_ik_params_t *)j.params;
            b = N * 4 + search(p->m_radiuses, p->m_max_neighbours, N);
            break;
        }
        case HPGL_AMD_JOB_SGS: {
            const auto *p = (const hpgl_sgs_params_t *)j.params;
            b = N * (30 + 12ll * p->m_max_neighbours + 32) + box(p->m_radiuses) * 8;
            break;
        }
        default:
            throw std::runtime_error("unknown job kind");
    }
    return b + b / 4;
}

struct Job {
    long long id;
    hpgl_amd_job_t desc;
    long long bytes;
    double t_submit = 0, t_start = 0, t_end = 0;
    bool done = false;
    int status = 0;
    long long nodes = 0;
    std::string message;
};

// Runs one job on the calling thread's stream through the C entry points.
void run_job(Job &job) {
    hpgl_amd_job_t &j = job.desc;
    switch (j.kind) {
        case HPGL_AMD_JOB_OK: {
            hpgl_cont_masked_array_t in{j.cont_data, j.cont_mask, j.shape};
            hpgl_cont_masked_array_t out{j.out_cont, j.out_cont_mask, j.shape};
            hpgl_ordinary_kriging(&in, (hpgl_ok_params_t *)j.params, &out);
            break;
        }
        case HPGL_AMD_JOB_SK:
            hpgl_simple_kriging(j.cont_data, j.cont_mask, &j.shape, (hpgl_sk_params_t *)j.params, j.out_cont,
                                j.out_cont_mask, &j.shape);
            break;
        case HPGL_AMD_JOB_LVM:
            hpgl_lvm_kriging(j.cont_data, j.cont_mask, &j.shape, j.mean, &j.shape, (hpgl_ok_params_t *)j.params,
                             j.out_cont, j.out_cont_mask, &j.shape);
            break;
        case HPGL_AMD_JOB_IK: {
            hpgl_ind_masked_array_t in{j.ind_data, j.ind_mask, j.shape, j.indicator_count};
            hpgl_ind_masked_array_t out{j.out_ind, j.out_ind_mask, j.shape, j.indicator_count};
            hpgl_indicator_kriging(&in, &out, (hpgl_ik_params_t *)j.params, j.indicator_count);
            break;
        }
        case HPGL_AMD_JOB_MEDIAN_IK: {
            hpgl_ind_masked_array_t in{j.ind_data, j.ind_mask, j.shape, 2};
            hpgl_ind_masked_array_t out{j.out_ind, j.out_ind_mask, j.shape, 2};
            hpgl_median_ik(&in, (hpgl_median_ik_params_t *)j.params, &out);
            break;
        }
        case HPGL_AMD_JOB_SGS: {
            hpgl_cont_masked_array_t prop{j.cont_data, j.cont_mask, j.shape};
            hpgl_ubyte_array_t sm{j.sim_mask, j.shape};
            if (j.mean) {
                hpgl_float_array_t mean{j.mean, j.shape};
                hpgl_sgs_lvm_simulation(&prop, (hpgl_sgs_params_t *)j.params, j.cdf, &mean, j.sim_mask ? &sm : nullptr);
            } else {
                hpgl_sgs_simulation(&prop, (hpgl_sgs_params_t *)j.params, j.cdf, nullptr, j.sim_mask ? &sm : nullptr);
            }
            break;
        }
        default:
            break;
    }
    job.status = hpgl_amd_get_last_error();
    if (job.status) job.message = hpgl_get_last_exception_message();
    hpgl_amd_stats_t st;
    hpgl_amd_get_last_stats(&st);
    job.nodes = st.nodes_target;
}

struct Server {
    double t_created = now_ms();
    int device = 0;
    long long budget = 0;
    long long in_use = 0;
    int running = 0;
    bool stop = false;
    long long next_id = 1;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<Job *> pending;
    std::map<long long, Job *> jobs;
    std::vector<std::thread> workers;

    // first waiting job that fits, unless the head has waited too long
    Job *pick() {
        if (pending.empty()) return nullptr;
        Job *head = pending.front();
        const bool head_fits = head->bytes <= budget - in_use || running == 0;
        if (head_fits) {
            pending.pop_front();
            return head;
        }
        if (now_ms() - head->t_submit > kHeadPatienceMs) return nullptr;
        for (auto it = pending.begin() + 1; it != pending.end(); ++it)
            if ((*it)->bytes <= budget - in_use) {
                Job *j = *it;
                pending.erase(it);
                return j;
            }
        return nullptr;
    }

    void worker() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            Job *j = nullptr;
            cv_sched.wait(lk, [&] { return stop || (j = pick()) != nullptr; });
            if (!j) return;  // stop requested with nothing admissible left
            in_use += j->bytes;
            ++running;
            j->t_start = now_ms();
            lk.unlock();
            run_job(*j);
            lk.lock();
            j->t_end = now_ms();
            in_use -= j->bytes;
            --running;
            j->done = true;
            cv_done.notify_all();
            cv_sched.notify_all();
        }
    }
};

}  // namespace

#define API extern "C" __attribute__((visibility("default")))

API void *hpgl_amd_server_create(int slots, long long hbm_budget_bytes) {
    try {
        if (slots < 1) throw std::runtime_error("at least one worker slot is needed");
        auto *s = new Server;
        hip_check(hipGetDevice(&s->device), "hipGetDevice");
        if (hbm_budget_bytes <= 0) {
            size_t free_b = 0, total_b = 0;
            hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            hbm_budget_bytes = (long long)(free_b / 2);
        }
        s->budget = hbm_budget_bytes;
        for (int i = 0; i < slots; ++i) s->workers.emplace_back([s] { s->worker(); });
        return s;
    } catch (const std::exception &e) {
        report_error(e.what());
        return nullptr;
    }
}

API long long hpgl_amd_server_submit(void *server, const hpgl_amd_job_t *job) {
    try {
        auto *s = (Server *)server;
        if (!s || !job) throw std::runtime_error("null server or job");
        auto *j = new Job;
        j->desc = *job;
        j->bytes = footprint(*job);
        std::lock_guard<std::mutex> lk(s->mu);
        j->id = s->next_id++;
        j->t_submit = now_ms();
        s->jobs[j->id] = j;
        s->pending.push_back(j);
        s->cv_sched.notify_all();
        return j->id;
    } catch (const std::exception &e) {
        report_error(e.what());
        return -1;
    }
}

API int hpgl_amd_server_wait(void *server, long long id, hpgl_amd_job_result_t *result) {
    auto *s = (Server *)server;
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return -1;
    Job *j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    if (result) {
        std::memset(result, 0, sizeof(*result));
        result->status = j->status;
        result->queue_ms = j->t_start - j->t_submit;
        result->latency_ms = j->t_end - j->t_submit;
        result->run_ms = j->t_end - j->t_start;
        result->nodes = j->nodes;
        result->hbm_bytes = j->bytes;
        result->start_ms = j->t_start - s->t_created;
        result->end_ms = j->t_end - s->t_created;
        std::strncpy(result->message, j->message.c_str(), sizeof(result->message) - 1);
    }
    s->jobs.erase(it);
    delete j;
    return 0;
}

API void hpgl_amd_server_destroy(void *server) {
    auto *s = (Server *)server;
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // let every submitted job finish first
        s->cv_done.wait(lk, [&] { return s->pending.empty() && s->running == 0; });
        s->stop = true;
        s->cv_sched.notify_all();
    }
    for (auto &t : s->workers) t.join();
    for (auto &kv : s->jobs) delete kv.second;
    delete s;
}
