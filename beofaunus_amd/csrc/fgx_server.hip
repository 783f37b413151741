// fgx_server.hip -- serving independent simulation jobs on one device.
//
// A job is one Faunus simulation (a Space, its Hamiltonian, one registered
// move, a seed) and a number of Propagator::move() calls (sweeps).  Submitted
// jobs wait in a FIFO queue on the host; fgx_server_step admits queued jobs
// while their estimated device footprint fits the HBM budget (creating their
// device state), advances every resident job by one tick of batched sweeps,
// and retires finished jobs, releasing their memory so that later arrivals
// are admitted.  Jobs of one kernel kind (move, precision, Hamiltonian form)
// share one launch; launches of different kinds run on different streams at
// the same time.  Each job gets a trial budget per tick from its own
// measured trial rate, so every launch lasts about one tick whatever the
// job sizes: a large job's sweep spans several ticks (the kernel resumes an
// unfinished Propagator::move() call), a small job runs several sweeps per
// tick, and arrivals join at the next tick (continuous batching).  A large
// job whose chain runs the block AtomicTranslation sweep is spread over
// several workgroups of one XCD, as many as its size alone fixes.  A job's
// result does not depend on the other jobs it shares launches with: its
// workgroups and launch shape are its own, and every reduction has a fixed
// order.  A job the device sweeps cannot run fails at admission, alone.
//
// The reference runs one simulation per process (src/examples/bulk.cpp) and
// has no counterpart to this component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "fgx_internal.hpp"

using namespace fgx;

namespace {

double now_s()
{
    using clk = std::chrono::steady_clock;
    static const clk::time_point t0 = clk::now();
    return std::chrono::duration<double>(clk::now() - t0).count();
}

struct JobSpec {
    fgx_hamiltonian h{};
    std::vector<double> tsig, teps, tq, tr, tdp;
    double box[3]{};
    int n = 0, ng = 0;
    std::vector<double> xyz, q, r, mw;
    std::vector<int> type, front, back, molsize, molid;
    int move_type = 0, move_molid = 0, peratom = 0, permol = 0, center_rotation = 0;
    double prob = 1, dp = 0, dprot = 0;
    double dir[3] = {1, 1, 1}, dir2[3] = {1, 1, 1};
    uint32_t seed = 0;
    int sweeps = 0;
    int keep_positions = 0;
};

struct Job {
    int id = 0;
    JobSpec spec;
    fgx_system *sys = nullptr;
    size_t est_bytes = 0, bytes = 0;
    int done = 0;                 // sweeps completed
    int state = 0;                // 0 queued, 1 resident, 2 finished, 3 failed
    double t_submit = 0, t_admit = 0, t_done = 0;
    double chain_sum = 0;          // device seconds over all launches
    double trial_max = 0;          // slowest launch, seconds per trial
    double trial_est = 0;          // current estimate, seconds per trial
    // block-sweep workgroups of the job's launches, fixed at admission (0:
    // one-trial sweep); a job with more than one runs only in ticks whose
    // XCD packing has room for it, and ran holds the last such tick
    int parts = 0;
    long ran = -1;
    fgx_stats st{};
    std::vector<double> final_xyz;
    int error = FGX_OK;
};

// device bytes fgx_create and a sweep allocate for a job, from its sizes
// (the padding, FP64/FP32 copies and Ewald tables of fgx_system)
size_t estimate_bytes(const JobSpec &j)
{
    const size_t n = (size_t)j.n;
    size_t b = n * (6 * 8 + 3 * 8 + 2 * 4 + 16 + 2 * 4);
    b += (size_t)PARTNER_PAD * (16 + 4 + 4);
    b += (size_t)j.ng * (16 + 48 + 4) + 4 * (size_t)j.h.ntypes * j.h.ntypes * 12;
    b += 2 * 625 * 4 + 64 * 1024;   // RNG, reduction scratch and small buffers
    b += n * 3 * 8;                  // staging of coordinate transfers (final positions)
    if (j.h.elec == FGX_ELEC_EWALD) {
        const double kc = j.h.ewald_kc;
        const size_t K = (size_t)(2.5 * kc * kc * kc + 64);
        b += K * (3 * 8 + 8 + 3 * 4 + 2 * 16) + (size_t)j.h.ntypes * 8;
    }
    if (j.move_type == FGX_MOVE_TR && j.h.precision == FGX_MIXED && j.ng <= 256)
        b += (size_t)j.ng * j.ng * 8;
    return b;
}

// kernel kind: chains of one kind share a launch
long kind_of(const JobSpec &j)
{
    return (((((long)j.move_type * 4 + j.h.precision) * 8 + j.h.elec) * 8 + j.h.sr) * 4 +
            j.h.nbkind) * 2 + (j.h.lj_cut > 0 ? 1 : 0);
}

}  // namespace

// most workgroups one job's block sweeps take (of an XCD's 32 compute units
// on MI355X): more parts per chain gain little against the exchange each
// block pays between them
constexpr int SERVER_PARTS_MAX = 24;

struct fgx_server {
    size_t budget = 0, resident = 0, peak_resident = 0;
    int max_active = 0;
    int cu_per_xcd = 0;
    double tick = 0.05;
    int next_id = 1;
    std::deque<int> queue;
    std::map<int, std::unique_ptr<Job>> jobs;
    std::vector<int> finished;   // not yet polled
    long ticks = 0;
};

static int admit(fgx_server *sv, Job &J)
{
    const JobSpec &j = J.spec;
    fgx_system *s = nullptr;
    int rc = fgx_create(&j.h, j.box, j.n, j.xyz.data(), j.q.data(), j.r.data(), j.mw.data(),
                        j.type.data(), j.ng, j.front.data(), j.back.data(), j.molsize.data(),
                        j.molid.data(), &s);
    if (!rc)
        rc = fgx_add_move(s, j.move_type, j.move_molid, j.prob, j.peratom, j.permol, j.dir, j.dir2,
                          j.dp, j.dprot, j.center_rotation);
    if (!rc) rc = fgx_rng_seed(s, j.seed);
    if (!rc) rc = fgx_propagator_init(s);
    // one fixed schedule for every job, so a result does not depend on what
    // else is resident: the launch shape, and the block sweep's workgroups
    // from the job alone (jobs without a block sweep keep the one-trial one)
    if (!rc) rc = fgx_set_launch(s, 512, 4);
    if (!rc) {
        J.parts = blk_job_parts(s, std::min(SERVER_PARTS_MAX, sv->cu_per_xcd));
        rc = fgx_set_block_sweep(s, J.parts ? 1 : 0, std::max(1, J.parts));
    }
    // a job the device sweeps cannot run fails here, alone, instead of
    // failing the shared launch of its kind
    if (!rc) rc = check_chain(s);
    if (rc) {
        if (s) fgx_destroy(s);
        return rc;
    }
    J.sys = s;
    J.bytes = fgx_device_bytes(s);
    J.state = 1;
    J.t_admit = now_s();
    sv->resident += J.bytes;
    sv->peak_resident = std::max(sv->peak_resident, sv->resident);
    // first trial-time guess: 1e9 pair evaluations per second per CU
    J.trial_est = 2.0 * (double)j.n / (1e9 * std::max(1, J.parts)) + 2e-6;
    return FGX_OK;
}

static void retire(fgx_server *sv, Job &J, int error)
{
    if (J.sys) {
        if (!error && J.spec.keep_positions) {
            J.final_xyz.resize(3 * (size_t)J.spec.n);
            if (fgx_get_positions(J.sys, 0, J.final_xyz.data())) J.final_xyz.clear();
        }
        sv->resident -= std::min(sv->resident, J.bytes);
        fgx_destroy(J.sys);
        J.sys = nullptr;
    }
    J.state = error ? 3 : 2;
    J.error = error;
    J.t_done = now_s();
    J.spec.xyz.clear();
    J.spec.xyz.shrink_to_fit();
    sv->finished.push_back(J.id);
}

extern "C" int fgx_server_create(size_t hbm_budget, int max_active, double tick_seconds,
                                 fgx_server **out)
{
    FGX_CHECK(out && hbm_budget > 0 && max_active > 0 && tick_seconds > 0, FGX_EARG,
              "fgx_server_create: bad arguments");
    fgx_server *sv = new fgx_server;
    sv->budget = hbm_budget;
    sv->max_active = max_active;
    sv->tick = tick_seconds;
    // compute units of one XCD (eight per device): the room one job's split
    // block sweeps may take
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 0;
    sv->cu_per_xcd = std::max(1, ncu / 8);
    *out = sv;
    return FGX_OK;
}

extern "C" int fgx_server_submit(fgx_server *sv, const fgx_job *job, int *id)
{
    FGX_CHECK(sv && job && job->h && job->n > 0 && job->ng > 0 && job->xyz && job->charge &&
                  job->radius && job->mw && job->type && job->front && job->back &&
                  job->molsize && job->molid && job->sweeps >= 0,
              FGX_EARG, "fgx_server_submit: bad job");
    auto J = std::make_unique<Job>();
    JobSpec &j = J->spec;
    j.h = *job->h;
    const int T = j.h.ntypes;
    auto copyt = [&](const double *src, std::vector<double> &dst, const double **field) {
        if (src) dst.assign(src, src + T);
        *field = src ? dst.data() : nullptr;
    };
    copyt(job->h->type_sigma, j.tsig, &j.h.type_sigma);
    copyt(job->h->type_eps, j.teps, &j.h.type_eps);
    copyt(job->h->type_charge, j.tq, &j.h.type_charge);
    copyt(job->h->type_radius, j.tr, &j.h.type_radius);
    copyt(job->h->type_dp, j.tdp, &j.h.type_dp);
    std::memcpy(j.box, job->box, sizeof(j.box));
    j.n = job->n;
    j.ng = job->ng;
    j.xyz.assign(job->xyz, job->xyz + 3 * (size_t)j.n);
    j.q.assign(job->charge, job->charge + j.n);
    j.r.assign(job->radius, job->radius + j.n);
    j.mw.assign(job->mw, job->mw + j.n);
    j.type.assign(job->type, job->type + j.n);
    j.front.assign(job->front, job->front + j.ng);
    j.back.assign(job->back, job->back + j.ng);
    j.molsize.assign(job->molsize, job->molsize + j.ng);
    j.molid.assign(job->molid, job->molid + j.ng);
    j.move_type = job->move_type;
    j.move_molid = job->move_molid;
    j.prob = job->prob;
    j.peratom = job->peratom;
    j.permol = job->permol;
    j.dp = job->dp;
    j.dprot = job->dprot;
    j.center_rotation = job->center_rotation;
    for (int k = 0; k < 3; ++k) {
        if (job->dir) j.dir[k] = job->dir[k];
        if (job->dir2) j.dir2[k] = job->dir2[k];
    }
    j.seed = job->seed;
    j.sweeps = job->sweeps;
    j.keep_positions = job->keep_positions;
    J->est_bytes = estimate_bytes(j);
    FGX_CHECK(J->est_bytes <= sv->budget, FGX_ENOMEM,
              "fgx_server_submit: job larger than the HBM budget");
    J->id = sv->next_id++;
    J->t_submit = now_s();
    if (id) *id = J->id;
    sv->queue.push_back(J->id);
    sv->jobs[J->id] = std::move(J);
    return FGX_OK;
}

extern "C" int fgx_server_step(fgx_server *sv, int *n_admitted, int *n_finished)
{
    FGX_CHECK(sv, FGX_EARG, "fgx_server_step: null server");
    int adm = 0, fin = 0;
    // ---- admission: FIFO while the footprint fits the budget
    int active = 0;
    for (auto &kv : sv->jobs) active += kv.second->state == 1;
    while (!sv->queue.empty() && active < sv->max_active) {
        Job &J = *sv->jobs[sv->queue.front()];
        if (sv->resident + J.est_bytes > sv->budget) break;
        sv->queue.pop_front();
        const int rc = admit(sv, J);
        if (rc) {
            retire(sv, J, rc);
            ++fin;
            continue;
        }
        ++adm;
        ++active;
    }
    // ---- one tick: a batched launch per kernel kind
    std::map<long, std::vector<Job *>> groups;
    std::vector<Job *> runnable;
    for (auto &kv : sv->jobs) {
        Job &J = *kv.second;
        if (J.state != 1) continue;
        if (J.done >= J.spec.sweeps) {   // zero-sweep jobs
            retire(sv, J, FGX_OK);
            ++fin;
            continue;
        }
        runnable.push_back(&J);
    }
    // Jobs of more than one workgroup wait on each other's workgroups: they
    // run in one launch of one kernel kind per tick (the longest-waiting
    // job's), packed onto XCDs, the longest-waiting first, at most one XCD's
    // compute units per XCD; a job that does not fit waits for a later tick.
    // Launches group jobs by kernel kind and by workgroups per chain.
    {
        const int nx = 8, cux = sv->cu_per_xcd;
        std::vector<Job *> wide;
        for (Job *J : runnable) {
            J->sys->blk_xcd = -1;
            if (J->parts > 1) wide.push_back(J);
        }
        std::stable_sort(wide.begin(), wide.end(),
                         [](const Job *a, const Job *b) { return a->ran < b->ran; });
        std::vector<int> room(nx, cux);
        long wide_kind = -1;
        for (Job *J : wide) {
            const long k = kind_of(J->spec);
            if (wide_kind >= 0 && k != wide_kind) continue;
            wide_kind = k;
            int x = 0;
            for (int y = 1; y < nx; ++y)
                if (room[y] > room[x]) x = y;
            if (room[x] < J->parts) continue;
            room[x] -= J->parts;
            J->sys->blk_xcd = x;
            J->ran = sv->ticks;
        }
        for (Job *J : runnable) {
            if (J->parts > 1 && J->sys->blk_xcd < 0) continue;
            groups[kind_of(J->spec) * 4 + (J->parts > 1 ? 2 : 0) + (J->parts > 0 ? 1 : 0)]
                .push_back(J);
        }
    }
    struct InFlight {
        std::vector<Job *> jobs;
        std::vector<int> ncalls;
        std::vector<long> budget;
        void *h = nullptr;
        int rc = FGX_OK;
    };
    // The split jobs' launch runs alone, so that all its workgroups are
    // resident together whatever XCDs the hardware deals them to; then the
    // other launches, side by side.
    int rc_all = FGX_OK;
    for (int round = 0; round < 2; ++round) {
        std::vector<InFlight> fl;
        for (auto &g : groups) {
            if (((g.first & 2) != 0) != (round == 0)) continue;
            InFlight f;
            f.jobs = g.second;
            std::vector<fgx_system *> sys;
            for (Job *J : f.jobs) {
                f.ncalls.push_back(J->spec.sweeps - J->done);
                const double b = sv->tick / std::max(J->trial_est, 1e-8);
                f.budget.push_back(std::max(64L, (long)std::min(b, 1e12)));
                sys.push_back(J->sys);
            }
            f.rc = batch_launch(sys, f.ncalls, f.budget, &f.h);
            if (f.rc && f.jobs.size() > 1) {
                // a shared launch that could not be prepared: run its jobs
                // one by one, so only the one that fails is retired (of split
                // jobs only the first this tick: they do not run side by side)
                (void)batch_finish(f.h, nullptr, nullptr, nullptr);
                const size_t nrun = round == 0 ? 1 : f.jobs.size();
                for (size_t c = nrun; c < f.jobs.size(); ++c) f.jobs[c]->ran = -1;   // first next tick
                for (size_t c = 0; c < nrun; ++c) {
                    InFlight g1;
                    g1.jobs = {f.jobs[c]};
                    g1.ncalls = {f.ncalls[c]};
                    g1.budget = {f.budget[c]};
                    std::vector<fgx_system *> one = {sys[c]};
                    g1.rc = batch_launch(one, g1.ncalls, g1.budget, &g1.h);
                    fl.push_back(std::move(g1));
                }
                continue;
            }
            fl.push_back(std::move(f));
        }
        for (auto &f : fl) {
            double secs = 0;
            std::vector<long> si;
            std::vector<double> sd;
            int rc = batch_finish(f.h, &secs, f.rc ? nullptr : &si, &sd);
            if (f.rc) rc = f.rc;
            for (size_t c = 0; c < f.jobs.size(); ++c) {
                Job &J = *f.jobs[c];
                if (rc) {
                    retire(sv, J, rc);
                    ++fin;
                    continue;
                }
                collect_stats(J.sys, &si[8 * c], &sd[8 * c], &J.st, secs);
                const long tr = si[8 * c];
                J.chain_sum += J.sys->chain_seconds;
                if (tr > 0) {
                    J.trial_est = J.sys->chain_seconds / tr;
                    J.trial_max = std::max(J.trial_max, J.trial_est);
                }
                J.done += (int)si[8 * c + 5];
                if (J.done >= J.spec.sweeps) {
                    retire(sv, J, FGX_OK);
                    ++fin;
                }
            }
            // a job's own failure is in its result; a runtime error is everyone's
            if (rc == FGX_EHIP && !rc_all) rc_all = rc;
        }
    }
    ++sv->ticks;
    if (n_admitted) *n_admitted = adm;
    if (n_finished) *n_finished = fin;
    return rc_all;
}

extern "C" int fgx_server_poll(fgx_server *sv, int *ids, int cap, int *count)
{
    FGX_CHECK(sv && count && (ids || cap == 0), FGX_EARG, "fgx_server_poll: bad arguments");
    const int n = std::min(cap, (int)sv->finished.size());
    for (int k = 0; k < n; ++k) ids[k] = sv->finished[k];
    sv->finished.erase(sv->finished.begin(), sv->finished.begin() + n);
    *count = n;
    return FGX_OK;
}

extern "C" int fgx_server_result(fgx_server *sv, int id, fgx_job_result *r, double *xyz)
{
    FGX_CHECK(sv && r, FGX_EARG, "fgx_server_result: bad arguments");
    auto it = sv->jobs.find(id);
    FGX_CHECK(it != sv->jobs.end(), FGX_EARG, "fgx_server_result: unknown job");
    const Job &J = *it->second;
    FGX_CHECK(J.state >= 2, FGX_ESTATE, "fgx_server_result: job not finished");
    std::memset(r, 0, sizeof(*r));
    r->status = J.error;
    r->trials = J.st.trials;
    r->accepted = J.st.accepted;
    r->utot = J.st.utot;
    r->sweeps = J.done;
    r->queue_seconds = J.t_admit > 0 ? J.t_admit - J.t_submit : J.t_done - J.t_submit;
    r->latency_seconds = J.t_done - J.t_submit;
    // per sweep: device time over completed sweeps; the slowest launch's
    // trial rate times the trials of one sweep
    r->mean_sweep_seconds = J.done ? J.chain_sum / J.done : 0.0;
    r->max_sweep_seconds = J.done ? J.trial_max * ((double)J.st.trials / J.done) : 0.0;
    r->device_bytes = J.bytes;
    r->workgroups = J.parts;
    if (xyz) {
        FGX_CHECK(!J.final_xyz.empty(), FGX_ESTATE,
                  "fgx_server_result: positions were not kept for this job");
        std::memcpy(xyz, J.final_xyz.data(), sizeof(double) * J.final_xyz.size());
    }
    return FGX_OK;
}

extern "C" int fgx_server_info(const fgx_server *sv, size_t *resident, size_t *peak, int *active,
                               int *queued)
{
    FGX_CHECK(sv, FGX_EARG, "fgx_server_info: null server");
    int a = 0;
    for (auto &kv : sv->jobs) a += kv.second->state == 1;
    if (resident) *resident = sv->resident;
    if (peak) *peak = sv->peak_resident;
    if (active) *active = a;
    if (queued) *queued = (int)sv->queue.size();
    return FGX_OK;
}

extern "C" int fgx_server_destroy(fgx_server *sv)
{
    if (!sv) return FGX_OK;
    for (auto &kv : sv->jobs)
        if (kv.second->sys) fgx_destroy(kv.second->sys);
    delete sv;
    return FGX_OK;
}
