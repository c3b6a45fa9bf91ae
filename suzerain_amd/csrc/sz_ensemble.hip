// An ensemble of channels advanced together on one GPU: every member's
// SMR91 step (suzerain/lowstorage.hpp:1461-1520) with a fixed time step,
// the explicit stages of the members on a few streams, and the implicit
// solves of all members in the grouped mixed precision solver (one launch
// per stage over every member's systems).  A step can be captured once in a
// hipGraph and replayed.  The members share their explicit-operator scratch
// per stream (one set sized for the largest member on that stream), so the
// resident memory is the members' states, references and factors.
//
// The reference runs one case per process (apps/perfect/main_advance.cpp
// with lowstorage::step per member); an uncertainty sweep is many such runs.
#include "sz_precision.hpp"
#include "../../include/suzerain_amd.h"
#include "../../include/suzerain_amd_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" sz_channel *sz_internal_channel_create(const sz_channel_spec *spec,
                                                  void *stream, int defer_scratch,
                                                  char *err, int errlen);
extern "C" void sz_internal_channel_scratch_bytes(const sz_channel *h,
                                                  long long out[5]);
extern "C" void sz_internal_channel_set_scratch(sz_channel *h, void *const p[5]);
extern "C" int sz_internal_channel_problem(sz_channel *h, double phi, int which,
                                           void *out, int *n, int *KL, int *KU,
                                           int *nsys);
extern "C" void sz_internal_channel_set_stream(sz_channel *h, hipStream_t s);
extern "C" long long sz_internal_channel_state_bytes(const sz_channel *h);

namespace {

// SMR91 (lowstorage.hpp:1241-1305): numerators over 480.
constexpr double a_[3] = {145. / 480, -36. / 480, 80. / 480};
constexpr double b_[3] = {111. / 480, 100. / 480, 80. / 480};
constexpr double g_[3] = {256. / 480, 200. / 480, 360. / 480};
constexpr double z_[3] = {0. / 480, -136. / 480, -200. / 480};

int set_err(char *err, int errlen, const std::string &m)
{
    if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
    return -1;
}

}  // namespace

struct sz_ensemble {
    // live members in id order (ids are never reused)
    std::vector<sz_channel *> ch;
    std::vector<int> ids;
    int next_id = 0;
    std::vector<double> dt, chi;
    std::vector<sz_channel_spec> specs;
    long long scr[5] = {0, 0, 0, 0, 0};   // bytes of each scratch buffer
    hipStream_t user = nullptr;           // the caller's stream
    hipStream_t main = nullptr;           // own stream (capturable), owned
    hipEvent_t sync_ev = nullptr;
    int maxstreams = 1, nstreams = 1;
    std::vector<hipStream_t> streams;     // explicit stage streams (owned)
    std::vector<hipEvent_t> ev;           // join events, one per stream
    hipEvent_t fork = nullptr;
    std::vector<void *> scratch;          // maxstreams sets of 5 buffers
    sz_prec *prec = nullptr;
    sz_prec_params pp{};
    int graph = 0;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    double tuned_ms = 0;
    std::vector<double> trial_ms;         // autotune: per candidate
    long long bytes = 0;
    std::vector<void *> owned;

    ~sz_ensemble()
    {
        if (exec) hipGraphExecDestroy(exec);
        if (g) hipGraphDestroy(g);
        sz_prec_free(prec);
        for (sz_channel *c : ch) sz_channel_free(c);
        for (void *p : owned) (void)hipFree(p);
        for (hipEvent_t e : ev) hipEventDestroy(e);
        if (fork) hipEventDestroy(fork);
        if (sync_ev) hipEventDestroy(sync_ev);
        for (hipStream_t s : streams) hipStreamDestroy(s);
        if (main) hipStreamDestroy(main);
    }

    void *dalloc(size_t bytes_)
    {
        void *p = nullptr;
        if (hipMalloc(&p, bytes_ ? bytes_ : 1) != hipSuccess) throw std::bad_alloc();
        owned.push_back(p);
        bytes += (long long)bytes_;
        return p;
    }

    void drop_graph()
    {
        if (exec) {
            hipGraphExecDestroy(exec);
            exec = nullptr;
        }
        if (g) {
            hipGraphDestroy(g);
            g = nullptr;
        }
    }

    // scratch sets big enough for channel c (grown, all members rebound)
    void fit_scratch(sz_channel *c)
    {
        long long need[5];
        sz_internal_channel_scratch_bytes(c, need);
        bool grow = scratch.empty();
        for (int q = 0; q < 5; ++q) grow = grow || need[q] > scr[q];
        if (!grow) return;
        (void)hipDeviceSynchronize();
        for (void *p : scratch) {
            (void)hipFree(p);
            owned.erase(std::find(owned.begin(), owned.end(), p));
        }
        scratch.clear();
        for (int q = 0; q < 5; ++q) {
            bytes -= scr[q] * maxstreams;
            scr[q] = std::max(scr[q], need[q]);
        }
        for (int k = 0; k < maxstreams; ++k)
            for (int q = 0; q < 5; ++q) scratch.push_back(dalloc((size_t)scr[q]));
        drop_graph();
    }

    // members on stream k share scratch set k
    void assign_streams(int ns)
    {
        nstreams = ns;
        for (size_t m = 0; m < ch.size(); ++m) {
            const int k = (int)(m % (size_t)ns);
            sz_internal_channel_set_stream(ch[m], streams[(size_t)k]);
            sz_internal_channel_set_scratch(ch[m], &scratch[(size_t)k * 5]);
        }
    }

    int fork_all()
    {
        if (hipEventRecord(fork, main) != hipSuccess) return SZ_ERR_LAUNCH;
        for (int k = 0; k < nstreams; ++k)
            if (hipStreamWaitEvent(streams[(size_t)k], fork, 0) != hipSuccess)
                return SZ_ERR_LAUNCH;
        return 0;
    }

    int join_all()
    {
        for (int k = 0; k < nstreams; ++k) {
            if (hipEventRecord(ev[(size_t)k], streams[(size_t)k]) != hipSuccess ||
                hipStreamWaitEvent(main, ev[(size_t)k], 0) != hipSuccess)
                return SZ_ERR_LAUNCH;
        }
        return 0;
    }

    int problems(int i, std::vector<sz_prec_problem> &pr)
    {
        pr.resize(ch.size());
        for (size_t m = 0; m < ch.size(); ++m) {
            const int rc = sz_internal_channel_problem(ch[m], -dt[m] * b_[i], 0,
                                                       &pr[m], nullptr, nullptr,
                                                       nullptr, nullptr);
            if (rc) return rc;
        }
        return 0;
    }

    int solve(int i)
    {
        std::vector<sz_prec_problem> pr;
        int rc = problems(i, pr);
        if (rc) return rc;
        if (!graph) {
            rc = sz_prec_prepare(prec, i, pr.data());
            if (rc) return rc;
        }
        return sz_prec_solve(prec, i, pr.data());
    }

    // one SMR91 step of every member, enqueued on main (and the streams)
    int enqueue_step()
    {
        int rc = fork_all();
        for (size_t m = 0; m < ch.size() && !rc; ++m) {
            sz_channel *c = ch[m];
            const double d = dt[m];
            rc = sz_channel_copy(c, 0, 1);
            if (!rc) rc = sz_channel_nonlinear(c, 1, 1, nullptr);
            if (!rc) rc = sz_channel_apply_operator(c, d * a_[0], 0);
            if (!rc) rc = sz_channel_add_scaled(c, chi[m] * d * g_[0], 1, 0);
        }
        if (!rc) rc = join_all();
        if (!rc) rc = solve(0);
        for (int i = 1; i < 3 && !rc; ++i) {
            rc = fork_all();
            for (size_t m = 0; m < ch.size() && !rc; ++m) {
                sz_channel *c = ch[m];
                const double d = dt[m];
                rc = sz_channel_accumulate_operator(c, d * a_[i], 0, chi[m] * d * z_[i],
                                                    1);
                if (rc) break;
                sz_channel_exchange(c);
                rc = sz_channel_nonlinear(c, 1, 0, nullptr);
                if (!rc) rc = sz_channel_add_scaled(c, chi[m] * d * g_[i], 1, 0);
            }
            if (!rc) rc = join_all();
            if (!rc) rc = solve(i);
        }
        return rc;
    }

    // the descriptors every slot's solve reads, for the state pointers the
    // step will have at substep i (one exchange per later substep)
    int prepare_all()
    {
        for (int i = 0; i < 3; ++i) {
            if (i > 0)
                for (sz_channel *c : ch) sz_channel_exchange(c);
            std::vector<sz_prec_problem> pr;
            int rc = problems(i, pr);
            if (!rc) rc = sz_prec_prepare(prec, i, pr.data());
            if (rc) {
                return rc;
            }
        }
        // two exchanges: back to the starting parity
        return 0;
    }

    // Device milliseconds of one eager step (after one untimed step that
    // refactors every system).
    double trial(int steps)
    {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        hipEventCreate(&t0);
        hipEventCreate(&t1);
        double ms = -1;
        if (enqueue_step() == 0) {
            hipEventRecord(t0, main);
            int rc = 0;
            for (int k = 0; k < steps && !rc; ++k) rc = enqueue_step();
            hipEventRecord(t1, main);
            float t = 0;
            if (!rc && hipEventSynchronize(t1) == hipSuccess &&
                hipEventElapsedTime(&t, t0, t1) == hipSuccess)
                ms = t / steps;
        }
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        return ms;
    }

    int capture()
    {
        drop_graph();
        int rc = prepare_all();
        if (rc) return rc;
        if (hipStreamBeginCapture(main, hipStreamCaptureModeRelaxed) != hipSuccess)
            return SZ_ERR_LAUNCH;
        rc = enqueue_step();
        hipGraph_t gg = nullptr;
        const hipError_t e = hipStreamEndCapture(main, &gg);
        if (rc) {
            if (gg) hipGraphDestroy(gg);
            return rc;
        }
        if (e != hipSuccess || !gg) return SZ_ERR_LAUNCH;
        g = gg;
        if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
            exec = nullptr;
            return SZ_ERR_LAUNCH;
        }
        return 0;
    }
};

extern "C" {

int sz_ensemble_add(sz_ensemble *e, const sz_channel_spec *spec, double dt,
                    char *err, int errlen);

sz_ensemble *sz_ensemble_create(int nmembers, const sz_channel_spec *specs,
                                const double *dt, const sz_ensemble_opts *opts,
                                void *stream, char *err, int errlen)
{
    if (nmembers < 0 || (nmembers > 0 && (!specs || !dt)) || !opts) {
        set_err(err, errlen, "sz_ensemble_create: bad arguments");
        return nullptr;
    }
    sz_ensemble *e = new sz_ensemble();
    (void)hipGetLastError();                  // start from a clean error state
    try {
        // steps run on a stream of the ensemble's own (the legacy default
        // stream cannot be captured), ordered after and before the caller's
        e->user = (hipStream_t)stream;
        if (hipStreamCreate(&e->main) != hipSuccess ||
            hipEventCreateWithFlags(&e->sync_ev, hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("stream setup");
        e->maxstreams = std::max(1, std::min(opts->max_streams, 8));
        for (int k = 0; k < e->maxstreams; ++k) {
            hipStream_t s = nullptr;
            hipEvent_t v = nullptr;
            if (hipStreamCreate(&s) != hipSuccess ||
                hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess)
                throw std::runtime_error("stream setup");
            e->streams.push_back(s);
            e->ev.push_back(v);
        }
        if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("event setup");
        const std::string spec = opts->solver ? opts->solver : "zcgbsvx,reuse=true";
        const suzerain_amd::specification_zgbsv zs(spec);
        if (zs.method() != suzerain_amd::specification_zgbsv::zcgbsvx)
            throw std::runtime_error("ensembles solve with zcgbsvx specifications");
        e->pp.nslots = 3;
        e->pp.smax = zs.siter();
        e->pp.dmax = zs.diter();
        e->pp.aiter = zs.aiter();
        e->pp.tolsc = zs.tolsc();
        e->pp.reuse = zs.reuse() ? 1 : 0;
        e->pp.graph = opts->graph ? 1 : 0;
        e->pp.lu_bs = opts->lu_bs;
        e->graph = e->pp.graph;
        e->nstreams = std::max(1, std::min(opts->nstreams, e->maxstreams));
        e->prec = sz_prec_create({}, e->pp, e->main);
        if (!e->prec) throw std::runtime_error("grouped solver setup");
        for (int m = 0; m < nmembers; ++m) {
            char cerr[256] = {0};
            if (sz_ensemble_add(e, &specs[m], dt[m], cerr, sizeof cerr) < 0)
                throw std::runtime_error(cerr);
        }
        if (hipStreamSynchronize(e->main) != hipSuccess)
            throw std::runtime_error("device failure during setup");
    } catch (const std::exception &ex) {
        set_err(err, errlen, std::string("sz_ensemble_create: ") + ex.what());
        delete e;
        return nullptr;
    }
    return e;
}

// A new member (its channel, resident state and factor storage); returns
// its id, or a negative error with a message (SZ_ERR_ALLOC when the device
// is out of memory).  The next step is not captured from the old graph.
int sz_ensemble_add(sz_ensemble *e, const sz_channel_spec *spec, double dt,
                    char *err, int errlen)
{
    if (!e || !spec) return -1;
    if (!(dt > 0)) {
        set_err(err, errlen, "time steps must be positive");
        return -3;
    }
    sz_channel_spec sp = *spec;
    sp.solver = "none";
    char cerr[256] = {0};
    sz_channel *c = sz_internal_channel_create(&sp, e->streams[0], 1, cerr,
                                               sizeof cerr);
    if (!c) {
        set_err(err, errlen, std::string("member: ") + cerr);
        return SZ_ERR_ALLOC;
    }
    sz_prec_member pm{};
    sz_internal_channel_problem(c, 0.0, 0, nullptr, &pm.n, &pm.KL, &pm.KU,
                                &pm.nsys);
    try {
        e->fit_scratch(c);
    } catch (...) {
        sz_channel_free(c);
        set_err(err, errlen, "member: out of device memory (scratch)");
        return SZ_ERR_ALLOC;
    }
    const int pid = sz_prec_add(e->prec, pm);
    if (pid < 0) {
        sz_channel_free(c);
        set_err(err, errlen, "member: out of device memory (factors)");
        return SZ_ERR_ALLOC;
    }
    e->ch.push_back(c);
    e->ids.push_back(e->next_id++);
    e->specs.push_back(*spec);
    e->dt.push_back(dt);
    e->chi.push_back(1.0 / ((double)sp.dNx * sp.dNz));
    e->assign_streams(e->nstreams);
    e->drop_graph();
    return e->ids.back();
}

// Retire member `id` (its storage is freed); waits for the device.
int sz_ensemble_remove(sz_ensemble *e, int id)
{
    if (!e) return -1;
    const auto it = std::find(e->ids.begin(), e->ids.end(), id);
    if (it == e->ids.end()) return -2;
    const size_t m = (size_t)(it - e->ids.begin());
    (void)hipDeviceSynchronize();
    const int rc = sz_prec_remove(e->prec, id);
    if (rc) return rc;
    sz_channel_free(e->ch[m]);
    e->ch.erase(e->ch.begin() + (long)m);
    e->ids.erase(e->ids.begin() + (long)m);
    e->specs.erase(e->specs.begin() + (long)m);
    e->dt.erase(e->dt.begin() + (long)m);
    e->chi.erase(e->chi.begin() + (long)m);
    e->assign_streams(e->nstreams);
    e->drop_graph();
    return 0;
}

// Position of member `id` among the live members (or -1).
int sz_ensemble_find(const sz_ensemble *e, int id)
{
    if (!e) return -1;
    const auto it = std::find(e->ids.begin(), e->ids.end(), id);
    return it == e->ids.end() ? -1 : (int)(it - e->ids.begin());
}

void sz_ensemble_free(sz_ensemble *e) { delete e; }

int sz_ensemble_size(const sz_ensemble *e) { return e ? (int)e->ch.size() : -1; }

sz_channel *sz_ensemble_member(sz_ensemble *e, int m)
{
    return e && m >= 0 && m < (int)e->ch.size() ? e->ch[(size_t)m] : nullptr;
}

long long sz_ensemble_bytes(const sz_ensemble *e)
{
    return e ? e->bytes + sz_prec_bytes(e->prec) : -1;
}

namespace {

// order stream b after the work enqueued so far on stream a
int order_after(hipEvent_t ev, hipStream_t a, hipStream_t b)
{
    if (hipEventRecord(ev, a) != hipSuccess || hipStreamWaitEvent(b, ev, 0) != hipSuccess)
        return SZ_ERR_LAUNCH;
    return 0;
}

}  // namespace

int sz_ensemble_step(sz_ensemble *e, int nsteps)
{
    if (!e) return -1;
    if (nsteps < 0) return -2;
    int rc0 = order_after(e->sync_ev, e->user, e->main);
    if (rc0) return rc0;
    for (int k = 0; k < nsteps; ++k) {
        int rc;
        if (e->graph) {
            if (!e->exec) {
                // the first step runs eagerly (lazy device tables, first
                // factorizations); the next one is captured
                rc = e->prepare_all();
                if (!rc) rc = e->enqueue_step();
                if (!rc) rc = e->capture();
                if (rc) return rc;
            } else {
                rc = hipGraphLaunch(e->exec, e->main) == hipSuccess ? 0 : SZ_ERR_LAUNCH;
            }
        } else {
            rc = e->enqueue_step();
        }
        if (rc) return rc;
        // each member's simulation time moves on by its step
        for (size_t m = 0; m < e->ch.size(); ++m)
            sz_channel_set_time(e->ch[m], sz_channel_time(e->ch[m]) + e->dt[m]);
    }
    return order_after(e->sync_ev, e->main, e->user);
}

int sz_ensemble_check(sz_ensemble *e)
{
    if (!e) return -1;
    return sz_prec_first_info(e->prec, true);
}

void *sz_ensemble_solver_stats(sz_ensemble *e, int which)
{
    return e ? sz_prec_stats(e->prec, which) : nullptr;
}

// The autotuner: every candidate mapping -- explicit-stage streams 1, 2, 4
// (members dealt round-robin) x factorization unroll 4 or 8 steps -- timed
// over `steps` eager steps after a warm step, from the same saved states,
// factors refactored for each; the fastest is kept and the states restored.
int sz_ensemble_autotune(sz_ensemble *e, int steps)
{
    if (!e) return -1;
    if (steps < 1) return -2;
    if (order_after(e->sync_ev, e->user, e->main)) return SZ_ERR_LAUNCH;
    // a captured step holds the old stream assignment: retire it (the
    // next step captures again with the tuned choice)
    if (e->exec) {
        if (hipStreamSynchronize(e->main) != hipSuccess) return SZ_ERR_LAUNCH;
        e->drop_graph();
    }
    const size_t nm = e->ch.size();
    std::vector<void *> saved(nm, nullptr);
    int rc = 0;
    for (size_t m = 0; m < nm && !rc; ++m) {
        const long long b = sz_internal_channel_state_bytes(e->ch[m]);
        if (hipMalloc(&saved[m], (size_t)b) != hipSuccess) { rc = SZ_ERR_ALLOC; break; }
        if (hipMemcpyAsync(saved[m], sz_channel_state(e->ch[m], 0), (size_t)b,
                           hipMemcpyDeviceToDevice, e->main) != hipSuccess)
            rc = SZ_ERR_LAUNCH;
    }
    auto restore = [&]() {
        for (size_t m = 0; m < nm; ++m)
            (void)hipMemcpyAsync(sz_channel_state(e->ch[m], 0), saved[m],
                                 (size_t)sz_internal_channel_state_bytes(e->ch[m]),
                                 hipMemcpyDeviceToDevice, e->main);
        sz_prec_invalidate(e->prec);
    };
    const int graph = e->graph;
    e->graph = 0;                      // trials run eagerly
    sz_prec_set_graph(e->prec, 0);
    double best = -1;
    int best_ns = e->nstreams, best_bs = e->pp.lu_bs ? e->pp.lu_bs : 8;
    e->trial_ms.clear();
    const int cand_ns[4] = {1, 2, 3, 4};
    const int cand_bs[2] = {4, 8};
    for (int a = 0; a < 4 && !rc; ++a) {
        if (cand_ns[a] > e->maxstreams) continue;
        for (int b = 0; b < 2 && !rc; ++b) {
            restore();
            e->assign_streams(cand_ns[a]);
            sz_prec_set_lu_bs(e->prec, cand_bs[b]);
            const double ms = e->trial(steps);
            e->trial_ms.push_back(ms);
            if (ms > 0 && (best < 0 || ms < best)) {
                best = ms;
                best_ns = cand_ns[a];
                best_bs = cand_bs[b];
            }
        }
    }
    restore();
    e->assign_streams(best_ns);
    e->pp.lu_bs = best_bs;
    sz_prec_set_lu_bs(e->prec, best_bs);
    e->graph = graph;
    sz_prec_set_graph(e->prec, graph);
    e->tuned_ms = best;
    if (hipStreamSynchronize(e->main) != hipSuccess) rc = SZ_ERR_LAUNCH;
    for (void *p : saved) if (p) (void)hipFree(p);
    return rc;
}

int sz_ensemble_tuning(const sz_ensemble *e, int *nstreams, int *lu_bs,
                       double *trial_ms, int max, int *ntrials)
{
    if (!e) return -1;
    if (nstreams) *nstreams = e->nstreams;
    if (lu_bs) *lu_bs = e->pp.lu_bs;
    int n = 0;
    for (double t : e->trial_ms) {
        if (trial_ms && n < max) trial_ms[n] = t;
        ++n;
    }
    if (ntrials) *ntrials = n;
    return 0;
}

}  // extern "C"
