// Serving loop: simulated scenes send energy, gradient, Hessian and step
// requests with arrival times; the loop keeps every admitted scene's
// topology and positions resident in one batched context and serves, per
// iteration, the earliest pending request of every admitted scene that has
// arrived (continuous batching: requests that arrive while an iteration runs
// join the next one).  Energy, gradient and Hessian requests share one
// collision build over the scenes that asked (the others are masked out of
// the broad phase); step requests share one swept broad phase and CCD.  A
// scene's results do not depend on which other scenes share its launches,
// so each result is the one the scene gets when it is served alone.
//
// Admission: scenes enter the batch in submission order while the sum of
// their estimated device footprints stays within the HBM budget; when the
// admitted scenes have no request left, the batch is rebuilt with the
// scenes still waiting (their positions carried on the host).
#include "context.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <string>
#include <thread>
#include <vector>

namespace ipcg {

namespace {

// Device bytes a batched context holds per triangle and per vertex of its
// scenes during a full step (collision build, energy, gradient, Hessian,
// CCD), measured on sheet pairs and icosphere rows.
constexpr int64_t BYTES_PER_TRI = 4096;
constexpr int64_t BYTES_PER_VERTEX = 1024;

uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// standard normal draw (Box-Muller over two 53-bit uniforms in (0, 1])
double normal01(uint64_t& st)
{
    const double u1 = ((splitmix64(st) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Scene {
    int32_t nV = 0, nE = 0, nF = 0;
    std::vector<double> rest, V; // rest and current positions (nV x 3)
    std::vector<int32_t> E, F;   // scene-local topology
    double sigma = 0;            // perturbation standard deviation of a step
    uint64_t rng = 0;
    int64_t footprint = 0;
    std::deque<int64_t> pending; // tickets in arrival order
    int32_t slot = -1;           // index in the current batch
};

struct Req {
    int32_t scene = 0, kind = 0;
    double arrival = 0, start = -1, finish = -1, value = 0;
    int64_t pairs = 0;
    int32_t accepted = 0, iteration = -1;
    std::vector<double> grad;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

} // namespace

struct Server {
    int device = 0;
    int64_t budget = 0;
    bool keep = false;
    double dhat = 0, kappa = 0;
    ipcg_ctx* ctx = nullptr;
    std::vector<Scene> scenes;
    std::vector<Req> reqs;
    std::vector<int32_t> resident; // scene ids in batch order
    std::vector<int64_t> voff;     // vertex offsets of the batch
    bool ran = false;
    ipcg_server_stats stats {};
    std::string err;
};

namespace {

std::vector<double> records(const double* V, int64_t n)
{
    std::vector<double> r(4 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        r[4 * i] = V[3 * i];
        r[4 * i + 1] = V[3 * i + 1];
        r[4 * i + 2] = V[3 * i + 2];
        r[4 * i + 3] = 0.0;
    }
    return r;
}

// rebuild the batch from the scenes in `ids` (positions from the host)
void rebuild(Server& S, const std::vector<int32_t>& ids)
{
    Ctx& c = S.ctx->c;
    for (int32_t id : S.resident) S.scenes[id].slot = -1;
    S.resident = ids;
    const size_t n = ids.size();
    std::vector<int64_t> voff(n + 1, 0), eoff(n + 1, 0), foff(n + 1, 0);
    for (size_t k = 0; k < n; k++) {
        const Scene& sc = S.scenes[ids[k]];
        voff[k + 1] = voff[k] + sc.nV;
        eoff[k + 1] = eoff[k] + sc.nE;
        foff[k + 1] = foff[k] + sc.nF;
    }
    std::vector<double> rest(3 * (size_t)voff[n]), V(3 * (size_t)voff[n]);
    std::vector<int32_t> E(2 * (size_t)eoff[n]), F(3 * (size_t)foff[n]);
    for (size_t k = 0; k < n; k++) {
        Scene& sc = S.scenes[ids[k]];
        sc.slot = (int32_t)k;
        std::copy(sc.rest.begin(), sc.rest.end(), rest.begin() + 3 * voff[k]);
        std::copy(sc.V.begin(), sc.V.end(), V.begin() + 3 * voff[k]);
        std::copy(sc.E.begin(), sc.E.end(), E.begin() + 2 * eoff[k]);
        std::copy(sc.F.begin(), sc.F.end(), F.begin() + 3 * foff[k]);
    }
    mesh_upload_batched(c, (int32_t)n, voff.data(), rest.data(), IPCG_LAYOUT_AOS, eoff.data(),
                        E.data(), foff.data(), F.data());
    const std::vector<double> r = records(V.data(), voff[n]);
    for (int slot = 0; slot < 2; slot++) {
        c.X[slot].reserve(sizeof(double) * r.size());
        IPCG_HIP(hipMemcpyAsync(c.X[slot].p, r.data(), sizeof(double) * r.size(),
                                hipMemcpyHostToDevice, c.stream));
        c.has_pos[slot] = true;
    }
    IPCG_HIP(hipStreamSynchronize(c.stream));
    S.voff = voff;
    S.stats.epochs++;
}

// admitted scenes with requests left, then waiting scenes in submission
// order while the budget holds
bool admit(Server& S, std::vector<int32_t>& waiting)
{
    std::vector<int32_t> next;
    int64_t used = 0;
    for (int32_t id : S.resident)
        if (!S.scenes[id].pending.empty()) {
            next.push_back(id);
            used += S.scenes[id].footprint;
        }
    const bool changed_before = next.size() != S.resident.size();
    std::vector<int32_t> still;
    for (int32_t id : waiting) {
        if (still.empty() && used + S.scenes[id].footprint <= S.budget) {
            next.push_back(id);
            used += S.scenes[id].footprint;
        } else {
            still.push_back(id); // keep submission order: no overtaking
        }
    }
    const bool added = still.size() != waiting.size();
    waiting = still;
    if (next.empty()) {
        if (!waiting.empty())
            throw Error(IPCG_ERR_OOM, "a scene's estimated footprint exceeds the HBM budget");
        return false;
    }
    if (added || changed_before || S.resident.empty()) rebuild(S, next);
    S.stats.estimated_bytes = std::max<int64_t>(S.stats.estimated_bytes, used);
    return true;
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void serve_potentials(Server& S, const std::vector<int64_t>& batch, double (*now)(void*), void* clk)
{
    Ctx& c = S.ctx->c;
    const size_t n = S.resident.size();
    std::vector<uint8_t> mask(n, 0);
    bool want_g = false, want_h = false;
    for (int64_t t : batch) {
        mask[S.scenes[S.reqs[t].scene].slot] = 1;
        want_g |= S.reqs[t].kind == IPCG_REQ_GRADIENT;
        want_h |= S.reqs[t].kind == IPCG_REQ_HESSIAN;
    }
    batch_set_active(c, mask.data());
    bp_build(c, false, 0.5 * S.dhat, IPCG_BP_LBVH);
    np_build(c, S.dhat, 0.0);
    std::vector<double> energy(n);
    pot_energy_scenes(c, S.dhat, S.kappa, energy.data());
    const size_t S1 = n + 1;
    int64_t total = 0;
    for (int k = 0; k < 4; k++) total += c.ncoll[k];
    // with no collision at all every gradient is zero and every Hessian
    // block empty: nothing to launch
    if (want_g && total > 0) pot_gradient(c, S.dhat, S.kappa);
    if (want_h && total > 0) pot_hessian(c, S.dhat, S.kappa, IPCG_PSD_CLAMP);
    for (int64_t t : batch) {
        Req& q = S.reqs[t];
        const Scene& sc = S.scenes[q.scene];
        const int32_t k = sc.slot;
        q.pairs = 0;
        for (int kind = 0; kind < 4; kind++)
            q.pairs += c.h_coff[kind * S1 + k + 1] - c.h_coff[kind * S1 + k];
        const int64_t r0 = 3 * S.voff[k], r1 = 3 * S.voff[k + 1];
        if (q.kind == IPCG_REQ_ENERGY) {
            q.value = energy[k];
        } else if (q.kind == IPCG_REQ_GRADIENT) {
            std::vector<double> g((size_t)(r1 - r0), 0.0);
            if (total > 0) {
                IPCG_HIP(hipMemcpyAsync(g.data(), c.grad.as<double>() + r0,
                                        sizeof(double) * g.size(), hipMemcpyDeviceToHost,
                                        c.stream));
                IPCG_HIP(hipStreamSynchronize(c.stream));
            }
            double s2 = 0;
            for (double x : g) s2 += x * x;
            q.value = std::sqrt(s2);
            if (S.keep) q.grad = std::move(g);
        } else if (total == 0) { // Hessian of no collision
            q.value = 0;
            if (S.keep) q.rowptr.assign((size_t)(r1 - r0 + 1), 0);
        } else { // Hessian: the scene's rows stay on the device unless kept
            int64_t rp[2];
            IPCG_HIP(hipMemcpyAsync(rp, c.rowptr.as<int64_t>() + r0, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, c.stream));
            IPCG_HIP(hipMemcpyAsync(rp + 1, c.rowptr.as<int64_t>() + r1, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, c.stream));
            IPCG_HIP(hipStreamSynchronize(c.stream));
            q.value = (double)(rp[1] - rp[0]);
            if (S.keep) {
                q.rowptr.resize((size_t)(r1 - r0 + 1));
                q.col.resize((size_t)(rp[1] - rp[0]));
                q.val.resize(q.col.size());
                IPCG_HIP(hipMemcpyAsync(q.rowptr.data(), c.rowptr.as<int64_t>() + r0,
                                        sizeof(int64_t) * q.rowptr.size(), hipMemcpyDeviceToHost,
                                        c.stream));
                if (!q.col.empty()) {
                    IPCG_HIP(hipMemcpyAsync(q.col.data(), c.col.as<int32_t>() + rp[0],
                                            sizeof(int32_t) * q.col.size(),
                                            hipMemcpyDeviceToHost, c.stream));
                    IPCG_HIP(hipMemcpyAsync(q.val.data(), c.val.as<double>() + rp[0],
                                            sizeof(double) * q.val.size(), hipMemcpyDeviceToHost,
                                            c.stream));
                }
                IPCG_HIP(hipStreamSynchronize(c.stream));
                for (int64_t& x : q.rowptr) x -= rp[0];
                for (int32_t& x : q.col) x -= (int32_t)r0;
            }
        }
    }
    const double f = now(clk);
    for (int64_t t : batch) S.reqs[t].finish = f;
}

void serve_steps(Server& S, const std::vector<int64_t>& batch, double (*now)(void*), void* clk)
{
    Ctx& c = S.ctx->c;
    const size_t n = S.resident.size();
    std::vector<uint8_t> mask(n, 0);
    std::vector<std::vector<double>> V1(batch.size());
    for (size_t b = 0; b < batch.size(); b++) {
        Scene& sc = S.scenes[S.reqs[batch[b]].scene];
        mask[sc.slot] = 1;
        V1[b] = sc.V;
        for (double& x : V1[b]) x += sc.sigma * normal01(sc.rng);
        const std::vector<double> r = records(V1[b].data(), sc.nV);
        IPCG_HIP(hipMemcpyAsync(c.X[1].as<double>() + 4 * S.voff[sc.slot], r.data(),
                                sizeof(double) * r.size(), hipMemcpyHostToDevice, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream)); // r goes out of scope
    }
    batch_set_active(c, mask.data());
    bp_build(c, true, 0.0, IPCG_BP_LBVH);
    std::vector<double> steps(n, 1.0);
    ccd_stepsize(c, 0.0, 1e-6, 10000000, 0.8, false, nullptr, steps.data());
    for (size_t b = 0; b < batch.size(); b++) {
        Req& q = S.reqs[batch[b]];
        Scene& sc = S.scenes[q.scene];
        q.value = steps[sc.slot];
        q.accepted = q.value >= 1.0;
        if (q.accepted) { // the perturbed positions become the scene's state
            IPCG_HIP(hipMemcpyAsync(c.X[0].as<double>() + 4 * S.voff[sc.slot],
                                    c.X[1].as<double>() + 4 * S.voff[sc.slot],
                                    sizeof(double) * 4 * (size_t)sc.nV, hipMemcpyDeviceToDevice,
                                    c.stream));
            sc.V = std::move(V1[b]);
        }
    }
    IPCG_HIP(hipStreamSynchronize(c.stream));
    const double f = now(clk);
    for (int64_t t : batch) S.reqs[t].finish = f;
}

// One untimed pass over every admitted scene before the clock starts: the
// batched context's buffers reach their working size and every kernel is
// loaded.  Positions and generators are untouched (the swept pass runs from
// the current positions to themselves), so no result changes.
void warm_up(Server& S)
{
    Ctx& c = S.ctx->c;
    batch_set_active(c, nullptr);
    bp_build(c, false, 0.5 * S.dhat, IPCG_BP_LBVH);
    np_build(c, S.dhat, 0.0);
    std::vector<double> e(S.resident.size()), t(S.resident.size());
    pot_energy_scenes(c, S.dhat, S.kappa, e.data());
    int64_t total = 0;
    for (int k = 0; k < 4; k++) total += c.ncoll[k];
    if (total > 0) {
        pot_gradient(c, S.dhat, S.kappa);
        pot_hessian(c, S.dhat, S.kappa, IPCG_PSD_CLAMP);
    }
    IPCG_HIP(hipMemcpyAsync(c.X[1].p, c.X[0].p, sizeof(double) * 4 * (size_t)c.nV,
                            hipMemcpyDeviceToDevice, c.stream));
    bp_build(c, true, 0.0, IPCG_BP_LBVH);
    ccd_stepsize(c, 0.0, 1e-6, 10000000, 0.8, false, nullptr, t.data());
    IPCG_HIP(hipStreamSynchronize(c.stream));
}

double clock_now(void* p)
{
    return seconds_since(*static_cast<std::chrono::steady_clock::time_point*>(p));
}

} // namespace

void server_run(Server& S)
{
    if (S.ran) throw Error(IPCG_ERR_STATE, "the server has already run");
    S.ran = true;
    for (Scene& sc : S.scenes) // serve each scene's requests in arrival order
        std::stable_sort(sc.pending.begin(), sc.pending.end(), [&](int64_t a, int64_t b) {
            return S.reqs[a].arrival < S.reqs[b].arrival;
        });
    std::vector<int32_t> waiting;
    for (int32_t id = 0; id < (int32_t)S.scenes.size(); id++)
        if (!S.scenes[id].pending.empty()) waiting.push_back(id);
    if (!admit(S, waiting)) return;
    warm_up(S);
    auto t0 = std::chrono::steady_clock::now();
    double busy = 0;
    for (;;) {
        const double now = seconds_since(t0);
        // per scene, the arrived requests up to and including its first
        // step: the energy, gradient and Hessian requests before the step see
        // the scene's current positions, so one evaluation serves them all
        std::vector<int64_t> pot, step;
        double next_arrival = INFINITY;
        bool any_pending = false;
        for (int32_t id : S.resident) {
            Scene& sc = S.scenes[id];
            if (sc.pending.empty()) continue;
            any_pending = true;
            size_t taken = 0;
            for (int64_t t : sc.pending) {
                const Req& q = S.reqs[t];
                if (q.arrival > now) {
                    if (taken == 0) next_arrival = std::min(next_arrival, q.arrival);
                    break;
                }
                taken++;
                if (q.kind == IPCG_REQ_STEP) {
                    step.push_back(t);
                    break;
                }
                pot.push_back(t);
            }
        }
        if (pot.empty() && step.empty()) {
            if (any_pending) {
                std::this_thread::sleep_for(std::chrono::duration<double>(
                    std::min(next_arrival - now, 1e-3)));
                continue;
            }
            if (waiting.empty()) break;
            if (!admit(S, waiting)) break;
            continue;
        }
        const double start = seconds_since(t0);
        const int32_t it = (int32_t)S.stats.iterations++;
        for (auto* L : { &pot, &step })
            for (int64_t t : *L) {
                S.reqs[t].start = start;
                S.reqs[t].iteration = it;
            }
        // the taken requests lead their scenes' queues (arrival order)
        for (auto* L : { &pot, &step })
            for (int64_t t : *L) {
                std::deque<int64_t>& q = S.scenes[S.reqs[t].scene].pending;
                q.erase(std::find(q.begin(), q.end(), t));
            }
        if (!pot.empty()) serve_potentials(S, pot, clock_now, &t0);
        if (!step.empty()) serve_steps(S, step, clock_now, &t0);
        S.stats.requests += (int64_t)(pot.size() + step.size());
        busy += seconds_since(t0) - start;
        S.stats.peak_device_bytes = std::max<int64_t>(S.stats.peak_device_bytes,
                                                      device_bytes(S.ctx->c));
    }
    S.stats.busy_seconds = busy;
    S.stats.wall_seconds = seconds_since(t0);
}

} // namespace ipcg

using namespace ipcg;

struct ipcg_server {
    Server s;
};

namespace {

template <class F> int sguard(ipcg_server* srv, F&& f)
{
    if (!srv) return IPCG_ERR_ARG;
    try {
        f(srv->s);
        srv->s.err.clear();
        return IPCG_OK;
    } catch (const Error& e) {
        srv->s.err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        srv->s.err = e.what();
        return IPCG_ERR_HIP;
    }
}

Req& request(Server& S, int64_t ticket)
{
    if (ticket < 0 || ticket >= (int64_t)S.reqs.size()) throw Error(IPCG_ERR_ARG, "unknown ticket");
    return S.reqs[ticket];
}

} // namespace

extern "C" {

int ipcg_server_create(int device, int64_t hbm_budget_bytes, int keep_results, ipcg_server** out)
{
    if (!out || hbm_budget_bytes <= 0) return IPCG_ERR_ARG;
    *out = nullptr;
    ipcg_ctx* ctx = nullptr;
    const int rc = ipcg_ctx_create(device, &ctx);
    if (rc != IPCG_OK) return rc;
    ipcg_server* srv = new ipcg_server();
    srv->s.device = device;
    srv->s.budget = hbm_budget_bytes;
    srv->s.keep = keep_results != 0;
    srv->s.ctx = ctx;
    *out = srv;
    return IPCG_OK;
}

int ipcg_server_destroy(ipcg_server* srv)
{
    if (!srv) return IPCG_ERR_ARG;
    if (srv->s.ctx) ipcg_ctx_destroy(srv->s.ctx);
    delete srv;
    return IPCG_OK;
}

const char* ipcg_server_last_error(const ipcg_server* srv)
{
    if (!srv) return "null server";
    if (!srv->s.err.empty()) return srv->s.err.c_str();
    return ipcg_last_error(srv->s.ctx);
}

int ipcg_server_add_scene(ipcg_server* srv, int32_t nV, const double* V, int32_t nE,
                          const int32_t* E, int32_t nF, const int32_t* F, double dhat,
                          double kappa, double sigma, uint64_t seed, int32_t* scene)
{
    return sguard(srv, [&](Server& S) {
        if (S.ran) throw Error(IPCG_ERR_STATE, "the server has already run");
        if (nV <= 0 || nE < 0 || nF < 0 || !V || (nE && !E) || (nF && !F) || !scene)
            throw Error(IPCG_ERR_ARG, "invalid scene");
        if (!(dhat > 0) || !(kappa > 0) || !(sigma >= 0))
            throw Error(IPCG_ERR_ARG, "invalid barrier or perturbation parameters");
        if (S.scenes.empty()) {
            S.dhat = dhat;
            S.kappa = kappa;
        } else if (dhat != S.dhat || kappa != S.kappa) {
            throw Error(IPCG_ERR_ARG, "every scene of a server shares dhat and kappa");
        }
        for (int64_t i = 0; i < 2 * (int64_t)nE; i++)
            if (E[i] < 0 || E[i] >= nV) throw Error(IPCG_ERR_ARG, "edge vertex id out of range");
        for (int64_t i = 0; i < 3 * (int64_t)nF; i++)
            if (F[i] < 0 || F[i] >= nV) throw Error(IPCG_ERR_ARG, "face vertex id out of range");
        Scene sc;
        sc.nV = nV;
        sc.nE = nE;
        sc.nF = nF;
        sc.rest.assign(V, V + 3 * (size_t)nV);
        sc.V = sc.rest;
        sc.E.assign(E, E + 2 * (size_t)nE);
        sc.F.assign(F, F + 3 * (size_t)nF);
        sc.sigma = sigma;
        sc.rng = seed;
        sc.footprint = BYTES_PER_TRI * nF + BYTES_PER_VERTEX * nV;
        *scene = (int32_t)S.scenes.size();
        S.scenes.push_back(std::move(sc));
    });
}

int ipcg_server_submit(ipcg_server* srv, int32_t scene, int kind, double arrival, int64_t* ticket)
{
    return sguard(srv, [&](Server& S) {
        if (S.ran) throw Error(IPCG_ERR_STATE, "the server has already run");
        if (scene < 0 || scene >= (int32_t)S.scenes.size() || kind < IPCG_REQ_ENERGY
            || kind > IPCG_REQ_STEP || !(arrival >= 0) || !ticket)
            throw Error(IPCG_ERR_ARG, "invalid request");
        Req q;
        q.scene = scene;
        q.kind = kind;
        q.arrival = arrival;
        *ticket = (int64_t)S.reqs.size();
        S.scenes[scene].pending.push_back(*ticket);
        S.reqs.push_back(std::move(q));
    });
}

int ipcg_server_run(ipcg_server* srv)
{
    return sguard(srv, [&](Server& S) { server_run(S); });
}

int ipcg_server_result(ipcg_server* srv, int64_t ticket, ipcg_request_result* out)
{
    return sguard(srv, [&](Server& S) {
        const Req& q = request(S, ticket);
        if (!out) throw Error(IPCG_ERR_ARG, "null output");
        out->arrival = q.arrival;
        out->start = q.start;
        out->finish = q.finish;
        out->value = q.value;
        out->pairs = q.pairs;
        out->accepted = q.accepted;
        out->iteration = q.iteration;
    });
}

int ipcg_server_gradient(ipcg_server* srv, int64_t ticket, double* grad)
{
    return sguard(srv, [&](Server& S) {
        const Req& q = request(S, ticket);
        if (q.kind != IPCG_REQ_GRADIENT || q.grad.empty())
            throw Error(IPCG_ERR_STATE, "no kept gradient for this ticket");
        std::copy(q.grad.begin(), q.grad.end(), grad);
    });
}

int ipcg_server_hessian(ipcg_server* srv, int64_t ticket, int64_t* nnz, int64_t* rowptr,
                        int32_t* col, double* val)
{
    return sguard(srv, [&](Server& S) {
        const Req& q = request(S, ticket);
        if (q.kind != IPCG_REQ_HESSIAN || q.rowptr.empty())
            throw Error(IPCG_ERR_STATE, "no kept Hessian for this ticket");
        if (nnz) *nnz = (int64_t)q.col.size();
        if (rowptr) std::copy(q.rowptr.begin(), q.rowptr.end(), rowptr);
        if (col) std::copy(q.col.begin(), q.col.end(), col);
        if (val) std::copy(q.val.begin(), q.val.end(), val);
    });
}

int ipcg_server_positions(ipcg_server* srv, int32_t scene, double* V)
{
    return sguard(srv, [&](Server& S) {
        if (scene < 0 || scene >= (int32_t)S.scenes.size() || !V)
            throw Error(IPCG_ERR_ARG, "invalid scene or output");
        std::copy(S.scenes[scene].V.begin(), S.scenes[scene].V.end(), V);
    });
}

ipcg_ctx* ipcg_server_context(ipcg_server* srv) { return srv ? srv->s.ctx : nullptr; }

int ipcg_server_get_stats(ipcg_server* srv, ipcg_server_stats* out)
{
    return sguard(srv, [&](Server& S) {
        if (!out) throw Error(IPCG_ERR_ARG, "null output");
        *out = S.stats;
    });
}

} // extern "C"
