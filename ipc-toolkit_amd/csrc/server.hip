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

// Step perturbation, drawn on the device from a counter-based generator:
// coordinate i of the scene's draw number k takes its two uniforms from
// the splitmix64 finalizer of (seed, k, i), so a draw depends on nothing but
// the scene's own seed and step count.  Box-Muller over two 53-bit
// uniforms, u1 in (0, 1].
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_perturb(const double* X0, double* X1, int64_t v0, int64_t nv, double sigma,
                          uint64_t seed, uint64_t draw)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 3 * nv) return;
    const int64_t v = i / 3, d = i - 3 * v;
    const uint64_t key = seed * 0x9E3779B97F4A7C15ull + draw * 0xD1B54A32D192ED03ull
                       + (uint64_t)i * 2ull;
    const double u1 = (double)((mix64(key + 1) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(mix64(key + 2) >> 11) * (1.0 / 9007199254740992.0);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const int64_t at = 4 * (v0 + v) + d;
    X1[at] = X0[at] + sigma * z;
}

// Per scene k: the squared 2-norm of the gradient rows [3 voff[k], 3
// voff[k+1]) and the Hessian nonzeros of those rows.  The rows of a scene
// are cut into chunks of NORM_CHUNK counted from its first row, each summed
// by one block in a fixed order, and the chunks of a scene are added in
// order, so the value is the same for the scene in any batch.
constexpr int64_t NORM_CHUNK = 8192;

__global__ __launch_bounds__(256) void k_norm_chunks(const double* g, const int64_t* voff,
                                                     const int64_t* coff, int n, double* part)
{
    __shared__ double red[256];
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n; // scene of this chunk: last k with coff[k] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (coff[mid] <= b) lo = mid;
        else hi = mid;
    }
    const int64_t r0 = 3 * voff[lo] + (b - coff[lo]) * NORM_CHUNK;
    const int64_t r1 = min(3 * voff[lo + 1], r0 + NORM_CHUNK);
    double acc = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) acc += g[r] * g[r];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[b] = red[0];
}

__global__ void k_scene_norms(const double* part, const int64_t* rowptr, const int64_t* voff,
                              const int64_t* coff, int n, double* norm2, int64_t* nnz)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    double acc = 0;
    if (part)
        for (int64_t b = coff[k]; b < coff[k + 1]; b++) acc += part[b];
    norm2[k] = acc;
    nnz[k] = rowptr ? rowptr[3 * voff[k + 1]] - rowptr[3 * voff[k]] : 0;
}

struct Scene {
    int32_t nV = 0, nE = 0, nF = 0;
    std::vector<double> rest, V; // rest and current positions (nV x 3)
    std::vector<int32_t> E, F;   // scene-local topology
    double sigma = 0;            // perturbation standard deviation of a step
    uint64_t seed = 0, draws = 0; // perturbation generator: seed, draws so far
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
    int64_t nchunks = 0;           // gradient-norm chunks of the batch
    DevBuf voff_d, coff_d, red;    // voff, chunk offsets per scene; reductions
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

// the resident scenes' current positions (device slot 0) into their host
// copies
void download_positions(Server& S)
{
    Ctx& c = S.ctx->c;
    if (S.resident.empty() || !c.has_pos[0]) return;
    std::vector<double> r(4 * (size_t)S.voff.back());
    IPCG_HIP(hipMemcpyAsync(r.data(), c.X[0].p, sizeof(double) * r.size(), hipMemcpyDeviceToHost,
                            c.stream));
    IPCG_HIP(hipStreamSynchronize(c.stream));
    for (int32_t id : S.resident) {
        Scene& sc = S.scenes[id];
        const int64_t v0 = S.voff[sc.slot];
        for (int64_t i = 0; i < sc.nV; i++)
            for (int d = 0; d < 3; d++) sc.V[3 * i + d] = r[4 * (v0 + i) + d];
    }
}

// rebuild the batch from the scenes in `ids` (positions from the host)
void rebuild(Server& S, const std::vector<int32_t>& ids)
{
    Ctx& c = S.ctx->c;
    download_positions(S);
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
    std::vector<int64_t> coff(n + 1, 0);
    for (size_t k = 0; k < n; k++)
        coff[k + 1] = coff[k] + (3 * (voff[k + 1] - voff[k]) + NORM_CHUNK - 1) / NORM_CHUNK;
    S.voff_d.reserve(sizeof(int64_t) * voff.size());
    S.coff_d.reserve(sizeof(int64_t) * coff.size());
    IPCG_HIP(hipMemcpyAsync(S.voff_d.p, voff.data(), sizeof(int64_t) * voff.size(),
                            hipMemcpyHostToDevice, c.stream));
    IPCG_HIP(hipMemcpyAsync(S.coff_d.p, coff.data(), sizeof(int64_t) * coff.size(),
                            hipMemcpyHostToDevice, c.stream));
    S.nchunks = coff[n];
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
    // gradient norms and Hessian nonzeros of every scene: one launch, one copy
    std::vector<double> norm2(n, 0.0);
    std::vector<int64_t> nnz(n, 0);
    if ((want_g || want_h) && total > 0) {
        S.red.reserve((sizeof(double) + sizeof(int64_t)) * n
                      + sizeof(double) * (size_t)std::max<int64_t>(S.nchunks, 1));
        double* d_norm = S.red.as<double>();
        int64_t* d_nnz = reinterpret_cast<int64_t*>(d_norm + n);
        double* d_part = reinterpret_cast<double*>(d_nnz + n);
        if (want_g && S.nchunks > 0)
            hipLaunchKernelGGL(k_norm_chunks, dim3((unsigned)S.nchunks), dim3(256), 0, c.stream,
                               c.grad.as<double>(), S.voff_d.as<int64_t>(),
                               S.coff_d.as<int64_t>(), (int)n, d_part);
        hipLaunchKernelGGL(k_scene_norms, dim3(grid_for((int64_t)n, 256)), dim3(256), 0, c.stream,
                           want_g ? d_part : nullptr, want_h ? c.rowptr.as<int64_t>() : nullptr,
                           S.voff_d.as<int64_t>(), S.coff_d.as<int64_t>(), (int)n, d_norm,
                           d_nnz);
        IPCG_CHECK_LAUNCH();
        IPCG_HIP(hipMemcpyAsync(norm2.data(), d_norm, sizeof(double) * n, hipMemcpyDeviceToHost,
                                c.stream));
        IPCG_HIP(hipMemcpyAsync(nnz.data(), d_nnz, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                                c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    }
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
            q.value = std::sqrt(norm2[k]);
            if (S.keep) { // the scene's rows on the host
                q.grad.assign((size_t)(r1 - r0), 0.0);
                if (total > 0) {
                    IPCG_HIP(hipMemcpyAsync(q.grad.data(), c.grad.as<double>() + r0,
                                            sizeof(double) * q.grad.size(),
                                            hipMemcpyDeviceToHost, c.stream));
                    IPCG_HIP(hipStreamSynchronize(c.stream));
                }
            }
        } else if (total == 0) { // Hessian of no collision
            q.value = 0;
            if (S.keep) q.rowptr.assign((size_t)(r1 - r0 + 1), 0);
        } else { // Hessian: the scene's rows stay on the device unless kept
            q.value = (double)nnz[k];
            if (S.keep) {
                int64_t rp0 = 0;
                IPCG_HIP(hipMemcpyAsync(&rp0, c.rowptr.as<int64_t>() + r0, sizeof(int64_t),
                                        hipMemcpyDeviceToHost, c.stream));
                IPCG_HIP(hipStreamSynchronize(c.stream));
                q.rowptr.resize((size_t)(r1 - r0 + 1));
                q.col.resize((size_t)nnz[k]);
                q.val.resize(q.col.size());
                IPCG_HIP(hipMemcpyAsync(q.rowptr.data(), c.rowptr.as<int64_t>() + r0,
                                        sizeof(int64_t) * q.rowptr.size(), hipMemcpyDeviceToHost,
                                        c.stream));
                if (!q.col.empty()) {
                    IPCG_HIP(hipMemcpyAsync(q.col.data(), c.col.as<int32_t>() + rp0,
                                            sizeof(int32_t) * q.col.size(),
                                            hipMemcpyDeviceToHost, c.stream));
                    IPCG_HIP(hipMemcpyAsync(q.val.data(), c.val.as<double>() + rp0,
                                            sizeof(double) * q.val.size(), hipMemcpyDeviceToHost,
                                            c.stream));
                }
                IPCG_HIP(hipStreamSynchronize(c.stream));
                for (int64_t& x : q.rowptr) x -= rp0;
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
    // the perturbed positions of every stepping scene into slot 1
    for (int64_t t : batch) {
        Scene& sc = S.scenes[S.reqs[t].scene];
        mask[sc.slot] = 1;
        hipLaunchKernelGGL(k_perturb, dim3(grid_for(3 * (int64_t)sc.nV, 256)), dim3(256), 0,
                           c.stream, c.X[0].as<double>(), c.X[1].as<double>(), S.voff[sc.slot],
                           (int64_t)sc.nV, sc.sigma, sc.seed, sc.draws++);
    }
    IPCG_CHECK_LAUNCH();
    batch_set_active(c, mask.data());
    bp_build(c, true, 0.0, IPCG_BP_LBVH);
    std::vector<double> steps(n, 1.0);
    ccd_stepsize(c, 0.0, 1e-6, 10000000, 0.8, false, nullptr, steps.data());
    for (int64_t t : batch) {
        Req& q = S.reqs[t];
        Scene& sc = S.scenes[q.scene];
        q.value = steps[sc.slot];
        q.accepted = q.value >= 1.0;
        if (q.accepted) // the perturbed positions become the scene's state
            IPCG_HIP(hipMemcpyAsync(c.X[0].as<double>() + 4 * S.voff[sc.slot],
                                    c.X[1].as<double>() + 4 * S.voff[sc.slot],
                                    sizeof(double) * 4 * (size_t)sc.nV, hipMemcpyDeviceToDevice,
                                    c.stream));
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
    download_positions(S);
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
        sc.seed = seed;
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
