// Many systems at once: nppe_batch_solve (a fixed set of committed systems solved together) and nppe_server
// (requests admitted against an HBM budget into solver slots as they arrive, continuous batching). The batch shares
// launches between systems where they qualify (batch_solve_krylov64: every block still does one system's work of
// its own launch); otherwise, and in the server, every system runs on its own handle and stream. Either way each
// result is bit-identical to the same solve run alone: no buffer or reduction is shared between systems.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "handle.h"
#include "../../include/nppe.h"

namespace
{

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct StoredGrid
{
    nppe_grid g{};
    std::vector<uint8_t> types;
    std::vector<double> wx, wy, wz;
    uint64_t estimate = 0;

    void bind()
    {
        g.cell_types = types.data();
        g.width_x = wx.data();
        g.width_y = wy.data();
        g.width_z = g.dim == 3 ? wz.data() : nullptr;
    }
};

struct Request
{
    int64_t id = 0;
    uint64_t key = 0;
    std::vector<double> b;
    double tol = 1e-8;
    double *x = nullptr;
    double submitted = 0.0, finished = 0.0;
    bool done = false;
    int rc = 0;
    std::string err;
    nppe_result res{};
};

struct Resident
{
    nppe_solver *h = nullptr;
    uint64_t bytes = 0;     // reserved (estimate) until committed, then actual
    bool busy = false;
    bool committed = false;
    double lastUse = 0.0;
};

}  // namespace

struct nppe_server
{
    nppe_options opt{};
    int slots = 1;
    uint64_t budget = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::map<uint64_t, StoredGrid> grids;
    std::map<uint64_t, Resident> resident;
    std::deque<Request *> queue;
    std::map<int64_t, Request *> requests;
    int64_t nextId = 0;
    uint64_t bytesNow = 0, bytesPeak = 0;
    bool stop = false;
    std::string lastError;
    std::vector<std::thread> workers;

    // Under mu: the first queued request that can start now, reserving its grid; evicts idle residents (least
    // recently used) to make room. Returns null when none can.
    Request *admit()
    {
        for (auto it = queue.begin(); it != queue.end(); ++it)
        {
            Request *r = *it;
            auto rs = resident.find(r->key);
            if (rs != resident.end())
            {
                if (rs->second.busy) continue;
                rs->second.busy = true;
                queue.erase(it);
                return r;
            }
            const uint64_t need = grids[r->key].estimate;
            if (budget && need > budget) continue;   // can never fit; left for submit-time rejection
            while (budget && bytesNow + need > budget)
            {
                // evict the least recently used idle resident
                auto victim = resident.end();
                for (auto q = resident.begin(); q != resident.end(); ++q)
                    if (!q->second.busy && (victim == resident.end() || q->second.lastUse < victim->second.lastUse))
                        victim = q;
                if (victim == resident.end()) break;
                nppe_destroy(victim->second.h);
                bytesNow -= victim->second.bytes;
                resident.erase(victim);
            }
            if (budget && bytesNow + need > budget) continue;
            Resident nr;
            nr.busy = true;
            nr.bytes = need;
            resident[r->key] = nr;
            bytesNow += need;
            bytesPeak = std::max(bytesPeak, bytesNow);
            queue.erase(it);
            return r;
        }
        return nullptr;
    }

    void worker()
    {
        for (;;)
        {
            Request *r = nullptr;
            {
                std::unique_lock<std::mutex> lock(mu);
                cvWork.wait(lock, [&] { return stop || (r = admit()) != nullptr; });
                if (!r) return;
            }
            Resident *rs;
            nppe_solver *h;
            bool needCommit;
            StoredGrid *sg;
            {
                std::lock_guard<std::mutex> lock(mu);
                rs = &resident[r->key];
                h = rs->h;
                needCommit = !rs->committed;
                sg = &grids[r->key];
            }
            int rc = NPPE_OK;
            if (needCommit)
            {
                nppe_options o = opt;
                o.hbm_budget = 0;
                rc = nppe_create(&h, &o);
                if (rc == NPPE_OK)
                {
                    // a commit here lies on the request's path: no tile timing (a shape timed before, or the model)
                    h->tileTiming = false;
                    const int64_t n = nppe_commit_grid(h, &sg->g);
                    if (n < 0) rc = (int)n;
                }
                uint64_t actual = 0;
                if (rc == NPPE_OK) nppe_memory_bytes(h, &actual);
                std::lock_guard<std::mutex> lock(mu);
                if (rc == NPPE_OK)
                {
                    bytesNow = bytesNow - rs->bytes + actual;
                    bytesPeak = std::max(bytesPeak, bytesNow);
                    rs->bytes = actual;
                    rs->h = h;
                    rs->committed = true;
                }
            }
            if (rc == NPPE_OK)
            {
                nppe_set_tolerance(h, r->tol);
                rc = nppe_solve(h, r->b.data(), r->x, nullptr, 0.0, &r->res);
            }
            std::string err;
            if (rc != NPPE_OK && rc != NPPE_NOT_CONVERGED) err = nppe_last_error(h);
            std::lock_guard<std::mutex> lock(mu);
            r->err = err;
            if (!rs->committed)
            {
                // commit failed: drop the reservation
                if (h) nppe_destroy(h);
                bytesNow -= rs->bytes;
                resident.erase(r->key);
            }
            else
            {
                rs->busy = false;
                rs->lastUse = now_ms();
            }
            r->rc = rc;
            r->finished = now_ms();
            r->done = true;
            cvDone.notify_all();
            cvWork.notify_all();
        }
    }
};

extern "C" {

int nppe_batch_solve(nppe_solver *const *hs, const double *const *b, double *const *x, int32_t count, int32_t workers,
                     int32_t device, nppe_result *res)
{
    if (!hs || !b || !x || count < 0) return NPPE_ERR_ARGUMENT;
    if (count == 0) return NPPE_OK;
    if (device)
    {
        // the inner iterations of all systems batched into shared launches, where every system qualifies
        std::vector<int> rcs(count, NPPE_OK);
        const int rc = nppe::batch_solve_krylov64(hs, b, x, count, res, rcs.data());
        if (rc < 0) return rc;
        if (rc == 1 && std::getenv("NPPE_BATCH_VERBOSE")) std::fprintf(stderr, "nppe: %s\n", hs[0]->err.c_str());
        if (rc == NPPE_OK)
        {
            int worst = NPPE_OK;
            for (int r : rcs)
                if (r < 0) return r;
                else if (r != NPPE_OK) worst = r;
            return worst;
        }
    }
    const int nw = std::max(1, std::min<int>(workers > 0 ? workers : count, std::min(count, 64)));
    std::atomic<int> next(0);
    std::vector<int> rcs(count, NPPE_OK);
    auto run = [&]() {
        for (;;)
        {
            const int i = next.fetch_add(1);
            if (i >= count) return;
            nppe_result r{};
            rcs[i] = device ? nppe_solve_device(hs[i], b[i], x[i], nullptr, 0.0, &r)
                            : nppe_solve(hs[i], b[i], x[i], nullptr, 0.0, &r);
            if (res) res[i] = r;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nw; t++) pool.emplace_back(run);
    run();
    for (std::thread &t : pool) t.join();
    int worst = NPPE_OK;
    for (int rc : rcs)
        if (rc < 0) return rc;
        else if (rc != NPPE_OK) worst = rc;
    return worst;
}

int nppe_server_create(nppe_server **out, const nppe_options *opt, int32_t slots, uint64_t hbm_budget)
{
    if (!out || slots < 1 || slots > 64) return NPPE_ERR_ARGUMENT;
    nppe_server *s = new nppe_server();
    if (opt) s->opt = *opt;
    else nppe_default_options(&s->opt);
    s->slots = slots;
    s->budget = hbm_budget;
    for (int t = 0; t < slots; t++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return NPPE_OK;
}

int nppe_server_destroy(nppe_server *s)
{
    if (!s) return NPPE_OK;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        // finish the queued work first
        s->cvDone.wait(lock, [&] {
            if (!s->queue.empty()) return false;
            for (auto &kv : s->resident)
                if (kv.second.busy) return false;
            return true;
        });
        s->stop = true;
    }
    s->cvWork.notify_all();
    for (std::thread &t : s->workers) t.join();
    for (auto &kv : s->resident)
        if (kv.second.h) nppe_destroy(kv.second.h);
    for (auto &kv : s->requests) delete kv.second;
    delete s;
    return NPPE_OK;
}

int nppe_server_add_grid(nppe_server *s, uint64_t key, const nppe_grid *g)
{
    if (!s || !g || (g->dim != 2 && g->dim != 3) || !g->cell_types || !g->width_x || !g->width_y ||
        (g->dim == 3 && !g->width_z))
        return NPPE_ERR_ARGUMENT;
    StoredGrid sg;
    sg.g = *g;
    const int nz = g->dim == 2 ? 1 : g->nz;
    sg.g.nz = nz;
    sg.types.assign(g->cell_types, g->cell_types + (size_t)g->nx * g->ny * nz);
    sg.wx.assign(g->width_x, g->width_x + g->nx);
    sg.wy.assign(g->width_y, g->width_y + g->ny);
    if (g->dim == 3) sg.wz.assign(g->width_z, g->width_z + nz);
    sg.estimate = nppe_estimate_bytes_for(g, &s->opt);
    if (s->budget && sg.estimate > s->budget) return NPPE_ERR_BUDGET;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->resident.count(key)) return NPPE_ERR_ARGUMENT;   // a resident grid cannot be replaced
    s->grids[key] = std::move(sg);
    s->grids[key].bind();
    return NPPE_OK;
}

int64_t nppe_server_submit(nppe_server *s, uint64_t key, const double *b, double tolerance, double *x)
{
    if (!s || !x || !(tolerance > 0.0)) return NPPE_ERR_ARGUMENT;
    Request *r = new Request();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        auto it = s->grids.find(key);
        if (it == s->grids.end())
        {
            delete r;
            return NPPE_ERR_NO_GRID;
        }
        // unknown count = interior FLUID cells of the stored grid
        const nppe_grid &g = it->second.g;
        int64_t n = 0;
        const int k0 = g.dim == 3 ? 1 : 0, k1 = g.dim == 3 ? g.nz - 2 : 0;
        for (int i = 1; i <= g.ny - 2; i++)
            for (int j = 1; j <= g.nx - 2; j++)
                for (int k = k0; k <= k1; k++)
                    if (g.cell_types[((size_t)i * g.nx + j) * g.nz + k] == NPPE_FLUID) n++;
        if (n > 0 && !b)
        {
            delete r;
            return NPPE_ERR_ARGUMENT;
        }
        r->b.assign(b, b + n);
        r->id = s->nextId++;
        r->key = key;
        r->tol = tolerance;
        r->x = x;
        r->submitted = now_ms();
        s->requests[r->id] = r;
        s->queue.push_back(r);
    }
    s->cvWork.notify_one();
    return r->id;
}

int nppe_server_wait(nppe_server *s, int64_t ticket, nppe_result *res, double *latency_ms)
{
    if (!s) return NPPE_ERR_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return NPPE_ERR_ARGUMENT;
    Request *r = it->second;
    s->cvDone.wait(lock, [&] { return r->done; });
    if (res) *res = r->res;
    if (latency_ms) *latency_ms = r->finished - r->submitted;
    const int rc = r->rc;
    s->lastError = r->err;
    s->requests.erase(it);
    delete r;
    return rc;
}

const char *nppe_server_last_error(const nppe_server *s)
{
    return s ? s->lastError.c_str() : "null server";
}

int nppe_server_memory(const nppe_server *s, uint64_t *now, uint64_t *peak)
{
    if (!s) return NPPE_ERR_ARGUMENT;
    nppe_server *m = const_cast<nppe_server *>(s);
    std::lock_guard<std::mutex> lock(m->mu);
    if (now) *now = m->bytesNow;
    if (peak) *peak = m->bytesPeak;
    return NPPE_OK;
}

}  // extern "C"
