// Continuous-batching server of steady cavity solves on one device.
//
// Requests (a grid, its physics, a tolerance and an iteration cap) queue on
// submission.  Each scheduling round (cfd_server_step):
//   1. admission, first come first served with skipping: a request takes a free
//      slot of the batch of its grid shape and multigrid depth (the coarsest
//      level its Reynolds number allows); a class without a batch gets one when
//      its bytes fit the HBM budget (empty batches are released first when that
//      makes room); a request that cannot be placed waits without blocking the
//      requests behind it, and one whose batch alone exceeds the budget fails;
//   2. one outer iteration of every batch with an active slot: all batches are
//      launched on their own streams before any is waited for;
//   3. retirement: a slot whose mass and energy residuals are at the tolerance,
//      that reached its cap or went non-finite hands its fields and statistics
//      to the request and is freed for the next admission.
// Slots of a batch are independent (own arrays, scalars and physics), so a
// request's result does not depend on what else is being served.
//
// Packed mode (cfd_server_create_packed): the server holds one pack (cfd_pack_*)
// instead of size classes; a request takes free slots of the pack's slabs from its
// grid size down to its coarsest level (it waits while they are taken) and every
// scheduling round is one FAS cycle of the whole pack, so requests of every size
// share each launch.
//
// The reference serves nothing (one hard-coded case per process,
// solver.cpp:2041-2062); this is the serving layer of the batched path.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/cfd_gpu.h"

namespace {

using clk = std::chrono::steady_clock;

constexpr int NARR = 12;  // x y xix xiy etx ety alph beta gamma ajac p1 q1

struct Request {
    long id = 0;
    int n0 = 0, n1 = 0;
    double dxi = 0, deta = 0;
    std::vector<double> grid;  // NARR arrays of n0 x n1
    bool has_p1 = false, has_q1 = false;
    cfd_params phys{};
    double tol = 0;
    int max_outer = 0;
    int state = CFD_REQ_QUEUED;
    int cls = -1, slot = -1;
    cfd_stats st{};
    std::vector<double> fields;  // u v T p when finished
    clk::time_point t_submit, t_done;
};

struct Class {
    int n0 = 0, n1 = 0;
    double dxi = 0, deta = 0;
    int coarsest = 0;  // nodes per side of the batch's coarsest multigrid level
    cfd_batch b = nullptr;
    size_t bytes = 0;
    std::vector<long> owner;  // request per slot, 0 = free
    std::vector<cfd_stats> st;
    int active() const {
        int k = 0;
        for (long o : owner) k += o != 0;
        return k;
    }
    bool fits(const Request &r, int rc) const {
        return r.n0 == n0 && r.n1 == n1 && r.dxi == dxi && r.deta == deta && rc == coarsest;
    }
};

struct Server {
    cfd_params prm{};
    size_t budget = 0;
    int slots = 0;
    std::vector<Class> classes;
    // packed mode: the pack, its statistics per case id and the request per case id
    cfd_pack pack = nullptr;
    size_t pack_bytes = 0;
    std::vector<cfd_stats> pst;
    std::vector<long> powner;
    std::deque<long> queue;
    std::map<long, Request> reqs;
    long next = 1;
    size_t used = 0;
    long finished = 0;
};

void drop_class(Server &s, size_t k) {
    cfd_batch_destroy(s.classes[k].b);
    s.used -= s.classes[k].bytes;
    s.classes.erase(s.classes.begin() + (long)k);
    for (auto &kv : s.reqs)
        if (kv.second.cls > (int)k) kv.second.cls--;
}

// the batch parameters of r's class: the server's, with r's Reynolds, Richardson and
// Prandtl numbers (which set the multigrid depth when fas_min_n is automatic)
cfd_params class_params(const Server &s, const Request &r) {
    cfd_params p = s.prm;
    p.Re = r.phys.Re;
    p.Ri = r.phys.Ri;
    p.Pr = r.phys.Pr;
    return p;
}
int coarsest_of(const Server &s, const Request &r) {
    const cfd_params p = class_params(s, r);
    return cfd_fas_coarsest(r.n0, r.n1, &p);
}

// A batch for r's class within the budget; empty batches are released only when
// that makes room.  -1: no room now; -2: never (the batch alone exceeds the budget,
// or it cannot be built).
int make_class(Server &s, const Request &r) {
    const cfd_params p = class_params(s, r);
    size_t need = 0;
    if (cfd_batch_bytes_needed(r.n0, r.n1, r.dxi, r.deta, s.slots, &p, &need) != CFD_OK) return -2;
    if (need > s.budget) return -2;
    size_t idle = 0;
    for (const Class &c : s.classes)
        if (c.active() == 0) idle += c.bytes;
    if (s.used + need > s.budget && s.used - idle + need > s.budget) return -1;
    for (size_t k = s.classes.size(); k-- > 0 && s.used + need > s.budget;)
        if (s.classes[k].active() == 0) drop_class(s, k);
    if (s.used + need > s.budget) return -1;
    Class c;
    c.n0 = r.n0; c.n1 = r.n1; c.dxi = r.dxi; c.deta = r.deta;
    c.coarsest = cfd_fas_coarsest(r.n0, r.n1, &p);
    if (cfd_batch_create(r.n0, r.n1, r.dxi, r.deta, s.slots, &p, &c.b) != CFD_OK) return -2;
    cfd_batch_bytes(c.b, &c.bytes);
    c.owner.assign(s.slots, 0);
    c.st.assign(s.slots, cfd_stats{});
    s.used += c.bytes;
    s.classes.push_back(std::move(c));
    return (int)s.classes.size() - 1;
}

void finish(Server &s, Request &r, int state) {
    r.state = state;
    r.t_done = clk::now();
    const size_t N = (size_t)r.n0 * r.n1;
    r.fields.assign(4 * N, 0.0);
    double *f = r.fields.data();
    if (s.pack) {
        r.st = s.pst[r.slot];
        if (cfd_pack_get_fields(s.pack, r.slot, f, f + N, f + 2 * N, f + 3 * N) != CFD_OK) r.state = CFD_REQ_FAILED;
        cfd_pack_release(s.pack, r.slot);
        s.powner[r.slot] = 0;
    } else {
        Class &c = s.classes[r.cls];
        r.st = c.st[r.slot];
        if (cfd_batch_get_fields(c.b, r.slot, f, f + N, f + 2 * N, f + 3 * N) != CFD_OK) r.state = CFD_REQ_FAILED;
        cfd_batch_release(c.b, r.slot);
        c.owner[r.slot] = 0;
    }
    r.grid.clear();
    r.grid.shrink_to_fit();
    s.finished++;
}

// packed mode: a request takes its chain of pack slots, or waits while they are taken
void admit_packed(Server &s) {
    for (auto it = s.queue.begin(); it != s.queue.end();) {
        Request &r = s.reqs[*it];
        const size_t N = (size_t)r.n0 * r.n1;
        const double *g = r.grid.data();
        int id = -1;
        const int rc = cfd_pack_add_case(s.pack, r.n0, r.n1, g, g + N, g + 2 * N, g + 3 * N, g + 4 * N, g + 5 * N,
                                         g + 6 * N, g + 7 * N, g + 8 * N, g + 9 * N,
                                         r.has_p1 ? g + 10 * N : nullptr, r.has_q1 ? g + 11 * N : nullptr, &r.phys,
                                         &id);
        if (rc == CFD_E_FULL) { ++it; continue; }
        if (rc != CFD_OK) {
            r.state = CFD_REQ_FAILED;
            r.t_done = clk::now();
            s.finished++;
        } else {
            r.state = CFD_REQ_RUNNING;
            r.slot = id;
            s.powner[id] = r.id;
        }
        it = s.queue.erase(it);
    }
}

void admit(Server &s) {
    if (s.pack) return admit_packed(s);
    for (auto it = s.queue.begin(); it != s.queue.end();) {
        Request &r = s.reqs[*it];
        const int depth = coarsest_of(s, r);
        int k = -1;
        for (size_t q = 0; q < s.classes.size(); q++)
            if (s.classes[q].fits(r, depth) && s.classes[q].active() < s.slots) { k = (int)q; break; }
        if (k < 0) {
            bool class_known = false;
            for (const Class &c : s.classes) class_known |= c.fits(r, depth);
            if (!class_known) k = make_class(s, r);
            if (k == -2) {  // the class cannot be served at all
                r.state = CFD_REQ_FAILED;
                r.t_done = clk::now();
                s.finished++;
                it = s.queue.erase(it);
                continue;
            }
        }
        if (k < 0) { ++it; continue; }
        Class &c = s.classes[k];
        const int slot = (int)(std::find(c.owner.begin(), c.owner.end(), 0L) - c.owner.begin());
        const size_t N = (size_t)r.n0 * r.n1;
        const double *g = r.grid.data();
        const int rc = cfd_batch_set_case(c.b, slot, g, g + N, g + 2 * N, g + 3 * N, g + 4 * N, g + 5 * N,
                                          g + 6 * N, g + 7 * N, g + 8 * N, g + 9 * N, r.has_p1 ? g + 10 * N : nullptr,
                                          r.has_q1 ? g + 11 * N : nullptr, &r.phys);
        if (rc != CFD_OK) {
            r.state = CFD_REQ_FAILED;
            r.t_done = clk::now();
            s.finished++;
        } else {
            r.state = CFD_REQ_RUNNING;
            r.cls = k;
            r.slot = slot;
            c.owner[slot] = r.id;
        }
        it = s.queue.erase(it);
    }
}

}  // namespace

extern "C" {

int cfd_server_create(const cfd_params *params, size_t hbm_budget, int slots_per_batch, cfd_server *out) {
    if (!params || !out || slots_per_batch < 1 || hbm_budget == 0) return CFD_E_ARG;
    if (params->precision != CFD_FP64 && params->precision != CFD_MIXED) return CFD_E_ARG;
    Server *s = new Server();
    s->prm = *params;
    s->budget = hbm_budget;
    s->slots = slots_per_batch;
    *out = reinterpret_cast<cfd_server>(s);
    return CFD_OK;
}

int cfd_server_create_packed(const cfd_params *params, int n0, int n1, double dxi, double deta, int nslabs,
                             const int *slots, cfd_server *out) {
    if (!params || !out) return CFD_E_ARG;
    cfd_pack pk = nullptr;
    const int rc = cfd_pack_create(n0, n1, dxi, deta, nslabs, slots, params, &pk);
    if (rc != CFD_OK) return rc;
    Server *s = new Server();
    s->prm = *params;
    s->pack = pk;
    int ids = 0;
    cfd_pack_capacity(pk, &ids);
    cfd_pack_bytes(pk, &s->pack_bytes);
    s->budget = s->pack_bytes;
    s->used = s->pack_bytes;
    s->pst.assign(ids, cfd_stats{});
    s->powner.assign(ids, 0);
    *out = reinterpret_cast<cfd_server>(s);
    return CFD_OK;
}

int cfd_server_destroy(cfd_server h) {
    Server *s = reinterpret_cast<Server *>(h);
    if (!s) return CFD_OK;
    if (s->pack) cfd_pack_destroy(s->pack);
    for (Class &c : s->classes) cfd_batch_destroy(c.b);
    delete s;
    return CFD_OK;
}

int cfd_server_submit(cfd_server h, int n0, int n1, double dxi, double deta, const double *x, const double *y,
                      const double *xix, const double *xiy, const double *etx, const double *ety,
                      const double *alph, const double *beta, const double *gamma, const double *ajac,
                      const double *p1, const double *q1, const cfd_params *params, double tol, int max_outer,
                      long *ticket) {
    Server *s = reinterpret_cast<Server *>(h);
    if (!s || !ticket || !params || !x || !y || !xix || !xiy || !etx || !ety || !alph || !beta || !gamma || !ajac)
        return CFD_E_ARG;
    if (n0 < 5 || n1 < 5 || !(dxi > 0) || !(deta > 0) || !(tol > 0) || max_outer < 1) return CFD_E_ARG;
    Request r;
    r.id = s->next++;
    r.n0 = n0; r.n1 = n1; r.dxi = dxi; r.deta = deta;
    const size_t N = (size_t)n0 * n1;
    r.grid.resize(NARR * N);
    const double *src[NARR] = {x, y, xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1};
    for (int k = 0; k < NARR; k++)
        if (src[k]) std::memcpy(r.grid.data() + k * N, src[k], N * sizeof(double));
    r.has_p1 = p1 != nullptr;
    r.has_q1 = q1 != nullptr;
    r.phys = *params;
    r.tol = tol;
    r.max_outer = max_outer;
    r.t_submit = clk::now();
    *ticket = r.id;
    s->queue.push_back(r.id);
    s->reqs.emplace(r.id, std::move(r));
    return CFD_OK;
}

int cfd_server_step(cfd_server h, int *running, int *queued) {
    Server *s = reinterpret_cast<Server *>(h);
    if (!s) return CFD_E_ARG;
    admit(*s);
    if (s->pack) {
        int run = 0;
        for (long o : s->powner) run += o != 0;
        int rc = CFD_OK;
        if (run > 0) {
            rc = cfd_pack_launch(s->pack);
            if (rc == CFD_OK) rc = cfd_pack_finish(s->pack, s->pst.data());
            for (size_t id = 0; rc == CFD_OK && id < s->powner.size(); id++) {
                if (s->powner[id] == 0) continue;
                Request &r = s->reqs[s->powner[id]];
                const cfd_stats &st = s->pst[id];
                if (!std::isfinite(st.mass_l1) || !std::isfinite(st.energy_l1)) finish(*s, r, CFD_REQ_FAILED);
                else if (st.mass_res <= r.tol && st.energy_res <= r.tol) finish(*s, r, CFD_REQ_DONE);
                else if (st.iter >= r.max_outer) finish(*s, r, CFD_REQ_NOTCONV);
            }
        }
        if (running) {
            int a = 0;
            for (long o : s->powner) a += o != 0;
            *running = a;
        }
        if (queued) *queued = (int)s->queue.size();
        return rc;
    }
    std::vector<size_t> live;
    for (size_t k = 0; k < s->classes.size(); k++)
        if (s->classes[k].active() > 0) live.push_back(k);
    int rc = CFD_OK;
    for (size_t k : live) {
        const int e = cfd_batch_launch(s->classes[k].b);
        if (e != CFD_OK) rc = e;
    }
    for (size_t k : live) {
        Class &c = s->classes[k];
        const int e = cfd_batch_finish(c.b, c.st.data());
        if (e != CFD_OK && e != CFD_E_NAN) rc = e;
        for (int z = 0; z < s->slots; z++) {
            if (c.owner[z] == 0) continue;
            Request &r = s->reqs[c.owner[z]];
            const cfd_stats &st = c.st[z];
            if (!std::isfinite(st.mass_l1) || !std::isfinite(st.energy_l1)) finish(*s, r, CFD_REQ_FAILED);
            else if (st.mass_res <= r.tol && st.energy_res <= r.tol) finish(*s, r, CFD_REQ_DONE);
            else if (st.iter >= r.max_outer) finish(*s, r, CFD_REQ_NOTCONV);
        }
    }
    if (running) {
        int a = 0;
        for (const Class &c : s->classes) a += c.active();
        *running = a;
    }
    if (queued) *queued = (int)s->queue.size();
    return rc;
}

int cfd_server_poll(cfd_server h, long ticket, int *state, cfd_stats *st, double *latency_s, double *u, double *v,
                    double *T, double *p) {
    Server *s = reinterpret_cast<Server *>(h);
    if (!s || !state) return CFD_E_ARG;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return CFD_E_ARG;
    Request &r = it->second;
    *state = r.state;
    if (r.state == CFD_REQ_QUEUED || r.state == CFD_REQ_RUNNING) return CFD_OK;
    if (st) *st = r.st;
    if (latency_s) *latency_s = std::chrono::duration<double>(r.t_done - r.t_submit).count();
    const size_t N = (size_t)r.n0 * r.n1;
    double *dst[4] = {u, v, T, p};
    for (int k = 0; k < 4; k++)
        if (dst[k] && r.fields.size() == 4 * N) std::memcpy(dst[k], r.fields.data() + k * N, N * sizeof(double));
    s->reqs.erase(it);
    return CFD_OK;
}

int cfd_server_info(cfd_server h, int *batches, size_t *bytes, long *finished) {
    Server *s = reinterpret_cast<Server *>(h);
    if (!s) return CFD_E_ARG;
    if (batches) *batches = s->pack ? 1 : (int)s->classes.size();
    if (bytes) *bytes = s->used;
    if (finished) *finished = s->finished;
    return CFD_OK;
}

}  // extern "C"
