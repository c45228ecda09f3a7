// server.hip -- resident request loop for batches of load cases on many
// meshes (host code; the device work goes through the contexts' C ABI and the
// packed solver).
//
// Reference workflow served here: a macro-element / RVE homogenisation
// (MultiFieldMacroElement::CalculateK, MultiFieldMacroElement.cpp:302-407) --
// assemble one mesh, then solve one system per affine boundary load case.
// The reference runs one such model at a time on the host; the server keeps
// `max_slots` of them resident on one GPU at once.
//
// Design
//  - Slots: each slot is a worker thread with a pool of contexts (seats, one
//    per request of a pack; own HIP streams, device buffers) and a packed
//    solver with its own stream, so the work of different slots overlaps on
//    the device and no state is shared between slots.
//  - Packing: a slot takes the request at the head of the queue and, when it
//    is small (fewer than 2048 node slices, at most 8 load cases), also the
//    small requests queued right behind it with the same number of cases, up
//    to max_pack and the HBM budget.  It assembles each on its own seat, then
//    solves all their cases in one packed batched PCG (pack.hip): one
//    persistent kernel set, graph-captured chunks, every request keeping its
//    own recurrences and stopping tests.
//  - Continuous batching: workers take the next queued requests as soon as
//    they are free; a request that arrives while others run starts as soon as
//    a slot and device memory allow, without waiting for the batch to drain.
//  - Admission by HBM footprint: every request carries an estimate of the
//    device bytes its context needs; a slot charges the estimates of the
//    requests it holds (its buffers stay resident for reuse after the requests
//    end) and requests are started only while the charges of all slots fit
//    the budget.  Idle slots give their buffers back when a request does not
//    fit otherwise.  Requests start in arrival order.
//  - Reuse: a seat whose last request had the same connectivity and
//    restraints keeps its pattern (node graph, CSR, SELL layout) and only
//    re-uploads coordinates, materials and values.
//  - Determinism: every kernel of a solve runs with fixed-order reductions
//    confined to its own request, so a request's result does not depend on
//    what else runs beside it or shares its pack.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bhip_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    long long id = 0;
    int n_nodes = 0, n_el = 0, mat_kind = 0, n_groups = 0, n_cases = 0, order = 2;
    std::vector<double> xyz, mat, presc, load;
    std::vector<int> conn, group;
    std::vector<unsigned char> rest;
    double footprint = 0;
    int state = 0;  // 0 queued, 1 running, 2 done
    std::vector<double> x;
    bhip_request_stats st{};
    Clock::time_point t_submit, t_start, t_end;
};

struct Seat {
    bhip_ctx *ctx = nullptr;
    double hw = 0;  // device bytes the context holds: its buffers only grow, so the
                    // largest footprint of the requests it has run
    // topology of the last request (pattern reuse)
    bool have_topo = false;
    int n_nodes = 0, n_el = 0;
    std::vector<int> conn;
    std::vector<unsigned char> rest;
};

struct Slot {
    std::thread th;
    bool busy = false;
    double charge = 0;  // device bytes attributed to this slot
    PackEngine *pe = nullptr;
    std::vector<Seat> seats;
};

}  // namespace

struct bhip_server {
    int device = 0;
    bhip_server_opts opts{};
    double budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<long long, Request *> reqs;
    long long next_id = 1;
    bool stop = false;
    std::vector<Slot> slots;
    long long completed = 0;
    Clock::time_point t0 = Clock::now();
};

namespace {

// Device bytes of one hex8 request with 3 DOFs per node and `cases` load
// cases, from the buffers the context allocates (per node: node graph and
// incidence ~360 B, FP64 CSR ~2.9 KB, FP32 SELL-64 ~1.1 KB, solver vectors
// ~0.3 KB, cases 120 B each; per element: 1 KB of pattern sort keys and
// 2.6 KB of element-stage blocks) plus 64 MB for tables, partials and
// allocator rounding.
double estimate_bytes(long long n_nodes, long long n_el, int cases) {
    const int r = std::min(cases, 8);
    return (double)n_nodes * (4700.0 + 120.0 * r) + 3648.0 * (double)n_el + 64.0 * (1 << 20);
}

bool topo_same(const Seat &s, const Request &r) {
    return s.have_topo && s.n_nodes == r.n_nodes && s.n_el == r.n_el && s.conn == r.conn && s.rest == r.rest;
}

// Requests that go into packs: fewer than 2048 node slices, at most 8 cases
// (one case group).
constexpr long long kPackMaxNodes = 2048LL * 64;
bool packable(const bhip_server *srv, const Request &r) {
    return srv->opts.max_pack > 1 && srv->opts.precision == BHIP_PREC_FP32_IR && r.n_cases <= 8 &&
           r.n_nodes < kPackMaxNodes;
}

bhip_solve_opts solve_opts(const bhip_server *srv) {
    bhip_solve_opts o{};
    o.tol = srv->opts.tol;
    o.maxit = srv->opts.maxit;
    o.stop_rule = BHIP_STOP_REL_RESID;
    o.precision = srv->opts.precision;
    o.inner_tol = srv->opts.inner_tol;
    o.max_outer = 0;
    o.check_every = 0;
    o.use_x0 = 0;
    return o;
}

// Mesh, pattern (reused when the seat's last request had the same topology),
// materials and stiffness assembly of one request on its seat.
int setup_request(bhip_server *srv, Seat &seat, Request &r) {
    int st = BHIP_OK;
    if (!seat.ctx) {
        st = bhip_create(srv->device, &seat.ctx);
        if (st != BHIP_OK) return st;
        seat.have_topo = false;
        seat.hw = 0;
    }
    seat.hw = std::max(seat.hw, r.footprint);
    bhip_ctx *ctx = seat.ctx;
    r.st.reused_pattern = 0;
    if (topo_same(seat, r)) {
        st = bhip_update_coords(ctx, r.xyz.data());
        r.st.reused_pattern = 1;
    } else {
        seat.have_topo = false;
        st = bhip_set_mesh(ctx, 3, r.n_nodes, r.xyz.data(), r.n_el, 8, r.conn.data(), nullptr, 3);
        if (st == BHIP_OK) st = bhip_set_restraints(ctx, r.rest.data(), nullptr);
        if (st == BHIP_OK) st = bhip_build_pattern(ctx);
        if (st == BHIP_OK) {
            seat.have_topo = true;
            seat.n_nodes = r.n_nodes;
            seat.n_el = r.n_el;
            seat.conn = r.conn;
            seat.rest = r.rest;
        }
    }
    if (st == BHIP_OK)
        st = bhip_set_materials(ctx, r.mat_kind, r.n_groups, r.mat.data(), r.group.empty() ? nullptr : r.group.data());
    if (st == BHIP_OK) st = bhip_set_integration(ctx, r.order, 1.0);
    if (st == BHIP_OK) st = bhip_assemble(ctx, BHIP_ASM_STIFFNESS, nullptr, nullptr);
    r.x.assign((size_t)r.n_cases * 3 * (size_t)r.n_nodes, 0.0);
    return st;
}

// A failed request leaves its seat's context in an unknown state: start afresh.
void fail_seat(Seat &seat, Request &r, int st) {
    r.st.status = st;
    if (seat.ctx) {
        std::strncpy(r.st.message, bhip_last_error(seat.ctx), sizeof(r.st.message) - 1);
        bhip_destroy(seat.ctx);
    }
    seat.ctx = nullptr;
    seat.have_topo = false;
    seat.hw = 0;
}

// Runs the requests of one slot's pack: set-up on each seat, one packed solve
// of all their cases (or, for a request that does not pack, its case groups
// through bhip_solve_cases), then residuals and solutions per request.
void run_pack(bhip_server *srv, Slot &slot, std::vector<Request *> &pk) {
    const int np = (int)pk.size();
    if ((int)slot.seats.size() < np) slot.seats.resize(np);
    const bhip_solve_opts o = solve_opts(srv);
    std::vector<bhip_ctx *> ctxs;
    std::vector<int> idx;
    for (int k = 0; k < np; k++) {
        Request &r = *pk[k];
        r.st.status = BHIP_OK;
        r.st.pack_size = np;
        const auto t0 = Clock::now();
        int st = setup_request(srv, slot.seats[k], r);
        if (st != BHIP_OK) {
            fail_seat(slot.seats[k], r, st);
            continue;
        }
        bhip_ctx *ctx = slot.seats[k].ctx;
        const long long n_dof = 3LL * r.n_nodes;
        if (np == 1 && !packable(srv, r)) {
            // alone: case groups of up to 8 through bhip_solve_cases
            const auto t1 = Clock::now();
            r.st.setup_ms = ms_between(t0, t1);
            for (int c0 = 0; c0 < r.n_cases && st == BHIP_OK; c0 += 8) {
                const int nc = std::min(8, r.n_cases - c0);
                bhip_stats s{};
                st = bhip_solve_cases(ctx, nc, r.presc.data() + (size_t)c0 * n_dof,
                                      r.load.empty() ? nullptr : r.load.data(), r.x.data() + (size_t)c0 * n_dof, &o, &s);
                r.st.iters += s.iters;
                r.st.case_iters += s.case_iters;
                r.st.outer_iters = std::max(r.st.outer_iters, s.outer_iters);
                r.st.rel_resid = std::max(r.st.rel_resid, s.rel_resid);
            }
            r.st.solve_ms = ms_between(t1, Clock::now());
            if (st != BHIP_OK) fail_seat(slot.seats[k], r, st);
            continue;
        }
        bool fused = false;
        st = bhip_cases_begin(ctx, r.n_cases, r.presc.data(), r.load.empty() ? nullptr : r.load.data(), &o, &fused);
        if (st == BHIP_OK && !fused) st = bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "request cannot join a pack");
        r.st.setup_ms = ms_between(t0, Clock::now());
        if (st != BHIP_OK) {
            fail_seat(slot.seats[k], r, st);
            continue;
        }
        ctxs.push_back(ctx);
        idx.push_back(k);
    }
    if (ctxs.empty()) return;
    const int n = (int)ctxs.size();
    const int R = pk[idx[0]]->n_cases;
    std::vector<bhip_stats> S(n);
    std::vector<int> status(n, BHIP_OK);
    const auto t1 = Clock::now();
    const int e = bhip_pack_run(slot.pe, ctxs.data(), n, R, o, S.data(), status.data());
    const double solve_ms = ms_between(t1, Clock::now());
    for (int j = 0; j < n; j++) {
        Request &r = *pk[idx[j]];
        Seat &seat = slot.seats[idx[j]];
        int st = e != BHIP_OK ? e : BHIP_OK;
        r.st.solve_ms = solve_ms;
        r.st.iters = S[j].iters;
        r.st.case_iters = S[j].case_iters;
        r.st.outer_iters = S[j].outer_iters;
        if (st == BHIP_OK) {
            bhip_stats fin{};
            st = bhip_cases_finish(ctxs[j], R, r.x.data(), &fin);
            r.st.rel_resid = fin.rel_resid;
        }
        if (st == BHIP_OK && status[j] != BHIP_OK)
            st = bhip_fail(ctxs[j], status[j], "multi-case solve not converged within the iteration caps");
        if (st != BHIP_OK) fail_seat(seat, r, st);
    }
}

// Charges of every slot but `self`.
double others_charge(bhip_server *srv, size_t self) {
    double t = 0;
    for (size_t i = 0; i < srv->slots.size(); i++)
        if (i != self) t += srv->slots[i].charge;
    return t;
}

// Device bytes slot `s` holds once a pack with footprints fp runs on its
// first fp.size() seats (the other seats keep what they hold).
double slot_charge(const Slot &s, const std::vector<double> &fp) {
    double t = 0;
    const size_t ns = std::max(s.seats.size(), fp.size());
    for (size_t k = 0; k < ns; k++) {
        const double hw = k < s.seats.size() && s.seats[k].ctx ? s.seats[k].hw : 0.0;
        const double f = k < fp.size() ? fp[k] : 0.0;
        t += std::max(hw, f);
    }
    return t;
}

void release_seat(Seat &seat) {
    if (seat.ctx) bhip_destroy(seat.ctx);
    seat.ctx = nullptr;
    seat.have_topo = false;
    seat.hw = 0;
}

void release_slot(Slot &s) {
    for (auto &seat : s.seats) release_seat(seat);
    s.charge = 0;
}

// Under the lock: can idle slot `self` start a pack with footprints fp now?
// Frees idle slots' buffers, then its own seats beyond the pack, when that
// makes the pack fit.
bool admit(bhip_server *srv, size_t self, const std::vector<double> &fp) {
    Slot &me = srv->slots[self];
    auto fits = [&] { return others_charge(srv, self) + slot_charge(me, fp) <= srv->budget; };
    if (fits()) return true;
    for (size_t i = 0; i < srv->slots.size(); i++) {
        Slot &s = srv->slots[i];
        if (i == self || s.busy || s.charge == 0) continue;
        release_slot(s);
        if (fits()) return true;
    }
    for (size_t k = fp.size(); k < me.seats.size(); k++) release_seat(me.seats[k]);
    me.charge = slot_charge(me, {});
    return fits();
}

void worker(bhip_server *srv, size_t self) {
    (void)hipSetDevice(srv->device);
    for (;;) {
        std::vector<Request *> pk;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            std::vector<double> fp;
            srv->cv_work.wait(lk, [&] {
                if (srv->stop) return true;
                if (srv->queue.empty()) return false;
                fp.assign(1, srv->queue.front()->footprint);
                return admit(srv, self, fp);
            });
            if (srv->stop) return;
            Request *head = srv->queue.front();
            srv->queue.pop_front();
            pk.push_back(head);
            if (packable(srv, *head)) {
                while (!srv->queue.empty() && (int)pk.size() < srv->opts.max_pack) {
                    Request *nx = srv->queue.front();
                    if (!packable(srv, *nx) || nx->n_cases != head->n_cases) break;
                    fp.push_back(nx->footprint);
                    if (!admit(srv, self, fp)) {
                        fp.pop_back();
                        break;
                    }
                    srv->queue.pop_front();
                    pk.push_back(nx);
                }
            }
            Slot &s = srv->slots[self];
            s.busy = true;
            s.charge = slot_charge(s, fp);
            const auto now = Clock::now();
            for (Request *r : pk) {
                r->state = 1;
                r->t_start = now;
                r->st.slot = (int)self;
            }
        }
        run_pack(srv, srv->slots[self], pk);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            Slot &s = srv->slots[self];
            s.busy = false;
            s.charge = slot_charge(s, {});
            const auto now = Clock::now();
            for (Request *r : pk) {
                r->t_end = now;
                r->st.queue_ms = ms_between(r->t_submit, r->t_start);
                r->st.service_ms = ms_between(r->t_start, r->t_end);
                r->st.latency_ms = ms_between(r->t_submit, r->t_end);
                r->st.start_ms = ms_between(srv->t0, r->t_start);
                r->st.end_ms = ms_between(srv->t0, r->t_end);
                r->state = 2;
                srv->completed++;
            }
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int bhip_server_create(int device, const bhip_server_opts *opts, bhip_server **out) {
    if (!out) return BHIP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return BHIP_ERR_HIP;
    bhip_server *srv = new bhip_server();
    srv->device = device;
    bhip_server_opts o{};
    if (opts) o = *opts;
    if (o.max_slots <= 0) o.max_slots = 4;
    if (o.max_slots > 16) o.max_slots = 16;
    if (!(o.tol > 0)) o.tol = 1e-8;
    if (o.maxit <= 0) o.maxit = 40000;
    if (!(o.inner_tol > 0)) o.inner_tol = 1e-3;
    if (o.precision != BHIP_PREC_FP64 && o.precision != BHIP_PREC_FP64_BLOCKED) o.precision = BHIP_PREC_FP32_IR;
    if (o.max_pack <= 0) o.max_pack = 8;
    if (o.max_pack > 64) o.max_pack = 64;
    srv->opts = o;
    if (o.hbm_budget > 0) {
        srv->budget = o.hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        (void)hipSetDevice(device);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete srv;
            return BHIP_ERR_HIP;
        }
        srv->budget = 0.8 * (double)fr;
    }
    srv->slots.resize(o.max_slots);
    (void)hipSetDevice(device);
    for (int i = 0; i < o.max_slots; i++) {
        if (bhip_pack_engine_create(device, nullptr, &srv->slots[i].pe) != BHIP_OK) {
            for (int j = 0; j < i; j++) bhip_pack_engine_destroy(srv->slots[j].pe);
            delete srv;
            return BHIP_ERR_HIP;
        }
    }
    for (int i = 0; i < o.max_slots; i++) srv->slots[i].th = std::thread(worker, srv, (size_t)i);
    *out = srv;
    return BHIP_OK;
}

void bhip_server_destroy(bhip_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &s : srv->slots)
        if (s.th.joinable()) s.th.join();
    for (auto &s : srv->slots) {
        release_slot(s);
        bhip_pack_engine_destroy(s.pe);
    }
    for (auto &kv : srv->reqs) delete kv.second;
    delete srv;
}

double bhip_server_budget(const bhip_server *srv) { return srv ? srv->budget : 0.0; }

double bhip_request_footprint(int n_nodes, int n_el, int n_cases) {
    return estimate_bytes(n_nodes, n_el, n_cases);
}

int bhip_server_submit(bhip_server *srv, const bhip_request *q, long long *id) {
    if (!srv || !q || !id) return BHIP_ERR_ARG;
    if (q->n_nodes <= 0 || q->n_el <= 0 || !q->xyz || !q->conn || !q->rest || !q->presc || q->n_cases < 1 ||
        !q->mat || q->n_groups < 1)
        return BHIP_ERR_ARG;
    if (!(q->mat_kind == BHIP_MAT_ISO_ELASTIC || q->mat_kind == BHIP_MAT_CMAT3D)) return BHIP_ERR_UNSUPPORTED;
    const long long n_dof = 3LL * q->n_nodes;
    Request *r = new Request();
    r->n_nodes = q->n_nodes;
    r->n_el = q->n_el;
    r->mat_kind = q->mat_kind;
    r->n_groups = q->n_groups;
    r->n_cases = q->n_cases;
    r->order = q->order > 0 ? q->order : 2;
    r->xyz.assign(q->xyz, q->xyz + 3 * (size_t)q->n_nodes);
    r->conn.assign(q->conn, q->conn + 8 * (size_t)q->n_el);
    r->rest.assign(q->rest, q->rest + (size_t)n_dof);
    for (auto &v : r->rest) v = v ? 1 : 0;
    const int stride = q->mat_kind == BHIP_MAT_ISO_ELASTIC ? 2 : 36;
    r->mat.assign(q->mat, q->mat + (size_t)stride * q->n_groups);
    if (q->elem_group) r->group.assign(q->elem_group, q->elem_group + (size_t)q->n_el);
    r->presc.assign(q->presc, q->presc + (size_t)q->n_cases * n_dof);
    if (q->load) r->load.assign(q->load, q->load + (size_t)n_dof);
    r->footprint = estimate_bytes(q->n_nodes, q->n_el, q->n_cases);
    if (r->footprint > srv->budget) {
        delete r;
        return BHIP_ERR_CAPACITY;
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        r->id = srv->next_id++;
        srv->reqs[r->id] = r;
        srv->queue.push_back(r);
        *id = r->id;
    }
    srv->cv_work.notify_all();
    return BHIP_OK;
}

int bhip_server_poll(bhip_server *srv, long long id, int *done) {
    if (!srv || !done) return BHIP_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return BHIP_ERR_ARG;
    *done = it->second->state == 2;
    return BHIP_OK;
}

int bhip_server_wait(bhip_server *srv, long long id, double *x, bhip_request_stats *stats) {
    if (!srv) return BHIP_ERR_ARG;
    Request *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->reqs.find(id);
        if (it == srv->reqs.end()) return BHIP_ERR_ARG;
        r = it->second;
        srv->cv_done.wait(lk, [&] { return r->state == 2; });
        srv->reqs.erase(it);
    }
    if (x) std::memcpy(x, r->x.data(), sizeof(double) * r->x.size());
    if (stats) *stats = r->st;
    const int st = r->st.status;
    delete r;
    return st;
}

}  // extern "C"
