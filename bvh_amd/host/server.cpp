// server.cpp — continuous-batching build+trace server over the C ABI
// (include/bvh_server.h).  Host C++ only: every device operation goes through
// bvh_hip.h, each slot on its own context (stream + workspace).
#include "bvh_server.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bvh_hip.h"

namespace {

// Device bytes of one request's mesh and tree (with traversal records), from
// the library's own allocation sizes.
uint64_t persistent_bytes(int32_t algo, uint32_t ntri) {
    bvh_footprint f;
    if (bvh_footprint_estimate(algo, ntri, 1, 0, &f) != BVH_OK) return 0;
    return f.mesh + f.tree;
}

// Context workspace a group needs: one build over its new meshes and the
// largest of its traces (traces of a group run one after another on the
// slot's context and reuse the same buffers).
void group_workspace(int32_t algo, uint64_t tris, uint32_t builds, uint32_t max_rays, uint64_t* build,
                     uint64_t* trace) {
    bvh_footprint f;
    std::memset(&f, 0, sizeof(f));
    const uint32_t t32 = (uint32_t)std::min<uint64_t>(tris, 0xFFFFFFFFull);
    if (bvh_footprint_estimate(algo, builds ? t32 : 0, builds ? builds : 1, max_rays, &f) != BVH_OK) {
        *build = *trace = ~0ull / 4;
        return;
    }
    *build = f.build_workspace;
    *trace = f.trace_workspace;
}

// A resident mesh taken out of service by eviction; freed outside the lock.
struct Evicted {
    bvh_mesh_t mesh = nullptr;
    bvh_tree_t tree = nullptr;
    uint64_t bytes = 0;
};

struct Resident {
    uint64_t key = 0;
    bvh_mesh_t mesh = nullptr;
    bvh_tree_t tree = nullptr;
    uint32_t ntri = 0;
    uint64_t bytes = 0;
    int users = 0;        // admitted requests that still need it
    bool ready = false;   // built and complete on the device
    uint64_t last_use = 0;
};

struct Req {
    bvh_request r;
    bool build = false;          // this request builds its mesh
    Resident* res = nullptr;     // keyed residency (built here or used)
    uint64_t transient = 0;      // bytes released at completion
    double admitted = 0.0;
    bvh_mesh_t mesh = nullptr;   // unkeyed build: freed at completion
    bvh_tree_t tree = nullptr;
};

}  // namespace

struct bvh_server_rec {
    bvh_server_config cfg;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req> queue;
    std::deque<bvh_completion> done;
    std::unordered_map<uint64_t, std::unique_ptr<Resident>> residents;
    std::vector<bvh_ctx_t> ctxs;
    std::vector<std::thread> threads;
    bool stop = false;
    // in_use: resident meshes/BVHs + in-flight unkeyed meshes/BVHs; held_*:
    // what each slot's grow-only context workspace holds (or has reserved for
    // the group it runs).  Free memory = budget - in_use - held_total.
    uint64_t in_use = 0, peak = 0, resident_bytes = 0, lru = 0, held_total = 0;
    uint64_t cache_limit = 0, admit_budget = 0;  // device-cache share and what admission may use
    uint64_t cache_limit_before = 0;             // the device cache's limit before this server (restored at destroy)
    std::vector<uint64_t> held_build, held_trace;
    uint32_t in_flight = 0, groups = 0, completed = 0, evictions = 0, rejected = 0, max_group_seen = 0;
    std::string err;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void run_slot(int slot);
    bool form_group(int slot, std::vector<Req>& group, Evicted* ev);
    void set_held(int slot, uint64_t build, uint64_t trace);
    void finish_one(int slot, Req& q, int status, uint32_t gsize);
    void publish_residents(int slot, std::vector<Req>& group, std::vector<int>& status);
    void complete_locked(const Req& q, int status, int slot, uint32_t gsize);
};

extern "C" void bvh_serve_footprint(uint32_t ntri, uint32_t nrays, int32_t algo, uint64_t* transient,
                                    uint64_t* resident) {
    bvh_footprint f;
    std::memset(&f, 0, sizeof(f));
    (void)bvh_footprint_estimate(algo, ntri, 1, nrays, &f);
    if (transient) *transient = (ntri ? f.build_workspace : 0) + f.trace_workspace;
    if (resident) *resident = ntri ? f.mesh + f.tree : 0;
}

extern "C" uint32_t bvh_serve_form_group(const uint64_t* bytes, const uint32_t* tris, const uint8_t* runnable,
                                         uint32_t pending, uint64_t free_bytes, uint32_t max_group,
                                         uint32_t max_group_tris) {
    uint64_t used = 0;
    uint64_t t = 0;
    uint32_t g = 0;
    while (g < pending && g < max_group) {
        if (!runnable[g]) break;
        if (used + bytes[g] > free_bytes) break;
        if (g > 0 && t + tris[g] > max_group_tris) break;
        used += bytes[g];
        t += tris[g];
        ++g;
        if (t > max_group_tris) break;  // an oversized head runs alone
    }
    return g;
}

void bvh_server_rec::complete_locked(const Req& q, int status, int slot, uint32_t gsize) {
    bvh_completion c;
    c.id = q.r.id;
    c.status = status;
    c.slot = slot;
    c.group_size = gsize;
    c.built = q.build ? 1u : 0u;
    c.arrival = q.r.arrival;
    c.admitted = q.admitted;
    c.done = now();
    done.push_back(c);
    ++completed;
    cv_done.notify_all();
}

void bvh_server_rec::set_held(int slot, uint64_t build, uint64_t trace) {
    held_total -= held_build[slot] + held_trace[slot];
    held_build[slot] = build;
    held_trace[slot] = trace;
    held_total += build + trace;
    peak = std::max(peak, in_use + held_total);
}

// Called with the lock held by `slot`.  Takes the next group off the queue,
// or rejects requests that can never run.  When memory is short it takes the
// least recently used idle resident mesh out of service instead and returns
// it in *ev (the caller frees it outside the lock, then calls again).
bool bvh_server_rec::form_group(int slot, std::vector<Req>& group, Evicted* ev) {
    while (!queue.empty()) {
        {  // a head that names no mesh (no vertices, key neither resident nor being built) is rejected
            const bvh_request& r = queue.front().r;
            if (r.soa[0] == nullptr && !(r.mesh_key && residents.count(r.mesh_key))) {
                Req q = queue.front();
                queue.pop_front();
                q.admitted = now();
                err = "request names a mesh_key that is not resident";
                ++rejected;
                complete_locked(q, BVH_E_STATE, -1, 0);
                continue;
            }
        }
        // Candidates, in queue order: the requests that can run now.  A
        // follow-up trace whose mesh is still being built, a second build of
        // a key first built earlier in the window, and a request that names
        // no mesh stay queued (the first two run once the mesh is resident;
        // the last is rejected when it reaches the head), so none of them
        // holds up the requests behind it.
        const uint32_t look = std::min<uint32_t>((uint32_t)queue.size(), 4u * (uint32_t)cfg.max_group);
        std::vector<uint32_t> idx;
        std::vector<uint64_t> bytes;
        std::vector<uint32_t> tris;
        std::vector<Resident*> res;
        std::vector<uint8_t> builds;
        std::unordered_map<uint64_t, uint32_t> new_keys;  // keys first built inside the window
        for (uint32_t i = 0; i < look; ++i) {
            const bvh_request& r = queue[i].r;
            Resident* rp = nullptr;
            if (r.mesh_key) {
                auto it = residents.find(r.mesh_key);
                if (it != residents.end()) rp = it->second.get();
            }
            if (rp) {  // trace against a resident, once it is ready
                if (!rp->ready) continue;
                idx.push_back(i);
                bytes.push_back(0);
                tris.push_back(0);
                res.push_back(rp);
                builds.push_back(0);
            } else if (r.soa[0] != nullptr) {
                if (r.mesh_key && new_keys.count(r.mesh_key)) continue;  // waits for the first build of its key
                if (r.mesh_key) new_keys[r.mesh_key] = i;
                idx.push_back(i);
                bytes.push_back(persistent_bytes(cfg.algo, r.ntri));
                tris.push_back(r.ntri);
                res.push_back(nullptr);
                builds.push_back(1);
            }
        }
        const uint32_t ncand = (uint32_t)idx.size();
        if (ncand == 0) return false;  // everything queued waits for a mesh being built
        const std::vector<uint8_t> run(ncand, 1);
        const uint64_t used = in_use + held_total;
        const uint64_t free_bytes = admit_budget > used ? admit_budget - used : 0;
        // A slot answers its group's requests one after another, so it takes
        // at most its share of the queue (all slots' worth when the queue is
        // deep): the rest stays for the next slot that frees up.
        const uint32_t slots = (uint32_t)std::max(1, cfg.slots);
        const uint32_t share = ((uint32_t)queue.size() + slots - 1) / slots;
        const uint32_t cap = std::min<uint32_t>((uint32_t)cfg.max_group, std::max<uint32_t>(1, share));
        uint32_t g = bvh_serve_form_group(bytes.data(), tris.data(), run.data(), ncand, free_bytes, cap,
                                          cfg.max_group_tris);
        // The group's meshes must fit together with what this slot's context
        // workspace has to grow by for the group's build and traces.
        uint64_t need_build = 0, need_trace = 0;
        while (g > 0) {
            uint64_t persist = 0, t = 0;
            uint32_t nb = 0, max_rays = 0;
            for (uint32_t k = 0; k < g; ++k) {
                persist += bytes[k];
                if (builds[k]) {
                    t += tris[k];
                    ++nb;
                }
                max_rays = std::max(max_rays, queue[idx[k]].r.nrays);
            }
            group_workspace(cfg.algo, t, nb, max_rays, &need_build, &need_trace);
            const uint64_t grow = (need_build > held_build[slot] ? need_build - held_build[slot] : 0) +
                                  (need_trace > held_trace[slot] ? need_trace - held_trace[slot] : 0);
            if (persist + grow <= free_bytes) break;
            --g;
        }
        if (g == 0) {
            // Memory: take the least recently used idle resident out of service.
            Resident* victim = nullptr;
            for (auto& kv : residents) {
                Resident* r = kv.second.get();
                if (r->ready && r->users == 0 && (!victim || r->last_use < victim->last_use)) victim = r;
            }
            if (victim) {
                ev->mesh = victim->mesh;
                ev->tree = victim->tree;
                ev->bytes = victim->bytes;
                ++evictions;
                residents.erase(victim->key);
                return false;
            }
            if (in_flight == 0) {  // the first candidate can never fit
                Req q = queue[idx[0]];
                queue.erase(queue.begin() + idx[0]);
                q.admitted = now();
                err = "request needs more device memory than the budget";
                ++rejected;
                complete_locked(q, BVH_E_BUDGET, -1, 0);
                continue;
            }
            return false;
        }
        const double t = now();
        for (uint32_t k = 0; k < g; ++k) {
            Req q = queue[idx[k]];
            q.admitted = t;
            q.build = builds[k] != 0;
            if (q.build) {
                const uint64_t pb = bytes[k];
                if (q.r.mesh_key) {
                    auto r = std::make_unique<Resident>();
                    r->key = q.r.mesh_key;
                    r->ntri = q.r.ntri;
                    r->bytes = pb;
                    r->users = 1;
                    r->last_use = ++lru;
                    q.res = r.get();
                    residents[q.r.mesh_key] = std::move(r);
                    q.transient = 0;
                    resident_bytes += pb;
                } else {
                    q.transient = pb;
                }
                in_use += pb;
            } else {
                q.res = res[k];
                q.res->users++;
                q.res->last_use = ++lru;
                q.transient = 0;
            }
            group.push_back(q);
        }
        for (uint32_t k = g; k-- > 0;) queue.erase(queue.begin() + idx[k]);  // highest index first
        // Reserve the workspace growth until the group reports what it holds.
        set_held(slot, std::max(held_build[slot], need_build), std::max(held_trace[slot], need_trace));
        in_flight += g;
        ++groups;
        max_group_seen = std::max(max_group_seen, g);
        return true;
    }
    return false;
}

// One request is done (its rays traced, or it failed): release what it held
// and publish its completion.
void bvh_server_rec::finish_one(int slot, Req& q, int status, uint32_t gsize) {
    bvh_ctx_t ctx = ctxs[slot];
    if (q.build && !q.res) {  // unkeyed meshes end with their request (nothing else uses them)
        if (q.tree) bvh_tree_release(ctx, q.tree);
        if (q.mesh) bvh_mesh_release(ctx, q.mesh);
    }
    std::lock_guard<std::mutex> lk(mu);
    in_use -= q.transient;
    if (q.res) q.res->users--;
    complete_locked(q, status, slot, gsize);
    in_flight -= 1;
    cv_work.notify_all();
}

// Publishes (or drops, on failure) the resident meshes this group built, so
// follow-up requests on other slots can start before the group's traces end.
void bvh_server_rec::publish_residents(int slot, std::vector<Req>& group, std::vector<int>& status) {
    bvh_ctx_t ctx = ctxs[slot];
    for (size_t i = 0; i < group.size(); ++i) {
        Req& q = group[i];
        if (q.build && q.res && status[i] == BVH_OK) {
            const int st = bvh_tree_prepare(ctx, q.tree);  // traversal records before other slots use it
            if (st != BVH_OK) status[i] = st;
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < group.size(); ++i) {
        Req& q = group[i];
        if (!(q.build && q.res)) continue;
        if (status[i] == BVH_OK) {
            q.res->mesh = q.mesh;
            q.res->tree = q.tree;
            q.res->ready = true;
            // Count what the mesh and tree really hold (a block reused from
            // the device cache may be larger than the estimate admitted).
            uint64_t mb = 0, tb = 0;
            if (bvh_mesh_bytes(q.mesh, &mb) == BVH_OK && bvh_tree_bytes(q.tree, &tb) == BVH_OK) {
                in_use = in_use - q.res->bytes + (mb + tb);
                resident_bytes = resident_bytes - q.res->bytes + (mb + tb);
                q.res->bytes = mb + tb;
                peak = std::max(peak, in_use + held_total);
            }
        } else {
            if (q.tree) bvh_tree_release(ctx, q.tree);
            if (q.mesh) bvh_mesh_release(ctx, q.mesh);
            q.mesh = nullptr;
            q.tree = nullptr;
            in_use -= q.res->bytes;
            resident_bytes -= q.res->bytes;
            residents.erase(q.res->key);
            q.res = nullptr;
        }
    }
    cv_work.notify_all();
}

void bvh_server_rec::run_slot(int slot) {
    bvh_ctx_t ctx = ctxs[slot];
    while (true) {
        std::vector<Req> group;
        {
            std::unique_lock<std::mutex> lk(mu);
            while (!stop) {
                Evicted ev;
                if (form_group(slot, group, &ev)) break;
                if (ev.mesh || ev.tree) {
                    // Release the evicted mesh without holding the lock; its
                    // bytes stay counted until then.  An idle resident has no
                    // queued work on any slot (its build and every trace of it
                    // have returned), so its blocks go to the device cache
                    // without waiting for the other slots' work.
                    lk.unlock();
                    if (ev.tree) bvh_tree_release(ctx, ev.tree);
                    if (ev.mesh) bvh_mesh_release(ctx, ev.mesh);
                    lk.lock();
                    in_use -= ev.bytes;
                    resident_bytes -= ev.bytes;
                    cv_work.notify_all();
                    continue;
                }
                cv_work.wait(lk);
            }
            if (group.empty()) return;  // stopping
        }
        const uint32_t gsize = (uint32_t)group.size();
        std::vector<int> status(group.size(), BVH_OK);
        // 1. upload the meshes this group builds
        std::vector<bvh_mesh_t> bm;
        std::vector<bvh_tree_t> bt;
        std::vector<size_t> who;
        for (size_t i = 0; i < group.size(); ++i) {
            Req& q = group[i];
            if (!q.build) continue;
            int st = bvh_mesh_upload(ctx, q.r.soa, q.r.ntri, &q.mesh);
            if (st == BVH_OK) st = bvh_mesh_check(q.mesh);  // a non-finite mesh fails alone
            if (st == BVH_OK) st = bvh_tree_create(ctx, &q.tree);
            if (st != BVH_OK) {
                status[i] = st;
                continue;
            }
            bm.push_back(q.mesh);
            bt.push_back(q.tree);
            who.push_back(i);
        }
        // 2. one segmented build for all of them
        if (!bm.empty()) {
            bvh_build_params p;
            std::memset(&p, 0, sizeof(p));
            p.algo = cfg.algo;
            p.ploc_radius = cfg.ploc_radius;
            int st = bm.size() == 1 ? bvh_build(ctx, bm[0], &p, bt[0], nullptr)
                                    : bvh_build_batched(ctx, bm.data(), bt.data(), (uint32_t)bm.size(), &p, nullptr);
            if (st != BVH_OK) {
                std::lock_guard<std::mutex> lk(mu);
                err = bvh_last_error(ctx);
                for (size_t i : who) status[i] = st;
            }
        }
        publish_residents(slot, group, status);
        // 3. each request's rays against its own tree, fewest rays first; a
        //    request completes as soon as its own rays are done
        std::vector<size_t> order(group.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return group[x].r.nrays < group[y].r.nrays; });
        for (size_t i : order) {
            Req& q = group[i];
            if (status[i] == BVH_OK && q.r.nrays) {
                bvh_mesh_t m = q.build ? q.mesh : q.res->mesh;
                bvh_tree_t t = q.build ? q.tree : q.res->tree;
                int st = bvh_trace_host(ctx, t, m, q.r.rays, q.r.nrays, q.r.t_out, q.r.prim_out, nullptr, nullptr);
                if (st != BVH_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    err = bvh_last_error(ctx);
                    status[i] = st;
                }
            }
            finish_one(slot, q, status[i], gsize);
        }
        // What the slot's context holds now replaces the group's reservation.
        uint64_t hb = 0, ht = 0;
        if (bvh_ctx_held_bytes(ctx, &hb, &ht) == BVH_OK) {
            std::lock_guard<std::mutex> lk(mu);
            set_held(slot, hb, ht);
            cv_work.notify_all();
        }
    }
}

extern "C" {

int bvh_server_create(const bvh_server_config* cfg, bvh_server_t* out) {
    if (!cfg || !out) return BVH_E_ARG;
    *out = nullptr;
    if (cfg->algo != BVH_ALGO_LBVH && cfg->algo != BVH_ALGO_PLOC) return BVH_E_UNSUPPORTED;
    if (cfg->algo == BVH_ALGO_PLOC && (cfg->ploc_radius < 1 || cfg->ploc_radius > 1024)) return BVH_E_ARG;
    if (cfg->slots < 0 || cfg->slots > 16 || cfg->max_group < 0) return BVH_E_ARG;
    if (cfg->trace_order != BVH_TRACE_REFERENCE && cfg->trace_order != BVH_TRACE_NEAREST) return BVH_E_ARG;
    auto s = std::make_unique<bvh_server_rec>();
    s->cfg = *cfg;
    if (s->cfg.slots == 0) s->cfg.slots = 4;
    if (s->cfg.max_group == 0) s->cfg.max_group = 64;
    if (s->cfg.max_group_tris == 0) s->cfg.max_group_tris = 1u << 24;
    s->t0 = std::chrono::steady_clock::now();
    s->held_build.assign(s->cfg.slots, 0);
    s->held_trace.assign(s->cfg.slots, 0);
    for (int i = 0; i < s->cfg.slots; ++i) {
        bvh_ctx_t c = nullptr;
        int st = bvh_ctx_create(cfg->device, &c);
        if (st == BVH_OK) st = bvh_ctx_set_trace_order(c, cfg->trace_order);
        if (st != BVH_OK) {
            if (c) bvh_ctx_destroy(c);
            for (bvh_ctx_t x : s->ctxs) bvh_ctx_destroy(x);
            return st;
        }
        s->ctxs.push_back(c);
    }
    if (s->cfg.memory_budget == 0) {
        uint64_t fr = 0, tot = 0;
        if (bvh_device_memory(s->ctxs[0], &fr, &tot) != BVH_OK) {
            for (bvh_ctx_t x : s->ctxs) bvh_ctx_destroy(x);
            return BVH_E_HIP;
        }
        s->cfg.memory_budget = fr / 10 * 9;
    }
    // Released meshes and trees keep their blocks in the library's device
    // cache (at most an eighth of the budget here: a smaller cache fills with
    // blocks no request fits and frees the rest with device-wide waits), and
    // a group's reused blocks may hold more than admission estimated until its
    // residents are counted at their real size (a sixteenth): admission keeps
    // both out of the budget.
    s->cache_limit = s->cfg.memory_budget / 8;
    s->admit_budget = s->cfg.memory_budget - s->cache_limit - s->cfg.memory_budget / 16;
    if (bvh_device_cache_set_limit(cfg->device, s->cache_limit, &s->cache_limit_before) != BVH_OK) {
        for (bvh_ctx_t x : s->ctxs) bvh_ctx_destroy(x);
        return BVH_E_HIP;
    }
    bvh_server_rec* p = s.get();
    for (int i = 0; i < p->cfg.slots; ++i) p->threads.emplace_back([p, i] { p->run_slot(i); });
    *out = s.release();
    return BVH_OK;
}

int bvh_server_submit(bvh_server_t s, const bvh_request* req) {
    if (!s || !req) return BVH_E_ARG;
    const bool has_soa = req->soa[0] != nullptr;
    if (has_soa) {
        if (req->ntri == 0) return BVH_E_ARG;
        for (int k = 0; k < 9; ++k)
            if (!req->soa[k]) return BVH_E_ARG;
    } else if (req->mesh_key == 0) {
        return BVH_E_ARG;
    }
    if (req->nrays && (!req->rays || !req->t_out || !req->prim_out)) return BVH_E_ARG;
    Req q;
    q.r = *req;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return BVH_E_STATE;
        s->queue.push_back(q);
    }
    s->cv_work.notify_all();
    return BVH_OK;
}

int bvh_server_poll(bvh_server_t s, bvh_completion* out, uint32_t max, uint32_t* count, int32_t wait_ms) {
    if (!s || !count || (max && !out)) return BVH_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [s] { return !s->done.empty(); });
    uint32_t n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    *count = n;
    return BVH_OK;
}

double bvh_server_now(bvh_server_t s) { return s ? s->now() : 0.0; }

int bvh_server_get_stats(bvh_server_t s, bvh_server_stats* st) {
    if (!s || !st) return BVH_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    st->budget = s->cfg.memory_budget;
    st->in_use = s->in_use + s->held_total;
    st->peak = s->peak;
    st->resident_bytes = s->resident_bytes;
    st->resident_meshes = (uint32_t)s->residents.size();
    st->queued = (uint32_t)s->queue.size();
    st->in_flight = s->in_flight;
    st->groups = s->groups;
    st->completed = s->completed;
    st->evictions = s->evictions;
    st->rejected = s->rejected;
    st->max_group_seen = s->max_group_seen;
    return BVH_OK;
}

int bvh_server_drain(bvh_server_t s, int32_t wait_ms) {
    if (!s) return BVH_E_ARG;
    const auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(wait_ms);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!(s->queue.empty() && s->in_flight == 0)) {
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout &&
            !(s->queue.empty() && s->in_flight == 0))
            return BVH_E_STATE;
    }
    return BVH_OK;
}

const char* bvh_server_last_error(bvh_server_t s) { return s ? s->err.c_str() : "null server"; }

int bvh_server_destroy(bvh_server_t s) {
    if (!s) return BVH_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->threads) t.join();
    bvh_ctx_t c = s->ctxs.empty() ? nullptr : s->ctxs[0];
    for (auto& kv : s->residents) {
        if (kv.second->tree) bvh_tree_free(c, kv.second->tree);
        if (kv.second->mesh) bvh_mesh_free(c, kv.second->mesh);
    }
    s->residents.clear();
    for (bvh_ctx_t x : s->ctxs) bvh_ctx_destroy(x);
    if (s->cache_limit_before) (void)bvh_device_cache_set_limit(s->cfg.device, s->cache_limit_before, nullptr);
    delete s;
    return BVH_OK;
}

}  // extern "C"
