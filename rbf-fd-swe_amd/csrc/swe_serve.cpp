// Continuous-batching forecast server (include/swe_serve.h) over the
// libswe_hip C ABI.  Host-only C++.
//
// Batches: one batched context per (mesh, batch) with `slots_per_batch`
// member slots, sharing the mesh's device static data (swe_hip_create_ex
// with `share`).  Slots [0, active) are in use; a finished member is read
// back and the last active slot is moved into its place, so the kernel only
// runs the member blocks that hold live members (swe_hip_set_active).
// A batch advances all its live members by the same chunk of steps
// (min(chunk_steps, smallest remaining budget)); members join and leave only
// at chunk boundaries, which never changes any member's arithmetic.
#include "swe_serve.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Slot {
    int req = -1;
    int member = -1;
    int remaining = 0;
};

struct Batch {
    swe_hip_ctx *ctx = nullptr;
    int mesh = -1;
    std::vector<Slot> slots;
    int active = 0;
    bool busy = false;
    int chunk = 0;
    long long bytes = 0;
};

struct Mesh {
    swe_hip_static desc{};
    std::vector<int> mapping;    // host node -> file node (empty: identity)
    int n_nodes = 0;             // real nodes (result files)
    long long static_est = 0;    // bytes of the device static data
    long long slot_state_est = 0;
    long long batch_extra_est = 0;   // per batch context: tile lists and node-major copies
    bool tiled = false;
    std::vector<int> batches;
    long long static_bytes = 0;  // charged while resident
    bool resident = false;
};

struct Request {
    long long tag = 0;
    int mesh = -1;
    int members = 0;
    int nsteps = 0;
    std::vector<const double *> init;
    std::vector<double *> result;
    std::string result_path;     // written at completion when not empty
    int started = 0;
    int done = 0;
    double t_submit = 0.0, t_admit = -1.0;
};

struct Done {
    long long tag;
    double latency, queue;
};

}  // namespace

struct swe_srv {
    swe_srv_config cfg{};
    std::vector<Mesh> meshes;
    std::vector<Batch> batches;   // never shrinks; an evicted batch keeps ctx == nullptr
    std::vector<Request> reqs;
    std::deque<int> pending;      // requests with members not yet admitted
    std::deque<Done> done;
    long long hbm_in_use = 0, hbm_peak = 0;
    long long member_steps = 0, chunks = 0, requests_done = 0, evictions = 0;
};

namespace {

void charge(swe_srv *s, long long bytes)
{
    s->hbm_in_use += bytes;
    s->hbm_peak = std::max(s->hbm_peak, s->hbm_in_use);
}

// The host-layout index of (node, variable) (SVM_LIN_ID, swe_config.h:50-54).
long svm_index(const swe_hip_static &d, int node, int v)
{
    const int S = d.simd;
    return d.cfdl ? 4L * node + v : (long)(node / S) * S * 4 + (long)v * S + node % S;
}

// A finished request's result file, in the history format of swe_run: int32
// N, then one record per member: int32 step (= nsteps), f64 time, f64
// H[N][4] (u, v, w, h) in the file's node order.
int write_result(const Mesh &m, const Request &r)
{
    FILE *fp = fopen(r.result_path.c_str(), "wb");
    if (!fp)
        return SWE_SRV_EIO;
    const int N = m.n_nodes;
    std::vector<double> row(4 * (size_t)N);
    const int32_t n32 = N, st = r.nsteps;
    const double t = r.nsteps * m.desc.dt;
    bool ok = fwrite(&n32, sizeof(n32), 1, fp) == 1;
    for (int k = 0; ok && k < r.members; ++k) {
        const double *H = r.result[k];
        for (int i = 0; i < N; ++i) {
            const int fn = m.mapping.empty() ? i : m.mapping[i];
            for (int v = 0; v < 4; ++v)
                row[4 * (size_t)fn + v] = H[svm_index(m.desc, i, v)];
        }
        ok = fwrite(&st, sizeof(st), 1, fp) == 1 && fwrite(&t, sizeof(t), 1, fp) == 1
             && fwrite(row.data(), sizeof(double), row.size(), fp) == row.size();
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? 0 : SWE_SRV_EIO;
}

// Reads back and releases members whose budget is spent (batch idle).
int retire(swe_srv *s, Batch &b, int *progress)
{
    for (int i = 0; i < b.active;) {
        Slot &sl = b.slots[i];
        if (sl.remaining > 0) {
            ++i;
            continue;
        }
        Request &r = s->reqs[sl.req];
        int rc = swe_hip_get_state(b.ctx, i, r.result[sl.member]);
        if (rc)
            return rc;
        if (++r.done == r.members) {
            if (!r.result_path.empty()) {
                rc = write_result(s->meshes[r.mesh], r);
                if (rc)
                    return rc;
            }
            const double t = now_s();
            s->done.push_back({r.tag, t - r.t_submit, r.t_admit - r.t_submit});
            s->requests_done++;
        }
        const int last = b.active - 1;
        if (i != last) {
            rc = swe_hip_move_member(b.ctx, last, i);
            if (rc)
                return rc;
            b.slots[i] = b.slots[last];
        }
        b.slots[last] = Slot{};
        b.active--;
        ++*progress;
    }
    return 0;
}

// Destroys an idle batch (no live member, no chunk in flight) and returns
// its bytes to the budget; the mesh's static data go too when this was the
// mesh's last batch (the device copy is freed with its last context).
void evict(swe_srv *s, int batch_id)
{
    Batch &b = s->batches[batch_id];
    Mesh &m = s->meshes[b.mesh];
    swe_hip_destroy(b.ctx);
    b.ctx = nullptr;
    s->hbm_in_use -= b.bytes;
    b.bytes = 0;
    m.batches.erase(std::find(m.batches.begin(), m.batches.end(), batch_id));
    if (m.batches.empty() && m.resident) {
        s->hbm_in_use -= m.static_bytes;
        m.static_bytes = 0;
        m.resident = false;
    }
    s->evictions++;
}

int new_batch(swe_srv *s, int mesh_id, int *batch_id)
{
    Mesh &m = s->meshes[mesh_id];
    auto need = [&] {
        return m.slot_state_est * s->cfg.slots_per_batch + m.batch_extra_est
               + (m.resident ? 0 : m.static_est);
    };
    if ((int)m.batches.size() >= s->cfg.max_batches_per_mesh)
        return 1;
    if (s->hbm_in_use + need() > s->cfg.hbm_cap_bytes) {
        // make room from idle batches of other meshes, oldest first
        for (int bi = 0; bi < (int)s->batches.size() && s->hbm_in_use + need() >
                                                            s->cfg.hbm_cap_bytes; ++bi) {
            const Batch &b = s->batches[bi];
            if (b.ctx && !b.busy && b.active == 0 && b.mesh != mesh_id)
                evict(s, bi);
        }
    }
    auto will_free = [&] {   // a live member or a chunk in flight
        for (const Batch &b : s->batches)
            if (b.ctx && (b.busy || b.active > 0))
                return true;
        return false;
    };
    if (s->hbm_in_use + need() > s->cfg.hbm_cap_bytes)
        return will_free() ? 1 : SWE_HIP_ENOMEM;   // waits, or can never fit the budget
    swe_hip_ctx *owner = nullptr;
    for (int bi : m.batches)
        if (s->batches[bi].ctx) {
            owner = s->batches[bi].ctx;
            break;
        }
    swe_hip_options o{s->cfg.slots_per_batch, s->cfg.flags, s->cfg.device, s->cfg.member_block,
                      owner, 0, 0, m.tiled ? SWE_HIP_ORDER_HILBERT : SWE_HIP_ORDER_HOST,
                      m.tiled ? 1 : 0, 0, 0, 0};
    swe_hip_ctx *ctx = nullptr;
    int rc = swe_hip_create_ex(owner ? nullptr : &m.desc, &o, &ctx);
    if (rc == SWE_HIP_ENOMEM)   // the device itself is full: wait while work
        return will_free() ? 1 : SWE_HIP_ENOMEM;   // is in flight, else fail
    if (rc)
        return rc;
    Batch b;
    b.ctx = ctx;
    b.mesh = mesh_id;
    b.slots.assign(s->cfg.slots_per_batch, Slot{});
    b.bytes = swe_hip_device_bytes(ctx);
    charge(s, b.bytes);
    if (!m.resident) {
        m.resident = true;
        m.static_bytes = swe_hip_static_bytes(ctx);
        charge(s, m.static_bytes);
    }
    s->batches.push_back(std::move(b));
    *batch_id = (int)s->batches.size() - 1;
    m.batches.push_back(*batch_id);
    return 0;
}

// Places queued members into free slots of idle batches (FIFO per mesh).
int admit(swe_srv *s, int *progress)
{
    std::vector<char> blocked(s->meshes.size(), 0);
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Request &r = s->reqs[*it];
        if (blocked[r.mesh]) {
            ++it;
            continue;
        }
        Mesh &m = s->meshes[r.mesh];
        size_t k = 0;
        while (r.started < r.members) {
            int target = -1;
            for (; k < m.batches.size(); ++k) {
                Batch &b = s->batches[m.batches[k]];
                if (b.ctx && !b.busy && b.active < (int)b.slots.size()) {
                    target = m.batches[k];
                    break;
                }
            }
            if (target < 0) {
                int nb = -1;
                int rc = new_batch(s, r.mesh, &nb);
                if (rc < 0)
                    return rc;
                if (rc > 0)
                    break;
                target = nb;
                k = m.batches.size() - 1;
            }
            Batch &b = s->batches[target];
            const int slot = b.active;
            int rc = swe_hip_set_state(b.ctx, slot, r.init[r.started]);
            if (rc)
                return rc;
            b.slots[slot] = Slot{*it, r.started, r.nsteps};
            b.active++;
            if (r.started == 0)
                r.t_admit = now_s();
            r.started++;
            ++*progress;
        }
        if (r.started < r.members) {
            blocked[r.mesh] = 1;
            ++it;
        } else {
            it = s->pending.erase(it);
        }
    }
    return 0;
}

int launch(swe_srv *s, Batch &b, int *progress)
{
    if (b.active == 0)
        return 0;
    int chunk = s->cfg.chunk_steps;
    for (int i = 0; i < b.active; ++i)
        chunk = std::min(chunk, b.slots[i].remaining);
    int rc = swe_hip_set_active(b.ctx, b.active);
    if (!rc && chunk > 0)
        rc = swe_hip_rk_steps(b.ctx, chunk);
    if (rc)
        return rc;
    for (int i = 0; i < b.active; ++i)
        b.slots[i].remaining -= chunk;
    s->member_steps += (long long)chunk * b.active;
    s->chunks++;
    b.busy = true;
    b.chunk = chunk;
    ++*progress;
    return 0;
}

}  // namespace

extern "C" {

void swe_srv_default_config(swe_srv_config *c)
{
    c->hbm_cap_bytes = 200LL << 30;
    c->slots_per_batch = 64;
    c->member_block = 16;
    c->chunk_steps = 16;
    c->max_batches_per_mesh = 64;
    c->flags = 0;
    c->device = 0;
    c->tiled = 0;
}

int swe_srv_create(const swe_srv_config *cfg, swe_srv **out)
{
    if (!out || !cfg)
        return SWE_HIP_EINVAL;
    *out = nullptr;
    if (cfg->tiled != 0 && (cfg->tiled != 1 || cfg->member_block != 32 ||
                            cfg->slots_per_batch % 32))
        return SWE_HIP_EINVAL;
    if (cfg->slots_per_batch < 2 || cfg->slots_per_batch % 2 || cfg->chunk_steps < 1
        || cfg->max_batches_per_mesh < 1 || cfg->hbm_cap_bytes <= 0)
        return SWE_HIP_EINVAL;
    swe_srv *s = new swe_srv();
    s->cfg = *cfg;
    *out = s;
    return 0;
}

void swe_srv_destroy(swe_srv *s)
{
    if (!s)
        return;
    // contexts sharing static data first, the owner of each mesh last
    for (auto it = s->batches.rbegin(); it != s->batches.rend(); ++it)
        if (it->ctx)
            swe_hip_destroy(it->ctx);
    delete s;
}

int swe_srv_add_mesh(swe_srv *s, const swe_hip_static *d, int *mesh_id)
{
    return swe_srv_add_mesh_ex(s, d, nullptr, d ? d->compute_size : 0, mesh_id);
}

int swe_srv_add_mesh_ex(swe_srv *s, const swe_hip_static *d, const int *mapping, int n_nodes,
                        int *mesh_id)
{
    if (!s || !d || !mesh_id || d->compute_size <= 0 || d->Nnbr <= 0 || n_nodes < 1
        || n_nodes > d->compute_size)
        return SWE_HIP_EINVAL;
    Mesh m;
    m.desc = *d;
    m.n_nodes = n_nodes;
    if (mapping) {
        m.mapping.assign(mapping, mapping + n_nodes);
        for (int v : m.mapping)
            if (v < 0 || v >= n_nodes)
                return SWE_HIP_EINVAL;
    }
    const long long n = d->compute_size;
    m.static_est = n * ((long long)d->Nnbr * (4 + 32) + 17 * 8);
    m.slot_state_est = n * 4 * 8 * 4;   // H, K (two), D per member slot
    m.tiled = s->cfg.tiled && d->Nnbr == 31;
    // tile lists (<= 70 rows per tile, at least one node) and the node-major
    // weight and geometry copies: at most 2 KB per node
    m.batch_extra_est = m.tiled ? n * 2048 : 0;
    s->meshes.push_back(m);
    *mesh_id = (int)s->meshes.size() - 1;
    return 0;
}

int swe_srv_submit(swe_srv *s, int mesh_id, int members, const double *const *init,
                   double *const *result, int nsteps, long long tag, int *req_id)
{
    return swe_srv_submit_ex(s, mesh_id, members, init, result, nsteps, tag, nullptr, req_id);
}

int swe_srv_submit_ex(swe_srv *s, int mesh_id, int members, const double *const *init,
                      double *const *result, int nsteps, long long tag, const char *result_path,
                      int *req_id)
{
    if (!s || mesh_id < 0 || mesh_id >= (int)s->meshes.size() || members < 1 || !init
        || !result || nsteps < 0 || !req_id)
        return SWE_HIP_EINVAL;
    Request r;
    if (result_path)
        r.result_path = result_path;
    r.tag = tag;
    r.mesh = mesh_id;
    r.members = members;
    r.nsteps = nsteps;
    r.init.assign(init, init + members);
    r.result.assign(result, result + members);
    r.t_submit = now_s();
    s->reqs.push_back(std::move(r));
    *req_id = (int)s->reqs.size() - 1;
    s->pending.push_back(*req_id);
    return 0;
}

int swe_srv_pump(swe_srv *s, int *progress)
{
    if (!s)
        return SWE_HIP_EINVAL;
    int prog = 0;
    for (Batch &b : s->batches) {
        if (!b.ctx || !b.busy)
            continue;
        const int q = swe_hip_query(b.ctx);
        if (q < 0)
            return q;
        if (q == 1) {
            b.busy = false;
            ++prog;
        }
    }
    for (Batch &b : s->batches)
        if (b.ctx && !b.busy) {
            int rc = retire(s, b, &prog);
            if (rc)
                return rc;
        }
    int rc = admit(s, &prog);
    if (rc)
        return rc;
    for (Batch &b : s->batches)
        if (b.ctx && !b.busy) {
            rc = launch(s, b, &prog);
            if (rc)
                return rc;
        }
    if (progress)
        *progress = prog;
    return 0;
}

int swe_srv_wait(swe_srv *s)
{
    if (!s)
        return SWE_HIP_EINVAL;
    for (;;) {
        bool any = false;
        for (Batch &b : s->batches) {
            if (!b.ctx || !b.busy)
                continue;
            any = true;
            const int q = swe_hip_query(b.ctx);
            if (q < 0)
                return q;
            if (q == 1)
                return 0;
        }
        if (!any)
            return 0;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int swe_srv_completed(swe_srv *s, long long *tags, double *latency, double *queue, int max,
                      int *count)
{
    if (!s || !count || max < 0)
        return SWE_HIP_EINVAL;
    int n = 0;
    while (n < max && !s->done.empty()) {
        const Done d = s->done.front();
        s->done.pop_front();
        if (tags) tags[n] = d.tag;
        if (latency) latency[n] = d.latency;
        if (queue) queue[n] = d.queue;
        ++n;
    }
    *count = n;
    return 0;
}

int swe_srv_get_stats(const swe_srv *s, swe_srv_stats *st)
{
    if (!s || !st)
        return SWE_HIP_EINVAL;
    st->member_steps_done = s->member_steps;
    st->requests_done = s->requests_done;
    st->requests_pending = (long long)s->reqs.size() - s->requests_done;
    st->hbm_in_use = s->hbm_in_use;
    st->hbm_peak = s->hbm_peak;
    int nb = 0, nm = 0;
    for (const Batch &b : s->batches)
        nb += b.ctx != nullptr;
    for (const Mesh &m : s->meshes)
        nm += m.resident;
    st->batches = nb;
    st->meshes_resident = nm;
    st->chunks_launched = s->chunks;
    st->evictions = s->evictions;
    return 0;
}

}  // extern "C"
