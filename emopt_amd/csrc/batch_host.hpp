// Host side of batched solves (emf_batch_* entry points).  Included at the end
// of emfdtd.hip: it uses the solver internals defined there.
//
// A batch owns small device arrays (work table, member descriptors, member
// states) and one hipGraph of `spg` steps per lane, captured on first use and
// replayed until every member has finished.  The graph's kernels read
// everything that varies (which solves, which step, what the next step does)
// from those device arrays, so admitting or retiring a solve only rewrites
// them between replays -- the graphs are never re-captured and record no
// memset or copy nodes.
//
// Lanes: the members are split into cache-sized groups, and `lanes` groups
// are stepped at once, each by its own stream, copy of the table and graph.
// Lanes never step the same group, so they share no data; the kernels of one
// lane fill the CUs that another lane's launch leaves idle at its tail, and
// one lane's small per-step kernels (the tick) overlap the other's passes.
#pragma once

namespace {

struct BatchSlot {
    emf_solver *s = nullptr;
    int slot = 0;
    bool finished = false;
    double *fac = nullptr;              // device per-step factors
    emf_solve_info info{};
    std::chrono::steady_clock::time_point t_add;
    long long seq = 0;                  // admission number (table order)
};

}  // namespace

constexpr int BATCH_LANES_MAX = 4;

struct emf_batch {
    std::mutex mtx;
    int dev = 0;
    hipStream_t stream = nullptr;       // lane 0; admissions and table uploads
    hipStream_t lstream[BATCH_LANES_MAX] = {};   // lstream[0] == stream
    int lanes = 2;                      // groups stepped at once
    int lnext[BATCH_LANES_MAX] = {};    // next group of each lane (rotate)
    int spg = 64;                       // steps per graph replay
    int zc = 2;                         // z planes per work-group
    int bpc = 16;                       // persistent work-groups per CU
    int use_graph = 1;                  // 0: launch each replay's kernels directly
    long long group_bytes = 0;          // hot-state bytes per group (0: one group)
    int rotate = 0;                     // 1: groups take turns; 0: first group until done
    long long next_seq = 0;             // admission counter
    int ncu = 0;
    BTable *T = nullptr;                // one copy of the table per lane
    BMember *M = nullptr;
    BState *S = nullptr;
    BTable hT{};
    BMember hM[BATCH_MAX]{};
    BState *hS = nullptr;               // pinned ring of state snapshots
    int ring = 0;
    BatchSlot sl[BATCH_MAX];
    hipGraph_t graph[BATCH_LANES_MAX] = {};
    hipGraphExec_t exec[BATCH_LANES_MAX] = {};
};

constexpr int BATCH_RING = 3 * BATCH_LANES_MAX;   // > replays in flight (2 per lane)

namespace {

// wait for every lane's stream
int batch_sync(emf_batch *b)
{
    for (int l = 0; l < BATCH_LANES_MAX; ++l)
        if (b->lstream[l]) HIPCHK(hipStreamSynchronize(b->lstream[l]));
    return EMF_OK;
}

int batch_upload_table(emf_batch *b)
{
    BTable &t = b->hT;
    t.nact = 0;
    t.total = 0;
    // entries in admission order (slots are reused, so slot order is not
    // arrival order): the first group then holds the oldest solves
    int order[BATCH_MAX], na = 0;
    for (int m = 0; m < BATCH_MAX; ++m)
        if (b->sl[m].s && !b->sl[m].finished) order[na++] = m;
    std::sort(order, order + na, [&](int x, int y) { return b->sl[x].seq < b->sl[y].seq; });
    for (int q = 0; q < na; ++q) {
        const int m = order[q];
        const BMember &B = b->hM[m];
        const long long tiles = (long long)B.ntx * B.nty * B.ntz;
        if ((long long)t.total + tiles > 0x7fffffffLL) EMF_FAIL(EMF_ERR_ARG, "batch too large");
        t.mem[t.nact] = m;
        t.off[t.nact] = t.total;
        t.total += (int)tiles;
        ++t.nact;
    }
    t.off[t.nact] = t.total;
    // groups: consecutive entries whose hot state (E, H, coefficients: ~80
    // B/cell) stays within group_bytes
    t.ngroups = 0;
    long long acc = 0;
    for (int e = 0; e < t.nact; ++e) {
        const long long hot = 80LL * b->hM[t.mem[e]].n;
        if (e == 0 || (b->group_bytes > 0 && acc + hot > b->group_bytes)) {
            t.gent[t.ngroups++] = e;
            acc = 0;
        }
        acc += hot;
    }
    if (t.ngroups == 0) t.gent[t.ngroups++] = 0;
    t.gent[t.ngroups] = t.nact;
    t.cur = 0;
    for (int l = 0; l < BATCH_LANES_MAX; ++l) {
        b->lnext[l] = l;
        HIPCHK(hipMemcpyAsync(b->T + l, &t, sizeof(BTable), hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));   // t is rewritten by the next call
    return EMF_OK;
}

// the group lane l steps in its next replay (-1: none).  Lane l steps the
// groups g with g % lanes == l: always the same one (the first `lanes`
// groups until their members finish), or those in turn with rotate.
int batch_lane_group(emf_batch *b, int l)
{
    const int ng = b->hT.ngroups;
    if (l >= ng) return -1;
    if (!b->rotate) return l;
    int g = b->lnext[l];
    if (g >= ng) g = l;
    b->lnext[l] = g + b->lanes < ng ? g + b->lanes : l;
    return g;
}

// the kernels of one replay of lane l: spg steps (H, E, tick), then the
// phasor pass and the finish marks, on the lane's stream and table
void batch_enqueue(emf_batch *b, int l)
{
    const unsigned blocks = (unsigned)(b->ncu * b->bpc);
    hipStream_t st = b->lstream[l];
    BTable *T = b->T + l;
    for (int q = 0; q < b->spg; ++q) {
        hipLaunchKernelGGL(k_batch_h, dim3(blocks), dim3(PX, BTY), 0, st, T, b->M, b->S);
        hipLaunchKernelGGL(k_batch_e, dim3(blocks), dim3(PX, BTY), 0, st, T, b->M, b->S);
        hipLaunchKernelGGL(k_batch_tick, dim3(BATCH_MAX), dim3(64), 0, st, T, b->M, b->S);
    }
    hipLaunchKernelGGL(k_batch_phasor, dim3(blocks), dim3(PX, BTY), 0, st, T, b->M, b->S);
    hipLaunchKernelGGL(k_batch_finish, dim3(BATCH_MAX / 64), dim3(64), 0, st, T, b->S);
}

int batch_graph(emf_batch *b)
{
    if (!b->use_graph) return EMF_OK;
    for (int l = 0; l < b->lanes; ++l) {
        if (b->exec[l]) continue;
        HIPCHK(hipStreamBeginCapture(b->lstream[l], hipStreamCaptureModeThreadLocal));
        batch_enqueue(b, l);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->lstream[l], &g);
        if (ec != hipSuccess || !g)
            EMF_FAIL(EMF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        b->graph[l] = g;
        HIPCHK(hipGraphInstantiate(&b->exec[l], b->graph[l], nullptr, nullptr, 0));
    }
    return EMF_OK;
}

void batch_drop_graph(emf_batch *b)
{
    for (int l = 0; l < BATCH_LANES_MAX; ++l) {
        if (b->exec[l]) (void)hipGraphExecDestroy(b->exec[l]);
        if (b->graph[l]) (void)hipGraphDestroy(b->graph[l]);
        b->exec[l] = nullptr;
        b->graph[l] = nullptr;
    }
}

// a member whose state shows it finished: fill its solve info and hand the
// result back to its solver handle
void batch_finish(emf_batch *b, int m, const BState &st)
{
    BatchSlot &q = b->sl[m];
    const BMember &B = b->hM[m];
    q.finished = true;
    q.info.n0 = st.n0;
    q.info.Tn = B.Tn;
    q.info.steps = st.n0 + 2 * B.Tn;
    q.info.status = st.status;
    if (st.status == 1 && st.nchecks > 0 && !(std::isfinite(st.a_change) && std::isfinite(st.phi_change)))
        q.info.status = 3;   // stopped at Nmax with non-finite samples, as emf_solve reports
    q.info.nchecks = st.nchecks;
    q.info.a_change = st.a_change;
    q.info.phi_change = st.phi_change;
    q.info.seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - q.t_add).count();
    std::lock_guard<std::mutex> lks(q.s->mtx);
    q.s->nhist = std::min(st.nchecks, q.s->hist_cap);
}

}  // namespace

extern "C" {

int emf_batch_new(int device, emf_batch **out)
{
    if (!out) EMF_FAIL(EMF_ERR_ARG, "null output");
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    emf_batch *b = new emf_batch();
    b->dev = device;
    int rc = EMF_OK;
    if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        rc = EMF_ERR_HIP;
    // every lane's stream at the default priority
    for (int l = 0; l < BATCH_LANES_MAX && rc == EMF_OK; ++l)
        if (hipStreamCreateWithFlags(&b->lstream[l], hipStreamNonBlocking) != hipSuccess)
            rc = EMF_ERR_HIP;
    b->stream = b->lstream[0];
    if (rc == EMF_OK) rc = dalloc(&b->T, BATCH_LANES_MAX);
    if (rc == EMF_OK) rc = dalloc(&b->M, BATCH_MAX);
    if (rc == EMF_OK) rc = dalloc(&b->S, BATCH_MAX);
    if (rc == EMF_OK &&
        hipHostMalloc((void **)&b->hS, BATCH_RING * BATCH_MAX * sizeof(BState)) != hipSuccess)
        rc = EMF_ERR_NOMEM;
    if (rc == EMF_OK) {
        memset(b->hS, 0, BATCH_RING * BATCH_MAX * sizeof(BState));
        rc = batch_upload_table(b);
    }
    if (rc != EMF_OK) {
        dfree(b->T); dfree(b->M); dfree(b->S);
        if (b->hS) (void)hipHostFree(b->hS);
        for (auto &st : b->lstream)
            if (st) (void)hipStreamDestroy(st);
        delete b;
        EMF_FAIL(rc, "batch creation failed");
    }
    *out = b;
    return EMF_OK;
}

int emf_batch_delete(emf_batch *b)
{
    if (!b) return EMF_OK;
    {
        std::lock_guard<std::mutex> lk(b->mtx);
        (void)hipSetDevice(b->dev);
        (void)batch_sync(b);
        batch_drop_graph(b);
        for (auto &q : b->sl) {
            if (q.s) {
                std::lock_guard<std::mutex> lks(q.s->mtx);
                q.s->in_batch = nullptr;
            }
            dfree(q.fac);
        }
        dfree(b->T); dfree(b->M); dfree(b->S);
        if (b->hS) (void)hipHostFree(b->hS);
        for (auto &st : b->lstream)
            if (st) (void)hipStreamDestroy(st);
    }
    delete b;
    return EMF_OK;
}

#define BLOCK(b)                                                    \
    if (!(b)) EMF_FAIL(EMF_ERR_ARG, "null batch handle");          \
    std::lock_guard<std::mutex> blk_((b)->mtx);                     \
    if (hipSetDevice((b)->dev) != hipSuccess)                       \
        EMF_FAIL(EMF_ERR_HIP, "hipSetDevice(%d) failed", (b)->dev);

int emf_batch_set_tuning(emf_batch *b, int steps_per_graph, int zc, int blocks_per_cu)
{
    BLOCK(b);
    if (steps_per_graph < 1 || steps_per_graph > 4096) EMF_FAIL(EMF_ERR_ARG, "steps_per_graph must be 1..4096");
    if (zc < 1) EMF_FAIL(EMF_ERR_ARG, "zc must be >= 1");
    if (blocks_per_cu < 1 || blocks_per_cu > 64) EMF_FAIL(EMF_ERR_ARG, "blocks_per_cu must be 1..64");
    for (auto &q : b->sl)
        if (q.s && !q.finished) EMF_FAIL(EMF_ERR_STATE, "set the tuning while no solve is running");
    CHECK(batch_sync(b));
    if (steps_per_graph != b->spg || blocks_per_cu != b->bpc) batch_drop_graph(b);
    b->spg = steps_per_graph;
    b->zc = zc;
    b->bpc = blocks_per_cu;
    return EMF_OK;
}

int emf_batch_set_graph(emf_batch *b, int use_graph)
{
    BLOCK(b);
    if (use_graph != 0 && use_graph != 1) EMF_FAIL(EMF_ERR_ARG, "use_graph must be 0 or 1");
    CHECK(batch_sync(b));
    b->use_graph = use_graph;
    return EMF_OK;
}

int emf_batch_set_lanes(emf_batch *b, int lanes)
{
    BLOCK(b);
    if (lanes < 1 || lanes > BATCH_LANES_MAX)
        EMF_FAIL(EMF_ERR_ARG, "lanes must be 1..%d", BATCH_LANES_MAX);
    CHECK(batch_sync(b));
    b->lanes = lanes;
    for (int l = 0; l < BATCH_LANES_MAX; ++l) b->lnext[l] = l;
    return EMF_OK;
}

int emf_batch_set_grouping(emf_batch *b, int64_t group_bytes, int rotate)
{
    BLOCK(b);
    if (group_bytes < 0) EMF_FAIL(EMF_ERR_ARG, "group_bytes must be >= 0");
    if (rotate != 0 && rotate != 1) EMF_FAIL(EMF_ERR_ARG, "rotate must be 0 or 1");
    CHECK(batch_sync(b));
    b->group_bytes = group_bytes;
    b->rotate = rotate;
    return batch_upload_table(b);
}

int emf_batch_add(emf_batch *b, emf_solver *s, int slot, double rtol, double nmax, int64_t Tn,
                  int64_t fixed_n0, int *member)
{
    BLOCK(b);
    if (!s || !member) EMF_FAIL(EMF_ERR_ARG, "null solver or output");
    std::lock_guard<std::mutex> lks(s->mtx);
    if (s->in_batch) EMF_FAIL(EMF_ERR_ARG, "solver is already a member of a batch");
    if (s->dev != b->dev) EMF_FAIL(EMF_ERR_ARG, "solver is on device %d, batch on %d", s->dev, b->dev);
    if (!(rtol > 0)) EMF_FAIL(EMF_ERR_ARG, "rtol must be > 0");
    if (Tn < 1) EMF_FAIL(EMF_ERR_ARG, "Tn must be >= 1 in a batch");
    int id = -1;
    for (int m = 0; m < BATCH_MAX && id < 0; ++m)
        if (!b->sl[m].s) id = m;
    if (id < 0) EMF_FAIL(EMF_ERR_STATE, "batch is full (%d solves)", BATCH_MAX);
    CHECK(check_ready(s));
    if (s->ceps || !s->mu_uniform)
        EMF_FAIL(EMF_ERR_ARG, "batched solves need real eps and uniform mu");
    if (s->wide)
        EMF_FAIL(EMF_ERR_ARG, "a batch member takes at most %llu cells (32-bit offsets)",
                 ((1ULL << 32) - 1) / 24ULL);
    const long long n_stop = nmax < 0 ? 0 : (long long)floor(std::min(nmax, 1e15)) + 1;
    const long long nlast = fixed_n0 >= 0 ? fixed_n0 : n_stop;
    const long long nfac = nlast + 2 * Tn + 1;
    if (nfac > (1LL << 22)) EMF_FAIL(EMF_ERR_ARG, "step limit %lld too large for a batch", nlast);

    // the solver's state as emf_solve sets it up (emopt/fdtd.py:1016-1018)
    CHECK(ensure_slot(s, slot));
    CHECK(build_samples(s));
    PassTab th, te;
    CHECK(refresh_tables(s, &th, &te));
    CHECK(refresh_sources(s));
    long long cap = std::min<long long>(nlast / Tn + 4, 1LL << 20);
    if (cap > s->hist_cap) {
        dfree(s->hist);
        CHECK(dalloc(&s->hist, 3 * (size_t)cap));
        s->hist_cap = (int)cap;
    }
    set_current(s, 0);
    CHECK(reset_psi(s));
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemsetAsync(s->E[c], 0, s->n * sizeof(double), s->stream));
        HIPCHK(hipMemsetAsync(s->H[c], 0, s->n * sizeof(double), s->stream));
    }
    HIPCHK(hipMemsetAsync(s->conv, 0, sizeof(ConvState), s->stream));
    s->nhist = 0;

    // per-step factors, computed exactly as the single-solve launches do
    std::vector<double> f((size_t)nfac * NFAC);
    for (long long q = 0; q < nfac; ++q) {
        const double t = (double)q * s->dt, tE = ((double)q + 0.5) * s->dt;
        double *r = &f[(size_t)q * NFAC];
        r[0] = sin(t); r[1] = cos(t); r[2] = env_at(s, t);
        r[3] = sin(tE); r[4] = cos(tE); r[5] = env_at(s, tE);
        const double h = t - 0.5 * s->dt;
        r[6] = sin(h); r[7] = cos(h);
    }
    BatchSlot &q = b->sl[id];
    CHECK(dalloc(&q.fac, f.size()));
    int rc = hipMemcpyAsync(q.fac, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice,
                            s->stream) == hipSuccess ? EMF_OK : EMF_ERR_HIP;
    if (rc == EMF_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = EMF_ERR_HIP;
    if (rc != EMF_OK) { dfree(q.fac); EMF_FAIL(rc, "upload of the step factors failed"); }

    BMember &B = b->hM[id];
    B.A = make_args(s, 0, 0);
    B.g = s->geo();
    B.g.zc = b->zc;
    B.g.order = 0;
    B.blk = s->src_desc;
    B.nsrc = (int)s->srcs.size();
    src_union(s, B.sub);
    B.ntx = (s->Nx + PX - 1) / PX;
    B.nty = (s->Ny + BTY - 1) / BTY;
    B.ntz = (s->Nz + b->zc - 1) / b->zc;
    B.fac = q.fac;
    B.nfac = nfac;
    B.res = reinterpret_cast<double *>(s->res[slot][0]);
    B.n = s->n;
    B.n_stop = n_stop;
    B.fixed_n0 = fixed_n0;
    B.Tn = Tn;
    B.amp_rtol = rtol;
    B.phi_rtol = sqrt(rtol);
    B.dt = s->dt;
    B.samples = s->samples;
    B.nsamples = s->nsamples;
    B.hist_cap = s->hist_cap;
    B.conv = s->conv;
    B.hist = s->hist;

    BState st{};
    st.mode = BM_STEP;
    st.phase = BP_CONV;
    if (nlast == 0) {                    // no steps before the t0 snapshot
        st.mode = BM_CAP0;
        st.phase = BP_RUN1;
        st.status = fixed_n0 >= 0 ? 2 : 1;
    }
    // the batch's kernels may still be reading M/S/T: wait for them
    CHECK(batch_sync(b));
    HIPCHK(hipMemcpyAsync(&b->M[id], &B, sizeof(BMember), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(&b->S[id], &st, sizeof(BState), hipMemcpyHostToDevice, b->stream));
    q.s = s;
    s->in_batch = b;
    q.slot = slot;
    q.finished = false;
    q.info = emf_solve_info{};
    q.t_add = std::chrono::steady_clock::now();
    q.seq = b->next_seq++;
    rc = batch_upload_table(b);           // synchronises the stream (B, st are locals)
    if (rc != EMF_OK) {
        q.s = nullptr;
        s->in_batch = nullptr;
        dfree(q.fac);
        return rc;
    }
    *member = id;
    return EMF_OK;
}

int emf_batch_run(emf_batch *b, int max_graphs, int until, int *finished)
{
    BLOCK(b);
    if (max_graphs < 0 || until < 0 || until > 2) EMF_FAIL(EMF_ERR_ARG, "bad arguments");
    int nfin = 0;
    auto running = [&]() {
        int r = 0;
        for (auto &q : b->sl) r += q.s && !q.finished;
        return r;
    };
    if (running() > 0) {
        CHECK(batch_graph(b));
        // a replay in flight: its lane, snapshot slot, group and end event
        struct Pend { int lane, r, g; hipEvent_t ev; };
        std::vector<Pend> pend;
        int inflight[BATCH_LANES_MAX] = {};
        int launched = 0;
        bool dirty = false, stop = false;
        int rc = EMF_OK;
        auto retire = [&](const Pend &p) {
            if (hipEventSynchronize(p.ev) != hipSuccess) rc = EMF_ERR_HIP;
            (void)hipEventDestroy(p.ev);
            --inflight[p.lane];
            // only the replay's own group: the snapshot was taken while other
            // lanes were stepping theirs.  The table is the one the replay
            // ran on (it is only rewritten when nothing is in flight).
            const BState *snap = b->hS + (size_t)p.r * BATCH_MAX;
            for (int e = b->hT.gent[p.g]; e < b->hT.gent[p.g + 1]; ++e) {
                const int m = b->hT.mem[e];
                BatchSlot &q = b->sl[m];
                if (q.s && !q.finished && snap[m].phase == BP_DONE) {
                    batch_finish(b, m, snap[m]);
                    ++nfin;
                    dirty = true;
                }
            }
        };
        while (rc == EMF_OK) {
            if (until == 1 && nfin > 0) stop = true;
            if (until == 2 && running() == 0) stop = true;
            if (dirty && !stop) {
                // compact the work table: wait for the replays in flight first
                for (auto &p : pend) retire(p);
                pend.clear();
                if (rc == EMF_OK) rc = batch_upload_table(b);
                dirty = false;
                if (running() == 0) stop = true;
                continue;
            }
            // up to two replays in flight per lane, each lane on its group
            for (int l = 0; l < b->lanes && !stop && rc == EMF_OK; ++l) {
                while (rc == EMF_OK && inflight[l] < 2 && (until == 2 || launched < max_graphs)) {
                    const int g = batch_lane_group(b, l);
                    if (g < 0) break;
                    hipStream_t st = b->lstream[l];
                    hipEvent_t ev;
                    hipLaunchKernelGGL(k_batch_set_group, dim3(1), dim3(64), 0, st, b->T + l, g);
                    hipError_t le = hipSuccess;
                    if (b->use_graph) le = hipGraphLaunch(b->exec[l], st);
                    else batch_enqueue(b, l);
                    if (hipGetLastError() != hipSuccess || le != hipSuccess ||
                        hipMemcpyAsync(b->hS + (size_t)b->ring * BATCH_MAX, b->S,
                                       BATCH_MAX * sizeof(BState), hipMemcpyDeviceToHost,
                                       st) != hipSuccess ||
                        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                        hipEventRecord(ev, st) != hipSuccess) {
                        rc = EMF_ERR_HIP;
                        break;
                    }
                    pend.push_back({ l, b->ring, g, ev });
                    ++inflight[l];
                    b->ring = (b->ring + 1) % BATCH_RING;
                    ++launched;
                }
            }
            if (pend.empty()) break;
            const Pend p = pend.front();
            pend.erase(pend.begin());
            retire(p);
        }
        for (auto &p : pend) retire(p);
        if (rc == EMF_OK && dirty) rc = batch_upload_table(b);
        if (rc != EMF_OK) EMF_FAIL(rc, "batch run failed");
    }
    if (finished) *finished = nfin;
    return EMF_OK;
}

int emf_batch_poll(emf_batch *b, int member, emf_solve_info *info)
{
    BLOCK(b);
    if (member < 0 || member >= BATCH_MAX || !b->sl[member].s)
        EMF_FAIL(EMF_ERR_ARG, "no member %d in this batch", member);
    const BatchSlot &q = b->sl[member];
    if (!q.finished) return 0;
    if (info) *info = q.info;
    return 1;
}

int emf_batch_remove(emf_batch *b, int member)
{
    BLOCK(b);
    if (member < 0 || member >= BATCH_MAX || !b->sl[member].s)
        EMF_FAIL(EMF_ERR_ARG, "no member %d in this batch", member);
    BatchSlot &q = b->sl[member];
    const bool was_running = !q.finished;
    CHECK(batch_sync(b));
    {
        std::lock_guard<std::mutex> lks(q.s->mtx);
        q.s->in_batch = nullptr;
    }
    dfree(q.fac);
    q = BatchSlot{};
    if (was_running) CHECK(batch_upload_table(b));
    return EMF_OK;
}

int emf_batch_count(emf_batch *b, int *running, int *finished)
{
    BLOCK(b);
    int r = 0, f = 0;
    for (auto &q : b->sl) {
        if (!q.s) continue;
        if (q.finished) ++f;
        else ++r;
    }
    if (running) *running = r;
    if (finished) *finished = f;
    return EMF_OK;
}

}  // extern "C"
