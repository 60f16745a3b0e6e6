// opo_batch.cpp -- batched solves behind the C ABI (opo_batch_* of
// include/opo3d.h): admission by HBM footprint, per-size-class work lists,
// the round trip of every resident solve captured as one HIP graph, its
// profiling and tuning.  Replaces the reference's one-process-per-point
// sweep (src/cuOPO3D.sh:56-86).
#include "opo_solve.hpp"

using namespace opo;

// ----------------------------------------------------------------- batch
// A batch keeps many solves resident and advances all of them one round trip
// per opo_batch_step.  Each pass is one launch per size class over a work
// list of (solve, tile) items; the trip is captured once as a HIP graph and
// re-captured only when the set of resident solves changes.
struct BatchEntry {
    uint64_t ticket = 0;
    opo_grid g{};
    opo_params p{};
    std::vector<unsigned char> api, as, ai;  // host copies while queued
    std::unique_ptr<SolveImpl> s;
    double *d_series = nullptr;
    int n_trips = 0, done_trips = 0;
    size_t footprint = 0;
    bool finished = false;
    int error = OPO_OK;
    std::string error_msg;
};

struct opo_batch {
    int device = 0;
    size_t budget = 0, used = 0;
    int prec = OPO_FP32;
    uint64_t next_ticket = 1;
    std::deque<BatchEntry *> queue;
    std::vector<BatchEntry *> resident;
    std::map<uint64_t, std::unique_ptr<BatchEntry>> all;
    std::mutex mu;
    hipStream_t stream = nullptr;
    // Concurrent size classes (opo_batch_concurrent_classes): within each
    // stage of the trip graph the launches of the different size classes
    // are captured as parallel branches on side streams (the classes of a
    // stage touch disjoint solves), joined before the next stage.
    bool concurrent = false;
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    std::vector<hipEvent_t> sync_ev;   // fork/join events, distinct per use in one capture
    size_t sync_used = 0;
    int next_sync_event(hipEvent_t &e) {
        if (sync_used >= sync_ev.size()) {
            hipEvent_t x;
            CK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            sync_ev.push_back(x);
        }
        e = sync_ev[sync_used++];
        return OPO_OK;
    }
    // device-side descriptors, rebuilt when the resident set changes
    void *d_desc = nullptr;
    int *d_status = nullptr;
    TripState *d_ts = nullptr;
    WorkItem *d_work = nullptr;
    double *d_partial = nullptr;
    struct Range {
        int cls, off, n;
        int v;   // launch variant the class's work list was built for
        // Work items of slice k: the class's solves are listed by slice
        // count, most first, so slice k runs the prefix nk[k] of the list
        // (the solves with more than k slices).
        std::vector<int> nk;
    };
    std::vector<Range> trange, xyrange, fsrange;
    Range prange{0, 0, 0, 0, {}};
    int n_slices = -1, max_chunk = 1;
    hipGraphExec_t graph = nullptr;
    bool dirty = true;
    std::vector<int> built_variants;  // tuned launch variants the graph uses

    // Variant of the fused pass over NT class nt: the tuned one when its
    // tile divides the NX of every resident solve of the class, else
    // variant 0 (its tile divides every supported NX).
    int class_time_variant(int cls) const {
        if (cls < 0) return kBatchPlaneVariant;
        const bool fp64 = prec == OPO_FP64;
        const int v = tune_get(fp64, 0, cls);
        for (const BatchEntry *e : resident)
            if (tclass(e) == cls && time_variant_for(fp64, cls, e->g.nx) != v) return 0;
        return v;
    }
    // Size class of a solve's t pass: NT, or -NT for plane-wave solves
    // (NX = NY = 1: one pixel per workgroup, kBatchPlaneVariant).
    static int tclass(const BatchEntry *e) { return e->g.nx == 1 ? -e->g.nt : e->g.nt; }
    // Whether a solve takes the xy pass of class nx (plane-wave solves skip it).
    static bool in_xy(const BatchEntry *e, int nx) { return e->g.nx == nx && !e->p.no_diffraction; }
    std::vector<int> current_variants() const {
        std::vector<int> v;
        const bool fp64 = prec == OPO_FP64;
        for (const Range &r : trange) v.push_back(class_time_variant(r.cls));
        for (const Range &r : xyrange) v.push_back(tune_get(fp64, 1, r.cls));
        return v;
    }
    std::vector<TripState> h_ts;
    std::vector<int> h_status;

    // Kernel timing inside the replayed trip graph (opo_batch_profile):
    // events recorded around every launch at capture, read after each step.
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_entry;    // entry index of the pair (2k, 2k + 1)
    size_t ev_used = 0;
    std::vector<opo_stage_time> prof;
    std::vector<double> prof_alg;   // algorithmic bytes summed over the timed launches of an entry
    std::vector<double> ev_alg;     // algorithmic bytes of the launch of pair k
    int prof_entry(int stage, int size, double alg) {
        for (size_t k = 0; k < prof.size(); ++k)
            if (prof[k].stage == stage && prof[k].size == size) {
                prof[k].alg_bytes = alg;
                return (int)k;
            }
        prof.push_back(opo_stage_time{stage, size, 0, 0.0, alg});
        prof_alg.push_back(0.0);
        return (int)prof.size() - 1;
    }
    void prof_reset() {
        for (opo_stage_time &x : prof) x.launches = 0, x.ms = 0.0;
        prof_alg.assign(prof.size(), 0.0);
    }
    int mark(int stage, int size, double alg, hipStream_t st) {
        if (!profiling) return OPO_OK;
        if (ev_used + 1 >= ev_pool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                CK(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
        }
        if (ev_used % 2 == 0) {
            ev_entry.push_back(prof_entry(stage, size, alg));
            ev_alg.push_back(alg);
        }
        // recorded as an event node of the captured trip graph
        CK(hipEventRecordWithFlags(ev_pool[ev_used++], st, hipEventRecordExternal));
        return OPO_OK;
    }
    // After a step: add the elapsed time of every bracketed launch.
    // A runtime that cannot time event nodes of a graph leaves the sums at
    // zero and the reason in prof_error; the step itself is unaffected.
    std::string prof_error;
    int collect() {
        if (!profiling || !prof_error.empty()) return OPO_OK;
        for (size_t k = 0; k + 1 < ev_used; k += 2) {
            float t = 0.f;
            const hipError_t e = hipEventElapsedTime(&t, ev_pool[k], ev_pool[k + 1]);
            if (e != hipSuccess) {
                prof_error = std::string("timing the events of the trip graph: ") + hipGetErrorString(e);
                (void)hipGetLastError();
                prof_reset();
                return OPO_OK;
            }
            prof_alg[ev_entry[k / 2]] += ev_alg[k / 2];
            opo_stage_time &x = prof[ev_entry[k / 2]];
            x.ms += t;
            x.launches += 1;
        }
        return OPO_OK;
    }
    // Algorithmic bytes of one launch over the resident solves selected.
    template <typename Pred> double class_bytes(int fields, int rw, Pred pick) const {
        const double cb = prec == OPO_FP64 ? 16.0 : 8.0;
        double cells = 0.0;
        for (const BatchEntry *e : resident)
            if (pick(e)) cells += (double)e->g.nx * e->g.ny * e->g.nt;
        return cells * fields * cb * rw;
    }

    ~opo_batch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto e : sync_ev) (void)hipEventDestroy(e);
        for (auto &x : side)
            if (x) (void)hipStreamDestroy(x);
        for (auto &kv : all)
            if (kv.second->d_series) (void)hipFree(kv.second->d_series);
        all.clear();
        (void)hipFree(d_desc);
        (void)hipFree(d_status);
        (void)hipFree(d_ts);
        (void)hipFree(d_work);
        (void)hipFree(d_partial);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Power partial blocks of a solve: as many as the solve uses alone, so
    // its per-trip powers are bitwise the same in the batch.
    static int chunks_for(const opo_grid &g) { return power_blocks((size_t)g.nx * g.ny * g.nt); }

    // Device buffers of the resident set, grown (never shrunk) as it needs:
    // an admission or a finished solve reuses them without freeing.
    size_t cap_desc = 0, cap_status = 0, cap_ts = 0, cap_work = 0, cap_partial = 0;
    template <typename U> int ensure_buf(U *&ptr, size_t &cap, size_t bytes) {
        if (bytes <= cap) return OPO_OK;
        (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, (size_t)4096);
        if (hipMalloc((void **)&ptr, want) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "batch descriptor allocation of %zu bytes failed", want);
        cap = want;
        return OPO_OK;
    }
    // Rebuild bookkeeping: count, graphs updated in place, host seconds.
    int n_rebuilds = 0, n_updated = 0;
    double rebuild_s = 0.0;

    int rebuild() {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = rebuild_impl();
        rebuild_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        n_rebuilds += 1;
        return rc;
    }

    int rebuild_impl() {
        const int n = (int)resident.size();
        const size_t dsz = prec == OPO_FP64 ? sizeof(SolveDesc<double>) : sizeof(SolveDesc<float>);
        std::vector<unsigned char> hdesc(dsz * n);
        max_chunk = 1;
        for (int i = 0; i < n; ++i) max_chunk = std::max(max_chunk, chunks_for(resident[i]->g));
        for (int i = 0; i < n; ++i) {
            BatchEntry *e = resident[i];
            e->s->fill_desc(hdesc.data() + dsz * i, e->d_series, e->n_trips, chunks_for(e->g));
        }
        // work lists grouped by size class
        std::vector<WorkItem> work;
        trange.clear();
        xyrange.clear();
        fsrange.clear();
        const bool fp64 = prec == OPO_FP64;
        std::map<int, std::vector<int>> by_nt, by_nx;
        n_slices = 0;
        for (int i = 0; i < n; ++i) {
            by_nt[tclass(resident[i])].push_back(i);
            if (!resident[i]->p.no_diffraction) by_nx[resident[i]->g.nx].push_back(i);
            n_slices = std::max(n_slices, resident[i]->p.n_slices);
        }
        // within a class, solves with more slices first (stable)
        auto by_slices = [&](std::vector<int> &v) {
            std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
                return resident[a]->p.n_slices > resident[b]->p.n_slices;
            });
        };
        // per-slice prefix lengths of a class's work list
        auto prefixes = [&](Range &r, const std::vector<std::pair<int, int>> &items) {
            r.nk.assign(n_slices, 0);
            for (const auto &si : items)   // (slices, work items) per solve
                for (int k = 0; k < si.first; ++k) r.nk[k] += si.second;
        };
        // Each class's variant is read once here: the work list and the
        // launches captured below both follow it, whatever the process-wide
        // choice becomes meanwhile (a change makes the next step rebuild).
        for (auto &kv : by_nt) {
            by_slices(kv.second);
            Range r{kv.first, (int)work.size(), 0, class_time_variant(kv.first), {}};
            const int tx = kv.first < 0 ? 1 : tune_tile_px(fp64, kv.first, r.v);
            std::vector<std::pair<int, int>> items;
            for (int i : kv.second) {
                const opo_grid &g = resident[i]->g;
                if (tx == 0 || g.nx % tx != 0)
                    return fail(OPO_ERR_UNSUPPORTED, "tuned tile of %d pixels does not divide NX=%d (NT=%d)",
                                tx, g.nx, kv.first);
                const int tiles = (g.nx / tx) * g.ny;
                for (int t = 0; t < tiles; ++t) work.push_back(WorkItem{i, t});
                items.push_back({resident[i]->p.n_slices, tiles});
            }
            r.n = (int)work.size() - r.off;
            prefixes(r, items);
            trange.push_back(r);
        }
        for (auto &kv : by_nx) {
            by_slices(kv.second);
            // one item per group of consecutive planes of a field (xy_group)
            const int xv = tune_get(fp64, 1, kv.first);
            const int pg = xy_group(kv.first, xv);
            Range r{kv.first, (int)work.size(), 0, xv, {}};
            std::vector<std::pair<int, int>> items;
            for (int i : kv.second) {
                if (resident[i]->g.nt % pg != 0)
                    return fail(OPO_ERR_UNSUPPORTED, "NT=%d is not a multiple of the %d-plane xy group (NX=%d)",
                                resident[i]->g.nt, pg, kv.first);
                for (int it = 0; it < 3 * resident[i]->g.nt / pg; ++it) work.push_back(WorkItem{i, it});
                items.push_back({resident[i]->p.n_slices, 3 * resident[i]->g.nt / pg});
            }
            r.n = (int)work.size() - r.off;
            prefixes(r, items);
            xyrange.push_back(r);
            Range q{kv.first, (int)work.size(), 0, xv, {}};
            for (int i : kv.second)
                if (resident[i]->p.fs_on)   // in the NX class: not a plane-wave solve
                    for (int it = 0; it < 2 * resident[i]->g.nt / pg; ++it) work.push_back(WorkItem{i, it});
            q.n = (int)work.size() - q.off;
            if (q.n) fsrange.push_back(q);
        }
        prange = Range{0, (int)work.size(), 0, 0, {}};
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < chunks_for(resident[i]->g); ++c) work.push_back(WorkItem{i, c});
        prange.n = (int)work.size() - prange.off;
        // Grow the buffers where needed.  A buffer that moves changes the
        // captured launch arguments, which the update below carries over.
        int rc0;
        if ((rc0 = ensure_buf(d_desc, cap_desc, std::max<size_t>(hdesc.size(), 1))) ||
            (rc0 = ensure_buf(d_status, cap_status, sizeof(int) * std::max(n, 1))) ||
            (rc0 = ensure_buf(d_ts, cap_ts, sizeof(TripState) * std::max(n, 1))) ||
            (rc0 = ensure_buf(d_work, cap_work, sizeof(WorkItem) * std::max<size_t>(work.size(), 1))) ||
            (rc0 = ensure_buf(d_partial, cap_partial, sizeof(double) * 3 * (size_t)std::max(n, 1) * max_chunk)))
            return rc0;
        CK(hipMemcpyAsync(d_desc, hdesc.data(), hdesc.size(), hipMemcpyHostToDevice, stream));
        CK(hipMemcpyAsync(d_work, work.data(), sizeof(WorkItem) * work.size(), hipMemcpyHostToDevice, stream));
        CK(hipMemsetAsync(d_status, 0, sizeof(int) * std::max(n, 1), stream));
        h_ts.assign(n, TripState{0, 0});
        h_status.assign(n, 0);
        // capture one round trip of every resident solve
        ev_used = 0;
        ev_entry.clear();
        ev_alg.clear();
        sync_used = 0;
        hipGraph_t gr = nullptr;
        CK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        int rc = fp64 ? enqueue_trip<double>() : enqueue_trip<float>();
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return fail(OPO_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
        // Update the instantiated graph in place when the new trip has the
        // same shape (same launches in the same order: the size classes did
        // not change, only their work lists); instantiate it otherwise.
        if (graph) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            const hipError_t eu = hipGraphExecUpdate(graph, gr, &err_node, &res);
            if (eu == hipSuccess && res == hipGraphExecUpdateSuccess) {
                n_updated += 1;
            } else {
                (void)hipGetLastError();   // a refused update is not an error of the batch
                (void)hipGraphExecDestroy(graph);
                graph = nullptr;
            }
        }
        if (!graph) {
            e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail(OPO_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
            }
        }
        (void)hipGraphDestroy(gr);
        built_variants.clear();
        for (const Range &r : trange) built_variants.push_back(r.v);
        for (const Range &r : xyrange) built_variants.push_back(r.v);
        dirty = false;
        return OPO_OK;
    }

    // Launch filter of enqueue_trip: with only_stage >= 0, only the launches
    // of that (stage, size class) are enqueued (the stage timing graphs).
    int only_stage = -1, only_size = -1, n_enqueued = 0;
    bool wanted(int stage, int size) const {
        return only_stage < 0 || (stage == only_stage && size == only_size);
    }

    // Enqueue one round trip of every resident solve: per slice the fused
    // t pass of each NT class then the xy pass of each NX class; then the
    // trip-end pass, free space and the power sums.  Each launch goes through
    // launch(): the filter above, and the profiling events around it.
    template <typename T> int enqueue_trip() {
        const SolveDesc<T> *sd = static_cast<const SolveDesc<T> *>(d_desc);
        n_enqueued = 0;
        // one launch of a stage on stream st: the filter, and the profiling
        // events around it
        auto launch = [&](int stage, int size, double alg, hipStream_t st, auto &&fn) -> int {
            if (!wanted(stage, size)) return OPO_OK;
            int rc = mark(stage, size, alg, st);
            if (rc) return rc;
            CK(fn(st));
            n_enqueued += 1;
            return mark(stage, size, alg, st);
        };
        // the launches of one stage, class c on stream c (side streams for
        // c >= 1 when the classes run concurrently), joined at the end
        auto stage_of = [&](int nclass, auto &&one) -> int {
            const bool par = concurrent && only_stage < 0 && nclass > 1;
            const int nside = par ? std::min(nclass - 1, (int)kSide) : 0;
            hipEvent_t fork = nullptr;
            int rc;
            if (par) {
                if ((rc = next_sync_event(fork))) return rc;
                CK(hipEventRecord(fork, stream));
                for (int k = 0; k < nside; ++k) CK(hipStreamWaitEvent(side[k], fork, 0));
            }
            for (int c = 0; c < nclass; ++c) {
                const hipStream_t st = (par && c > 0) ? side[(c - 1) % kSide] : stream;
                if ((rc = one(c, st))) return rc;
            }
            for (int k = 0; k < nside; ++k) {
                hipEvent_t join = nullptr;
                if ((rc = next_sync_event(join))) return rc;
                CK(hipEventRecord(join, side[k]));
                CK(hipStreamWaitEvent(stream, join, 0));
            }
            return OPO_OK;
        };
        std::vector<double> tb, fb;
        for (const Range &r : trange) tb.push_back(class_bytes(3, 2, [&](const BatchEntry *e) { return tclass(e) == r.cls; }));
        // bytes of slice k's launch: the class's solves with more than k slices
        auto slice_bytes = [&](bool t_pass, int cls, int k) {
            return class_bytes(3, 2, [&](const BatchEntry *e) {
                return (t_pass ? tclass(e) == cls : in_xy(e, cls)) && e->p.n_slices > k;
            });
        };
        for (const Range &r : fsrange)
            fb.push_back(class_bytes(2, 2, [&](const BatchEntry *e) { return in_xy(e, r.cls) && e->p.fs_on; }));
        const double pb = class_bytes(3, 1, [](const BatchEntry *) { return true; });
        int rc;
        for (int k = 0; k < n_slices; ++k) {
            const int mode = k == 0 ? BATCH_FIRST : BATCH_MID;
            if ((rc = stage_of((int)trange.size(), [&](int c, hipStream_t st) {
                     const Range &r = trange[c];
                     const int nw = r.nk[k];
                     if (nw == 0) return (int)OPO_OK;   // every solve of the class has fewer slices
                     return launch(OPO_STAGE_TIME_OP, r.cls, slice_bytes(true, r.cls, k), st, [&](hipStream_t s2) {
                         return launch_batch_time<T>(s2, std::abs(r.cls), nw, sd, d_status, d_work + r.off, mode, r.v);
                     });
                 })))
                return rc;
            if ((rc = stage_of((int)xyrange.size(), [&](int c, hipStream_t st) {
                     const Range &r = xyrange[c];
                     const int nw = r.nk[k];
                     if (nw == 0) return (int)OPO_OK;
                     return launch(OPO_STAGE_XY_OP, r.cls, slice_bytes(false, r.cls, k), st, [&](hipStream_t s2) {
                         return launch_batch_xy<T>(s2, r.cls, nw, sd, d_ts, d_status, d_work + r.off, 0, r.v);
                     });
                 })))
                return rc;
        }
        if ((rc = stage_of((int)trange.size(), [&](int c, hipStream_t st) {
                 const Range &r = trange[c];
                 return launch(OPO_STAGE_GDD, r.cls, tb[c], st, [&](hipStream_t s2) {
                     return launch_batch_time<T>(s2, std::abs(r.cls), r.n, sd, d_status, d_work + r.off, BATCH_LAST,
                                                 r.v);
                 });
             })))
            return rc;
        if ((rc = stage_of((int)fsrange.size(), [&](int c, hipStream_t st) {
                 const Range &r = fsrange[c];
                 return launch(OPO_STAGE_FREE_SPACE, r.cls, fb[c], st, [&](hipStream_t s2) {
                     return launch_batch_xy<T>(s2, r.cls, r.n, sd, d_ts, d_status, d_work + r.off, 1, r.v);
                 });
             })))
            return rc;
        return launch(OPO_STAGE_POWER, 0, pb, stream, [&](hipStream_t s2) {
            return launch_batch_power<T>(s2, prange.n, (int)resident.size(), sd, d_ts, d_status,
                                         d_work + prange.off, d_partial, max_chunk);
        });
    }

    // Time each (stage, size class) of the trip by itself: a graph holding
    // that class's launches of one round trip, in trip order and with the
    // trip's arguments, replayed reps times between two HIP events on the
    // batch stream.  The solves' fields go through these launches too.
    // One (stage, size class) of the trip by itself: its launches of one
    // round trip captured as a graph and replayed reps times (after one
    // untimed replay) between two events on the batch stream.
    int time_key(int stage, int size, int reps, float &ms, int &launches) {
        ms = 0.f;
        launches = 0;
        const bool keep_prof = profiling;
        profiling = false;
        only_stage = stage;
        only_size = size;
        hipGraph_t gr = nullptr;
        hipGraphExec_t ge = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        int rc = OPO_OK;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = fail(OPO_ERR_HIP, "stage timing: capture failed");
        } else {
            rc = prec == OPO_FP64 ? enqueue_trip<double>() : enqueue_trip<float>();
            launches = n_enqueued;
            const hipError_t e = hipStreamEndCapture(stream, &gr);
            if (!rc && e != hipSuccess) rc = fail(OPO_ERR_HIP, "stage timing capture: %s", hipGetErrorString(e));
        }
        only_stage = only_size = -1;
        profiling = keep_prof;
        if (!rc && hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) != hipSuccess)
            rc = fail(OPO_ERR_HIP, "stage timing: instantiate failed");
        if (gr) (void)hipGraphDestroy(gr);
        if (!rc) {
            hipError_t e2 = hipEventCreate(&ev[0]);
            if (e2 == hipSuccess) e2 = hipEventCreate(&ev[1]);
            if (e2 == hipSuccess) e2 = hipGraphLaunch(ge, stream);   // warm-up
            if (e2 == hipSuccess) e2 = hipEventRecord(ev[0], stream);
            for (int r = 0; r < reps && e2 == hipSuccess; ++r) e2 = hipGraphLaunch(ge, stream);
            if (e2 == hipSuccess) e2 = hipEventRecord(ev[1], stream);
            if (e2 == hipSuccess) e2 = hipEventSynchronize(ev[1]);
            if (e2 == hipSuccess) e2 = hipEventElapsedTime(&ms, ev[0], ev[1]);
            if (e2 != hipSuccess) rc = fail(OPO_ERR_HIP, "stage timing: %s", hipGetErrorString(e2));
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        return rc;
    }

    // Time each (stage, size class) of the trip by itself (time_key).
    // The solves' fields go through these launches too.
    int time_stages(int reps, std::vector<opo_stage_time> &out) {
        out.clear();
        // the (stage, size) entries of the trip, in trip order
        std::vector<std::pair<int, int>> keys;
        for (const Range &r : trange) keys.push_back({OPO_STAGE_TIME_OP, r.cls});
        for (const Range &r : xyrange) keys.push_back({OPO_STAGE_XY_OP, r.cls});
        for (const Range &r : trange) keys.push_back({OPO_STAGE_GDD, r.cls});
        for (const Range &r : fsrange) keys.push_back({OPO_STAGE_FREE_SPACE, r.cls});
        keys.push_back({OPO_STAGE_POWER, 0});
        for (const auto &key : keys) {
            float ms = 0.f;
            int launches = 0;
            const int rc = time_key(key.first, key.second, reps, ms, launches);
            if (rc) return rc;
            // per-slice stages: the mean over the trip's launches (solves
            // with fewer slices drop out of the later launches)
            auto per_slice = [&](bool t_pass) {
                double sum = 0.0;
                int nl = 0;
                for (int k = 0; k < n_slices; ++k) {
                    const double b = class_bytes(3, 2, [&](const BatchEntry *e) {
                        return (t_pass ? tclass(e) == key.second : in_xy(e, key.second)) && e->p.n_slices > k;
                    });
                    if (b > 0) sum += b, nl += 1;
                }
                return nl ? sum / nl : 0.0;
            };
            const double alg = key.first == OPO_STAGE_POWER
                                   ? class_bytes(3, 1, [](const BatchEntry *) { return true; })
                               : key.first == OPO_STAGE_XY_OP ? per_slice(false)
                               : key.first == OPO_STAGE_FREE_SPACE
                                   ? class_bytes(2, 2, [&](const BatchEntry *e) { return in_xy(e, key.second) && e->p.fs_on; })
                               : key.first == OPO_STAGE_TIME_OP ? per_slice(true)
                                   : class_bytes(3, 2, [&](const BatchEntry *e) { return tclass(e) == key.second; });
            out.push_back(opo_stage_time{key.first, key.second, launches * reps, (double)ms, alg});
        }
        return OPO_OK;
    }

    // Tune the launch variants in this batch's own trip: for each NT class
    // the t-pass variants whose tile divides the NX of every resident solve
    // of the class, for each NX class the xy-pass variants, each timed as
    // that class's launches of one trip over all its solves (time_key), the
    // fastest kept in the process-wide tables.  The resident fields are
    // saved first and put back after, so the solves continue from where
    // they were; this needs free device memory of the size of those fields.
    int autotune(int reps, std::string &report) {
        const bool fp64 = prec == OPO_FP64;
        int rc;
        if (dirty && (rc = rebuild())) return rc;
        size_t total = 0;
        for (const BatchEntry *e : resident) total += 3 * e->s->field_bytes();
        unsigned char *save = nullptr;
        if (hipMalloc(&save, total) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "batch autotune: no room to save the resident fields (%zu bytes)", total);
        size_t off = 0;
        for (const BatchEntry *e : resident)
            for (int f = 0; f < 3; ++f) {
                const size_t fb = e->s->field_bytes();
                if (hipMemcpyAsync(save + off, e->s->dev_ptr(f), fb, hipMemcpyDeviceToDevice, stream) != hipSuccess) {
                    (void)hipFree(save);
                    return fail(OPO_ERR_HIP, "batch autotune: saving the fields failed");
                }
                off += fb;
            }
        std::vector<int> nts, nxs;
        for (const Range &r : trange)
            if (r.cls > 0) nts.push_back(r.cls);   // the plane-wave classes have one launch shape
        for (const Range &r : xyrange) nxs.push_back(r.cls);
        rc = OPO_OK;
        char buf[160];
        for (int axis = 0; axis < 2 && !rc; ++axis) {
            for (int n : axis == 0 ? nts : nxs) {
                const int stage = axis == 0 ? OPO_STAGE_TIME_OP : OPO_STAGE_XY_OP;
                int best = tune_get(fp64, axis, n);
                double best_us = 1e30;
                report += axis == 0 ? "batch t pass NT=" : "batch xy pass NX=";
                report += std::to_string(n) + " (us per launch):";
                for (int v = 0; v < tune_variant_count(axis) && !rc; ++v) {
                    if (!tune_variant_valid(fp64, axis, n, v)) continue;
                    bool fits = true;
                    if (axis == 0)
                        for (const BatchEntry *e : resident)
                            if (tclass(e) == n && e->g.nx % tune_tile_px(fp64, n, v) != 0) fits = false;
                    if (!fits) continue;
                    tune_set(fp64, axis, n, v);
                    if ((rc = rebuild())) break;
                    float ms = 0.f;
                    int launches = 0;
                    if ((rc = time_key(stage, n, reps, ms, launches))) break;
                    const double us = launches ? 1e3 * ms / ((double)launches * reps) : 1e30;
                    std::snprintf(buf, sizeof(buf), " v%d %.2f;", v, us);
                    report += buf;
                    if (us < best_us) {
                        best_us = us;
                        best = v;
                    }
                }
                tune_set(fp64, axis, n, best);
                report += " -> v" + std::to_string(best) + "\n";
                if (rc) break;
            }
        }
        off = 0;
        for (const BatchEntry *e : resident)
            for (int f = 0; f < 3; ++f) {
                const size_t fb = e->s->field_bytes();
                if (hipMemcpyAsync(e->s->dev_ptr(f), save + off, fb, hipMemcpyDeviceToDevice, stream) != hipSuccess &&
                    !rc)
                    rc = fail(OPO_ERR_HIP, "batch autotune: restoring the fields failed");
                off += fb;
            }
        const hipError_t es = hipStreamSynchronize(stream);
        (void)hipFree(save);
        if (!rc && es != hipSuccess) rc = fail(OPO_ERR_HIP, "batch autotune: %s", hipGetErrorString(es));
        dirty = true;
        return rc;
    }

    // Admit queued solves in arrival order while the head of the queue fits
    // the budget.
    void admit() {
        while (!queue.empty() && used + queue.front()->footprint <= budget) {
            BatchEntry *e = queue.front();
            std::unique_ptr<SolveImpl> s;
            int rc = make_solve(&e->g, &e->p, device, s);
            if (rc == OPO_ERR_NOMEM && !resident.empty()) break;  // wait for memory to free up
            queue.pop_front();
            if (!rc) rc = s->set_fields(e->api.data(), e->as.data(), e->ai.data());
            if (!rc && hipMalloc(&e->d_series, sizeof(double) * 3 * (size_t)e->n_trips) != hipSuccess)
                rc = fail(OPO_ERR_NOMEM, "series allocation failed");
            if (rc) {
                e->finished = true;
                e->error = rc;
                e->error_msg = last_error_text();
                continue;
            }
            e->api = {};
            e->as = {};
            e->ai = {};
            e->s = std::move(s);
            used += e->footprint;
            resident.push_back(e);
            dirty = true;
        }
    }
};

extern "C" {

// ------------------------------------------------------------- batching
size_t opo_batch_footprint(const opo_grid *g, int precision, int n_trips) {
    if (!g || n_trips < 0) return 0;
    const size_t n = (size_t)g->nx * g->ny * g->nt;
    const size_t cb = precision == OPO_FP64 ? 16 : 8;
    // Every device allocation of a resident solve, each rounded up to the
    // allocation granularity assumed here: Api, Ap, As, Ai; the table block
    // (twiddles, D, GDD, Q, free-space and modulator tables as hi/lo pairs);
    // the power partials (one per 4096 cells, 1 to 1024, three fields) and
    // counters; the per-trip power series of the batch.
    const size_t kGran = 64 * 1024;
    auto up = [&](size_t b) { return (b + kGran - 1) / kGran * kGran; };
    const size_t tabs = 2 * (size_t)(6 * g->nx + 5 * g->ny + 6 * g->nt) * cb;
    size_t nblk = (n + 4095) / 4096;
    nblk = nblk < 1 ? 1 : (nblk > 1024 ? 1024 : nblk);
    const size_t dbuf = sizeof(double) * (3 * nblk + 4);
    const size_t series = sizeof(double) * 3 * (size_t)(n_trips > 0 ? n_trips : 1);
    return 4 * up(n * cb) + up(tabs) + up(dbuf) + up(4 * sizeof(int)) + up(series);
}

int opo_batch_create(int device, size_t budget, int precision, opo_batch **out) {
    if (!out) return fail(OPO_ERR_ARG, "null output");
    *out = nullptr;
    if (precision != OPO_FP32 && precision != OPO_FP64) return fail(OPO_ERR_ARG, "bad precision");
    int ndev = opo_device_count();
    if (ndev < 1) return fail(OPO_ERR_NOGPU, "no HIP device");
    if (device < 0) device = 0;
    if (device >= ndev) return fail(OPO_ERR_ARG, "device out of range");
    std::unique_ptr<opo_batch> b(new opo_batch);
    b->device = device;
    b->budget = budget;
    b->prec = precision;
    CK(hipSetDevice(device));
    CK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    *out = b.release();
    return OPO_OK;
}

void opo_batch_destroy(opo_batch *b) { delete b; }

int opo_batch_submit(opo_batch *b, const opo_grid *g, const opo_params *p, const void *api,
                     const void *as, const void *ai, int n_trips, uint64_t *ticket) {
    if (!b || !g || !p || !api || !as || !ai || n_trips < 1)
        return fail(OPO_ERR_ARG, "opo_batch_submit: bad argument");
    int rc = validate(g, p);
    if (rc) return rc;
    if (p->n_slices < 1) return fail(OPO_ERR_ARG, "round trips need n_slices >= 1");
    if (p->precision != b->prec) return fail(OPO_ERR_ARG, "precision differs from the batch's");
    if (!batch_supported(g->nx, g->ny, g->nt, p->precision == OPO_FP64))
        return fail(OPO_ERR_UNSUPPORTED, "grid %dx%dx%d is not supported in a batch", g->nx, g->ny, g->nt);
    if (p->dk_phase_scale != 0.0 && p->dk != 0.0)
        return fail(OPO_ERR_UNSUPPORTED, "a batch runs solves without the dk phase term");
    if (g->nx == 1 && !p->no_diffraction)
        return fail(OPO_ERR_UNSUPPORTED, "a 1x1 grid runs in a batch as the plane-wave model (no_diffraction)");
    const size_t fp = opo_batch_footprint(g, p->precision, n_trips);
    if (fp > b->budget) return fail(OPO_ERR_BUDGET, "solve needs %zu bytes, budget is %zu", fp, b->budget);
    std::lock_guard<std::mutex> lk(b->mu);
    auto e = std::make_unique<BatchEntry>();
    e->ticket = b->next_ticket++;
    e->g = *g;
    e->p = *p;
    e->p.omega = nullptr;
    e->n_trips = n_trips;
    e->footprint = fp;
    const size_t fb = (size_t)g->nx * g->ny * g->nt * (p->precision == OPO_FP64 ? 16 : 8);
    e->api.assign((const unsigned char *)api, (const unsigned char *)api + fb);
    e->as.assign((const unsigned char *)as, (const unsigned char *)as + fb);
    e->ai.assign((const unsigned char *)ai, (const unsigned char *)ai + fb);
    if (ticket) *ticket = e->ticket;
    b->queue.push_back(e.get());
    b->all[e->ticket] = std::move(e);
    return OPO_OK;
}

int opo_batch_step(opo_batch *b, int *n_active) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    if (n_active) *n_active = 0;
    b->admit();
    const int n = (int)b->resident.size();
    if (n == 0) return OPO_OK;
    int rc;
    if (!b->dirty && b->current_variants() != b->built_variants) b->dirty = true;  // re-tuned
    if (b->dirty && (rc = b->rebuild())) return rc;
    for (int i = 0; i < n; ++i) {
        const BatchEntry *e = b->resident[i];
        const int rt = e->done_trips;
        b->h_ts[i].trip = rt;
        b->h_ts[i].fs = e->p.fs_on && !(e->p.skip_fs_on_last_trip && rt == e->p.n_round_trips - 1);
    }
    CK(hipMemcpyAsync(b->d_ts, b->h_ts.data(), sizeof(TripState) * n, hipMemcpyHostToDevice, b->stream));
    CK(hipGraphLaunch(b->graph, b->stream));
    CK(hipMemcpyAsync(b->h_status.data(), b->d_status, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    CK(hipStreamSynchronize(b->stream));
    if ((rc = b->collect())) return rc;
    std::vector<BatchEntry *> keep;
    for (int i = 0; i < n; ++i) {
        BatchEntry *e = b->resident[i];
        if (b->h_status[i]) {
            e->done_trips += 1;
            e->error = OPO_ERR_NAN;
            e->error_msg = "non-finite signal power at round trip " + std::to_string(e->done_trips - 1);
        } else {
            e->done_trips += 1;
        }
        if (e->error || e->done_trips >= e->n_trips) e->finished = true;
        else keep.push_back(e);
    }
    if (keep.size() != b->resident.size()) b->dirty = true;
    b->resident.swap(keep);
    if (n_active) *n_active = n;
    return OPO_OK;
}

int opo_batch_poll(opo_batch *b, uint64_t ticket, int *done, double *power, void *ap, void *as,
                   void *ai) {
    if (!b || !done) return fail(OPO_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->all.find(ticket);
    if (it == b->all.end()) return fail(OPO_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    BatchEntry *e = it->second.get();
    *done = e->finished ? 1 : 0;
    if (!e->finished) return OPO_OK;
    int rc = e->error;
    std::string msg = e->error_msg;
    if (e->s) {
        CK(hipSetDevice(b->device));
        if (power && e->done_trips > 0)
            CK(hipMemcpy(power, e->d_series, sizeof(double) * 3 * (size_t)e->done_trips,
                         hipMemcpyDeviceToHost));
        int r2 = e->s->get_fields(ap, as, ai);
        if (!rc) rc = r2;
        b->used -= e->footprint;
        e->s.reset();
        (void)hipFree(e->d_series);
        e->d_series = nullptr;
    }
    b->all.erase(it);
    if (rc) return fail(rc, "%s", msg.c_str());
    return OPO_OK;
}

int opo_batch_rebuilds(opo_batch *b, int *rebuilds, int *updated, double *seconds) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    if (rebuilds) *rebuilds = b->n_rebuilds;
    if (updated) *updated = b->n_updated;
    if (seconds) *seconds = b->rebuild_s;
    return OPO_OK;
}

int opo_batch_concurrent_classes(opo_batch *b, int on) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    const bool want = on != 0;
    if (want && !b->side[0]) {
        CK(hipSetDevice(b->device));
        for (auto &x : b->side) CK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    }
    if (want != b->concurrent) {
        b->concurrent = want;
        b->dirty = true;   // re-capture with or without the branches
    }
    return OPO_OK;
}

int opo_batch_autotune(opo_batch *b, int reps, char *report, size_t report_len) {
    if (!b || reps < 1) return fail(OPO_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    b->admit();
    if (b->resident.empty()) return fail(OPO_ERR_STATE, "no resident solves to tune on");
    std::string rep;
    const int rc = b->autotune(reps, rep);
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return rc;
}

int opo_batch_time_stages(opo_batch *b, int reps, opo_stage_time *out, int cap, int *n) {
    if (!b || !n || reps < 1 || (cap > 0 && !out)) return fail(OPO_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    if (b->resident.empty()) return fail(OPO_ERR_STATE, "no resident solves to time");
    int rc;
    if (b->dirty && (rc = b->rebuild())) return rc;
    std::vector<opo_stage_time> v;
    if ((rc = b->time_stages(reps, v))) return rc;
    *n = (int)v.size();
    for (int k = 0; k < cap && k < *n; ++k) out[k] = v[k];
    return OPO_OK;
}

int opo_batch_profile(opo_batch *b, int on) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    const bool want = on != 0;
    if (want != b->profiling) {   // re-capture with or without the events
        b->profiling = want;
        b->dirty = true;
        b->prof.clear();
        b->prof_alg.clear();
        b->prof_error.clear();
    }
    b->prof_reset();   // start the sums again
    return OPO_OK;
}

int opo_batch_profile_read(opo_batch *b, opo_stage_time *out, int cap, int *n) {
    if (!b || !n || (cap > 0 && !out)) return fail(OPO_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    if (!b->prof_error.empty()) return fail(OPO_ERR_HIP, "%s", b->prof_error.c_str());
    *n = (int)b->prof.size();
    for (int k = 0; k < cap && k < *n; ++k) {
        out[k] = b->prof[k];
        if (out[k].launches) out[k].alg_bytes = b->prof_alg[k] / out[k].launches;
    }
    return OPO_OK;
}

int opo_batch_stats(opo_batch *b, int *resident, int *queued, size_t *bytes_used) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    if (resident) *resident = (int)b->resident.size();
    if (queued) *queued = (int)b->queue.size();
    if (bytes_used) *bytes_used = b->used;
    return OPO_OK;
}

}  // extern "C"

