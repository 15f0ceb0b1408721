// run_writer.cpp — run output in the reference's formats
// (cpp/io/snapshot_writer.cpp): metadata.json, diagnostics.csv and one
// snapshot_NNNNNN.csv per snapshot, byte for byte as the reference writes
// them, so its Python tools (python/utils/snapshots.py) read them unchanged.
//
// Snapshots and diagnostics rows of a context's resident state are staged
// on the context's stream: one device-to-host copy of the state block (or
// the diagnostics kernels and a copy of their 13 results) into pinned
// slots, and an event.  The writer's own thread waits for each event in
// call order, formats and writes, while the device steps on.  Text follows
// iostream's default float format at the reference's precisions: %.17g for
// snapshot and diagnostics values (they set std::setprecision(17)), %g for
// metadata.json (default precision 6).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <exception>
#include <filesystem>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "abi_guard.hpp"
#include "gcsim.h"

namespace gc {
// capi.hip: copies x y z vx vy vz ax ay az m (stride *cap) into dst on the
// context's stream and records ev after the copy; dst == nullptr only
// reports n and cap.
gc_status stage_state(gc_ctx* c, double* dst, size_t dst_doubles, size_t* n, size_t* cap,
                      hipEvent_t ev);
// capi.hip: launches the diagnostics kernels and copies their 13 results,
// in gc_diagnostics field order, into dst on the context's stream; records
// ev after the copy.
gc_status stage_diagnostics(gc_ctx* c, double* dst, hipEvent_t ev, size_t* n);
}  // namespace gc

namespace {

constexpr int kSnapSlots = 2;   // staged snapshots in flight
constexpr int kDiagSlots = 16;  // staged diagnostics rows in flight
constexpr size_t kChunkRows = 8192;   // rows one formatting task writes
constexpr size_t kMaxRowChars = 320;  // id, group and 10 values of <= 24 chars

static_assert(sizeof(gc_diagnostics) == 13 * sizeof(double), "gc_diagnostics is 13 doubles");

int default_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(8u, hw)));
}

// %.17g text of v: std::to_chars in general form at precision 17 is
// specified as printf's %.17g (and runs about 3x faster); at most 24 chars.
char* put17(char* p, double v) {
    return std::to_chars(p, p + 32, v, std::chars_format::general, 17).ptr;
}

char* comma(char* p) {
    *p = ',';
    return p + 1;
}

template <class T>
char* put_int(char* p, T v) {
    return std::to_chars(p, p + 24, v).ptr;
}

std::string json_escaped(const char* s) {
    std::string out;
    for (const char* p = s ? s : ""; *p; ++p) {
        if (*p == '"' || *p == '\\') out.push_back('\\');
        out.push_back(*p);
    }
    return out;
}

// Pinned host slots with one event each; a job holds its slot until the
// writer thread has consumed it.
struct SlotPool {
    double* mem = nullptr;  // count * stride doubles
    size_t stride = 0;
    std::vector<hipEvent_t> ev;
    std::vector<char> busy;

    explicit SlotPool(int count) : ev(count, nullptr), busy(count, 0) {}
    double* slot(int k) const { return mem + k * stride; }
    bool any_free() const { return std::find(busy.begin(), busy.end(), 0) != busy.end(); }
    bool all_free() const { return std::find(busy.begin(), busy.end(), 1) == busy.end(); }
    void release() {
        if (mem) (void)hipHostFree(mem);
        mem = nullptr;
        stride = 0;
        for (auto& e : ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
};

struct Job {
    enum Kind { kSnapshot, kDiagnostics } kind = kSnapshot;
    int step = 0;
    double time = 0.0;
    size_t n = 0, stride = 0;
    const double* staged = nullptr;  // pinned slot filled by the device, or
    std::vector<double> host;        // snapshot arrays copied in (stride n)
    std::vector<std::uint32_t> groups;
    gc_diagnostics diag{};           // row values given by the caller
    SlotPool* pool = nullptr;        // slot to release when done
    int slot = -1;
};

// Rows [r0, r1) of a snapshot: id,group_id,mass,x,y,z,vx,vy,vz,ax,ay,az.
void format_rows(const Job& j, size_t r0, size_t r1, std::vector<char>& out) {
    out.resize((r1 - r0) * kMaxRowChars);
    const double* b = j.staged ? j.staged : j.host.data();
    const size_t s = j.stride;
    char* p = out.data();
    for (size_t i = r0; i < r1; ++i) {
        p = put_int(p, i);
        p = put_int(comma(p), j.groups.empty() ? 0u : j.groups[i]);
        p = put17(comma(p), b[9 * s + i]);  // mass
        for (int c = 0; c < 9; ++c) p = put17(comma(p), b[c * s + i]);
        *p++ = '\n';
    }
    out.resize(p - out.data());
}

// snapshot_writer.cpp:72-96.  Rows are formatted by up to `threads` tasks
// at a time, kChunkRows each, and written in row order.
bool write_snapshot_file(const std::filesystem::path& dir, const Job& j, int threads,
                         std::string* err) {
    char name[40];
    std::snprintf(name, sizeof(name), "snapshot_%06d.csv", j.step);
    FILE* f = std::fopen((dir / name).string().c_str(), "wb");
    if (!f) {
        *err = "Could not write snapshot output in " + dir.string();
        return false;
    }
    char head[96] = "# time=";
    char* p = put17(head + 7, j.time);
    const char cols[] = "\nid,group_id,mass,x,y,z,vx,vy,vz,ax,ay,az\n";
    std::memcpy(p, cols, sizeof(cols) - 1);
    p += sizeof(cols) - 1;
    bool ok = std::fwrite(head, 1, p - head, f) == static_cast<size_t>(p - head);
    const size_t nchunks = (j.n + kChunkRows - 1) / kChunkRows;
    const size_t width = std::max<size_t>(1, std::min<size_t>(threads, nchunks));
    std::vector<std::vector<char>> bufs(width);
    std::vector<char> formatted(width);
    for (size_t c0 = 0; ok && c0 < nchunks; c0 += width) {
        const size_t cw = std::min(width, nchunks - c0);
        auto task = [&](size_t t) {
            try {
                const size_t r0 = (c0 + t) * kChunkRows;
                format_rows(j, r0, std::min(j.n, r0 + kChunkRows), bufs[t]);
                formatted[t] = 1;
            } catch (...) {
                formatted[t] = 0;  // out of host memory: reported below
            }
        };
        std::vector<std::thread> pool;
        size_t started = 1;
        try {
            for (; started < cw; ++started) pool.emplace_back(task, started);
        } catch (...) {  // fewer threads: this thread formats the rest
        }
        for (size_t t = started; t < cw; ++t) task(t);
        task(0);
        for (auto& th : pool) th.join();
        for (size_t t = 0; ok && t < cw; ++t)
            ok = formatted[t] && std::fwrite(bufs[t].data(), 1, bufs[t].size(), f) == bufs[t].size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) *err = "Could not write snapshot output in " + dir.string();
    return ok;
}

// snapshot_writer.cpp:98-121
bool write_diagnostics_row(FILE* f, const Job& j) {
    const gc_diagnostics& d = j.diag;
    const double v[13] = {d.total_mass,        d.kinetic_energy,     d.potential_energy,
                          d.total_energy,      d.momentum[0],        d.momentum[1],
                          d.momentum[2],       d.center_of_mass[0],  d.center_of_mass[1],
                          d.center_of_mass[2], d.angular_momentum[0], d.angular_momentum[1],
                          d.angular_momentum[2]};
    char line[512];
    char* p = put_int(line, j.step);
    p = put17(comma(p), j.time);
    p = put_int(comma(p), j.n);
    for (double x : v) p = put17(comma(p), x);
    *p++ = '\n';
    const size_t len = static_cast<size_t>(p - line);
    return std::fwrite(line, 1, len, f) == len;
}

}  // namespace

struct gc_writer {
    std::filesystem::path dir;
    FILE* diag = nullptr;  // written by the worker only, once it runs
    SlotPool snaps{kSnapSlots};
    SlotPool diags{kDiagSlots};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;
    bool stop = false;
    bool working = false;  // the worker holds a job
    std::string err;       // first write error
    int threads = default_threads();

    void run() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv.wait(lock, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                j = std::move(jobs.front());
                jobs.pop_front();
                working = true;
            }
            std::string e;
            if (j.pool && hipEventSynchronize(j.pool->ev[j.slot]) != hipSuccess)
                e = "device staging for step " + std::to_string(j.step) + " failed";
            try {
                if (e.empty() && j.kind == Job::kSnapshot) {
                    write_snapshot_file(dir, j, threads, &e);
                } else if (e.empty()) {
                    if (j.staged) std::memcpy(&j.diag, j.staged, sizeof(gc_diagnostics));
                    if (!write_diagnostics_row(diag, j))
                        e = "Could not write diagnostics output in " + dir.string();
                }
            } catch (const std::exception& x) {  // host memory or threads ran out
                e = std::string("snapshot for step ") + std::to_string(j.step) + ": " + x.what();
            }
            {
                std::lock_guard<std::mutex> lock(mu);
                if (!e.empty() && err.empty()) err = e;
                if (j.pool) j.pool->busy[j.slot] = 0;
                working = false;
            }
            cv.notify_all();
        }
    }

    // queues j; if queueing throws, j's slot is freed (and waiters woken)
    // before the exception leaves.  A copy may still be in flight into that
    // slot: take() waits on the slot's event before handing it out again.
    void push(Job&& j) {
        std::exception_ptr failed;
        {
            std::lock_guard<std::mutex> lock(mu);
            try {
                jobs.push_back(std::move(j));
            } catch (...) {
                if (j.pool) j.pool->busy[j.slot] = 0;
                failed = std::current_exception();
            }
        }
        cv.notify_all();
        if (failed) std::rethrow_exception(failed);
    }

    // waits for a free slot of pool (the worker frees one per job), then
    // for the last copy recorded into it
    int take(SlotPool& pool) {
        int k = 0;
        {
            std::unique_lock<std::mutex> lock(mu);
            cv.wait(lock, [&] { return pool.any_free(); });
            k = static_cast<int>(std::find(pool.busy.begin(), pool.busy.end(), 0) -
                                 pool.busy.begin());
            pool.busy[k] = 1;
        }
        if (pool.ev[k]) (void)hipEventSynchronize(pool.ev[k]);
        return k;
    }

    void give_back(SlotPool& pool, int k) {
        {
            std::lock_guard<std::mutex> lock(mu);
            pool.busy[k] = 0;
        }
        cv.notify_all();
    }

    // every queued job done and every slot free
    void drain() {
        std::unique_lock<std::mutex> lock(mu);
        cv.wait(lock, [&] { return jobs.empty() && !working && snaps.all_free() && diags.all_free(); });
    }

    // pinned slots of `doubles` each (and their events) for pool
    gc_status ensure(SlotPool& pool, size_t doubles, const char* what) {
        if (doubles > pool.stride) {
            drain();  // no copy or formatting reads the old slots
            if (pool.mem) (void)hipHostFree(pool.mem);
            pool.mem = nullptr;
            pool.stride = 0;
            if (hipHostMalloc(&pool.mem, sizeof(double) * doubles * pool.busy.size()) != hipSuccess) {
                pool.mem = nullptr;
                return gc::set_error(std::string(what) + ": pinned staging memory", GC_ERR_OUT_OF_MEMORY);
            }
            pool.stride = doubles;
        }
        for (auto& e : pool.ev)
            if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                e = nullptr;
                return gc::set_error(std::string(what) + ": event", GC_ERR_HIP);
            }
        return GC_OK;
    }
};

extern "C" {

gc_status gc_writer_open(gc_writer** out, const char* directory) {
    GC_ABI_BEGIN
    if (!out || !directory) return gc::set_error("writer_open: null argument", GC_ERR_INVALID_ARGUMENT);
    *out = nullptr;
    gc_writer* w = new (std::nothrow) gc_writer();
    if (!w) return gc::set_error("writer_open: out of memory", GC_ERR_OUT_OF_MEMORY);
    w->dir = directory;
    // snapshot_writer.cpp:31-48: the directory, diagnostics.csv and its header
    std::error_code ec;
    std::filesystem::create_directories(w->dir, ec);
    w->diag = std::fopen((w->dir / "diagnostics.csv").string().c_str(), "wb");
    if (!w->diag) {
        delete w;
        return gc::set_error(std::string("Could not open diagnostics output in ") + directory,
                             GC_ERR_IO);
    }
    std::fputs("step,time,n,total_mass,kinetic_energy,potential_energy,total_energy,"
               "momentum_x,momentum_y,momentum_z,"
               "center_of_mass_x,center_of_mass_y,center_of_mass_z,"
               "angular_momentum_x,angular_momentum_y,angular_momentum_z\n",
               w->diag);
    try {
        w->worker = std::thread([w] { w->run(); });
    } catch (...) {
        std::fclose(w->diag);
        delete w;
        throw;
    }
    *out = w;
    return GC_OK;
    GC_ABI_END("gc_writer_open")
}

gc_status gc_writer_metadata(gc_writer* w, const gc_run_info* info) {
    GC_ABI_BEGIN
    if (!w || !info) return gc::set_error("writer_metadata: null argument", GC_ERR_INVALID_ARGUMENT);
    // snapshot_writer.cpp:50-70, at iostream's default precision (%g)
    FILE* m = std::fopen((w->dir / "metadata.json").string().c_str(), "wb");
    if (!m) return gc::set_error("Could not write metadata output in " + w->dir.string(), GC_ERR_IO);
    std::fprintf(m,
                 "{\n  \"name\": \"%s\",\n  \"solver\": \"%s\",\n  \"particle_count\": %zu,\n"
                 "  \"steps\": %d,\n  \"dt\": %g,\n  \"snapshot_every\": %d,\n  \"dim\": %d,\n"
                 "  \"gravitational_constant\": %g,\n  \"softening\": %g,\n  \"tree_theta\": %g,\n"
                 "  \"tree_leaf_capacity\": %zu,\n  \"fmm_expansion_order\": %d\n}\n",
                 json_escaped(info->name).c_str(), json_escaped(info->solver).c_str(),
                 info->particle_count, info->steps, info->dt, info->snapshot_every, info->dim,
                 info->gravitational_constant, info->softening, info->tree_theta,
                 info->tree_leaf_capacity, info->fmm_expansion_order);
    if (std::fclose(m) != 0)
        return gc::set_error("Could not write metadata output in " + w->dir.string(), GC_ERR_IO);
    return GC_OK;
    GC_ABI_END("gc_writer_metadata")
}

gc_status gc_writer_snapshot_host(gc_writer* w, int step, double time, size_t n, const double* pos,
                                  const double* vel, const double* acc, const double* mass,
                                  const std::uint32_t* group_ids) {
    GC_ABI_BEGIN
    if (!w || (n > 0 && (!pos || !vel || !acc || !mass)))
        return gc::set_error("writer_snapshot_host: null argument", GC_ERR_INVALID_ARGUMENT);
    Job j;
    j.step = step;
    j.time = time;
    j.n = n;
    j.stride = n;
    j.host.resize(10 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            j.host[d * n + i] = pos[3 * i + d];
            j.host[(3 + d) * n + i] = vel[3 * i + d];
            j.host[(6 + d) * n + i] = acc[3 * i + d];
        }
        j.host[9 * n + i] = mass[i];
    }
    if (group_ids) j.groups.assign(group_ids, group_ids + n);
    w->push(std::move(j));
    return GC_OK;
    GC_ABI_END("gc_writer_snapshot_host")
}

gc_status gc_writer_snapshot(gc_writer* w, gc_ctx* ctx, int step, double time,
                             const std::uint32_t* group_ids) {
    GC_ABI_BEGIN
    if (!w || !ctx) return gc::set_error("writer_snapshot: null argument", GC_ERR_INVALID_ARGUMENT);
    size_t n = 0, cap = 0;
    gc_status st = gc::stage_state(ctx, nullptr, 0, &n, &cap, nullptr);  // sizes only
    if (st == GC_OK) st = w->ensure(w->snaps, 10 * cap, "writer_snapshot");
    if (st != GC_OK) return st;
    Job j;  // everything that allocates happens before a slot is taken
    j.step = step;
    j.time = time;
    if (group_ids) j.groups.assign(group_ids, group_ids + n);
    const int k = w->take(w->snaps);
    st = gc::stage_state(ctx, w->snaps.slot(k), w->snaps.stride, &n, &cap, w->snaps.ev[k]);
    if (st != GC_OK) {
        w->give_back(w->snaps, k);
        return st;
    }
    j.n = n;
    j.stride = cap;
    j.staged = w->snaps.slot(k);
    j.pool = &w->snaps;
    j.slot = k;
    w->push(std::move(j));
    return GC_OK;
    GC_ABI_END("gc_writer_snapshot")
}

gc_status gc_writer_diagnostics_values(gc_writer* w, int step, double time, size_t n,
                                       const gc_diagnostics* d) {
    GC_ABI_BEGIN
    if (!w || !d) return gc::set_error("writer_diagnostics: null argument", GC_ERR_INVALID_ARGUMENT);
    Job j;
    j.kind = Job::kDiagnostics;
    j.step = step;
    j.time = time;
    j.n = n;
    j.diag = *d;
    w->push(std::move(j));
    return GC_OK;
    GC_ABI_END("gc_writer_diagnostics_values")
}

gc_status gc_writer_diagnostics(gc_writer* w, gc_ctx* ctx, int step, double time) {
    GC_ABI_BEGIN
    if (!w || !ctx) return gc::set_error("writer_diagnostics: null argument", GC_ERR_INVALID_ARGUMENT);
    gc_status st = w->ensure(w->diags, 16, "writer_diagnostics");
    if (st != GC_OK) return st;
    const int k = w->take(w->diags);
    size_t n = 0;
    st = gc::stage_diagnostics(ctx, w->diags.slot(k), w->diags.ev[k], &n);
    if (st != GC_OK) {
        w->give_back(w->diags, k);
        return st;
    }
    Job j;
    j.kind = Job::kDiagnostics;
    j.step = step;
    j.time = time;
    j.n = n;
    j.staged = w->diags.slot(k);
    j.pool = &w->diags;
    j.slot = k;
    w->push(std::move(j));
    return GC_OK;
    GC_ABI_END("gc_writer_diagnostics")
}

gc_status gc_writer_set_threads(gc_writer* w, int threads) {
    GC_ABI_BEGIN
    if (!w || threads < 0 || threads > 256)
        return gc::set_error("writer_set_threads: threads must be in [0, 256]", GC_ERR_INVALID_ARGUMENT);
    w->drain();  // the worker reads the count per snapshot
    w->threads = threads ? threads : default_threads();
    return GC_OK;
    GC_ABI_END("gc_writer_set_threads")
}

gc_status gc_writer_flush(gc_writer* w) {
    GC_ABI_BEGIN
    if (!w) return gc::set_error("writer_flush: null argument", GC_ERR_INVALID_ARGUMENT);
    w->drain();
    if (std::fflush(w->diag) != 0)
        return gc::set_error("Could not write diagnostics output in " + w->dir.string(), GC_ERR_IO);
    std::lock_guard<std::mutex> lock(w->mu);
    if (!w->err.empty()) return gc::set_error(w->err, GC_ERR_IO);
    return GC_OK;
    GC_ABI_END("gc_writer_flush")
}

gc_status gc_writer_close(gc_writer* w) {
    GC_ABI_BEGIN
    if (!w) return GC_OK;
    w->drain();
    {
        std::lock_guard<std::mutex> lock(w->mu);
        w->stop = true;
    }
    w->cv.notify_all();
    if (w->worker.joinable()) w->worker.join();
    gc_status st = GC_OK;
    if (w->diag && std::fclose(w->diag) != 0)
        st = gc::set_error("Could not write diagnostics output in " + w->dir.string(), GC_ERR_IO);
    if (st == GC_OK && !w->err.empty()) st = gc::set_error(w->err, GC_ERR_IO);
    w->snaps.release();
    w->diags.release();
    delete w;
    return st;
    GC_ABI_END("gc_writer_close")
}

}  // extern "C"
