// Pooled engine: many independent periodic systems ("jobs") share one set of
// device arrays and advance together, one launch sequence per pass for the
// whole pool.  It serves both the batched mode (submit many, run to the end)
// and the served mode (jobs arrive while others run, are admitted as particle
// slots free up, and leave when done).
//
// Reference counterparts: the ensemble mode of the reference runs independent
// cases as separate MPI communicators, one process (and launch stream) per
// case (src/main.cpp:49-117); here they are packed into one launch per pass.
// Each job follows the same per-step algorithm as a plain run
// (ModifiedVerlet::run, src/USER-MESO/mvv_meso.cu:309-572), including the
// reneighbor decision of Neighbor::decide (src/neighbor.cpp:1216-1290), made
// per system on the device.
//
// Device side of one pass (all systems at once, every launch skips systems it
// has nothing to do for):
//   re-sort + neighbor rows of the stale systems, commit,
//   density / compact rows, force with the fused velocity-Verlet epilogue
//   (MODE_PHASED: each system's own lifecycle), step commit + decide.
// A pass is captured once in a hipGraph and replayed.  The host reads a
// per-system status word every few passes (asynchronously, one batch behind):
// finished systems are harvested (energy/virial pass + copy-out), their
// particle and cell ranges freed, and queued jobs admitted into the space.
#include "system.hpp"

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <set>

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// first fit over [0, cap) with coalescing of freed ranges
struct RangeAlloc {
    std::map<long long, long long> fr;   // start -> length
    long long cap = 0, used = 0;
    void reset(long long c)
    {
        fr.clear();
        cap = c;
        used = 0;
        if (c > 0) fr[0] = c;
    }
    long long take(long long len)
    {
        for (auto it = fr.begin(); it != fr.end(); ++it) {
            if (it->second < len) continue;
            const long long st = it->first, rest = it->second - len;
            fr.erase(it);
            if (rest > 0) fr[st + len] = rest;
            used += len;
            return st;
        }
        return -1;
    }
    // end of the highest range in use (0 when empty)
    long long high() const
    {
        if (fr.empty()) return cap;
        auto last = std::prev(fr.end());
        return last->first + last->second == cap ? last->first : cap;
    }
    void give(long long st, long long len)
    {
        used -= len;
        auto it = fr.emplace(st, len).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
};

// page-locked staging buffers, reused by size (asynchronous copies need them)
struct PinnedCache {
    std::multimap<size_t, void *> free_;
    ~PinnedCache()
    {
        for (auto &f : free_) (void)hipHostFree(f.second);
    }
    void *get(size_t bytes, size_t &got)
    {
        auto it = free_.lower_bound(bytes);
        if (it != free_.end()) {
            got = it->first;
            void *p = it->second;
            free_.erase(it);
            return p;
        }
        size_t sz = (size_t)1 << 20;
        while (sz < bytes) sz <<= 1;
        void *p = nullptr;
        hip_check(hipHostMalloc(&p, sz, hipHostMallocDefault), "hipHostMalloc");
        got = sz;
        return p;
    }
    void put(void *p, size_t sz) { free_.emplace(sz, p); }
};

enum JobState { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2 };

struct Job {
    int id = 0, state = JOB_QUEUED;
    mdpd_system::Desc ds;
    int n = 0, npad = 0, ncell = 0, ntypes = 0;
    long long nsteps = 0;
    double volume = 0;
    double box[3] = {0, 0, 0};            // edge lengths
    double mass[MAXTYPES + 1] = {0};
    std::vector<float4> hp, hv;           // packed input, npad slots
    std::vector<int> ht, hm, ho, himg;
    std::vector<float> hph;               // wall fraction carried in (fix aboundary/rho/meso)
    std::vector<FixDesc> rdf;             // its rdf/fast/meso fixes (their counts follow the
    double rdf_cut = 0;                   // harvested block), over pairs r < rdf_cut
    std::vector<FixDesc> dens;            // its density/meso fix (at most one): windows
    long long dens_m0 = 0;                // centred at dens_m0 + w * every, w < dens_nwin,
    int dens_nwin = 0;                    // whose counts follow the rdf counts
    size_t rdf_bytes() const
    {
        size_t b = 0;
        for (auto &f : rdf) b += sizeof(unsigned) * (size_t)(f.nbin + 1);
        return b;
    }
    size_t dens_bytes() const { return dens.empty() ? 0 : sizeof(unsigned) * (size_t)dens_nwin * (dens[0].nbin + 1); }
    size_t result_bytes() const { return harvest_bytes(npad) + rdf_bytes() + dens_bytes(); }
    // placement while admitted
    int slot = -1;
    long long part_base = -1, cell_base = -1;
    // harvest: the packed results (launch_harvest's layout, slot order) in a
    // pinned stage, in flight until ev completes, then kept until release
    // (moved to pageable memory when the pool already holds too much pinned)
    bool harvesting = false;
    void *stage = nullptr;
    size_t stage_bytes = 0;
    std::vector<char> kept;
    hipEvent_t ev = nullptr;
    const void *results() const { return stage ? stage : (const void *)kept.data(); }
    long long step_end = 0;
    int nbuild = 0;
    long long pass_admit = 0;             // index of the job's first pass (its setup pass)
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// typed views of a finished job's stage (slot k holds the particle with
// creation index origin[k], k < n)
struct HarvestView {
    const float4 *x, *v, *f;
    const int *tag, *mask, *origin, *image;
    const float *rho, *vir;   // vir: six planes of npad
    const char *desc;
    explicit HarvestView(const Job &j)
    {
        const size_t np = j.npad;
        x = (const float4 *)j.results();
        v = x + np;
        f = x + 2 * np;
        tag = (const int *)(x + 3 * np);
        mask = tag + np;
        origin = tag + 2 * np;
        image = tag + 3 * np;
        rho = (const float *)(tag + 4 * np);
        vir = rho + np;
        desc = (const char *)j.results() + np * HARVEST_SLOT_BYTES;
    }
};
static_assert(HARVEST_SLOT_BYTES == 92, "stage layout");
// the packed results fit the finished system's neighbor-row range, whose
// width is at least NBR_ALIGN entries for at least 256 slots
static_assert(sizeof(SysDev) <= 256 * (4 * NBR_ALIGN - HARVEST_SLOT_BYTES), "harvest block fits the rows");

// bytes of device memory per particle slot (state and scratch), per cell
// slot (rank, coordinates, system; worklist entry, linear range; count and
// start per sub-cell), and per neighbor-row entry (nlist + clist);
// CELLS_PER_SLOT cell slots are kept per particle slot
constexpr double BYTES_PER_SLOT = 256.0;
constexpr double BYTES_PER_CELL = 12.0 + 16.0 + 8.0 + 8.0 * SUBCELLS;
constexpr double BYTES_PER_ENTRY = 8.0;
constexpr int CELLS_PER_SLOT = 2;
constexpr long long RETUNE_PASSES = 256;
constexpr size_t ADMIT_LOOKAHEAD = 64;    // queued jobs considered per admission round
// observable words per descriptor slot: the rdf histograms, then the
// density windows
constexpr size_t SLOT_OBS_WORDS = (size_t)POOL_RDF * (OBS_MAXBIN + 1) + POOL_DENS_WORDS;
constexpr int MAX_BATCH = 16;             // passes captured in one batch graph
constexpr int MAX_BYPASS = 64;            // admissions past one waiting head job

}  // namespace

struct mdpd_pool {
    int device = 0;
    std::string err;
    hipStream_t stream = nullptr;
    double budget = 0;
    int max_sys = 0;

    // ---- device arrays (allocated at the first admission) ----
    bool allocated = false;
    int cap = 0, nslots = 0, maxnbr = 0;
    int nact = 0;                         // launch extent: high-water slot of the resident systems
    int ncact = 0;                        // per-cell launch extent: high-water cell slot
    DevBuf<float4> pos[2], vel[2], xhold, frc, vend, pr;
    DevBuf<int> tag[2], mask[2], origin[2], image[2];
    DevBuf<float> rho, vir, coef, minv, mmass, phi[2];
    DevBuf<int> pcell, pslot, psub, tmp, cell_count, cell_start, cell_xyz, cell_rank, cell_sys, work_count;
    DevBuf<int4> cell_work;
    DevBuf<int2> cell_lse;
    DevBuf<int> nlist, ncount, clist, ccount, wave_sys, moved, status, cl, wall_near;
    DevBuf<SysDev> dsys;
    DevBuf<unsigned> obs_arena;           // per descriptor slot: POOL_RDF histograms of OBS_MAXBIN + 1
                                          // words, then POOL_DENS_WORDS of density windows
    RangeAlloc parts, cells;              // parts in units of 256 slots
    std::vector<int> free_slots, slot_job;

    // ---- jobs ----
    std::map<int, std::unique_ptr<Job>> jobs;
    std::deque<int> queue;
    int head_bypass = 0;                  // jobs admitted past the current head of the queue
    std::vector<int> finished;            // done since the last mdpd_pool_done call
    int next_id = 1, nrunning = 0, nharvest = 0;

    // ---- launch shape ----
    int split = 2;
    bool autotune = true, tuned = false, retuned = false;
    long long retune_at = 0;              // passes after which the split may be tuned again
    double tuned_load = 0;                // resident particle slots when the split was last tuned
    double tune_ms[5] = {0, 0, 0, 0, 0};
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;       // one pass
    hipGraphExec_t gbatch = nullptr;      // batch_len passes, one launch
    int batch_len = 0;
    bool graph_valid = false;
    int nsys_act = 64;                    // per-system launch extent (descriptor slots)
    int maxsub = 16384;                   // sub-cell slots of the largest resident system
    long long captures = 0;           // graphs of the pass captured so far
    int passes_per_poll = 8;
    std::set<std::pair<int, int>> combos;  // (style, single type) of admitted jobs
    bool with_rho = false, with_arb = false, with_obs = false;

    // ---- status polling and staging ----
    int *pin_status[2] = {nullptr, nullptr};
    hipEvent_t ev_poll[2] = {nullptr, nullptr};
    bool poll_pending[2] = {false, false};
    int poll_cur = 0;
    bool need_regrow = false;
    PinnedCache pinned;
    size_t held_pinned = 0;               // stage bytes kept by finished jobs
    size_t max_held = (size_t)1 << 30;    // beyond this their results move to pageable memory
    std::vector<std::pair<hipEvent_t, std::pair<void *, size_t>>> uploads;
    std::vector<hipEvent_t> ev_free;

    // ---- counters ----
    long long passes = 0, regrows = 0;
    double particle_steps = 0;

    // ---- sampled kernel timing: every prof_every-th batch of passes is
    // launched without the graph, with events between the phases ----
    int prof_every = 0;
    long long batches = 0;
    enum { PP_REBUILD = 0, PP_FIRST, PP_FORCE, PP_COMMIT, PP_COUNT };
    struct PassEvents { hipEvent_t e[PP_COUNT + 1]; };
    std::vector<PassEvents> prof_open;
    std::vector<hipEvent_t> prof_free;
    double prof_ms[PP_COUNT] = {0, 0, 0, 0};
    long long prof_passes = 0;
    double prof_particles = 0;            // sum over sampled passes of the particles they evaluate
    double particle_builds = 0;           // finished jobs: sum of particles x neighbor builds

    ~mdpd_pool()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (gbatch) (void)hipGraphExecDestroy(gbatch);
        for (auto &u : uploads) {
            (void)hipEventDestroy(u.first);
            (void)hipHostFree(u.second.first);
        }
        for (auto &j : jobs) {
            if (j.second->ev) (void)hipEventDestroy(j.second->ev);
            if (j.second->stage) (void)hipHostFree(j.second->stage);
        }
        for (auto e : ev_free) (void)hipEventDestroy(e);
        for (auto &pe : prof_open)
            for (auto e : pe.e) (void)hipEventDestroy(e);
        for (auto e : prof_free) (void)hipEventDestroy(e);
        for (int k = 0; k < 2; ++k) {
            if (ev_poll[k]) (void)hipEventDestroy(ev_poll[k]);
            if (pin_status[k]) (void)hipHostFree(pin_status[k]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t take_event()
    {
        if (!ev_free.empty()) {
            hipEvent_t e = ev_free.back();
            ev_free.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        return e;
    }

    void ensure_stream()
    {
        if (stream) return;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev)
            fail(MDPD_E_HIP, "HIP device %d is not available", device);
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (int k = 0; k < 2; ++k) {
            hip_check(hipHostMalloc((void **)&pin_status[k], sizeof(int) * max_sys, hipHostMallocDefault),
                      "hipHostMalloc");
            hip_check(hipEventCreateWithFlags(&ev_poll[k], hipEventDisableTiming), "hipEventCreate");
        }
    }

    PoolDev view()
    {
        PoolDev P{};
        for (int b = 0; b < 2; ++b) {
            P.pos[b] = pos[b].p; P.vel[b] = vel[b].p;
            P.tag[b] = tag[b].p; P.mask[b] = mask[b].p; P.origin[b] = origin[b].p; P.image[b] = image[b].p;
        }
        P.xhold = xhold.p; P.rho = rho.p; P.pr = pr.p; P.frc = frc.p; P.vir = vir.p; P.vend = vend.p;
        P.phi[0] = phi[0].p; P.phi[1] = phi[1].p;
        P.pcell = pcell.p; P.pslot = pslot.p; P.psub = psub.p; P.tmp = tmp.p; P.key = nullptr;
        P.cell_count = cell_count.p; P.cell_start = cell_start.p; P.cell_xyz = cell_xyz.p;
        P.cell_rank = cell_rank.p; P.cell_sys = cell_sys.p;
        P.work_count = work_count.p; P.cell_work = cell_work.p; P.cell_lse = cell_lse.p;
        P.nlist = nlist.p; P.ncount = ncount.p; P.wave_sys = wave_sys.p;
        P.clist = clist.p; P.ccount = ccount.p; P.moved = moved.p; P.status = status.p;
        P.cl = cl.p; P.wall_near = wall_near.p;
        P.sys = dsys.p; P.coef = coef.p; P.minv = minv.p; P.mass = mmass.p;
        P.cap = cap; P.nact = nact; P.ncellslots = nslots; P.ncact = ncact; P.nsys = nsys_act; P.maxsub = maxsub;
        P.maxnbr = maxnbr;
        return P;
    }

    double fixed_bytes() const
    {
        return (double)max_sys * (sizeof(SysDev) + sizeof(float) * (MAXTYPES * MAXTYPES * NCOEF + 2 * MAXTYPES) + 4);
    }
    // particle slots a budget holds at a given row width
    long long capacity_for(int mnb) const
    {
        const double per = BYTES_PER_SLOT + CELLS_PER_SLOT * BYTES_PER_CELL + BYTES_PER_ENTRY * mnb;
        const double c = (budget - fixed_bytes()) / per;
        return c < 256 ? 0 : (long long)(c / 256) * 256;
    }
    double device_bytes() const
    {
        return fixed_bytes() + (BYTES_PER_SLOT + CELLS_PER_SLOT * BYTES_PER_CELL + BYTES_PER_ENTRY * maxnbr) * (double)cap;
    }

    // ------------------------------------------------------------------
    // submit: copy a configured system's state and atoms
    // ------------------------------------------------------------------
    int submit(mdpd_system *sys, long long nsteps)
    {
        if (!sys) fail(MDPD_E_ARG, "null system");
        if (nsteps < 1) fail(MDPD_E_ARG, "a job needs at least one step");
        if (!sys->run_style_ok) fail(MDPD_E_STATE, "run_style mvv/meso must be set for the USER-MESO styles");
        std::vector<FixDesc> rdf, dens;
        for (auto &f : sys->fixes) {
            if (f.style == "rdf/fast/meso") rdf.push_back(f);
            else if (f.style == "density/meso") dens.push_back(f);
        }
        if (rdf.size() > (size_t)POOL_RDF)
            fail(MDPD_E_UNSUPPORTED, "a pool job takes at most %d rdf/fast/meso fixes", POOL_RDF);
        if (dens.size() > 1) fail(MDPD_E_UNSUPPORTED, "a pool job takes at most one density/meso fix");
        // the density windows the job samples: steps t0+1 .. t0+nsteps, each
        // window of `every` steps holds samples (window >= 1)
        long long m0 = 0;
        int nwin = 0;
        if (!dens.empty()) {
            const FixDesc &f = dens[0];
            const long long t0 = sys->ntimestep, t1 = t0 + nsteps, ev = f.every;
            auto centre = [&](long long t) { return (t + ev / 2) / ev * ev; };
            long long a = t0 + 1, b = t1;
            while (a <= t1 && !mdpd_system::density_sample(f, a)) ++a;
            while (b > a && !mdpd_system::density_sample(f, b)) --b;
            if (a <= t1) {
                m0 = centre(a);
                const long long nw = (centre(b) - m0) / ev + 1;
                if (nw * (f.nbin + 1) > POOL_DENS_WORDS)
                    fail(MDPD_E_UNSUPPORTED, "fix %s: %lld windows of %d bins exceed the %d words a pool job keeps",
                         f.id.c_str(), nw, f.nbin, POOL_DENS_WORDS);
                nwin = (int)nw;
            }
        }
        if (sys->setup_done) sys->pull_state();
        auto j = std::make_unique<Job>();
        j->rdf = rdf;
        j->rdf_cut = sys->max_cut();
        j->dens = dens;
        j->dens_m0 = m0;
        j->dens_nwin = nwin;
        sys->describe(j->ds);
        j->n = j->ds.d.n;
        j->npad = j->ds.d.npad;
        j->ncell = j->ds.d.ncell;
        j->ntypes = sys->ntypes;
        j->nsteps = nsteps;
        j->volume = (sys->hi[0] - sys->lo[0]) * (sys->hi[1] - sys->lo[1]) * (sys->hi[2] - sys->lo[2]);
        for (int k = 0; k < 3; ++k) j->box[k] = sys->hi[k] - sys->lo[k];
        for (int t = 0; t <= MAXTYPES; ++t) j->mass[t] = sys->mass[t];
        const int mnb = std::max(maxnbr, j->ds.want_maxnbr);
        const long long c = allocated ? cap : capacity_for(mnb);
        const long long cs = allocated ? nslots : c * CELLS_PER_SLOT;
        if (j->npad > c || j->ncell + 1 > cs)
            fail(MDPD_E_NOMEM, "job of %d atoms / %d cells does not fit a pool of %lld slots / %lld cells", j->n,
                 j->ncell, c, cs);
        j->hp.resize(j->npad); j->hv.resize(j->npad);
        j->ht.resize(j->npad); j->hm.resize(j->npad); j->ho.resize(j->npad); j->himg.resize(j->npad);
        sys->pack(j->npad, j->hp.data(), j->hv.data(), j->ht.data(), j->hm.data(), j->ho.data(), j->himg.data());
        j->hph.assign(j->npad, 0.f);
        for (size_t i = 0; i < sys->h_phi.size() && i < (size_t)j->n; ++i) j->hph[i] = (float)sys->h_phi[i];
        j->id = next_id++;
        if (next_id >= (1 << 27)) next_id = 1;
        j->t_submit = now_s();
        const int id = j->id;
        jobs[id] = std::move(j);
        queue.push_back(id);
        return id;
    }

    // ------------------------------------------------------------------
    // device arrays
    // ------------------------------------------------------------------
    // row width of the first allocation: the widest any queued job asks for
    int planned_maxnbr()
    {
        int mnb = 32;
        for (int id : queue) mnb = std::max(mnb, jobs[id]->ds.want_maxnbr);
        return mnb;
    }

    void allocate()
    {
        ensure_stream();
        maxnbr = planned_maxnbr();
        cap = (int)std::min<long long>(capacity_for(maxnbr), (1LL << IMG_SHIFT) - 256);
        if (cap < 256) fail(MDPD_E_NOMEM, "HBM budget %.3g bytes holds no particles", budget);
        nslots = cap * CELLS_PER_SLOT;
        hip_check(hipSetDevice(device), "hipSetDevice");
        for (int b = 0; b < 2; ++b) {
            pos[b].alloc(cap); vel[b].alloc(cap);
            tag[b].alloc(cap); mask[b].alloc(cap); origin[b].alloc(cap); image[b].alloc(cap);
        }
        xhold.alloc(cap); frc.alloc(cap); vend.alloc(cap);
        rho.alloc(cap); pr.alloc(cap); vir.alloc((size_t)6 * cap); phi[0].alloc(cap); phi[1].alloc(cap);
        pcell.alloc(cap); pslot.alloc(cap); psub.alloc(cap); tmp.alloc(cap);
        ncount.alloc(cap); ccount.alloc((size_t)MAXSPLIT * cap);
        nlist.alloc((size_t)cap * maxnbr);
        clist.alloc((size_t)cap * maxnbr);
        wave_sys.alloc(cap / WAVE); moved.alloc(cap / WAVE); status.alloc(max_sys);
        cl.alloc(cap); wall_near.alloc(cap);
        cell_count.alloc((size_t)nslots * SUBCELLS); cell_start.alloc((size_t)nslots * SUBCELLS);
        cell_xyz.alloc(nslots);
        cell_rank.alloc(nslots); cell_sys.alloc(nslots);
        cell_work.alloc(nslots); cell_lse.alloc(nslots); work_count.alloc(64);
        coef.alloc((size_t)max_sys * MAXTYPES * MAXTYPES * NCOEF);
        minv.alloc((size_t)max_sys * MAXTYPES); mmass.alloc((size_t)max_sys * MAXTYPES);
        dsys.alloc(max_sys);
        hip_check(hipMemsetAsync(cell_count.p, 0, sizeof(int) * (size_t)nslots * SUBCELLS, stream), "memset");
        hip_check(hipMemsetAsync(cell_sys.p, 0xFF, sizeof(int) * nslots, stream), "memset");
        hip_check(hipMemsetAsync(work_count.p, 0, sizeof(int), stream), "memset");
        hip_check(hipMemsetAsync(wave_sys.p, 0xFF, sizeof(int) * (cap / WAVE), stream), "memset");
        hip_check(hipMemsetAsync(status.p, 0, sizeof(int) * max_sys, stream), "memset");
        hip_check(hipMemsetAsync(dsys.p, 0, sizeof(SysDev) * max_sys, stream), "memset");
        hip_check(hipMemsetAsync(ncount.p, 0, sizeof(int) * cap, stream), "memset");
        hip_check(hipStreamSynchronize(stream), "allocate");
        parts.reset(cap / 256);
        cells.reset(nslots);
        free_slots.clear();
        for (int s = max_sys - 1; s >= 0; --s) free_slots.push_back(s);
        slot_job.assign(max_sys, 0);
        allocated = true;
        graph_valid = false;
    }

    // Systems that finished on the device but whose status word the host has
    // not handled yet (it reads them one batch behind) are harvested now,
    // from the device descriptors: their harvest pass (energy and virial of
    // the last pass) reads the compact rows of that pass, which a regrow
    // reallocates.  The stream is drained on return, so the harvest pass has
    // run before any row array is freed.
    void harvest_finished_now()
    {
        std::vector<SysDev> h(max_sys);
        hip_check(hipMemcpyAsync(h.data(), dsys.p, sizeof(SysDev) * max_sys, hipMemcpyDeviceToHost, stream),
                  "readback");
        hip_check(hipStreamSynchronize(stream), "readback");
        bool any = false;
        for (int s = 0; s < max_sys; ++s) {
            const int jid = slot_job[s];
            if (!jid || h[s].job != jid || h[s].phase != PH_DONE) continue;
            Job &j = *jobs[jid];
            if (j.harvesting) continue;
            start_harvest(j);
            any = true;
        }
        if (any) hip_check(hipStreamSynchronize(stream), "harvest before regrow");
    }

    // wider neighbor rows for every system, rows rebuilt in place (no re-sort,
    // so results do not change); the stream is drained first
    void regrow(int want)
    {
        hip_check(hipStreamSynchronize(stream), "regrow");
        harvest_finished_now();
        // status words read before the regrow are stale (overflow flags are
        // cleared below; finished systems keep their word and show up again)
        poll_pending[0] = poll_pending[1] = false;
        for (int attempt = 0;; ++attempt) {
            if (attempt > 6) fail(MDPD_E_OVERFLOW, "<MESO> Pair table overflow; local density too high");
            maxnbr = std::max(want, (int)std::ceil(maxnbr * 1.5 / NBR_ALIGN) * NBR_ALIGN);
            want = 0;
            nlist.alloc((size_t)cap * maxnbr);
            clist.alloc((size_t)cap * maxnbr);
            graph_valid = false;
            ++regrows;
            PoolDev P = view();
                hip_check(launch_relist(P, stream), "relist");
                std::vector<SysDev> h(max_sys);
            hip_check(hipMemcpyAsync(h.data(), dsys.p, sizeof(SysDev) * max_sys, hipMemcpyDeviceToHost, stream),
                      "readback");
            hip_check(hipStreamSynchronize(stream), "relist");
            bool any = false;
            for (int s = 0; s < max_sys; ++s) any |= slot_job[s] && h[s].active && h[s].overflow;
            if (!any) break;
        }
    }

    // ------------------------------------------------------------------
    // admission: FIFO, a job waits until its particle and cell ranges fit
    // ------------------------------------------------------------------
    // Admission in arrival order with bounded backfill: when the oldest
    // queued job does not fit (its particle or cell range is not free in one
    // piece), later jobs among the next ADMIT_LOOKAHEAD that do fit are
    // admitted past it, at most MAX_BYPASS times for the same head job; after
    // that only the head is admitted until it fits.
    bool admit()
    {
        bool any = false;
        if (!queue.empty() && !allocated) allocate();
        size_t k = 0;
        while (k < queue.size() && k < ADMIT_LOOKAHEAD && !free_slots.empty()) {
            if (k > 0 && head_bypass >= MAX_BYPASS) break;
            Job &j = *jobs[queue[k]];
            const long long pu = j.npad / 256;
            const long long pb = parts.take(pu);
            long long cb = -1;
            if (pb >= 0) {
                cb = cells.take(j.ncell + 1);
                if (cb < 0) parts.give(pb, pu);
            }
            if (pb < 0 || cb < 0) {
                ++k;
                continue;
            }
            if (j.ds.want_maxnbr > maxnbr) regrow(j.ds.want_maxnbr);
            const int s = free_slots.back();
            free_slots.pop_back();
            if (k == 0) head_bypass = 0;
            else ++head_bypass;
            queue.erase(queue.begin() + (long)k);
            upload(j, s, pb * 256, cb);
            any = true;
        }
        if (any) update_extent();
        return any;
    }

    void upload(Job &j, int s, long long pbase, long long cbase)
    {
        SysDev d = j.ds.d;
        d.part_base = (int)pbase;
        d.cell_base = (int)cbase;
        d.coef_base = s * MAXTYPES * MAXTYPES * NCOEF;
        d.mass_base = s * MAXTYPES;
        d.maxnbr = maxnbr;
        d.phase = PH_FRESH;
        d.steps_left = (int)std::min<long long>(j.nsteps, 0x7FFFFFFF);
        d.stale = 1;
        d.active = 1;
        d.overflow = 0;
        d.ago = 0;
        d.job = j.id;
        d.sel = 0;
        d.sela = 0;
        d.split_last = split;
        d.nrdf = (int)j.rdf.size();
        d.ndens = j.dens_nwin > 0 ? 1 : 0;
        if (d.nrdf || d.ndens) {
            // the slot's histograms start from zero (ordered before the
            // system's first pass on the stream)
            const size_t per = (size_t)OBS_MAXBIN + 1;
            if (!obs_arena.p) obs_arena.alloc((size_t)max_sys * SLOT_OBS_WORDS);
            unsigned *base = obs_arena.p + (size_t)s * SLOT_OBS_WORDS;
            hip_check(hipMemsetAsync(base, 0, sizeof(unsigned) * SLOT_OBS_WORDS, stream), "memset");
            if (d.ndens) {
                const FixDesc &f = j.dens[0];
                d.dens_nbin = f.nbin;
                d.dens_every = f.every;
                d.dens_window = f.window;
                d.dens_along = f.along;
                d.dens_g = f.groupbit;
                d.dens_nwin = j.dens_nwin;
                d.dens_m0 = j.dens_m0;
                d.dens_hist = base + POOL_RDF * per;
            }
            for (int k = 0; k < d.nrdf; ++k) {
                const FixDesc &f = j.rdf[k];
                d.rdf_nbin[k] = f.nbin;
                d.rdf_every[k] = f.every;
                d.rdf_gi[k] = f.groupbit;
                d.rdf_gj[k] = f.other_bit;
                d.rdf_rc[k] = (float)j.rdf_cut;
                d.rdf_hist[k] = base + (size_t)k * per;
            }
        }
        const int npad = j.npad, nc1 = j.ncell + 1, ntw = npad / WAVE;
        const size_t ncoef = j.ds.coef.size(), nt = j.ntypes;
        // staging layout (every piece 16-byte aligned)
        auto al = [](size_t b) { return (b + 15) / 16 * 16; };
        size_t off[14], at = 0;
        const size_t sz[14] = {sizeof(float4) * npad, sizeof(float4) * npad, sizeof(int) * npad, sizeof(int) * npad,
                               sizeof(int) * npad, sizeof(int) * npad, sizeof(int) * nc1, sizeof(int) * nc1,
                               sizeof(int) * nc1, sizeof(int) * ntw, sizeof(float) * ncoef, sizeof(float) * 2 * nt,
                               sizeof(SysDev), sizeof(float) * npad};
        for (int k = 0; k < 14; ++k) { off[k] = at; at += al(sz[k]); }
        size_t got = 0;
        char *st = (char *)pinned.get(at, got);
        memcpy(st + off[0], j.hp.data(), sz[0]);
        memcpy(st + off[1], j.hv.data(), sz[1]);
        memcpy(st + off[2], j.ht.data(), sz[2]);
        memcpy(st + off[3], j.hm.data(), sz[3]);
        memcpy(st + off[4], j.ho.data(), sz[4]);
        memcpy(st + off[5], j.himg.data(), sz[5]);
        memcpy(st + off[13], j.hph.data(), sz[13]);
        memcpy(st + off[6], j.ds.rank.data(), sz[6]);
        memcpy(st + off[7], j.ds.xyz.data(), sz[7]);
        int *cs = (int *)(st + off[8]);
        for (int k = 0; k < nc1; ++k) cs[k] = s;
        int *ws = (int *)(st + off[9]);
        for (int k = 0; k < ntw; ++k) ws[k] = s;
        memcpy(st + off[10], j.ds.coef.data(), sz[10]);
        float *ms = (float *)(st + off[11]);
        for (size_t t = 0; t < nt; ++t) { ms[t] = j.ds.mass[t]; ms[nt + t] = j.ds.minv[t]; }
        memcpy(st + off[12], &d, sizeof(SysDev));
        auto up = [&](void *dst, int k) {
            hip_check(hipMemcpyAsync(dst, st + off[k], sz[k], hipMemcpyHostToDevice, stream), "upload");
        };
        up(pos[0].p + pbase, 0);
        up(vel[0].p + pbase, 1);
        up(tag[0].p + pbase, 2);
        up(mask[0].p + pbase, 3);
        up(origin[0].p + pbase, 4);
        up(image[0].p + pbase, 5);
        up(cell_rank.p + cbase, 6);
        up(cell_xyz.p + cbase, 7);
        up(cell_sys.p + cbase, 8);
        up(wave_sys.p + pbase / WAVE, 9);
        up(coef.p + d.coef_base, 10);
        hip_check(hipMemcpyAsync(mmass.p + d.mass_base, ms, sizeof(float) * nt, hipMemcpyHostToDevice, stream), "upload");
        hip_check(hipMemcpyAsync(minv.p + d.mass_base, ms + nt, sizeof(float) * nt, hipMemcpyHostToDevice, stream),
                  "upload");
        hip_check(hipMemsetAsync(cell_count.p + (size_t)cbase * SUBCELLS, 0, sizeof(int) * (size_t)nc1 * SUBCELLS, stream),
                  "memset");
        up(phi[0].p + pbase, 13);
        up(dsys.p + s, 12);
        hipEvent_t e = take_event();
        hip_check(hipEventRecord(e, stream), "hipEventRecord");
        uploads.push_back({e, {st, got}});

        j.slot = s;
        j.part_base = pbase;
        j.cell_base = cbase;
        j.state = JOB_RUNNING;
        j.t_admit = now_s();
        j.pass_admit = passes;
        slot_job[s] = j.id;
        ++nrunning;
        const auto combo = std::make_pair(d.style, j.ntypes == 1 ? 1 : 0);
        if (!combos.count(combo)) { combos.insert(combo); graph_valid = false; }
        if (d.style == STYLE_MDPD && d.rho_bit && !with_rho) { with_rho = true; graph_valid = false; }
        if (d.fix_arb && !with_arb) { with_arb = true; graph_valid = false; }
        if ((d.nrdf || d.ndens) && !with_obs) { with_obs = true; graph_valid = false; }
        // the host copy of the input is no longer needed
        std::vector<float4>().swap(j.hp); std::vector<float4>().swap(j.hv);
        std::vector<int>().swap(j.ht); std::vector<int>().swap(j.hm);
        std::vector<int>().swap(j.ho); std::vector<int>().swap(j.himg);
        std::vector<float>().swap(j.hph);
    }

    // launches cover particle slots [0, nact) and cell slots [0, ncact): the
    // highest resident ranges, rounded up to 64K slots so the captured pass is
    // rebuilt only when they move by that
    // Per-system launches cover descriptor slots [0, nsys_act): the highest
    // slot holding a job, rounded up to 64 (free slots are reused last in,
    // first out, so the occupied ones stay low); the chunked scan takes as
    // many workgroups per system as the largest resident system needs.
    void update_extent()
    {
        const long long hw = parts.high() * 256;
        const long long q = 65536;
        const int want = (int)std::min<long long>(cap, (hw + q - 1) / q * q);
        const long long hc = cells.high();
        const int wantc = (int)std::min<long long>(nslots, (hc + q - 1) / q * q);
        int hs = 0;
        long long mc = 0;
        for (int s = 0; s < max_sys; ++s)
            if (slot_job[s]) {
                hs = s + 1;
                const auto it = jobs.find(slot_job[s]);
                if (it != jobs.end()) mc = std::max<long long>(mc, it->second->ncell + 1);
            }
        const int wants = std::min(max_sys, std::max(64, (hs + 63) / 64 * 64));
        // sub-cell slots of the largest system, rounded to 16K so that the
        // captured pass changes only when it moves by that much
        const long long qs = 16384;
        const int wantm = (int)std::max<long long>(qs, (mc * SUBCELLS + qs - 1) / qs * qs);
        if (want != nact || wantc != ncact || wants != nsys_act || wantm != maxsub) {
            nact = want;
            ncact = wantc;
            nsys_act = wants;
            maxsub = wantm;
            graph_valid = false;
        }
    }

    void recycle_uploads(bool block)
    {
        for (size_t k = 0; k < uploads.size();) {
            auto &u = uploads[k];
            const hipError_t q = block ? hipEventSynchronize(u.first) : hipEventQuery(u.first);
            if (q == hipSuccess) {
                pinned.put(u.second.first, u.second.second);
                ev_free.push_back(u.first);
                uploads[k] = uploads.back();
                uploads.pop_back();
            } else if (q == hipErrorNotReady) {
                ++k;
            } else {
                hip_check(q, "upload event");
            }
        }
    }

    // ------------------------------------------------------------------
    // passes
    // ------------------------------------------------------------------
    hipEvent_t timing_event()
    {
        if (!prof_free.empty()) {
            hipEvent_t e = prof_free.back();
            prof_free.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    // one pass; with `ev`, events are recorded between its phases
    void launch_pass(int sp, PassEvents *ev = nullptr)
    {
        PoolDev P = view();
        auto mark = [&](int k) {
            if (ev) hip_check(hipEventRecord(ev->e[k], stream), "hipEventRecord");
        };
        mark(PP_REBUILD);
        hip_check(launch_rebuild(P, stream), "rebuild");
        hip_check(launch_commit_rebuild(P, stream), "commit rebuild");
        mark(PP_FIRST);
        hip_check(launch_first(P, with_rho, sp, stream), "first pass");
        mark(PP_FORCE);
        for (auto &c : combos) hip_check(launch_force(P, c.first, MODE_PHASED, c.second != 0, sp, stream), "force");
        if (with_arb) hip_check(launch_aboundary(P, 1, stream), "aboundary");
        if (with_obs) hip_check(launch_obs_pool(P, stream), "observables");
        mark(PP_COMMIT);
        hip_check(launch_commit_step(P, COMMIT_PHASED, sp, stream), "commit step");
        mark(PP_COUNT);
    }

    // fold finished sampled passes into the per-phase totals
    void collect_profile(bool block)
    {
        for (size_t k = 0; k < prof_open.size();) {
            PassEvents &pe = prof_open[k];
            const hipError_t q = block ? hipEventSynchronize(pe.e[PP_COUNT]) : hipEventQuery(pe.e[PP_COUNT]);
            if (q == hipErrorNotReady) { ++k; continue; }
            hip_check(q, "profile event");
            for (int ph = 0; ph < PP_COUNT; ++ph) {
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, pe.e[ph], pe.e[ph + 1]), "hipEventElapsedTime");
                prof_ms[ph] += ms;
            }
            ++prof_passes;
            for (auto e : pe.e) prof_free.push_back(e);
            prof_open[k] = prof_open.back();
            prof_open.pop_back();
        }
    }

    // npass passes of the current launch shape as one executable graph
    hipGraphExec_t capture_passes(int npass)
    {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            for (int k = 0; k < npass; ++k) launch_pass(split);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
        hipGraphExec_t x = nullptr;
        const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "hipGraphInstantiate");
        return x;
    }

    // the pass as a graph, and a batch of passes (one poll interval, at most
    // MAX_BATCH) as another: one graph launch per batch instead of per pass
    void capture()
    {
        for (hipGraphExec_t *x : {&gexec, &gbatch})
            if (*x) {
                hip_check(hipGraphExecDestroy(*x), "hipGraphExecDestroy");
                *x = nullptr;
            }
        gexec = capture_passes(1);
        batch_len = std::min(passes_per_poll, MAX_BATCH);
        gbatch = batch_len > 1 ? capture_passes(batch_len) : nullptr;
        graph_valid = true;
        ++captures;
    }

    // particles of the systems that pass number p evaluates: a job admitted
    // before pass a runs passes a .. a + nsteps (its setup pass, its steps,
    // the last half-kick).  Scales the sampled kernel times to bytes per
    // launch.  (A system waiting for wider rows is skipped by the passes but
    // counted here; its window then ends later than counted.)
    double active_particles(long long p) const
    {
        double sum = 0;
        for (int sl = 0; sl < max_sys; ++sl) {
            const int jid = slot_job[sl];
            if (!jid) continue;
            const Job &j = *jobs.at(jid);
            if (p >= j.pass_admit && p <= j.pass_admit + j.nsteps) sum += j.n;
        }
        return sum;
    }

    void run_passes(int m)
    {
        if (m <= 0) return;
        const bool sample = prof_every > 0 && batches++ % prof_every == 0;
        if (sample) {
            for (int k = 0; k < m; ++k) {
                PassEvents pe;
                for (auto &e : pe.e) e = timing_event();
                launch_pass(split, &pe);
                prof_open.push_back(pe);
                prof_particles += active_particles(passes + k);
            }
        } else if (use_graph) {
            if (!graph_valid) capture();
            int k = 0;
            if (gbatch)
                for (; k + batch_len <= m; k += batch_len) hip_check(hipGraphLaunch(gbatch, stream), "hipGraphLaunch");
            for (; k < m; ++k) hip_check(hipGraphLaunch(gexec, stream), "hipGraphLaunch");
        } else {
            for (int k = 0; k < m; ++k) launch_pass(split);
        }
        passes += m;
    }

    // time each split on a few real passes (the work is not repeated: every
    // pass advances the pool); results differ between splits only by FP32
    // summation order.  Only the split-dependent part of a pass is compared
    // (first pass + force pass): the rebuild kernels do not depend on the
    // split, and the share of systems re-sorting varies from pass to pass.
    // the resident load moved by 2x either way since the split was tuned
    bool load_moved() const
    {
        const double l = (double)parts.used;
        return l > 2.0 * tuned_load || l < 0.5 * tuned_load;
    }

    void tune()
    {
        tuned = true;
        if (!autotune) return;
        tuned_load = (double)parts.used;
        const int cand[3] = {1, 2, 4};
        double ms[3] = {0, 0, 0};
        std::vector<PassEvents> evs;
        for (int rep = 0; rep < 2; ++rep)
            for (int k = 0; k < 3; ++k) {
                const int c = rep ? cand[2 - k] : cand[k];
                evs.clear();
                for (int p = 0; p < 4; ++p) {
                    PassEvents pe;
                    for (auto &e : pe.e) e = timing_event();
                    launch_pass(c, &pe);
                    evs.push_back(pe);
                }
                hip_check(hipEventSynchronize(evs.back().e[PP_COUNT]), "tune");
                for (auto &pe : evs) {
                    float t = 0;
                    hip_check(hipEventElapsedTime(&t, pe.e[PP_FIRST], pe.e[PP_COMMIT]), "hipEventElapsedTime");
                    ms[rep ? 2 - k : k] += t;
                    for (auto e : pe.e) prof_free.push_back(e);
                }
                passes += 4;
            }
        int best = 0;
        for (int k = 1; k < 3; ++k)
            if (ms[k] < ms[best]) best = k;
        for (int k = 0; k < 3; ++k) tune_ms[cand[k]] = ms[k] / 8.0;
        if (split != cand[best]) graph_valid = false;
        split = cand[best];
    }

    // ------------------------------------------------------------------
    // status, harvest, completion
    // ------------------------------------------------------------------
    void poll()
    {
        const int c = poll_cur;
        hip_check(hipMemcpyAsync(pin_status[c], status.p, sizeof(int) * max_sys, hipMemcpyDeviceToHost, stream),
                  "status");
        hip_check(hipEventRecord(ev_poll[c], stream), "hipEventRecord");
        poll_pending[c] = true;
        const int p = c ^ 1;
        if (poll_pending[p]) {
            hip_check(hipEventSynchronize(ev_poll[p]), "status");
            poll_pending[p] = false;
            handle(pin_status[p]);
        }
        poll_cur = p;
    }

    void drain_polls()
    {
        hip_check(hipStreamSynchronize(stream), "drain");
        for (int k = 0; k < 2; ++k) {
            const int p = poll_cur ^ 1 ^ k;   // older first
            if (poll_pending[p]) {
                poll_pending[p] = false;
                handle(pin_status[p]);
            }
        }
    }

    void handle(const int *st)
    {
        collect_profile(false);
        for (int s = 0; s < max_sys; ++s) {
            const int jid = slot_job[s];
            if (!jid) continue;
            const int w = st[s];
            if ((w >> 3) != jid) continue;          // no commit yet for this occupant
            Job &j = *jobs[jid];
            if (j.harvesting) continue;
            if (w & 2) start_harvest(j);
            if (w & 4) need_regrow = true;
        }
        if (need_regrow) {
            need_regrow = false;
            regrow(0);
        }
    }

    void start_harvest(Job &j)
    {
        PoolDev P = view();
        const int t0 = (int)(j.part_base / WAVE), nt = j.npad / WAVE;
        const int style = j.ds.d.style;
        const bool uni = j.ntypes == 1;
        // energy / virial of the last pass; the launch whose split matches it runs
        for (int sp : {1, 2, 4}) hip_check(launch_force_range(P, style, MODE_HARVEST, uni, sp, t0, nt, stream), "harvest");
        // the results are packed into the system's own neighbor-row range
        // (no pass reads it once the system has stopped) and copied as one
        // block; they stay in the pinned stage until the job is released
        const size_t bytes = harvest_bytes(j.npad);
        int *dst = nlist.p + (size_t)t0 * maxnbr * WAVE;
        const size_t room = (size_t)j.npad * maxnbr * sizeof(int);
        hip_check(launch_harvest(P, j.slot, (int)j.part_base, j.npad, dst, room, stream), "harvest");
        j.stage = pinned.get(j.result_bytes(), j.stage_bytes);
        hip_check(hipMemcpyAsync(j.stage, dst, bytes, hipMemcpyDeviceToHost, stream), "harvest copy");
        size_t at = bytes;
        const unsigned *base = obs_arena.p + (size_t)j.slot * SLOT_OBS_WORDS;
        for (size_t k = 0; k < j.rdf.size(); ++k) {
            const size_t b = sizeof(unsigned) * (size_t)(j.rdf[k].nbin + 1);
            const unsigned *src = base + k * ((size_t)OBS_MAXBIN + 1);
            hip_check(hipMemcpyAsync((char *)j.stage + at, src, b, hipMemcpyDeviceToHost, stream), "harvest copy");
            at += b;
        }
        if (j.dens_bytes())
            hip_check(hipMemcpyAsync((char *)j.stage + at, base + (size_t)POOL_RDF * (OBS_MAXBIN + 1), j.dens_bytes(),
                                     hipMemcpyDeviceToHost, stream),
                      "harvest copy");
        j.ev = take_event();
        hip_check(hipEventRecord(j.ev, stream), "hipEventRecord");
        j.harvesting = true;
        --nrunning;
        ++nharvest;
    }

    // landed copies: the job is done (its packed results stay in the stage),
    // its ranges and slot are freed
    int finish_harvests(bool block)
    {
        int done = 0;
        for (auto &kv : jobs) {
            Job &j = *kv.second;
            if (!j.harvesting) continue;
            const hipError_t q = block ? hipEventSynchronize(j.ev) : hipEventQuery(j.ev);
            if (q == hipErrorNotReady) continue;
            hip_check(q, "harvest event");
            const HarvestView h(j);
            for (int k = 0; k < j.n; ++k)
                if (h.origin[k] < 0 || h.origin[k] >= j.n)
                    fail(MDPD_E_HIP, "harvest of job %d: bad particle origin %d", j.id, h.origin[k]);
            SysDev d;
            memcpy(&d, h.desc, sizeof(SysDev));
            j.step_end = d.step;
            if (held_pinned + j.stage_bytes > max_held) {
                const char *b = (const char *)j.stage;
                j.kept.assign(b, b + j.result_bytes());
                pinned.put(j.stage, j.stage_bytes);
                j.stage = nullptr;
            } else {
                held_pinned += j.stage_bytes;
            }
            ev_free.push_back(j.ev);
            j.ev = nullptr;
            // free the slot and ranges (the harvest kernel has unmapped its
            // tiles and cell slots, ahead of any new occupant's upload)
            parts.give(j.part_base / 256, j.npad / 256);
            cells.give(j.cell_base, j.ncell + 1);
            slot_job[j.slot] = 0;
            free_slots.push_back(j.slot);
            j.slot = -1;
            j.harvesting = false;
            j.state = JOB_DONE;
            j.t_done = now_s();
            --nharvest;
            particle_steps += (double)j.n * (double)j.nsteps;
            j.nbuild = d.builds;
            particle_builds += (double)j.n * (double)d.builds;
            finished.push_back(j.id);
            ++done;
        }
        if (done) update_extent();
        return done;
    }

    // a finished job's density windows: centre step and the mean count per
    // bin over the window's samples (fix_density_profile_meso.cpp:192-215)
    void density_windows(const Job &j, long long *centre, double *value) const
    {
        const int nb = j.dens[0].nbin;
        const unsigned *c = (const unsigned *)((const char *)j.results() + harvest_bytes(j.npad) + j.rdf_bytes());
        for (int w = 0; w < j.dens_nwin; ++w) {
            const unsigned *h = c + (size_t)w * (nb + 1);
            const unsigned samples = h[nb];
            if (centre) centre[w] = j.dens_m0 + (long long)w * j.dens[0].every;
            for (int b = 0; b < nb; ++b) value[(size_t)w * nb + b] = samples ? h[b] / double((float)samples) : 0.0;
        }
    }

    // ------------------------------------------------------------------
    // drive the device.  Returns after max_passes passes (0: no limit), when
    // stop_on_done and a job finished, or when every job is done.
    // ------------------------------------------------------------------
    int advance(long long max_passes, bool stop_on_done)
    {
        int ndone = 0;
        long long launched = 0;
        if (queue.empty() && nrunning == 0 && nharvest == 0) return 0;
        ensure_stream();
        hip_check(hipSetDevice(device), "hipSetDevice");
        for (;;) {
            ndone += finish_harvests(false);
            recycle_uploads(false);
            admit();
            if (stop_on_done && ndone > 0) break;
            if (nrunning == 0) {
                if (nharvest > 0) {
                    drain_polls();
                    ndone += finish_harvests(true);
                    continue;
                }
                if (poll_pending[0] || poll_pending[1]) {
                    drain_polls();
                    continue;
                }
                if (queue.empty()) break;
                fail(MDPD_E_NOMEM, "job %d does not fit the free pool", queue.front());
            }
            if (max_passes > 0 && launched >= max_passes) break;
            // the first passes after a start from raw input are not typical
            // (rows relax, most systems re-sort every pass): the split is
            // tuned again once the pool has run RETUNE_PASSES passes, and
            // after that whenever the resident load has doubled or halved
            // (a served pool filling up or draining), at most once every
            // RETUNE_PASSES passes
            if (!tuned || (autotune && passes >= retune_at && (!retuned || load_moved()))) {
                if (tuned) retuned = true;
                tune();
                retune_at = passes + RETUNE_PASSES;
                launched += 24;
                poll();
                continue;
            }
            long long m = passes_per_poll;
            if (max_passes > 0) m = std::min(m, max_passes - launched);
            run_passes((int)m);
            launched += m;
            poll();
        }
        if (nrunning == 0 && nharvest == 0 && queue.empty()) {
            drain_polls();
            recycle_uploads(true);
            collect_profile(true);
        }
        return ndone;
    }

    // admit what fits now and wait until those jobs are resident
    int admit_now()
    {
        ensure_stream();
        hip_check(hipSetDevice(device), "hipSetDevice");
        const int before = nrunning;
        admit();
        hip_check(hipStreamSynchronize(stream), "admit");
        recycle_uploads(true);
        return nrunning - before;
    }

    Job &job(int id)
    {
        auto it = jobs.find(id);
        if (it == jobs.end()) fail(MDPD_E_ARG, "unknown job %d", id);
        return *it->second;
    }
};

#define POOL_GUARD(pool, ...)                                  \
    do {                                                       \
        if (!(pool)) return MDPD_E_ARG;                        \
        try {                                                  \
            (pool)->err.clear();                               \
            __VA_ARGS__;                                       \
            return MDPD_OK;                                    \
        } catch (const Fail &f) {                              \
            (pool)->err = f.msg;                               \
            return f.code;                                     \
        } catch (const std::exception &e) {                    \
            (pool)->err = e.what();                            \
            return MDPD_E_ARG;                                 \
        }                                                      \
    } while (0)

extern "C" {

int mdpd_pool_open(int device, double hbm_budget_bytes, int max_systems, mdpd_pool **out)
{
    if (!out || device < 0 || max_systems < 1 || max_systems > 65536 || !(hbm_budget_bytes > 0)) return MDPD_E_ARG;
    *out = nullptr;
    mdpd_pool *p = new (std::nothrow) mdpd_pool();
    if (!p) return MDPD_E_NOMEM;
    p->device = device;
    p->budget = hbm_budget_bytes;
    p->max_sys = max_systems;
    *out = p;
    return MDPD_OK;
}

void mdpd_pool_close(mdpd_pool *pool)
{
    if (!pool) return;
    (void)hipSetDevice(pool->device);
    delete pool;
}

const char *mdpd_pool_last_error(const mdpd_pool *pool) { return pool ? pool->err.c_str() : "null handle"; }

int mdpd_pool_submit(mdpd_pool *pool, mdpd_system *sys, long long nsteps, int *job)
{
    POOL_GUARD(pool, {
        if (!job) fail(MDPD_E_ARG, "null job output");
        *job = pool->submit(sys, nsteps);
    });
}

int mdpd_pool_advance(mdpd_pool *pool, long long max_passes, int stop_on_done, int *ndone)
{
    POOL_GUARD(pool, {
        const int d = pool->advance(max_passes, stop_on_done != 0);
        if (ndone) *ndone = d;
    });
}

int mdpd_pool_run(mdpd_pool *pool) { POOL_GUARD(pool, { pool->advance(0, false); }); }

int mdpd_pool_admit(mdpd_pool *pool, int *admitted)
{
    POOL_GUARD(pool, {
        const int a = pool->admit_now();
        if (admitted) *admitted = a;
    });
}

int mdpd_pool_job_state(mdpd_pool *pool, int job, int *state)
{
    POOL_GUARD(pool, {
        if (!state) fail(MDPD_E_ARG, "null output");
        *state = pool->job(job).state;
    });
}

int mdpd_pool_done(mdpd_pool *pool, int *jobs, int max, int *count)
{
    POOL_GUARD(pool, {
        if (!count || (max > 0 && !jobs)) fail(MDPD_E_ARG, "null output");
        const int k = std::min<int>(max, (int)pool->finished.size());
        for (int i = 0; i < k; ++i) jobs[i] = pool->finished[i];
        pool->finished.erase(pool->finished.begin(), pool->finished.begin() + k);
        *count = k;
    });
}

int mdpd_pool_gather(mdpd_pool *pool, int job, const char *name, int type, int count, void *data)
{
    POOL_GUARD(pool, {
        if (!name || !data) fail(MDPD_E_ARG, "null argument");
        Job &j = pool->job(job);
        if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
        const std::string nm(name);
        const HarvestView h(j);
        const int n = j.n;
        // checks the caller's type and width, then clears the output
        auto want = [&](int t, int c) {
            if (type != t || count != c) fail(MDPD_E_ARG, "gather of %s needs type %d count %d", name, t, c);
            memset(data, 0, (size_t)n * c * (t == 1 ? sizeof(double) : sizeof(int)));
        };
        if (nm == "x" || nm == "v" || nm == "f") {
            want(1, 3);
            const float4 *src = nm == "x" ? h.x : (nm == "v" ? h.v : h.f);
            double *o = (double *)data;
            for (int k = 0; k < n; ++k) {
                double *r = o + 3 * (size_t)h.origin[k];
                r[0] = src[k].x; r[1] = src[k].y; r[2] = src[k].z;
            }
        } else if (nm == "e_pair" || nm == "rho") {
            want(1, 1);
            double *o = (double *)data;
            if (nm == "rho")
                for (int k = 0; k < n; ++k) o[h.origin[k]] = h.rho[k];
            else
                for (int k = 0; k < n; ++k) o[h.origin[k]] = h.f[k].w;
        } else if (nm == "virial") {
            want(1, 6);
            double *o = (double *)data;
            for (int k = 0; k < n; ++k)
                for (int c = 0; c < 6; ++c) o[6 * (size_t)h.origin[k] + c] = h.vir[(size_t)c * j.npad + k];
        } else if (nm == "tag" || nm == "type" || nm == "mask") {
            want(0, 1);
            int *o = (int *)data;
            for (int k = 0; k < n; ++k) {
                int val;
                if (nm == "tag") val = h.tag[k];
                else if (nm == "mask") val = h.mask[k];
                else { memcpy(&val, &h.x[k].w, 4); val += 1; }
                o[h.origin[k]] = val;
            }
        } else if (nm == "image") {
            want(0, 3);
            int *o = (int *)data;
            for (int k = 0; k < n; ++k) {
                const int im = h.image[k];
                int *r = o + 3 * (size_t)h.origin[k];
                r[0] = (im & 1023) - 512; r[1] = ((im >> 10) & 1023) - 512; r[2] = ((im >> 20) & 1023) - 512;
            }
        } else {
            fail(MDPD_E_ARG, "unknown per-atom quantity '%s'", name);
        }
    });
}

int mdpd_pool_thermo(mdpd_pool *pool, int job, double out[4])
{
    POOL_GUARD(pool, {
        if (!out) fail(MDPD_E_ARG, "null output");
        Job &j = pool->job(job);
        if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
        // the per-atom arrays in creation order, as a handle sums them
        const HarvestView h(j);
        const int n = j.n;
        std::vector<int> ty(n);
        std::vector<double> v(3 * (size_t)n), vir(6 * (size_t)n), e(n);
        for (int k = 0; k < n; ++k) {
            const size_t i = h.origin[k];
            int t;
            memcpy(&t, &h.x[k].w, 4);
            ty[i] = t + 1;
            v[3 * i] = h.v[k].x; v[3 * i + 1] = h.v[k].y; v[3 * i + 2] = h.v[k].z;
            for (int c = 0; c < 6; ++c) vir[6 * i + c] = h.vir[(size_t)c * j.npad + k];
            e[i] = h.f[k].w;
        }
        thermo_of(n, j.mass, ty.data(), v.data(), vir.data(), e.data(), j.volume, out);
    });
}

int mdpd_pool_fix_result(mdpd_pool *pool, int job, const char *fix_id, int n, double *x, double *value, int *nbin)
{
    POOL_GUARD(pool, {
        if (!fix_id || !nbin) fail(MDPD_E_ARG, "null argument");
        Job &j = pool->job(job);
        if (!j.dens.empty() && j.dens[0].id == fix_id) {
            // density/meso: the job's last window
            *nbin = j.dens[0].nbin;
            if (n > 0) {
                if (n < *nbin || !x || !value) fail(MDPD_E_ARG, "result needs %d bins", *nbin);
                if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
                std::vector<double> v((size_t)std::max(j.dens_nwin, 1) * *nbin, 0.0);
                if (j.dens_nwin) pool->density_windows(j, nullptr, v.data());
                const double bin = j.box[j.dens[0].along] / *nbin;
                for (int b = 0; b < *nbin; ++b) {
                    x[b] = (b + 0.5) * bin;
                    value[b] = v[(size_t)std::max(j.dens_nwin - 1, 0) * *nbin + b];
                }
            }
            return MDPD_OK;
        }
        size_t at = harvest_bytes(j.npad);
        const FixDesc *f = nullptr;
        for (auto &g : j.rdf) {
            if (g.id == fix_id) { f = &g; break; }
            at += sizeof(unsigned) * (size_t)(g.nbin + 1);
        }
        if (!f) fail(MDPD_E_ARG, "job %d has no rdf/fast/meso or density/meso fix '%s'", job, fix_id);
        *nbin = f->nbin;
        if (n > 0) {
            if (n < f->nbin || !x || !value) fail(MDPD_E_ARG, "result needs %d bins", f->nbin);
            if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
            const HarvestView h(j);
            const unsigned *cnt = (const unsigned *)((const char *)j.results() + at);
            std::vector<double> xs, vs;
            mdpd_system::rdf_values(*f, cnt, cnt[f->nbin], h.mask, (size_t)j.n, j.volume, j.rdf_cut, xs, vs);
            std::copy(xs.begin(), xs.end(), x);
            std::copy(vs.begin(), vs.end(), value);
        }
    });
}

int mdpd_pool_fix_windows(mdpd_pool *pool, int job, const char *fix_id, int max_windows, long long *centre,
                          double *value, int *nwin, int *nbin)
{
    POOL_GUARD(pool, {
        if (!fix_id || !nwin || !nbin) fail(MDPD_E_ARG, "null argument");
        Job &j = pool->job(job);
        if (j.dens.empty() || j.dens[0].id != fix_id) fail(MDPD_E_ARG, "job %d has no density/meso fix '%s'", job, fix_id);
        *nwin = j.dens_nwin;
        *nbin = j.dens[0].nbin;
        if (max_windows > 0) {
            if (max_windows < j.dens_nwin || !centre || !value) fail(MDPD_E_ARG, "result needs %d windows", j.dens_nwin);
            if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
            pool->density_windows(j, centre, value);
        }
    });
}

int mdpd_pool_job_info(mdpd_pool *pool, int job, long long *natoms, long long *timestep, double times[3])
{
    POOL_GUARD(pool, {
        Job &j = pool->job(job);
        if (natoms) *natoms = j.n;
        if (timestep) *timestep = j.step_end;
        if (times) { times[0] = j.t_submit; times[1] = j.t_admit; times[2] = j.t_done; }
    });
}

int mdpd_pool_release(mdpd_pool *pool, int job)
{
    POOL_GUARD(pool, {
        Job &j = pool->job(job);
        if (j.state != JOB_DONE) fail(MDPD_E_STATE, "job %d has not finished", job);
        if (j.stage) {
            pool->held_pinned -= j.stage_bytes;
            pool->pinned.put(j.stage, j.stage_bytes);
        }
        j.stage = nullptr;
        pool->jobs.erase(job);
    });
}

int mdpd_pool_tune(mdpd_pool *pool, const char *knob, int value)
{
    POOL_GUARD(pool, {
        if (!knob) fail(MDPD_E_ARG, "null knob");
        const std::string k(knob);
        if (k == "split") {
            if (value != 0 && value != 1 && value != 2 && value != 4) fail(MDPD_E_ARG, "split must be 0 (auto), 1, 2 or 4");
            pool->autotune = value == 0;
            pool->tuned = value != 0;
            pool->retuned = false;
            if (value) pool->split = value;
            pool->graph_valid = false;
        } else if (k == "graph") {
            pool->use_graph = value != 0;
        } else if (k == "hold") {
            if (value < 0) fail(MDPD_E_ARG, "hold must be >= 0 MiB");
            pool->max_held = (size_t)value << 20;
        } else if (k == "profile") {
            if (value < 0) fail(MDPD_E_ARG, "profile must be >= 0");
            pool->prof_every = value;
            for (double &m : pool->prof_ms) m = 0;
            pool->prof_passes = 0;
            pool->prof_particles = 0;
        } else if (k == "poll") {
            if (value < 1 || value > 4096) fail(MDPD_E_ARG, "poll must be 1..4096 passes");
            pool->passes_per_poll = value;
            pool->graph_valid = false;
        } else {
            fail(MDPD_E_ARG, "unknown tuning knob '%s'", knob);
        }
    });
}

int mdpd_pool_stat(mdpd_pool *pool, const char *name, double *value)
{
    POOL_GUARD(pool, {
        if (!name || !value) fail(MDPD_E_ARG, "null argument");
        const std::string k(name);
        if (k == "passes") *value = (double)pool->passes;
        else if (k == "particle_steps") *value = pool->particle_steps;
        else if (k == "split") *value = pool->split;
        else if (k == "capacity") *value = pool->allocated ? pool->cap : (double)pool->capacity_for(pool->planned_maxnbr());
        else if (k == "maxnbr") *value = pool->allocated ? pool->maxnbr : pool->planned_maxnbr();
        else if (k == "bytes") *value = pool->allocated ? pool->device_bytes() : 0.0;
        else if (k == "regrows") *value = (double)pool->regrows;
        else if (k == "captures") *value = (double)pool->captures;
        else if (k == "running") *value = pool->nrunning + pool->nharvest;
        else if (k == "queued") *value = (double)pool->queue.size();
        else if (k == "used_slots") *value = (double)pool->parts.used * 256.0;
        else if (k == "prof_passes") *value = (double)pool->prof_passes;
        else if (k == "prof_particles") *value = pool->prof_particles;
        else if (k == "particle_builds") *value = pool->particle_builds;
        else if (k == "held_pinned") *value = (double)pool->held_pinned;
        else if (k == "prof_rebuild_ms") *value = pool->prof_ms[mdpd_pool::PP_REBUILD];
        else if (k == "prof_first_ms") *value = pool->prof_ms[mdpd_pool::PP_FIRST];
        else if (k == "prof_force_ms") *value = pool->prof_ms[mdpd_pool::PP_FORCE];
        else if (k == "prof_commit_ms") *value = pool->prof_ms[mdpd_pool::PP_COMMIT];
        else if (k == "tune_ms_1") *value = pool->tune_ms[1];
        else if (k == "tune_ms_2") *value = pool->tune_ms[2];
        else if (k == "tune_ms_4") *value = pool->tune_ms[4];
        else fail(MDPD_E_ARG, "unknown statistic '%s'", name);
    });
}

double mdpd_clock(void) { return now_s(); }

}  // extern "C"
