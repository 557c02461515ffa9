// C ABI of the plain engine (include/mdpd355.h): handles, script commands,
// atoms, run, gather / scatter, thermo, tuning, timing and parity hooks.
//
// Reference counterpart: the LAMMPS library interface src/library.h:27-42
// (lammps_open / lammps_command / lammps_file / lammps_get_natoms /
// lammps_gather_atoms / lammps_scatter_atoms).
#include "system.hpp"

#include <mutex>
#include <new>

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#define MDPD_GUARD(sys, ...)                                   \
    do {                                                       \
        if (!(sys)) return MDPD_E_ARG;                         \
        try {                                                  \
            (sys)->err.clear();                                \
            __VA_ARGS__;                                       \
            return MDPD_OK;                                    \
        } catch (const Fail &f) {                              \
            (sys)->err = f.msg;                                \
            return f.code;                                     \
        } catch (const std::exception &e) {                    \
            (sys)->err = e.what();                             \
            return MDPD_E_ARG;                                 \
        }                                                      \
    } while (0)

extern "C" {

const char *mdpd_version(void) { return "mdpd355 0.1 (gfx950)"; }

int mdpd_open(int device, mdpd_system **out)
{
    // the device is first touched by setup (ensure_stream), so a handle can
    // parse and validate a script on a host without a GPU
    if (!out || device < 0) return MDPD_E_ARG;
    *out = nullptr;
    mdpd_system *s = new (std::nothrow) mdpd_system();
    if (!s) return MDPD_E_NOMEM;
    s->device = device;
    *out = s;
    return MDPD_OK;
}

void mdpd_close(mdpd_system *sys)
{
    if (!sys) return;
    (void)hipSetDevice(sys->device);
    if (sys->stream) (void)hipStreamSynchronize(sys->stream);
    delete sys;
}

const char *mdpd_last_error(const mdpd_system *sys) { return sys ? sys->err.c_str() : "null handle"; }

int mdpd_command(mdpd_system *sys, const char *line)
{
    MDPD_GUARD(sys, { if (!line) fail(MDPD_E_ARG, "null command"); sys->command(line); });
}

int mdpd_commands_string(mdpd_system *sys, const char *text)
{
    MDPD_GUARD(sys, {
        if (!text) fail(MDPD_E_ARG, "null text");
        std::istringstream is(text);
        std::string line, acc;
        while (std::getline(is, line)) {
            // '&' continues a line, as in the LAMMPS input parser
            size_t amp = line.find_last_not_of(" \t\r");
            if (amp != std::string::npos && line[amp] == '&') {
                acc += line.substr(0, amp) + " ";
                continue;
            }
            sys->command(acc + line);
            acc.clear();
        }
        if (!acc.empty()) sys->command(acc);
    });
}

int mdpd_create_box(mdpd_system *sys, int ntypes, const double lo[3], const double hi[3])
{
    MDPD_GUARD(sys, {
        if (sys->box_set) fail(MDPD_E_STATE, "Cannot create_box after simulation box is defined");
        if (ntypes < 1 || ntypes > MAXTYPES) fail(MDPD_E_ARG, "number of atom types must be 1..%d", MAXTYPES);
        if (!lo || !hi) fail(MDPD_E_ARG, "null box bounds");
        for (int d = 0; d < 3; ++d)
            if (!(hi[d] > lo[d])) fail(MDPD_E_ARG, "Create_box region does not support a bounding box");
        sys->ntypes = ntypes;
        for (int d = 0; d < 3; ++d) { sys->lo[d] = lo[d]; sys->hi[d] = hi[d]; }
        sys->box_set = true;
        sys->dirty = true;
    });
}

int mdpd_create_atoms(mdpd_system *sys, int n, const int *tag, const int *type, const int *mask,
                      const double *x, const double *v)
{
    MDPD_GUARD(sys, {
        if (!sys->box_set) fail(MDPD_E_STATE, "Create_atoms command before simulation box is defined");
        if (sys->setup_done) fail(MDPD_E_STATE, "atoms can only be added before the first setup (use mdpd_clear_atoms)");
        if (n < 0 || !type || !x) fail(MDPD_E_ARG, "bad create_atoms arguments");
        int next = 0;
        for (int t : sys->h_tag) next = std::max(next, t);
        for (int i = 0; i < n; ++i) {
            if (type[i] < 1 || type[i] > sys->ntypes) fail(MDPD_E_ARG, "Invalid atom type %d", type[i]);
            if (tag && tag[i] <= 0) fail(MDPD_E_ARG, "Invalid atom ID %d", tag[i]);
        }
        for (int i = 0; i < n; ++i) {
            sys->h_tag.push_back(tag ? tag[i] : ++next);
            sys->h_type.push_back(type[i]);
            sys->h_mask.push_back(mask ? (mask[i] | 1) : 1);
            for (int d = 0; d < 3; ++d) sys->h_image.push_back(0);
            for (int d = 0; d < 3; ++d) {
                sys->h_x.push_back(x[3 * i + d]);
                sys->h_v.push_back(v ? v[3 * i + d] : 0.0);
            }
        }
        sys->dirty = true;
    });
}

int mdpd_clear_atoms(mdpd_system *sys)
{
    MDPD_GUARD(sys, {
        sys->h_tag.clear(); sys->h_type.clear(); sys->h_mask.clear(); sys->h_x.clear(); sys->h_v.clear();
        sys->h_image.clear(); sys->h_phi.clear();
        sys->setup_done = false;
        sys->host_stale = false;
        sys->dirty = true;
    });
}

int mdpd_get_natoms(const mdpd_system *sys) { return sys ? (int)sys->h_tag.size() : 0; }

int mdpd_group_bit(const mdpd_system *sys, const char *name) { return sys && name ? sys->group_bit(name) : 0; }

int mdpd_setup(mdpd_system *sys) { MDPD_GUARD(sys, { sys->setup(); }); }

int mdpd_run(mdpd_system *sys, long long nsteps)
{
    MDPD_GUARD(sys, {
        if (!sys->run_style_ok) fail(MDPD_E_STATE, "run_style mvv/meso must be set for the USER-MESO styles");
        sys->run(nsteps);
    });
}

long long mdpd_get_timestep(const mdpd_system *sys) { return sys ? sys->ntimestep : 0; }
long long mdpd_get_nbuild(const mdpd_system *sys) { return sys ? sys->nbuild : 0; }

int mdpd_gather_atoms(mdpd_system *sys, const char *name, int type, int count, void *data)
{
    MDPD_GUARD(sys, {
        if (!name || !data) fail(MDPD_E_ARG, "null argument");
        sys->gather(name, type, count, data);
    });
}

int mdpd_scatter_atoms(mdpd_system *sys, const char *name, int type, int count, const void *data)
{
    MDPD_GUARD(sys, {
        if (!name || !data) fail(MDPD_E_ARG, "null argument");
        std::string nm(name);
        if (type != 1 || count != 3 || (nm != "x" && nm != "v"))
            fail(MDPD_E_ARG, "scatter supports x and v (type 1, count 3)");
        sys->pull_state();
        std::vector<double> &dst = nm == "x" ? sys->h_x : sys->h_v;
        memcpy(dst.data(), data, sizeof(double) * dst.size());
        sys->dirty = true;
    });
}

int mdpd_thermo(mdpd_system *sys, double out[4])
{
    MDPD_GUARD(sys, {
        if (!out) fail(MDPD_E_ARG, "null output");
        if (!sys->setup_done) fail(MDPD_E_STATE, "thermo requires setup");
        const int n = (int)sys->h_tag.size();
        std::vector<double> v(3 * n), vir(6 * n), e(n);
        std::vector<int> ty(n);
        sys->gather("v", 1, 3, v.data());
        sys->gather("virial", 1, 6, vir.data());
        sys->gather("e_pair", 1, 1, e.data());
        sys->gather("type", 0, 1, ty.data());
        const double V = (sys->hi[0] - sys->lo[0]) * (sys->hi[1] - sys->lo[1]) * (sys->hi[2] - sys->lo[2]);
        thermo_of(n, sys->mass, ty.data(), v.data(), vir.data(), e.data(), V, out);
    });
}

int mdpd_fix_result(mdpd_system *sys, const char *fix_id, int n, double *x, double *value, int *nbin)
{
    MDPD_GUARD(sys, {
        if (!fix_id || !nbin) fail(MDPD_E_ARG, "null argument");
        const FixDesc *f = nullptr;
        for (auto &g : sys->fixes)
            if (g.id == fix_id) f = &g;
        if (!f) fail(MDPD_E_ARG, "could not find fix ID '%s'", fix_id);
        if (!mdpd_system::is_obs(*f)) fail(MDPD_E_ARG, "fix %s has no result vector", fix_id);
        *nbin = f->nbin;
        if (n > 0) {
            if (n < f->nbin || !x || !value) fail(MDPD_E_ARG, "result needs %d bins", f->nbin);
            std::vector<double> xs, vs;
            sys->obs_result(*f, xs, vs);
            std::copy(xs.begin(), xs.end(), x);
            std::copy(vs.begin(), vs.end(), value);
        }
    });
}

int mdpd_tune(mdpd_system *sys, const char *knob, int value)
{
    MDPD_GUARD(sys, {
        if (!knob) fail(MDPD_E_ARG, "null knob");
        if (!strcmp(knob, "split")) {
            if (value != 1 && value != 2 && value != 4) fail(MDPD_E_ARG, "split must be 1, 2 or 4");
            sys->split = value;
            sys->graph_valid = false;
        } else if (!strcmp(knob, "graph")) {
            sys->use_graph = value != 0;
        } else {
            fail(MDPD_E_ARG, "unknown tuning knob '%s'", knob);
        }
    });
}

int mdpd_profile(mdpd_system *sys, int enable)
{
    MDPD_GUARD(sys, {
        sys->prof = enable != 0;
        for (int k = 0; k < PH_COUNT; ++k) { sys->prof_ms[k] = 0; sys->prof_n[k] = 0; }
    });
}

int mdpd_profile_get(mdpd_system *sys, const char *phase, double *total_ms, long long *launches)
{
    MDPD_GUARD(sys, {
        if (!phase || !total_ms || !launches) fail(MDPD_E_ARG, "null argument");
        int k = 0;
        while (k < PH_COUNT && strcmp(PHASE_NAMES[k], phase) != 0) ++k;
        if (k == PH_COUNT) fail(MDPD_E_ARG, "unknown phase '%s'", phase);
        *total_ms = sys->prof_ms[k];
        *launches = sys->prof_n[k];
    });
}

int mdpd_debug_sorted(mdpd_system *sys, int *origin_out, unsigned long long *key_out, float *xw)
{
    MDPD_GUARD(sys, {
        if (!sys->setup_done) fail(MDPD_E_STATE, "requires setup");
        sys->sync_sys();
        const int n = (int)sys->h_tag.size();
        if (origin_out) {
            auto h = sys->download(sys->origin[sys->hsys.sela].p, n);
            hip_check(hipStreamSynchronize(sys->stream), "sync");
            memcpy(origin_out, h.data(), sizeof(int) * n);
        }
        if (key_out) {
            // k_sortgather copies each particle's key to key[cap + sorted slot]
            auto k = sys->download(sys->key.p + sys->cap, n);
            hip_check(hipStreamSynchronize(sys->stream), "sync");
            memcpy(key_out, k.data(), sizeof(unsigned long long) * n);
        }
        if (xw) {
            // positions at the last build: the ones keys and neighbor rows were made from
            auto h = sys->download(sys->xhold.p, n);
            hip_check(hipStreamSynchronize(sys->stream), "sync");
            for (int i = 0; i < n; ++i) { xw[3 * i] = h[i].x; xw[3 * i + 1] = h[i].y; xw[3 * i + 2] = h[i].z; }
        }
    });
}

int mdpd_debug_neighbors(mdpd_system *sys, int *maxnbr, int *counts, int *rows)
{
    MDPD_GUARD(sys, {
        if (!sys->setup_done) fail(MDPD_E_STATE, "requires setup");
        if (!maxnbr) fail(MDPD_E_ARG, "null maxnbr");
        *maxnbr = sys->maxnbr;
        const int n = (int)sys->h_tag.size();
        if (counts) {
            auto h = sys->download(sys->ncount.p, n);
            hip_check(hipStreamSynchronize(sys->stream), "sync");
            memcpy(counts, h.data(), sizeof(int) * n);
        }
        if (rows) {
            const int m = sys->maxnbr;
            auto h = sys->download(sys->nlist.p, (size_t)sys->cap * m);
            auto c = sys->download(sys->ncount.p, n);
            hip_check(hipStreamSynchronize(sys->stream), "sync");
            for (int i = 0; i < n; ++i) {
                const int cnt = std::min(c[i], m);
                for (int p = 0; p < m; ++p)
                    rows[(size_t)i * m + p] = p < cnt ? h[mdpd::nbr_at(i, p, m)] : -1;
            }
        }
    });
}

int mdpd_debug_signatures(mdpd_system *sys, unsigned int *sig)
{
    MDPD_GUARD(sys, {
        if (!sys->setup_done) fail(MDPD_E_STATE, "requires setup");
        if (!sig) fail(MDPD_E_ARG, "null output");
        sys->sync_sys();
        const int n = (int)sys->h_tag.size();
        auto v = sys->download(sys->vel[sys->hsys.sel].p, n);
        auto o = sys->download(sys->origin[sys->hsys.sela].p, n);
        hip_check(hipStreamSynchronize(sys->stream), "sync");
        for (int k = 0; k < n; ++k) {
            unsigned int u;
            memcpy(&u, &v[k].w, 4);
            sig[o[k]] = u;
        }
    });
}

static int stateless(int device, int n, const unsigned int *a, const unsigned int *b, void *out, size_t osz,
                     int rounds, bool gauss)
{
    if (n < 0 || (n > 0 && (!a || !b || !out))) return MDPD_E_ARG;
    if (n == 0) return MDPD_OK;
    if (hipSetDevice(device) != hipSuccess) return MDPD_E_HIP;
    uint32_t *da = nullptr, *db = nullptr;
    void *dout = nullptr;
    int rc = MDPD_OK;
    if (hipMalloc((void **)&da, n * 4) != hipSuccess || hipMalloc((void **)&db, n * 4) != hipSuccess ||
        hipMalloc(&dout, n * osz) != hipSuccess) {
        rc = MDPD_E_NOMEM;
    } else {
        hipError_t e = hipMemcpy(da, a, n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(db, b, n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = gauss ? launch_pair_gaussian(da, db, (float *)dout, n, nullptr)
                      : launch_premix(da, db, (uint32_t *)dout, n, rounds, nullptr);
        if (e == hipSuccess) e = hipMemcpy(out, dout, n * osz, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = MDPD_E_HIP;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return rc;
}

int mdpd_tea_gaussian(int device, int n, const unsigned int *a, const unsigned int *b, float *out)
{
    return stateless(device, n, a, b, out, sizeof(float), 4, true);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pair Gaussians without any allocation: a file-scope device array stages
// the inputs and the outputs (MDPD_TEA_SCRATCH_MAX words each), for callers
// that may not allocate device memory.  Each GPU holds its own copy of the
// array until the process ends; host code uses the device address that
// hipGetSymbolAddress returns for the current device, never the array's
// name.  Every call on a GPU uses the same array: calls that overlap on one
// GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint32_t scratch[3 * (size_t)MDPD_TEA_SCRATCH_MAX];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
uint32_t *scratch_dev[SCRATCH_DEVICES] = {nullptr};

// device address of `scratch` on `device` (made current by the caller),
// looked up on the first call for that device
int scratch_on(int device, uint32_t **p)
{
    if (device < 0 || device >= SCRATCH_DEVICES) return MDPD_E_ARG;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[device]) {
        uint32_t *ptr = nullptr;
        if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MDPD_E_HIP;
        scratch_dev[device] = ptr;
    }
    *p = scratch_dev[device];
    return MDPD_OK;
}
}  // namespace

extern "C" {

int mdpd_tea_gaussian_noalloc(int device, int n, const unsigned int *a, const unsigned int *b, float *out)
{
    if (n < 0 || (n > 0 && (!a || !b || !out))) return MDPD_E_ARG;
    if (n > MDPD_TEA_SCRATCH_MAX) return MDPD_E_OVERFLOW;
    if (n == 0) return MDPD_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MDPD_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return MDPD_E_HIP;
    uint32_t *s = nullptr;
    const int rc = scratch_on(device, &s);
    if (rc != MDPD_OK) return rc;
    uint32_t *da = s, *db = s + MDPD_TEA_SCRATCH_MAX;
    float *dout = (float *)(s + 2 * (size_t)MDPD_TEA_SCRATCH_MAX);
    const size_t bytes = (size_t)n * 4;
    hipError_t e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_pair_gaussian(da, db, dout, n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? MDPD_OK : MDPD_E_HIP;
}

int mdpd_tea_premix(int device, int n, const unsigned int *a, const unsigned int *b, int rounds, unsigned int *out)
{
    if (rounds < 0 || rounds > 256) return MDPD_E_ARG;
    return stateless(device, n, a, b, out, sizeof(unsigned int), rounds, false);
}

}  // extern "C"
