// Host side of one mdpd355 system: the input-script surface of the mDPD/DPD
// styles, device memory ownership and the plain (one system per launch) step
// loop.  The C ABI over it is in engine.cpp; the pooled engine (pool.cpp)
// reads a system's script state and atoms through describe() / pack().
//
// Reference counterparts:
//   script commands    src/input.cpp:542-579, pair_mdpd_meso.cu:299-397,
//                      pair_dpd_meso.cu:272-357, fix_nve_meso.cu:29-40,
//                      fix_aboundary_rho_meso.cu:30-104, fix_addforce_meso.cu:38-47,
//                      atom_vec_mdpd_atomic_meso.cu:19-54, src/force.cpp (bounds)
//   setup / run        src/USER-MESO/mvv_meso.cu:203-283, 309-572
//   bins               src/USER-MESO/neighbor_meso.cu:771-853
//   gather / scatter   src/library.cpp (lammps_gather_atoms / lammps_scatter_atoms)
#pragma once

#include "common.hpp"
#include "kernels.hpp"
#include "../../include/mdpd355.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

namespace mdpd {

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{code, buf};
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) fail(MDPD_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            fail(MDPD_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

inline double parse_num(const std::string &s)
{
    char *end = nullptr;
    double v = strtod(s.c_str(), &end);
    if (!end || *end != '\0' || s.empty()) fail(MDPD_E_ARG, "expected a number, got '%s'", s.c_str());
    return v;
}

inline long long parse_int(const std::string &s)
{
    char *end = nullptr;
    long long v = strtoll(s.c_str(), &end, 10);
    if (!end || *end != '\0' || s.empty()) fail(MDPD_E_ARG, "expected an integer, got '%s'", s.c_str());
    return v;
}

// Force::bounds: "*", "n*", "*n", "m*n", "n" over 1..nmax
inline void type_bounds(const std::string &s, int nmax, int &lo, int &hi)
{
    size_t star = s.find('*');
    if (star == std::string::npos) {
        lo = hi = (int)parse_int(s);
    } else if (s.size() == 1) {
        lo = 1; hi = nmax;
    } else if (star == 0) {
        lo = 1; hi = (int)parse_int(s.substr(1));
    } else if (star == s.size() - 1) {
        lo = (int)parse_int(s.substr(0, star)); hi = nmax;
    } else {
        lo = (int)parse_int(s.substr(0, star)); hi = (int)parse_int(s.substr(star + 1));
    }
    if (lo < 1 || hi > nmax || lo > hi) fail(MDPD_E_ARG, "numeric type range '%s' out of 1..%d", s.c_str(), nmax);
}

enum Phase { PH_REBUILD = 0, PH_DENSITY, PH_FORCE, PH_INITIAL, PH_WALL, PH_COUNT };
const char *const PHASE_NAMES[PH_COUNT] = {"rebuild", "density", "force", "initial", "wall"};

struct FixDesc {
    std::string id, style;
    int groupbit = 0;
    double rd = 0;       // aboundary/rho/meso
    int lucy = 0;
    // aboundary/rho/meso wall branch (fix_aboundary_rho_meso.cu:46-101)
    double cut_phi = 0, phi_c = 0, rho_wall = 0;
    int wall_bit = 0, mobile_bit = 0, arb = 0;
    double f[3] = {0, 0, 0};   // addforce/meso
    // rdf/fast/meso (fix_rdf_fast_meso.cu:37-75) and density/meso
    // (fix_density_profile_meso.cpp:30-83)
    std::string output;
    int nbin = 0, every = 1, window = 0, along = -1, other_bit = 0;
    int exact_pi = 0;    // rdf: shell volumes with pi instead of the reference's 3.1415
};

// dump ID group custom N file columns... (src/dump_custom.cpp)
struct DumpDesc {
    std::string id, file;
    int groupbit = 1;
    long long every = 1;
    std::vector<std::string> cols;
    bool started = false;
};

// device histogram and sample counts of an observable fix
struct ObsState {
    DevBuf<unsigned> hist;       // nbin counts, then the samples taken (rdf: sampled steps;
                                 // density: frames of the current window)
    long long last_dump = -1;
};

// Thermo scalars of one system (compute_temp_meso.cu:49-101: dof = 3N - 3;
// compute_pressure_meso.cu:115-155: P = (dof kT + sum of virial traces) / 3V):
// out = {T, P, pe, ke}.  mass is 1-based by type.
inline void thermo_of(int n, const double *mass, const int *type, const double *v, const double *vir,
                      const double *e, double volume, double out[4])
{
    double ke = 0, vs = 0, pe = 0;
    for (int i = 0; i < n; ++i) {
        ke += mass[type[i]] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        vs += vir[6 * i] + vir[6 * i + 1] + vir[6 * i + 2];
        pe += e[i];
    }
    const double dof = 3.0 * n - 3.0;
    const double T = dof > 0 ? ke / dof : 0;
    out[0] = T;
    out[1] = (dof * T + vs) / 3.0 / volume;
    out[2] = pe;
    out[3] = 0.5 * ke;
}

}  // namespace mdpd

using namespace mdpd;

struct mdpd_system {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // ---- script state ----
    int atom_style = -1;              // STYLE_MDPD / STYLE_DPD
    double rho_est = 0;
    int pair_style = -1;
    double temperature = 1.0, cut_global = 1.0;
    int pair_mobile_bit = 0, pair_wall_bit = 0;   // pair_style mdpd/meso ... mobile wall
    int seed = 0;
    int ntypes = 0;
    bool box_set = false;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double mass[MAXTYPES + 1] = {0};
    bool mass_set[MAXTYPES + 1] = {false};
    // pair coefficients (1-based, full matrix after init_one)
    double cut[MAXTYPES + 1][MAXTYPES + 1] = {{0}}, cut_r[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
    double A[MAXTYPES + 1][MAXTYPES + 1] = {{0}}, B[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
    double gamma[MAXTYPES + 1][MAXTYPES + 1] = {{0}}, sigma[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
    double expw[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
    int setflag[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
    double dt = 0.005;                // LAMMPS lj default
    double skin = 0.3;
    int every = 1, delay = 10, check = 1;
    bool run_style_ok = false;
    std::vector<FixDesc> fixes;
    std::map<std::string, ObsState> obs;
    std::vector<DumpDesc> dumps;
    std::vector<std::pair<std::string, int>> groups{{"all", 1}};
    long long ntimestep = 0;
    int thermo_every = 0;

    // ---- atoms (creation order) ----
    std::vector<int> h_tag, h_type, h_mask, h_image;
    std::vector<double> h_x, h_v;
    std::vector<double> h_phi;        // wall fraction carried between runs (empty: all zero)
    bool host_stale = false;          // a run moved the device state past the host copy

    // ---- device state ----
    bool setup_done = false;
    bool dirty = true;                // configuration changed since setup
    int cap = 0, maxnbr = 0, ncell = 0, nslots = 0;
    long long nbuild = 0;
    SysDev hsys{};
    DevBuf<float4> pos[2], vel[2], xhold, frc, vend, pr;
    DevBuf<int> tag[2], mask[2], origin[2], image[2];
    DevBuf<float> rho, vir, coef, minv, mmass, phi[2];
    DevBuf<int> pcell, pslot, psub, tmp, cell_count, cell_start, cell_xyz, cell_rank, cell_sys, work_count;
    DevBuf<int4> cell_work;
    DevBuf<int2> cell_lse;
    DevBuf<int> nlist, ncount, clist, ccount, wave_sys, moved, status, cl, wall_near;
    DevBuf<unsigned long long> key;
    DevBuf<SysDev> dsys;
    int *pinned = nullptr;            // host-visible readback of flags

    // ---- kernel timing with HIP events (phase -> total ms, launches) ----
    bool prof = false;
    int split = 2;                    // waves per 64-particle tile in the per-step passes
    std::vector<hipEvent_t> ev_free;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
    double prof_ms[PH_COUNT] = {0};
    long long prof_n[PH_COUNT] = {0};

    // ---- step replay: a fused step (rebuild phase, passes, commit with the
    // reneighbor decision) captured once as a hipGraph and replayed; the host
    // reads the device step counter once per batch of STEP_BATCH steps ----
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    std::map<uint32_t, hipGraphExec_t> gobs;   // steps with observable samples, by sample mask
    bool graph_valid = false;
    static constexpr int STEP_BATCH = 64;

    ~mdpd_system()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto &v : gobs) (void)hipGraphExecDestroy(v.second);
        for (auto &f : fixes) {
            try { finish_obs(f); } catch (...) {}
        }
        for (auto &e : ev_open) { (void)hipEventDestroy(e.second.first); (void)hipEventDestroy(e.second.second); }
        for (auto e : ev_free) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ------------------------------------------------------------------
    int group_bit(const std::string &name) const
    {
        for (auto &g : groups)
            if (g.first == name) return g.second;
        return 0;
    }
    int need_group(const std::string &name) const
    {
        int b = group_bit(name);
        if (!b) fail(MDPD_E_ARG, "could not find group ID '%s'", name.c_str());
        return b;
    }
    const FixDesc *find_fix(const char *style) const
    {
        for (auto &f : fixes)
            if (f.style == style) return &f;
        return nullptr;
    }

    // ------------------------------------------------------------------
    // input-script surface
    // ------------------------------------------------------------------
    // ------------------------------------------------------------------
    // equal-style variables (src/variable.cpp): `variable NAME equal EXPR`
    // stores the formula; ${NAME} / $x in later lines is replaced by its
    // value (numbers, + - * / ^, parentheses, unary minus, other variables)
    // ------------------------------------------------------------------
    std::map<std::string, std::string> variables;

    double eval_formula(const std::string &text, int depth = 0)
    {
        if (depth > 32) fail(MDPD_E_ARG, "variable formula nests too deeply");
        const std::string e = substitute(text, depth + 1);
        size_t k = 0;
        auto skip = [&]() { while (k < e.size() && isspace((unsigned char)e[k])) ++k; };
        std::function<double()> expr, term, power, unary, atom;
        atom = [&]() -> double {
            skip();
            if (k < e.size() && e[k] == '(') {
                ++k;
                const double v = expr();
                skip();
                if (k >= e.size() || e[k] != ')') fail(MDPD_E_ARG, "unbalanced parentheses in '%s'", e.c_str());
                ++k;
                return v;
            }
            const char *b = e.c_str() + k;
            char *end = nullptr;
            const double v = strtod(b, &end);
            if (end == b) fail(MDPD_E_ARG, "invalid variable formula '%s'", e.c_str());
            k += (size_t)(end - b);
            return v;
        };
        unary = [&]() -> double {
            skip();
            if (k < e.size() && (e[k] == '-' || e[k] == '+')) {
                const bool neg = e[k] == '-';
                ++k;
                const double v = unary();
                return neg ? -v : v;
            }
            return atom();
        };
        power = [&]() -> double {
            const double b = unary();
            skip();
            if (k < e.size() && e[k] == '^') { ++k; return std::pow(b, power()); }
            return b;
        };
        term = [&]() -> double {
            double v = power();
            for (;;) {
                skip();
                if (k < e.size() && (e[k] == '*' || e[k] == '/')) {
                    const char op = e[k++];
                    const double r = power();
                    v = op == '*' ? v * r : v / r;
                } else {
                    return v;
                }
            }
        };
        expr = [&]() -> double {
            double v = term();
            for (;;) {
                skip();
                if (k < e.size() && (e[k] == '+' || e[k] == '-')) {
                    const char op = e[k++];
                    const double r = term();
                    v = op == '+' ? v + r : v - r;
                } else {
                    return v;
                }
            }
        };
        const double v = expr();
        skip();
        if (k != e.size()) fail(MDPD_E_ARG, "invalid variable formula '%s'", e.c_str());
        return v;
    }

    std::string substitute(const std::string &line, int depth = 0)
    {
        std::string out;
        for (size_t k = 0; k < line.size(); ++k) {
            if (line[k] != '$' || k + 1 >= line.size()) { out += line[k]; continue; }
            std::string name;
            if (line[k + 1] == '{') {
                const size_t close = line.find('}', k + 2);
                if (close == std::string::npos) fail(MDPD_E_ARG, "invalid variable name in '%s'", line.c_str());
                name = line.substr(k + 2, close - k - 2);
                k = close;
            } else {
                name = line.substr(k + 1, 1);
                k += 1;
            }
            auto it = variables.find(name);
            if (it == variables.end()) fail(MDPD_E_ARG, "Substitution for illegal variable %s", name.c_str());
            char buf[64];
            snprintf(buf, sizeof(buf), "%.15g", eval_formula(it->second, depth));
            out += buf;
        }
        return out;
    }

    void command(const std::string &raw)
    {
        std::string line = raw.substr(0, raw.find('#'));
        {
            // `variable NAME equal FORMULA` keeps its formula text unevaluated
            std::istringstream vs(line);
            std::string c0, nm, style;
            vs >> c0 >> nm >> style;
            if (c0 == "variable") {
                if (style != "equal") fail(MDPD_E_UNSUPPORTED, "only equal-style variables are supported");
                std::string rest;
                std::getline(vs, rest);
                if (rest.find_first_not_of(" \t\r") == std::string::npos) fail(MDPD_E_ARG, "Illegal variable command");
                variables[nm] = rest;
                return;
            }
        }
        if (line.find('$') != std::string::npos) line = substitute(line);
        std::istringstream is(line);
        std::vector<std::string> w;
        for (std::string t; is >> t;) w.push_back(t);
        if (w.empty()) return;
        const std::string &c = w[0];
        auto need = [&](size_t n) {
            if (w.size() < n) fail(MDPD_E_ARG, "illegal %s command", c.c_str());
        };
        if (c == "echo" || c == "log" || c == "print" || c == "newton" || c == "comm_modify" ||
            c == "atom_modify" || c == "processors" || c == "thermo_style" || c == "thermo_modify") {
            return;
        } else if (c == "units") {
            need(2);
            if (w[1] != "lj") fail(MDPD_E_UNSUPPORTED, "only units lj is supported");
        } else if (c == "dimension") {
            need(2);
            if (w[1] != "3") fail(MDPD_E_UNSUPPORTED, "only dimension 3 is supported");
        } else if (c == "boundary") {
            need(4);
            if (w[1] != "p" || w[2] != "p" || w[3] != "p") fail(MDPD_E_UNSUPPORTED, "only boundary p p p is supported");
        } else if (c == "atom_style") {
            need(2);
            if (w[1] == "mdpd/atomic/meso") {
                // atom_vec_mdpd_atomic_meso.cu:27-30: a density estimate is required
                if (w.size() < 3)
                    fail(MDPD_E_ARG, "For atom_style mdpd/atomic/meso, an estimation to maximum local particle density must be given. Or input 0 for default.");
                atom_style = STYLE_MDPD;
                rho_est = parse_num(w[2]);
            } else if (w[1] == "dpd/atomic/meso") {
                atom_style = STYLE_DPD;
            } else {
                fail(MDPD_E_UNSUPPORTED, "atom_style %s is outside this engine", w[1].c_str());
            }
            dirty = true;
        } else if (c == "mass") {
            need(3);
            if (!box_set) fail(MDPD_E_STATE, "mass command before simulation box is defined");
            int a, b;
            type_bounds(w[1], ntypes, a, b);
            double m = parse_num(w[2]);
            if (m <= 0) fail(MDPD_E_ARG, "invalid mass value");
            for (int t = a; t <= b; ++t) { mass[t] = m; mass_set[t] = true; }
            dirty = true;
        } else if (c == "group") {
            cmd_group(w);
        } else if (c == "pair_style") {
            cmd_pair_style(w);
        } else if (c == "pair_coeff") {
            cmd_pair_coeff(w);
        } else if (c == "fix") {
            cmd_fix(w);
        } else if (c == "unfix") {
            need(2);
            auto it = std::find_if(fixes.begin(), fixes.end(), [&](const FixDesc &f) { return f.id == w[1]; });
            if (it == fixes.end()) fail(MDPD_E_ARG, "could not find fix ID '%s' to delete", w[1].c_str());
            const bool o = is_obs(*it);
            finish_obs(*it);
            fixes.erase(it);
            if (!o) dirty = true;
        } else if (c == "timestep") {
            need(2);
            dt = parse_num(w[1]);
            if (dt <= 0) fail(MDPD_E_ARG, "illegal timestep");
            dirty = true;
        } else if (c == "neighbor") {
            need(3);
            skin = parse_num(w[1]);
            if (skin < 0) fail(MDPD_E_ARG, "illegal neighbor command");
            if (w[2] != "bin") fail(MDPD_E_UNSUPPORTED, "only neighbor style bin is supported");
            dirty = true;
        } else if (c == "neigh_modify") {
            for (size_t k = 1; k < w.size(); k += 2) {
                if (k + 1 >= w.size()) fail(MDPD_E_ARG, "illegal neigh_modify command");
                if (w[k] == "every") every = (int)parse_int(w[k + 1]);
                else if (w[k] == "delay") delay = (int)parse_int(w[k + 1]);
                else if (w[k] == "check") check = w[k + 1] == "yes";
                else fail(MDPD_E_UNSUPPORTED, "neigh_modify %s is not supported", w[k].c_str());
            }
            if (every < 1 || delay < 0) fail(MDPD_E_ARG, "illegal neigh_modify command");
        } else if (c == "run_style") {
            need(2);
            if (w[1] != "mvv/meso") fail(MDPD_E_UNSUPPORTED, "run_style %s is outside this engine", w[1].c_str());
            run_style_ok = true;
        } else if (c == "reset_timestep") {
            need(2);
            ntimestep = parse_int(w[1]);
            dirty = true;
        } else if (c == "read_data") {
            need(2);
            read_data(w[1]);
        } else if (c == "write_data") {
            need(2);
            write_data(w[1]);
        } else if (c == "write_restart") {
            need(2);
            write_restart(w[1]);
        } else if (c == "read_restart") {
            need(2);
            read_restart(w[1]);
        } else if (c == "velocity") {
            cmd_velocity(w);
        } else if (c == "dump") {
            cmd_dump(w);
        } else if (c == "undump") {
            need(2);
            auto it = std::find_if(dumps.begin(), dumps.end(), [&](const DumpDesc &d) { return d.id == w[1]; });
            if (it == dumps.end()) fail(MDPD_E_ARG, "Could not find undump ID %s", w[1].c_str());
            dumps.erase(it);
        } else if (c == "dump_modify") {
        } else if (c == "thermo") {
            need(2);
            thermo_every = (int)parse_int(w[1]);
        } else if (c == "run") {
            need(2);
            run(parse_int(w[1]));
        } else {
            fail(MDPD_E_UNSUPPORTED, "unknown or unsupported command: %s", c.c_str());
        }
    }

    // read_data for atom_style mdpd/atomic/meso or dpd/atomic/meso: header
    // (atoms, atom types, xlo xhi, ylo yhi, zlo zhi), Masses, Atoms
    // (id type x y z [ix iy iz]; AtomVecMDPDAtomic::data_atom,
    // atom_vec_mdpd_atomic_meso.cu:609-637) and Velocities (id vx vy vz);
    // src/read_data.cpp
    void read_data(const std::string &path)
    {
        if (box_set) fail(MDPD_E_STATE, "Cannot read_data after simulation box is defined");
        if (atom_style < 0) fail(MDPD_E_STATE, "atom_style must be set before read_data");
        std::ifstream in(path);
        if (!in) fail(MDPD_E_ARG, "Cannot open file %s", path.c_str());
        auto strip = [](std::string l) {
            l = l.substr(0, l.find('#'));
            const size_t a = l.find_first_not_of(" \t\r"), b = l.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : l.substr(a, b - a + 1);
        };
        std::string line;
        std::getline(in, line);                       // title line
        long long natoms = -1;
        int nt = -1;
        double blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
        bool have[3] = {false, false, false};
        std::string section;
        // header
        while (std::getline(in, line)) {
            const std::string t = strip(line);
            if (t.empty()) continue;
            std::istringstream is(t);
            std::vector<std::string> f;
            for (std::string x; is >> x;) f.push_back(x);
            if (f.size() == 2 && f[1] == "atoms") natoms = parse_int(f[0]);
            else if (f.size() == 3 && f[1] == "atom" && f[2] == "types") nt = (int)parse_int(f[0]);
            else if (f.size() == 4 && f[2] == "xlo" && f[3] == "xhi") { blo[0] = parse_num(f[0]); bhi[0] = parse_num(f[1]); have[0] = true; }
            else if (f.size() == 4 && f[2] == "ylo" && f[3] == "yhi") { blo[1] = parse_num(f[0]); bhi[1] = parse_num(f[1]); have[1] = true; }
            else if (f.size() == 4 && f[2] == "zlo" && f[3] == "zhi") { blo[2] = parse_num(f[0]); bhi[2] = parse_num(f[1]); have[2] = true; }
            else if (f.size() == 6 && f[3] == "xy") fail(MDPD_E_UNSUPPORTED, "triclinic boxes are outside this engine");
            else if (f.size() >= 2 && (f[1] == "bonds" || f[1] == "angles" || f[1] == "dihedrals" || f[1] == "impropers")) {
                if (parse_int(f[0]) != 0) fail(MDPD_E_UNSUPPORTED, "molecular topology is outside this engine");
            } else if (f.size() >= 3 && f[2] == "types") {
            } else { section = f[0]; break; }
        }
        if (natoms < 0 || nt < 1) fail(MDPD_E_ARG, "Invalid data file header in %s", path.c_str());
        if (nt > MAXTYPES) fail(MDPD_E_ARG, "number of atom types must be 1..%d", MAXTYPES);
        for (int d = 0; d < 3; ++d)
            if (!have[d] || !(bhi[d] > blo[d])) fail(MDPD_E_ARG, "Invalid box bounds in data file %s", path.c_str());
        ntypes = nt;
        for (int d = 0; d < 3; ++d) { lo[d] = blo[d]; hi[d] = bhi[d]; }
        box_set = true;
        std::vector<double> vx;
        std::map<int, size_t> where;
        const size_t first = h_tag.size();
        while (!section.empty()) {
            const std::string sec = section;
            section.clear();
            long long rows = 0;
            const long long want = sec == "Masses" ? nt : natoms;
            if (sec != "Masses" && sec != "Atoms" && sec != "Velocities")
                fail(MDPD_E_UNSUPPORTED, "read_data section %s is outside this engine", sec.c_str());
            while (rows < want && std::getline(in, line)) {
                const std::string t = strip(line);
                if (t.empty()) continue;
                std::istringstream is(t);
                std::vector<std::string> f;
                for (std::string x; is >> x;) f.push_back(x);
                if (sec == "Masses") {
                    if (f.size() < 2) fail(MDPD_E_ARG, "Invalid Masses line in data file");
                    const int ty = (int)parse_int(f[0]);
                    if (ty < 1 || ty > nt) fail(MDPD_E_ARG, "Invalid type for mass set");
                    mass[ty] = parse_num(f[1]);
                    if (mass[ty] <= 0) fail(MDPD_E_ARG, "Invalid mass value");
                    mass_set[ty] = true;
                } else if (sec == "Atoms") {
                    if (f.size() != 5 && f.size() != 8) fail(MDPD_E_ARG, "Incorrect atom format in data file");
                    const int tg = (int)parse_int(f[0]), ty = (int)parse_int(f[1]);
                    if (tg <= 0) fail(MDPD_E_ARG, "Invalid atom ID in Atoms section of data file");
                    if (ty <= 0 || ty > nt) fail(MDPD_E_ARG, "Invalid atom type in Atoms section of data file");
                    where[tg] = h_tag.size();
                    h_tag.push_back(tg);
                    h_type.push_back(ty);
                    h_mask.push_back(1);
                    for (int d = 0; d < 3; ++d) h_x.push_back(parse_num(f[2 + d]));
                    for (int d = 0; d < 3; ++d) h_image.push_back(f.size() == 8 ? (int)parse_int(f[5 + d]) : 0);
                    for (int d = 0; d < 3; ++d) h_v.push_back(0.0);
                } else {
                    if (f.size() != 4) fail(MDPD_E_ARG, "Incorrect velocity format in data file");
                    auto it = where.find((int)parse_int(f[0]));
                    if (it == where.end()) fail(MDPD_E_ARG, "Invalid atom ID in Velocities section of data file");
                    for (int d = 0; d < 3; ++d) h_v[3 * it->second + d] = parse_num(f[1 + d]);
                }
                ++rows;
            }
            if (rows < want) fail(MDPD_E_ARG, "Unexpected end of data file in section %s", sec.c_str());
            while (std::getline(in, line)) {
                const std::string t = strip(line);
                if (t.empty()) continue;
                std::istringstream is(t);
                is >> section;
                break;
            }
        }
        if ((long long)(h_tag.size() - first) != natoms) fail(MDPD_E_ARG, "Did not assign all atoms correctly");
        dirty = true;
    }

    // write_data: header, Masses, Atoms (id type x y z ix iy iz, positions
    // wrapped into the box with the image flags adjusted) and Velocities, in
    // creation order; src/write_data.cpp
    void write_data(const std::string &path)
    {
        if (!box_set) fail(MDPD_E_STATE, "Write_data command before simulation box is defined");
        pull_state();
        FILE *f = fopen(path.c_str(), "w");
        if (!f) fail(MDPD_E_ARG, "Cannot open data file %s", path.c_str());
        const size_t n = h_tag.size();
        fprintf(f, "LAMMPS data file via write_data (mdpd355), timestep = %lld\n\n", ntimestep);
        fprintf(f, "%zu atoms\n%d atom types\n\n", n, ntypes);
        const char *ax[3] = {"x", "y", "z"};
        for (int d = 0; d < 3; ++d) fprintf(f, "%.16e %.16e %slo %shi\n", lo[d], hi[d], ax[d], ax[d]);
        fprintf(f, "\nMasses\n\n");
        for (int t = 1; t <= ntypes; ++t) fprintf(f, "%d %.16g\n", t, mass_set[t] ? mass[t] : 0.0);
        fprintf(f, "\nAtoms # %s\n\n", atom_style == STYLE_DPD ? "dpd/atomic/meso" : "mdpd/atomic/meso");
        for (size_t i = 0; i < n; ++i) {
            double x[3];
            int im[3];
            for (int d = 0; d < 3; ++d) {
                const double L = hi[d] - lo[d];
                const double k = std::floor((h_x[3 * i + d] - lo[d]) / L);
                x[d] = h_x[3 * i + d] - k * L;
                if (x[d] >= hi[d]) x[d] = lo[d];
                im[d] = h_image[3 * i + d] + (int)k;
            }
            fprintf(f, "%d %d %.16e %.16e %.16e %d %d %d\n", h_tag[i], h_type[i], x[0], x[1], x[2], im[0], im[1], im[2]);
        }
        fprintf(f, "\nVelocities\n\n");
        for (size_t i = 0; i < n; ++i)
            fprintf(f, "%d %.16e %.16e %.16e\n", h_tag[i], h_v[3 * i], h_v[3 * i + 1], h_v[3 * i + 2]);
        const bool ok = fclose(f) == 0;
        if (!ok) fail(MDPD_E_ARG, "error writing data file %s", path.c_str());
    }

    // ------------------------------------------------------------------
    // write_restart / read_restart: checkpoint of one system in this
    // engine's own binary file (little-endian, version 1).  Each atom is the
    // record AtomVecMDPDAtomic::pack_restart writes
    // (atom_vec_mdpd_atomic_meso.cu:548-570), 13 doubles
    //     [13, x, y, z, vx, vy, vz, rho, phi, tag, type, mask, image]
    // with the integers as double values and image packed as
    // (ix+512) | (iy+512) << 10 | (iz+512) << 20; atoms in creation order.
    // Before the records the file holds what the LAMMPS restart header and
    // the pair section hold for these styles (src/write_restart.cpp,
    // MesoPairMDPD::write_restart pair_mdpd_meso.cu:403-418 and
    // MesoPairDPD::write_restart pair_dpd_meso.cu:363-378):
    //     char[8] "MDPD355R", int32 version, int32 0x01020304,
    //     int32 atom_style, double rho_est, int64 ntimestep, double dt,
    //     int32 ntypes, double lo[3], double hi[3],
    //     ntypes x (int32 mass_set, double mass),
    //     int32 ngroups, ngroups x (int32 bit, int32 len, char[len] name),
    //     int32 pair_style (-1: none), double temperature, double cut_global,
    //     int32 seed, int32 mobile_bit, int32 wall_bit,
    //     for 1 <= i <= j <= ntypes: int32 setflag, and when set
    //         double A, B, gamma, sigma, expw, cut, cut_r,
    //     int64 natoms, natoms x double[13].
    // Fixes, dumps, neighbor settings and run_style are not stored: the
    // input script gives them again after read_restart, as in LAMMPS.
    // rho is recomputed by the setup pass before the first force
    // (FixArbitraryBoundaryRho::setup_pre_force, fix_aboundary_rho_meso.cu:214);
    // phi is state the force pass reads, so it is carried over.
    // ------------------------------------------------------------------
    static constexpr int RESTART_VERSION = 1;
    static constexpr int RESTART_RECORD = 13;

    void write_restart(const std::string &path)
    {
        if (!box_set || atom_style < 0) fail(MDPD_E_STATE, "Write_restart command before simulation box is defined");
        pull_state();
        const size_t n = h_tag.size();
        std::vector<double> rh(n, 0.0);
        if (setup_done && n) gather("rho", 1, 1, rh.data());
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) fail(MDPD_E_ARG, "Cannot open restart file %s", path.c_str());
        bool ok = true;
        auto put = [&](const void *p, size_t b) { ok = ok && fwrite(p, 1, b, f) == b; };
        auto i32 = [&](int v) { put(&v, 4); };
        auto i64 = [&](long long v) { int64_t w = v; put(&w, 8); };
        auto f64 = [&](double v) { put(&v, 8); };
        put("MDPD355R", 8);
        i32(RESTART_VERSION);
        i32(0x01020304);
        i32(atom_style);
        f64(rho_est);
        i64(ntimestep);
        f64(dt);
        i32(ntypes);
        for (int d = 0; d < 3; ++d) f64(lo[d]);
        for (int d = 0; d < 3; ++d) f64(hi[d]);
        for (int t = 1; t <= ntypes; ++t) { i32(mass_set[t] ? 1 : 0); f64(mass[t]); }
        i32((int)groups.size());
        for (auto &g : groups) {
            i32(g.second);
            i32((int)g.first.size());
            put(g.first.data(), g.first.size());
        }
        i32(pair_style);
        f64(temperature);
        f64(cut_global);
        i32(seed);
        i32(pair_mobile_bit);
        i32(pair_wall_bit);
        for (int i = 1; i <= ntypes; ++i)
            for (int j = i; j <= ntypes; ++j) {
                i32(setflag[i][j]);
                if (setflag[i][j])
                    for (double v : {A[i][j], B[i][j], gamma[i][j], sigma[i][j], expw[i][j], cut[i][j], cut_r[i][j]})
                        f64(v);
            }
        i64((long long)n);
        std::vector<double> rec((size_t)RESTART_RECORD * n);
        for (size_t i = 0; i < n; ++i) {
            double *r = &rec[(size_t)RESTART_RECORD * i];
            const int im = ((h_image[3 * i] + 512) & 1023) | (((h_image[3 * i + 1] + 512) & 1023) << 10) |
                           (((h_image[3 * i + 2] + 512) & 1023) << 20);
            r[0] = RESTART_RECORD;
            for (int d = 0; d < 3; ++d) { r[1 + d] = h_x[3 * i + d]; r[4 + d] = h_v[3 * i + d]; }
            r[7] = rh[i];
            r[8] = i < h_phi.size() ? h_phi[i] : 0.0;
            r[9] = h_tag[i];
            r[10] = h_type[i];
            r[11] = h_mask[i];
            r[12] = im;
        }
        put(rec.data(), rec.size() * sizeof(double));
        ok = (fclose(f) == 0) && ok;
        if (!ok) fail(MDPD_E_ARG, "error writing restart file %s", path.c_str());
    }

    void read_restart(const std::string &path)
    {
        if (box_set) fail(MDPD_E_STATE, "Cannot read_restart after simulation box is defined");
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) fail(MDPD_E_ARG, "Cannot open restart file %s", path.c_str());
        std::unique_ptr<FILE, int (*)(FILE *)> hold(f, fclose);
        auto get = [&](void *p, size_t b) {
            if (fread(p, 1, b, f) != b) fail(MDPD_E_ARG, "Unexpected end of restart file %s", path.c_str());
        };
        auto i32 = [&]() { int v; get(&v, 4); return v; };
        auto i64 = [&]() { int64_t v; get(&v, 8); return (long long)v; };
        auto f64 = [&]() { double v; get(&v, 8); return v; };
        char magic[8];
        get(magic, 8);
        if (memcmp(magic, "MDPD355R", 8) != 0) fail(MDPD_E_ARG, "%s is not a restart file of this engine", path.c_str());
        if (i32() != RESTART_VERSION) fail(MDPD_E_UNSUPPORTED, "restart file %s has an unknown version", path.c_str());
        if (i32() != 0x01020304) fail(MDPD_E_UNSUPPORTED, "restart file %s has the other byte order", path.c_str());
        const int style = i32();
        if (style != STYLE_MDPD && style != STYLE_DPD) fail(MDPD_E_ARG, "Invalid atom style in restart file");
        const double rest = f64();
        const long long step = i64();
        const double tstep = f64();
        const int nt = i32();
        if (nt < 1 || nt > MAXTYPES) fail(MDPD_E_ARG, "number of atom types must be 1..%d", MAXTYPES);
        double blo[3], bhi[3];
        for (int d = 0; d < 3; ++d) blo[d] = f64();
        for (int d = 0; d < 3; ++d) bhi[d] = f64();
        for (int d = 0; d < 3; ++d)
            if (!(bhi[d] > blo[d])) fail(MDPD_E_ARG, "Invalid box bounds in restart file");
        if (!(tstep > 0) || step < 0) fail(MDPD_E_ARG, "Invalid time step in restart file");
        int mset[MAXTYPES + 1] = {0};
        double mval[MAXTYPES + 1] = {0};
        for (int t = 1; t <= nt; ++t) { mset[t] = i32(); mval[t] = f64(); }
        const int ng = i32();
        if (ng < 1 || ng > 32) fail(MDPD_E_ARG, "Invalid group count in restart file");
        std::vector<std::pair<std::string, int>> gs;
        for (int k = 0; k < ng; ++k) {
            const int bit = i32(), len = i32();
            if (len < 1 || len > 4096 || bit != (1 << k)) fail(MDPD_E_ARG, "Invalid group in restart file");
            std::string name((size_t)len, '\0');
            get(&name[0], (size_t)len);
            gs.push_back({name, bit});
        }
        if (gs[0].first != "all") fail(MDPD_E_ARG, "Invalid group in restart file");
        const int ps = i32();
        if (ps != -1 && ps != STYLE_MDPD && ps != STYLE_DPD) fail(MDPD_E_ARG, "Invalid pair style in restart file");
        const double temp = f64(), cg = f64();
        const int sd = i32(), mob = i32(), wal = i32();
        const int allbits = (int)((1ull << ng) - 1);
        if ((mob & ~allbits) || (wal & ~allbits)) fail(MDPD_E_ARG, "Invalid pair groups in restart file");
        int sf[MAXTYPES + 1][MAXTYPES + 1] = {{0}};
        double cv[MAXTYPES + 1][MAXTYPES + 1][7] = {{{0}}};
        for (int i = 1; i <= nt; ++i)
            for (int j = i; j <= nt; ++j) {
                sf[i][j] = i32();
                if (sf[i][j])
                    for (int c = 0; c < 7; ++c) cv[i][j][c] = f64();
            }
        const long long natoms = i64();
        if (natoms < 0 || natoms > 0x7FFFFFFF) fail(MDPD_E_ARG, "Invalid atom count in restart file");
        {
            // the records must be exactly what is left of the file
            const long at = ftell(f);
            if (at < 0 || fseek(f, 0, SEEK_END) != 0) fail(MDPD_E_ARG, "Cannot read restart file %s", path.c_str());
            const long end = ftell(f);
            if (end < 0 || fseek(f, at, SEEK_SET) != 0) fail(MDPD_E_ARG, "Cannot read restart file %s", path.c_str());
            if ((long long)(end - at) != natoms * (long long)(RESTART_RECORD * sizeof(double)))
                fail(MDPD_E_ARG, "restart file %s holds %ld bytes of atom records for %lld atoms", path.c_str(),
                     end - at, natoms);
        }
        const size_t n = (size_t)natoms;
        std::vector<double> rec((size_t)RESTART_RECORD * n);
        get(rec.data(), rec.size() * sizeof(double));
        {
            char extra;
            if (fread(&extra, 1, 1, f) != 0) fail(MDPD_E_ARG, "restart file %s has trailing data", path.c_str());
        }
        // an integer field of a record: a whole number in int range
        auto whole = [&](double v) {
            if (!(v >= -2147483648.0 && v <= 2147483647.0) || v != std::floor(v))
                fail(MDPD_E_ARG, "Invalid atom record in restart file");
            return (int)v;
        };
        std::vector<int> tg(n), ty(n), mk(n), im(3 * n);
        for (size_t i = 0; i < n; ++i) {
            const double *r = &rec[(size_t)RESTART_RECORD * i];
            if (r[0] != RESTART_RECORD) fail(MDPD_E_ARG, "Invalid atom record in restart file");
            for (int c = 1; c <= 8; ++c)
                if (!std::isfinite(r[c])) fail(MDPD_E_ARG, "Invalid atom record in restart file");
            tg[i] = whole(r[9]);
            ty[i] = whole(r[10]);
            mk[i] = whole(r[11]);
            const int pk = whole(r[12]);
            if (tg[i] <= 0 || ty[i] < 1 || ty[i] > nt || !(mk[i] & 1) || (mk[i] & ~allbits) || pk < 0 ||
                pk >= (1 << 30))
                fail(MDPD_E_ARG, "Invalid atom record in restart file");
            for (int d = 0; d < 3; ++d) im[3 * i + d] = ((pk >> (10 * d)) & 1023) - 512;
        }
        // everything read and checked: install it
        atom_style = style;
        rho_est = rest;
        ntimestep = step;
        dt = tstep;
        ntypes = nt;
        for (int d = 0; d < 3; ++d) { lo[d] = blo[d]; hi[d] = bhi[d]; }
        box_set = true;
        for (int t = 1; t <= nt; ++t) { mass_set[t] = mset[t] != 0; mass[t] = mval[t]; }
        groups = gs;
        if (ps != -1) {
            pair_style = ps;
            temperature = temp;
            cut_global = cg;
            seed = sd;
            pair_mobile_bit = mob;
            pair_wall_bit = wal;
            for (int i = 1; i <= nt; ++i)
                for (int j = i; j <= nt; ++j) {
                    setflag[i][j] = sf[i][j];
                    A[i][j] = cv[i][j][0]; B[i][j] = cv[i][j][1]; gamma[i][j] = cv[i][j][2];
                    sigma[i][j] = cv[i][j][3]; expw[i][j] = cv[i][j][4]; cut[i][j] = cv[i][j][5];
                    cut_r[i][j] = cv[i][j][6];
                }
        }
        h_tag = tg;
        h_type = ty;
        h_mask = mk;
        h_image = im;
        h_x.resize(3 * n);
        h_v.resize(3 * n);
        h_phi.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const double *r = &rec[(size_t)RESTART_RECORD * i];
            for (int d = 0; d < 3; ++d) { h_x[3 * i + d] = r[1 + d]; h_v[3 * i + d] = r[4 + d]; }
            h_phi[i] = r[8];
        }
        setup_done = false;
        host_stale = false;
        dirty = true;
    }

    void cmd_dump(const std::vector<std::string> &w)
    {
        if (w.size() < 7) fail(MDPD_E_ARG, "Illegal dump command");
        if (w[3] != "custom") fail(MDPD_E_UNSUPPORTED, "dump style %s is not supported (custom only)", w[3].c_str());
        DumpDesc d;
        d.id = w[1];
        d.groupbit = need_group(w[2]);
        d.every = parse_int(w[4]);
        if (d.every <= 0) fail(MDPD_E_ARG, "Illegal dump command");
        d.file = w[5];
        static const char *known[] = {"id", "type", "x", "y", "z", "xs", "ys", "zs", "xu", "yu", "zu",
                                      "ix", "iy", "iz", "vx", "vy", "vz", "fx", "fy", "fz"};
        for (size_t k = 6; k < w.size(); ++k) {
            bool ok = false;
            for (const char *q : known) ok |= w[k] == q;
            if (!ok) fail(MDPD_E_UNSUPPORTED, "dump custom column %s is not supported", w[k].c_str());
            d.cols.push_back(w[k]);
        }
        for (auto &e : dumps)
            if (e.id == d.id) fail(MDPD_E_ARG, "Reuse of dump ID %s", d.id.c_str());
        dumps.push_back(d);
    }

    bool dump_due(long long t) const
    {
        for (auto &d : dumps)
            if (t % d.every == 0) return true;
        return false;
    }

    // the state of timestep t to every dump due at t.  `prev`: right after a
    // fused pass with output (x(t) in the buffer the pass read, v(t) and f(t)
    // written by the pass); otherwise the current buffer (setup, end of run).
    void write_dumps(long long t, bool prev)
    {
        const int n = (int)h_tag.size();
        sync_sys();
        const int sel = prev ? hsys.sel ^ 1 : hsys.sel, sela = hsys.sela;
        std::vector<int> org = download(origin[sela].p, cap);
        std::vector<int> im = download(image[sela].p, cap);
        std::vector<int> mk = download(mask[sela].p, cap);
        std::vector<int> tg = download(tag[sela].p, cap);
        std::vector<float4> px = download(pos[sel].p, cap);
        std::vector<float4> pv = download(prev ? vend.p : vel[sel].p, cap);
        std::vector<float4> pf = download(frc.p, cap);
        hip_check(hipStreamSynchronize(stream), "dump");
        std::vector<int> slot(n, -1);
        for (int k = 0; k < n; ++k) slot[org[k]] = k;
        for (auto &d : dumps) {
            if (t % d.every != 0) continue;
            FILE *f = fopen(d.file.c_str(), d.started ? "a" : "w");
            if (!f) fail(MDPD_E_ARG, "Cannot open dump file %s", d.file.c_str());
            d.started = true;
            int cnt = 0;
            for (int k = 0; k < n; ++k) cnt += (mk[k] & d.groupbit) != 0;
            fprintf(f, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", t, cnt);
            for (int a = 0; a < 3; ++a) fprintf(f, "%.16e %.16e\n", lo[a], hi[a]);
            fprintf(f, "ITEM: ATOMS");
            for (auto &c : d.cols) fprintf(f, " %s", c.c_str());
            fprintf(f, "\n");
            for (int i = 0; i < n; ++i) {
                const int k = slot[i];
                if (k < 0 || !(mk[k] & d.groupbit)) continue;
                const float xv[3] = {px[k].x, px[k].y, px[k].z};
                const int iv[3] = {(im[k] & 1023) - 512, ((im[k] >> 10) & 1023) - 512, ((im[k] >> 20) & 1023) - 512};
                for (size_t q = 0; q < d.cols.size(); ++q) {
                    const std::string &c = d.cols[q];
                    const char *sp = q ? " " : "";
                    const int a = c.back() == 'x' ? 0 : (c.back() == 'y' ? 1 : 2);
                    if (c == "id") fprintf(f, "%s%d", sp, tg[k]);
                    else if (c == "type") { int ty; memcpy(&ty, &px[k].w, 4); fprintf(f, "%s%d", sp, ty + 1); }
                    else if (c == "x" || c == "y" || c == "z") fprintf(f, "%s%.9g", sp, xv[a]);
                    else if (c == "xs" || c == "ys" || c == "zs") fprintf(f, "%s%.9g", sp, (xv[c[0] == 'x' ? 0 : (c[0] == 'y' ? 1 : 2)] - lo[c[0] == 'x' ? 0 : (c[0] == 'y' ? 1 : 2)]) / (hi[c[0] == 'x' ? 0 : (c[0] == 'y' ? 1 : 2)] - lo[c[0] == 'x' ? 0 : (c[0] == 'y' ? 1 : 2)]));
                    else if (c == "xu" || c == "yu" || c == "zu") {
                        const int b = c[0] == 'x' ? 0 : (c[0] == 'y' ? 1 : 2);
                        fprintf(f, "%s%.9g", sp, xv[b] + iv[b] * (hi[b] - lo[b]));
                    } else if (c == "ix" || c == "iy" || c == "iz") fprintf(f, "%s%d", sp, iv[c[1] == 'x' ? 0 : (c[1] == 'y' ? 1 : 2)]);
                    else if (c == "vx" || c == "vy" || c == "vz") fprintf(f, "%s%.9g", sp, c[1] == 'x' ? pv[k].x : (c[1] == 'y' ? pv[k].y : pv[k].z));
                    else fprintf(f, "%s%.9g", sp, c[1] == 'x' ? pf[k].x : (c[1] == 'y' ? pf[k].y : pf[k].z));
                }
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }

    // velocity G set vx vy vz (NULL keeps a component); src/velocity.cpp
    void cmd_velocity(const std::vector<std::string> &w)
    {
        if (w.size() < 3) fail(MDPD_E_ARG, "Illegal velocity command");
        const int bit = need_group(w[1]);
        if (w[2] != "set") fail(MDPD_E_UNSUPPORTED, "velocity style %s is not supported", w[2].c_str());
        if (w.size() < 6) fail(MDPD_E_ARG, "Illegal velocity command");
        pull_state();
        for (size_t i = 0; i < h_tag.size(); ++i) {
            if (!(h_mask[i] & bit)) continue;
            for (int d = 0; d < 3; ++d)
                if (w[3 + d] != "NULL") h_v[3 * i + d] = parse_num(w[3 + d]);
        }
        dirty = true;
    }

    void cmd_group(const std::vector<std::string> &w)
    {
        if (w.size() < 3) fail(MDPD_E_ARG, "illegal group command");
        int bit = group_bit(w[1]);
        if (!bit) {
            if (groups.size() >= 32) fail(MDPD_E_ARG, "too many groups");
            bit = 1 << groups.size();
            groups.push_back({w[1], bit});
        }
        const size_t n = h_tag.size();
        if (w[2] == "type" || w[2] == "id") {
            for (size_t k = 3; k < w.size(); ++k) {
                long long a, b;
                size_t col = w[k].find(':');
                if (col == std::string::npos) a = b = parse_int(w[k]);
                else { a = parse_int(w[k].substr(0, col)); b = parse_int(w[k].substr(col + 1)); }
                for (size_t i = 0; i < n; ++i) {
                    long long v = w[2] == "type" ? h_type[i] : h_tag[i];
                    if (v >= a && v <= b) h_mask[i] |= bit;
                }
            }
        } else if (w[2] == "empty") {
        } else if (w[2] == "union") {
            // group ID union G1 G2 ... (src/group.cpp)
            int u = 0;
            for (size_t k = 3; k < w.size(); ++k) u |= need_group(w[k]);
            for (size_t i = 0; i < n; ++i)
                if (h_mask[i] & u) h_mask[i] |= bit;
        } else {
            fail(MDPD_E_UNSUPPORTED, "group style %s is not supported", w[2].c_str());
        }
        dirty = true;
    }

    void cmd_pair_style(const std::vector<std::string> &w)
    {
        if (w.size() < 2) fail(MDPD_E_ARG, "illegal pair_style command");
        if (w[1] == "mdpd/meso") {
            // MesoPairMDPD::settings, pair_mdpd_meso.cu:299-328
            if (w.size() < 5)
                fail(MDPD_E_ARG, "Illegal pair_style command:\n temperature cut_global seed fluid_group wall_group, or temperature cut_global seed ");
            if (w.size() != 5 && w.size() != 7)
                fail(MDPD_E_ARG, "Illegal pair_style command:\n temperature cut_global seed fluid_group wall_group, or temperature cut_global seed ");
            pair_style = STYLE_MDPD;
            temperature = parse_num(w[2]);
            cut_global = parse_num(w[3]);
            seed = (int)parse_int(w[4]);
            pair_mobile_bit = pair_wall_bit = 0;
            if (w.size() == 7) {
                // arbitrary-wall form: fluid and wall groups (pair_mdpd_meso.cu:313-321)
                pair_mobile_bit = group_bit(w[5]);
                pair_wall_bit = group_bit(w[6]);
                if (!pair_mobile_bit) fail(MDPD_E_ARG, "<MESO> Undefined fluid group id in pair_style mdpd/meso");
                if (!pair_wall_bit) fail(MDPD_E_ARG, "<MESO> Undefined wall group id in pair_style mdpd/meso");
            }
        } else if (w[1] == "dpd/meso") {
            // MesoPairDPD::settings, pair_dpd_meso.cu:272-287
            if (w.size() != 4) fail(MDPD_E_ARG, "Illegal pair_style command");
            pair_style = STYLE_DPD;
            cut_global = parse_num(w[2]);
            seed = (int)parse_int(w[3]);
        } else {
            fail(MDPD_E_UNSUPPORTED, "pair_style %s is outside this engine", w[1].c_str());
        }
        for (int i = 1; i <= MAXTYPES; ++i)
            for (int j = i + 1; j <= MAXTYPES; ++j)
                if (setflag[i][j]) cut[i][j] = cut_global;
        dirty = true;
    }

    void cmd_pair_coeff(const std::vector<std::string> &w)
    {
        if (pair_style < 0) fail(MDPD_E_STATE, "pair_coeff command before pair_style is defined");
        if (!box_set) fail(MDPD_E_STATE, "pair_coeff command before simulation box is defined");
        std::vector<std::string> a(w.begin() + 1, w.end());
        int ilo, ihi, jlo, jhi;
        int count = 0;
        if (pair_style == STYLE_MDPD) {
            // MesoPairMDPD::coeff, pair_mdpd_meso.cu:330-365
            if (a.size() != 7) fail(MDPD_E_ARG, "Incorrect args for mdpd pair coefficients: \n itype jtype A B gamma cutA cutB");
            type_bounds(a[0], ntypes, ilo, ihi);
            type_bounds(a[1], ntypes, jlo, jhi);
            double a0 = parse_num(a[2]), b0 = parse_num(a[3]), g = parse_num(a[4]);
            double c1 = parse_num(a[5]), c2 = parse_num(a[6]);
            if (c1 < c2) fail(MDPD_E_ARG, "Incorrect args for pair coefficients:\n cutA should be larger than cutB.");
            for (int i = ilo; i <= ihi; ++i)
                for (int j = std::max(jlo, i); j <= jhi; ++j) {
                    A[i][j] = a0; B[i][j] = b0; gamma[i][j] = g; cut[i][j] = c1; cut_r[i][j] = c2;
                    setflag[i][j] = 1;
                    ++count;
                }
        } else {
            // MesoPairDPD::coeff, pair_dpd_meso.cu:289-323
            if (a.size() < 6 || a.size() > 7) fail(MDPD_E_ARG, "Incorrect args for pair coefficients");
            type_bounds(a[0], ntypes, ilo, ihi);
            type_bounds(a[1], ntypes, jlo, jhi);
            double a0 = parse_num(a[2]), g = parse_num(a[3]), s = parse_num(a[4]), e = parse_num(a[5]);
            double c1 = a.size() == 7 ? parse_num(a[6]) : cut_global;
            for (int i = ilo; i <= ihi; ++i)
                for (int j = std::max(jlo, i); j <= jhi; ++j) {
                    A[i][j] = a0; gamma[i][j] = g; sigma[i][j] = s; expw[i][j] = e; cut[i][j] = c1;
                    setflag[i][j] = 1;
                    ++count;
                }
        }
        if (count == 0) fail(MDPD_E_ARG, "Incorrect args for pair coefficients");
        dirty = true;
    }

    void cmd_fix(const std::vector<std::string> &w)
    {
        if (w.size() < 4) fail(MDPD_E_ARG, "illegal fix command");
        FixDesc f;
        f.id = w[1];
        f.groupbit = need_group(w[2]);
        f.style = w[3];
        if (f.style == "nve/meso") {
        } else if (f.style == "aboundary/rho/meso") {
            // FixArbitraryBoundaryRho ctor, fix_aboundary_rho_meso.cu:46-104
            for (size_t k = 4; k < w.size(); ++k) {
                if (w[k] == "cut_rho" && k + 1 < w.size()) f.rd = parse_num(w[++k]);
                else if (w[k] == "weight" && k + 1 < w.size()) {
                    const std::string &m = w[++k];
                    if (m == "lucy") f.lucy = 1;
                    else if (m == "quadratic") f.lucy = 0;
                    else fail(MDPD_E_ARG, "fix aboundary/rho/meso weight must be quadratic or lucy");
                } else if (w[k] == "cut_phi" && k + 1 < w.size()) {
                    f.cut_phi = parse_num(w[++k]);
                } else if (w[k] == "phi_c" && k + 1 < w.size()) {
                    f.phi_c = parse_num(w[++k]);
                } else if (w[k] == "rho_wall" && k + 1 < w.size()) {
                    f.rho_wall = parse_num(w[++k]);
                } else if (w[k] == "wall" && k + 1 < w.size()) {
                    f.wall_bit = group_bit(w[++k]);
                    if (!f.wall_bit) fail(MDPD_E_ARG, "<MESO> Undefined wall group id in fix aboundary/rho/meso");
                    f.arb = 1;
                } else if (w[k] == "mobile" && k + 1 < w.size()) {
                    f.mobile_bit = group_bit(w[++k]);
                    if (!f.mobile_bit) fail(MDPD_E_ARG, "<MESO> Undefined fluid group id in fix aboundary/rho/meso");
                    f.arb = 1;
                }
            }
            // fix_aboundary_rho_meso.cu:96-101
            if (!f.arb && f.rd == 0) fail(MDPD_E_ARG, "<MESO> Fix aboundary/rho/meso parameters are not set properly.");
            if (f.arb && (f.rd == 0 || f.cut_phi == 0 || f.phi_c == 0 || f.rho_wall == 0))
                fail(MDPD_E_ARG, "<MESO> Fix aboundary/rho/meso parameters are not set properly.");
        } else if (f.style == "rdf/fast/meso") {
            // MesoFixRDFFast ctor: output FILE nbin N [every M] [other G];
            // [pi 3.1415|exact] chooses the shell-volume constant (an option
            // of this library: the reference always uses 3.1415)
            for (size_t k = 4; k < w.size(); ++k) {
                if (w[k] == "output" && k + 1 < w.size()) f.output = w[++k];
                else if (w[k] == "nbin" && k + 1 < w.size()) f.nbin = (int)parse_int(w[++k]);
                else if (w[k] == "every" && k + 1 < w.size()) f.every = (int)parse_int(w[++k]);
                else if (w[k] == "other" && k + 1 < w.size()) {
                    f.other_bit = group_bit(w[++k]);
                    if (!f.other_bit) fail(MDPD_E_ARG, "<MESO> Undefined other group id in fix rdf/meso");
                } else if (w[k] == "pi" && k + 1 < w.size()) {
                    const std::string &v = w[++k];
                    if (v != "3.1415" && v != "exact") fail(MDPD_E_ARG, "fix rdf/fast/meso: pi must be 3.1415 or exact");
                    f.exact_pi = v == "exact";
                }
            }
            if (f.output.empty() || f.nbin <= 0) fail(MDPD_E_ARG, "Incomplete compute rdf command: insufficient arguments");
            if (f.nbin > 4096 || f.every < 1) fail(MDPD_E_ARG, "fix rdf/fast/meso: nbin must be 1..4096 and every >= 1");
            if (!f.other_bit) f.other_bit = f.groupbit;
        } else if (f.style == "density/meso") {
            // MesoFixDensityProf ctor: output FILE along x|y|z nbin N every M window W
            for (size_t k = 4; k < w.size(); ++k) {
                if (w[k] == "output" && k + 1 < w.size()) f.output = w[++k];
                else if (w[k] == "along" && k + 1 < w.size()) {
                    const char a = w[++k][0];
                    f.along = a == 'x' ? 0 : (a == 'y' ? 1 : (a == 'z' ? 2 : -1));
                } else if (w[k] == "nbin" && k + 1 < w.size()) f.nbin = (int)parse_int(w[++k]);
                else if (w[k] == "every" && k + 1 < w.size()) f.every = (int)parse_int(w[++k]);
                else if (w[k] == "window" && k + 1 < w.size()) f.window = (int)parse_int(w[++k]);
            }
            if (f.output.empty() || f.along < 0 || f.nbin <= 0 || f.every <= 0 || f.window <= 0)
                fail(MDPD_E_ARG, "Incomplete fix density/meso command: insufficient arguments");
            if (f.nbin > 4096) fail(MDPD_E_ARG, "fix density/meso: nbin must be 1..4096");
        } else if (f.style == "addforce/meso") {
            // MesoFixAddForce ctor, fix_addforce_meso.cu:38-47
            if (w.size() < 7) fail(MDPD_E_ARG, "Illegal fix addforce/meso command");
            for (int d = 0; d < 3; ++d) f.f[d] = parse_num(w[4 + d]);
        } else {
            fail(MDPD_E_UNSUPPORTED, "fix style %s is outside this engine", f.style.c_str());
        }
        auto it = std::find_if(fixes.begin(), fixes.end(), [&](const FixDesc &g) { return g.id == f.id; });
        if (it != fixes.end()) {
            finish_obs(*it);
            *it = f;
        } else {
            fixes.push_back(f);
        }
        if (is_obs(f)) {
            obs.erase(f.id);
            ObsState &o = obs[f.id];
            (void)o;
            graph_valid = false;         // the captured sampling steps change
            return;                      // observables do not change the step: no new setup
        }
        dirty = true;
    }

    // ------------------------------------------------------------------
    // observables: rdf/fast/meso, density/meso
    // ------------------------------------------------------------------
    static bool is_obs(const FixDesc &f) { return f.style == "rdf/fast/meso" || f.style == "density/meso"; }

    ObsState &obs_state(const FixDesc &f)
    {
        ObsState &o = obs[f.id];
        if (!o.hist.p) {
            ensure_stream();
            hip_check(hipSetDevice(device), "hipSetDevice");
            o.hist.alloc(f.nbin + 1);
            hip_check(hipMemsetAsync(o.hist.p, 0, sizeof(unsigned) * (f.nbin + 1), stream), "memset");
        }
        return o;
    }

    // the nbin counts and the sample count
    std::vector<unsigned> obs_counts(ObsState &o, int nbin)
    {
        std::vector<unsigned> h(nbin + 1, 0u);
        if (o.hist.p) {
            hip_check(hipMemcpyAsync(h.data(), o.hist.p, sizeof(unsigned) * (nbin + 1), hipMemcpyDeviceToHost, stream),
                      "download");
            hip_check(hipStreamSynchronize(stream), "sync");
        }
        return h;
    }

    // g(r) of fix rdf/fast/meso from its pair counts h[0, nbin) over `samples`
    // samples, the group sizes counted over the n masks, box volume V and
    // range rc (fix_rdf_fast_meso.cu:187-217)
    static void rdf_values(const FixDesc &f, const unsigned *h, long long samples, const int *mask, size_t n,
                           double V, double rc, std::vector<double> &x, std::vector<double> &val)
    {
        const double bin = rc / f.nbin;
        long long ni = 0, nj = 0;
        for (size_t i = 0; i < n; ++i) {
            ni += (mask[i] & f.groupbit) != 0;
            nj += (mask[i] & f.other_bit) != 0;
        }
        const double rho_j = nj / V;
        x.assign(f.nbin, 0.0);
        val.assign(f.nbin, 0.0);
        for (int b = 0; b < f.nbin; ++b) {
            x[b] = (b + 0.5) * bin;
            if (!samples || !ni || !nj) continue;
            const double freq = h[b] / double(ni) / double(samples);
            const double pi = f.exact_pi ? 3.14159265358979323846 : 3.1415;
            const double shell = 4. / 3. * pi * (std::pow(bin * (b + 1), 3.) - std::pow(bin * b, 3.));
            val[b] = freq / shell / rho_j;
        }
    }

    // current result of an observable fix: bin centres and values (rdf: g(r)
    // with the reference's normalization, 4/3 * 3.1415 * shell volume,
    // fix_rdf_fast_meso.cu:187-217; density: counts per bin averaged over
    // the frames of the current window, fix_density_profile_meso.cpp:192-215)
    void obs_result(const FixDesc &f, std::vector<double> &x, std::vector<double> &val)
    {
        ObsState &o = obs[f.id];
        std::vector<unsigned> h = obs_counts(o, f.nbin);
        const long long samples = h[f.nbin];
        x.assign(f.nbin, 0.0);
        val.assign(f.nbin, 0.0);
        if (f.style == "rdf/fast/meso") {
            const double V = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
            rdf_values(f, h.data(), samples, h_mask.data(), h_mask.size(), V, max_cut(), x, val);
        } else {
            const double bin = (hi[f.along] - lo[f.along]) / f.nbin;
            for (int b = 0; b < f.nbin; ++b) {
                x[b] = (b + 0.5) * bin;
                val[b] = samples ? h[b] / double((float)samples) : 0.0;
            }
        }
    }

    void write_obs(const FixDesc &f, const std::string &path)
    {
        std::vector<double> x, val;
        obs_result(f, x, val);
        FILE *fp = fopen(path.c_str(), "w");
        if (!fp) fail(MDPD_E_ARG, "Cannot open %s", path.c_str());
        for (int b = 0; b < f.nbin; ++b) fprintf(fp, "%.15g\t%.15g\t\n", x[b], val[b]);
        fclose(fp);
    }

    // Observables per step t of a run: density/meso samples x(t) before the
    // force pass (post_integrate), rdf/fast/meso after it (post_force); a
    // density window is written on the step after it.  Bit k of a sample
    // mask stands for the k-th observable fix (in definition order).
    static bool density_sample(const FixDesc &f, long long t)
    {
        const long long m = (t + f.every / 2) / f.every * f.every;
        return t - m >= -f.window && t - m < f.window;
    }
    static bool density_dump(const FixDesc &f, long long t)
    {
        const long long m = (t + f.every / 2) / f.every * f.every;
        return t - m == f.window;
    }
    static bool obs_samples(const FixDesc &f, long long t)
    {
        return f.style == "rdf/fast/meso" ? t % f.every == 0 : density_sample(f, t);
    }
    uint32_t obs_sample_mask(long long t) const     // the first 32 fixes
    {
        uint32_t mask = 0;
        int k = 0;
        for (const auto &f : fixes) {
            if (!is_obs(f)) continue;
            if (k < 32 && obs_samples(f, t)) mask |= 1u << k;
            ++k;
        }
        return mask;
    }
    // a step on which a density window is written (host work: the direct path)
    bool obs_dump_due(long long t) const
    {
        for (const auto &f : fixes)
            if (f.style == "density/meso" && density_dump(f, t)) return true;
        return false;
    }
    // sampling launches of the fixes that sample step t (t < 0: the fixes
    // in `mask`), the density ones (pre) or the rdf ones (post); the kernels
    // count their samples on the device
    void obs_launch(long long t, uint32_t mask, bool pre)
    {
        int k = 0;
        for (auto &f : fixes) {
            if (!is_obs(f)) continue;
            const bool on = t >= 0 ? obs_samples(f, t) : k < 32 && ((mask >> k) & 1u);
            ++k;
            if (!on || (f.style == "density/meso") != pre) continue;
            ObsState &o = obs_state(f);
            PoolDev P = view();
            if (pre)
                hip_check(launch_profile(P, o.hist.p, f.nbin, f.along, f.groupbit, stream), "density profile");
            else
                hip_check(launch_rdf(P, o.hist.p, f.nbin, (float)max_cut(), f.groupbit, f.other_bit, stream), "rdf");
        }
    }
    void obs_pre(long long t)
    {
        for (auto &f : fixes)
            if (f.style == "density/meso" && density_dump(f, t)) {
                const long long m = (t + f.every / 2) / f.every * f.every;
                dump_profile(f, obs_state(f), m, t);
            }
        obs_launch(t, 0, true);
    }
    void obs_post(long long t) { obs_launch(t, 0, false); }

    void dump_profile(const FixDesc &f, ObsState &o, long long m, long long t)
    {
        char fn[512];
        snprintf(fn, sizeof(fn), "%s.%09lld", f.output.c_str(), m);
        write_obs(f, fn);
        o.last_dump = t;
        hip_check(hipMemsetAsync(o.hist.p, 0, sizeof(unsigned) * (f.nbin + 1), stream), "memset");
    }

    // end of an observable fix (unfix, redefinition, handle closed): the
    // reference's destructors write the accumulated result
    void finish_obs(const FixDesc &f)
    {
        if (!is_obs(f)) return;
        auto it = obs.find(f.id);
        if (it == obs.end() || !it->second.hist.p) return;
        ObsState &o = it->second;
        if (f.style == "rdf/fast/meso") {
            write_obs(f, f.output);
        } else {
            const long long t = ntimestep, m = (t + f.every / 2) / f.every * f.every;
            if (density_sample(f, t) && o.last_dump < t && obs_counts(o, f.nbin)[f.nbin]) dump_profile(f, o, m, t);
        }
        obs.erase(it);
        graph_valid = false;
    }

    // ------------------------------------------------------------------
    // setup: device tables and the first build + force pass
    // ------------------------------------------------------------------
    void init_coeffs(std::vector<float> &tab)
    {
        // init_one: pair_mdpd_meso.cu:382-397 (sigma = sqrt(2 kB T gamma)),
        // pair_dpd_meso.cu:343-357 (sigma given)
        for (int i = 1; i <= ntypes; ++i)
            for (int j = i; j <= ntypes; ++j) {
                if (!setflag[i][j]) fail(MDPD_E_STATE, "All pair coeffs are not set");
                if (pair_style == STYLE_MDPD) sigma[i][j] = std::sqrt(2.0 * temperature * gamma[i][j]);
                cut[j][i] = cut[i][j]; cut_r[j][i] = cut_r[i][j]; A[j][i] = A[i][j]; B[j][i] = B[i][j];
                gamma[j][i] = gamma[i][j]; sigma[j][i] = sigma[i][j]; expw[j][i] = expw[i][j];
            }
        tab.assign((size_t)ntypes * ntypes * NCOEF, 0.f);
        for (int i = 1; i <= ntypes; ++i)
            for (int j = 1; j <= ntypes; ++j) {
                float *c = &tab[((size_t)(i - 1) * ntypes + (j - 1)) * NCOEF];
                c[C_RC] = (float)cut[i][j];
                c[C_RCINV] = (float)(1.0 / cut[i][j]);
                c[C_RD] = (float)cut_r[i][j];
                c[C_RDINV] = cut_r[i][j] > 0 ? (float)(1.0 / cut_r[i][j]) : 0.f;
                c[C_A] = (float)A[i][j];
                c[C_B] = (float)B[i][j];
                c[C_GAMMA] = (float)gamma[i][j];
                c[C_SIGMA] = (float)sigma[i][j];
                c[C_EXPW] = (float)expw[i][j];
                c[C_RC2] = (float)(cut[i][j] * cut[i][j]);
            }
    }

    double max_cut() const
    {
        double m = 0;
        for (int i = 1; i <= ntypes; ++i)
            for (int j = 1; j <= ntypes; ++j) m = std::max(m, cut[i][j]);
        return m;
    }

    void ensure_stream()
    {
        if (stream) return;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev)
            fail(MDPD_E_HIP, "HIP device %d is not available", device);
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipHostMalloc((void **)&pinned, 64 * sizeof(int), hipHostMallocDefault), "hipHostMalloc");
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
        P.pcell = pcell.p; P.pslot = pslot.p; P.psub = psub.p; P.tmp = tmp.p; P.key = key.p;
        P.cell_count = cell_count.p; P.cell_start = cell_start.p; P.cell_xyz = cell_xyz.p;
        P.cell_rank = cell_rank.p; P.cell_sys = cell_sys.p;
        P.work_count = work_count.p; P.cell_work = cell_work.p; P.cell_lse = cell_lse.p;
        P.nlist = nlist.p; P.ncount = ncount.p; P.wave_sys = wave_sys.p;
        P.clist = clist.p; P.ccount = ccount.p; P.moved = moved.p; P.status = status.p;
        P.cl = cl.p; P.wall_near = wall_near.p;
        P.sys = dsys.p; P.coef = coef.p; P.minv = minv.p; P.mass = mmass.p;
        P.cap = cap; P.nact = cap; P.ncellslots = nslots; P.ncact = nslots; P.nsys = 1; P.maxsub = 0;
        P.maxnbr = maxnbr;
        return P;
    }

    void validate()
    {
        if (!box_set) fail(MDPD_E_STATE, "run command before simulation box is defined");
        if (atom_style < 0) fail(MDPD_E_STATE, "atom_style mdpd/atomic/meso or dpd/atomic/meso must be set");
        if (pair_style < 0) fail(MDPD_E_STATE, "pair style must be set");
        if (pair_style == STYLE_MDPD && atom_style != STYLE_MDPD)
            fail(MDPD_E_STATE, "pair_style mdpd/meso requires atom_style mdpd/atomic/meso");
        if (h_tag.empty()) fail(MDPD_E_STATE, "no atoms");
        for (int t = 1; t <= ntypes; ++t)
            if (!mass_set[t]) fail(MDPD_E_STATE, "Not all per-type masses are set");
        if ((long long)h_tag.size() >= (1LL << IMG_SHIFT)) fail(MDPD_E_OVERFLOW, "too many atoms for one handle");
    }

    // Device-free description of this system for a launch: the per-system
    // descriptor (bases left at 0), the FP32 coefficient table, per-type
    // masses, the Morton rank of each interior cell and the packed coordinates
    // of the cell at each rank, and the neighbor row width it asks for.
    struct Desc {
        SysDev d{};
        std::vector<float> coef, mass, minv;
        std::vector<int> rank, xyz;
        int want_maxnbr = 0;
    };

    void describe(Desc &out)
    {
        validate();
        init_coeffs(out.coef);
        const int n = (int)h_tag.size();
        const double rcut = max_cut();
        const double cutneigh = rcut + skin;

        // ---- bins (neighbor_meso.cu:837-853) ----
        SysDev &d = out.d;
        d = SysDev{};
        for (int a = 0; a < 3; ++a) {
            const double L = hi[a] - lo[a];
            int nb = (int)(L * (1.0 / cutneigh));
            if (nb < 1) nb = 1;
            if (L < cutneigh) fail(MDPD_E_ARG, "Cannot use neighbor bins - box size << cutoff");
            if (nb > 1000) fail(MDPD_E_OVERFLOW, "more than 1000 cells along one axis");
            d.lo[a] = (float)lo[a];
            d.hi[a] = (float)hi[a];
            d.L[a] = (float)L;
            d.lo_d[a] = lo[a];
            d.binsize_d[a] = L / nb;
            d.bininv_d[a] = 1.0 / d.binsize_d[a];
            d.mbin[a] = nb + 2;
            d.nb[a] = nb;
        }
        d.ncell = d.nb[0] * d.nb[1] * d.nb[2];
        d.l1_shift = 12;
        d.n = n;
        d.npad = (n + 255) / 256 * 256;
        d.rc2_list = (float)(cutneigh * cutneigh);
        d.rc2_max = (float)(rcut * rcut);
        d.trig2 = (float)(0.25 * skin * skin);
        d.dt = (float)dt;
        d.dtf = (float)(0.5 * dt);
        d.dtinvsqrt = (float)(1.0 / std::sqrt(dt));
        d.seed = (uint32_t)seed;
        d.style = pair_style;
        d.ntypes = ntypes;
        const FixDesc *fn = find_fix("nve/meso"), *fr = find_fix("aboundary/rho/meso"), *fa = find_fix("addforce/meso");
        d.nve_bit = fn ? fn->groupbit : 0;
        d.rho_bit = fr ? fr->groupbit : 0;
        d.add_bit = fa ? fa->groupbit : 0;
        for (int a = 0; a < 3; ++a) d.add[a] = fa ? (float)fa->f[a] : 0.f;
        if (fr) {
            d.rd = (float)fr->rd;
            d.lucy = fr->lucy;
            const double r = fr->rd;
            // fix_aboundary_rho_meso.cu:88-90
            d.rho_factor = fr->lucy ? (float)(105.0 / (16.0 * M_PI * r * r * r * r * r * r * r))
                                    : (float)(15.0 / (2.0 * M_PI * r * r * r * r * r));
        }
        if (pair_style == STYLE_MDPD && pair_mobile_bit && pair_wall_bit) {
            d.pair_arb = 1;
            d.pair_mobile_bit = pair_mobile_bit;
            d.pair_wall_bit = pair_wall_bit;
        }
        if (fr && fr->arb) {
            // fix_aboundary_rho_meso.cu:92-94
            const double c = fr->cut_phi, c7 = c * c * c * c * c * c * c;
            d.fix_arb = 1;
            d.fix_mobile_bit = fr->mobile_bit;
            d.fix_wall_bit = fr->wall_bit;
            d.cut_phi = (float)c;
            d.phi_c = (float)fr->phi_c;
            d.phi_factor = (float)(105.0 / (16.0 * M_PI * fr->rho_wall * c7));
            d.dw_factor = (float)(-315.0 / (4.0 * M_PI * fr->rho_wall * c7));
        }
        d.stale = 1;
        d.active = 1;
        d.phase = PH_RUN;      // signatures come from launch_signatures, not the re-sort
        d.step = ntimestep;
        d.seed_cur = premix_host((uint32_t)seed, (uint32_t)ntimestep, 64);
        d.seed_next = premix_host((uint32_t)seed, (uint32_t)(ntimestep + 1), 64);
        d.n_every = every;
        d.n_delay = delay;
        d.n_check = check;

        // ---- neighbor row width: 2x the expected count at rho_est ----
        const double dens = rho_est > 0 ? rho_est : n / ((hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]));
        const double expect = dens * 4.0 / 3.0 * M_PI * cutneigh * cutneigh * cutneigh;
        int want = std::max(32, (int)std::ceil(2.0 * expect));
        out.want_maxnbr = (want + NBR_ALIGN - 1) / NBR_ALIGN * NBR_ALIGN;

        // ---- Morton rank of interior cells (order of the sort key's bin part) ----
        const int nc = d.ncell;
        std::vector<std::pair<uint64_t, int>> order(nc);
        for (int z = 0; z < d.nb[2]; ++z)
            for (int y = 0; y < d.nb[1]; ++y)
                for (int x = 0; x < d.nb[0]; ++x) {
                    int lin = x + d.nb[0] * (y + d.nb[1] * z);
                    order[lin] = {morton3(x + 1, y + 1, z + 1), lin};
                }
        std::sort(order.begin(), order.end());
        out.rank.assign(nc + 1, 0);
        out.xyz.assign(nc + 1, 0);
        for (int r = 0; r < nc; ++r) {
            int lin = order[r].second;
            out.rank[lin] = r;
            int x = lin % d.nb[0], y = (lin / d.nb[0]) % d.nb[1], z = lin / (d.nb[0] * d.nb[1]);
            out.xyz[r] = x | (y << 10) | (z << 20);
        }
        out.mass.resize(ntypes);
        out.minv.resize(ntypes);
        for (int t = 0; t < ntypes; ++t) { out.mass[t] = (float)mass[t + 1]; out.minv[t] = (float)(1.0 / mass[t + 1]); }
    }

    // Particle arrays in the device layout, creation order, `count` slots
    // (slots past n are inert padding).  origin = creation index.
    void pack(int count, float4 *hp, float4 *hv, int *ht, int *hm, int *ho, int *himg) const
    {
        const int n = (int)h_tag.size();
        for (int i = 0; i < count; ++i) {
            if (i < n) {
                int ty = h_type[i] - 1;
                float tyf;
                memcpy(&tyf, &ty, 4);
                hp[i] = make_float4((float)h_x[3 * i], (float)h_x[3 * i + 1], (float)h_x[3 * i + 2], tyf);
                hv[i] = make_float4((float)h_v[3 * i], (float)h_v[3 * i + 1], (float)h_v[3 * i + 2], 0.f);
                ht[i] = h_tag[i];
                hm[i] = h_mask[i];
                ho[i] = i;
                himg[i] = ((h_image[3 * i] + 512) & 1023) | (((h_image[3 * i + 1] + 512) & 1023) << 10) |
                          (((h_image[3 * i + 2] + 512) & 1023) << 20);
            } else {
                hp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                hv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                ht[i] = 0; hm[i] = 0; ho[i] = -1;
                himg[i] = 512 | (512 << 10) | (512 << 20);
            }
        }
    }

    void setup()
    {
        pull_state();                 // a re-setup continues from where the last run ended
        Desc ds;
        describe(ds);
        ensure_stream();
        SysDev &d = ds.d;
        maxnbr = std::max(maxnbr, ds.want_maxnbr);
        d.maxnbr = maxnbr;
        const int nc = d.ncell;
        std::vector<int> csys(nc + 1, 0);

        // ---- allocate ----
        hip_check(hipSetDevice(device), "hipSetDevice");
        cap = d.npad;
        ncell = nc;
        nslots = nc + 1;
        for (int b = 0; b < 2; ++b) {
            pos[b].alloc(cap); vel[b].alloc(cap);
            tag[b].alloc(cap); mask[b].alloc(cap); origin[b].alloc(cap); image[b].alloc(cap);
        }
        xhold.alloc(cap); frc.alloc(cap); vend.alloc(cap);
        rho.alloc(cap); pr.alloc(cap); vir.alloc((size_t)6 * cap); phi[0].alloc(cap); phi[1].alloc(cap);
        pcell.alloc(cap); pslot.alloc(cap); psub.alloc(cap); tmp.alloc(cap); key.alloc(2 * (size_t)cap);
        ncount.alloc(cap); ccount.alloc((size_t)MAXSPLIT * cap);
        nlist.alloc((size_t)cap * maxnbr);
        clist.alloc((size_t)cap * maxnbr);
        wave_sys.alloc(cap / WAVE); moved.alloc(cap / WAVE); status.alloc(1);
        cl.alloc(cap); wall_near.alloc(cap);
        cell_count.alloc((size_t)nslots * SUBCELLS); cell_start.alloc((size_t)nslots * SUBCELLS);
        cell_xyz.alloc(nslots);
        cell_rank.alloc(nslots); cell_sys.alloc(nslots);
        cell_work.alloc(nslots); cell_lse.alloc(nslots); work_count.alloc(64);
        coef.alloc(ds.coef.size()); minv.alloc(ntypes); mmass.alloc(ntypes); dsys.alloc(1);

        // ---- upload ----
        std::vector<float4> hp(cap), hv(cap);
        std::vector<int> ht(cap), hm(cap), ho(cap), himg(cap);
        pack(cap, hp.data(), hv.data(), ht.data(), hm.data(), ho.data(), himg.data());
        std::vector<int> ws(cap / WAVE, 0);
        auto up = [&](void *dst, const void *src, size_t bytes) {
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "upload");
        };
        up(pos[0].p, hp.data(), cap * sizeof(float4));
        up(vel[0].p, hv.data(), cap * sizeof(float4));
        up(tag[0].p, ht.data(), cap * sizeof(int));
        up(mask[0].p, hm.data(), cap * sizeof(int));
        up(origin[0].p, ho.data(), cap * sizeof(int));
        up(image[0].p, himg.data(), cap * sizeof(int));
        up(wave_sys.p, ws.data(), ws.size() * sizeof(int));
        up(cell_rank.p, ds.rank.data(), nslots * sizeof(int));
        up(cell_xyz.p, ds.xyz.data(), nslots * sizeof(int));
        up(cell_sys.p, csys.data(), nslots * sizeof(int));
        up(coef.p, ds.coef.data(), ds.coef.size() * sizeof(float));
        up(minv.p, ds.minv.data(), ntypes * sizeof(float));
        up(mmass.p, ds.mass.data(), ntypes * sizeof(float));
        hip_check(hipMemsetAsync(cell_count.p, 0, (size_t)nslots * SUBCELLS * sizeof(int), stream), "memset");
        hip_check(hipMemsetAsync(work_count.p, 0, sizeof(int), stream), "memset");
        hip_check(hipMemsetAsync(ncount.p, 0, cap * sizeof(int), stream), "memset");
        hip_check(hipMemsetAsync(rho.p, 0, cap * sizeof(float), stream), "memset");
        std::vector<float> hph(cap, 0.f);   // per-atom state of fix aboundary/rho/meso
        for (size_t i = 0; i < h_phi.size() && i < h_tag.size(); ++i) hph[i] = (float)h_phi[i];
        up(phi[0].p, hph.data(), cap * sizeof(float));
        hip_check(hipMemsetAsync(phi[1].p, 0, cap * sizeof(float), stream), "memset");
        hip_check(hipMemsetAsync(frc.p, 0, cap * sizeof(float4), stream), "memset");
        hip_check(hipMemsetAsync(vend.p, 0, cap * sizeof(float4), stream), "memset");
        hip_check(hipMemsetAsync(vir.p, 0, 6 * cap * sizeof(float), stream), "memset");
        d.builds = (int)nbuild;       // the device counts builds from here on
        hsys = d;
        up(dsys.p, &hsys, sizeof(SysDev));
        graph_valid = false;

        rebuild();
        PoolDev P = view();
        hip_check(launch_signatures(P, stream), "signatures");
        force_pass(MODE_SETUP);
        hip_check(hipStreamSynchronize(stream), "setup");
        collect_events();
        nbuild = read_field(offsetof(SysDev, builds));
        setup_done = true;
        dirty = false;
    }

    // rebuild phase for a stale system; grows the neighbor rows on overflow
    void rebuild()
    {
        PoolDev P = view();
        timed(PH_REBUILD, [&]() { hip_check(launch_rebuild(P, stream), "rebuild"); });
        for (int attempt = 0;; ++attempt) {
            int ovf = read_field(offsetof(SysDev, overflow));
            if (!ovf) break;
            if (attempt > 4) fail(MDPD_E_OVERFLOW, "<MESO> Pair table overflow; local density too high");
            // grow rows and rebuild the rows only (sorted data is in place)
            maxnbr = (int)std::ceil(maxnbr * 1.5 / NBR_ALIGN) * NBR_ALIGN;
            nlist.alloc((size_t)cap * maxnbr);
            clist.alloc((size_t)cap * maxnbr);
            int zero = 0;
            hip_check(hipMemcpyAsync((char *)dsys.p + offsetof(SysDev, overflow), &zero, sizeof(int),
                                     hipMemcpyHostToDevice, stream), "clear overflow");
            hip_check(hipMemcpyAsync((char *)dsys.p + offsetof(SysDev, maxnbr), &maxnbr, sizeof(int),
                                     hipMemcpyHostToDevice, stream), "maxnbr");
            P = view();
            hip_check(launch_nlist_only(P, stream), "nlist");
        }
        hip_check(launch_commit_rebuild(P, stream), "commit");
        graph_valid = false;
    }

    // Rows of a step's rebuild overflowed on the device (that step's passes
    // and commit skipped the system): widen the rows and re-list the sorted
    // system in place from the positions of its last build, which clears
    // the flag; the skipped step then runs again.  Stream drained on entry.
    void regrow_rows()
    {
        for (int attempt = 0;; ++attempt) {
            if (attempt > 4) fail(MDPD_E_OVERFLOW, "<MESO> Pair table overflow; local density too high");
            maxnbr = (int)std::ceil(maxnbr * 1.5 / NBR_ALIGN) * NBR_ALIGN;
            nlist.alloc((size_t)cap * maxnbr);
            clist.alloc((size_t)cap * maxnbr);
            hsys.maxnbr = maxnbr;
            hip_check(hipMemcpyAsync((char *)dsys.p + offsetof(SysDev, maxnbr), &maxnbr, sizeof(int),
                                     hipMemcpyHostToDevice, stream), "maxnbr");
            graph_valid = false;
            PoolDev P = view();
            hip_check(launch_relist(P, stream), "relist");
            if (!read_field(offsetof(SysDev, overflow))) return;
        }
    }

    hipEvent_t take_event()
    {
        if (!ev_free.empty()) { hipEvent_t e = ev_free.back(); ev_free.pop_back(); return e; }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    template <typename F>
    void timed(int phase, F launch)
    {
        if (!prof) { launch(); return; }
        hipEvent_t a = take_event(), b = take_event();
        hip_check(hipEventRecord(a, stream), "hipEventRecord");
        launch();
        hip_check(hipEventRecord(b, stream), "hipEventRecord");
        ev_open.push_back({phase, {a, b}});
    }
    // fold finished event pairs into the per-phase totals (stream synchronized)
    void collect_events()
    {
        for (auto &e : ev_open) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e.second.first, e.second.second), "hipEventElapsedTime");
            prof_ms[e.first] += ms;
            prof_n[e.first] += 1;
            ev_free.push_back(e.second.first);
            ev_free.push_back(e.second.second);
        }
        ev_open.clear();
    }

    int read_field(size_t offset)
    {
        hip_check(hipMemcpyAsync(pinned, (char *)dsys.p + offset, sizeof(int), hipMemcpyDeviceToHost, stream),
                  "readback");
        hip_check(hipStreamSynchronize(stream), "readback sync");
        return pinned[0];
    }

    // t: timestep this pass evaluates (observables sample it), -1 at setup
    // and in a captured step, which takes the samples of `mask` instead
    void force_pass(int mode, long long t = -1, uint32_t mask = 0)
    {
        PoolDev P = view();
        const bool with_rho = pair_style == STYLE_MDPD && hsys.rho_bit;
        if (t >= 0 && !obs.empty()) obs_pre(t);
        else if (mask) obs_launch(-1, mask, true);
        timed(PH_DENSITY, [&]() { hip_check(launch_first(P, with_rho, split, stream), "first pass"); });
        timed(PH_FORCE, [&]() { hip_check(launch_force(P, pair_style, mode, ntypes == 1, split, stream), "force"); });
        if (t >= 0 && !obs.empty()) obs_post(t);
        else if (mask) obs_launch(-1, mask, false);
        if ((mode == MODE_FUSED || mode == MODE_FUSED_OUT) && hsys.fix_arb)
            timed(PH_WALL, [&]() { hip_check(launch_aboundary(P, 0, stream), "aboundary"); });   // post_integrate
        if (mode != MODE_SETUP)
            hip_check(launch_commit_step(P, mode == MODE_FINAL ? COMMIT_FINAL : COMMIT_DECIDE, split, stream), "commit");
    }

    // The rebuild phase of a step, launched every step: its kernels re-sort
    // and re-list the system only when the previous commit flagged it stale
    // (Neighbor::decide made on the device), otherwise they exit at once.
    void rebuild_phase()
    {
        PoolDev P = view();
        timed(PH_REBUILD, [&]() { hip_check(launch_rebuild(P, stream), "rebuild"); });
        hip_check(launch_commit_rebuild(P, stream), "commit rebuild");
    }

    // the graph of one fused step whose observables take the samples of
    // `mask` (0: none); all are captured again after a change
    hipGraphExec_t step_graph(uint32_t mask)
    {
        if (!graph_valid) {
            if (gexec) hip_check(hipGraphExecDestroy(gexec), "hipGraphExecDestroy");
            gexec = nullptr;
            for (auto &v : gobs) hip_check(hipGraphExecDestroy(v.second), "hipGraphExecDestroy");
            gobs.clear();
            gexec = capture_step(0);
            graph_valid = true;
        }
        if (!mask) return gexec;
        auto it = gobs.find(mask);
        if (it != gobs.end()) return it->second;
        for (auto &f : fixes)
            if (is_obs(f)) obs_state(f);           // histograms allocated outside the capture
        return gobs[mask] = capture_step(mask);
    }

    // one fused step (rebuild phase, passes, commit) as a hipGraph
    hipGraphExec_t capture_step(uint32_t mask)
    {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            rebuild_phase();
            force_pass(MODE_FUSED, -1, mask);
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

    // a step the graph does not replay because of its observables: a density
    // window is written, more than 32 observable fixes, or a sample pattern
    // beyond the OBS_VARIANTS graphs kept; `fresh` collects the patterns the
    // current batch will capture
    static constexpr size_t OBS_VARIANTS = 16;
    bool obs_direct(long long t, std::vector<uint32_t> &fresh) const
    {
        if (obs.empty()) return false;
        if (obs.size() > 32 || obs_dump_due(t)) return true;
        const uint32_t mask = obs_sample_mask(t);
        if (!mask || (graph_valid && gobs.count(mask))) return false;
        if (std::find(fresh.begin(), fresh.end(), mask) != fresh.end()) return false;
        if ((graph_valid ? gobs.size() : 0) + fresh.size() >= OBS_VARIANTS) return true;
        fresh.push_back(mask);
        return false;
    }

    // Drain the stream and read the device step counter; rows that
    // overflowed in a rebuild are widened (the step that found them was
    // skipped and runs again).  Returns the step of the next force pass.
    long long settle()
    {
        hip_check(hipMemcpyAsync(pinned, (char *)dsys.p + offsetof(SysDev, step), sizeof(long long),
                                 hipMemcpyDeviceToHost, stream), "readback");
        hip_check(hipMemcpyAsync(pinned + 2, (char *)dsys.p + offsetof(SysDev, overflow), sizeof(int),
                                 hipMemcpyDeviceToHost, stream), "readback");
        hip_check(hipStreamSynchronize(stream), "readback sync");
        long long step;
        memcpy(&step, pinned, sizeof(step));
        if (pinned[2]) regrow_rows();
        return step;
    }

    // run N (ModifiedVerlet::run, mvv_meso.cu:309-572).  The reneighbor
    // decision (Neighbor::decide, src/neighbor.cpp:1216-1290) is made by each
    // step's commit on the device, counting steps since the last build across
    // runs; the host does not wait on the device inside the step loop except
    // once per batch of replayed steps, and at steps with host work (dumps,
    // observables, timed phases, the final half-kick).
    void run(long long nsteps)
    {
        if (nsteps < 0) fail(MDPD_E_ARG, "invalid run length");
        if (!run_style_ok) fail(MDPD_E_STATE, "run_style mvv/meso must be set for the USER-MESO styles");
        if (!setup_done || dirty) setup();
        if (!dumps.empty() && dump_due(ntimestep)) write_dumps(ntimestep, false);   // Verlet::setup output
        if (nsteps == 0) return;
        PoolDev P = view();
        timed(PH_INITIAL, [&]() { hip_check(launch_initial(P, stream), "initial"); });
        if (hsys.fix_arb) timed(PH_WALL, [&]() { hip_check(launch_aboundary(P, 0, stream), "aboundary"); });
        hip_check(launch_commit_step(P, COMMIT_DECIDE, split, stream), "commit");
        const long long t0 = ntimestep;
        long long k = 1;                  // the step whose force pass runs next
        while (k < nsteps) {
            const long long t = t0 + k;
            const bool out = !dumps.empty() && dump_due(t);
            std::vector<uint32_t> fresh;
            if (out || obs_direct(t, fresh) || prof || !use_graph) {
                // a step with host work: its rebuild settles before the passes
                rebuild_phase();
                if (settle() != t) fail(MDPD_E_HIP, "device step counter out of order at step %lld", t);
                force_pass(out ? MODE_FUSED_OUT : MODE_FUSED, t);
                if (out) write_dumps(t, true);
                ++k;
                continue;
            }
            long long m = 1;
            while (m < STEP_BATCH && k + m < nsteps && !(!dumps.empty() && dump_due(t + m)) && !obs_direct(t + m, fresh))
                ++m;
            // step t+q replays the graph whose observables sample t+q; a step
            // that overflows skips the rest of the batch, samples included
            for (long long q = 0; q < m; ++q)
                hip_check(hipGraphLaunch(step_graph(obs.empty() ? 0 : obs_sample_mask(t + q)), stream), "hipGraphLaunch");
            k = settle() - t0;
            if (k < 1 || k > nsteps) fail(MDPD_E_HIP, "device step counter out of range (%lld)", k + t0);
        }
        // the final half-kick: its rebuild settles first, so the pass runs
        rebuild_phase();
        if (settle() != t0 + nsteps) fail(MDPD_E_HIP, "device step counter out of order at the last step");
        const bool out = !dumps.empty() && dump_due(t0 + nsteps);
        force_pass(MODE_FINAL, t0 + nsteps);
        if (out) write_dumps(t0 + nsteps, false);
        hip_check(hipStreamSynchronize(stream), "run");
        collect_events();
        nbuild = read_field(offsetof(SysDev, builds));
        ntimestep += nsteps;
        host_stale = true;
    }

    // ------------------------------------------------------------------
    // gather
    // ------------------------------------------------------------------
    template <typename T>
    std::vector<T> download(const T *src, size_t count)
    {
        std::vector<T> h(count);
        if (count)
            hip_check(hipMemcpyAsync(h.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, stream), "download");
        return h;
    }

    void sync_sys()
    {
        hip_check(hipMemcpyAsync(&hsys, dsys.p, sizeof(SysDev), hipMemcpyDeviceToHost, stream), "sys");
        hip_check(hipStreamSynchronize(stream), "sync");
    }

    void gather(const std::string &name, int type, int count, void *data)
    {
        const int n = (int)h_tag.size();
        if (!setup_done) {
            // creation-order host copies before the first setup
            if (type == 1 && count == 3 && (name == "x" || name == "v")) {
                memcpy(data, (name == "x" ? h_x : h_v).data(), sizeof(double) * 3 * n);
                return;
            }
            if (type == 0 && count == 1 && (name == "tag" || name == "type" || name == "mask")) {
                const std::vector<int> &s = name == "tag" ? h_tag : (name == "type" ? h_type : h_mask);
                memcpy(data, s.data(), sizeof(int) * n);
                return;
            }
            fail(MDPD_E_STATE, "%s is only available after setup", name.c_str());
        }
        sync_sys();
        const int sel = hsys.sel, sela = hsys.sela;
        std::vector<int> org = download(origin[sela].p, cap);
        hip_check(hipStreamSynchronize(stream), "sync");
        auto put_d = [&](int k, int c, double v) { ((double *)data)[(size_t)org[k] * count + c] = v; };
        auto put_i = [&](int k, int c, int v) { ((int *)data)[(size_t)org[k] * count + c] = v; };
        auto want = [&](int t, int c) {
            if (type != t || count != c) fail(MDPD_E_ARG, "gather of %s needs type %d count %d", name.c_str(), t, c);
        };
        if (name == "x" || name == "v" || name == "f") {
            want(1, 3);
            std::vector<float4> h = download(name == "x" ? pos[sel].p : (name == "v" ? vel[sel].p : frc.p), cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) { put_d(k, 0, h[k].x); put_d(k, 1, h[k].y); put_d(k, 2, h[k].z); }
        } else if (name == "e_pair") {
            want(1, 1);
            std::vector<float4> h = download(frc.p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) put_d(k, 0, h[k].w);
        } else if (name == "rho") {
            want(1, 1);
            std::vector<float> h = download(rho.p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) put_d(k, 0, h[k]);
        } else if (name == "phi") {
            want(1, 1);
            std::vector<float> h = download(phi[sela].p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) put_d(k, 0, h[k]);
        } else if (name == "virial") {
            want(1, 6);
            std::vector<float> h = download(vir.p, (size_t)6 * cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k)
                for (int c = 0; c < 6; ++c) put_d(k, c, h[(size_t)c * cap + k]);
        } else if (name == "tag" || name == "mask") {
            want(0, 1);
            std::vector<int> h = download(name == "tag" ? tag[sela].p : mask[sela].p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) put_i(k, 0, h[k]);
        } else if (name == "type") {
            want(0, 1);
            std::vector<float4> h = download(pos[sel].p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) { int t; memcpy(&t, &h[k].w, 4); put_i(k, 0, t + 1); }
        } else if (name == "image") {
            want(0, 3);
            std::vector<int> h = download(image[sela].p, cap);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (int k = 0; k < n; ++k) {
                put_i(k, 0, (h[k] & 1023) - 512);
                put_i(k, 1, ((h[k] >> 10) & 1023) - 512);
                put_i(k, 2, ((h[k] >> 20) & 1023) - 512);
            }
        } else {
            fail(MDPD_E_ARG, "unknown per-atom quantity '%s'", name.c_str());
        }
    }

    // pull the device state back into the creation-order host arrays
    // (only when a run has advanced the device past the host copy, so that
    // host edits made since, such as scattered positions, are kept)
    void pull_state()
    {
        if (!setup_done || !host_stale) return;
        std::vector<double> x(3 * h_tag.size()), v(3 * h_tag.size()), ph(h_tag.size());
        std::vector<int> im(3 * h_tag.size());
        gather("x", 1, 3, x.data());
        gather("v", 1, 3, v.data());
        gather("image", 0, 3, im.data());
        gather("phi", 1, 1, ph.data());
        h_x = x;
        h_v = v;
        h_image = im;
        h_phi = ph;
        host_stale = false;
    }
};
