// lsm_api.cpp - host side of the C-ABI (include/laminarsmoke_amd.h).
//
// Replaces, for the chemistry step, what laminarSMOKE does on the host:
//   * solver selection / parameter dictionary: LS/readOptions.H:474-530 and
//     OS/math/external-ode-solvers/ODE_Parameters.hpp:39-256;
//   * map construction from kinetics.xml: LS/readOptions.H:68-78 (here: flat
//     tables already in memory, include/lsm_mechanism.h);
//   * the per-cell loop of LS/unsteady/chemistry_DI.H:92-328, which becomes one
//     batched launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/laminarsmoke_amd.h"
#include "lsm_internal.hpp"

using namespace lsm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_OK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(LSM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// dictionary parsing
// ---------------------------------------------------------------------------
struct Tok {
    std::string s;
    bool quoted;
};

std::vector<Tok> tokenize(const char *text, std::string &err) {
    std::vector<Tok> out;
    const char *p = text;
    while (*p) {
        if (isspace((unsigned char)*p)) { p++; continue; }
        if (p[0] == '/' && p[1] == '/') { while (*p && *p != '\n') p++; continue; }
        if (p[0] == '/' && p[1] == '*') {
            const char *e = strstr(p + 2, "*/");
            if (!e) { err = "unterminated comment"; return {}; }
            p = e + 2;
            continue;
        }
        if (*p == '{' || *p == '}' || *p == ';') { out.push_back({std::string(1, *p), false}); p++; continue; }
        if (*p == '"') {
            const char *e = strchr(p + 1, '"');
            if (!e) { err = "unterminated string"; return {}; }
            out.push_back({std::string(p + 1, e), true});
            p = e + 1;
            continue;
        }
        const char *b = p;
        while (*p && !isspace((unsigned char)*p) && *p != ';' && *p != '{' && *p != '}' && *p != '"') p++;
        out.push_back({std::string(b, p), false});
    }
    return out;
}

bool parse_double(const std::string &s, double &v) {
    if (s.empty()) return false;
    char *end = nullptr;
    v = strtod(s.c_str(), &end);
    return end && *end == '\0' && std::isfinite(v);
}

bool parse_int(const std::string &s, int &v) {
    if (s.empty()) return false;
    char *end = nullptr;
    long x = strtol(s.c_str(), &end, 10);
    if (!end || *end != '\0' || x < -2147483647L || x > 2147483647L) return false;
    v = (int)x;
    return true;
}

// OpenFOAM Switch words and OpenSMOKE booleans
bool parse_bool(std::string s, int &v) {
    for (auto &c : s) c = (char)tolower((unsigned char)c);
    if (s == "true" || s == "on" || s == "yes" || s == "y" || s == "t" || s == "1") { v = 1; return true; }
    if (s == "false" || s == "off" || s == "no" || s == "n" || s == "f" || s == "none" || s == "0") { v = 0; return true; }
    return false;
}

// names accepted by ODE_Parameters::SetupFromDictionary (ODE_Parameters.hpp:43-120)
// and readOptions.H:495-513, plus the GPU integrator
bool solver_type(const std::string &name, int &type) {
    static const struct { const char *n; int t; } tab[] = {
        {"OpenSMOKE", LSM_ODE_OPENSMOKE}, {"BzzOde", LSM_ODE_BZZODE}, {"CVODE", LSM_ODE_CVODE},
        {"DVODE", LSM_ODE_DVODE},         {"DASPK", LSM_ODE_DASPK},   {"DLSODE", LSM_ODE_DLSODE},
        {"DLSODA", LSM_ODE_DLSODA},       {"RADAU5", LSM_ODE_RADAU5}, {"MEBDF", LSM_ODE_MEBDF},
        {"CHEMEQ2", LSM_ODE_CHEMEQ2},     {"OpenSMOKE_GPU", LSM_ODE_OPENSMOKE_GPU},
    };
    for (auto &e : tab)
        if (name == e.n) { type = e.t; return true; }
    return false;
}

int apply_key(const std::string &key, const std::string &val, lsm_ode_parameters *p) {
    double d;
    int i;
    auto bad = [&](void) { return fail(LSM_E_PARSE, "invalid value '" + val + "' for key '" + key + "'"); };
    if (key == "odeSolver" || key == "@OdeSolver") {
        if (!solver_type(val, i)) return fail(LSM_E_PARSE, "unknown ODE solver '" + val + "'");
        p->type = i;
    } else if (key == "relTolerance" || key == "@RelativeTolerance") {
        if (!parse_double(val, d)) return bad();
        p->relative_tolerance = d;
    } else if (key == "absTolerance" || key == "@AbsoluteTolerance") {
        if (!parse_double(val, d)) return bad();
        p->absolute_tolerance = d;
    } else if (key == "maximumOrder" || key == "@MaximumOrder") {
        if (!parse_int(val, i)) return bad();
        p->maximum_order = i;
    } else if (key == "fullPivoting" || key == "@FullPivoting") {
        if (!parse_bool(val, i)) return bad();
        p->full_pivoting = i;
    } else if (key == "@MaximumNumberOfSteps") {
        if (!parse_int(val, i)) return bad();
        p->maximum_number_of_steps = i;
    } else if (key == "@MaximumStep") {
        if (!parse_double(val, d)) return bad();
        p->maximum_step = d;
    } else if (key == "@MinimumStep") {
        if (!parse_double(val, d)) return bad();
        p->minimum_step = d;
    } else if (key == "@InitialStep") {
        if (!parse_double(val, d)) return bad();
        p->initial_step = d;
    } else if (key == "@LinearAlgebra") {
        if (val.size() >= sizeof(p->linear_algebra)) return bad();
        snprintf(p->linear_algebra, sizeof(p->linear_algebra), "%s", val.c_str());
    } else if (key == "@SparseSolver" || key == "@SparsePreconditioner") {
        // sparse options of the grammar (ODE_Parameters.h:182-192): accepted, no effect on the dense kernel
    } else if (key == "@SparsePreconditionerDropTolerance") {
        if (!parse_double(val, d)) return bad();
    } else if (key == "@SparsePreconditionerFillFactor") {
        if (!parse_int(val, i)) return bad();
    } else {
        return fail(LSM_E_PARSE, "unknown key '" + key + "'");
    }
    return LSM_OK;
}

// ---------------------------------------------------------------------------
// Gear l-vectors, MethodGear::MemoryAllocationMethod (MethodGear.hpp:155-178)
// ---------------------------------------------------------------------------
double factorial(unsigned n) {
    double f = 1.;
    for (unsigned i = 2; i <= n; i++) f *= (double)i;
    return f;
}

void gear_r(double r[MAX_ORDER + 1][MAX_ORDER]) {
    for (int i = 0; i <= MAX_ORDER; i++)
        for (int j = 0; j < MAX_ORDER; j++) r[i][j] = 0.;
    r[0][0] = 1.;
    r[1][0] = 1.;
    for (int j = 2; j <= MAX_ORDER; j++) {
        r[0][j - 1] = factorial(j);
        for (int i = 2; i <= j; i++) r[i - 1][j - 1] = (double)j * r[i - 1][j - 2] + r[i - 2][j - 2];
        r[j][j - 1] = 1.;
    }
    double sum = 1.;
    for (int j = 2; j <= MAX_ORDER; j++) {
        sum += 1. / (double)j;
        for (int i = 1; i <= j + 1; i++) r[i - 1][j - 1] /= (factorial(j) * sum);
    }
}

int nu_class(double nu) {
    if (nu == 1.) return 1;
    if (nu == 2.) return 2;
    if (nu == 3.) return 3;
    if (nu == 0.5) return 4;
    return 5;
}

// host image of the device tables: one aligned blob, offsets patched afterwards
struct Blob {
    std::vector<char> bytes;
    template <class T>
    size_t add(const std::vector<T> &v) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + std::max<size_t>(v.size(), 1) * sizeof(T), 0);
        if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

}  // namespace

struct lsm_handle {
    int device = 0;
    int num_cu = 0;
    lsm_ode_parameters ode{};
    lsm_reactor_options reactor{};
    int method = LSM_METHOD_AUTO;
    DevMech dm{};
    KParams kp{};
    void *d_tables = nullptr;
    unsigned *d_queue = nullptr;
    double *d_work = nullptr;
    int64_t work_slots = 0;
    // staging for host-pointer calls
    double *d_y = nullptr, *d_dt = nullptr, *d_q = nullptr;
    int *d_status = nullptr, *d_cnt = nullptr;
    int64_t cap = 0;
    int team_grid = 0;
    int team_wpb = 0;
    int team_cpw = 1;           // cells per wavefront of the team kernel (1, 2 or 4)
    int pack_grid = 0, pack_wpb = 0;   // launch shape of the packed kernel (team_cpw > 1)
    size_t lds_cu = 0;          // LDS bytes per CU
    int wide_grid = 0;
    double *d_wjws = nullptr;   // wide kernel workspaces, [wide_grid][2][ne][ne]: Jacobian, FP64 LU
    unsigned long long *d_fallbacks = nullptr;   // wide kernel: factorisations redone in FP64
    int team_max_waves = 0;     // waves per CU x CUs the team kernel's register cap allows
    size_t tables_bytes = 0;
    // cost-order sort buffers
    unsigned *d_keys = nullptr, *d_keys_out = nullptr;
    int *d_vals = nullptr;
    void *d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int64_t sort_cap = 0;
    // transport fits (lsm_set_transport)
    DevTransport tr{};
    void *d_transport = nullptr;
    void *d_transport_opts = nullptr;        // Lewis numbers and soot corrections (lsm_set_transport_options)
    size_t transport_bytes = 0, transport_opts_bytes = 0;
};

int lsm_fail(int code, const char *msg) { return fail(code, msg); }

// Default team launch shape: the most waves per CU within the register cap and the LDS, in
// the fewest workgroups (each copies the mechanism image); false when the team kernel cannot
// run this mechanism.
static bool team_shape_default(lsm_handle *h) {
    const int cap = team_waves_per_cu(h->dm);
    int best_w = 0, best_wpb = 0, best_bpc = 0;
    for (int wpb = 4; wpb >= 1; wpb--)
        for (int bpc = 1; wpb * bpc <= cap; bpc++) {
            if (!team_shape_fits(h->dm, wpb, bpc, h->lds_cu)) continue;
            if (wpb * bpc > best_w) { best_w = wpb * bpc; best_wpb = wpb; best_bpc = bpc; }
        }
    if (best_w == 0) return false;
    h->team_wpb = best_wpb;
    h->team_grid = best_bpc * h->num_cu;
    h->team_max_waves = cap * h->num_cu;
    return true;
}

// Default launch shape of the packed team kernel with cpw cells per wavefront (as above);
// false when the mechanism does not allow it.
static bool pack_shape_default(lsm_handle *h, int cpw) {
    const int cap = team_waves_per_cu_cpw(h->dm, cpw);
    int best_w = 0, best_wpb = 0, best_bpc = 0;
    for (int wpb = 4; wpb >= 1; wpb--)
        for (int bpc = 1; wpb * bpc <= cap; bpc++) {
            if (!team_shape_fits_cpw(h->dm, cpw, wpb, bpc, h->lds_cu)) continue;
            if (wpb * bpc > best_w) { best_w = wpb * bpc; best_wpb = wpb; best_bpc = bpc; }
        }
    if (best_w == 0) return false;
    h->pack_wpb = best_wpb;
    h->pack_grid = best_bpc * h->num_cu;
    return true;
}

// Device bytes the handle holds now: tables, queue, kernel workspaces, and the buffers grown on
// demand (plain-kernel workspace, host-call staging, cost-order sort, transport tables).
size_t lsm_handle_device_bytes(const lsm_handle *h) {
    size_t b = h->tables_bytes + 256;
    b += sizeof(double) * wide_jac_doubles(h->dm) * (size_t)h->wide_grid + (h->d_fallbacks ? 8 : 0);
    b += h->d_work ? (size_t)h->work_slots * plain_work_doubles(h->dm, 1) * sizeof(double) : 0;
    b += (size_t)h->cap * (sizeof(double) * (h->dm.ne + 2) + sizeof(int) * (1 + LSM_NCOUNTERS));
    b += (size_t)h->sort_cap * (2 * sizeof(unsigned) + sizeof(int)) + h->sort_tmp_bytes;
    b += h->transport_bytes + h->transport_opts_bytes;
    return b;
}

int lsm_handle_neq(const lsm_handle *h) { return h->dm.ne; }
double lsm_handle_pressure(const lsm_handle *h) { return h->reactor.pressure; }

extern "C" {

const char *lsm_last_error(void) { return g_err.c_str(); }

const char *lsm_version(void) { return "laminarsmoke_amd 0.1 (gfx950)"; }

int lsm_ode_parameters_default(lsm_ode_parameters *p) {
    if (!p) return fail(LSM_E_INVALID, "null parameters");
    memset(p, 0, sizeof(*p));
    p->type = LSM_ODE_OPENSMOKE_GPU;
    p->relative_tolerance = 100. * (double)1.192092896e-07F;   // 100*MachEpsFloat()
    p->absolute_tolerance = 1.e-10;
    p->minimum_step = -1.;
    p->maximum_step = -1.;
    p->initial_step = -1.;
    p->maximum_order = -1;
    p->maximum_number_of_steps = -1;
    p->full_pivoting = 0;
    snprintf(p->linear_algebra, sizeof(p->linear_algebra), "Eigen");
    return LSM_OK;
}

int lsm_ode_parameters_parse(const char *text, lsm_ode_parameters *p) {
    if (!text || !p) return fail(LSM_E_INVALID, "null argument");
    std::string err;
    std::vector<Tok> t = tokenize(text, err);
    if (!err.empty()) return fail(LSM_E_PARSE, err);
    lsm_ode_parameters q = *p;
    size_t i = 0;
    int depth = 0;
    while (i < t.size()) {
        if (!t[i].quoted && t[i].s == "}") {
            if (--depth < 0) return fail(LSM_E_PARSE, "unbalanced '}'");
            i++;
            continue;
        }
        if (!t[i].quoted && t[i].s == ";") { i++; continue; }
        if (i + 1 < t.size() && !t[i + 1].quoted && t[i + 1].s == "{") {   // sub-dictionary name
            depth++;
            i += 2;
            continue;
        }
        const std::string key = t[i].s;
        if (i + 1 >= t.size() || (!t[i + 1].quoted && (t[i + 1].s == ";" || t[i + 1].s == "}")))
            return fail(LSM_E_PARSE, "missing value for key '" + key + "'");
        const std::string val = t[i + 1].s;
        int rc = apply_key(key, val, &q);
        if (rc != LSM_OK) return rc;
        i += 2;
        if (i < t.size() && !t[i].quoted && t[i].s == ";") i++;
        else if (key[0] != '@' && !(i < t.size() && t[i].s == "}"))
            return fail(LSM_E_PARSE, "missing ';' after key '" + key + "'");
    }
    if (depth != 0) return fail(LSM_E_PARSE, "unbalanced '{'");
    *p = q;
    return LSM_OK;
}

int lsm_reactor_options_default(lsm_reactor_options *r) {
    if (!r) return fail(LSM_E_INVALID, "null options");
    r->pressure = 101325.;
    r->min_temperature = 0.;
    r->energy_equation = 1;
    r->mass_fractions_tol = 1.e-4;
    r->constant_volume = 0;
    r->strang_compact = 0;
    return LSM_OK;
}

static int check_mechanism(const lsm_mechanism *m) {
    if (!m) return fail(LSM_E_INVALID, "null mechanism");
    if (m->ns < 1 || m->nr < 0) return fail(LSM_E_INVALID, "mechanism needs ns >= 1 and nr >= 0");
    if (!m->nasa || (m->nr > 0 && (!m->lnA || !m->beta || !m->E_over_R || !m->rflags || !m->reac_ptr ||
                                   !m->reac_sp || !m->reac_nu || !m->prod_ptr || !m->prod_sp || !m->prod_nu)))
        return fail(LSM_E_INVALID, "mechanism table pointer is null");
    for (int j = 0; j < m->nr; j++) {
        for (int side = 0; side < 2; side++) {
            const int32_t *ptr = side ? m->prod_ptr : m->reac_ptr;
            const int32_t *sp = side ? m->prod_sp : m->reac_sp;
            if (ptr[j] > ptr[j + 1]) return fail(LSM_E_INVALID, "reaction CSR pointers decrease");
            for (int q = ptr[j]; q < ptr[j + 1]; q++)
                if (sp[q] < 0 || sp[q] >= m->ns) return fail(LSM_E_INVALID, "species index out of range");
        }
    }
    for (int s = 0; s < m->ntb; s++) {
        if (m->tb_rxn[s] < 0 || m->tb_rxn[s] >= m->nr) return fail(LSM_E_INVALID, "third-body reaction index");
        for (int q = m->tb_ptr[s]; q < m->tb_ptr[s + 1]; q++)
            if (m->tb_sp[q] < 0 || m->tb_sp[q] >= m->ns) return fail(LSM_E_INVALID, "third-body species index");
    }
    for (int j = 0; j < m->nr; j++) {
        if (m->rflags[j] & ~(LSM_RXN_REVERSIBLE | LSM_RXN_NEGATIVE_A | LSM_RXN_EXPLICIT_REV))
            return fail(LSM_E_UNSUPPORTED, "unknown reaction flag");
        if (m->rflags[j] & LSM_RXN_EXPLICIT_REV) {
            if (!(m->rflags[j] & LSM_RXN_REVERSIBLE))
                return fail(LSM_E_INVALID, "explicit reverse parameters on an irreversible reaction");
            if (!m->rev_arr) return fail(LSM_E_INVALID, "explicitly reversible reaction without rev_arr");
        }
    }
    for (int s = 0; s < m->nfo; s++) {
        if (m->fo_rxn[s] < 0 || m->fo_rxn[s] >= m->nr) return fail(LSM_E_INVALID, "fall-off reaction index");
        const int form = m->fo_type[s] & ~LSM_FALLOFF_CHEMACT;
        if (form != LSM_FALLOFF_LINDEMANN && form != LSM_FALLOFF_TROE && form != LSM_FALLOFF_SRI)
            return fail(LSM_E_UNSUPPORTED, "fall-off type (Lindemann, Troe or SRI)");
        // ReactionPolicy_CHEMKIN.hpp:1392-1407
        if (m->rflags[m->fo_rxn[s]] & LSM_RXN_EXPLICIT_REV)
            return fail(LSM_E_INVALID, "explicitly reversible fall-off / CABR reactions are not allowed");
        if (form == LSM_FALLOFF_SRI) {
            if (!m->fo_sri) return fail(LSM_E_INVALID, "SRI fall-off reaction without fo_sri parameters");
            const double *t = m->fo_sri + 5 * s;
            if (!(t[3] > 0.) || t[2] == 0.) return fail(LSM_E_INVALID, "SRI parameters need d > 0 and c != 0");
        }
        if (m->fo_single[s] >= m->ns) return fail(LSM_E_INVALID, "fall-off collider species index");
        for (int q = m->fo_ptr[s]; q < m->fo_ptr[s + 1]; q++)
            if (m->fo_sp[q] < 0 || m->fo_sp[q] >= m->ns) return fail(LSM_E_INVALID, "fall-off species index");
    }
    // a CHEMKIN reaction is either "+M" (third body) or "(+M)" (fall-off), once
    std::vector<char> seen(m->nr, 0);
    for (int s = 0; s < m->ntb + m->nfo; s++) {
        const int j = s < m->ntb ? m->tb_rxn[s] : m->fo_rxn[s - m->ntb];
        if (seen[j]) return fail(LSM_E_INVALID, "reaction listed twice among third-body / fall-off reactions");
        seen[j] = 1;
    }
    if (m->npl < 0 || (m->npl > 0 && (!m->pl_rxn || !m->pl_ptr || !m->pl_par)))
        return fail(LSM_E_INVALID, "PLOG tables");
    {
        std::vector<char> pl(m->nr, 0);
        for (int s = 0; s < m->npl; s++) {
            const int j = m->pl_rxn[s];
            if (j < 0 || j >= m->nr) return fail(LSM_E_INVALID, "PLOG reaction index");
            if (pl[j]) return fail(LSM_E_INVALID, "reaction listed twice among PLOG reactions");
            pl[j] = 1;
            if (m->pl_ptr[s + 1] - m->pl_ptr[s] < 1) return fail(LSM_E_INVALID, "PLOG reaction without levels");
            for (int q = m->pl_ptr[s]; q < m->pl_ptr[s + 1]; q++) {
                if (!(m->pl_par[4 * q] > 0.)) return fail(LSM_E_INVALID, "PLOG pressure must be positive");
                // PressureLogarithmicRateExpression.hpp:97-101
                if (q > m->pl_ptr[s] && !(m->pl_par[4 * q] > m->pl_par[4 * (q - 1)]))
                    return fail(LSM_E_INVALID, "PLOG pressures must increase");
            }
        }
        // PLOG cannot be combined with LOW / HIGH (ReactionPolicy_CHEMKIN.hpp:778-785)
        for (int s = 0; s < m->nfo; s++)
            if (pl[m->fo_rxn[s]]) return fail(LSM_E_INVALID, "PLOG on a fall-off / CABR reaction");
        // Chebyshev: not with LOW / HIGH / PLOG (ReactionPolicy_CHEMKIN.hpp:456-474)
        if (m->nch < 0 || (m->nch > 0 && (!m->ch_rxn || !m->ch_dims || !m->ch_lim || !m->ch_ptr || !m->ch_a)))
            return fail(LSM_E_INVALID, "Chebyshev tables");
        for (int s = 0; s < m->nch; s++) {
            const int j = m->ch_rxn[s];
            if (j < 0 || j >= m->nr) return fail(LSM_E_INVALID, "Chebyshev reaction index");
            if (pl[j]) return fail(LSM_E_INVALID, "reaction listed twice among PLOG / Chebyshev reactions");
            pl[j] = 1;
            const int N = m->ch_dims[2 * s], M = m->ch_dims[2 * s + 1];
            if (N < 1 || M < 1 || N > 64 || M > 64 || m->ch_ptr[s + 1] - m->ch_ptr[s] != N * M)
                return fail(LSM_E_INVALID, "Chebyshev coefficient table is not N x M (1 <= N, M <= 64)");
            const double *L = m->ch_lim + 5 * s;
            if (!(L[0] > 0. && L[1] > L[0] && L[2] > 0. && L[3] > L[2]))
                return fail(LSM_E_INVALID, "Chebyshev limits need 0 < Tmin < Tmax and 0 < Pmin < Pmax");
        }
        for (int s = 0; s < m->nfo; s++)
            if (pl[m->fo_rxn[s]]) return fail(LSM_E_INVALID, "Chebyshev on a fall-off / CABR reaction");
    }
    if (m->ord_ptr) {
        if (m->ord_ptr[0] != 0) return fail(LSM_E_INVALID, "ord_ptr[0] must be 0");
        for (int j = 0; j < m->nr; j++) {
            if (m->ord_ptr[j + 1] < m->ord_ptr[j]) return fail(LSM_E_INVALID, "ord_ptr must not decrease");
            if (m->ord_ptr[j + 1] > m->ord_ptr[j] && (!m->ord_sp || !m->ord_val))
                return fail(LSM_E_INVALID, "explicit orders without ord_sp / ord_val");
            for (int q = m->ord_ptr[j]; q < m->ord_ptr[j + 1]; q++)
                if (m->ord_sp[q] < 0 || m->ord_sp[q] >= m->ns) return fail(LSM_E_INVALID, "order species index");
        }
    }
    return LSM_OK;
}


}  // extern "C"

// Compact LDS image of the mechanism for the team kernel (layout: lsm_internal.hpp).
struct Image {
    std::vector<char> b;
    template <class T>
    int put(const std::vector<T> &v) {
        const size_t off = (b.size() + 15) & ~size_t(15);
        b.resize(off + v.size() * sizeof(T), 0);
        if (!v.empty()) memcpy(b.data() + off, v.data(), v.size() * sizeof(T));
        return (int)off;
    }
};

// Forward orders of reaction j: the explicit orders when given, else the reactant
// coefficients (sorted by species); and whether each side is non-elementary (an
// order below 1 takes the smoothed form, StoichiometricMap.hpp:256-300, 617-654).
static std::vector<std::pair<int, double>> forward_orders(const lsm_mechanism *m, int j) {
    std::vector<std::pair<int, double>> o;
    if (m->ord_ptr && m->ord_ptr[j + 1] > m->ord_ptr[j])
        for (int q = m->ord_ptr[j]; q < m->ord_ptr[j + 1]; q++) o.push_back({m->ord_sp[q], m->ord_val[q]});
    else
        for (int q = m->reac_ptr[j]; q < m->reac_ptr[j + 1]; q++) o.push_back({m->reac_sp[q], m->reac_nu[q]});
    std::stable_sort(o.begin(), o.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
        return a.first < b.first;
    });
    return o;
}
static bool explicit_orders(const lsm_mechanism *m, int j) {
    return m->ord_ptr && m->ord_ptr[j + 1] > m->ord_ptr[j];
}
static bool smooth_forward(const lsm_mechanism *m, int j) {
    for (auto &e : forward_orders(m, j)) if (e.second < 1.) return true;
    return false;
}
static int plog_slot_of(const lsm_mechanism *m, int j) {
    for (int s = 0; s < m->npl; s++) if (m->pl_rxn[s] == j) return s;
    return -1;
}
static int cheb_slot_of(const lsm_mechanism *m, int j) {
    for (int s = 0; s < m->nch; s++) if (m->ch_rxn[s] == j) return s;
    return -1;
}
// Chebyshev block of slot s: N, M, Tmin, Tmax, log10 Pmin, log10 Pmax, conv, a[N*M]
static void cheb_block(const lsm_mechanism *m, int s, std::vector<double> &out) {
    const int N = m->ch_dims[2 * s], M = m->ch_dims[2 * s + 1];
    const double *L = m->ch_lim + 5 * s;
    out.push_back(N); out.push_back(M); out.push_back(L[0]); out.push_back(L[1]);
    out.push_back(std::log10(L[2])); out.push_back(std::log10(L[3])); out.push_back(L[4]);
    for (int q = m->ch_ptr[s]; q < m->ch_ptr[s] + N * M; q++) out.push_back(m->ch_a[q]);
}
// A reaction the team kernel's paired loop can take: no collider, no tabulated or explicit reverse
// rate, no explicit orders, integer coefficients 1-3 with at most three species (counted with
// multiplicity) on each side.  It is then a fixed-format record of the image (build_image).
static bool smooth_reverse(const lsm_mechanism *m, int j);
static bool simple_record(const lsm_mechanism *m, int j, const std::vector<int> &tb_of, const std::vector<int> &fo_of) {
    if (tb_of[j] >= 0 || fo_of[j] >= 0 || plog_slot_of(m, j) >= 0 || cheb_slot_of(m, j) >= 0) return false;
    if ((m->rflags[j] & LSM_RXN_EXPLICIT_REV) || explicit_orders(m, j) || smooth_forward(m, j) || smooth_reverse(m, j))
        return false;
    for (int side = 0; side < 2; side++) {
        const int *ptr = side ? m->prod_ptr : m->reac_ptr;
        const double *nu = side ? m->prod_nu : m->reac_nu;
        double mult = 0.;
        for (int q = ptr[j]; q < ptr[j + 1]; q++) {
            if (nu[q] != std::floor(nu[q]) || nu[q] < 1. || nu[q] > 3.) return false;
            mult += nu[q];
        }
        if (mult > 3.) return false;
    }
    return true;
}

static bool smooth_reverse(const lsm_mechanism *m, int j) {
    if (!(m->rflags[j] & LSM_RXN_REVERSIBLE)) return false;
    for (int q = m->prod_ptr[j]; q < m->prod_ptr[j + 1]; q++) if (m->prod_nu[q] < 1.) return true;
    return false;
}

static bool build_image(const lsm_mechanism *m, const std::vector<int> &perm, const std::vector<int> &tb_of,
                        const std::vector<int> &fo_of, int ns_pad, LdsImage &L, std::vector<char> &out) {
    const int ns = m->ns, nr = m->nr;
    L = LdsImage{};
    if (ns > 254 || m->ntb > 254 || m->nfo > 254) return false;
    const unsigned dummy = ns <= 63 ? (unsigned)LSM_DUMMY_SP : 255u;
    Image im;
    std::vector<double> sp((size_t)LSP_ROWS * ns_pad, 0.);
    for (int k = 0; k < ns; k++) {
        const double *c = m->nasa + LSM_NASA_STRIDE * k;
        for (int i = 0; i < 7; i++) { sp[(size_t)i * ns_pad + k] = c[i]; sp[(size_t)(7 + i) * ns_pad + k] = c[7 + i]; }
        sp[(size_t)14 * ns_pad + k] = c[16];
        sp[(size_t)15 * ns_pad + k] = c[17];
    }
    std::vector<double> P((size_t)4 * nr), nuv, rvp, pll;
    std::vector<unsigned> W((size_t)4 * nr), ent, plp;
    for (int s = 0; s < m->npl; s++) {
        plp.push_back((unsigned)(pll.size() / 4));
        for (int q = m->pl_ptr[s]; q < m->pl_ptr[s + 1]; q++) {
            pll.push_back(std::log(m->pl_par[4 * q]));
            for (int i = 1; i < 4; i++) pll.push_back(m->pl_par[4 * q + i]);
        }
    }
    plp.push_back((unsigned)(pll.size() / 4));
    std::vector<double> chd;
    std::vector<unsigned> chp;
    for (int s = 0; s < m->nch; s++) { chp.push_back((unsigned)chd.size()); cheb_block(m, s, chd); }
    chp.push_back((unsigned)chd.size());
    auto code_of = [&](double nu) -> unsigned {
        const double t = 2. * nu;
        if (t == std::floor(t) && t >= 1. && t <= 15.) return (unsigned)t;
        for (size_t i = 0; i < nuv.size(); i++) if (nuv[i] == nu) return 16u + (unsigned)i;
        nuv.push_back(nu);
        return 16u + (unsigned)(nuv.size() - 1);
    };
    for (int jn = 0; jn < nr; jn++) {
        const int j = perm[jn];
        P[4 * jn + 0] = m->lnA[j]; P[4 * jn + 1] = m->beta[j]; P[4 * jn + 2] = m->E_over_R[j];
        double d = 0.;
        std::vector<std::pair<int, double>> re, pr;
        for (int q = m->reac_ptr[j]; q < m->reac_ptr[j + 1]; q++) { re.push_back({m->reac_sp[q], m->reac_nu[q]}); d -= m->reac_nu[q]; }
        for (int q = m->prod_ptr[j]; q < m->prod_ptr[j + 1]; q++) { pr.push_back({m->prod_sp[q], m->prod_nu[q]}); d += m->prod_nu[q]; }
        P[4 * jn + 3] = d;
        const int pls = plog_slot_of(m, j), chs = cheb_slot_of(m, j);
        if (pls >= 0) { P[4 * jn + 0] = (double)pls; P[4 * jn + 1] = 0.; }
        if (chs >= 0) { P[4 * jn + 0] = (double)chs; P[4 * jn + 1] = 1.; }
        const bool tab = pls >= 0 || chs >= 0;
        const bool xrev = (m->rflags[j] & LSM_RXN_EXPLICIT_REV) != 0;
        if (xrev) {
            P[4 * jn + 3] = (double)(rvp.size() / 4);
            for (int i = 0; i < 3; i++) rvp.push_back(m->rev_arr[3 * j + i]);
            rvp.push_back(0.);
        }
        auto bysp = [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; };
        std::stable_sort(re.begin(), re.end(), bysp);
        std::stable_sort(pr.begin(), pr.end(), bysp);
        const bool xo = explicit_orders(m, j), sf = smooth_forward(m, j), sb = smooth_reverse(m, j);
        const std::vector<std::pair<int, double>> od = xo ? forward_orders(m, j) : std::vector<std::pair<int, double>>();
        if (re.size() > 255 || pr.size() > 255 || od.size() > 255 ||
            ent.size() + re.size() + pr.size() + od.size() > 65535) return false;
        W[4 * jn + 3] = (unsigned)ent.size() | ((unsigned)re.size() << 16) | ((unsigned)pr.size() << 24);
        for (auto &e : re) ent.push_back((unsigned)e.first | (code_of(e.second) << 12));
        for (auto &e : pr) ent.push_back((unsigned)e.first | (code_of(e.second) << 12));
        for (auto &e : od) ent.push_back((unsigned)e.first | (code_of(e.second) << 12));
        // fixed-format record: integer orders, at most 3 species per side (multiplicity counted)
        auto pack = [&](const std::vector<std::pair<int, double>> &side, unsigned &word) {
            std::vector<unsigned> sl;
            for (auto &e : side) {
                if (e.second != std::floor(e.second) || e.second < 1. || e.second > 3.) return false;
                for (int r = 0; r < (int)e.second; r++) sl.push_back((unsigned)e.first);
            }
            if (sl.size() > 3) return false;
            while (sl.size() < 3) sl.push_back(dummy);
            word = sl[0] | (sl[1] << 8) | (sl[2] << 16);
            return true;
        };
        unsigned wr = 0, wp = 0;
        const bool fast = !xo && !sf && !sb && !xrev && pack(re, wr) && pack(pr, wp);
        W[4 * jn + 1] = wr;
        W[4 * jn + 2] = wp;
        // bits: 0 reversible, 1 negative A, 2 fixed format, 3/4 non-elementary forward/reverse
        // side, 5 explicit reverse, 6 chemically activated, 7 PLOG / Chebyshev, 8-15 third-body
        // slot + 1, 16-23 fall-off slot + 1, 24-31 explicit orders.  A PLOG or Chebyshev
        // constant replaces the sign-flipped Arrhenius one (KineticsMap_CHEMKIN.hpp:1208-1210,
        // 1333-1349).
        const bool cabr = fo_of[j] >= 0 && (m->fo_type[fo_of[j]] & LSM_FALLOFF_CHEMACT);
        W[4 * jn + 0] = (unsigned)(m->rflags[j] & (tab ? 1 : 3)) | (fast ? 4u : 0u) | (sf ? 8u : 0u) |
                        (sb ? 16u : 0u) | (xrev ? 32u : 0u) | (cabr ? 64u : 0u) | (tab ? 128u : 0u) |
                        ((unsigned)(tb_of[j] + 1) << 8) | ((unsigned)(fo_of[j] + 1) << 16) |
                        ((unsigned)od.size() << 24);
    }
    if (nuv.size() > (1u << 19)) return false;
    std::vector<unsigned> tbp(m->ntb + 1), tbs;
    std::vector<double> tbe;
    for (int s = 0; s < m->ntb; s++) {
        tbp[s] = (unsigned)tbs.size();
        for (int q = m->tb_ptr[s]; q < m->tb_ptr[s + 1]; q++) { tbs.push_back((unsigned)m->tb_sp[q]); tbe.push_back(m->tb_eff_m1[q]); }
    }
    tbp[m->ntb] = (unsigned)tbs.size();
    std::vector<double> fop((size_t)8 * m->nfo);
    std::vector<int> fos(m->nfo);
    std::vector<unsigned> fol(m->nfo + 1), fos2;
    std::vector<double> foe;
    for (int s = 0; s < m->nfo; s++) {
        // [0..2] k_inf; Troe: [3..6] a, T***, T*, T**, [7] = 1; SRI: [3..6] a, b, c, e,
        // [7] = -d (d > 0); Lindemann: [7] = 0
        for (int i = 0; i < 3; i++) fop[8 * s + i] = m->fo_inf[3 * s + i];
        const int form = m->fo_type[s] & 3;
        if (form == LSM_FALLOFF_SRI) {
            const double *t = m->fo_sri + 5 * s;
            fop[8 * s + 3] = t[0]; fop[8 * s + 4] = t[1]; fop[8 * s + 5] = t[2]; fop[8 * s + 6] = t[4];
            fop[8 * s + 7] = -t[3];
        } else {
            for (int i = 0; i < 4; i++) fop[8 * s + 3 + i] = m->fo_troe[4 * s + i];
            fop[8 * s + 7] = (double)form;
        }
        fos[s] = m->fo_single[s];
        fol[s] = (unsigned)fos2.size();
        for (int q = m->fo_ptr[s]; q < m->fo_ptr[s + 1]; q++) { fos2.push_back((unsigned)m->fo_sp[q]); foe.push_back(m->fo_eff_m1[q]); }
    }
    fol[m->nfo] = (unsigned)fos2.size();
    L.o_sp = im.put(sp);
    L.o_P = im.put(P); L.o_W = im.put(W);
    L.o_ent = im.put(ent); L.o_nuv = im.put(nuv);
    L.o_tbp = im.put(tbp); L.o_tbs = im.put(tbs); L.o_tbe = im.put(tbe);
    L.o_fop = im.put(fop); L.o_fos = im.put(fos); L.o_fol = im.put(fol); L.o_fos2 = im.put(fos2); L.o_foe = im.put(foe);
    L.o_rvp = im.put(rvp); L.o_plp = im.put(plp); L.o_pll = im.put(pll);
    L.o_chp = im.put(chp); L.o_chd = im.put(chd);
    im.b.resize((im.b.size() + 15) & ~size_t(15), 0);
    L.bytes = (int)im.b.size();
    L.team_ok = ns <= 63 ? 1 : 0;
    L.wide_ok = 1;
    L.dummy = (int)dummy;
    out.swap(im.b);
    return true;
}

static int upload_tables(lsm_handle *h, const lsm_mechanism *m) {
    const int ns = m->ns, nr = m->nr;
    DevMech &d = h->dm;
    d.ns = ns; d.nr = nr; d.ne = ns + 1; d.ntb = m->ntb; d.nfo = m->nfo;
    d.xk = (m->npl > 0 || m->nch > 0) ? 1 : 0;
    for (int j = 0; j < nr; j++) if (m->rflags[j] & LSM_RXN_EXPLICIT_REV) d.xk = 1;
    for (int s = 0; s < m->nfo; s++) if (m->fo_type[s] & LSM_FALLOFF_CHEMACT) d.xk = 1;
    d.ns_pad = (ns + 63) / 64 * 64;
    // species coefficient rows (SetCoefficients, ThermodynamicsMap_CHEMKIN.hpp:263-343)
    std::vector<double> sp((size_t)NSP_ROWS * d.ns_pad, 0.);
    const double one_third = 1. / 3.;
    for (int k = 0; k < ns; k++) {
        const double *c = m->nasa + LSM_NASA_STRIDE * k;
        auto S = [&](int row) -> double & { return sp[(size_t)row * d.ns_pad + k]; };
        for (int hi = 0; hi < 2; hi++) {
            const double *a = hi ? c : c + 7;
            const double a6 = hi ? c[5] : c[12], a7 = hi ? c[6] : c[13];
            const int rc = hi ? SP_CP_HT : SP_CP_LT, rh = hi ? SP_DH_HT : SP_DH_LT, rs = hi ? SP_DS_HT : SP_DS_LT;
            for (int i = 0; i < 5; i++) S(rc + i) = a[i];
            S(rh + 0) = a[0]; S(rh + 1) = 0.50 * a[1]; S(rh + 2) = one_third * a[2]; S(rh + 3) = 0.25 * a[3];
            S(rh + 4) = 0.20 * a[4]; S(rh + 5) = a6;
            S(rs + 0) = a[0]; S(rs + 1) = a[1]; S(rs + 2) = 0.50 * a[2]; S(rs + 3) = one_third * a[3];
            S(rs + 4) = 0.25 * a[4]; S(rs + 5) = a7;
        }
        S(SP_TMID) = c[16];
        S(SP_MW) = c[17];
    }
    // Reaction renumbering: fall-off, then third-body, then elementary.  The
    // reaction loops hand reaction j to lane j mod 64 (team) or thread j mod 256
    // (wide); the fall-off work (Troe: several transcendental functions) is a branch
    // the whole wave waits for, so the fall-off reactions are packed into as few
    // passes of the loop as possible (team), or dealt round-robin to the four waves
    // of the first pass (wide, which waits for its slowest wave at a barrier).
    std::vector<int> tb_of(nr, -1), fo_of(nr, -1);
    for (int s = 0; s < m->ntb; s++) tb_of[m->tb_rxn[s]] = s;
    for (int s = 0; s < m->nfo; s++) fo_of[m->fo_rxn[s]] = s;
    std::vector<int> by_class, perm, inv(nr);
    // The team layout (ns + 1 <= 64) puts the simple records (simple_record) last: its kernel
    // takes them two per lane in a loop of their own.
    const bool team_layout = ns + 1 <= 64;
    int n_simple = 0;
    for (int pass = 3; pass >= 0; pass--)
        for (int j = 0; j < nr; j++) {
            const bool simple = team_layout && simple_record(m, j, tb_of, fo_of);
            const int cls = fo_of[j] >= 0 ? 3 : (tb_of[j] >= 0 ? 2 : (simple ? 0 : 1));
            if (cls == pass) by_class.push_back(j);
            if (cls == pass && pass == 0) n_simple++;
        }
    if (team_layout) {
        perm = by_class;
    } else {
        std::vector<std::pair<int64_t, int>> slot(nr);   // (loop position, reaction)
        for (int i = 0; i < nr; i++) {
            const int64_t b = i / 256, r = i % 256;
            slot[i] = {256 * b + (r % 4) * 64 + r / 4, by_class[i]};
        }
        std::sort(slot.begin(), slot.end());
        for (auto &e : slot) perm.push_back(e.second);
    }
    for (int jn = 0; jn < nr; jn++) inv[perm[jn]] = jn;
    std::vector<double> rA(nr), rB(nr), rE(nr), rdn(nr), fw_nu, bw_nu, rx_nu;
    std::vector<int> rinfo(nr), rtb(nr), rfo(nr), fw_ptr(nr + 1), bw_ptr(nr + 1), rx_ptr(nr + 1), fw_sp, bw_sp, rx_sp;
    std::vector<int> rpl(nr), rch(nr), pl_ptr(m->npl + 1), ch_ptr(m->nch + 1);
    std::vector<double> ch_dat;
    for (int s = 0; s < m->nch; s++) { ch_ptr[s] = (int)ch_dat.size(); cheb_block(m, s, ch_dat); }
    ch_ptr[m->nch] = (int)ch_dat.size();
    std::vector<double> rv((size_t)3 * nr, 0.), pl_lev;
    for (int s = 0; s < m->npl; s++) {
        pl_ptr[s] = (int)(pl_lev.size() / 4);
        for (int q = m->pl_ptr[s]; q < m->pl_ptr[s + 1]; q++) {
            pl_lev.push_back(std::log(m->pl_par[4 * q]));
            for (int i = 1; i < 4; i++) pl_lev.push_back(m->pl_par[4 * q + i]);
        }
    }
    pl_ptr[m->npl] = (int)(pl_lev.size() / 4);
    for (int jn = 0; jn < nr; jn++) {
        const int j = perm[jn];
        rA[jn] = m->lnA[j]; rB[jn] = m->beta[j]; rE[jn] = m->E_over_R[j];
        rinfo[jn] = (m->rflags[j] & 7) | (smooth_forward(m, j) ? 8 : 0) | (smooth_reverse(m, j) ? 16 : 0);
        rpl[jn] = plog_slot_of(m, j);
        rch[jn] = cheb_slot_of(m, j);
        if (m->rflags[j] & LSM_RXN_EXPLICIT_REV)
            for (int i = 0; i < 3; i++) rv[3 * jn + i] = m->rev_arr[3 * j + i];
        rtb[jn] = tb_of[j]; rfo[jn] = fo_of[j];
        std::vector<std::pair<int, double>> re, pr;
        double dn = 0.;
        for (int q = m->reac_ptr[j]; q < m->reac_ptr[j + 1]; q++) { re.push_back({m->reac_sp[q], m->reac_nu[q]}); }
        for (int q = m->prod_ptr[j]; q < m->prod_ptr[j + 1]; q++) { pr.push_back({m->prod_sp[q], m->prod_nu[q]}); }
        for (auto &e : pr) dn += e.second;
        for (auto &e : re) dn -= e.second;
        rdn[jn] = dn;
        auto bysp = [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; };
        std::stable_sort(re.begin(), re.end(), bysp);
        std::stable_sort(pr.begin(), pr.end(), bysp);
        fw_ptr[jn] = (int)fw_sp.size();
        for (auto &e : forward_orders(m, j)) { fw_sp.push_back(e.first); fw_nu.push_back(e.second); }
        bw_ptr[jn] = (int)bw_sp.size();
        for (auto &e : pr) { bw_sp.push_back(e.first); bw_nu.push_back(e.second); }
        std::vector<std::pair<int, double>> mx;
        for (auto &e : re) mx.push_back({e.first, -e.second});
        for (auto &e : pr) mx.push_back(e);
        std::stable_sort(mx.begin(), mx.end(), bysp);
        rx_ptr[jn] = (int)rx_sp.size();
        for (auto &e : mx) { rx_sp.push_back(e.first); rx_nu.push_back(e.second); }
    }
    fw_ptr[nr] = (int)fw_sp.size(); bw_ptr[nr] = (int)bw_sp.size(); rx_ptr[nr] = (int)rx_sp.size();
    // species-major formation list (StoichiometricMap.hpp:656-730 class order)
    std::vector<int> f_ptr(ns + 1), f_rx;
    std::vector<double> f_nu;
    {
        std::vector<std::vector<std::pair<int, double>>> lists(ns);
        for (int side = 0; side < 2; side++)
            for (int cls = 1; cls <= 5; cls++)
                for (int j = 0; j < nr; j++) {
                    const int32_t *ptr = side ? m->prod_ptr : m->reac_ptr;
                    const int32_t *spp = side ? m->prod_sp : m->reac_sp;
                    const double *nup = side ? m->prod_nu : m->reac_nu;
                    for (int q = ptr[j]; q < ptr[j + 1]; q++)
                        if (nu_class(nup[q]) == cls)
                            lists[spp[q]].push_back({inv[j], side ? nup[q] : -nup[q]});
                }
        // per species: reactant classes then product classes, as built above per side
        std::vector<std::vector<std::pair<int, double>>> ordered(ns);
        for (int k = 0; k < ns; k++) {
            for (auto &e : lists[k]) if (e.second < 0) ordered[k].push_back(e);
            for (auto &e : lists[k]) if (e.second > 0) ordered[k].push_back(e);
        }
        for (int k = 0; k < ns; k++) {
            f_ptr[k] = (int)f_rx.size();
            for (auto &e : ordered[k]) { f_rx.push_back(e.first); f_nu.push_back(e.second); }
        }
        f_ptr[ns] = (int)f_rx.size();
    }
    std::vector<int> tb_ptr(m->ntb + 1), tb_sp;
    std::vector<double> tb_eff;
    for (int s = 0; s < m->ntb; s++) {
        tb_ptr[s] = (int)tb_sp.size();
        for (int q = m->tb_ptr[s]; q < m->tb_ptr[s + 1]; q++) { tb_sp.push_back(m->tb_sp[q]); tb_eff.push_back(m->tb_eff_m1[q]); }
    }
    tb_ptr[m->ntb] = (int)tb_sp.size();
    std::vector<int> fo_type(m->nfo), fo_single(m->nfo), fo_ptr(m->nfo + 1), fo_sp;
    std::vector<double> fo_par((size_t)8 * m->nfo, 0.), fo_eff;
    for (int s = 0; s < m->nfo; s++) {
        fo_type[s] = m->fo_type[s];   // form | LSM_FALLOFF_CHEMACT
        fo_single[s] = m->fo_single[s];
        for (int i = 0; i < 3; i++) fo_par[8 * s + i] = m->fo_inf[3 * s + i];
        if ((m->fo_type[s] & 3) == LSM_FALLOFF_SRI) {        // a, b, c, e, d
            const double *t = m->fo_sri + 5 * s;
            fo_par[8 * s + 3] = t[0]; fo_par[8 * s + 4] = t[1]; fo_par[8 * s + 5] = t[2]; fo_par[8 * s + 6] = t[4];
            fo_par[8 * s + 7] = t[3];
        } else {
            for (int i = 0; i < 4; i++) fo_par[8 * s + 3 + i] = m->fo_troe[4 * s + i];
        }
        fo_ptr[s] = (int)fo_sp.size();
        for (int q = m->fo_ptr[s]; q < m->fo_ptr[s + 1]; q++) { fo_sp.push_back(m->fo_sp[q]); fo_eff.push_back(m->fo_eff_m1[q]); }
    }
    fo_ptr[m->nfo] = (int)fo_sp.size();

    Blob b;
    size_t o_sp = b.add(sp), o_rA = b.add(rA), o_rB = b.add(rB), o_rE = b.add(rE), o_rdn = b.add(rdn);
    size_t o_rinfo = b.add(rinfo), o_rtb = b.add(rtb), o_rfo = b.add(rfo);
    size_t o_fwp = b.add(fw_ptr), o_fws = b.add(fw_sp), o_fwn = b.add(fw_nu);
    size_t o_bwp = b.add(bw_ptr), o_bws = b.add(bw_sp), o_bwn = b.add(bw_nu);
    size_t o_rxp = b.add(rx_ptr), o_rxs = b.add(rx_sp), o_rxn = b.add(rx_nu);
    size_t o_tbp = b.add(tb_ptr), o_tbs = b.add(tb_sp), o_tbe = b.add(tb_eff);
    size_t o_fot = b.add(fo_type), o_fos1 = b.add(fo_single), o_fopar = b.add(fo_par);
    size_t o_fop = b.add(fo_ptr), o_fos = b.add(fo_sp), o_foe = b.add(fo_eff);
    size_t o_fp = b.add(f_ptr), o_frx = b.add(f_rx), o_fnu = b.add(f_nu);
    size_t o_rv = b.add(rv), o_rpl = b.add(rpl), o_plptr = b.add(pl_ptr), o_pllev = b.add(pl_lev);
    size_t o_rch = b.add(rch), o_chptr = b.add(ch_ptr), o_chdat = b.add(ch_dat);
    std::vector<char> image;
    const bool img_ok = build_image(m, perm, tb_of, fo_of, d.ns_pad, d.img, image);
    d.img.simple0 = nr - n_simple;
    size_t o_img = img_ok ? b.add(image) : 0;
    HIP_OK(hipMalloc(&h->d_tables, b.bytes.size()));
    h->tables_bytes = b.bytes.size();
    HIP_OK(hipMemcpy(h->d_tables, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice));
    char *base = (char *)h->d_tables;
    auto D = [&](size_t o) { return (const double *)(base + o); };
    auto I = [&](size_t o) { return (const int *)(base + o); };
    d.sp = D(o_sp); d.rA = D(o_rA); d.rB = D(o_rB); d.rE = D(o_rE); d.rdn = D(o_rdn);
    d.rinfo = I(o_rinfo); d.rtb = I(o_rtb); d.rfo = I(o_rfo);
    d.fw_ptr = I(o_fwp); d.fw_sp = I(o_fws); d.fw_nu = D(o_fwn);
    d.bw_ptr = I(o_bwp); d.bw_sp = I(o_bws); d.bw_nu = D(o_bwn);
    d.rx_ptr = I(o_rxp); d.rx_sp = I(o_rxs); d.rx_nu = D(o_rxn);
    d.tb_ptr = I(o_tbp); d.tb_sp = I(o_tbs); d.tb_eff = D(o_tbe);
    d.fo_type = I(o_fot); d.fo_single = I(o_fos1); d.fo_par = D(o_fopar);
    d.fo_ptr = I(o_fop); d.fo_sp = I(o_fos); d.fo_eff = D(o_foe);
    d.f_ptr = I(o_fp); d.f_rx = I(o_frx); d.f_nu = D(o_fnu);
    d.rv = D(o_rv); d.rpl = I(o_rpl); d.pl_ptr = I(o_plptr); d.pl_lev = D(o_pllev);
    d.rch = I(o_rch); d.ch_ptr = I(o_chptr); d.ch_dat = D(o_chdat);
    d.img.src = img_ok ? (const void *)(base + o_img) : nullptr;
    return LSM_OK;
}

extern "C" {

int lsm_create(const lsm_mechanism *mech, const lsm_ode_parameters *ode, const lsm_reactor_options *reactor,
               lsm_handle **out) {
    if (!out) return fail(LSM_E_INVALID, "null output handle");
    *out = nullptr;
    int rc = check_mechanism(mech);
    if (rc) return rc;
    lsm_ode_parameters p;
    if (ode) p = *ode; else lsm_ode_parameters_default(&p);
    lsm_reactor_options r;
    if (reactor) r = *reactor; else lsm_reactor_options_default(&r);
    if (p.type != LSM_ODE_OPENSMOKE_GPU && p.type != LSM_ODE_OPENSMOKE)
        return fail(LSM_E_UNSUPPORTED, "this library integrates with the OpenSMOKE Gear/BDF method only "
                                       "(odeSolver OpenSMOKE_GPU or OpenSMOKE)");
    // MultiValueSolver::SetAbsoluteTolerances / SetRelativeTolerances (:196-214)
    if (!(p.absolute_tolerance >= 1.e-32)) return fail(LSM_E_INVALID, "The user-defined absolute tolerance is too small");
    if (!(p.relative_tolerance >= 1.e-32)) return fail(LSM_E_INVALID, "The user-defined relative tolerance is too small");
    if (strcmp(p.linear_algebra, "Eigen") != 0 && p.linear_algebra[0] != '\0')
        return fail(LSM_E_UNSUPPORTED, "Requested linear algebra is not supported!");   // KernelDense.hpp:73-81
    if (!(r.pressure > 0.)) return fail(LSM_E_INVALID, "pressure must be positive");
    lsm_handle *h = new lsm_handle();
    h->ode = p;
    h->reactor = r;
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return fail(LSM_E_HIP, "no HIP device"); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) { delete h; return fail(LSM_E_HIP, "hipGetDeviceProperties failed"); }
    h->num_cu = prop.multiProcessorCount;
    rc = upload_tables(h, mech);
    if (rc) { lsm_destroy(h); return rc; }
    if (hipMalloc(&h->d_queue, 256) != hipSuccess) { lsm_destroy(h); return fail(LSM_E_HIP, "hipMalloc queue"); }
    KParams &k = h->kp;
    gear_r(k.bk.r);
    k.bk.abs_tol = p.absolute_tolerance;
    k.bk.rel_tol = p.relative_tolerance;
    k.bk.t0 = 0.;                                // set per call (lsm_integrate_batch*)
    k.bk.max_steps = p.maximum_number_of_steps > 0 ? (unsigned)p.maximum_number_of_steps : 500000u;
    k.pressure = r.pressure;
    k.min_temperature = r.min_temperature;
    k.energy = r.energy_equation ? 1 : 0;
    k.full_pivoting = p.full_pivoting ? 1 : 0;
    k.jac_mode = LSM_JACOBIAN_NUMERICAL;
    k.refine = 4;
    k.const_volume = r.constant_volume ? 1 : 0;
    k.compact = r.strang_compact ? 1 : 0;
    k.rho = nullptr;
    k.phase_cycles = nullptr;
    k.fp64_fallbacks = nullptr;
#ifdef LSM_PHASE_TIMERS
    if (hipMalloc(&k.phase_cycles, 12 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(k.phase_cycles, 0, 12 * sizeof(unsigned long long)) != hipSuccess) {
        lsm_destroy(h);
        return fail(LSM_E_HIP, "phase timer buffer");
    }
#endif
    // persistent team grid: one cell per wave, the waves of a workgroup sharing the LDS copy
    // of the mechanism; as many waves per CU as the kernel's register cap allows (its LU rows
    // live in registers), in the workgroup shape that leaves the widest Jacobian tile.
    const size_t lds_block = std::max<size_t>(prop.sharedMemPerBlock, 65536);
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor > 0 ? prop.maxSharedMemoryPerMultiProcessor : lds_block;
    h->lds_cu = std::min(lds_block, lds_cu);
    team_shape_default(h);
    // wide grid: one 256-thread workgroup per CU (its FP32 LU takes most of the LDS)
    if (h->dm.img.wide_ok && h->dm.ne <= WIDE_MAX_NEQ && wide_lds_bytes(h->dm) <= std::min(lds_block, lds_cu)) {
        h->wide_grid = h->num_cu;
        if (hipMalloc(&h->d_wjws, sizeof(double) * wide_jac_doubles(h->dm) * (size_t)h->wide_grid) != hipSuccess ||
            hipMalloc(&h->d_fallbacks, sizeof(unsigned long long)) != hipSuccess ||
            hipMemset(h->d_fallbacks, 0, sizeof(unsigned long long)) != hipSuccess) {
            lsm_destroy(h);
            return fail(LSM_E_NOMEM, "wide Jacobian workspace");
        }
        k.fp64_fallbacks = h->d_fallbacks;
    }
    *out = h;
    return LSM_OK;
}

int lsm_set_method(lsm_handle *h, int32_t method) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (method < LSM_METHOD_AUTO || method > LSM_METHOD_WIDE) return fail(LSM_E_INVALID, "unknown method");
    if (method == LSM_METHOD_TEAM && h->team_wpb < 1)
        return fail(LSM_E_UNSUPPORTED, "team kernel needs ns+1 <= 64 and a mechanism image that fits in LDS");
    if (method == LSM_METHOD_WIDE && h->wide_grid < 1)
        return fail(LSM_E_UNSUPPORTED, "wide kernel needs ns+1 <= 176");
    h->method = method;
    return LSM_OK;
}

int lsm_set_cells_per_wave(lsm_handle *h, int32_t cells) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (cells != 1 && cells != 2 && cells != 4) return fail(LSM_E_INVALID, "cells per wavefront must be 1, 2 or 4");
    if (h->team_wpb < 1)
        return fail(LSM_E_UNSUPPORTED, "team kernel needs ns+1 <= 64 and a mechanism image that fits in LDS");
    if (cells > team_max_cells_per_wave(h->dm))
        return fail(LSM_E_UNSUPPORTED, cells == 4 ? "4 cells per wavefront need ns+1 <= 16"
                                                  : "2 cells per wavefront need ns+1 <= 32");
    if (cells > 1 && !pack_shape_default(h, cells))
        return fail(LSM_E_UNSUPPORTED, "packed team kernel: the per-cell LDS does not fit");
    h->team_cpw = cells;
    return LSM_OK;
}

int lsm_get_cells_per_wave(lsm_handle *h, int32_t *cells) {
    if (!h || !cells) return fail(LSM_E_INVALID, "null argument");
    *cells = h->team_cpw;
    return LSM_OK;
}

int lsm_set_jacobian(lsm_handle *h, int32_t type) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (type != LSM_JACOBIAN_NUMERICAL && type != LSM_JACOBIAN_ANALYTIC) return fail(LSM_E_INVALID, "unknown Jacobian type");
    if (type == LSM_JACOBIAN_ANALYTIC && h->kp.const_volume)
        return fail(LSM_E_UNSUPPORTED, "the analytic Jacobian is built for the constant-pressure reactor; "
                                       "constant-volume handles use the numerical Jacobian");
    h->kp.jac_mode = type;
    return LSM_OK;
}

int lsm_set_refinement(lsm_handle *h, int32_t sweeps) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (sweeps < 0 || sweeps > 8) return fail(LSM_E_INVALID, "refinement sweeps must be 0-8");
    h->kp.refine = sweeps;
    return LSM_OK;
}

int lsm_destroy(lsm_handle *h) {
    if (!h) return LSM_OK;
    if (h->d_keys) (void)hipFree(h->d_keys);
    if (h->d_keys_out) (void)hipFree(h->d_keys_out);
    if (h->d_vals) (void)hipFree(h->d_vals);
    if (h->d_sort_tmp) (void)hipFree(h->d_sort_tmp);
    if (h->d_wjws) (void)hipFree(h->d_wjws);
    if (h->d_fallbacks) (void)hipFree(h->d_fallbacks);
    if (h->d_tables) (void)hipFree(h->d_tables);
    if (h->d_queue) (void)hipFree(h->d_queue);
    if (h->d_work) (void)hipFree(h->d_work);
    if (h->kp.phase_cycles) (void)hipFree(h->kp.phase_cycles);
    if (h->d_y) (void)hipFree(h->d_y);
    if (h->d_dt) (void)hipFree(h->d_dt);
    if (h->d_q) (void)hipFree(h->d_q);
    if (h->d_status) (void)hipFree(h->d_status);
    if (h->d_cnt) (void)hipFree(h->d_cnt);
    if (h->d_transport) (void)hipFree(h->d_transport);
    if (h->d_transport_opts) (void)hipFree(h->d_transport_opts);
    delete h;
    return LSM_OK;
}

static int pick_method(const lsm_handle *h) {
    if (h->method != LSM_METHOD_AUTO) return h->method;
    if (h->team_wpb > 0) return LSM_METHOD_TEAM;
    return h->wide_grid > 0 ? LSM_METHOD_WIDE : LSM_METHOD_PLAIN;
}

static int integrate_device(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                            const double *d_rho, const double *d_y_in, double *d_y_out, int32_t *d_status,
                            double *d_q, int32_t *d_counters, const int32_t *d_order, void *stream);

int lsm_integrate_batch_device(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                               const double *d_y_in, double *d_y_out, int32_t *d_status, double *d_q,
                               int32_t *d_counters, void *stream) {
    return integrate_device(h, n, t0, pressure, d_dt, nullptr, d_y_in, d_y_out, d_status, d_q, d_counters, nullptr,
                            stream);
}

int lsm_integrate_batch_device_ordered(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                                       const double *d_y_in, double *d_y_out, int32_t *d_status, double *d_q,
                                       int32_t *d_counters, const int32_t *d_order, void *stream) {
    if (!d_order) return fail(LSM_E_INVALID, "null order");
    return integrate_device(h, n, t0, pressure, d_dt, nullptr, d_y_in, d_y_out, d_status, d_q, d_counters, d_order,
                            stream);
}

int lsm_integrate_batch_device_rho(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                                   const double *d_rho, const double *d_y_in, double *d_y_out, int32_t *d_status,
                                   double *d_q, int32_t *d_counters, const int32_t *d_order, void *stream) {
    return integrate_device(h, n, t0, pressure, d_dt, d_rho, d_y_in, d_y_out, d_status, d_q, d_counters, d_order,
                            stream);
}

// Per-call start time and pressure of a chemistry step (chemistry_DI.H:134,138,157).
static int check_step_args(double t0, double pressure) {
    if (!std::isfinite(t0)) return fail(LSM_E_INVALID, "t0 must be finite");
    if (!(pressure > 0.) || !std::isfinite(pressure)) return fail(LSM_E_INVALID, "pressure must be positive");
    return LSM_OK;
}

// Device memory the kernels of one chemistry-step call use besides its inputs and outputs.
struct StepScratch {
    unsigned *queue = nullptr;                 // work-queue counter (team, wide)
    double *wide_ws = nullptr;                 // wide: [wide_grid][2][ne][ne] Jacobian, FP64 LU
    int wide_grid = 0;
    unsigned long long *fallbacks = nullptr;   // wide: FP64 refactorisation count, or null
    double *plain_ws = nullptr;                // plain: [plain_slots] per-cell workspaces
    int64_t plain_slots = 0;
};

// Checks of a chemistry-step call and its kernel parameters; *skip when there is nothing to do.
static int prepare_step(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                        const double *d_rho, const double *d_y_in, double *d_y_out, KParams &kp, int &method,
                        bool &skip) {
    skip = false;
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (int rc = check_step_args(t0, pressure)) return rc;
    if (n == 0) { skip = true; return LSM_OK; }
    if (!d_dt || !d_y_in || !d_y_out) return fail(LSM_E_INVALID, "null dt / y pointer");
    if (n > (int64_t)0x7fffffff) return fail(LSM_E_INVALID, "more than 2^31-1 cells in one call");
    method = pick_method(h);
    kp = h->kp;                                  // per call: start time, pressure, constant-volume densities
    kp.bk.t0 = t0;
    kp.pressure = pressure;
    kp.rho = h->kp.const_volume ? d_rho : nullptr;
    return LSM_OK;
}

// The kernels of one call, enqueued on `stream` with the given scratch.
static int launch_step(lsm_handle *h, int method, const KParams &kp_in, int64_t n, const double *d_dt,
                       const double *d_y_in, double *d_y_out, int32_t *d_status, double *d_q, int32_t *d_counters,
                       const int32_t *d_order, const StepScratch &sc, void *stream) {
    KParams kp = kp_in;
    if (method == LSM_METHOD_TEAM) {
        if (launch_reset_queue(sc.queue, stream)) return fail(LSM_E_HIP, "queue reset launch failed");
        const bool packed = h->team_cpw > 1;
        int rc = launch_team_cpw(h->dm, kp, n, d_dt, d_y_in, d_y_out, d_status, d_q, d_counters, sc.queue, d_order,
                                 packed ? h->pack_grid : h->team_grid, packed ? h->pack_wpb : h->team_wpb,
                                 h->team_cpw, stream);
        if (rc) return fail(LSM_E_HIP, std::string("team kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
        return LSM_OK;
    }
    if (method == LSM_METHOD_WIDE) {
        kp.fp64_fallbacks = sc.fallbacks;
        if (launch_reset_queue(sc.queue, stream)) return fail(LSM_E_HIP, "queue reset launch failed");
        int rc = launch_wide(h->dm, kp, n, d_dt, d_y_in, d_y_out, d_status, d_q, d_counters, sc.queue, d_order,
                             sc.wide_grid, sc.wide_ws, stream);
        if (rc) return fail(LSM_E_HIP, std::string("wide kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
        return LSM_OK;
    }
    int rc = launch_plain(h->dm, kp, n, d_dt, d_y_in, d_y_out, d_status, d_q, d_counters, sc.plain_ws, sc.plain_slots,
                          stream);
    if (rc) return fail(LSM_E_HIP, std::string("plain kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    return LSM_OK;
}

static int integrate_device(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                            const double *d_rho, const double *d_y_in, double *d_y_out, int32_t *d_status,
                            double *d_q, int32_t *d_counters, const int32_t *d_order, void *stream) {
    KParams kp;
    int method = 0;
    bool skip = false;
    if (int rc = prepare_step(h, n, t0, pressure, d_dt, d_rho, d_y_in, d_y_out, kp, method, skip)) return rc;
    if (skip) return LSM_OK;
    StepScratch sc;
    sc.queue = h->d_queue;
    sc.wide_ws = h->d_wjws;
    sc.wide_grid = h->wide_grid;
    sc.fallbacks = h->d_fallbacks;
    if (method == LSM_METHOD_PLAIN) {
        // one thread per cell slot, workspace in HBM, grown on demand
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const size_t per = plain_work_doubles(h->dm, 1) * sizeof(double);
        int64_t slots = std::min<int64_t>(n, (int64_t)h->num_cu * 1024);
        const size_t budget = std::min<size_t>((size_t)32 << 30, free_b / 4 + (h->d_work ? (size_t)h->work_slots * per : 0));
        slots = std::max<int64_t>(1, std::min<int64_t>(slots, (int64_t)(budget / per)));
        if (slots > h->work_slots) {
            if (h->d_work) { HIP_OK(hipStreamSynchronize((hipStream_t)stream)); (void)hipFree(h->d_work); h->d_work = nullptr; }
            HIP_OK(hipMalloc(&h->d_work, (size_t)slots * per));
            h->work_slots = slots;
        }
        sc.plain_ws = h->d_work;
        sc.plain_slots = h->work_slots;
    }
    return launch_step(h, method, kp, n, d_dt, d_y_in, d_y_out, d_status, d_q, d_counters, d_order, sc, stream);
}

// ---- Chemistry step that allocates nothing ----------------------------------------------------
// The scratch of lsm_integrate_batch_device_noalloc, one copy per GPU, held until the process
// ends: the work-queue counter in its first 256 bytes, then the wide kernel's per-workgroup
// Jacobian / FP64 LU workspaces or the plain kernel's per-cell-slot workspaces.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in, not a device
// address): the address of each device's copy comes from hipGetSymbolAddress, once per
// device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[LSM_NOALLOC_SCRATCH_BYTES];

enum { NOALLOC_MAX_DEVICES = 64, NOALLOC_QUEUE_BYTES = 256 };
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `dev` (the current device); nullptr when it cannot be had.
static unsigned char *scratch_on_device(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[dev] = (unsigned char *)ptr;
    }
    return g_scratch_ptr[dev];
}

int lsm_integrate_batch_device_noalloc(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt,
                                       const double *d_y_in, double *d_y_out, int32_t *d_status, double *d_q,
                                       int32_t *d_counters, const int32_t *d_order, void *stream) {
    KParams kp;
    int method = 0;
    bool skip = false;
    if (int rc = prepare_step(h, n, t0, pressure, d_dt, nullptr, d_y_in, d_y_out, kp, method, skip)) return rc;
    if (skip) return LSM_OK;
    if (h->kp.const_volume) return fail(LSM_E_UNSUPPORTED, "constant-volume handles need per-cell densities: "
                                                           "use lsm_integrate_batch_device_rho");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail(LSM_E_HIP, "hipGetDevice failed");
    if (dev != h->device) return fail(LSM_E_INVALID, "the current device is not the handle's device");
    unsigned char *base = scratch_on_device(dev);
    if (!base) return fail(LSM_E_HIP, "hipGetSymbolAddress of the no-allocation scratch failed");
    const size_t room = (size_t)LSM_NOALLOC_SCRATCH_BYTES - NOALLOC_QUEUE_BYTES;
    StepScratch sc;
    sc.queue = (unsigned *)base;
    if (method == LSM_METHOD_WIDE) {
        const size_t per = sizeof(double) * wide_jac_doubles(h->dm);
        const int64_t fit = (int64_t)(room / per);
        if (fit < 1) return fail(LSM_E_INVALID, "one wide workgroup's workspace exceeds the no-allocation scratch");
        sc.wide_ws = (double *)(base + NOALLOC_QUEUE_BYTES);
        sc.wide_grid = (int)std::min<int64_t>(h->wide_grid, fit);
        sc.fallbacks = nullptr;                  // not counted: lsm_fp64_fallbacks reports the other entry points
    } else if (method == LSM_METHOD_PLAIN) {
        const size_t per = plain_work_doubles(h->dm, 1) * sizeof(double);
        const int64_t fit = (int64_t)(room / per);
        if (fit < 1) return fail(LSM_E_INVALID, "one plain-kernel workspace exceeds the no-allocation scratch");
        sc.plain_ws = (double *)(base + NOALLOC_QUEUE_BYTES);
        sc.plain_slots = std::min<int64_t>(std::min<int64_t>(n, (int64_t)h->num_cu * 1024), fit);
    }
    return launch_step(h, method, kp, n, d_dt, d_y_in, d_y_out, d_status, d_q, d_counters, d_order, sc, stream);
}

int lsm_integrate_batch(lsm_handle *h, int64_t n, double t0, double pressure, const double *dt, const double *y_in,
                        double *y_out, int32_t *status, double *q, int32_t *counters) {
    return lsm_integrate_batch_rho(h, n, t0, pressure, dt, nullptr, y_in, y_out, status, q, counters);
}

int lsm_integrate_batch_rho(lsm_handle *h, int64_t n, double t0, double pressure, const double *dt, const double *rho,
                            const double *y_in, double *y_out, int32_t *status, double *q, int32_t *counters) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (int rc = check_step_args(t0, pressure)) return rc;
    if (n == 0) return LSM_OK;
    if (!dt || !y_in || !y_out) return fail(LSM_E_INVALID, "null dt / y pointer");
    const int ne = h->dm.ne;
    if (n > h->cap) {
        if (h->d_y) (void)hipFree(h->d_y);
        if (h->d_dt) (void)hipFree(h->d_dt);
        if (h->d_q) (void)hipFree(h->d_q);
        if (h->d_status) (void)hipFree(h->d_status);
        if (h->d_cnt) (void)hipFree(h->d_cnt);
        h->d_y = h->d_dt = h->d_q = nullptr;
        h->d_status = h->d_cnt = nullptr;
        h->cap = 0;
        HIP_OK(hipMalloc(&h->d_y, sizeof(double) * ne * n));
        HIP_OK(hipMalloc(&h->d_dt, sizeof(double) * n));
        HIP_OK(hipMalloc(&h->d_q, sizeof(double) * n));
        HIP_OK(hipMalloc(&h->d_status, sizeof(int) * n));
        HIP_OK(hipMalloc(&h->d_cnt, sizeof(int) * LSM_NCOUNTERS * n));
        h->cap = n;
    }
    HIP_OK(hipMemcpy(h->d_y, y_in, sizeof(double) * ne * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h->d_dt, dt, sizeof(double) * n, hipMemcpyHostToDevice));
    if (q) HIP_OK(hipMemcpy(h->d_q, q, sizeof(double) * n, hipMemcpyHostToDevice));
    double *d_rho = nullptr;
    if (rho && h->kp.const_volume) {
        HIP_OK(hipMalloc(&d_rho, sizeof(double) * n));
        if (hipMemcpy(d_rho, rho, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d_rho);
            return fail(LSM_E_HIP, "copy rho");
        }
    }
    int rc = integrate_device(h, n, t0, pressure, h->d_dt, d_rho, h->d_y, h->d_y, status ? h->d_status : nullptr,
                              q ? h->d_q : nullptr, counters ? h->d_cnt : nullptr, nullptr, nullptr);
    if (rc) { if (d_rho) (void)hipFree(d_rho); return rc; }
    const hipError_t sync = hipDeviceSynchronize();
    if (d_rho) (void)hipFree(d_rho);
    HIP_OK(sync);
    HIP_OK(hipMemcpy(y_out, h->d_y, sizeof(double) * ne * n, hipMemcpyDeviceToHost));
    if (status) HIP_OK(hipMemcpy(status, h->d_status, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (q) HIP_OK(hipMemcpy(q, h->d_q, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (counters) HIP_OK(hipMemcpy(counters, h->d_cnt, sizeof(int) * LSM_NCOUNTERS * n, hipMemcpyDeviceToHost));
    return LSM_OK;
}

int lsm_properties_device(lsm_handle *h, int64_t n, const double *d_y, const double *d_p, double *d_props, double *d_x,
                          double *d_cpk, void *stream) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0) return LSM_OK;
    if (!d_y || !d_props) return fail(LSM_E_INVALID, "null y / props pointer");
    if (launch_properties(h->dm, h->kp.pressure, n, d_y, d_p, d_props, d_x, d_cpk, h->num_cu, stream))
        return fail(LSM_E_HIP, std::string("properties kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    return LSM_OK;
}

int lsm_properties(lsm_handle *h, int64_t n, const double *y, const double *p, double *props, double *x, double *cpk) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0) return LSM_OK;
    if (!y || !props) return fail(LSM_E_INVALID, "null y / props pointer");
    const size_t ns = (size_t)h->dm.ns;
    double *dy = nullptr, *dp = nullptr, *dout = nullptr, *dx = nullptr, *dc = nullptr;
    auto release = [&]() {
        for (double *q : {dy, dp, dout, dx, dc}) if (q) (void)hipFree(q);
    };
    int rc = LSM_OK;
    if (hipMalloc(&dy, sizeof(double) * (ns + 1) * n) != hipSuccess ||
        hipMalloc(&dout, sizeof(double) * LSM_NPROPS * n) != hipSuccess ||
        (p && hipMalloc(&dp, sizeof(double) * n) != hipSuccess) ||
        (x && hipMalloc(&dx, sizeof(double) * ns * n) != hipSuccess) ||
        (cpk && hipMalloc(&dc, sizeof(double) * ns * n) != hipSuccess)) {
        release();
        return fail(LSM_E_NOMEM, "properties buffers");
    }
    if (hipMemcpy(dy, y, sizeof(double) * (ns + 1) * n, hipMemcpyHostToDevice) != hipSuccess ||
        (p && hipMemcpy(dp, p, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(LSM_E_HIP, "copy inputs");
    if (!rc) rc = lsm_properties_device(h, n, dy, dp, dout, dx, dc, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(LSM_E_HIP, "properties kernel failed");
    if (!rc && hipMemcpy(props, dout, sizeof(double) * LSM_NPROPS * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy props");
    if (!rc && x && hipMemcpy(x, dx, sizeof(double) * ns * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy x");
    if (!rc && cpk && hipMemcpy(cpk, dc, sizeof(double) * ns * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy cp_species");
    release();
    return rc;
}

int lsm_set_transport(lsm_handle *h, const lsm_transport *t) {
    if (!h || !t) return fail(LSM_E_INVALID, "null handle / transport");
    const int ns = h->dm.ns;
    if (t->ns != ns) return fail(LSM_E_INVALID, "transport ns differs from the mechanism's");
    if (!t->mw || !t->lambda_fit || !t->eta_fit || (ns > 1 && !t->gamma_fit))
        return fail(LSM_E_INVALID, "transport table pointer is null");
    if (t->nteta < 0 || t->nteta > ns || (t->nteta > 0 && (!t->teta_sp || !t->teta_fit)))
        return fail(LSM_E_INVALID, "thermal diffusion ratio tables");
    if (t->viscosity_model != LSM_VISCOSITY_WILKE && t->viscosity_model != LSM_VISCOSITY_HERNING &&
        t->viscosity_model != LSM_VISCOSITY_MATHUR_SAXENA)
        return fail(LSM_E_INVALID, "viscosity model");
    if ((size_t)3 * ns * sizeof(double) * 4 > 65536) return fail(LSM_E_UNSUPPORTED, "transport needs ns <= 682");
    for (int k = 0; k < ns; k++)
        if (!(t->mw[k] > 0.)) return fail(LSM_E_INVALID, "transport molecular weights must be positive");
    std::vector<int> slot(ns, -1);
    for (int q = 0; q < t->nteta; q++) {
        const int k = t->teta_sp[q];
        if (k < 0 || k >= ns || slot[k] >= 0) return fail(LSM_E_INVALID, "thermal diffusion species index");
        slot[k] = q;
    }
    const size_t np = (size_t)ns * (ns - 1) / 2;
    // Wilke constants (TransportPropertiesMap_CHEMKIN::CompleteInitialization, :655-670)
    std::vector<double> sup(np), inf(np), r14(np);
    {
        const double *M = t->mw;
        size_t i = 0;
        for (int k = 0; k < ns; k++)
            for (int j = k + 1; j < ns; j++, i++) {
                sup[i] = 1. / std::sqrt(8. * (1. + M[k] / M[j]));
                inf[i] = 1. / std::sqrt(8. * (1. + M[j] / M[k]));
                r14[i] = std::sqrt(std::sqrt(M[j] / M[k]));
            }
    }
    // The same constants as full partner-major matrices for the kernel's branch-free loop:
    // species k sums x_q ww[q][k] (1 + e)^2 with e = sqrt(eta_k / eta_q) wr[q][k], which is
    // d = sqrt(eta_k/eta_j) r14 for a later partner and 1/d for an earlier one (:988-993).
    std::vector<double> ww((size_t)ns * ns, 0.), wr((size_t)ns * ns, 0.);
    {
        size_t i = 0;
        for (int k = 0; k < ns; k++)
            for (int j = k + 1; j < ns; j++, i++) {
                ww[(size_t)j * ns + k] = sup[i];          // species k, later partner j
                wr[(size_t)j * ns + k] = r14[i];
                ww[(size_t)k * ns + j] = inf[i];          // species j, earlier partner k
                wr[(size_t)k * ns + j] = 1. / r14[i];
            }
    }
    Blob b;
    const size_t o_mw = b.add(std::vector<double>(t->mw, t->mw + ns));
    const size_t o_l = b.add(std::vector<double>(t->lambda_fit, t->lambda_fit + 4 * ns));
    const size_t o_e = b.add(std::vector<double>(t->eta_fit, t->eta_fit + 4 * ns));
    const size_t o_g = b.add(np ? std::vector<double>(t->gamma_fit, t->gamma_fit + 4 * np) : std::vector<double>());
    const size_t o_s = b.add(sup), o_i = b.add(inf), o_r = b.add(r14), o_sl = b.add(slot);
    const size_t o_ww = b.add(ww), o_wr = b.add(wr);
    const size_t o_t = b.add(t->nteta ? std::vector<double>(t->teta_fit, t->teta_fit + (size_t)4 * t->nteta * ns)
                                      : std::vector<double>());
    void *d = nullptr;
    if (hipMalloc(&d, b.bytes.size()) != hipSuccess) return fail(LSM_E_NOMEM, "transport tables");
    if (hipMemcpy(d, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return fail(LSM_E_HIP, "copy transport tables");
    }
    if (h->d_transport) (void)hipFree(h->d_transport);
    h->d_transport = d;
    h->transport_bytes = b.bytes.size();
    char *base = (char *)d;
    DevTransport &T = h->tr;
    T.ns = ns; T.nteta = t->nteta; T.model = t->viscosity_model;
    T.mw = (const double *)(base + o_mw); T.lam = (const double *)(base + o_l); T.eta = (const double *)(base + o_e);
    T.gam = (const double *)(base + o_g);
    T.wsup = (const double *)(base + o_s); T.winf = (const double *)(base + o_i); T.wr14 = (const double *)(base + o_r);
    T.teta_slot = (const int *)(base + o_sl); T.teta = (const double *)(base + o_t);
    T.ww = (const double *)(base + o_ww); T.wr = (const double *)(base + o_wr);
    return LSM_OK;
}

int lsm_set_transport_options(lsm_handle *h, const lsm_transport_options *o) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    DevTransport &T = h->tr;
    if (!o) {
        if (h->d_transport_opts) (void)hipFree(h->d_transport_opts);
        h->d_transport_opts = nullptr;
        h->transport_opts_bytes = 0;
        T.dmodel = LSM_DIFFUSIVITY_MULTICOMPONENT; T.simplified = 0; T.soret = 0; T.soot_ref = -1;
        T.lewis = nullptr; T.soot_slot = nullptr; T.soot_corr = nullptr;
        return LSM_OK;
    }
    const int ns = h->dm.ns;
    if (o->diffusivity_model != LSM_DIFFUSIVITY_MULTICOMPONENT && o->diffusivity_model != LSM_DIFFUSIVITY_LEWIS_NUMBERS)
        return fail(LSM_E_INVALID, "diffusivity model");
    const bool lewis = o->diffusivity_model == LSM_DIFFUSIVITY_LEWIS_NUMBERS;
    if (lewis) {
        if (!o->lewis) return fail(LSM_E_INVALID, "Lewis numbers pointer is null");
        for (int k = 0; k < ns; k++)
            if (!(o->lewis[k] > 0.) || !std::isfinite(o->lewis[k]))
                return fail(LSM_E_INVALID, "Lewis numbers must be positive and finite");
    }
    std::vector<int> slot(ns, -1);
    const bool soot = o->soot_reference >= 0;
    if (soot) {
        if (o->soot_reference >= ns) return fail(LSM_E_INVALID, "soot reference species index");
        if (o->n_soot < 0 || o->n_soot > ns || (o->n_soot > 0 && (!o->soot_sp || !o->soot_corr)))
            return fail(LSM_E_INVALID, "soot correction tables");
        for (int i = 0; i < o->n_soot; i++) {
            const int k = o->soot_sp[i];
            if (k < 0 || k >= ns || slot[k] >= 0) return fail(LSM_E_INVALID, "soot species index");
            if (!std::isfinite(o->soot_corr[i])) return fail(LSM_E_INVALID, "soot correction factor");
            slot[k] = i;
        }
    } else if (o->soot_reference != -1 || o->n_soot != 0) {
        return fail(LSM_E_INVALID, "soot corrections need a reference species");
    }
    Blob b;
    const size_t o_l = b.add(lewis ? std::vector<double>(o->lewis, o->lewis + ns) : std::vector<double>());
    const size_t o_s = b.add(soot ? slot : std::vector<int>());
    const size_t o_c = b.add(soot && o->n_soot ? std::vector<double>(o->soot_corr, o->soot_corr + o->n_soot)
                                               : std::vector<double>());
    void *d = nullptr;
    if (b.bytes.size()) {
        if (hipMalloc(&d, b.bytes.size()) != hipSuccess) return fail(LSM_E_NOMEM, "transport option tables");
        if (hipMemcpy(d, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            return fail(LSM_E_HIP, "copy transport option tables");
        }
    }
    if (h->d_transport_opts) (void)hipFree(h->d_transport_opts);
    h->d_transport_opts = d;
    h->transport_opts_bytes = d ? b.bytes.size() : 0;
    char *base = (char *)d;
    T.dmodel = o->diffusivity_model;
    T.simplified = o->simplified ? 1 : 0;
    T.soret = o->soret ? 1 : 0;
    T.soot_ref = soot ? o->soot_reference : -1;
    T.lewis = lewis ? (const double *)(base + o_l) : nullptr;
    T.soot_slot = soot ? (const int *)(base + o_s) : nullptr;
    T.soot_corr = soot && o->n_soot ? (const double *)(base + o_c) : nullptr;
    return LSM_OK;
}

int lsm_transport_properties_device(lsm_handle *h, int64_t n, const double *d_y, const double *d_p, double *d_mu,
                                    double *d_lambda, double *d_dmix, double *d_teta, void *stream) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (!h->d_transport) return fail(LSM_E_INVALID, "no transport fits (lsm_set_transport)");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0) return LSM_OK;
    if (!d_y || !d_mu || !d_lambda) return fail(LSM_E_INVALID, "null y / mu / lambda pointer");
    if (launch_transport(h->dm, h->tr, h->kp.pressure, n, d_y, d_p, d_mu, d_lambda, d_dmix, d_teta, h->num_cu,
                         stream))
        return fail(LSM_E_HIP, std::string("transport kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    return LSM_OK;
}

int lsm_transport_properties(lsm_handle *h, int64_t n, const double *y, const double *p, double *mu, double *lambda,
                             double *dmix, double *teta) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (!h->d_transport) return fail(LSM_E_INVALID, "no transport fits (lsm_set_transport)");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0) return LSM_OK;
    if (!y || !mu || !lambda) return fail(LSM_E_INVALID, "null y / mu / lambda pointer");
    const size_t ns = (size_t)h->dm.ns;
    double *dy = nullptr, *dp = nullptr, *dmu = nullptr, *dl = nullptr, *dd = nullptr, *dt = nullptr;
    auto release = [&]() {
        for (double *q : {dy, dp, dmu, dl, dd, dt}) if (q) (void)hipFree(q);
    };
    int rc = LSM_OK;
    if (hipMalloc(&dy, sizeof(double) * (ns + 1) * n) != hipSuccess || hipMalloc(&dmu, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&dl, sizeof(double) * n) != hipSuccess || (p && hipMalloc(&dp, sizeof(double) * n) != hipSuccess) ||
        (dmix && hipMalloc(&dd, sizeof(double) * ns * n) != hipSuccess) ||
        (teta && hipMalloc(&dt, sizeof(double) * ns * n) != hipSuccess)) {
        release();
        return fail(LSM_E_NOMEM, "transport buffers");
    }
    if (hipMemcpy(dy, y, sizeof(double) * (ns + 1) * n, hipMemcpyHostToDevice) != hipSuccess ||
        (p && hipMemcpy(dp, p, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(LSM_E_HIP, "copy inputs");
    if (!rc) rc = lsm_transport_properties_device(h, n, dy, dp, dmu, dl, dd, dt, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(LSM_E_HIP, "transport kernel failed");
    if (!rc && (hipMemcpy(mu, dmu, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(lambda, dl, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess))
        rc = fail(LSM_E_HIP, "copy mu / lambda");
    if (!rc && dmix && hipMemcpy(dmix, dd, sizeof(double) * ns * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy dmix");
    if (!rc && teta && hipMemcpy(teta, dt, sizeof(double) * ns * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy teta");
    release();
    return rc;
}

int lsm_equations_device(lsm_handle *h, int64_t n, const double *d_y, double *d_f, double *d_jac, void *stream) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0 || (!d_f && !d_jac)) return LSM_OK;
    if (!d_y) return fail(LSM_E_INVALID, "null y pointer");
    if (h->team_wpb < 1) return fail(LSM_E_UNSUPPORTED, "equations need ns+1 <= 64 and a mechanism image that fits in LDS");
    if (d_jac && h->kp.const_volume)
        return fail(LSM_E_UNSUPPORTED, "the analytic Jacobian is built for the constant-pressure reactor");
    if (launch_equations(h->dm, h->kp, n, d_y, d_f, d_jac, h->team_grid, h->team_wpb, stream))
        return fail(LSM_E_HIP, std::string("equations kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    return LSM_OK;
}

int lsm_equations(lsm_handle *h, int64_t n, const double *y, double *f, double *jac) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0) return fail(LSM_E_INVALID, "negative cell count");
    if (n == 0 || (!f && !jac)) return LSM_OK;
    if (!y) return fail(LSM_E_INVALID, "null y pointer");
    const size_t ne = (size_t)h->dm.ne;
    double *dy = nullptr, *df = nullptr, *dj = nullptr;
    auto release = [&]() {
        if (dy) (void)hipFree(dy);
        if (df) (void)hipFree(df);
        if (dj) (void)hipFree(dj);
    };
    int rc = LSM_OK;
    if (hipMalloc(&dy, sizeof(double) * ne * n) != hipSuccess || (f && hipMalloc(&df, sizeof(double) * ne * n) != hipSuccess) ||
        (jac && hipMalloc(&dj, sizeof(double) * ne * ne * n) != hipSuccess)) {
        release();
        return fail(LSM_E_NOMEM, "equations buffers");
    }
    if (hipMemcpy(dy, y, sizeof(double) * ne * n, hipMemcpyHostToDevice) != hipSuccess) rc = fail(LSM_E_HIP, "copy y");
    if (!rc) rc = lsm_equations_device(h, n, dy, df, dj, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(LSM_E_HIP, "equations kernel failed");
    if (!rc && f && hipMemcpy(f, df, sizeof(double) * ne * n, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(LSM_E_HIP, "copy f");
    if (!rc && jac && hipMemcpy(jac, dj, sizeof(double) * ne * ne * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LSM_E_HIP, "copy J");
    release();
    return rc;
}

int lsm_cost_order(lsm_handle *h, int64_t n, const int32_t *d_counters, int32_t *d_order, void *stream) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n < 0 || n > (int64_t)0x7fffffff) return fail(LSM_E_INVALID, "cell count out of range");
    if (n == 0) return LSM_OK;
    if (!d_counters || !d_order) return fail(LSM_E_INVALID, "null counters / order");
    hipStream_t s = (hipStream_t)stream;
    if (n > h->sort_cap) {
        if (h->d_keys) (void)hipFree(h->d_keys);
        if (h->d_keys_out) (void)hipFree(h->d_keys_out);
        if (h->d_vals) (void)hipFree(h->d_vals);
        if (h->d_sort_tmp) (void)hipFree(h->d_sort_tmp);
        h->d_keys = h->d_keys_out = nullptr;
        h->d_vals = nullptr;
        h->d_sort_tmp = nullptr;
        h->sort_cap = 0;
        HIP_OK(hipMalloc(&h->d_keys, sizeof(unsigned) * n));
        HIP_OK(hipMalloc(&h->d_keys_out, sizeof(unsigned) * n));
        HIP_OK(hipMalloc(&h->d_vals, sizeof(int) * n));
        size_t bytes = 0;
        HIP_OK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, h->d_keys, h->d_keys_out, h->d_vals,
                                                             (int *)d_order, (int)n, 0, 32, s));
        HIP_OK(hipMalloc(&h->d_sort_tmp, bytes));
        h->sort_tmp_bytes = bytes;
        h->sort_cap = n;
    }
    if (launch_order_keys(d_counters, n, h->d_keys, h->d_vals, stream)) return fail(LSM_E_HIP, "order keys launch failed");
    size_t bytes = h->sort_tmp_bytes;
    HIP_OK(hipcub::DeviceRadixSort::SortPairsDescending(h->d_sort_tmp, bytes, h->d_keys, h->d_keys_out, h->d_vals,
                                                         (int *)d_order, (int)n, 0, 32, s));
    return LSM_OK;
}

int lsm_autotune(lsm_handle *h, int64_t n, double t0, double pressure, const double *d_dt, const double *d_y,
                 int64_t sample, void *stream, int32_t *result) {
    if (!h) return fail(LSM_E_INVALID, "null handle");
    if (n <= 0 || !d_dt || !d_y) return fail(LSM_E_INVALID, "autotune needs cells");
    if (int rc = check_step_args(t0, pressure)) return rc;
    KParams kp = h->kp;
    kp.bk.t0 = t0;
    kp.pressure = pressure;
    if (pick_method(h) != LSM_METHOD_TEAM) {
        if (result) { result[0] = h->team_wpb; result[1] = h->team_grid; result[2] = -1; }
        return LSM_OK;                                     // only the team kernel has launch shapes to choose
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t m = std::max<int64_t>(1, std::min<int64_t>(n, sample > 0 ? sample : 65536));
    const int64_t stride = n / m;
    const int ne = h->dm.ne;
    double *ys = nullptr, *dts = nullptr, *yo = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (ys) (void)hipFree(ys);
        if (dts) (void)hipFree(dts);
        if (yo) (void)hipFree(yo);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    if (hipMalloc(&ys, sizeof(double) * ne * m) != hipSuccess || hipMalloc(&dts, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&yo, sizeof(double) * ne * m) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(LSM_E_NOMEM, "autotune buffers");
    }
    int rc = launch_gather_cells(d_y, d_dt, n, ne, m, stride, ys, dts, stream);
    int best_wpb = h->team_wpb, best_grid = h->team_grid, best_cpw = 1;
    float best_ms = 3.4e38f;
    bool warmed = false;
    // cells per wavefront x waves per workgroup x workgroups per CU
    for (int cpw = 1; cpw <= team_max_cells_per_wave(h->dm) && !rc; cpw *= 2) {
        const int cap = team_waves_per_cu_cpw(h->dm, cpw);
        for (int wpb = 1; wpb <= 4 && !rc; wpb++) {
            for (int bpc = 1; wpb * bpc <= cap; bpc++) {
                const int grid = bpc * h->num_cu;
                if (!team_shape_fits_cpw(h->dm, cpw, wpb, bpc, h->lds_cu)) continue;
                for (int rep = warmed ? 0 : -1; rep < 1 && !rc; rep++) {   // the first candidate runs twice (warm-up)
                    if (launch_reset_queue(h->d_queue, stream)) { rc = -1; break; }
                    if (hipEventRecord(e0, s) != hipSuccess) { rc = -1; break; }
                    if (launch_team_cpw(h->dm, kp, m, dts, ys, yo, nullptr, nullptr, nullptr, h->d_queue, nullptr,
                                        grid, wpb, cpw, stream)) { rc = -1; break; }
                    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = -1; break; }
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = -1; break; }
                    warmed = true;
                    if (rep == 0 && ms < best_ms) { best_ms = ms; best_wpb = wpb; best_grid = grid; best_cpw = cpw; }
                }
            }
        }
    }
    cleanup();
    if (rc) return fail(LSM_E_HIP, "autotune launch failed");
    h->team_cpw = best_cpw;
    if (best_cpw > 1) {
        h->pack_wpb = best_wpb;
        h->pack_grid = best_grid;
    } else {
        h->team_wpb = best_wpb;
        h->team_grid = best_grid;
    }
    if (result) { result[0] = best_wpb; result[1] = best_grid; result[2] = (int32_t)(best_ms * 1000.f); }
    return LSM_OK;
}

int lsm_fp64_fallbacks(lsm_handle *h, uint64_t *count) {
    if (!h || !count) return fail(LSM_E_INVALID, "null argument");
    *count = 0;
    if (!h->d_fallbacks) return LSM_OK;
    unsigned long long v = 0;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(&v, h->d_fallbacks, sizeof(v), hipMemcpyDeviceToHost));
    *count = v;
    return LSM_OK;
}

int lsm_device_bytes(lsm_handle *h, uint64_t *bytes) {
    if (!h || !bytes) return fail(LSM_E_INVALID, "null argument");
    *bytes = (uint64_t)lsm_handle_device_bytes(h);
    return LSM_OK;
}

int lsm_get_launch(lsm_handle *h, int32_t *out) {
    if (!h || !out) return fail(LSM_E_INVALID, "null argument");
    const int method = pick_method(h);
    out[0] = method;
    const bool packed = method == LSM_METHOD_TEAM && h->team_cpw > 1;
    out[1] = method == LSM_METHOD_TEAM ? (packed ? h->pack_wpb : h->team_wpb) : 4;
    out[2] = method == LSM_METHOD_TEAM ? (packed ? h->pack_grid : h->team_grid)
                                       : (method == LSM_METHOD_WIDE ? h->wide_grid : 0);
    out[3] = h->kp.jac_mode;
    out[4] = h->kp.refine;
    return LSM_OK;
}

#ifdef LSM_PHASE_TIMERS
// Instrumented builds only: cumulative per-phase cycles (rhs, jacobian, factor,
// solve, J*v, cells, cell count, RHS parts) since lsm_create; out[12].
int lsm_debug_phase_cycles(lsm_handle *h, unsigned long long *out) {
    if (!h || !out || !h->kp.phase_cycles) return fail(LSM_E_INVALID, "no phase timers");
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out, h->kp.phase_cycles, 12 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return LSM_OK;
}
#endif

}  // extern "C"
