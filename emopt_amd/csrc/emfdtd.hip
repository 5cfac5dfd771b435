// libemfdtd: C ABI over the gfx950 CW-FDTD kernels (declarations and the
// reference function each entry point replaces: include/emfdtd.h).
//
// One emf_solver owns every buffer of one solve on one GPU: the six real
// time-domain fields, the raw complex materials and the derived update
// coefficients, the CPML psi slabs and tables, the source blocks, the
// convergence state and two complex result slots (forward, adjoint).  All
// work is queued on the handle's own HIP stream (default priority).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/emfdtd.h"
#include "batch.hpp"
#include "material.hpp"

using namespace emf;

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define EMF_FAIL(code, ...)   \
    do {                      \
        set_err(__VA_ARGS__); \
        return code;          \
    } while (0)

#define HIPCHK(x)                                                                \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess)                                                    \
            EMF_FAIL(EMF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define CHECK(x)                  \
    do {                          \
        int r_ = (x);             \
        if (r_ != EMF_OK) return r_; \
    } while (0)

template <class T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return EMF_OK;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        EMF_FAIL(EMF_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                 hipGetErrorString(e));
    }
    return EMF_OK;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

struct SourceDev {
    int i0, j0, k0, I, J, K;
    double2 *f[6];   // device (amplitude, phase) per component Jx..Mz
};

// Closed forms shared with the exported scalar helpers.
double phase_3t(double t0, double t1, double t2, double f0, double f1, double f2)
{
    double f10 = f1 - f0, f21 = f2 - f1;
    if (f10 == 0 && f21 == 0) return 0.0;
    return atan2(f10 * (sin(t2) - sin(t1)) - f21 * (sin(t1) - sin(t0)),
                 f21 * (cos(t1) - cos(t0)) - f10 * (cos(t2) - cos(t1)));
}

double amplitude_3t(double t0, double t1, double t2, double f0, double f1, double f2, double ph)
{
    double f21 = f2 - f1, f10 = f1 - f0;
    if (f21 == 0 && f10 == 0) return 0.0;
    if (f21 * f21 >= f10 * f10)
        return f21 / (cos(ph) * (sin(t2) - sin(t1)) + sin(ph) * (cos(t2) - cos(t1)));
    return f10 / (cos(ph) * (sin(t1) - sin(t0)) + sin(ph) * (cos(t1) - cos(t0)));
}

double phase_2t(double t0, double t1, double f0, double f1)
{
    if (f0 == 0.0 && f1 == 0) return 0.0;
    return atan((f1 * sin(t0) - f0 * sin(t1)) / (f0 * cos(t1) - f1 * cos(t0)));
}

double amplitude_2t(double t0, double t1, double f0, double f1, double ph)
{
    if (f0 * f0 > f1 * f1) return f1 / (sin(t1) * cos(ph) + cos(t1) * sin(ph));
    return f0 / (sin(t0) * cos(ph) + cos(t0) * sin(ph));
}

// kappa, b, c of one CPML layer at ramp factor f (src/fdtd.cpp:992-997)
void pml_layer(double sigma0, double alpha0, double kappa0, double dt, double f, double *kap,
               double *b, double *c)
{
    double sig = sigma0 * f;
    double alp = alpha0 * (1 - f);
    double k = (kappa0 - 1.0) * f + 1.0;
    double bb = exp(-dt * (sig / k + alp));
    *kap = k;
    *b = bb;
    *c = (bb == 1) ? 0.0 : (bb - 1) * sig / (sig * k + k * k * alp);
}

}  // namespace

struct emf_solver {
    std::mutex mtx;
    int dev = 0;
    hipStream_t stream = nullptr;
    // when set, the next two-pass launch records its own start and stop time
    // in these events (the dispatch's timestamps, no marker packets around it)
    hipEvent_t tev0 = nullptr, tev1 = nullptr;

    // geometry / time
    int Nx = 0, Ny = 0, Nz = 0;
    long long n = 0;
    double X = 0, Y = 0, Z = 0, dx = 0, dy = 0, dz = 0;
    double wl = 0, R = 0, dt = 0, odt = 0;
    double src_T = 0, src_min = 1e-4, src_k = 0;
    int ceps = 0;
    int w[6] = { 0, 0, 0, 0, 0, 0 };
    double sigma = 3.0, alpha = 0.0, kappa = 1.0, pw = 3.0;
    int bc[3] = { '0', '0', '0' };
    int nconv = 10;
    int zc = 2;
    int order = 3;                 // two-pass work-group order (pass_tile)
    int psi_stream = -1;           // streaming psi access: -1 by size, 0 off, 1 on

    // device state.  E/H (and the H-pass psi set) exist twice for the fused
    // step, which reads one set and writes the other; E, H and psi[0] always
    // point at the set holding the current state (index cur).
    // Every 3-component quantity is one allocation of 3n doubles (components
    // at a fixed stride), so kernels take one base pointer per quantity; the
    // per-component pointers below are views into those allocations.
    double *E[3] = {}, *H[3] = {};
    double *Ev[2] = {}, *Hv[2] = {};
    double *Eb[2][3] = {}, *Hb[2][3] = {};
    int cur = 0;
    double2 *eps[3] = {};
    double2 *mu[3] = {};           // only when mu is not uniform
    bool mat_set = false;
    bool mu_uniform = true;
    double mu_re = 1.0;            // uniform mu.re
    double *bE[3] = {}, *aE[3] = {};
    double *mH[3] = {}, *muR[3] = {};
    double *bEv = nullptr, *aEv = nullptr, *mHv = nullptr, *muRv = nullptr;
    bool coef_dirty = true;

    double *tab = nullptr;         // CPML tables: per axis kH bH cH kE bE cE
    std::vector<double> tab_host;  // host copy of the same tables
    bool tab_dirty = true;
    double *psi[2][6][2] = {};     // [pass H/E][face][component]; psi[0] = psiH[cur]
    double *psiH[2][6][2] = {};    // the two H-pass sets
    double *psiHv[2] = {}, *psiEv = nullptr;   // one allocation per psi set
    long long psi_off[6][2] = {};  // element offset of each slab array in a set
    long long psi_tot = 0;
    long long psi_n[6] = {};
    bool pml_built = false;

    double2 *res[2][6] = {};       // result slots: fwd, adj (one 6n block per slot)

    std::vector<SourceDev> srcs;
    std::vector<SourceDev> staged; // adjoint sources built on the device (emf_stage_*)
    SrcBlk *src_desc = nullptr;    // device copy of the source descriptors
    size_t src_desc_cap = 0;
    bool src_dirty = true;

    int kind = 1;                  // 0, 2 = fused step kernels; 1, 3 = separate H / E passes
    int tyg = 4;                   // waves (rows) per work-group
    int tyg_h = 0, zc_h = 0;       // the H pass's rows and planes per work-group (0: as tyg, zc)
    bool wide = false;             // the passes address cells by 64-bit offsets (set by check_ready)
    int wide_force = 0;            // 1: 64-bit offsets at any size (emf_set_wide_offsets)

    ConvState *conv = nullptr;
    long long *samples = nullptr;
    int nsamples = 0;
    double *hist = nullptr;
    int hist_cap = 0;
    int *done_host = nullptr;      // pinned ring of convergence flags
    int nhist = 0;

    double2 *scratch = nullptr;
    size_t scratch_n = 0;

    // weight planes of emf_interp_dot: per id, six optional complex planes
    static constexpr int kWeightSets = 8;
    double2 *wplane[kWeightSets][6] = {};
    long long wcount[kWeightSets] = {};

    void *in_batch = nullptr;      // the emf_batch this handle is a member of

    // Coefficient compression (real eps): a per-cell index into a table of
    // the distinct (eps_x, eps_y, eps_z).re triples; the E pass then reads
    // 2 or 4 bytes per cell instead of three 8-byte coefficients.
    int compress = 1;              // allowed (emf_set_compression)
    bool cidx_valid = false;       // index built from the current materials
    int cidx_bytes = 0;            // width of the index in use (0: off)
    void *cidx = nullptr;          // device index, n entries
    double *ctabE = nullptr;       // device table of eps triples
    double *ctabB = nullptr;       // device table of b = dt/eps triples (4 doubles each)
    long long ctab_n = 0;

    Geo geo() const
    {
        Geo g;
        g.Nx = Nx; g.Ny = Ny; g.Nz = Nz;
        g.nxy = (long long)Nx * Ny;
        g.rx = R / dx; g.ry = R / dy; g.rz = R / dz;
        g.dt = dt;
        for (int q = 0; q < 6; ++q) g.w[q] = w[q];
        for (int q = 0; q < 3; ++q) g.bc[q] = bc[q];
        g.zc = zc;
        g.order = order;
        return g;
    }
};

namespace {

int ensure_fields(emf_solver *s)
{
    if (s->Nx <= 0 || s->Ny <= 0 || s->Nz <= 0) EMF_FAIL(EMF_ERR_STATE, "grid dims not set");
    if (s->E[0]) return EMF_OK;
    CHECK(dalloc(&s->Ev[0], 3 * s->n));
    CHECK(dalloc(&s->Hv[0], 3 * s->n));
    CHECK(dalloc(&s->bEv, 3 * s->n));
    HIPCHK(hipMemsetAsync(s->Ev[0], 0, 3 * s->n * sizeof(double), s->stream));
    HIPCHK(hipMemsetAsync(s->Hv[0], 0, 3 * s->n * sizeof(double), s->stream));
    for (int c = 0; c < 3; ++c) {
        CHECK(dalloc(&s->eps[c], s->n));
        s->Eb[0][c] = s->Ev[0] + c * s->n;
        s->Hb[0][c] = s->Hv[0] + c * s->n;
        s->bE[c] = s->bEv + c * s->n;
        s->E[c] = s->Eb[0][c];
        s->H[c] = s->Hb[0][c];
    }
    s->cur = 0;
    return EMF_OK;
}

// Make buffer set `par` the current state.
void set_current(emf_solver *s, int par)
{
    s->cur = par;
    for (int c = 0; c < 3; ++c) {
        s->E[c] = s->Eb[par][c];
        s->H[c] = s->Hb[par][c];
    }
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < 2; ++c) s->psi[0][f][c] = s->psiH[par][f][c];
}

// Second buffer set for the fused step (allocated on first use).
int ensure_alt(emf_solver *s)
{
    if (!s->Ev[1]) {
        CHECK(dalloc(&s->Ev[1], 3 * s->n));
        CHECK(dalloc(&s->Hv[1], 3 * s->n));
        for (int c = 0; c < 3; ++c) {
            s->Eb[1][c] = s->Ev[1] + c * s->n;
            s->Hb[1][c] = s->Hv[1] + c * s->n;
        }
    }
    if (s->pml_built && !s->psiHv[1] && s->psi_tot > 0) {
        CHECK(dalloc(&s->psiHv[1], (size_t)s->psi_tot));
        for (int f = 0; f < 6; ++f)
            for (int c = 0; c < 2; ++c)
                s->psiH[1][f][c] = s->psi_n[f] ? s->psiHv[1] + s->psi_off[f][c] : nullptr;
    }
    return EMF_OK;
}

void free_fields(emf_solver *s)
{
    for (int q = 0; q < 2; ++q) { dfree(s->Ev[q]); dfree(s->Hv[q]); }
    dfree(s->bEv); dfree(s->aEv); dfree(s->mHv); dfree(s->muRv);
    for (int c = 0; c < 3; ++c) {
        s->E[c] = s->H[c] = nullptr;
        s->Eb[0][c] = s->Eb[1][c] = s->Hb[0][c] = s->Hb[1][c] = nullptr;
        s->bE[c] = s->aE[c] = s->mH[c] = s->muR[c] = nullptr;
        dfree(s->eps[c]); dfree(s->mu[c]);
    }
    for (int sl = 0; sl < 2; ++sl) {
        dfree(s->res[sl][0]);   // one block per slot; [1..5] point into it
        for (int c = 1; c < 6; ++c) s->res[sl][c] = nullptr;
    }
    dfree(s->cidx); dfree(s->ctabE); dfree(s->ctabB);
    s->cidx_valid = false;
    s->cidx_bytes = 0;
    s->ctab_n = 0;
    s->mat_set = false;
    s->coef_dirty = true;
    s->cur = 0;
}

int ensure_slot(emf_solver *s, int slot)
{
    if (slot < 0 || slot > 1) EMF_FAIL(EMF_ERR_ARG, "result slot must be 0 or 1, got %d", slot);
    if (s->res[slot][0]) return EMF_OK;
    // six components in one block: component c at res[slot][0] + c * n
    CHECK(dalloc(&s->res[slot][0], 6 * s->n));
    for (int c = 1; c < 6; ++c) s->res[slot][c] = s->res[slot][0] + (size_t)c * s->n;
    HIPCHK(hipMemsetAsync(s->res[slot][0], 0, 6 * s->n * sizeof(double2), s->stream));
    return EMF_OK;
}

int ensure_scratch(emf_solver *s, size_t n)
{
    if (s->scratch_n >= n) return EMF_OK;
    dfree(s->scratch);
    s->scratch_n = 0;
    CHECK(dalloc(&s->scratch, n));
    s->scratch_n = n;
    return EMF_OK;
}

unsigned grid1d(long long n, int block)
{
    long long b = (n + block - 1) / block;
    return (unsigned)std::min<long long>(std::max<long long>(b, 1), 65535LL * 16);
}

int refresh_coeffs(emf_solver *s)
{
    if (!s->mat_set) EMF_FAIL(EMF_ERR_STATE, "materials not set (emf_set_mat_arrays)");
    if (!s->coef_dirty) return EMF_OK;
    if (s->dt <= 0) EMF_FAIL(EMF_ERR_STATE, "dt not set");
    if (s->ceps && !s->aEv) {
        CHECK(dalloc(&s->aEv, 3 * s->n));
        for (int c = 0; c < 3; ++c) s->aE[c] = s->aEv + c * s->n;
    }
    if (!s->mu_uniform && !s->mHv) {
        CHECK(dalloc(&s->mHv, 3 * s->n));
        CHECK(dalloc(&s->muRv, 3 * s->n));
        for (int c = 0; c < 3; ++c) { s->mH[c] = s->mHv + c * s->n; s->muR[c] = s->muRv + c * s->n; }
    }
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_eps_coeffs, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream,
                           s->eps[c], s->bE[c], s->aE[c], s->n, s->dt, s->odt, s->ceps);
        if (!s->mu_uniform) {
            hipLaunchKernelGGL(k_mu_coeffs, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream,
                               s->mu[c], s->mH[c], s->muR[c], s->n, s->dt);
        }
    }
    s->cidx_bytes = 0;
    if (s->compress && s->cidx_valid && !s->ceps && s->ctab_n > 0) {
        hipLaunchKernelGGL(k_b_table, dim3(grid1d(4 * s->ctab_n, 256)), dim3(256), 0, s->stream,
                           s->ctabE, s->ctabB, s->ctab_n, s->dt);
        s->cidx_bytes = s->ctab_n <= 65535 ? 2 : 4;
    }
    HIPCHK(hipGetLastError());
    s->coef_dirty = false;
    return EMF_OK;
}

// Index of distinct real eps triples, built on the host from the arrays
// being uploaded.  Runs of equal cells along x are the common case, so the
// previous cell's triple is checked before the hash map.
struct EpsKey {
    uint64_t a, b, c;
    bool operator==(const EpsKey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct EpsKeyHash {
    size_t operator()(const EpsKey &k) const
    {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ULL;
        h ^= k.b + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
        h ^= k.c + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

int build_coeff_index(emf_solver *s, const double *ex, const double *ey, const double *ez)
{
    dfree(s->cidx); dfree(s->ctabE); dfree(s->ctabB);
    s->cidx_valid = false;
    s->ctab_n = 0;
    const long long n = s->n, limit = 1LL << 20;
    std::unordered_map<EpsKey, uint32_t, EpsKeyHash> map;
    std::vector<double> tab;
    std::vector<uint32_t> idx((size_t)n);
    EpsKey last{ ~0ULL, ~0ULL, ~0ULL };
    uint32_t last_i = 0;
    for (long long g = 0; g < n; ++g) {
        EpsKey k;
        memcpy(&k.a, &ex[2 * g], 8); memcpy(&k.b, &ey[2 * g], 8); memcpy(&k.c, &ez[2 * g], 8);
        if (!(k == last)) {
            auto it = map.find(k);
            if (it == map.end()) {
                if ((long long)map.size() >= limit) return EMF_OK;   // too many: stay off
                last_i = (uint32_t)map.size();
                map.emplace(k, last_i);
                tab.push_back(ex[2 * g]); tab.push_back(ey[2 * g]); tab.push_back(ez[2 * g]);
            } else {
                last_i = it->second;
            }
            last = k;
        }
        idx[(size_t)g] = last_i;
    }
    const long long nt = (long long)map.size();
    CHECK(dalloc(&s->ctabE, 3 * (size_t)nt));
    CHECK(dalloc(&s->ctabB, 4 * (size_t)nt));
    HIPCHK(hipMemcpyAsync(s->ctabE, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                          s->stream));
    if (nt <= 65535) {
        std::vector<uint16_t> i16((size_t)n);
        for (long long g = 0; g < n; ++g) i16[(size_t)g] = (uint16_t)idx[(size_t)g];
        uint16_t *d = nullptr;
        CHECK(dalloc(&d, (size_t)n + 2));   // the E pass reads aligned dwords
        s->cidx = d;
        HIPCHK(hipMemcpyAsync(d, i16.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    } else {
        uint32_t *d = nullptr;
        CHECK(dalloc(&d, (size_t)n));
        s->cidx = d;
        HIPCHK(hipMemcpyAsync(d, idx.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    s->ctab_n = nt;
    s->cidx_valid = true;
    return EMF_OK;
}

// CPML tables of compute_pml_params (src/fdtd.cpp:950-1181) into one device
// buffer: for axis a, six arrays (kH bH cH kE bE cE) of w_min + w_max entries.
int refresh_tables(emf_solver *s, PassTab *th, PassTab *te)
{
    size_t off[3], len[3], total = 0;
    for (int a = 0; a < 3; ++a) {
        len[a] = std::max(1, s->w[2 * a] + s->w[2 * a + 1]);
        off[a] = total;
        total += 6 * len[a];
    }
    if (s->tab_dirty) {
        std::vector<double> h(total, 0.0);
        for (int a = 0; a < 3; ++a) {
            double *kH = &h[off[a]], *bH = kH + len[a], *cH = bH + len[a];
            double *kE = cH + len[a], *bE = kE + len[a], *cE = bE + len[a];
            for (int side = 0; side < 2; ++side) {
                int wd = s->w[2 * a + side];
                for (int l = 0; l < wd; ++l) {
                    double dH, dE;
                    bool clampH = false, clampE = false;
                    if (a < 2) {   // x, y faces
                        dH = side == 0 ? double(l - 0.5) / wd : double(l + 0.5) / wd;
                        dE = double(l) / wd;
                        clampH = side == 0;
                    } else {       // z faces
                        dH = double(l) / wd;
                        dE = side == 0 ? double(l + 0.5) / wd : double(l - 0.5) / wd;
                        clampE = side == 1;
                    }
                    double fH = std::pow(dH, s->pw), fE = std::pow(dE, s->pw);
                    if (clampH && fH < 0) fH = 0;
                    if (clampE && fE < 0) fE = 0;
                    int q = side == 0 ? l : s->w[2 * a] + l;
                    pml_layer(s->sigma, s->alpha, s->kappa, s->dt, fH, &kH[q], &bH[q], &cH[q]);
                    pml_layer(s->sigma, s->alpha, s->kappa, s->dt, fE, &kE[q], &bE[q], &cE[q]);
                }
            }
        }
        dfree(s->tab);
        CHECK(dalloc(&s->tab, total));
        HIPCHK(hipMemcpyAsync(s->tab, h.data(), total * sizeof(double), hipMemcpyHostToDevice,
                              s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));   // h is a temporary
        s->tab_host = h;
        s->tab_dirty = false;
    }
    for (int a = 0; a < 3; ++a) {
        double *base = s->tab + off[a];
        th->kap[a] = base; th->b[a] = base + len[a]; th->c[a] = base + 2 * len[a];
        te->kap[a] = base + 3 * len[a]; te->b[a] = base + 4 * len[a]; te->c[a] = base + 5 * len[a];
    }
    return EMF_OK;
}

void free_psi(emf_solver *s)
{
    dfree(s->psiHv[0]); dfree(s->psiHv[1]); dfree(s->psiEv);
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < 2; ++c)
            s->psi[0][f][c] = s->psi[1][f][c] = s->psiH[0][f][c] = s->psiH[1][f][c] = nullptr;
    s->psi_tot = 0;
    s->pml_built = false;
}

int build_psi(emf_solver *s)
{
    CHECK(ensure_fields(s));
    free_psi(s);
    long long nx = s->Nx, ny = s->Ny, nz = s->Nz;
    for (int f = 0; f < 6; ++f) {
        int wd = s->w[f];
        long long cells = f < 2 ? nz * ny * wd : (f < 4 ? nz * nx * wd : ny * nx * wd);
        s->psi_n[f] = cells;
    }
    long long tot = 0;
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < 2; ++c) { s->psi_off[f][c] = tot; tot += s->psi_n[f]; }
    s->psi_tot = tot;
    if (tot > 0) {
        CHECK(dalloc(&s->psiHv[0], (size_t)tot));
        CHECK(dalloc(&s->psiEv, (size_t)tot));
        if (s->Ev[1]) CHECK(dalloc(&s->psiHv[1], (size_t)tot));
        HIPCHK(hipMemsetAsync(s->psiHv[0], 0, tot * sizeof(double), s->stream));
        HIPCHK(hipMemsetAsync(s->psiEv, 0, tot * sizeof(double), s->stream));
    }
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < 2; ++c) {
            const bool has = s->psi_n[f] > 0;
            s->psiH[0][f][c] = has ? s->psiHv[0] + s->psi_off[f][c] : nullptr;
            s->psiH[1][f][c] = has && s->psiHv[1] ? s->psiHv[1] + s->psi_off[f][c] : nullptr;
            s->psi[1][f][c] = has ? s->psiEv + s->psi_off[f][c] : nullptr;
        }
    s->pml_built = true;   // set_current below needs the views
    set_current(s, s->cur == 1 && !s->psiHv[1] && tot > 0 ? 0 : s->cur);
    s->tab_dirty = true;
    return EMF_OK;
}

int reset_psi(emf_solver *s)
{
    // zero the current H-pass set and the E-pass set
    if (s->psi_tot > 0) {
        HIPCHK(hipMemsetAsync(s->psiHv[s->cur] ? s->psiHv[s->cur] : s->psiHv[0], 0,
                              s->psi_tot * sizeof(double), s->stream));
        HIPCHK(hipMemsetAsync(s->psiEv, 0, s->psi_tot * sizeof(double), s->stream));
    }
    return EMF_OK;
}

// Device descriptors of the registered sources (rebuilt after changes).
int refresh_sources(emf_solver *s)
{
    if (!s->src_dirty) return EMF_OK;
    size_t n = s->srcs.size();
    if (n > s->src_desc_cap) {
        dfree(s->src_desc);
        s->src_desc_cap = 0;
        CHECK(dalloc(&s->src_desc, n));
        s->src_desc_cap = n;
    }
    if (n) {
        std::vector<SrcBlk> h(n);
        for (size_t q = 0; q < n; ++q) {
            const SourceDev &d = s->srcs[q];
            h[q].i0 = d.i0; h[q].j0 = d.j0; h[q].k0 = d.k0; h[q].I = d.I; h[q].J = d.J; h[q].K = d.K;
            for (int c = 0; c < 6; ++c) h[q].f[c] = d.f[c];
        }
        HIPCHK(hipMemcpyAsync(s->src_desc, h.data(), n * sizeof(SrcBlk), hipMemcpyHostToDevice,
                              s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));   // h is a temporary
    }
    s->src_dirty = false;
    return EMF_OK;
}

int check_ready(emf_solver *s)
{
    CHECK(ensure_fields(s));
    if (!s->pml_built) EMF_FAIL(EMF_ERR_STATE, "CPML not built (emf_build_pml)");
    for (int f = 0; f < 6; ++f) {
        int lim = f < 2 ? s->Nx : (f < 4 ? s->Ny : s->Nz);
        if (s->w[f] < 0 || s->w[f] > lim)
            EMF_FAIL(EMF_ERR_ARG, "pml width %d out of range for face %d", s->w[f], f);
    }
    if (s->w[0] + s->w[1] > s->Nx || s->w[2] + s->w[3] > s->Ny || s->w[4] + s->w[5] > s->Nz)
        EMF_FAIL(EMF_ERR_ARG, "opposite pml slabs overlap");
    if (s->R <= 0 || s->dx <= 0 || s->dy <= 0 || s->dz <= 0)
        EMF_FAIL(EMF_ERR_STATE, "wavelength / physical dims not set");
    // The step passes address a cell of a field by a byte offset from the
    // field's base: 32-bit while three components fit in 4 GiB (178 M cells),
    // 64-bit above.  The fused step kernels have only the 32-bit form; the
    // psi slabs always use 32-bit offsets.
    const bool over = 3ULL * (unsigned long long)s->n * 8ULL >= (1ULL << 32);
    if ((unsigned long long)s->psi_tot * 8ULL >= (1ULL << 32))
        EMF_FAIL(EMF_ERR_ARG, "CPML slabs of %lld cells exceed the kernels' 32-bit offset range",
                 (long long)s->psi_tot);
    if ((over || s->wide_force) && (s->kind == 0 || s->kind == 2))
        EMF_FAIL(EMF_ERR_ARG,
                 "the fused step kernels take at most %llu cells and no 64-bit offsets; use the "
                 "two passes (emf_set_tuning kind 1)", ((1ULL << 32) - 1) / 24ULL);
    s->wide = over || s->wide_force;
    CHECK(refresh_coeffs(s));
    return EMF_OK;
}

double env_at(const emf_solver *s, double t)
{
    // ramp envelope of src_func_t (src/fdtd.cpp:1595-1602)
    if (t <= s->src_T)
        return (1 + s->src_min) * exp(-(t - s->src_T) * (t - s->src_T) / s->src_k) - s->src_min;
    return 1.0;
}

// union box (i0, i1, j0, j1, k0, k1) of the solver's source blocks
void src_union(const emf_solver *s, int ub[6])
{
    ub[0] = ub[2] = ub[4] = INT_MAX;
    ub[1] = ub[3] = ub[5] = INT_MIN;
    for (const auto &d : s->srcs) {
        ub[0] = std::min(ub[0], d.i0); ub[1] = std::max(ub[1], d.i0 + d.I);
        ub[2] = std::min(ub[2], d.j0); ub[3] = std::max(ub[3], d.j0 + d.J);
        ub[4] = std::min(ub[4], d.k0); ub[5] = std::max(ub[5], d.k0 + d.K);
    }
}

PassSrc pass_sources(const emf_solver *s, int comp0, double t)
{
    PassSrc ps{};
    ps.n = (int)s->srcs.size();
    ps.comp0 = comp0;
    ps.st = sin(t);
    ps.ct = cos(t);
    ps.env = env_at(s, t);
    ps.blk = s->src_desc;
    src_union(s, ps.ub);
    ps.ninl = std::min(ps.n, SRC_INLINE);
    for (int q = 0; q < ps.ninl; ++q) {
        const SourceDev &d = s->srcs[q];
        const int b[6] = { d.i0, d.I, d.j0, d.J, d.k0, d.K };
        for (int a = 0; a < 6; ++a) ps.ib[q][a] = b[a];
    }
    return ps;
}



// Kernel arguments for the state in buffer set `from` (written to `to`).
FusedArgs make_args(const emf_solver *s, int from, int to)
{
    FusedArgs A;
    A.Eo = s->Ev[from]; A.Ho = s->Hv[from];
    A.En = s->Ev[to]; A.Hn = s->Hv[to];
    A.bE = s->bEv; A.aE = s->aEv;
    A.mH = s->mu_uniform ? nullptr : s->mHv;
    A.muR = s->mu_uniform ? nullptr : s->muRv;
    A.pho = s->psiHv[from]; A.phn = s->psiHv[to]; A.pe = s->psiEv;
    A.tab = s->tab;
    A.cs = (unsigned)(s->n * 8);
    A.csl = (unsigned long long)s->n * 8ULL;
    A.m_s = s->dt / s->mu_re;
    A.mu_s = s->mu_re;
    const bool comp = s->cidx_bytes != 0 && s->compress && s->cidx_valid && !s->ceps;
    A.bI = comp ? s->cidx : nullptr;
    A.bT = s->ctabB;
    A.bIw = s->cidx_bytes;
    return A;
}

// Whether the passes stream the psi arrays past the caches (non-temporal):
// when a step's data -- fields, coefficients and psi -- exceeds the 256 MB
// Infinity Cache, nothing of it survives to the next step and psi lines would
// only displace the fields the next pass re-reads; a smaller solve keeps its
// psi cached from step to step.  emf_set_streaming overrides.
inline bool stream_psi(const emf_solver *s)
{
    if (s->psi_stream >= 0) return s->psi_stream == 1;
    const double bytes = (double)s->n * 100.0 + 2.0 * 2.0 * 8.0 * (double)s->psi_tot;
    return bytes > 256e6;
}

// rows per work-group of the two passes (tyg of a fused kind may be 16)
int pass_ty(const emf_solver *s)
{
    return s->tyg == 1 || s->tyg == 2 || s->tyg == 8 ? s->tyg : 4;
}

// a two-pass launch, through hipExtLaunchKernelGGL with the solver's timing
// events when they are set
#define PASS_LAUNCH(K, GR, BLK, ...)                                                         \
    do {                                                                                     \
        if (s->tev0)                                                                         \
            hipExtLaunchKernelGGL(K, GR, BLK, 0, s->stream, s->tev0, s->tev1, 0, __VA_ARGS__); \
        else                                                                                 \
            hipLaunchKernelGGL(K, GR, BLK, 0, s->stream, __VA_ARGS__);                       \
    } while (0)

dim3 pass_grid(const emf_solver *s)
{
    const int ty = pass_ty(s);
    return dim3((s->Nx + PX - 1) / PX, (s->Ny + ty - 1) / ty, (s->Nz + s->zc - 1) / s->zc);
}

// the H pass's work-group shape: its own when set, else the E pass's
int pass_ty_h(const emf_solver *s)
{
    const int t = s->tyg_h > 0 ? s->tyg_h : s->tyg;
    return t == 1 || t == 2 || t == 8 ? t : 4;
}

int pass_zc_h(const emf_solver *s) { return s->zc_h > 0 ? s->zc_h : s->zc; }

template <bool MU, int OCC, bool NT>
void launch_h_occ(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                  const int *done)
{
    const int ty = pass_ty_h(s);
    const dim3 gr((s->Nx + PX - 1) / PX, (s->Ny + ty - 1) / ty, (s->Nz + g.zc - 1) / g.zc);
    switch (ty) {
    case 1: PASS_LAUNCH((k_pass_h<MU, 1, OCC, NT>), gr, dim3(PX, 1), A, g, ps, done); break;
    case 2: PASS_LAUNCH((k_pass_h<MU, 2, OCC, NT>), gr, dim3(PX, 2), A, g, ps, done); break;
    case 8: PASS_LAUNCH((k_pass_h<MU, 8, OCC, NT>), gr, dim3(PX, 8), A, g, ps, done); break;
    default: PASS_LAUNCH((k_pass_h<MU, 4, OCC, NT>), gr, dim3(PX, 4), A, g, ps, done);
    }
}

template <int CM, int OCC, int ZU, bool NT>
void launch_e_zu(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                 const int *done)
{
    const dim3 gr = pass_grid(s);
    switch (pass_ty(s)) {
    case 1: PASS_LAUNCH((k_pass_e<CM, 1, OCC, ZU, NT>), gr, dim3(PX, 1), A, g, ps, done); break;
    case 2: PASS_LAUNCH((k_pass_e<CM, 2, OCC, ZU, NT>), gr, dim3(PX, 2), A, g, ps, done); break;
    case 8: PASS_LAUNCH((k_pass_e<CM, 8, OCC, ZU, NT>), gr, dim3(PX, 8), A, g, ps, done); break;
    default: PASS_LAUNCH((k_pass_e<CM, 4, OCC, ZU, NT>), gr, dim3(PX, 4), A, g, ps, done);
    }
}

// columns of at most 8 planes take the unrolled E pass (indices loaded first)
template <int CM, int OCC, bool NT>
void launch_e_occ(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                  const int *done)
{
    if (CM == CM_INDEX && s->zc <= 2) launch_e_zu<CM, OCC, 2, NT>(s, A, g, ps, done);
    else if (CM == CM_INDEX && s->zc <= 4) launch_e_zu<CM, OCC, 4, NT>(s, A, g, ps, done);
    else if (CM == CM_INDEX && s->zc <= 8) launch_e_zu<CM, OCC, 8, NT>(s, A, g, ps, done);
    else launch_e_zu<CM, OCC, 0, NT>(s, A, g, ps, done);
}

// the passes with 64-bit cell offsets (grids above 178 M cells): 64 x 4
// work-groups, 4 waves per SIMD, the E pass's column unrolled as for 32 bits
template <bool MU, bool NT>
void launch_h_wide(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                   const int *done)
{
    const dim3 gr((s->Nx + PX - 1) / PX, (s->Ny + 3) / 4, (s->Nz + g.zc - 1) / g.zc);
    PASS_LAUNCH((k_pass_h<MU, 4, 4, NT, true>), gr, dim3(PX, 4), A, g, ps, done);
}

template <int CM, int ZU, bool NT>
void launch_e_wide_zu(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                      const int *done)
{
    const dim3 gr((s->Nx + PX - 1) / PX, (s->Ny + 3) / 4, (s->Nz + s->zc - 1) / s->zc);
    PASS_LAUNCH((k_pass_e<CM, 4, 4, ZU, NT, true>), gr, dim3(PX, 4), A, g, ps, done);
}

template <int CM, bool NT>
void launch_e_wide(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                   const int *done)
{
    if (CM == CM_INDEX && s->zc <= 2) launch_e_wide_zu<CM, 2, NT>(s, A, g, ps, done);
    else if (CM == CM_INDEX && s->zc <= 4) launch_e_wide_zu<CM, 4, NT>(s, A, g, ps, done);
    else if (CM == CM_INDEX && s->zc <= 8) launch_e_wide_zu<CM, 8, NT>(s, A, g, ps, done);
    else launch_e_wide_zu<CM, 0, NT>(s, A, g, ps, done);
}

// kind 4: the H pass with its y-neighbour row staged in LDS (uniform mu)
template <bool NT>
void launch_h_lds(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                  const int *done)
{
    const int ty = pass_ty_h(s);
    const dim3 gr((s->Nx + PX - 1) / PX, (s->Ny + ty - 1) / ty, (s->Nz + g.zc - 1) / g.zc);
    switch (ty) {
    case 2: PASS_LAUNCH((k_pass_h_lds<2, NT>), gr, dim3(PX, 2), A, g, ps, done); break;
    case 8: PASS_LAUNCH((k_pass_h_lds<8, NT>), gr, dim3(PX, 8), A, g, ps, done); break;
    default: PASS_LAUNCH((k_pass_h_lds<4, NT>), gr, dim3(PX, 4), A, g, ps, done);
    }
}

// kind 3 = the two passes compiled for 5 waves per SIMD
template <bool MU>
void launch_h_t(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                const int *done)
{
    if (s->kind == 4 && !MU && !s->wide && pass_ty_h(s) > 1) {
        if (stream_psi(s)) launch_h_lds<true>(s, A, g, ps, done);
        else launch_h_lds<false>(s, A, g, ps, done);
    } else if (s->wide) {
        if (stream_psi(s)) launch_h_wide<MU, true>(s, A, g, ps, done);
        else launch_h_wide<MU, false>(s, A, g, ps, done);
    } else if (s->kind == 3) launch_h_occ<MU, 5, false>(s, A, g, ps, done);
    else if (stream_psi(s)) launch_h_occ<MU, 4, true>(s, A, g, ps, done);
    else launch_h_occ<MU, 4, false>(s, A, g, ps, done);
}

template <int CM>
void launch_e_t(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &ps,
                const int *done)
{
    if (s->wide) {
        if (stream_psi(s)) launch_e_wide<CM, true>(s, A, g, ps, done);
        else launch_e_wide<CM, false>(s, A, g, ps, done);
    } else if (s->kind == 3) launch_e_occ<CM, 5, false>(s, A, g, ps, done);
    else if (stream_psi(s)) launch_e_occ<CM, 4, true>(s, A, g, ps, done);
    else launch_e_occ<CM, 4, false>(s, A, g, ps, done);
}

// H pass in place on the current set (update_H, src/fdtd.cpp:121-367)
int launch_h(emf_solver *s, double t, const int *done)
{
    PassTab th, te;
    CHECK(refresh_tables(s, &th, &te));
    CHECK(refresh_sources(s));
    PassSrc ps = pass_sources(s, 3, t);
    FusedArgs A = make_args(s, s->cur, s->cur);
    Geo g = s->geo();
    g.zc = pass_zc_h(s);
    if (s->mu_uniform) launch_h_t<false>(s, A, g, ps, done);
    else launch_h_t<true>(s, A, g, ps, done);
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

// E pass in place on the current set (update_E, src/fdtd.cpp:369-755)
int launch_e(emf_solver *s, double t, const int *done)
{
    PassTab th, te;
    CHECK(refresh_tables(s, &th, &te));
    CHECK(refresh_sources(s));
    PassSrc ps = pass_sources(s, 0, t);
    FusedArgs A = make_args(s, s->cur, s->cur);
    Geo g = s->geo();
    if (s->ceps) launch_e_t<CM_COMPLEX>(s, A, g, ps, done);
    else if (A.bI) launch_e_t<CM_INDEX>(s, A, g, ps, done);
    else launch_e_t<CM_REAL>(s, A, g, ps, done);
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

template <bool MU, int CM>
void launch_lds_cm(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &sh,
                   const PassSrc &se, const int *done)
{
    const int ty = s->tyg < 4 ? 4 : s->tyg;
    dim3 grid((s->Nx + FXO - 1) / FXO, (s->Ny + ty - 2) / (ty - 1), (s->Nz + s->zc - 1) / s->zc);
    switch (ty) {
    case 16:
        hipLaunchKernelGGL((k_step_lds<MU, CM, 16>), grid, dim3(FX, 16), 0, s->stream, A, g, sh, se, done);
        break;
    case 8:
        hipLaunchKernelGGL((k_step_lds<MU, CM, 8>), grid, dim3(FX, 8), 0, s->stream, A, g, sh, se, done);
        break;
    default:
        hipLaunchKernelGGL((k_step_lds<MU, CM, 4>), grid, dim3(FX, 4), 0, s->stream, A, g, sh, se, done);
    }
}

template <bool MU, bool CE>
void launch_lds_t(const emf_solver *s, const FusedArgs &A, const Geo &g, const PassSrc &sh,
                  const PassSrc &se, const int *done)
{
    if (CE) launch_lds_cm<MU, CM_COMPLEX>(s, A, g, sh, se, done);
    else if (A.bI) launch_lds_cm<MU, CM_INDEX>(s, A, g, sh, se, done);
    else launch_lds_cm<MU, CM_REAL>(s, A, g, sh, se, done);
}

template <bool MU, bool CE>
void launch_fused_t(const emf_solver *s, dim3 grid, const FusedArgs &A, const Geo &g,
                    const PassSrc &sh, const PassSrc &se, const int *done)
{
    if (s->kind == 2) {
        launch_lds_t<MU, CE>(s, A, g, sh, se, done);
        return;
    }
    switch (s->tyg) {
    case 1:
        hipLaunchKernelGGL((k_step_fused<MU, CE, 1>), grid, dim3(FX, 1), 0, s->stream, A, g, sh, se, done);
        break;
    case 2:
        hipLaunchKernelGGL((k_step_fused<MU, CE, 2>), grid, dim3(FX, 2), 0, s->stream, A, g, sh, se, done);
        break;
    case 8:
        hipLaunchKernelGGL((k_step_fused<MU, CE, 8>), grid, dim3(FX, 8), 0, s->stream, A, g, sh, se, done);
        break;
    default:
        hipLaunchKernelGGL((k_step_fused<MU, CE, 4>), grid, dim3(FX, 4), 0, s->stream, A, g, sh, se, done);
    }
}

// One fused time step n reading buffer set `from` and writing set 1-from.
int launch_fused(emf_solver *s, long long nidx, int from, const int *done)
{
    CHECK(ensure_alt(s));
    PassTab th, te;
    CHECK(refresh_tables(s, &th, &te));
    CHECK(refresh_sources(s));
    FusedArgs A = make_args(s, from, 1 - from);
    PassSrc sh = pass_sources(s, 3, (double)nidx * s->dt);
    PassSrc se = pass_sources(s, 0, ((double)nidx + 0.5) * s->dt);
    Geo g = s->geo();
    // x tiles overlap by one cell (63 owned lanes + one halo lane); one row per wave
    dim3 grid((s->Nx + FXO - 1) / FXO, (s->Ny + s->tyg - 1) / s->tyg, (s->Nz + s->zc - 1) / s->zc);
    if (s->mu_uniform) {
        if (s->ceps) launch_fused_t<false, true>(s, grid, A, g, sh, se, done);
        else launch_fused_t<false, false>(s, grid, A, g, sh, se, done);
    } else {
        if (s->ceps) launch_fused_t<true, true>(s, grid, A, g, sh, se, done);
        else launch_fused_t<true, false>(s, grid, A, g, sh, se, done);
    }
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

inline bool fused(const emf_solver *s);

// One time step.  Fused: reads set `from`, writes 1-from (the caller tracks
// which set is current).  Separate passes: in place on the current set.
int step(emf_solver *s, long long nidx, const int *done, int from)
{
    if (fused(s)) return launch_fused(s, nidx, from, done);
    CHECK(launch_h(s, (double)nidx * s->dt, done));
    CHECK(launch_e(s, ((double)nidx + 0.5) * s->dt, done));
    return EMF_OK;
}

// Whether steps use the fused kernel: selected, and every 3-component buffer
// and psi set below 4 GiB (the kernel addresses with 32-bit byte offsets).
inline bool fused(const emf_solver *s)
{
    return (s->kind == 0 || s->kind == 2) && 3ULL * (unsigned long long)s->n * 8ULL < (1ULL << 32) &&
           (unsigned long long)s->psi_tot * 8ULL < (1ULL << 32);
}

// Step n of a run that started at n_start with set `start_set` current.
inline int from_set(const emf_solver *s, long long n, long long n_start, int start_set)
{
    return fused(s) ? (int)((start_set + (n - n_start)) & 1) : start_set;
}

int capture(emf_solver *s, int slot, int which)
{
    CHECK(ensure_fields(s));
    CHECK(ensure_slot(s, slot));
    double2 **r = s->res[slot];
    hipLaunchKernelGGL(k_capture, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream, s->E[0],
                       s->E[1], s->E[2], s->H[0], s->H[1], s->H[2], r[0], r[1], r[2], r[3], r[4],
                       r[5], s->n, which);
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

int phasor3(emf_solver *s, int slot, double t0, double t1, double t2)
{
    CHECK(ensure_fields(s));
    CHECK(ensure_slot(s, slot));
    PhTimes te, thh;
    double h0 = t0 - 0.5 * s->dt, h1 = t1 - 0.5 * s->dt, h2 = t2 - 0.5 * s->dt;
    te.s21 = sin(t2) - sin(t1); te.s10 = sin(t1) - sin(t0);
    te.c10 = cos(t1) - cos(t0); te.c21 = cos(t2) - cos(t1);
    thh.s21 = sin(h2) - sin(h1); thh.s10 = sin(h1) - sin(h0);
    thh.c10 = cos(h1) - cos(h0); thh.c21 = cos(h2) - cos(h1);
    double2 **r = s->res[slot];
    hipLaunchKernelGGL(k_phasor3, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream, s->E[0],
                       s->E[1], s->E[2], s->H[0], s->H[1], s->H[2], r[0], r[1], r[2], r[3], r[4],
                       r[5], s->n, te, thh);
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

// Convergence sample points: arange(0, I*J*K, int(I*J*K/nconv)) over the
// padded array (emopt/fdtd.py:411-413).  On one rank a ghost-face sample reads
// the opposite face's outermost layer (the periodic self-exchange of
// emopt/fdtd.py:1733-1873) and edge/corner ghosts stay 0.
int build_samples(emf_solver *s)
{
    long long I = s->Nz, J = s->Ny, K = s->Nx, IJK = I * J * K;
    long long stepv = IJK / std::max(1, s->nconv);
    if (stepv <= 0) EMF_FAIL(EMF_ERR_ARG, "nconv larger than the grid");
    std::vector<long long> cells;
    for (long long p = 0; p < IJK; p += stepv) {
        long long pk = p % (K + 2), pj = (p / (K + 2)) % (J + 2), pi = p / ((K + 2) * (J + 2));
        int ghosts = 0;
        auto wrap = [&](long long v, long long N) -> long long {
            if (v == 0) { ++ghosts; return N - 1; }
            if (v == N + 1) { ++ghosts; return 0; }
            return v - 1;
        };
        long long k = wrap(pk, K), j = wrap(pj, J), i = wrap(pi, I);
        cells.push_back(ghosts > 1 ? -1 : (i * J + j) * K + k);
    }
    if ((int)cells.size() > NCONV_MAX)
        EMF_FAIL(EMF_ERR_ARG, "%zu convergence points exceed the limit %d", cells.size(), NCONV_MAX);
    dfree(s->samples);
    CHECK(dalloc(&s->samples, cells.size()));
    HIPCHK(hipMemcpyAsync(s->samples, cells.data(), cells.size() * sizeof(long long),
                          hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));   // cells is a temporary
    s->nsamples = (int)cells.size();
    return EMF_OK;
}

}  // namespace

// ---- allocation-free gradient assembly ----------------------------------
// emf_grad_assemble_noalloc stages its entries, their per-parameter segments
// and the gradient in this file-scope device array instead of allocating:
// every GPU the library runs on holds its own copy until the process ends.
// In host code the name is a stand-in, not a device address, so the address
// is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that pointer is used afterwards.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu[kMaxDevices];     // one call at a time per device
unsigned char *g_noalloc_ptr[kMaxDevices];

// Device address of `scratch` on the current device `dev` (caller holds
// g_noalloc_mu[dev]).
int noalloc_scratch(int dev, unsigned char **p)
{
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) EMF_FAIL(EMF_ERR_HIP, "no device address for the scratch array");
        g_noalloc_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *p = g_noalloc_ptr[dev];
    return EMF_OK;
}

// seg[p] = first entry of parameter p in the sorted pidx (seg[nparam] = n)
__global__ void k_segments(const int *__restrict__ pidx, long long n, int nparam,
                           long long *__restrict__ seg)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nparam) return;
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) / 2;
        if (pidx[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    seg[p] = lo;
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

// =========================================================================
// C ABI
// =========================================================================

#define LOCK(s)                                                     \
    if (!(s)) EMF_FAIL(EMF_ERR_ARG, "null solver handle");         \
    std::lock_guard<std::mutex> lk_((s)->mtx);                      \
    if ((s)->in_batch)                                              \
        EMF_FAIL(EMF_ERR_STATE, "handle is a member of a batch (emf_batch_remove it first)"); \
    if (hipSetDevice((s)->dev) != hipSuccess)                       \
        EMF_FAIL(EMF_ERR_HIP, "hipSetDevice(%d) failed", (s)->dev);

extern "C" {

const char *emf_last_error(void) { return g_err.c_str(); }

const char *emf_version(void)
{
    return "emfdtd 0.1 gfx950 fp64 (hipcc -O3, packed-fp32-ops disabled)";
}

int emf_device_count(void)
{
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    return n;
}

int emf_new(int device, emf_solver **out)
{
    if (!out) EMF_FAIL(EMF_ERR_ARG, "null output pointer");
    *out = nullptr;
    int nd = 0;
    HIPCHK(hipGetDeviceCount(&nd));
    if (device < 0 || device >= nd) EMF_FAIL(EMF_ERR_ARG, "device %d not present (%d visible)", device, nd);
    HIPCHK(hipSetDevice(device));
    emf_solver *s = new emf_solver();
    s->dev = device;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        EMF_FAIL(EMF_ERR_HIP, "hipStreamCreate failed");
    }
    if (hipHostMalloc((void **)&s->done_host, 64 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        EMF_FAIL(EMF_ERR_HIP, "hipHostMalloc failed");
    }
    if (dalloc(&s->conv, 1) != EMF_OK) {
        (void)hipHostFree(s->done_host);
        (void)hipStreamDestroy(s->stream);
        delete s;
        return EMF_ERR_NOMEM;
    }
    *out = s;
    return EMF_OK;
}

int emf_delete(emf_solver *s)
{
    if (!s) return EMF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mtx);
        if (s->in_batch)
            EMF_FAIL(EMF_ERR_STATE, "handle is a member of a batch (emf_batch_remove it first)");
        (void)hipSetDevice(s->dev);
        (void)hipStreamSynchronize(s->stream);
        free_fields(s);
        free_psi(s);
        dfree(s->tab);
        for (auto &d : s->srcs)
            for (int c = 0; c < 6; ++c) dfree(d.f[c]);
        s->srcs.clear();
        for (auto &d : s->staged)
            for (int c = 0; c < 6; ++c) dfree(d.f[c]);
        s->staged.clear();
        dfree(s->src_desc);
        dfree(s->conv);
        dfree(s->samples);
        dfree(s->hist);
        dfree(s->scratch);
        for (int w = 0; w < emf_solver::kWeightSets; ++w)
            for (int q = 0; q < 6; ++q) dfree(s->wplane[w][q]);
        if (s->done_host) (void)hipHostFree(s->done_host);
        (void)hipStreamDestroy(s->stream);
    }
    delete s;
    return EMF_OK;
}

int emf_set_wavelength(emf_solver *s, double wl)
{
    LOCK(s);
    if (!(wl > 0)) EMF_FAIL(EMF_ERR_ARG, "wavelength must be > 0");
    s->wl = wl;
    s->R = wl / (2 * M_PI);
    return EMF_OK;
}

int emf_set_physical_dims(emf_solver *s, double X, double Y, double Z, double dx, double dy,
                          double dz)
{
    LOCK(s);
    if (!(dx > 0 && dy > 0 && dz > 0)) EMF_FAIL(EMF_ERR_ARG, "grid spacings must be > 0");
    s->X = X; s->Y = Y; s->Z = Z; s->dx = dx; s->dy = dy; s->dz = dz;
    return EMF_OK;
}

int emf_set_grid_dims(emf_solver *s, int Nx, int Ny, int Nz)
{
    LOCK(s);
    if (Nx <= 0 || Ny <= 0 || Nz <= 0) EMF_FAIL(EMF_ERR_ARG, "grid dims must be positive");
    if ((long long)Nx * Ny * Nz > (1LL << 40)) EMF_FAIL(EMF_ERR_ARG, "grid too large");
    if (Nx != s->Nx || Ny != s->Ny || Nz != s->Nz) {
        (void)hipStreamSynchronize(s->stream);
        free_fields(s);
        free_psi(s);
        s->Nx = Nx; s->Ny = Ny; s->Nz = Nz;
        s->n = (long long)Nx * Ny * Nz;
    }
    return ensure_fields(s);
}

int emf_set_local_grid(emf_solver *s, int k0, int j0, int i0, int K, int J, int I)
{
    LOCK(s);
    if (k0 != 0 || j0 != 0 || i0 != 0 || K != s->Nx || J != s->Ny || I != s->Nz)
        EMF_FAIL(EMF_ERR_ARG, "one GPU owns the whole grid: local grid must be (0,0,0,Nx,Ny,Nz)");
    return EMF_OK;
}

int emf_set_dt(emf_solver *s, double dt)
{
    LOCK(s);
    if (!(dt > 0)) EMF_FAIL(EMF_ERR_ARG, "dt must be > 0");
    s->dt = dt;
    s->odt = 1.0 / dt;
    s->coef_dirty = true;
    s->tab_dirty = true;
    return EMF_OK;
}

int emf_set_complex_eps(emf_solver *s, int c)
{
    LOCK(s);
    s->ceps = c ? 1 : 0;
    s->coef_dirty = true;
    return EMF_OK;
}

int emf_set_mat_arrays(emf_solver *s, const double *ex, const double *ey, const double *ez,
                       const double *mx, const double *my, const double *mz)
{
    LOCK(s);
    if (!ex || !ey || !ez || !mx || !my || !mz) EMF_FAIL(EMF_ERR_ARG, "null material array");
    CHECK(ensure_fields(s));
    const double *e[3] = { ex, ey, ez }, *m[3] = { mx, my, mz };
    for (int c = 0; c < 3; ++c)
        HIPCHK(hipMemcpyAsync(s->eps[c], e[c], s->n * sizeof(double2), hipMemcpyHostToDevice,
                              s->stream));
    // mu enters only through mu.re (src/fdtd.cpp:186-188); a uniform value
    // becomes a scalar and no per-cell mu traffic is generated.
    double v = m[0][0];
    bool uni = true;
    for (int c = 0; c < 3 && uni; ++c)
        for (long long g = 0; g < s->n; ++g)
            if (m[c][2 * g] != v) { uni = false; break; }
    s->mu_uniform = uni;
    s->mu_re = v;
    if (!uni)
        for (int c = 0; c < 3; ++c) {
            if (!s->mu[c]) CHECK(dalloc(&s->mu[c], s->n));
            HIPCHK(hipMemcpyAsync(s->mu[c], m[c], s->n * sizeof(double2), hipMemcpyHostToDevice,
                                  s->stream));
        }
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->compress) CHECK(build_coeff_index(s, ex, ey, ez));
    s->mat_set = true;
    s->coef_dirty = true;
    return EMF_OK;
}

int emf_update_eps_box(emf_solver *s, int c, const double *vals, int k1, int k2, int j1, int j2,
                       int i1, int i2)
{
    LOCK(s);
    if (c < 0 || c > 2 || !vals) EMF_FAIL(EMF_ERR_ARG, "bad component or null array");
    if (!s->mat_set) EMF_FAIL(EMF_ERR_STATE, "materials not set");
    if (k1 < 0 || j1 < 0 || i1 < 0 || k2 > s->Nx || j2 > s->Ny || i2 > s->Nz || k1 >= k2 ||
        j1 >= j2 || i1 >= i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range");
    long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    CHECK(ensure_scratch(s, cnt));
    HIPCHK(hipMemcpyAsync(s->scratch, vals, cnt * sizeof(double2), hipMemcpyHostToDevice, s->stream));
    Geo g = s->geo();
    hipLaunchKernelGGL(k_scatter_box, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream, s->eps[c],
                       s->scratch, g, k1, k2 - k1, j1, j2 - j1, i1, i2 - i1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cidx_valid = false;         // the index no longer describes eps: full coefficients
    s->coef_dirty = true;
    return EMF_OK;
}

int emf_update_mu_box(emf_solver *s, int c, const double *vals, int k1, int k2, int j1, int j2,
                      int i1, int i2)
{
    LOCK(s);
    if (c < 0 || c > 2 || !vals) EMF_FAIL(EMF_ERR_ARG, "bad component or null array");
    if (!s->mat_set) EMF_FAIL(EMF_ERR_STATE, "materials not set");
    if (k1 < 0 || j1 < 0 || i1 < 0 || k2 > s->Nx || j2 > s->Ny || i2 > s->Nz || k1 >= k2 ||
        j1 >= j2 || i1 >= i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range");
    long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (s->mu_uniform) {
        bool same = true;
        for (long long m = 0; m < cnt && same; ++m)
            same = vals[2 * m] == s->mu_re && vals[2 * m + 1] == 0.0;
        if (same) return EMF_OK;              // still the uniform value: nothing changes
        // the grid's mu becomes per cell: every component starts at the uniform value
        for (int q = 0; q < 3; ++q) {
            if (!s->mu[q]) CHECK(dalloc(&s->mu[q], s->n));
            hipLaunchKernelGGL(k_fill2, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream, s->mu[q],
                               s->n, make_double2(s->mu_re, 0.0));
        }
        HIPCHK(hipGetLastError());
        s->mu_uniform = false;
    }
    CHECK(ensure_scratch(s, cnt));
    HIPCHK(hipMemcpyAsync(s->scratch, vals, cnt * sizeof(double2), hipMemcpyHostToDevice, s->stream));
    Geo g = s->geo();
    hipLaunchKernelGGL(k_scatter_box, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream, s->mu[c],
                       s->scratch, g, k1, k2 - k1, j1, j2 - j1, i1, i2 - i1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->coef_dirty = true;
    return EMF_OK;
}

int emf_set_compression(emf_solver *s, int on)
{
    LOCK(s);
    s->compress = on ? 1 : 0;
    s->coef_dirty = true;          // the next refresh switches the E pass over
    return EMF_OK;
}

int emf_get_mat_arrays(emf_solver *s, double *ex, double *ey, double *ez, double *mx, double *my,
                       double *mz)
{
    LOCK(s);
    if (!s->mat_set) EMF_FAIL(EMF_ERR_STATE, "materials not set");
    double *e[3] = { ex, ey, ez }, *m[3] = { mx, my, mz };
    for (int c = 0; c < 3; ++c) {
        if (e[c])
            HIPCHK(hipMemcpyAsync(e[c], s->eps[c], s->n * sizeof(double2), hipMemcpyDeviceToHost,
                                  s->stream));
        if (m[c]) {
            if (s->mu_uniform) {
                HIPCHK(hipStreamSynchronize(s->stream));
                for (long long g = 0; g < s->n; ++g) { m[c][2 * g] = s->mu_re; m[c][2 * g + 1] = 0.0; }
            } else {
                HIPCHK(hipMemcpyAsync(m[c], s->mu[c], s->n * sizeof(double2),
                                      hipMemcpyDeviceToHost, s->stream));
            }
        }
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_set_bc(emf_solver *s, const char *bc)
{
    LOCK(s);
    if (!bc || strlen(bc) < 3) EMF_FAIL(EMF_ERR_ARG, "bc must have 3 characters");
    for (int a = 0; a < 3; ++a)
        if (bc[a] != '0' && bc[a] != 'E' && bc[a] != 'H')
            EMF_FAIL(EMF_ERR_ARG, "unsupported boundary condition '%c' (use 0, E or H)", bc[a]);
    for (int a = 0; a < 3; ++a) s->bc[a] = bc[a];
    return EMF_OK;
}

int emf_set_pml_widths(emf_solver *s, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax)
{
    LOCK(s);
    int v[6] = { xmin, xmax, ymin, ymax, zmin, zmax };
    for (int q = 0; q < 6; ++q)
        if (v[q] < 0) EMF_FAIL(EMF_ERR_ARG, "negative pml width");
    memcpy(s->w, v, sizeof(v));
    s->pml_built = false;
    s->tab_dirty = true;
    return EMF_OK;
}

int emf_set_pml_properties(emf_solver *s, double sigma, double alpha, double kappa, double pw)
{
    LOCK(s);
    s->sigma = sigma; s->alpha = alpha; s->kappa = kappa; s->pw = pw;
    s->tab_dirty = true;
    return EMF_OK;
}

int emf_build_pml(emf_solver *s)
{
    LOCK(s);
    return build_psi(s);
}

int emf_reset_pml(emf_solver *s)
{
    LOCK(s);
    return reset_psi(s);
}

int emf_add_source(emf_solver *s, const double *Jx, const double *Jy, const double *Jz,
                   const double *Mx, const double *My, const double *Mz, int i0, int j0, int k0,
                   int I, int J, int K, int calc_phase)
{
    LOCK(s);
    const double *f[6] = { Jx, Jy, Jz, Mx, My, Mz };
    for (int c = 0; c < 6; ++c)
        if (!f[c]) EMF_FAIL(EMF_ERR_ARG, "null source array");
    if (I <= 0 || J <= 0 || K <= 0) EMF_FAIL(EMF_ERR_ARG, "empty source block");
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + I > s->Nz || j0 + J > s->Ny || k0 + K > s->Nx)
        EMF_FAIL(EMF_ERR_ARG, "source block outside the grid");
    long long cnt = (long long)I * J * K;
    SourceDev d;
    d.i0 = i0; d.j0 = j0; d.k0 = k0; d.I = I; d.J = J; d.K = K;
    for (int c = 0; c < 6; ++c) d.f[c] = nullptr;
    std::vector<double> buf(2 * cnt);
    for (int c = 0; c < 6; ++c) {
        for (long long m = 0; m < cnt; ++m) {
            const double re = f[c][2 * m], im = f[c][2 * m + 1];
            // stored as (A cos(phi), A sin(phi)); see SrcBlk
            if (calc_phase) {
                // the reference keeps (|v|, -atan2(im, re)) (src/fdtd.cpp:1517-1576),
                // whose (A cos(phi), A sin(phi)) is (re, -im)
                buf[2 * m] = re;
                buf[2 * m + 1] = -im;
            } else {
                buf[2 * m] = re * cos(im);
                buf[2 * m + 1] = re * sin(im);
            }
        }
        int rc = dalloc(&d.f[c], cnt);
        if (rc != EMF_OK) {
            for (int q = 0; q < c; ++q) dfree(d.f[q]);
            return rc;
        }
        HIPCHK(hipMemcpy(d.f[c], buf.data(), cnt * sizeof(double2), hipMemcpyHostToDevice));
    }
    s->srcs.push_back(d);
    s->src_dirty = true;
    return EMF_OK;
}

int emf_clear_sources(emf_solver *s)
{
    LOCK(s);
    (void)hipStreamSynchronize(s->stream);
    for (auto &d : s->srcs)
        for (int c = 0; c < 6; ++c) dfree(d.f[c]);
    s->srcs.clear();
    s->src_dirty = true;
    return EMF_OK;
}

int emf_set_source_properties(emf_solver *s, double T, double m)
{
    LOCK(s);
    if (!(T > 0) || !(m > 0)) EMF_FAIL(EMF_ERR_ARG, "ramp time and minimum must be > 0");
    s->src_T = T;
    s->src_min = m;
    s->src_k = T * T / log((1 + m) / m);
    return EMF_OK;
}

double emf_src_func_t(emf_solver *s, int n, double t, double phase)
{
    (void)n;
    if (!s) return NAN;
    if (t <= s->src_T)
        return sin(t + phase) * ((1 + s->src_min) * exp(-(t - s->src_T) * (t - s->src_T) / s->src_k) -
                                 s->src_min);
    return sin(t + phase);
}

int emf_update_H(emf_solver *s, int n, double t)
{
    LOCK(s);
    (void)n;
    CHECK(check_ready(s));
    return launch_h(s, t, nullptr);
}

int emf_update_E(emf_solver *s, int n, double t)
{
    LOCK(s);
    (void)n;
    CHECK(check_ready(s));
    return launch_e(s, t, nullptr);
}

int emf_zero_fields(emf_solver *s)
{
    LOCK(s);
    CHECK(ensure_fields(s));
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemsetAsync(s->E[c], 0, s->n * sizeof(double), s->stream));
        HIPCHK(hipMemsetAsync(s->H[c], 0, s->n * sizeof(double), s->stream));
    }
    return EMF_OK;
}

int emf_get_time_fields(emf_solver *s, double *ex, double *ey, double *ez, double *hx, double *hy,
                        double *hz)
{
    LOCK(s);
    CHECK(ensure_fields(s));
    double *o[6] = { ex, ey, ez, hx, hy, hz };
    for (int c = 0; c < 6; ++c)
        if (o[c])
            HIPCHK(hipMemcpyAsync(o[c], c < 3 ? s->E[c] : s->H[c - 3], s->n * sizeof(double),
                                  hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_set_time_fields(emf_solver *s, const double *ex, const double *ey, const double *ez,
                        const double *hx, const double *hy, const double *hz)
{
    LOCK(s);
    CHECK(ensure_fields(s));
    const double *o[6] = { ex, ey, ez, hx, hy, hz };
    for (int c = 0; c < 6; ++c)
        if (o[c])
            HIPCHK(hipMemcpyAsync(c < 3 ? s->E[c] : s->H[c - 3], o[c], s->n * sizeof(double),
                                  hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_capture_t0_fields(emf_solver *s, int slot)
{
    LOCK(s);
    return capture(s, slot, 0);
}

int emf_capture_t1_fields(emf_solver *s, int slot)
{
    LOCK(s);
    return capture(s, slot, 1);
}

int emf_calc_complex_fields_3T(emf_solver *s, int slot, double t0, double t1, double t2)
{
    LOCK(s);
    return phasor3(s, slot, t0, t1, t2);
}

int emf_calc_complex_fields_2T(emf_solver *s, int slot, double t0, double t1)
{
    LOCK(s);
    CHECK(ensure_fields(s));
    CHECK(ensure_slot(s, slot));
    PhTimes2 te, thh;
    double h0 = t0 - 0.5 * s->dt, h1 = t1 - 0.5 * s->dt;
    te.st0 = sin(t0); te.st1 = sin(t1); te.ct0 = cos(t0); te.ct1 = cos(t1);
    thh.st0 = sin(h0); thh.st1 = sin(h1); thh.ct0 = cos(h0); thh.ct1 = cos(h1);
    double2 **r = s->res[slot];
    hipLaunchKernelGGL(k_phasor2, dim3(grid1d(s->n, 256)), dim3(256), 0, s->stream, s->E[0],
                       s->E[1], s->E[2], s->H[0], s->H[1], s->H[2], r[0], r[1], r[2], r[3], r[4],
                       r[5], s->n, te, thh);
    HIPCHK(hipGetLastError());
    return EMF_OK;
}

double emf_calc_phase_2T(double t0, double t1, double f0, double f1)
{
    return phase_2t(t0, t1, f0, f1);
}

double emf_calc_amplitude_2T(double t0, double t1, double f0, double f1, double ph)
{
    return amplitude_2t(t0, t1, f0, f1, ph);
}

double emf_calc_phase_3T(double t0, double t1, double t2, double f0, double f1, double f2)
{
    return phase_3t(t0, t1, t2, f0, f1, f2);
}

double emf_calc_amplitude_3T(double t0, double t1, double t2, double f0, double f1, double f2,
                             double ph)
{
    return amplitude_3t(t0, t1, t2, f0, f1, f2, ph);
}

int emf_set_nconv(emf_solver *s, int nconv)
{
    LOCK(s);
    if (nconv <= 0) EMF_FAIL(EMF_ERR_ARG, "nconv must be > 0");
    s->nconv = nconv;
    return EMF_OK;
}

int emf_solve(emf_solver *s, int slot, double rtol, double nmax, int64_t Tn, int64_t fixed_n0,
              emf_solve_info *info)
{
    LOCK(s);
    if (!(rtol > 0)) EMF_FAIL(EMF_ERR_ARG, "rtol must be > 0");
    if (Tn < 0) EMF_FAIL(EMF_ERR_ARG, "Tn must be >= 0");
    CHECK(check_ready(s));
    CHECK(ensure_slot(s, slot));
    CHECK(build_samples(s));
    if (fused(s)) CHECK(ensure_alt(s));
    long long cap = (Tn > 0 ? (long long)(std::min(std::max(nmax, 0.0), 1e15) / (double)Tn) : 0) + 4;
    cap = std::min<long long>(cap, 1LL << 20);
    if (cap > s->hist_cap) {
        dfree(s->hist);
        CHECK(dalloc(&s->hist, 3 * (size_t)cap));
        s->hist_cap = (int)cap;
    }
    auto t_start = std::chrono::steady_clock::now();
    // reset CPML and fields (emopt/fdtd.py:1016-1018); set 0 holds step 0
    set_current(s, 0);
    CHECK(reset_psi(s));
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemsetAsync(s->E[c], 0, s->n * sizeof(double), s->stream));
        HIPCHK(hipMemsetAsync(s->H[c], 0, s->n * sizeof(double), s->stream));
    }
    HIPCHK(hipMemsetAsync(s->conv, 0, sizeof(ConvState), s->stream));
    const double amp_rtol = rtol, phi_rtol = sqrt(rtol);
    int *dflag = &s->conv->done;
    // the state after n steps lives in set from_set(n) (fused) or set 0

    long long n = 0;
    int status = 0;
    if (fixed_n0 >= 0) {
        for (long long q = 0; q < fixed_n0; ++q) CHECK(step(s, q, nullptr, from_set(s, q, 0, 0)));
        n = fixed_n0;
        status = 2;
    } else {
        // The loop breaks at the first step index n > nmax (emopt/fdtd.py:1054).
        long long n_stop = nmax < 0 ? 0 : (long long)floor(nmax) + 1;
        const int LOOK = 2;               // chunks in flight before the first flag is read
        std::vector<std::pair<long long, hipEvent_t>> pend;
        int ring = 0;
        while (true) {
            while ((int)pend.size() < LOOK && Tn > 0 && n + Tn <= n_stop) {
                // steps after convergence are gated off by dflag and leave the state as is
                for (long long q = 0; q < Tn; ++q) CHECK(step(s, n + q, dflag, from_set(s, n + q, 0, 0)));
                long long nchk = n + Tn - 1;
                int slotf = ring++ % 64;
                const int ps = from_set(s, n + Tn, 0, 0);   // set holding E after step nchk
                hipLaunchKernelGGL(k_conv_check, dim3(1), dim3(64), 0, s->stream, s->Eb[ps][0],
                                   s->Eb[ps][1], s->Eb[ps][2], s->samples, s->nsamples,
                                   (double)nchk * s->dt, amp_rtol, phi_rtol, nchk, s->conv,
                                   s->hist, s->hist_cap, (int *)nullptr);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(&s->done_host[slotf], dflag, sizeof(int),
                                      hipMemcpyDeviceToHost, s->stream));
                hipEvent_t ev;
                HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                HIPCHK(hipEventRecord(ev, s->stream));
                pend.push_back({ n, ev });
                n += Tn;
            }
            if (pend.empty()) {
                // fewer than Tn steps remain before nmax: run them and stop
                for (; n < n_stop; ++n) CHECK(step(s, n, dflag, from_set(s, n, 0, 0)));
                status = 1;
                break;
            }
            auto front = pend.front();
            pend.erase(pend.begin());
            HIPCHK(hipEventSynchronize(front.second));
            (void)hipEventDestroy(front.second);
            int slotf = (ring - 1 - (int)pend.size()) % 64;
            if (s->done_host[slotf]) {
                n = front.first + Tn;
                break;
            }
        }
        for (auto &pe : pend) {
            (void)hipEventSynchronize(pe.second);
            (void)hipEventDestroy(pe.second);
        }
    }
    // capture / two more Tn-step runs / phasor (emopt/fdtd.py:1156-1194)
    long long n0 = n;
    set_current(s, from_set(s, n0, 0, 0));
    CHECK(capture(s, slot, 0));
    for (long long q = n0; q < n0 + Tn; ++q) CHECK(step(s, q, nullptr, from_set(s, q, 0, 0)));
    set_current(s, from_set(s, n0 + Tn, 0, 0));
    CHECK(capture(s, slot, 1));
    for (long long q = n0 + Tn; q < n0 + 2 * Tn; ++q) CHECK(step(s, q, nullptr, from_set(s, q, 0, 0)));
    set_current(s, from_set(s, n0 + 2 * Tn, 0, 0));
    CHECK(phasor3(s, slot, (double)n0 * s->dt, (double)(n0 + Tn) * s->dt,
                  (double)(n0 + 2 * Tn) * s->dt));
    ConvState cs;
    HIPCHK(hipMemcpyAsync(&cs, s->conv, sizeof(ConvState), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    auto t_end = std::chrono::steady_clock::now();
    s->nhist = std::min(cs.nchecks, s->hist_cap);
    // the loop runs on past non-finite samples (emopt/fdtd.py:1050-1052,
    // 1137-1138) and stops at Nmax: say so when the last check saw them
    if (status == 1 && cs.nchecks > 0 && !(std::isfinite(cs.a_change) && std::isfinite(cs.phi_change)))
        status = 3;
    if (info) {
        info->n0 = n0;
        info->Tn = Tn;
        info->steps = n0 + 2 * Tn;
        info->status = status;
        info->nchecks = cs.nchecks;
        info->a_change = cs.a_change;
        info->phi_change = cs.phi_change;
        info->seconds = std::chrono::duration<double>(t_end - t_start).count();
    }
    return EMF_OK;
}

int emf_get_history(emf_solver *s, double *out, int cap)
{
    LOCK(s);
    int nrec = s->nhist;
    if (out && cap > 0 && nrec > 0) {
        int m = std::min(cap, nrec);
        HIPCHK(hipMemcpy(out, s->hist, 3 * (size_t)m * sizeof(double), hipMemcpyDeviceToHost));
    }
    return nrec;
}

int emf_get_field(emf_solver *s, int slot, int comp, int k1, int k2, int j1, int j2, int i1,
                  int i2, double *out)
{
    LOCK(s);
    if (slot < 0 || slot > 1 || !s->res[slot][0]) EMF_FAIL(EMF_ERR_STATE, "result slot %d empty", slot);
    if (comp < 0 || comp > 5 || !out) EMF_FAIL(EMF_ERR_ARG, "bad component or null output");
    if (k1 < 0 || j1 < 0 || i1 < 0 || k2 > s->Nx || j2 > s->Ny || i2 > s->Nz || k1 > k2 ||
        j1 > j2 || i1 > i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range");
    long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (cnt == 0) return EMF_OK;
    if (cnt == s->n) {
        HIPCHK(hipMemcpyAsync(out, s->res[slot][comp], cnt * sizeof(double2),
                              hipMemcpyDeviceToHost, s->stream));
    } else {
        CHECK(ensure_scratch(s, cnt));
        Geo g = s->geo();
        hipLaunchKernelGGL(k_gather_box, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream,
                           s->res[slot][comp], s->scratch, g, k1, k2 - k1, j1, j2 - j1, i1, i2 - i1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, s->scratch, cnt * sizeof(double2), hipMemcpyDeviceToHost,
                              s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

// The averaging stencil of get_field_interp for component comp (Ex..Hz) on a
// grown box starting at (ek1, ej1): offsets and the 'E'/'H' mirrors of
// emopt/fdtd.py:1472-1547.
static IStencil interp_stencil(const emf_solver *s, int comp, int ek1, int ej1)
{
    static const int nst[6] = {4, 4, 1, 2, 2, 8};
    static const int off[6][8][3] = {
        {{0, 0, 0}, {0, 0, -1}, {1, 0, 0}, {1, 0, -1}},
        {{0, 0, 0}, {0, -1, 0}, {1, 0, 0}, {1, -1, 0}},
        {{0, 0, 0}},
        {{0, 0, 0}, {0, -1, 0}},
        {{0, 0, 0}, {0, 0, -1}},
        {{0, 0, 0}, {0, 0, -1}, {0, -1, 0}, {0, -1, -1}, {1, 0, 0}, {1, 0, -1}, {1, -1, 0}, {1, -1, -1}}};
    static const int mirx[6] = {1, 0, 0, 0, 1, 1}, miry[6] = {0, 1, 0, 1, 0, 1};
    IStencil st{};
    st.n = nst[comp];
    for (int q = 0; q < st.n; ++q)
        for (int a = 0; a < 3; ++a) st.d[q][a] = off[comp][q][a];
    const char bx = s->bc[0], by = s->bc[1];
    st.mx = mirx[comp] && ek1 == 0 && (bx == 'E' || bx == 'H');
    st.my = miry[comp] && ej1 == 0 && (by == 'E' || by == 'H');
    st.sx = bx == 'E' ? -1.0 : 1.0;
    st.sy = by == 'E' ? -1.0 : 1.0;
    return st;
}

// checks shared by the interpolating entry points
static int check_interp_boxes(const emf_solver *s, int slot, int k1, int k2, int j1, int j2, int i1,
                              int i2, int ek1, int ek2, int ej1, int ej2, int ei1, int ei2)
{
    if (slot < 0 || slot > 1 || !s->res[slot][0]) EMF_FAIL(EMF_ERR_STATE, "result slot %d empty", slot);
    if (ek1 < 0 || ej1 < 0 || ei1 < 0 || ek2 > s->Nx || ej2 > s->Ny || ei2 > s->Nz ||
        k1 < ek1 || j1 < ej1 || i1 < ei1 || k2 > ek2 || j2 > ej2 || i2 > ei2 || k1 > k2 ||
        j1 > j2 || i1 > i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range (the domain must lie in the grown box, and it in "
                 "the grid)");
    return EMF_OK;
}

int emf_get_field_interp(emf_solver *s, int slot, int comp, int k1, int k2, int j1, int j2,
                         int i1, int i2, int ek1, int ek2, int ej1, int ej2, int ei1, int ei2,
                         double *out)
{
    LOCK(s);
    if (slot < 0 || slot > 1 || !s->res[slot][0]) EMF_FAIL(EMF_ERR_STATE, "result slot %d empty", slot);
    if (comp < 0 || comp > 5 || !out) EMF_FAIL(EMF_ERR_ARG, "bad component or null output");
    if (ek1 < 0 || ej1 < 0 || ei1 < 0 || ek2 > s->Nx || ej2 > s->Ny || ei2 > s->Nz ||
        k1 < ek1 || j1 < ej1 || i1 < ei1 || k2 > ek2 || j2 > ej2 || i2 > ei2 || k1 > k2 ||
        j1 > j2 || i1 > i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range (the domain must lie in the grown box, and it in "
                 "the grid)");
    const long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (cnt == 0) return EMF_OK;
    IStencil st = interp_stencil(s, comp, ek1, ej1);
    const IBox dom{k1, k2, j1, j2, i1, i2}, ext{ek1, ek2, ej1, ej2, ei1, ei2};
    CHECK(ensure_scratch(s, cnt));
    hipLaunchKernelGGL(k_field_interp, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream,
                       s->res[slot][comp], s->Nx, s->Ny, dom, ext, st, s->scratch);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->scratch, cnt * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_interp_flux(emf_solver *s, int slot, int axis, int k1, int k2, int j1, int j2, int i1,
                    int i2, int ek1, int ek2, int ej1, int ej2, int ei1, int ei2, double *out)
{
    LOCK(s);
    if (axis < 0 || axis > 2 || !out) EMF_FAIL(EMF_ERR_ARG, "axis must be 0..2, out non-null");
    CHECK(check_interp_boxes(s, slot, k1, k2, j1, j2, i1, i2, ek1, ek2, ej1, ej2, ei1, ei2));
    *out = 0.0;
    const long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (cnt == 0) return EMF_OK;
    // the component pairs of emopt/fdtd.py:1611-1665: x faces Ey Hz* - Ez Hy*,
    // y faces Ex Hz* - Ez Hx*, z faces Ex Hy* - Ey Hx*
    static const int pairs[3][4] = {{1, 5, 2, 4}, {0, 5, 2, 3}, {0, 4, 1, 3}};
    FluxArgs fa{};
    for (int q = 0; q < 4; ++q) {
        fa.F[q] = s->res[slot][pairs[axis][q]];
        fa.st[q] = interp_stencil(s, pairs[axis][q], ek1, ej1);
    }
    const int nblk = (int)std::min<long long>(256, (cnt + RED_THREADS - 1) / RED_THREADS);
    CHECK(ensure_scratch(s, nblk));
    const IBox dom{k1, k2, j1, j2, i1, i2}, ext{ek1, ek2, ej1, ej2, ei1, ei2};
    hipLaunchKernelGGL(k_interp_flux, dim3(nblk), dim3(RED_THREADS), 0, s->stream, fa, s->Nx,
                       s->Ny, dom, ext, s->scratch);
    HIPCHK(hipGetLastError());
    double2 part[256];
    HIPCHK(hipMemcpyAsync(part, s->scratch, nblk * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[b].x;
    *out = acc;
    return EMF_OK;
}

int emf_set_plane_weights(emf_solver *s, int id, int64_t count, const double *w0, const double *w1,
                          const double *w2, const double *w3, const double *w4, const double *w5)
{
    LOCK(s);
    if (id < 0 || id >= emf_solver::kWeightSets) EMF_FAIL(EMF_ERR_ARG, "weight set id must be 0..%d",
                                                        emf_solver::kWeightSets - 1);
    if (count < 0) EMF_FAIL(EMF_ERR_ARG, "count must be >= 0");
    const double *w[6] = { w0, w1, w2, w3, w4, w5 };
    for (int q = 0; q < 6; ++q) {
        if (s->wcount[id] != count || !w[q]) { dfree(s->wplane[id][q]); s->wplane[id][q] = nullptr; }
        if (!w[q] || count == 0) continue;
        if (!s->wplane[id][q]) CHECK(dalloc(&s->wplane[id][q], (size_t)count));
        HIPCHK(hipMemcpyAsync(s->wplane[id][q], w[q], count * sizeof(double2), hipMemcpyHostToDevice,
                              s->stream));
    }
    s->wcount[id] = count;
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_interp_dot(emf_solver *s, int slot, int id, int k1, int k2, int j1, int j2, int i1, int i2,
                   int ek1, int ek2, int ej1, int ej2, int ei1, int ei2, double *out)
{
    LOCK(s);
    if (id < 0 || id >= emf_solver::kWeightSets || !out) EMF_FAIL(EMF_ERR_ARG, "bad weight set or null out");
    CHECK(check_interp_boxes(s, slot, k1, k2, j1, j2, i1, i2, ek1, ek2, ej1, ej2, ei1, ei2));
    const long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (cnt != s->wcount[id])
        EMF_FAIL(EMF_ERR_ARG, "domain has %lld cells, weight set %d has %lld", cnt, id,
                 (long long)s->wcount[id]);
    for (int q = 0; q < 4; ++q) out[q] = 0.0;
    if (cnt == 0) return EMF_OK;
    DotArgs da{};
    for (int q = 0; q < 6; ++q) {
        da.F[q] = s->res[slot][q];
        da.st[q] = interp_stencil(s, q, ek1, ej1);
        da.W[q] = s->wplane[id][q];
    }
    const int nblk = (int)std::min<long long>(256, (cnt + RED_THREADS - 1) / RED_THREADS);
    CHECK(ensure_scratch(s, 2 * nblk));
    const IBox dom{k1, k2, j1, j2, i1, i2}, ext{ek1, ek2, ej1, ej2, ei1, ei2};
    hipLaunchKernelGGL(k_interp_dot, dim3(nblk), dim3(RED_THREADS), 0, s->stream, da, s->Nx, s->Ny,
                       dom, ext, s->scratch);
    HIPCHK(hipGetLastError());
    double2 part[512];
    HIPCHK(hipMemcpyAsync(part, s->scratch, 2 * nblk * sizeof(double2), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int b = 0; b < nblk; ++b) {
        out[0] += part[2 * b].x; out[1] += part[2 * b].y;
        out[2] += part[2 * b + 1].x; out[3] += part[2 * b + 1].y;
    }
    return EMF_OK;
}

int emf_field_products(emf_solver *s, int which, int k1, int k2, int j1, int j2, int i1, int i2,
                       double *out)
{
    LOCK(s);
    if (which != 0 && which != 1) EMF_FAIL(EMF_ERR_ARG, "which must be 0 (E) or 1 (H)");
    if (!out) EMF_FAIL(EMF_ERR_ARG, "null output");
    if (!s->res[0][0] || !s->res[1][0])
        EMF_FAIL(EMF_ERR_STATE, "needs a forward and an adjoint solution");
    if (k1 < 0 || j1 < 0 || i1 < 0 || k2 > s->Nx || j2 > s->Ny || i2 > s->Nz || k1 > k2 ||
        j1 > j2 || i1 > i2)
        EMF_FAIL(EMF_ERR_ARG, "box outside the grid");
    const long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    if (cnt == 0) return EMF_OK;
    ProdArgs pa{};
    for (int c = 0; c < 3; ++c) {
        pa.f[c] = s->res[0][3 * which + c];
        pa.a[c] = s->res[1][3 * which + c];
    }
    CHECK(ensure_scratch(s, 3 * cnt));
    pa.out = s->scratch;
    const IBox dom{k1, k2, j1, j2, i1, i2};
    hipLaunchKernelGGL(k_field_products, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream, pa,
                       s->Nx, s->Ny, dom);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->scratch, 3 * cnt * sizeof(double2), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

// ------------------------------------------------ staged (device-built) sources

namespace {

void free_blocks(std::vector<SourceDev> &v)
{
    for (auto &d : v)
        for (int c = 0; c < 6; ++c) dfree(d.f[c]);
    v.clear();
}

// The transposed interpolation of interpolated_dFdx_3D
// (emopt/fomutils.py:1282-1429) of six planes P (device arrays of the
// domain's Ni x Nj x Nk cells; null = zero) of a domain whose min x and y
// indices are k1 and j1: each plane is added, scaled, into a zero array
// padded by one cell on every side, in the reference's order of terms; the
// padded arrays, cropped to the grown domain and then to place = (i0, j0,
// k0, id1, jd1, kd1, sI, sJ, sK) (the grid block and its offset into the
// grown domain), become one staged source block.
int stage_transposed(emf_solver *s, const double2 *const P[6], int Ni, int Nj, int Nk, int k1,
                     int j1, const int *place)
{
    const int Pz = Ni + 2, Py = Nj + 2, Px = Nk + 2;
    const long long np = (long long)Pz * Py * Px;
    double2 *pad[6] = {};
    int rc = EMF_OK;
    for (int c = 0; c < 6 && rc == EMF_OK; ++c) {
        rc = dalloc(&pad[c], (size_t)np);
        if (rc == EMF_OK)
            hipLaunchKernelGGL(k_fill2, dim3(grid1d(np, 256)), dim3(256), 0, s->stream, pad[c], np,
                               make_double2(0.0, 0.0));
    }
    // one term: pad[c][dz.., dy.., dx..] += w * P[c][sz.., sy.., sx..] over nz x ny x nx
    auto term = [&](int c, int dz, int dy, int dx, int sz, int sy, int sx, int nz, int ny, int nx,
                    double w, int assign) {
        if (rc != EMF_OK || !P[c] || nz <= 0 || ny <= 0 || nx <= 0) return;
        BoxMap b{Py, Px, dz, dy, dx, Nj, Nk, sz, sy, sx, nz, ny, nx};
        const long long n = (long long)nz * ny * nx;
        hipLaunchKernelGGL(k_box_axpy, dim3(grid1d(n, 256)), dim3(256), 0, s->stream, pad[c], P[c],
                           b, w, assign);
    };
    // the whole plane at slice origins (z, y, x): I = 1, O = 0, M = 2
    auto full = [&](int c, int z, int y, int x, double w) { term(c, z, y, x, 0, 0, 0, Ni, Nj, Nk, w, 0); };
    const char bx = s->bc[0], by = s->bc[1];
    const double sgn = by == 'H' ? 1.0 : -1.0;
    const bool ymir = j1 == 0 && (by == 'H' || by == 'E');
    // Ex
    full(0, 1, 1, 1, 0.25); full(0, 1, 1, 0, 0.25); full(0, 2, 1, 1, 0.25); full(0, 2, 1, 0, 0.25);
    if (k1 == 0 && bx == 'H') {
        term(0, 1, 1, 1, 0, 0, 0, Ni, Nj, 1, 0.25, 0);
        term(0, 2, 1, 1, 0, 0, 0, Ni, Nj, 1, 0.25, 0);
    } else if (k1 == 0 && bx == 'E') {
        // as in the reference, this branch indexes the y axis
        term(0, 1, 1, 1, 0, 0, 0, Ni, 1, Nk, -0.25, 0);
        term(0, 2, 1, 1, 0, 0, 0, Ni, 1, Nk, -0.25, 0);
    }
    // Ey
    full(1, 1, 1, 1, 0.25); full(1, 1, 0, 1, 0.25); full(1, 2, 1, 1, 0.25); full(1, 2, 0, 1, 0.25);
    if (ymir) {
        term(1, 1, 1, 1, 0, 0, 0, Ni, 1, Nk, sgn * 0.25, 0);
        term(1, 2, 1, 1, 0, 0, 0, Ni, 1, Nk, sgn * 0.25, 0);
    }
    // Ez: assigned, not added
    term(2, 1, 1, 1, 0, 0, 0, Ni, Nj, Nk, 1.0, 1);
    // Hx
    full(3, 1, 1, 1, 0.5); full(3, 1, 0, 1, 0.5);
    if (ymir) term(3, 1, 1, 1, 0, 0, 0, Ni, 1, Nk, sgn * 0.5, 0);
    // Hy
    full(4, 1, 1, 1, 0.5); full(4, 1, 1, 0, 0.5);
    // Hz
    full(5, 1, 1, 1, 0.125); full(5, 1, 0, 1, 0.125); full(5, 1, 1, 0, 0.125);
    full(5, 1, 0, 0, 0.125); full(5, 2, 1, 1, 0.125); full(5, 2, 0, 1, 0.125);
    full(5, 2, 1, 0, 0.125); full(5, 2, 0, 0, 0.125);
    if (ymir) {
        term(5, 1, 1, 1, 0, 0, 0, Ni, 1, Nk, sgn * 0.125, 0);
        term(5, 1, 1, 0, 0, 0, 0, Ni, 1, Nk, sgn * 0.125, 0);
        term(5, 2, 1, 1, 0, 0, 0, Ni, 1, Nk, sgn * 0.125, 0);
        term(5, 2, 1, 0, 0, 0, 0, Ni, 1, Nk, sgn * 0.125, 0);
    }
    if (rc == EMF_OK && hipGetLastError() != hipSuccess) rc = EMF_ERR_HIP;
    // crop: grown cell (a, b, c) is pad[a + 1][b][c]; the block starts at
    // (id1, jd1, kd1) of the grown domain
    const int sI = place[6], sJ = place[7], sK = place[8];
    const long long cnt = (long long)sI * sJ * sK;
    if (rc == EMF_OK && cnt > 0) {
        if (place[3] < 0 || place[4] < 0 || place[5] < 0 || place[3] + sI > Ni + 1 ||
            place[4] + sJ > Nj + 1 || place[5] + sK > Nk + 1 || place[0] < 0 || place[1] < 0 ||
            place[2] < 0 || place[0] + sI > s->Nz || place[1] + sJ > s->Ny || place[2] + sK > s->Nx) {
            rc = EMF_ERR_ARG;
        } else {
            SourceDev d;
            d.i0 = place[0]; d.j0 = place[1]; d.k0 = place[2]; d.I = sI; d.J = sJ; d.K = sK;
            for (int c = 0; c < 6; ++c) d.f[c] = nullptr;
            for (int c = 0; c < 6 && rc == EMF_OK; ++c) {
                rc = dalloc(&d.f[c], (size_t)cnt);
                if (rc == EMF_OK)
                    hipLaunchKernelGGL(k_stage_block, dim3(grid1d(cnt, 256)), dim3(256), 0,
                                       s->stream, d.f[c], pad[c], Py, Px, place[3] + 1, place[4],
                                       place[5], sI, sJ, sK);
            }
            if (rc == EMF_OK && hipGetLastError() != hipSuccess) rc = EMF_ERR_HIP;
            if (rc == EMF_OK) s->staged.push_back(d);
            else for (int c = 0; c < 6; ++c) dfree(d.f[c]);
        }
    }
    (void)hipStreamSynchronize(s->stream);
    for (int c = 0; c < 6; ++c) dfree(pad[c]);
    if (rc == EMF_ERR_ARG) EMF_FAIL(rc, "source placement outside the grown domain or the grid");
    if (rc != EMF_OK) EMF_FAIL(rc, "staging a source block failed");
    return EMF_OK;
}

}  // namespace

int emf_stage_clear(emf_solver *s)
{
    LOCK(s);
    (void)hipStreamSynchronize(s->stream);
    free_blocks(s->staged);
    return EMF_OK;
}

int emf_stage_count(emf_solver *s, int *count)
{
    LOCK(s);
    if (!count) EMF_FAIL(EMF_ERR_ARG, "null output");
    *count = (int)s->staged.size();
    return EMF_OK;
}

int emf_stage_interp_T(emf_solver *s, int id, const double *scale, int mask, double rdiv, int k1,
                       int k2, int j1, int j2, int i1, int i2, const int *place)
{
    LOCK(s);
    if (id < 0 || id >= emf_solver::kWeightSets || !scale || !place)
        EMF_FAIL(EMF_ERR_ARG, "bad weight set or null argument");
    if (k1 > k2 || j1 > j2 || i1 > i2) EMF_FAIL(EMF_ERR_ARG, "empty or inverted domain");
    const int Nk = k2 - k1, Nj = j2 - j1, Ni = i2 - i1;
    const long long cnt = (long long)Nk * Nj * Ni;
    if (cnt != s->wcount[id])
        EMF_FAIL(EMF_ERR_ARG, "domain has %lld cells, weight set %d has %lld", cnt, id,
                 (long long)s->wcount[id]);
    if (cnt == 0) return EMF_OK;
    double2 *P[6] = {};
    int rc = EMF_OK;
    for (int q = 0; q < 6 && rc == EMF_OK; ++q) {
        if (!(mask >> q & 1) || !s->wplane[id][q]) continue;
        rc = dalloc(&P[q], (size_t)cnt);
        if (rc == EMF_OK)
            hipLaunchKernelGGL(k_plane_scale, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream, P[q],
                               s->wplane[id][q], scale[2 * q], scale[2 * q + 1], rdiv, cnt);
    }
    if (rc == EMF_OK && hipGetLastError() != hipSuccess) rc = EMF_ERR_HIP;
    if (rc == EMF_OK) rc = stage_transposed(s, P, Ni, Nj, Nk, k1, j1, place);
    (void)hipStreamSynchronize(s->stream);
    for (int q = 0; q < 6; ++q) dfree(P[q]);
    return rc;
}

int emf_stage_power_face(emf_solver *s, int slot, const int *comps, double s1, double mf,
                         double rp2, int k1, int k2, int j1, int j2, int i1, int i2, int ek1,
                         int ek2, int ej1, int ej2, int ei1, int ei2, const int *place)
{
    LOCK(s);
    if (!comps || !place) EMF_FAIL(EMF_ERR_ARG, "null argument");
    for (int q = 0; q < 4; ++q)
        if (comps[q] < 0 || comps[q] > 5) EMF_FAIL(EMF_ERR_ARG, "component must be 0..5");
    CHECK(check_interp_boxes(s, slot, k1, k2, j1, j2, i1, i2, ek1, ek2, ej1, ej2, ei1, ei2));
    const int Nk = k2 - k1, Nj = j2 - j1, Ni = i2 - i1;
    const long long cnt = (long long)Nk * Nj * Ni;
    if (cnt == 0) return EMF_OK;
    FaceArgs fa{};
    double2 *P[6] = {};
    int rc = EMF_OK;
    for (int q = 0; q < 4 && rc == EMF_OK; ++q) {
        fa.F[q] = s->res[slot][comps[q]];
        fa.st[q] = interp_stencil(s, comps[q], ek1, ej1);
        if (P[comps[q]]) { rc = EMF_ERR_ARG; break; }
        rc = dalloc(&P[comps[q]], (size_t)cnt);
        fa.out[q] = P[comps[q]];
    }
    if (rc == EMF_OK) {
        fa.s1 = s1; fa.mf = mf; fa.rp2 = rp2;
        const IBox dom{k1, k2, j1, j2, i1, i2}, ext{ek1, ek2, ej1, ej2, ei1, ei2};
        hipLaunchKernelGGL(k_face_dP, dim3(grid1d(cnt, 256)), dim3(256), 0, s->stream, fa, s->Nx,
                           s->Ny, dom, ext);
        if (hipGetLastError() != hipSuccess) rc = EMF_ERR_HIP;
    }
    if (rc == EMF_OK) rc = stage_transposed(s, P, Ni, Nj, Nk, k1, j1, place);
    (void)hipStreamSynchronize(s->stream);
    for (int q = 0; q < 6; ++q) dfree(P[q]);
    if (rc == EMF_ERR_ARG) EMF_FAIL(rc, "the four components must differ");
    return rc;
}

int emf_stage_register(emf_solver *s)
{
    LOCK(s);
    (void)hipStreamSynchronize(s->stream);
    free_blocks(s->srcs);
    for (const auto &b : s->staged) {
        SourceDev d = b;
        const long long cnt = (long long)b.I * b.J * b.K;
        int rc = EMF_OK;
        for (int c = 0; c < 6; ++c) d.f[c] = nullptr;
        for (int c = 0; c < 6 && rc == EMF_OK; ++c) {
            rc = dalloc(&d.f[c], (size_t)cnt);
            if (rc == EMF_OK &&
                hipMemcpyAsync(d.f[c], b.f[c], cnt * sizeof(double2), hipMemcpyDeviceToDevice,
                               s->stream) != hipSuccess)
                rc = EMF_ERR_HIP;
        }
        if (rc != EMF_OK) {
            for (int c = 0; c < 6; ++c) dfree(d.f[c]);
            s->src_dirty = true;
            EMF_FAIL(rc, "registering the staged sources failed");
        }
        s->srcs.push_back(d);
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->src_dirty = true;
    return EMF_OK;
}

int emf_stage_get(emf_solver *s, int idx, int *geom, double *out)
{
    LOCK(s);
    if (idx < 0 || idx >= (int)s->staged.size() || !geom)
        EMF_FAIL(EMF_ERR_ARG, "no staged block %d", idx);
    const SourceDev &d = s->staged[idx];
    geom[0] = d.i0; geom[1] = d.j0; geom[2] = d.k0; geom[3] = d.I; geom[4] = d.J; geom[5] = d.K;
    if (!out) return EMF_OK;
    const long long cnt = (long long)d.I * d.J * d.K;
    for (int c = 0; c < 6; ++c)
        HIPCHK(hipMemcpyAsync(out + 2 * c * cnt, d.f[c], cnt * sizeof(double2),
                              hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

static YdaxArgs ydax_args(const emf_solver *s)
{
    YdaxArgs y{};
    for (int c = 0; c < 6; ++c) { y.a[c] = s->res[1][c]; y.f[c] = s->res[0][c]; }
    for (int c = 0; c < 3; ++c) {
        y.eps[c] = s->eps[c];
        y.mu[c] = s->mu_uniform ? nullptr : s->mu[c];
    }
    y.mu_re = s->mu_re;
    return y;
}

int emf_calc_ydAx(emf_solver *s, const double *eps0_x, const double *eps0_y,
                  const double *eps0_z, const double *mu0_x, const double *mu0_y,
                  const double *mu0_z, int k1, int k2, int j1, int j2, int i1, int i2, double *out)
{
    LOCK(s);
    if (!out || !eps0_x || !eps0_y || !eps0_z) EMF_FAIL(EMF_ERR_ARG, "null argument");
    const bool with_mu = mu0_x || mu0_y || mu0_z;
    if (with_mu && !(mu0_x && mu0_y && mu0_z))
        EMF_FAIL(EMF_ERR_ARG, "give all three mu0 arrays or none");
    if (!s->res[0][0] || !s->res[1][0]) EMF_FAIL(EMF_ERR_STATE, "forward and adjoint results needed");
    if (k1 < 0 || j1 < 0 || i1 < 0 || k2 > s->Nx || j2 > s->Ny || i2 > s->Nz || k1 > k2 ||
        j1 > j2 || i1 > i2)
        EMF_FAIL(EMF_ERR_ARG, "box out of range");
    long long cnt = (long long)(k2 - k1) * (j2 - j1) * (i2 - i1);
    out[0] = out[1] = 0.0;
    if (cnt == 0) return EMF_OK;
    const int nblk = 1024;
    const int nz = with_mu ? 6 : 3;
    CHECK(ensure_scratch(s, nz * cnt + nblk));
    double2 *z = s->scratch, *part = s->scratch + nz * cnt;
    const double *e0[6] = { eps0_x, eps0_y, eps0_z, mu0_x, mu0_y, mu0_z };
    for (int c = 0; c < nz; ++c)
        HIPCHK(hipMemcpyAsync(z + c * cnt, e0[c], cnt * sizeof(double2), hipMemcpyHostToDevice,
                              s->stream));
    Geo g = s->geo();
    hipLaunchKernelGGL(k_ydAx, dim3(nblk), dim3(RED_THREADS), 0, s->stream, ydax_args(s), z,
                       z + cnt, z + 2 * cnt, with_mu ? z + 3 * cnt : nullptr,
                       with_mu ? z + 4 * cnt : nullptr, with_mu ? z + 5 * cnt : nullptr, g, k1,
                       k2 - k1, j1, j2 - j1, i1, i2 - i1, part);
    HIPCHK(hipGetLastError());
    std::vector<double2> h(nblk);
    HIPCHK(hipMemcpyAsync(h.data(), part, nblk * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double re = 0, im = 0;
    for (int b = 0; b < nblk; ++b) { re += h[b].x; im += h[b].y; }
    out[0] = re;
    out[1] = im;
    return EMF_OK;
}

int emf_grad_assemble(emf_solver *s, int64_t n, const int32_t *pidx, const int32_t *comp,
                      const int64_t *cell, const double *deps, int nparam, double step,
                      double *grad)
{
    LOCK(s);
    if (nparam <= 0 || !grad) EMF_FAIL(EMF_ERR_ARG, "nparam must be > 0");
    if (n < 0 || (n > 0 && (!pidx || !comp || !cell || !deps))) EMF_FAIL(EMF_ERR_ARG, "null entries");
    if (!(step != 0)) EMF_FAIL(EMF_ERR_ARG, "step must be non-zero");
    if (!s->res[0][0] || !s->res[1][0]) EMF_FAIL(EMF_ERR_STATE, "forward and adjoint results needed");
    std::vector<long long> seg(nparam + 1, 0);
    for (int64_t e = 0; e < n; ++e) {
        if (pidx[e] < 0 || pidx[e] >= nparam) EMF_FAIL(EMF_ERR_ARG, "parameter index out of range");
        if (e > 0 && pidx[e] < pidx[e - 1]) EMF_FAIL(EMF_ERR_ARG, "entries must be grouped by parameter");
        if (comp[e] < 0 || comp[e] > 5) EMF_FAIL(EMF_ERR_ARG, "component must be 0..5");
        if (cell[e] < 0 || cell[e] >= s->n) EMF_FAIL(EMF_ERR_ARG, "cell index out of range");
        seg[pidx[e] + 1]++;
    }
    for (int p = 0; p < nparam; ++p) seg[p + 1] += seg[p];
    long long *dseg = nullptr, *dcell = nullptr;
    int *dcomp = nullptr;
    double2 *ddeps = nullptr;
    double *dgrad = nullptr;
    int rc = EMF_OK;
    do {
        if ((rc = dalloc(&dseg, seg.size())) != EMF_OK) break;
        if ((rc = dalloc(&dgrad, (size_t)nparam)) != EMF_OK) break;
        if (n > 0) {
            if ((rc = dalloc(&dcell, (size_t)n)) != EMF_OK) break;
            if ((rc = dalloc(&dcomp, (size_t)n)) != EMF_OK) break;
            if ((rc = dalloc(&ddeps, (size_t)n)) != EMF_OK) break;
        }
        hipError_t e = hipMemcpy(dseg, seg.data(), seg.size() * sizeof(long long), hipMemcpyHostToDevice);
        if (e == hipSuccess && n > 0) e = hipMemcpy(dcell, cell, n * sizeof(long long), hipMemcpyHostToDevice);
        if (e == hipSuccess && n > 0) e = hipMemcpy(dcomp, comp, n * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess && n > 0) e = hipMemcpy(ddeps, deps, n * sizeof(double2), hipMemcpyHostToDevice);
        if (e != hipSuccess) { set_err("hipMemcpy failed: %s", hipGetErrorString(e)); rc = EMF_ERR_HIP; break; }
        hipLaunchKernelGGL(k_grad, dim3(nparam), dim3(RED_THREADS), 0, s->stream, ydax_args(s),
                           dseg, dcomp, dcell, ddeps, step, dgrad);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(grad, dgrad, nparam * sizeof(double), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) { set_err("gradient kernel failed: %s", hipGetErrorString(e)); rc = EMF_ERR_HIP; }
    } while (0);
    dfree(dseg); dfree(dcell); dfree(dcomp); dfree(ddeps); dfree(dgrad);
    return rc;
}

int emf_grad_assemble_noalloc(emf_solver *s, int64_t n, const int32_t *pidx, const int32_t *comp,
                              const int64_t *cell, const double *deps, int nparam, double step,
                              double *grad)
{
    LOCK(s);
    if (nparam <= 0 || !grad) EMF_FAIL(EMF_ERR_ARG, "nparam must be > 0");
    if (n < 0 || (n > 0 && (!pidx || !comp || !cell || !deps))) EMF_FAIL(EMF_ERR_ARG, "null entries");
    if (!(step != 0)) EMF_FAIL(EMF_ERR_ARG, "step must be non-zero");
    if (!s->res[0][0] || !s->res[1][0]) EMF_FAIL(EMF_ERR_STATE, "forward and adjoint results needed");
    if (s->dev < 0 || s->dev >= kMaxDevices) EMF_FAIL(EMF_ERR_ARG, "device %d out of range", s->dev);
    for (int64_t e = 0; e < n; ++e) {
        if (pidx[e] < 0 || pidx[e] >= nparam) EMF_FAIL(EMF_ERR_ARG, "parameter index out of range");
        if (e > 0 && pidx[e] < pidx[e - 1]) EMF_FAIL(EMF_ERR_ARG, "entries must be grouped by parameter");
        if (comp[e] < 0 || comp[e] > 5) EMF_FAIL(EMF_ERR_ARG, "component must be 0..5");
        if (cell[e] < 0 || cell[e] >= s->n) EMF_FAIL(EMF_ERR_ARG, "cell index out of range");
    }
    // layout: segments, gradient, cells, d_material, parameter ids, components
    const size_t o_seg = 0;
    const size_t o_grad = align_up(o_seg + (size_t(nparam) + 1) * sizeof(long long));
    const size_t o_cell = align_up(o_grad + size_t(nparam) * sizeof(double));
    const size_t o_deps = align_up(o_cell + size_t(n) * sizeof(long long));
    const size_t o_pidx = align_up(o_deps + size_t(n) * sizeof(double2));
    const size_t o_comp = align_up(o_pidx + size_t(n) * sizeof(int));
    const size_t need = o_comp + size_t(n) * sizeof(int);
    if (need > kNoAllocScratchBytes)
        EMF_FAIL(EMF_ERR_ARG, "%lld entries and %d parameters need %zu bytes of scratch; the "
                 "allocation-free entry holds %zu (use emf_grad_assemble)", (long long)n, nparam,
                 need, kNoAllocScratchBytes);
    std::lock_guard<std::mutex> g(g_noalloc_mu[s->dev]);
    unsigned char *base = nullptr;
    CHECK(noalloc_scratch(s->dev, &base));
    long long *dseg = reinterpret_cast<long long *>(base + o_seg);
    double *dgrad = reinterpret_cast<double *>(base + o_grad);
    long long *dcell = reinterpret_cast<long long *>(base + o_cell);
    double2 *ddeps = reinterpret_cast<double2 *>(base + o_deps);
    int *dpidx = reinterpret_cast<int *>(base + o_pidx);
    int *dcomp = reinterpret_cast<int *>(base + o_comp);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(dcell, cell, n * sizeof(long long), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(ddeps, deps, n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(dpidx, pidx, n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(dcomp, comp, n * sizeof(int), hipMemcpyHostToDevice, s->stream));
    }
    // every segment and every gradient entry is written below: nothing left
    // from an earlier call is read
    hipLaunchKernelGGL(k_segments, dim3(grid1d(nparam + 1, 256)), dim3(256), 0, s->stream, dpidx,
                       (long long)n, nparam, dseg);
    hipLaunchKernelGGL(k_grad, dim3(nparam), dim3(RED_THREADS), 0, s->stream, ydax_args(s), dseg,
                       dcomp, dcell, ddeps, step, dgrad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, dgrad, nparam * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EMF_OK;
}

int emf_pml_stretch(emf_solver *s, int64_t n, const int32_t *comp, const int64_t *cell,
                    double *out)
{
    LOCK(s);
    if (n < 0 || (n > 0 && (!comp || !cell || !out))) EMF_FAIL(EMF_ERR_ARG, "null entries");
    CHECK(check_ready(s));
    PassTab th, te;
    CHECK(refresh_tables(s, &th, &te));
    size_t off[3], len[3], total = 0;
    for (int a = 0; a < 3; ++a) {
        len[a] = std::max(1, s->w[2 * a] + s->w[2 * a + 1]);
        off[a] = total;
        total += 6 * len[a];
    }
    const Geo g = s->geo();
    const std::complex<double> ph = std::exp(std::complex<double>(0.0, s->dt));
    // 1/s = 1/kappa + c / (1 - b e^{i dt}): the CPML recursion psi = c D + b psi_prev
    // at steady state (phasors in the e^{-i w t} sense, one step = e^{i dt})
    auto stretch = [&](int a, bool htab, int par) {
        const double *t = &s->tab_host[off[a] + (htab ? 0 : 3 * len[a])];
        const double kap = t[par], b = t[len[a] + par], c = t[2 * len[a] + par];
        return 1.0 / (1.0 / kap + c / (1.0 - b * ph));
    };
    for (int64_t e = 0; e < n; ++e) {
        if (comp[e] < 0 || comp[e] > 5) EMF_FAIL(EMF_ERR_ARG, "component must be 0..5");
        if (cell[e] < 0 || cell[e] >= s->n) EMF_FAIL(EMF_ERR_ARG, "cell index out of range");
        const long long q = cell[e];
        const int k = (int)(q % s->Nx), j = (int)((q / s->Nx) % s->Ny), i = (int)(q / g.nxy);
        const Slab sl[3] = { slab_x(g, i, j, k), slab_y(g, i, j, k), slab_z(g, i, j, k) };
        std::complex<double> P = 1.0;
        for (int a = 0; a < 3; ++a)
            if (sl[a].in) P *= stretch(a, a == comp[e], sl[a].par);
        out[2 * e] = P.real();
        out[2 * e + 1] = P.imag();
    }
    return EMF_OK;
}

int emf_time_steps(emf_solver *s, int64_t n_start, int64_t nsteps, int64_t check_every,
                   int per_kernel, double *out)
{
    LOCK(s);
    if (!out || nsteps < 0 || n_start < 0) EMF_FAIL(EMF_ERR_ARG, "bad arguments");
    CHECK(check_ready(s));
    if (check_every > 0) CHECK(build_samples(s));
    if (fused(s)) CHECK(ensure_alt(s));
    // ev[0] start, ev[1] end.  With per_kernel, two passes: each launch of step q
    // records its own start and stop (hipExtLaunchKernelGGL), H in ev[2+4q],
    // ev[3+4q] and E in ev[4+4q], ev[5+4q]; fused kinds: ev[2+2q] / ev[3+2q]
    // bracket the launch of step q.
    const bool ext = per_kernel && !fused(s);
    std::vector<hipEvent_t> ev(ext ? 4 * nsteps + 2 : (per_kernel ? 2 * nsteps + 3 : 2));
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipMemsetAsync(s->conv, 0, sizeof(ConvState), s->stream));
    HIPCHK(hipEventRecord(ev[0], s->stream));
    const int start = s->cur;
    int rc = EMF_OK;
    auto mark = [&](size_t idx) {
        if (rc == EMF_OK && per_kernel)
            rc = hipEventRecord(ev[idx], s->stream) == hipSuccess ? EMF_OK : EMF_ERR_HIP;
    };
    for (int64_t q = 0; q < nsteps && rc == EMF_OK; ++q) {
        long long n = n_start + q;
        if (ext) {
            s->tev0 = ev[2 + 4 * q];
            s->tev1 = ev[3 + 4 * q];
            rc = launch_h(s, (double)n * s->dt, nullptr);
            s->tev0 = ev[4 + 4 * q];
            s->tev1 = ev[5 + 4 * q];
            if (rc == EMF_OK) rc = launch_e(s, ((double)n + 0.5) * s->dt, nullptr);
            s->tev0 = s->tev1 = nullptr;
        } else if (fused(s)) {
            mark(2 + 2 * q);
            rc = launch_fused(s, n, from_set(s, n, n_start, start), nullptr);
            mark(3 + 2 * q);
        } else {
            rc = launch_h(s, (double)n * s->dt, nullptr);
            if (rc == EMF_OK) rc = launch_e(s, ((double)n + 0.5) * s->dt, nullptr);
        }
        if (rc == EMF_OK && check_every > 0 && (q + 1) % check_every == 0) {
            const int ps = from_set(s, n + 1, n_start, start);
            hipLaunchKernelGGL(k_conv_check, dim3(1), dim3(64), 0, s->stream, s->Eb[ps][0],
                               s->Eb[ps][1], s->Eb[ps][2], s->samples, s->nsamples,
                               (double)n * s->dt, 1e-300, 1e-300, n, s->conv, (double *)nullptr, 0,
                               (int *)nullptr);
        }
    }
    if (rc == EMF_OK) rc = hipEventRecord(ev[1], s->stream) == hipSuccess ? EMF_OK : EMF_ERR_HIP;
    if (rc == EMF_OK) set_current(s, from_set(s, n_start + nsteps, n_start, start));
    if (rc == EMF_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = EMF_ERR_HIP;
    double tot = 0, th = 0, te = 0;
    if (rc == EMF_OK) {
        float ms;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        tot = ms;
        for (int64_t q = 0; ext && q < nsteps; ++q) {
            (void)hipEventElapsedTime(&ms, ev[2 + 4 * q], ev[3 + 4 * q]); th += ms;
            (void)hipEventElapsedTime(&ms, ev[4 + 4 * q], ev[5 + 4 * q]); te += ms;
        }
        for (int64_t q = 0; per_kernel && !ext && q < nsteps; ++q) {
            (void)hipEventElapsedTime(&ms, ev[2 + 2 * q], ev[3 + 2 * q]); th += ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc != EMF_OK) EMF_FAIL(rc, "timed stepping failed");
    out[0] = tot; out[1] = th; out[2] = te;
    return EMF_OK;
}

int emf_set_streaming(emf_solver *s, int mode)
{
    LOCK(s);
    if (mode < -1 || mode > 1) EMF_FAIL(EMF_ERR_ARG, "mode must be -1, 0 or 1");
    s->psi_stream = mode;
    return EMF_OK;
}

int emf_set_tuning(emf_solver *s, int kind, int tyg, int zc, int order)
{
    LOCK(s);
    if (kind < 0 || kind > 4)
        EMF_FAIL(EMF_ERR_ARG, "kind must be 0 (fused, row recompute), 1 (two passes), 2 (fused, LDS), "
                              "3 (two passes, 5 waves per SIMD) or 4 (two passes, the H pass's "
                              "y-neighbour row through LDS)");
    const bool ok = kind == 2 ? (tyg == 4 || tyg == 8 || tyg == 16)
                              : (tyg == 1 || tyg == 2 || tyg == 4 || tyg == 8);
    if (!ok) EMF_FAIL(EMF_ERR_ARG, "tyg must be 1, 2, 4 or 8 (kinds 0, 1) or 4, 8, 16 (kind 2)");
    if (zc < 1) EMF_FAIL(EMF_ERR_ARG, "zc must be >= 1");
    if (order < 0 || order > 3) EMF_FAIL(EMF_ERR_ARG, "order must be 0, 1, 2 or 3");
    s->kind = kind;
    s->tyg = tyg;
    s->zc = zc;
    s->order = order;
    return EMF_OK;
}

int emf_set_wide_offsets(emf_solver *s, int force)
{
    LOCK(s);
    if (force != 0 && force != 1) EMF_FAIL(EMF_ERR_ARG, "force must be 0 or 1");
    s->wide_force = force;
    return EMF_OK;
}

int emf_set_h_shape(emf_solver *s, int tyg, int zc)
{
    LOCK(s);
    if (!(tyg == 0 || tyg == 1 || tyg == 2 || tyg == 4 || tyg == 8))
        EMF_FAIL(EMF_ERR_ARG, "tyg must be 0 (as the E pass), 1, 2, 4 or 8");
    if (zc < 0) EMF_FAIL(EMF_ERR_ARG, "zc must be >= 0 (0: as the E pass)");
    s->tyg_h = tyg;
    s->zc_h = zc;
    return EMF_OK;
}

int emf_sample_structured(int device, const emf_prim *prims, int nprims, const double *vx,
                          const double *vy, int nverts, double dx, double dy, double dz,
                          int supersample, double bg_re, double bg_im, int k1, int k2, int j1,
                          int j2, int i1, int i2, double sx, double sy, double sz, double *out)
{
    static_assert(sizeof(emf_prim) == sizeof(MatPrim), "emf_prim and MatPrim layouts differ");
    if (!out || nprims < 0 || (nprims > 0 && !prims) || nverts < 0 || supersample < 0 ||
        supersample > 64 || !(dx > 0) || !(dy > 0) || !(dz > 0) || k2 <= k1 || j2 <= j1 ||
        i2 <= i1)
        EMF_FAIL(EMF_ERR_ARG, "bad arguments");
    for (int q = 0; q < nprims; ++q) {
        const emf_prim &p = prims[q];
        if (p.kind < 0 || p.kind > 2) EMF_FAIL(EMF_ERR_ARG, "primitive %d: unknown kind", q);
        if (supersample == 0 && p.kind == 1)
            EMF_FAIL(EMF_ERR_ARG, "primitive %d: the exact overlap rule takes ellipses as polygons", q);
        if (p.kind == 2 && (p.v0 < 0 || p.nv < 0 || (long long)p.v0 + p.nv > nverts || !vx || !vy))
            EMF_FAIL(EMF_ERR_ARG, "primitive %d: vertices out of range", q);
    }
    const int nk = k2 - k1, nj = j2 - j1, ni = i2 - i1;
    const long long ncell = (long long)ni * nj * nk;
    // z segments as the host sampler forms them: consecutive distinct primitive
    // z bounds, skipped when no cell of the box overlaps them
    std::vector<double> b;
    for (int q = 0; q < nprims; ++q) {
        if (!(prims[q].z2 > prims[q].z1)) continue;
        b.push_back(prims[q].z1);
        b.push_back(prims[q].z2);
    }
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    std::vector<double> zab;
    std::vector<int> off(1, 0), sp;
    for (size_t s = 0; s + 1 < b.size(); ++s) {
        const double za = b[s], zb = b[s + 1];
        bool used = false;
        for (int i = i1; i < i2 && !used; ++i) {
            const double zc = ((double)i + sz) * dz;
            const double zlo = zc - dz / 2, zhi = zc + dz / 2;
            used = std::max(std::min(zhi, zb) - std::max(zlo, za), 0.0) / dz > 0;
        }
        if (!used) continue;
        zab.push_back(za);
        zab.push_back(zb);
        for (int q = 0; q < nprims; ++q)
            if (prims[q].z2 > prims[q].z1 && prims[q].z1 <= za && prims[q].z2 >= zb) sp.push_back(q);
        off.push_back((int)sp.size());
    }
    const int nseg = (int)zab.size() / 2;
    HIPCHK(hipSetDevice(device));
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    MatPrim *dp = nullptr;
    double *dvx = nullptr, *dvy = nullptr, *dz_ab = nullptr;
    int *doff = nullptr, *dsp = nullptr, *derr = nullptr;
    double2 *v2 = nullptr, *dout = nullptr;
    int rc = EMF_OK;
    auto ok = [&](int r) { if (rc == EMF_OK && r != EMF_OK) rc = r; return rc == EMF_OK; };
    if (ok(dalloc(&dout, (size_t)ncell)) && nseg > 0 && ok(dalloc(&dp, (size_t)nprims)) &&
        ok(dalloc(&doff, off.size())) && ok(dalloc(&dsp, std::max<size_t>(1, sp.size()))) &&
        ok(dalloc(&dz_ab, zab.size())) && ok(dalloc(&v2, (size_t)nseg * nj * nk)) &&
        ok(dalloc(&derr, 1)) &&
        (nverts == 0 || (ok(dalloc(&dvx, (size_t)nverts)) && ok(dalloc(&dvy, (size_t)nverts))))) {
        if (hipMemcpyAsync(dp, prims, nprims * sizeof(MatPrim), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(doff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
            (!sp.empty() && hipMemcpyAsync(dsp, sp.data(), sp.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess) ||
            hipMemcpyAsync(dz_ab, zab.data(), zab.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
            (nverts > 0 && (hipMemcpyAsync(dvx, vx, nverts * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
                            hipMemcpyAsync(dvy, vy, nverts * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)))
            rc = EMF_ERR_HIP;
        if (rc == EMF_OK && hipMemsetAsync(derr, 0, sizeof(int), st) != hipSuccess) rc = EMF_ERR_HIP;
        if (rc == EMF_OK) {
            const long long nv2 = (long long)nseg * nj * nk;
            hipLaunchKernelGGL(k_mat_inplane, dim3(grid1d(nv2, 256)), dim3(256), 0, st, dp, dvx, dvy,
                               doff, dsp, nseg, k1, nk, j1, nj, sx, sy, dx, dy, supersample, bg_re,
                               bg_im, v2, derr);
            if (hipGetLastError() != hipSuccess) rc = EMF_ERR_HIP;
        }
    }
    int cell_err = 0;
    if (rc == EMF_OK && derr &&
        (hipMemcpyAsync(&cell_err, derr, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
        rc = EMF_ERR_HIP;
    if (rc == EMF_OK) {
        hipLaunchKernelGGL(k_mat_combine, dim3(grid1d(ncell, 256)), dim3(256), 0, st, v2, dz_ab, nseg,
                           i1, ni, nj, nk, sz, dz, bg_re, bg_im, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, (size_t)ncell * sizeof(double2), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = EMF_ERR_HIP;
    }
    (void)hipStreamSynchronize(st);
    dfree(dp); dfree(dvx); dfree(dvy); dfree(dz_ab); dfree(doff); dfree(dsp); dfree(v2); dfree(dout);
    dfree(derr);
    (void)hipStreamDestroy(st);
    if (rc != EMF_OK) EMF_FAIL(rc, "material sampling failed");
    if (cell_err)
        EMF_FAIL(EMF_ERR_ARG, "a cell meets more than %d primitives or %d polygon edges "
                 "(exact overlap rule)", MAT_MAXU, MAT_MAXE);
    return EMF_OK;
}

int emf_mem_info(int device, int64_t *free_bytes, int64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) EMF_FAIL(EMF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return EMF_OK;
}

int emf_copy_bandwidth(int device, int64_t bytes, int reps, double *gbps)
{
    if (!gbps || bytes < 1024 || reps < 1) EMF_FAIL(EMF_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(device));
    long long n = bytes / (long long)sizeof(double2);
    if ((n + 255) / 256 > 0x7fffffffLL) EMF_FAIL(EMF_ERR_ARG, "bytes too large");
    double2 *a = nullptr, *b = nullptr;
    CHECK(dalloc(&a, n));
    if (dalloc(&b, n) != EMF_OK) { dfree(a); return EMF_ERR_NOMEM; }
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = EMF_OK;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { dfree(a); dfree(b); EMF_FAIL(EMF_ERR_HIP, "stream"); }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 0, n * sizeof(double2), st) != hipSuccess)
        rc = EMF_ERR_HIP;
    // one warm-up launch, then `reps` launches back to back between two
    // events: the average per launch, without per-launch event overhead (a
    // copy of a few MB takes a few microseconds)
    const unsigned grid = (unsigned)((n + 255) / 256);
    float ms = 0;
    if (rc == EMF_OK) {
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, a, b, n);
        if (hipEventRecord(e0, st) != hipSuccess) rc = EMF_ERR_HIP;
        for (int r = 0; r < reps && rc == EMF_OK; ++r)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, a, b, n);
        if (rc == EMF_OK &&
            (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
             hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = EMF_ERR_HIP;
    }
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    dfree(a); dfree(b);
    if (rc != EMF_OK) EMF_FAIL(rc, "copy kernel failed");
    *gbps = 2.0 * n * sizeof(double2) * reps / (ms * 1e-3) / 1e9;
    return EMF_OK;
}

}  // extern "C"

#include "batch_host.hpp"
