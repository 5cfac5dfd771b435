// Batched solves: many independent CW solves of different sizes advance
// together.  One launch per half step covers every solve in the batch: a
// persistent grid walks a device work table of tiles (all solves' columns,
// back to back).  Each solve's whole __solve loop (emopt/fdtd.py:994-1194) runs
// on the device as a small state machine advanced once per step by
// k_batch_tick, so a chunk of steps can be captured once into a hipGraph and
// replayed with no host work per step; the host only reads the solves' states
// between replays, retires finished solves and admits new ones.
//
// Per step:  k_batch_h   H half step of every tile (storing the t0 / t1
//                        snapshot for solves at those points of their loop)
//            k_batch_e   E half step of every tile
//            k_batch_tick  per solve: step count, convergence check every Tn
//                        steps, choice of what the next step does
// and once per graph replay, after its steps:
//            k_batch_phasor  phasor pass of the solves that reached it (they
//                        wait, skipped by every kernel, until then)
//            k_batch_finish  marks those solves finished
//
// Arithmetic per cell is that of k_pass_h / k_pass_e / k_phasor3 with the same
// host-computed source factors, so a solve's result agrees with the
// single-solve path to roundoff and does not depend at all on what else is in
// the batch.
#pragma once
#include "twopass.hpp"

namespace emf {

constexpr int BATCH_MAX = 128;   // solves in one batch
constexpr int BTY = 4;           // rows per work-group of the batch passes
// host-computed factors per step n (t = n dt, te = (n + 1/2) dt):
// sin t, cos t, env(t), sin te, cos te, env(te), sin(t - dt/2), cos(t - dt/2)
constexpr int NFAC = 8;

struct BMember {                  // written by the host when the solve is admitted
    FusedArgs A;                  // in place: Eo == En, Ho == Hn, pho == phn
    Geo g;
    const SrcBlk *blk;
    int nsrc;
    int sub[6];                   // union box of the source blocks (PassSrc::ub)
    int ntx, nty, ntz;            // tile grid of this solve
    const double *fac;            // [nfac][NFAC]
    long long nfac;
    double *res;                  // result slot: double2[6 n] viewed as doubles
    long long n;                  // cells
    long long n_stop;             // first step index past Nmax
    long long fixed_n0;           // >= 0: run exactly this many steps before the snapshots
    long long Tn;
    double amp_rtol, phi_rtol, dt;
    const long long *samples;
    int nsamples;
    int hist_cap;
    ConvState *conv;
    double *hist;
};

enum { BM_STEP = 0, BM_CAP0 = 1, BM_CAP1 = 2, BM_PHASOR = 3, BM_IDLE = 4 };
enum { BP_CONV = 0, BP_RUN1 = 1, BP_RUN2 = 2, BP_DONE = 3 };

struct BState {                   // advanced on the device by k_batch_tick
    long long step;               // steps completed
    long long n0;                 // step count at the t0 snapshot
    int mode;                     // what the next H pass does (BM_*)
    int phase;                    // where the solve is in its loop (BP_*)
    int status;                   // 0 converged, 1 Nmax, 2 fixed step count
    int nchecks;                  // convergence checks so far
    double a_change, phi_change;  // last check's changes
    PhTimes te, th;               // phasor sample times (set with BM_PHASOR)
};

struct BTable {                   // the tiles of the solves in the batch
    int nact;                     // entries
    int total;                    // tiles over all entries
    int mem[BATCH_MAX];           // member slot of entry e
    int off[BATCH_MAX + 1];       // first tile of entry e
    // Entries are grouped (consecutive runs) so that a group's hot state fits
    // the Infinity Cache; a graph replay steps the group `cur` only.
    int ngroups;
    int cur;
    int gent[BATCH_MAX + 1];      // first entry of group g
};

// the tile range [t_lo, t_hi) and entry range [e_lo, e_hi) of the current group
__device__ __forceinline__ void batch_group(const BTable *__restrict__ T, int &t_lo, int &t_hi,
                                            int &e_lo, int &e_hi)
{
    const int g = T->cur;
    e_lo = T->gent[g];
    e_hi = T->gent[g + 1];
    t_lo = T->off[e_lo];
    t_hi = T->off[e_hi];
}

__global__ void k_batch_set_group(BTable *T, int g)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) T->cur = g;
}

// entry e with off[e] <= t < off[e + 1] (t < total)
__device__ __forceinline__ int batch_entry(const BTable *__restrict__ T, int t)
{
    int lo = 0, hi = T->nact - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T->off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// tile -> column of member B; false when this thread's column is off the grid
__device__ __forceinline__ bool batch_column(const BMember &B, int lt, int &k, int &j, int &z0,
                                             int &z1)
{
    const int bx = lt % B.ntx, by = (lt / B.ntx) % B.nty, bz = lt / (B.ntx * B.nty);
    k = bx * PX + threadIdx.x;
    j = __builtin_amdgcn_readfirstlane(by * BTY + (int)threadIdx.y);   // one row per wave
    z0 = bz * B.g.zc;
    z1 = min(z0 + B.g.zc, B.g.Nz);
    return k < B.g.Nx && j < B.g.Ny;
}

// 3-sample phasors of one column (k_phasor3's arithmetic), in place in the
// result slot, whose halves hold the t0 and t1 snapshots
__device__ __forceinline__ void phasor_column(const BMember &B, const BState &S, int k, int j,
                                              int z0, int z1)
{
    double2 *r = reinterpret_cast<double2 *>(B.res);
    const long long n = B.n;
    for (int i = z0; i < z1; ++i) {
        const long long gc = ((long long)i * B.g.Ny + j) * B.g.Nx + k;
        const unsigned o = (unsigned)(gc * 8);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double f = c < 3 ? ldo(B.A.Eo, o + (unsigned)c * B.A.cs)
                                   : ldo(B.A.Ho, o + (unsigned)(c - 3) * B.A.cs);
            const double2 v = ld2(r + c * n + gc);
            st2(r + c * n + gc, phasor3(v.x, v.y, f, c < 3 ? S.te : S.th));
        }
    }
}

// Each work-group takes one contiguous chunk of the tile range, so the member
// of a tile changes a handful of times per work-group: its descriptor and
// state are looked up when it changes, not per tile, and the tiles of an idle
// member are skipped in one jump.
struct BCursor {
    int e, next, m, mode;
    long long step;
};

__device__ __forceinline__ void batch_seek(const BTable *__restrict__ T,
                                           const BState *__restrict__ S, int t, BCursor &c)
{
    c.e = batch_entry(T, t);
    c.next = T->off[c.e + 1];
    c.m = T->mem[c.e];
    c.mode = S[c.m].mode;
    c.step = S[c.m].step;
}

__device__ __forceinline__ void batch_advance(const BTable *__restrict__ T,
                                              const BState *__restrict__ S, int t, BCursor &c)
{
    while (t >= c.next) {
        ++c.e;
        c.next = T->off[c.e + 1];
        c.m = T->mem[c.e];
        c.mode = S[c.m].mode;
        c.step = S[c.m].step;
    }
}

// backwards: the entry holding tile t when t only decreases
__device__ __forceinline__ void batch_retreat(const BTable *__restrict__ T,
                                              const BState *__restrict__ S, int t, BCursor &c)
{
    while (t < T->off[c.e]) {
        --c.e;
        c.next = T->off[c.e + 1];
        c.m = T->mem[c.e];
        c.mode = S[c.m].mode;
        c.step = S[c.m].step;
    }
}

__global__ __launch_bounds__(PX * BTY) void k_batch_h(const BTable *__restrict__ T,
                                                      const BMember *__restrict__ M,
                                                      const BState *__restrict__ S)
{
    int t_lo, t_hi, e_lo, e_hi;
    batch_group(T, t_lo, t_hi, e_lo, e_hi);
    const int chunk = (t_hi - t_lo + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = t_lo + (int)blockIdx.x * chunk, t1 = min(t0 + chunk, t_hi);
    if (t0 >= t1) return;
    BCursor c;
    batch_seek(T, S, t0, c);
    for (int t = t0; t < t1; ++t) {
        batch_advance(T, S, t, c);
        if (c.mode == BM_IDLE || c.mode == BM_PHASOR) { t = c.next - 1; continue; }
        const BMember &B = M[c.m];
        int k, j, z0, z1;
        const bool on = batch_column(B, t - T->off[c.e], k, j, z0, z1);
        const unsigned fo = (unsigned)(c.step * NFAC) * 8u;   // step factors, read-only here
        PassSrc ps;
        ps.n = B.nsrc; ps.comp0 = 3; ps.blk = B.blk; ps.ninl = 0;
        for (int a = 0; a < 6; ++a) ps.ub[a] = B.sub[a];
        ps.st = ldc(B.fac, fo + 0u); ps.ct = ldc(B.fac, fo + 8u); ps.env = ldc(B.fac, fo + 16u);
        if (!on) continue;
        const unsigned long long sm = col_src_mask(ps, k, j, z0, z1);
        double *cap = c.mode == BM_CAP0 ? B.res : (c.mode == BM_CAP1 ? B.res + 1 : nullptr);
        h_column<false>(B.A, B.g, ps, k, j, z0, z1, cap, B.n, sm);
    }
}

// The E pass takes the chunks in reverse order and walks each backwards, so
// it starts on the tiles the H pass wrote last (still in the Infinity Cache),
// as the single-solve E pass does with work-group order 3.
__global__ __launch_bounds__(PX * BTY) void k_batch_e(const BTable *__restrict__ T,
                                                      const BMember *__restrict__ M,
                                                      const BState *__restrict__ S)
{
    int t_lo, t_hi, e_lo, e_hi;
    batch_group(T, t_lo, t_hi, e_lo, e_hi);
    const int chunk = (t_hi - t_lo + (int)gridDim.x - 1) / (int)gridDim.x;
    // chunks in reverse dispatch order: the first work-groups to run take the
    // chunks whose H pass ran last
    const int w = (int)gridDim.x - 1 - (int)blockIdx.x;
    const int t0 = t_lo + w * chunk, t1 = min(t0 + chunk, t_hi);
    if (t0 >= t1) return;
    BCursor c;
    batch_seek(T, S, t1 - 1, c);
    for (int t = t1 - 1; t >= t0; --t) {
        batch_retreat(T, S, t, c);
        if (c.mode == BM_IDLE || c.mode == BM_PHASOR) { t = T->off[c.e]; continue; }
        const BMember &B = M[c.m];
        int k, j, z0, z1;
        const bool on = batch_column(B, t - T->off[c.e], k, j, z0, z1);
        const unsigned fo = (unsigned)(c.step * NFAC) * 8u;   // step factors, read-only here
        PassSrc ps;
        ps.n = B.nsrc; ps.comp0 = 0; ps.blk = B.blk; ps.ninl = 0;
        for (int a = 0; a < 6; ++a) ps.ub[a] = B.sub[a];
        ps.st = ldc(B.fac, fo + 24u); ps.ct = ldc(B.fac, fo + 32u); ps.env = ldc(B.fac, fo + 40u);
        if (!on) continue;
        const unsigned long long sm = col_src_mask(ps, k, j, z0, z1);
        if (B.A.bI && z1 - z0 <= 2) e_column<CM_INDEX, 2>(B.A, B.g, ps, k, j, z0, z1, sm);
        else if (B.A.bI) e_column<CM_INDEX, 0>(B.A, B.g, ps, k, j, z0, z1, sm);
        else e_column<CM_REAL, 0>(B.A, B.g, ps, k, j, z0, z1, sm);
    }
}

// The phasor pass of the solves that reached it (the H and E kernels skip
// them), once at the end of a graph replay.  A kernel of its own so that its
// transcendental math does not set the register budget, and so the
// occupancy, of the H and E passes.
__global__ __launch_bounds__(PX * BTY) void k_batch_phasor(const BTable *__restrict__ T,
                                                           const BMember *__restrict__ M,
                                                           const BState *__restrict__ S)
{
    int t_lo, t_hi, e_lo, e_hi;
    batch_group(T, t_lo, t_hi, e_lo, e_hi);
    const int chunk = (t_hi - t_lo + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = t_lo + (int)blockIdx.x * chunk, t1 = min(t0 + chunk, t_hi);
    if (t0 >= t1) return;
    BCursor c;
    batch_seek(T, S, t0, c);
    for (int t = t0; t < t1; ++t) {
        batch_advance(T, S, t, c);
        if (c.mode != BM_PHASOR) { t = c.next - 1; continue; }
        const BMember &B = M[c.m];
        int k, j, z0, z1;
        if (!batch_column(B, t - T->off[c.e], k, j, z0, z1)) continue;
        phasor_column(B, S[c.m], k, j, z0, z1);
    }
}

// One step of every solve's loop (emopt/fdtd.py:1049-1194), one wave per
// table entry: every lane follows the member's state machine (the same
// values), the convergence check is shared by the wave (conv_check), and
// lane 0 stores the new state.
__global__ void k_batch_tick(const BTable *__restrict__ T, const BMember *__restrict__ M,
                             BState *__restrict__ S)
{
    int t_lo, t_hi, e_lo, e_hi;
    batch_group(T, t_lo, t_hi, e_lo, e_hi);
    const int e = e_lo + (int)blockIdx.x;
    if (e >= e_hi || threadIdx.x >= 64) return;
    const int m = T->mem[e];
    const BMember &B = M[m];
    BState st = S[m];
    if (st.mode == BM_IDLE || st.mode == BM_PHASOR) return;   // PHASOR: waits for k_batch_finish
    const long long nd = ++st.step;  // steps completed
    int mode = BM_STEP;
    if (st.phase == BP_CONV) {
        bool fin = false;
        if (B.fixed_n0 >= 0) {
            if (nd >= B.fixed_n0) { fin = true; st.status = 2; }
        } else {
            if (nd % B.Tn == 0) {
                const double *Ex = B.A.Eo;
                const double *Ey = reinterpret_cast<const double *>(
                    reinterpret_cast<const char *>(B.A.Eo) + B.A.cs);
                const double *Ez = reinterpret_cast<const double *>(
                    reinterpret_cast<const char *>(B.A.Eo) + 2 * (size_t)B.A.cs);
                int nck;
                double ac, pc;
                const int conv = conv_check(Ex, Ey, Ez, B.samples, B.nsamples,
                                            (double)(nd - 1) * B.dt, B.amp_rtol, B.phi_rtol,
                                            nd - 1, B.conv, B.hist, B.hist_cap, &nck, &ac, &pc);
                st.nchecks = nck;
                st.a_change = ac;
                st.phi_change = pc;
                if (conv) { fin = true; st.status = 0; }
            }
            if (!fin && nd >= B.n_stop) { fin = true; st.status = 1; }
        }
        if (fin) { st.n0 = nd; st.phase = BP_RUN1; mode = BM_CAP0; }
    } else if (st.phase == BP_RUN1) {
        if (nd == st.n0 + B.Tn) { st.phase = BP_RUN2; mode = BM_CAP1; }
    } else if (st.phase == BP_RUN2) {
        if (nd == st.n0 + 2 * B.Tn) {
            const auto *f0 = gptr(B.fac + st.n0 * NFAC), *f1 = gptr(B.fac + (st.n0 + B.Tn) * NFAC),
                       *f2 = gptr(B.fac + (st.n0 + 2 * B.Tn) * NFAC);
            st.te.s21 = f2[0] - f1[0]; st.te.s10 = f1[0] - f0[0];
            st.te.c10 = f1[1] - f0[1]; st.te.c21 = f2[1] - f1[1];
            st.th.s21 = f2[6] - f1[6]; st.th.s10 = f1[6] - f0[6];
            st.th.c10 = f1[7] - f0[7]; st.th.c21 = f2[7] - f1[7];
            mode = BM_PHASOR;
        }
    }
    st.mode = mode;
    if (threadIdx.x == 0) S[m] = st;
}

// After the replay's phasor pass: the solves that took it are finished.
__global__ void k_batch_finish(const BTable *__restrict__ T, BState *__restrict__ S)
{
    int t_lo, t_hi, e_lo, e_hi;
    batch_group(T, t_lo, t_hi, e_lo, e_hi);
    const int e = e_lo + (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
    if (e >= e_hi) return;
    const int m = T->mem[e];
    if (S[m].mode == BM_PHASOR) {
        S[m].mode = BM_IDLE;
        S[m].phase = BP_DONE;
    }
}

}  // namespace emf
