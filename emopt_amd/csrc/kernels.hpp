// Device kernels of the CW-FDTD hot path for gfx950 (MI355X), FP64.
//
// Layout in HBM: every array is unpadded, [Nz][Ny][Nx] with x fastest, cell
// g = (i*Ny + j)*Nx + k, the reference's "global" index (src/fdtd.cpp:183).
// There are no ghost layers: the boundary values the reference keeps in its
// padded arrays are produced inside the kernels:
//   * H pass: E beyond a max face reads as 0 (PEC ghosts, src/fdtd.cpp:142-173)
//   * E pass: H beyond a min face reads as 0 / -H(first) / +H(first) for
//     bc '0' / 'E' / 'H' (src/fdtd.cpp:397-501)
//
// Thread mapping: a 64x4 (x, y) work-group owns a tile of columns and sweeps
// a chunk of z planes, keeping the z-neighbour it needs in registers (the H
// pass carries E(i+1) -> E(i); the E pass carries H(i) -> H(i-1)), so every
// HBM byte of the interior stencil is loaded once per pass.  In-plane
// neighbours (x+1/x-1, y+1/y-1) are re-read through L1/L2.  Work-group ids are
// remapped so that each XCD (own L2) receives a contiguous range of tiles.
//
// Arithmetic follows the reference's operation order term by term, so the
// only differences to the CPU are FMA contraction and libm last-bit effects.
#pragma once
#include <hip/hip_runtime.h>

namespace emf {

constexpr int TX = 64;          // lanes along x: one wavefront row
constexpr int TY = 4;           // rows along y per work-group
constexpr int NCONV_MAX = 1024;
constexpr int NUM_XCD = 8;

struct Geo {
    int Nx, Ny, Nz;
    long long nxy;
    double rx, ry, rz;   // R/dx, R/dy, R/dz
    double dt;
    int w[6];            // CPML widths xmin xmax ymin ymax zmin zmax
    int bc[3];           // '0', 'E', 'H' per axis x, y, z
    int zc;              // z planes per work-group
    int order;           // work-group order of the two passes (see pass_tile)
};

// CPML coefficient tables of one pass (H uses the *_H_* set, E the *_E_*).
struct PassTab {
    const double *kap[3], *b[3], *c[3];
};



// One source block (the reference's SourceArray, src/fdtd.hpp:158-161) as a
// device-resident descriptor.  Per cell and component it holds
// (A cos(phi), A sin(phi)) of the reference's amplitude A and phase phi, so
// that the time factor sin(t + phi) * A = sin(t) A cos(phi) + cos(t) A sin(phi)
// needs no trigonometry on the device.
struct SrcBlk {
    int i0, j0, k0, I, J, K;
    const double2 *f[6];
};

// Global-memory views of device pointers (every array here lives in device
// global memory): accesses through them are global, not flat, also where the
// pointer itself was loaded from memory rather than passed as an argument.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// double2 through the global address space (a native two-element vector,
// which binds where HIP_vector_type's copy constructor would not)
typedef double gvec2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 ld2(const double2 *p)
{
    const gvec2 v = *(const __attribute__((address_space(1))) gvec2 *)p;
    return make_double2(v.x, v.y);
}

__device__ __forceinline__ void st2(double2 *p, double2 v)
{
    gvec2 w;
    w.x = v.x;
    w.y = v.y;
    *(__attribute__((address_space(1))) gvec2 *)p = w;
}

// Source block q, read through the constant address space: descriptors are
// written before a launch and only read while it runs, and all lanes read the
// same one, so the backend loads it with scalar loads.
__device__ __forceinline__ SrcBlk src_blk(const SrcBlk *blk, int q)
{
    typedef const __attribute__((address_space(4))) int cint;
    typedef const double2 *dptr;
    typedef const __attribute__((address_space(4))) dptr cdptr;
    const SrcBlk *p = blk + q;
    SrcBlk s;
    s.i0 = *(cint *)&p->i0; s.j0 = *(cint *)&p->j0; s.k0 = *(cint *)&p->k0;
    s.I = *(cint *)&p->I; s.J = *(cint *)&p->J; s.K = *(cint *)&p->K;
#pragma unroll
    for (int c = 0; c < 6; ++c) s.f[c] = *(cdptr *)&p->f[c];
    return s;
}

// Sources of one pass: n descriptors in device memory evaluated at time t:
// st = sin(t), ct = cos(t), env = ramp envelope (exactly 1 after the ramp)
// of src_func_t (src/fdtd.cpp:1595-1602).  Blocks are applied in order, as
// the reference loops over its source list.
// The boxes of the first SRC_INLINE blocks, and the union of all blocks,
// travel with the kernel arguments, so that a column decides which blocks
// can meet it without loading a descriptor.
constexpr int SRC_INLINE = 8;

struct PassSrc {
    int n;
    int comp0;
    double st, ct;
    double env;
    const SrcBlk *blk;
    int ninl;                     // inline boxes (the first min(n, SRC_INLINE) blocks, or 0)
    int ub[6];                    // union of all blocks: i0, i1, j0, j1, k0, k1 (ends excluded)
    int ib[SRC_INLINE][6];        // i0, I, j0, J, k0, K of the first ninl blocks
};

// src_func_t(t, phi) * A for one cell
__device__ __forceinline__ double src_val(const PassSrc &p, double2 v)
{
    return p.env * (p.st * v.x + p.ct * v.y);
}

// The source blocks that can meet column (k, j) over planes [z0, z1): bit q
// for block q < 64.  Nothing outside the union box; among the inline boxes
// the ones that meet the column; every block from ninl on (tested cell by
// cell against its descriptor).  Register compares only.
__device__ __forceinline__ unsigned long long col_src_mask(const PassSrc &ps, int k, int j,
                                                           int z0, int z1)
{
    if (ps.n == 0 || k < ps.ub[4] || k >= ps.ub[5] || j < ps.ub[2] || j >= ps.ub[3] ||
        z1 <= ps.ub[0] || z0 >= ps.ub[1])
        return 0ULL;
    unsigned long long m = 0;
#pragma unroll
    for (int q = 0; q < SRC_INLINE; ++q) {
        if (q < ps.ninl) {
            const int *b = ps.ib[q];
            if (k >= b[4] && k < b[4] + b[5] && j >= b[2] && j < b[2] + b[3] && b[0] < z1 &&
                b[0] + b[1] > z0)
                m |= 1ULL << q;
        }
    }
    const int hi = ps.n < 64 ? ps.n : 64;
    if (hi > ps.ninl)
        m |= (hi >= 64 ? ~0ULL : ((1ULL << hi) - 1ULL)) & ~((1ULL << ps.ninl) - 1ULL);
    return m;
}

// f(q) for the blocks a cell may receive, in block order: those of `mask`
// among the first 64, then (with a nonzero mask) every block from 64 on
template <typename F>
__device__ __forceinline__ void src_each(const PassSrc &ps, unsigned long long mask, F &&f)
{
    const unsigned long long lim = ps.n >= 64 ? ~0ULL : ((1ULL << ps.n) - 1ULL);
    for (unsigned long long m = mask & lim; m; m &= m - 1ULL) f(__builtin_ctzll(m));
    if (mask)
        for (int q = 64; q < ps.n; ++q) f(q);
}

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ unsigned xcd_remap(unsigned lin, unsigned nb)
{
    // Consecutive work-group ids are dispatched round-robin over the XCDs;
    // give XCD x the contiguous logical range of tiles [start_x, start_x+count_x).
    unsigned q = nb / NUM_XCD, r = nb % NUM_XCD;
    unsigned x = lin % NUM_XCD, slot = lin / NUM_XCD;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// psi <- c*d + b*psi ; returns the new psi
__device__ __forceinline__ double psi_step(double *psi, long long ip, double c, double b, double d)
{
    double v = c * d + b * psi[ip];
    psi[ip] = v;
    return v;
}

__device__ __forceinline__ bool in_box(const SrcBlk &s, int i, int j, int k, long long &m)
{
    const int a = i - s.i0, b = j - s.j0, c = k - s.k0;
    if ((unsigned)a < (unsigned)s.I && (unsigned)b < (unsigned)s.J && (unsigned)c < (unsigned)s.K) {
        m = ((long long)a * s.J + b) * s.K + c;
        return true;
    }
    return false;
}

// --------------------------------------------------------- field interpolation
// get_field_interp (emopt/fdtd.py:1401-1553) of one component on the box dom:
// each cell is the mean of the raw field over the component's stencil (offsets
// in order, own value first, then one division), where the raw field is that of
// the box ext -- dom grown by one cell as the reference grows it -- and zero
// outside it, except that ghosts one cell below ext's x / y start, when ext
// starts at 0 and the face has an 'E'/'H' symmetry, mirror the first cell with
// sign sx / sy (x applied before y, as the reference fills its padded copy).
struct IBox {
    int k1, k2, j1, j2, i1, i2;
};

struct IStencil {
    int n;                 // offsets used (1..8)
    int d[8][3];           // (di, dj, dk)
    int mx, my;            // mirror at the x / y min face
    double sx, sy;         // their signs
};

// the interpolated value of one cell (k, j, i) of the domain
__device__ __forceinline__ double2 interp_at(const double2 *__restrict__ F, int Nx, int Ny,
                                             const IBox &ext, const IStencil &st, int k, int j,
                                             int i)
{
    double re = 0.0, im = 0.0;
    for (int q = 0; q < st.n; ++q) {
        const int a = i + st.d[q][0];
        int b = j + st.d[q][1], c = k + st.d[q][2];
        const bool gy = st.my && b == ext.j1 - 1, gx = st.mx && c == ext.k1 - 1;
        if (gy) ++b;
        if (gx) ++c;
        double vr = 0.0, vi = 0.0;
        if (a >= ext.i1 && a < ext.i2 && b >= ext.j1 && b < ext.j2 && c >= ext.k1 && c < ext.k2) {
            const double2 v = F[((long long)a * Ny + b) * Nx + c];
            vr = v.x;
            vi = v.y;
        }
        if (gx) { vr = st.sx * vr; vi = st.sx * vi; }
        if (gy) { vr = st.sy * vr; vi = st.sy * vi; }
        if (q == 0) { re = vr; im = vi; }
        else { re += vr; im += vi; }
    }
    return make_double2(re / (double)st.n, im / (double)st.n);
}

__global__ void k_field_interp(const double2 *__restrict__ F, int Nx, int Ny, IBox dom, IBox ext,
                               IStencil st, double2 *__restrict__ out)
{
    const int nk = dom.k2 - dom.k1, nj = dom.j2 - dom.j1;
    const long long total = (long long)nk * nj * (dom.i2 - dom.i1);
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = dom.k1 + (int)(g % nk), j = dom.j1 + (int)((g / nk) % nj);
        const int i = dom.i1 + (int)(g / ((long long)nk * nj));
        out[g] = interp_at(F, Nx, Ny, ext, st, k, j, i);
    }
}

// ------------------------------------------------------ convergence check
//
// State of the check of emopt/fdtd.py:1072-1144 for up to NCONV_MAX sample
// points (the reference's nconv=200 of its 3D examples fits with room).
struct ConvState {
    double eh[3][3][NCONV_MAX];   // [sample slot 0..2][component][point]
    double phi1[NCONV_MAX], a1[NCONV_MAX];
    double t0, t1, t2;
    double a_change, phi_change;
    int done;
    int nchecks;
};

// The 3T phase and amplitude of src/fdtd.cpp:1465-1493, given the sample
// times' differences ds21 = sin t2 - sin t1, ds10 = sin t1 - sin t0,
// dc10 = cos t1 - cos t0, dc21 = cos t2 - cos t1 (the same for every point
// of a check, so computed once per check).
__device__ inline double phase3(double ds21, double ds10, double dc10, double dc21, double f0,
                                double f1, double f2)
{
    double f10 = f1 - f0, f21 = f2 - f1;
    if (f10 == 0 && f21 == 0) return 0.0;
    return atan2(f10 * ds21 - f21 * ds10, f21 * dc10 - f10 * dc21);
}

__device__ inline double amp3(double ds21, double ds10, double dc10, double dc21, double f0,
                              double f1, double f2, double ph)
{
    double f21 = f2 - f1, f10 = f1 - f0;
    if (f21 == 0 && f10 == 0) return 0.0;
    if (f21 * f21 >= f10 * f10) return f21 / (cos(ph) * ds21 + sin(ph) * dc21);
    return f10 / (cos(ph) * ds10 + sin(ph) * dc10);
}

// One check of emopt/fdtd.py:1072-1144 by one 64-lane wave (every lane of it
// must call): lane l handles sample points l, l + 64, ... (nc <= NCONV_MAX),
// the four norms are summed per lane and then across the wave by xor
// butterflies (each lane ends with the same sums, in a fixed order), and
// lane 0 writes the shared state.  Returns the
// convergence flag; *nchecks, *a_change, *phi_change receive the check's
// count and changes in every lane.
__device__ inline int conv_check(const double *__restrict__ Ex, const double *__restrict__ Ey,
                                 const double *__restrict__ Ez,
                                 const long long *__restrict__ sample, int nc, double tnow,
                                 double amp_rtol, double phi_rtol, long long nstep, ConvState *st,
                                 double *hist, int hist_cap, int *nchecks, double *a_change,
                                 double *phi_change)
{
    // all through the global address space (the pointers may come from memory)
    typedef __attribute__((address_space(1))) ConvState gconv;
    gconv *const S = (gconv *)st;
    const int lane = (int)(threadIdx.x & 63);
    const double t0 = S->t1, t1 = S->t2, t2 = tnow;
    const int h = S->nchecks;
    const double st1 = sin(t1), ct1 = cos(t1);
    const double ds21 = sin(t2) - st1, ds10 = st1 - sin(t0);
    const double dc10 = ct1 - cos(t0), dc21 = cos(t2) - ct1;
    double sa = 0, s0 = 0, sp = 0, sp0 = 0;
    // lane l handles sample points l, l + 64, l + 128, ...
    for (int q = lane; q < nc; q += 64) {
        const double *E[3] = { Ex, Ey, Ez };
        const long long at = *gptr(sample + q);
        double ph[3], am[3];
        for (int c = 0; c < 3; ++c) {
            const double f0 = S->eh[1][c][q], f1 = S->eh[2][c][q];
            const double f2 = at >= 0 ? *gptr(E[c] + at) : 0.0;
            S->eh[0][c][q] = f0;
            S->eh[1][c][q] = f1;
            S->eh[2][c][q] = f2;
            ph[c] = phase3(ds21, ds10, dc10, dc21, f0, f1, f2);
            am[c] = amp3(ds21, ds10, dc10, dc21, f0, f1, f2, ph[c]);
        }
        for (int c = 0; c < 3; ++c)
            if (am[c] < 0) { am[c] *= -1; ph[c] += M_PI; }
        const double phi0 = S->phi1[q], a0 = S->a1[q];
        const double phi1 = ph[0] + ph[1] + ph[2];
        const double a1 = am[0] + am[1] + am[2];
        S->phi1[q] = phi1; S->a1[q] = a1;
        sa += (a1 - a0) * (a1 - a0); s0 += a0 * a0;
        sp += (phi1 - phi0) * (phi1 - phi0); sp0 += phi0 * phi0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o, 64);
        s0 += __shfl_xor(s0, o, 64);
        sp += __shfl_xor(sp, o, 64);
        sp0 += __shfl_xor(sp0, o, 64);
    }
    const double ac = sqrt(sa) / sqrt(s0), pc = sqrt(sp) / sqrt(sp0);
    const int conv = (ac <= amp_rtol) && (pc <= phi_rtol);
    if (lane == 0) {
        S->t0 = t0; S->t1 = t1; S->t2 = t2;
        S->a_change = ac; S->phi_change = pc;
        S->nchecks = h + 1;
        if (hist && h < hist_cap) {
            *gptr(hist + 3 * h) = (double)nstep; *gptr(hist + 3 * h + 1) = ac;
            *gptr(hist + 3 * h + 2) = pc;
        }
        S->done = conv;
    }
    *nchecks = h + 1;
    *a_change = ac;
    *phi_change = pc;
    return conv;
}

// one wave (64 threads): the check of the single-solve loop
__global__ void k_conv_check(const double *__restrict__ Ex, const double *__restrict__ Ey,
                             const double *__restrict__ Ez, const long long *__restrict__ sample,
                             int nc, double tnow, double amp_rtol, double phi_rtol,
                             long long nstep, ConvState *st, double *hist, int hist_cap,
                             int *done_out)
{
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    if (st->done) return;
    int nck;
    double ac, pc;
    const int conv = conv_check(Ex, Ey, Ez, sample, nc, tnow, amp_rtol, phi_rtol, nstep, st, hist,
                                hist_cap, &nck, &ac, &pc);
    if (done_out && threadIdx.x == 0) *done_out = conv;
}

// -------------------------------------------------------- capture / phasor

// t0 snapshot -> .x, t1 snapshot -> .y of the result slot (src/fdtd.cpp:1205-1251)
__global__ void k_capture(const double *__restrict__ f0, const double *__restrict__ f1,
                          const double *__restrict__ f2, const double *__restrict__ f3,
                          const double *__restrict__ f4, const double *__restrict__ f5,
                          double2 *r0, double2 *r1, double2 *r2, double2 *r3, double2 *r4,
                          double2 *r5, long long n, int which)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) {
        double *d0 = which ? &r0[g].y : &r0[g].x;
        double *d1 = which ? &r1[g].y : &r1[g].x;
        double *d2 = which ? &r2[g].y : &r2[g].x;
        double *d3 = which ? &r3[g].y : &r3[g].x;
        double *d4 = which ? &r4[g].y : &r4[g].x;
        double *d5 = which ? &r5[g].y : &r5[g].x;
        *d0 = f0[g]; *d1 = f1[g]; *d2 = f2[g]; *d3 = f3[g]; *d4 = f4[g]; *d5 = f5[g];
    }
}

// Differences of sin/cos of the three sample times, precomputed on the host
// with the same expressions the reference evaluates per cell.
struct PhTimes {
    double s21, s10, c10, c21;   // 3T: sin(t2)-sin(t1), sin(t1)-sin(t0), cos(t1)-cos(t0), cos(t2)-cos(t1)
};

__device__ __forceinline__ double2 phasor3(double f0, double f1, double f2, const PhTimes &T)
{
    const double f10 = f1 - f0, f21 = f2 - f1;
    double ph, A;
    if (f10 == 0 && f21 == 0) {
        ph = 0.0; A = 0.0;
    } else {
        ph = atan2(f10 * T.s21 - f21 * T.s10, f21 * T.c10 - f10 * T.c21);
        if (f21 * f21 >= f10 * f10) A = f21 / (cos(ph) * T.s21 + sin(ph) * T.c21);
        else A = f10 / (cos(ph) * T.s10 + sin(ph) * T.c10);
    }
    if (A < 0) { A *= -1; ph += M_PI; }
    return make_double2(A * cos(ph), -A * sin(ph));
}

// calc_complex_fields 3T (src/fdtd.cpp:1346-1443): slot holds (t0, t1) per cell
__global__ void k_phasor3(const double *__restrict__ f0, const double *__restrict__ f1,
                          const double *__restrict__ f2, const double *__restrict__ f3,
                          const double *__restrict__ f4, const double *__restrict__ f5,
                          double2 *r0, double2 *r1, double2 *r2, double2 *r3, double2 *r4,
                          double2 *r5, long long n, PhTimes TE, PhTimes TH)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) {
        double2 v;
        v = r0[g]; r0[g] = phasor3(v.x, v.y, f0[g], TE);
        v = r1[g]; r1[g] = phasor3(v.x, v.y, f1[g], TE);
        v = r2[g]; r2[g] = phasor3(v.x, v.y, f2[g], TE);
        v = r3[g]; r3[g] = phasor3(v.x, v.y, f3[g], TH);
        v = r4[g]; r4[g] = phasor3(v.x, v.y, f4[g], TH);
        v = r5[g]; r5[g] = phasor3(v.x, v.y, f5[g], TH);
    }
}

// 2T variant (src/fdtd.cpp:1253-1343,1445-1463)
struct PhTimes2 {
    double st0, st1, ct0, ct1;
};

__device__ __forceinline__ double2 phasor2(double f0, double f1, const PhTimes2 &T)
{
    double ph, A;
    if (f0 == 0.0 && f1 == 0) ph = 0.0;
    else ph = atan((f1 * T.st0 - f0 * T.st1) / (f0 * T.ct1 - f1 * T.ct0));
    if (f0 * f0 > f1 * f1) A = f1 / (T.st1 * cos(ph) + T.ct1 * sin(ph));
    else A = f0 / (T.st0 * cos(ph) + T.ct0 * sin(ph));
    if (A < 0) { A *= -1; ph += M_PI; }
    return make_double2(A * cos(ph), -A * sin(ph));
}

__global__ void k_phasor2(const double *__restrict__ f0, const double *__restrict__ f1,
                          const double *__restrict__ f2, const double *__restrict__ f3,
                          const double *__restrict__ f4, const double *__restrict__ f5,
                          double2 *r0, double2 *r1, double2 *r2, double2 *r3, double2 *r4,
                          double2 *r5, long long n, PhTimes2 TE, PhTimes2 TH)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) {
        r0[g] = phasor2(r0[g].x, f0[g], TE);
        r1[g] = phasor2(r1[g].x, f1[g], TE);
        r2[g] = phasor2(r2[g].x, f2[g], TE);
        r3[g] = phasor2(r3[g].x, f3[g], TH);
        r4[g] = phasor2(r4[g].x, f4[g], TH);
        r5[g] = phasor2(r5[g].x, f5[g], TH);
    }
}

// ------------------------------------------------------------- utilities

__global__ void k_fill(double *a, long long n, double v)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) a[g] = v;
}

__global__ void k_fill2(double2 *a, long long n, double2 v)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) a[g] = v;
}

// coefficient arrays from raw complex eps (real: b = dt/eps.re;
// complex: a = (er - ei)/(er + ei), b = 1/(er + ei), er = eps.re/dt, ei = eps.im/2)
__global__ void k_eps_coeffs(const double2 *__restrict__ eps, double *__restrict__ b,
                             double *__restrict__ a, long long n, double dt, double odt, int ceps)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) {
        const double2 e = eps[g];
        if (!ceps) {
            b[g] = dt / e.x;
        } else {
            const double er = e.x * odt, ei = e.y * 0.5;
            a[g] = (er - ei) / (er + ei);
            b[g] = 1.0 / (er + ei);
        }
    }
}

// b = dt/eps for a table of nt distinct real eps triples (the same division
// as k_eps_coeffs, so the same doubles).  eps holds 3 values per triple, b 4
// (b_x, b_y, b_z, 0): a 32-byte entry is gathered with one 16-byte and one
// 8-byte load.
__global__ void k_b_table(const double *__restrict__ eps, double *__restrict__ b, long long nt,
                          double dt)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < 4 * nt; g += (long long)gridDim.x * blockDim.x) {
        const long long t = g >> 2;
        const int c = (int)(g & 3);
        b[g] = c < 3 ? dt / eps[3 * t + c] : 0.0;
    }
}

// dt/mu.re and mu.re from raw complex mu
__global__ void k_mu_coeffs(const double2 *__restrict__ mu, double *__restrict__ m,
                            double *__restrict__ mr, long long n, double dt)
{
    long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; g < n; g += (long long)gridDim.x * blockDim.x) {
        m[g] = dt / mu[g].x;
        mr[g] = mu[g].x;
    }
}

// gather a box of a complex [Nz][Ny][Nx] array into a dense box array
__global__ void k_gather_box(const double2 *__restrict__ src, double2 *__restrict__ dst, Geo g,
                             int k1, int nk, int j1, int nj, int i1, int ni)
{
    long long m = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)nk * nj * ni;
    for (; m < n; m += (long long)gridDim.x * blockDim.x) {
        int k = (int)(m % nk), j = (int)((m / nk) % nj), i = (int)(m / ((long long)nk * nj));
        dst[m] = src[(long long)(i + i1) * g.nxy + (long long)(j + j1) * g.Nx + (k + k1)];
    }
}

__global__ void k_scatter_box(double2 *__restrict__ dst, const double2 *__restrict__ src, Geo g,
                              int k1, int nk, int j1, int nj, int i1, int ni)
{
    long long m = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)nk * nj * ni;
    for (; m < n; m += (long long)gridDim.x * blockDim.x) {
        int k = (int)(m % nk), j = (int)((m / nk) % nj), i = (int)(m / ((long long)nk * nj));
        dst[(long long)(i + i1) * g.nxy + (long long)(j + j1) * g.Nx + (k + k1)] = src[m];
    }
}

// plain vector copy, one 16-byte element per thread (bandwidth reference;
// on MI355X this form streams faster than a grid-stride loop)
__global__ void k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long long n)
{
    const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (g < n) b[g] = a[g];
}

// ------------------------------------------------------ gradient assembly

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// adj * 1j * d * fwd evaluated left to right like numpy (emopt/fdtd.py:1710-1712)
__device__ __forceinline__ double2 ydx_term(double2 adj, double2 d, double2 fwd)
{
    return cmul(cmul(cmul(adj, make_double2(0.0, 1.0)), d), fwd);
}

constexpr int RED_THREADS = 256;

// Sum over a work-group of RED_THREADS threads: each 64-lane wave reduces its
// values by cross-lane butterflies, then one thread adds the waves' sums in
// wave order (a fixed order, so results are reproducible).
__device__ __forceinline__ double2 block_sum2(double2 v)
{
    constexpr int W = RED_THREADS / 64;
    __shared__ double2 sh[W];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o, 64);
        v.y += __shfl_xor(v.y, o, 64);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double2 t = sh[0];
    for (int w = 1; w < W; ++w) { t.x += sh[w].x; t.y += sh[w].y; }
    return t;
}

// adj * (-1j) * d * fwd, the H terms of emopt/fdtd.py:1713-1715 (numpy's
// -1j is (-0.0, -1.0))
__device__ __forceinline__ double2 ydx_term_h(double2 adj, double2 d, double2 fwd)
{
    return cmul(cmul(cmul(adj, make_double2(-0.0, -1.0)), d), fwd);
}

// The fields and materials y^T dA x reads: result slots of the adjoint (a) and
// forward (f) solves, Ex..Hz; eps per component; mu per component, or null
// with the uniform value mu_re.
struct YdaxArgs {
    const double2 *a[6];
    const double2 *f[6];
    const double2 *eps[3];
    const double2 *mu[3];
    double mu_re;
};

// partial sums of sum_c adj_c * 1j * (eps_c - eps0_c) * fwd_c
//               + sum_c adj_Hc * -1j * (mu_c - mu0_c) * fwd_Hc   (z_mu non-null)
// over a box, the reference's left-to-right order per cell; one partial per
// work-group (fixed grid => fixed summation order)
__global__ __launch_bounds__(RED_THREADS) void k_ydAx(
    YdaxArgs y, const double2 *__restrict__ z0, const double2 *__restrict__ z1,
    const double2 *__restrict__ z2, const double2 *__restrict__ zm0,
    const double2 *__restrict__ zm1, const double2 *__restrict__ zm2, Geo g, int k1, int nk,
    int j1, int nj, int i1, int ni, double2 *partial)
{
    long long n = (long long)nk * nj * ni;
    const double2 *z[3] = { z0, z1, z2 }, *zm[3] = { zm0, zm1, zm2 };
    double2 acc = make_double2(0.0, 0.0);
    for (long long m = blockIdx.x * (long long)blockDim.x + threadIdx.x; m < n;
         m += (long long)gridDim.x * blockDim.x) {
        int k = (int)(m % nk), j = (int)((m / nk) % nj), i = (int)(m / ((long long)nk * nj));
        long long c = (long long)(i + i1) * g.nxy + (long long)(j + j1) * g.Nx + (k + k1);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double2 e = y.eps[q][c], e0 = z[q][m];
            const double2 d = make_double2(e.x - e0.x, e.y - e0.y);
            const double2 t = ydx_term(y.a[q][c], d, y.f[q][c]);
            acc.x += t.x; acc.y += t.y;
        }
        if (zm0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double2 mu = y.mu[q] ? y.mu[q][c] : make_double2(y.mu_re, 0.0);
                const double2 m0 = zm[q][m];
                const double2 d = make_double2(mu.x - m0.x, mu.y - m0.y);
                const double2 t = ydx_term_h(y.a[3 + q][c], d, y.f[3 + q][c]);
                acc.x += t.x; acc.y += t.y;
            }
        }
    }
    double2 tot = block_sum2(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// grad[p] = -2 Re(sum over the entries of parameter p of adj*1j*deps*fwd)/step
// for E components (comp 0..2) and of adj*(-1j)*dmu*fwd for H components
// (comp 3..5, a change of mu_x..mu_z); entries are grouped by parameter,
// seg[p]..seg[p+1] (one work-group each)
__global__ __launch_bounds__(RED_THREADS) void k_grad(
    YdaxArgs y, const long long *__restrict__ seg, const int *__restrict__ comp,
    const long long *__restrict__ cell, const double2 *__restrict__ deps, double step,
    double *grad)
{
    const int p = blockIdx.x;
    double2 acc = make_double2(0.0, 0.0);
    for (long long e = seg[p] + threadIdx.x; e < seg[p + 1]; e += blockDim.x) {
        const int c = comp[e];
        const long long g = cell[e];
        const double2 t = c < 3 ? ydx_term(y.a[c][g], deps[e], y.f[c][g])
                                : ydx_term_h(y.a[c][g], deps[e], y.f[c][g]);
        acc.x += t.x; acc.y += t.y;
    }
    double2 tot = block_sum2(acc);
    if (threadIdx.x == 0) grad[p] = -2 * (tot.x / step);
}

// ------------------------------------------- reductions over interpolated fields

// Four interpolated components (a, b, c, d) of one result slot on a domain.
struct FluxArgs {
    const double2 *F[4];
    IStencil st[4];
};

// partial sums of Re(a conj(b) - c conj(d)) over the domain (one per
// work-group): the Poynting flux integrand of get_source_power
// (emopt/fdtd.py:1611-1665) with each field interpolated onto the Ez grid
__global__ __launch_bounds__(RED_THREADS) void k_interp_flux(FluxArgs fa, int Nx, int Ny,
                                                             IBox dom, IBox ext,
                                                             double2 *__restrict__ partial)
{
    const int nk = dom.k2 - dom.k1, nj = dom.j2 - dom.j1;
    const long long total = (long long)nk * nj * (dom.i2 - dom.i1);
    double acc = 0.0;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = dom.k1 + (int)(g % nk), j = dom.j1 + (int)((g / nk) % nj);
        const int i = dom.i1 + (int)(g / ((long long)nk * nj));
        const double2 a = interp_at(fa.F[0], Nx, Ny, ext, fa.st[0], k, j, i);
        const double2 b = interp_at(fa.F[1], Nx, Ny, ext, fa.st[1], k, j, i);
        const double2 c = interp_at(fa.F[2], Nx, Ny, ext, fa.st[2], k, j, i);
        const double2 d = interp_at(fa.F[3], Nx, Ny, ext, fa.st[3], k, j, i);
        // Re(a conj(b)) - Re(c conj(d))
        acc += (a.x * b.x + a.y * b.y) - (c.x * d.x + c.y * d.y);
    }
    const double2 tot = block_sum2(make_double2(acc, 0.0));
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// The six interpolated components of one result slot, and six weight planes
// of the domain's shape (a null plane weighs 0).
struct DotArgs {
    const double2 *F[6];
    IStencil st[6];
    const double2 *W[6];
};

// partial sums over the domain of sum_q W_q * field_q, separately for the E
// components (q = 0..2, partial[2 b]) and the H components (q = 3..5,
// partial[2 b + 1]): the overlap integrals of ModeMatch
// (emopt/fomutils.py:1047-1076) with the fields interpolated onto the Ez grid
__global__ __launch_bounds__(RED_THREADS) void k_interp_dot(DotArgs da, int Nx, int Ny, IBox dom,
                                                            IBox ext,
                                                            double2 *__restrict__ partial)
{
    const int nk = dom.k2 - dom.k1, nj = dom.j2 - dom.j1;
    const long long total = (long long)nk * nj * (dom.i2 - dom.i1);
    double2 ae = make_double2(0.0, 0.0), ah = make_double2(0.0, 0.0);
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = dom.k1 + (int)(g % nk), j = dom.j1 + (int)((g / nk) % nj);
        const int i = dom.i1 + (int)(g / ((long long)nk * nj));
        for (int q = 0; q < 6; ++q) {
            if (!da.W[q]) continue;
            const double2 w = da.W[q][g];
            const double2 f = interp_at(da.F[q], Nx, Ny, ext, da.st[q], k, j, i);
            const double2 t = cmul(w, f);
            if (q < 3) { ae.x += t.x; ae.y += t.y; }
            else { ah.x += t.x; ah.y += t.y; }
        }
    }
    const double2 te = block_sum2(ae);
    __syncthreads();                       // block_sum2's LDS is reused
    const double2 th = block_sum2(ah);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = te;
        partial[2 * blockIdx.x + 1] = th;
    }
}

// ------------------------------------------- forward x adjoint field products
//
// out[c][g] = fwd_c * adj_c over a box of cells for three components (the E
// or the H ones), the per-cell factors of the topology gradient and of the
// torch pseudo-loss (emopt/experimental/fdtd.py:37-223).  The product is
// rounded as numpy rounds an elementwise complex product on an FMA-capable
// host: fma(ar, br, -(ai bi)) + i fma(ar, bi, ai br).
struct ProdArgs {
    const double2 *f[3], *a[3];
    double2 *out;
};

__global__ void k_field_products(ProdArgs pa, int Nx, int Ny, IBox dom)
{
    const int nk = dom.k2 - dom.k1, nj = dom.j2 - dom.j1;
    const long long total = (long long)nk * nj * (dom.i2 - dom.i1);
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = dom.k1 + (int)(g % nk), j = dom.j1 + (int)((g / nk) % nj);
        const int i = dom.i1 + (int)(g / ((long long)nk * nj));
        const long long c0 = ((long long)i * Ny + j) * Nx + k;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 x = pa.f[c][c0], y = pa.a[c][c0];
            pa.out[c * total + g] = make_double2(fma(x.x, y.x, -(x.y * y.y)), fma(x.x, y.y, x.y * y.x));
        }
    }
}

// ------------------------------------------- adjoint sources built on the device
//
// The reference forms adjoint sources on the host (emopt/fomutils.py:1282-1429,
// 1654-1920): derivative planes on the Ez grid, mapped back onto the Yee
// grid by the transpose of the interpolation stencils.  Here the planes are
// formed and transposed on the device, and the result is appended to the
// solver's staged sources without crossing to the host.

// P = (k x W) / div for a plane W of n cells (k complex, div real), rounded
// as numpy rounds the host formula k * W / div on an FMA-capable host: the
// product as fma(kr, wr, -(ki wi)) + i fma(kr, wi, ki wr), the quotient as a
// product with the reciprocal rdiv = 1 / div
__global__ void k_plane_scale(double2 *__restrict__ out, const double2 *__restrict__ W, double kr,
                              double ki, double rdiv, long long n)
{
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < n;
         g += (long long)gridDim.x * blockDim.x) {
        const double2 w = W[g];
        const double tr = fma(kr, w.x, -(ki * w.y)), ti = fma(kr, w.y, ki * w.x);
        out[g] = make_double2(tr * rdiv, ti * rdiv);
    }
}

// The derivative planes of one face of the source-power box
// (emopt/fomutils.py:1654-1920): with S = 0.5 Re(E1 H2* - E2 H1*) on the
// face, dS/dE1 = s1 H2*, dS/dE2 = -s1 H1*, dS/dH1 = -s1 E2*, dS/dH2 = s1 E1*
// (s1 = +-area/4 for the face's outward sign), each times mf = -f and
// divided by Psrc^2 (as a product with rp2 = 1 / Psrc^2, the way numpy
// divides a complex array by a real scalar).  F: the slot's E1, E2, H1, H2;
// out: their planes.
struct FaceArgs {
    const double2 *F[4];
    IStencil st[4];
    double2 *out[4];
    double s1, mf, rp2;
};

__global__ void k_face_dP(FaceArgs fa, int Nx, int Ny, IBox dom, IBox ext)
{
    const int nk = dom.k2 - dom.k1, nj = dom.j2 - dom.j1;
    const long long total = (long long)nk * nj * (dom.i2 - dom.i1);
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = dom.k1 + (int)(g % nk), j = dom.j1 + (int)((g / nk) % nj);
        const int i = dom.i1 + (int)(g / ((long long)nk * nj));
        double2 f[4];
        for (int q = 0; q < 4; ++q) f[q] = interp_at(fa.F[q], Nx, Ny, ext, fa.st[q], k, j, i);
        // (coefficient, conjugated field) of E1, E2, H1, H2
        const double c[4] = { fa.s1, -fa.s1, -fa.s1, fa.s1 };
        const int src[4] = { 3, 2, 1, 0 };
        for (int q = 0; q < 4; ++q) {
            const double2 v = f[src[q]];
            const double dr = c[q] * v.x, di = c[q] * (-v.y);
            fa.out[q][g] = make_double2((fa.mf * dr) * fa.rp2, (fa.mf * di) * fa.rp2);
        }
    }
}

// dst[box] += w * src[box] (or dst[box] = src[box] with assign) over an
// nz x ny x nx box of two (z, y, x) row-major arrays: one term of the
// transposed interpolation.  w is +-1/2^m, so w * src is exact and the sum
// rounds once, as src * w then += does.
struct BoxMap {
    int dNy, dNx, dz, dy, dx;     // destination row lengths and box origin
    int sNy, sNx, sz, sy, sx;     // source row lengths and box origin
    int nz, ny, nx;
};

__global__ void k_box_axpy(double2 *__restrict__ dst, const double2 *__restrict__ src, BoxMap b,
                           double w, int assign)
{
    const long long total = (long long)b.nz * b.ny * b.nx;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(g % b.nx), r = (int)((g / b.nx) % b.ny);
        const int a = (int)(g / ((long long)b.nx * b.ny));
        const long long d = ((long long)(b.dz + a) * b.dNy + (b.dy + r)) * b.dNx + (b.dx + c);
        const long long q = ((long long)(b.sz + a) * b.sNy + (b.sy + r)) * b.sNx + (b.sx + c);
        const double2 v = src[q];
        if (assign) {
            dst[d] = v;
        } else {
            const double2 o = dst[d];
            dst[d] = make_double2(o.x + v.x * w, o.y + v.y * w);
        }
    }
}

// One source block from a padded transposed array: block cell (a, b, c) is
// pad[oz + a][oy + b][ox + c], stored as the update reads sources: a complex
// value v becomes (Re v, -Im v) = (|v| cos phi, |v| sin phi) with
// phi = -arg v (src/fdtd.cpp:1517-1576).
__global__ void k_stage_block(double2 *__restrict__ blk, const double2 *__restrict__ pad, int Py,
                              int Px, int oz, int oy, int ox, int I, int J, int K)
{
    const long long total = (long long)I * J * K;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(g % K), b = (int)((g / K) % J), a = (int)(g / ((long long)J * K));
        const double2 v = pad[((long long)(oz + a) * Py + (oy + b)) * Px + (ox + c)];
        blk[g] = make_double2(v.x, -v.y);
    }
}

}  // namespace emf
