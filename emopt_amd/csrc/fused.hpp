// Fused Yee step for gfx950: one launch advances H and then E by a full time
// step (update_H + update_E of src/fdtd.cpp:121-755 with the same arithmetic),
// reading the old state and writing a second ("new") set of buffers.
//
// HBM traffic per cell: read E(3) H(3) b_eps(3), write E(3) H(3) = 120 B
// (the two separate passes need 168 B: E and H are each read twice), plus
// the CPML psi slabs.
//
// E_new(i,j,k) needs H_new at (k-1), (j-1) and (i-1).  Each wavefront owns 63
// consecutive x cells of one row j and sweeps a chunk of z planes:
//   * (i-1): its own H_new of the previous plane, kept in registers;
//   * (k-1): the neighbouring lane's H_new (lane shuffle); lane 0 is a halo
//     column at k0-1 that computes H_new but owns nothing;
//   * (j-1): the wave recomputes H_new of row j-1 itself (the row's owner
//     computes the same value from the same inputs).
// No LDS, no barriers: every wave is independent.  Because halo cells read
// E_old, H_old and psi_old that another wave is replacing, old and new state
// live in separate buffers (ping-pong).  The E-side psi is read and written by
// its owner only and stays in place.  A chunk that starts inside the grid
// first recomputes H_new of plane z0-1 (prologue, nothing written).
#pragma once
#include "kernels.hpp"

namespace emf {

constexpr int FX = 64;     // lanes per row (one wavefront)
constexpr int FXO = 63;    // owned cells per row

// Arguments of the fused step.  Each 3-component quantity is one allocation
// with components `cs` bytes apart; the psi slabs of a set share one
// allocation and the CPML tables another, at offsets derived from the grid
// dimensions and CPML widths (Layout below), so the kernel needs only a
// dozen uniform pointers.
struct FusedArgs {
    const double *Eo, *Ho;   // state at the start of the step
    double *En, *Hn;         // state after the step
    const double *bE, *aE;   // E coefficients (aE only for complex eps)
    const double *mH, *muR;  // dt/mu.re and mu.re per cell (non-uniform mu only)
    const double *pho;       // H-pass psi, old set
    double *phn;             // H-pass psi, new set
    double *pe;              // E-pass psi (in place)
    const double *tab;       // CPML tables
    unsigned cs;             // byte stride between components (32-bit offset kernels)
    unsigned long long csl;  // the same stride for the 64-bit offset kernels
    double m_s, mu_s;        // uniform dt/mu.re and mu.re
    const void *bI;          // real eps only, optional: per-cell index into bT
    const double *bT;        // (b_x, b_y, b_z, 0) per distinct eps triple
    int bIw;                 // index width in bytes (2 or 4)
};

// Coefficient modes of the E update: real eps with full b arrays, complex eps
// (a and b arrays), real eps with b gathered through a per-cell index.
enum { CM_REAL = 0, CM_COMPLEX = 1, CM_INDEX = 2 };

// Byte offsets of cells within a field component are 32-bit (OT =
// unsigned) while 3 components of the grid fit in 4 GiB (178 M cells), and
// 64-bit (OT = unsigned long long) above; the two forms of a kernel differ
// only in that arithmetic.  cstride gives the component stride in OT.
template <typename OT>
__device__ __forceinline__ OT cstride(const FusedArgs &A)
{
    if constexpr (sizeof(OT) == 8) return (OT)A.csl;
    else return (OT)A.cs;
}

// Entry of the per-cell coefficient index at byte offset o of a field
// component: one aligned dword holds the 2- or 4-byte entry (the 2-byte index
// is allocated with two spare entries for the last dword).
template <typename OT = unsigned>
__device__ __forceinline__ unsigned coef_index(const FusedArgs &A, OT o)
{
    const OT e = o >> 3;
    const unsigned w2 = A.bIw == 2;
    const unsigned v = *gptr((const unsigned *)((const char *)A.bI +
                                                (w2 ? (e << 1) & ~(OT)3 : e << 2)));
    return w2 ? (v >> (((unsigned)e & 1u) << 4)) & 0xffffu : v;
}

// (b_x, b_y, b_z) of table entry ci: 32-byte entries, one 16-byte and one
// 8-byte load
__device__ __forceinline__ void coef_gather(const FusedArgs &A, unsigned ci, double &bx,
                                            double &by, double &bz)
{
    const double2 xy = ld2(reinterpret_cast<const double2 *>(A.bT + 4 * ci));
    bx = xy.x; by = xy.y;
    bz = *gptr(A.bT + 4 * ci + 2);
}

// Byte offsets of the psi slab arrays of a set ([face][component], faces
// xmin xmax ymin ymax zmin zmax, each face's two arrays consecutive) and of
// the CPML tables (per axis: kH bH cH kE bE cE, w_min + w_max entries each).
struct Layout {
    unsigned pso[6][2];
    unsigned tbo[3];     // start of each axis' tables
    unsigned tbl[3];     // bytes per table of that axis
};

__device__ __forceinline__ Layout make_layout(const Geo &g)
{
    Layout L;
    unsigned off = 0;
    for (int f = 0; f < 6; ++f) {
        const unsigned plane = f < 2 ? (unsigned)g.Nz * (unsigned)g.Ny
                                     : (f < 4 ? (unsigned)g.Nz * (unsigned)g.Nx : (unsigned)g.Ny * (unsigned)g.Nx);
        const unsigned bytes = plane * (unsigned)g.w[f] * 8u;
        L.pso[f][0] = off;
        L.pso[f][1] = off + bytes;
        off += 2 * bytes;
    }
    unsigned t = 0;
    for (int a = 0; a < 3; ++a) {
        const int len = max(1, g.w[2 * a] + g.w[2 * a + 1]);
        L.tbo[a] = t;
        L.tbl[a] = (unsigned)len * 8u;
        t += 6u * (unsigned)len * 8u;
    }
    return L;
}

// Byte offset of psi array c (0, 1) of the min (lo) or max face of an axis.
// The select between two constant-indexed entries keeps the Layout in
// registers (a runtime face index would send the array to memory).
__device__ __forceinline__ unsigned psof(const Layout &L, int axis, bool lo, int c)
{
    return lo ? L.pso[2 * axis][c] : L.pso[2 * axis + 1][c];
}

// Accesses by 32-bit byte offset from a uniform base pointer: the backend then
// addresses with an SGPR base plus one VGPR offset shared by all arrays,
// instead of a 64-bit VGPR address per array.  The fused path is used only
// when every array is smaller than 4 GiB.  Every array lives in device global
// memory, and the accesses say so: where a base pointer comes from memory (a
// batch member's descriptor) rather than from the kernel arguments, the
// backend could not infer it and would emit flat accesses, which count on the
// scalar-memory counter too and make every wait for a scalar load also wait
// for the field loads in flight.
typedef __attribute__((address_space(1))) double gdouble;

__device__ __forceinline__ double ldo(const double *base, unsigned off)
{
    return *(const gdouble *)(reinterpret_cast<const char *>(base) + off);
}

__device__ __forceinline__ double ldo(const double *base, unsigned long long off)
{
    return *(const gdouble *)(reinterpret_cast<const char *>(base) + off);
}

// Load from data that no kernel writes while a step runs (the CPML tables),
// through the constant address space: where the offset is the same for the
// whole wave the backend then loads it with a scalar load, which keeps it off
// the vector memory path.
__device__ __forceinline__ double ldc(const double *base, unsigned off)
{
    typedef const __attribute__((address_space(4))) double cdouble;
    return *(cdouble *)(reinterpret_cast<const char *>(base) + off);
}

// Streaming access for data that is touched once per step (the CPML psi
// arrays): non-temporal loads and stores keep these lines from displacing the
// fields that the next pass re-reads from L2 and the Infinity Cache.  Worth
// it only when a step's data is larger than that cache; a smaller solve keeps
// its psi in the cache from step to step.
__device__ __forceinline__ double ldn(const double *base, unsigned off)
{
    return __builtin_nontemporal_load((const gdouble *)(reinterpret_cast<const char *>(base) + off));
}

__device__ __forceinline__ void stn(double *base, unsigned off, double v)
{
    __builtin_nontemporal_store(v, (gdouble *)(reinterpret_cast<char *>(base) + off));
}

__device__ __forceinline__ void sto(double *base, unsigned off, double v)
{
    *(gdouble *)(reinterpret_cast<char *>(base) + off) = v;
}

__device__ __forceinline__ void sto(double *base, unsigned long long off, double v)
{
    *(gdouble *)(reinterpret_cast<char *>(base) + off) = v;
}

// psi access, streaming (NT) or plain
template <bool NT>
__device__ __forceinline__ double ldp(const double *base, unsigned off)
{
    return NT ? ldn(base, off) : ldo(base, off);
}

template <bool NT>
__device__ __forceinline__ void stp(double *base, unsigned off, double v)
{
    if (NT) stn(base, off, v);
    else sto(base, off, v);
}

// CPML slab of one cell along one axis: which face, psi byte offset and
// table index (the index formulas of src/fdtd.cpp:212-306 / 573-666).
struct Slab {
    bool in, lo;
    unsigned io;   // psi byte offset
    int par;       // table index
};

__host__ __device__ __forceinline__ Slab slab_x(const Geo &g, int i, int j, int k)
{
    Slab s;
    const int xlo = g.w[0], xhi = g.Nx - g.w[1];
    s.in = k < xlo || k >= xhi;
    s.lo = k < xlo;
    const int wl = s.lo ? g.w[0] : g.w[1];
    const int kk = s.lo ? k : k - xhi;
    s.par = s.lo ? xlo - k - 1 : k - xhi + g.w[0];
    s.io = (((unsigned)i * (unsigned)g.Ny + (unsigned)j) * (unsigned)wl + (unsigned)kk) * 8u;
    return s;
}

__host__ __device__ __forceinline__ Slab slab_y(const Geo &g, int i, int j, int k)
{
    Slab s;
    const int ylo = g.w[2], yhi = g.Ny - g.w[3];
    s.in = j < ylo || j >= yhi;
    s.lo = j < ylo;
    const int wl = s.lo ? g.w[2] : g.w[3];
    const int jj = s.lo ? j : j - yhi;
    s.par = s.lo ? ylo - j - 1 : j - yhi + g.w[2];
    s.io = (((unsigned)i * (unsigned)wl + (unsigned)jj) * (unsigned)g.Nx + (unsigned)k) * 8u;
    return s;
}

__host__ __device__ __forceinline__ Slab slab_z(const Geo &g, int i, int j, int k)
{
    Slab s;
    const int zlo = g.w[4], zhi = g.Nz - g.w[5];
    s.in = i < zlo || i > zhi;   // zmax is strict: its first layer is skipped
    s.lo = i < zlo;
    const int ii = s.lo ? i : i - zhi;
    s.par = s.lo ? zlo - i - 1 : i - zhi + g.w[4];
    s.io = (((unsigned)ii * (unsigned)g.Ny + (unsigned)j) * (unsigned)g.Nx + (unsigned)k) * 8u;
    return s;
}

// everything H_new at one cell depends on, loaded before any store
struct HIn {
    double ex, ey, ez;         // E_old(i, j, k)
    double exi, eyi;           // E_old(i+1, j, k)
    double ezj, exj;           // E_old(i, j+1, k)
    double ezk, eyk;           // E_old(i, j, k+1)
    double hx, hy, hz;         // H_old(i, j, k)
    double mx, my, mz;         // dt/mu.re
    double px0, px1, py0, py1, pz0, pz1;     // psi_old of the slabs the cell is in
    double kx, bx, cx, ky, by, cy, kz, bz, cz;   // CPML table entries
    Slab sx, sy, sz;
};

template <bool MU_ARRAY, bool NT = false, typename OT = unsigned>
__device__ __forceinline__ void h_load_aux(const Geo &g, const Layout &L, const FusedArgs &A,
                                           int i, int j, int k, OT o, HIn &d)
{
    const OT cs = cstride<OT>(A);
    d.mx = MU_ARRAY ? ldo(A.mH, o) : A.m_s;
    d.my = MU_ARRAY ? ldo(A.mH, o + cs) : A.m_s;
    d.mz = MU_ARRAY ? ldo(A.mH, o + 2 * cs) : A.m_s;
    d.sx = slab_x(g, i, j, k);
    d.sy = slab_y(g, i, j, k);
    d.sz = slab_z(g, i, j, k);
    d.px0 = d.px1 = d.py0 = d.py1 = d.pz0 = d.pz1 = 0;
    d.kx = d.bx = d.cx = d.ky = d.by = d.cy = d.kz = d.bz = d.cz = 0;
    if (d.sx.in) {
        d.px0 = ldp<NT>(A.pho, psof(L, 0, d.sx.lo, 0) + d.sx.io);
        d.px1 = ldp<NT>(A.pho, psof(L, 0, d.sx.lo, 1) + d.sx.io);
        const unsigned t = L.tbo[0] + (unsigned)d.sx.par * 8u;
        d.kx = ldc(A.tab, t); d.bx = ldc(A.tab, t + L.tbl[0]); d.cx = ldc(A.tab, t + 2 * L.tbl[0]);
    }
    if (d.sy.in) {
        d.py0 = ldp<NT>(A.pho, psof(L, 1, d.sy.lo, 0) + d.sy.io);
        d.py1 = ldp<NT>(A.pho, psof(L, 1, d.sy.lo, 1) + d.sy.io);
        const unsigned t = L.tbo[1] + (unsigned)d.sy.par * 8u;
        d.ky = ldc(A.tab, t); d.by = ldc(A.tab, t + L.tbl[1]); d.cy = ldc(A.tab, t + 2 * L.tbl[1]);
    }
    if (d.sz.in) {
        d.pz0 = ldp<NT>(A.pho, psof(L, 2, d.sz.lo, 0) + d.sz.io);
        d.pz1 = ldp<NT>(A.pho, psof(L, 2, d.sz.lo, 1) + d.sz.io);
        const unsigned t = L.tbo[2] + (unsigned)d.sz.par * 8u;
        d.kz = ldc(A.tab, t); d.bz = ldc(A.tab, t + L.tbl[2]); d.cz = ldc(A.tab, t + 2 * L.tbl[2]);
    }
}

// H_new from loaded inputs; the new psi values are returned for the owner to store
__device__ __forceinline__ void h_compute(const Geo &g, const HIn &d, double &hx, double &hy,
                                          double &hz, double np[6])
{
    const double dEzdy = g.ry * (d.ezj - d.ez);
    const double dEydz = g.rz * (d.eyi - d.ey);
    hx = d.hx + d.mx * (dEydz - dEzdy);
    const double dExdz = g.rz * (d.exi - d.ex);
    const double dEzdx = g.rx * (d.ezk - d.ez);
    hy = d.hy + d.my * (dEzdx - dExdz);
    const double dEydx = g.rx * (d.eyk - d.ey);
    const double dExdy = g.ry * (d.exj - d.ex);
    hz = d.hz + d.mz * (dExdy - dEydx);
    if (d.sx.in) {
        np[0] = d.cx * dEydx + d.bx * d.px0;
        np[1] = d.cx * dEzdx + d.bx * d.px1;
        hz = hz - d.mz * (np[0] - dEydx + dEydx / d.kx);
        hy = hy + d.my * (np[1] - dEzdx + dEzdx / d.kx);
    }
    if (d.sy.in) {
        np[2] = d.cy * dExdy + d.by * d.py0;
        np[3] = d.cy * dEzdy + d.by * d.py1;
        hz = hz + d.mz * (np[2] - dExdy + dExdy / d.ky);
        hx = hx - d.mx * (np[3] - dEzdy + dEzdy / d.ky);
    }
    if (d.sz.in) {
        np[4] = d.cz * dExdz + d.bz * d.pz0;
        np[5] = d.cz * dEydz + d.bz * d.pz1;
        hx = hx + d.mx * (np[5] - dEydz + dEydz / d.kz);
        hy = hy - d.my * (np[4] - dExdz + dExdz / d.kz);
    }
}

// magnetic sources of src/fdtd.cpp:313-366 at one cell: the blocks of mask
// (wave_src_mask; all blocks by default)
template <bool MU_ARRAY, typename OT = unsigned>
__device__ __forceinline__ void h_sources(const Geo &g, const PassSrc &sh, const FusedArgs &A,
                                          int i, int j, int k, OT o, double &hx,
                                          double &hy, double &hz, unsigned long long mask = ~0ULL)
{
    const OT cs = cstride<OT>(A);
    src_each(sh, mask, [&](int q) {
        long long m;
        const SrcBlk sb = src_blk(sh.blk, q);
        if (in_box(sb, i, j, k, m)) {
            const double2 vx = ld2(sb.f[3] + m), vy = ld2(sb.f[4] + m), vz = ld2(sb.f[5] + m);
            const double ux = MU_ARRAY ? ldo(A.muR, o) : A.mu_s;
            const double uy = MU_ARRAY ? ldo(A.muR, o + cs) : A.mu_s;
            const double uz = MU_ARRAY ? ldo(A.muR, o + 2 * cs) : A.mu_s;
            hx = hx + src_val(sh, vx) * g.dt / ux;
            hy = hy + src_val(sh, vy) * g.dt / uy;
            hz = hz + src_val(sh, vz) * g.dt / uz;
        }
    });
}

template <bool MU_ARRAY, bool CEPS, int TY>
__global__ __launch_bounds__(FX * TY) void k_step_fused(FusedArgs A, Geo g, PassSrc sh, PassSrc se,
                                                       const int *__restrict__ done)
{
    if (done && *done) return;
    const Layout L = make_layout(g);
    const unsigned cs = A.cs, cs2 = 2 * A.cs;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const unsigned nbx = gridDim.x, nby = gridDim.y;
    const unsigned l = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int bx = (int)(l % nbx), by = (int)((l / nbx) % nby), bz = (int)(l / (nbx * nby));

    const int lane = threadIdx.x;
    const int k = bx * FXO - 1 + lane, j = by * TY + threadIdx.y;
    const int z0 = bz * g.zc, z1 = min(z0 + g.zc, Nz);
    const bool hcell = k >= 0 && k < Nx && j < Ny;   // computes H_new(i, j, k)
    const bool own = hcell && lane >= 1;             // owns the cell (E + writes)
    const bool low = hcell && j > 0;                 // also computes H_new(i, j-1, k)
    const bool jn = j + 1 < Ny, kn = k + 1 < Nx;

    // does this wave's column range touch any source block? (uniform per wave)
    bool row_src = false;
    {
        const int tk0 = bx * FXO - 1, tk1 = tk0 + FX;
        const int tz0 = z0 > 0 ? z0 - 1 : z0;
        for (int q = 0; q < sh.n; ++q) {
            const SrcBlk s = src_blk(sh.blk, q);
            if (s.k0 < tk1 && s.k0 + s.K > tk0 && s.j0 <= j && s.j0 + s.J > j - 1 && s.i0 < z1 &&
                s.i0 + s.I > tz0)
                row_src = true;
        }
    }
    const double sgx = g.bc[0] == 'E' ? -1.0 : (g.bc[0] == 'H' ? 1.0 : 0.0);
    const double sgy = g.bc[1] == 'E' ? -1.0 : (g.bc[1] == 'H' ? 1.0 : 0.0);
    const double sgz = g.bc[2] == 'E' ? -1.0 : (g.bc[2] == 'H' ? 1.0 : 0.0);

    const int zs = z0 > 0 ? z0 - 1 : z0;
    const unsigned oz = (unsigned)g.nxy * 8u, oy = (unsigned)Nx * 8u;
    const unsigned o0 = hcell ? ((unsigned)j * (unsigned)Nx + (unsigned)k) * 8u : 0u;

    // E_old(i) of this row and of row j-1 (register queues along z)
    double ex = 0, ey = 0, ez = 0, lx = 0, ly = 0, lz = 0;
    if (hcell) {
        const unsigned o = (unsigned)zs * oz + o0;
        ex = ldo(A.Eo, o); ey = ldo(A.Eo, o + cs); ez = ldo(A.Eo, o + cs2);
        if (low) { lx = ldo(A.Eo, o - oy); ly = ldo(A.Eo, o - oy + cs); lz = ldo(A.Eo, o - oy + cs2); }
    }
    double hxm = 0, hym = 0;   // H_new(i-1, j, k)

    for (int i = zs; i < z1; ++i) {
        const unsigned o = (unsigned)i * oz + o0;
        const bool in_ = i + 1 < Nz;
        const bool doE = own && i >= z0;
        const bool doL = low && i >= z0;
        // ---------------- all loads of this plane
        HIn a, b;   // a: this row, b: row j-1
        a.ex = ex; a.ey = ey; a.ez = ez;
        b.ex = lx; b.ey = ly; b.ez = lz;
        a.exi = a.eyi = b.exi = b.eyi = 0;
        double ezi = 0, lzi = 0;
        a.ezj = a.exj = 0;
        a.hx = a.hy = a.hz = b.hx = b.hy = b.hz = 0;
        double bex = 0, bey = 0, bez = 0;
        if (hcell) {
            if (in_) { a.exi = ldo(A.Eo, o + oz); a.eyi = ldo(A.Eo, o + oz + cs); ezi = ldo(A.Eo, o + oz + cs2); }
            if (jn) { a.ezj = ldo(A.Eo, o + oy + cs2); a.exj = ldo(A.Eo, o + oy); }
            a.hx = ldo(A.Ho, o); a.hy = ldo(A.Ho, o + cs); a.hz = ldo(A.Ho, o + cs2);
            h_load_aux<MU_ARRAY>(g, L, A, i, j, k, o, a);
        }
        if (doL) {
            const unsigned ol = o - oy;
            if (in_) { b.exi = ldo(A.Eo, ol + oz); b.eyi = ldo(A.Eo, ol + oz + cs); lzi = ldo(A.Eo, ol + oz + cs2); }
            b.hx = ldo(A.Ho, ol); b.hy = ldo(A.Ho, ol + cs); b.hz = ldo(A.Ho, ol + cs2);
            h_load_aux<MU_ARRAY>(g, L, A, i, j - 1, k, ol, b);
        } else if (low && in_) {   // prologue plane: only the z-queue of row j-1 advances
            lzi = ldo(A.Eo, o - oy + oz + cs2);
            b.exi = ldo(A.Eo, o - oy + oz); b.eyi = ldo(A.Eo, o - oy + oz + cs);
        }
        if (doE) { bex = ldo(A.bE, o); bey = ldo(A.bE, o + cs); bez = ldo(A.bE, o + cs2); }
        b.ezj = ez; b.exj = ex;                    // E_old(i, j, k) is row j-1's j+1 neighbour
        // E_old(i, j', k+1) from the next lane; lane 63 loads it (zero = PEC ghost)
        a.ezk = __shfl_down(ez, 1); a.eyk = __shfl_down(ey, 1);
        b.ezk = __shfl_down(lz, 1); b.eyk = __shfl_down(ly, 1);
        if (lane == FX - 1) {
            a.ezk = a.eyk = b.ezk = b.eyk = 0;
            if (hcell && kn) {
                a.ezk = ldo(A.Eo, o + 8u + cs2); a.eyk = ldo(A.Eo, o + 8u + cs);
                if (low) { b.ezk = ldo(A.Eo, o - oy + 8u + cs2); b.eyk = ldo(A.Eo, o - oy + 8u + cs); }
            }
        }
        // ---------------- H_new of this row and of row j-1
        double hx = 0, hy = 0, hz = 0, npa[6];
        if (hcell) {
            h_compute(g, a, hx, hy, hz, npa);
            if (row_src) h_sources<MU_ARRAY>(g, sh, A, i, j, k, o, hx, hy, hz);
        }
        double gx = 0, gy = 0, gz = 0, npb[6];
        if (doL) {
            h_compute(g, b, gx, gy, gz, npb);
            if (row_src) h_sources<MU_ARRAY>(g, sh, A, i, j - 1, k, o - oy, gx, gy, gz);
        }
        (void)gy;
        // H_new(i, j, k-1) from the previous lane (lane 0 is the halo column)
        double hz_k = __shfl_up(hz, 1), hy_k = __shfl_up(hy, 1);

        // ---------------- E_new and all stores (owner only)
        if (doE) {
            double hx_j = gx, hz_j = gz;   // H_new(i, j-1, k)
            if (j == 0) {
                hx_j = g.bc[1] == '0' ? 0.0 : sgy * hx;
                hz_j = g.bc[1] == '0' ? 0.0 : sgy * hz;
            }
            if (k == 0) {
                hz_k = g.bc[0] == '0' ? 0.0 : sgx * hz;
                hy_k = g.bc[0] == '0' ? 0.0 : sgx * hy;
            }
            double hx_m = hxm, hy_m = hym;   // H_new(i-1, j, k)
            if (i == 0) {
                hx_m = g.bc[2] == '0' ? 0.0 : sgz * hx;
                hy_m = g.bc[2] == '0' ? 0.0 : sgz * hy;
            }
            // E-side psi and tables of the owned cell
            const Slab &sx = a.sx, &sy = a.sy, &sz = a.sz;
            double qx0 = 0, qx1 = 0, qy0 = 0, qy1 = 0, qz0 = 0, qz1 = 0;
            if (sx.in) { qx0 = ldo(A.pe, psof(L, 0, sx.lo, 0) + sx.io); qx1 = ldo(A.pe, psof(L, 0, sx.lo, 1) + sx.io); }
            if (sy.in) { qy0 = ldo(A.pe, psof(L, 1, sy.lo, 0) + sy.io); qy1 = ldo(A.pe, psof(L, 1, sy.lo, 1) + sy.io); }
            if (sz.in) { qz0 = ldo(A.pe, psof(L, 2, sz.lo, 0) + sz.io); qz1 = ldo(A.pe, psof(L, 2, sz.lo, 1) + sz.io); }

            const double dHzdy = g.ry * (hz - hz_j);
            const double dHydz = g.rz * (hy - hy_m);
            double nex, ney, nez;
            if (CEPS) nex = ldo(A.aE, o) * ex + (dHzdy - dHydz) * bex;
            else nex = ex + (dHzdy - dHydz) * bex;
            const double dHxdz = g.rz * (hx - hx_m);
            const double dHzdx = g.rx * (hz - hz_k);
            if (CEPS) ney = ldo(A.aE, o + cs) * ey + (dHxdz - dHzdx) * bey;
            else ney = ey + (dHxdz - dHzdx) * bey;
            const double dHydx = g.rx * (hy - hy_k);
            const double dHxdy = g.ry * (hx - hx_j);
            if (CEPS) nez = ldo(A.aE, o + cs2) * ez + (dHydx - dHxdy) * bez;
            else nez = ez + (dHydx - dHxdy) * bez;
            if (sx.in) {
                const unsigned t = L.tbo[0] + 3 * L.tbl[0] + (unsigned)sx.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[0]), cc = ldc(A.tab, t + 2 * L.tbl[0]);
                qx0 = cc * dHydx + bb * qx0;
                qx1 = cc * dHzdx + bb * qx1;
                nez = nez + (qx0 - dHydx + dHydx / kap) * bez;
                ney = ney - (qx1 - dHzdx + dHzdx / kap) * bey;
            }
            if (sy.in) {
                const unsigned t = L.tbo[1] + 3 * L.tbl[1] + (unsigned)sy.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[1]), cc = ldc(A.tab, t + 2 * L.tbl[1]);
                qy0 = cc * dHxdy + bb * qy0;
                qy1 = cc * dHzdy + bb * qy1;
                nez = nez - (qy0 - dHxdy + dHxdy / kap) * bez;
                nex = nex + (qy1 - dHzdy + dHzdy / kap) * bex;
            }
            if (sz.in) {
                const unsigned t = L.tbo[2] + 3 * L.tbl[2] + (unsigned)sz.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[2]), cc = ldc(A.tab, t + 2 * L.tbl[2]);
                qz0 = cc * dHxdz + bb * qz0;
                qz1 = cc * dHydz + bb * qz1;
                nex = nex - (qz1 - dHydz + dHydz / kap) * bex;
                ney = ney + (qz0 - dHxdz + dHxdz / kap) * bey;
            }
            if (row_src) {
                for (int q = 0; q < se.n; ++q) {
                    long long m;
                    const SrcBlk sb = src_blk(se.blk, q);
                    if (in_box(sb, i, j, k, m)) {
                        const double2 vx = ld2(sb.f[0] + m), vy = ld2(sb.f[1] + m), vz = ld2(sb.f[2] + m);
                        nex = nex - src_val(se, vx) * bex;
                        ney = ney - src_val(se, vy) * bey;
                        nez = nez - src_val(se, vz) * bez;
                    }
                }
            }
            sto(A.Hn, o, hx); sto(A.Hn, o + cs, hy); sto(A.Hn, o + cs2, hz);
            sto(A.En, o, nex); sto(A.En, o + cs, ney); sto(A.En, o + cs2, nez);
            if (sx.in) { sto(A.phn, psof(L, 0, sx.lo, 0) + sx.io, npa[0]); sto(A.phn, psof(L, 0, sx.lo, 1) + sx.io, npa[1]);
                         sto(A.pe, psof(L, 0, sx.lo, 0) + sx.io, qx0); sto(A.pe, psof(L, 0, sx.lo, 1) + sx.io, qx1); }
            if (sy.in) { sto(A.phn, psof(L, 1, sy.lo, 0) + sy.io, npa[2]); sto(A.phn, psof(L, 1, sy.lo, 1) + sy.io, npa[3]);
                         sto(A.pe, psof(L, 1, sy.lo, 0) + sy.io, qy0); sto(A.pe, psof(L, 1, sy.lo, 1) + sy.io, qy1); }
            if (sz.in) { sto(A.phn, psof(L, 2, sz.lo, 0) + sz.io, npa[4]); sto(A.phn, psof(L, 2, sz.lo, 1) + sz.io, npa[5]);
                         sto(A.pe, psof(L, 2, sz.lo, 0) + sz.io, qz0); sto(A.pe, psof(L, 2, sz.lo, 1) + sz.io, qz1); }
        }
        hxm = hx; hym = hy;
        ex = a.exi; ey = a.eyi; ez = ezi;
        lx = b.exi; ly = b.eyi; lz = lzi;
    }
}

}  // namespace emf

namespace emf {

// ---------------------------------------------------------------------------
// Fused step, LDS variant: a 64 x TY work-group owns 63 x (TY-1) cells of a
// tile (lane 0 = halo column k0-1, wave 0 = halo row j0-1; both compute H_new
// only) and sweeps a chunk of z planes, starting one plane below the chunk
// (H_new of plane z0-1 is recomputed, nothing written).  H_new of row j-1
// comes from LDS (one barrier per plane, buffers alternate by plane parity),
// H_new of k-1 from the neighbouring lane, H_new of i-1 from registers.
//
// Per plane every load goes out before any compute: the coefficient index
// first, then the E and H inputs of H_new, the CPML psi of both updates, the
// E coefficients, and the coefficient-table gather last; the thread then
// waits for memory once per plane.  Neighbour loads past the grid's max faces
// read the cell itself and are replaced by the zero ghost, so the loads need
// no branches.  Cells outside the grid (halo lanes at the edges) load cell
// (i, 0, 0) and hold zeros.
template <bool MU_ARRAY, int CM, int TY>
__global__ __launch_bounds__(FX * TY) void k_step_lds(FusedArgs A, Geo g, PassSrc sh, PassSrc se,
                                                     const int *__restrict__ done)
{
    if (done && *done) return;
    constexpr bool CEPS = CM == CM_COMPLEX;
    __shared__ double lds[2][TY][FX][2];   // (Hx, Hz) of H_new per row
    const Layout L = make_layout(g);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const unsigned nbx = gridDim.x, nby = gridDim.y;
    const unsigned l = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int bx = (int)(l % nbx), by = (int)((l / nbx) % nby), bz = (int)(l / (nbx * nby));
    const int lane = threadIdx.x, row = threadIdx.y;
    const int k = bx * FXO - 1 + lane, j = by * (TY - 1) - 1 + row;
    const int z0 = bz * g.zc, z1 = min(z0 + g.zc, Nz);
    const bool hcell = k >= 0 && k < Nx && j >= 0 && j < Ny;
    const bool own = hcell && lane >= 1 && row >= 1;
    const bool jn = hcell && j + 1 < Ny, kn = hcell && k + 1 < Nx;
    const unsigned cs = A.cs, cs2 = 2 * A.cs;
    const unsigned oz = (unsigned)g.nxy * 8u, oy = (unsigned)Nx * 8u;
    const unsigned o0 = hcell ? ((unsigned)j * (unsigned)Nx + (unsigned)k) * 8u : 0u;
    const unsigned djp = jn ? oy : 0u, dkp = kn ? 8u : 0u;
    bool tsh = false, tse = false;   // does the tile meet an H / E source block?
    {
        const int tk0 = bx * FXO - 1, tk1 = tk0 + FX, tj0 = by * (TY - 1) - 1, tj1 = tj0 + TY;
        const int tz0 = z0 > 0 ? z0 - 1 : z0;
        for (int q = 0; q < sh.n; ++q) {
            const SrcBlk s = src_blk(sh.blk, q);
            if (s.k0 < tk1 && s.k0 + s.K > tk0 && s.j0 < tj1 && s.j0 + s.J > tj0 && s.i0 < z1 &&
                s.i0 + s.I > tz0)
                tsh = true;
        }
        for (int q = 0; q < se.n; ++q) {
            const SrcBlk s = src_blk(se.blk, q);
            if (s.k0 < tk1 && s.k0 + s.K > tk0 && s.j0 < tj1 && s.j0 + s.J > tj0 && s.i0 < z1 &&
                s.i0 + s.I > z0)
                tse = true;
        }
    }
    const double sgx = g.bc[0] == 'E' ? -1.0 : (g.bc[0] == 'H' ? 1.0 : 0.0);
    const double sgy = g.bc[1] == 'E' ? -1.0 : (g.bc[1] == 'H' ? 1.0 : 0.0);
    const double sgz = g.bc[2] == 'E' ? -1.0 : (g.bc[2] == 'H' ? 1.0 : 0.0);
    const int zs = z0 > 0 ? z0 - 1 : z0;
    double ex, ey, ez;   // E_old(i)
    {
        const unsigned o = (unsigned)zs * oz + o0;
        ex = ldo(A.Eo, o); ey = ldo(A.Eo, o + cs); ez = ldo(A.Eo, o + cs2);
        if (!hcell) ex = ey = ez = 0;
    }
    double hxm = 0, hym = 0;   // H_new(i-1)
    for (int i = zs; i < z1; ++i) {
        const unsigned o = (unsigned)i * oz + o0;
        const bool doE = own && i >= z0;
        const bool up = hcell && i + 1 < Nz;
        const unsigned dip = up ? oz : 0u;
        // ---------------- all loads of this plane
        unsigned ci = 0;
        if (CM == CM_INDEX) ci = coef_index(A, o);
        HIn d;
        d.ex = ex; d.ey = ey; d.ez = ez;
        d.exi = ldo(A.Eo, o + dip); d.eyi = ldo(A.Eo, o + dip + cs);
        double ezi = ldo(A.Eo, o + dip + cs2);
        d.ezj = ldo(A.Eo, o + djp + cs2); d.exj = ldo(A.Eo, o + djp);
        d.hx = ldo(A.Ho, o); d.hy = ldo(A.Ho, o + cs); d.hz = ldo(A.Ho, o + cs2);
        double ezk = 0, eyk = 0;   // E_old(k+1) of lane 63 (the others shuffle)
        if (lane == FX - 1) { ezk = ldo(A.Eo, o + dkp + cs2); eyk = ldo(A.Eo, o + dkp + cs); }
        if (hcell) {
            h_load_aux<MU_ARRAY>(g, L, A, i, j, k, o, d);
        } else {
            d.mx = d.my = d.mz = 0;
            d.sx.in = d.sy.in = d.sz.in = false;
        }
        double bex = 0, bey = 0, bez = 0, aex = 1, aey = 1, aez = 1;
        if (CM != CM_INDEX) { bex = ldo(A.bE, o); bey = ldo(A.bE, o + cs); bez = ldo(A.bE, o + cs2); }
        if (CEPS) { aex = ldo(A.aE, o); aey = ldo(A.aE, o + cs); aez = ldo(A.aE, o + cs2); }
        double qx0 = 0, qx1 = 0, qy0 = 0, qy1 = 0, qz0 = 0, qz1 = 0;
        if (doE) {
            if (d.sx.in) { qx0 = ldo(A.pe, psof(L, 0, d.sx.lo, 0) + d.sx.io); qx1 = ldo(A.pe, psof(L, 0, d.sx.lo, 1) + d.sx.io); }
            if (d.sy.in) { qy0 = ldo(A.pe, psof(L, 1, d.sy.lo, 0) + d.sy.io); qy1 = ldo(A.pe, psof(L, 1, d.sy.lo, 1) + d.sy.io); }
            if (d.sz.in) { qz0 = ldo(A.pe, psof(L, 2, d.sz.lo, 0) + d.sz.io); qz1 = ldo(A.pe, psof(L, 2, d.sz.lo, 1) + d.sz.io); }
        }
        if (CM == CM_INDEX) coef_gather(A, ci, bex, bey, bez);
        // ---------------- H_new(i, j, k)
        if (!up) { d.exi = d.eyi = ezi = 0; }
        if (!jn) { d.ezj = d.exj = 0; }
        {
            const double zk = __shfl_down(d.ez, 1), yk = __shfl_down(d.ey, 1);
            if (lane == FX - 1) {
                d.ezk = kn ? ezk : 0.0; d.eyk = kn ? eyk : 0.0;
            } else {
                d.ezk = zk; d.eyk = yk;   // zero from a lane outside the grid
            }
        }
        double hx = 0, hy = 0, hz = 0, np[6];
        if (hcell) {
            h_compute(g, d, hx, hy, hz, np);
            if (tsh) h_sources<MU_ARRAY>(g, sh, A, i, j, k, o, hx, hy, hz);
        }
        // ---------------- exchange H_new with the neighbours
        const int par = i & 1;
        lds[par][row][lane][0] = hx;
        lds[par][row][lane][1] = hz;
        __syncthreads();
        double hz_k = __shfl_up(hz, 1), hy_k = __shfl_up(hy, 1);
        // ---------------- E_new and all stores (owner only)
        if (doE) {
            double hx_j = lds[par][row - 1][lane][0], hz_j = lds[par][row - 1][lane][1];
            if (j == 0) { hx_j = g.bc[1] == '0' ? 0.0 : sgy * hx; hz_j = g.bc[1] == '0' ? 0.0 : sgy * hz; }
            if (k == 0) { hz_k = g.bc[0] == '0' ? 0.0 : sgx * hz; hy_k = g.bc[0] == '0' ? 0.0 : sgx * hy; }
            double hx_m = hxm, hy_m = hym;
            if (i == 0) { hx_m = g.bc[2] == '0' ? 0.0 : sgz * hx; hy_m = g.bc[2] == '0' ? 0.0 : sgz * hy; }
            const double dHzdy = g.ry * (hz - hz_j);
            const double dHydz = g.rz * (hy - hy_m);
            double nex = CEPS ? aex * d.ex + (dHzdy - dHydz) * bex : d.ex + (dHzdy - dHydz) * bex;
            const double dHxdz = g.rz * (hx - hx_m);
            const double dHzdx = g.rx * (hz - hz_k);
            double ney = CEPS ? aey * d.ey + (dHxdz - dHzdx) * bey : d.ey + (dHxdz - dHzdx) * bey;
            const double dHydx = g.rx * (hy - hy_k);
            const double dHxdy = g.ry * (hx - hx_j);
            double nez = CEPS ? aez * d.ez + (dHydx - dHxdy) * bez : d.ez + (dHydx - dHxdy) * bez;
            if (d.sx.in) {
                const unsigned t = L.tbo[0] + 3 * L.tbl[0] + (unsigned)d.sx.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[0]), cc = ldc(A.tab, t + 2 * L.tbl[0]);
                qx0 = cc * dHydx + bb * qx0;
                qx1 = cc * dHzdx + bb * qx1;
                nez = nez + (qx0 - dHydx + dHydx / kap) * bez;
                ney = ney - (qx1 - dHzdx + dHzdx / kap) * bey;
            }
            if (d.sy.in) {
                const unsigned t = L.tbo[1] + 3 * L.tbl[1] + (unsigned)d.sy.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[1]), cc = ldc(A.tab, t + 2 * L.tbl[1]);
                qy0 = cc * dHxdy + bb * qy0;
                qy1 = cc * dHzdy + bb * qy1;
                nez = nez - (qy0 - dHxdy + dHxdy / kap) * bez;
                nex = nex + (qy1 - dHzdy + dHzdy / kap) * bex;
            }
            if (d.sz.in) {
                const unsigned t = L.tbo[2] + 3 * L.tbl[2] + (unsigned)d.sz.par * 8u;
                const double kap = ldc(A.tab, t), bb = ldc(A.tab, t + L.tbl[2]), cc = ldc(A.tab, t + 2 * L.tbl[2]);
                qz0 = cc * dHxdz + bb * qz0;
                qz1 = cc * dHydz + bb * qz1;
                nex = nex - (qz1 - dHydz + dHydz / kap) * bex;
                ney = ney + (qz0 - dHxdz + dHxdz / kap) * bey;
            }
            if (tse) {
                for (int q = 0; q < se.n; ++q) {
                    long long m;
                    const SrcBlk sb = src_blk(se.blk, q);
                    if (in_box(sb, i, j, k, m)) {
                        const double2 vx = ld2(sb.f[0] + m), vy = ld2(sb.f[1] + m), vz = ld2(sb.f[2] + m);
                        nex = nex - src_val(se, vx) * bex;
                        ney = ney - src_val(se, vy) * bey;
                        nez = nez - src_val(se, vz) * bez;
                    }
                }
            }
            sto(A.Hn, o, hx); sto(A.Hn, o + cs, hy); sto(A.Hn, o + cs2, hz);
            sto(A.En, o, nex); sto(A.En, o + cs, ney); sto(A.En, o + cs2, nez);
            if (d.sx.in) { sto(A.phn, psof(L, 0, d.sx.lo, 0) + d.sx.io, np[0]); sto(A.phn, psof(L, 0, d.sx.lo, 1) + d.sx.io, np[1]);
                           sto(A.pe, psof(L, 0, d.sx.lo, 0) + d.sx.io, qx0); sto(A.pe, psof(L, 0, d.sx.lo, 1) + d.sx.io, qx1); }
            if (d.sy.in) { sto(A.phn, psof(L, 1, d.sy.lo, 0) + d.sy.io, np[2]); sto(A.phn, psof(L, 1, d.sy.lo, 1) + d.sy.io, np[3]);
                           sto(A.pe, psof(L, 1, d.sy.lo, 0) + d.sy.io, qy0); sto(A.pe, psof(L, 1, d.sy.lo, 1) + d.sy.io, qy1); }
            if (d.sz.in) { sto(A.phn, psof(L, 2, d.sz.lo, 0) + d.sz.io, np[4]); sto(A.phn, psof(L, 2, d.sz.lo, 1) + d.sz.io, np[5]);
                           sto(A.pe, psof(L, 2, d.sz.lo, 0) + d.sz.io, qz0); sto(A.pe, psof(L, 2, d.sz.lo, 1) + d.sz.io, qz1); }
        }
        hxm = hx; hym = hy;
        ex = d.exi; ey = d.eyi; ez = ezi;
    }
}

}  // namespace emf
