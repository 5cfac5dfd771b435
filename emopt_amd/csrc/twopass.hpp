// Separate H and E passes (update_H / update_E of src/fdtd.cpp:121-755), in
// place on one buffer set, with the same argument layout as the fused step.
//
// HBM traffic per cell: H pass reads E(3) H(3) and writes H(3) = 72 B; E pass
// reads H(3) E(3) b_eps(3) and writes E(3) = 96 B; plus the psi slabs.  With
// real eps the E pass can read a 2- or 4-byte index per cell into a small
// table of distinct (b_x, b_y, b_z) instead of the three coefficients (the
// same doubles, so the same results): 74 or 76 B per cell.  The index loads
// go out before a plane's other loads and the table gather after them.
//
// A 64 x TY work-group owns a tile of (x, y) columns and sweeps zc planes
// (zc = 1 is one cell per thread); each wave is one row, so the row index is
// wave-uniform and row-dependent values (y-slab tests, CPML table entries)
// live in scalar registers.  Every load a plane needs -- fields,
// coefficients, psi and CPML table entries -- is issued before any store, so
// a thread waits for memory once per plane.  The x neighbour comes from the
// adjacent lane (the wave's edge lane loads it); row and plane neighbours
// come through L1/L2.
#pragma once
#include <type_traits>

#include "fused.hpp"

namespace emf {

constexpr int PX = 64;

// Tile of work-group blockIdx.  order 0: x fastest, then y, then z.  order 1:
// z fastest, then x, then y.  order 2: as 1, after regrouping the linear id
// by XCD (the dispatcher deals consecutive work-groups to the 8 XCDs in turn),
// so that z-neighbouring tiles, which share a halo plane, run on one XCD and
// share its L2.  order 3: as 0, but the E pass (rev) walks the tiles
// backwards, so that each pass starts where the previous one ended and
// first reads what that pass wrote last, while it is still in the Infinity
// Cache.  The mapping is a bijection for any work-group count.
__device__ __forceinline__ void pass_tile(const Geo &g, int ty, int &k, int &j, int &z0, int &z1,
                                          bool rev = false)
{
    const unsigned nbx = gridDim.x, nby = gridDim.y, nbz = gridDim.z;
    unsigned l = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    if (rev && g.order == 3) l = nbx * nby * nbz - 1 - l;
    unsigned bx, by, bz;
    if (g.order == 0 || g.order == 3) {
        bx = l % nbx; by = (l / nbx) % nby; bz = l / (nbx * nby);
    } else {
        if (g.order == 2) {
            const unsigned nb = nbx * nby * nbz, q = nb / 8, r = nb % 8;
            const unsigned xcd = l % 8, slot = l / 8;
            l = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
        }
        bz = l % nbz; bx = (l / nbz) % nbx; by = l / (nbz * nbx);
    }
    k = (int)(bx * PX + threadIdx.x);
    j = (int)(by * ty + threadIdx.y);
    z0 = (int)bz * g.zc;
    z1 = min(z0 + g.zc, g.Nz);
}

// ------------------------------------------------------------------ H pass
// Cells (i, j, k), z0 <= i < z1, of one (k, j) column.  With cap != nullptr
// the pre-update E and H (the state after the steps taken so far) are also
// stored as a snapshot: component c of cell g goes to cap[2 * (c * n + g)]
// (cap points at the .x or .y half of a six-component result block).
template <bool MU_ARRAY, bool NT = false, typename OT = unsigned>
__device__ __forceinline__ void h_column(const FusedArgs &A, const Geo &g, const PassSrc &sh,
                                         int k, int j, int z0, int z1, double *cap,
                                         long long capn, unsigned long long smask)
{
    const Layout L = make_layout(g);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const OT cs = cstride<OT>(A), cs2 = 2 * cstride<OT>(A);
    const OT oz = (OT)g.nxy * 8u, oy = (OT)Nx * 8u;
    const bool jn = j + 1 < Ny, kn = k + 1 < Nx;
    // may a source block meet this column?  (smask: col_src_mask)
    const bool col_src = smask != 0;
    OT o = ((OT)z0 * (OT)Ny + (OT)j) * (OT)Nx * 8u + (OT)k * 8u;
    double ex = ldo(A.Eo, o), ey = ldo(A.Eo, o + cs);
    for (int i = z0; i < z1; ++i, o += oz) {
        HIn d;
        d.ex = ex; d.ey = ey;
        d.exi = d.eyi = d.ezj = d.exj = d.ezk = d.eyk = 0;
        if (i + 1 < Nz) { d.exi = ldo(A.Eo, o + oz); d.eyi = ldo(A.Eo, o + oz + cs); }
        d.ez = ldo(A.Eo, o + cs2);
        if (jn) { d.ezj = ldo(A.Eo, o + oy + cs2); d.exj = ldo(A.Eo, o + oy); }
        d.hx = ldo(A.Ho, o); d.hy = ldo(A.Ho, o + cs); d.hz = ldo(A.Ho, o + cs2);
        // E(k+1) comes from the next lane; the last lane loads it
        double ezk = 0, eyk = 0;
        if (threadIdx.x == PX - 1 && kn) { ezk = ldo(A.Eo, o + 8u + cs2); eyk = ldo(A.Eo, o + 8u + cs); }
        h_load_aux<MU_ARRAY, NT, OT>(g, L, A, i, j, k, o, d);
        d.ezk = __shfl_down(d.ez, 1);
        d.eyk = __shfl_down(d.ey, 1);
        if (threadIdx.x == PX - 1) { d.ezk = ezk; d.eyk = eyk; }
        if (!kn) d.ezk = d.eyk = 0;   // zero ghost past the max face
        if (cap) {
            const long long gc = (long long)(o >> 3);
            *gptr(cap + 2 * gc) = d.ex; *gptr(cap + 2 * (capn + gc)) = d.ey;
            *gptr(cap + 2 * (2 * capn + gc)) = d.ez; *gptr(cap + 2 * (3 * capn + gc)) = d.hx;
            *gptr(cap + 2 * (4 * capn + gc)) = d.hy; *gptr(cap + 2 * (5 * capn + gc)) = d.hz;
        }
        double hx, hy, hz, np[6];
        h_compute(g, d, hx, hy, hz, np);
        if (col_src) h_sources<MU_ARRAY, OT>(g, sh, A, i, j, k, o, hx, hy, hz, smask);
        sto(A.Hn, o, hx); sto(A.Hn, o + cs, hy); sto(A.Hn, o + cs2, hz);
        if (d.sx.in) { stp<NT>(A.phn, psof(L, 0, d.sx.lo, 0) + d.sx.io, np[0]); stp<NT>(A.phn, psof(L, 0, d.sx.lo, 1) + d.sx.io, np[1]); }
        if (d.sy.in) { stp<NT>(A.phn, psof(L, 1, d.sy.lo, 0) + d.sy.io, np[2]); stp<NT>(A.phn, psof(L, 1, d.sy.lo, 1) + d.sy.io, np[3]); }
        if (d.sz.in) { stp<NT>(A.phn, psof(L, 2, d.sz.lo, 0) + d.sz.io, np[4]); stp<NT>(A.phn, psof(L, 2, d.sz.lo, 1) + d.sz.io, np[5]); }
        ex = d.exi; ey = d.eyi;
    }
}

// OCC = minimum waves per SIMD the register allocation must allow (4 fits
// the passes without spilling; 5 trades a few spilled registers for 25 %
// more waves in flight).  NT: streaming psi access (see ldn).  BIG: 64-bit
// cell offsets (grids above 178 M cells).
template <bool MU_ARRAY, int TY, int OCC = 4, bool NT = false, bool BIG = false>
__global__ __launch_bounds__(PX * TY, OCC) void k_pass_h(FusedArgs A, Geo g, PassSrc sh,
                                                        const int *__restrict__ done)
{
    if (done && *done) return;
    int k, j, z0, z1;
    pass_tile(g, TY, k, j, z0, z1);
    j = __builtin_amdgcn_readfirstlane(j);   // one row per wave
    if (k >= g.Nx || j >= g.Ny) return;
    const unsigned long long smask = col_src_mask(sh, k, j, z0, z1);
    using OT = typename std::conditional<BIG, unsigned long long, unsigned>::type;
    h_column<MU_ARRAY, NT, OT>(A, g, sh, k, j, z0, z1, nullptr, 0, smask);
}

// The H pass (uniform mu, 32-bit offsets) with the y-neighbour row staged in
// LDS: each wave (one row) puts its own row's E_x, E_z of the plane into a
// double-buffered LDS slot, the work-group passes one barrier per plane, and
// rows 0 .. TY-2 read E_x, E_z at j + 1 from there; the last row loads them
// from memory as h_column does.  The values are those h_column loads (the H
// pass does not write E), so the results are identical.  Every wave of the
// work-group runs every plane of the column range, which depends on
// blockIdx only: rows past the grid and lanes past Nx skip their loads,
// arithmetic and stores but not the barriers.  A fast wave cannot overwrite a
// buffer before the slow ones have read it: it writes a plane's buffer only
// after passing the previous plane's barrier.
template <int TY, bool NT>
__global__ __launch_bounds__(PX * TY, 4) void k_pass_h_lds(FusedArgs A, Geo g, PassSrc sh,
                                                          const int *__restrict__ done)
{
    __shared__ double lx[2][TY][PX], lz[2][TY][PX];
    if (done && *done) return;                 // the same for every thread of the grid
    int k, j, z0, z1;
    pass_tile(g, TY, k, j, z0, z1);
    j = __builtin_amdgcn_readfirstlane(j);   // one row per wave
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const bool act = k < g.Nx && j < g.Ny;
    const unsigned long long smask = act ? col_src_mask(sh, k, j, z0, z1) : 0ULL;
    const Layout L = make_layout(g);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const unsigned cs = cstride<unsigned>(A), cs2 = 2 * cstride<unsigned>(A);
    const unsigned oz = (unsigned)g.nxy * 8u, oy = (unsigned)Nx * 8u;
    const bool jn = j + 1 < Ny, kn = k + 1 < Nx;
    const bool col_src = smask != 0;
    unsigned o = act ? ((unsigned)z0 * (unsigned)Ny + (unsigned)j) * (unsigned)Nx * 8u +
                           (unsigned)k * 8u
                     : 0u;
    double ex = 0, ey = 0;
    if (act) { ex = ldo(A.Eo, o); ey = ldo(A.Eo, o + cs); }
    for (int i = z0; i < z1; ++i, o += oz) {
        const int b = (i - z0) & 1;
        HIn d;
        d.ex = ex; d.ey = ey; d.ez = 0;
        d.exi = d.eyi = d.ezj = d.exj = d.ezk = d.eyk = 0;
        double ezk = 0, eyk = 0;
        if (act) {
            if (i + 1 < Nz) { d.exi = ldo(A.Eo, o + oz); d.eyi = ldo(A.Eo, o + oz + cs); }
            d.ez = ldo(A.Eo, o + cs2);
            if (jn && ty == TY - 1) { d.ezj = ldo(A.Eo, o + oy + cs2); d.exj = ldo(A.Eo, o + oy); }
            d.hx = ldo(A.Ho, o); d.hy = ldo(A.Ho, o + cs); d.hz = ldo(A.Ho, o + cs2);
            if (tx == PX - 1 && kn) { ezk = ldo(A.Eo, o + 8u + cs2); eyk = ldo(A.Eo, o + 8u + cs); }
            h_load_aux<false, NT, unsigned>(g, L, A, i, j, k, o, d);
        }
        lx[b][ty][tx] = d.ex;
        lz[b][ty][tx] = d.ez;
        __syncthreads();
        if (act) {
            if (jn && ty < TY - 1) { d.exj = lx[b][ty + 1][tx]; d.ezj = lz[b][ty + 1][tx]; }
            d.ezk = __shfl_down(d.ez, 1);
            d.eyk = __shfl_down(d.ey, 1);
            if (tx == PX - 1) { d.ezk = ezk; d.eyk = eyk; }
            if (!kn) d.ezk = d.eyk = 0;   // zero ghost past the max face
            double hx, hy, hz, np[6];
            h_compute(g, d, hx, hy, hz, np);
            if (col_src) h_sources<false, unsigned>(g, sh, A, i, j, k, o, hx, hy, hz, smask);
            sto(A.Hn, o, hx); sto(A.Hn, o + cs, hy); sto(A.Hn, o + cs2, hz);
            if (d.sx.in) { stp<NT>(A.phn, psof(L, 0, d.sx.lo, 0) + d.sx.io, np[0]); stp<NT>(A.phn, psof(L, 0, d.sx.lo, 1) + d.sx.io, np[1]); }
            if (d.sy.in) { stp<NT>(A.phn, psof(L, 1, d.sy.lo, 0) + d.sy.io, np[2]); stp<NT>(A.phn, psof(L, 1, d.sy.lo, 1) + d.sy.io, np[3]); }
            if (d.sz.in) { stp<NT>(A.phn, psof(L, 2, d.sz.lo, 0) + d.sz.io, np[4]); stp<NT>(A.phn, psof(L, 2, d.sz.lo, 1) + d.sz.io, np[5]); }
        }
        ex = d.exi; ey = d.eyi;
    }
}

// ------------------------------------------------------------------ E pass
// One cell (i, j, k) at byte offset o; hxm/hym hold H_x, H_y at i-1 and are
// advanced to plane i.  With CM_INDEX, ci is the cell's coefficient index,
// loaded by the caller ahead of the plane's other loads: the table gather is
// issued last, after every other load of the plane (psi included).
template <int CM, bool NT = false, typename OT = unsigned>
__device__ __forceinline__ void e_plane(const FusedArgs &A, const Geo &g, const PassSrc &se,
                                        const Layout &L, int k, int j, int i, OT o,
                                        unsigned ci, unsigned long long smask, double &hxm,
                                        double &hym)
{
    const OT cs = cstride<OT>(A), cs2 = 2 * cstride<OT>(A);
    const OT oy = (OT)g.Nx * 8u;
    // min-face ghost rule: '0' -> 0, 'E' -> -H(first), 'H' -> +H(first)
    const double sgx = g.bc[0] == 'E' ? -1.0 : (g.bc[0] == 'H' ? 1.0 : 0.0);
    const double sgy = g.bc[1] == 'E' ? -1.0 : (g.bc[1] == 'H' ? 1.0 : 0.0);
    const bool bx0 = g.bc[0] == '0', by0 = g.bc[1] == '0';
    constexpr bool CEPS = CM == CM_COMPLEX;
    // ---- all loads of this plane
    const double hx = ldo(A.Ho, o), hy = ldo(A.Ho, o + cs), hz = ldo(A.Ho, o + cs2);
    // row j-1; at j == 0 the own cell stands in (an unconditional load keeps
    // the count of loads behind the index fixed) and the ghost rule below
    // replaces it
    const OT dj = j > 0 ? oy : (OT)0;
    double hz_j = ldo(A.Ho, o - dj + cs2), hx_j = ldo(A.Ho, o - dj);
    const double ex = ldo(A.Eo, o), ey = ldo(A.Eo, o + cs), ez = ldo(A.Eo, o + cs2);
    double bx = 0, by = 0, bz = 0;
    if (CM != CM_INDEX) { bx = ldo(A.bE, o); by = ldo(A.bE, o + cs); bz = ldo(A.bE, o + cs2); }
    double ax = 1, ay = 1, az = 1;
    if (CEPS) { ax = ldo(A.aE, o); ay = ldo(A.aE, o + cs); az = ldo(A.aE, o + cs2); }
    const Slab sx = slab_x(g, i, j, k), sy = slab_y(g, i, j, k), sz = slab_z(g, i, j, k);
    double qx0 = 0, qx1 = 0, qy0 = 0, qy1 = 0, qz0 = 0, qz1 = 0;
    double kx = 1, tbx = 0, tcx = 0, ky = 1, tby = 0, tcy = 0, kz = 1, tbz = 0, tcz = 0;
    if (sx.in) {
        qx0 = ldp<NT>(A.pe, psof(L, 0, sx.lo, 0) + sx.io); qx1 = ldp<NT>(A.pe, psof(L, 0, sx.lo, 1) + sx.io);
        const unsigned t = L.tbo[0] + 3 * L.tbl[0] + (unsigned)sx.par * 8u;
        kx = ldc(A.tab, t); tbx = ldc(A.tab, t + L.tbl[0]); tcx = ldc(A.tab, t + 2 * L.tbl[0]);
    }
    if (sy.in) {
        qy0 = ldp<NT>(A.pe, psof(L, 1, sy.lo, 0) + sy.io); qy1 = ldp<NT>(A.pe, psof(L, 1, sy.lo, 1) + sy.io);
        const unsigned t = L.tbo[1] + 3 * L.tbl[1] + (unsigned)sy.par * 8u;
        ky = ldc(A.tab, t); tby = ldc(A.tab, t + L.tbl[1]); tcy = ldc(A.tab, t + 2 * L.tbl[1]);
    }
    if (sz.in) {
        qz0 = ldp<NT>(A.pe, psof(L, 2, sz.lo, 0) + sz.io); qz1 = ldp<NT>(A.pe, psof(L, 2, sz.lo, 1) + sz.io);
        const unsigned t = L.tbo[2] + 3 * L.tbl[2] + (unsigned)sz.par * 8u;
        kz = ldc(A.tab, t); tbz = ldc(A.tab, t + L.tbl[2]); tcz = ldc(A.tab, t + 2 * L.tbl[2]);
    }
    // H(k-1) comes from the previous lane; lane 0 loads it
    double hz_k0 = 0, hy_k0 = 0;
    if (threadIdx.x == 0 && k > 0) { hz_k0 = ldo(A.Ho, o - 8u + cs2); hy_k0 = ldo(A.Ho, o - 8u + cs); }
    if (CM == CM_INDEX) {
        // a wave whose cells share one table entry (bulk material) reads it
        // with scalar loads; the others gather per lane
        const unsigned c0 = __builtin_amdgcn_readfirstlane(ci);
        if (__all((int)(ci == c0))) {
            const unsigned t = c0 * 32u;
            bx = ldc(A.bT, t); by = ldc(A.bT, t + 8u); bz = ldc(A.bT, t + 16u);
        } else {
            coef_gather(A, ci, bx, by, bz);
        }
    }
    double hz_k = __shfl_up(hz, 1), hy_k = __shfl_up(hy, 1);
    if (threadIdx.x == 0) { hz_k = hz_k0; hy_k = hy_k0; }
    // ---- boundary ghosts at the min faces
    if (j == 0) { hz_j = by0 ? 0.0 : sgy * hz; hx_j = by0 ? 0.0 : sgy * hx; }
    if (k == 0) { hz_k = bx0 ? 0.0 : sgx * hz; hy_k = bx0 ? 0.0 : sgx * hy; }
    // ---- update (src/fdtd.cpp:554-666)
    const double dHzdy = g.ry * (hz - hz_j);
    const double dHydz = g.rz * (hy - hym);
    double nex = CEPS ? ax * ex + (dHzdy - dHydz) * bx : ex + (dHzdy - dHydz) * bx;
    const double dHxdz = g.rz * (hx - hxm);
    const double dHzdx = g.rx * (hz - hz_k);
    double ney = CEPS ? ay * ey + (dHxdz - dHzdx) * by : ey + (dHxdz - dHzdx) * by;
    const double dHydx = g.rx * (hy - hy_k);
    const double dHxdy = g.ry * (hx - hx_j);
    double nez = CEPS ? az * ez + (dHydx - dHxdy) * bz : ez + (dHydx - dHxdy) * bz;
    if (sx.in) {
        qx0 = tcx * dHydx + tbx * qx0;
        qx1 = tcx * dHzdx + tbx * qx1;
        nez = nez + (qx0 - dHydx + dHydx / kx) * bz;
        ney = ney - (qx1 - dHzdx + dHzdx / kx) * by;
    }
    if (sy.in) {
        qy0 = tcy * dHxdy + tby * qy0;
        qy1 = tcy * dHzdy + tby * qy1;
        nez = nez - (qy0 - dHxdy + dHxdy / ky) * bz;
        nex = nex + (qy1 - dHzdy + dHzdy / ky) * bx;
    }
    if (sz.in) {
        qz0 = tcz * dHxdz + tbz * qz0;
        qz1 = tcz * dHydz + tbz * qz1;
        nex = nex - (qz1 - dHydz + dHydz / kz) * bx;
        ney = ney + (qz0 - dHxdz + dHxdz / kz) * by;
    }
    if (smask != 0) {
        src_each(se, smask, [&](int q) {
            long long m;
            const SrcBlk sb = src_blk(se.blk, q);
            if (in_box(sb, i, j, k, m)) {
                const double2 vx = ld2(sb.f[0] + m), vy = ld2(sb.f[1] + m), vz = ld2(sb.f[2] + m);
                nex = nex - src_val(se, vx) * bx;
                ney = ney - src_val(se, vy) * by;
                nez = nez - src_val(se, vz) * bz;
            }
        });
    }
    sto(A.En, o, nex); sto(A.En, o + cs, ney); sto(A.En, o + cs2, nez);
    if (sx.in) { stp<NT>(A.pe, psof(L, 0, sx.lo, 0) + sx.io, qx0); stp<NT>(A.pe, psof(L, 0, sx.lo, 1) + sx.io, qx1); }
    if (sy.in) { stp<NT>(A.pe, psof(L, 1, sy.lo, 0) + sy.io, qy0); stp<NT>(A.pe, psof(L, 1, sy.lo, 1) + sy.io, qy1); }
    if (sz.in) { stp<NT>(A.pe, psof(L, 2, sz.lo, 0) + sz.io, qz0); stp<NT>(A.pe, psof(L, 2, sz.lo, 1) + sz.io, qz1); }
    hxm = hx; hym = hy;
}

// Cells z0 <= i < z1 of one (k, j) column.  ZU > 0: the column has at most
// ZU planes and the loop is unrolled, so the coefficient indices of all its
// planes are loaded first, ahead of everything else -- a plane's gather then
// never waits for its own index.  ZU = 0: any number of planes, each plane
// loading its index first.
template <int CM, int ZU, bool NT = false, typename OT = unsigned>
__device__ __forceinline__ void e_column(const FusedArgs &A, const Geo &g, const PassSrc &se,
                                         int k, int j, int z0, int z1, unsigned long long smask)
{
    const Layout L = make_layout(g);
    const OT cs = cstride<OT>(A);
    const OT oz = (OT)g.nxy * 8u;
    const double sgz = g.bc[2] == 'E' ? -1.0 : (g.bc[2] == 'H' ? 1.0 : 0.0);
    const bool bz0 = g.bc[2] == '0';
    OT o = ((OT)z0 * (OT)g.Ny + (OT)j) * (OT)g.Nx * 8u + (OT)k * 8u;
    if (ZU > 0) {
        unsigned ci[ZU > 0 ? ZU : 1];
#pragma unroll
        for (int p = 0; p < ZU; ++p)   // a missing plane re-reads the last one
            ci[p] = CM == CM_INDEX ? coef_index<OT>(A, o + (OT)min(p, z1 - z0 - 1) * oz) : 0u;
        double hxm, hym;   // H at i-1
        if (z0 > 0) { hxm = ldo(A.Ho, o - oz); hym = ldo(A.Ho, o - oz + cs); }
        else { hxm = bz0 ? 0.0 : sgz * ldo(A.Ho, o); hym = bz0 ? 0.0 : sgz * ldo(A.Ho, o + cs); }
#pragma unroll
        for (int p = 0; p < ZU; ++p) {
            if (z0 + p >= z1) break;
            e_plane<CM, NT, OT>(A, g, se, L, k, j, z0 + p, o + (OT)p * oz, ci[p], smask, hxm,
                                hym);
        }
    } else {
        double hxm, hym;
        if (z0 > 0) { hxm = ldo(A.Ho, o - oz); hym = ldo(A.Ho, o - oz + cs); }
        else { hxm = bz0 ? 0.0 : sgz * ldo(A.Ho, o); hym = bz0 ? 0.0 : sgz * ldo(A.Ho, o + cs); }
        for (int i = z0; i < z1; ++i, o += oz) {
            const unsigned ci = CM == CM_INDEX ? coef_index<OT>(A, o) : 0u;
            e_plane<CM, NT, OT>(A, g, se, L, k, j, i, o, ci, smask, hxm, hym);
        }
    }
}

// ZU: see e_column (the launch picks the smallest ZU of 2, 4, 8 that holds
// zc).  BIG: as for k_pass_h.
template <int CM, int TY, int OCC = 4, int ZU = 0, bool NT = false, bool BIG = false>
__global__ __launch_bounds__(PX * TY, OCC) void k_pass_e(FusedArgs A, Geo g, PassSrc se,
                                                        const int *__restrict__ done)
{
    if (done && *done) return;
    int k, j, z0, z1;
    pass_tile(g, TY, k, j, z0, z1, true);
    j = __builtin_amdgcn_readfirstlane(j);   // one row per wave
    if (k >= g.Nx || j >= g.Ny) return;
    const unsigned long long smask = col_src_mask(se, k, j, z0, z1);
    using OT = typename std::conditional<BIG, unsigned long long, unsigned>::type;
    e_column<CM, ZU, NT, OT>(A, g, se, k, j, z0, z1, smask);
}

}  // namespace emf
