/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded FP64 restatement of EMopt's CW-FDTD kernel
 * (src/fdtd.cpp, src/fdtd.hpp).  It exports the same `FDTD_*` C ABI as the
 * reference's FDTD.so (src/fdtd.hpp:571-643) so that the oracle driver
 * (oracle/driver.py) can run either this file or the reference compiled from
 * its own sources (oracle/_ref/) and diff them.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline; the product
 * path (emopt_amd) never calls it.
 *
 * Semantics restated, with the reference lines they follow:
 *   update_H            src/fdtd.cpp:121-367   (max-face PEC ghosts, forward-difference
 *                                               curl E, CPML slabs, M sources)
 *   update_E            src/fdtd.cpp:369-755   (min-face '0'/'E'/'H' ghosts, backward-
 *                                               difference curl H, real/complex eps,
 *                                               CPML slabs, J sources)
 *   CPML tables         src/fdtd.cpp:950-1186  (kappa, b, c per layer; zmax strict '>')
 *   slab alloc/reset    src/fdtd.cpp:779-948
 *   snapshots           src/fdtd.cpp:1205-1251 (writes .real only)
 *   phasor 2T / 3T      src/fdtd.cpp:1253-1493
 *   sources             src/fdtd.cpp:1498-1602
 *   ghost pack/unpack   src/fdtd.cpp:1789-1929
 *
 * Arithmetic is written in the same operation order as the reference so the
 * two agree to rounding; the loop structure and data organisation are this
 * file's own (one table struct per axis, one slab struct per face).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double re, im; } c128;   /* numpy complex128 memory layout */

/* 1-D CPML coefficient tables of one axis: [w_min + w_max] entries each. */
typedef struct {
    double *kH, *bH, *cH;   /* used by update_H (derivatives of E) */
    double *kE, *bE, *cE;   /* used by update_E (derivatives of H) */
} pml_axis;

/* Auxiliary (psi) arrays on one face.  Index [0] belongs to the update_H pass
 * (psi of E derivatives), [1] to update_E (psi of H derivatives); each pass
 * keeps two arrays, one per transverse component. */
typedef struct {
    double *psi[2][2];
    long n;
} pml_face;

typedef struct {
    int Nx, Ny, Nz;
    int I, J, K, i0, j0, k0;
    double X, Y, Z, dx, dy, dz;
    double wavelength, R, dt, odt;
    double src_T, src_min, src_k;
    double *E[3], *H[3];
    bool complex_eps;
    c128 *eps[3], *mu[3];
    c128 *t0[6], *t1[6];
    int w[6];                /* xmin xmax ymin ymax zmin zmax */
    double sigma, alpha, kappa, pw;
    char bc[3];
    pml_axis tab[3];         /* x, y, z */
    pml_face face[6];        /* xmin xmax ymin ymax zmin zmax */
    struct src_block { c128 *f[6]; int i0, j0, k0, I, J, K; } *src;
    int nsrc, src_cap;
} cfdtd;

/* ---------------------------------------------------------------- helpers */

static inline long pidx(const cfdtd *s, int i, int j, int k)
{   /* padded index of cell (i,j,k) with one ghost layer on every face */
    return ((long)(i + 1) * (s->J + 2) + (j + 1)) * (s->K + 2) + (k + 1);
}

static inline long gidx(const cfdtd *s, int i, int j, int k)
{
    return ((long)i * s->J + j) * s->K + k;
}

static double ramp(const cfdtd *s, double d) { return pow(d, s->pw); }

/* kappa, b, c of one CPML layer at ramp factor f (src/fdtd.cpp:992-997). */
static void pml_coeffs(const cfdtd *s, double f, double *kap, double *b, double *c)
{
    double sig = s->sigma * f;
    double alp = s->alpha * (1 - f);
    double k = (s->kappa - 1.0) * f + 1.0;
    double bb = exp(-s->dt * (sig / k + alp));
    *kap = k;
    *b = bb;
    *c = (bb == 1) ? 0.0 : (bb - 1) * sig / (sig * k + k * k * alp);
}

static void free_tables(cfdtd *s)
{
    for (int a = 0; a < 3; a++) {
        pml_axis *t = &s->tab[a];
        free(t->kH); free(t->bH); free(t->cH);
        free(t->kE); free(t->bE); free(t->cE);
        memset(t, 0, sizeof(*t));
    }
}

/* Distances from the onset of the PML for layer index n of a face.
 * x and y faces: H at (n -/+ 0.5)/w (min clamps at 0), E at n/w.
 * z faces:       H at n/w,  E at (n + 0.5)/w on zmin and (n - 0.5)/w on zmax.
 * (src/fdtd.cpp:986-1180)  Returns 1 when the ramp must be clamped at 0. */
static void layer_dist(int axis, int side, int n, int w, double *dH, double *dE,
                       int *clampH, int *clampE)
{
    *clampH = *clampE = 0;
    if (axis < 2) {
        *dH = side == 0 ? (n - 0.5) / w : (n + 0.5) / w;
        *dE = (double)n / w;
        *clampH = (side == 0);
    } else {
        *dH = (double)n / w;
        *dE = side == 0 ? (n + 0.5) / w : (n - 0.5) / w;
        *clampE = (side == 1);
    }
}

static void compute_tables(cfdtd *s)
{
    free_tables(s);
    for (int a = 0; a < 3; a++) {
        int w0 = s->w[2 * a], w1 = s->w[2 * a + 1], n = w0 + w1;
        pml_axis *t = &s->tab[a];
        size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
        t->kH = malloc(bytes); t->bH = malloc(bytes); t->cH = malloc(bytes);
        t->kE = malloc(bytes); t->bE = malloc(bytes); t->cE = malloc(bytes);
        for (int side = 0; side < 2; side++) {
            int w = side == 0 ? w0 : w1;
            for (int l = 0; l < w; l++) {
                double dH, dE; int cH, cE;
                layer_dist(a, side, l, w, &dH, &dE, &cH, &cE);
                double fH = ramp(s, dH), fE = ramp(s, dE);
                if (cH && fH < 0) fH = 0;
                if (cE && fE < 0) fE = 0;
                int slot = side == 0 ? l : w0 + l;
                pml_coeffs(s, fH, &t->kH[slot], &t->bH[slot], &t->cH[slot]);
                pml_coeffs(s, fE, &t->kE[slot], &t->bE[slot], &t->cE[slot]);
            }
        }
    }
}

/* Number of slab cells this chunk holds on each face (src/fdtd.cpp:779-876). */
static long face_cells(const cfdtd *s, int f)
{
    int lo_x = s->w[0], hi_x = s->Nx - s->w[1];
    int lo_y = s->w[2], hi_y = s->Ny - s->w[3];
    int lo_z = s->w[4], hi_z = s->Nz - s->w[5];
    switch (f) {
    case 0: return s->k0 < lo_x ? (long)s->I * s->J * (lo_x - s->k0) : 0;
    case 1: return s->k0 + s->K > hi_x ? (long)s->I * s->J * (s->k0 + s->K - hi_x) : 0;
    case 2: return s->j0 < lo_y ? (long)s->I * s->K * (lo_y - s->j0) : 0;
    case 3: return s->j0 + s->J > hi_y ? (long)s->I * s->K * (s->j0 + s->J - hi_y) : 0;
    case 4: return s->i0 < lo_z ? (long)s->J * s->K * (lo_z - s->i0) : 0;
    default: return s->i0 + s->I > hi_z ? (long)s->J * s->K * (s->i0 + s->I - hi_z) : 0;
    }
}

static void free_faces(cfdtd *s)
{
    for (int f = 0; f < 6; f++) {
        for (int p = 0; p < 2; p++)
            for (int c = 0; c < 2; c++) { free(s->face[f].psi[p][c]); s->face[f].psi[p][c] = NULL; }
        s->face[f].n = 0;
    }
}

/* ------------------------------------------------------------- curl passes
 *
 * A CPML slab contribution along axis `a` for one pass: psi <- c*d + b*psi,
 * then F <- F +/- m*(psi - d + d/kappa).  The caller supplies the sign and the
 * multiplier m (dt/mu for H, b_eps for E) in the reference's order.
 */
static inline double psi_step(double *psi, double c, double b, double d)
{
    *psi = c * d + b * (*psi);
    return *psi;
}

/* Which slab (if any) a cell sits in along one axis and its psi/param index. */
typedef struct { int side; long ip; int par; } slab_hit;

static inline slab_hit hit_x(const cfdtd *s, int i, int j, int k)
{
    slab_hit h = { -1, 0, 0 };
    int lo = s->w[0], hi = s->Nx - s->w[1], kg = k + s->k0;
    if (kg < lo) {
        int wl = lo - s->k0;
        h.side = 0; h.ip = ((long)i * s->J + j) * wl + k; h.par = lo - k - s->k0 - 1;
    } else if (kg >= hi) {
        int wl = s->k0 + s->K - hi;
        h.side = 1; h.ip = ((long)i * s->J + j) * wl + (kg - hi); h.par = kg - hi + s->w[0];
    }
    return h;
}

static inline slab_hit hit_y(const cfdtd *s, int i, int j, int k)
{
    slab_hit h = { -1, 0, 0 };
    int lo = s->w[2], hi = s->Ny - s->w[3], jg = j + s->j0;
    if (jg < lo) {
        int wl = lo - s->j0;
        h.side = 0; h.ip = ((long)i * wl + j) * s->K + k; h.par = lo - j - s->j0 - 1;
    } else if (jg >= hi) {
        int wl = s->j0 + s->J - hi;
        h.side = 1; h.ip = ((long)i * wl + (jg - hi)) * s->K + k; h.par = jg - hi + s->w[2];
    }
    return h;
}

static inline slab_hit hit_z(const cfdtd *s, int i, int j, int k)
{
    slab_hit h = { -1, 0, 0 };
    int lo = s->w[4], hi = s->Nz - s->w[5], ig = i + s->i0;
    if (ig < lo) {
        h.side = 0; h.ip = ((long)i * s->J + j) * s->K + k; h.par = lo - i - s->i0 - 1;
    } else if (ig > hi) {            /* strict: the first zmax layer is skipped */
        h.side = 1; h.ip = ((long)(ig - hi) * s->J + j) * s->K + k; h.par = ig - hi + s->w[4];
    }
    return h;
}

static inline double src_env(const cfdtd *s, double t, double phase)
{
    if (t <= s->src_T)
        return sin(t + phase) * ((1 + s->src_min) * exp(-(t - s->src_T) * (t - s->src_T) / s->src_k) - s->src_min);
    return sin(t + phase);
}

static void h_pass(cfdtd *s, int n, double t)
{
    (void)n;
    double *Ex = s->E[0], *Ey = s->E[1], *Ez = s->E[2];
    double *Hx = s->H[0], *Hy = s->H[1], *Hz = s->H[2];
    const double rx = s->R / s->dx, ry = s->R / s->dy, rz = s->R / s->dz;
    const int I = s->I, J = s->J, K = s->K;
    const long sj = K + 2, si = (long)(J + 2) * (K + 2);

    /* PEC ghosts on the max faces (the only faces this pass reads past). */
    if (s->bc[0] != 'P' && s->k0 + K == s->Nx)
        for (int i = 0; i < I; i++) for (int j = 0; j < J; j++) {
            long p = pidx(s, i, j, K);
            Ey[p] = 0.0; Ez[p] = 0.0;
        }
    if (s->bc[1] != 'P' && s->j0 + J == s->Ny)
        for (int i = 0; i < I; i++) for (int k = 0; k < K; k++) {
            long p = pidx(s, i, J, k);
            Ex[p] = 0.0; Ez[p] = 0.0;
        }
    if (s->bc[2] != 'P' && s->i0 + I == s->Nz)
        for (int j = 0; j < J; j++) for (int k = 0; k < K; k++) {
            long p = pidx(s, I, j, k);
            Ex[p] = 0.0; Ey[p] = 0.0;
        }

    /* Cells are independent within a pass (each writes only its own H and
     * psi entries), so the z loop may run on several threads (built with
     * -fopenmp); the arithmetic per cell is unchanged. */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < I; i++) for (int j = 0; j < J; j++) for (int k = 0; k < K; k++) {
        long p = pidx(s, i, j, k), g = gidx(s, i, j, k);
        double mx = s->dt / s->mu[0][g].re, my = s->dt / s->mu[1][g].re, mz = s->dt / s->mu[2][g].re;

        double dEzdy = ry * (Ez[p + sj] - Ez[p]);
        double dEydz = rz * (Ey[p + si] - Ey[p]);
        Hx[p] = Hx[p] + mx * (dEydz - dEzdy);
        double dExdz = rz * (Ex[p + si] - Ex[p]);
        double dEzdx = rx * (Ez[p + 1] - Ez[p]);
        Hy[p] = Hy[p] + my * (dEzdx - dExdz);
        double dEydx = rx * (Ey[p + 1] - Ey[p]);
        double dExdy = ry * (Ex[p + sj] - Ex[p]);
        Hz[p] = Hz[p] + mz * (dExdy - dEydx);

        slab_hit h = hit_x(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[h.side];
            const pml_axis *T = &s->tab[0];
            double kap = T->kH[h.par], b = T->bH[h.par], c = T->cH[h.par];
            double p0 = psi_step(&f->psi[0][0][h.ip], c, b, dEydx);
            double p1 = psi_step(&f->psi[0][1][h.ip], c, b, dEzdx);
            Hz[p] = Hz[p] - mz * (p0 - dEydx + dEydx / kap);
            Hy[p] = Hy[p] + my * (p1 - dEzdx + dEzdx / kap);
        }
        h = hit_y(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[2 + h.side];
            const pml_axis *T = &s->tab[1];
            double kap = T->kH[h.par], b = T->bH[h.par], c = T->cH[h.par];
            double p0 = psi_step(&f->psi[0][0][h.ip], c, b, dExdy);
            double p1 = psi_step(&f->psi[0][1][h.ip], c, b, dEzdy);
            Hz[p] = Hz[p] + mz * (p0 - dExdy + dExdy / kap);
            Hx[p] = Hx[p] - mx * (p1 - dEzdy + dEzdy / kap);
        }
        h = hit_z(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[4 + h.side];
            const pml_axis *T = &s->tab[2];
            double kap = T->kH[h.par], b = T->bH[h.par], c = T->cH[h.par];
            double p0 = psi_step(&f->psi[0][0][h.ip], c, b, dExdz);
            double p1 = psi_step(&f->psi[0][1][h.ip], c, b, dEydz);
            Hx[p] = Hx[p] + mx * (p1 - dEydz + dEydz / kap);
            Hy[p] = Hy[p] - my * (p0 - dExdz + dExdz / kap);
        }
    }

    /* magnetic current sources: Mx, My, Mz in that order, per block */
    for (int q = 0; q < s->nsrc; q++) {
        const struct src_block *b = &s->src[q];
        for (int comp = 0; comp < 3; comp++) {
            const c128 *M = b->f[3 + comp];
            double *F = s->H[comp];
            const c128 *mu = s->mu[comp];
            for (int i = 0; i < b->I; i++) for (int j = 0; j < b->J; j++) for (int k = 0; k < b->K; k++) {
                long p = pidx(s, i + b->i0, j + b->j0, k + b->k0);
                long g = gidx(s, i + b->i0, j + b->j0, k + b->k0);
                long m = ((long)i * b->J + j) * b->K + k;
                double st = src_env(s, t, M[m].im);
                F[p] = F[p] + st * M[m].re * s->dt / mu[g].re;
            }
        }
    }
}

static void e_pass(cfdtd *s, int n, double t)
{
    (void)n;
    double *Ex = s->E[0], *Ey = s->E[1], *Ez = s->E[2];
    double *Hx = s->H[0], *Hy = s->H[1], *Hz = s->H[2];
    const double rx = s->R / s->dx, ry = s->R / s->dy, rz = s->R / s->dz;
    const int I = s->I, J = s->J, K = s->K;
    const long sj = K + 2, si = (long)(J + 2) * (K + 2);

    /* Min-face H ghosts: '0' -> 0, 'E' -> -H(first), 'H' -> +H(first). */
    if (s->k0 == 0)
        for (int i = 0; i < I; i++) for (int j = 0; j < J; j++) {
            long p = pidx(s, i, j, -1);
            if (s->bc[0] == '0') { Hy[p] = 0.0; Hz[p] = 0.0; }
            else if (s->bc[0] == 'E') { Hy[p] = -1 * Hy[p + 1]; Hz[p] = -1 * Hz[p + 1]; }
            else if (s->bc[0] == 'H') { Hy[p] = Hy[p + 1]; Hz[p] = Hz[p + 1]; }
        }
    if (s->j0 == 0)
        for (int i = 0; i < I; i++) for (int k = 0; k < K; k++) {
            long p = pidx(s, i, -1, k);
            if (s->bc[1] == '0') { Hx[p] = 0.0; Hz[p] = 0.0; }
            else if (s->bc[1] == 'E') { Hx[p] = -1 * Hx[p + sj]; Hz[p] = -1 * Hz[p + sj]; }
            else if (s->bc[1] == 'H') { Hx[p] = Hx[p + sj]; Hz[p] = Hz[p + sj]; }
        }
    if (s->i0 == 0)
        for (int j = 0; j < J; j++) for (int k = 0; k < K; k++) {
            long p = pidx(s, -1, j, k);
            if (s->bc[2] == '0') { Hx[p] = 0.0; Hy[p] = 0.0; }
            else if (s->bc[2] == 'E') { Hx[p] = -1 * Hx[p + si]; Hy[p] = -1 * Hy[p + si]; }
            else if (s->bc[2] == 'H') { Hx[p] = Hx[p + si]; Hy[p] = Hy[p + si]; }
        }

#pragma omp parallel for schedule(static)
    for (int i = 0; i < I; i++) for (int j = 0; j < J; j++) for (int k = 0; k < K; k++) {
        long p = pidx(s, i, j, k), g = gidx(s, i, j, k);
        double ax, ay, az, bx, by, bz;
        if (!s->complex_eps) {
            ax = ay = az = 1.0;
            bx = s->dt / s->eps[0][g].re; by = s->dt / s->eps[1][g].re; bz = s->dt / s->eps[2][g].re;
        } else {
            double a[3], b[3];
            for (int c = 0; c < 3; c++) {
                double er = s->eps[c][g].re * s->odt, ei = s->eps[c][g].im * 0.5;
                a[c] = (er - ei) / (er + ei);
                b[c] = 1.0 / (er + ei);
            }
            ax = a[0]; ay = a[1]; az = a[2]; bx = b[0]; by = b[1]; bz = b[2];
        }

        double dHzdy = ry * (Hz[p] - Hz[p - sj]);
        double dHydz = rz * (Hy[p] - Hy[p - si]);
        Ex[p] = ax * Ex[p] + (dHzdy - dHydz) * bx;
        double dHxdz = rz * (Hx[p] - Hx[p - si]);
        double dHzdx = rx * (Hz[p] - Hz[p - 1]);
        Ey[p] = ay * Ey[p] + (dHxdz - dHzdx) * by;
        double dHydx = rx * (Hy[p] - Hy[p - 1]);
        double dHxdy = ry * (Hx[p] - Hx[p - sj]);
        Ez[p] = az * Ez[p] + (dHydx - dHxdy) * bz;

        slab_hit h = hit_x(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[h.side];
            const pml_axis *T = &s->tab[0];
            double kap = T->kE[h.par], b = T->bE[h.par], c = T->cE[h.par];
            double p0 = psi_step(&f->psi[1][0][h.ip], c, b, dHydx);
            double p1 = psi_step(&f->psi[1][1][h.ip], c, b, dHzdx);
            Ez[p] = Ez[p] + (p0 - dHydx + dHydx / kap) * bz;
            Ey[p] = Ey[p] - (p1 - dHzdx + dHzdx / kap) * by;
        }
        h = hit_y(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[2 + h.side];
            const pml_axis *T = &s->tab[1];
            double kap = T->kE[h.par], b = T->bE[h.par], c = T->cE[h.par];
            double p0 = psi_step(&f->psi[1][0][h.ip], c, b, dHxdy);
            double p1 = psi_step(&f->psi[1][1][h.ip], c, b, dHzdy);
            Ez[p] = Ez[p] - (p0 - dHxdy + dHxdy / kap) * bz;
            Ex[p] = Ex[p] + (p1 - dHzdy + dHzdy / kap) * bx;
        }
        h = hit_z(s, i, j, k);
        if (h.side >= 0) {
            pml_face *f = &s->face[4 + h.side];
            const pml_axis *T = &s->tab[2];
            double kap = T->kE[h.par], b = T->bE[h.par], c = T->cE[h.par];
            double p0 = psi_step(&f->psi[1][0][h.ip], c, b, dHxdz);
            double p1 = psi_step(&f->psi[1][1][h.ip], c, b, dHydz);
            Ex[p] = Ex[p] - (p1 - dHydz + dHydz / kap) * bx;
            Ey[p] = Ey[p] + (p0 - dHxdz + dHxdz / kap) * by;
        }
    }

    /* electric current sources: Jx, Jy, Jz in that order, per block */
    for (int q = 0; q < s->nsrc; q++) {
        const struct src_block *b = &s->src[q];
        for (int comp = 0; comp < 3; comp++) {
            const c128 *Jc = b->f[comp];
            double *F = s->E[comp];
            const c128 *eps = s->eps[comp];
            for (int i = 0; i < b->I; i++) for (int j = 0; j < b->J; j++) for (int k = 0; k < b->K; k++) {
                long p = pidx(s, i + b->i0, j + b->j0, k + b->k0);
                long g = gidx(s, i + b->i0, j + b->j0, k + b->k0);
                long m = ((long)i * b->J + j) * b->K + k;
                double be = !s->complex_eps ? s->dt / eps[g].re
                                            : 1.0 / (eps[g].re * s->odt + eps[g].im * 0.5);
                double st = src_env(s, t, Jc[m].im);
                F[p] = F[p] - st * Jc[m].re * be;
            }
        }
    }
}

/* ------------------------------------------------------ phasor closed forms */

double FDTD_calc_phase_2T(double t0, double t1, double f0, double f1)
{
    if (f0 == 0.0 && f1 == 0) return 0.0;
    return atan((f1 * sin(t0) - f0 * sin(t1)) / (f0 * cos(t1) - f1 * cos(t0)));
}

double FDTD_calc_amplitude_2T(double t0, double t1, double f0, double f1, double phase)
{
    if (f0 * f0 > f1 * f1) return f1 / (sin(t1) * cos(phase) + cos(t1) * sin(phase));
    return f0 / (sin(t0) * cos(phase) + cos(t0) * sin(phase));
}

double FDTD_calc_phase_3T(double t0, double t1, double t2, double f0, double f1, double f2)
{
    double f10 = f1 - f0, f21 = f2 - f1;
    if (f10 == 0 && f21 == 0) return 0.0;
    return atan2(f10 * (sin(t2) - sin(t1)) - f21 * (sin(t1) - sin(t0)),
                 f21 * (cos(t1) - cos(t0)) - f10 * (cos(t2) - cos(t1)));
}

double FDTD_calc_amplitude_3T(double t0, double t1, double t2, double f0, double f1, double f2,
                              double phase)
{
    double f21 = f2 - f1, f10 = f1 - f0;
    if (f21 == 0 && f10 == 0) return 0.0;
    if (f21 * f21 >= f10 * f10)
        return f21 / (cos(phase) * (sin(t2) - sin(t1)) + sin(phase) * (cos(t2) - cos(t1)));
    return f10 / (cos(phase) * (sin(t1) - sin(t0)) + sin(phase) * (cos(t1) - cos(t0)));
}

/* ------------------------------------------------------------------ C ABI */

void *FDTD_new(void)
{
    cfdtd *s = calloc(1, sizeof(cfdtd));
    return s;
}

void FDTD_delete(void *h)
{
    cfdtd *s = h;
    if (!s) return;
    free_tables(s);
    free_faces(s);
    free(s->src);
    free(s);
}

void FDTD_set_wavelength(void *h, double wl)
{
    cfdtd *s = h; s->wavelength = wl; s->R = wl / (2 * M_PI);
}

void FDTD_set_physical_dims(void *h, double X, double Y, double Z, double dx, double dy, double dz)
{
    cfdtd *s = h; s->X = X; s->Y = Y; s->Z = Z; s->dx = dx; s->dy = dy; s->dz = dz;
}

void FDTD_set_grid_dims(void *h, int Nx, int Ny, int Nz)
{
    cfdtd *s = h; s->Nx = Nx; s->Ny = Ny; s->Nz = Nz;
}

void FDTD_set_local_grid(void *h, int k0, int j0, int i0, int K, int J, int I)
{
    cfdtd *s = h; s->k0 = k0; s->j0 = j0; s->i0 = i0; s->K = K; s->J = J; s->I = I;
}

void FDTD_set_dt(void *h, double dt)
{
    cfdtd *s = h; s->dt = dt; s->odt = 1.0 / dt;
}

void FDTD_set_complex_eps(void *h, bool c) { ((cfdtd *)h)->complex_eps = c; }

void FDTD_set_field_arrays(void *h, double *Ex, double *Ey, double *Ez, double *Hx, double *Hy, double *Hz)
{
    cfdtd *s = h;
    s->E[0] = Ex; s->E[1] = Ey; s->E[2] = Ez; s->H[0] = Hx; s->H[1] = Hy; s->H[2] = Hz;
}

void FDTD_set_mat_arrays(void *h, c128 *ex, c128 *ey, c128 *ez, c128 *mx, c128 *my, c128 *mz)
{
    cfdtd *s = h;
    s->eps[0] = ex; s->eps[1] = ey; s->eps[2] = ez; s->mu[0] = mx; s->mu[1] = my; s->mu[2] = mz;
}

void FDTD_update_H(void *h, int n, double t) { h_pass(h, n, t); }
void FDTD_update_E(void *h, int n, double t) { e_pass(h, n, t); }

void FDTD_set_pml_widths(void *h, int xmin, int xmax, int ymin, int ymax, int zmin, int zmax)
{
    cfdtd *s = h;
    s->w[0] = xmin; s->w[1] = xmax; s->w[2] = ymin; s->w[3] = ymax; s->w[4] = zmin; s->w[5] = zmax;
}

void FDTD_set_pml_properties(void *h, double sigma, double alpha, double kappa, double pw)
{
    cfdtd *s = h;
    s->sigma = sigma; s->alpha = alpha; s->kappa = kappa; s->pw = pw;
    compute_tables(s);
}

void FDTD_build_pml(void *h)
{
    cfdtd *s = h;
    free_faces(s);
    for (int f = 0; f < 6; f++) {
        long n = face_cells(s, f);
        s->face[f].n = n;
        if (n > 0)
            for (int p = 0; p < 2; p++) for (int c = 0; c < 2; c++)
                s->face[f].psi[p][c] = calloc((size_t)n, sizeof(double));
    }
    compute_tables(s);
}

void FDTD_reset_pml(void *h)
{
    cfdtd *s = h;
    for (int f = 0; f < 6; f++)
        for (int p = 0; p < 2; p++) for (int c = 0; c < 2; c++)
            if (s->face[f].psi[p][c]) memset(s->face[f].psi[p][c], 0, (size_t)s->face[f].n * sizeof(double));
}

void FDTD_set_t0_arrays(void *h, c128 *a, c128 *b, c128 *c, c128 *d, c128 *e, c128 *f)
{
    cfdtd *s = h; c128 *v[6] = { a, b, c, d, e, f }; memcpy(s->t0, v, sizeof(v));
}

void FDTD_set_t1_arrays(void *h, c128 *a, c128 *b, c128 *c, c128 *d, c128 *e, c128 *f)
{
    cfdtd *s = h; c128 *v[6] = { a, b, c, d, e, f }; memcpy(s->t1, v, sizeof(v));
}

static void capture(cfdtd *s, c128 **dst)
{
    for (int i = 0; i < s->I; i++) for (int j = 0; j < s->J; j++) for (int k = 0; k < s->K; k++) {
        long p = pidx(s, i, j, k), g = gidx(s, i, j, k);
        for (int c = 0; c < 3; c++) { dst[c][g].re = s->E[c][p]; dst[3 + c][g].re = s->H[c][p]; }
    }
}

void FDTD_capture_t0_fields(void *h) { cfdtd *s = h; capture(s, s->t0); }
void FDTD_capture_t1_fields(void *h) { cfdtd *s = h; capture(s, s->t1); }

/* A*cos(phi) - i*A*sin(phi) after folding a negative amplitude into phi. */
static inline void store_phasor(c128 *out, double A, double phi)
{
    if (A < 0) { A *= -1; phi += M_PI; }
    out->re = A * cos(phi);
    out->im = -A * sin(phi);
}

void FDTD_calc_complex_fields_2T(void *h, double t0, double t1)
{
    cfdtd *s = h;
    double tH0 = t0 - 0.5 * s->dt, tH1 = t1 - 0.5 * s->dt;
    for (int i = 0; i < s->I; i++) for (int j = 0; j < s->J; j++) for (int k = 0; k < s->K; k++) {
        long p = pidx(s, i, j, k), g = gidx(s, i, j, k);
        for (int c = 0; c < 6; c++) {
            double a = c < 3 ? t0 : tH0, b = c < 3 ? t1 : tH1;
            double f0 = s->t0[c][g].re, f1 = c < 3 ? s->E[c][p] : s->H[c - 3][p];
            double phi = FDTD_calc_phase_2T(a, b, f0, f1);
            double A = FDTD_calc_amplitude_2T(a, b, f0, f1, phi);
            store_phasor(&s->t0[c][g], A, phi);
        }
    }
}

void FDTD_calc_complex_fields_3T(void *h, double t0, double t1, double t2)
{
    cfdtd *s = h;
    double tH0 = t0 - 0.5 * s->dt, tH1 = t1 - 0.5 * s->dt, tH2 = t2 - 0.5 * s->dt;
    for (int i = 0; i < s->I; i++) for (int j = 0; j < s->J; j++) for (int k = 0; k < s->K; k++) {
        long p = pidx(s, i, j, k), g = gidx(s, i, j, k);
        for (int c = 0; c < 6; c++) {
            double a = c < 3 ? t0 : tH0, b = c < 3 ? t1 : tH1, d = c < 3 ? t2 : tH2;
            double f0 = s->t0[c][g].re, f1 = s->t1[c][g].re;
            double f2 = c < 3 ? s->E[c][p] : s->H[c - 3][p];
            double phi = FDTD_calc_phase_3T(a, b, d, f0, f1, f2);
            double A = FDTD_calc_amplitude_3T(a, b, d, f0, f1, f2, phi);
            store_phasor(&s->t0[c][g], A, phi);
        }
    }
}

void FDTD_add_source(void *h, c128 *Jx, c128 *Jy, c128 *Jz, c128 *Mx, c128 *My, c128 *Mz,
                     int i0, int j0, int k0, int I, int J, int K, bool calc_phase)
{
    cfdtd *s = h;
    c128 *f[6] = { Jx, Jy, Jz, Mx, My, Mz };
    if (calc_phase) {
        long n = (long)I * J * K;
        for (long m = 0; m < n; m++)
            for (int c = 0; c < 6; c++) {
                double re = f[c][m].re, im = f[c][m].im;
                f[c][m].re = sqrt(re * re + im * im);
                f[c][m].im = (im == 0 && re == 0) ? 0.0 : -1 * atan2(im, re);
            }
    }
    if (s->nsrc == s->src_cap) {
        s->src_cap = s->src_cap ? 2 * s->src_cap : 4;
        s->src = realloc(s->src, (size_t)s->src_cap * sizeof(*s->src));
    }
    struct src_block *b = &s->src[s->nsrc++];
    memcpy(b->f, f, sizeof(f));
    b->i0 = i0; b->j0 = j0; b->k0 = k0; b->I = I; b->J = J; b->K = K;
}

void FDTD_clear_sources(void *h) { ((cfdtd *)h)->nsrc = 0; }

void FDTD_set_source_properties(void *h, double src_T, double src_min)
{
    cfdtd *s = h;
    s->src_T = src_T; s->src_min = src_min;
    s->src_k = src_T * src_T / log((1 + src_min) / src_min);
}

double FDTD_src_func_t(void *h, int n, double t, double phase)
{
    (void)n;
    return src_env(h, t, phase);
}

void FDTD_set_bc(void *h, char *bc) { cfdtd *s = h; s->bc[0] = bc[0]; s->bc[1] = bc[1]; s->bc[2] = bc[2]; }

/* Ghost pack/unpack: 6 face blocks xmin,xmax,ymin,ymax,zmin,zmax of the
 * padded array's outermost interior layers into `ghost[0:nb]`, and back from
 * `ghost[nb:2nb]` into the ghost layers (src/fdtd.cpp:1789-1929). */
void FDTD_copy_to_ghost_comm(double *src, c128 *ghost, int I, int J, int K)
{
    long sj = K + 2, si = (long)(J + 2) * (K + 2), o = 0;
    for (int side = 0; side < 2; side++, o += (long)I * J)
        for (int i = 0; i < I; i++) for (int j = 0; j < J; j++)
            ghost[o + j + (long)i * J].re = src[(i + 1) * si + (j + 1) * sj + (side ? K : 1)];
    for (int side = 0; side < 2; side++, o += (long)I * K)
        for (int i = 0; i < I; i++) for (int k = 0; k < K; k++)
            ghost[o + k + (long)i * K].re = src[(i + 1) * si + (side ? J : 1) * sj + k + 1];
    for (int side = 0; side < 2; side++, o += (long)J * K)
        for (int j = 0; j < J; j++) for (int k = 0; k < K; k++)
            ghost[o + k + (long)j * K].re = src[(side ? I : 1) * si + (j + 1) * sj + k + 1];
}

void FDTD_copy_from_ghost_comm(double *dst, c128 *ghost, int I, int J, int K)
{
    long sj = K + 2, si = (long)(J + 2) * (K + 2);
    long o = 2L * I * J + 2L * I * K + 2L * J * K;
    for (int side = 0; side < 2; side++, o += (long)I * J)
        for (int i = 0; i < I; i++) for (int j = 0; j < J; j++)
            dst[(i + 1) * si + (j + 1) * sj + (side ? K + 1 : 0)] = ghost[o + j + (long)i * J].re;
    for (int side = 0; side < 2; side++, o += (long)I * K)
        for (int i = 0; i < I; i++) for (int k = 0; k < K; k++)
            dst[(i + 1) * si + (side ? J + 1 : 0) * sj + k + 1] = ghost[o + k + (long)i * K].re;
    for (int side = 0; side < 2; side++, o += (long)J * K)
        for (int j = 0; j < J; j++) for (int k = 0; k < K; k++)
            dst[(side ? I + 1 : 0) * si + (j + 1) * sj + k + 1] = ghost[o + k + (long)j * K].re;
}
