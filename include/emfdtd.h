/*
 * emfdtd — C ABI of the MI355X (gfx950) CW-FDTD library `libemfdtd.so`.
 *
 * This header is the drop-in boundary that replaces EMopt's FDTD.so ABI
 * (src/fdtd.hpp:571-643, bound by emopt/fdtd_ctypes.py:26-149).  Every entry
 * point below names the reference function it replaces.  Differences that
 * follow from device residency:
 *
 *  - Arrays are COPIED into device-owned buffers instead of being borrowed by
 *    pointer (the reference keeps numpy/PETSc pointers, src/fdtd.cpp:107-119).
 *    Field state never leaves HBM during a solve.
 *  - Every call returns an int status (EMF_OK = 0, negative on error) instead
 *    of undefined behaviour on bad input; emf_last_error() gives the message.
 *  - Complex arrays are numpy complex128 ({re, im} interleaved doubles),
 *    C-contiguous, shape [Nz][Ny][Nx] (x fastest), exactly the reference's
 *    unpadded "global" layout (src/fdtd.cpp:183).
 *  - The grid is never padded with ghost layers: boundary conditions are
 *    applied inside the kernels.
 *  - One handle = one solver on one GPU; calls on one handle are serialised
 *    by a per-handle lock.  Different handles may be used from different
 *    threads.
 *
 * Component order everywhere: 0..5 = Ex, Ey, Ez, Hx, Hy, Hz (and Jx..Mz).
 * Result slots: 0 = forward phasors, 1 = adjoint phasors.
 */
#ifndef EMFDTD_H
#define EMFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Symbols are exported explicitly; the library builds with -fvisibility=hidden. */
#if defined(__GNUC__)
#define EMF_API __attribute__((visibility("default")))
#else
#define EMF_API
#endif

#define EMF_OK 0
#define EMF_ERR_ARG (-1)    /* invalid argument or call order */
#define EMF_ERR_HIP (-2)    /* a HIP runtime call failed */
#define EMF_ERR_NOMEM (-3)  /* device allocation failed */
#define EMF_ERR_STATE (-4)  /* handle not ready (dims/materials/pml missing) */

typedef struct emf_solver emf_solver;

/* Outcome of one emf_solve call (the loop of emopt/fdtd.py:1049-1194). */
typedef struct {
    int64_t n0;         /* step index at which the t0 snapshot was taken */
    int64_t Tn;         /* steps between snapshots, int(0.75 * Ncycle) */
    int64_t steps;      /* total time steps executed (n0 + 2*Tn) */
    int32_t status;     /* 0 converged, 1 hit Nmax, 2 fixed step count, 3 hit Nmax with
                           non-finite fields at the last check (the reference's loop runs on
                           past NaN/Inf, emopt/fdtd.py:1050-1052,1137-1138) */
    int32_t nchecks;    /* convergence checks performed */
    double a_change;    /* last relative amplitude change */
    double phi_change;  /* last relative phase change */
    double seconds;     /* wall time of the stepping on the device */
} emf_solve_info;

/* Message of the last failing call on this thread ("" if none). */
EMF_API const char *emf_last_error(void);

/* Library build string (target, compile flags). */
EMF_API const char *emf_version(void);

/* Number of visible HIP devices, or a negative status. */
EMF_API int emf_device_count(void);

/* ---- lifecycle -------------------------------------------------------- */

/* Replaces FDTD_new (src/fdtd.hpp:572).  Binds the handle to `device`. */
EMF_API int emf_new(int device, emf_solver **out);

/* No counterpart: the reference never frees its object.  Releases every
 * device buffer of the handle. */
EMF_API int emf_delete(emf_solver *s);

/* ---- geometry and time ------------------------------------------------ */

/* Replaces FDTD_set_wavelength (src/fdtd.hpp:575): R = wavelength / 2pi. */
EMF_API int emf_set_wavelength(emf_solver *s, double wavelength);

/* Replaces FDTD_set_physical_dims (src/fdtd.hpp:576-578). */
EMF_API int emf_set_physical_dims(emf_solver *s, double X, double Y, double Z,
                          double dx, double dy, double dz);

/* Replaces FDTD_set_grid_dims (src/fdtd.hpp:579).  (Re)allocates the
 * device field state for Nx*Ny*Nz cells. */
EMF_API int emf_set_grid_dims(emf_solver *s, int Nx, int Ny, int Nz);

/* Replaces FDTD_set_local_grid (src/fdtd.hpp:580-582).  One GPU owns the
 * whole grid, so the chunk must be (0,0,0,Nx,Ny,Nz). */
EMF_API int emf_set_local_grid(emf_solver *s, int k0, int j0, int i0, int K, int J, int I);

/* Replaces FDTD_set_dt (src/fdtd.hpp:583). */
EMF_API int emf_set_dt(emf_solver *s, double dt);

/* Replaces FDTD_set_complex_eps (src/fdtd.hpp:584). */
EMF_API int emf_set_complex_eps(emf_solver *s, int complex_eps);

/* Replaces FDTD_set_mat_arrays (src/fdtd.hpp:589-591).  Copies six complex
 * [Nz][Ny][Nx] arrays (eps_x, eps_y, eps_z, mu_x, mu_y, mu_z) to the device;
 * only the real parts of mu are used, as in the reference. */
EMF_API int emf_set_mat_arrays(emf_solver *s, const double *eps_x, const double *eps_y,
                       const double *eps_z, const double *mu_x, const double *mu_y,
                       const double *mu_z);

/* Overwrites eps_c (c = 0,1,2) on the sub-box [k1,k2)x[j1,j2)x[i1,i2) with a
 * complex array of that box's shape.  Used by FDTD.update(bbox)
 * (emopt/fdtd.py:882-992). */
EMF_API int emf_update_eps_box(emf_solver *s, int c, const double *vals, int k1, int k2,
                       int j1, int j2, int i1, int i2);

/* Overwrites mu_c (c = 0,1,2) on a sub-box, as emf_update_eps_box does for
 * eps: the mu half of FDTD.update(bbox) (emopt/fdtd.py:958-992).  Values
 * equal to the grid's uniform mu (real, imaginary part 0) change nothing;
 * any other value makes mu per cell (the H pass then reads dt/mu per cell). */
EMF_API int emf_update_mu_box(emf_solver *s, int c, const double *vals, int k1, int k2,
                      int j1, int j2, int i1, int i2);

/* Coefficient compression for real eps (no counterpart in the reference):
 * on (default), emf_set_mat_arrays also builds a per-cell index into the
 * table of distinct (eps_x, eps_y, eps_z) triples, and the E pass reads that
 * 2- or 4-byte index instead of three 8-byte coefficients b = dt/eps (the
 * table holds the same doubles, so results are identical).  It stays off
 * after emf_update_eps_box until the next emf_set_mat_arrays, and when there
 * are more than 2^20 distinct triples. */
EMF_API int emf_set_compression(emf_solver *s, int on);

/* Copies the device materials back (six complex arrays, any may be NULL).
 * Supports get_A_diag (emopt/fdtd.py:1669-1692). */
EMF_API int emf_get_mat_arrays(emf_solver *s, double *eps_x, double *eps_y, double *eps_z,
                       double *mu_x, double *mu_y, double *mu_z);

/* ---- boundary conditions and CPML -------------------------------------- */

/* Replaces FDTD_set_bc (src/fdtd.hpp:637): three chars from "0EH". */
EMF_API int emf_set_bc(emf_solver *s, const char *bc);

/* Replaces FDTD_set_pml_widths (src/fdtd.hpp:597-599). */
EMF_API int emf_set_pml_widths(emf_solver *s, int xmin, int xmax, int ymin, int ymax,
                       int zmin, int zmax);

/* Replaces FDTD_set_pml_properties (src/fdtd.hpp:600-601). */
EMF_API int emf_set_pml_properties(emf_solver *s, double sigma, double alpha, double kappa,
                           double pow);

/* Replaces FDTD_build_pml (src/fdtd.hpp:602): allocates the psi slabs. */
EMF_API int emf_build_pml(emf_solver *s);

/* Replaces FDTD_reset_pml (src/fdtd.hpp:603): zeroes the psi slabs. */
EMF_API int emf_reset_pml(emf_solver *s);

/* ---- sources ---------------------------------------------------------- */

/* Replaces FDTD_add_source (src/fdtd.hpp:625-629).  Six complex arrays of
 * shape [I][J][K] placed at (i0,j0,k0).  With calc_phase != 0 each value is
 * converted to (|v|, -atan2(im, re)) as the reference does in place
 * (src/fdtd.cpp:1517-1576); the caller's arrays are never modified.  With
 * calc_phase == 0 the arrays already hold (amplitude, phase). */
EMF_API int emf_add_source(emf_solver *s, const double *Jx, const double *Jy, const double *Jz,
                   const double *Mx, const double *My, const double *Mz,
                   int i0, int j0, int k0, int I, int J, int K, int calc_phase);

/* Replaces FDTD_clear_sources (src/fdtd.hpp:630). */
EMF_API int emf_clear_sources(emf_solver *s);

/* Replaces FDTD_set_source_properties (src/fdtd.hpp:632). */
EMF_API int emf_set_source_properties(emf_solver *s, double src_T, double src_min);

/* Replaces FDTD_src_func_t (src/fdtd.hpp:634). Host evaluation. */
EMF_API double emf_src_func_t(emf_solver *s, int n, double t, double phase);

/* ---- stepping --------------------------------------------------------- */

/* Replaces FDTD_update_H / FDTD_update_E (src/fdtd.hpp:593-594): one half
 * step on the device (asynchronous on the handle's stream). */
EMF_API int emf_update_H(emf_solver *s, int n, double t);
EMF_API int emf_update_E(emf_solver *s, int n, double t);

/* Zeroes the six time-domain field arrays (the reference's numpy fill(0),
 * emopt/fdtd.py:1017-1018). */
EMF_API int emf_zero_fields(emf_solver *s);

/* Copies the six real time-domain fields ([Nz][Ny][Nx] doubles each, any may
 * be NULL) to / from the host.  For step-level parity checks. */
EMF_API int emf_get_time_fields(emf_solver *s, double *Ex, double *Ey, double *Ez,
                        double *Hx, double *Hy, double *Hz);
EMF_API int emf_set_time_fields(emf_solver *s, const double *Ex, const double *Ey,
                        const double *Ez, const double *Hx, const double *Hy,
                        const double *Hz);

/* ---- snapshots and phasors -------------------------------------------- */

/* Replace FDTD_capture_t0_fields / FDTD_capture_t1_fields (src/fdtd.hpp:619-620)
 * for result slot `slot`. */
EMF_API int emf_capture_t0_fields(emf_solver *s, int slot);
EMF_API int emf_capture_t1_fields(emf_solver *s, int slot);

/* Replace FDTD_calc_complex_fields_2T / _3T (src/fdtd.hpp:621-622): recovers
 * the complex phasors of all six components into result slot `slot`. */
EMF_API int emf_calc_complex_fields_2T(emf_solver *s, int slot, double t0, double t1);
EMF_API int emf_calc_complex_fields_3T(emf_solver *s, int slot, double t0, double t1, double t2);

/* Replace FDTD_calc_phase_2T/3T and FDTD_calc_amplitude_2T/3T
 * (src/fdtd.hpp:613-617): scalar closed forms evaluated on the host. */
EMF_API double emf_calc_phase_2T(double t0, double t1, double f0, double f1);
EMF_API double emf_calc_amplitude_2T(double t0, double t1, double f0, double f1, double phase);
EMF_API double emf_calc_phase_3T(double t0, double t1, double t2, double f0, double f1, double f2);
EMF_API double emf_calc_amplitude_3T(double t0, double t1, double t2, double f0, double f1,
                             double f2, double phase);

/* ---- whole solve ------------------------------------------------------- */

/* Number of convergence sample points (nconv, emopt/fdtd.py:240-241,411-413);
 * the solve refuses a grid on which it yields more than 1024 points. */
EMF_API int emf_set_nconv(emf_solver *s, int nconv);

/* Runs the complete CW solve of emopt/fdtd.py:994-1194 on the device:
 * reset CPML and fields, step with a convergence check every Tn steps until
 * the amplitude change <= rtol and the phase change <= sqrt(rtol) (or the step
 * index exceeds nmax), capture t0, Tn steps, capture t1, Tn steps, and recover
 * the phasors into `slot`.  With fixed_n0 >= 0 the convergence loop is
 * replaced by exactly fixed_n0 steps.  Tn = int(0.75 * Ncycle) is computed by
 * the caller (emopt/fdtd.py:1021).  Sources currently registered are used.
 * Replaces the per-step Python loop plus FDTD_update_H/E, the 6 ghost
 * exchanges per step and the 6*nconv ctypes calls per check. */
EMF_API int emf_solve(emf_solver *s, int slot, double rtol, double nmax, int64_t Tn,
              int64_t fixed_n0, emf_solve_info *info);

/* Convergence history of the last emf_solve: up to `cap` records of
 * (n, a_change, phi_change).  Returns the number of records available. */
EMF_API int emf_get_history(emf_solver *s, double *out, int cap);

/* Copies the complex phasor of component `comp` in result slot `slot` on the
 * box [k1,k2)x[j1,j2)x[i1,i2) into `out` (complex [i2-i1][j2-j1][k2-k1]).
 * Replaces __get_field's gather (emopt/fdtd.py:1322-1359). */
EMF_API int emf_get_field(emf_solver *s, int slot, int comp, int k1, int k2, int j1, int j2,
                  int i1, int i2, double *out);

/* Component `comp` of result slot `slot` interpolated onto the Ez grid on the
 * box [k1,k2)x[j1,j2)x[i1,i2) (get_field_interp, emopt/fdtd.py:1401-1553):
 * each cell is the mean of the raw field over the component's stencil
 * (:1472-1547), the raw field being that of the grown box [ek1,ek2)x... --
 * the domain grown by one cell where the reference grows it (:1443-1457) --
 * and zero outside it; with an 'E'/'H' symmetry at the x / y min face and the
 * grown box starting there, the ghost below mirrors the first cell.
 * The domain must lie in the grown box and the grown box in the grid.  out:
 * complex [i2-i1][j2-j1][k2-k1] in host memory. */
EMF_API int emf_get_field_interp(emf_solver *s, int slot, int comp, int k1, int k2, int j1,
                                 int j2, int i1, int i2, int ek1, int ek2, int ej1, int ej2,
                                 int ei1, int ei2, double *out);

/* Poynting flux integrand of one face of get_source_power
 * (emopt/fdtd.py:1611-1665) reduced on the device: out = sum over the domain
 * of Re(a conj(b) - c conj(d)) with the fields interpolated as in
 * emf_get_field_interp (same domain and grown box), where (a, b, c, d) is
 * (Ey, Hz, Ez, Hy) for axis 0, (Ex, Hz, Ez, Hx) for axis 1 and
 * (Ex, Hy, Ey, Hx) for axis 2.  The caller applies the face's sign and
 * 0.5 * area.  Replaces the host sums over four downloaded planes. */
EMF_API int emf_interp_flux(emf_solver *s, int slot, int axis, int k1, int k2, int j1, int j2,
                    int i1, int i2, int ek1, int ek2, int ej1, int ej2, int ei1, int ei2,
                    double *out);

/* Stores up to six complex weight planes (NULL: weight 0) of `count` cells
 * on the device as set `id` (0..7) of this solver, for emf_interp_dot.  No
 * counterpart in the reference. */
EMF_API int emf_set_plane_weights(emf_solver *s, int id, int64_t count, const double *w_ex,
                          const double *w_ey, const double *w_ez, const double *w_hx,
                          const double *w_hy, const double *w_hz);

/* sum over the domain of sum_q W_q * field_q, the fields interpolated as in
 * emf_get_field_interp and W the weight set `id` (its cell count must be the
 * domain's): out[0..1] = the E components' sum (re, im), out[2..3] = the H
 * components'.  The overlap integrals of ModeMatch.compute
 * (emopt/fomutils.py:1047-1076) reduced on the device, with weights formed
 * once from the reference mode. */
EMF_API int emf_interp_dot(emf_solver *s, int slot, int id, int k1, int k2, int j1, int j2, int i1,
                   int i2, int ek1, int ek2, int ej1, int ej2, int ei1, int ei2, double *out);

/* out (host) = fwd_c * adj_c (complex, elementwise) over the box k1..k2 x
 * j1..j2 x i1..i2 for the three E components (which = 0) or H components
 * (which = 1) of the two result slots: 3 consecutive blocks of
 * (i2-i1)(j2-j1)(k2-k1) complex values, z-major.  The per-cell factors of
 * calc_ydAx_topology and calc_ydAx_autograd (emopt/experimental/fdtd.py:
 * 37-223), which take them from get_field and get_adjoint_field: 3 products
 * come back instead of 6 fields. */
EMF_API int emf_field_products(emf_solver *s, int which, int k1, int k2, int j1, int j2, int i1,
                               int i2, double *out);

/* ---- adjoint sources built on the device ------------------------------ */
/* The reference forms adjoint sources on the host: derivative planes on the
 * Ez grid (ModeMatch.get_dFd*, emopt/fomutils.py:1118-1196; the face terms
 * of power_norm_dFdx_3D, :1654-1920) mapped back onto the Yee grid by
 * interpolated_dFdx_3D (:1282-1429), then passed to FDTD_add_source
 * (src/fdtd.hpp:628).  These entry points build the same blocks on the
 * device, into a staged list, and emf_stage_register makes them the
 * solver's sources, so no plane crosses to the host.  A staged block holds
 * what emf_add_source would store for the same complex arrays. */

/* Empties the staged list. */
EMF_API int emf_stage_clear(emf_solver *s);

/* *count = blocks in the staged list. */
EMF_API int emf_stage_count(emf_solver *s, int *count);

/* Stages interpolated_dFdx_3D of the planes P_q = (scale_q x W_q) / div,
 * q = Ex..Hz, with W the weight set `id` (emf_set_plane_weights; its cell
 * count must be the domain's), scale[2q], scale[2q+1] the re and im of
 * scale_q, rdiv = 1 / div, and P_q = 0 where bit q of mask is clear.  The
 * domain (k1..i2, Ez-grid indices) fixes the transposed stencils and the
 * symmetric-boundary terms (at k1 = 0 or j1 = 0); place = (i0, j0, k0, id1,
 * jd1, kd1, I, J, K) is the grid block the result covers and its offset in
 * the grown domain (one cell more at -x, -y and +z).  For the PNF3D form:
 * scale = ModeMatch's dF/dfield factor, div = the source power. */
EMF_API int emf_stage_interp_T(emf_solver *s, int id, const double *scale, int mask, double rdiv,
                               int k1, int k2, int j1, int j2, int i1, int i2, const int *place);

/* Stages the source of one face of the source-power box
 * (power_norm_dFdx_3D, emopt/fomutils.py:1654-1920): with E1, E2, H1, H2 =
 * components comps[0..3] of result slot `slot` interpolated on the face's
 * domain (k1..i2; grown box ek1..ei2 as for emf_get_field_interp), the planes
 * dE1 = s1 H2*, dE2 = -s1 H1*, dH1 = -s1 E2*, dH2 = s1 E1*, each times mf and
 * rp2 (the reference's -f dS/dx / Psrc^2 with s1 = +-area/4, mf = -f,
 * rp2 = 1 / Psrc^2), transposed and placed as for emf_stage_interp_T. */
EMF_API int emf_stage_power_face(emf_solver *s, int slot, const int *comps, double s1, double mf,
                                 double rp2, int k1, int k2, int j1, int j2, int i1, int i2,
                                 int ek1, int ek2, int ej1, int ej2, int ei1, int ei2,
                                 const int *place);

/* Replaces the solver's sources with copies of the staged blocks (the
 * staged list is kept). */
EMF_API int emf_stage_register(emf_solver *s);

/* geom = (i0, j0, k0, I, J, K) of staged block idx; with out != NULL also
 * its six stored components (complex, I*J*K each, Jx..Mz). */
EMF_API int emf_stage_get(emf_solver *s, int idx, int *geom, double *out);

/* ---- measurement (no counterpart in the reference) --------------------- */

/* Runs `nsteps` time steps starting at step index n_start on the current
 * state (sources, CPML and fields as they are), with a convergence check
 * every `check_every` steps (0 = none), exactly as emf_solve's loop does, and
 * times them with HIP events on the handle's stream.  out[0] = total device
 * ms of the region.  With per_kernel != 0 every launch is timed as well: each
 * H- and E-pass launch records its own start and stop event through
 * hipExtLaunchKernelGGL (the dispatch's own timestamps, no marker packets
 * between the kernels), and an event pair brackets each fused-step launch.
 * out[1] = summed ms of the H-pass (or fused-step) launches, out[2] = summed
 * ms of the E-pass launches (0 for fused kinds); otherwise both are 0. */
EMF_API int emf_time_steps(emf_solver *s, int64_t n_start, int64_t nsteps, int64_t check_every,
                           int per_kernel, double *out);

/* Step kernel selection: kind 1 = separate H and E passes (default), 3 = the
 * same passes compiled for 5 waves per SIMD (a few registers spilled), 4 =
 * the two passes with the H pass taking its y-neighbour row of E through LDS
 * (uniform mu; the plain H pass otherwise), 0 = one fused H+E launch per step
 * recomputing the H row it needs, 2 = one fused launch sharing H through
 * LDS; tyg = rows (wavefronts) per work-group (1, 2,
 * 4 or 8; 4, 8 or 16 for kind 2); zc = z planes per work-group; order =
 * work-group order of the two passes (0 x-fastest, 1 z-fastest, 2 z-fastest
 * grouped by XCD, 3 x-fastest with the E pass walking backwards).  Defaults kind 1, tyg 4, zc 2, order 3.  Results are
 * identical; only speed differs. */
EMF_API int emf_set_tuning(emf_solver *s, int kind, int tyg, int zc, int order);

/* force = 1: the H and E passes address cells by 64-bit byte offsets at any
 * grid size (the same arithmetic; for tests).  Grids of more than 178 M cells
 * (three components above 4 GiB) use them regardless, with the two passes'
 * default shape (64 x 4 work-groups); the fused kernels (kinds 0, 2) and
 * batch members take only smaller grids.  No counterpart in the reference. */
EMF_API int emf_set_wide_offsets(emf_solver *s, int force);

/* Work-group shape of the H pass of the two-pass kernels (kinds 1 and 3):
 * rows (waves) per work-group tyg in {1, 2, 4, 8} and z planes per work-group
 * zc >= 1; 0 keeps the E pass's value (emf_set_tuning).  A launch shape only:
 * results are the same.  No counterpart in the reference. */
EMF_API int emf_set_h_shape(emf_solver *s, int tyg, int zc);

/* Streaming (non-temporal) access to the CPML psi arrays in the two passes:
 * -1 (default) on when a step's data -- fields, coefficients and psi --
 * exceeds the 256 MB Infinity Cache, 0 off, 1 on.  A cache policy only:
 * results are the same either way.  No counterpart in the reference. */
EMF_API int emf_set_streaming(emf_solver *s, int mode);

/* Free and total device memory of `device` in bytes (hipMemGetInfo).  No
 * counterpart in the reference; the solve server admits requests against
 * it. */
EMF_API int emf_mem_info(int device, int64_t *free_bytes, int64_t *total_bytes);

/* Device copy bandwidth (GB/s, read + write bytes) of a plain vector copy of
 * `bytes` bytes: one warm-up launch, then the average of `reps` launches back
 * to back over the same two buffers; a roofline reference point (and, for
 * buffers smaller than the Infinity Cache, the cache's copy rate). */
EMF_API int emf_copy_bandwidth(int device, int64_t bytes, int reps, double *gbps);

/* ---- batched solves (no counterpart in the reference) ------------------ */

/* A batch advances many configured solver handles on one device together:
 * one launch per half step covers every member (a persistent grid over a
 * device table of all members' tiles), and each member's whole solve loop --
 * the emf_solve loop of emopt/fdtd.py:994-1194, with its own grid size,
 * wavelength, sources, Tn and convergence test -- runs on the device, so a
 * chunk of steps is one captured hipGraph replayed with no host work per
 * step.  Members are admitted and retired between replays (continuous
 * batching).  A member's result agrees with emf_solve on the same handle to
 * roundoff, with the same n0 and step counts (per cell the arithmetic and the
 * source factors are the same), and does not depend on the other members.
 * Members need real eps and uniform mu.  While a handle is a running member
 * it must not be used through any other call. */
/* ----------------------------------------------------- material sampling
 * Cell-averaged values of z-stacked primitives at one Yee offset, on the GPU
 * (replaces the CPU sampling of StructuredMaterial3D, src/Grid.cpp:736-874 /
 * emopt/grid.py:874-961, that EMopt's FDTD.build/update call,
 * emopt/fdtd.py:842-992).  The arithmetic is that of the host sampler in
 * emopt_amd/grid.py, operation for operation, by one of two in-plane rules:
 *   supersample == 0: the reference's rule (StructuredMaterial2D::get_value,
 *     src/Grid.cpp:460-525): a primitive counts in a cell only when a cell
 *     corner lies strictly inside it; four corners inside with nothing of the
 *     cell taken yet give the primitive's value; otherwise it takes the exact
 *     area of the part of the cell not yet taken that it covers
 *     (GridCell::intersect, src/Grid.cpp:74-108), in double precision where
 *     the reference clips in __float128; the rest is background.  Rectangles
 *     and polygons only (ellipses as polygons); a cell may meet at most 8
 *     counted primitives and 48 edges of them, else EMF_ERR_ARG;
 *   supersample = ss in 1..64: rectangles by exact overlap, ellipses and
 *     polygons (even-odd rule) by ss x ss sample points per cell, each
 *     primitive taking at most what is left of the cell.
 * prims are given in priority order (layer, then insertion); a z segment
 * between consecutive primitive z bounds takes the primitives spanning it.
 * out: complex128 [i2-i1][j2-j1][k2-k1] in host memory. */
typedef struct {
    int kind;          /* 0 rectangle, 1 ellipse, 2 polygon */
    int clip;          /* ellipse: clipped to the box q[6..9] */
    int v0, nv;        /* polygon: vertices [v0, v0 + nv) of vx, vy */
    double re, im;     /* material value */
    double z1, z2;     /* extent along z */
    double bbox[4];    /* xmin, xmax, ymin, ymax (rectangle: the rectangle) */
    double q[10];      /* ellipse: x0, y0, a, b, cos(angle), sin(angle), clip xmin, xmax,
                          ymin, ymax */
} emf_prim;

EMF_API int emf_sample_structured(int device, const emf_prim *prims, int nprims,
                                  const double *vx, const double *vy, int nverts, double dx,
                                  double dy, double dz, int supersample, double bg_re,
                                  double bg_im, int k1, int k2, int j1, int j2, int i1, int i2,
                                  double sx, double sy, double sz, double *out);

typedef struct emf_batch emf_batch;

EMF_API int emf_batch_new(int device, emf_batch **out);
EMF_API int emf_batch_delete(emf_batch *b);

/* steps_per_graph = steps in one graph replay (default 64); zc = z planes per
 * work-group for members admitted from now on (default 2); blocks_per_cu =
 * persistent work-groups per CU (default 16: more, smaller chunks of the
 * tile range balance better than one resident round).  Only while nothing
 * runs. */
EMF_API int emf_batch_set_tuning(emf_batch *b, int steps_per_graph, int zc, int blocks_per_cu);

/* Cache-aware stepping: members are grouped (in admission order) so that a
 * group's hot state -- E, H and coefficients, ~80 B/cell -- stays within
 * group_bytes, and each graph replay steps one group.  Sized to the Infinity
 * Cache, a group's E pass re-reads from cache what its H pass just wrote.
 * rotate = 1: the groups take turns, replay by replay; rotate = 0: the first
 * group is stepped until its members finish, then the next (same throughput,
 * earlier results for the first admitted).  group_bytes = 0 (default): one
 * group, every replay steps all members.  Results never depend on either. */
EMF_API int emf_batch_set_grouping(emf_batch *b, int64_t group_bytes, int rotate);

/* Groups stepped at once (1..4, default 2), each by its own stream with its
 * own copy of the work table and graph: lane l steps the groups g with
 * g % lanes == l (the first `lanes` groups until their members finish, or
 * those in turn with rotate).  Lanes share no data, so results do not depend
 * on the number of lanes; the hot state of all lanes' groups together should
 * fit the Infinity Cache (lanes x group_bytes).  No counterpart in the
 * reference (emopt/fdtd.py runs one solve at a time). */
EMF_API int emf_batch_set_lanes(emf_batch *b, int lanes);

/* use_graph = 1 (default): each replay is one launch of the captured graph;
 * 0: the same kernels are launched one by one on the batch's stream (for
 * per-kernel tracing; slower by the launch overhead).  Results are the same
 * either way.  No counterpart in the reference (emopt/fdtd.py runs one solve
 * at a time). */
EMF_API int emf_batch_set_graph(emf_batch *b, int use_graph);

/* Admits solver s: resets its fields and CPML as emf_solve does and queues
 * the solve (result slot, rtol, nmax, Tn >= 1, fixed_n0 as for emf_solve).
 * *member receives the member id.  The solver handle stays owned by the
 * caller. */
EMF_API int emf_batch_add(emf_batch *b, emf_solver *s, int slot, double rtol, double nmax,
                          int64_t Tn, int64_t fixed_n0, int *member);

/* Replays the graph: until = 0 runs max_graphs replays; 1 returns as soon as
 * a member finishes (or after max_graphs replays); 2 runs until every member
 * has finished.  *finished (may be NULL) = members that finished during the
 * call. */
EMF_API int emf_batch_run(emf_batch *b, int max_graphs, int until, int *finished);

/* 1 and *info filled when the member has finished, 0 while it runs. */
EMF_API int emf_batch_poll(emf_batch *b, int member, emf_solve_info *info);

/* Detaches a member (finished or not); its handle is free for other calls,
 * and after a finished solve its result slot holds the phasors. */
EMF_API int emf_batch_remove(emf_batch *b, int member);

/* Numbers of running and finished (not yet removed) members. */
EMF_API int emf_batch_count(emf_batch *b, int *running, int *finished);

/* ---- gradient assembly ------------------------------------------------- */

/* y^T dA x of emopt/fdtd.py:1694-1717 on the device:
 *   sum_c sum_cells E_adj_c * 1j * (eps_c - eps0_c) * E_c
 *                 + H_adj_c * -1j * (mu_c - mu0_c) * H_c
 * over the box [k1,k2)x[j1,j2)x[i1,i2), where eps0 and mu0 are given as
 * complex arrays of the box's shape (the three mu0 arrays may all be NULL:
 * mu unchanged, no mu term).  Result written to out[0] (re), out[1] (im). */
EMF_API int emf_calc_ydAx(emf_solver *s, const double *eps0_x, const double *eps0_y,
                  const double *eps0_z, const double *mu0_x, const double *mu0_y,
                  const double *mu0_z, int k1, int k2, int j1, int j2, int i1, int i2,
                  double *out);

/* Hand-written gradient assembly for many design parameters at once:
 *   grad[p] = -2 Re( sum_c sum_cells E_adj_c * 1j * deps[p,c,cell] * E_c
 *                   + H_adj_c * -1j * dmu[p,c,cell] * H_c ) / step
 * where deps holds, per parameter, sparse (cell, component, d_material)
 * triples (complex); component 0..2 is eps_x..eps_z against Ex..Ez, 3..5 is
 * mu_x..mu_z against Hx..Hz.  This is the reference's per-parameter loop
 * (emopt/adjoint_method.py:504-533) collapsed into one device pass.
 * entries: n triples grouped by parameter; cell index g = (i*Ny + j)*Nx + k;
 * param ids in pidx. */
EMF_API int emf_grad_assemble(emf_solver *s, int64_t n, const int32_t *pidx,
                      const int32_t *comp, const int64_t *cell, const double *deps,
                      int nparam, double step, double *grad);

/* emf_grad_assemble without allocating device memory: the same kernel and
 * the same result (bit for bit), for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The entries, segments and gradient are
 * staged in a file-scope device array of 32 MiB per GPU, held until the
 * process ends; calls needing more (32 B per entry plus 16 B per parameter,
 * about 1 M entries) are rejected with EMF_ERR_ARG.  Calls on one GPU share that
 * array: the library serialises them within the process (a per-device lock),
 * and nothing a call reads is left over from an earlier one. */
EMF_API int emf_grad_assemble_noalloc(emf_solver *s, int64_t n, const int32_t *pidx,
                              const int32_t *comp, const int64_t *cell, const double *deps,
                              int nparam, double step, double *grad);

/* CPML stretch product s_x*s_y*s_z at the position of field component comp
 * (0..2 = Ex, Ey, Ez) of each cell, as complex (re, im) pairs in out[2n].
 * No counterpart in the reference.  In the frequency domain the CPML
 * recursion of the step kernels acts as a coordinate stretch
 * 1/s = 1/kappa + c/(1 - b e^{i dt}); the CPML operator is complex symmetric
 * only after scaling by these factors, so the adjoint gradient term of an eps
 * perturbation inside a CPML slab is -2 Re(s_x s_y s_z E_adj i d_eps E).
 * Outside the slabs the product is 1.  (The reference's gradient omits this
 * factor; adjoint sources are assumed to lie outside the slabs.) */
EMF_API int emf_pml_stretch(emf_solver *s, int64_t n, const int32_t *comp, const int64_t *cell,
                            double *out);

#ifdef __cplusplus
}
#endif

#endif /* EMFDTD_H */
