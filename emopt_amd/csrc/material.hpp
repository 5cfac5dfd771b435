// Structured material sampling on the device: the cell-averaged values of
// z-stacked primitives at one Yee offset (StructuredMaterial3D, the C++
// class of src/Grid.cpp:736-874 behind emopt/grid.py:874-961), computed with
// exactly the arithmetic of the host sampler in emopt_amd/grid.py -- same
// sample points, same operation order, no fused multiply-adds -- so the two
// give identical doubles.
//
//   value(i, j, k) = sum over z segments s of fz(i, s) * v(s, j, k)
//                    + (1 - sum fz) * background
//   v(s, j, k)     = sum over the primitives active in segment s, in priority
//                    order, of min(coverage, remaining) * value, the
//                    remaining part of the cell taking the background
//
// Rectangle coverage is the exact overlap of the cell with the rectangle;
// ellipse and polygon coverage is the fraction of ss x ss sample points
// inside (cells outside the primitive's bounding box: 0).
#pragma once
#include "kernels.hpp"

namespace emf {

struct MatPrim {
    int kind;                 // 0 rectangle, 1 ellipse, 2 polygon
    int clip;                 // ellipse clipped to q[6..9]
    int v0, nv;               // polygon vertices
    double re, im;
    double z1, z2;
    double bbox[4];           // xmin, xmax, ymin, ymax
    double q[10];             // ellipse: x0, y0, a, b, cos, sin, clip box
};

__device__ inline bool mat_inside(const MatPrim &p, const double *__restrict__ vx,
                                  const double *__restrict__ vy, double x, double y)
{
#pragma clang fp contract(off)
    if (p.kind == 1) {
        const double c = p.q[4], s = p.q[5];
        const double u = (x - p.q[0]) * c + (y - p.q[1]) * s;
        const double v = -(x - p.q[0]) * s + (y - p.q[1]) * c;
        const double ua = u / p.q[2], vb = v / p.q[3];
        bool in = ua * ua + vb * vb <= 1.0;
        if (p.clip) in = in && x >= p.q[6] && x < p.q[7] && y >= p.q[8] && y < p.q[9];
        return in;
    }
    bool res = false;   // polygon, even-odd rule
    for (int a = 0; a < p.nv; ++a) {
        const int b = a + 1 < p.nv ? a + 1 : 0;
        const double xa = vx[p.v0 + a], ya = vy[p.v0 + a];
        const double xb = vx[p.v0 + b], yb = vy[p.v0 + b];
        if (ya == yb) continue;
        const bool cond = (ya > y) != (yb > y);
        const double xint = xa + (y - ya) * (xb - xa) / (yb - ya);
        res = res != (cond && x < xint);
    }
    return res;
}

__device__ inline double mat_coverage(const MatPrim &p, const double *__restrict__ vx,
                                      const double *__restrict__ vy, double xc, double yc,
                                      double dx, double dy, int ss)
{
#pragma clang fp contract(off)
    const double hx = dx / 2, hy = dy / 2;
    if (p.kind == 0) {
        const double fx = fmax(fmin(xc + hx, p.bbox[1]) - fmax(xc - hx, p.bbox[0]), 0.0) / dx;
        const double fy = fmax(fmin(yc + hy, p.bbox[3]) - fmax(yc - hy, p.bbox[2]), 0.0) / dy;
        return fy * fx;
    }
    if (!(xc + hx > p.bbox[0] && xc - hx < p.bbox[1] && yc + hy > p.bbox[2] && yc - hy < p.bbox[3]))
        return 0.0;
    int count = 0;
    for (int a = 0; a < ss; ++a) {
        const double y = yc + (((double)a + 0.5) / ss - 0.5) * dy;
        for (int b = 0; b < ss; ++b) {
            const double x = xc + (((double)b + 0.5) / ss - 0.5) * dx;
            count += mat_inside(p, vx, vy, x, y) ? 1 : 0;
        }
    }
    return (double)count / (double)(ss * ss);
}

// v(s, j, k) for the used segments: v2[(s * nj + jj) * nk + kk]
__global__ void k_mat_inplane(const MatPrim *__restrict__ prims, const double *__restrict__ vx,
                              const double *__restrict__ vy, const int *__restrict__ seg_off,
                              const int *__restrict__ seg_prim, int nseg, int k1, int nk, int j1,
                              int nj, double sx, double sy, double dx, double dy, int ss,
                              double bg_re, double bg_im, double2 *__restrict__ v2)
{
#pragma clang fp contract(off)
    const long long total = (long long)nseg * nj * nk;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int kk = (int)(g % nk), jj = (int)((g / nk) % nj), s = (int)(g / ((long long)nk * nj));
        const double xc = ((double)(k1 + kk) + sx) * dx;
        const double yc = ((double)(j1 + jj) + sy) * dy;
        double rem = 1.0, vr = 0.0, vi = 0.0;
        for (int q = seg_off[s]; q < seg_off[s + 1]; ++q) {
            const MatPrim &p = prims[seg_prim[q]];
            const double cov = fmin(mat_coverage(p, vx, vy, xc, yc, dx, dy, ss), rem);
            vr += cov * p.re;
            vi += cov * p.im;
            rem -= cov;
        }
        vr += rem * bg_re;
        vi += rem * bg_im;
        v2[g] = make_double2(vr, vi);
    }
}

// out[(ii * nj + jj) * nk + kk] = sum_s fz * v(s) + (1 - sum fz) * background
__global__ void k_mat_combine(const double2 *__restrict__ v2, const double *__restrict__ zab,
                              int nseg, int i1, int ni, int nj, int nk, double sz, double dz,
                              double bg_re, double bg_im, double2 *__restrict__ out)
{
#pragma clang fp contract(off)
    const long long total = (long long)ni * nj * nk;
    const long long plane = (long long)nj * nk;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int ii = (int)(g / plane);
        const long long jk = g % plane;
        const double zc = ((double)(i1 + ii) + sz) * dz;
        const double zlo = zc - dz / 2, zhi = zc + dz / 2;
        double orr = 0.0, oi = 0.0, covered = 0.0;
        for (int s = 0; s < nseg; ++s) {
            const double fz = fmax(fmin(zhi, zab[2 * s + 1]) - fmax(zlo, zab[2 * s]), 0.0) / dz;
            if (fz > 0) {
                const double2 v = v2[s * plane + jk];
                orr += fz * v.x;
                oi += fz * v.y;
            }
            covered += fz;
        }
        orr += (1.0 - covered) * bg_re;
        oi += (1.0 - covered) * bg_im;
        out[g] = make_double2(orr, oi);
    }
}

}  // namespace emf
