// Lagrangian particles: velocity interpolation from the physical velocity
// slots and Euler/AB2 advection in grid units (m_particles.f90:815-1230,
// cint_3d/cint :1451-1508), with the periodic wrap the slab exchange
// performs on one rank (particles_update_slabs :1236-1300).
//
// One thread per particle; positions and velocities stay in the reference's
// (3, np) order.  The arithmetic follows the reference's operation order
// without contraction, so results reproduce a plain IEEE evaluation of the
// same expressions bit for bit.
#include "particles.hpp"

namespace hit3d {

namespace {

constexpr int PTB = 256;

// row-major physical slot: value at 0-based (x, y, z)
struct Phys {
    const double* p;
    int n;
    int64_t rp;  // reals per row (2P)
    __device__ __forceinline__ double at(int x, int y, int z) const {
        return p[((int64_t)z * n + y) * rp + x];
    }
};

#pragma clang fp contract(off)

// cint (m_particles.f90:1485-1508): Lagrange cubic through nodes -1,0,1,2.
struct CintR {
    double r1, r2, r3, r4;
    __device__ __forceinline__ explicit CintR(double x) {
        double s = (x - 1.0) * (x - 2.0);
        r1 = -x * s;
        r2 = 3.0 * (x + 1.0) * s;
        s = (x + 1) * x;
        r3 = -3.0 * s * (x - 2.0);
        r4 = s * (x - 1.0);
    }
    __device__ __forceinline__ double operator()(double p1, double p2, double p3, double p4) const {
        return (r1 * p1 + r2 * p2 + r3 * p3 + r4 * p4) / 6.0;
    }
};

__device__ __forceinline__ int wrap1(int i, int n) {  // 1-based index into 1..n
    if (i < 1) i += n;
    if (i > n) i -= n;
    return i;
}

// Euler (first move, uvwp(1) == -999) or AB2, then the periodic wrap of
// m_particles.f90:938-1006 (x, y) and :1300-1337 (z on one rank).
__device__ __forceinline__ void advance_wrap(double* xyz, double* uvw, double up, double vp,
                                             double wp, int ip, int jp, int kp, double xp,
                                             double yp, double zp, double dt, double dx, int n) {
    if (uvw[0] == -999.0) {
        xp = xp + dt * up / dx;
        yp = yp + dt * vp / dx;
        zp = zp + dt * wp / dx;
    } else {
        xp = xp + (1.50 * up - 0.50 * uvw[0]) * dt / dx;
        yp = yp + (1.50 * vp - 0.50 * uvw[1]) * dt / dx;
        zp = zp + (1.50 * wp - 0.50 * uvw[2]) * dt / dx;
    }
    uvw[0] = up;
    uvw[1] = vp;
    uvw[2] = wp;
    if (xp > 1.0) {
        xp = xp - 1.0;
        ip = ip + 1;
        if (ip > n) ip = ip - n;
    }
    if (xp < 0.0) {
        xp = xp + 1.0;
        ip = ip - 1;
        if (ip < 1) ip = ip + n;
    }
    if (yp > 1.0) {
        yp = yp - 1.0;
        jp = jp + 1;
        if (jp > n) jp = jp - n;
    }
    if (yp < 0.0) {
        yp = yp + 1.0;
        jp = jp - 1;
        if (jp < 1) jp = jp + n;
    }
    if (zp > 1.0) {
        zp = zp - 1.0;
        kp = kp + 1;
    }
    if (zp < 0.0) {
        zp = zp + 1.0;
        kp = kp - 1;
    }
    double z = zp + (double)kp;
    if (z >= (double)(n + 1)) z = z - (double)n;
    if (z < 1.0) z = z + (double)n;
    xyz[0] = xp + (double)ip;
    xyz[1] = yp + (double)jp;
    xyz[2] = z;
}

// cell (1-based) and fraction of one coordinate; the cell is taken into
// 1..n so that a coordinate sitting exactly on the upper face stays in bounds
__device__ __forceinline__ int cell_of(double c, int n, double* frac) {
    int i = (int)floor(c);
    *frac = c - (double)i;
    return wrap1(i, n);
}

__global__ void __launch_bounds__(PTB) k_trilinear(Phys u, Phys v, Phys w, double* xyzp,
                                                   double* uvwp, int64_t np, double dt, double dx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int n = u.n;
    double* xyz = xyzp + 3 * p;
    double* uvw = uvwp + 3 * p;
    double xp, yp, zp;
    const int i = cell_of(xyz[0], n, &xp);
    const int j = cell_of(xyz[1], n, &yp);
    const int k = cell_of(xyz[2], n, &zp);
    const double c1 = (1.0 - xp) * (1.0 - yp) * (1.0 - zp);
    const double c2 = (xp) * (1.0 - yp) * (1.0 - zp);
    const double c3 = (1.0 - xp) * (yp) * (1.0 - zp);
    const double c4 = (xp) * (yp) * (1.0 - zp);
    const double c5 = (1.0 - xp) * (1.0 - yp) * (zp);
    const double c6 = (xp) * (1.0 - yp) * (zp);
    const double c7 = (1.0 - xp) * (yp) * (zp);
    const double c8 = (xp) * (yp) * (zp);
    // 0-based node indices; the upper neighbours wrap periodically
    const int x0 = i - 1, x1 = (i == n) ? 0 : i;
    const int y0 = j - 1, y1 = (j == n) ? 0 : j;
    const int z0 = k - 1, z1 = (k == n) ? 0 : k;
    double vel[3];
    const Phys* f[3] = {&u, &v, &w};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const Phys& a = *f[c];
        double s = c1 * a.at(x0, y0, z0) + c2 * a.at(x1, y0, z0) + c3 * a.at(x0, y1, z0) +
                   c4 * a.at(x1, y1, z0);
        s = s + c5 * a.at(x0, y0, z1) + c6 * a.at(x1, y0, z1) + c7 * a.at(x0, y1, z1) +
            c8 * a.at(x1, y1, z1);
        vel[c] = s;
    }
    advance_wrap(xyz, uvw, vel[0], vel[1], vel[2], i, j, k, xp, yp, zp, dt, dx, n);
}

__global__ void __launch_bounds__(PTB) k_cint(Phys u, Phys v, Phys w, double* xyzp, double* uvwp,
                                              int64_t np, double dt, double dx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int n = u.n;
    double* xyz = xyzp + 3 * p;
    double* uvw = uvwp + 3 * p;
    double xp, yp, zp;
    const int ip = cell_of(xyz[0], n, &xp);
    const int jp = cell_of(xyz[1], n, &yp);
    const int kp = cell_of(xyz[2], n, &zp);
    const CintR rx(xp), ry(yp), rz(zp);
    // 0-based stencil nodes ip-2..ip+1 (1-based ip-1..ip+2), periodic
    int xs[4], ys[4], zs[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        xs[t] = wrap1(ip - 2 + t + 1, n) - 1;
        ys[t] = wrap1(jp - 2 + t + 1, n) - 1;
        zs[t] = wrap1(kp - 2 + t + 1, n) - 1;
    }
    double vel[3];
    const Phys* f[3] = {&u, &v, &w};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const Phys& a = *f[c];
        double t1[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            double t2[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const double* row = a.p + ((int64_t)zs[kk] * n + ys[jj]) * a.rp;
                t2[jj] = rx(row[xs[0]], row[xs[1]], row[xs[2]], row[xs[3]]);
            }
            t1[kk] = ry(t2[0], t2[1], t2[2], t2[3]);
        }
        vel[c] = rz(t1[0], t1[1], t1[2], t1[3]);
    }
    advance_wrap(xyz, uvw, vel[0], vel[1], vel[2], ip, jp, kp, xp, yp, zp, dt, dx, n);
}

}  // namespace

void launch_particles_move(const Geo& g, const double* u, const double* v, const double* w,
                           double* xyzp, double* uvwp, int64_t np, double dt, double dx,
                           int scheme, hipStream_t s) {
    if (np <= 0) return;
    const Phys pu{u, g.n, 2 * g.P}, pv{v, g.n, 2 * g.P}, pw{w, g.n, 2 * g.P};
    const dim3 grid((unsigned)((np + PTB - 1) / PTB));
    if (scheme == 0)
        hipLaunchKernelGGL(k_trilinear, grid, dim3(PTB), 0, s, pu, pv, pw, xyzp, uvwp, np, dt, dx);
    else
        hipLaunchKernelGGL(k_cint, grid, dim3(PTB), 0, s, pu, pv, pw, xyzp, uvwp, np, dt, dx);
}

}  // namespace hit3d
