// k_mk.hip -- Monte-Carlo neutral densities: the flight loop of Setka::MK_go / MK_fly_immit
// (Setka_MK.cpp:1225-2316) for the interstellar hydrogen that enters through the outer boundary.
//
// One thread per particle.  A particle is born on an outer boundary face (Gran::Get_Random_pozition,
// Setka::Velosity_initial), flies straight through the cells -- the exit face from the two
// triangles of each face (Setka::Time_to_vilet, Gran::Luch_crossing, Gran::rayTriangleIntersect) --
// while its optical depth grows by l / (|V| / nu), nu = rho_Th Velosity_1(u, cp) sigma / Kn, and
// charge-exchanges where the depth passes -ln(1 - xi): its velocity becomes the proton's
// (Setka::M_K_Change_Velosity, in the spherical frame of its position) and its sort the zone of the
// cell.  Every cell it crosses gains mu * time in its sort's slot, one atomic per segment; the
// sum over sorts (MK_n_H of Cell::MK_Add_particle) is formed from the slots at the end.  The
// particle ends at the outer boundary.  A particle that comes within 1.01 R0 of the centre after
// crossing a face is carried along its velocity to where it leaves that sphere and goes on in the
// cell there (Setka_MK.cpp:2138-2160): the cell on the inner boundary face the ray from the centre
// through its new position crosses, found by a walk over the inner faces (mk_locate_inner).
// Differences from the reference, which keep every loop bounded: the birth sampler gives up after
// 10000 tries and drops the particle (only on faces where the flow leaves the domain, whose inflow
// is ~exp(-V_inf^2) of the upwind faces'); a particle that finds no exit after 11 nudges towards
// the cell centre, or exceeds the step cap, is counted lost.  The tallies
// are FP64 atomic sums (the order of additions varies; the estimate is statistical).
//
// The flight is bound by memory latency at low occupancy (245+ VGPRs).  Measured choices: each
// cell's nodes, face-node indices and neighbours packed into one 240-byte record (one round of
// independent loads per segment instead of 18 dependent ones); exchanges held until a quarter of
// the wavefront's lanes have one due; a register cap of two waves per SIMD (a small spill beats
// one wave).  Persistent lanes that take a new particle when theirs ends were slower (2.5x): a
// wavefront then runs births, crossings and exchanges side by side.
#define TITAN_FLUX_ONLY 1
#include "step_kernels.h"
#include "bera.h"
#include "launch.h"

namespace titan {

namespace {
inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

constexpr double kMkSqrtPi = 1.77245385;  // sqrtpi_, Header.h:90 (the reference's literal)

__device__ __forceinline__ uint64_t mk_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// per-particle counter stream, values in (0, 1)
__device__ __forceinline__ double mk_rand(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    return ((double)(mk_mix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Velosity_1 (Setka_MK.cpp:40-50)
__device__ __forceinline__ double mk_velosity_1(double u, double cp) {
    if (u < 0.00001)
        return 2.0 * cp / kMkSqrtPi + 2.0 * u * u / (3.0 * cp * kMkSqrtPi) -
               u * u * u * u / (15.0 * cp * cp * cp * kMkSqrtPi);
    return exp(-u * u / (cp * cp)) * cp / kMkSqrtPi + (u + (cp * cp) / (2.0 * u)) * erf(u / cp);
}

__device__ __forceinline__ double mk_sigma(double a2, double x) {
    const double t = 1.0 - a2 * log(x);
    return t * t;
}

// Gran::rayTriangleIntersect (Gran.cpp:522-561)
__device__ __forceinline__ bool mk_ray_tri(const double o[3], const double d[3], const double *v0, const double *v1,
                                           const double *v2, double &t) {
    const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const double det = (e1[0] * pv[0] + e1[1] * pv[1]) + e1[2] * pv[2];
    if (fabs(det) < 1e-6) return false;
    const double inv = 1.0 / det;
    const double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const double u = ((tv[0] * pv[0] + tv[1] * pv[1]) + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const double v = ((d[0] * qv[0] + d[1] * qv[1]) + d[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    t = ((e2[0] * qv[0] + e2[1] * qv[1]) + e2[2] * qv[2]) * inv;
    return t > 1e-6;
}

// Setka::Time_to_vilet / Gran::Luch_crossing: the earliest crossing of the cell's faces, each
// split into the triangles (0,1,2) and (0,2,3) of its node order.  The cell's packed record
// (MkCell: its 8 node positions, each face's 4 nodes as indices into them) comes in with one
// round of independent loads; the positions go to this lane's LDS column (stride kMkThreads
// doubles), from which each face reads its nodes by index.  jx = the exit face's slot.
constexpr int kMkThreads = 256;
__device__ __forceinline__ bool mk_exit(const MkCell *__restrict__ pack, int64_t c, double *__restrict__ col,
                                        const double P[3], const double V[3], double &time, int &jx) {
    const double2 *src = reinterpret_cast<const double2 *>(pack[c].x);
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double2 v = src[k];
        col[(2 * k) * kMkThreads] = v.x;
        col[(2 * k + 1) * kMkThreads] = v.y;
    }
    const uint4 lw = *reinterpret_cast<const uint4 *>(pack[c].loc);
    const uint32_t loc[3] = {lw.x, lw.y, lw.z};
    const double nv = sqrt((V[0] * V[0] + V[1] * V[1]) + V[2] * V[2]);
    const double d[3] = {V[0] / nv, V[1] / nv, V[2] / nv};
    double tmin = 1e300;
    bool found = false;
#pragma unroll 1
    for (int j = 0; j < 6; j++) {
        const uint32_t q = (loc[j >> 1] >> (16 * (j & 1))) & 0xffffu;  // 4 nodes, 4 bits each
        double p0[3], p1[3], p2[3], p3[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            p0[i] = col[(3 * (q & 7u) + i) * kMkThreads];
            p1[i] = col[(3 * ((q >> 4) & 7u) + i) * kMkThreads];
            p2[i] = col[(3 * ((q >> 8) & 7u) + i) * kMkThreads];
            p3[i] = col[(3 * ((q >> 12) & 7u) + i) * kMkThreads];
        }
        double t1 = 0.0, t2 = 0.0;
        const bool b1 = mk_ray_tri(P, d, p0, p1, p2, t1);
        const bool b2 = mk_ray_tri(P, d, p0, p2, p3, t2);
        if (!b1 && !b2) continue;
        // distances along the unit direction: dividing each by |V| (as the reference does) keeps
        // their order, so the division is taken once, for the nearest
        const double tf = (b1 && b2) ? smin(t1, t2) : (b1 ? t1 : t2);
        if (tmin > tf) { tmin = tf; jx = j; found = true; }
    }
    time = found ? tmin / nv : 1e10;
    if (!(time < 1e10)) {  // the reference accepts only crossings sooner than 1e10
        time = 1e10;
        found = false;
    }
    return found;
}

// spherical components in the frame with polar axis x at the position X (spherical_skorost with
// (z, x, y) = (x0, x1, x2), Help_funk.cpp:108-121) and back (dekard_skorost, :123-135); the frame's
// sines and cosines are taken once per exchange (sincos gives the values of sin and cos)
struct MkFrame {
    double st, ct, sp, cp;
};
__device__ __forceinline__ MkFrame mk_frame(const double X[3]) {
    const double rr = sqrt(X[1] * X[1] + X[2] * X[2] + X[0] * X[0]);
    const double the = acos(X[0] / rr);
    const double phi = polar_angle(X[1], X[2]);
    MkFrame F;
    sincos(the, &F.st, &F.ct);
    sincos(phi, &F.sp, &F.cp);
    return F;
}
__device__ __forceinline__ void mk_to_sph(const MkFrame &F, const double Vc[3], double &r, double &th, double &ph) {
    const double Vz = Vc[0], Vx = Vc[1], Vy = Vc[2];
    r = Vx * F.st * F.cp + Vy * F.st * F.sp + Vz * F.ct;
    th = Vx * F.ct * F.cp + Vy * F.ct * F.sp - Vz * F.st;
    ph = -Vx * F.sp + Vy * F.cp;
}
__device__ __forceinline__ void mk_from_sph(const MkFrame &F, double r, double th, double ph, double Vc[3]) {
    Vc[1] = r * F.st * F.cp + th * F.ct * F.cp - ph * F.sp;
    Vc[2] = r * F.st * F.sp + th * F.ct * F.sp + ph * F.cp;
    Vc[0] = r * F.ct - th * F.st;
}

// Setka::M_K_Change_Velosity (Setka_MK.cpp:2788-2850), in units of cp; capped at 10000 tries
__device__ __forceinline__ void mk_change_velosity(uint64_t &rs, double a2, const double U[3], const double V[3],
                                                   double W[3], double cp, bool &capped) {
    const double X = sqrt((V[0] - U[0]) * (V[0] - U[0]) + (V[1] - U[1]) * (V[1] - U[1]) + (V[2] - U[2]) * (V[2] - U[2]));
    const double p4 = 0.5 * kMkSqrtPi * X / (1.0 + 0.5 * kMkSqrtPi * X);
    const double sX = mk_sigma(a2, X * cp);
    double v1 = 0.0, v2 = 0.0, v3 = 0.0;
#pragma unroll 1
    for (int tries = 0; tries < 10000; tries++) {
        const double k1 = mk_rand(rs), k2 = mk_rand(rs), k3 = mk_rand(rs), k4 = mk_rand(rs), k5 = mk_rand(rs),
                     k6 = mk_rand(rs);
        double y1, y2, y3;
        double s5, c5;
        sincos(2.0 * kPi * k5, &s5, &c5);
        if (p4 < k1) {
            const double om1 = 1.0 - 2.0 * k4;
            const double om2 = sqrt(1.0 - om1 * om1) * c5;
            const double om3 = sqrt(1.0 - om1 * om1) * s5;
            const double lo = sqrt(-log(k2 * k3));
            y1 = lo * om1; y2 = lo * om2; y3 = lo * om3;
        } else {
            y1 = sqrt(-log(k2)) * cos(kPi * k3);
            y2 = sqrt(-log(k4)) * c5;
            y3 = sqrt(-log(k4)) * s5;
        }
        v1 = y1 + U[0]; v2 = y2 + U[1]; v3 = y3 + U[2];
        const double u1 = V[0] - v1, u2 = V[1] - v2, u3 = V[2] - v3;
        const double uuu = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
        const double yy = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
        const double h = (uuu * mk_sigma(a2, uuu * cp)) / (sX * (X + yy));
        if (!(h < k6)) { W[0] = v1; W[1] = v2; W[2] = v3; return; }
    }
    capped = true;
    W[0] = v1; W[1] = v2; W[2] = v3;
}

// Setka::Velosity_initial (Setka_MK.cpp:2671-2723); false when the sampler gives up
__device__ __forceinline__ bool mk_velosity_initial(uint64_t &rs, double v_inf, const double n[3], double V[3]) {
    double t[3], m[3];
    get_bazis<double>(n, t, m);
    const double Vi[3] = {v_inf * n[0] + 0.0 * n[1] + 0.0 * n[2], v_inf * t[0] + 0.0 * t[1] + 0.0 * t[2],
                          v_inf * m[0] + 0.0 * m[1] + 0.0 * m[2]};
    const double p1 = fabs(Vi[0]) * kMkSqrtPi / (1.0 + fabs(Vi[0]) * kMkSqrtPi);
    double z = 0.0;
    bool ok = false;
#pragma unroll 1
    for (int tries = 0; tries < 10000; tries++) {
        const double k3 = mk_rand(rs), k4 = mk_rand(rs), k5 = mk_rand(rs), k6 = mk_rand(rs);
        if (p1 > k3) z = cos(kPi * k5) * sqrt(-log(k4));
        else if (k4 <= 0.5) z = -sqrt(-log(2.0 * k4));
        else z = sqrt(-log(2.0 * (1.0 - k4)));
        if (!(fabs(z + Vi[0]) / (fabs(Vi[0]) + fabs(z)) < k6 || z > -Vi[0])) { ok = true; break; }
    }
    if (!ok) return false;
    const double V1 = z + Vi[0];
    const double k1 = mk_rand(rs), k2 = mk_rand(rs);
    const double a = sqrt(-log(k2));
    double s1, c1;
    sincos(2.0 * kPi * k1, &s1, &c1);
    const double V2 = Vi[1] + a * c1;
    const double V3 = Vi[2] + a * s1;
    for (int i = 0; i < 3; i++) V[i] = V1 * n[i] + V2 * t[i] + V3 * m[i];
    return true;
}

// findSphereIntersectionTime (Setka_MK.cpp:4-38): the first non-negative time at which X + t V is at
// radius R.
__device__ __forceinline__ bool mk_sphere_time(const double X[3], const double V[3], double R, double &time) {
    const double a = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
    const double b = 2.0 * (X[0] * V[0] + X[1] * V[1] + X[2] * V[2]);
    const double c = (X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) - R * R;
    const double D = b * b - 4 * a * c;
    if (D < 0) return false;
    const double sqrtD = sqrt(D);
    const double t1 = (-b - sqrtD) / (2 * a);
    const double t2 = (-b + sqrtD) / (2 * a);
    if (t1 >= 0) { time = t1; return true; }
    if (t2 >= 0) { time = t2; return true; }
    return false;
}

__device__ __forceinline__ void mk_cross(const double *a, const double *b, double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Whether the direction d lies inside inner face i as seen from the centre; if not, the edge to
// cross next (the first whose plane through the centre puts d and the face's centre on opposite
// sides).  Returns -1 inside, else the edge index.
__device__ __forceinline__ int mk_inner_side(const DevMesh &M, const MkDev &K, const double *__restrict__ X, int i,
                                             const double d[3]) {
    const int64_t f = K.inner_face[i];
    const double *a[4];
    double cen[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 4; k++) {
        a[k] = X + 3 * (int64_t)M.face_nodes[4 * f + k];
        for (int e = 0; e < 3; e++) cen[e] += a[k][e];
    }
    for (int k = 0; k < 4; k++) {
        double ne[3];
        mk_cross(a[k], a[(k + 1) & 3], ne);
        const double sc = ne[0] * cen[0] + ne[1] * cen[1] + ne[2] * cen[2];
        const double sd = ne[0] * d[0] + ne[1] * d[1] + ne[2] * d[2];
        if ((sc > 0.0 && sd < 0.0) || (sc < 0.0 && sd > 0.0)) return k;
    }
    return -1;
}

// The cell on the inner boundary face the ray from the centre through P crosses: a visibility walk
// over the inner faces from face `start` (a few steps across neighbouring faces towards P); past
// the step cap every inner face is tried in turn.  -1 if none holds it.
__device__ int mk_locate_inner(const DevMesh &M, const MkDev &K, const double *__restrict__ X, int start,
                               const double P[3]) {
    int i = start;
    const int cap = 64 + 4 * (int)sqrt((double)K.n_inner);
    for (int step = 0; step < cap && i >= 0; step++) {
        const int k = mk_inner_side(M, K, X, i, P);
        if (k < 0) return M.face_cells[2 * (int64_t)K.inner_face[i]];
        i = K.inner_nb[4 * i + k];
    }
    for (int j = 0; j < K.n_inner; j++)
        if (mk_inner_side(M, K, X, j, P) < 0) return M.face_cells[2 * (int64_t)K.inner_face[j]];
    return -1;
}

// Index of device face f among the inner faces (ascending), or 0 if it is not one.
__device__ __forceinline__ int mk_inner_index(const MkDev &K, int64_t f) {
    int lo = 0, hi = K.n_inner - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (K.inner_face[mid] < f) lo = mid + 1;
        else hi = mid;
    }
    return (K.n_inner > 0 && K.inner_face[lo] == f) ? lo : 0;
}

// ---- zone borders: the velocity distributions of AMR_f (AMR_f.cpp:4-50,240-316) at the unrefined
// grid of Gran::Read_AMR (Gran.cpp:66-99) refined r times in each direction: (0, 20) x (-20, 20)^2 in
// the face's frame (Vn, get_bazis(Vn)), Vn = the normal (side 0, what leaves cells[0]) or its
// negative (side 1), bins (i, j, k) with k fastest, each holding the weight deposited in it.
__device__ __forceinline__ void mk_dist_frame(const double n[3], int side, double e[3], double t[3], double m[3]) {
    for (int i = 0; i < 3; i++) e[i] = side ? -n[i] : n[i];
    get_bazis<double>(e, t, m);
}

// AMR_f::Add_particle: the bin of V in the face's frame, coordinates clipped 1e-7 inside the box
__device__ __forceinline__ void mk_deposit(double *w, int r, const double n[3], int side, const double V[3], double mu) {
    double e[3], t[3], m[3];
    mk_dist_frame(n, side, e, t, m);
    const int nx = 3 * r, ny = 6 * r, nz = 6 * r;
    const double xL = 0.0, xR = 20.0, yL = -20.0, yR = 20.0, zL = -20.0, zR = 20.0;
    double x = (V[0] * e[0] + V[1] * e[1]) + V[2] * e[2];
    double y = (V[0] * t[0] + V[1] * t[1]) + V[2] * t[2];
    double z = (V[0] * m[0] + V[1] * m[1]) + V[2] * m[2];
    if (x < xL) x = xL + 1e-7;
    if (x > xR) x = xR - 1e-7;
    if (y < yL) y = yL + 1e-7;
    if (y > yR) y = yR - 1e-7;
    if (z < zL) z = zL + 1e-7;
    if (z > zR) z = zR - 1e-7;
    const double dx = (xR - xL) / nx, dy = (yR - yL) / ny, dz = (zR - zL) / nz;
    // a coordinate that is not a number (none is expected) goes to bin 0 rather than out of the grid
    int i = (x >= xL && x <= xR) ? (int)((x - xL) / dx) : 0;
    int j = (y >= yL && y <= yR) ? (int)((y - yL) / dy) : 0;
    int k = (z >= zL && z <= zR) ? (int)((z - zL) / dz) : 0;
    if (i >= nx) i = nx - 1;
    if (j >= ny) j = ny - 1;
    if (k >= nz) k = nz - 1;
    atomicAdd(&w[((int64_t)i * ny + j) * nz + k], mu);
}

// AMR_f::Get_random_velosity with AMR_cell::Get_random_velosity_in_cell: the first bin whose running
// sum passes ksi = xi * (the total), then Vn with density ~ Vn across it, Vt and Vm uniform
__device__ __forceinline__ void mk_dist_sample(uint64_t &rs, int r, const double *__restrict__ cdf, const double n[3],
                                               int side, double V[3]) {
    const int nx = 3 * r, ny = 6 * r, nz = 6 * r;
    const int64_t nb = (int64_t)nx * ny * nz;
    const double ksi = mk_rand(rs) * cdf[nb - 1];
    int64_t lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > ksi) hi = mid;
        else lo = mid + 1;
    }
    if (!(cdf[lo] > ksi))  // rounding at the top of the sum: the last bin with weight
        while (lo > 0 && cdf[lo] == cdf[lo - 1]) lo--;
    const int i = (int)(lo / ((int64_t)ny * nz)), j = (int)((lo / nz) % ny), k = (int)(lo % nz);
    const double dx = 20.0 / nx, dy = 40.0 / ny, dz = 40.0 / nz;
    const double xl = 0.0 + i * dx, xr = 0.0 + (i + 1) * dx, yl = -20.0 + j * dy, yr = -20.0 + (j + 1) * dy,
                 zl = -20.0 + k * dz, zr = -20.0 + (k + 1) * dz;
    const double cx = (xr + xl) / 2.0, cy = (yr + yl) / 2.0, cz = (zr + zl) / 2.0;
    const double hx = xr - xl, hy = yr - yl, hz = zr - zl;
    const double L = cx - hx / 2.0, R = cx + hx / 2.0;
    const double v0 = sqrt(L * L + mk_rand(rs) * (R * R - L * L));
    const double v1 = (cy - hy / 2.0) + mk_rand(rs) * hy;
    const double v2 = (cz - hz / 2.0) + mk_rand(rs) * hz;
    double e[3], t[3], m[3];
    mk_dist_frame(n, side, e, t, m);
    for (int q = 0; q < 3; q++) V[q] = (v0 * e[q] + v1 * t[q]) + v2 * m[q];
}

// running sums of each incoming distribution's bins, in bin order (the order Culk_SpotokV and
// Get_random_velosity add them in): one thread per distribution
__global__ void __launch_bounds__(64) k_mkz_cdf(int64_t n, const int64_t *__restrict__ off, const double *__restrict__ dist,
                                                int64_t nb, double *__restrict__ cdf) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const double *w = dist + off[g];
    double *o = cdf + g * nb;
    double s = 0.0;
    for (int64_t b = 0; b < nb; b++) {
        s = s + w[b];
        o[b] = s;
    }
}

// zero n distributions (their outgoing side, before a pass: MK_go's Fill_null)
__global__ void __launch_bounds__(256) k_mkz_clear(int64_t n, const int64_t *__restrict__ off, double *__restrict__ dist,
                                                   int64_t nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nb) return;
    const int64_t g = i / nb;
    dist[off[g] + (i - g * nb)] = 0.0;
}

// One record per cell for the flights: the 8 node positions of the cell, each face's 4 nodes (in
// the face's own order) as indices into them, and the cell across each face (-1 - boundary code
// on a boundary face), faces in the device order of cell_faces.
__global__ void __launch_bounds__(256) k_mk_pack(DevMesh M, const double *__restrict__ X, MkCell *__restrict__ pack) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.n_cells) return;
    MkCell &R = pack[c];
    int32_t cn[8];
    for (int k = 0; k < 8; k++) {
        cn[k] = M.cell_nodes[8 * c + k];
        for (int i = 0; i < 3; i++) R.x[3 * k + i] = X[3 * (int64_t)cn[k] + i];
    }
    uint32_t loc[3] = {0u, 0u, 0u};
    for (int j = 0; j < 6; j++) {
        const int64_t f = M.cell_faces[6 * c + j];
        uint32_t q = 0u;
        for (int m = 0; m < 4; m++) {
            const int32_t node = M.face_nodes[4 * f + m];
            uint32_t k8 = 0u;
            for (int k = 0; k < 8; k++)
                if (cn[k] == node) k8 = (uint32_t)k;
            q |= k8 << (4 * m);
        }
        loc[j >> 1] |= q << (16 * (j & 1));
        const int32_t a = M.face_cells[2 * f], b = M.face_cells[2 * f + 1];
        R.nbr[j] = b < 0 ? -1 - (int32_t)(M.face_code[f] & 15u) : (a == c ? b : a);
    }
    R.loc[0] = loc[0];
    R.loc[1] = loc[1];
    R.loc[2] = loc[2];
    R.loc[3] = 0u;
}

// normal and area of each candidate injection face, for the host's particle counts
__global__ void __launch_bounds__(256) k_mk_gather(int64_t n, const int32_t *__restrict__ faces, DevGeom G,
                                                   double *__restrict__ na) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t f = faces[i];
    na[4 * i] = G.face_normal[3 * f];
    na[4 * i + 1] = G.face_normal[3 * f + 1];
    na[4 * i + 2] = G.face_normal[3 * f + 2];
    na[4 * i + 3] = G.face_area[f];
}

// tallies per unit volume (Cell::MK_normir_Moments, Cell.cpp:1175-1215), written in reference
// cell order for the caller
__global__ void __launch_bounds__(256) k_mk_finish(int64_t n_cells, const double *__restrict__ vol,
                                                   const int32_t *__restrict__ cell_ref,
                                                   const double *__restrict__ tally, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * n_cells) return;
    const int64_t d = i / 5, k = i - 5 * d;
    const double *t = tally + 5 * d;
    // slot 0 holds the sorts outside 1..4; the total adds every slot
    const double v = k == 0 ? (((t[0] + t[1]) + t[2]) + t[3]) + t[4] : t[k];
    out[5 * (int64_t)cell_ref[d] + k] = v / vol[d];
}

// Cell::MK_Add_pui_source (Cell.cpp:1087-1133) with parent 0 (a thermal proton), at the zone of
// Setka::determ_zone(P.cel, 0) (Setka_MK.cpp:2010, 2088): Q routes it when the scalars are carried,
// else the cell's type stands in
__device__ __forceinline__ void mk_pui_add(const MkDev &K, const DevMesh &M, const double *__restrict__ prim,
                                           int64_t c, int sort, double u, double nu_ex, double mu, double t) {
    const int nw = K.pui_nw;
    const double xi = u / K.pui_wr * nw;
    const int iw = xi >= (double)nw ? nw - 1 : (xi > 0.0 ? (int)xi : 0);
    atomicAdd(&K.pui_sm[c * nw + iw], mu * t);
    int zone = (int)((M.cell_code[c] >> 8) & 15u);
    if (K.scal) zone = bera_zone(K.scal[kScal * c], prim[8 * c], zone);
    if (zone < 1 || zone > 4 || sort < 1 || sort > 4) return;
    const int k = K.pui_arise[(zone - 1) * 4 + (sort - 1)];
    if (k > 0) atomicAdd(&K.pui_sp[(c * nw + iw) * 2 + (k - 1)], mu * t * nu_ex);
}

// The angle integral of Cell::MK_calc_Sm (Cell.cpp:1135-1169) does not depend on the cell: for the
// speed bins ij (w) and j (Vh), tt[j*nw + ij] = sum over the 40 angles th_k = k pi / 40 of
// d sigma(d) sin(th_k) dth 2 pi, d = |w - Vh| at th_k (terms with d <= 1e-9 left out); the
// angles' sines and cosines come in trig[0..39], trig[40..79].  One thread per (j, ij).
__global__ void __launch_bounds__(256) k_mk_pui_table(int32_t nw, double wr, double a2,
                                                      const double *__restrict__ trig, double *__restrict__ tt) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nw * nw) return;
    const int j = (int)(q / nw), ij = (int)(q - (int64_t)j * nw);
    const double pi = 3.14159265358979323846;
    const double dthe = pi / 40.0;
    const double w = ((ij + 0.5) * wr / nw), Vh = ((j + 0.5) * wr / nw);
    double acc = 0.0;
    for (int k = 1; k <= 40; k++) {
        const double d = sqrt(Vh * Vh + w * w - 2.0 * w * Vh * trig[40 + k - 1]);
        if (d > 0.000000001) {
            const double lg = log(d);
            const double sg = (1.0 - a2 * lg) * (1.0 - a2 * lg);
            acc += d * sg * trig[k - 1] * dthe * 2.0 * pi;
        }
    }
    tt[q] = acc;
}

// Cell::MK_normir_Moments (Cell.cpp:1225-1241) and Cell::MK_calc_Sm of the pickup-ion tallies, one
// workgroup per cell and one thread per speed bin, in place: S- becomes the loss rate at w_ij,
// sum over the bins j with ff_j > 0 of ff_j tt[j][ij], over 4 pi and par_Kn, from the
// volume-normalised tally ff (staged in LDS); S+ is divided by V 4 pi / 3 (w2^3 - w1^3).  The
// reference sums ff_j d sigma sin dth 2 pi term by term; taking ff_j out of the angle sum changes
// the result only by rounding.  vol in device cell order.
__global__ void __launch_bounds__(256) k_mk_pui_finish(int64_t n, int32_t nw, double wr, double kn,
                                                       const double *__restrict__ vol, const double *__restrict__ tt,
                                                       double *__restrict__ sm, double *__restrict__ sp) {
    extern __shared__ double ff[];  // [nw] normalised S- of the cell
    const double pi = 3.14159265358979323846;
    for (int64_t c = blockIdx.x; c < n; c += gridDim.x) {
        const double V = vol[c];
        bool any = false;
        for (int ij = threadIdx.x; ij < nw; ij += blockDim.x) {
            const double w1 = ij * wr / nw, w2 = (ij + 1) * wr / nw;
            const double f = sm[c * nw + ij] / V;
            ff[ij] = f;
            any |= f > 0.0;
            for (int ki = 0; ki < 2; ki++)
                sp[(c * nw + ij) * 2 + ki] =
                    sp[(c * nw + ij) * 2 + ki] / (V * 4.0 * pi * (1.0 / 3.0) * (w2 * w2 * w2 - w1 * w1 * w1));
        }
        any = __syncthreads_or(any);
        for (int ij = threadIdx.x; ij < nw; ij += blockDim.x) {
            double acc = 0.0;
            if (any) {
                for (int j = 0; j < nw; j++) {
                    const double f = ff[j];
                    if (f <= 0.0) continue;
                    acc += f * tt[(int64_t)j * nw + ij];
                }
                acc = acc / (4.0 * pi);
            }
            sm[c * nw + ij] = acc / kn;
        }
        __syncthreads();
    }
}

// ZONE: the zone form (titan_mk_go_zone); PUI: the pickup-ion tallies.  The one-zone flight
// without tallies is its own instantiation, so the zone tests, deposits and tallies add nothing
// to its registers.
template <bool ZONE, bool PUI>
__global__ void __launch_bounds__(kMkThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_mk_particles(DevMesh M, MkDev K, const double *__restrict__ X, const double *__restrict__ prim, DevGeom G) {
    __shared__ double lds[24 * kMkThreads];
    double *col = lds + threadIdx.x;
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= K.n_particles) return;
    // injection face: the last q with first[q] <= pid
    int lo = 0, hi = (int)K.n_in - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (K.first[mid] <= pid) lo = mid;
        else hi = mid - 1;
    }
    const int q = lo;
    const int64_t f = K.face[q];
    const double n[3] = {G.face_normal[3 * f], G.face_normal[3 * f + 1], G.face_normal[3 * f + 2]};
    const int32_t *fn = M.face_nodes + 4 * f;
    uint64_t rs = mk_mix(K.seed ^ mk_mix((uint64_t)pid + 0x9E3779B97F4A7C15ull));
    // Gran::Get_Random_pozition (Gran.cpp:220-285)
    const double *A = X + 3 * (int64_t)fn[0], *B1 = X + 3 * (int64_t)fn[1], *Cc = X + 3 * (int64_t)fn[2],
                 *B2 = X + 3 * (int64_t)fn[3];
    double s0, s1;
    {
        double v1x = B1[0] - A[0], v1y = B1[1] - A[1], v1z = B1[2] - A[2];
        const double v2x = Cc[0] - A[0], v2y = Cc[1] - A[1], v2z = Cc[2] - A[2];
        double cx = v1y * v2z - v1z * v2y, cy = v1z * v2x - v1x * v2z, cz = v1x * v2y - v1y * v2x;
        s0 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        v1x = B2[0] - A[0]; v1y = B2[1] - A[1]; v1z = B2[2] - A[2];
        cx = v1y * v2z - v1z * v2y; cy = v1z * v2x - v1x * v2z; cz = v1x * v2y - v1y * v2x;
        s1 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    }
    const double S = s0 + s1;
    const double ksi = mk_rand(rs);
    const double *T1 = (ksi <= s0 / S) ? B1 : B2;
    const double r1 = mk_rand(rs), r2 = mk_rand(rs);
    double P[3], V[3];
    for (int i = 0; i < 3; i++) P[i] = (1.0 - sqrt(r1)) * A[i] + (sqrt(r1) * (1.0 - r2)) * T1[i] + (r2 * sqrt(r1)) * Cc[i];
    unsigned long long *st = K.stats;
    const int in_side = ZONE ? (int)K.in_side[q] : -1;
    if (ZONE && in_side >= 0) {  // from the face's distribution (AMR_f::Get_random_velosity)
        mk_dist_sample(rs, K.r, K.cdf + K.in_cdf[q], n, in_side, V);
    } else if (!mk_velosity_initial(rs, K.v_inf, n, V)) {  // no admissible velocity within the cap: dropped
        atomicAdd(&st[0], 1ull);
        atomicAdd(&st[5], 1ull);
        return;
    }
    unsigned long long n_pass = 0;
    int64_t c = in_side == 0 ? M.face_cells[2 * f + 1] : M.face_cells[2 * f];
    {
        const double *cc = G.cell_center + 3 * c;
        for (int i = 0; i < 3; i++) P[i] = P[i] + (cc[i] - P[i]) / 1000.0;  // into the cell
    }
    double KSI = -log(1.0 - mk_rand(rs)), I_do = 0.0;
    const double mu = K.mu[q];
    int sort = ZONE ? (int)K.in_sort[q] : 4;
    int slot = (sort >= 1 && sort <= 4) ? sort : 0;  // tally slot: the sort, or 0 for one outside 1..4
    bool capped = false;
    unsigned long long n_ex = 0, n_seg = 0;
    int outcome = 4;  // 2 left, 4 lost, 8 ended on the zone's border
    const double me = 1.0 / 1836.15;
    // An exchange that falls due is held (pend) until a quarter of the wavefront's live lanes hold
    // one, or all of them do; the lanes then draw their new velocities together, so the costly
    // exchange path runs once for many lanes instead of on nearly every segment.  Each lane keeps
    // its own random stream, so the particle's history does not change.
    bool pend = false;
    double cp_ex = 1.0;
    int64_t pass_face = -2;  // -2: no pass due; else the inner face crossed (-1: an interior face)
    int it = 0;
#pragma unroll 1
    while (true) {
        if (!pend) {
            if (it >= K.max_steps) break;
            it++;
            double time = 0.0;
            int jx = 0;
            bool found = false;
#pragma unroll 1
            for (int k1 = 0; k1 < 12 && !found; k1++) {
                found = mk_exit(K.pack, c, col, P, V, time, jx);
                if (!found) {
                    const double *cn = G.cell_center + 3 * c;
                    for (int i = 0; i < 3; i++) P[i] = (k1 < 10) ? P[i] + (cn[i] - P[i]) / 800.0 : cn[i];
                }
            }
            if (!found) break;
            n_seg++;
            const double *s = prim + 8 * c;
            // the thermal protons: Plasma_components_1 (Phys_param.cpp:641-660) at the cell's type,
            // as the flight calls it (Setka_MK.cpp:1944-1966)
            const double rho_He = K.scal ? K.scal[kScal * c + 1] : 0.0;
            const int al = ((M.cell_code[c] >> 8) & 15u) <= 2u ? 2 : 1;
            double rho_Th = -(me * al * rho_He + 4.0 * (-s[0] + rho_He)) / (4.0 * (1.0 + me));
            double p_Th = s[4] * (4.0 * s[0] - (4.0 + al * me) * rho_He) /
                          (8.0 * s[0] + (-7.0 + al + (1.0 - al) * me) * rho_He);
            if (rho_Th <= 1e-8) rho_Th = 1e-8;
            if (p_Th <= 1e-8 / 2.0) p_Th = 1e-8 / 2.0;
            const double ro = rho_Th, cp = sqrt(2.0 * p_Th / rho_Th);
            const double l =
                sqrt((time * V[0]) * (time * V[0]) + (time * V[1]) * (time * V[1]) + (time * V[2]) * (time * V[2]));
            double I = I_do;
            const double u =
                sqrt((V[0] - s[1]) * (V[0] - s[1]) + (V[1] - s[2]) * (V[1] - s[2]) + (V[2] - s[3]) * (V[2] - s[3]));
            const double vn = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
            const double uz = mk_velosity_1(u, cp);
            const double nu_ex = ro * uz * mk_sigma(K.a2, uz) / K.kn;
            const double sig = vn / nu_ex;
            I += l / sig;
            if (I < KSI) {
                I_do = I;
                atomicAdd(&K.tally[5 * c + slot], time * mu);
                if (PUI) mk_pui_add(K, M, prim, c, sort, u, nu_ex, mu, time);
                for (int i = 0; i < 3; i++) P[i] += 1.000001 * time * V[i];
                const int32_t nb = K.pack[c].nbr[jx];
                if (nb < 0 && (-1 - nb) != 1) {  // out through the outer boundary
                    outcome = 2;
                    break;
                }
                if (K.R0 > 0.0 && sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) < 1.01 * K.R0) {
                    pass_face = nb < 0 ? M.cell_faces[6 * c + jx] : -1;  // the pass follows below
                } else if (nb < 0) {  // an inner face without the sphere: nowhere to go (lost)
                    break;
                } else if (ZONE) {
                    const int64_t fx = M.cell_faces[6 * c + jx];
                    if ((K.face_mask[fx] >> (K.zone - 1)) & 1u) {
                        // the zone's border (Setka_MK.cpp:2165-2184): into the face's outgoing
                        // distribution, on the side the particle leaves by
                        const int nn = K.cell_mk[M.face_cells[2 * fx]] == K.zone ? 0 : 1;
                        const int32_t ds = K.dslot[fx];
                        if (slot > 0 && ds >= 0) {
                            const double nf[3] = {G.face_normal[3 * fx], G.face_normal[3 * fx + 1],
                                                  G.face_normal[3 * fx + 2]};
                            const int64_t nbins = 108 * (int64_t)K.r * K.r * K.r;
                            mk_deposit(K.dist + (((int64_t)ds * 4 + (slot - 1)) * 2 + nn) * nbins, K.r, nf, nn, V, mu);
                        }
                        outcome = 8;
                        break;
                    }
                    if (K.cell_mk[nb] != K.zone) break;  // into another zone through a face no list holds: lost
                    c = nb;
                } else {
                    c = nb;
                }
            } else {
                const double kk = (KSI - I_do) * sig;
                const double t_ex = kk / vn;
                I_do = 0.0;
                for (int i = 0; i < 3; i++) P[i] += t_ex * V[i];
                atomicAdd(&K.tally[5 * c + slot], t_ex * mu);
                if (PUI) mk_pui_add(K, M, prim, c, sort, u, nu_ex, mu, t_ex);
                cp_ex = cp;
                pend = true;
            }
        }
        if (pass_face != -2) {
            // through the Sun's sphere to its far side, on into the cell there (its own block, so
            // that its temporaries do not add to the segment's)
            double t_;
            if (mk_sphere_time(P, V, 1.01 * K.R0, t_)) {
                for (int i = 0; i < 3; i++) P[i] += t_ * V[i];
            } else {
                const double nv = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
                for (int i = 0; i < 3; i++) P[i] += (2.02 * K.R0) * V[i] / nv;
            }
            n_pass++;
            const int cn = mk_locate_inner(M, K, X, pass_face >= 0 ? mk_inner_index(K, pass_face) : 0, P);
            pass_face = -2;
            if (cn < 0) break;  // lost
            c = cn;
        }
        const unsigned long long held = __ballot(pend), live = __ballot(true);
        if (pend && (4 * __popcll(held) >= __popcll(live) || held == live)) {
            const double *s = prim + 8 * c;
            const double cp = cp_ex;
            double U[3], Vs[3], W[3];
            const double Vp[3] = {s[1], s[2], s[3]};
            const MkFrame F = mk_frame(P);
            mk_to_sph(F, Vp, U[0], U[1], U[2]);
            mk_to_sph(F, V, Vs[0], Vs[1], Vs[2]);
            for (int i = 0; i < 3; i++) { U[i] /= cp; Vs[i] /= cp; }
            mk_change_velosity(rs, K.a2, U, Vs, W, cp, capped);
            for (int i = 0; i < 3; i++) W[i] *= cp;
            mk_from_sph(F, W[0], W[1], W[2], V);
            // P.sort = determ_zone(P.cel, 0) (Setka_MK.cpp:2088, 2121): Q routes it when the scalars
            // are carried; without them the cell's type stands in (what Q gives where the marker
            // follows the types, Q/rho >= 50 exactly past the HP)
            sort = (int)((M.cell_code[c] >> 8) & 15u);
            if (K.scal) sort = bera_zone(K.scal[kScal * c], prim[8 * c], sort);
            slot = (sort >= 1 && sort <= 4) ? sort : 0;
            KSI = -log(1.0 - mk_rand(rs));
            n_ex++;
            pend = false;
        }
    }
    atomicAdd(&st[0], 1ull);
    if (n_ex) atomicAdd(&st[1], n_ex);
    if (n_pass) atomicAdd(&st[3], n_pass);
    atomicAdd(&st[6], n_seg);
    atomicAdd(&st[outcome], 1ull);
    if (capped) atomicAdd(&st[7], 1ull);
}

void launch_mk_pack(hipStream_t s, const DevMesh &M, const double *X, MkCell *pack) {
    if (M.n_cells > 0) hipLaunchKernelGGL(k_mk_pack, dim3(grid_of(M.n_cells, 256)), dim3(256), 0, s, M, X, pack);
}

void launch_mk_gather(hipStream_t s, int64_t n, const int32_t *faces, const DevGeom &G, double *na) {
    if (n > 0) hipLaunchKernelGGL(k_mk_gather, dim3(grid_of(n, 256)), dim3(256), 0, s, n, faces, G, na);
}

void launch_mk_particles(hipStream_t s, const DevMesh &M, const MkDev &K, const double *X, const double *prim,
                         const DevGeom &G) {
    if (K.n_particles > 0) {
        const dim3 grid(grid_of(K.n_particles, kMkThreads)), block(kMkThreads);
        const bool pui = K.pui_nw > 0;
        if (K.zone > 0 && pui)
            hipLaunchKernelGGL((k_mk_particles<true, true>), grid, block, 0, s, M, K, X, prim, G);
        else if (K.zone > 0)
            hipLaunchKernelGGL((k_mk_particles<true, false>), grid, block, 0, s, M, K, X, prim, G);
        else if (pui)
            hipLaunchKernelGGL((k_mk_particles<false, true>), grid, block, 0, s, M, K, X, prim, G);
        else
            hipLaunchKernelGGL((k_mk_particles<false, false>), grid, block, 0, s, M, K, X, prim, G);
    }
    if (M.n_cells > 0)
        hipLaunchKernelGGL(k_mk_finish, dim3(grid_of(5 * M.n_cells, 256)), dim3(256), 0, s, M.n_cells,
                           (const double *)G.cell_volume, M.cell_ref, (const double *)K.tally, K.out);
}

void launch_mk_pui_finish(hipStream_t s, int64_t n, int32_t nw, double wr, double a2, double kn, const double *vol,
                          const double *trig, double *tt, double *sm, double *sp) {
    if (n > 0 && nw > 0) {
        const int64_t nt = (int64_t)nw * nw;
        hipLaunchKernelGGL(k_mk_pui_table, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nw, wr, a2, trig, tt);
        const int64_t blocks = n < 65536 ? n : 65536;
        hipLaunchKernelGGL(k_mk_pui_finish, dim3((unsigned)blocks), dim3(nw < 256 ? 64 * ((nw + 63) / 64) : 256),
                           sizeof(double) * nw, s, n, nw, wr, kn, vol, (const double *)tt, sm, sp);
    }
}

void launch_mkz_cdf(hipStream_t s, int64_t n, const int64_t *off, const double *dist, int64_t nb, double *cdf) {
    if (n > 0) hipLaunchKernelGGL(k_mkz_cdf, dim3(grid_of(n, 64)), dim3(64), 0, s, n, off, dist, nb, cdf);
}

void launch_mkz_clear(hipStream_t s, int64_t n, const int64_t *off, double *dist, int64_t nb) {
    if (n > 0) hipLaunchKernelGGL(k_mkz_clear, dim3(grid_of(n * nb, 256)), dim3(256), 0, s, n, off, dist, nb);
}

}  // namespace titan
