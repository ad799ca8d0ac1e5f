// titan_fit.cpp -- host set-up of the fitted surfaces (TS, HP, BS) and of the A_Luch rays that
// carry the other nodes: the surface lists of Setka::New_append_surfaces, the surface neighbours
// of each face (macros4, Setka.cpp:58-73), the node neighbours and sphere centres of
// Smooth_head_TS3 / Smooth_head_HP3 / Find_Yzel_Sosed_for_BS (Setka_move.cpp:2046-2397), and the
// rays' angles and indices for Luch::dvigenie (Luch.cpp:6-152).  The per-step arithmetic runs in
// k_fit.hip.
#include "titan_fit.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace titan {

namespace {

constexpr double kConstPi = 3.14159265358979323846;  // const_pi, Header.h:91

// polar_angle, Help_funk.cpp:78-107
double polar_angle_h(double x, double y) {
    if (std::fabs(x) + std::fabs(y) < 0.000001) return 0.0;
    if (x < 0) return std::atan(y / x) + 1.0 * kConstPi;
    if (x > 0 && y >= 0) return std::atan(y / x);
    if (x > 0 && y < 0) return std::atan(y / x) + 2.0 * kConstPi;
    if (y > 0 && x >= 0 && x <= 0) return kConstPi / 2.0;
    if (y < 0 && x >= 0 && x <= 0) return 3.0 * kConstPi / 2.0;
    return 0.0;
}

// cos(the), sin(the), cos(phi), sin(phi) of a ray from one glibc sincos call per angle: glibc's
// sincos and cos differ in the last bit for some arguments (cos(1.3646395517420515)), and whether a
// compiler turns cos(a) beside sin(a) into sincos is its own choice, so the call is made explicit.
void ray_trig(const double *the_phi, double *t) {
    ::sincos(the_phi[0], &t[1], &t[0]);
    ::sincos(the_phi[1], &t[3], &t[2]);
}

template <typename T>
int upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    if (*dst) { hipFree(*dst); *dst = nullptr; }
    if (hipMalloc((void **)dst, std::max<size_t>(1, src.size()) * sizeof(T)) != hipSuccess) return TITAN_E_HIP;
    if (!src.empty() &&
        hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
        return TITAN_E_HIP;
    return TITAN_OK;
}

// Minimum-norm least squares min |A x - b| for A [n x 3] (row-major) through a one-sided Jacobi
// SVD; singular values below 3 eps times the largest count as zero (the rank rule of Eigen's
// SVD solve, which Smooth_head_TS3 calls through bdcSvd).
void lsq3(int64_t n, const std::vector<double> &A, const std::vector<double> &b, double x[3]) {
    std::vector<double> U(A);  // columns rotated in place
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 60; sweep++) {
        bool rotated = false;
        for (int p = 0; p < 2; p++) {
            for (int q = p + 1; q < 3; q++) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int64_t i = 0; i < n; i++) {
                    const double up = U[3 * i + p], uq = U[3 * i + q];
                    al += up * up; be += uq * uq; ga += up * uq;
                }
                if (ga == 0.0 || std::fabs(ga) <= eps * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int64_t i = 0; i < n; i++) {
                    const double up = U[3 * i + p], uq = U[3 * i + q];
                    U[3 * i + p] = c * up - s * uq;
                    U[3 * i + q] = s * up + c * uq;
                }
                for (int k = 0; k < 3; k++) {
                    const double vp = V[k][p], vq = V[k][q];
                    V[k][p] = c * vp - s * vq;
                    V[k][q] = s * vp + c * vq;
                }
            }
        }
        if (!rotated) break;
    }
    double sig[3], smax = 0.0;
    for (int k = 0; k < 3; k++) {
        double s2 = 0.0;
        for (int64_t i = 0; i < n; i++) s2 += U[3 * i + k] * U[3 * i + k];
        sig[k] = std::sqrt(s2);
        smax = std::max(smax, sig[k]);
    }
    const double thr = std::max(smax * (3.0 * eps), std::numeric_limits<double>::min());
    x[0] = x[1] = x[2] = 0.0;
    for (int k = 0; k < 3; k++) {
        if (!(sig[k] >= thr) || sig[k] == 0.0) continue;
        double ub = 0.0;
        for (int64_t i = 0; i < n; i++) ub += U[3 * i + k] * b[i];
        const double coef = ub / (sig[k] * sig[k]);  // (u_k . b) / sigma_k with u_k = U_k / sigma_k
        for (int j = 0; j < 3; j++) x[j] += coef * V[j][k];
    }
}

}  // namespace

// Setka_move.cpp:2257-2355: least squares for 2 p . c = |p|^2 over the points, then up to 8
// reweighted passes (Tukey biweights of |dist - mean dist| with a cutoff of 2.5 standard
// deviations) until the rms residual changes by less than 1e-6.
int sphere_fit(int64_t n, const double *pts, double centre[3]) {
    if (n < 3 || !pts || !centre) return TITAN_E_ARG;
    std::vector<double> A(3 * n), b(n), WA(3 * n), Wb(n), w(n, 1.0), res(n);
    for (int64_t i = 0; i < n; i++) {
        const double *p = pts + 3 * i;
        for (int k = 0; k < 3; k++) A[3 * i + k] = 2.0 * p[k];
        b[i] = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    }
    double c[3];
    lsq3(n, A, b, c);
    double prev = std::numeric_limits<double>::max();
    for (int iter = 0; iter < 8; iter++) {
        double radius = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const double *p = pts + 3 * i;
            const double d0 = p[0] - c[0], d1 = p[1] - c[1], d2 = p[2] - c[2];
            radius += std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        }
        radius /= (double)n;
        double total = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const double *p = pts + 3 * i;
            const double d0 = p[0] - c[0], d1 = p[1] - c[1], d2 = p[2] - c[2];
            res[i] = std::fabs(std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) - radius);
            total += res[i] * res[i];
        }
        const double rms = std::sqrt(total / (double)n);
        if (std::fabs(prev - rms) < 1e-6) break;
        prev = rms;
        double mean = 0.0;
        for (int64_t i = 0; i < n; i++) mean += res[i];
        mean /= (double)n;
        double var = 0.0;
        for (int64_t i = 0; i < n; i++) var += (res[i] - mean) * (res[i] - mean);
        const double cutoff = 2.5 * std::sqrt(var / (double)n);
        for (int64_t i = 0; i < n; i++) {
            if (res[i] < cutoff) {
                const double t = res[i] / cutoff;
                w[i] = (1.0 - t * t) * (1.0 - t * t);
            } else {
                w[i] = 0.0;
            }
            for (int k = 0; k < 3; k++) WA[3 * i + k] = w[i] * A[3 * i + k];
            Wb[i] = w[i] * b[i];
        }
        lsq3(n, WA, Wb, c);
    }
    centre[0] = c[0]; centre[1] = c[1]; centre[2] = c[2];
    return TITAN_OK;
}

struct FitState {
    int64_t nn = 0, nf = 0;
    // surface faces in ascending reference id: kind 1 TS, 2 HP, 3 BS
    std::vector<int32_t> sf_ref, sf_dev;
    std::vector<uint8_t> sf_kind;
    // surface nodes in ascending reference id
    std::vector<int32_t> sn_ref, sn_dev, node_sn;
    std::vector<uint8_t> sn_kind;
    std::vector<int32_t> sn_face_off, sn_face;  // faces that move the node (indices into sf_*)
    std::vector<int32_t> cand_off, cand;        // the node and the nodes of its faces of its kind
    std::vector<std::vector<int32_t>> nb;       // smoothing neighbours (reference ids)
    std::vector<double> centre;
    int n_kind[4] = {0, 0, 0, 0};
    bool smoothed[4] = {false, false, false, false};
    bool bad = false;  // a face with two surface flags, or a node on two kinds of surface
    // rays
    bool have_rays = false;
    int64_t n_rays = 0;
    int32_t npr = 0;
    double R0 = 0, R1 = 0, R5 = 0, da[5] = {0, 0, 0, 0, 0};
    int32_t M[6] = {0, 0, 0, 0, 0, 0};
    double phi_lim = 0.0;
    std::vector<int32_t> ray_dev;
    std::vector<double> ray_trig;
    std::vector<int2> node_ray;  // device node -> (ray, index), ray -1 off the rays
    // the other ray families
    RayExtDev ext{};             // host copy of the scalars; pointers filled at upload
    std::vector<int8_t> ext_type;
    std::vector<int64_t> ext_off;
    std::vector<int32_t> ext_nodes, ext_opor, ext_order;
    std::vector<double> ext_trig;
    bool up_ext = false;
    // HP grid smoothing and the corrections after the move
    std::vector<double> cc;             // [ns*2] circle centres, NaN: none
    bool have_cc = false, have_grid = false;
    std::vector<int32_t> grid;          // [ns*8] reference ids, -1: none
    double hp_k[4] = {0, 0, 0, 0};      // k_angle, k, angle, along
    int32_t n_tail = 0, tail_len = 0, N4 = 0;
    std::vector<int32_t> tail;          // reference ids
    std::vector<int32_t> bs_ref, bs_off, bs_nodes;
    bool up_grid = false;
    double *d_cc = nullptr;
    int32_t *d_grid = nullptr, *d_tail = nullptr, *d_bs_ref = nullptr, *d_bs_off = nullptr, *d_bs_nodes = nullptr;
    int32_t *d_sn_ref = nullptr;
    int8_t *d_ext_type = nullptr;
    int64_t *d_ext_off = nullptr;
    int32_t *d_ext_nodes = nullptr, *d_ext_opor = nullptr, *d_ext_order = nullptr;
    double *d_ext_trig = nullptr;
    // device copies
    bool up_static = false, up_nb = false, up_rays = false;
    int32_t *d_sf_face = nullptr, *d_sn_node = nullptr, *d_sn_face_off = nullptr, *d_sn_face = nullptr;
    int32_t *d_nb_off = nullptr, *d_nb = nullptr, *d_ray_nodes = nullptr;
    int2 *d_node_ray = nullptr;
    double *d_ray_R = nullptr;
    uint8_t *d_sf_kind = nullptr, *d_sn_kind = nullptr;
    double *d_sf_vel = nullptr, *d_sn_vel = nullptr, *d_centre = nullptr, *d_ray_trig = nullptr;
    const DeviceOrder *ord = nullptr;
};

int fit_create(FitState **out, int64_t nn, int64_t nf, const int32_t *face_nodes, const int32_t *face_cells,
               const int32_t *cell_faces, const uint8_t *face_flags, const DeviceOrder &o) {
    FitState *F = new FitState();
    F->nn = nn;
    F->nf = nf;
    F->ord = &o;
    std::vector<uint8_t> kind(nf, 0);
    if (face_flags) {
        for (int64_t f = 0; f < nf; f++) {
            const uint8_t fl = face_flags[f];
            const int ts = (fl >> 3) & 1, hp = (fl >> 2) & 1, bs = (fl >> 4) & 1;
            if (ts + hp + bs > 1) { F->bad = true; continue; }
            // a fitted surface lies between two cells: boundary faces with a surface flag are not
            // moved (the flag still selects the TVD rule of Snos_on_HP)
            if (face_cells[2 * f + 1] >= 0) kind[f] = ts ? 1 : (hp ? 2 : (bs ? 3 : 0));
        }
    }
    std::vector<int32_t> face_sf(nf, -1);
    for (int64_t f = 0; f < nf; f++) {
        if (!kind[f]) continue;
        face_sf[f] = (int32_t)F->sf_ref.size();
        F->sf_ref.push_back((int32_t)f);
        F->sf_dev.push_back(o.face_r2d[f]);
        F->sf_kind.push_back(kind[f]);
    }
    // HP faces with fewer than 4 surface neighbours add nothing (Setka_move.cpp:197); the
    // neighbours are the same-surface faces of the cells beside cells[0] (macros4)
    std::vector<uint8_t> moves(F->sf_ref.size(), 1);
    for (size_t i = 0; i < F->sf_ref.size(); i++) {
        if (F->sf_kind[i] != 2) continue;
        const int64_t f = F->sf_ref[i];
        const int64_t A = face_cells[2 * f];
        int cnt = 0;
        for (int j = 0; j < 6; j++) {
            const int64_t g = cell_faces[6 * A + j];
            if (g == f) continue;
            const int64_t B = face_cells[2 * g] == A ? face_cells[2 * g + 1] : face_cells[2 * g];
            if (B < 0) continue;
            for (int k = 0; k < 6; k++) {
                const int64_t h = cell_faces[6 * B + k];
                if (kind[h] == 2 && h != f) cnt++;
            }
        }
        if (cnt < 4) moves[i] = 0;
    }
    std::vector<uint8_t> nkind(nn, 0);
    for (size_t i = 0; i < F->sf_ref.size(); i++) {
        const int64_t f = F->sf_ref[i];
        for (int k = 0; k < 4; k++) {
            const int32_t v = face_nodes[4 * f + k];
            if (nkind[v] && nkind[v] != F->sf_kind[i]) F->bad = true;
            nkind[v] = F->sf_kind[i];
        }
    }
    F->node_sn.assign(nn, -1);
    for (int64_t v = 0; v < nn; v++) {
        if (!nkind[v]) continue;
        F->node_sn[v] = (int32_t)F->sn_ref.size();
        F->sn_ref.push_back((int32_t)v);
        F->sn_dev.push_back(o.node_r2d[v]);
        F->sn_kind.push_back(nkind[v]);
        F->n_kind[nkind[v]]++;
    }
    const size_t ns = F->sn_ref.size();
    std::vector<std::vector<int32_t>> nf_list(ns), cand(ns);
    for (size_t i = 0; i < F->sf_ref.size(); i++) {  // ascending reference face id
        const int64_t f = F->sf_ref[i];
        for (int k = 0; k < 4; k++) {
            const int32_t s = F->node_sn[face_nodes[4 * f + k]];
            if (moves[i]) nf_list[s].push_back((int32_t)i);
            for (int q = 0; q < 4; q++) cand[s].push_back(face_nodes[4 * f + q]);
        }
    }
    F->sn_face_off.assign(1, 0);
    F->cand_off.assign(1, 0);
    for (size_t s = 0; s < ns; s++) {
        for (int32_t i : nf_list[s]) F->sn_face.push_back(i);
        F->sn_face_off.push_back((int32_t)F->sn_face.size());
        std::vector<int32_t> &c = cand[s];
        c.push_back(F->sn_ref[s]);
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        for (int32_t v : c) F->cand.push_back(v);
        F->cand_off.push_back((int32_t)F->cand.size());
    }
    F->nb.assign(ns, {});
    F->centre.assign(3 * ns, 0.0);
    F->cc.assign(2 * ns, std::numeric_limits<double>::quiet_NaN());
    F->grid.assign(8 * ns, -1);
    *out = F;
    return TITAN_OK;
}

void fit_destroy(FitState *F) {
    if (!F) return;
    void *p[] = {F->d_sf_face, F->d_sn_node, F->d_sn_face_off, F->d_sn_face, F->d_nb_off, F->d_nb, F->d_ray_nodes,
                 F->d_sf_kind, F->d_sn_kind, F->d_sf_vel, F->d_sn_vel, F->d_centre, F->d_ray_trig, F->d_node_ray,
                 F->d_ray_R, F->d_ext_type, F->d_ext_off, F->d_ext_nodes, F->d_ext_opor, F->d_ext_order, F->d_ext_trig,
                 F->d_cc, F->d_grid, F->d_tail, F->d_bs_ref, F->d_bs_off, F->d_bs_nodes, F->d_sn_ref};
    for (void *q : p)
        if (q) hipFree(q);
    delete F;
}

bool fit_has_rays(const FitState *F) { return F && F->have_rays; }

int fit_set_rays(FitState *F, const titan_rays *r, const DeviceOrder &o) {
    if (!F || !r || r->n_rays < 1 || !r->ray_nodes || !r->ray_angles) return TITAN_E_ARG;
    const int32_t M0 = r->M0, M1 = r->M1, M11 = r->M11, M2 = r->M2, M3 = r->M3, M4 = r->M4;
    if (M0 < 1 || M1 < 0 || M11 < 0 || M2 < M11 || M3 < 0 || M4 < 0) return TITAN_E_ARG;
    if (r->nodes_per_ray != M0 + M1 + M2 + M3 + M4 + 5) return TITAN_E_ARG;
    const int64_t tot = r->n_rays * (int64_t)r->nodes_per_ray;
    std::vector<uint8_t> seen(F->nn, 0);
    std::vector<int32_t> dev(tot);
    for (int64_t i = 0; i < tot; i++) {
        const int32_t v = r->ray_nodes[i];
        if (v < 0 || v >= F->nn || seen[v]) return TITAN_E_ARG;  // each node on one ray
        seen[v] = 1;
        dev[i] = o.node_r2d[v];
    }
    // ray-major on the device: node j of ray k at [k*npr + j]; and each device node's (ray, j)
    F->node_ray.assign(F->nn, make_int2(-1, -1));
    for (int64_t i = 0; i < tot; i++)
        F->node_ray[dev[i]] = make_int2((int)(i / r->nodes_per_ray), (int)(i % r->nodes_per_ray));
    F->ray_dev.swap(dev);
    F->ray_trig.resize(4 * r->n_rays);
    for (int64_t k = 0; k < r->n_rays; k++) ray_trig(r->ray_angles + 2 * k, &F->ray_trig[4 * k]);
    F->n_rays = r->n_rays;
    F->npr = r->nodes_per_ray;
    F->R0 = r->R0; F->R1 = r->R1; F->R5 = r->R5;
    F->da[0] = r->da1; F->da[1] = r->da2; F->da[2] = r->da3; F->da[3] = r->da4; F->da[4] = r->da5;
    F->M[0] = M0; F->M[1] = M1; F->M[2] = M11; F->M[3] = M2; F->M[4] = M3; F->M[5] = M4;
    F->phi_lim = r->phi_a + kConstPi / 180.0;
    F->have_rays = true;
    F->up_rays = false;
    F->ext = RayExtDev{};  // the other families were checked against the old M0..M4
    return TITAN_OK;
}

namespace {
// Indices on a ray of type t (1 B .. 6 G) that Luch::dvigenie writes (Luch.cpp:153-739), and the
// node count it needs.
int ext_written(int t, const int32_t M[6], int32_t N4, int32_t N5, int32_t MF, std::vector<int> &w) {
    const int M0 = M[0], M1 = M[1], M11 = M[2], M2 = M[3], M3 = M[4], M4 = M[5];
    w.clear();
    auto run = [&](int a, int n) { for (int j = 0; j < n; j++) w.push_back(a + j); };
    switch (t) {
    case 1:
        run(0, M0 + 1);
        run(M0 + 1, M1);
        run(M0 + M1 + 2, M11);
        run(M0 + M1 + M11 + 2, M2 - M11);
        run(M0 + M1 + M2 + 3, M3);
        run(M0 + M1 + M2 + M3 + 4, M4 + 1);
        return M0 + M1 + M2 + M3 + M4 + 5;
    case 2:
        run(0, M0);
        run(M0 + 1, M1);
        run(M0 + M1 + 2, M11 + N4 + N5);
        return M0 + M1 + M11 + N4 + N5 + 2;
    case 3:
        run(1, M2 - M11);
        run(M2 - M11 + 2, M3);
        run(M2 - M11 + M3 + 3, M4 + 1);
        return M2 - M11 + M3 + M4 + 4;
    case 4:
        run(0, M3 + M4 + 3);
        return M3 + M4 + 3;
    case 5:
        run(1, 4);
        return 5;
    case 6:
        run(1, M2 - M11 - MF + 1);
        return M2 - M11 - MF + 2;
    }
    return -1;
}
const int kExtSlots[7] = {0, 0xE, 0x2, 0x7, 0x7, 0xF, 0x7};  // support slots each type reads
}  // namespace

int fit_set_rays_ext(FitState *F, const titan_rays_ext *r, const DeviceOrder &o) {
    if (!F || !r || !F->have_rays || r->n_rays < 0) return TITAN_E_ARG;
    const int64_t n = r->n_rays;
    if (n > 0 && (!r->ray_type || !r->ray_off || !r->ray_nodes || !r->ray_opor || !r->ray_angles)) return TITAN_E_ARG;
    if (n > 0 && r->ray_off[0] != 0) return TITAN_E_ARG;
    const int32_t *M = F->M;
    if (r->N4 < 0 || r->N5 < 0 || r->MF < 0) return TITAN_E_ARG;
    std::vector<int> w;
    std::vector<int64_t> fam_n(8, 0);
    for (int64_t k = 0; k < n; k++) {
        const int t = r->ray_type[k];
        if (t < 1 || t > 6) return TITAN_E_ARG;
        if (t == 2 && r->N5 < 2) return TITAN_E_ARG;  // the tail spacing divides by N5 - 1
        if (t == 6 && M[3] - M[2] - r->MF + 1 < 0) return TITAN_E_ARG;
        const int need = ext_written(t, M, r->N4, r->N5, r->MF, w);
        const int64_t len = r->ray_off[k + 1] - r->ray_off[k];
        if (len < need || len < 1) return TITAN_E_ARG;
        for (int64_t i = r->ray_off[k]; i < r->ray_off[k + 1]; i++)
            if (r->ray_nodes[i] < 0 || r->ray_nodes[i] >= F->nn) return TITAN_E_ARG;
        for (int q = 0; q < 4; q++)
            if (((kExtSlots[t] >> q) & 1) && (r->ray_opor[4 * k + q] < 0 || r->ray_opor[4 * k + q] >= F->nn))
                return TITAN_E_ARG;
        fam_n[t]++;
    }
    RayExtDev &E = F->ext;
    E = RayExtDev{};
    E.n = n;
    E.fam_off[0] = E.fam_off[1] = 0;
    for (int f = 1; f <= 6; f++) E.fam_off[f + 1] = E.fam_off[f] + fam_n[f];
    F->ext_order.assign(n, 0);
    std::vector<int64_t> fill(E.fam_off, E.fam_off + 8);
    for (int64_t k = 0; k < n; k++) F->ext_order[fill[r->ray_type[k]]++] = (int32_t)k;
    // a family runs ray after ray when one of its rays reads or writes a node another writes
    std::vector<int64_t> writer(F->nn, -1);
    for (int f = 1; f <= 6; f++) {
        bool serial = false;
        std::vector<int32_t> touched;
        for (int64_t i = E.fam_off[f]; i < E.fam_off[f + 1] && !serial; i++) {
            const int64_t k = F->ext_order[i];
            ext_written(f, M, r->N4, r->N5, r->MF, w);
            std::vector<int32_t> wr;
            for (int j : w) wr.push_back(r->ray_nodes[r->ray_off[k] + j]);
            if (f == 1) { wr.push_back(r->ray_opor[4 * k + 2]); wr.push_back(r->ray_opor[4 * k + 3]); }
            if (f == 3) { wr.push_back(r->ray_opor[4 * k + 1]); wr.push_back(r->ray_opor[4 * k + 2]); }
            for (int32_t v : wr) {
                if (writer[v] >= 0 && writer[v] != k) serial = true;
                writer[v] = k;
                touched.push_back(v);
            }
        }
        for (int64_t i = E.fam_off[f]; i < E.fam_off[f + 1] && !serial; i++) {
            const int64_t k = F->ext_order[i];
            for (int q = 0; q < 4; q++) {
                if (!((kExtSlots[f] >> q) & 1)) continue;
                const int64_t wk = writer[r->ray_opor[4 * k + q]];
                if (wk >= 0 && wk != k) serial = true;
            }
        }
        if (serial) E.serial |= 1 << f;
        for (int32_t v : touched) writer[v] = -1;
    }
    F->ext_type.assign(r->ray_type, r->ray_type + n);
    F->ext_off.assign(r->ray_off, r->ray_off + n + 1);
    const int64_t tot = n > 0 ? r->ray_off[n] : 0;
    F->ext_nodes.resize(tot);
    for (int64_t i = 0; i < tot; i++) F->ext_nodes[i] = o.node_r2d[r->ray_nodes[i]];
    F->ext_opor.resize(4 * n);
    for (int64_t i = 0; i < 4 * n; i++) F->ext_opor[i] = r->ray_opor[i] >= 0 ? o.node_r2d[r->ray_opor[i]] : -1;
    F->ext_trig.resize(4 * n);
    for (int64_t k = 0; k < n; k++) ray_trig(r->ray_angles + 2 * k, &F->ext_trig[4 * k]);
    E.N4 = r->N4; E.N5 = r->N5; E.MF = r->MF;
    E.L6 = r->L6; E.L7 = r->L7;
    E.dd1 = r->dd1; E.dd2 = r->dd2; E.dd3 = r->dd3; E.dd4 = r->dd4;
    E.dd5i = (double)(int)r->dd5; E.dd6i = (double)(int)r->dd6;  // `int dd5 = geo->dd5` (Luch.cpp:166-167)
    E.dd7 = r->dd7; E.dd8 = r->dd8;
    E.ba1 = r->ba1; E.ba2 = r->ba2; E.ba3 = r->ba3; E.ba4 = r->ba4; E.ba5 = r->ba5;
    E.md1 = r->md1; E.md2 = r->md2; E.md3 = r->md3; E.ea1 = r->ea1; E.ea2 = r->ea2; E.ea3 = r->ea3;
    F->up_ext = false;
    return TITAN_OK;
}

// 3x3 solve by elimination with partial pivoting; false when singular.
static bool solve3(double H[3][3], const double g[3], double d[3]) {
    double M[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i][j] = H[i][j];
        M[i][3] = g[i];
    }
    for (int c = 0; c < 3; c++) {
        int piv = c;
        for (int r = c + 1; r < 3; r++)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 4; j++) std::swap(M[c][j], M[piv][j]);
        for (int r = c + 1; r < 3; r++) {
            const double f = M[r][c] / M[c][c];
            for (int j = c; j < 4; j++) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double t = M[i][3];
        for (int j = i + 1; j < 3; j++) t -= M[i][j] * d[j];
        d[i] = t / M[i][i];
    }
    return true;
}

// fitCircleRobust with max_iter 8, tol 1e-8 (Setka_move.cpp:2544-2620): Levenberg-Marquardt on the
// centre and radius, from the origin and the points' mean distance from it.  The reference sets
// prev_error to the new error before comparing the two, so the damping grows tenfold every step; a
// step whose update is not finite only raises the damping.
int circle_fit(int64_t n, const double *pts, double *out) {
    if (n < 3 || !pts || !out) return TITAN_E_ARG;
    double P[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; i++) P[2] += std::sqrt(pts[2 * i] * pts[2 * i] + pts[2 * i + 1] * pts[2 * i + 1]);
    P[2] /= (double)n;
    double lambda = 0.001, prev = std::numeric_limits<double>::max();
    for (int it = 0; it < 8; it++) {
        double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, g[3] = {0, 0, 0}, err = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const double dx = pts[2 * i] - P[0], dy = pts[2 * i + 1] - P[1];
            const double dist = std::sqrt(dx * dx + dy * dy);
            double res, J[3] = {0.0, 0.0, -1.0};
            if (dist < 1e-10) {
                res = -P[2];
            } else {
                res = dist - P[2];
                J[0] = -dx / dist;
                J[1] = -dy / dist;
            }
            err += res * res;
            for (int a = 0; a < 3; a++) {
                g[a] -= J[a] * res;
                for (int b = 0; b < 3; b++) H[a][b] += J[a] * J[b];
            }
        }
        err = std::sqrt(err / (double)n);
        if (std::fabs(prev - err) < 1e-8) break;
        prev = err;
        for (int a = 0; a < 3; a++) H[a][a] *= 1.0 + lambda;
        double d[3];
        if (!solve3(H, g, d) || !std::isfinite(d[0]) || !std::isfinite(d[1]) || !std::isfinite(d[2])) {
            lambda *= 10.0;
            continue;
        }
        for (int a = 0; a < 3; a++) P[a] += d[a];
        lambda = err < prev ? lambda / 2.0 : lambda * 10.0;
    }
    out[0] = P[0];
    out[1] = P[1];
    out[2] = P[2];
    return TITAN_OK;
}

int fit_smooth_angle(FitState *F, int64_t n, const int32_t *nodes, const int32_t *nb6, const double *x0) {
    if (!F || n < 0 || (n > 0 && (!nodes || !nb6)) || !x0) return TITAN_E_ARG;
    for (int64_t k = 0; k < n; k++) {
        if (nodes[k] < 0 || nodes[k] >= F->nn) return TITAN_E_ARG;
        const int32_t s = F->node_sn[nodes[k]];
        if (s < 0 || F->sn_kind[s] != 2) return TITAN_E_ARG;
        for (int q = 0; q < 6; q++)
            if (nb6[6 * k + q] < 0 || nb6[6 * k + q] >= F->nn) return TITAN_E_ARG;
    }
    for (int64_t k = 0; k < n; k++) {
        const int32_t s = F->node_sn[nodes[k]];
        double pts[14], cr[3];
        pts[0] = x0[3 * (int64_t)nodes[k] + 1];
        pts[1] = x0[3 * (int64_t)nodes[k] + 2];
        for (int q = 0; q < 6; q++) {
            pts[2 + 2 * q] = x0[3 * (int64_t)nb6[6 * k + q] + 1];
            pts[3 + 2 * q] = x0[3 * (int64_t)nb6[6 * k + q] + 2];
        }
        circle_fit(7, pts, cr);
        F->cc[2 * s] = cr[0];
        F->cc[2 * s + 1] = cr[1];
        std::vector<int32_t> &nb = F->nb[s];  // a set: each neighbour once, ascending id
        for (int q = 0; q < 6; q++) nb.push_back(nb6[6 * k + q]);
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    }
    F->have_cc = F->have_cc || n > 0;
    F->up_nb = false;
    F->up_grid = false;
    return TITAN_OK;
}

int fit_set_hp_grid(FitState *F, const titan_hp_grid *g, const DeviceOrder &o) {
    if (!F || !g || g->n_grid < 0 || g->n_tail < 0 || g->n_bs < 0) return TITAN_E_ARG;
    if (g->n_grid > 0 && (!g->grid_node || !g->grid_nb)) return TITAN_E_ARG;
    if (g->n_tail > 0 && (!g->tail || g->N4 < 6 || g->tail_len < g->N4 - 3)) return TITAN_E_ARG;
    if (g->n_bs > 0 && (!g->bs_ref || !g->bs_off || !g->bs_nodes || g->bs_off[0] != 0)) return TITAN_E_ARG;
    auto kind_of = [&](int32_t v) { return (v < 0 || v >= F->nn || F->node_sn[v] < 0) ? 0 : F->sn_kind[F->node_sn[v]]; };
    std::vector<int32_t> grid(8 * F->sn_ref.size(), -1);
    for (int64_t k = 0; k < g->n_grid; k++) {
        if (kind_of(g->grid_node[k]) != 2) return TITAN_E_ARG;
        const int32_t s = F->node_sn[g->grid_node[k]];
        if (grid[8 * s] >= 0) return TITAN_E_ARG;
        for (int q = 0; q < 8; q++) {
            const int32_t v = g->grid_nb[8 * k + q];
            if (v < 0 || v >= F->nn) return TITAN_E_ARG;
            grid[8 * s + q] = v;
        }
    }
    const int64_t nt = (int64_t)g->n_tail * g->tail_len;
    for (int64_t i = 0; i < nt; i++)
        if (kind_of(g->tail[i]) != 2) return TITAN_E_ARG;
    for (int32_t a = 0; a < g->n_bs; a++) {
        if (kind_of(g->bs_ref[a]) != 3 || g->bs_off[a + 1] < g->bs_off[a]) return TITAN_E_ARG;
        for (int32_t j = g->bs_off[a]; j < g->bs_off[a + 1]; j++)
            if (kind_of(g->bs_nodes[j]) != 3) return TITAN_E_ARG;
    }
    F->grid.swap(grid);
    F->have_grid = g->n_grid > 0;
    F->hp_k[0] = g->sglag_HP_k_angle; F->hp_k[1] = g->sglag_HP_k;
    F->hp_k[2] = g->sglag_HP_angle; F->hp_k[3] = g->sglag_HP_along;
    F->n_tail = g->n_tail; F->tail_len = g->n_tail > 0 ? g->tail_len : 0; F->N4 = g->N4;
    F->tail.assign(g->tail, g->tail + (g->n_tail > 0 ? nt : 0));
    F->bs_ref.assign(g->bs_ref, g->bs_ref + g->n_bs);
    F->bs_off.assign(g->bs_off, g->bs_off + (g->n_bs > 0 ? g->n_bs + 1 : 0));
    F->bs_nodes.assign(g->bs_nodes, g->bs_nodes + (g->n_bs > 0 ? g->bs_off[g->n_bs] : 0));
    (void)o;
    F->up_grid = false;
    return TITAN_OK;
}

int fit_smooth(FitState *F, int kind, const double *x0) {
    if (!F || kind < 1 || kind > 3 || !x0) return TITAN_E_ARG;
    if (kind == 2 && !F->have_rays) return TITAN_E_ARG;  // phi_a comes with the rays
    std::vector<double> pts;
    for (size_t s = 0; s < F->sn_ref.size(); s++) {
        if (F->sn_kind[s] != kind) continue;
        const double *a = x0 + 3 * (int64_t)F->sn_ref[s];
        if (kind == 2) {
            const double phi = polar_angle_h(a[0], std::sqrt(0.0 * 0.0 + a[1] * a[1] + a[2] * a[2]));
            if (phi > F->phi_lim) continue;
        }
        const int32_t c0 = F->cand_off[s], c1 = F->cand_off[s + 1];
        if (c1 - c0 < 3) continue;
        F->nb[s].clear();
        pts.clear();
        for (int32_t j = c0; j < c1; j++) {
            const int32_t v = F->cand[j];
            if (v != F->sn_ref[s]) F->nb[s].push_back(v);
            for (int k = 0; k < 3; k++) pts.push_back(x0[3 * (int64_t)v + k]);
        }
        if (kind != 3) {
            const int rc = sphere_fit(c1 - c0, pts.data(), &F->centre[3 * s]);
            if (rc != TITAN_OK) return rc;
        }
    }
    F->smoothed[kind] = true;
    F->up_nb = false;
    return TITAN_OK;
}

int fit_prepare(FitState *F, const titan_params &p, hipStream_t s, FitDev *dev) {
    if (!F || !dev || F->bad) return TITAN_E_ARG;
    const int32_t sg[4] = {0, p.sglag_TS, p.sglag_HP, p.sglag_BS};
    for (int k = 1; k <= 3; k++)
        if (sg[k] && F->n_kind[k] > 0 && !F->smoothed[k]) return TITAN_E_ARG;
    int rc = TITAN_OK;
    auto up = [&](int r) { if (rc == TITAN_OK) rc = r; };
    if (!F->up_static) {
        up(upload(&F->d_sf_face, F->sf_dev, s));
        up(upload(&F->d_sf_kind, F->sf_kind, s));
        up(upload(&F->d_sn_node, F->sn_dev, s));
        up(upload(&F->d_sn_kind, F->sn_kind, s));
        up(upload(&F->d_sn_face_off, F->sn_face_off, s));
        up(upload(&F->d_sn_face, F->sn_face, s));
        up(upload(&F->d_sf_vel, std::vector<double>(3 * F->sf_ref.size(), 0.0), s));
        up(upload(&F->d_sn_vel, std::vector<double>(3 * F->sn_ref.size(), 0.0), s));
        if (rc != TITAN_OK) return rc;
        F->up_static = true;
    }
    if (!F->up_nb) {
        std::vector<int32_t> off(1, 0), nbd;
        for (size_t i = 0; i < F->nb.size(); i++) {
            for (int32_t v : F->nb[i]) nbd.push_back(F->ord->node_r2d[v]);
            off.push_back((int32_t)nbd.size());
        }
        up(upload(&F->d_nb_off, off, s));
        up(upload(&F->d_nb, nbd, s));
        up(upload(&F->d_centre, F->centre, s));
        if (rc != TITAN_OK) return rc;
        F->up_nb = true;
    }
    if (F->have_rays && !F->up_rays) {
        up(upload(&F->d_ray_nodes, F->ray_dev, s));
        up(upload(&F->d_ray_trig, F->ray_trig, s));
        up(upload(&F->d_node_ray, F->node_ray, s));
        up(upload(&F->d_ray_R, std::vector<double>(3 * F->n_rays, 0.0), s));
        if (rc != TITAN_OK) return rc;
        F->up_rays = true;
    }
    if (F->ext.n > 0 && !F->up_ext) {
        up(upload(&F->d_ext_type, F->ext_type, s));
        up(upload(&F->d_ext_off, F->ext_off, s));
        up(upload(&F->d_ext_nodes, F->ext_nodes, s));
        up(upload(&F->d_ext_opor, F->ext_opor, s));
        up(upload(&F->d_ext_order, F->ext_order, s));
        up(upload(&F->d_ext_trig, F->ext_trig, s));
        if (rc != TITAN_OK) return rc;
        F->up_ext = true;
    }
    if (!F->up_grid) {
        const DeviceOrder &o = *F->ord;
        auto dev_of = [&](const std::vector<int32_t> &v) {
            std::vector<int32_t> d(v.size());
            for (size_t i = 0; i < v.size(); i++) d[i] = v[i] >= 0 ? o.node_r2d[v[i]] : -1;
            return d;
        };
        up(upload(&F->d_cc, F->cc, s));
        up(upload(&F->d_grid, dev_of(F->grid), s));
        up(upload(&F->d_tail, dev_of(F->tail), s));
        up(upload(&F->d_bs_ref, dev_of(F->bs_ref), s));
        up(upload(&F->d_bs_off, F->bs_off, s));
        up(upload(&F->d_bs_nodes, dev_of(F->bs_nodes), s));
        up(upload(&F->d_sn_ref, F->sn_ref, s));
        if (rc != TITAN_OK) return rc;
        F->up_grid = true;
    }
    FitDev &D = *dev;
    D = FitDev{};
    D.n_sf = (int32_t)F->sf_ref.size();
    D.n_sn = (int32_t)F->sn_ref.size();
    D.sf_face = F->d_sf_face; D.sf_kind = F->d_sf_kind; D.sf_vel = F->d_sf_vel;
    D.sn_node = F->d_sn_node; D.sn_kind = F->d_sn_kind;
    D.sn_face_off = F->d_sn_face_off; D.sn_face = F->d_sn_face;
    D.sn_nb_off = F->d_nb_off; D.sn_nb = F->d_nb; D.sn_centre = F->d_centre; D.sn_vel = F->d_sn_vel;
    D.n_rays = F->have_rays ? F->n_rays : 0;
    D.npr = F->npr;
    D.ray_nodes = F->d_ray_nodes; D.ray_trig = F->d_ray_trig;
    D.n_nodes = F->nn; D.node_ray = F->d_node_ray; D.ray_R = F->d_ray_R;
    D.R0 = F->R0; D.R1 = F->R1; D.R5 = F->R5;
    D.da1 = F->da[0]; D.da2 = F->da[1]; D.da3 = F->da[2]; D.da4 = F->da[3]; D.da5 = F->da[4];
    D.M0 = F->M[0]; D.M1 = F->M[1]; D.M11 = F->M[2]; D.M2 = F->M[3]; D.M3 = F->M[4]; D.M4 = F->M[5];
    D.move[1] = p.move_TS; D.move[2] = p.move_HP; D.move[3] = p.move_BS;
    D.sglag[1] = p.sglag_TS; D.sglag[2] = p.sglag_HP; D.sglag[3] = p.sglag_BS;
    D.velocity[1] = p.velocity_TS; D.velocity[2] = p.velocity_HP; D.velocity[3] = p.velocity_BS;
    D.k_smooth[1] = p.sglag_TS_k_sphere; D.k_smooth[2] = p.sglag_HP_k_sphere; D.k_smooth[3] = p.sglag_BS_k;
    D.phi_lim = F->phi_lim;
    D.sn_cc = F->have_cc ? F->d_cc : nullptr;
    D.sn_grid = F->have_grid ? F->d_grid : nullptr;
    D.k_angle = F->hp_k[0]; D.k_hp = F->hp_k[1]; D.hp_angle = F->hp_k[2]; D.hp_along = F->hp_k[3];
    D.n_tail = F->n_tail; D.tail_len = F->tail_len; D.N4 = F->N4; D.tail = F->d_tail;
    D.n_bs = (int32_t)F->bs_ref.size(); D.bs_ref = F->d_bs_ref; D.bs_off = F->d_bs_off; D.bs_nodes = F->d_bs_nodes;
    D.sn_ref = F->d_sn_ref;
    if (F->have_rays && F->ext.n > 0) {
        D.ext = F->ext;
        D.ext.type = F->d_ext_type; D.ext.off = F->d_ext_off; D.ext.nodes = F->d_ext_nodes;
        D.ext.opor = F->d_ext_opor; D.ext.order = F->d_ext_order; D.ext.trig = F->d_ext_trig;
    }
    return TITAN_OK;
}

int fit_node_velocity(FitState *F, hipStream_t s, int64_t nn, double *v) {
    if (!F || !v || nn != F->nn) return TITAN_E_ARG;
    std::vector<double> h(3 * F->sn_ref.size());
    if (!h.empty() && (!F->d_sn_vel || hipMemcpyAsync(h.data(), F->d_sn_vel, h.size() * sizeof(double),
                                                      hipMemcpyDeviceToHost, s) != hipSuccess))
        return TITAN_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return TITAN_E_HIP;
    std::memset(v, 0, sizeof(double) * 3 * nn);
    for (size_t i = 0; i < F->sn_ref.size(); i++)
        for (int k = 0; k < 3; k++) v[3 * (int64_t)F->sn_ref[i] + k] = h[3 * i + k];
    return TITAN_OK;
}

}  // namespace titan
