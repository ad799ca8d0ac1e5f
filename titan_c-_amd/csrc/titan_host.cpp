// titan_host.cpp -- C ABI of the MI355X step: context, device buffers, numbering, step driver.
//
// Host counterpart of Setka::Go (TITAN/source/Setka_physics.cpp:1484-2278).  Everything that is
// per face or per cell runs in HIP kernels on one stream; the host only sequences launches.  The
// lagged time step lives on the device (StepScalars), so a titan_go call issues all its steps
// without a host synchronisation and synchronises once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/titan_gpu.h"
#include "device_mesh.h"
#include "launch.h"
#include "mlp.h"
#include "titan_mesh.h"
#include "titan_fit.h"

using namespace titan;

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "titan_gpu: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                                 \
            return TITAN_E_HIP;                                                               \
        }                                                                                     \
    } while (0)

namespace {

struct Timer {
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, int>> pending;  // (family, event pair index)
    size_t used = 0;
    double ms[4] = {0, 0, 0, 0};
    long cnt[4] = {0, 0, 0, 0};
};

}  // namespace

struct titan_ctx {
    hipStream_t stream = nullptr;
    DeviceOrder ord;
    int64_t nn = 0, nc = 0, nf = 0, nb = 0;
    titan_params params{};
    DevParams dp{};
    // device topology
    int32_t *d_cell_nodes = nullptr, *d_cell_faces = nullptr, *d_face_nodes = nullptr, *d_face_cells = nullptr;
    int32_t *d_face_bidx = nullptr, *d_face_ref = nullptr, *d_cell_ref = nullptr;
    uint32_t *d_cell_code = nullptr, *d_face_code = nullptr;
    double *d_bstate = nullptr, *d_disp = nullptr;
    double *d_X[2] = {nullptr, nullptr};
    double *d_prim[2] = {nullptr, nullptr};
    double *d_scal[2] = {nullptr, nullptr};  // passive scalars (Q, rho_He) per layer, when carried
    double *d_bscal = nullptr;               // their fixed boundary values [Nb*2]
    double *d_geom[2] = {nullptr, nullptr};  // one allocation per layer, carved into DevGeom
    DevGeom geom[2];
    double *d_flux = nullptr, *d_src = nullptr, *d_neutrals = nullptr;
    double *d_hfluid = nullptr;  // hydrogen fluids of the closed-form source [Nc][n_h][5]
    int n_hfluid = 0;
    StepScalars *d_S = nullptr, *d_S2 = nullptr;
    // tiles of the fused step kernel
    TileMesh tiles{};
    int64_t tile_info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t *d_t_cell_off = nullptr, *d_t_node_off = nullptr, *d_t_nodes = nullptr, *d_t_own = nullptr, *d_t_face_off = nullptr,
            *d_t_face_ref = nullptr;
    uint64_t *d_t_faces = nullptr;
    uint16_t *d_t_cell_lnodes = nullptr, *d_t_cell_lfaces = nullptr;
    double *d_vol = nullptr;  // cell volumes at the newest node layer (step kernel's cache)
    // second-order reconstruction: Init_TVD picks (device numbering) and the per-face / per-record
    // pairs the kernels use
    std::vector<int32_t> tvd_raw;
    int32_t *d_face_tvd = nullptr, *d_t_face_tvd = nullptr;
    bool started = false;
    bool geom_dirty = false;  // nodes replaced since the geometry was last measured
    long long step = 0;
    int now1 = 1, now2 = 0;
    int last_err = 0;
    int64_t last_err_index = -1;
    bool timing = false;
    Timer tm;
    MlpState mlp;
    FitState *fit = nullptr;  // fitted surfaces and rays (move_mesh 2)
    bool has_area = false;    // titan_set_inner_area was called
    std::vector<int32_t> mk_cand;  // outer boundary faces (device ids, ascending reference id) for titan_mk_go
    // the Sun's sphere of titan_mk_go: inner boundary faces (device ids ascending), the inner face
    // across each of their 4 edges, and the sphere's radius (0 without inner faces)
    std::vector<int32_t> mk_inner, mk_inner_nb;
    double mk_R0 = 0.0;
    void *d_mk = nullptr;          // titan_mk_go: flight records, tallies, densities, counters
    bool mk_cand_ready = false;
    // Monte-Carlo zones (titan_set_mk_zone): MK zone per cell, zone-list mask and distribution slot
    // per face (device numbering), and the distributions themselves
    int mkz_r = 0;  // 0: not set
    std::vector<int8_t> mkz_cell;
    std::vector<uint8_t> mkz_mask;
    std::vector<int32_t> mkz_slot;
    std::vector<uint32_t> mkz_fcode;
    std::vector<int32_t> mkz_fcells;
    int64_t mkz_slots = 0;
    int8_t *d_mkz_cell = nullptr;
    uint8_t *d_mkz_mask = nullptr;
    int32_t *d_mkz_slot = nullptr;
    double *d_mkz_w = nullptr;
    // pickup-ion source tallies of MK_source_S (titan_set_mk_source_s): bins, range, arise table;
    // S- [Nc][nw] then S+ [Nc][nw][2] (device cell order), the 40 angles' sines and cosines
    int32_t pui_nw = 0;
    double pui_wr = 0.0;
    int8_t pui_arise[16] = {};
    double *d_pui = nullptr;
    double *d_pui_trig = nullptr;
    bool pui_valid = false;  // the arrays hold the last Monte-Carlo call's values
    DevMesh mesh() const {
        DevMesh M;
        M.n_nodes = nn; M.n_cells = nc; M.n_faces = nf;
        M.cell_nodes = d_cell_nodes; M.cell_faces = d_cell_faces; M.cell_code = d_cell_code;
        M.face_nodes = d_face_nodes; M.face_cells = d_face_cells; M.face_code = d_face_code;
        M.face_bidx = d_face_bidx; M.bstate = d_bstate; M.bscal = d_bscal; M.face_ref = d_face_ref; M.cell_ref = d_cell_ref;
        M.face_tvd = d_face_tvd;
        return M;
    }
};

namespace {

template <typename T>
int dev_upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    HIPCHK(hipMalloc((void **)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
    if (!src.empty()) HIPCHK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return TITAN_OK;
}

int set_err(titan_ctx *c, int code, int64_t idx) {
    c->last_err = code;
    c->last_err_index = idx;
    return code;
}

void timer_mark(titan_ctx *c, int fam, bool begin) {
    if (!c->timing) return;
    Timer &t = c->tm;
    if (t.used >= t.pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        t.pool.push_back(e);
    }
    hipEventRecord(t.pool[t.used], c->stream);
    if (!begin) t.pending.push_back({fam, (int)t.used - 1});
    t.used++;
}

void timer_collect(titan_ctx *c) {
    Timer &t = c->tm;
    for (auto &p : t.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.pool[p.second], t.pool[p.second + 1]) == hipSuccess) {
            t.ms[p.first] += ms;
            t.cnt[p.first] += 1;
        }
    }
    t.pending.clear();
    t.used = 0;
}

int check_scalars(titan_ctx *c, StepScalars *dS, StepScalars *host) {
    HIPCHK(hipMemcpyAsync(host, dS, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (host->err_key != ~0ull) {
        static const int kCode[6] = {TITAN_E_SURFACE, TITAN_E_RIEMANN, TITAN_E_SURFACE, TITAN_E_RIEMANN,
                                     TITAN_E_NONFINITE_STATE, TITAN_E_ARG};
        const unsigned rank = (unsigned)(host->err_key >> 48);
        const int64_t idx = (int64_t)(host->err_key & ((1ull << 48) - 1));
        const int code = rank < 6 ? kCode[rank] : TITAN_E_ARG;
        // clear for later calls
        const unsigned long long none = ~0ull;
        HIPCHK(hipMemcpy(&dS->err_key, &none, sizeof(none), hipMemcpyHostToDevice));
        return set_err(c, code, idx);
    }
    return TITAN_OK;
}

int init_scalars(StepScalars *d) {
    StepScalars h;
    std::memset(&h, 0, sizeof(h));
    h.err_key = ~0ull;
    h.dtmin_bits[0] = h.dtmin_bits[1] = 0x7ff0000000000000ull;
    HIPCHK(hipMemcpy(d, &h, sizeof(h), hipMemcpyHostToDevice));
    return TITAN_OK;
}

}  // namespace

// Setka::Init_TVD from the geometry of node layer `layer`, then the pairs of the faces that reconstruct to second
// order (Setka_gran_snos.cpp:188-194): TVD on, both cells is_TVD, both neighbours present, not a
// heliopause face unless Snos_on_HP.  Arrays in device numbering; rec_face = device face of each
// tile face record.  Replaces the context's pairs (per face and per record).
static int setup_tvd(titan_ctx *c, const std::vector<int32_t> &cell_faces, const std::vector<int32_t> &face_cells,
                     const std::vector<uint32_t> &cell_code, const std::vector<uint32_t> &face_code,
                     const std::vector<int32_t> &rec_face, int layer) {
    const int64_t nc = c->nc, nf = c->nf;
    const titan_params *p = &c->params;
    std::vector<double> cc(3 * nc), fc(3 * nf), fn(3 * nf);
    HIPCHK(hipMemcpy(cc.data(), c->geom[layer].cell_center, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fc.data(), c->geom[layer].face_center, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fn.data(), c->geom[layer].face_normal, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost));
    c->tvd_raw.assign(2 * nf, -1);
    init_tvd(nf, cell_faces.data(), face_cells.data(), cc.data(), fc.data(), fn.data(), c->tvd_raw.data());
    std::vector<int32_t> pair(2 * nf, -1);
    int64_t n_active = 0;
    if (p->tvd) {
        for (int64_t d = 0; d < nf; d++) {
            const int32_t a = face_cells[2 * d], b = face_cells[2 * d + 1];
            const int32_t AA = c->tvd_raw[2 * d], BB = c->tvd_raw[2 * d + 1];
            if (b < 0 || AA < 0 || BB < 0) continue;
            if (!((cell_code[a] >> 6) & 1u) || !((cell_code[b] >> 6) & 1u)) continue;
            const uint32_t fflags = (face_code[d] >> 8) & 255u;
            if ((fflags & 4u) && !p->snos_on_hp) continue;
            pair[2 * d] = AA;
            pair[2 * d + 1] = BB;
            n_active++;
        }
    }
    std::vector<int32_t> rec_pair(2 * rec_face.size());
    for (size_t r = 0; r < rec_face.size(); r++) {
        rec_pair[2 * r] = pair[2 * (int64_t)rec_face[r]];
        rec_pair[2 * r + 1] = pair[2 * (int64_t)rec_face[r] + 1];
    }
    if (!c->d_face_tvd) HIPCHK(hipMalloc((void **)&c->d_face_tvd, sizeof(int32_t) * std::max<size_t>(1, pair.size())));
    if (!c->d_t_face_tvd)
        HIPCHK(hipMalloc((void **)&c->d_t_face_tvd, sizeof(int32_t) * std::max<size_t>(1, rec_pair.size())));
    HIPCHK(hipMemcpy(c->d_face_tvd, pair.data(), sizeof(int32_t) * pair.size(), hipMemcpyHostToDevice));
    if (!rec_pair.empty())
        HIPCHK(hipMemcpy(c->d_t_face_tvd, rec_pair.data(), sizeof(int32_t) * rec_pair.size(), hipMemcpyHostToDevice));
    c->tiles.face_tvd = reinterpret_cast<const int2 *>(c->d_t_face_tvd);
    c->dp.tvd = n_active > 0 ? 1 : 0;
    c->tile_info[7] = n_active;
    return TITAN_OK;
}

extern "C" int titan_create(titan_ctx **out, const titan_mesh *m, const titan_params *p) {
    if (!out || !m || !p) return TITAN_E_ARG;
    *out = nullptr;
    if (m->n_nodes <= 0 || m->n_cells <= 0 || m->n_faces <= 0 || !m->cell_nodes || !m->face_nodes ||
        !m->face_cells || !m->cell_faces || !m->face_bc || !m->face_method || !m->cell_zone || !m->cell_is_tvd ||
        !m->node_xyz)
        return TITAN_E_ARG;
    if (m->n_cells > INT32_MAX || m->n_faces > INT32_MAX || m->n_nodes > INT32_MAX) return TITAN_E_ARG;
    if (p->move_mesh < 0 || p->move_mesh > 2 || (p->move_mesh == 1 && !m->node_disp)) return TITAN_E_ARG;
    if (p->source_mode < TITAN_SRC_NONE || p->source_mode > TITAN_SRC_BERA) return TITAN_E_ARG;
    if (p->source_mode == TITAN_SRC_BERA && !(p->par_Kn > 0.0)) return TITAN_E_ARG;
    if (p->passive_scalars < 0 || p->passive_scalars > 1) return TITAN_E_ARG;
    const bool scal = p->passive_scalars != 0;
    const int64_t nn = m->n_nodes, nc = m->n_cells, nf = m->n_faces;
    // ---- topology checks (Setka::Test_geometr, Setka.cpp:475-617) ----
    for (int64_t f = 0; f < nf; f++) {
        const int32_t a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
        if (a < 0 || a >= nc || b >= nc || a == b) return TITAN_E_TOPOLOGY;
        if ((b < 0) != (m->face_bc[f] != TITAN_BC_INTERIOR)) return TITAN_E_TOPOLOGY;
        if (m->face_bc[f] < 0 || m->face_bc[f] > 4 || m->face_method[f] < 0 || m->face_method[f] > 3)
            return TITAN_E_ARG;
        if ((m->face_bc[f] == TITAN_BC_INNER_HARD || m->face_bc[f] == TITAN_BC_OUTER_HARD) &&
            (!m->face_bstate || (scal && !m->face_bscalar)))
            return TITAN_E_ARG;
        for (int i = 0; i < 4; i++)
            if (m->face_nodes[4 * f + i] < 0 || m->face_nodes[4 * f + i] >= nn) return TITAN_E_TOPOLOGY;
    }
    for (int64_t c = 0; c < nc; c++) {
        for (int j = 0; j < 6; j++) {
            const int32_t f = m->cell_faces[6 * c + j];
            if (f < 0 || f >= nf) return TITAN_E_TOPOLOGY;
            if (j > 0 && f <= m->cell_faces[6 * c + j - 1]) return TITAN_E_TOPOLOGY;  // ascending, distinct
            if (m->face_cells[2 * f] != c && m->face_cells[2 * f + 1] != c) return TITAN_E_TOPOLOGY;
        }
        for (int i = 0; i < 8; i++)
            if (m->cell_nodes[8 * c + i] < 0 || m->cell_nodes[8 * c + i] >= nn) return TITAN_E_TOPOLOGY;
    }

    titan_ctx *c = new titan_ctx();
    c->nn = nn; c->nc = nc; c->nf = nf;
    c->params = *p;
    c->dp.gamma = p->gamma; c->dp.kfl = p->kfl; c->dp.v_inf = p->v_inf;
    c->dp.move_mesh = p->move_mesh; c->dp.n_disc = p->n_disc; c->dp.geom_mode = p->geom_mode;
    c->dp.tvd = 0;
    c->dp.motion_omega = p->motion_omega; c->dp.motion_phase = p->motion_phase;
    c->dp.n_scal = scal ? kScal : 0;
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return TITAN_E_HIP; }

    MeshView mv{nn, nc, nf, m->cell_nodes, m->face_cells, m->node_xyz};
    build_order(mv, c->ord);
    const DeviceOrder &o = c->ord;
    {
        const int frc = fit_create(&c->fit, nn, nf, m->face_nodes, m->face_cells, m->cell_faces, m->face_flags, o);
        if (frc != TITAN_OK) { titan_destroy(c); return frc; }
    }

    std::vector<int32_t> cell_nodes(8 * nc), cell_faces(6 * nc), face_nodes(4 * nf), face_cells(2 * nf);
    std::vector<int32_t> face_bidx(nf, -1), face_ref(nf), cell_ref(nc);
    std::vector<uint32_t> cell_code(nc), face_code(nf);
    std::vector<double> bstate, bscal;
    for (int64_t d = 0; d < nc; d++) {
        const int64_t r = o.cell_d2r[d];
        cell_ref[d] = (int32_t)r;
        for (int i = 0; i < 8; i++) cell_nodes[8 * d + i] = o.node_r2d[m->cell_nodes[8 * r + i]];
        uint32_t fsign = 0;
        for (int j = 0; j < 6; j++) {
            const int32_t fr = m->cell_faces[6 * r + j];
            cell_faces[6 * d + j] = o.face_r2d[fr];
            if (m->face_cells[2 * fr] == r) fsign |= 1u << j;
        }
        cell_code[d] = cell_code_make(fsign, m->cell_is_tvd[r] ? 1u : 0u, (uint32_t)m->cell_zone[r]);
    }
    for (int64_t d = 0; d < nf; d++) {
        const int64_t r = o.face_d2r[d];
        face_ref[d] = (int32_t)r;
        for (int i = 0; i < 4; i++) face_nodes[4 * d + i] = o.node_r2d[m->face_nodes[4 * r + i]];
        face_cells[2 * d] = o.cell_r2d[m->face_cells[2 * r]];
        const int32_t b = m->face_cells[2 * r + 1];
        face_cells[2 * d + 1] = b >= 0 ? o.cell_r2d[b] : -1;
        const int bc = m->face_bc[r];
        face_code[d] = face_code_make((uint32_t)bc, (uint32_t)m->face_method[r], m->face_flags ? m->face_flags[r] : 0u);
        if (bc == TITAN_BC_INNER_HARD || bc == TITAN_BC_OUTER_HARD) {
            face_bidx[d] = (int32_t)(bstate.size() / 8);
            for (int i = 0; i < 8; i++) bstate.push_back(m->face_bstate[8 * r + i]);
            if (scal)
                for (int i = 0; i < kScal; i++) bscal.push_back(m->face_bscalar[kScal * r + i]);
        }
    }
    c->nb = (int64_t)bstate.size() / 8;
    {
        // inner boundary faces, their edge neighbours among them, and the radius of their nodes
        std::unordered_map<uint64_t, int32_t> edge;  // node pair -> first (face index * 4 + edge)
        for (int64_t d = 0; d < nf; d++)
            if ((face_code[d] & 15u) == (uint32_t)TITAN_BC_INNER_HARD) c->mk_inner.push_back((int32_t)d);
        c->mk_inner_nb.assign(4 * c->mk_inner.size(), -1);
        for (size_t i = 0; i < c->mk_inner.size(); i++) {
            const int64_t d = c->mk_inner[i];
            for (int k = 0; k < 4; k++) {
                const uint32_t a = (uint32_t)face_nodes[4 * d + k], b = (uint32_t)face_nodes[4 * d + ((k + 1) & 3)];
                const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
                const auto it = edge.find(key);
                if (it == edge.end()) {
                    edge.emplace(key, (int32_t)(4 * i + k));
                } else {
                    c->mk_inner_nb[4 * i + k] = it->second / 4;
                    c->mk_inner_nb[it->second] = (int32_t)i;
                }
                const double *x = m->node_xyz + 3 * (int64_t)o.node_d2r[face_nodes[4 * d + k]];
                c->mk_R0 = std::max(c->mk_R0, std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]));
            }
        }
    }
    std::vector<double> X(3 * nn), disp;
    for (int64_t d = 0; d < nn; d++)
        for (int k = 0; k < 3; k++) X[3 * d + k] = m->node_xyz[3 * (int64_t)o.node_d2r[d] + k];
    if (m->node_disp) {
        disp.resize(3 * nn);
        for (int64_t d = 0; d < nn; d++)
            for (int k = 0; k < 3; k++) disp[3 * d + k] = m->node_disp[3 * (int64_t)o.node_d2r[d] + k];
    }
    int rc = TITAN_OK;
    auto up = [&](int r) { if (rc == TITAN_OK) rc = r; };
    std::vector<int32_t> rec_face;  // device face of each tile face record
    up(dev_upload(&c->d_cell_nodes, cell_nodes, c->stream));
    up(dev_upload(&c->d_cell_faces, cell_faces, c->stream));
    up(dev_upload(&c->d_face_nodes, face_nodes, c->stream));
    up(dev_upload(&c->d_face_cells, face_cells, c->stream));
    up(dev_upload(&c->d_face_bidx, face_bidx, c->stream));
    up(dev_upload(&c->d_face_ref, face_ref, c->stream));
    up(dev_upload(&c->d_cell_ref, cell_ref, c->stream));
    up(dev_upload(&c->d_cell_code, cell_code, c->stream));
    up(dev_upload(&c->d_face_code, face_code, c->stream));
    up(dev_upload(&c->d_bstate, bstate, c->stream));
    if (scal) up(dev_upload(&c->d_bscal, bscal, c->stream));
    up(dev_upload(&c->d_X[0], X, c->stream));
    up(dev_upload(&c->d_X[1], X, c->stream));
    if (!disp.empty()) up(dev_upload(&c->d_disp, disp, c->stream));
    if (rc != TITAN_OK) { titan_destroy(c); return rc; }
    {
        // tile size: the largest of 128, 112, ... cells whose LDS footprint lets two workgroups
        // share a CU (the step kernel is built for 2 waves per SIMD, and a CU's LDS equals the
        // per-workgroup limit); failing that, the largest that fits one workgroup
        // (TITAN_TILE_CELLS overrides the starting size)
        int lds_limit = 65536;
        {
            int dev = 0, v = 0;
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
                lds_limit = v;
        }
        int tile_start = 128;
        if (const char *e = std::getenv("TITAN_TILE_CELLS")) tile_start = std::max(16, std::min(1024, std::atoi(e)));
        // at most two face slots per thread: a tile with more faces would hold its workgroup for a
        // third pass of the face phase that only a few lanes use
        const int face_cap = 2 * kTileThreads;
        TileData td;
        bool fits = false;
        for (int pass = 0; pass < 2 && !fits; pass++) {
            // 256 bytes stay free for the kernel's static LDS (the dt reduction's wavefront minima)
            const size_t budget = pass == 0 ? (size_t)lds_limit / 2 - 256 : (size_t)lds_limit - 256;
            for (int tile_cells = tile_start; tile_cells >= 16; tile_cells -= 16) {
                if (!build_tiles(nc, nf, nn, tile_cells, face_cap, cell_nodes.data(), cell_faces.data(),
                                 face_nodes.data(), face_cells.data(), face_code.data(), cell_code.data(),
                                 face_bidx.data(), face_ref.data(), td))
                    continue;
                TileMesh probe{};
                probe.tile_cells = td.tile_cells;
                probe.max_nodes = td.max_nodes;
                probe.max_faces = td.max_faces;
                if (step_tile_lds_bytes(probe, p->move_mesh != 0, scal) <= budget) { fits = true; break; }
            }
        }
        if (!fits) { titan_destroy(c); return TITAN_E_TOPOLOGY; }
        up(dev_upload(&c->d_t_cell_off, td.cell_off, c->stream));
        up(dev_upload(&c->d_t_node_off, td.node_off, c->stream));
        up(dev_upload(&c->d_t_nodes, td.nodes, c->stream));
        up(dev_upload(&c->d_t_own, td.own, c->stream));
        up(dev_upload(&c->d_t_face_off, td.face_off, c->stream));
        up(dev_upload(&c->d_t_face_ref, td.face_ref, c->stream));
        up(dev_upload(&c->d_t_faces, td.faces, c->stream));
        up(dev_upload(&c->d_t_cell_lnodes, td.cell_lnodes, c->stream));
        up(dev_upload(&c->d_t_cell_lfaces, td.cell_lfaces, c->stream));
        if (rc != TITAN_OK) { titan_destroy(c); return rc; }
        TileMesh &T = c->tiles;
        T.n_tiles = (int64_t)td.node_off.size() - 1;
        T.tile_cells = td.tile_cells;
        T.max_nodes = td.max_nodes;
        T.max_faces = td.max_faces;
        T.cell_off = c->d_t_cell_off; T.node_off = c->d_t_node_off; T.nodes = c->d_t_nodes; T.own = c->d_t_own; T.face_off = c->d_t_face_off;
        T.faces = reinterpret_cast<const ulonglong2 *>(c->d_t_faces); T.face_ref = c->d_t_face_ref;
        T.cell_lnodes = c->d_t_cell_lnodes; T.cell_lfaces = c->d_t_cell_lfaces;
        if (hipMalloc((void **)&c->d_vol, sizeof(double) * nc) != hipSuccess) { titan_destroy(c); return TITAN_E_HIP; }
        T.vol = c->d_vol;
        T.vol_cached = 0;
        c->tile_info[0] = T.tile_cells;
        c->tile_info[1] = T.max_nodes;
        c->tile_info[2] = T.max_faces;
        c->tile_info[3] = (int64_t)step_tile_lds_bytes(T, p->move_mesh != 0, scal);
        c->tile_info[4] = lds_limit;
        c->tile_info[5] = T.n_tiles;
        c->tile_info[6] = (int64_t)td.face_ref.size();
        rec_face.swap(td.face_dev);
    }
    const size_t geom_n = (size_t)(4 * nc + 7 * nf);
    for (int l = 0; l < 2; l++) {
        if (hipMalloc((void **)&c->d_prim[l], sizeof(double) * 8 * nc) != hipSuccess ||
            hipMalloc((void **)&c->d_geom[l], sizeof(double) * geom_n) != hipSuccess) {
            titan_destroy(c);
            return TITAN_E_HIP;
        }
        double *g = c->d_geom[l];
        c->geom[l].cell_center = g;
        c->geom[l].cell_volume = g + 3 * nc;
        c->geom[l].face_center = g + 4 * nc;
        c->geom[l].face_normal = g + 4 * nc + 3 * nf;
        c->geom[l].face_area = g + 4 * nc + 6 * nf;
        hipMemsetAsync(c->d_prim[l], 0, sizeof(double) * 8 * nc, c->stream);
        if (scal) {
            if (hipMalloc((void **)&c->d_scal[l], sizeof(double) * kScal * nc) != hipSuccess) {
                titan_destroy(c);
                return TITAN_E_HIP;
            }
            hipMemsetAsync(c->d_scal[l], 0, sizeof(double) * kScal * nc, c->stream);
        }
    }
    if (hipMalloc((void **)&c->d_flux, sizeof(double) * (9 + (scal ? kScal : 0)) * nf) != hipSuccess ||
        hipMalloc((void **)&c->d_S, sizeof(StepScalars)) != hipSuccess ||
        hipMalloc((void **)&c->d_S2, sizeof(StepScalars)) != hipSuccess) {
        titan_destroy(c);
        return TITAN_E_HIP;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess || init_scalars(c->d_S) != TITAN_OK ||
        init_scalars(c->d_S2) != TITAN_OK) {
        titan_destroy(c);
        return TITAN_E_HIP;
    }
    // geometry of both layers from the initial nodes
    launch_measure(c->stream, c->mesh(), c->d_X[0], c->geom[0], p->geom_mode);
    launch_measure(c->stream, c->mesh(), c->d_X[1], c->geom[1], p->geom_mode);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { titan_destroy(c); return TITAN_E_HIP; }
    {
        const int trc = setup_tvd(c, cell_faces, face_cells, cell_code, face_code, rec_face, 0);
        if (trc != TITAN_OK) { titan_destroy(c); return trc; }
    }
    *out = c;
    return TITAN_OK;
}

extern "C" int titan_destroy(titan_ctx *c) {
    if (!c) return TITAN_E_ARG;
    if (c->stream) hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->d_cell_nodes, c->d_cell_faces, c->d_face_nodes, c->d_face_cells, c->d_face_bidx,
                    c->d_face_ref, c->d_cell_ref, c->d_cell_code, c->d_face_code, c->d_bstate, c->d_disp,
                    c->d_X[0], c->d_X[1], c->d_prim[0], c->d_prim[1], c->d_scal[0], c->d_scal[1], c->d_bscal,
                    c->d_geom[0], c->d_geom[1], c->d_flux,
                    c->d_src, c->d_neutrals, c->d_hfluid, c->d_S, c->d_S2, c->d_t_cell_off, c->d_t_node_off,
                    c->d_t_nodes, c->d_t_own,
                    c->d_t_face_off, c->d_t_face_ref, c->d_t_faces, c->d_t_cell_lnodes, c->d_t_cell_lfaces,
                    c->d_vol, c->d_face_tvd, c->d_t_face_tvd, c->d_mk, c->d_mkz_cell, c->d_mkz_mask,
                    c->d_mkz_slot, c->d_mkz_w, c->d_pui, c->d_pui_trig};
    for (void *q : ptrs)
        if (q) hipFree(q);
    mlp_free(c->mlp);
    fit_destroy(c->fit);
    for (auto e : c->tm.pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return TITAN_OK;
}

extern "C" int titan_last_error(titan_ctx *c, int64_t *index) {
    if (!c) return TITAN_E_ARG;
    if (index) *index = c->last_err_index;
    return c->last_err;
}

extern "C" int titan_set_state(titan_ctx *c, const double *prim, int layer) {
    if (!c || !prim || layer < 0 || layer > 1) return TITAN_E_ARG;
    std::vector<double> h(8 * c->nc);
    for (int64_t d = 0; d < c->nc; d++) std::memcpy(&h[8 * d], prim + 8 * (int64_t)c->ord.cell_d2r[d], 64);
    HIPCHK(hipMemcpyAsync(c->d_prim[layer], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_get_state(titan_ctx *c, double *prim, int layer) {
    if (!c || !prim || layer < 0 || layer > 1) return TITAN_E_ARG;
    std::vector<double> h(8 * c->nc);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_prim[layer], h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t d = 0; d < c->nc; d++) std::memcpy(prim + 8 * (int64_t)c->ord.cell_d2r[d], &h[8 * d], 64);
    return TITAN_OK;
}

extern "C" int titan_set_scalars(titan_ctx *c, const double *qhe, int layer) {
    if (!c || !qhe || layer < 0 || layer > 1 || !c->d_scal[layer]) return TITAN_E_ARG;
    std::vector<double> h(kScal * c->nc);
    for (int64_t d = 0; d < c->nc; d++)
        std::memcpy(&h[kScal * d], qhe + kScal * (int64_t)c->ord.cell_d2r[d], sizeof(double) * kScal);
    HIPCHK(hipMemcpyAsync(c->d_scal[layer], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_get_scalars(titan_ctx *c, double *qhe, int layer) {
    if (!c || !qhe || layer < 0 || layer > 1 || !c->d_scal[layer]) return TITAN_E_ARG;
    std::vector<double> h(kScal * c->nc);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_scal[layer], h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t d = 0; d < c->nc; d++)
        std::memcpy(qhe + kScal * (int64_t)c->ord.cell_d2r[d], &h[kScal * d], sizeof(double) * kScal);
    return TITAN_OK;
}

extern "C" int titan_set_nodes(titan_ctx *c, const double *xyz, int layer) {
    if (!c || !xyz || layer < 0 || layer > 1) return TITAN_E_ARG;
    std::vector<double> h(3 * c->nn);
    for (int64_t d = 0; d < c->nn; d++)
        for (int k = 0; k < 3; k++) h[3 * d + k] = xyz[3 * (int64_t)c->ord.node_d2r[d] + k];
    HIPCHK(hipMemcpyAsync(c->d_X[layer], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tiles.vol_cached = 0;  // the cached volumes belong to the old coordinates
    return TITAN_OK;
}

extern "C" int titan_get_nodes(titan_ctx *c, double *xyz, int layer) {
    if (!c || !xyz || layer < 0 || layer > 1) return TITAN_E_ARG;
    std::vector<double> h(3 * c->nn);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_X[layer], h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t d = 0; d < c->nn; d++)
        for (int k = 0; k < 3; k++) xyz[3 * (int64_t)c->ord.node_d2r[d] + k] = h[3 * d + k];
    return TITAN_OK;
}

extern "C" int titan_set_neutrals(titan_ctx *c, const double *nh5) {
    if (!c || !nh5) return TITAN_E_ARG;
    if (!c->d_neutrals) HIPCHK(hipMalloc((void **)&c->d_neutrals, sizeof(double) * 5 * c->nc));
    std::vector<double> h(5 * c->nc);
    for (int64_t d = 0; d < c->nc; d++)
        for (int k = 0; k < 5; k++) h[5 * d + k] = nh5[5 * (int64_t)c->ord.cell_d2r[d] + k];
    HIPCHK(hipMemcpyAsync(c->d_neutrals, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_set_h_fluids(titan_ctx *c, int n_h, const double *h) {
    if (!c || !h || n_h < 1 || n_h > 9) return TITAN_E_ARG;
    if (c->d_hfluid && n_h != c->n_hfluid) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->d_hfluid));
        c->d_hfluid = nullptr;
    }
    const int64_t w = 5 * (int64_t)n_h;
    if (!c->d_hfluid) HIPCHK(hipMalloc((void **)&c->d_hfluid, sizeof(double) * w * c->nc));
    c->n_hfluid = n_h;
    std::vector<double> v(w * c->nc);
    for (int64_t d = 0; d < c->nc; d++)
        std::memcpy(&v[w * d], h + w * (int64_t)c->ord.cell_d2r[d], sizeof(double) * w);
    HIPCHK(hipMemcpyAsync(c->d_hfluid, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_calculating_measure(titan_ctx *c, int layer) {
    if (!c || layer < 0 || layer > 1) return TITAN_E_ARG;
    launch_measure(c->stream, c->mesh(), c->d_X[layer], c->geom[layer], c->params.geom_mode);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_get_geometry(titan_ctx *c, int layer, double *cell_center, double *cell_volume,
                                  double *face_center, double *face_normal, double *face_area) {
    if (!c || layer < 0 || layer > 1) return TITAN_E_ARG;
    const int64_t nc = c->nc, nf = c->nf;
    // the fused step keeps only cell centres current: rebuild the layer's geometry from its nodes
    launch_measure(c->stream, c->mesh(), c->d_X[layer], c->geom[layer], c->params.geom_mode);
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)(4 * nc + 7 * nf));
    HIPCHK(hipMemcpyAsync(h.data(), c->d_geom[layer], h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double *cc = h.data(), *cv = cc + 3 * nc, *fc = cc + 4 * nc, *fn = fc + 3 * nf, *fa = fc + 6 * nf;
    for (int64_t d = 0; d < nc; d++) {
        const int64_t r = c->ord.cell_d2r[d];
        if (cell_center) for (int k = 0; k < 3; k++) cell_center[3 * r + k] = cc[3 * d + k];
        if (cell_volume) cell_volume[r] = cv[d];
    }
    for (int64_t d = 0; d < nf; d++) {
        const int64_t r = c->ord.face_d2r[d];
        if (face_center) for (int k = 0; k < 3; k++) face_center[3 * r + k] = fc[3 * d + k];
        if (face_normal) for (int k = 0; k < 3; k++) face_normal[3 * r + k] = fn[3 * d + k];
        if (face_area) face_area[r] = fa[d];
    }
    return TITAN_OK;
}

extern "C" int titan_gran_potok(titan_ctx *c, int now, double dt, double *face_flux, double *face_kflux,
                                double *dt_min, uint8_t *branch) {
    if (!c || now < 0 || now > 1 || !(dt > 0.0)) return TITAN_E_ARG;
    if (face_kflux && !c->dp.n_scal) return TITAN_E_ARG;
    if (init_scalars(c->d_S2) != TITAN_OK) return TITAN_E_HIP;
    // slot (step & 1) = 1 is reset to the reference's initial 1e8
    unsigned long long init = 0x4197D78400000000ull;  // bits of 1.0e8
    HIPCHK(hipMemcpy(&c->d_S2->dtmin_bits[1], &init, sizeof(init), hipMemcpyHostToDevice));
    uint8_t *d_branch = nullptr;
    if (branch) HIPCHK(hipMalloc((void **)&d_branch, c->nf));
    launch_measure(c->stream, c->mesh(), c->d_X[now], c->geom[now], c->params.geom_mode);
    double *d_kflux = c->d_flux + 9 * c->nf;  // allocated with the flux when the scalars are carried
    launch_face_flux(c->stream, c->params.precision, c->mesh(), c->dp, c->d_prim[now], c->d_scal[now], c->geom[now],
                     c->d_X[now], c->d_X[1 - now], c->d_S2, c->d_flux, d_kflux, d_branch, 1, dt);
    HIPCHK(hipGetLastError());
    StepScalars hs;
    int rc = check_scalars(c, c->d_S2, &hs);
    if (dt_min) *dt_min = __builtin_bit_cast(double, hs.dtmin_bits[1]);
    if (face_flux) {
        std::vector<double> h(9 * c->nf);
        HIPCHK(hipMemcpy(h.data(), c->d_flux, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t d = 0; d < c->nf; d++) std::memcpy(face_flux + 9 * (int64_t)c->ord.face_d2r[d], &h[9 * d], 72);
    }
    if (face_kflux) {
        std::vector<double> h(kScal * c->nf);
        HIPCHK(hipMemcpy(h.data(), d_kflux, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t d = 0; d < c->nf; d++)
            std::memcpy(face_kflux + kScal * (int64_t)c->ord.face_d2r[d], &h[kScal * d], sizeof(double) * kScal);
    }
    if (branch) {
        std::vector<uint8_t> h(c->nf);
        HIPCHK(hipMemcpy(h.data(), d_branch, c->nf, hipMemcpyDeviceToHost));
        for (int64_t d = 0; d < c->nf; d++) branch[c->ord.face_d2r[d]] = h[d];
        hipFree(d_branch);
    }
    return rc;
}

// The inputs a source mode reads are present, and its output buffer exists.
static int prepare_sources(titan_ctx *c, int mode) {
    if (mode == TITAN_SRC_NONE) return TITAN_OK;
    if (mode != TITAN_SRC_BERA && mode != TITAN_SRC_EXACT && mode != TITAN_SRC_MLP) return TITAN_E_ARG;
    if (mode == TITAN_SRC_BERA ? !c->d_hfluid : !c->d_neutrals) return TITAN_E_ARG;
    if (mode == TITAN_SRC_MLP && !c->mlp.ready) return TITAN_E_NOT_TRAINED;
    if (!c->d_src) HIPCHK(hipMalloc((void **)&c->d_src, sizeof(double) * 5 * c->nc));
    return TITAN_OK;
}

static int compute_sources(titan_ctx *c, int layer, int mode) {
    if (mode == TITAN_SRC_NONE) return TITAN_OK;
    const int prc = prepare_sources(c, mode);
    if (prc != TITAN_OK) return prc;
    if (mode == TITAN_SRC_BERA) {
        const titan_params &p = c->params;
        launch_bera_cells(c->stream, c->nc, c->d_prim[layer], c->d_scal[layer], c->d_cell_code, c->d_hfluid,
                          c->n_hfluid, p.par_a_2, p.par_n_H_LISM / p.par_Kn, c->d_src);
        HIPCHK(hipGetLastError());
        return TITAN_OK;
    }
    const int nodes = c->params.cx_nodes > 0 ? c->params.cx_nodes : 100;
    if (mode == TITAN_SRC_EXACT) {
        launch_cx_cells(c->stream, c->nc, c->d_prim[layer], c->d_neutrals, nodes, c->d_src);
    } else if (mode == TITAN_SRC_MLP) {
        if (!c->mlp.ready) return TITAN_E_NOT_TRAINED;
        int rc = mlp_cell_sources(c->mlp, c->stream, c->nc, c->d_prim[layer], c->d_neutrals, nodes, c->d_src,
                                  &c->d_S->n_fallback);
        if (rc != TITAN_OK) return rc;
    } else {
        return TITAN_E_ARG;
    }
    HIPCHK(hipGetLastError());
    return TITAN_OK;
}

// The two events that time a titan_go call, destroyed on every return path.
struct CallEvents {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~CallEvents() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// The step loop: `area` selects the cells updated (cell_in_area), `restart` starts the dt schedule
// again at 1e-5 as each Setka::Go call does.  Every argument and input check comes before the first
// change to the context (layers, step count, dt slots), so a call that fails them leaves the
// context as it was.
static int go_impl(titan_ctx *c, int64_t steps, titan_step_stats *stats, int area, bool restart) {
    if (!c || steps < 0) return TITAN_E_ARG;
    // surrogate source: fused into the step kernel, out-of-box cells redone by the fixup kernel
    const bool fused = c->params.source_mode == TITAN_SRC_MLP;
    SurArgs sur{};
    if (fused) {
        if (!c->d_neutrals || c->tiles.tile_cells > kTileThreads / 2) return TITAN_E_ARG;
        const int rc = mlp_fused_args(c->mlp, c->nc, c->d_neutrals, &sur);
        if (rc != TITAN_OK) return rc;
    } else {
        const int rc = prepare_sources(c, c->params.source_mode);
        if (rc != TITAN_OK) return rc;
    }
    const int nodes = c->params.cx_nodes > 0 ? c->params.cx_nodes : 100;
    // fitted surfaces: Culc_Velocity_surface and dvigenie of every ray write the new node layer
    // before each step (Setka_physics.cpp:1616-1627)
    const bool fitting = c->params.move_mesh == 2;
    FitDev fd{};
    if (fitting) {
        const int rc = fit_prepare(c->fit, c->params, c->stream, &fd);
        if (rc != TITAN_OK) return rc;
    }
    c->dp.area = area;
    const DevMesh M = c->mesh();
    const int64_t nn3 = 3 * c->nn;
    CallEvents ev;
    HIPCHK(hipEventCreate(&ev.e0));
    HIPCHK(hipEventCreate(&ev.e1));
    HIPCHK(hipEventRecord(ev.e0, c->stream));
    if (c->started && c->geom_dirty) {
        // nodes were replaced (titan_download_cell_parameters): measure both layers again, as
        // Setka::Go does on every entry (Setka_physics.cpp:1495-1510)
        c->tiles.vol_cached = 0;
        launch_measure(c->stream, M, c->d_X[0], c->geom[0], c->params.geom_mode);
        launch_measure(c->stream, M, c->d_X[1], c->geom[1], c->params.geom_mode);
    }
    c->geom_dirty = false;
    if (!c->started) {
        c->tiles.vol_cached = 0;
        // Setka_physics.cpp:1486-1502: copy node and cell layers 0 -> 1, measure both
        HIPCHK(hipMemcpyAsync(c->d_X[1], c->d_X[0], sizeof(double) * nn3, hipMemcpyDeviceToDevice, c->stream));
        launch_measure(c->stream, M, c->d_X[0], c->geom[0], c->params.geom_mode);
        launch_measure(c->stream, M, c->d_X[1], c->geom[1], c->params.geom_mode);
        HIPCHK(hipMemcpyAsync(c->d_prim[1], c->d_prim[0], sizeof(double) * 8 * c->nc, hipMemcpyDeviceToDevice,
                              c->stream));
        if (c->dp.n_scal)
            HIPCHK(hipMemcpyAsync(c->d_scal[1], c->d_scal[0], sizeof(double) * kScal * c->nc, hipMemcpyDeviceToDevice,
                                  c->stream));
        c->now1 = 1;
        c->now2 = 0;
        c->started = true;
    }
    for (int64_t s = 0; s < steps; s++) {
        c->now2 = c->now1;
        c->now1 = (c->now1 + 1) % 2;
        c->step++;
        const int n1 = c->now1, n2 = c->now2;
        launch_begin_step(c->stream, c->d_S, c->step, c->dp, restart && s == 0 ? 1 : 0);
        if (fitting) {
            timer_mark(c, 2, true);
            launch_fit_surface(c->stream, M, c->dp, fd, c->d_prim[n1], c->d_X[n1], c->d_X[n2],
                               c->geom[n1].cell_center, c->d_S, 0.0);
            launch_fit_rays(c->stream, fd, c->d_X[n2]);
            timer_mark(c, 2, false);
        }
        const double *src = nullptr;
        if (c->params.source_mode != TITAN_SRC_NONE && !fused) {
            timer_mark(c, 3, true);
            int rc = compute_sources(c, n1, c->params.source_mode);
            timer_mark(c, 3, false);
            if (rc != TITAN_OK) return rc;
            src = c->d_src;
        }
        timer_mark(c, 0, true);
        launch_step_tile(c->stream, c->params.precision, c->tiles, M, c->dp, c->d_prim[n1], c->d_prim[n2],
                         c->d_scal[n1], c->d_scal[n2], c->d_X[n1], c->d_X[n2], c->d_disp, c->geom[n1].cell_center,
                         c->geom[n2].cell_center, src, fused ? &sur : nullptr, c->d_S, c->step);
        c->tiles.vol_cached = 1;
        timer_mark(c, 0, false);
        if (fused) {
            timer_mark(c, 3, true);
            launch_cx_fixup(c->stream, M, c->d_prim[n1], c->d_neutrals, nodes, sur.list, sur.stash, sur.cap,
                            sur.fixpart, c->d_S, c->d_prim[n2]);
            timer_mark(c, 3, false);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev.e1, c->stream));
    StepScalars hs;
    int rc = check_scalars(c, c->d_S, &hs);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
    if (c->timing) timer_collect(c);
    if (stats) {
        stats->steps_done = c->step;
        stats->dt_last = hs.dt;
        stats->dt_next = __builtin_bit_cast(double, hs.dtmin_bits[c->step & 1]);
        stats->t_sim = hs.t + (c->step > 0 ? hs.dt : 0.0);
        stats->n_rho_floor = (int64_t)hs.n_rho_floor;
        stats->n_p_floor = (int64_t)hs.n_p_floor;
        stats->n_fallback = (int64_t)hs.n_fallback;
        stats->n_fp64_faces = (int64_t)hs.n_fp64_faces;
        stats->ms_total = ms;
    }
    return rc;
}

extern "C" int titan_go(titan_ctx *c, int64_t steps, titan_step_stats *stats) {
    const int rc = go_impl(c, steps, stats, 0, false);
    if (c) c->dp.area = 0;
    return rc;
}

extern "C" int titan_go_area(titan_ctx *c, int is_inner_area, int64_t steps, titan_step_stats *stats) {
    if (!c || !c->has_area) return TITAN_E_ARG;
    const int rc = go_impl(c, steps, stats, is_inner_area ? 1 : 2, true);
    c->dp.area = 0;
    return rc;
}

extern "C" int titan_set_inner_area(titan_ctx *c, const uint8_t *cell_is_inner) {
    if (!c || !cell_is_inner) return TITAN_E_ARG;
    std::vector<uint32_t> code(c->nc);
    HIPCHK(hipMemcpy(code.data(), c->d_cell_code, sizeof(uint32_t) * c->nc, hipMemcpyDeviceToHost));
    for (int64_t d = 0; d < c->nc; d++) {
        const bool in = cell_is_inner[c->ord.cell_d2r[d]] != 0;
        code[d] = (code[d] & ~(1u << 7)) | (in ? (1u << 7) : 0u);
    }
    HIPCHK(hipMemcpy(c->d_cell_code, code.data(), sizeof(uint32_t) * c->nc, hipMemcpyHostToDevice));
    c->has_area = true;
    return TITAN_OK;
}

extern "C" int titan_current_layer(titan_ctx *c) {
    if (!c) return TITAN_E_ARG;
    return c->started ? c->now2 : 0;
}

extern "C" int titan_init_tvd(titan_ctx *c) {
    if (!c) return TITAN_E_ARG;
    const int64_t nc = c->nc, nf = c->nf;
    const int layer = c->started ? c->now2 : 0;  // the newest nodes
    launch_measure(c->stream, c->mesh(), c->d_X[layer], c->geom[layer], c->params.geom_mode);
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int32_t> cell_faces(6 * nc), face_cells(2 * nf), rec_ref((size_t)c->tile_info[6]), rec_face;
    std::vector<uint32_t> cell_code(nc), face_code(nf);
    HIPCHK(hipMemcpy(cell_faces.data(), c->d_cell_faces, sizeof(int32_t) * cell_faces.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(face_cells.data(), c->d_face_cells, sizeof(int32_t) * face_cells.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cell_code.data(), c->d_cell_code, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(face_code.data(), c->d_face_code, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rec_ref.data(), c->d_t_face_ref, sizeof(int32_t) * rec_ref.size(), hipMemcpyDeviceToHost));
    rec_face.resize(rec_ref.size());
    for (size_t r = 0; r < rec_ref.size(); r++) rec_face[r] = c->ord.face_r2d[rec_ref[r]];
    return setup_tvd(c, cell_faces, face_cells, cell_code, face_code, rec_face, layer);
}

extern "C" int titan_set_rays(titan_ctx *c, const titan_rays *rays) {
    if (!c) return TITAN_E_ARG;
    return fit_set_rays(c->fit, rays, c->ord);
}

extern "C" int titan_set_rays_ext(titan_ctx *c, const titan_rays_ext *rays) {
    if (!c) return TITAN_E_ARG;
    return fit_set_rays_ext(c->fit, rays, c->ord);
}

// node layer 0 in reference numbering (Yzel::coord[0], which the set-up functions read)
static int nodes_layer0(titan_ctx *c, std::vector<double> &x) {
    x.resize(3 * c->nn);
    return titan_get_nodes(c, x.data(), 0);
}

static int fit_smooth_ctx(titan_ctx *c, int kind) {
    if (!c) return TITAN_E_ARG;
    std::vector<double> x;
    const int rc = nodes_layer0(c, x);
    if (rc != TITAN_OK) return rc;
    return fit_smooth(c->fit, kind, x.data());
}

extern "C" int titan_smooth_angle_hp(titan_ctx *c, int64_t n, const int32_t *nodes, const int32_t *nb6) {
    if (!c) return TITAN_E_ARG;
    std::vector<double> x;
    const int rc = nodes_layer0(c, x);
    if (rc != TITAN_OK) return rc;
    return fit_smooth_angle(c->fit, n, nodes, nb6, x.data());
}

extern "C" int titan_set_hp_grid(titan_ctx *c, const titan_hp_grid *g) {
    if (!c) return TITAN_E_ARG;
    return fit_set_hp_grid(c->fit, g, c->ord);
}

extern "C" int titan_circle_fit(int64_t n, const double *points, double *centre_radius) {
    return circle_fit(n, points, centre_radius);
}

extern "C" int titan_smooth_head_ts3(titan_ctx *c) { return fit_smooth_ctx(c, 1); }
extern "C" int titan_smooth_head_hp3(titan_ctx *c) { return fit_smooth_ctx(c, 2); }
extern "C" int titan_find_yzel_sosed_for_bs(titan_ctx *c) { return fit_smooth_ctx(c, 3); }

extern "C" int titan_culc_velocity_surface(titan_ctx *c, int now, double time, double *node_velocity) {
    if (!c || now < 0 || now > 1 || !(time > 0.0)) return TITAN_E_ARG;
    FitDev fd{};
    int rc = fit_prepare(c->fit, c->params, c->stream, &fd);
    if (rc != TITAN_OK) return rc;
    if (init_scalars(c->d_S2) != TITAN_OK) return TITAN_E_HIP;
    // the face and cell geometry of layer `now` (Setka::Go measures it after each move)
    launch_measure(c->stream, c->mesh(), c->d_X[now], c->geom[now], c->params.geom_mode);
    launch_fit_surface(c->stream, c->mesh(), c->dp, fd, c->d_prim[now], c->d_X[now], c->d_X[1 - now],
                       c->geom[now].cell_center, c->d_S2, time);
    HIPCHK(hipGetLastError());
    StepScalars hs;
    rc = check_scalars(c, c->d_S2, &hs);
    c->tiles.vol_cached = 0;
    if (rc != TITAN_OK) return rc;
    if (node_velocity) return fit_node_velocity(c->fit, c->stream, c->nn, node_velocity);
    return TITAN_OK;
}

extern "C" int titan_dvigenie(titan_ctx *c, int layer) {
    if (!c || layer < 0 || layer > 1 || !fit_has_rays(c->fit)) return TITAN_E_ARG;
    FitDev fd{};
    titan_params p = c->params;
    p.sglag_TS = p.sglag_HP = p.sglag_BS = 0;  // the rays need no smoothing data
    const int rc = fit_prepare(c->fit, p, c->stream, &fd);
    if (rc != TITAN_OK) return rc;
    launch_fit_rays(c->stream, fd, c->d_X[layer]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tiles.vol_cached = 0;
    return TITAN_OK;
}

extern "C" int titan_sphere_fit(int64_t n, const double *points, double *centre) {
    return sphere_fit(n, points, centre);
}

static const char *const kPlasmaKeys[10] = {"rho", "Vx", "Vy", "Vz", "p", "Bx", "By", "Bz", "Q", "rho_He"};

extern "C" int titan_save_cell_parameters(titan_ctx *c, const char *path, int n_extra,
                                          const char *const *extra_keys, const double *extra_values) {
    if (!c || !path || n_extra < 0 || (n_extra > 0 && (!extra_keys || !extra_values))) return TITAN_E_ARG;
    const int layer = c->started ? c->now2 : 0;
    const int64_t nc = c->nc;
    const int np = 8 + c->dp.n_scal;  // the plasma keys, with Q and rho_He when the step carries them
    std::vector<double> prim(8 * nc), qhe(kScal * nc), xyz(3 * c->nn);
    int rc = titan_get_state(c, prim.data(), layer);
    if (rc == TITAN_OK && c->dp.n_scal) rc = titan_get_scalars(c, qhe.data(), layer);
    if (rc == TITAN_OK) rc = titan_get_nodes(c, xyz.data(), layer);
    if (rc != TITAN_OK) return rc;
    const int nk = np + n_extra;
    std::vector<const char *> keys(nk);
    for (int k = 0; k < np; k++) keys[k] = kPlasmaKeys[k];
    for (int k = 0; k < n_extra; k++) keys[np + k] = extra_keys[k];
    std::vector<double> vals((size_t)nk * nc);
    for (int64_t i = 0; i < nc; i++) {
        std::memcpy(&vals[(size_t)nk * i], &prim[8 * i], 64);
        if (c->dp.n_scal) std::memcpy(&vals[(size_t)nk * i + 8], &qhe[kScal * i], sizeof(double) * kScal);
        if (n_extra) std::memcpy(&vals[(size_t)nk * i + np], extra_values + (size_t)n_extra * i, 8 * (size_t)n_extra);
    }
    return titan_checkpoint_write(path, nc, nk, keys.data(), vals.data(), c->nn, xyz.data());
}

extern "C" int titan_download_cell_parameters(titan_ctx *c, const char *path) {
    if (!c || !path) return TITAN_E_ARG;
    const int64_t nc = c->nc;
    const int np = 8 + c->dp.n_scal;
    std::vector<double> vals((size_t)np * nc), xyz(3 * c->nn);
    int64_t nodes_read = 0;
    int rc = titan_checkpoint_read(path, nc, np, kPlasmaKeys, vals.data(), c->nn, xyz.data(), &nodes_read);
    if (rc != TITAN_OK) return rc;
    // a key the file gives a cell replaces it in both layers (Setka_physics.cpp:4238-4247)
    std::vector<double> prim(8 * nc), qhe(kScal * nc);
    for (int layer = 0; layer < 2; layer++) {
        rc = titan_get_state(c, prim.data(), layer);
        if (rc == TITAN_OK && c->dp.n_scal) rc = titan_get_scalars(c, qhe.data(), layer);
        if (rc != TITAN_OK) return rc;
        for (int64_t i = 0; i < nc; i++) {
            for (int k = 0; k < 8; k++)
                if (!std::isnan(vals[(size_t)np * i + k])) prim[8 * i + k] = vals[(size_t)np * i + k];
            for (int k = 0; k < c->dp.n_scal; k++)
                if (!std::isnan(vals[(size_t)np * i + 8 + k])) qhe[kScal * i + k] = vals[(size_t)np * i + 8 + k];
        }
        rc = titan_set_state(c, prim.data(), layer);
        if (rc == TITAN_OK && c->dp.n_scal) rc = titan_set_scalars(c, qhe.data(), layer);
        if (rc != TITAN_OK) return rc;
    }
    if (nodes_read > 0) {  // Setka_physics.cpp:4257-4282: both node layers
        rc = titan_set_nodes(c, xyz.data(), 0);
        if (rc == TITAN_OK) rc = titan_set_nodes(c, xyz.data(), 1);
        if (rc != TITAN_OK) return rc;
        c->geom_dirty = true;
    }
    return TITAN_OK;
}

extern "C" int titan_tvd_cells(titan_ctx *c, int32_t *tvd) {
    if (!c || !tvd) return TITAN_E_ARG;
    for (int64_t d = 0; d < c->nf; d++) {
        const int64_t r = c->ord.face_d2r[d];
        for (int k = 0; k < 2; k++) {
            const int32_t v = c->tvd_raw[2 * d + k];
            tvd[2 * r + k] = v < 0 ? -1 : c->ord.cell_d2r[v];
        }
    }
    return TITAN_OK;
}

extern "C" int titan_set_timing(titan_ctx *c, int enable) {
    if (!c) return TITAN_E_ARG;
    c->timing = enable != 0;
    for (int i = 0; i < 4; i++) { c->tm.ms[i] = 0; c->tm.cnt[i] = 0; }
    return TITAN_OK;
}

extern "C" double titan_kernel_ms(titan_ctx *c, int which) {
    if (!c || which < 0 || which > 3 || c->tm.cnt[which] == 0) return -1.0;
    return c->tm.ms[which] / (double)c->tm.cnt[which];
}

// ---- standalone batches ------------------------------------------------------------------------
namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};
int dup(DevBuf &b, const void *src, size_t bytes) {
    if (!src || bytes == 0) return TITAN_OK;
    HIPCHK(hipMalloc(&b.p, bytes));
    HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return TITAN_OK;
}
int alloc(DevBuf &b, size_t bytes) {
    if (bytes == 0) return TITAN_OK;
    HIPCHK(hipMalloc(&b.p, bytes));
    return TITAN_OK;
}
}  // namespace

extern "C" int titan_chlld_batch(int64_t n, const int8_t *n_state, const int8_t *n_disc, const uint8_t *flags,
                                 const double *normal, const double *w, const double *qL, const double *qR, int K,
                                 const double *kL, const double *kR, double *q, double *k, double *ds,
                                 uint8_t *branch, int precision) {
    if (n < 0 || K < 0 || K > 4 || !n_state || !normal || !qL || !qR || !q) return TITAN_E_ARG;
    if (K > 0 && (!kL || !kR || !k)) return TITAN_E_ARG;
    if (n == 0) return TITAN_OK;
    for (int64_t i = 0; i < n; i++)
        if (n_state[i] < 0 || n_state[i] > 3 || (n_disc && (n_disc[i] < 0 || n_disc[i] > 3))) return TITAN_E_ARG;
    DevBuf dns, dnd, dfl, dn, dw, dqL, dqR, dkL, dkR, dq, dk, dds, dbr, derr;
    int rc = TITAN_OK;
    auto ck = [&](int r) { if (rc == TITAN_OK) rc = r; };
    ck(dup(dns, n_state, n));
    ck(dup(dnd, n_disc, n_disc ? n : 0));
    ck(dup(dfl, flags, flags ? n : 0));
    ck(dup(dn, normal, 24 * n));
    ck(dup(dw, w, w ? 8 * n : 0));
    ck(dup(dqL, qL, 64 * n));
    ck(dup(dqR, qR, 64 * n));
    ck(dup(dkL, kL, 8 * K * n));
    ck(dup(dkR, kR, 8 * K * n));
    ck(alloc(dq, 64 * n));
    ck(alloc(dk, 8 * K * n));
    ck(alloc(dds, 24 * n));
    ck(alloc(dbr, n));
    ck(alloc(derr, sizeof(int)));
    if (rc != TITAN_OK) return rc;
    HIPCHK(hipMemset(derr.p, 0, sizeof(int)));
    launch_chlld_batch(nullptr, precision, n, K, (int8_t *)dns.p, (int8_t *)dnd.p, (uint8_t *)dfl.p, (double *)dn.p,
                       (double *)dw.p, (double *)dqL.p, (double *)dqR.p, (double *)dkL.p, (double *)dkR.p,
                       (double *)dq.p, (double *)dk.p, (double *)dds.p, (uint8_t *)dbr.p, (int *)derr.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(q, dq.p, 64 * n, hipMemcpyDeviceToHost));
    if (K > 0) HIPCHK(hipMemcpy(k, dk.p, 8 * K * n, hipMemcpyDeviceToHost));
    if (ds) HIPCHK(hipMemcpy(ds, dds.p, 24 * n, hipMemcpyDeviceToHost));
    if (branch) HIPCHK(hipMemcpy(branch, dbr.p, n, hipMemcpyDeviceToHost));
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost));
    return herr ? TITAN_E_RIEMANN : TITAN_OK;
}

extern "C" int titan_cx_exact_batch(int64_t n, const double *in10, int nodes, double *out5) {
    if (n < 0 || !in10 || !out5 || nodes < 1 || nodes > 4096) return TITAN_E_ARG;
    if (n == 0) return TITAN_OK;
    DevBuf din, dout;
    int rc = dup(din, in10, 80 * n);
    if (rc == TITAN_OK) rc = alloc(dout, 40 * n);
    if (rc != TITAN_OK) return rc;
    launch_cx_exact(nullptr, n, (double *)din.p, nodes, (double *)dout.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(out5, dout.p, 40 * n, hipMemcpyDeviceToHost));
    return TITAN_OK;
}

extern "C" int titan_sources(titan_ctx *c, int layer, int mode, double *out) {
    if (!c || layer < 0 || layer > 1 || !out) return TITAN_E_ARG;
    int rc = compute_sources(c, layer, mode);
    if (rc != TITAN_OK) return rc;
    std::vector<double> h(5 * c->nc);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_src, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t d = 0; d < c->nc; d++)
        for (int k = 0; k < 5; k++) out[5 * (int64_t)c->ord.cell_d2r[d] + k] = h[5 * d + k];
    return TITAN_OK;
}

extern "C" int titan_sources_h(titan_ctx *c, int layer, const int8_t *arise, double *out, double *out_h) {
    if (!c || layer < 0 || layer > 1 || !out_h || !c->d_hfluid) return TITAN_E_ARG;
    const int n_h = c->n_hfluid;
    if (n_h < 1 || n_h > TITAN_MAX_H_FLUIDS) return TITAN_E_ARG;
    if (arise)
        for (int i = 0; i < 4 * n_h; i++)
            if (arise[i] < 1 || arise[i] > n_h) return TITAN_E_ARG;
    const int prc = prepare_sources(c, TITAN_SRC_BERA);
    if (prc != TITAN_OK) return prc;
    const int64_t nh5 = (int64_t)5 * n_h * c->nc;
    DevBuf d_h, d_ar;
    int rc = alloc(d_h, sizeof(double) * nh5);
    if (rc == TITAN_OK && arise) rc = dup(d_ar, arise, 4 * n_h);
    if (rc != TITAN_OK) return rc;
    const titan_params &p = c->params;
    launch_bera_cells_h(c->stream, c->nc, c->d_prim[layer], c->d_scal[layer], c->d_cell_code, c->d_hfluid, n_h,
                        p.par_a_2, p.par_n_H_LISM / p.par_Kn, p.par_n_H_LISM, (const int8_t *)d_ar.p, c->d_src,
                        (double *)d_h.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nh5), s5(5 * c->nc);
    HIPCHK(hipMemcpyAsync(h.data(), d_h.p, nh5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(s5.data(), c->d_src, s5.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t d = 0; d < c->nc; d++) {
        const int64_t r = c->ord.cell_d2r[d];
        for (int k = 0; k < 5 * n_h; k++) out_h[5 * n_h * r + k] = h[5 * n_h * d + k];
        if (out)
            for (int k = 0; k < 5; k++) out[5 * r + k] = s5[5 * d + k];
    }
    return TITAN_OK;
}

// ---- Monte-Carlo neutral densities (kernels in k_mk.hip) --------------------------------------
// Get_Spotok_inf (Setka_MK.cpp:2316-2324): inflow of the unit-temperature LISM Maxwellian drifting
// at (v_inf, 0, 0) through a face of outward normal n, per unit area
static double mk_spotok_inf(double v_inf, const double *n) {
    const double sqrtpi = 1.77245385;  // sqrtpi_, Header.h:90
    const double Ux = v_inf * n[0] + 0.0 * n[1] + 0.0 * n[2];
    return -(-std::exp(-(Ux * Ux)) + sqrtpi * Ux * std::erfc(Ux)) / (2.0 * sqrtpi);
}

// The injections of one MK_go call: faces (device ids), particles per face and weights, and in the
// zone form each one's sort, side (-1: the boundary Maxwellian) and the offset of its distribution's
// running sums.
struct MkInj {
    std::vector<int32_t> face;
    std::vector<int64_t> first{0};
    std::vector<double> mu;
    std::vector<int8_t> sort, side;
    std::vector<int64_t> cdf;
};

// Particles per face and weight (Setka_MK.cpp:1233,1380-1390 with N_per_gran = n_per_face)
static int mk_count(double sv, double mu_expect, int32_t n_per_face, int32_t *n) {
    const double q = sv / mu_expect;
    if (!(q < 2.0e9)) return TITAN_E_ARG;
    *n = (int32_t)q + 1;
    const int32_t lo = n_per_face < 100 ? n_per_face : 100;
    if (*n < lo) *n = lo;
    return TITAN_OK;
}

// Normal and area of the given faces at the geometry of `layer` (measured here: the fused step
// keeps only the cell centres current).
static int mk_face_na(titan_ctx *c, int layer, const std::vector<int32_t> &faces, std::vector<double> &na) {
    const int64_t nb = (int64_t)faces.size();
    na.assign(4 * nb, 0.0);
    if (nb == 0) return TITAN_OK;
    DevBuf dcand, dna;
    int rc = dup(dcand, faces.data(), sizeof(int32_t) * nb);
    if (rc == TITAN_OK) rc = alloc(dna, sizeof(double) * 4 * nb);
    if (rc != TITAN_OK) return rc;
    launch_mk_gather(c->stream, nb, (const int32_t *)dcand.p, c->geom[layer], (double *)dna.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(na.data(), dna.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

// The flights of one MK_go call (k_mk_particles), densities per unit volume in reference order.
static int mk_flights(titan_ctx *c, int layer, const MkInj &inj, int zone, const double *d_cdf, uint64_t seed,
                      int32_t max_steps, std::vector<double> &dens, int64_t st[9]) {
    const int64_t nc = c->nc;
    MkDev K{};
    K.n_in = (int64_t)inj.face.size();
    K.n_particles = inj.first.back();
    if (K.n_particles > (int64_t)1 << 38) return TITAN_E_ARG;  // grid of at most 2^30 workgroups
    K.seed = seed;
    K.v_inf = c->params.v_inf;
    K.a2 = c->params.par_a_2;
    K.kn = c->params.par_Kn;
    K.max_steps = max_steps;
    K.R0 = c->mk_R0;
    K.n_inner = (int32_t)c->mk_inner.size();
    DevBuf dface, dfirst, dmu, dinner, dinner_nb, dsort, dside, dcdf;
    int rc = TITAN_OK;
    if (K.n_inner > 0) {
        rc = dup(dinner, c->mk_inner.data(), sizeof(int32_t) * c->mk_inner.size());
        if (rc == TITAN_OK) rc = dup(dinner_nb, c->mk_inner_nb.data(), sizeof(int32_t) * c->mk_inner_nb.size());
        if (rc != TITAN_OK) return rc;
        K.inner_face = (const int32_t *)dinner.p;
        K.inner_nb = (const int32_t *)dinner_nb.p;
    }
    if (K.n_in > 0) {
        rc = dup(dface, inj.face.data(), sizeof(int32_t) * inj.face.size());
        if (rc == TITAN_OK) rc = dup(dfirst, inj.first.data(), sizeof(int64_t) * inj.first.size());
        if (rc == TITAN_OK) rc = dup(dmu, inj.mu.data(), sizeof(double) * inj.mu.size());
        if (rc == TITAN_OK && zone > 0) rc = dup(dsort, inj.sort.data(), inj.sort.size());
        if (rc == TITAN_OK && zone > 0) rc = dup(dside, inj.side.data(), inj.side.size());
        if (rc == TITAN_OK && zone > 0) rc = dup(dcdf, inj.cdf.data(), sizeof(int64_t) * inj.cdf.size());
    }
    if (rc != TITAN_OK) return rc;
    if (!c->d_mk) {  // flight records, tallies, densities and counters: kept for later calls
        const size_t bytes = sizeof(MkCell) * nc + sizeof(double) * 10 * nc + sizeof(unsigned long long) * 9;
        HIPCHK(hipMalloc((void **)&c->d_mk, bytes));
    }
    MkCell *pack = (MkCell *)c->d_mk;
    double *tally = (double *)(pack + nc);
    double *out = tally + 5 * nc;
    unsigned long long *dst = (unsigned long long *)(out + 5 * nc);
    HIPCHK(hipMemsetAsync(tally, 0, sizeof(double) * 5 * nc, c->stream));
    HIPCHK(hipMemsetAsync(dst, 0, sizeof(unsigned long long) * 9, c->stream));
    K.face = (const int32_t *)dface.p;
    K.first = (const int64_t *)dfirst.p;
    K.mu = (const double *)dmu.p;
    K.tally = tally;
    K.out = out;
    K.stats = dst;
    K.pack = pack;
    K.scal = c->d_scal[layer];  // null unless the passive scalars are carried
    c->pui_valid = false;
    if (c->pui_nw > 0) {  // MK_source_S: the tallies start from zero (Cell::Init_S)
        const size_t cnt = (size_t)3 * c->pui_nw * nc;  // S-, S+; then the angle-integral table
        if (!c->d_pui)
            HIPCHK(hipMalloc((void **)&c->d_pui, sizeof(double) * (cnt + (size_t)c->pui_nw * c->pui_nw)));
        if (!c->d_pui_trig) {
            // the 40 angles of Cell::MK_calc_Sm, sin and cos as the host's libm gives them
            double tr[80];
            const double dthe = 3.14159265358979323846 / 40.0;
            for (int k = 1; k <= 40; k++) {
                const double the = dthe * k;
                tr[k - 1] = std::sin(the);
                tr[40 + k - 1] = std::cos(the);
            }
            HIPCHK(hipMalloc((void **)&c->d_pui_trig, sizeof(tr)));
            HIPCHK(hipMemcpyAsync(c->d_pui_trig, tr, sizeof(tr), hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK(hipMemsetAsync(c->d_pui, 0, sizeof(double) * cnt, c->stream));
        K.pui_nw = c->pui_nw;
        K.pui_wr = c->pui_wr;
        for (int i = 0; i < 16; i++) K.pui_arise[i] = c->pui_arise[i];
        K.pui_sm = c->d_pui;
        K.pui_sp = c->d_pui + (size_t)c->pui_nw * nc;
    }
    if (zone > 0) {
        K.zone = zone;
        K.r = c->mkz_r;
        K.cell_mk = c->d_mkz_cell;
        K.face_mask = c->d_mkz_mask;
        K.dslot = c->d_mkz_slot;
        K.dist = c->d_mkz_w;
        K.in_sort = (const int8_t *)dsort.p;
        K.in_side = (const int8_t *)dside.p;
        K.in_cdf = (const int64_t *)dcdf.p;
        K.cdf = d_cdf;
    }
    launch_mk_pack(c->stream, c->mesh(), c->d_X[layer], pack);
    launch_mk_particles(c->stream, c->mesh(), K, c->d_X[layer], c->d_prim[layer], c->geom[layer]);
    if (K.pui_nw > 0)
        launch_mk_pui_finish(c->stream, nc, K.pui_nw, K.pui_wr, c->params.par_a_2, c->params.par_Kn,
                             c->geom[layer].cell_volume, c->d_pui_trig, c->d_pui + (size_t)3 * K.pui_nw * nc, K.pui_sm,
                             K.pui_sp);
    HIPCHK(hipGetLastError());
    dens.resize(5 * nc);
    unsigned long long h[9];
    HIPCHK(hipMemcpyAsync(dens.data(), out, sizeof(double) * 5 * nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h, dst, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 9; i++) st[i] = (int64_t)h[i];
    c->pui_valid = K.pui_nw > 0;
    return TITAN_OK;
}

constexpr int32_t kMaxPuiBins = 4096;

extern "C" int titan_set_mk_source_s(titan_ctx *c, int32_t n_w, double w_r, const int8_t *arise) {
    if (!c || n_w < 0 || n_w > kMaxPuiBins || (n_w > 0 && !(w_r > 0.0 && std::isfinite(w_r)))) return TITAN_E_ARG;
    // proton_arise_1..4 of the run with pickup ions (Phys_param.cpp:93-126), parent 0, sorts 1..4
    static const int8_t kArise[16] = {0, 1, 1, 1, 2, 0, 2, 2, 1, 1, 0, 1, 1, 1, 1, 0};
    const int8_t *a = arise ? arise : kArise;
    for (int i = 0; i < 16; i++)
        if (a[i] < 0 || a[i] > 2) return TITAN_E_ARG;
    if (n_w != c->pui_nw && c->d_pui) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->d_pui));
        c->d_pui = nullptr;
    }
    c->pui_nw = n_w;
    c->pui_wr = w_r;
    for (int i = 0; i < 16; i++) c->pui_arise[i] = a[i];
    c->pui_valid = false;
    return TITAN_OK;
}

extern "C" int titan_mk_get_s(titan_ctx *c, double *sm, double *sp) {
    if (!c || !sm || !sp || c->pui_nw <= 0) return TITAN_E_ARG;
    if (!c->pui_valid) return TITAN_E_ARG;  // no Monte-Carlo call with the tallies on yet
    const int64_t nc = c->nc, nw = c->pui_nw;
    std::vector<double> h((size_t)3 * nw * nc);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_pui, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double *hs = h.data(), *hp = h.data() + nw * nc;
    for (int64_t d = 0; d < nc; d++) {
        const int64_t r = c->ord.cell_d2r[d];
        std::memcpy(sm + r * nw, hs + d * nw, sizeof(double) * nw);
        std::memcpy(sp + r * nw * 2, hp + d * nw * 2, sizeof(double) * nw * 2);
    }
    return TITAN_OK;
}

extern "C" int titan_mk_s_finish_batch(int64_t n, int32_t n_w, double w_r, double a2, double kn, const double *vol,
                                       const double *sm, const double *sp, double *sm_out, double *sp_out) {
    if (n < 0 || n_w < 1 || n_w > kMaxPuiBins || !(w_r > 0.0) || !vol || !sm || !sp || !sm_out || !sp_out)
        return TITAN_E_ARG;
    if (n == 0) return TITAN_OK;
    double tr[80];
    const double dthe = 3.14159265358979323846 / 40.0;
    for (int k = 1; k <= 40; k++) {
        const double the = dthe * k;
        tr[k - 1] = std::sin(the);
        tr[40 + k - 1] = std::cos(the);
    }
    DevBuf dv, dsm, dsp, dtr, dtt;
    int rc = dup(dv, vol, sizeof(double) * n);
    if (rc == TITAN_OK) rc = dup(dsm, sm, sizeof(double) * n * n_w);
    if (rc == TITAN_OK) rc = dup(dsp, sp, sizeof(double) * n * n_w * 2);
    if (rc == TITAN_OK) rc = dup(dtr, tr, sizeof(tr));
    if (rc == TITAN_OK) rc = alloc(dtt, sizeof(double) * n_w * n_w);
    if (rc != TITAN_OK) return rc;
    launch_mk_pui_finish(nullptr, n, n_w, w_r, a2, kn, (const double *)dv.p, (const double *)dtr.p, (double *)dtt.p,
                         (double *)dsm.p, (double *)dsp.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(sm_out, dsm.p, sizeof(double) * n * n_w, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sp_out, dsp.p, sizeof(double) * n * n_w * 2, hipMemcpyDeviceToHost));
    return TITAN_OK;
}

extern "C" int titan_mk_go(titan_ctx *c, int layer, int32_t n_per_face, uint64_t seed, int32_t max_steps,
                           double *density, int64_t *stats) {
    if (!c || layer < 0 || layer > 1 || n_per_face < 1 || max_steps < 1 || !density) return TITAN_E_ARG;
    if (!(c->params.par_Kn > 0.0)) return TITAN_E_ARG;
    const int64_t nc = c->nc, nf = c->nf;
    // candidate injection faces: the outer boundary (bc 2, 3), ascending reference face id (the
    // boundary codes are fixed at titan_create, so the list is found once per context)
    if (!c->mk_cand_ready) {
        std::vector<uint32_t> code(nf);
        HIPCHK(hipMemcpyAsync(code.data(), c->d_face_code, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->mk_cand.clear();
        for (int64_t r = 0; r < nf; r++) {
            const int32_t d = c->ord.face_r2d[r];
            const uint32_t bc = code[d] & 15u;
            if (bc == TITAN_BC_OUTER_HARD || bc == TITAN_BC_OUTER_SOFT) c->mk_cand.push_back(d);
        }
        c->mk_cand_ready = true;
    }
    const std::vector<int32_t> &cand = c->mk_cand;
    int64_t st[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!cand.empty()) {
        const int64_t nb = (int64_t)cand.size();
        launch_measure(c->stream, c->mesh(), c->d_X[layer], c->geom[layer], c->params.geom_mode);
        std::vector<double> na;
        int rc = mk_face_na(c, layer, cand, na);
        if (rc != TITAN_OK) return rc;
        const double v_inf = c->params.v_inf;
        double total = 0.0;
        for (int64_t i = 0; i < nb; i++) total += mk_spotok_inf(v_inf, &na[4 * i]) * na[4 * i + 3];
        const double mu_expect = total / (1.0 * n_per_face * nb);
        MkInj inj;
        for (int64_t i = 0; i < nb; i++) {
            const double sv = mk_spotok_inf(v_inf, &na[4 * i]) * na[4 * i + 3];
            if (sv < 0.0000001 * total / nb) continue;
            int32_t n = 0;
            if (mk_count(sv, mu_expect, n_per_face, &n) != TITAN_OK) return TITAN_E_ARG;
            inj.face.push_back(cand[i]);
            inj.first.push_back(inj.first.back() + n);
            inj.mu.push_back(sv / n);
        }
        std::vector<double> dens;
        rc = mk_flights(c, layer, inj, 0, nullptr, seed, max_steps, dens, st);
        if (rc != TITAN_OK) return rc;
        std::memcpy(density, dens.data(), sizeof(double) * 5 * nc);
    } else {
        std::fill(density, density + 5 * nc, 0.0);
    }
    if (stats)
        for (int i = 0; i < 8; i++) stats[i] = st[i];
    return TITAN_OK;
}

// ---- Monte-Carlo zones ------------------------------------------------------------------------
// Setka::Set_MK_Zone (Setka_MK.cpp:142-683) from the centres of node layer 0: the MK zone of each
// cell and the zones whose lists hold each face; then the distributions of the listed interior
// faces (zeroed).
extern "C" int titan_set_mk_zone(titan_ctx *c, int32_t r, int8_t *cell_mk, uint8_t *face_mask) {
    if (!c || r < 1 || r > 4) return TITAN_E_ARG;
    const int64_t nc = c->nc, nf = c->nf;
    launch_measure(c->stream, c->mesh(), c->d_X[0], c->geom[0], c->params.geom_mode);
    std::vector<double> cc(3 * nc), fcen(3 * nf);
    std::vector<uint32_t> ccode(nc), fcode(nf);
    std::vector<int32_t> fcells(2 * nf);
    HIPCHK(hipMemcpyAsync(cc.data(), c->geom[0].cell_center, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fcen.data(), c->geom[0].face_center, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ccode.data(), c->d_cell_code, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fcode.data(), c->d_face_code, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fcells.data(), c->d_face_cells, sizeof(int32_t) * 2 * nf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int8_t> cm(nc);
    for (int64_t d = 0; d < nc; d++) {
        const uint32_t z = (ccode[d] >> 8) & 15u;
        const double x = cc[3 * d];
        if (z == 1) cm[d] = 1;
        else if (z == 2) cm[d] = x > 0 ? 2 : 3;
        else if (z == 3) cm[d] = x > 0 ? 4 : 5;
        else if (z == 4) cm[d] = x > 0 ? 6 : 7;
        else return set_err(c, TITAN_E_TOPOLOGY, c->ord.cell_d2r[d]);  // Error 86743207564
    }
    std::vector<uint8_t> mask(nf, 0);
    std::vector<int32_t> slot(nf, -1);
    int64_t n_slots = 0;
    for (int64_t d = 0; d < nf; d++) {
        const uint32_t bc = fcode[d] & 15u, fl = (fcode[d] >> 8) & 255u;
        const bool ts = (fl >> 3) & 1u, hp = (fl >> 2) & 1u, bs = (fl >> 4) & 1u;
        const bool us2 = !ts && !hp && !bs;  // Gran::type2 == Us
        const int64_t a = fcells[2 * d], b = fcells[2 * d + 1];
        const bool two = b >= 0;  // gr->cells.size() == 2
        const uint32_t za = (ccode[a] >> 8) & 15u, zb = two ? (ccode[b] >> 8) & 15u : 0u;
        const double X = fcen[3 * d];
        const double xa = cc[3 * a], xb = two ? cc[3 * b] : 0.0;
        const bool across = two && xa * xb < 0.0;
        int n[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ts) n[0]++;                               // zone 1: Gran_TS (Setka_MK.cpp:356-361)
        if (!(X < -0.000001)) {                      // zone 2 (:364-409)
            if (hp) n[1]++;
            if (ts) n[1]++;
            if (bc == TITAN_BC_OUTER_SOFT && za == 2) n[1]++;
            if (two && za == 2 && zb == 2 && across) n[1]++;
        }
        if (!(X > 0.000001)) {                       // zone 3 (:412-462)
            if (hp) n[2]++;
            if (ts) n[2]++;
            if (bc == TITAN_BC_OUTER_SOFT && za == 2) n[2]++;
            if (two && us2) {
                if (za == 2 && zb == 3) n[2]++;
                if (za == 2 && zb == 2 && across) n[2]++;
            }
        }
        if (hp && X > 0) n[3]++;                      // zone 4 (:465-500)
        if (bs) n[3]++;
        if (two && za == 3 && zb == 3 && across) n[3]++;
        if (hp && X < 0.0) n[4]++;                    // zone 5 (:503-557)
        if (bc == TITAN_BC_OUTER_SOFT && za == 3) n[4]++;
        if (two && us2) {
            if (za == 2 && zb == 3) n[4]++;
            if (za == 3 && zb == 3 && across) n[4]++;
            if (za == 3 && zb == 4 && xa < 0) n[4]++;
        }
        if (bs) n[5]++;                               // zone 6 (:560-592)
        if (bc == TITAN_BC_OUTER_HARD && X > 0) n[5]++;
        if (two && za == 4 && zb == 4 && across) n[5]++;
        if (bc != TITAN_BC_INTERIOR && X < 0 && za == 4) n[6]++;  // zone 7 (:595-631)
        if (two && us2) {
            if (za == 3 && zb == 4 && xa < 0) n[6]++;
            if (za == 4 && zb == 4 && across) n[6]++;
        }
        for (int z = 0; z < 7; z++) {
            if (n[z] > 1) return set_err(c, TITAN_E_TOPOLOGY, c->ord.face_d2r[d]);  // ERROR 6435856408
            if (n[z]) mask[d] |= (uint8_t)(1u << z);
        }
        if (mask[d] && bc == TITAN_BC_INTERIOR && two) slot[d] = (int32_t)n_slots++;
    }
    const size_t nbin = 108 * (size_t)r * r * r;
    const size_t wbytes = sizeof(double) * 8 * nbin * (size_t)std::max<int64_t>(n_slots, 1);
    for (void *q : {(void *)c->d_mkz_cell, (void *)c->d_mkz_mask, (void *)c->d_mkz_slot, (void *)c->d_mkz_w})
        if (q) hipFree(q);
    c->d_mkz_cell = nullptr; c->d_mkz_mask = nullptr; c->d_mkz_slot = nullptr; c->d_mkz_w = nullptr;
    c->mkz_r = 0;
    HIPCHK(hipMalloc((void **)&c->d_mkz_w, wbytes));
    HIPCHK(hipMemsetAsync(c->d_mkz_w, 0, wbytes, c->stream));
    if (dev_upload(&c->d_mkz_cell, cm, c->stream) != TITAN_OK || dev_upload(&c->d_mkz_mask, mask, c->stream) != TITAN_OK ||
        dev_upload(&c->d_mkz_slot, slot, c->stream) != TITAN_OK)
        return TITAN_E_HIP;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mkz_cell = std::move(cm);
    c->mkz_mask = std::move(mask);
    c->mkz_slot = std::move(slot);
    c->mkz_fcode = std::move(fcode);
    c->mkz_fcells = std::move(fcells);
    c->mkz_slots = n_slots;
    c->mkz_r = r;
    if (cell_mk)
        for (int64_t d = 0; d < nc; d++) cell_mk[c->ord.cell_d2r[d]] = c->mkz_cell[d];
    if (face_mask)
        for (int64_t d = 0; d < nf; d++) face_mask[c->ord.face_d2r[d]] = c->mkz_mask[d];
    return TITAN_OK;
}

// MK_prepare + MK_go(zone) + MK_delete (Setka_MK.cpp:686-1752, Setka.cpp:221-226) on the state and
// nodes of `layer`.
extern "C" int titan_mk_go_zone(titan_ctx *c, int layer, int32_t zone, int32_t n_per_face, uint64_t seed,
                                int32_t max_steps, double *density, int64_t *stats) {
    if (!c || layer < 0 || layer > 1 || zone < 1 || zone > 7 || n_per_face < 1 || max_steps < 1 || !density)
        return TITAN_E_ARG;
    if (!(c->params.par_Kn > 0.0) || c->mkz_r == 0) return TITAN_E_ARG;
    const int64_t nc = c->nc, nf = c->nf;
    const int64_t nbin = 108 * (int64_t)c->mkz_r * c->mkz_r * c->mkz_r;
    const uint8_t bit = (uint8_t)(1u << (zone - 1));
    // the zone's list in ascending reference face id, each face's side entering the zone
    std::vector<int32_t> list, bfaces;
    std::vector<int8_t> ni2;
    std::vector<int64_t> grp;       // running-sum groups: offsets of the incoming distributions
    std::vector<int64_t> clr;       // the outgoing ones, cleared before the flights
    for (int64_t r = 0; r < nf; r++) {
        const int32_t d = c->ord.face_r2d[r];
        if (!(c->mkz_mask[d] & bit)) continue;
        const int32_t a = c->mkz_fcells[2 * d], b = c->mkz_fcells[2 * d + 1];
        const int8_t side = c->mkz_cell[a] == zone ? 1 : 0;
        const int32_t enter = side ? a : b;
        if (enter < 0 || c->mkz_cell[enter] != zone) return set_err(c, TITAN_E_TOPOLOGY, r);  // Error 9767653421
        list.push_back(d);
        ni2.push_back(side);
        if ((c->mkz_fcode[d] & 15u) != TITAN_BC_INTERIOR) bfaces.push_back(d);
        const int32_t sl = c->mkz_slot[d];
        if (sl >= 0)
            for (int s = 0; s < 4; s++) {
                grp.push_back((((int64_t)sl * 4 + s) * 2 + side) * nbin);
                clr.push_back((((int64_t)sl * 4 + s) * 2 + (1 - side)) * nbin);
            }
    }
    int64_t st[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (list.empty()) {
        for (int64_t d = 0; d < nc; d++)
            if (c->mkz_cell[d] == zone)
                for (int k = 0; k < 5; k++) density[5 * (int64_t)c->ord.cell_d2r[d] + k] = 0.0;
        if (stats)
            for (int i = 0; i < 9; i++) stats[i] = 0;
        return TITAN_OK;
    }
    launch_measure(c->stream, c->mesh(), c->d_X[layer], c->geom[layer], c->params.geom_mode);
    // running sums of the incoming distributions; their totals are the fluxes (Culk_SpotokV)
    const int64_t ng = (int64_t)grp.size();
    DevBuf dgrp, dclr, dcdf;
    int rc = dup(dgrp, grp.data(), sizeof(int64_t) * ng);
    if (rc == TITAN_OK) rc = dup(dclr, clr.data(), sizeof(int64_t) * clr.size());
    if (rc == TITAN_OK) rc = alloc(dcdf, sizeof(double) * (size_t)std::max<int64_t>(ng, 1) * nbin);
    if (rc != TITAN_OK) return rc;
    std::vector<double> gsum(ng);
    if (ng > 0) {
        launch_mkz_cdf(c->stream, ng, (const int64_t *)dgrp.p, c->d_mkz_w, nbin, (double *)dcdf.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy2DAsync(gsum.data(), sizeof(double), (const double *)dcdf.p + (nbin - 1), sizeof(double) * nbin,
                                sizeof(double), ng, hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<double> na;
    rc = mk_face_na(c, layer, bfaces, na);  // synchronises the stream
    if (rc != TITAN_OK) return rc;
    // fluxes into the zone per face and sort (MK_prepare, Setka_MK.cpp:851-898)
    const int64_t nl = (int64_t)list.size();
    std::vector<double> ff(nl, 0.0), fs(4 * nl, 0.0);
    std::vector<int64_t> gof(nl, -1);
    const double v_inf = c->params.v_inf;
    double S = 0.0;
    for (int64_t i = 0, ib = 0, ig = 0; i < nl; i++) {
        const int32_t d = list[i];
        if ((c->mkz_fcode[d] & 15u) == TITAN_BC_INTERIOR) {
            if (c->mkz_slot[d] >= 0) {
                gof[i] = ig;
                for (int s = 0; s < 4; s++) {
                    fs[4 * i + s] = gsum[ig + s];
                    ff[i] += gsum[ig + s];
                }
                ig += 4;
            }
        } else {
            const double v = mk_spotok_inf(v_inf, &na[4 * ib]) * na[4 * ib + 3];
            fs[4 * i + 3] = v;
            ff[i] = v;
            ib++;
        }
        S += ff[i];
    }
    const double mu_expect = S / (1.0 * n_per_face * nl);
    MkInj inj;
    for (int s = 1; s <= 4 && S > 0.0; s++)  // no inflow at all (no pass has fed the zone yet): no particles
        for (int64_t i = 0; i < nl; i++) {
            if (ff[i] < 0.0000001 * S / nl) continue;
            const double sv = fs[4 * i + s - 1];
            if (sv < 0.0000001 * ff[i]) continue;
            int32_t n = 0;
            if (mk_count(sv, mu_expect, n_per_face, &n) != TITAN_OK) return TITAN_E_ARG;
            const bool interior = (c->mkz_fcode[list[i]] & 15u) == TITAN_BC_INTERIOR;
            inj.face.push_back(list[i]);
            inj.first.push_back(inj.first.back() + n);
            inj.mu.push_back(sv / n);
            inj.sort.push_back((int8_t)s);
            inj.side.push_back(interior ? ni2[i] : (int8_t)-1);
            inj.cdf.push_back(interior ? (gof[i] + s - 1) * nbin : (int64_t)-1);
        }
    // the outgoing distributions of the zone's faces start empty (MK_go's Fill_null)
    launch_mkz_clear(c->stream, (int64_t)clr.size(), (const int64_t *)dclr.p, c->d_mkz_w, nbin);
    HIPCHK(hipGetLastError());
    std::vector<double> dens;
    rc = mk_flights(c, layer, inj, zone, (const double *)dcdf.p, seed, max_steps, dens, st);
    if (rc != TITAN_OK) return rc;
    // the zone's cells get their densities (Cell::MK_normir_Moments over the zone, Setka_MK.cpp:1674-1694)
    for (int64_t d = 0; d < nc; d++)
        if (c->mkz_cell[d] == zone) {
            const int64_t r = c->ord.cell_d2r[d];
            for (int k = 0; k < 5; k++) density[5 * r + k] = dens[5 * r + k];
        }
    if (stats)
        for (int i = 0; i < 9; i++) stats[i] = st[i];
    return TITAN_OK;
}

// One face distribution (reference face id, sort 1..4, side 0 / 1) to or from the caller: w [108 r^3].
static int mk_dist_copy(titan_ctx *c, int64_t face, int32_t sort, int32_t side, double *w, bool set) {
    if (!c || !w || c->mkz_r == 0 || face < 0 || face >= c->nf || sort < 1 || sort > 4 || side < 0 || side > 1)
        return TITAN_E_ARG;
    const int32_t sl = c->mkz_slot[c->ord.face_r2d[face]];
    if (sl < 0) return TITAN_E_ARG;
    const int64_t nbin = 108 * (int64_t)c->mkz_r * c->mkz_r * c->mkz_r;
    double *p = c->d_mkz_w + (((int64_t)sl * 4 + (sort - 1)) * 2 + side) * nbin;
    if (set) HIPCHK(hipMemcpyAsync(p, w, sizeof(double) * nbin, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemcpyAsync(w, p, sizeof(double) * nbin, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TITAN_OK;
}

extern "C" int titan_mk_dist_get(titan_ctx *c, int64_t face, int32_t sort, int32_t side, double *w) {
    return mk_dist_copy(c, face, sort, side, w, false);
}

extern "C" int titan_mk_dist_set(titan_ctx *c, int64_t face, int32_t sort, int32_t side, const double *w) {
    return mk_dist_copy(c, face, sort, side, const_cast<double *>(w), true);
}

// ---- surrogate network entry points (kernels in k_mlp.hip) ----------------------------------
extern "C" int64_t titan_mlp_n_params(void) { return kMlpParams; }

extern "C" int titan_mlp_set(titan_ctx *c, const float *params, const float *in_norm16, const float *out_scale5,
                             const float *box16) {
    if (!c || !params || !in_norm16 || !out_scale5 || !box16) return TITAN_E_ARG;
    return mlp_set(c->mlp, c->stream, params, in_norm16, out_scale5, box16);
}

extern "C" int titan_mlp_get(titan_ctx *c, float *params) {
    if (!c || !params) return TITAN_E_ARG;
    return mlp_get(c->mlp, c->stream, params);
}

extern "C" int titan_mlp_forward(titan_ctx *c, int64_t n, const float *x, float *y) {
    if (!c || n < 0 || !x || !y) return TITAN_E_ARG;
    return mlp_forward_host(c->mlp, c->stream, n, x, y);
}

extern "C" int titan_mlp_loss_grad(titan_ctx *c, int64_t n, const float *x, const float *y, float *loss,
                                   float *grad) {
    if (!c || n <= 0 || !x || !y || !loss || !grad) return TITAN_E_ARG;
    return mlp_loss_grad_host(c->mlp, c->stream, n, x, y, loss, grad);
}

extern "C" int titan_mlp_train(titan_ctx *c, int64_t n, const float *x, const float *y, int epochs, int batch,
                               float lr, uint64_t seed, float *loss_hist) {
    if (!c || n <= 0 || !x || !y || epochs < 0 || batch <= 0) return TITAN_E_ARG;
    return mlp_train_host(c->mlp, c->stream, n, x, y, epochs, batch, lr, seed, loss_hist);
}

extern "C" int titan_tile_info(titan_ctx *c, int64_t *info8) {
    if (!c || !info8) return TITAN_E_ARG;
    for (int i = 0; i < 8; i++) info8[i] = c->tile_info[i];
    return TITAN_OK;
}
