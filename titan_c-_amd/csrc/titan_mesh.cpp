// titan_mesh.cpp -- host-side mesh indexing: New_connect face numbering, synthetic generators,
// Morton (space-filling-curve) renumbering of cells, faces and nodes for the device.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/titan_gpu.h"
#include "device_mesh.h"
#include "titan_mesh.h"

namespace titan {

// Local faces of a hexahedron in the reference's order and node order (Setka.cpp:2086-2130).
static const int kLocalFaces[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                      {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};

int64_t connect_faces(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes, int32_t *face_nodes,
                      int32_t *face_cells, int32_t *cell_faces) {
    // Bucket the 6*Nc local faces by their smallest node id (counting sort), find each face's
    // twin (same node set) inside its bucket, then number faces in creation order keeping the
    // first copy: cells[0] = first creator, cells[1] = second (Setka.cpp:2154-2201).
    const int64_t nt = 6 * n_cells;
    std::vector<int32_t> minnode(nt);
    std::vector<uint64_t> key_hi(nt);  // sorted ids packed: (k1,k2,k3) of the sorted quadruple
    for (int64_t c = 0; c < n_cells; c++) {
        for (int l = 0; l < 6; l++) {
            int32_t k[4];
            for (int i = 0; i < 4; i++) {
                k[i] = cell_nodes[8 * c + kLocalFaces[l][i]];
                if (k[i] < 0 || k[i] >= n_nodes) return TITAN_E_TOPOLOGY;
            }
            std::sort(k, k + 4);
            if (k[0] == k[1] || k[1] == k[2] || k[2] == k[3]) return TITAN_E_TOPOLOGY;  // Test_geometr 6
            minnode[6 * c + l] = k[0];
            key_hi[6 * c + l] = ((uint64_t)(uint32_t)k[1] << 42) ^ ((uint64_t)(uint32_t)k[2] << 21) ^ (uint64_t)(uint32_t)k[3];
            // exact comparison is done below on the full quadruple; the key only speeds it up
        }
    }
    std::vector<int64_t> start(n_nodes + 1, 0);
    for (int64_t i = 0; i < nt; i++) start[minnode[i] + 1]++;
    for (int64_t v = 0; v < n_nodes; v++) start[v + 1] += start[v];
    std::vector<int64_t> order(nt);
    {
        std::vector<int64_t> pos(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < nt; i++) order[pos[minnode[i]]++] = i;
    }
    auto same = [&](int64_t a, int64_t b) {
        int32_t ka[4], kb[4];
        for (int i = 0; i < 4; i++) {
            ka[i] = cell_nodes[8 * (a / 6) + kLocalFaces[a % 6][i]];
            kb[i] = cell_nodes[8 * (b / 6) + kLocalFaces[b % 6][i]];
        }
        std::sort(ka, ka + 4);
        std::sort(kb, kb + 4);
        return std::memcmp(ka, kb, sizeof(ka)) == 0;
    };
    std::vector<int64_t> twin(nt, -1);
    for (int64_t v = 0; v < n_nodes; v++) {
        for (int64_t i = start[v]; i < start[v + 1]; i++) {
            const int64_t a = order[i];
            for (int64_t j = i + 1; j < start[v + 1]; j++) {
                const int64_t b = order[j];
                if (key_hi[a] != key_hi[b] || !same(a, b)) continue;
                if (twin[a] >= 0 || twin[b] >= 0) return TITAN_E_TOPOLOGY;  // >2 cells on a face
                twin[a] = b;
                twin[b] = a;
            }
        }
    }
    std::vector<int64_t> gid(nt);
    int64_t nf = 0;
    for (int64_t i = 0; i < nt; i++) {
        if (twin[i] >= 0 && twin[i] < i) { gid[i] = gid[twin[i]]; continue; }
        const int64_t c = i / 6;
        const int l = (int)(i % 6);
        gid[i] = nf;
        for (int j = 0; j < 4; j++) face_nodes[4 * nf + j] = cell_nodes[8 * c + kLocalFaces[l][j]];
        face_cells[2 * nf] = (int32_t)c;
        face_cells[2 * nf + 1] = twin[i] >= 0 ? (int32_t)(twin[i] / 6) : -1;
        nf++;
    }
    for (int64_t c = 0; c < n_cells; c++) {
        int64_t g[6];
        for (int l = 0; l < 6; l++) g[l] = gid[6 * c + l];
        std::sort(g, g + 6);
        for (int l = 0; l < 6; l++) cell_faces[6 * c + l] = (int32_t)g[l];
    }
    return nf;
}

// SplitMix64 (Steele, Lea, Flood 2014): counter-based uniform doubles.
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double uniform01(uint64_t seed, uint64_t index) {
    return (double)(splitmix64(seed * 0x2545F4914F6CDD1Dull + index) >> 11) * (1.0 / 9007199254740992.0);
}

int mesh_cube(int32_t n, double jitter, uint64_t seed, double *xyz, int32_t *cell_nodes) {
    if (n < 1) return TITAN_E_ARG;
    const int64_t m = n + 1;
    const double h = 1.0 / n;
    for (int64_t k = 0; k < m; k++)
        for (int64_t j = 0; j < m; j++)
            for (int64_t i = 0; i < m; i++) {
                const int64_t v = i + m * (j + m * k);
                double p[3] = {i * h, j * h, k * h};
                const bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                if (interior && jitter != 0.0) {
                    for (int d = 0; d < 3; d++) p[d] += jitter * h * (2.0 * uniform01(seed, 3 * v + d) - 1.0);
                }
                xyz[3 * v] = p[0]; xyz[3 * v + 1] = p[1]; xyz[3 * v + 2] = p[2];
            }
    for (int64_t k = 0; k < n; k++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++) {
                const int64_t c = i + n * (j + n * k);
                const int64_t b = i + m * (j + m * k);
                int32_t *cn = cell_nodes + 8 * c;
                cn[0] = (int32_t)b;
                cn[1] = (int32_t)(b + 1);
                cn[2] = (int32_t)(b + 1 + m);
                cn[3] = (int32_t)(b + m);
                cn[4] = (int32_t)(b + m * m);
                cn[5] = (int32_t)(b + 1 + m * m);
                cn[6] = (int32_t)(b + 1 + m + m * m);
                cn[7] = (int32_t)(b + m + m * m);
            }
    return TITAN_OK;
}

// Equiangular cube-sphere: 6 panels of n_edge x n_edge quads on the unit sphere; nodes are
// identified by integer lattice points of the cube surface, so panel edges share nodes.
int mesh_shell(int32_t ne, int32_t nl, const double *radii, double *xyz, int32_t *cell_nodes) {
    if (ne < 1 || nl < 1) return TITAN_E_ARG;
    const int64_t nq = 6LL * ne * ne;
    const int64_t ns = nq + 2;
    std::unordered_map<int64_t, int32_t> id;
    id.reserve((size_t)ns * 2);
    std::vector<double> unit;
    unit.reserve((size_t)ns * 3);
    auto node = [&](int a, int b, int c) -> int32_t {
        const int64_t key = ((int64_t)(a + ne) * (2 * ne + 1) + (b + ne)) * (2 * ne + 1) + (c + ne);
        auto it = id.find(key);
        if (it != id.end()) return it->second;
        const double q = 0.78539816339744830962 / ne;  // pi/4 per lattice unit
        double x = std::tan(q * a), y = std::tan(q * b), z = std::tan(q * c);
        if (a == ne || a == -ne) x = a > 0 ? 1.0 : -1.0;
        if (b == ne || b == -ne) y = b > 0 ? 1.0 : -1.0;
        if (c == ne || c == -ne) z = c > 0 ? 1.0 : -1.0;
        const double r = std::sqrt(x * x + y * y + z * z);
        const int32_t v = (int32_t)(unit.size() / 3);
        unit.push_back(x / r); unit.push_back(y / r); unit.push_back(z / r);
        id.emplace(key, v);
        return v;
    };
    std::vector<int32_t> quads((size_t)nq * 4);
    int64_t q = 0;
    for (int face = 0; face < 6; face++) {
        const int axis = face / 2, sgn = (face % 2) ? 1 : -1;
        for (int j = 0; j < ne; j++)
            for (int i = 0; i < ne; i++) {
                int uv[4][2] = {{i, j}, {i + 1, j}, {i + 1, j + 1}, {i, j + 1}};
                for (int k = 0; k < 4; k++) {
                    // lattice coords in [-ne, ne] with step 2 -> use 2*i - ne
                    int u = 2 * uv[k][0] - ne, w = 2 * uv[k][1] - ne;
                    int L[3];
                    L[axis] = sgn * ne;
                    L[(axis + 1) % 3] = u;
                    L[(axis + 2) % 3] = w;
                    quads[4 * q + k] = node(L[0], L[1], L[2]);
                }
                q++;
            }
    }
    if ((int64_t)unit.size() / 3 != ns) return TITAN_E_TOPOLOGY;
    for (int64_t k = 0; k <= nl; k++)
        for (int64_t v = 0; v < ns; v++)
            for (int d = 0; d < 3; d++) xyz[3 * (k * ns + v) + d] = radii[k] * unit[3 * v + d];
    for (int64_t k = 0; k < nl; k++)
        for (int64_t qq = 0; qq < nq; qq++) {
            int32_t *cn = cell_nodes + 8 * (k * nq + qq);
            for (int i = 0; i < 4; i++) {
                cn[i] = (int32_t)(k * ns + quads[4 * qq + i]);
                cn[4 + i] = (int32_t)((k + 1) * ns + quads[4 * qq + i]);
            }
        }
    return TITAN_OK;
}

static inline uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

void build_order(const MeshView &m, DeviceOrder &o) {
    const int64_t nc = m.n_cells, nf = m.n_faces, nn = m.n_nodes;
    // cell centres (initial) and their bounding box
    std::vector<double> cc(3 * nc);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t c = 0; c < nc; c++) {
        for (int d = 0; d < 3; d++) {
            double s = 0.0;
            for (int i = 0; i < 8; i++) s += m.xyz[3 * (int64_t)m.cell_nodes[8 * c + i] + d];
            cc[3 * c + d] = s / 8;
            lo[d] = std::min(lo[d], cc[3 * c + d]);
            hi[d] = std::max(hi[d], cc[3 * c + d]);
        }
    }
    double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(span > 0)) span = 1.0;
    std::vector<std::pair<uint64_t, int32_t>> key(nc);
    for (int64_t c = 0; c < nc; c++) {
        uint64_t q[3];
        for (int d = 0; d < 3; d++) q[d] = (uint64_t)((cc[3 * c + d] - lo[d]) / span * 2097151.0);
        key[c] = {spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2), (int32_t)c};
    }
    std::sort(key.begin(), key.end());
    o.cell_d2r.resize(nc);
    o.cell_r2d.resize(nc);
    for (int64_t d = 0; d < nc; d++) { o.cell_d2r[d] = key[d].second; o.cell_r2d[key[d].second] = (int32_t)d; }
    // faces by the lower device id of their cells (ties: reference id)
    std::vector<std::pair<int64_t, int32_t>> fk(nf);
    for (int64_t f = 0; f < nf; f++) {
        int64_t a = o.cell_r2d[m.face_cells[2 * f]];
        const int32_t b = m.face_cells[2 * f + 1];
        if (b >= 0) a = std::min<int64_t>(a, o.cell_r2d[b]);
        fk[f] = {a * 8 + 0, (int32_t)f};
    }
    std::stable_sort(fk.begin(), fk.end());
    o.face_d2r.resize(nf);
    o.face_r2d.resize(nf);
    for (int64_t d = 0; d < nf; d++) { o.face_d2r[d] = fk[d].second; o.face_r2d[fk[d].second] = (int32_t)d; }
    // nodes by first use in device cell order
    o.node_d2r.assign(nn, -1);
    o.node_r2d.assign(nn, -1);
    int64_t next = 0;
    for (int64_t d = 0; d < nc; d++) {
        const int64_t c = o.cell_d2r[d];
        for (int i = 0; i < 8; i++) {
            const int32_t v = m.cell_nodes[8 * c + i];
            if (o.node_r2d[v] < 0) { o.node_r2d[v] = (int32_t)next; o.node_d2r[next] = v; next++; }
        }
    }
    for (int64_t v = 0; v < nn; v++)
        if (o.node_r2d[v] < 0) { o.node_r2d[v] = (int32_t)next; o.node_d2r[next] = (int32_t)v; next++; }
}

namespace {

// Which of the hexahedron's three face directions face f is, seen from its cells[0] a: 0 for the
// local faces (0,3,2,1)/(4,5,6,7), 1 for (0,1,5,4)/(2,3,7,6), 2 for (1,2,6,5)/(0,4,7,3) (the
// local face table of Setka::New_connect); 3 if its nodes match none.
int face_direction(const int32_t *cell_nodes, const int32_t *face_nodes, int64_t f, int64_t a) {
    static const int lf[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};
    int32_t fv[4];
    for (int i = 0; i < 4; i++) fv[i] = face_nodes[4 * f + i];
    std::sort(fv, fv + 4);
    for (int j = 0; j < 6; j++) {
        int32_t cv[4];
        for (int i = 0; i < 4; i++) cv[i] = cell_nodes[8 * a + lf[j][i]];
        std::sort(cv, cv + 4);
        if (std::equal(fv, fv + 4, cv)) return j / 2;
    }
    return 3;
}

}  // namespace

bool build_tiles(int64_t nc, int64_t nf, int64_t nn, int tile_cells, int face_cap, const int32_t *cell_nodes,
                 const int32_t *cell_faces, const int32_t *face_nodes, const int32_t *face_cells,
                 const uint32_t *face_code, const uint32_t *cell_code, const int32_t *face_bidx,
                 const int32_t *face_ref, TileData &t) {
    // Face slots of a tile are ordered by the branches their solve takes -- boundary or interior,
    // Riemann method, Zone_1 rescaling or not, face direction -- so that a wavefront's faces
    // mostly take the same path (the direction separates, e.g., the radial faces of a supersonic
    // wind, all upwind, from the lateral ones, which fall inside the wave fan).  The flux sum of
    // a cell reads its faces through cell_lfaces in reference order, so the order of the slots
    // changes no result.
    std::vector<uint32_t> fkey(nf);
    for (int64_t f = 0; f < nf; f++) {
        const int32_t a = face_cells[2 * f], b = face_cells[2 * f + 1];
        const uint32_t bc = face_code[f] & 15u, method = (face_code[f] >> 4) & 15u;
        const bool z1 = ((cell_code[a] >> 8) & 15u) == 1u || (b >= 0 && ((cell_code[b] >> 8) & 15u) == 1u);
        fkey[f] = ((bc != 0u ? 1u : 0u) << 8) | (method << 4) | ((z1 ? 1u : 0u) << 2) |
                  (uint32_t)face_direction(cell_nodes, face_nodes, f, a);
    }
    t = TileData();
    t.tile_cells = tile_cells;
    std::vector<int32_t> nstamp(nn, -1), nlocal(nn, 0), fstamp(nf, -1), flocal(nf, 0);
    // tile boundaries: runs of consecutive cells, closed before a cell that would take the tile past
    // tile_cells cells, face_cap faces or 4095 nodes
    t.cell_off.push_back(0);
    {
        int64_t ti = 0, ncell = 0, nface = 0, nnode = 0;
        for (int64_t c = 0; c < nc; c++) {
            int64_t add_f = 0, add_n = 0;
            for (int j = 0; j < 6; j++) {
                const int32_t f = cell_faces[6 * c + j];
                if (fstamp[f] != ti) add_f++;
            }
            for (int i = 0; i < 8; i++) {
                const int32_t v = cell_nodes[8 * c + i];
                if (nstamp[v] != ti) {
                    bool again = false;  // a node repeated within this cell counts once
                    for (int k = 0; k < i; k++) again |= cell_nodes[8 * c + k] == v;
                    add_n += again ? 0 : 1;
                }
            }
            if (ncell > 0 && (ncell + 1 > tile_cells || nface + add_f > face_cap || nnode + add_n > 4095)) {
                t.cell_off.push_back((int32_t)c);
                ti++;
                ncell = nface = nnode = 0;
                add_f = 6;
                add_n = 0;
                for (int i = 0; i < 8; i++) {
                    bool again = false;
                    for (int k = 0; k < i; k++) again |= cell_nodes[8 * c + k] == cell_nodes[8 * c + i];
                    add_n += again ? 0 : 1;
                }
            }
            for (int j = 0; j < 6; j++) fstamp[cell_faces[6 * c + j]] = (int32_t)ti;
            for (int i = 0; i < 8; i++) nstamp[cell_nodes[8 * c + i]] = (int32_t)ti;
            ncell++;
            nface += add_f;
            nnode += add_n;
        }
        t.cell_off.push_back((int32_t)nc);
        std::fill(nstamp.begin(), nstamp.end(), -1);
        std::fill(fstamp.begin(), fstamp.end(), -1);
    }
    const int64_t nt = (int64_t)t.cell_off.size() - 1;
    t.node_off.assign(nt + 1, 0);
    t.face_off.assign(nt + 1, 0);
    t.own.assign(2 * nt, 0);
    t.cell_lnodes.assign(8 * nc, 0);
    t.cell_lfaces.assign(6 * nc, 0);
    std::vector<int32_t> lst;
    int64_t seen_max = -1;
    for (int64_t ti = 0; ti < nt; ti++) {
        const int64_t c0 = t.cell_off[ti], c1 = t.cell_off[ti + 1];
        t.own[2 * ti] = (int32_t)(seen_max + 1);
        // nodes
        lst.clear();
        for (int64_t c = c0; c < c1; c++)
            for (int i = 0; i < 8; i++) {
                const int32_t v = cell_nodes[8 * c + i];
                if (nstamp[v] != ti) { nstamp[v] = (int32_t)ti; lst.push_back(v); }
                seen_max = std::max<int64_t>(seen_max, v);
            }
        std::sort(lst.begin(), lst.end());
        if (lst.size() > 4095) return false;
        for (size_t i = 0; i < lst.size(); i++) nlocal[lst[i]] = (int32_t)i;
        t.node_off[ti + 1] = t.node_off[ti] + (int32_t)lst.size();
        t.nodes.insert(t.nodes.end(), lst.begin(), lst.end());
        t.max_nodes = std::max<int32_t>(t.max_nodes, (int32_t)lst.size());
        t.own[2 * ti + 1] = (int32_t)(seen_max + 1);
        for (int64_t c = c0; c < c1; c++)
            for (int i = 0; i < 8; i++) t.cell_lnodes[8 * c + i] = (uint16_t)nlocal[cell_nodes[8 * c + i]];
        // faces
        lst.clear();
        for (int64_t c = c0; c < c1; c++)
            for (int j = 0; j < 6; j++) {
                const int32_t f = cell_faces[6 * c + j];
                if (fstamp[f] != ti) { fstamp[f] = (int32_t)ti; lst.push_back(f); }
            }
        std::sort(lst.begin(), lst.end(), [&](int32_t x, int32_t y) {
            return fkey[x] != fkey[y] ? fkey[x] < fkey[y] : x < y;
        });
        for (size_t i = 0; i < lst.size(); i++) flocal[lst[i]] = (int32_t)i;
        t.face_off[ti + 1] = t.face_off[ti] + (int32_t)lst.size();
        t.max_faces = std::max<int32_t>(t.max_faces, (int32_t)lst.size());
        for (int32_t f : lst) {
            int ln[4];
            for (int i = 0; i < 4; i++) {
                const int32_t v = face_nodes[4 * (int64_t)f + i];
                if (nstamp[v] != ti) return false;  // a face node outside the tile's cells
                ln[i] = nlocal[v];
            }
            const int32_t a = face_cells[2 * (int64_t)f];
            int32_t b = face_cells[2 * (int64_t)f + 1];
            if (b < 0) b = face_bidx[f] >= 0 ? -(face_bidx[f] + 2) : -1;
            t.faces.push_back(tile_face_pack(ln, face_code[f]));
            t.faces.push_back((uint64_t)(uint32_t)a | ((uint64_t)(uint32_t)b << 32));
            t.face_ref.push_back(face_ref[f]);
            t.face_dev.push_back(f);
        }
        for (int64_t c = c0; c < c1; c++)
            for (int j = 0; j < 6; j++) t.cell_lfaces[6 * c + j] = (uint16_t)flocal[cell_faces[6 * c + j]];
    }
    return true;
}

namespace {

// Eigen's Vector3d::normalized(): divide by the norm unless it is zero; dot and squared norm summed
// as (x0 y0 + x1 y1) + x2 y2
void tvd_unit(double v[3]) {
    const double n = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (n > 0.0) { v[0] /= n; v[1] /= n; v[2] /= n; }
}

// Among cell C's other neighbours, the one whose centre, seen from the centre of face f, is most
// opposite to vec (smallest cosine, the first one on ties); -1 if C has no other neighbour.
int32_t tvd_pick(const int32_t *cell_faces, const int32_t *face_cells, const double *cc, const double *G,
                 int64_t f, int32_t C, const double vec[3]) {
    double best = 2.0;
    int32_t pick = -1;
    for (int k = 0; k < 6; k++) {
        const int32_t gr = cell_faces[6 * (int64_t)C + k];
        if (gr == f) continue;
        const int32_t a = face_cells[2 * (int64_t)gr], b = face_cells[2 * (int64_t)gr + 1];
        if (b < 0) continue;  // Get_Sosed across a boundary face: none
        const int32_t cel = a == C ? b : a;
        double v[3] = {cc[3 * (int64_t)cel] - G[0], cc[3 * (int64_t)cel + 1] - G[1], cc[3 * (int64_t)cel + 2] - G[2]};
        tvd_unit(v);
        const double cosine = (vec[0] * v[0] + vec[1] * v[1]) + vec[2] * v[2];
        if (cosine < best) { best = cosine; pick = cel; }
    }
    return pick;
}

}  // namespace

void init_tvd(int64_t nf, const int32_t *cell_faces, const int32_t *face_cells, const double *cc,
              const double *fc, const double *fn, int32_t *tvd) {
    for (int64_t f = 0; f < nf; f++) {
        const double *n = fn + 3 * f, *G = fc + 3 * f;
        const int32_t AA = tvd_pick(cell_faces, face_cells, cc, G, f, face_cells[2 * f], n);
        tvd[2 * f] = AA;
        tvd[2 * f + 1] = -1;
        if (face_cells[2 * f + 1] < 0) continue;  // boundary face: one pick, no check
        const double mn[3] = {-n[0], -n[1], -n[2]};
        const int32_t BB = tvd_pick(cell_faces, face_cells, cc, G, f, face_cells[2 * f + 1], mn);
        tvd[2 * f + 1] = BB;
        // a second neighbour that does not lie behind its cell is dropped (Setka_physics.cpp:728-738)
        if (AA >= 0) {
            double k3[3] = {cc[3 * (int64_t)AA] - G[0], cc[3 * (int64_t)AA + 1] - G[1], cc[3 * (int64_t)AA + 2] - G[2]};
            tvd_unit(k3);
            if (-((n[0] * k3[0] + n[1] * k3[1]) + n[2] * k3[2]) < 0.1) tvd[2 * f] = -1;
        }
        if (BB >= 0) {
            double k3[3] = {cc[3 * (int64_t)BB] - G[0], cc[3 * (int64_t)BB + 1] - G[1], cc[3 * (int64_t)BB + 2] - G[2]};
            tvd_unit(k3);
            if (((n[0] * k3[0] + n[1] * k3[1]) + n[2] * k3[2]) < 0.1) tvd[2 * f + 1] = -1;
        }
    }
}

}  // namespace titan

extern "C" int64_t titan_mesh_connect(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes,
                                      int32_t *face_nodes, int32_t *face_cells, int32_t *cell_faces) {
    if (n_cells <= 0 || !cell_nodes || !face_nodes || !face_cells || !cell_faces) return TITAN_E_ARG;
    return titan::connect_faces(n_cells, cell_nodes, n_nodes, face_nodes, face_cells, cell_faces);
}

extern "C" int titan_mesh_cube(int32_t n, double jitter, uint64_t seed, double *xyz, int32_t *cell_nodes) {
    if (!xyz || !cell_nodes) return TITAN_E_ARG;
    return titan::mesh_cube(n, jitter, seed, xyz, cell_nodes);
}

extern "C" int titan_mesh_shell(int32_t n_edge, int32_t n_layers, const double *radii, double *xyz,
                                int32_t *cell_nodes) {
    if (!radii || !xyz || !cell_nodes) return TITAN_E_ARG;
    return titan::mesh_shell(n_edge, n_layers, radii, xyz, cell_nodes);
}
