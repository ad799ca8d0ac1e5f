// titan_mesh.h -- host-side mesh indexing helpers.
#pragma once
#include <cstdint>
#include <vector>

namespace titan {

struct MeshView {
    int64_t n_nodes, n_cells, n_faces;
    const int32_t *cell_nodes;
    const int32_t *face_cells;
    const double *xyz;
};

// Permutations between reference numbering (r) and device numbering (d).
struct DeviceOrder {
    std::vector<int32_t> cell_d2r, cell_r2d, face_d2r, face_r2d, node_d2r, node_r2d;
};

int64_t connect_faces(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes, int32_t *face_nodes,
                      int32_t *face_cells, int32_t *cell_faces);
int mesh_cube(int32_t n, double jitter, uint64_t seed, double *xyz, int32_t *cell_nodes);
int mesh_shell(int32_t n_edge, int32_t n_layers, const double *radii, double *xyz, int32_t *cell_nodes);
double uniform01(uint64_t seed, uint64_t index);
// Cells along a Morton curve of their centres, faces by their lower device cell, nodes by first use.
void build_order(const MeshView &m, DeviceOrder &o);

// Host image of the tiles of the fused step kernel (see TileMesh in device_mesh.h).
struct TileData {
    int32_t tile_cells = 0, max_nodes = 0, max_faces = 0;
    std::vector<int32_t> cell_off, node_off, nodes, own, face_off, face_ref, face_dev;  // face_dev: device face id
    std::vector<uint64_t> faces;  // 2 words per record
    std::vector<uint16_t> cell_lnodes, cell_lfaces;
};
// Tiles are runs of consecutive device cells of at most tile_cells cells and face_cap faces (and
// 4095 nodes).  All arrays in device numbering; face_code/face_bidx as uploaded; returns false if a
// tile's face has a node outside the tile's cells.
bool build_tiles(int64_t nc, int64_t nf, int64_t nn, int tile_cells, int face_cap, const int32_t *cell_nodes,
                 const int32_t *cell_faces, const int32_t *face_nodes, const int32_t *face_cells,
                 const uint32_t *face_code, const uint32_t *cell_code, const int32_t *face_bidx,
                 const int32_t *face_ref, TileData &t);

// Second neighbours of every face (Setka::Init_TVD, Setka_physics.cpp:648-744) from the geometry of
// the initial nodes, all in device numbering: tvd[2f] = AA behind cells[0], tvd[2f+1] = BB behind
// cells[1], -1 = none.  cell_faces lists each cell's faces in ascending reference face id, the
// order of Cell::grans, so ties go to the same face as in the reference.
void init_tvd(int64_t nf, const int32_t *cell_faces, const int32_t *face_cells, const double *cell_center,
              const double *face_center, const double *face_normal, int32_t *tvd);

}  // namespace titan
