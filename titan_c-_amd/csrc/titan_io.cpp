// titan_io.cpp -- TITAN's cell-parameter checkpoint file (Setka::Save_cell_parameters /
// Setka::Download_cell_parameters, Setka_physics.cpp:4107-4285).
//
// Byte layout (little-endian, the reference's x86-64 build: size_t = 8 bytes, bool = 1 byte):
//   for every cell in All_Cell order:
//     u64 count; count x { u64 key_size; key_size bytes of key; f64 value }
//   u8 nodes_follow (1 from Save_cell_parameters);
//   if nodes_follow: u64 n_nodes; n_nodes x { f64 x, y, z }   (Yzel::coord[0])
//   1000 x u8 0                                                 (room the reference reserves)
// The reference iterates an unordered_map, so the key order inside a cell is unspecified; the
// reader here matches keys by name and accepts any order, and expects each cell's key list to be
// mostly the same as the previous cell's (it is: every cell holds the same keys), which it
// exploits to match a key with one comparison.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/titan_gpu.h"

namespace {

constexpr uint64_t kMaxKey = 4096;   // longer keys mean a damaged file (the reference's are < 32)
constexpr int kTrailer = 1000;       // bool false bytes after the node block

struct Reader {
    const uint8_t *p, *end;
    bool ok = true;
    bool take(void *dst, size_t n) {
        if (!ok || (size_t)(end - p) < n) { ok = false; return false; }
        std::memcpy(dst, p, n);
        p += n;
        return true;
    }
    bool skip(size_t n) {
        if (!ok || (size_t)(end - p) < n) { ok = false; return false; }
        p += n;
        return true;
    }
};

int load_file(const char *path, std::vector<uint8_t> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return TITAN_E_IO;
    if (std::fseek(f, 0, SEEK_END) != 0) { std::fclose(f); return TITAN_E_IO; }
    const long sz = std::ftell(f);
    if (sz < 0 || std::fseek(f, 0, SEEK_SET) != 0) { std::fclose(f); return TITAN_E_IO; }
    buf.resize((size_t)sz);
    const size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return got == (size_t)sz ? TITAN_OK : TITAN_E_IO;
}

}  // namespace

extern "C" int titan_checkpoint_write(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                                      const double *values, int64_t n_nodes, const double *xyz) {
    if (!path || n_cells < 0 || n_keys < 0 || (n_keys > 0 && (!keys || !values)) || (n_nodes > 0 && !xyz))
        return TITAN_E_ARG;
    std::vector<uint64_t> ksz(n_keys);
    size_t cell_bytes = 8;
    for (int k = 0; k < n_keys; k++) {
        if (!keys[k]) return TITAN_E_ARG;
        ksz[k] = std::strlen(keys[k]);
        if (ksz[k] == 0 || ksz[k] > kMaxKey) return TITAN_E_ARG;
        cell_bytes += 16 + ksz[k];
    }
    // one cell's record with the key names in place; only the values change from cell to cell
    std::vector<uint8_t> rec(cell_bytes);
    std::vector<size_t> voff(n_keys);
    {
        uint8_t *q = rec.data();
        const uint64_t cnt = (uint64_t)n_keys;
        std::memcpy(q, &cnt, 8); q += 8;
        for (int k = 0; k < n_keys; k++) {
            std::memcpy(q, &ksz[k], 8); q += 8;
            std::memcpy(q, keys[k], ksz[k]); q += ksz[k];
            voff[k] = (size_t)(q - rec.data());
            q += 8;
        }
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return TITAN_E_IO;
    std::vector<uint8_t> chunk;
    const int64_t per_chunk = std::max<int64_t>(1, (int64_t)((8u << 20) / cell_bytes));
    chunk.reserve((size_t)per_chunk * cell_bytes);
    bool ok = true;
    for (int64_t c0 = 0; c0 < n_cells && ok; c0 += per_chunk) {
        const int64_t c1 = std::min(n_cells, c0 + per_chunk);
        chunk.clear();
        for (int64_t c = c0; c < c1; c++) {
            for (int k = 0; k < n_keys; k++) std::memcpy(&rec[voff[k]], &values[c * n_keys + k], 8);
            chunk.insert(chunk.end(), rec.begin(), rec.end());
        }
        ok = std::fwrite(chunk.data(), 1, chunk.size(), f) == chunk.size();
    }
    // node block (Setka_physics.cpp:4136-4150)
    const uint8_t yes = 1, no = 0;
    if (ok && n_nodes >= 0) {
        const uint64_t nn = (uint64_t)n_nodes;
        ok = std::fwrite(&yes, 1, 1, f) == 1 && std::fwrite(&nn, 8, 1, f) == 1 &&
             (n_nodes == 0 || std::fwrite(xyz, 24, (size_t)n_nodes, f) == (size_t)n_nodes);
    } else if (ok) {
        ok = std::fwrite(&no, 1, 1, f) == 1;
    }
    // Setka_physics.cpp:4155-4160
    if (ok) {
        const std::vector<uint8_t> zeros(kTrailer, 0);
        ok = std::fwrite(zeros.data(), 1, zeros.size(), f) == zeros.size();
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? TITAN_OK : TITAN_E_IO;
}

extern "C" int titan_checkpoint_read(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                                     double *values, int64_t n_nodes, double *xyz, int64_t *nodes_read) {
    if (!path || n_cells < 0 || n_keys < 0 || (n_keys > 0 && (!keys || !values)) || (n_nodes > 0 && !xyz) ||
        (n_nodes < 0 && xyz))
        return TITAN_E_ARG;
    if (nodes_read) *nodes_read = -1;
    std::unordered_map<std::string, int> want;
    for (int k = 0; k < n_keys; k++) {
        if (!keys[k]) return TITAN_E_ARG;
        want.emplace(keys[k], k);
    }
    std::vector<uint8_t> buf;
    int rc = load_file(path, buf);
    if (rc != TITAN_OK) return rc;
    Reader r{buf.data(), buf.data() + buf.size()};
    const double nan = std::nan("");
    for (int64_t i = 0; i < n_cells * n_keys; i++) values[i] = nan;
    // key positions of the previous cell: key j of a cell is usually key j of the one before
    std::vector<std::pair<std::string, int>> last;
    for (int64_t c = 0; c < n_cells; c++) {
        uint64_t cnt = 0;
        if (!r.take(&cnt, 8)) return TITAN_E_IO;
        if (cnt > (uint64_t)(r.end - r.p) / 16) return TITAN_E_IO;  // each pair takes >= 16 bytes
        if (last.size() < cnt) last.resize(cnt, {std::string(), -2});
        for (uint64_t j = 0; j < cnt; j++) {
            uint64_t ks = 0;
            if (!r.take(&ks, 8) || ks > kMaxKey) return TITAN_E_IO;
            const char *kp = reinterpret_cast<const char *>(r.p);
            if (!r.skip(ks)) return TITAN_E_IO;
            double v;
            if (!r.take(&v, 8)) return TITAN_E_IO;
            auto &hint = last[j];
            int slot;
            if (hint.second != -2 && hint.first.size() == ks && std::memcmp(hint.first.data(), kp, ks) == 0) {
                slot = hint.second;
            } else {
                hint.first.assign(kp, ks);
                auto it = want.find(hint.first);
                slot = hint.second = it == want.end() ? -1 : it->second;
            }
            if (slot >= 0) values[c * n_keys + slot] = v;  // Download_cell_parameters: last write wins
        }
    }
    // node block (Setka_physics.cpp:4257-4282)
    uint8_t follow = 0;
    if (!r.take(&follow, 1)) return TITAN_E_IO;
    if (follow) {
        uint64_t nn = 0;
        if (!r.take(&nn, 8)) return TITAN_E_IO;
        if (nn > (uint64_t)(r.end - r.p) / 24) return TITAN_E_IO;
        if (n_nodes >= 0 && nn != (uint64_t)n_nodes) return TITAN_E_ARG;  // another mesh
        if (xyz && !r.take(xyz, 24 * (size_t)nn)) return TITAN_E_IO;
        if (nodes_read) *nodes_read = (int64_t)nn;
    } else if (nodes_read) {
        *nodes_read = 0;
    }
    return TITAN_OK;
}

// TITAN's Monte-Carlo parameter file (Setka::Save_cell_MK_parameters /
// Setka::Download_cell_MK_parameters, Setka_physics.cpp:4165-4209, 4287-4336): u64 key count, the
// keys (u64 length, bytes) once, then per cell in All_Cell order one f64 per key in that order,
// then 1000 zero bytes.
extern "C" int titan_mk_parameters_write(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                                         const double *values) {
    if (!path || n_cells < 0 || n_keys < 0 || (n_keys > 0 && (!keys || !values))) return TITAN_E_ARG;
    std::vector<uint8_t> head(8);
    const uint64_t cnt = (uint64_t)n_keys;
    std::memcpy(head.data(), &cnt, 8);
    for (int k = 0; k < n_keys; k++) {
        if (!keys[k]) return TITAN_E_ARG;
        const uint64_t ks = std::strlen(keys[k]);
        if (ks == 0 || ks > kMaxKey) return TITAN_E_ARG;
        const size_t at = head.size();
        head.resize(at + 8 + ks);
        std::memcpy(&head[at], &ks, 8);
        std::memcpy(&head[at + 8], keys[k], ks);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return TITAN_E_IO;
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    if (ok && n_cells > 0 && n_keys > 0)
        ok = std::fwrite(values, sizeof(double) * (size_t)n_keys, (size_t)n_cells, f) == (size_t)n_cells;
    if (ok) {
        const std::vector<uint8_t> zeros(kTrailer, 0);
        ok = std::fwrite(zeros.data(), 1, zeros.size(), f) == zeros.size();
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? TITAN_OK : TITAN_E_IO;
}

extern "C" int titan_mk_parameters_read(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                                        double *values, const int8_t *cell_mk, int32_t zone_except,
                                        int32_t *keys_found) {
    if (!path || n_cells < 0 || n_keys < 0 || (n_keys > 0 && (!keys || !values))) return TITAN_E_ARG;
    std::unordered_map<std::string, int> want;
    for (int k = 0; k < n_keys; k++) {
        if (!keys[k]) return TITAN_E_ARG;
        want.emplace(keys[k], k);
    }
    std::vector<uint8_t> buf;
    int rc = load_file(path, buf);
    if (rc != TITAN_OK) return rc;
    Reader r{buf.data(), buf.data() + buf.size()};
    uint64_t cnt = 0;
    if (!r.take(&cnt, 8) || cnt > (uint64_t)(r.end - r.p) / 8) return TITAN_E_IO;
    std::vector<int> slot(cnt, -1);
    int found = 0;
    for (uint64_t j = 0; j < cnt; j++) {
        uint64_t ks = 0;
        if (!r.take(&ks, 8) || ks > kMaxKey) return TITAN_E_IO;
        const char *kp = reinterpret_cast<const char *>(r.p);
        if (!r.skip(ks)) return TITAN_E_IO;
        auto it = want.find(std::string(kp, ks));
        if (it != want.end()) {
            slot[j] = it->second;
            found++;
        }
    }
    // the file does not state its cell count: the values and the trailer must fill it exactly
    const uint64_t left = (uint64_t)(r.end - r.p);
    if (cnt > 0 && left / (8 * cnt) < (uint64_t)n_cells) return TITAN_E_IO;
    if (left != 8 * cnt * (uint64_t)n_cells + kTrailer) return TITAN_E_ARG;
    for (int64_t c = 0; c < n_cells; c++) {
        const bool keep = cell_mk && cell_mk[c] == zone_except;  // the zone being computed keeps its values
        for (uint64_t j = 0; j < cnt; j++) {
            double v;
            r.take(&v, 8);
            if (!keep && slot[j] >= 0) values[c * n_keys + slot[j]] = v;
        }
    }
    if (keys_found) *keys_found = found;
    return TITAN_OK;
}

// One cell's pickup-ion source file of Cell::write_S_ToFile / Cell::read_S_FromFile (Cell.cpp:89-125,
// 374-420): int32 rows (2), int32 cols (n_w), int32 size (n_w), S- [n_w], then S+ as the
// column-major (2, n_w) matrix, i.e. [n_w][2].
extern "C" int titan_mk_s_write(const char *path, int32_t n_w, const double *sm, const double *sp) {
    if (!path || n_w < 1 || !sm || !sp) return TITAN_E_ARG;
    FILE *f = std::fopen(path, "wb");
    if (!f) return TITAN_E_IO;
    const int32_t head[3] = {2, n_w, n_w};
    bool ok = std::fwrite(head, sizeof(int32_t), 3, f) == 3;
    if (ok) ok = std::fwrite(sm, sizeof(double), (size_t)n_w, f) == (size_t)n_w;
    if (ok) ok = std::fwrite(sp, sizeof(double), 2 * (size_t)n_w, f) == 2 * (size_t)n_w;
    if (std::fclose(f) != 0) ok = false;
    return ok ? TITAN_OK : TITAN_E_IO;
}

extern "C" int titan_mk_s_read(const char *path, int32_t n_w, double *sm, double *sp) {
    if (!path || n_w < 1 || !sm || !sp) return TITAN_E_ARG;
    std::vector<uint8_t> buf;
    int rc = load_file(path, buf);
    if (rc != TITAN_OK) return rc;
    Reader r{buf.data(), buf.data() + buf.size()};
    int32_t head[3];
    if (!r.take(head, sizeof(head))) return TITAN_E_IO;
    if (head[0] != 2 || head[1] != n_w || head[2] != n_w) return TITAN_E_ARG;  // other sizes (read_S_FromFile exits)
    if (!r.take(sm, sizeof(double) * (size_t)n_w) || !r.take(sp, sizeof(double) * 2 * (size_t)n_w)) return TITAN_E_IO;
    return TITAN_OK;
}
